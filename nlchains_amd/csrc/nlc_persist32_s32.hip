// FP32-compensated persistent step kernels of ragged and mixed FPUT / Toda
// ensembles with 32 (and 16) sites per lane (see nlc_step_wave.inl).
#define NLC_PERSIST_S 32
#define NLC_PERSIST_C32 1
#define NLC_PERSIST_LOOKUP lookup_persistent32_s32
#include "nlc_step_wave.inl"
