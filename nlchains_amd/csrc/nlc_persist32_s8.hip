// FP32-compensated persistent step kernels of ragged and mixed FPUT / Toda
// ensembles with 8 (and 4) sites per lane (see nlc_step_wave.inl).
#define NLC_PERSIST_S 8
#define NLC_PERSIST_C32 1
#define NLC_PERSIST_LOOKUP lookup_persistent32_s8
#include "nlc_step_wave.inl"
