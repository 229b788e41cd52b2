// Persistent step kernels of ragged and mixed ensembles with 8 (and 4)
// sites per lane (see nlc_step_wave.inl).
#define NLC_PERSIST_S 8
#define NLC_PERSIST_LOOKUP lookup_persistent_s8
#include "nlc_step_wave.inl"
