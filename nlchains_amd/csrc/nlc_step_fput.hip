// Wave-resident step kernels, model FPUT (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::FPUT
#define NLC_STEP_LOOKUP lookup_step_fput
#include "nlc_step_wave.inl"
