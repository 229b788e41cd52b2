// FP32-compensated wave-resident step kernels, model FPUT (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::FPUT
#define NLC_STEP_LOOKUP32 lookup_step32_fput
#include "nlc_step_wave.inl"
