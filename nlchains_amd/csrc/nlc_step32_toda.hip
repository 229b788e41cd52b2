// FP32-compensated wave-resident step kernels, model Toda (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::TODA
#define NLC_STEP_LOOKUP32 lookup_step32_toda
#include "nlc_step_wave.inl"
