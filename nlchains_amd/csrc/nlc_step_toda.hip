// Wave-resident step kernels, model TODA (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::TODA
#define NLC_STEP_LOOKUP lookup_step_toda
#include "nlc_step_wave.inl"
