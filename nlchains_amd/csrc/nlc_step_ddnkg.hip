// Wave-resident step kernels, model DDNKG (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::DDNKG
#define NLC_STEP_LOOKUP lookup_step_ddnkg
#include "nlc_step_wave.inl"
