// Wave-resident step kernels, model DNKG (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::DNKG
#define NLC_STEP_LOOKUP lookup_step_dnkg
#include "nlc_step_wave.inl"
