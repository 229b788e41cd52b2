// Wave-resident step kernels, model Toda: 16-lane shapes with ragged lanes for
// chains of 65..512 sites (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::TODA
#define NLC_STEP_LOOKUP16 lookup_step16_toda
#include "nlc_step_wave.inl"
