// Wave-resident step kernels, model dDNKG: 16-lane shapes with ragged lanes for
// chains of 65..512 sites (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::DDNKG
#define NLC_STEP_LOOKUP16 lookup_step16_ddnkg
#include "nlc_step_wave.inl"
