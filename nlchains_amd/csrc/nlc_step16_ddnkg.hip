// Wave-resident step kernels, model dDNKG: 16-lane shapes with ragged lanes for
// chains of 65..512 sites, and the temporally blocked kernel for long chains
// (see nlc_step_wave.inl).
#define NLC_STEP_MODEL ::nlc::DDNKG
#define NLC_STEP_LOOKUP16 lookup_step16_ddnkg
#define NLC_STEP_TILED lookup_tiled_ddnkg
#include "nlc_step_wave.inl"
