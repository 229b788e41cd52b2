// Workgroup-resident step kernels, model dDNKG (see nlc_step_group.inl).
#define NLC_STEP_MODEL ::nlc::DDNKG
#define NLC_GROUP_LOOKUP lookup_group_ddnkg
#include "nlc_step_group.inl"
