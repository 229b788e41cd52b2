// Workgroup-resident step kernels, model FPUT (see nlc_step_group.inl).
#define NLC_STEP_MODEL ::nlc::FPUT
#define NLC_GROUP_LOOKUP lookup_group_fput
#include "nlc_step_group.inl"
