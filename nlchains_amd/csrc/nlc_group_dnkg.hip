// Workgroup-resident step kernels, model DNKG (see nlc_step_group.inl).
#define NLC_STEP_MODEL ::nlc::DNKG
#define NLC_GROUP_LOOKUP lookup_group_dnkg
#include "nlc_step_group.inl"
