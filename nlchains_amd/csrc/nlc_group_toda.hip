// Workgroup-resident step kernels, model Toda (see nlc_step_group.inl).
#define NLC_STEP_MODEL ::nlc::TODA
#define NLC_GROUP_LOOKUP lookup_group_toda
#include "nlc_step_group.inl"
