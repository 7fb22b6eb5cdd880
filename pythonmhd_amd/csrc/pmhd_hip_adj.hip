// Adjoint handles of libpmhd.so (FP64 and FP32), with IEEE semantics throughout:
// each phase's adjoint evaluates the phase's kernel body on dual numbers, and the
// contributions add up in a fixed order (pmhd_hip_adjdet.h).
#define PMHD_TU_NS adj
// the adjoint differentiates the Roe step only (its cycle instantiates no other solver)
#define PMHD_ROE_ONLY
#include "pmhd_hip_adjdet.h"

PMHD_DEFINE_ADJOINT_FACTORIES(pmhd::HipAdjBackend)
