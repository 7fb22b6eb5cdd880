// The C ABI of libpmhd.so (include/pmhd.h): argument checks and dispatch to the
// handles of the requested precision, which the step and adjoint units construct.
// Host code only; shared by the HIP library and the host test build.
#define PMHD_TU_NS abi
#include "pmhd_adjoint.h"

PMHD_DEFINE_C_ABI()
PMHD_DEFINE_ADJOINT_ABI()
