// Host test build: the adjoint handles.
#define PMHD_TU_NS adj
#define PMHD_ROE_ONLY
#include "pmhd_host.h"

PMHD_DEFINE_ADJOINT_FACTORIES(pmhd::HostBackend)
