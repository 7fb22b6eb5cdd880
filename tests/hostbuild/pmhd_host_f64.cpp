// Host test build: the FP64 step (same flags as the HIP FP64 unit).
#define PMHD_TU_NS f64
#include "pmhd_host.h"

PMHD_DEFINE_FACTORIES(f64, double, pmhd::HostBackend)
