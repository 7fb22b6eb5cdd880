// Host test build: the FP32-bulk step.
#define PMHD_TU_NS f32
#include "pmhd_host.h"

PMHD_DEFINE_FACTORIES(f32, float, pmhd::HostBackend)
