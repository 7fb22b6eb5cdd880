// Host test build: the FP32-bulk step (split faces, as in the HIP unit).
#define PMHD_TU_NS f32
#define PMHD_FACES_SPLIT
#define PMHD_FAST_ORDER
#include "pmhd_host.h"

PMHD_DEFINE_FACTORIES(f32, float, pmhd::HostBackend)

// the HIP library's non-temporal FP32 unit differs only in its store instructions:
// on the host it is the same handle
pmhd_handle* pmhd_new_handle_f32nt() { return pmhd_new_handle_f32(); }
