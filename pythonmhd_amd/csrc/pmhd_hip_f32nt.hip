// FP32-bulk single-grid step of libpmhd.so for large grids: the FP32 unit's code
// (same flags and arithmetic, pmhd_hip_f32.hip) with every launch's results stored
// by non-temporal stores (PMHD_NT_STORES): on grids whose outputs are far larger
// than the caches, streaming them past the L2 leaves it to the neighbour re-reads
// of each launch's inputs (the constrained-transport faces, which the next launch
// reads at once, keep ordinary stores).  pmhd_create picks this unit by grid size.
#define PMHD_TU_NS f32nt
#define PMHD_FACES_SPLIT
#define PMHD_FAST_ORDER
#define PMHD_NT_STORES
#include "pmhd_hip.h"

pmhd_handle* pmhd_new_handle_f32nt() { return new (std::nothrow) pmhd::Handle<float, pmhd::HipBackend>(); }
