// FP32-bulk step of libpmhd.so (single grids and batches).  Compiled for speed:
// FP32 multiply-adds are contracted, FP32 division and square root use the
// hardware's approximate instructions (a few ulp), the characteristic tracing and
// slope limiter use reassociated projections and maxima/minima are single
// instructions (PMHD_FAST_ORDER); the double-precision wave-speed and
// eigenvector-normalisation islands keep IEEE division and square root.  The
// face-centred field is kept as high and low float parts of a double value and
// updated in double (div B to double rounding); consumers read the high part.
#define PMHD_TU_NS f32
#define PMHD_FACES_SPLIT
#define PMHD_FAST_ORDER
#include "pmhd_hip.h"

PMHD_DEFINE_FACTORIES(f32, float, pmhd::HipBackend)
