// FP64 step of libpmhd.so (single grids and batches).  Compiled with
// -ffp-contract=off and IEEE division and square root, so that every expression
// keeps the reference's operation order and rounding (bit-exact with NumPy).
#define PMHD_TU_NS f64
#include "pmhd_hip.h"

PMHD_DEFINE_FACTORIES(f64, double, pmhd::HipBackend)
