// pmhd_cfl_dt_noalloc (include/pmhd.h): the CFL timestep of pmhd_compute_dt for a
// state held in the caller's device arrays, with no allocation -- for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops).
// Its only device memory beyond the inputs and the output is `scratch`, a
// file-scope array of 24 KiB (1024 blocks' partial maxima) that every GPU the
// library runs on holds until the process ends.  Compiled like the FP64 unit
// (IEEE, no contraction), so the timestep equals an FP64 handle's bit for bit.
#define PMHD_TU_NS noalloc
#include <mutex>

#include "pmhd_hip.h"

namespace {
constexpr int kNoallocBlocks = 1024;                   // blocks of the grid-stride maxima launch
constexpr int kScratchWords = 3 * kNoallocBlocks;      // three maxima per block
}  // namespace

// Shared by every call on one GPU: calls that overlap on one GPU race.  Host code
// never uses this name as an address; it uses the per-device address below.
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per
// device, with that device current); nullptr and a message on failure
unsigned long long* scratch_on_current_device(const char** err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    *err = "pmhd_cfl_dt_noalloc: no current HIP device";
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    *err = "pmhd_cfl_dt_noalloc: device index beyond 63";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      *err = "pmhd_cfl_dt_noalloc: hipGetSymbolAddress of the scratch array failed";
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<unsigned long long*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

PMHD_NS_BEGIN

// per-block maxima of the three CFL speeds over a grid-stride subset of the cells
// (the speeds of CflSpeeds, ordered as IEEE bit patterns like k_phase_max)
__global__ void __launch_bounds__(256) k_cfl_stride(const Ctx<double> c, unsigned long long* part) {
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  const int64_t n = c.ncell();
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < n; t += int64_t(gridDim.x) * 256) {
    double o[3];
    CflSpeeds()(c, unflat(t, c.ny, c.nx, c.nz), o);
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], speed_bits(o[d]));
  }
  block_max3(m);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) part[3 * blockIdx.x + d] = m[d];
  }
}

// one block: the maxima of the nparts partials, then the timestep into *dt
__global__ void __launch_bounds__(256) k_dt_final(const unsigned long long* part, int nparts, const Ctx<double> c,
                                                  double cfl, double* dt) {
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  for (int b = int(threadIdx.x); b < nparts; b += 256) {
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], part[3 * b + d]);
  }
  block_max3(m);
  if (threadIdx.x == 0) {
    double mx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) mx[d] = __longlong_as_double((long long)m[d]);
    *dt = dt_of(c, cfl, mx);
  }
}

PMHD_NS_END

extern "C" int pmhd_cfl_dt_noalloc(const pmhd_grid* grid, const pmhd_params* params, const double* cons,
                                   const double* fbx, const double* fby, const double* fbz, double* dt,
                                   void* stream) {
  using namespace pmhd;
  if (!grid || !params || !cons || !dt) {
    pmhd_note_null_error("pmhd_cfl_dt_noalloc: grid, params, cons and dt are required");
    return PMHD_EINVAL;
  }
  const char* problem = grid_problem(grid);
  if (!problem) problem = params_problem(params);
  if (!problem && params->precision != 64)
    problem = "pmhd_cfl_dt_noalloc computes in FP64 from float64 arrays: precision must be 64";
  const bool mhd = params->physics == 0;
  if (!problem && mhd && grid->ndim >= 2 && (!fbx || !fby || (grid->ndim == 3 && !fbz)))
    problem = "pmhd_cfl_dt_noalloc: the face arrays of the grid's dimensions are required";
  if (problem) {
    pmhd_note_null_error(problem);
    return PMHD_EINVAL;
  }
  unsigned long long* part = scratch_on_current_device(&problem);
  if (!part) {
    pmhd_note_null_error(problem);
    return PMHD_EHIP;
  }
  Ctx<double> c;
  init_ctx(c, normal_grid(*grid), *params);
  c.U = cons;
  c.fb[0] = fbx;
  c.fb[1] = fby;
  c.fb[2] = fbz;
  const int64_t n = c.ncell();
  const int64_t want = (n + 255) / 256;
  const int nb = int(want < kNoallocBlocks ? want : kNoallocBlocks);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_cfl_stride, dim3(unsigned(nb)), dim3(256), 0, s, c, part);
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, s, part, nb, c, params->cfl, dt);
  if (hipGetLastError() != hipSuccess) {
    pmhd_note_null_error("pmhd_cfl_dt_noalloc: kernel launch failed");
    return PMHD_EHIP;
  }
  return PMHD_OK;
}
