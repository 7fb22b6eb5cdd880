// HIP backend of libpmhd.so for gfx950 (MI355X): one thread per element, wave64
// shuffle reductions for the CFL maximum, one HIP stream per handle.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pmhd_api.h"

namespace pmhd {

// one thread per element of an (n0, n1, n2) box: x covers the contiguous axis
template <class F>
__global__ void __launch_bounds__(256) k_each3(int n0, int n1, int n2, F f) {
  const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = int(blockIdx.y * blockDim.y + threadIdx.y);
  const int j = int(blockIdx.z);
  if (k < n2 && i < n1) f(I3{j, i, k});
  (void)n0;
}

__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}

// Maximum of three non-negative per-element values.  The IEEE bit patterns of
// non-negative doubles order like the values (NaN above +inf, as np.max does), so
// the reduction is an integer max: deterministic in any order.
template <class F>
__global__ void __launch_bounds__(256) k_max3(int64_t n, F f, unsigned long long* out) {
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    double o[3];
    f(i, o);
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], (unsigned long long)__double_as_longlong(o[d]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], __shfl_xor(m[d], off, 64));
  }
  __shared__ unsigned long long sm[3][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) sm[d][wv] = m[d];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      unsigned long long v = umax64(umax64(sm[d][0], sm[d][1]), umax64(sm[d][2], sm[d][3]));
      atomicMax(&out[d], v);
    }
  }
}

// Deterministic two-level reduction of 9 sums and 2 maxima in double: each
// thread accumulates a fixed grid-stride subset, the block combines in a fixed
// tree, and the host adds the per-block partials in block order.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;
template <class F>
__global__ void __launch_bounds__(kRedThreads) k_reduce(int64_t n, F f, double* part) {
  double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double m[2] = {0, 0};
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    double a[9], b[2];
    f(i, a, b);
#pragma unroll
    for (int q = 0; q < 9; ++q) s[q] += a[q];
    m[0] = npmax(m[0], b[0]);
    m[1] = npmax(m[1], b[1]);
  }
  __shared__ double sh[11][kRedThreads];
#pragma unroll
  for (int q = 0; q < 9; ++q) sh[q][threadIdx.x] = s[q];
  sh[9][threadIdx.x] = m[0];
  sh[10][threadIdx.x] = m[1];
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
#pragma unroll
      for (int q = 0; q < 9; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
      sh[9][threadIdx.x] = npmax(sh[9][threadIdx.x], sh[9][threadIdx.x + w]);
      sh[10][threadIdx.x] = npmax(sh[10][threadIdx.x], sh[10][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 11; ++q) part[blockIdx.x * 11 + q] = sh[q][0];
  }
}

struct HipExec {
  hipStream_t s = nullptr;
  unsigned long long* red = nullptr;
  hipError_t last = hipSuccess;
  // optional per-launch timing with events on the launch stream
  bool timing = false;
  struct Rec {
    int id;
    int64_t n;
    hipEvent_t a, b;
  };
  std::vector<Rec> pend;
  std::vector<hipEvent_t> pool;
  double tot_ms[K_COUNT] = {};
  int64_t launches[K_COUNT] = {};
  int64_t elems[K_COUNT] = {};

  void note(hipError_t e) {
    if (e != hipSuccess && last == hipSuccess) last = e;
  }
  hipEvent_t event() {
    hipEvent_t e = nullptr;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else {
      note(hipEventCreate(&e));
    }
    return e;
  }
  void collect() {
    if (pend.empty()) return;
    note(hipStreamSynchronize(s));
    for (const Rec& r : pend) {
      float ms = 0.f;
      note(hipEventElapsedTime(&ms, r.a, r.b));
      tot_ms[r.id] += ms;
      launches[r.id] += 1;
      elems[r.id] += r.n;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pend.clear();
  }
  void reset_profile() {
    collect();
    for (int k = 0; k < K_COUNT; ++k) {
      tot_ms[k] = 0;
      launches[k] = 0;
      elems[k] = 0;
    }
  }
  void destroy_events() {
    collect();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }

  template <class L>
  void timed(int id, int64_t n, L launch) {
    if (!timing) {
      launch();
      return;
    }
    if (pend.size() > 4096) collect();
    hipEvent_t a = event(), b = event();
    note(hipEventRecord(a, s));
    launch();
    note(hipEventRecord(b, s));
    pend.push_back(Rec{id, n, a, b});
  }

  template <class F>
  void run3(int id, int64_t n0, int64_t n1, int64_t n2, F f) {
    if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
    unsigned bx = 1;
    while (bx < 64 && int64_t(bx) < n2) bx <<= 1;
    const unsigned by = 256 / bx;
    dim3 grid(unsigned((n2 + bx - 1) / bx), unsigned((n1 + by - 1) / by), unsigned(n0));
    timed(id, n0 * n1 * n2, [&] {
      hipLaunchKernelGGL(k_each3<F>, grid, dim3(bx, by, 1), 0, s, int(n0), int(n1), int(n2), f);
    });
    note(hipGetLastError());
  }

  double* part = nullptr;
  template <class F>
  void reduce(int id, int64_t n, F f, double sums[9], double maxes[2]) {
    if (!part) note(hipMalloc(reinterpret_cast<void**>(&part), sizeof(double) * 11 * kRedBlocks));
    int64_t nb = (n + kRedThreads - 1) / kRedThreads;
    if (nb > kRedBlocks) nb = kRedBlocks;
    if (nb < 1) nb = 1;
    timed(id, n, [&] { hipLaunchKernelGGL(k_reduce<F>, dim3(unsigned(nb)), dim3(kRedThreads), 0, s, n, f, part); });
    note(hipGetLastError());
    std::vector<double> h(size_t(11 * nb));
    note(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    note(hipStreamSynchronize(s));
    for (int q = 0; q < 9; ++q) sums[q] = 0;
    maxes[0] = maxes[1] = 0;
    for (int64_t b = 0; b < nb; ++b) {
      for (int q = 0; q < 9; ++q) sums[q] += h[size_t(b * 11 + q)];
      maxes[0] = npmax(maxes[0], h[size_t(b * 11 + 9)]);
      maxes[1] = npmax(maxes[1], h[size_t(b * 11 + 10)]);
    }
  }

  template <class F>
  void max3(int id, int64_t n, F f, double out[3]) {
    note(hipMemsetAsync(red, 0, 3 * sizeof(unsigned long long), s));
    const int tb = 256;
    int64_t nb = (n + tb - 1) / tb;
    if (nb > 8192) nb = 8192;
    if (nb < 1) nb = 1;
    timed(id, n, [&] { hipLaunchKernelGGL(k_max3<F>, dim3(unsigned(nb)), dim3(tb), 0, s, n, f, red); });
    note(hipGetLastError());
    unsigned long long h[3] = {0, 0, 0};
    note(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, s));
    note(hipStreamSynchronize(s));
    for (int d = 0; d < 3; ++d) std::memcpy(&out[d], &h[d], sizeof(double));
  }
};

struct HipBackend {
  HipExec ex;
  std::string msg;

  HipBackend() {
    ex.note(hipStreamCreate(&ex.s));
    ex.note(hipMalloc(reinterpret_cast<void**>(&ex.red), 3 * sizeof(unsigned long long)));
  }
  ~HipBackend() {
    ex.destroy_events();
    if (ex.s) hipStreamSynchronize(ex.s);
    if (ex.red) hipFree(ex.red);
    if (ex.part) hipFree(ex.part);
    if (ex.s) hipStreamDestroy(ex.s);
  }
  HipExec& exec() { return ex; }
  void* alloc(size_t bytes) {
    void* ptr = nullptr;
    if (hipMalloc(&ptr, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return ptr;
  }
  void release(void* ptr) { (void)hipFree(ptr); }
  void h2d(void* dst, const void* src, size_t bytes) {
    ex.note(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
  }
  void d2h(void* dst, const void* src, size_t bytes) {
    ex.note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
  }
  void sync() { ex.note(hipStreamSynchronize(ex.s)); }
  void set_profiling(bool on) {
    if (on) ex.reset_profile();
    else ex.collect();
    ex.timing = on;
  }
  int profile(int maxn, char* names, double* ms, int64_t* launches, int64_t* elems) {
    ex.collect();
    int k = 0;
    for (int id = 0; id < K_COUNT && k < maxn; ++id) {
      if (ex.launches[id] == 0) continue;
      std::strncpy(names + 32 * k, kernel_name(id), 31);
      names[32 * k + 31] = 0;
      ms[k] = ex.tot_ms[id];
      launches[k] = ex.launches[id];
      elems[k] = ex.elems[id];
      ++k;
    }
    return k;
  }
  const char* error() {
    if (ex.last == hipSuccess) return nullptr;
    msg = std::string("HIP error: ") + hipGetErrorString(ex.last);
    return msg.c_str();
  }
};

}  // namespace pmhd

PMHD_DEFINE_C_ABI(pmhd::HipBackend)
