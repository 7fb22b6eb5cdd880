// HIP backend of libpmhd.so for gfx950 (MI355X): one thread per element, wave64
// shuffle reductions for the CFL maximum, one HIP stream per handle.  A single
// grid launches each phase over its own index box; a batch launches each phase
// once over the tiles of every box (a persistent grid walks the tile list).
// Included by the library's translation units (pmhd_hip_f64.hip, pmhd_hip_f32.hip,
// pmhd_hip_adj.hip), each of which instantiates the handles of one kind.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "pmhd_adjoint.h"

PMHD_NS_BEGIN

// ---------------------------------------------------------------------------
// single-grid kernels
// ---------------------------------------------------------------------------
// one thread per element of an (n0, n1, n2) box: x covers the contiguous axis
template <class Ph, class T>
__global__ void __launch_bounds__(256) k_phase(const Ctx<T> c, int n1, int n2) {
  const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = int(blockIdx.y * blockDim.y + threadIdx.y);
  const int j = int(blockIdx.z);
  if (k < n2 && i < n1) Ph()(c, I3{j, i, k});
}

// one thread per element of an (n0, n1, n2) box numbered flat with n2 fastest:
// launches over runs of faces, whose fastest axis may be short (the runs along z)
// Workgroups are handed to the 8 XCDs in turn, and each XCD has its own L2: number
// them so that each XCD walks a contiguous eighth of the index box, whose
// neighbouring elements (the stencils' shared reads) then meet in one L2.
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb / 8;
  return (per >= 8 && b < per * 8) ? (b % 8) * per + b / 8 : b;
}

// the run sweeps keep a window of cell vectors in registers: hold them to the
// 168 registers at which three waves share each SIMD
template <class Ph>
constexpr int phase_waves() { return Ph::kind == EXT_SEG ? 3 : 1; }
template <class Ph, class T>
__global__ void __launch_bounds__(256, phase_waves<Ph>()) k_phase_flat(const Ctx<T> c, unsigned n1, unsigned n2, unsigned total) {
  // (the sweeps' runs of faces share little between neighbouring threads)
  const unsigned b = Ph::kind == EXT_SEG ? unsigned(blockIdx.x) : xcd_block(blockIdx.x, gridDim.x);
  const unsigned t = b * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const unsigned k = t % n2;
  const unsigned r = t / n2;
  Ph()(c, I3{int64_t(r / n1), int64_t(r % n1), int64_t(k)});
}

// Fused predictor sweep over one line along the contiguous axis D (z in 3D, x in
// 2D) by one workgroup, in chunks of blockDim.x cells: each thread reconstructs one
// cell and keeps its left edge state; the right edge states go through LDS to the
// thread of the next face.  Faces 1..n-1 only: the end faces 0 and n need the edge
// states of both ends of the line, which are kept (store_line_end) for the
// SweepEnds launch, so every thread of every wave does the same work.  The face
// states are those of k_sweep_run, so the results are identical.  q: the line's
// first cell.  Every thread of the workgroup calls it (it synchronises).
template <int D, int S, class T>
__device__ inline void sweep_line(const Ctx<T>& c, const I3 q) {
  // right edge states and reconstruction source terms of the chunk's cells; prev:
  // those of the last cell of the previous chunk
  __shared__ T sra[9][256];
  __shared__ T prev[9];
  const int t = int(threadIdx.x), B = int(blockDim.x);
  const int64_t n = c.n_along(D);
  const int bc = c.bc[D];
  const bool from_s = c.order == 2;
  for (int64_t k0 = 0; k0 < n; k0 += B) {
    const int64_t k = k0 + t;
    T la[7], ra[7], src[2];
    if (k < n) {
      cell_edge_states(c, D, bc, with(q, D, k), from_s, la, ra);
      cell_recon_src(c, D, with(q, D, k), k, src);
#pragma unroll
      for (int v = 0; v < 7; ++v) sra[v][t] = ra[v];
      sra[7][t] = src[0];
      sra[8][t] = src[1];
      if (k == 0) store_line_end(c, D, q, 0, la);
      if (k == n - 1) store_line_end(c, D, q, 1, ra);
    }
    __syncthreads();
    if (k > 0 && k < n) {
      T wl[7], sl[2];
#pragma unroll
      for (int v = 0; v < 7; ++v) wl[v] = t == 0 ? prev[v] : sra[v][t - 1];
      sl[0] = t == 0 ? prev[7] : sra[7][t - 1];
      sl[1] = t == 0 ? prev[8] : sra[8][t - 1];
      predict_face_src<S>(c, D, with(q, D, k), wl, la, sl, src);
    }
    __syncthreads();
    if (k0 + B < n) {
      if (t == 0) {
#pragma unroll
        for (int v = 0; v < 9; ++v) prev[v] = sra[v][B - 1];
      }
      __syncthreads();
    }
  }
}

// one workgroup per line: (j, i) = (blockIdx.y, blockIdx.x) along z in 3D, j =
// blockIdx.x along x in 2D
template <int D, int S, class T>
__global__ void __launch_bounds__(256) k_sweep_line(const Ctx<T> c) {
  static_assert(D == 0 || D == 2, "line sweeps run along the contiguous axis");
  const I3 q = D == 2 ? I3{int64_t(blockIdx.y), int64_t(blockIdx.x), 0} : I3{int64_t(blockIdx.x), 0, 0};
  sweep_line<D, S>(c, q);
}

__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}

// Maxima of three non-negative values.  The IEEE bit patterns of non-negative
// doubles order like the values (NaN above +inf, as np.max propagates it), so the
// reductions are integer maxima: deterministic in any order.
// block_max3: reduce over a block of at most 256 threads; the result is valid in
// thread 0.  Every thread of the block must call it.
__device__ inline void block_max3(unsigned long long m[3]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], __shfl_xor(m[d], off, 64));
  }
  __shared__ unsigned long long sm[3][4];
  const int tid = int(threadIdx.y * blockDim.x + threadIdx.x);
  const int lane = tid & 63, wv = tid >> 6;
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) sm[d][wv] = m[d];
  }
  __syncthreads();
  if (tid == 0) {
    const int nw = int(blockDim.x * blockDim.y + 63) / 64;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      unsigned long long v = sm[d][0];
      for (int w = 1; w < nw; ++w) v = umax64(v, sm[d][w]);
      m[d] = v;
    }
  }
  __syncthreads();
}

__device__ inline unsigned long long speed_bits(double x) { return (unsigned long long)__double_as_longlong(x); }

// a phase whose elements also yield three speeds: each block writes its maxima
// to part[3 * block] (no atomics; reduced by k_max_parts)
// (flat numbering with n2 fastest, XCD-contiguous like k_phase_flat)
template <class Ph, class T>
__global__ void __launch_bounds__(256) k_phase_max(const Ctx<T> c, unsigned n1, unsigned n2, unsigned total,
                                                   unsigned long long* part) {
  const unsigned t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if constexpr (Ph::id == K_UPDATE) {
    // the cycle's floor counts are complete (every flooring launch precedes the
    // update on the stream): one thread folds them
    if (blockIdx.x == 0 && threadIdx.x == 0 && c.fcnt) fold_floor_counts(c.fcnt);
  }
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  if (t < total) {
    const unsigned k = t % n2;
    const unsigned r = t / n2;
    double o[3];
    Ph()(c, I3{int64_t(r / n1), int64_t(r % n1), int64_t(k)}, o);
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = speed_bits(o[d]);
  }
  block_max3(m);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) part[3 * size_t(blockIdx.x) + d] = m[d];
  }
}

// out[d] = max over the nparts partial maxima: each block reduces a strided
// subset, then one atomic per block and component (out cleared beforehand)
__global__ void __launch_bounds__(256) k_max_parts(const unsigned long long* part, int64_t nparts,
                                                   unsigned long long* out) {
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < nparts; b += stride) {
#pragma unroll
    for (int d = 0; d < 3; ++d) m[d] = umax64(m[d], part[3 * b + d]);
  }
  block_max3(m);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicMax(&out[d], m[d]);
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

// ---------------------------------------------------------------------------
// batch kernels: the tiles of all boxes form one list; box b owns tiles
// [prefix[b], prefix[b+1]); a tile is bk (along z) x bi (along x) elements of one
// y row.  The grid walks the list with a stride (persistent when smaller).
// ---------------------------------------------------------------------------
template <class T>
__device__ inline TileRef tile_of(const Ctx<T>* __restrict__ ctxs, const int* __restrict__ prefix, int nbox,
                                  int vb, int kind, int dir, int bk, int bi) {
  return locate_tile(ctxs, prefix, nbox, vb, kind, dir, bk, bi, int(threadIdx.x));
}

template <class Ph, class T>
__global__ void __launch_bounds__(256) k_batch(const Ctx<T>* __restrict__ ctxs, const int* __restrict__ prefix,
                                               int nbox, int bk, int bi) {
  const int total = prefix[nbox];
  for (int vb = int(blockIdx.x); vb < total; vb += int(gridDim.x)) {
    const TileRef r = tile_of(ctxs, prefix, nbox, vb, Ph::kind, Ph::dir, bk, bi);
    const Ctx<T>& c = ctxs[r.box];
    if (!c.active) continue;
    if (r.in) Ph()(c, I3{r.j, r.i, r.k});
  }
}

// line sweeps of a batch: one workgroup per line of one box (prefix: lines per box)
template <int D, int S, class T>
__global__ void __launch_bounds__(256) k_batch_line(const Ctx<T>* __restrict__ ctxs, const int* __restrict__ prefix,
                                                    int nbox) {
  const int total = prefix[nbox];
  for (int vb = int(blockIdx.x); vb < total; vb += int(gridDim.x)) {
    const TileRef r = locate_tile(ctxs, prefix, nbox, vb, EXT_LINE, D, 1, 1, 0);
    const Ctx<T>& c = ctxs[r.box];
    if (!c.active) continue;
    sweep_line<D, S>(c, I3{r.j, r.i, r.k});
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_batch_cfl(const Ctx<T>* __restrict__ ctxs, const int* __restrict__ prefix,
                                                   int nbox, int bk, int bi, unsigned long long* red) {
  const int total = prefix[nbox];
  for (int vb = int(blockIdx.x); vb < total; vb += int(gridDim.x)) {
    const TileRef r = tile_of(ctxs, prefix, nbox, vb, EXT_CELL, 0, bk, bi);
    const Ctx<T>& c = ctxs[r.box];
    if (!c.active) continue;
    unsigned long long m[3] = {0ull, 0ull, 0ull};
    if (r.in) {
      double o[3];
      CflSpeeds()(c, I3{r.j, r.i, r.k}, o);
#pragma unroll
      for (int d = 0; d < 3; ++d) m[d] = speed_bits(o[d]);
    }
    block_max3(m);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicMax(&red[3 * r.box + d], m[d]);
    }
  }
}

// a batch phase that also yields speeds: per-box maxima into red (3 per box)
template <class Ph, class T>
__global__ void __launch_bounds__(256) k_batch_max(const Ctx<T>* __restrict__ ctxs, const int* __restrict__ prefix,
                                                   int nbox, int bk, int bi, unsigned long long* red) {
  const int total = prefix[nbox];
  for (int vb = int(blockIdx.x); vb < total; vb += int(gridDim.x)) {
    const TileRef r = tile_of(ctxs, prefix, nbox, vb, Ph::kind, Ph::dir, bk, bi);
    const Ctx<T>& c = ctxs[r.box];
    if (!c.active) continue;
    if constexpr (Ph::id == K_UPDATE) {
      // one thread of the box's first tile folds the box's floor counts
      if (threadIdx.x == 0 && vb == prefix[r.box] && c.fcnt) fold_floor_counts(c.fcnt);
    }
    unsigned long long m[3] = {0ull, 0ull, 0ull};
    if (r.in) {
      double o[3];
      Ph()(c, I3{r.j, r.i, r.k}, o);
#pragma unroll
      for (int d = 0; d < 3; ++d) m[d] = speed_bits(o[d]);
    }
    block_max3(m);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicMax(&red[3 * r.box + d], m[d]);
    }
  }
}


template <class T>
__global__ void k_box_begin(Ctx<T>* ctxs, BoxClock* clk, unsigned long long* red, int nbox, double cfl,
                            double t_end, int64_t max_cycles, int commit) {
  const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= nbox) return;
  double m[3];
  for (int d = 0; d < 3; ++d) {
    m[d] = __longlong_as_double((long long)red[3 * b + d]);
    red[3 * b + d] = 0ull;
  }
  box_begin(ctxs[b], clk[b], m, cfl, t_end, max_cycles, commit);
}

template <class T>
__global__ void k_box_end(Ctx<T>* ctxs, BoxClock* clk, int nbox) {
  const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < nbox) box_end(ctxs[b], clk[b]);
}

template <class T>
__global__ void k_box_activate(Ctx<T>* ctxs, int nbox) {
  const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (b < nbox) ctxs[b].active = 1;
}

// ---------------------------------------------------------------------------
// device float64 arrays of the reference layout <-> the state's storage type
// (grid-stride loops; the FP32 face field as high and low float parts)
// ---------------------------------------------------------------------------
template <class D, class S>
__global__ void __launch_bounds__(256) k_convert(D* __restrict__ dst, const S* __restrict__ src, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) dst[i] = D(src[i]);
}
template <class T>
__global__ void __launch_bounds__(256) k_split(T* __restrict__ hi, T* __restrict__ lo, const double* __restrict__ src,
                                               int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    const double v = src[i];
    const T h = T(v);
    hi[i] = h;
    lo[i] = T(v - double(h));
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_join(double* __restrict__ dst, const T* __restrict__ hi,
                                              const T* __restrict__ lo, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    dst[i] = double(hi[i]) + double(lo[i]);
}
inline unsigned stride_blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return unsigned(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// ---------------------------------------------------------------------------
// the Riemann flux of n faces (pmhd_riemann_flux): host float64 arrays staged
// through device buffers, one thread per face
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_riemann(int d, int solver, int64_t n, const double* wl, const double* wr,
                                                 const double* ul, const double* ur, const double* bn, T g,
                                                 int efix, T eps, double* F) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    riemann_face<T>(d, solver, n, i, wl, wr, ul, ur, bn, g, efix != 0, eps, F);
}

template <class T>
int riemann_run(int d, int solver, int64_t n, const double* wl, const double* wr, const double* ul,
                const double* ur, const double* bn, double gamma, int efix, double eps, double* F,
                std::string& err) {
  if (n <= 0) return PMHD_OK;
  const size_t plane = size_t(n) * sizeof(double);
  double* buf = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&buf), 36 * plane) != hipSuccess) {
    (void)hipGetLastError();
    err = "device allocation failed";
    return PMHD_ENOMEM;
  }
  const double* src[5] = {wl, wr, ul, ur, bn};
  hipError_t e = hipSuccess;
  for (int k = 0; k < 5 && e == hipSuccess; ++k)
    e = hipMemcpy(buf + size_t(7 * k) * size_t(n), src[k], (k < 4 ? 7 : 1) * plane, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const double* dv = buf;
    hipLaunchKernelGGL((k_riemann<T>), dim3(stride_blocks(n)), dim3(256), 0, nullptr, d, solver, n, dv,
                       dv + 7 * n, dv + 14 * n, dv + 21 * n, dv + 28 * n, T(gamma), efix, T(eps), buf + 29 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(F, buf + 29 * n, 7 * plane, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) {
    err = std::string("HIP error: ") + hipGetErrorString(e);
    return PMHD_EHIP;
  }
  return PMHD_OK;
}

// ---------------------------------------------------------------------------
// adjoint of one phase: one thread per output element (pmhd_adjoint.h)
// ---------------------------------------------------------------------------
struct DevAtomicAdd {
  __device__ void operator()(double* p, double v) const {
    if (v != 0.0) atomicAdd(p, v);
  }
};

template <class Ph, class T>
__global__ void __launch_bounds__(64) k_adjoint(const Ctx<T> c, const Ctx<double> a, int n1, int n2, int* flag) {
  const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = int(blockIdx.y * blockDim.y + threadIdx.y);
  const int j = int(blockIdx.z);
  if (k < n2 && i < n1) {
    AdState st;
    if (!adjoint_element<Ph>(c, a, I3{j, i, k}, st, DevAtomicAdd())) atomicOr(flag, 1);
  }
}

// ---------------------------------------------------------------------------
// executors
// ---------------------------------------------------------------------------
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

  int flat_block = 256;   // threads per block of the flat launches (64..256)
  bool flat_all = true;   // every phase as a flat, XCD-numbered launch

  // A second stream (default priority) for launches that overlap the following
  // ones: run_side forks after the work queued so far, join_side makes the main
  // stream wait for it.  With per-launch timing on, side launches run in order.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool side_pending = false;
  template <class L>
  void on_side(L launch) {
    if (timing || !side_ready()) {
      launch();
      return;
    }
    note(hipEventRecord(fork_ev, s));
    note(hipStreamWaitEvent(side, fork_ev, 0));
    std::swap(s, side);
    launch();
    std::swap(s, side);
    note(hipEventRecord(join_ev, side));
    side_pending = true;
  }
  bool side_ready() {
    if (side) return true;
    if (hipStreamCreate(&side) != hipSuccess) {
      side = nullptr;
      (void)hipGetLastError();
      return false;
    }
    note(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    note(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    return true;
  }
  void join_side() {
    if (!side_pending) return;
    note(hipStreamWaitEvent(s, join_ev, 0));
    side_pending = false;
  }
  void destroy_side() {
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    side = nullptr;
    fork_ev = join_ev = nullptr;
  }
  template <class Ph, class T>
  void run_side(const Ctx<T>& c) {
    on_side([&] { run<Ph>(c); });
  }

  static void grid_of(const int64_t n[3], dim3& grid, dim3& block) {
    unsigned bx = 1;
    while (bx < 64 && int64_t(bx) < n[2]) bx <<= 1;
    const unsigned by = 256 / bx;
    grid = dim3(unsigned((n[2] + bx - 1) / bx), unsigned((n[1] + by - 1) / by), unsigned(n[0]));
    block = dim3(bx, by, 1);
  }

  template <class Ph, class T>
  void run(const Ctx<T>& c) {
    int64_t n[3];
    c.extent(Ph::kind, Ph::dir, n);
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return;
    const int64_t total = n[0] * n[1] * n[2];
    // profiled work: faces for the launches over runs or lines of faces
    const int64_t work = Ph::kind == EXT_SEG || Ph::kind == EXT_LINE ? c.nface(Ph::dir) : total;
    if constexpr (Ph::kind == EXT_LINE) {
      // one workgroup per line: the smallest of 64, 128, 256 threads that covers
      // the line, else 256 in chunks
      unsigned b = 64;
      while (b < 256 && int64_t(b) < c.n_along(Ph::dir)) b <<= 1;
      // lines (j, i) along z in 3D, j along x in 2D (nz = 1)
      const dim3 grid = Ph::dir == 2 ? dim3(unsigned(c.nx), unsigned(c.ny)) : dim3(unsigned(c.ny));
      timed(Ph::id, work, [&] {
        hipLaunchKernelGGL((k_sweep_line<Ph::dir, Ph::solver, T>), grid, dim3(b), 0, s, c);
      });
      note(hipGetLastError());
      return;
    }
    dim3 grid, block;
    grid_of(n, grid, block);
    // flat numbering when the box launch would leave many lanes idle (e.g. the
    // n + 1 faces or edge points along z: 257 = 4 x 64 + 1)
    const double launched = double(grid.x) * block.x * double(grid.y) * block.y * double(grid.z);
    const bool sparse = launched > 1.05 * double(total);
    if ((Ph::kind == EXT_SEG || sparse || flat_all) && total < (int64_t(1) << 31)) {
      const unsigned nb = unsigned((total + flat_block - 1) / flat_block);
      timed(Ph::id, work, [&] {
        hipLaunchKernelGGL((k_phase_flat<Ph, T>), dim3(nb), dim3(unsigned(flat_block)), 0, s, c, unsigned(n[1]),
                           unsigned(n[2]), unsigned(total));
      });
      note(hipGetLastError());
      return;
    }
    timed(Ph::id, total, [&] {
      hipLaunchKernelGGL((k_phase<Ph, T>), grid, block, 0, s, c, int(n[1]), int(n[2]));
    });
    note(hipGetLastError());
  }

  // partial maxima of the last run_max launch (one triple per block)
  unsigned long long* parts = nullptr;
  int64_t parts_cap = 0;
  int64_t parts_n = 0;

  template <class Ph, class T>
  void run_max(const Ctx<T>& c) {
    int64_t n[3];
    c.extent(Ph::kind, Ph::dir, n);
    const int64_t total = n[0] * n[1] * n[2];   // cells: below 2^29
    const int64_t nb = (total + flat_block - 1) / flat_block;
    if (nb > parts_cap) {
      if (parts) note(hipFree(parts));
      parts = nullptr;
      note(hipMalloc(reinterpret_cast<void**>(&parts), sizeof(unsigned long long) * 3 * size_t(nb)));
      parts_cap = nb;
    }
    parts_n = nb;
    timed(Ph::id, total, [&] {
      hipLaunchKernelGGL((k_phase_max<Ph, T>), dim3(unsigned(nb)), dim3(unsigned(flat_block)), 0, s, c,
                         unsigned(n[1]), unsigned(n[2]), unsigned(total), parts);
    });
    note(hipGetLastError());
  }

  // the maxima of the last run_max (waits for it)
  void max_partials(double out[3]) {
    note(hipMemsetAsync(red, 0, 3 * sizeof(unsigned long long), s));
    int64_t nb = (parts_n + 255) / 256;
    nb = nb < 1 ? 1 : (nb > 128 ? 128 : nb);
    hipLaunchKernelGGL(k_max_parts, dim3(unsigned(nb)), dim3(256), 0, s, parts, parts_n, red);
    note(hipGetLastError());
    unsigned long long h[3] = {0, 0, 0};
    note(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, s));
    note(hipStreamSynchronize(s));
    for (int d = 0; d < 3; ++d) std::memcpy(&out[d], &h[d], sizeof(double));
  }

  template <class T>
  void cfl(const Ctx<T>& c, double out[3]) {
    run_max<CflSpeeds>(c);
    max_partials(out);
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
};

// one launch per phase over every box of a batch
struct HipBatchExec {
  HipExec* ex = nullptr;
  BatchLayout* lay = nullptr;    // device contexts, clocks, tile tables
  // when set, run() times every tile candidate of each phase and keeps the best
  bool tuning = false;
  int tune_reps = 3;

  void launch_cfg(int id, int& bk, int& bi, int& grid) const {
    bk = lay->tiles[id].bk;
    bi = lay->tiles[id].bi;
    const int total = lay->total[id];
    grid = lay->tiles[id].grid > 0 && lay->tiles[id].grid < total ? lay->tiles[id].grid : total;
    if (grid < 1) grid = 1;
  }

  template <class Ph, class T>
  void launch(int bk, int bi, int grid) {
    const Ctx<T>* dctx = static_cast<const Ctx<T>*>(lay->dctx);
    const int* prefix = lay->dprefix + Ph::id * (lay->nbox + 1);
    ex->timed(Ph::id, lay->elems[Ph::id], [&] {
      hipLaunchKernelGGL((k_batch<Ph, T>), dim3(unsigned(grid)), dim3(unsigned(bk * bi)), 0, ex->s, dctx, prefix,
                         lay->nbox, bk, bi);
    });
    ex->note(hipGetLastError());
  }

  template <class Ph, class T>
  void run_side(const Ctx<T>& c) {
    if (tuning) {
      run<Ph>(c);
      return;
    }
    ex->on_side([&] { run<Ph>(c); });
  }
  void join_side() { ex->join_side(); }
  template <class Ph, class T>
  void run(const Ctx<T>&) {
    if constexpr (Ph::kind == EXT_LINE) {
      // one workgroup per line (64 threads in a batch: lines of uneven boxes waste
      // at most the tail of their last chunk); the layout's tile is one line
      const Ctx<T>* dctx = static_cast<const Ctx<T>*>(lay->dctx);
      const int* prefix = lay->dprefix + Ph::id * (lay->nbox + 1);
      const int total = lay->total[Ph::id];
      if (total <= 0) return;
      ex->timed(Ph::id, lay->elems[Ph::id], [&] {
        hipLaunchKernelGGL((k_batch_line<Ph::dir, Ph::solver, T>), dim3(unsigned(total)),
                           dim3(unsigned(lay->line_block)), 0, ex->s, dctx, prefix, lay->nbox);
      });
      ex->note(hipGetLastError());
      return;
    }
    if (tuning) tune<Ph, T>();
    int bk, bi, grid;
    launch_cfg(Ph::id, bk, bi, grid);
    launch<Ph, T>(bk, bi, grid);
  }

  // time each candidate tile (and grid) of launch id and keep the fastest;
  // launch_with(bk, bi, grid) issues one launch
  template <class L>
  void tune_id(int id, L&& launch_with) {
    const std::vector<TileCfg> cands = tile_candidates(*lay, id);
    hipEvent_t a = ex->event(), b = ex->event();
    float best = -1.f;
    TileCfg keep = lay->tiles[id];
    for (const TileCfg& t : cands) {
      if (!lay->apply_tile(id, t)) continue;
      int bk, bi, grid;
      launch_cfg(id, bk, bi, grid);
      launch_with(bk, bi, grid);   // warm
      ex->note(hipEventRecord(a, ex->s));
      for (int r = 0; r < tune_reps; ++r) launch_with(bk, bi, grid);
      ex->note(hipEventRecord(b, ex->s));
      ex->note(hipEventSynchronize(b));
      float ms = 0.f;
      ex->note(hipEventElapsedTime(&ms, a, b));
      if (best < 0.f || ms < best) {
        best = ms;
        keep = t;
      }
    }
    (void)lay->apply_tile(id, keep);
    lay->tuned_ms[id] = best / float(tune_reps);
    ex->pool.push_back(a);
    ex->pool.push_back(b);
  }

  template <class Ph, class T>
  void tune() {
    tune_id(Ph::id, [&](int bk, int bi, int grid) { launch<Ph, T>(bk, bi, grid); });
  }

  template <class Ph, class T>
  void launch_max(int bk, int bi, int grid) {
    const Ctx<T>* dctx = static_cast<const Ctx<T>*>(lay->dctx);
    const int* prefix = lay->dprefix + Ph::id * (lay->nbox + 1);
    ex->timed(Ph::id, lay->elems[Ph::id], [&] {
      hipLaunchKernelGGL((k_batch_max<Ph, T>), dim3(unsigned(grid)), dim3(unsigned(bk * bi)), 0, ex->s, dctx,
                         prefix, lay->nbox, bk, bi, lay->dred);
    });
    ex->note(hipGetLastError());
  }

  // per-box speed maxima accumulate into lay->dred for the next cycle's begin()
  template <class Ph, class T>
  void run_max(const Ctx<T>&) {
    if (tuning) tune_id(Ph::id, [&](int bk, int bi, int grid) { launch_max<Ph, T>(bk, bi, grid); });
    int bk, bi, grid;
    launch_cfg(Ph::id, bk, bi, grid);
    launch_max<Ph, T>(bk, bi, grid);
  }

  template <class T>
  void cfl_launch(int bk, int bi, int grid) {
    const Ctx<T>* dctx = static_cast<const Ctx<T>*>(lay->dctx);
    const int* prefix = lay->dprefix + K_CFL * (lay->nbox + 1);
    ex->timed(K_CFL, lay->elems[K_CFL], [&] {
      hipLaunchKernelGGL((k_batch_cfl<T>), dim3(unsigned(grid)), dim3(unsigned(bk * bi)), 0, ex->s, dctx, prefix,
                         lay->nbox, bk, bi, lay->dred);
    });
    ex->note(hipGetLastError());
  }

  // per-box CFL maxima into lay->dred (3 per box; cleared by begin())
  template <class T>
  void cfl(const Ctx<T>&) {
    if (tuning) tune_id(K_CFL, [&](int bk, int bi, int grid) { cfl_launch<T>(bk, bi, grid); });
    int bk, bi, grid;
    launch_cfg(K_CFL, bk, bi, grid);
    cfl_launch<T>(bk, bi, grid);
  }

  template <class T>
  void begin(double cfl, double t_end, int64_t max_cycles, int commit) {
    const unsigned nb = unsigned((lay->nbox + 63) / 64);
    hipLaunchKernelGGL((k_box_begin<T>), dim3(nb), dim3(64), 0, ex->s, static_cast<Ctx<T>*>(lay->dctx), lay->dclk,
                       lay->dred, lay->nbox, cfl, t_end, max_cycles, commit);
    ex->note(hipGetLastError());
  }
  template <class T>
  void end() {
    const unsigned nb = unsigned((lay->nbox + 63) / 64);
    hipLaunchKernelGGL((k_box_end<T>), dim3(nb), dim3(64), 0, ex->s, static_cast<Ctx<T>*>(lay->dctx), lay->dclk,
                       lay->nbox);
    ex->note(hipGetLastError());
  }
  template <class T>
  void activate_all() {
    const unsigned nb = unsigned((lay->nbox + 63) / 64);
    hipLaunchKernelGGL((k_box_activate<T>), dim3(nb), dim3(64), 0, ex->s, static_cast<Ctx<T>*>(lay->dctx),
                       lay->nbox);
    ex->note(hipGetLastError());
  }
};

struct HipBackend {
  HipExec ex;
  HipBatchExec bex;
  std::string msg;
  int cus = 256;

  HipBackend() {
    ex.note(hipStreamCreate(&ex.s));
    ex.note(hipMalloc(reinterpret_cast<void**>(&ex.red), 3 * sizeof(unsigned long long)));
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      int n = 0;
      if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) cus = n;
    }
    bex.ex = &ex;
  }
  ~HipBackend() {
    ex.destroy_events();
    ex.destroy_side();
    if (ex.s) hipStreamSynchronize(ex.s);
    if (own) ex.s = own;   // destroy the handle's own stream, never the caller's
    if (ex.red) hipFree(ex.red);
    if (ex.part) hipFree(ex.part);
    if (ex.parts) hipFree(ex.parts);
    if (adj_flag) hipFree(adj_flag);
    if (ex.s) hipStreamDestroy(ex.s);
  }
  HipExec& exec() { return ex; }
  HipBatchExec& batch_exec(BatchLayout* lay) {
    bex.lay = lay;
    return bex;
  }
  int compute_units() const { return cus; }
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
  void zero(void* dst, size_t bytes) {
    ex.note(hipMemsetAsync(dst, 0, bytes, ex.s));
  }
  void d2d(void* dst, const void* src, size_t bytes) {
    ex.note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ex.s));
  }
  template <class T>
  void dev_in(T* dst, const double* src, int64_t n) {
    if (n <= 0) return;
    if constexpr (sizeof(T) == sizeof(double)) {
      d2d(dst, src, size_t(n) * sizeof(double));
    } else {
      hipLaunchKernelGGL((k_convert<T, double>), dim3(stride_blocks(n)), dim3(256), 0, ex.s, dst, src, n);
      ex.note(hipGetLastError());
    }
  }
  template <class T>
  void dev_out(double* dst, const T* src, int64_t n) {
    if (n <= 0) return;
    if constexpr (sizeof(T) == sizeof(double)) {
      d2d(dst, src, size_t(n) * sizeof(double));
    } else {
      hipLaunchKernelGGL((k_convert<double, T>), dim3(stride_blocks(n)), dim3(256), 0, ex.s, dst, src, n);
      ex.note(hipGetLastError());
    }
  }
  template <class T>
  void dev_split(T* hi, T* lo, const double* src, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_split<T>), dim3(stride_blocks(n)), dim3(256), 0, ex.s, hi, lo, src, n);
    ex.note(hipGetLastError());
  }
  template <class T>
  void dev_join(double* dst, const T* hi, const T* lo, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_join<T>), dim3(stride_blocks(n)), dim3(256), 0, ex.s, dst, hi, lo, n);
    ex.note(hipGetLastError());
  }
  // launch on the caller's stream (nullptr: the handle's own); the work queued so
  // far is finished first
  hipStream_t own = nullptr;
  void set_stream(void* s) {
    if (!own) own = ex.s;
    ex.note(hipStreamSynchronize(ex.s));
    ex.s = s ? static_cast<hipStream_t>(s) : own;
  }
  int* adj_flag = nullptr;
  template <class Ph, class T>
  void adjoint_phase(const Ctx<T>& c, const Ctx<double>& a) {
    if (!adj_flag) {
      ex.note(hipMalloc(reinterpret_cast<void**>(&adj_flag), sizeof(int)));
      ex.note(hipMemsetAsync(adj_flag, 0, sizeof(int), ex.s));
    }
    int64_t n[3];
    c.extent(Ph::kind, Ph::dir, n);
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return;
    unsigned bx = 1;
    while (bx < 64 && int64_t(bx) < n[2]) bx <<= 1;
    const unsigned by = 64 / bx;
    dim3 grid(unsigned((n[2] + bx - 1) / bx), unsigned((n[1] + by - 1) / by), unsigned(n[0]));
    ex.timed(K_ADJOINT, n[0] * n[1] * n[2], [&] {
      hipLaunchKernelGGL((k_adjoint<Ph, T>), grid, dim3(bx, by, 1), 0, ex.s, c, a, int(n[1]), int(n[2]), adj_flag);
    });
    ex.note(hipGetLastError());
  }
  bool adjoint_overflow() {
    if (!adj_flag) return false;
    int h = 0;
    ex.note(hipMemcpyAsync(&h, adj_flag, sizeof(int), hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
    return h != 0;
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

PMHD_NS_END
