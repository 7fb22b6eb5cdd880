// Deterministic accumulation of the adjoint's cotangent contributions (HIP).
//
// The element adjoints of a phase add their contributions to shared cotangent
// arrays; with FP64 atomics the order of those additions, and so the rounding of
// the sums, changes from run to run.  Here a launch writes every contribution into
// its own place instead -- contribution q of element e at q * E + e of a value
// buffer -- and a second launch sums, for each target, its contributions in the
// order of their places and adds the sum to the target.  The result is the same
// bit for bit on every run.
//
// Which places feed which target depends only on the element positions and the
// configuration (each element body adds to the same targets in the same order
// every time), so it is worked out once per launch of the tape and kept as a plan:
// the first time, the launch also writes each place's target index (the cotangent
// arrays are one block, so a target is an index into it); a stable radix sort of
// (target, place), a run-length encoding and a scan give the targets, the start of
// each target's run and the places in ascending order.  Later one-thread launches
// check that each element still makes as many contributions, to targets with the
// same hash.
//
// The generic element adjoints also keep their recording pass's slot lists (the
// input slots to seed and the outputs whose cotangents they consume) with the
// launch, by array id and index, so later evaluations skip the recording pass.
// Later evaluations then split an element's seeded passes over two lanes of one
// wave (k_adjoint_kept2: the share of slot k goes to place k, so the places are the
// one-thread ones by construction), and the predictor and corrector faces split
// their 15 flux passes likewise (k_adjoint_corr2: the face's first lane then adds
// every contribution in the one-thread order).
#pragma once
#include <hipcub/hipcub.hpp>

#include "pmhd_hip.h"

PMHD_NS_BEGIN

struct DetSink {
  double* vals;          // [kAdjMaxContrib][E]
  uint32_t* keys;        // target index per place (the plan's first launch), else null
  int32_t* cnt;          // contributions per element: written by the first launch, then checked
  uint32_t* hash;        // hash of the element's target sequence: same
  const double* arena;   // the block of cotangent arrays
  int64_t arena_len;
  int64_t E;
  int* flag;
};

// the sink one element adds through: n and h are the element's running count and
// target hash (shared by the copies the element body makes)
struct DevDetAdd {
  DetSink s;
  int64_t e;
  int* n;
  uint32_t* h;
  __device__ void operator()(double* p, double v) const {
    const int q = (*n)++;
    const int64_t id = p - s.arena;
    const bool inside = id >= 0 && id < s.arena_len;
    *h = *h * 0x01000193u ^ uint32_t(id);
    if (q >= kAdjMaxContrib) return;
    const int64_t at = int64_t(q) * s.E + e;
    s.vals[at] = v;
    if (s.keys) {
      if (!inside) atomicOr(s.flag, 4);
      s.keys[at] = inside ? uint32_t(id) : uint32_t(s.arena_len);
    }
  }
};

// the slot lists of a launch: written by its first evaluation (rows of E elements,
// outputs from row rin), read by the later ones
struct SlotBufs {
  uint64_t* slots;
  int32_t* nin;
  int32_t* nout;
  int rin;
  bool write;
  bool use;
};

template <class Ph, class T>
__global__ void __launch_bounds__(64) k_adjoint_det(const Ctx<T> c, const Ctx<double> a, int n1, int n2, DetSink s,
                                                    SlotBufs sb) {
  const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int i = int(blockIdx.y * blockDim.y + threadIdx.y);
  const int j = int(blockIdx.z);
  if (k < n2 && i < n1) {
    const int64_t e = (int64_t(j) * n1 + i) * n2 + k;
    int n = 0;
    uint32_t h = 0x811c9dc5u;
    AdState st;
    const SlotView sv{sb.slots, sb.nin, sb.nout, s.E, e, sb.rin, sb.write};
    const bool ok = adjoint_element<Ph>(c, a, I3{j, i, k}, st, DevDetAdd{s, e, &n, &h}, sb.use ? &sv : nullptr);
    if (!ok || n > kAdjMaxContrib) atomicOr(s.flag, 1);
    if (s.keys) {
      s.cnt[e] = n;
      s.hash[e] = h;
    } else if (n != s.cnt[e] || h != s.hash[e]) {
      atomicOr(s.flag, 2);
    }
  }
}

// kept slot lists: two adjacent threads (lanes of one wave) per element
template <class Ph, class T>
__global__ void __launch_bounds__(64) k_adjoint_kept2(const Ctx<T> c, const Ctx<double> a, int n1, int n2, int64_t E,
                                                      double* vals, SlotBufs sb) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t e = t >> 1;
  if (e < E) {
    const int k = int(e % n2);
    const int64_t r = e / n2;
    const int i = int(r % n1);
    const int j = int(r / n1);
    const SlotView sv{sb.slots, sb.nin, sb.nout, E, e, sb.rin, false};
    adjoint_generic_lanes<Ph>(c, a, I3{j, i, k}, sv, int(t & 1), 2, vals);
  }
}

// corrector and predictor faces: two adjacent threads per face split the 15 flux
// passes, then the face's first thread adds every contribution (in the one-thread
// order)
struct PairShare {
  __device__ void operator()(double* g) const {
#pragma unroll
    for (int q = 0; q < 15; ++q) g[q] += __shfl_xor(g[q], 1);
  }
};
template <class Ph, class T>
__global__ void __launch_bounds__(64) k_adjoint_corr2(const Ctx<T> c, const Ctx<double> a, int n1, int n2, int64_t E,
                                                      DetSink s) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t e = t >> 1;
  if (e < E) {
    const int k = int(e % n2);
    const int64_t r = e / n2;
    const int i = int(r % n1);
    const int j = int(r / n1);
    int n = 0;
    uint32_t h = 0x811c9dc5u;
    adjoint_staged<Ph>(c, a, I3{j, i, k}, DevDetAdd{s, e, &n, &h}, int(t & 1), 2, PairShare());
  }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void __launch_bounds__(256) k_iota_u32(uint32_t* p, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = uint32_t(i);
}

__global__ void __launch_bounds__(256) k_max_pair(const int32_t* a, const int32_t* b, int64_t n, int* out) {
  int ma = 0, mb = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    ma = a[i] > ma ? a[i] : ma;
    mb = b[i] > mb ? b[i] : mb;
  }
  atomicMax(out, ma);
  atomicMax(out + 1, mb);
}

// one thread per target: its contributions summed in the order of their places
__global__ void __launch_bounds__(256) k_det_gather(double* arena, const double* vals, const uint32_t* uid,
                                                    const uint32_t* off, const uint32_t* pos, int64_t nuniq) {
  const int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u < nuniq) {
    double s = 0.0;
    const uint32_t q1 = off[u + 1];
    for (uint32_t q = off[u]; q < q1; ++q) s += vals[pos[q]];
    arena[uid[u]] += s;
  }
}

// the HIP backend of the adjoint handles: the base backend's launches, with the
// element adjoints accumulating through plans (or through FP64 atomics when
// deterministic accumulation is switched off)
struct HipAdjBackend : HipBackend {
  struct Plan {
    int id = -1;
    int64_t E = 0;
    int64_t nuniq = 0;
    uint32_t* uid = nullptr;
    uint32_t* off = nullptr;
    uint32_t* pos = nullptr;
    int32_t* cnt = nullptr;
    uint32_t* hash = nullptr;
  };
  struct SlotPlan {
    int id = -1;
    int64_t E = 0;
    uint64_t* slots = nullptr;
    int32_t* nin = nullptr;
    int32_t* nout = nullptr;
    int rin = 0;
  };
  std::vector<SlotPlan> splans;   // by tape position (the same for both parities)
  uint64_t* slot_scratch = nullptr;
  int64_t slot_cap = 0;
  int* pair = nullptr;
  bool det = true;
  double* arena = nullptr;
  int64_t arena_len = 0;
  std::vector<Plan> plans[2];
  int sel_par = 0;
  size_t sel_t = 0;
  std::vector<void*> mine;
  // the value buffer, and the scratch of a plan's construction
  double* vals = nullptr;
  int64_t vals_cap = 0;
  uint32_t* sk[6] = {};   // keys in/out, places in/out, run targets, run lengths (+1)
  int64_t sk_cap = 0;
  int* nruns = nullptr;
  void* temp = nullptr;
  size_t temp_cap = 0;

  ~HipAdjBackend() {
    if (ex.s) hipStreamSynchronize(ex.s);
    for (void* p : mine) (void)hipFree(p);
  }
  void* grab(size_t bytes) {
    void* p = nullptr;
    ex.note(hipMalloc(&p, bytes > 0 ? bytes : 8));
    if (p) mine.push_back(p);
    return p;
  }
  void drop(void* p) {
    if (!p) return;
    for (auto& q : mine)
      if (q == p) {
        (void)hipFree(p);
        q = mine.back();
        mine.pop_back();
        return;
      }
  }
  void set_deterministic(bool on) { det = on; }
  void set_arena(double* p, int64_t n) {
    arena = p;
    arena_len = n;
  }
  void adjoint_select(int par, size_t t) {
    sel_par = par & 1;
    sel_t = t;
  }
  // the end of a backward pass: the plans are built, their scratch can go
  void adjoint_end() {
    ex.note(hipStreamSynchronize(ex.s));
    for (int q = 0; q < 6; ++q) {
      drop(sk[q]);
      sk[q] = nullptr;
    }
    drop(temp);
    temp = nullptr;
    temp_cap = 0;
    sk_cap = 0;
    drop(slot_scratch);
    slot_scratch = nullptr;
    slot_cap = 0;
  }

  // the kept slot lists of a launch whose first evaluation wrote them into the
  // scratch rows (inputs from row 0, outputs from row kAdMaxIn): trimmed to the
  // largest counts
  void finish_slots(SlotPlan& sp, int64_t E) {
    if (!pair) pair = static_cast<int*>(grab(2 * sizeof(int)));
    if (!pair) return;
    ex.note(hipMemsetAsync(pair, 0, 2 * sizeof(int), ex.s));
    hipLaunchKernelGGL(k_max_pair, dim3(stride_blocks(E)), dim3(256), 0, ex.s, sp.nin, sp.nout, E, pair);
    ex.note(hipGetLastError());
    int m[2] = {0, 0};
    ex.note(hipMemcpyAsync(m, pair, sizeof(m), hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
    if (ex.last != hipSuccess) return;
    sp.slots = static_cast<uint64_t*>(grab(sizeof(uint64_t) * size_t(E) * size_t(m[0] + m[1] > 0 ? m[0] + m[1] : 1)));
    if (!sp.slots) return;
    if (m[0] > 0)
      ex.note(hipMemcpyAsync(sp.slots, slot_scratch, sizeof(uint64_t) * size_t(E) * size_t(m[0]),
                             hipMemcpyDeviceToDevice, ex.s));
    if (m[1] > 0)
      ex.note(hipMemcpyAsync(sp.slots + int64_t(m[0]) * E, slot_scratch + int64_t(kAdMaxIn) * E,
                             sizeof(uint64_t) * size_t(E) * size_t(m[1]), hipMemcpyDeviceToDevice, ex.s));
    sp.rin = m[0];
  }
  std::string det_err;
  const char* error() {
    if (!det_err.empty()) return det_err.c_str();
    return HipBackend::error();
  }

  void ensure_flag() {
    if (!adj_flag) {
      ex.note(hipMalloc(reinterpret_cast<void**>(&adj_flag), sizeof(int)));
      ex.note(hipMemsetAsync(adj_flag, 0, sizeof(int), ex.s));
    }
  }

  // the plan's targets, run starts and places from the keys the first launch wrote
  // into sk[0]
  // bits of a key: target indices and the unused-place key arena_len
  int key_bits() const {
    int bits = 1;
    while (bits < 32 && (int64_t(1) << bits) <= arena_len) ++bits;
    return bits;
  }
  void build_plan(Plan& pl, int64_t n) {
    const int bits = key_bits();
    const int ni = int(n);
    hipLaunchKernelGGL(k_iota_u32, dim3(stride_blocks(n)), dim3(256), 0, ex.s, sk[2], n);
    ex.note(hipGetLastError());
    size_t tb = temp_cap;
    ex.note(hipcub::DeviceRadixSort::SortPairs(temp, tb, sk[0], sk[1], sk[2], sk[3], ni, 0, bits, ex.s));
    tb = temp_cap;
    ex.note(hipcub::DeviceRunLengthEncode::Encode(temp, tb, sk[1], sk[4], sk[5], nruns, ni, ex.s));
    int runs = 0;
    ex.note(hipMemcpyAsync(&runs, nruns, sizeof(int), hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
    if (ex.last != hipSuccess || runs <= 0) return;
    uint32_t lastkey = 0;
    ex.note(hipMemcpyAsync(&lastkey, sk[4] + (runs - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
    // places no element used carry the key arena_len and sort last
    const int64_t nu = lastkey == uint32_t(arena_len) ? runs - 1 : runs;
    pl.nuniq = nu;
    pl.off = static_cast<uint32_t*>(grab(sizeof(uint32_t) * size_t(nu + 1)));
    pl.uid = static_cast<uint32_t*>(grab(sizeof(uint32_t) * size_t(nu > 0 ? nu : 1)));
    if (!pl.off || !pl.uid) return;
    tb = temp_cap;
    ex.note(hipcub::DeviceScan::ExclusiveSum(temp, tb, sk[5], pl.off, int(nu + 1), ex.s));
    uint32_t total = 0;
    ex.note(hipMemcpyAsync(&total, pl.off + nu, sizeof(uint32_t), hipMemcpyDeviceToHost, ex.s));
    ex.note(hipStreamSynchronize(ex.s));
    pl.pos = static_cast<uint32_t*>(grab(sizeof(uint32_t) * size_t(total > 0 ? total : 1)));
    if (!pl.pos) return;
    if (nu > 0) ex.note(hipMemcpyAsync(pl.uid, sk[4], sizeof(uint32_t) * size_t(nu), hipMemcpyDeviceToDevice, ex.s));
    if (total > 0)
      ex.note(hipMemcpyAsync(pl.pos, sk[3], sizeof(uint32_t) * size_t(total), hipMemcpyDeviceToDevice, ex.s));
  }

  // scratch for n places (sorts of up to n pairs)
  void ensure_scratch(int64_t n) {
    if (n <= sk_cap) return;
    for (int q = 0; q < 6; ++q) {
      drop(sk[q]);
      sk[q] = static_cast<uint32_t*>(grab(sizeof(uint32_t) * size_t(n + 1)));
    }
    if (!nruns) nruns = static_cast<int*>(grab(sizeof(int)));
    size_t a = 0, b = 0, c = 0;
    const int ni = int(n);
    ex.note(hipcub::DeviceRadixSort::SortPairs(nullptr, a, sk[0], sk[1], sk[2], sk[3], ni, 0, key_bits(), ex.s));
    ex.note(hipcub::DeviceRunLengthEncode::Encode(nullptr, b, sk[1], sk[4], sk[5], nruns, ni, ex.s));
    ex.note(hipcub::DeviceScan::ExclusiveSum(nullptr, c, sk[5], sk[4], ni + 1, ex.s));
    const size_t need = std::max(a, std::max(b, c));
    if (need > temp_cap) {
      drop(temp);
      temp = grab(need);
      temp_cap = need;
    }
    sk_cap = n;
  }

  template <class Ph, class T>
  void adjoint_phase(const Ctx<T>& c, const Ctx<double>& a) {
    if (!det) {
      HipBackend::adjoint_phase<Ph>(c, a);
      return;
    }
    ensure_flag();
    int64_t n[3];
    c.extent(Ph::kind, Ph::dir, n);
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return;
    const int64_t E = n[0] * n[1] * n[2];
    const int64_t places = E * kAdjMaxContrib;
    if (places >= (int64_t(1) << 31) || arena_len >= (int64_t(1) << 32) - 1) {
      det_err = "deterministic adjoint accumulation supports up to 13.4M elements per launch (switch it off "
                "with pmhd_adjoint_set_deterministic for larger grids)";
      return;
    }
    std::vector<Plan>& pv = plans[sel_par];
    if (pv.size() <= sel_t) pv.resize(sel_t + 1);
    Plan& pl = pv[sel_t];
    const bool fresh = pl.id != int(Ph::id) || pl.E != E;
    if (fresh) {
      drop(pl.uid);
      drop(pl.off);
      drop(pl.pos);
      drop(pl.cnt);
      drop(pl.hash);
      pl = Plan();
      pl.id = int(Ph::id);
      pl.E = E;
      pl.cnt = static_cast<int32_t*>(grab(sizeof(int32_t) * size_t(E)));
      pl.hash = static_cast<uint32_t*>(grab(sizeof(uint32_t) * size_t(E)));
      ensure_scratch(places);
    }
    if (places > vals_cap) {
      drop(vals);
      vals = static_cast<double*>(grab(sizeof(double) * size_t(places)));
      vals_cap = places;
    }
    if (ex.last != hipSuccess || !vals || !pl.cnt || !pl.hash) return;
    // the launch's slot lists: kept from its first evaluation, else written now
    SlotBufs sb{nullptr, nullptr, nullptr, 0, false, false};
    SlotPlan* sp = nullptr;
    bool sfresh = false;
    if constexpr (uses_slots<Ph>()) {
      if (splans.size() <= sel_t) splans.resize(sel_t + 1);
      sp = &splans[sel_t];
      sfresh = sp->id != int(Ph::id) || sp->E != E;
      if (sfresh) {
        drop(sp->slots);
        drop(sp->nin);
        drop(sp->nout);
        *sp = SlotPlan();
        sp->nin = static_cast<int32_t*>(grab(sizeof(int32_t) * size_t(E)));
        sp->nout = static_cast<int32_t*>(grab(sizeof(int32_t) * size_t(E)));
        const int64_t need = E * (kAdMaxIn + kAdMaxOut);
        if (need > slot_cap) {
          drop(slot_scratch);
          slot_scratch = static_cast<uint64_t*>(grab(sizeof(uint64_t) * size_t(need)));
          slot_cap = need;
        }
        if (ex.last != hipSuccess || !sp->nin || !sp->nout || !slot_scratch) return;
        sb = SlotBufs{slot_scratch, sp->nin, sp->nout, kAdMaxIn, true, true};
      } else {
        sb = SlotBufs{sp->slots, sp->nin, sp->nout, sp->rin, false, true};
      }
    }
    DetSink s{vals, nullptr, pl.cnt, pl.hash, arena, arena_len, E, adj_flag};
    if (fresh) {
      s.keys = sk[0];
      hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(places)), dim3(256), 0, ex.s, sk[0], uint32_t(arena_len),
                         places);
      ex.note(hipGetLastError());
    }
    unsigned bx = 1;
    while (bx < 64 && int64_t(bx) < n[2]) bx <<= 1;
    const unsigned by = 64 / bx;
    dim3 grid(unsigned((n[2] + bx - 1) / bx), unsigned((n[1] + by - 1) / by), unsigned(n[0]));
    bool launched = false;
    if constexpr ((is_correct<Ph>::value || is_predict<Ph>::value) && !uses_slots<Ph>()) {
      if (!fresh) {
        ex.timed(K_ADJOINT, E, [&] {
          hipLaunchKernelGGL((k_adjoint_corr2<Ph, T>), dim3(unsigned((2 * E + 63) / 64)), dim3(64), 0, ex.s, c, a,
                             int(n[1]), int(n[2]), E, s);
        });
        launched = true;
      }
    }
    if constexpr (uses_slots<Ph>() && !is_edges_own<Ph>::value) {
      if (!fresh && !sfresh) {
        ex.timed(K_ADJOINT, E, [&] {
          hipLaunchKernelGGL((k_adjoint_kept2<Ph, T>), dim3(unsigned((2 * E + 63) / 64)), dim3(64), 0, ex.s, c, a,
                             int(n[1]), int(n[2]), E, vals, sb);
        });
        launched = true;
      }
    }
    if (!launched) {
      ex.timed(K_ADJOINT, E, [&] {
        hipLaunchKernelGGL((k_adjoint_det<Ph, T>), grid, dim3(bx, by, 1), 0, ex.s, c, a, int(n[1]), int(n[2]), s, sb);
      });
    }
    ex.note(hipGetLastError());
    if (sfresh) {
      finish_slots(*sp, E);
      if (ex.last != hipSuccess) return;
      if (!sp->slots) {
        det_err = "could not keep the adjoint's slot lists";
        return;
      }
      sp->id = int(Ph::id);
      sp->E = E;
    }
    if (fresh) build_plan(pl, places);
    if (ex.last != hipSuccess) return;
    if (!pl.pos) {
      det_err = "could not build the adjoint's accumulation plan";
      return;
    }
    if (pl.nuniq > 0) {
      ex.timed(K_ADJOINT, E, [&] {
        hipLaunchKernelGGL(k_det_gather, dim3(unsigned((pl.nuniq + 255) / 256)), dim3(256), 0, ex.s, arena, vals,
                           pl.uid, pl.off, pl.pos, pl.nuniq);
      });
      ex.note(hipGetLastError());
    }
  }
};

PMHD_NS_END
