// Host (CPU) build of the product's kernel bodies and C ABI -- TEST
// INFRASTRUCTURE ONLY.  It compiles pythonmhd_amd/csrc/*.h with g++ so the
// FP64 arithmetic of the HIP kernels can be checked against the oracle in a
// container without a GPU.  The product package never loads this library.
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "pmhd_api.h"

namespace pmhd {

struct HostExec {
  template <class F>
  void run3(int, int64_t n0, int64_t n1, int64_t n2, F f) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t j = 0; j < n0; ++j)
      for (int64_t i = 0; i < n1; ++i)
        for (int64_t k = 0; k < n2; ++k) f(I3{j, i, k});
  }
  template <class F>
  void reduce(int, int64_t n, F f, double sums[9], double maxes[2]) {
    for (int q = 0; q < 9; ++q) sums[q] = 0;
    maxes[0] = maxes[1] = 0;
    for (int64_t i = 0; i < n; ++i) {
      double a[9], b[2];
      f(i, a, b);
      for (int q = 0; q < 9; ++q) sums[q] += a[q];
      maxes[0] = npmax(maxes[0], b[0]);
      maxes[1] = npmax(maxes[1], b[1]);
    }
  }
  template <class F>
  void max3(int, int64_t n, F f, double out[3]) {
    uint64_t m[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
      double o[3];
      f(i, o);
      for (int d = 0; d < 3; ++d) {
        uint64_t b;
        std::memcpy(&b, &o[d], 8);
        if (b > m[d]) m[d] = b;
      }
    }
    for (int d = 0; d < 3; ++d) std::memcpy(&out[d], &m[d], 8);
  }
};

struct HostBackend {
  HostExec ex;
  HostExec& exec() { return ex; }
  void* alloc(size_t bytes) { return std::calloc(bytes, 1); }
  void release(void* p) { std::free(p); }
  void h2d(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
  void d2h(void* d, const void* s, size_t b) { std::memcpy(d, s, b); }
  void sync() {}
  const char* error() { return nullptr; }
  void set_profiling(bool) {}
  int profile(int, char*, double*, int64_t*, int64_t*) { return 0; }
};

}  // namespace pmhd

PMHD_DEFINE_C_ABI(pmhd::HostBackend)
