// Forward-mode dual numbers and the memory interface that lets the step's own
// kernel bodies form local vector-Jacobian products for the adjoint.
//
// A phase of the step computes each output element from a small stencil of
// input array values.  Run on Ctx<ADual<V>>, every array access of a body goes
// through an AdPtr: a load yields the stored value with tangent 1 when it is the
// element's current seed slot (0 otherwise), and in recording mode it notes the
// (array, index) it read; a store accumulates  cotangent(output) * tangent  into
// the element's accumulator.  One recording pass lists the element's input slots;
// one seeded pass per slot then gives  sum_outputs cotangent * d output / d slot,
// the slot's share of the vector-Jacobian product.  The value part of every
// operation is the plain V arithmetic in the forward order, so each pass sees the
// forward values exactly.
#pragma once
#include <cmath>
#include <cstdint>

#include "pmhd_math.h"

PMHD_NS_BEGIN

template <class V>
struct ADual {
  V v, t;
  PMHD_HD ADual() : v(0), t(0) {}
  PMHD_HD ADual(double x) : v(V(x)), t(0) {}
  PMHD_HD ADual(V x, V dx, int) : v(x), t(dx) {}
  template <class W>
  PMHD_HD explicit ADual(const ADual<W>& o) : v(V(o.v)), t(V(o.t)) {}
  PMHD_HD explicit operator double() const { return double(v); }

  friend PMHD_FI ADual operator+(const ADual& a, const ADual& b) { return ADual(a.v + b.v, a.t + b.t, 0); }
  friend PMHD_FI ADual operator-(const ADual& a, const ADual& b) { return ADual(a.v - b.v, a.t - b.t, 0); }
  friend PMHD_FI ADual operator*(const ADual& a, const ADual& b) {
    return ADual(a.v * b.v, a.t * b.v + a.v * b.t, 0);
  }
  friend PMHD_FI ADual operator/(const ADual& a, const ADual& b) {
    const V q = a.v / b.v;
    return ADual(q, (a.t - q * b.t) / b.v, 0);
  }
  friend PMHD_FI ADual operator-(const ADual& a) { return ADual(-a.v, -a.t, 0); }
  friend PMHD_FI bool operator<(const ADual& a, const ADual& b) { return a.v < b.v; }
  friend PMHD_FI bool operator<=(const ADual& a, const ADual& b) { return a.v <= b.v; }
  friend PMHD_FI bool operator>(const ADual& a, const ADual& b) { return a.v > b.v; }
  friend PMHD_FI bool operator>=(const ADual& a, const ADual& b) { return a.v >= b.v; }
  friend PMHD_FI bool operator==(const ADual& a, const ADual& b) { return a.v == b.v; }
  friend PMHD_FI bool operator!=(const ADual& a, const ADual& b) { return a.v != b.v; }
  PMHD_FI ADual& operator+=(const ADual& b) { return *this = *this + b; }
  PMHD_FI ADual& operator-=(const ADual& b) { return *this = *this - b; }
};

// at 0 the derivative is taken as 0 (the step's square roots vanish only at its
// degenerate points, e.g. |B_t| = 0, where it switches to a fixed branch)
template <class V>
PMHD_FI ADual<V> sqrt_(const ADual<V>& a) {
  const V s = sqrt_(a.v);
  return ADual<V>(s, s > V(0) ? a.t / (V(2) * s) : V(0), 0);
}
template <class V>
PMHD_FI ADual<V> fabs_(const ADual<V>& a) {
  return a.v < V(0) ? ADual<V>(-a.v, -a.t, 0) : ADual<V>(fabs_(a.v), a.t, 0);
}

// the wave-speed islands of a dual body run on dual doubles
template <class V>
struct Island<ADual<V>> {
  using type = ADual<double>;
};

// ---------------------------------------------------------------------------
// element state of one adjoint evaluation
// ---------------------------------------------------------------------------
constexpr int kAdMaxIn = 160;
constexpr int kAdMaxOut = 40;

struct AdSlot {
  const void* base;   // value array (identity)
  double* adj;        // its cotangent array
  int64_t i;
};

// the slot tables of one element (indexed at run time: they live in scratch memory)
struct AdState {
  int nin, nout;
  int overflow;
  AdSlot in[kAdMaxIn];
  AdSlot out[kAdMaxOut];
};

// what every load and store of a pass reads or updates: kept apart from the slot
// tables, as a few scalars the compiler can hold in registers
struct AdHot {
  int mode;                // 0: record the accessed slots; 1: seeded pass
  const void* seed_base;   // the seeded slot (array identity and index)
  int64_t seed_i;
  double acc;              // sum over stores of cotangent * tangent
};

PMHD_FI bool slot_has(const AdSlot* s, int n, const void* base, int64_t i) {
  for (int k = 0; k < n; ++k)
    if (s[k].base == base && s[k].i == i) return true;
  return false;
}

template <class V>
struct AdRef;

// an array of a dual context: values, cotangents and the element state
template <class V>
struct AdPtr {
  const V* val;
  double* adj;
  AdState* st;
  AdHot* hot;
  int64_t off;
  PMHD_FI AdRef<V> operator[](int64_t i) const { return AdRef<V>(*this, off + i); }
  PMHD_FI AdPtr operator+(int64_t n) const {
    AdPtr r = *this;
    r.off += n;
    return r;
  }
};

template <class V>
struct AdRef : ADual<V> {
  AdPtr<V> p;
  int64_t i;
  PMHD_FI AdRef(const AdPtr<V>& p_, int64_t i_) : ADual<V>(), p(p_), i(i_) {
    const AdHot& h = *p.hot;
    this->v = p.val[i];
    this->t = V(0);
    if (h.mode == 0) {
      AdState& s = *p.st;
      if (!slot_has(s.in, s.nin, p.val, i)) {
        if (s.nin < kAdMaxIn) s.in[s.nin++] = AdSlot{p.val, p.adj, i};
        else s.overflow = 1;
      }
    } else if (h.seed_base == p.val && h.seed_i == i) {
      this->t = V(1);
    }
  }
  PMHD_FI AdRef& operator=(const ADual<V>& x) {
    AdHot& h = *p.hot;
    if (h.mode == 0) {
      AdState& s = *p.st;
      if (!slot_has(s.out, s.nout, p.val, i)) {
        if (s.nout < kAdMaxOut) s.out[s.nout++] = AdSlot{p.val, p.adj, i};
        else s.overflow = 1;
      }
    } else {
      h.acc += p.adj[i] * double(x.t);
    }
    return *this;
  }
  PMHD_FI AdRef& operator=(const AdRef& x) { return *this = static_cast<const ADual<V>&>(x); }
};

template <class V>
struct Mem<ADual<V>> {
  using cp = AdPtr<V>;
  using p = AdPtr<V>;
};

// after the recording pass: an element's input slots are the slots it read and
// did not itself write (a body may read back its own outputs)
PMHD_FI void ad_prune(AdState& s) {
  int n = 0;
  for (int k = 0; k < s.nin; ++k)
    if (!slot_has(s.out, s.nout, s.in[k].base, s.in[k].i)) s.in[n++] = s.in[k];
  s.nin = n;
}

PMHD_NS_END
