// Forward-mode dual numbers and the memory interface that lets the step's own
// kernel bodies form local vector-Jacobian products for the adjoint.
//
// A phase of the step computes each output element from a small stencil of
// input array values.  One recording pass runs the body on Ctx<ARec<V>> (plain
// values), whose array accesses note the (array, index) slots the element reads
// and writes.  Then one seeded pass per input slot runs it on Ctx<ADual<V>>, whose
// array accesses go through an AdPtr: a load yields the stored value with tangent 1
// when it is the pass's seed slot (0 otherwise), and a store accumulates
// cotangent(output) * tangent into the pass's accumulator, which gives
// sum_outputs cotangent * d output / d slot, the slot's share of the
// vector-Jacobian product.  The value part of every
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

// The eigen-system of a dual body.  When none of its inputs carries a tangent (a
// seeded pass whose slot is not one of the cell's own values), the dual evaluation
// would give the plain-value eigen-system with zero tangents: computed so, at the
// plain cost.  Otherwise the generic dual evaluation.  (Not in the all-dual test
// build, PMHD_ADJ_DUAL_CORRECTOR, which checks this against the dual evaluation.)
#ifndef PMHD_ADJ_DUAL_CORRECTOR
template <class V>
PMHD_FI void eigen_prim(const ADual<V> w[7], ADual<V> bn, ADual<V> gamma, PrimEig<ADual<V>>& E) {
  bool seeded = bn.t != V(0) || gamma.t != V(0);
#pragma unroll
  for (int k = 0; k < 7; ++k) seeded = seeded || w[k].t != V(0);
  if (seeded) {
    eigen_prim<ADual<V>>(w, bn, gamma, E);
    return;
  }
  V wv[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) wv[k] = w[k].v;
  PrimEig<V> P;
  eigen_prim<V>(wv, bn.v, gamma.v, P);
  constexpr int kFields = sizeof(PrimEig<V>) / sizeof(V);
  static_assert(sizeof(PrimEig<V>) == kFields * sizeof(V), "PrimEig holds only numbers");
  static_assert(sizeof(PrimEig<ADual<V>>) == kFields * sizeof(ADual<V>), "PrimEig holds only numbers");
  const V* src = reinterpret_cast<const V*>(&P);
  ADual<V>* dst = reinterpret_cast<ADual<V>*>(&E);
#pragma unroll
  for (int k = 0; k < kFields; ++k) dst[k] = ADual<V>(src[k], V(0), 0);
}
#endif

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

// the arrays of a context carry ids 0 .. kAdIds-1 (make_ctx_of); a slot is an
// array id and an index into that array
constexpr int kAdIds = 32;

struct AdSlot {
  int id;
  int64_t i;
};

// the slot tables of one element (indexed at run time: they live in scratch memory)
struct AdState {
  int nin, nout;
  int overflow;
  AdSlot in[kAdMaxIn];
  AdSlot out[kAdMaxOut];
};

// what every load and store of a seeded pass reads or updates: kept apart from the
// slot tables, as a few scalars the compiler can hold in registers
struct AdHot {
  int seed_id;      // the seeded slot (array id and index)
  int64_t seed_i;
  double acc;       // sum over stores of cotangent * tangent
};

PMHD_FI bool slot_has(const AdSlot* s, int n, int id, int64_t i) {
  for (int k = 0; k < n; ++k)
    if (s[k].id == id && s[k].i == i) return true;
  return false;
}

template <class V>
struct AdRef;

// an array of a dual context: values, cotangents and the element state
template <class V>
struct AdPtr {
  const V* val;
  double* adj;
  AdHot* hot;
  int64_t off;
  int id;
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
    this->t = (h.seed_id == p.id && h.seed_i == i) ? V(1) : V(0);
  }
  PMHD_FI AdRef& operator=(const ADual<V>& x) {
    p.hot->acc += p.adj[i] * double(x.t);
    return *this;
  }
  PMHD_FI AdRef& operator=(const AdRef& x) { return *this = static_cast<const ADual<V>&>(x); }
};

template <class V>
struct Mem<ADual<V>> {
  using cp = AdPtr<V>;
  using p = AdPtr<V>;
};

// ---------------------------------------------------------------------------
// the recording pass: the body runs on plain values (operation for operation the
// value parts of the seeded passes, so it takes the same branches and reads the
// same slots), and every array access notes its slot in the element state
// ---------------------------------------------------------------------------
template <class V>
struct ARec {
  V v;
  PMHD_HD ARec() : v(0) {}
  PMHD_HD ARec(double x) : v(V(x)) {}
  template <class W>
  PMHD_HD explicit ARec(const ARec<W>& o) : v(V(o.v)) {}
  PMHD_HD explicit operator double() const { return double(v); }
  static PMHD_FI ARec of(V x) {
    ARec r;
    r.v = x;
    return r;
  }
  friend PMHD_FI ARec operator+(const ARec& a, const ARec& b) { return of(a.v + b.v); }
  friend PMHD_FI ARec operator-(const ARec& a, const ARec& b) { return of(a.v - b.v); }
  friend PMHD_FI ARec operator*(const ARec& a, const ARec& b) { return of(a.v * b.v); }
  friend PMHD_FI ARec operator/(const ARec& a, const ARec& b) { return of(a.v / b.v); }
  friend PMHD_FI ARec operator-(const ARec& a) { return of(-a.v); }
  friend PMHD_FI bool operator<(const ARec& a, const ARec& b) { return a.v < b.v; }
  friend PMHD_FI bool operator<=(const ARec& a, const ARec& b) { return a.v <= b.v; }
  friend PMHD_FI bool operator>(const ARec& a, const ARec& b) { return a.v > b.v; }
  friend PMHD_FI bool operator>=(const ARec& a, const ARec& b) { return a.v >= b.v; }
  friend PMHD_FI bool operator==(const ARec& a, const ARec& b) { return a.v == b.v; }
  friend PMHD_FI bool operator!=(const ARec& a, const ARec& b) { return a.v != b.v; }
  PMHD_FI ARec& operator+=(const ARec& b) { return *this = *this + b; }
  PMHD_FI ARec& operator-=(const ARec& b) { return *this = *this - b; }
};
template <class V>
PMHD_FI ARec<V> sqrt_(const ARec<V>& a) {
  return ARec<V>::of(sqrt_(a.v));
}
template <class V>
PMHD_FI ARec<V> fabs_(const ARec<V>& a) {
  return ARec<V>::of(fabs_(a.v));
}
template <class V>
struct Island<ARec<V>> {
  using type = ARec<double>;
};

template <class V>
struct RecRef;

template <class V>
struct RecPtr {
  const V* val;
  AdState* st;
  int64_t off;
  int id;
  PMHD_FI RecRef<V> operator[](int64_t i) const { return RecRef<V>(*this, off + i); }
  PMHD_FI RecPtr operator+(int64_t n) const {
    RecPtr r = *this;
    r.off += n;
    return r;
  }
};

template <class V>
struct RecRef : ARec<V> {
  RecPtr<V> p;
  int64_t i;
  PMHD_FI RecRef(const RecPtr<V>& p_, int64_t i_) : ARec<V>(), p(p_), i(i_) {
    this->v = p.val[i];
    AdState& s = *p.st;
    if (!slot_has(s.in, s.nin, p.id, i)) {
      if (s.nin < kAdMaxIn) s.in[s.nin++] = AdSlot{p.id, i};
      else s.overflow = 1;
    }
  }
  PMHD_FI RecRef& operator=(const ARec<V>&) {
    AdState& s = *p.st;
    if (!slot_has(s.out, s.nout, p.id, i)) {
      if (s.nout < kAdMaxOut) s.out[s.nout++] = AdSlot{p.id, i};
      else s.overflow = 1;
    }
    return *this;
  }
  PMHD_FI RecRef& operator=(const RecRef& x) { return *this = static_cast<const ARec<V>&>(x); }
};

template <class V>
struct Mem<ARec<V>> {
  using cp = RecPtr<V>;
  using p = RecPtr<V>;
};

// after the recording pass: an element's input slots are the slots it read and
// did not itself write (a body may read back its own outputs)
PMHD_FI void ad_prune(AdState& s) {
  int n = 0;
  for (int k = 0; k < s.nin; ++k)
    if (!slot_has(s.out, s.nout, s.in[k].id, s.in[k].i)) s.in[n++] = s.in[k];
  s.nin = n;
}

PMHD_NS_END
