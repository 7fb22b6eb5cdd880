// Per-cell / per-face arithmetic of the CTU+CT step, shared by the HIP kernels
// (compiled by hipcc for gfx950) and the host test build.
//
// Every expression keeps the operation order of PythonMHD's NumPy code so that the
// FP64 instantiation reproduces it bit for bit (compiled with -ffp-contract=off).
// The FP32 instantiation uses the same code with T = float; the wave-speed and
// eigenvector-normalisation block ("islands") is always evaluated in double.
//
// Vectors are in the rotated frame (rho, v_n, v_t1, v_t2, B_t1, B_t2, p) with
// (n, t1, t2) cyclic in (x, y, z); the reference writes the x, y and z variants
// out with the same operation order in that frame
// (Source/Reconstruction_PPM_MHD_HelperModule.py:341-373 / :668-700 / :990-1005,
//  Source/RiemannSolvers_MHD_Roe_HelperModule.py:505-512 / :1015-1022 / :1503-1510).
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PMHD_HD __host__ __device__
#else
#define PMHD_HD
#endif
#define PMHD_FI PMHD_HD inline __attribute__((always_inline))

// Each translation unit of the library (FP64 step, FP32 step, adjoint, C ABI) is
// compiled with its own floating-point flags; its copy of the code lives in its
// own inline namespace so that no inline function of one unit is merged with the
// same-named function of another at link time.
#ifndef PMHD_TU_NS
#define PMHD_TU_NS unit
#endif
#define PMHD_NS_BEGIN namespace pmhd { inline namespace PMHD_TU_NS {
#define PMHD_NS_END } }

PMHD_NS_BEGIN

constexpr double kSmall = 0.000000000001;        // Source/PythonMHD_Constants.py:7
constexpr double kFourThirds = 1.333333333333333;  // Source/PythonMHD_Constants.py:8

using N = double;  // precision of the wave-speed / normalisation islands
// island type of a value type (dual numbers carry their island as dual doubles)
template <class T>
struct Island {
  using type = double;
};
template <class T>
using island_t = typename Island<T>::type;
// true for plain double (not for a dual number of the same size)
template <class T>
PMHD_FI constexpr bool is_f64() { return false; }
template <>
PMHD_FI constexpr bool is_f64<double>() { return true; }

// NumPy element semantics: maximum/minimum propagate NaN and return the second
// argument on ties (so the sign of zero follows NumPy); sign(0) is +0.
template <class T> PMHD_FI T npmax(T a, T b) { return (a > b || a != a) ? a : b; }
template <class T> PMHD_FI T npmin(T a, T b) { return (a < b || a != a) ? a : b; }
template <class T> PMHD_FI T npsign(T x) {
  return x > T(0) ? T(1) : (x < T(0) ? T(-1) : (x == T(0) ? T(0) : x));
}
#ifdef PMHD_FAST_ORDER
// the FP32-bulk unit: single-instruction maxima and minima (they differ from NumPy's
// only for NaN operands and the sign of a zero result)
PMHD_FI float npmax(float a, float b) { return fmaxf(a, b); }
PMHD_FI float npmin(float a, float b) { return fminf(a, b); }
#endif
template <class T> PMHD_FI T sq(T x) { return x * x; }
PMHD_FI double sqrt_(double x) { return sqrt(x); }
PMHD_FI float sqrt_(float x) { return sqrtf(x); }
// reciprocal square root: the FP32 fast unit uses the hardware estimate where one
// root and one quotient of the same value are both needed
template <class T> PMHD_FI T rsqrt_(T x) { return T(1) / sqrt_(x); }
PMHD_FI float rsqrt_(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / std::sqrt(x);
#endif
}
// true in the FP32 fast unit for float arithmetic (reassociated forms allowed)
template <class T> PMHD_FI constexpr bool fast_f32() {
#ifdef PMHD_FAST_ORDER
  return sizeof(T) == 4;
#else
  return false;
#endif
}
PMHD_FI double fabs_(double x) { return fabs(x); }
PMHD_FI float fabs_(float x) { return fabsf(x); }
template <class T> PMHD_FI T npabs(T x) { return fabs_(x); }

// ---------------------------------------------------------------------------
// primitive <-> conservative on a directional 7-vector
// (Source/PrimCons_HelperModule.py:208-452).  ke selects the order of the
// kinetic-energy sum: 0 -> (1,2,3), 1 -> (2,3,1), 2 -> (3,1,2).
// ---------------------------------------------------------------------------
// |v|^2 of slots 1..3 summed from slot a = 1 + ke cyclically (the order is uniform
// over a launch: written as branches, not as a run-time index into the registers)
template <class T>
PMHD_FI T ke_sum(const T v[7], int ke) {
  if (ke == 0) return v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
  if (ke == 1) return v[2] * v[2] + v[3] * v[3] + v[1] * v[1];
  return v[3] * v[3] + v[1] * v[1] + v[2] * v[2];
}

template <class T>
PMHD_FI void prim_to_cons(const T w[7], T bn, T gamma, int ke, T u[7]) {
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[0] * w[3];
  u[4] = w[4];
  u[5] = w[5];
  T e = w[6] / (gamma - T(1)) + T(0.5) * w[0] * ke_sum(w, ke);
  u[6] = e + T(0.5) * (bn * bn + w[4] * w[4] + w[5] * w[5]);
}

template <class T>
PMHD_FI void cons_to_prim(const T u[7], T bn, T gamma, int ke, T w[7]) {
  T inv = T(1) / u[0];
  w[0] = u[0];
  w[1] = u[1] * inv;
  w[2] = u[2] * inv;
  w[3] = u[3] * inv;
  w[4] = u[4];
  w[5] = u[5];
  T p = u[6] - T(0.5) * ke_sum(u, ke) * inv;
  p = p - T(0.5) * (bn * bn + w[4] * w[4] + w[5] * w[5]);
  w[6] = p * (gamma - T(1));
}

// floors on density and on the last component (Source/PrimCons_HelperModule.py:499-554)
template <class T>
PMHD_FI void apply_floors(T v[7], T min_dens, T min_last) {
  v[0] = v[0] < min_dens ? min_dens : v[0];
  v[6] = v[6] < min_last ? min_last : v[6];
}

// ---------------------------------------------------------------------------
// CFL: per-cell fast speeds (Source/TimestepCalculations_HelperModule.py:305-361)
// b_cell: cell B (cons[4..6]); b_face: left-face B of the cell (ignored when
// use_face == false, as in the 1D step).
// ---------------------------------------------------------------------------
namespace hydro {
// per-cell hydro CFL speeds |v_d| + c_s (Source/TimestepCalculations_HelperModule.py:24-170)
template <class T>
PMHD_FI void cell_speeds(const T u[5], T gamma, T out[3]) {
  const T inv = T(1) / u[0];
  const T vx = u[1] * inv, vy = u[2] * inv, vz = u[3] * inv;
  const T vsq = vx * vx + vy * vy + vz * vz;
  const T pres = npmax((gamma - T(1)) * (u[4] - T(0.5) * u[0] * vsq), T(kSmall));
  const T cs = sqrt_(gamma * pres * inv);
  out[0] = npabs(vx) + cs;
  out[1] = npabs(vy) + cs;
  out[2] = npabs(vz) + cs;
}
}  // namespace hydro

template <class T>
PMHD_FI T fast_speed_sq(T bsq, T inv, T cs2, T btsq) {
  T ssum = bsq * inv + cs2;
  T sdiff = bsq * inv - cs2;
  return T(0.5) * (ssum + sqrt_(sdiff * sdiff + T(4) * cs2 * btsq * inv));
}

template <class T>
PMHD_FI void cell_speeds(const T u[8], const T bf[3], int ndim, T gamma, T out[3]) {
  T g1 = gamma - T(1);
  T inv = T(1) / u[0];
  T vx = u[1] * inv, vy = u[2] * inv, vz = u[3] * inv;
  T vsq = vx * vx + vy * vy + vz * vz;
  T bx = u[4], by = u[5], bz = u[6];
  if (ndim >= 2) {
    bx = u[4] + fabs_(bf[0] - u[4]);
    by = u[5] + fabs_(bf[1] - u[5]);
  }
  if (ndim == 3) bz = u[6] + fabs_(bf[2] - u[6]);
  T bsq = bx * bx + by * by + bz * bz;
  T pres = npmax(g1 * (u[7] - T(0.5) * u[0] * vsq - T(0.5) * bsq), T(kSmall));
  T cs2 = gamma * pres * inv;
  out[0] = fabs_(vx) + sqrt_(fast_speed_sq(bsq, inv, cs2, by * by + bz * bz));
  out[1] = fabs_(vy) + sqrt_(fast_speed_sq(bsq, inv, cs2, bx * bx + bz * bz));
  out[2] = fabs_(vz) + sqrt_(fast_speed_sq(bsq, inv, cs2, bx * bx + by * by));
}

// ---------------------------------------------------------------------------
// primitive eigen-system, rotated frame
// (Source/ReconstructionEigenmatrices_MHD_HelperModule.py:27-207).
// Only the structurally non-zero entries are stored; rows of L are the waves.
// ---------------------------------------------------------------------------
template <class T>
struct PrimEig {
  T lam[7];
  // right eigenvectors (rows = variables, cols = waves)
  T r00, r02, r10, r12, r20, r21, r22, r30, r31, r32, r60, r62, r40, r41, r42, r50, r51, r52;
  // left eigenvectors (rows = waves)
  T l01, l02, l03, l06, l04, l05, l12, l13, l14, l15, l21, l22, l23, l26, l24, l25, l36;
};

template <class T>
PMHD_FI void eigen_prim(const T w[7], T bn, T gamma, PrimEig<T>& E) {
  using N = island_t<T>;
  const T rho = w[0], vn = w[1], bt1 = w[4], bt2 = w[5], p = w[6];
  T btsq = bt1 * bt1 + bt2 * bt2;
  T gp = gamma * p;
  // FP32 fast unit: one reciprocal root each gives sqrt(rho), 1/rho, 1/sqrt(rho)
  // and 1/|B_t|
  T inv, sqd, isqd = T(0), ibt = T(0);
  if constexpr (fast_f32<T>()) {
    isqd = rsqrt_(rho);
    sqd = rho * isqd;
    inv = isqd * isqd;
    ibt = rsqrt_(btsq);
  } else {
    inv = T(1) / rho;
    sqd = sqrt_(rho);
  }
  N n_cs2, n_ca2, n_cf2, n_csl2, af, as;
  {
    // wave speeds and normalisation in N (double) precision
#ifdef PMHD_FAST_ORDER
    // a relative error of 1/rho scales cs2, ca2 and bt2/rho alike, so every
    // difference below keeps its relative accuracy with the float reciprocal
    N n_inv = N(inv);
#else
    N n_inv = N(1) / N(rho);
#endif
    N n_rbtsq = N(btsq) * n_inv;
    n_cs2 = N(gp) * n_inv;
    n_ca2 = N(bn) * N(bn) * n_inv;
    N ssum = n_ca2 + n_rbtsq + n_cs2;
    N sdiff = n_ca2 + n_rbtsq - n_cs2;
    N ssqrt = sqrt_(sdiff * sdiff + N(4) * n_cs2 * n_rbtsq);
    n_cf2 = N(0.5) * (ssum + ssqrt);
    n_csl2 = n_cs2 * n_ca2 / n_cf2;
    if (ssqrt == N(0)) {
      af = N(1); as = N(0);
    } else if ((n_cs2 - n_csl2) <= N(0)) {
      af = N(0); as = N(1);
    } else if ((n_cf2 - n_cs2) <= N(0)) {
      af = N(1); as = N(0);
    } else {
#ifdef PMHD_FAST_ORDER
      if constexpr (sizeof(T) == 4) {
        // the differences carry the cancellation near the degeneracies and stay
        // double; quotient and root of these double-accurate values round in float
        const T iss = T(1) / T(ssqrt);
        af = N(sqrt_(T(n_cs2 - n_csl2) * iss));
        as = N(sqrt_(T(n_cf2 - n_cs2) * iss));
      } else {
        af = sqrt_((n_cs2 - n_csl2) / ssqrt);
        as = sqrt_((n_cf2 - n_cs2) / ssqrt);
      }
#else
      af = sqrt_((n_cs2 - n_csl2) / ssqrt);
      as = sqrt_((n_cf2 - n_cs2) / ssqrt);
#endif
    }
  }
  const T cs2 = T(n_cs2);
  const T cs = sqrt_(cs2);
  const T ca = sqrt_(T(n_ca2));
  const T cf = sqrt_(T(n_cf2));
  const T csl = sqrt_(T(n_csl2));
  const T Af_ = T(af), As_ = T(as);
  T b1, b2;
  if constexpr (fast_f32<T>()) {
    b1 = btsq == T(0) ? T(1) : (btsq > T(0) ? bt1 * ibt : T(0));
    b2 = btsq > T(0) ? bt2 * ibt : T(0);
  } else {
    T bt = sqrt_(btsq);
    b1 = bt == T(0) ? T(1) : (bt > T(0) ? bt1 / bt : T(0));
    b2 = bt > T(0) ? bt2 / bt : T(0);
  }
  T sg = npsign(bn);
  sg = sg == T(0) ? T(1) : sg;
  T qf = cf * Af_ * sg;
  T qs = csl * As_ * sg;
  T a_f = cs * Af_ * sqd;
  T a_s = cs * As_ * sqd;
  E.lam[0] = vn - cf; E.lam[1] = vn - ca; E.lam[2] = vn - csl; E.lam[3] = vn;
  E.lam[4] = vn + csl; E.lam[5] = vn + ca; E.lam[6] = vn + cf;
  E.r00 = rho * Af_;
  E.r02 = rho * As_;
  E.r10 = -cf * Af_;
  E.r12 = -csl * As_;
  E.r20 = qs * b1;
  E.r21 = -b2;
  E.r22 = -qf * b1;
  E.r30 = qs * b2;
  E.r31 = b1;
  E.r32 = -qf * b2;
  E.r60 = rho * cs2 * Af_;
  E.r62 = rho * cs2 * As_;
  E.r40 = a_s * b1;
  E.r41 = -b2 * sg * sqd;
  E.r42 = -a_f * b1;
  E.r50 = a_s * b2;
  E.r51 = b1 * sg * sqd;
  E.r52 = -a_f * b2;
  T ics2 = T(0);
  T nf;
  if constexpr (fast_f32<T>()) {
    ics2 = T(1) / cs2;
    nf = T(0.5) * ics2;
  } else {
    nf = T(0.5) / cs2;
  }
  T qf2 = nf * qf;
  T qs2 = nf * qs;
  T afp = nf * a_f * inv;
  T asp = nf * a_s * inv;
  E.l01 = -nf * cf * Af_;
  E.l02 = qs2 * b1;
  E.l03 = qs2 * b2;
  E.l06 = nf * Af_ * inv;
  E.l04 = asp * b1;
  E.l05 = asp * b2;
  E.l12 = T(-0.5) * b2;
  E.l13 = T(0.5) * b1;
  if constexpr (fast_f32<T>()) {
    E.l14 = T(-0.5) * b2 * sg * isqd;
    E.l15 = T(0.5) * b1 * sg * isqd;
  } else {
    E.l14 = T(-0.5) * b2 * sg / sqd;
    E.l15 = T(0.5) * b1 * sg / sqd;
  }
  E.l21 = T(-1) * nf * csl * As_;
  E.l22 = T(-1) * qf2 * b1;
  E.l23 = T(-1) * qf2 * b2;
  E.l26 = nf * As_ * inv;
  E.l24 = T(-1) * afp * b1;
  E.l25 = T(-1) * afp * b2;
  if constexpr (fast_f32<T>()) E.l36 = -ics2;
  else E.l36 = T(-1) / cs2;
}

// L . d with the reference's summation order (rho, v_n, v_t1, v_t2, p, B_t1, B_t2)
template <class T>
PMHD_FI void eig_project(const PrimEig<T>& E, const T d[7], T c[7]) {
  c[0] = T(0) + E.l01 * d[1] + E.l02 * d[2] + E.l03 * d[3] + E.l06 * d[6] + E.l04 * d[4] + E.l05 * d[5];
  c[1] = T(0) + E.l12 * d[2] + E.l13 * d[3] + E.l14 * d[4] + E.l15 * d[5];
  c[2] = T(0) + E.l21 * d[1] + E.l22 * d[2] + E.l23 * d[3] + E.l26 * d[6] + E.l24 * d[4] + E.l25 * d[5];
  c[3] = T(0) + d[0] + E.l36 * d[6];
  c[4] = T(0) + (-E.l21) * d[1] + (-E.l22) * d[2] + (-E.l23) * d[3] + E.l26 * d[6] + E.l24 * d[4] + E.l25 * d[5];
  c[5] = T(0) + (-E.l12) * d[2] + (-E.l13) * d[3] + E.l14 * d[4] + E.l15 * d[5];
  c[6] = T(0) + (-E.l01) * d[1] + (-E.l02) * d[2] + (-E.l03) * d[3] + E.l06 * d[6] + E.l04 * d[4] + E.l05 * d[5];
}

// row i of L (wave i) dotted with a vector, only the reference's term order
template <class T>
PMHD_FI T eig_row_dot(const PrimEig<T>& E, int i, const T v[7]) {
  switch (i) {
    case 0: return T(0) + E.l01 * v[1] + E.l02 * v[2] + E.l03 * v[3] + E.l06 * v[6] + E.l04 * v[4] + E.l05 * v[5];
    case 1: return T(0) + E.l12 * v[2] + E.l13 * v[3] + E.l14 * v[4] + E.l15 * v[5];
    case 2: return T(0) + E.l21 * v[1] + E.l22 * v[2] + E.l23 * v[3] + E.l26 * v[6] + E.l24 * v[4] + E.l25 * v[5];
    case 3: return T(0) + v[0] + E.l36 * v[6];
    case 4: return T(0) + (-E.l21) * v[1] + (-E.l22) * v[2] + (-E.l23) * v[3] + E.l26 * v[6] + E.l24 * v[4] + E.l25 * v[5];
    case 5: return T(0) + (-E.l12) * v[2] + (-E.l13) * v[3] + E.l14 * v[4] + E.l15 * v[5];
    default: return T(0) + (-E.l01) * v[1] + (-E.l02) * v[2] + (-E.l03) * v[3] + E.l06 * v[6] + E.l04 * v[4] + E.l05 * v[5];
  }
}

// add f * (column i of R) to x
template <class T>
PMHD_FI void eig_add_col(const PrimEig<T>& E, int i, T f, T x[7]) {
  switch (i) {
    case 0:
      x[0] += f * E.r00; x[1] += f * E.r10; x[2] += f * E.r20; x[3] += f * E.r30;
      x[4] += f * E.r40; x[5] += f * E.r50; x[6] += f * E.r60; break;
    case 1:
      x[2] += f * E.r21; x[3] += f * E.r31; x[4] += f * E.r41; x[5] += f * E.r51; break;
    case 2:
      x[0] += f * E.r02; x[1] += f * E.r12; x[2] += f * E.r22; x[3] += f * E.r32;
      x[4] += f * E.r42; x[5] += f * E.r52; x[6] += f * E.r62; break;
    case 3:
      x[0] += f * T(1); break;
    case 4:
      x[0] += f * E.r02; x[1] += f * (-E.r12); x[2] += f * (-E.r22); x[3] += f * (-E.r32);
      x[4] += f * E.r42; x[5] += f * E.r52; x[6] += f * E.r62; break;
    case 5:
      x[2] += f * (-E.r21); x[3] += f * (-E.r31); x[4] += f * E.r41; x[5] += f * E.r51; break;
    default:
      x[0] += f * E.r00; x[1] += f * (-E.r10); x[2] += f * (-E.r20); x[3] += f * (-E.r30);
      x[4] += f * E.r40; x[5] += f * E.r50; x[6] += f * E.r60; break;
  }
}

#ifdef PMHD_FAST_ORDER
// L . d from the waves' symmetric pairs: rows 0/6, 2/4 and 1/5 share their entries
// up to the sign of the velocity terms, so each pair is one sum and one difference
// of two partial dot products (23 operations instead of 34; rounds differently)
template <class T>
PMHD_FI void eig_project_pairs(const PrimEig<T>& E, const T d[7], T c[7]) {
  const T a0 = E.l01 * d[1] + E.l02 * d[2] + E.l03 * d[3];
  const T b0 = E.l06 * d[6] + E.l04 * d[4] + E.l05 * d[5];
  const T a1 = E.l12 * d[2] + E.l13 * d[3];
  const T b1 = E.l14 * d[4] + E.l15 * d[5];
  const T a2 = E.l21 * d[1] + E.l22 * d[2] + E.l23 * d[3];
  const T b2 = E.l26 * d[6] + E.l24 * d[4] + E.l25 * d[5];
  c[0] = b0 + a0;
  c[1] = b1 + a1;
  c[2] = b2 + a2;
  c[3] = d[0] + E.l36 * d[6];
  c[4] = b2 - a2;
  c[5] = b1 - a1;
  c[6] = b0 - a0;
}

// x += sum_i f[i] * (column i of R), columns 0/6, 2/4 and 1/5 paired like the rows
// of L (25 operations instead of 37; rounds differently)
template <class T>
PMHD_FI void eig_add_cols(const PrimEig<T>& E, const T f[7], T x[7]) {
  const T s06 = f[0] + f[6], d06 = f[0] - f[6];
  const T s24 = f[2] + f[4], d24 = f[2] - f[4];
  const T s15 = f[1] + f[5], d15 = f[1] - f[5];
  x[0] = x[0] + (s06 * E.r00 + s24 * E.r02 + f[3]);
  x[1] = x[1] + (d06 * E.r10 + d24 * E.r12);
  x[2] = x[2] + (d06 * E.r20 + d15 * E.r21 + d24 * E.r22);
  x[3] = x[3] + (d06 * E.r30 + d15 * E.r31 + d24 * E.r32);
  x[4] = x[4] + (s06 * E.r40 + s15 * E.r41 + s24 * E.r42);
  x[5] = x[5] + (s06 * E.r50 + s15 * E.r51 + s24 * E.r52);
  x[6] = x[6] + (s06 * E.r60 + s24 * E.r62);
}
#endif

// characteristic-limited primitive slope of one cell (Colella 1990; Stone+ 2008 sec 4;
// Source/Reconstruction_PPM_MHD_HelperModule.py:320-389)
template <class T>
PMHD_FI void limited_slope(const PrimEig<T>& E, const T wl[7], const T wc[7], const T wr[7], T ps[7]) {
  T dc[7], dl[7], dr[7], wg[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    dc[v] = wr[v] - wl[v];
    dl[v] = wc[v] - wl[v];
    dr[v] = wr[v] - wc[v];
    wg[v] = (dl[v] * dr[v] > T(0)) ? (T(2) * dl[v] * dr[v]) / (dl[v] + dr[v]) : T(0);
  }
  T cc[7], lc[7], rc[7], wc_[7];
#ifdef PMHD_FAST_ORDER
  // dc = dl + dr, so L.dc = L.dl + L.dr (one projection fewer; rounds differently)
  (void)dc;
  eig_project_pairs(E, dl, lc);
  eig_project_pairs(E, dr, rc);
#pragma unroll
  for (int v = 0; v < 7; ++v) cc[v] = lc[v] + rc[v];
  eig_project_pairs(E, wg, wc_);
#else
  eig_project(E, dc, cc);
  eig_project(E, dl, lc);
  eig_project(E, dr, rc);
  eig_project(E, wg, wc_);
#endif
  T s[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    T lim1 = npmin(npabs(lc[v]), npabs(rc[v]));
    T lim2 = npmin(T(0.5) * npabs(cc[v]), npabs(wc_[v]));
    s[v] = (lc[v] * rc[v] > T(0)) ? npsign(cc[v]) * npmin(T(2) * lim1, lim2) : T(0);
  }
#ifdef PMHD_FAST_ORDER
#pragma unroll
  for (int v = 0; v < 7; ++v) ps[v] = T(0);
  eig_add_cols(E, s, ps);
  return;
#endif
  // R . s, columns in wave order 0..6
  ps[0] = T(0) + s[0] * E.r00 + s[2] * E.r02 + s[3] * T(1) + s[4] * E.r02 + s[6] * E.r00;
  ps[1] = T(0) + s[0] * E.r10 + s[2] * E.r12 + s[4] * (-E.r12) + s[6] * (-E.r10);
  ps[2] = T(0) + s[0] * E.r20 + s[1] * E.r21 + s[2] * E.r22 + s[4] * (-E.r22) + s[5] * (-E.r21) + s[6] * (-E.r20);
  ps[3] = T(0) + s[0] * E.r30 + s[1] * E.r31 + s[2] * E.r32 + s[4] * (-E.r32) + s[5] * (-E.r31) + s[6] * (-E.r30);
  ps[4] = T(0) + s[0] * E.r40 + s[1] * E.r41 + s[2] * E.r42 + s[4] * E.r42 + s[5] * E.r41 + s[6] * E.r40;
  ps[5] = T(0) + s[0] * E.r50 + s[1] * E.r51 + s[2] * E.r52 + s[4] * E.r52 + s[5] * E.r51 + s[6] * E.r50;
  ps[6] = T(0) + s[0] * E.r60 + s[2] * E.r62 + s[4] * E.r62 + s[6] * E.r60;
}

// edge values + characteristic tracing of one cell (PPM: order 2; PLM: order 1,
// the linear member of the same family with a zero parabolic coefficient).
// Outputs la (left edge, traced) and ra (right edge, traced), rotated frame.
template <class T>
PMHD_FI void trace_cell(const PrimEig<T>& E, const T wl[7], const T wc[7], const T wr[7],
                        const T sl[7], const T sc[7], const T sr[7], T dtdx, int order,
                        T la[7], T ra[7]) {
  T c1[7], c2[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    T l, r;
    if (order == 2) {
      l = T(0.5) * (wc[v] + wl[v]) - (sc[v] - sl[v]) / T(6);
      r = T(0.5) * (wr[v] + wc[v]) - (sr[v] - sc[v]) / T(6);
      T dprod = (r - wc[v]) * (wc[v] - l);
      T idiff = r - l;
      T adiff = T(6) * (wc[v] - T(0.5) * (l + r));
      if (dprod <= T(0)) {
        l = wc[v];
        r = wc[v];
      } else if (dprod > T(0)) {
        if ((idiff * adiff) > idiff * idiff) {
          l = (T(6) * wc[v] - T(4) * r) / T(2);
        } else if ((idiff * adiff) < (T(-1) * (idiff * idiff))) {
          r = (T(6) * wc[v] - T(4) * l) / T(2);
        }
      }
    } else {
      l = wc[v] - T(0.5) * sc[v];
      r = wc[v] + T(0.5) * sc[v];
    }
    l = npmax(npmin(wc[v], wl[v]), l);
    l = npmin(npmax(wc[v], wl[v]), l);
    r = npmax(npmin(wc[v], wr[v]), r);
    r = npmin(npmax(wc[v], wr[v]), r);
    la[v] = l;
    ra[v] = r;
    c1[v] = r - l;
    c2[v] = order == 2 ? T(6) * (wc[v] - T(0.5) * (l + r)) : T(0);
  }
  T c3 = T(0.5) * npmax(E.lam[6], T(0)) * dtdx;
  T c4 = T(-0.5) * npmin(E.lam[0], T(0)) * dtdx;
  const T hdt = T(0.5) * dtdx;
#ifdef PMHD_FAST_ORDER
  // L_i . (v3 (c1 + c2) + v4 c2) = v3 (L_i.c1 + L_i.c2) + v4 L_i.c2: project c1 and c2
  // once and weight the projections per wave, without branches (rounds differently
  // from the per-wave projections below).  PLM has c2 = 0: its terms are left out
  // (a product with zero is not folded without finite-math assumptions).
  T p1[7];
  const T vl = hdt * E.lam[0], vr = hdt * E.lam[6];
  if (order == 2) {
#pragma unroll
    for (int v = 0; v < 7; ++v) {
      la[v] = la[v] + c4 * (c1[v] + (T(1) - T(kFourThirds) * c4) * c2[v]);
      ra[v] = ra[v] - c3 * (c1[v] - (T(1) - T(kFourThirds) * c3) * c2[v]);
    }
    T p2[7], fl[7], fr[7];
    eig_project_pairs(E, c1, p1);
    eig_project_pairs(E, c2, p2);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const T v2 = hdt * E.lam[i];
      const T v3l = vl - v2, v4l = T(kFourThirds) * (vl * vl - v2 * v2);
      const T v3r = vr - v2, v4r = T(kFourThirds) * (vr * vr - v2 * v2);
      fl[i] = E.lam[i] <= T(0) ? v3l * (p1[i] + p2[i]) + v4l * p2[i] : T(0);
      fr[i] = E.lam[i] >= T(0) ? v3r * (p1[i] - p2[i]) + v4r * p2[i] : T(0);
    }
    eig_add_cols(E, fl, la);
    eig_add_cols(E, fr, ra);
  } else {
#pragma unroll
    for (int v = 0; v < 7; ++v) {
      la[v] = la[v] + c4 * c1[v];
      ra[v] = ra[v] - c3 * c1[v];
    }
    T fl[7], fr[7];
    eig_project_pairs(E, c1, p1);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const T v2 = hdt * E.lam[i];
      fl[i] = E.lam[i] <= T(0) ? (vl - v2) * p1[i] : T(0);
      fr[i] = E.lam[i] >= T(0) ? (vr - v2) * p1[i] : T(0);
    }
    eig_add_cols(E, fl, la);
    eig_add_cols(E, fr, ra);
  }
#else
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    la[v] = la[v] + c4 * (c1[v] + (T(1) - T(kFourThirds) * c4) * c2[v]);
    ra[v] = ra[v] - c3 * (c1[v] - (T(1) - T(kFourThirds) * c3) * c2[v]);
  }
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (E.lam[i] <= T(0)) {
      T v1 = hdt * E.lam[0];
      T v2 = hdt * E.lam[i];
      T v3 = v1 - v2;
      T v4 = T(kFourThirds) * (v1 * v1 - v2 * v2);
      T q[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) q[j] = v3 * (c1[j] + c2[j]) + v4 * c2[j];
      T fac = eig_row_dot(E, i, q);
      eig_add_col(E, i, fac, la);
    }
  }
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (E.lam[i] >= T(0)) {
      T v1 = hdt * E.lam[6];
      T v2 = hdt * E.lam[i];
      T v3 = v1 - v2;
      T v4 = T(kFourThirds) * (v1 * v1 - v2 * v2);
      T q[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) q[j] = v3 * (c1[j] - c2[j]) + v4 * c2[j];
      T fac = eig_row_dot(E, i, q);
      eig_add_col(E, i, fac, ra);
    }
  }
#endif
}

// ---------------------------------------------------------------------------
// Roe flux for ideal MHD, rotated frame
// (Source/RiemannSolvers_MHD_Roe_HelperModule.py:74-576)
// wl, wr: primitive; ul, ur: conserved; bn: face-normal B; f: flux (rotated)
// ---------------------------------------------------------------------------
template <class T>
PMHD_FI void cell_flux(const T w[7], const T u[7], T bn, T hb, T f[7]) {
  const T vn = w[1], vt1 = w[2], vt2 = w[3], b1 = w[4], b2 = w[5], p = w[6];
  const T mn = u[1];
  f[0] = mn;
  T f1 = mn * vn;
  f[2] = mn * vt1;
  f[3] = mn * vt2;
  f1 = f1 + p;
  T f6 = (u[6] + p) * vn;
  f[1] = f1 - T(0.5) * (bn * bn - b1 * b1 - b2 * b2);
  f[2] = f[2] - bn * b1;
  f[3] = f[3] - bn * b2;
  f[6] = f6 + (hb * vn - bn * (bn * vn + b1 * vt1 + b2 * vt2));
  f[4] = b1 * vn - bn * vt1;
  f[5] = b2 * vn - bn * vt2;
}

template <class T>
PMHD_FI T entropy_fixed(T l, T eps) {
  T s = npsign(l) * ((npabs(l) < T(2) * eps) ? T(1) : T(0));
  return l + s * (eps + l * l / (T(4) * eps));
}

// Riemann solvers of the step: the reference's Roe flux; HLLE (no counterpart in
// the reference, which names it as missing at
// Source/RiemannSolvers_MHD_Roe_HelperModule.py:24-30); Roe that falls back to HLLE
// at a face whose Roe intermediate states have a non-positive density or pressure
// (the switch of Athena's Roe solver that the reference's header describes)
enum { RIEMANN_ROE = 0, RIEMANN_HLLE = 1, RIEMANN_ROE_HLLE = 2 };

// fast magnetosonic speed of a rotated primitive state with normal field bn
template <class T>
PMHD_FI T fast_speed_of(const T w[7], T bn, T g) {
  const T inv = T(1) / w[0];
  const T btsq = w[4] * w[4] + w[5] * w[5];
  const T cs2 = g * w[6] * inv;
  const T bsq = bn * bn + btsq;
  const T ssum = bsq * inv + cs2;
  const T sdiff = bsq * inv - cs2;
  return sqrt_(T(0.5) * (ssum + sqrt_(sdiff * sdiff + T(4) * cs2 * btsq * inv)));
}

// HLLE flux with Einfeldt's wave-speed bounds: the slowest of the left state's and
// the Roe average's left-going fast speeds, the fastest of the right state's and
// the Roe average's right-going ones; the upwind flux when both bounds have one sign
template <class T>
PMHD_FI void hlle_flux(const T wl[7], const T wr[7], const T ul[7], const T ur[7], T bn, T g, T lam_lo,
                       T lam_hi, const T FL[7], const T FR[7], T F[7]) {
  const T sl = npmin(wl[1] - fast_speed_of(wl, bn, g), lam_lo);
  const T sr = npmax(wr[1] + fast_speed_of(wr, bn, g), lam_hi);
  if (sl >= T(0)) {
#pragma unroll
    for (int v = 0; v < 7; ++v) F[v] = FL[v];
    return;
  }
  if (sr <= T(0)) {
#pragma unroll
    for (int v = 0; v < 7; ++v) F[v] = FR[v];
    return;
  }
  const T inv = T(1) / (sr - sl);
#pragma unroll
  for (int v = 0; v < 7; ++v) F[v] = (sr * FL[v] - sl * FR[v] + sl * sr * (ur[v] - ul[v])) * inv;
}

// pressure of a rotated conserved state (normal field bn)
template <class T>
PMHD_FI T cons_pressure(const T u[7], T bn, T g) {
  const T ke = T(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0];
  return (g - T(1)) * (u[6] - ke - T(0.5) * (bn * bn + u[4] * u[4] + u[5] * u[5]));
}

// S: the solver (RIEMANN_*), fixed at compile time so that each grid's kernels
// carry the code of its own solver only
template <int S = RIEMANN_ROE, class T>
PMHD_HD inline void roe_flux(const T wl[7], const T wr[7], const T ul[7], const T ur[7], T bn,
                             T g, bool efix, T eps, T F[7]) {
  using N = island_t<T>;
  const T rl = wl[0], rr = wr[0];
  // FP32 fast unit: reciprocal roots of the two densities, of the Cargo-Gallice
  // factor and of B_t^2 replace the roots and the quotients by them
  T sl, sr, isl = T(0), isr = T(0);
  if constexpr (fast_f32<T>()) {
    isl = rsqrt_(rl);
    isr = rsqrt_(rr);
    sl = rl * isl;
    sr = rr * isr;
  } else {
    sl = sqrt_(rl);
    sr = sqrt_(rr);
  }
  T isum = T(1) / (sl + sr);
  T rd = sl * sr;
  T vn = (sl * wl[1] + sr * wr[1]) * isum;
  T vt1 = (sl * wl[2] + sr * wr[2]) * isum;
  T vt2 = (sl * wl[3] + sr * wr[3]) * isum;
  T vsq = vn * vn + vt1 * vt1 + vt2 * vt2;
  T bt1 = (sr * wl[4] + sl * wr[4]) * isum;
  T bt2 = (sr * wl[5] + sl * wr[5]) * isum;
  T hbl = T(0.5) * (bn * bn + wl[4] * wl[4] + wl[5] * wl[5]);
  T hbr = T(0.5) * (bn * bn + wr[4] * wr[4] + wr[5] * wr[5]);
  T pl = wl[6], pr = wr[6];
  T H, ird, btc, rhoc, bt, bts, ibt = T(0), ifac = T(0);
  T d4 = wl[4] - wr[4], d5 = wl[5] - wr[5];
  T btsq = bt1 * bt1 + bt2 * bt2;
  T btsqs;
  if constexpr (fast_f32<T>()) {
    H = ((ul[6] + pl + hbl) * isl + (ur[6] + pr + hbr) * isr) * isum;
    ird = isl * isr;
    btc = T(0.5) * (d4 * d4 + d5 * d5) * (isum * isum);
    rhoc = T(0.5) * (rl + rr) * ird;
    const T fac = (g - T(1)) - (g - T(2)) * rhoc;
    ifac = rsqrt_(fac);
    ibt = rsqrt_(btsq);
    bt = btsq > T(0) ? btsq * ibt : T(0);
    btsqs = fac * btsq;
    bts = (fac * ifac) * bt;
  } else {
    H = ((ul[6] + pl + hbl) / sl + (ur[6] + pr + hbr) / sr) * isum;
    ird = T(1) / rd;
    T ssd = sl + sr;
    btc = T(0.5) * (d4 * d4 + d5 * d5) / (ssd * ssd);
    rhoc = T(0.5) * (rl + rr) / rd;
    bt = sqrt_(btsq);
    btsqs = ((g - T(1)) - (g - T(2)) * rhoc) * btsq;
    bts = sqrt_(btsqs);
  }
  T hmb, cs2, cf, ca, csl, cs;
  N af, as;
  {
    // wave speeds and normalisation in N precision
    N n_ird = N(ird);
    N n_rbtsqs = N(btsqs) * n_ird;
    N n_ca2 = N(bn) * N(bn) * n_ird;
    N n_hmb = N(H) - (n_ca2 + N(btsq) * n_ird);
    N n_cs2 = npmax((N(g) - N(1)) * (n_hmb - N(0.5) * N(vsq)) - (N(g) - N(2)) * N(btc), N(kSmall));
    N ssum = n_ca2 + n_rbtsqs + n_cs2;
    N sdiff = n_ca2 + n_rbtsqs - n_cs2;
    N ssqrt = sqrt_(sdiff * sdiff + N(4) * n_cs2 * n_rbtsqs);
    N n_cf2 = N(0.5) * (ssum + ssqrt);
    N n_csl2 = n_cs2 * n_ca2 / n_cf2;
    if ((n_cf2 - n_csl2) == N(0)) {
      af = N(1); as = N(0);
    } else if ((n_cs2 - n_csl2) <= N(0)) {
      af = N(0); as = N(1);
    } else if ((n_cf2 - n_cs2) <= N(0)) {
      af = N(1); as = N(0);
    } else {
#ifdef PMHD_FAST_ORDER
      if constexpr (sizeof(T) == 4) {
        // as in eigen_prim: double differences, float quotient and root
        const T iq = T(1) / T(n_cf2 - n_csl2);
        af = N(sqrt_(T(n_cs2 - n_csl2) * iq));
        as = N(sqrt_(T(n_cf2 - n_cs2) * iq));
      } else {
        af = sqrt_((n_cs2 - n_csl2) / (n_cf2 - n_csl2));
        as = sqrt_((n_cf2 - n_cs2) / (n_cf2 - n_csl2));
      }
#else
      af = sqrt_((n_cs2 - n_csl2) / (n_cf2 - n_csl2));
      as = sqrt_((n_cf2 - n_cs2) / (n_cf2 - n_csl2));
#endif
    }
    hmb = T(n_hmb);
    cs2 = T(n_cs2);
    ca = sqrt_(T(n_ca2));
    cs = sqrt_(cs2);
    cf = sqrt_(T(n_cf2));
    csl = sqrt_(T(n_csl2));
  }
  T lam[7] = {vn - cf, vn - ca, vn - csl, vn, vn + csl, vn + ca, vn + cf};
  T FL[7], FR[7];
  cell_flux(wl, ul, bn, hbl, FL);
  cell_flux(wr, ur, bn, hbr, FR);
  if constexpr (S == RIEMANN_HLLE) {
    hlle_flux(wl, wr, ul, ur, bn, g, lam[0], lam[6], FL, FR, F);
    return;
  }
  const T lam_lo = lam[0], lam_hi = lam[6];
  (void)lam_lo;
  (void)lam_hi;
  if (efix) {
#pragma unroll
    for (int k = 0; k < 7; ++k) lam[k] = entropy_fixed(lam[k], eps);
  }
  if (lam[0] >= T(0)) {
    // supersonic to the right; the reference still evaluates the Roe sum and then
    // overwrites it, so only the selection matters
    if (lam[6] <= T(0)) {
#pragma unroll
      for (int v = 0; v < 7; ++v) F[v] = FR[v];
    } else {
#pragma unroll
      for (int v = 0; v < 7; ++v) F[v] = FL[v];
    }
    return;
  }
  if (lam[6] <= T(0)) {
#pragma unroll
    for (int v = 0; v < 7; ++v) F[v] = FR[v];
    return;
  }
  T Af_ = T(af), As_ = T(as);
  T sg = npsign(bn);
  sg = sg == T(0) ? T(1) : sg;
  T srd, isrd;
  if constexpr (fast_f32<T>()) {
    isrd = rsqrt_(rd);
    srd = rd * isrd;
  } else {
    srd = sqrt_(rd);
    isrd = T(1) / srd;
  }
  T qf = cf * Af_ * sg;
  T qs = csl * As_ * sg;
  T afp = cs * Af_ * isrd;
  T asp = cs * As_ * isrd;
  T b1, b2, b1s, b2s;
  if constexpr (fast_f32<T>()) {
    b1 = btsq == T(0) ? T(1) : (btsq > T(0) ? bt1 * ibt : T(0));
    b2 = btsq > T(0) ? bt2 * ibt : T(0);
    b1s = b1 * ifac;
    b2s = b2 * ifac;
  } else {
    b1 = bt == T(0) ? T(1) : (bt > T(0) ? bt1 / bt : T(0));
    b2 = bt > T(0) ? bt2 / bt : T(0);
    b1s = b1 / sqrt_((g - T(1)) - (g - T(2)) * rhoc);
    b2s = b2 / sqrt_((g - T(1)) - (g - T(2)) * rhoc);
  }
  T sb = vt1 * b1s + vt2 * b2s;
  T bss = b1s * b1s + b2s * b2s;
  if constexpr (S == RIEMANN_ROE && fast_f32<T>()) {
    // FP32 fast unit: the wave pairs (0, 6), (2, 4) and (1, 5) have eigenvectors
    // that differ only in the sign of their velocity terms, so the projections are
    // sums and differences of shared partial dot products and the eigenvector sum
    // runs over the pair sums and differences; the eigenvector entries themselves
    // are never formed (the same flux in exact arithmetic; rounds differently)
    T dU[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) dU[v] = ur[v] - ul[v];
    const T nf = T(0.5) / cs2;
    const T cff = nf * Af_ * cf, css = nf * As_ * csl;
    const T qfn = qf * nf, qsn = qs * nf;
    const T a_f = nf * afp * rd, a_s = nf * asp * rd;
    const T afbn = nf * afp * bts, asbn = nf * asp * bts;
    const T nfg = nf * (g - T(1));
    const T afn = Af_ * nfg, asn = As_ * nfg;
    const T ibss = T(1) / bss;
    const T b1c = b1s * ibss, b2c = b2s * ibss;
    const T sbs = vt1 * b1c + vt2 * b2c;
    const T G = (vsq - hmb) * dU[0] - vn * dU[1] - vt1 * dU[2] - vt2 * dU[3] + dU[6];
    const T Hd = vn * dU[0] - dU[1];
    const T K = b1c * dU[2] + b2c * dU[3] - sbs * dU[0];
    const T P0 = afn * G + (cff * cf - asbn) * dU[0] + (a_s * b1c - afn * bt1) * dU[4] +
                 (a_s * b2c - afn * bt2) * dU[5];
    const T Q0 = cff * Hd + qsn * K;
    const T P2 = asn * G + (css * csl + afbn) * dU[0] + (-a_f * b1c - asn * bt1) * dU[4] +
                 (-a_f * b2c - asn * bt2) * dU[5];
    const T Q2 = css * Hd - qfn * K;
    const T A1 = T(0.5) * (vt1 * b2 - vt2 * b1) * dU[0] + T(-0.5) * b2 * dU[2] + T(0.5) * b1 * dU[3];
    const T B1 = T(0.5) * srd * sg * (b1 * dU[5] - b2 * dU[4]);
    const T nf2 = T(2) * nfg;
    const T r63 = T(0.5) * vsq + (g - T(2)) * btc / (g - T(1));
    const T p3 = (T(1) - nf2 * (T(0.5) * vsq - (g - T(2)) * btc / (g - T(1)))) * dU[0] +
                 nf2 * (vn * dU[1] + vt1 * dU[2] + vt2 * dU[3] - dU[6] + bt1 * dU[4] + bt2 * dU[5]);
    const T a0 = T(0.5) * npabs(lam[0]) * (P0 + Q0);
    const T a1 = T(0.5) * npabs(lam[1]) * (B1 + A1);
    const T a2 = T(0.5) * npabs(lam[2]) * (P2 + Q2);
    const T a3 = T(0.5) * npabs(lam[3]) * p3;
    const T a4 = T(0.5) * npabs(lam[4]) * (P2 - Q2);
    const T a5 = T(0.5) * npabs(lam[5]) * (B1 - A1);
    const T a6 = T(0.5) * npabs(lam[6]) * (P0 - Q0);
    const T s06 = a0 + a6, d06 = a0 - a6, s24 = a2 + a4, d24 = a2 - a4, s15 = a1 + a5, d15 = a1 - a5;
    const T X0 = Af_ * s06 + As_ * s24 + a3;
    const T Y = qs * d06 - qf * d24;
    const T Z = asp * s06 - afp * s24;
    const T r41 = -b2 * sg * isrd, r51 = b1 * sg * isrd;
    const T r61 = -(vt1 * b2 - vt2 * b1);
    F[0] = T(0.5) * (FL[0] + FR[0]) - X0;
    F[1] = T(0.5) * (FL[1] + FR[1]) -
           (a0 * (Af_ * lam[0]) + a2 * (As_ * lam[2]) + a3 * vn + a4 * (As_ * lam[4]) + a6 * (Af_ * lam[6]));
    F[2] = T(0.5) * (FL[2] + FR[2]) - (vt1 * X0 + b1s * Y - b2 * d15);
    F[3] = T(0.5) * (FL[3] + FR[3]) - (vt2 * X0 + b2s * Y + b1 * d15);
    F[4] = T(0.5) * (FL[4] + FR[4]) - (b1s * Z + r41 * s15);
    F[5] = T(0.5) * (FL[5] + FR[5]) - (b2s * Z + r51 * s15);
    F[6] = T(0.5) * (FL[6] + FR[6]) - (hmb * (X0 - a3) - vn * (Af_ * cf * d06 + As_ * csl * d24) + sb * Y +
                                       bts * bss * Z + r61 * d15 + a3 * r63);
    return;
  }
  T afb = afp * bts * bss;
  T asb = asp * bts * bss;
  // right eigenvectors (conserved), structurally non-zero entries
  T R10 = Af_ * lam[0], R12 = As_ * lam[2], R14 = As_ * lam[4], R16 = Af_ * lam[6];
  T fav = Af_ * vt1, sav = As_ * vt1, qsb = qs * b1s, qfb = qf * b1s;
  T R20 = fav + qsb, R21 = -b2, R22 = sav - qfb, R23 = vt1, R24 = sav + qfb, R25 = b2, R26 = fav - qsb;
  fav = Af_ * vt2; sav = As_ * vt2; qsb = qs * b2s; qfb = qf * b2s;
  T R30 = fav + qsb, R31 = b1, R32 = sav - qfb, R33 = vt2, R34 = sav + qfb, R35 = -b1, R36 = fav - qsb;
  T R60 = Af_ * (hmb - vn * cf) + qs * sb + asb;
  T R61 = -(vt1 * b2 - vt2 * b1);
  T R62 = As_ * (hmb - vn * csl) - qf * sb - afb;
  T R63 = T(0.5) * vsq + (g - T(2)) * btc / (g - T(1));
  T R64 = As_ * (hmb + vn * csl) + qf * sb - afb;
  T R65 = -R61;
  T R66 = Af_ * (hmb + vn * cf) - qs * sb + asb;
  T R40 = asp * b1s, R41 = -b2 * sg * isrd, R42 = -afp * b1s;
  T R50 = asp * b2s, R51 = b1 * sg * isrd, R52 = -afp * b2s;
  // left eigenvectors
  T nf = T(0.5) / cs2;
  T cff = nf * Af_ * cf;
  T css = nf * As_ * csl;
  qf = qf * nf;
  qs = qs * nf;
  T a_f = nf * afp * rd;
  T a_s = nf * asp * rd;
  afb = nf * afp * bts;
  asb = nf * asp * bts;
  nf = nf * (g - T(1));
  T afn = Af_ * nf;
  T asn = As_ * nf;
  T b1c, b2c;
  if constexpr (fast_f32<T>()) {
    const T ibss = T(1) / bss;
    b1c = b1s * ibss;
    b2c = b2s * ibss;
  } else {
    b1c = b1s / bss;
    b2c = b2s / bss;
  }
  T sbs = vt1 * b1c + vt2 * b2c;
  nf = nf * T(2);
  T L00 = afn * (vsq - hmb) + cff * (cf + vn) - qs * sbs - asb;
  T L01 = T(-1) * afn * vn - cff;
  T L02 = T(-1) * afn * vt1 + qs * b1c;
  T L03 = T(-1) * afn * vt2 + qs * b2c;
  T L06 = afn;
  T L04 = a_s * b1c - afn * bt1;
  T L05 = a_s * b2c - afn * bt2;
  T L10 = T(0.5) * (vt1 * b2 - vt2 * b1);
  T L12 = T(-0.5) * b2;
  T L13 = T(0.5) * b1;
  T L14 = T(-0.5) * srd * b2 * sg;
  T L15 = T(0.5) * srd * b1 * sg;
  T L20 = asn * (vsq - hmb) + css * (csl + vn) + qf * sbs + afb;
  T L21 = -asn * vn - css;
  T L22 = -asn * vt1 - qf * b1c;
  T L23 = -asn * vt2 - qf * b2c;
  T L26 = asn;
  T L24 = -a_f * b1c - asn * bt1;
  T L25 = -a_f * b2c - asn * bt2;
  T L30 = T(1) - nf * (T(0.5) * vsq - (g - T(2)) * btc / (g - T(1)));
  T L31 = nf * vn, L32 = nf * vt1, L33 = nf * vt2, L36 = T(-1) * nf, L34 = nf * bt1, L35 = nf * bt2;
  T L40 = asn * (vsq - hmb) + css * (csl - vn) - qf * sbs + afb;
  T L41 = -asn * vn + css;
  T L42 = -asn * vt1 + qf * b1c;
  T L43 = -asn * vt2 + qf * b2c;
  T L46 = asn;
  T L60 = afn * (vsq - hmb) + cff * (cf - vn) + qs * sbs - asb;
  T L61 = -afn * vn + cff;
  T L62 = -afn * vt1 - qs * b1c;
  T L63 = -afn * vt2 - qs * b2c;
  T L66 = afn;
  T dU[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) dU[v] = ur[v] - ul[v];
  // projections, summation order (0,1,2,3,6,4,5)
  T p0 = T(0) + L00 * dU[0] + L01 * dU[1] + L02 * dU[2] + L03 * dU[3] + L06 * dU[6] + L04 * dU[4] + L05 * dU[5];
  T p1 = T(0) + L10 * dU[0] + L12 * dU[2] + L13 * dU[3] + L14 * dU[4] + L15 * dU[5];
  T p2 = T(0) + L20 * dU[0] + L21 * dU[1] + L22 * dU[2] + L23 * dU[3] + L26 * dU[6] + L24 * dU[4] + L25 * dU[5];
  T p3 = T(0) + L30 * dU[0] + L31 * dU[1] + L32 * dU[2] + L33 * dU[3] + L36 * dU[6] + L34 * dU[4] + L35 * dU[5];
  T p4 = T(0) + L40 * dU[0] + L41 * dU[1] + L42 * dU[2] + L43 * dU[3] + L46 * dU[6] + L24 * dU[4] + L25 * dU[5];
  T p5 = T(0) + (-L10) * dU[0] + (-L12) * dU[2] + (-L13) * dU[3] + L14 * dU[4] + L15 * dU[5];
  T p6 = T(0) + L60 * dU[0] + L61 * dU[1] + L62 * dU[2] + L63 * dU[3] + L66 * dU[6] + L04 * dU[4] + L05 * dU[5];
  if constexpr (S == RIEMANN_ROE_HLLE) {
    // the Roe intermediate states U_L + sum_{j<=k} p_j R_j, k = 0..5: a non-positive
    // density or pressure in any of them selects the HLLE flux
    const T pk[6] = {p0, p1, p2, p3, p4, p5};
    const T col[6][7] = {{Af_, R10, R20, R30, R40, R50, R60}, {T(0), T(0), R21, R31, R41, R51, R61},
                         {As_, R12, R22, R32, R42, R52, R62}, {T(1), vn, R23, R33, T(0), T(0), R63},
                         {As_, R14, R24, R34, R42, R52, R64}, {T(0), T(0), R25, R35, R41, R51, R65}};
    T us[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) us[v] = ul[v];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
      for (int v = 0; v < 7; ++v) us[v] = us[v] + pk[k] * col[k][v];
      bad = bad || !(us[0] > T(0)) || !(cons_pressure(us, bn, g) > T(0));
    }
    if (bad) {
      hlle_flux(wl, wr, ul, ur, bn, g, lam_lo, lam_hi, FL, FR, F);
      return;
    }
  }
  T a0 = T(0.5) * npabs(lam[0]) * p0;
  T a1 = T(0.5) * npabs(lam[1]) * p1;
  T a2 = T(0.5) * npabs(lam[2]) * p2;
  T a3 = T(0.5) * npabs(lam[3]) * p3;
  T a4 = T(0.5) * npabs(lam[4]) * p4;
  T a5 = T(0.5) * npabs(lam[5]) * p5;
  T a6 = T(0.5) * npabs(lam[6]) * p6;
  T f;
  f = T(0.5) * (FL[0] + FR[0]);
  f = f - a0 * Af_; f = f - a2 * As_; f = f - a3 * T(1); f = f - a4 * As_; f = f - a6 * Af_;
  F[0] = f;
  f = T(0.5) * (FL[1] + FR[1]);
  f = f - a0 * R10; f = f - a2 * R12; f = f - a3 * vn; f = f - a4 * R14; f = f - a6 * R16;
  F[1] = f;
  f = T(0.5) * (FL[2] + FR[2]);
  f = f - a0 * R20; f = f - a1 * R21; f = f - a2 * R22; f = f - a3 * R23; f = f - a4 * R24; f = f - a5 * R25; f = f - a6 * R26;
  F[2] = f;
  f = T(0.5) * (FL[3] + FR[3]);
  f = f - a0 * R30; f = f - a1 * R31; f = f - a2 * R32; f = f - a3 * R33; f = f - a4 * R34; f = f - a5 * R35; f = f - a6 * R36;
  F[3] = f;
  f = T(0.5) * (FL[4] + FR[4]);
  f = f - a0 * R40; f = f - a1 * R41; f = f - a2 * R42; f = f - a4 * R42; f = f - a5 * R41; f = f - a6 * R40;
  F[4] = f;
  f = T(0.5) * (FL[5] + FR[5]);
  f = f - a0 * R50; f = f - a1 * R51; f = f - a2 * R52; f = f - a4 * R52; f = f - a5 * R51; f = f - a6 * R50;
  F[5] = f;
  f = T(0.5) * (FL[6] + FR[6]);
  f = f - a0 * R60; f = f - a1 * R61; f = f - a2 * R62; f = f - a3 * R63; f = f - a4 * R64; f = f - a5 * R65; f = f - a6 * R66;
  F[6] = f;
}

// the flux of a solver chosen at run time (the Riemann-flux entry point)
template <class T>
PMHD_FI void roe_flux_solver(const T wl[7], const T wr[7], const T ul[7], const T ur[7], T bn, T g, bool efix,
                             T eps, T F[7], int solver) {
  if (solver == RIEMANN_HLLE) roe_flux<RIEMANN_HLLE>(wl, wr, ul, ur, bn, g, efix, eps, F);
  else if (solver == RIEMANN_ROE_HLLE) roe_flux<RIEMANN_ROE_HLLE>(wl, wr, ul, ur, bn, g, efix, eps, F);
  else roe_flux<RIEMANN_ROE>(wl, wr, ul, ur, bn, g, efix, eps, F);
}

// ---------------------------------------------------------------------------
// The reference forms the x-direction CTU factor 0.5*dt/dy in np.longdouble
// (80-bit x87, 64-bit significand; Source/CTU_HelperModule.py:455), so
// u -= h*x is RN53(RN64(u - RN64(h*x))).  ld_fms reproduces that in double-double
// arithmetic; h is passed as the exact pair hh + hl.
// ---------------------------------------------------------------------------
PMHD_FI void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
PMHD_FI void fast_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  e = b - (s - a);
}
// rounding of lo such that hi + lo becomes the nearest 64-bit-significand value
PMHD_FI double round64_lo(double hi, double lo) {
  if (lo == 0.0 || hi == 0.0 || !(fabs(hi) < 1.0e300)) return lo;
  int e;
  (void)frexp(hi, &e);  // |hi| in [2^(e-1), 2^e)
  double q = ldexp(1.0, e - 64);
  if (fabs(hi) == ldexp(1.0, e - 1) && ((hi > 0.0) != (lo > 0.0))) q *= 0.5;
  return rint(lo / q) * q;
}
PMHD_FI double ld_fms(double u, double hh, double hl, double x, bool add) {
  double p1 = hh * x, e1 = fma(hh, x, -p1);
  double p2 = hl * x, e2 = fma(hl, x, -p2);
  double s, t;
  two_sum(p1, p2, s, t);
  t = t + (e1 + e2);
  double ph, pl;
  fast_two_sum(s, t, ph, pl);
  pl = round64_lo(ph, pl);
  if (!add) { ph = -ph; pl = -pl; }
  double dh, dl;
  two_sum(u, ph, dh, dl);
  dl = dl + pl;
  double rh, rl;
  fast_two_sum(dh, dl, rh, rl);
  rl = round64_lo(rh, rl);
  return rh + rl;
}

// minmod as coded in the reconstruction source terms
// (Source/Reconstruction_PPM_MHD_HelperModule.py:1398-1425, with the y-direction
//  zero-difference case of :1614-1618 when zero_fix is set)
template <class T>
PMHD_FI T minmod_recon(T dn, T dt_, bool zero_fix) {
  T t;
  bool took_t;
  if (dn >= T(0)) {
    took_t = !(dn < -dt_);
    t = took_t ? -dt_ : dn;
  } else {
    took_t = !(dn > -dt_);
    t = took_t ? -dt_ : dn;
  }
  if (zero_fix && took_t && dt_ == T(0)) t = T(0);
  if (dn >= T(0) && t < T(0)) t = T(0);
  if (!(dn >= T(0)) && t > T(0)) t = T(0);
  return t;
}

// minmod of (t, -n) as coded in the 3D CTU source terms (Source/CTU_HelperModule.py:624-646)
template <class T>
PMHD_FI T minmod_ctu(T t, T n) {
  if (n > T(0) && t < T(0)) return (t > -n) ? t : -n;
  if (n < T(0) && t > T(0)) return (t < -n) ? t : -n;
  return T(0);
}

// upwind selector of the corner-EMF terms (Source/CornerEMF_HelperModule.py:497-525)
template <class T>
PMHD_FI T upwind(T m, T pos, T neg, T zero) {
  return m > T(0) ? pos : (m < T(0) ? neg : (m == T(0) ? zero : T(0)));
}

PMHD_NS_END
