// Device-side data layout and helpers for the VSVB energy path on gfx950.
//
// Layout in HBM (all FP64 unless noted; one copy per geometry):
//   Shell      {atom, l, nprim, prim0, ao}                     static per molecule
//   Pair       {A, B, pp0, npp, cls, offdiag, AB, Q}           one per unordered shell pair
//   PrimPair   {p, 1/(2p), P, P-A, K}                          one per primitive pair
// Shell pairs are stored class-major (pair classes ss, ps, pp, ds, dp, dd) and,
// inside a class, by decreasing primitive-pair count so a wavefront's threads
// run primitive loops of similar length.  A is the shell of higher angular
// momentum, so every quartet maps onto one canonical generated class.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsvb {

constexpr int kLMax = 2;            // s, p, d
constexpr int kLTot = 4 * kLMax;    // highest Boys order for (dd|dd)
constexpr int kNPairClass = 6;
constexpr int kNQuartetClass = 21;

// Boys table: F_k(T_i) for T_i = i*dT, i < kBoysN, k < kBoysK (k = L + 0..8).
constexpr double kBoysDT = 0.05;
constexpr double kBoysInvDT = 20.0;
constexpr double kBoysTMax = 60.0;  // above: asymptotic F_0, upward recursion (e^-T < 1e-26 is dropped)
constexpr int kBoysN = 1202;
constexpr int kBoysK = kLTot + 9;
// Compact tables, one per L = 0..kLTot, for the straight-line quartet kernels
// (staged in LDS): c_k(T_i) = F_{L+k}(T_i) / k!, T_i = i * kBoysCDT, k < kBoysCK.
// F_L(T) = sum_k c_k d^k and F_{L+1}(T) = its derivative in d, d = T_i - T,
// |d| <= 0.1: the first omitted terms are below 2.5e-19 F_{L+11}(T_i) (value)
// and 2.8e-17 F_{L+11}(T_i) (derivative).  Rows are padded to 14 doubles: with
// 12 the 64 lanes' random rows start on only 4 distinct LDS bank groups
// (measured: the C8H18 J/K pass 6.99 -> 6.84 ms with the padding).
constexpr double kBoysCDT = 0.2;
constexpr double kBoysCInvDT = 5.0;
constexpr int kBoysCN = 301;  // T_i = 0 .. 60
constexpr int kBoysCK = 11;
constexpr int kBoysCRow = 14;  // row stride (doubles)
constexpr int kBoysCLen = kBoysCN * kBoysCRow;  // doubles per L table
constexpr size_t kBoysCOff = (size_t)kBoysN * kBoysK;  // compact tables follow the full one

struct PrimPair {
  double p, oo2p, Px, Py, Pz, PAx, PAy, PAz, K, pad;
};

// Primitive pairs that every lane of a wavefront reads at the same address
// (the ket list of a quartet kernel, the pair list of a Schwarz wave): through
// the constant address space the compiler may load them with scalar loads into
// SGPRs.  The data are written by an earlier kernel and read-only afterwards.
typedef const __attribute__((address_space(4))) PrimPair* UniformPrimPairs;
__device__ __forceinline__ UniformPrimPairs as_uniform(const PrimPair* p) { return (UniformPrimPairs)p; }
// Field-wise copy out of the constant address space (the implicit copy
// constructor binds only generic references); unused fields are dropped.
__device__ __forceinline__ PrimPair load_uniform(UniformPrimPairs q) {
  PrimPair r;
  r.p = q->p;
  r.oo2p = q->oo2p;
  r.Px = q->Px;
  r.Py = q->Py;
  r.Pz = q->Pz;
  r.PAx = q->PAx;
  r.PAy = q->PAy;
  r.PAz = q->PAz;
  r.K = q->K;
  r.pad = 0.0;
  return r;
}

struct Shell {
  int atom, l, nprim, prim0, ao;
};

struct Pair {
  int A, B;        // shells; l(A) >= l(B)
  int pp0, npp;    // primitive pairs
  int offdiag;     // 1 if A != B
  int pad;
  double ABx, ABy, ABz;
  double Q;        // sqrt(max |(ab|ab)|)
};

// 1/sqrt(x) to full FP64 precision: the hardware estimate (relative error
// below 2^-24 on gfx950, measured over 4M inputs in [2^-40, 2^40)) and one
// third-order step y <- y (1 + e/2 + 3e^2/8), e = 1 - x y^2, whose error is
// O(e^3) ~ 2^-70 before rounding.  Measured on the same inputs: at most 0.99 ulp
// from the exact value, as with two Newton steps, in 5 instructions instead of
// 8.  x is a positive normal number here.
__device__ __forceinline__ double rsqrt_fp64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  const double t = __builtin_fma(e, 0.375, 0.5);
  return __builtin_fma(y * e, t, y);
}

// Boys function F_m(T), m = 0..L, without divisions or exp().
// T < kBoysTMax: 8-term Taylor expansions of F_L and F_{L+1} about the nearest
// grid point (dF_m/dT = -F_{m+1}, |T - T_i| <= dT/2, remainder < 4e-18 F),
// e^{-T} = (2L+1) F_L - 2T F_{L+1} (from the recursion; its absolute error
// ~eps (2L+1) F_L is damped by the downward steps), then the stable downward
// recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1).
// T >= kBoysTMax: F_0 = sqrt(pi/T)/2 (erfc(sqrt 60) ~ 1e-28) and the upward
// recursion F_{m+1} = (2m+1) F_m / (2T) (the e^{-T} term is below 1e-26).
// The two paths branch: a wavefront whose lanes all take one path skips the
// other entirely, which measured faster than evaluating both on every lane.
template <int L>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ tab, double* F) {
  if (T < kBoysTMax) {
    const int i = __double2int_rn(T * kBoysInvDT);
    const double d = i * kBoysDT - T;
    const double* t = tab + i * kBoysK + L;
    double f = t[7], g = t[8];
#pragma unroll
    for (int k = 6; k >= 0; --k) {
      const double dk = d * (1.0 / (k + 1));
      f = __builtin_fma(f, dk, t[k]);
      if constexpr (L > 0) g = __builtin_fma(g, dk, t[k + 1]);
    }
    F[L] = f;
    if constexpr (L > 0) {
      const double t2 = 2.0 * T;
      const double e = __builtin_fma(-t2, g, (2 * L + 1) * f);
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = __builtin_fma(t2, F[m + 1], e) * (1.0 / (2 * m + 1));
    }
  } else {
    const double r = rsqrt_fp64(T);
    F[0] = 0.88622692545275801365 * r;  // sqrt(pi)/2
    if constexpr (L > 0) {
      const double h = 0.5 * r * r;
#pragma unroll
      for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * h) * F[m];
    }
  }
}

// Boys function from the compact table of this L (tab = its first row): one
// polynomial in d for F_L and its derivative for F_{L+1}, then as boys_eval.
// Every F_m comes out multiplied by w (the primitive quartet's prefactor): the
// recursions are linear, so scaling their two starting values (or F_0 of the
// asymptotic branch) scales them all.
template <int L>
__device__ __forceinline__ void boys_eval_c(double T, double w, const double* __restrict__ tab, double* F) {
  if (T < kBoysTMax) {
    const double x = __builtin_rint(T * kBoysCInvDT);
    const int i = (int)x;
    const double d = __builtin_fma(x, kBoysCDT, -T);
    const double* c = tab + i * kBoysCRow;
    // Horner for the polynomial (f) and, one step behind, its derivative (g)
    double f = c[kBoysCK - 1];
    double g = f;
    f = __builtin_fma(f, d, c[kBoysCK - 2]);
#pragma unroll
    for (int k = kBoysCK - 3; k >= 0; --k) {
      if constexpr (L > 0) g = __builtin_fma(g, d, f);
      f = __builtin_fma(f, d, c[k]);
    }
    f *= w;
    F[L] = f;
    if constexpr (L > 0) {
      g *= w;
      const double t2 = 2.0 * T;
      const double e = __builtin_fma(-t2, g, (2 * L + 1) * f);
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = __builtin_fma(t2, F[m + 1], e) * (1.0 / (2 * m + 1));
    }
  } else {
    const double r = rsqrt_fp64(T);
    F[0] = (0.88622692545275801365 * w) * r;  // sqrt(pi)/2
    if constexpr (L > 0) {
      const double h = 0.5 * r * r;
#pragma unroll
      for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * h) * F[m];
    }
  }
}

__host__ __device__ __forceinline__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

}  // namespace vsvb
