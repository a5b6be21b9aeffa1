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

// Boys table: F_k(T_i) for T_i = i*dT, i < kBoysN, k < kBoysK (k = L + 0..7).
constexpr double kBoysDT = 0.05;
constexpr double kBoysInvDT = 20.0;
constexpr double kBoysTMax = 36.0;  // above: asymptotic F_0 and upward recursion
constexpr int kBoysN = 722;
constexpr int kBoysK = kLTot + 8;
constexpr int kBoysTerms = 8;

struct PrimPair {
  double p, oo2p, Px, Py, Pz, PAx, PAy, PAz, K, pad;
};

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

// Boys function F_m(T), m = 0..L.  For T < kBoysTMax: Taylor expansion of
// F_L about the nearest grid point (F_m' = -F_{m+1}), 8 terms, then the
// downward recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1).  Above: F_0 =
// sqrt(pi/T)/2 (erfc(6) < 3e-17) and the upward recursion, stable for T > L.
template <int L>
__device__ __forceinline__ void boys_eval(double T, const double* __restrict__ tab, double* F) {
  if (T < kBoysTMax) {
    const int i = __double2int_rn(T * kBoysInvDT);
    const double d = i * kBoysDT - T;
    const double* t = tab + i * kBoysK + L;
    double f = t[7];
    f = t[6] + f * d * (1.0 / 7.0);
    f = t[5] + f * d * (1.0 / 6.0);
    f = t[4] + f * d * (1.0 / 5.0);
    f = t[3] + f * d * (1.0 / 4.0);
    f = t[2] + f * d * (1.0 / 3.0);
    f = t[1] + f * d * (1.0 / 2.0);
    f = t[0] + f * d;
    F[L] = f;
    if constexpr (L > 0) {
      const double e = exp(-T);
      const double t2 = 2.0 * T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
    }
  } else {
    const double rT = 1.0 / T;
    F[0] = 0.88622692545275801365 * sqrt(rT);  // sqrt(pi)/2
    if constexpr (L > 0) {
      const double e = exp(-T);
      const double h = 0.5 * rT;
#pragma unroll
      for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
    }
  }
}

__host__ __device__ __forceinline__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

}  // namespace vsvb
