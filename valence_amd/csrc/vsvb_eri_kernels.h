// Shell-quartet ERI kernels for the VSVB energy path (gfx950, FP64).
//
// One thread evaluates one unique AO shell quartet (bra pair >= ket pair) of a
// fixed angular-momentum class with the generated Obara-Saika / Head-Gordon-
// Pople code, then contracts the cartesian block on the fly with the AO
// densities of every wave function that shares the geometry; no ERI is ever
// stored.  Work is tiled 16 bra pairs x 16 ket pairs per 256-thread
// workgroup; a fixed grid of workgroups walks the tiles (tile = block + k*grid),
// which keeps the per-workgroup partial sums few and the reduction
// deterministic.  Pairs inside a class are sorted by primitive-pair count, so
// the threads of a wavefront run primitive loops of similar length.
//
// Density contraction (closed or open shell; RT = Coulomb density, D_s the
// exchange densities with weight x_s), per unique quartet of degeneracy deg:
//   E2 += deg * sum (mn|ls) [ RT_mn RT_ls / 2 - sum_s x_s (D_ml D_ns + D_ms D_nl) / 4 ]
// This is the reference's sum over spin-orbital quartets of second-order
// cofactors times integrals (src/valence.F90:1163-1433) once the cofactors are
// written in closed form, D = C S^-1 C^T (Jacobi's complementary-minor
// identity applied to the strikes of det, src/valence.F90:1895-2056).
#pragma once
#include "vsvb_device.h"

namespace vsvb {

#include "eri_gen.inc"

struct QuartetArgs {
  const Pair* pairs;
  const PrimPair* pp;
  const Shell* shells;
  const double* boys;
  int bra0, nbra;       // pair index range of the bra class
  int ket0, nket;       // pair index range of the ket class
  int same;             // bra class == ket class: triangle only
  int ntk;              // ket tiles
  int ntiles;           // tiles of this class pair
  double tau;           // skip quartet if Q_bra * Q_ket < tau
  int nbf;
  int nwfn;             // wave functions sharing this geometry
  int nspin;            // 1 = closed shell (D = R, RT = 2R, x = 2); 2 = open (alpha, beta)
  const double* dens;   // [nwfn][nspin][nbf][nbf]
  double* partial;      // energy: [gridDim.x * 4][nwfn] per-wave partial sums
  double* jacc;         // J/K mode: Coulomb accumulators [nwfn][nbf][nbf]
  double* kacc;         // J/K mode: exchange accumulators [nwfn][nspin][nbf][nbf]
  unsigned long long* counters;  // [0] quartets evaluated, [1] primitive quartets, [2] skipped
};

constexpr int kTile = 16;

__device__ __forceinline__ void tile_of(int t, int same, int ntk, int& tb, int& tk) {
  if (same) {  // triangular enumeration: tb >= tk
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    tb = i;
    tk = t - i * (i + 1) / 2;
  } else {
    tb = t / ntk;
    tk = t % ntk;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NA, int NB, int NC, int ND>
struct DBlocks {
  double ac[NA * NC], bd[NB * ND], ad[NA * ND], bc[NB * NC];
  __device__ __forceinline__ void load(const double* __restrict__ D, int n, int ma, int mb, int mc, int md) {
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) ac[i * NC + k] = D[(size_t)(ma + i) * n + mc + k];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int l = 0; l < ND; ++l) bd[j * ND + l] = D[(size_t)(mb + j) * n + md + l];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int l = 0; l < ND; ++l) ad[i * ND + l] = D[(size_t)(ma + i) * n + md + l];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int k = 0; k < NC; ++k) bc[j * NC + k] = D[(size_t)(mb + j) * n + mc + k];
  }
};

template <int LA, int LB, int LC, int LD, bool kJK>
__global__ void __launch_bounds__(256) quartet_kernel(QuartetArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  const int n = a.nbf;
  const size_t nn = (size_t)n * n;
  const double cfac = a.nspin == 1 ? 2.0 : 1.0;  // RT = cfac * sum_s D_s
  const double xfac = a.nspin == 1 ? 2.0 : 1.0;  // exchange weight x_s
  constexpr int kMaxW = 8;
  double esum[kMaxW];
#pragma unroll
  for (int w = 0; w < kMaxW; ++w) esum[w] = 0.0;
  unsigned long long nq = 0, nprimq = 0, nskip = 0;

  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    int tb, tk;
    tile_of(tile, a.same, a.ntk, tb, tk);
    const int ib = tb * kTile + (threadIdx.x >> 4);
    const int ik = tk * kTile + (threadIdx.x & 15);
    if (!(ib < a.nbra && ik < a.nket && (!a.same || ik <= ib))) continue;
    const Pair P = a.pairs[a.bra0 + ib];
    const Pair K = a.pairs[a.ket0 + ik];
    if (P.Q * K.Q < a.tau) {
      ++nskip;
      continue;
    }
    double v[NAB * NCD];
    eri_block<LA, LB, LC, LD>(a.pp + P.pp0, P.npp, a.pp + K.pp0, K.npp, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz,
                              a.boys, v);
    ++nq;
    nprimq += (unsigned long long)P.npp * K.npp;
    const double deg = (P.offdiag ? 2.0 : 1.0) * (K.offdiag ? 2.0 : 1.0) * ((ib != ik || !a.same) ? 2.0 : 1.0);
    const int ma = a.shells[P.A].ao, mb = a.shells[P.B].ao, mc = a.shells[K.A].ao, md = a.shells[K.B].ao;
    for (int w = 0; w < a.nwfn; ++w) {
      const double* D0 = a.dens + (size_t)w * a.nspin * nn;
      double rab[NAB], rcd[NCD];
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          double t = D0[(size_t)(ma + i) * n + mb + j];
          if (a.nspin == 2) t += D0[nn + (size_t)(ma + i) * n + mb + j];
          rab[i * NB + j] = cfac * t;
        }
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int l = 0; l < ND; ++l) {
          double t = D0[(size_t)(mc + k) * n + md + l];
          if (a.nspin == 2) t += D0[nn + (size_t)(mc + k) * n + md + l];
          rcd[k * ND + l] = cfac * t;
        }
      if constexpr (!kJK) {
        double ej = 0.0;
#pragma unroll
        for (int ij = 0; ij < NAB; ++ij) {
          double t = 0.0;
#pragma unroll
          for (int kl = 0; kl < NCD; ++kl) t += v[ij * NCD + kl] * rcd[kl];
          ej += rab[ij] * t;
        }
        double ek = 0.0;
        for (int s = 0; s < a.nspin; ++s) {
          DBlocks<NA, NB, NC, ND> d;
          d.load(D0 + s * nn, n, ma, mb, mc, md);
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
              for (int k = 0; k < NC; ++k)
#pragma unroll
                for (int l = 0; l < ND; ++l)
                  ek += v[((i * NB + j) * NC + k) * ND + l] *
                        (d.ac[i * NC + k] * d.bd[j * ND + l] + d.ad[i * ND + l] * d.bc[j * NC + k]);
        }
        const double e = deg * (0.5 * ej - 0.25 * xfac * ek);
#pragma unroll
        for (int ww = 0; ww < kMaxW; ++ww)
          if (ww == w) esum[ww] += e;
      } else {
        // Coulomb: A_mn += deg/4 sum v RT_ls ; A_ls += deg/4 sum v RT_mn      (J = A + A^T)
        double* J = a.jacc + (size_t)w * nn;
        const double fj = 0.25 * deg, fk = 0.125 * deg * xfac;
#pragma unroll
        for (int ij = 0; ij < NAB; ++ij) {
          double t = 0.0;
#pragma unroll
          for (int kl = 0; kl < NCD; ++kl) t += v[ij * NCD + kl] * rcd[kl];
          unsafeAtomicAdd(&J[(size_t)(ma + ij / NB) * n + mb + ij % NB], fj * t);
        }
#pragma unroll
        for (int kl = 0; kl < NCD; ++kl) {
          double t = 0.0;
#pragma unroll
          for (int ij = 0; ij < NAB; ++ij) t += v[ij * NCD + kl] * rab[ij];
          unsafeAtomicAdd(&J[(size_t)(mc + kl / ND) * n + md + kl % ND], fj * t);
        }
        // Exchange (carries x_s): B_ml += v D_ns ; B_nl += v D_ms ; B_ms += v D_nl ; B_ns += v D_ml  (K = B + B^T)
        for (int s = 0; s < a.nspin; ++s) {
          DBlocks<NA, NB, NC, ND> d;
          d.load(D0 + s * nn, n, ma, mb, mc, md);
          double* Kx = a.kacc + ((size_t)w * a.nspin + s) * nn;
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
              double t = 0.0;
#pragma unroll
              for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int l = 0; l < ND; ++l) t += v[((i * NB + j) * NC + k) * ND + l] * d.bd[j * ND + l];
              unsafeAtomicAdd(&Kx[(size_t)(ma + i) * n + mc + k], fk * t);
            }
#pragma unroll
          for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int k = 0; k < NC; ++k) {
              double t = 0.0;
#pragma unroll
              for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int l = 0; l < ND; ++l) t += v[((i * NB + j) * NC + k) * ND + l] * d.ad[i * ND + l];
              unsafeAtomicAdd(&Kx[(size_t)(mb + j) * n + mc + k], fk * t);
            }
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
              double t = 0.0;
#pragma unroll
              for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int k = 0; k < NC; ++k) t += v[((i * NB + j) * NC + k) * ND + l] * d.bc[j * NC + k];
              unsafeAtomicAdd(&Kx[(size_t)(ma + i) * n + md + l], fk * t);
            }
#pragma unroll
          for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
              double t = 0.0;
#pragma unroll
              for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int k = 0; k < NC; ++k) t += v[((i * NB + j) * NC + k) * ND + l] * d.ac[i * NC + k];
              unsafeAtomicAdd(&Kx[(size_t)(mb + j) * n + md + l], fk * t);
            }
        }
      }
    }
  }
  if constexpr (!kJK) {
    // deterministic: shuffle-reduce each wave, one partial per (wave, wave function)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = a.nwfn < kMaxW ? a.nwfn : kMaxW;
    for (int w = 0; w < nw; ++w) {
      double s = 0.0;
#pragma unroll
      for (int ww = 0; ww < kMaxW; ++ww)
        if (ww == w) s = esum[ww];
      s = wave_sum(s);
      if (lane == 0) a.partial[((size_t)blockIdx.x * 4 + wave) * a.nwfn + w] = s;
    }
  }
  if (a.counters) {
    unsigned long long q = nq, pq = nprimq, sk = nskip;
    for (int o = 32; o > 0; o >>= 1) {
      q += __shfl_xor(q, o, 64);
      pq += __shfl_xor(pq, o, 64);
      sk += __shfl_xor(sk, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&a.counters[0], q);
      atomicAdd(&a.counters[1], pq);
      atomicAdd(&a.counters[2], sk);
    }
  }
}

// Schwarz bound of every pair of one class: Q = sqrt(max_ab |(ab|ab)|).
template <int LA, int LB>
__global__ void __launch_bounds__(64) schwarz_kernel(Pair* pairs, const PrimPair* pp, const double* boys, int p0,
                                                     int np) {
  constexpr int NAB = ncart(LA) * ncart(LB);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const Pair P = pairs[p0 + i];
  double v[NAB * NAB];
  eri_block<LA, LB, LA, LB>(pp + P.pp0, P.npp, pp + P.pp0, P.npp, P.ABx, P.ABy, P.ABz, P.ABx, P.ABy, P.ABz, boys, v);
  double m = 0.0;
#pragma unroll
  for (int k = 0; k < NAB; ++k) m = fmax(m, fabs(v[k * NAB + k]));
  pairs[p0 + i].Q = sqrt(m);
}

template <int LA, int LB, int LC, int LD>
static hipError_t launch_q(const QuartetArgs& a, bool jk, int nblocks, hipStream_t st) {
  if (nblocks <= 0 || a.ntiles <= 0) return hipSuccess;
  if (jk)
    hipLaunchKernelGGL((quartet_kernel<LA, LB, LC, LD, true>), dim3(nblocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((quartet_kernel<LA, LB, LC, LD, false>), dim3(nblocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Per-bra-class launchers (one translation unit each).
hipError_t launch_quartet_b0(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_quartet_b1(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_quartet_b2(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_quartet_b3(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_quartet_b4(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_quartet_b5(int kc, const QuartetArgs& a, bool jk, int nblocks, hipStream_t st);
hipError_t launch_schwarz_b0(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b1(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b2(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b3(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b4(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b5(Pair*, const PrimPair*, const double*, int, int, hipStream_t);

}  // namespace vsvb
