// Shell-quartet ERI kernels for the VSVB energy path (gfx950, FP64).
//
// One thread evaluates one unique AO shell quartet (bra pair >= ket pair) of a
// fixed angular-momentum class with the generated Obara-Saika / Head-Gordon-
// Pople code, then contracts the cartesian block on the fly with the AO
// densities of every wave function that shares the geometry; no ERI is ever
// stored.  In energy mode a wavefront takes a unit of 64 bra pairs (one per
// lane) and a chunk of ket pairs, visited one at a time (wave-uniform: the
// ket's pair and primitive data are scalar loads).  A grid sized to the
// occupancy takes units from an atomic counter; each wave keeps its own
// partial sum, so the partials are few (the final sum is over a fixed-size
// array).  Pairs
// inside a class are sorted by primitive-pair count, then by centre distance,
// so a wavefront's lanes run primitive loops of the same length and the
// Schwarz-skipped quartets gather in the same wavefronts.
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
  int same;             // bra class == ket class: only ket <= bra
  int ntk;              // cooperative kernel: units claimed per wavefront (1..1024)
  int ntiles;           // work units: energy: bra tiles (64 pairs) x ket chunks; J/K: strips x claims
  int kchunk;           // pairs of the wave-uniform side per claim (1..64): kets; bras in transposed J/K units
  int nkc;              // energy: ket chunks, ceil(nket / kchunk)
  int seg_tiles;        // energy: bra tiles per counter segment (nstrip segments)
  double tau;           // skip quartet if Q_bra * Q_ket < tau
  int nbf;
  int nwfn;             // wave functions sharing this geometry
  int nspin;            // 1 = closed shell (D = R, RT = 2R, x = 2); 2 = open (alpha, beta);
                        // 0 = Coulomb only (J/K mode): J of each of nwfn general symmetric densities
  const double* dens;   // [nwfn][nspin][nbf][nbf]
  double* partial;      // energy: [gridDim.x * 4][nwfn] per-wave partial sums
  double* jacc;         // J/K mode: Coulomb accumulators [nwfn][nbf][nbf]
  double* kacc;         // J/K mode: exchange accumulators [nwfn][nspin][nbf][nbf]
  unsigned long long* counters;  // [0] quartets evaluated, [1] primitive quartets, [2] skipped
  int* tile_ctr;        // dynamic tile counter(s), zero before the launch: one, or one per strip (J/K)
  int nstrip;           // straight-line units: energy segments of kSegTiles bra tiles, J/K strips (bra, or ket
                        // when transposed), each with its own work counter
  int lds_k;            // J/K mode: accumulate exchange rows in LDS (nbf * (NC + ND) doubles of dynamic LDS)
  unsigned long long* stamp;  // per-class timing (null when off): kStampSlots pairs {earliest wavefront
                              // start, latest wavefront end} in wall-clock ticks, ~0 and 0 before the launch
};

// Per-class device timing that also holds inside a replayed HIP graph: every
// wavefront stamps its start and end on the constant-rate wall clock, into one
// of kStampSlots slot pairs (by wavefront index, so that the atomics of a
// launch spread over many addresses); a reduction after the stage takes the
// minimum start and maximum end of the class.
constexpr int kStampSlots = 256;
__device__ __forceinline__ void stamp_wave(unsigned long long* st, bool end) {
  if (st && (threadIdx.x & 63) == 0) {
    const unsigned slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kStampSlots - 1);
    const unsigned long long t = (unsigned long long)wall_clock64();
    if (end)
      atomicMax(st + 2 * slot + 1, t);
    else
      atomicMin(st + 2 * slot, t);
  }
}

// Bra pairs of an energy unit (one per lane of a wavefront), and bra tiles of
// an energy counter segment when the class has several.
constexpr int kTileB = 64;
constexpr int kSegTiles = 4;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
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

// The compact Boys table of total angular momentum kLq (F_{kLq+k}(T_i)/k!),
// copied into the workgroup's LDS by all its threads before any use.
template <int kLq>
__device__ __forceinline__ void stage_boys(const double* __restrict__ boys, double* btab) {
  const double2* src = reinterpret_cast<const double2*>(boys + kBoysCOff + (size_t)kLq * kBoysCLen);
  double2* dst = reinterpret_cast<double2*>(btab);
  for (int e = threadIdx.x; e < kBoysCLen / 2; e += blockDim.x) dst[e] = src[e];
  __syncthreads();
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void quartet_energy_body(const QuartetArgs& a, const double* __restrict__ btab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  const int n = a.nbf;
  const size_t nn = (size_t)n * n;
  const double cfac = a.nspin == 1 ? 2.0 : 1.0;  // RT = cfac * sum_s D_s
  const double xfac = a.nspin == 1 ? 2.0 : 1.0;  // exchange weight x_s
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int kMaxW = 8;
  double esum[kMaxW];
#pragma unroll
  for (int w = 0; w < kMaxW; ++w) esum[w] = 0.0;
  unsigned long long nq = 0, nprimq = 0, nskip = 0;
  stamp_wave(a.stamp, false);

  // Work units (64 bra pairs x a chunk of kchunk ket pairs) are handed to
  // wavefronts one at a time from atomic counters, with no barrier, so
  // wavefronts never wait for each other.  The unit indices form nstrip
  // segments of seg_tiles * nkc units, each with its own counter: a wavefront
  // starts on segment (global wave index) mod nstrip and moves to the next one
  // only when its segment is exhausted, leaving after it has found every
  // segment exhausted once (counters only grow, so every wavefront ends);
  // spreading the claims over several counters keeps them from queueing on
  // one address.  Pairs sort heaviest first (most primitive pairs), so low
  // unit indices hold the heaviest quartets and the tail is short.  With
  // chunks of several kets, one lane per ket tests it
  // against the largest bra bound of the tile, and the wavefront visits only
  // the kets that can pass the Schwarz test for some lane: sparse systems,
  // where most (tile, ket) slots are screened out, then cost one counter claim
  // per chunk instead of one per ket.
  const int nbt = (a.nbra + kTileB - 1) / kTileB;
  int sg = (int)((blockIdx.x * 4u + (unsigned)wave) % (unsigned)a.nstrip), left = a.nstrip;
  for (;;) {
    int ul = 0;
    if (lane == 0) ul = atomicAdd(a.tile_ctr + sg, 1);
    ul = __builtin_amdgcn_readfirstlane(__shfl(ul, 0, 64));
    const int t0 = sg * a.seg_tiles;
    if (ul >= min(a.seg_tiles, nbt - t0) * a.nkc) {  // this segment is exhausted
      if (--left == 0) break;
      sg = sg + 1 == a.nstrip ? 0 : sg + 1;
      continue;
    }
    // Bra-tile-major order, except for one-ket units of different bra and ket
    // classes: ket-major there, so that the wavefronts in flight share few
    // kets and their primitive data stay in the scalar caches (2% faster on
    // C8H18's off-diagonal classes; slower for LiF128's 64-ket chunks, and a
    // same-class tile's kets end at its diagonal).
    const int u = t0 * a.nkc + ul;
    const bool km = !a.same && a.kchunk == 1;
    const int kci = km ? u / nbt : 0;
    const int tb = km ? u - kci * nbt : u / a.nkc;
    const int k0 = __builtin_amdgcn_readfirstlane((km ? kci : u - tb * a.nkc) * a.kchunk);
    const int klast = tb * kTileB + kTileB - 1;  // same class: last ket any lane of the tile takes
    if (a.same && k0 > klast) continue;  // whole wave above the diagonal
    const int ib = tb * kTileB + lane;
    const bool bra_ok = ib < a.nbra;
    Pair P{};
    if (bra_ok) P = a.pairs[a.bra0 + ib];
    unsigned long long live = 1;
    if (a.kchunk > 1) {
      const double qmax = wave_max(bra_ok ? P.Q : 0.0);
      const int ikl = k0 + lane;
      bool lv = false;
      if (lane < a.kchunk && ikl < a.nket && (!a.same || ikl <= klast)) {
        lv = a.pairs[a.ket0 + ikl].Q * qmax >= a.tau;
        if (!lv) {  // every bra lane that takes this ket is screened out
          const int lo = a.same ? max(ikl, tb * kTileB) : tb * kTileB;
          const int hi = min(a.nbra, tb * kTileB + kTileB);
          nskip += hi > lo ? hi - lo : 0;
        }
      }
      live = __ballot(lv);
    }
    while (live) {
      const int ik = __builtin_amdgcn_readfirstlane(k0 + __builtin_ctzll(live));
      live &= live - 1;
      const Pair K = a.pairs[a.ket0 + ik];
      bool act = bra_ok && (!a.same || ik <= ib);
      if (act && P.Q * K.Q < a.tau) {
        act = false;
        ++nskip;
      }
      double v[NAB * NCD];
      if (act) {
        eri_block<LA, LB, LC, LD>(a.pp + P.pp0, P.npp, a.pp + K.pp0, K.npp, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy,
                                  K.ABz, btab, v);
        ++nq;
        nprimq += (unsigned long long)P.npp * K.npp;
      } else {
#pragma unroll
        for (int k = 0; k < NAB * NCD; ++k) v[k] = 0.0;
      }
      if (!__any(act)) continue;  // wave-uniform: nothing to contract
      const double deg =
          act ? (P.offdiag ? 2.0 : 1.0) * (K.offdiag ? 2.0 : 1.0) * ((ib != ik || !a.same) ? 2.0 : 1.0) : 0.0;
      // inactive lanes read valid blocks and contribute zero
      const int ma = a.shells[P.A].ao, mb = a.shells[P.B].ao;
      const int mc = a.shells[K.A].ao, md = a.shells[K.B].ao;
      for (int w = 0; w < a.nwfn; ++w) {
        const double* D0 = a.dens + (size_t)w * (a.nspin > 0 ? a.nspin : 1) * nn;
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
      }
    }
  }
  {
    // deterministic: shuffle-reduce each wave, one partial per (wave, wave function)
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
  stamp_wave(a.stamp, true);
  if (a.counters) {
    unsigned long long q = nq, pq = nprimq, sk = nskip;
    for (int o = 32; o > 0; o >>= 1) {
      q += __shfl_xor(q, o, 64);
      pq += __shfl_xor(pq, o, 64);
      sk += __shfl_xor(sk, o, 64);
    }
    if (lane == 0) {
      atomicAdd(&a.counters[0], q);
      atomicAdd(&a.counters[1], pq);
      atomicAdd(&a.counters[2], sk);
    }
  }
}


// J/K (adjoint) contraction.  A workgroup takes one ket pair and a strip of
// four 64-pair bra tiles (one per wavefront) from a counter.  Coulomb:
// A_mn += deg/4 sum v RT_ls per lane and A_ls += deg/4 sum v RT_mn reduced over
// the wave (J = A + A^T).  Exchange (carries x_s): B_ml += v D_ns, B_nl += v D_ms,
// B_ms += v D_nl, B_ns += v D_ml (K = B + B^T): every update lands in a row of
// the bra AOs and in one of the NC + ND columns of the unit's ket AOs, so with
// lds_k the unit accumulates them in an nbf x (NC + ND) LDS block (LDS atomics)
// and writes each nonzero entry to HBM once per unit, instead of one global
// atomic per quartet element; without lds_k (block too large) they go straight
// to the global accumulators.
constexpr int kStripB = 4 * kTileB;

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void quartet_jk_body(const QuartetArgs& a, const double* __restrict__ btab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NK = NC + ND;
  extern __shared__ double kbuf[];
  __shared__ int s_unit;
  __shared__ double s_qmax[4];  // per-wavefront largest Schwarz bound of the strip
  const int n = a.nbf;
  const size_t nn = (size_t)n * n;
  const int ns = a.nspin;
  const int dstride = ns > 0 ? ns : 1;
  const double cfac = ns == 1 ? 2.0 : 1.0;
  const double xfac = ns == 1 ? 2.0 : 1.0;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool lds = a.lds_k != 0;
  const int kb_len = lds ? n * NK : 0;
  for (int e = threadIdx.x; e < kb_len; e += blockDim.x) kbuf[e] = 0.0;
  unsigned long long nq = 0, nprimq = 0, nskip = 0;
  stamp_wave(a.stamp, false);
  // Units are (bra strip, ket pair).  Every strip has its own counter of ket
  // pairs: a workgroup starts on strip blockIdx mod nstrip and moves to the
  // next strip only when its strip is exhausted, leaving after it has found
  // every strip exhausted once (counters only grow, so every workgroup ends).
  // For one density, a lane keeps the Coulomb update of its bra block in
  // registers while its workgroup stays on one strip.
  constexpr bool kRegJ = NAB <= 9;
  const bool racc = kRegJ && a.nwfn == 1;
  double jreg[kRegJ ? NAB : 1];
#pragma unroll
  for (int k = 0; k < (kRegJ ? NAB : 1); ++k) jreg[k] = 0.0;
  int jst = -1;  // strip whose bra block jreg holds
  auto flush_j = [&](int s_) {
    const int ibf = s_ * kStripB + wave * kTileB + lane;
    if (s_ < 0 || ibf >= a.nbra) return;
    const Pair Pf = a.pairs[a.bra0 + ibf];
    const int fa = a.shells[Pf.A].ao, fb = a.shells[Pf.B].ao;
#pragma unroll
    for (int ij = 0; ij < (kRegJ ? NAB : 1); ++ij) {
      if (jreg[ij] != 0.0) unsafeAtomicAdd(&a.jacc[(size_t)(fa + ij / NB) * n + fb + ij % NB], jreg[ij]);
      jreg[ij] = 0.0;
    }
  };
  // A claim takes a chunk of kchunk consecutive ket pairs of the strip; with
  // more than one, kets that fail the Schwarz test against the strip's largest
  // bra bound are skipped without a visit (one lane tests each), which spares
  // sparse systems a claim and an exchange flush per screened-out ket.
  int st = (int)(blockIdx.x % (unsigned)a.nstrip), left = a.nstrip;
  for (;;) {
    __syncthreads();  // the previous unit is flushed and s_unit, s_qmax are free
    if (threadIdx.x == 0) s_unit = atomicAdd(a.tile_ctr + st, a.kchunk);
    __syncthreads();
    const int k0 = __builtin_amdgcn_readfirstlane(s_unit);
    if (k0 >= a.nket) {  // this strip is exhausted
      if (--left == 0) break;
      st = st + 1 == a.nstrip ? 0 : st + 1;
      continue;
    }
    const int klast = st * kStripB + kStripB - 1;  // same class: last ket any lane of the strip takes
    if (a.same && k0 > klast) continue;            // the whole strip is above the diagonal
    if (racc && jst != st) {
      flush_j(jst);
      jst = st;
    }
    const int tile0 = st * kStripB + wave * kTileB;
    const int ib = tile0 + lane;
    // every lane holds a pair of the bra class (the last one past the end), so
    // that the density blocks read below stay inside the matrix
    const Pair P = a.pairs[a.bra0 + (ib < a.nbra ? ib : a.nbra - 1)];
    unsigned long long live = 1;
    if (a.kchunk > 1) {
      const double qw = wave_max(ib < a.nbra ? P.Q : 0.0);
      if (lane == 0) s_qmax[wave] = qw;
      __syncthreads();
      const double qmax = fmax(fmax(s_qmax[0], s_qmax[1]), fmax(s_qmax[2], s_qmax[3]));
      const int ikl = k0 + lane;
      bool lv = false;
      if (lane < a.kchunk && ikl < a.nket && (!a.same || ikl <= klast)) {
        lv = a.pairs[a.ket0 + ikl].Q * qmax >= a.tau;
        if (!lv) {  // every bra lane of this wavefront that takes the ket is screened out
          const int lo = a.same ? max(ikl, tile0) : tile0, hi = min(a.nbra, tile0 + kTileB);
          nskip += hi > lo ? hi - lo : 0;
        }
      }
      live = __ballot(lv);  // the same in every wavefront of the workgroup
    }
    while (live) {
      const int ik = __builtin_amdgcn_readfirstlane(k0 + __builtin_ctzll(live));
      live &= live - 1;
      const Pair K = a.pairs[a.ket0 + ik];
      bool act = ib < a.nbra && (!a.same || ik <= ib);
      if (act && P.Q * K.Q < a.tau) {
        act = false;
        ++nskip;
      }
      double v[NAB * NCD];
      if (act) {
        eri_block<LA, LB, LC, LD>(a.pp + P.pp0, P.npp, a.pp + K.pp0, K.npp, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy,
                                  K.ABz, btab, v);
        ++nq;
        nprimq += (unsigned long long)P.npp * K.npp;
      } else {
#pragma unroll
        for (int k = 0; k < NAB * NCD; ++k) v[k] = 0.0;
      }
      const bool wact = __any(act);  // wave-uniform
      const double deg = act ? (P.offdiag ? 2.0 : 1.0) * (K.offdiag ? 2.0 : 1.0) * ((ib != ik || !a.same) ? 2.0 : 1.0)
                             : 0.0;
      // inactive lanes read a valid (ket) block and contribute zero
      const int ma = act ? a.shells[P.A].ao : a.shells[K.A].ao, mb = act ? a.shells[P.B].ao : a.shells[K.B].ao;
      const int mc = a.shells[K.A].ao, md = a.shells[K.B].ao;
      const double fj = 0.25 * deg, fk = 0.125 * deg * xfac;
      for (int w = 0; w < a.nwfn; ++w) {
        const double* D0 = a.dens + (size_t)w * dstride * nn;
        if (wact) {
          double rab[NAB], rcd[NCD];
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
              double t = D0[(size_t)(ma + i) * n + mb + j];
              if (ns == 2) t += D0[nn + (size_t)(ma + i) * n + mb + j];
              rab[i * NB + j] = cfac * t;
            }
#pragma unroll
          for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
              double t = D0[(size_t)(mc + k) * n + md + l];
              if (ns == 2) t += D0[nn + (size_t)(mc + k) * n + md + l];
              rcd[k * ND + l] = cfac * t;
            }
          double* J = a.jacc + (size_t)w * nn;
          if (act) {
#pragma unroll
            for (int ij = 0; ij < NAB; ++ij) {
              double t = 0.0;
#pragma unroll
              for (int kl = 0; kl < NCD; ++kl) t += v[ij * NCD + kl] * rcd[kl];
              if constexpr (kRegJ) {
                if (racc) {
                  jreg[ij] += fj * t;
                  continue;
                }
              }
              unsafeAtomicAdd(&J[(size_t)(ma + ij / NB) * n + mb + ij % NB], fj * t);
            }
          }
#pragma unroll
          for (int kl = 0; kl < NCD; ++kl) {
            double t = 0.0;
#pragma unroll
            for (int ij = 0; ij < NAB; ++ij) t += v[ij * NCD + kl] * rab[ij];
            t = wave_sum(fj * t);
            if (lane == 0 && t != 0.0) unsafeAtomicAdd(&J[(size_t)(mc + kl / ND) * n + md + kl % ND], t);
          }
        }
        for (int s = 0; s < ns; ++s) {
          double* Kx = a.kacc + ((size_t)w * ns + s) * nn;
          if (act) {
            DBlocks<NA, NB, NC, ND> d;
            d.load(D0 + s * nn, n, ma, mb, mc, md);
            // target (row, ket column c < NK): LDS block or the global accumulator
            auto put = [&](int row, int c, double x) {
              if (lds)
                unsafeAtomicAdd(&kbuf[row * NK + c], x);
              else
                unsafeAtomicAdd(&Kx[(size_t)row * n + (c < NC ? mc + c : md + c - NC)], x);
            };
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
              for (int k = 0; k < NC; ++k) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                  for (int l = 0; l < ND; ++l) t += v[((i * NB + j) * NC + k) * ND + l] * d.bd[j * ND + l];
                put(ma + i, k, fk * t);
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
                put(mb + j, k, fk * t);
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
                put(ma + i, NC + l, fk * t);
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
                put(mb + j, NC + l, fk * t);
              }
          }
          if (lds) {
            __syncthreads();
            for (int e = threadIdx.x; e < kb_len; e += blockDim.x) {
              const double x = kbuf[e];
              if (x != 0.0) {
                kbuf[e] = 0.0;
                const int row = e / NK, c = e - row * NK;
                unsafeAtomicAdd(&Kx[(size_t)row * n + (c < NC ? mc + c : md + c - NC)], x);
              }
            }
            __syncthreads();
          }
        }
      }
    }
  }
  if (racc) flush_j(jst);
  stamp_wave(a.stamp, true);
  if (a.counters) {
    unsigned long long q = nq, pq = nprimq, sk = nskip;
    for (int o = 32; o > 0; o >>= 1) {
      q += __shfl_xor(q, o, 64);
      pq += __shfl_xor(pq, o, 64);
      sk += __shfl_xor(sk, o, 64);
    }
    if (lane == 0) {
      atomicAdd(&a.counters[0], q);
      atomicAdd(&a.counters[1], pq);
      atomicAdd(&a.counters[2], sk);
    }
  }
}

// Transposed J/K unit, for classes whose bra and ket pair classes differ: one
// bra pair for the whole workgroup and a strip of 256 ket pairs, one per lane.
// Units are ket strips x bra pairs (bra pair fastest).  The bra pair's
// primitive data are scalar loads; the exchange rows of the bra's AOs (NA + NB
// rows of nbf) are accumulated in LDS and flushed once per unit; the Coulomb
// block of the bra is a wave sum, those of the kets per-lane atomics.  Suits
// classes with few bra pairs (a 256-pair bra strip would leave most lanes
// idle) or large bra blocks (fewer global atomics per quartet).
constexpr int kStripK = 256;

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void quartet_jkt_body(const QuartetArgs& a, const double* __restrict__ btab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NR = NA + NB;
  extern __shared__ double kbuf[];
  __shared__ int s_unit;
  __shared__ double s_qmax[4];  // per-wavefront largest Schwarz bound of the strip
  const int n = a.nbf;
  const size_t nn = (size_t)n * n;
  const int ns = a.nspin;
  const int dstride = ns > 0 ? ns : 1;
  const double cfac = ns == 1 ? 2.0 : 1.0;
  const double xfac = ns == 1 ? 2.0 : 1.0;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool lds = a.lds_k != 0;
  const int kb_len = lds ? n * NR : 0;
  for (int e = threadIdx.x; e < kb_len; e += blockDim.x) kbuf[e] = 0.0;
  unsigned long long nq = 0, nprimq = 0, nskip = 0;
  stamp_wave(a.stamp, false);
  // One counter of bra pairs per ket strip, as in the normal units: a
  // workgroup stays on its ket strip until it is exhausted, and for one
  // density a lane keeps the Coulomb update of its ket block in registers.
  constexpr bool kRegJ = NCD <= 9;
  const bool racc = kRegJ && a.nwfn == 1;
  double jreg[kRegJ ? NCD : 1];
#pragma unroll
  for (int k = 0; k < (kRegJ ? NCD : 1); ++k) jreg[k] = 0.0;
  int jst = -1;  // ket strip whose ket block jreg holds
  auto flush_j = [&](int s_) {
    const int ikf = s_ * kStripK + wave * kTileB + lane;
    if (s_ < 0 || ikf >= a.nket) return;
    const Pair Kf = a.pairs[a.ket0 + ikf];
    const int fc = a.shells[Kf.A].ao, fd = a.shells[Kf.B].ao;
#pragma unroll
    for (int kl = 0; kl < (kRegJ ? NCD : 1); ++kl) {
      if (jreg[kl] != 0.0) unsafeAtomicAdd(&a.jacc[(size_t)(fc + kl / ND) * n + fd + kl % ND], jreg[kl]);
      jreg[kl] = 0.0;
    }
  };
  // A claim takes a chunk of kchunk consecutive bra pairs; with more than one,
  // bra pairs that fail the Schwarz test against the strip's largest ket bound
  // are skipped without a visit.
  int st = (int)(blockIdx.x % (unsigned)a.nstrip), left = a.nstrip;
  for (;;) {
    __syncthreads();  // the previous unit is flushed and s_unit, s_qmax are free
    if (threadIdx.x == 0) s_unit = atomicAdd(a.tile_ctr + st, a.kchunk);
    __syncthreads();
    const int b0 = __builtin_amdgcn_readfirstlane(s_unit);
    if (b0 >= a.nbra) {  // this ket strip is exhausted
      if (--left == 0) break;
      st = st + 1 == a.nstrip ? 0 : st + 1;
      continue;
    }
    if (racc && jst != st) {
      flush_j(jst);
      jst = st;
    }
    const int tile0 = st * kStripK + wave * kTileB;
    const int ik = tile0 + lane;
    const bool kin = ik < a.nket;
    // every lane holds a pair of the ket class (the last one past the end)
    const Pair K = a.pairs[a.ket0 + (kin ? ik : a.nket - 1)];
    unsigned long long live = 1;
    if (a.kchunk > 1) {
      const double qw = wave_max(kin ? K.Q : 0.0);
      if (lane == 0) s_qmax[wave] = qw;
      __syncthreads();
      const double qmax = fmax(fmax(s_qmax[0], s_qmax[1]), fmax(s_qmax[2], s_qmax[3]));
      const int ibl = b0 + lane;
      bool lv = false;
      if (lane < a.kchunk && ibl < a.nbra) {
        lv = a.pairs[a.bra0 + ibl].Q * qmax >= a.tau;
        if (!lv) {  // every ket lane of this wavefront is screened out
          const int hi = min(a.nket, tile0 + kTileB);
          nskip += hi > tile0 ? hi - tile0 : 0;
        }
      }
      live = __ballot(lv);  // the same in every wavefront of the workgroup
    }
    while (live) {
      const int ib = __builtin_amdgcn_readfirstlane(b0 + __builtin_ctzll(live));
      live &= live - 1;
      const Pair P = a.pairs[a.bra0 + ib];
      bool act = kin;
      if (act && P.Q * K.Q < a.tau) {
        act = false;
        ++nskip;
      }
      double v[NAB * NCD];
      if (act) {
        eri_block_ub<LA, LB, LC, LD>(a.pp + P.pp0, P.npp, a.pp + K.pp0, K.npp, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy,
                                     K.ABz, btab, v);
        ++nq;
        nprimq += (unsigned long long)P.npp * K.npp;
      } else {
#pragma unroll
        for (int k = 0; k < NAB * NCD; ++k) v[k] = 0.0;
      }
      const bool wact = __any(act);  // wave-uniform
      // bra and ket classes differ, so every quartet stands for its (cd|ab) image too
      const double deg = act ? (P.offdiag ? 2.0 : 1.0) * (K.offdiag ? 2.0 : 1.0) * 2.0 : 0.0;
      const int ma = a.shells[P.A].ao, mb = a.shells[P.B].ao;
      const int mc = a.shells[K.A].ao, md = a.shells[K.B].ao;
      const double fj = 0.25 * deg, fk = 0.125 * deg * xfac;
      for (int w = 0; w < a.nwfn; ++w) {
        const double* D0 = a.dens + (size_t)w * dstride * nn;
        if (wact) {
          double rab[NAB], rcd[NCD];
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
              double t = D0[(size_t)(ma + i) * n + mb + j];
              if (ns == 2) t += D0[nn + (size_t)(ma + i) * n + mb + j];
              rab[i * NB + j] = cfac * t;
            }
#pragma unroll
          for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int l = 0; l < ND; ++l) {
              double t = D0[(size_t)(mc + k) * n + md + l];
              if (ns == 2) t += D0[nn + (size_t)(mc + k) * n + md + l];
              rcd[k * ND + l] = cfac * t;
            }
          double* J = a.jacc + (size_t)w * nn;
          if (act) {
#pragma unroll
            for (int kl = 0; kl < NCD; ++kl) {
              double t = 0.0;
#pragma unroll
              for (int ij = 0; ij < NAB; ++ij) t += v[ij * NCD + kl] * rab[ij];
              if constexpr (kRegJ) {
                if (racc) {
                  jreg[kl] += fj * t;
                  continue;
                }
              }
              unsafeAtomicAdd(&J[(size_t)(mc + kl / ND) * n + md + kl % ND], fj * t);
            }
          }
#pragma unroll
          for (int ij = 0; ij < NAB; ++ij) {
            double t = 0.0;
#pragma unroll
            for (int kl = 0; kl < NCD; ++kl) t += v[ij * NCD + kl] * rcd[kl];
            t = wave_sum(fj * t);
            if (lane == 0 && t != 0.0) unsafeAtomicAdd(&J[(size_t)(ma + ij / NB) * n + mb + ij % NB], t);
          }
        }
        for (int s = 0; s < ns; ++s) {
          double* Kx = a.kacc + ((size_t)w * ns + s) * nn;
          if (act) {
            DBlocks<NA, NB, NC, ND> d;
            d.load(D0 + s * nn, n, ma, mb, mc, md);
            // target (bra row r < NR, column col): LDS block or the global accumulator
            auto put = [&](int r, int col, double x) {
              if (lds)
                unsafeAtomicAdd(&kbuf[r * n + col], x);
              else
                unsafeAtomicAdd(&Kx[(size_t)(r < NA ? ma + r : mb + r - NA) * n + col], x);
            };
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
              for (int k = 0; k < NC; ++k) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
#pragma unroll
                  for (int l = 0; l < ND; ++l) t += v[((i * NB + j) * NC + k) * ND + l] * d.bd[j * ND + l];
                put(i, mc + k, fk * t);
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
                put(NA + j, mc + k, fk * t);
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
                put(i, md + l, fk * t);
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
                put(NA + j, md + l, fk * t);
              }
          }
          if (lds) {
            __syncthreads();
            for (int e = threadIdx.x; e < kb_len; e += blockDim.x) {
              const double x = kbuf[e];
              if (x != 0.0) {
                kbuf[e] = 0.0;
                const int r = e / n, col = e - r * n;
                unsafeAtomicAdd(&Kx[(size_t)(r < NA ? ma + r : mb + r - NA) * n + col], x);
              }
            }
            __syncthreads();
          }
        }
      }
    }
  }
  if (racc) flush_j(jst);
  stamp_wave(a.stamp, true);
  if (a.counters) {
    unsigned long long q = nq, pq = nprimq, sk = nskip;
    for (int o = 32; o > 0; o >>= 1) {
      q += __shfl_xor(q, o, 64);
      pq += __shfl_xor(pq, o, 64);
      sk += __shfl_xor(sk, o, 64);
    }
    if (lane == 0) {
      atomicAdd(&a.counters[0], q);
      atomicAdd(&a.counters[1], pq);
      atomicAdd(&a.counters[2], sk);
    }
  }
}

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) quartet_jkt_kernel(QuartetArgs a) {
  __shared__ __attribute__((aligned(16))) double btab[kBoysCLen];  // this class's compact Boys table
  stage_boys<LA + LB + LC + LD>(a.boys, btab);
  quartet_jkt_body<LA, LB, LC, LD>(a, btab);
}

template <int LA, int LB, int LC, int LD, bool kJK>
__global__ void __launch_bounds__(256) quartet_kernel(QuartetArgs a) {
  __shared__ __attribute__((aligned(16))) double btab[kBoysCLen];  // this class's compact Boys table
  stage_boys<LA + LB + LC + LD>(a.boys, btab);
  if constexpr (kJK)
    quartet_jk_body<LA, LB, LC, LD>(a, btab);
  else
    quartet_energy_body<LA, LB, LC, LD>(a, btab);
}

// One J/K launch of a class over several groups (molecules) of a batch, with
// persistent workgroups: they drain group 0's work units, then group 1's, and
// so on.  Each group's units come from its own counters, so a workgroup that
// finds one group exhausted starts on the next while others finish the
// previous one: no launch ends with a tail, and one launch replaces one per
// group.  Every group's arguments hold its own buffers; the Boys table values
// are the same in every context, so the first group's is staged.  stamp: the
// launch's wall-clock slots (profiling) or null.
template <int LA, int LB, int LC, int LD, bool kTr>
__global__ void __launch_bounds__(256) quartet_jk_groups_kernel(const QuartetArgs* __restrict__ gs, int ng,
                                                               unsigned long long* stamp) {
  __shared__ __attribute__((aligned(16))) double btab[kBoysCLen];
  stage_boys<LA + LB + LC + LD>(gs[0].boys, btab);
  stamp_wave(stamp, false);
  for (int g = 0; g < ng; ++g) {
    const QuartetArgs a = gs[g];
    if (a.ntiles > 0) {
      if constexpr (kTr)
        quartet_jkt_body<LA, LB, LC, LD>(a, btab);
      else
        quartet_jk_body<LA, LB, LC, LD>(a, btab);
    }
    __syncthreads();  // the group's LDS blocks are flushed before the next group's zeroing
  }
  stamp_wave(stamp, true);
}

// Schwarz bound of every pair of one class, Q = sqrt(max_ab |(ab|ab)|): one
// wavefront per pair, the lanes split the bra primitive pairs and the partial
// diagonal integrals are summed across the wave.
template <int LA, int LB>
__global__ void __launch_bounds__(256) schwarz_kernel(Pair* pairs, const PrimPair* pp, const double* boys, int p0,
                                                      int np) {
  constexpr int NAB = ncart(LA) * ncart(LB);
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (i >= np) return;
  const Pair P = pairs[p0 + i];
  double dsum[NAB];
#pragma unroll
  for (int k = 0; k < NAB; ++k) dsum[k] = 0.0;
  for (int b = lane; b < P.npp; b += 64) {
    double v[NAB * NAB];
    eri_block<LA, LB, LA, LB>(pp + P.pp0 + b, 1, pp + P.pp0, P.npp, P.ABx, P.ABy, P.ABz, P.ABx, P.ABy, P.ABz,
                              boys + kBoysCOff + (size_t)(2 * (LA + LB)) * kBoysCLen, v);
#pragma unroll
    for (int k = 0; k < NAB; ++k) dsum[k] += v[k * NAB + k];
  }
  double m = 0.0;
#pragma unroll
  for (int k = 0; k < NAB; ++k) m = fmax(m, fabs(wave_sum(dsum[k])));
  if (lane == 0) pairs[p0 + i].Q = sqrt(m);
}

template <int LA, int LB, int LC, int LD>
static int occupancy_q(bool jk, bool tr, size_t lds_bytes) {
  int n = 0;
  if constexpr (LA != LC || LB != LD) {
    if (jk && tr) {
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, quartet_jkt_kernel<LA, LB, LC, LD>, 256, lds_bytes);
      return n > 0 ? n : 1;
    }
  }
  if (jk)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, quartet_kernel<LA, LB, LC, LD, true>, 256, lds_bytes);
  else
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, quartet_kernel<LA, LB, LC, LD, false>, 256, 0);
  return n > 0 ? n : 1;
}

// Private (scratch) bytes per lane of the straight-line kernel of a class in
// one mode and orientation: the register spills of its generated code.
template <int LA, int LB, int LC, int LD>
static size_t scratch_q(bool jk, bool tr) {
  hipFuncAttributes at{};
  if constexpr (LA != LC || LB != LD) {
    if (jk && tr) {
      (void)hipFuncGetAttributes(&at, reinterpret_cast<const void*>(quartet_jkt_kernel<LA, LB, LC, LD>));
      return at.localSizeBytes;
    }
  }
  if (jk)
    (void)hipFuncGetAttributes(&at, reinterpret_cast<const void*>(quartet_kernel<LA, LB, LC, LD, true>));
  else
    (void)hipFuncGetAttributes(&at, reinterpret_cast<const void*>(quartet_kernel<LA, LB, LC, LD, false>));
  return at.localSizeBytes;
}

template <int LA, int LB, int LC, int LD>
static hipError_t launch_q(const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds_bytes, hipStream_t st) {
  if (nblocks <= 0 || a.ntiles <= 0) return hipSuccess;
  if constexpr (LA != LC || LB != LD) {
    if (jk && tr) {
      hipLaunchKernelGGL((quartet_jkt_kernel<LA, LB, LC, LD>), dim3(nblocks), dim3(256), lds_bytes, st, a);
      return hipGetLastError();
    }
  } else {
    if (tr) return hipErrorInvalidValue;  // no transposed kernel for a class paired with itself
  }
  if (jk)
    hipLaunchKernelGGL((quartet_kernel<LA, LB, LC, LD, true>), dim3(nblocks), dim3(256), lds_bytes, st, a);
  else
    hipLaunchKernelGGL((quartet_kernel<LA, LB, LC, LD, false>), dim3(nblocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
static hipError_t launch_q_groups(const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds_bytes,
                                  unsigned long long* stamp, hipStream_t st) {
  if (nblocks <= 0 || ng <= 0) return hipSuccess;
  if constexpr (LA != LC || LB != LD) {
    if (tr) {
      hipLaunchKernelGGL((quartet_jk_groups_kernel<LA, LB, LC, LD, true>), dim3(nblocks), dim3(256), lds_bytes, st, gs,
                         ng, stamp);
      return hipGetLastError();
    }
  } else {
    if (tr) return hipErrorInvalidValue;  // no transposed kernel for a class paired with itself
  }
  hipLaunchKernelGGL((quartet_jk_groups_kernel<LA, LB, LC, LD, false>), dim3(nblocks), dim3(256), lds_bytes, st, gs, ng,
                     stamp);
  return hipGetLastError();
}

// Per-bra-class launchers (one translation unit each).
hipError_t launch_quartet_b0(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b0(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
hipError_t launch_quartet_b1(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b1(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
hipError_t launch_quartet_b2(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b2(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
hipError_t launch_quartet_b3(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b3(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
hipError_t launch_quartet_b4(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b4(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
hipError_t launch_quartet_b5(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st);
hipError_t launch_groups_b5(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st);
int occupancy_b0(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b0(int kc, bool jk, bool tr);
int occupancy_b1(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b1(int kc, bool jk, bool tr);
int occupancy_b2(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b2(int kc, bool jk, bool tr);
int occupancy_b3(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b3(int kc, bool jk, bool tr);
int occupancy_b4(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b4(int kc, bool jk, bool tr);
int occupancy_b5(int kc, bool jk, bool tr, size_t lds);
size_t scratch_b5(int kc, bool jk, bool tr);
hipError_t launch_schwarz_b0(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b1(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b2(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b3(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b4(Pair*, const PrimPair*, const double*, int, int, hipStream_t);
hipError_t launch_schwarz_b5(Pair*, const PrimPair*, const double*, int, int, hipStream_t);

}  // namespace vsvb
