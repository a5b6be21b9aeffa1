// Cooperative shell-quartet kernel for the classes whose one-lane-per-quartet
// code does not fit the register file ((dd|pp), (dd|dp), (dd|dd)): one
// wavefront evaluates one contracted quartet.
//
// The same Obara-Saika VRR and Head-Gordon-Pople HRR as the straight-line
// kernels (csrc/gen_eri.py), executed from step tables: every intermediate
// lives in LDS, the steps of one dependency level are spread over the 64 lanes
// and a barrier separates the levels.  The contracted block is then contracted
// with the densities with the lanes over its elements (energy), or over the
// index pairs of each Coulomb / exchange target with one atomic per element
// (J/K mode).  Degeneracy factors and accumulation conventions are those of
// the straight-line kernels (vsvb_eri_kernels.h).
#include "vsvb_internal.h"

#include <mutex>

namespace vsvb {

struct CoopDesc {
  int la, lb, lc, ld;
  int nT, nV, nH, lds;  // targets, VRR slots, HRR slots, LDS doubles
  int nbase, base;      // base values: (slot, m) pairs
  int nvl, voff, vsteps;
  int targ;
  int nhl, hoff, hsteps;
  int out;
  int dflt;  // used by default (the batch tuner may switch any class with tables)
};

#include "coop_tables.inc"

namespace {

__device__ __forceinline__ void boys_any(int L, double T, const double* __restrict__ tab, double* F) {
  switch (L) {
    case 0: boys_eval<0>(T, tab, F); break;
    case 1: boys_eval<1>(T, tab, F); break;
    case 2: boys_eval<2>(T, tab, F); break;
    case 3: boys_eval<3>(T, tab, F); break;
    case 4: boys_eval<4>(T, tab, F); break;
    case 5: boys_eval<5>(T, tab, F); break;
    case 6: boys_eval<6>(T, tab, F); break;
    case 7: boys_eval<7>(T, tab, F); break;
    default: boys_eval<8>(T, tab, F); break;
  }
}

__device__ __forceinline__ double pick3(int d, double x, double y, double z) { return d == 0 ? x : (d == 1 ? y : z); }

template <bool kJK>
__device__ __forceinline__ void coop_body(const QuartetArgs& a, const CoopDesc& d, const int* __restrict__ tab,
                                          double* S) {
  const int lane = threadIdx.x;
  const int NA = ncart(d.la), NB = ncart(d.lb), NC = ncart(d.lc), ND = ncart(d.ld);
  const int NAB = NA * NB, NCD = NC * ND;
  const int L = d.la + d.lb + d.lc + d.ld;
  const int* base = tab + d.base;
  const int* voff = tab + d.voff;
  const int* vst = tab + d.vsteps;
  const int* targ = tab + d.targ;
  const int* hoff = tab + d.hoff;
  const int* hst = tab + d.hsteps;
  const int* outs = tab + d.out;
  const int n = a.nbf;
  const size_t nn = (size_t)n * n;
  const int ns = a.nspin;
  const int dstride = ns > 0 ? ns : 1;
  const double cfac = ns == 1 ? 2.0 : 1.0;
  const double xfac = ns == 1 ? 2.0 : 1.0;
  constexpr int kMaxW = 8;
  double esum[kMaxW];
#pragma unroll
  for (int w = 0; w < kMaxW; ++w) esum[w] = 0.0;
  unsigned long long nq = 0, nprimq = 0, nskip = 0;
  stamp_wave(a.stamp, false);

  // Units are claimed a.ntk (1..1024) at a time and tested 64 at a time: each
  // lane applies the Schwarz test to one of them, and the wavefront then walks
  // the surviving quartets one by one (large, mostly screened classes claim
  // many, small ones 1).  A claim of up to 64 is a run of consecutive units; a
  // larger one is claim number c of nclaims and takes units c, c + nclaims,
  // c + 2 nclaims, ..., so that the dense corners of the (bra, ket) square
  // (the heaviest, closest pairs sort first) spread over all claims.
  const int claim = a.ntk;
  const bool strided = claim > 64;
  const int nclaims = strided ? (a.ntiles + claim - 1) / claim : 0;
  for (;;) {
    int uc = 0;
    if (lane == 0) uc = atomicAdd(a.tile_ctr, strided ? 1 : claim);
    uc = __builtin_amdgcn_readfirstlane(__shfl(uc, 0, 64));
    if (uc >= (strided ? nclaims : a.ntiles)) break;
    for (int c0 = 0; c0 < claim; c0 += 64) {
      const long long stride = strided ? nclaims : 1;
      const long long uml = (long long)uc + (long long)(c0 + lane) * stride;
      const int um = uml < a.ntiles ? (int)uml : a.ntiles;
      bool keep = false;
      if (c0 + lane < claim && um < a.ntiles) {
        const int b0 = um / a.nket, k0 = um - b0 * a.nket;
        if (!a.same || k0 <= b0) {
          keep = a.pairs[a.bra0 + b0].Q * a.pairs[a.ket0 + k0].Q >= a.tau;
          if (!keep) ++nskip;
        }
      }
      unsigned long long live = __ballot(keep);
      while (live) {
      const int bit = __builtin_ctzll(live);
      live &= live - 1;
      const int u = (int)((long long)uc + (long long)(c0 + bit) * stride);
      const int ib = u / a.nket;
      const int ik = u - ib * a.nket;
      const Pair P = a.pairs[a.bra0 + ib];
      const Pair K = a.pairs[a.ket0 + ik];
      __syncthreads();  // the previous unit is done with the LDS
      for (int s = lane; s < d.nT; s += 64) S[s] = 0.0;
      const UniformPrimPairs bp = as_uniform(a.pp + P.pp0);
      const UniformPrimPairs kp = as_uniform(a.pp + K.pp0);
      for (int ipb = 0; ipb < P.npp; ++ipb) {
        const PrimPair B = load_uniform(bp + ipb);
        for (int ipk = 0; ipk < K.npp; ++ipk) {
          const PrimPair Q = load_uniform(kp + ipk);
          const double r = rsqrt_fp64(B.p + Q.p), rs = r * r;
          const double PQx = B.Px - Q.Px, PQy = B.Py - Q.Py, PQz = B.Pz - Q.Pz;
          const double T = B.p * Q.p * rs * (PQx * PQx + PQy * PQy + PQz * PQz);
          const double pref = B.K * Q.K * r;
          double F[9];
          boys_any(L, T, a.boys, F);
          const double WPx = -Q.p * rs * PQx, WPy = -Q.p * rs * PQy, WPz = -Q.p * rs * PQz;
          const double WQx = B.p * rs * PQx, WQy = B.p * rs * PQy, WQz = B.p * rs * PQz;
          const double rop = Q.p * rs, roq = B.p * rs, oo2s = 0.5 * rs;
          __syncthreads();  // the previous primitive quartet's VRR slots are consumed
          for (int i = lane; i < d.nbase; i += 64) S[base[2 * i]] = pref * F[base[2 * i + 1]];
          __syncthreads();
          for (int l = 0; l < d.nvl; ++l) {
            for (int i = voff[l] + lane; i < voff[l + 1]; i += 64) {
              const int4 s0 = *reinterpret_cast<const int4*>(vst + 8 * i);
              const int4 s1 = *reinterpret_cast<const int4*>(vst + 8 * i + 4);
              const int code = s1.z;
              const int kind = code & 3, dir = (code >> 2) & 3, n1 = (code >> 4) & 15, n2 = (code >> 8) & 15;
              double v;
              if (kind == 1) {
                v = pick3(dir, B.PAx, B.PAy, B.PAz) * S[s0.y] + pick3(dir, WPx, WPy, WPz) * S[s0.z];
                if (n1) v += (double)n1 * B.oo2p * (S[s0.w] - rop * S[s1.x]);
              } else {
                v = pick3(dir, Q.PAx, Q.PAy, Q.PAz) * S[s0.y] + pick3(dir, WQx, WQy, WQz) * S[s0.z];
                if (n1) v += (double)n1 * Q.oo2p * (S[s0.w] - roq * S[s1.x]);
                if (n2) v += (double)n2 * oo2s * S[s1.y];
              }
              S[s0.x] = v;
            }
            __syncthreads();
          }
          for (int i = lane; i < d.nT; i += 64) S[targ[2 * i]] += S[targ[2 * i + 1]];
        }
      }
      __syncthreads();
      for (int l = 0; l < d.nhl; ++l) {
        for (int i = hoff[l] + lane; i < hoff[l + 1]; i += 64) {
          const int4 h = *reinterpret_cast<const int4*>(hst + 4 * i);
          const int dir = h.w & 3;
          const double cf = (h.w & 4) ? pick3(dir, K.ABx, K.ABy, K.ABz) : pick3(dir, P.ABx, P.ABy, P.ABz);
          S[h.x] = S[h.y] + cf * S[h.z];
        }
        __syncthreads();
      }
      ++nq;
      nprimq += (unsigned long long)P.npp * K.npp;
      const double deg = (P.offdiag ? 2.0 : 1.0) * (K.offdiag ? 2.0 : 1.0) * ((ib != ik || !a.same) ? 2.0 : 1.0);
      const int ma = a.shells[P.A].ao, mb = a.shells[P.B].ao;
      const int mc = a.shells[K.A].ao, md = a.shells[K.B].ao;
      for (int w = 0; w < a.nwfn; ++w) {
        const double* D0 = a.dens + (size_t)w * dstride * nn;
        auto rho = [&](int x, int y) {  // cfac * sum_s D_s[x][y]
          double t = D0[(size_t)x * n + y];
          if (ns == 2) t += D0[nn + (size_t)x * n + y];
          return cfac * t;
        };
        if constexpr (!kJK) {
          double e = 0.0;
          for (int idx = lane; idx < NAB * NCD; idx += 64) {
            const int ab = idx / NCD, cd = idx - ab * NCD;
            const int i = ab / NB, j = ab - i * NB, k = cd / ND, l = cd - k * ND;
            const double v = S[outs[idx]];
            double ex = 0.0;
            for (int s = 0; s < ns; ++s) {
              const double* Ds = D0 + s * nn;
              ex += Ds[(size_t)(ma + i) * n + mc + k] * Ds[(size_t)(mb + j) * n + md + l] +
                    Ds[(size_t)(ma + i) * n + md + l] * Ds[(size_t)(mb + j) * n + mc + k];
            }
            e += v * (0.5 * rho(ma + i, mb + j) * rho(mc + k, md + l) - 0.25 * xfac * ex);
          }
#pragma unroll
          for (int ww = 0; ww < kMaxW; ++ww)
            if (ww == w) esum[ww] += deg * e;
        } else {
          // The contracted block V (contiguous, [ab][cd]) and the density blocks
          // the contraction reads go to LDS first: the loops below then read LDS
          // only, instead of a step-table index and a global density element per
          // term.  V and the blocks follow the kernel's VRR/HRR slots (coop_lds_bytes).
          double* V = S + d.lds;
          double* Rab = V + NAB * NCD;  // cfac * sum_s D_s[a][b], then [c][d]
          double* Rcd = Rab + NAB;
          double* Dbd = Rcd + NCD;      // one spin's D[b][d], D[a][d], D[b][c], D[a][c]
          double* Dad = Dbd + NB * ND;
          double* Dbc = Dad + NA * ND;
          double* Dac = Dbc + NB * NC;
          if (w == 0)
            for (int idx = lane; idx < NAB * NCD; idx += 64) V[idx] = S[outs[idx]];
          __syncthreads();  // V is written; the previous density's blocks are consumed
          for (int x = lane; x < NAB + NCD; x += 64) {
            if (x < NAB)
              Rab[x] = rho(ma + x / NB, mb + x % NB);
            else
              Rcd[x - NAB] = rho(mc + (x - NAB) / ND, md + (x - NAB) % ND);
          }
          __syncthreads();
          const double fj = 0.25 * deg, fk = 0.125 * deg * xfac;
          double* J = a.jacc + (size_t)w * nn;
          for (int ab = lane; ab < NAB; ab += 64) {
            const int i = ab / NB, j = ab - i * NB;
            double t = 0.0;
            for (int cd = 0; cd < NCD; ++cd) t += V[ab * NCD + cd] * Rcd[cd];
            unsafeAtomicAdd(&J[(size_t)(ma + i) * n + mb + j], fj * t);
          }
          for (int cd = lane; cd < NCD; cd += 64) {
            const int k = cd / ND, l = cd - k * ND;
            double t = 0.0;
            for (int ab = 0; ab < NAB; ++ab) t += V[ab * NCD + cd] * Rab[ab];
            unsafeAtomicAdd(&J[(size_t)(mc + k) * n + md + l], fj * t);
          }
          for (int s = 0; s < ns; ++s) {
            const double* Ds = D0 + s * nn;
            double* Kx = a.kacc + ((size_t)w * ns + s) * nn;
            if (s > 0) __syncthreads();  // the previous spin's blocks are consumed
            for (int x = lane; x < NB * ND + NA * ND + NB * NC + NA * NC; x += 64) {
              int y = x;
              if (y < NB * ND) {
                Dbd[y] = Ds[(size_t)(mb + y / ND) * n + md + y % ND];
                continue;
              }
              y -= NB * ND;
              if (y < NA * ND) {
                Dad[y] = Ds[(size_t)(ma + y / ND) * n + md + y % ND];
                continue;
              }
              y -= NA * ND;
              if (y < NB * NC) {
                Dbc[y] = Ds[(size_t)(mb + y / NC) * n + mc + y % NC];
                continue;
              }
              y -= NB * NC;
              Dac[y] = Ds[(size_t)(ma + y / NC) * n + mc + y % NC];
            }
            __syncthreads();
            // B[a][c] += v D[b][d]
            for (int x = lane; x < NA * NC; x += 64) {
              const int i = x / NC, k = x - i * NC;
              double t = 0.0;
              for (int j = 0; j < NB; ++j)
                for (int l = 0; l < ND; ++l) t += V[(i * NB + j) * NCD + k * ND + l] * Dbd[j * ND + l];
              unsafeAtomicAdd(&Kx[(size_t)(ma + i) * n + mc + k], fk * t);
            }
            // B[b][c] += v D[a][d]
            for (int x = lane; x < NB * NC; x += 64) {
              const int j = x / NC, k = x - j * NC;
              double t = 0.0;
              for (int i = 0; i < NA; ++i)
                for (int l = 0; l < ND; ++l) t += V[(i * NB + j) * NCD + k * ND + l] * Dad[i * ND + l];
              unsafeAtomicAdd(&Kx[(size_t)(mb + j) * n + mc + k], fk * t);
            }
            // B[a][d] += v D[b][c]
            for (int x = lane; x < NA * ND; x += 64) {
              const int i = x / ND, l = x - i * ND;
              double t = 0.0;
              for (int j = 0; j < NB; ++j)
                for (int k = 0; k < NC; ++k) t += V[(i * NB + j) * NCD + k * ND + l] * Dbc[j * NC + k];
              unsafeAtomicAdd(&Kx[(size_t)(ma + i) * n + md + l], fk * t);
            }
            // B[b][d] += v D[a][c]
            for (int x = lane; x < NB * ND; x += 64) {
              const int j = x / ND, l = x - j * ND;
              double t = 0.0;
              for (int i = 0; i < NA; ++i)
                for (int k = 0; k < NC; ++k) t += V[(i * NB + j) * NCD + k * ND + l] * Dac[i * NC + k];
              unsafeAtomicAdd(&Kx[(size_t)(mb + j) * n + md + l], fk * t);
            }
          }
        }
      }
      }  // surviving units of the batch of 64
    }  // batches of the claim
  }
  if constexpr (!kJK) {
    // one wavefront per workgroup: slot 0 of the block's four wave slots holds the sum
    const int nw = a.nwfn < kMaxW ? a.nwfn : kMaxW;
    for (int w = 0; w < nw; ++w) {
      double s = 0.0;
#pragma unroll
      for (int ww = 0; ww < kMaxW; ++ww)
        if (ww == w) s = esum[ww];
      s = wave_sum(s);
      if (lane == 0)
        for (int k = 0; k < 4; ++k) a.partial[((size_t)blockIdx.x * 4 + k) * a.nwfn + w] = k == 0 ? s : 0.0;
    }
  }
  stamp_wave(a.stamp, true);
  if (a.counters) {
    unsigned long long sk = nskip;  // per lane; nq and nprimq are wave-uniform
    for (int o = 32; o > 0; o >>= 1) sk += __shfl_xor(sk, o, 64);
    if (lane == 0) {
      atomicAdd(&a.counters[0], nq);
      atomicAdd(&a.counters[1], nprimq);
      atomicAdd(&a.counters[2], sk);
    }
  }
}

template <bool kJK>
__global__ void __launch_bounds__(64) coop_kernel(QuartetArgs a, CoopDesc d, const int* __restrict__ tab) {
  extern __shared__ double S[];
  coop_body<kJK>(a, d, tab, S);
}

// J/K mode over several groups (molecules) of a batch in one launch: the
// persistent wavefronts drain group 0's units, then group 1's, and so on (the
// same scheme as the straight-line kernels' multi-group launch).
__global__ void __launch_bounds__(64) coop_groups_kernel(const QuartetArgs* __restrict__ gs, int ng, CoopDesc d,
                                                         const int* __restrict__ tab, unsigned long long* stamp) {
  extern __shared__ double S[];
  stamp_wave(stamp, false);
  for (int g = 0; g < ng; ++g) {
    const QuartetArgs a = gs[g];
    if (a.ntiles > 0) coop_body<true>(a, d, tab, S);
    __syncthreads();  // the group's last unit is done with the LDS
  }
  stamp_wave(stamp, true);
}

std::mutex g_tab_mu;
int* g_tab[64] = {};

}  // namespace

int coop_class(int la, int lb, int lc, int ld) {
  for (int i = 0; i < kNCoop; ++i)
    if (kCoopDesc[i].la == la && kCoopDesc[i].lb == lb && kCoopDesc[i].lc == lc && kCoopDesc[i].ld == ld) return i;
  return -1;
}

// LDS of one workgroup: the VRR/HRR slots and, in J/K mode, the contiguous
// contracted block, the two Coulomb density blocks and one spin's four
// exchange density blocks.
size_t coop_lds_bytes(int idx, bool jk) {
  const CoopDesc& d = kCoopDesc[idx];
  size_t n = (size_t)d.lds;
  if (jk) {
    const size_t NA = ncart(d.la), NB = ncart(d.lb), NC = ncart(d.lc), ND = ncart(d.ld);
    n += NA * NB * NC * ND + NA * NB + NC * ND + NB * ND + NA * ND + NB * NC + NA * NC;
  }
  return n * sizeof(double);
}

bool coop_default(int idx) { return idx >= 0 && kCoopDesc[idx].dflt != 0; }

int coop_occupancy(int idx, bool jk) {
  int nb = 0;
  const size_t lds = coop_lds_bytes(idx, jk);
  if (jk)
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, coop_kernel<true>, 64, lds);
  else
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, coop_kernel<false>, 64, lds);
  return nb > 0 ? nb : 1;
}

// The step tables on a device: a device array initialized in the code object,
// whose address is looked up once per device (with that device current).
static hipError_t coop_tables(int device, const int** out) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  if (!g_tab[device]) {
    int* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(kCoopData));
    if (e != hipSuccess) return e;
    g_tab[device] = p;
  }
  *out = g_tab[device];
  return hipSuccess;
}

hipError_t coop_prepare(int device) {
  const int* t = nullptr;
  return coop_tables(device, &t);
}

hipError_t launch_coop(int idx, const QuartetArgs& a, bool jk, int nblocks, int device, hipStream_t st) {
  if (nblocks <= 0 || a.ntiles <= 0) return hipSuccess;
  const int* t = nullptr;
  hipError_t e = coop_tables(device, &t);
  if (e != hipSuccess) return e;
  const size_t lds = coop_lds_bytes(idx, jk);
  if (jk)
    hipLaunchKernelGGL(coop_kernel<true>, dim3(nblocks), dim3(64), lds, st, a, kCoopDesc[idx], t);
  else
    hipLaunchKernelGGL(coop_kernel<false>, dim3(nblocks), dim3(64), lds, st, a, kCoopDesc[idx], t);
  return hipGetLastError();
}

hipError_t launch_coop_groups(int idx, const QuartetArgs* gs, int ng, int nblocks, int device,
                              unsigned long long* stamp, hipStream_t st) {
  if (nblocks <= 0 || ng <= 0) return hipSuccess;
  const int* t = nullptr;
  hipError_t e = coop_tables(device, &t);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(coop_groups_kernel, dim3(nblocks), dim3(64), coop_lds_bytes(idx, true), st, gs, ng, kCoopDesc[idx], t,
                     stamp);
  return hipGetLastError();
}

}  // namespace vsvb
