// Spin-coupled VSVB energy and adjoint gradient (npair > 0) on MI355X.
//
// The reference sums, for every orbital quartet, second-order cofactors of
// the spin-orbital overlap over all pairs of determinants of the spin-coupled
// expansion (density_sc, dbra, dket, det: src/valence.F90:1612-2056).  Here
// the same sum is regrouped by determinant pair (S, T):
//
//   A_ST[r][c] = <alpha_S(r)|alpha_T(c)>,  B_ST likewise for beta, as
//   build_abket orders them (src/valence.F90:2524-2586);
//   G_a = A^-1, G_b = B^-1, omega = c_S c_T det A det B;
//   Gamma^s[p][q] = G_s[c][r] at p = orbital of bra row r, q = of ket column c
//   (Jacobi: a first-order cofactor is det * an inverse element, a second-order
//   one the 2x2 determinant of inverse elements);
//   Phi = sum h Gamma^t + 1/2 sum (pq|rs) Gamma^t_pq Gamma^t_rs
//         - 1/2 sum_s sum (pq|rs) Gamma^s_ps Gamma^s_rq;
//   E = sum omega Phi / sum omega + E_nuc.
//
// The orbital integrals (pq|rs) come from one Coulomb-only pass of the
// shell-quartet kernels over the n(n+1)/2 orbital-pair densities
// D^rs = (C_r C_s^T + C_s C_r^T)/2, which gives J^rs = (..|rs) in the AO
// basis; W^rs = J^rs C^T is kept and reused by the gradient.  The Coulomb and
// exchange contractions of every determinant pair are two dense products
// [pairs x n^2] x [n^2 x n^2].
//
// Adjoint (one reverse sweep over the same determinant pairs and the kept
// W^rs; no integral is stored or recomputed twice):
//   dE/dGamma^s = (omega/N) (h + J(Gamma^t) - K(Gamma^s)),
//   dE/dA = ((Phi - E_el)/N) omega G^T - G^T Gbar G^T   -> dE/dS_orb,
//   dE/dh_orb = sum (omega/N) Gamma^t,  dE/d(pq|rs) = D2 = sum (omega/N)
//   [Gamma^t (x) Gamma^t / 2 - sum_s Gamma^s (x) Gamma^s (exchange order) / 2],
//   dE/dC = (Shat + Shat^T) C S_ao + (hhat + hhat^T) C h_ao + 4-index term
//   through W^rs.
#include "vsvb_internal.h"

using namespace vsvb;

namespace {

constexpr int kMaxSpinDim = 48;  // largest alpha / beta block (three n x n LDS arrays per workgroup)

// D[rs][mu][nu] = (C_r[mu] C_s[nu] + C_s[mu] C_r[nu]) / 2 for the unique orbital pairs.
__global__ void k_pair_dens(const double* __restrict__ Cs, const int* __restrict__ pq, int npq, int nbf,
                            double* __restrict__ D) {
  const size_t nn = (size_t)nbf * nbf;
  const size_t tot = (size_t)npq * nn;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const int w = (int)(t / nn);
    const int mu = (int)((t % nn) / nbf), nu = (int)(t % nbf);
    const int r = pq[2 * w], s = pq[2 * w + 1];
    const double* cr = Cs + (size_t)r * nbf;
    const double* cs = Cs + (size_t)s * nbf;
    D[t] = 0.5 * (cr[mu] * cs[nu] + cs[mu] * cr[nu]);
  }
}

// J = A + A^T for each of nmat n x n accumulators, in place.
__global__ void k_sym_batch(double* __restrict__ A, int nmat, int n) {
  const size_t nn = (size_t)n * n;
  const size_t tot = (size_t)nmat * nn;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const size_t w = t / nn;
    const int i = (int)((t % nn) / n), j = (int)(t % n);
    if (j > i) continue;
    double* M = A + w * nn;
    const double v = M[(size_t)i * n + j] + M[(size_t)j * n + i];
    M[(size_t)i * n + j] = v;
    M[(size_t)j * n + i] = v;
  }
}

// (pq|rs) = sum_mu C[p][mu] W[rs][mu][q], stored for both (r,s) and (s,r):
// eri[(p n + q) n^2 + r n + s].
__global__ void k_orb_eri(const double* __restrict__ Cs, const double* __restrict__ W, const int* __restrict__ pq,
                          int npq, int ns, int nbf, double* __restrict__ eri) {
  const size_t n2 = (size_t)ns * ns;
  const size_t tot = (size_t)npq * n2;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const int w = (int)(t / n2);
    const int p = (int)((t % n2) / ns), q = (int)(t % ns);
    const double* Wr = W + (size_t)w * nbf * ns;
    const double* cp = Cs + (size_t)p * nbf;
    double v = 0.0;
    for (int mu = 0; mu < nbf; ++mu) v += cp[mu] * Wr[(size_t)mu * ns + q];
    const int r = pq[2 * w], s = pq[2 * w + 1];
    eri[((size_t)p * ns + q) * n2 + (size_t)r * ns + s] = v;
    eri[((size_t)p * ns + q) * n2 + (size_t)s * ns + r] = v;
  }
}

// Exchange ordering: erik[(a n + b) n^2 + c n + d] = (cb|ad).
__global__ void k_orb_erik(const double* __restrict__ eri, int ns, double* __restrict__ erik) {
  const size_t n2 = (size_t)ns * ns;
  const size_t tot = n2 * n2;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const int ab = (int)(t / n2), cd = (int)(t % n2);
    const int a = ab / ns, b = ab % ns, c = cd / ns, d = cd % ns;
    erik[t] = eri[((size_t)c * ns + b) * n2 + (size_t)a * ns + d];
  }
}

__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o, 64);
    const int oi = __shfl_xor(idx, o, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

struct DetArgs {
  const double* Sorb;
  const int* alist;    // [ndet][na]
  const int* blist;    // [ndet][nb]
  const double* cdet;  // [ndet]
  int ndet, na, nb, ns;
  double* Y;           // [pairs][2 spins][3][ns^2]: Gamma_R, g1, g2 in orbital space
  double* q4;          // [pairs][2 spins][4]: s, P, d1, d2
  int* status;         // status[3]: a block with a pivot ratio below 1e-9 (gradient refused)
};

// Cofactors of a possibly singular block without an inverse.  Complete-
// pivoting LU, P A Q = L D U~ (L, U~ unit triangular), so that
//   A^-1 = Q U~^-1 D^-1 L^-1 P   wherever it exists.
// With z1 = n-1, z2 = n-2 (the two smallest pivots under complete pivoting),
// R = the other pivots, P_R = prod_R d_k, s = sign of P and Q:
//   Gamma_R = sum_{k in R} U~^-1[:,k] L^-1[k,:] / d_k,  g_z = U~^-1[:,z] L^-1[z,:],
//   det = s d1 d2 P_R,  first-order cofactors = s P_R (d1 d2 Gamma_R + d2 g1 + d1 g2),
//   second-order cofactors = s P_R [d1 d2 Gamma_R(x)Gamma_R + d2 (g1(x)Gamma_R + Gamma_R(x)g1)
//                                   + d1 (g2(x)Gamma_R + Gamma_R(x)g2) + g1(x)g2 + g2(x)g1]
// (antisymmetrized pairs; the self products of the rank-one g_z cancel).  These
// are the cofactors of the diagonal factor carried through X = P^T L and
// Y^T = U~ Q^T, polynomial in the pivots, so they hold for d1 = 0 or d2 = 0;
// with a zero pivot in R every cofactor up to second order vanishes.
// One workgroup (one wavefront) per ordered determinant pair t = S * ndet + T.
__global__ void __launch_bounds__(64) k_det_pairs(DetArgs a, long p0) {
  extern __shared__ double lds[];
  __shared__ int prow[kMaxSpinDim], pcol[kMaxSpinDim];
  __shared__ double piv[kMaxSpinDim];
  const long t = p0 + blockIdx.x;
  const int S = (int)(t / a.ndet), T = (int)(t % a.ndet);
  const int ns = a.ns, n2 = ns * ns;
  const int dmax = a.na > a.nb ? a.na : a.nb;
  const int tid = threadIdx.x;
  double* M = lds;                     // n x n, factored in place
  double* Ui = lds + dmax * dmax;      // U~^-1
  double* Li = lds + 2 * dmax * dmax;  // L^-1
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin == 0 ? a.na : a.nb;
    double* Yr = a.Y + ((size_t)t * 2 + spin) * 3 * n2;
    double* Y1 = Yr + n2;
    double* Y2 = Y1 + n2;
    double* q = a.q4 + ((size_t)t * 2 + spin) * 4;
    for (int k = tid; k < 3 * n2; k += 64) Yr[k] = 0.0;
    if (n == 0) {
      if (tid == 0) {
        q[0] = 1.0; q[1] = 1.0; q[2] = 1.0; q[3] = 1.0;
      }
      continue;
    }
    const int* lb = (spin == 0 ? a.alist : a.blist) + (size_t)S * n;
    const int* lk = (spin == 0 ? a.alist : a.blist) + (size_t)T * n;
    __syncthreads();  // the previous spin block is done with the LDS
    for (int k = tid; k < n * n; k += 64) M[k] = a.Sorb[(size_t)lb[k / n] * ns + lk[k % n]];
    for (int k = tid; k < n; k += 64) {
      prow[k] = k;
      pcol[k] = k;
    }
    __syncthreads();
    double sign = 1.0;
    for (int k = 0; k < n; ++k) {
      const int m = n - k;
      double bv = -1.0;
      int bi = k * n + k;
      for (int e = tid; e < m * m; e += 64) {
        const int i = k + e / m, j = k + e % m;
        const double v = fabs(M[i * n + j]);
        if (v > bv || (v == bv && i * n + j < bi)) {
          bv = v;
          bi = i * n + j;
        }
      }
      wave_argmax(bv, bi);
      const int pi = bi / n, pj = bi % n;
      if (pi != k) {
        for (int j = tid; j < n; j += 64) {
          const double x = M[k * n + j];
          M[k * n + j] = M[pi * n + j];
          M[pi * n + j] = x;
        }
        sign = -sign;
      }
      __syncthreads();
      if (pj != k) {
        for (int i = tid; i < n; i += 64) {
          const double x = M[i * n + k];
          M[i * n + k] = M[i * n + pj];
          M[i * n + pj] = x;
        }
        sign = -sign;
      }
      if (tid == 0) {
        int x = prow[k]; prow[k] = prow[pi]; prow[pi] = x;
        x = pcol[k]; pcol[k] = pcol[pj]; pcol[pj] = x;
      }
      __syncthreads();
      const double d = M[k * n + k];
      if (tid == 0) piv[k] = d;
      if (d != 0.0) {  // d == 0: the trailing block is exactly zero, nothing to eliminate
        for (int i = k + 1 + tid; i < n; i += 64) M[i * n + k] /= d;
        __syncthreads();
        const int r = m - 1;
        for (int e = tid; e < r * r; e += 64) {
          const int i = k + 1 + e / r, j = k + 1 + e % r;
          M[i * n + j] -= M[i * n + k] * M[k * n + j];
        }
      }
      __syncthreads();
    }
    // unit-triangular inverses, one column per thread (n <= 64)
    if (tid < n) {
      const int j = tid;
      for (int i = 0; i < n; ++i) Ui[i * n + j] = i == j ? 1.0 : 0.0;
      for (int i = j - 1; i >= 0; --i) {
        const double di = piv[i];
        double x = 0.0;
        if (di != 0.0)
          for (int mm = i + 1; mm <= j; ++mm) x -= M[i * n + mm] / di * Ui[mm * n + j];
        Ui[i * n + j] = x;
      }
      for (int i = 0; i < n; ++i) Li[i * n + j] = i == j ? 1.0 : 0.0;
      for (int i = j + 1; i < n; ++i) {
        double x = 0.0;
        for (int mm = j; mm < i; ++mm) x -= M[i * n + mm] * Li[mm * n + j];
        Li[i * n + j] = x;
      }
    }
    __syncthreads();
    const int z1 = n - 1, z2 = n - 2;
    const double d1 = piv[z1], d2 = n >= 2 ? piv[z2] : 1.0;
    double P = 1.0;
    for (int k = 0; k < n - 2; ++k) P *= piv[k];
    const double pmax = fabs(piv[0]);
    if (tid == 0) {
      q[0] = sign; q[1] = P; q[2] = d1; q[3] = d2;
      if (!(fabs(d1) > 1e-9 * pmax) || !(fabs(d2) > 1e-9 * pmax)) atomicOr(a.status + 3, 1);
    }
    // Gamma_R, g1, g2 at (bra orbital of row prow[i], ket orbital of column pcol[j])
    for (int e = tid; e < n * n; e += 64) {
      const int j = e / n, i = e % n;
      double gr = 0.0;
      if (P != 0.0)
        for (int k = 0; k < n - 2; ++k) gr += Ui[j * n + k] * Li[k * n + i] / piv[k];
      const size_t pos = (size_t)lb[prow[i]] * ns + lk[pcol[j]];
      Yr[pos] = gr;
      Y1[pos] = Ui[j * n + z1] * Li[z1 * n + i];
      if (n >= 2) Y2[pos] = Ui[j * n + z2] * Li[z2 * n + i];
    }
  }
}

__device__ __forceinline__ double block_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per determinant pair, from Y = (Gamma_R, g1, g2) per spin and their Coulomb
// (JY) and exchange (KY) images:
//   f = c_S c_T [ sum_s h.C1_s det_s' + 1/2 sum_s T_s det_s' + C1_a . J(C1_b) ],  omega = c_S c_T det_a det_b
// with C1_s, T_s, det_s the split cofactors above (one wavefront per pair).
// With ham != nullptr the unweighted f / (c_S c_T) and omega / (c_S c_T) are
// also added into the spin-coupling matrices at (I, J) = (S >> npair, T >> npair)
// for I >= J (bra coupling I, ket coupling J).
__global__ void __launch_bounds__(64) k_pair_energy(const double* __restrict__ h, const double* __restrict__ Y,
                                                    const double* __restrict__ JY, const double* __restrict__ KY,
                                                    const double* __restrict__ q4, const double* __restrict__ cdet,
                                                    int ndet, int n2, long p0, double* __restrict__ phi,
                                                    double* __restrict__ num, double* __restrict__ wout,
                                                    double* __restrict__ e1w, int npair, int nspinc,
                                                    double* __restrict__ ham, double* __restrict__ ovl) {
  const long t = p0 + blockIdx.x;
  const size_t o = (size_t)t * 6 * n2;
  // sums: hY[6], B[2][4] = (rr, 1r, 2r, 12), X[3][3] = Y_a(alpha) . JY_b(beta)
  double acc[23];
#pragma unroll
  for (int k = 0; k < 23; ++k) acc[k] = 0.0;
  for (int k = threadIdx.x; k < n2; k += 64) {
    double y[6], jy[6], ky[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      y[m] = Y[o + (size_t)m * n2 + k];
      jy[m] = JY[o + (size_t)m * n2 + k];
      ky[m] = KY[o + (size_t)m * n2 + k];
    }
    const double hk = h[k];
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] += hk * y[m];
#pragma unroll
    for (int sp = 0; sp < 2; ++sp) {
      const int b = 3 * sp;
      const double gr = jy[b] - ky[b];
      acc[6 + 4 * sp + 0] += y[b] * gr;
      acc[6 + 4 * sp + 1] += y[b + 1] * gr;
      acc[6 + 4 * sp + 2] += y[b + 2] * gr;
      acc[6 + 4 * sp + 3] += y[b + 1] * (jy[b + 2] - ky[b + 2]);
    }
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) acc[14 + 3 * aa + bb] += y[aa] * jy[3 + bb];
  }
#pragma unroll
  for (int k = 0; k < 23; ++k) acc[k] = block_sum64(acc[k]);
  if (threadIdx.x == 0) {
    const int S = (int)(t / ndet), T = (int)(t % ndet);
    const double cst = cdet[S] * cdet[T];
    double det[2], e1[2], tt[2], cf[2][3];
    for (int sp = 0; sp < 2; ++sp) {
      const double* q = q4 + ((size_t)t * 2 + sp) * 4;
      const double sP = q[0] * q[1], d1 = q[2], d2 = q[3];
      det[sp] = sP * d1 * d2;
      cf[sp][0] = sP * d1 * d2;
      cf[sp][1] = sP * d2;
      cf[sp][2] = sP * d1;
      e1[sp] = cf[sp][0] * acc[3 * sp] + cf[sp][1] * acc[3 * sp + 1] + cf[sp][2] * acc[3 * sp + 2];
      const double* B = acc + 6 + 4 * sp;
      tt[sp] = sP * (d1 * d2 * B[0] + 2.0 * d2 * B[1] + 2.0 * d1 * B[2] + 2.0 * B[3]);
    }
    double cross = 0.0;
    for (int aa = 0; aa < 3; ++aa)
      for (int bb = 0; bb < 3; ++bb) cross += cf[0][aa] * cf[1][bb] * acc[14 + 3 * aa + bb];
    const double one = e1[0] * det[1] + e1[1] * det[0];
    const double fu = one + 0.5 * (tt[0] * det[1] + tt[1] * det[0]) + cross, omu = det[0] * det[1];
    const double f = cst * fu;
    const double om = cst * omu;
    num[t] = f;
    wout[t] = om;
    e1w[t] = cst * one;
    phi[t] = om != 0.0 ? f / om : 0.0;
    if (ham) {
      const int I = S >> npair, J = T >> npair;
      if (I >= J) {
        unsafeAtomicAdd(&ham[I * nspinc + J], fu);
        unsafeAtomicAdd(&ovl[I * nspinc + J], omu);
      }
    }
  }
}

// Regular blocks (the gradient path): Gamma_s = Gamma_R + g1/d1 + g2/d2 is the
// full inverse, with its Coulomb and exchange images by linearity; slot-space
// inverses for the adjoint.
__global__ void __launch_bounds__(64) k_pair_regular(const double* __restrict__ Y, const double* __restrict__ JY,
                                                     const double* __restrict__ KY, const double* __restrict__ q4,
                                                     const int* __restrict__ alist, const int* __restrict__ blist,
                                                     int ndet, int na, int nb, int ns, long p0,
                                                     double* __restrict__ Xa, double* __restrict__ Xb,
                                                     double* __restrict__ Xt, double* __restrict__ JX,
                                                     double* __restrict__ KXa, double* __restrict__ KXb,
                                                     double* __restrict__ Ga, double* __restrict__ Gb) {
  const long t = p0 + blockIdx.x;
  const int n2 = ns * ns;
  const size_t o = (size_t)t * 6 * n2, ot = (size_t)t * n2;
  double w[2][3];
  for (int sp = 0; sp < 2; ++sp) {
    const double* q = q4 + ((size_t)t * 2 + sp) * 4;
    w[sp][0] = 1.0;
    w[sp][1] = 1.0 / q[2];
    w[sp][2] = 1.0 / q[3];
  }
  for (int k = threadIdx.x; k < n2; k += 64) {
    double x[2], jx = 0.0, kx[2];
    for (int sp = 0; sp < 2; ++sp) {
      x[sp] = 0.0;
      kx[sp] = 0.0;
      for (int m = 0; m < 3; ++m) {
        const size_t idx = o + (size_t)(3 * sp + m) * n2 + k;
        x[sp] += w[sp][m] * Y[idx];
        jx += w[sp][m] * JY[idx];
        kx[sp] += w[sp][m] * KY[idx];
      }
    }
    Xa[ot + k] = x[0];
    Xb[ot + k] = x[1];
    Xt[ot + k] = x[0] + x[1];
    JX[ot + k] = jx;
    KXa[ot + k] = kx[0];
    KXb[ot + k] = kx[1];
  }
  __syncthreads();
  const int S = (int)(t / ndet), T = (int)(t % ndet);
  for (int sp = 0; sp < 2; ++sp) {
    const int n = sp == 0 ? na : nb;
    const int* lb = (sp == 0 ? alist : blist) + (size_t)S * n;
    const int* lk = (sp == 0 ? alist : blist) + (size_t)T * n;
    const double* X = sp == 0 ? Xa + ot : Xb + ot;
    double* G = sp == 0 ? Ga + (size_t)t * na * na : Gb + (size_t)t * nb * nb;
    for (int k = threadIdx.x; k < n * n; k += 64) {
      const int c = k / n, r = k % n;
      G[k] = X[(size_t)lb[r] * ns + lk[c]];
    }
  }
}

// scal[0] = sum omega Phi, scal[1] = N = sum omega, scal[4] = sum omega e1 (from the reductions);
// scal[2] = E, scal[3] = E_el, scal[5] = E1.
__global__ void k_sc_final(double* __restrict__ scal, double enuc) {
  const double el = scal[0] / scal[1];
  scal[3] = el;
  scal[2] = el + enuc;
  scal[5] = scal[4] / scal[1];
}

struct AdjArgs {
  const double* h;
  const double* JX;
  const double* KXa;
  const double* KXb;
  const double* Ga;
  const double* Gb;
  const double* w;
  const double* phi;
  const double* scal;
  const int* alist;
  const int* blist;
  int ndet, na, nb, ns;
  double* Shat;  // [ns^2] dE/dS_orb (atomic)
  double* wv;    // [pairs] omega / N
};

// Adjoint of one determinant pair: Abar = ((Phi - E_el)/N) omega G^T - G^T Gbar G^T,
// Gbar[c][r] = (omega/N) (h + J - K_s)[p_r][q_c]; scattered into Shat.
__global__ void __launch_bounds__(64) k_pair_adjoint(AdjArgs a, long p0) {
  extern __shared__ double lds[];
  const long t = p0 + blockIdx.x;
  const int S = (int)(t / a.ndet), T = (int)(t % a.ndet);
  const int ns = a.ns, n2 = ns * ns;
  const double N = a.scal[1], el = a.scal[3];
  const double om = a.w[t];
  const double f = om / N, g = (a.phi[t] - el) / N * om;
  if (threadIdx.x == 0) a.wv[t] = f;
  if (om == 0.0) return;
  const int dmax = a.na > a.nb ? a.na : a.nb;
  double* G = lds;                    // n x n
  double* Gb = lds + dmax * dmax;     // Gbar
  double* M1 = lds + 2 * dmax * dmax; // Gbar G^T
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin == 0 ? a.na : a.nb;
    if (n == 0) continue;
    const int* lb = (spin == 0 ? a.alist : a.blist) + (size_t)S * n;
    const int* lk = (spin == 0 ? a.alist : a.blist) + (size_t)T * n;
    const double* Gg = (spin == 0 ? a.Ga + (size_t)t * a.na * a.na : a.Gb + (size_t)t * a.nb * a.nb);
    const double* KX = (spin == 0 ? a.KXa : a.KXb) + (size_t)t * n2;
    const double* JX = a.JX + (size_t)t * n2;
    __syncthreads();
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
      const int c = k / n, r = k % n;
      const size_t pos = (size_t)lb[r] * ns + lk[c];
      G[k] = Gg[k];
      Gb[k] = f * (a.h[pos] + JX[pos] - KX[pos]);
    }
    __syncthreads();
    // M1[i][c] = sum_j Gbar[i][j] G[c][j]
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
      const int i = k / n, c = k % n;
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += Gb[i * n + j] * G[c * n + j];
      M1[k] = s;
    }
    __syncthreads();
    // Abar[r][c] = g G[c][r] - sum_i G[i][r] M1[i][c]
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
      const int r = k / n, c = k % n;
      double s = g * G[c * n + r];
      for (int i = 0; i < n; ++i) s -= G[i * n + r] * M1[i * n + c];
      atomicAdd(&a.Shat[(size_t)lb[r] * ns + lk[c]], s);
    }
  }
}

// Rows of X scaled by wv: Y[t][k] = wv[t] X[t][k].
__global__ void k_scale_rows(const double* __restrict__ X, const double* __restrict__ wv, long rows, int n2,
                             double* __restrict__ Y) {
  const size_t tot = (size_t)rows * n2;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x)
    Y[t] = wv[t / n2] * X[t];
}

// D2[(p q),(r s)] -= 1/2 Dk[(p s),(r q)]
__global__ void k_exchange_fold(const double* __restrict__ Dk, int ns, double* __restrict__ D2) {
  const size_t n2 = (size_t)ns * ns;
  const size_t tot = n2 * n2;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const int pq = (int)(t / n2), rs = (int)(t % n2);
    const int p = pq / ns, q = pq % ns, r = rs / ns, s = rs % ns;
    D2[t] -= 0.5 * Dk[((size_t)p * ns + s) * n2 + (size_t)r * ns + q];
  }
}

// Dh[t][q][w] for the unique orbital pair w = (r, s): the coefficient of
// W^rs[mu][q] in dE/dC[t][mu]; Dtil[tq][rs] = D2[tq,rs] + D2[qt,rs] + D2[rs,tq] + D2[rs,qt],
// summed over both orders of (r, s) (W^rs = W^sr).
__global__ void k_grad_coef(const double* __restrict__ D2, const int* __restrict__ pq, int npq, int ns,
                            double* __restrict__ Dh) {
  const size_t n2 = (size_t)ns * ns;
  const size_t tot = n2 * npq;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot; x += (size_t)gridDim.x * blockDim.x) {
    const int tq = (int)(x / npq), w = (int)(x % npq);
    const int t = tq / ns, q = tq % ns;
    const int r = pq[2 * w], s = pq[2 * w + 1];
    auto dt = [&](int a, int b) {
      const size_t ab = (size_t)a * ns + b;
      const size_t tq1 = (size_t)t * ns + q, qt1 = (size_t)q * ns + t;
      return D2[tq1 * n2 + ab] + D2[qt1 * n2 + ab] + D2[ab * n2 + tq1] + D2[ab * n2 + qt1];
    };
    double v = dt(r, s);
    if (r != s) v += dt(s, r);
    Dh[x] = v;
  }
}

// gradC[t][mu] = sum_w sum_q Dh[t][q][w] W[w][mu][q]
__global__ void k_grad_eri(const double* __restrict__ Dh, const double* __restrict__ W, int npq, int ns, int nbf,
                           double* __restrict__ gradC) {
  const size_t tot = (size_t)ns * nbf;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < tot; x += (size_t)gridDim.x * blockDim.x) {
    const int t = (int)(x / nbf), mu = (int)(x % nbf);
    double v = 0.0;
    for (int w = 0; w < npq; ++w) {
      const double* Wm = W + ((size_t)w * nbf + mu) * ns;
      for (int q = 0; q < ns; ++q) v += Dh[((size_t)t * ns + q) * npq + w] * Wm[q];
    }
    gradC[x] = v;
  }
}

// M = X + X^T (n x n)
__global__ void k_sym_into(const double* __restrict__ X, int n, double* __restrict__ M) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t % n;
  M[t] = X[t] + X[(size_t)j * n + i];
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(4096, (n + 255) / 256 + 1); }

// op(A)[M x K] * op(B)[K x N] with M split into chunks the 2-D grid can hold.
hipError_t gemm_rows(hipStream_t st, long M, int N, int K, double alpha, const double* A, int lda,
                     const double* B, int ldb, bool tb, double beta, double* C, int ldc) {
  const long chunk = 65535L * 16;  // rows per launch: the grid's y extent of 16-row tiles
  for (long r0 = 0; r0 < M; r0 += chunk) {
    const int m = (int)std::min(chunk, M - r0);
    const hipError_t e = gemm(st, m, N, K, alpha, A + (size_t)r0 * lda, lda, false, B, ldb, tb, beta,
                              C + (size_t)r0 * ldc, ldc);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

int vsvb::sc_setup(vsvb_context* c, const vsvb_system* s) {
  SpinCoupled& sc = c->sc;
  if (s->nspinc < 1 || !s->pair_sc || !s->coeff_sc) return fail(VSVB_ERR_ARG, "spin-coupled pairs without spin couplings");
  if (s->npair > 16) return fail(VSVB_ERR_UNSUPPORTED, "more than 16 spin-coupled pairs");
  const int np = s->npair, nnd = 2 * s->npair + s->nunpd;
  // doubly occupied orbitals with an alpha electron (all, or all but the last: beta_only_last)
  const int nda = s->ndocc - (c->nalpha_rows < c->nsp ? 1 : 0);
  sc.na = s->npair + s->nunpd + nda;
  sc.nb = s->npair + s->ndocc;
  if (sc.na > kMaxSpinDim) return fail(VSVB_ERR_UNSUPPORTED, "alpha block larger than 48 orbitals");
  sc.ndet = s->nspinc << np;
  sc.nspinc = s->nspinc;
  sc.npair = np;
  sc.npairs = (long)sc.ndet * sc.ndet;
  const int ns = c->nsp;
  sc.npq = ns * (ns + 1) / 2;
  std::vector<int> al((size_t)sc.ndet * sc.na), bl((size_t)sc.ndet * std::max(sc.nb, 1)), pq(2 * (size_t)sc.npq);
  std::vector<double> cd(sc.ndet);
  for (int isc = 0; isc < s->nspinc; ++isc) {
    std::vector<int> seen(nnd, 0);
    for (int k = 0; k < np; ++k)
      for (int ab = 0; ab < 2; ++ab) {
        const int slot = s->pair_sc[((size_t)isc * 2 + ab) * np + k];  // pair_sc(k, ab, isc)
        if (slot < 1 || slot > 2 * np) return fail(VSVB_ERR_ARG, "pair_sc slot outside the spin-coupled orbitals");
        if (seen[slot - 1]++) return fail(VSVB_ERR_ARG, "pair_sc uses an orbital twice in one spin coupling");
      }
    for (int mask = 0; mask < (1 << np); ++mask) {
      const int d = (isc << np) + mask;
      int* A = al.data() + (size_t)d * sc.na;
      int* B = bl.data() + (size_t)d * sc.nb;
      for (int k = 0; k < np; ++k) {
        int x = s->pair_sc[((size_t)isc * 2) * np + k] - 1, y = s->pair_sc[((size_t)isc * 2 + 1) * np + k] - 1;
        if (mask >> k & 1) std::swap(x, y);
        A[k] = x;
        B[k] = y;
      }
      for (int u = 0; u < s->nunpd; ++u) A[np + u] = 2 * np + u;
      for (int o = 0; o < s->ndocc; ++o) {
        if (o < nda) A[np + s->nunpd + o] = nnd + o;
        B[np + o] = nnd + o;
      }
      cd[d] = s->coeff_sc[isc];
    }
  }
  for (int r = 0, w = 0; r < ns; ++r)
    for (int q = 0; q <= r; ++q, ++w) {
      pq[2 * w] = r;
      pq[2 * w + 1] = q;
    }
  const size_t n2 = (size_t)ns * ns, nn = (size_t)c->nbf * c->nbf, P = (size_t)sc.npairs;
  const double bytes = 8.0 * (2.0 * sc.npq * nn + (double)sc.npq * c->nbf * ns + 4.0 * n2 * n2 +
                              P * (sc.na * sc.na + sc.nb * sc.nb + 25.0 * n2 + 12.0) + n2 * ns * sc.npq);
  if (bytes > 64e9) return fail(VSVB_ERR_UNSUPPORTED, "spin-coupled work arrays would exceed 64 GB");
  HIP_OK(sc.d_alist.alloc(al.size()));
  HIP_OK(sc.d_blist.alloc(bl.size()));
  HIP_OK(sc.d_cdet.alloc(cd.size()));
  HIP_OK(sc.d_pq.alloc(pq.size()));
  HIP_OK(sc.d_pdens.alloc(sc.npq * nn));
  HIP_OK(sc.d_pj.alloc(sc.npq * nn));
  HIP_OK(sc.d_W.alloc((size_t)sc.npq * c->nbf * ns));
  HIP_OK(sc.d_eri.alloc(n2 * n2));
  HIP_OK(sc.d_erik.alloc(n2 * n2));
  HIP_OK(sc.d_Ga.alloc(P * sc.na * sc.na));
  HIP_OK(sc.d_Gb.alloc(P * std::max(sc.nb * sc.nb, 1)));
  HIP_OK(sc.d_Xa.alloc(P * n2));
  HIP_OK(sc.d_Xb.alloc(P * n2));
  HIP_OK(sc.d_Xt.alloc(P * n2));
  HIP_OK(sc.d_JX.alloc(P * n2));
  HIP_OK(sc.d_KXa.alloc(P * n2));
  HIP_OK(sc.d_KXb.alloc(P * n2));
  HIP_OK(sc.d_w.alloc(P));
  HIP_OK(sc.d_num.alloc(P));
  HIP_OK(sc.d_phi.alloc(P));
  HIP_OK(sc.d_e1w.alloc(P));
  HIP_OK(sc.d_wv.alloc(P));
  HIP_OK(sc.d_Shat.alloc(n2));
  HIP_OK(sc.d_hhat.alloc(n2));
  HIP_OK(sc.d_hsym.alloc(n2));
  HIP_OK(sc.d_D2.alloc(n2 * n2));
  HIP_OK(sc.d_Dk.alloc(n2 * n2));
  HIP_OK(sc.d_Dh.alloc(n2 * sc.npq));
  HIP_OK(sc.d_Xw.alloc(P * n2));
  HIP_OK(sc.d_Y.alloc(6 * P * n2));
  HIP_OK(sc.d_JY.alloc(6 * P * n2));
  HIP_OK(sc.d_KY.alloc(6 * P * n2));
  HIP_OK(sc.d_q4.alloc(8 * P));
  HIP_OK(sc.d_ham.alloc((size_t)sc.nspinc * sc.nspinc));
  HIP_OK(sc.d_ovl.alloc((size_t)sc.nspinc * sc.nspinc));
  HIP_OK(hipMemcpy(sc.d_alist.p, al.data(), al.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sc.d_blist.p, bl.data(), bl.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sc.d_cdet.p, cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(sc.d_pq.p, pq.data(), pq.size() * sizeof(int), hipMemcpyHostToDevice));
  return VSVB_OK;
}

// Energy (and gradient when grad_dev) of a spin-coupled wave function; the
// geometry and orbital basis (C, C_eri, S_orb, h_orb) are prepared by the caller.
int vsvb::sc_enqueue(vsvb_context* c, bool want_grad, double* grad_dev) {
  SpinCoupled& sc = c->sc;
  hipStream_t st = c->stream;
  const int ns = c->nsp, nbf = c->nbf, n2 = ns * ns;
  const size_t nn = (size_t)nbf * nbf;
  const long P = sc.npairs;
  int rc;
  // orbital-pair Coulomb matrices from one pass over the screened shell quartets
  hipLaunchKernelGGL(k_pair_dens, dim3(grid_for(sc.npq * nn)), dim3(256), 0, st, c->d_Cs.p, sc.d_pq.p, sc.npq, nbf,
                     sc.d_pdens.p);
  HIP_OK(hipGetLastError());
  if ((rc = run_quartets(c, true, nullptr, sc.npq, sc.d_pdens.p, sc.d_pj.p, c->d_K.p, 0))) return rc;
  hipLaunchKernelGGL(k_sym_batch, dim3(grid_for(sc.npq * nn)), dim3(256), 0, st, sc.d_pj.p, sc.npq, nbf);
  HIP_OK(hipGetLastError());
  // W^rs = J^rs C^T  ([npq nbf] x nbf times nbf x ns), then (pq|rs) and its exchange ordering
  HIP_OK(gemm_rows(st, (long)sc.npq * nbf, ns, nbf, 1.0, sc.d_pj.p, nbf, c->d_Cs.p, nbf, true, 0.0, sc.d_W.p, ns));
  hipLaunchKernelGGL(k_orb_eri, dim3(grid_for((size_t)sc.npq * n2)), dim3(256), 0, st, c->d_Cs.p, sc.d_W.p,
                     sc.d_pq.p, sc.npq, ns, nbf, sc.d_eri.p);
  hipLaunchKernelGGL(k_orb_erik, dim3(grid_for((size_t)n2 * n2)), dim3(256), 0, st, sc.d_eri.p, ns, sc.d_erik.p);
  HIP_OK(hipGetLastError());
  // determinant pairs: split cofactor data (exact for singular blocks)
  DetArgs da{c->d_Sorb.p, sc.d_alist.p, sc.d_blist.p, sc.d_cdet.p, sc.ndet, sc.na, sc.nb, ns,
             sc.d_Y.p, sc.d_q4.p, c->d_status.p};
  const int dmax = std::max(sc.na, sc.nb);
  const size_t lds_det = 3 * (size_t)dmax * dmax * sizeof(double);
  const long chunk = 1L << 30;
  for (long p0 = 0; p0 < P; p0 += chunk) {
    const unsigned nb = (unsigned)std::min(chunk, P - p0);
    hipLaunchKernelGGL(k_det_pairs, dim3(nb), dim3(64), lds_det, st, da, p0);
    HIP_OK(hipGetLastError());
  }
  if (sc.want_mats) {
    HIP_OK(launch_fill(st, sc.d_ham.p, (size_t)sc.nspinc * sc.nspinc, 0.0));
    HIP_OK(launch_fill(st, sc.d_ovl.p, (size_t)sc.nspinc * sc.nspinc, 0.0));
  }
  // Coulomb and exchange images of the 6 matrices of every pair: [6P x n^2] x [n^2 x n^2]
  HIP_OK(gemm_rows(st, 6 * P, n2, n2, 1.0, sc.d_Y.p, n2, sc.d_eri.p, n2, false, 0.0, sc.d_JY.p, n2));
  HIP_OK(gemm_rows(st, 6 * P, n2, n2, 1.0, sc.d_Y.p, n2, sc.d_erik.p, n2, false, 0.0, sc.d_KY.p, n2));
  for (long p0 = 0; p0 < P; p0 += chunk) {
    const unsigned nb = (unsigned)std::min(chunk, P - p0);
    hipLaunchKernelGGL(k_pair_energy, dim3(nb), dim3(64), 0, st, c->d_Horb.p, sc.d_Y.p, sc.d_JY.p, sc.d_KY.p,
                       sc.d_q4.p, sc.d_cdet.p, sc.ndet, n2, p0, sc.d_phi.p, sc.d_num.p, sc.d_w.p, sc.d_e1w.p,
                       sc.npair, sc.nspinc, sc.want_mats ? sc.d_ham.p : nullptr, sc.d_ovl.p);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(launch_reduce(st, sc.d_num.p, P, 1, c->d_scal.p + 0));
  HIP_OK(launch_reduce(st, sc.d_w.p, P, 1, c->d_scal.p + 1));
  HIP_OK(launch_reduce(st, sc.d_e1w.p, P, 1, c->d_scal.p + 4));
  hipLaunchKernelGGL(k_sc_final, dim3(1), dim3(1), 0, st, c->d_scal.p, c->enuc);
  HIP_OK(hipGetLastError());
  if (!want_grad) return VSVB_OK;

  // ---- adjoint (every block regular: checked at the end of the call)
  for (long p0 = 0; p0 < P; p0 += chunk) {
    const unsigned nb = (unsigned)std::min(chunk, P - p0);
    hipLaunchKernelGGL(k_pair_regular, dim3(nb), dim3(64), 0, st, sc.d_Y.p, sc.d_JY.p, sc.d_KY.p, sc.d_q4.p,
                       sc.d_alist.p, sc.d_blist.p, sc.ndet, sc.na, sc.nb, ns, p0, sc.d_Xa.p, sc.d_Xb.p, sc.d_Xt.p,
                       sc.d_JX.p, sc.d_KXa.p, sc.d_KXb.p, sc.d_Ga.p, sc.d_Gb.p);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(launch_fill(st, sc.d_Shat.p, n2, 0.0));
  AdjArgs aa{c->d_Horb.p, sc.d_JX.p, sc.d_KXa.p, sc.d_KXb.p, sc.d_Ga.p, sc.d_Gb.p, sc.d_w.p, sc.d_phi.p,
             c->d_scal.p, sc.d_alist.p, sc.d_blist.p, sc.ndet, sc.na, sc.nb, ns, sc.d_Shat.p, sc.d_wv.p};
  const size_t lds_adj = 3 * (size_t)dmax * dmax * sizeof(double);
  for (long p0 = 0; p0 < P; p0 += chunk) {
    const unsigned nb = (unsigned)std::min(chunk, P - p0);
    hipLaunchKernelGGL(k_pair_adjoint, dim3(nb), dim3(64), lds_adj, st, aa, p0);
    HIP_OK(hipGetLastError());
  }
  // hhat = sum_t wv_t Gamma^t_t ; D2 = 1/2 Xw^T Xt - 1/2 (exchange-ordered Xwa^T Xa + Xwb^T Xb)
  HIP_OK(gemm(st, 1, n2, (int)P, 1.0, sc.d_wv.p, (int)P, false, sc.d_Xt.p, n2, false, 0.0, sc.d_hhat.p, n2));
  hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(P * n2)), dim3(256), 0, st, sc.d_Xt.p, sc.d_wv.p, P, n2, sc.d_Xw.p);
  HIP_OK(gemm(st, n2, n2, (int)P, 0.5, sc.d_Xw.p, n2, true, sc.d_Xt.p, n2, false, 0.0, sc.d_D2.p, n2));
  for (int spin = 0; spin < 2; ++spin) {
    const double* X = spin == 0 ? sc.d_Xa.p : sc.d_Xb.p;
    hipLaunchKernelGGL(k_scale_rows, dim3(grid_for(P * n2)), dim3(256), 0, st, X, sc.d_wv.p, P, n2, sc.d_Xw.p);
    HIP_OK(gemm(st, n2, n2, (int)P, 1.0, sc.d_Xw.p, n2, true, X, n2, false, 0.0, sc.d_Dk.p, n2));
    hipLaunchKernelGGL(k_exchange_fold, dim3(grid_for((size_t)n2 * n2)), dim3(256), 0, st, sc.d_Dk.p, ns, sc.d_D2.p);
    HIP_OK(hipGetLastError());
  }
  // four-index term through W^rs, then the overlap and one-electron terms
  hipLaunchKernelGGL(k_grad_coef, dim3(grid_for((size_t)n2 * sc.npq)), dim3(256), 0, st, sc.d_D2.p, sc.d_pq.p,
                     sc.npq, ns, sc.d_Dh.p);
  hipLaunchKernelGGL(k_grad_eri, dim3(grid_for((size_t)ns * nbf)), dim3(256), 0, st, sc.d_Dh.p, sc.d_W.p, sc.npq, ns,
                     nbf, c->d_gradC.p);
  HIP_OK(hipGetLastError());
  for (int k = 0; k < 2; ++k) {
    const double* X = k == 0 ? sc.d_Shat.p : sc.d_hhat.p;
    const double* M = k == 0 ? c->d_S.p : c->d_H.p;
    hipLaunchKernelGGL(k_sym_into, dim3((n2 + 255) / 256), dim3(256), 0, st, X, ns, sc.d_hsym.p);
    HIP_OK(gemm(st, ns, nbf, ns, 1.0, sc.d_hsym.p, ns, false, c->d_Ct.p, nbf, false, 0.0, c->d_X.p, nbf));
    HIP_OK(gemm(st, ns, nbf, nbf, 1.0, c->d_X.p, nbf, false, M, nbf, false, 1.0, c->d_gradC.p, nbf));
  }
  return launch_gather_grad(c, c->d_gradC.p, grad_dev);
}

int vsvb::sc_set_coupling(vsvb_context* c, const double* coeff_sc) {
  SpinCoupled& sc = c->sc;
  std::vector<double> cd(sc.ndet);
  for (int d = 0; d < sc.ndet; ++d) cd[d] = coeff_sc[d >> sc.npair];
  HIP_OK(hipStreamSynchronize(c->stream));
  HIP_OK(hipMemcpy(sc.d_cdet.p, cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice));
  return VSVB_OK;
}

int vsvb::sc_matrices(vsvb_context* c, double* ham, double* ovl) {
  SpinCoupled& sc = c->sc;
  int rc;
  sc.want_mats = true;
  rc = sc_enqueue(c, false, nullptr);
  sc.want_mats = false;
  if (rc) return rc;
  const int m = sc.nspinc, ns = c->nsp;
  std::vector<double> h((size_t)m * m), o((size_t)m * m), sorb((size_t)ns * ns);
  HIP_OK(hipMemcpyAsync(h.data(), sc.d_ham.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(o.data(), sc.d_ovl.p, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(sorb.data(), c->d_Sorb.p, sorb.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if ((rc = sc_finish(c, nullptr, nullptr, nullptr))) return rc;  // synchronizes and checks the status words
  // The reference normalizes every orbital in place (normal, src/valence.F90:2157-2182); the
  // energy does not depend on it, these matrices do: a singly occupied orbital enters every
  // determinant pair twice (bra and ket), a doubly occupied one four times.
  const int nnd = 2 * sc.npair + c->nunpd;
  double logf = 0.0;
  for (int p = 0; p < ns; ++p)
    logf -= (p < nnd || p >= c->nalpha_rows ? 1.0 : 2.0) * std::log(sorb[(size_t)p * ns + p]);
  const double f = std::exp(logf);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const int a = std::max(i, j), b = std::min(i, j);
      ham[(size_t)i * m + j] = f * h[(size_t)a * m + b];
      ovl[(size_t)i * m + j] = f * o[(size_t)a * m + b];
    }
  return VSVB_OK;
}

int vsvb::sc_finish(vsvb_context* c, double* energy, double* parts, double* grad_host) {
  int rc;
  double h[6];
  HIP_OK(hipMemcpyAsync(h, c->d_scal.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  if (grad_host)
    HIP_OK(hipMemcpyAsync(grad_host, c->d_grad.p, c->totlen * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if ((rc = check_status(c, grad_host != nullptr))) return rc;
  if (grad_host) {
    int st3 = 0;
    HIP_OK(hipMemcpy(&st3, c->d_status.p + 3, sizeof(int), hipMemcpyDeviceToHost));
    if (st3)
      return fail(VSVB_ERR_UNSUPPORTED, "gradient at a singular or nearly singular transition overlap of the "
                                        "spin-coupled expansion (pivot ratio below 1e-9); the energy is exact there");
  }
  if ((rc = finish_counters(c))) return rc;
  if (!std::isfinite(h[2])) return fail(VSVB_ERR_NUMERIC, "non-finite energy");
  if (energy) *energy = h[2];
  if (parts) {
    parts[0] = h[3];
    parts[1] = 1.0;
    parts[2] = h[5];
    parts[3] = c->enuc;
  }
  return VSVB_OK;
}
