/*
 * CPU oracle for the VSVB energy path -- TEST INFRASTRUCTURE ONLY (see header).
 *
 * Each function names the reference routine it restates.  The integral
 * engine underneath (McMurchie-Davidson) stands in for SIMINT, which the
 * reference fetches at build time (install-simint.sh) and which is not
 * available offline; its conventions are the ones VALENCE relies on:
 * no internal normalization (VALENCE pre-normalizes, norm_prim/setangn),
 * CCA cartesian order, bra-pair-major ERI blocks with i slowest, potential
 * integrals returning -Z <a|1/r_C|b>, no screening (tolerance 0.0).
 */
#include "vsvb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAX 2
#define LTOT (4 * LMAX)
#define NCART(l) (((l) + 1) * ((l) + 2) / 2)

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* Boys function F_n(T) = int_0^1 t^{2n} exp(-T t^2) dt                 */
/* ------------------------------------------------------------------ */
void orc_boys(int nmax, double T, double *F) {
    if (T < 30.0) {
        /* F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)): all terms
         * positive, so the sum is accurate to a few ulp; then the downward
         * recursion F_n = (2T F_{n+1} + e^{-T}) / (2n+1), which is stable. */
        double term = 1.0 / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * T / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        double ex = exp(-T);
        F[nmax] = ex * sum;
        for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + ex) / (2 * n + 1);
    } else {
        /* F_0 = (1/2) sqrt(pi/T) erf(sqrt T); upward recursion is stable for T >> n. */
        double ex = exp(-T);
        F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - ex) / (2.0 * T);
    }
}

/* ------------------------------------------------------------------ */
/* Cartesian tables: CCA order (valence.F90:2364-2378) and angn          */
/* (setangn, valence.F90:2395-2421).                                    */
/* ------------------------------------------------------------------ */
static int cart_xyz[LMAX + 1][6][3];
static double cart_angn[LMAX + 1][6];

static void init_cart(void) {
    static int done = 0;
    if (done) return;
    for (int l = 0; l <= LMAX; ++l) {
        int n = 0;
        for (int i = 0; i <= l; ++i)
            for (int j = 0; j <= i; ++j) {
                cart_xyz[l][n][0] = l - i;
                cart_xyz[l][n][1] = i - j;
                cart_xyz[l][n][2] = j;
                ++n;
            }
    }
    double ashl[2 * LMAX + 1], ashi[2 * LMAX + 1];
    for (int i = 0; i <= LMAX; ++i) {
        double fii = 1.0;
        for (int j = 1; j <= 2 * i - 1; j += 2) fii *= (double)j;
        fii = sqrt(fii);
        ashl[i] = fii;
        ashi[i] = 1.0 / fii;
    }
    for (int l = 0; l <= LMAX; ++l)
        for (int c = 0; c < NCART(l); ++c)
            cart_angn[l][c] = ashl[l] * ashi[cart_xyz[l][c][0]] * ashi[cart_xyz[l][c][1]] * ashi[cart_xyz[l][c][2]];
    done = 1;
}

/* norm_prim (valence.F90:2264-2306) */
static double dblfac(int n) {
    double d = 1.0;
    for (int i = 3; i <= n; i += 2) d *= (double)i;
    return d;
}

void orc_norm_prim(int l, int n, const double *exps, const double *unnorm, double *out) {
    double pi32 = pow(acos(-1.0), 1.5);
    double fac = pi32 * dblfac(2 * l - 1) * pow(2.0, -l);
    double fax = -1.5 - l;
    for (int i = 0; i < n; ++i) {
        double sovl = fac * pow(2.0 * exps[i], fax);
        out[i] = unnorm[i] * pow(sovl, -0.5);
    }
    double sovl = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) sovl += fac * out[i] * out[j] * pow(exps[i] + exps[j], fax);
    sovl = pow(sovl, -0.5);
    for (int i = 0; i < n; ++i) out[i] *= sovl;
}

/* ------------------------------------------------------------------ */
/* McMurchie-Davidson building blocks                                   */
/* ------------------------------------------------------------------ */
/* E[i][j][t] for one direction, i<=la(+2), j<=lb(+2), t<=i+j. */
#define EMAX (LMAX + 3)
typedef double Etab[EMAX][EMAX][2 * EMAX];

static void hermite_e(int la, int lb, double a, double b, double xab, Etab E) {
    double p = a + b, mu = a * b / p;
    double xpa = -b / p * xab, xpb = a / p * xab;
    memset(E, 0, sizeof(Etab));
    E[0][0][0] = exp(-mu * xab * xab);
    for (int i = 0; i <= la; ++i) {
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = xpa * E[i - 1][j][t];
                    if (t > 0) v += E[i - 1][j][t - 1] / (2 * p);
                    if (t + 1 <= i - 1 + j) v += (t + 1) * E[i - 1][j][t + 1];
                } else {
                    v = xpb * E[i][j - 1][t];
                    if (t > 0) v += E[i][j - 1][t - 1] / (2 * p);
                    if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
                }
                E[i][j][t] = v;
            }
        }
    }
}

/* R^0_{tuv} for t+u+v <= L, alpha, PC = (X,Y,Z).  R[t][u][v]. */
#define RDIM (LTOT + 1)
typedef double Rtab[RDIM][RDIM][RDIM];

static void hermite_r(int L, double alpha, double X, double Y, double Z, Rtab R) {
    static __thread double Rn[RDIM + 1][RDIM][RDIM][RDIM];
    double F[RDIM + 1];
    double T = alpha * (X * X + Y * Y + Z * Z);
    orc_boys(L, T, F);
    double m2a = -2.0 * alpha, pw = 1.0;
    for (int n = 0; n <= L; ++n) {
        Rn[n][0][0][0] = pw * F[n];
        pw *= m2a;
    }
    for (int n = L - 1; n >= 0; --n) {
        int top = L - n;
        for (int s = 1; s <= top; ++s)
            for (int t = 0; t <= s; ++t)
                for (int u = 0; u <= s - t; ++u) {
                    int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * Rn[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = Y * Rn[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        val = Z * Rn[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

/* ------------------------------------------------------------------ */
/* Contracted integrals over shells (raw: normalized contraction coefs, */
/* no angn).  Blocks are row-major with the first index slowest.        */
/* ------------------------------------------------------------------ */
static void shell_1e(const double *A, int la, int na, const double *ea, const double *ca,
                     const double *B, int lb, int nb, const double *eb, const double *cb,
                     int natom, const double *coords, const double *charges,
                     double *S, double *T, double *V) {
    int nca = NCART(la), ncb = NCART(lb);
    if (S) memset(S, 0, sizeof(double) * nca * ncb);
    if (T) memset(T, 0, sizeof(double) * nca * ncb);
    if (V) memset(V, 0, sizeof(double) * nca * ncb);
    Etab Ex, Ey, Ez;
    Rtab R;
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            double a = ea[ia], b = eb[ib], p = a + b;
            double cc = ca[ia] * cb[ib];
            hermite_e(la, lb + 2, a, b, A[0] - B[0], Ex);
            hermite_e(la, lb + 2, a, b, A[1] - B[1], Ey);
            hermite_e(la, lb + 2, a, b, A[2] - B[2], Ez);
            double sq = sqrt(M_PI / p);
            double P[3] = {(a * A[0] + b * B[0]) / p, (a * A[1] + b * B[1]) / p, (a * A[2] + b * B[2]) / p};
            for (int i = 0; i < nca; ++i)
                for (int j = 0; j < ncb; ++j) {
                    const int *ai = cart_xyz[la][i], *bj = cart_xyz[lb][j];
                    double s1[3], t1[3];
                    const Etab *Ed[3] = {&Ex, &Ey, &Ez};
                    for (int d = 0; d < 3; ++d) {
                        int li = ai[d], lj = bj[d];
                        const Etab *e = Ed[d];
                        s1[d] = (*e)[li][lj][0] * sq;
                        double km = lj >= 2 ? (*e)[li][lj - 2][0] * sq : 0.0;
                        double kp = (*e)[li][lj + 2][0] * sq;
                        t1[d] = -0.5 * (lj * (lj - 1) * km - 2.0 * b * (2 * lj + 1) * s1[d] + 4.0 * b * b * kp);
                    }
                    if (S) S[i * ncb + j] += cc * s1[0] * s1[1] * s1[2];
                    if (T) T[i * ncb + j] += cc * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                }
            if (V) {
                int L = la + lb;
                for (int c = 0; c < natom; ++c) {
                    double Z = charges[c];
                    if (fabs(Z) <= 1e-12) continue; /* int1e, valence.F90:3146 */
                    const double *C = coords + 3 * c;
                    hermite_r(L, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], R);
                    double pref = -Z * 2.0 * M_PI / p * cc;
                    for (int i = 0; i < nca; ++i)
                        for (int j = 0; j < ncb; ++j) {
                            const int *ai = cart_xyz[la][i], *bj = cart_xyz[lb][j];
                            double sum = 0.0;
                            for (int t = 0; t <= ai[0] + bj[0]; ++t)
                                for (int u = 0; u <= ai[1] + bj[1]; ++u)
                                    for (int v = 0; v <= ai[2] + bj[2]; ++v)
                                        sum += Ex[ai[0]][bj[0]][t] * Ey[ai[1]][bj[1]][u] * Ez[ai[2]][bj[2]][v] * R[t][u][v];
                            V[i * ncb + j] += pref * sum;
                        }
                }
            }
        }
}

int orc_eri_block(const double *A, int la, int na, const double *ea, const double *ca,
                  const double *B, int lb, int nb, const double *eb, const double *cb,
                  const double *C, int lc, int nc, const double *ec, const double *cc,
                  const double *D, int ld, int nd, const double *ed, const double *cd,
                  double *out) {
    init_cart();
    int nA = NCART(la), nB = NCART(lb), nC = NCART(lc), nD = NCART(ld);
    int nab = nA * nB, ncd = nC * nD;
    int Lab = la + lb, Lcd = lc + ld, L = Lab + Lcd;
    memset(out, 0, sizeof(double) * nab * ncd);
    /* per-pair Hermite tables */
    Etab Ex1, Ey1, Ez1, Ex2, Ey2, Ez2;
    Rtab R;
    int nh1 = (Lab + 1) * (Lab + 1) * (Lab + 1);
    double *X = (double *)malloc(sizeof(double) * nh1 * ncd);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            double a = ea[ia], b = eb[ib], p = a + b;
            hermite_e(la, lb, a, b, A[0] - B[0], Ex1);
            hermite_e(la, lb, a, b, A[1] - B[1], Ey1);
            hermite_e(la, lb, a, b, A[2] - B[2], Ez1);
            double P[3] = {(a * A[0] + b * B[0]) / p, (a * A[1] + b * B[1]) / p, (a * A[2] + b * B[2]) / p};
            for (int ic = 0; ic < nc; ++ic)
                for (int id = 0; id < nd; ++id) {
                    double c = ec[ic], d = ed[id], q = c + d;
                    hermite_e(lc, ld, c, d, C[0] - D[0], Ex2);
                    hermite_e(lc, ld, c, d, C[1] - D[1], Ey2);
                    hermite_e(lc, ld, c, d, C[2] - D[2], Ez2);
                    double Q[3] = {(c * C[0] + d * D[0]) / q, (c * C[1] + d * D[1]) / q, (c * C[2] + d * D[2]) / q};
                    double alpha = p * q / (p + q);
                    hermite_r(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
                    double pref = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * ca[ia] * cb[ib] * cc[ic] * cd[id];
                    /* X[tuv][kl] = sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,...} */
                    for (int t = 0; t <= Lab; ++t)
                        for (int u = 0; u <= Lab - t; ++u)
                            for (int v = 0; v <= Lab - t - u; ++v) {
                                int h = (t * (Lab + 1) + u) * (Lab + 1) + v;
                                for (int k = 0; k < nC; ++k)
                                    for (int l = 0; l < nD; ++l) {
                                        const int *ck = cart_xyz[lc][k], *dl = cart_xyz[ld][l];
                                        double s = 0.0;
                                        for (int tau = 0; tau <= ck[0] + dl[0]; ++tau) {
                                            double ex = Ex2[ck[0]][dl[0]][tau];
                                            for (int nu = 0; nu <= ck[1] + dl[1]; ++nu) {
                                                double exy = ex * Ey2[ck[1]][dl[1]][nu];
                                                for (int ph = 0; ph <= ck[2] + dl[2]; ++ph) {
                                                    double sg = ((tau + nu + ph) & 1) ? -1.0 : 1.0;
                                                    s += sg * exy * Ez2[ck[2]][dl[2]][ph] * R[t + tau][u + nu][v + ph];
                                                }
                                            }
                                        }
                                        X[h * ncd + k * nD + l] = s;
                                    }
                            }
                    for (int i = 0; i < nA; ++i)
                        for (int j = 0; j < nB; ++j) {
                            const int *ai = cart_xyz[la][i], *bj = cart_xyz[lb][j];
                            double *o = out + (i * nB + j) * ncd;
                            for (int t = 0; t <= ai[0] + bj[0]; ++t) {
                                double ex = Ex1[ai[0]][bj[0]][t];
                                for (int u = 0; u <= ai[1] + bj[1]; ++u) {
                                    double exy = ex * Ey1[ai[1]][bj[1]][u];
                                    for (int v = 0; v <= ai[2] + bj[2]; ++v) {
                                        double e = pref * exy * Ez1[ai[2]][bj[2]][v];
                                        const double *x = X + ((t * (Lab + 1) + u) * (Lab + 1) + v) * ncd;
                                        for (int kl = 0; kl < ncd; ++kl) o[kl] += e * x[kl];
                                    }
                                }
                            }
                        }
                }
        }
    free(X);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Molecule-level state                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int atom, l, nprim, ao, tshell; /* tshell: index into the type shell list */
    const double *A;
    const double *exps;
    double *coef; /* normalized con_coeff (shared per type shell) */
} Shell;

typedef struct {
    const orc_system *s;
    int nsh, nbf, nsp;          /* shells, AOs, spatial (electron-carrying) orbitals */
    Shell *sh;
    int *atom_sh0;              /* first global shell of each atom */
    int *ao_shell;              /* AO -> shell */
    double *con_coeff;          /* per primitive (normalized) */
    double *angn_ao;            /* per AO */
    double *coeff;              /* normalized orbital weights (totlen) */
    double *Sao, *Hao;          /* nbf x nbf raw (no angn) */
    double dtol, itol;
    /* orbital-level */
    double *cvec;               /* [norbs][nbf] raw effective weights (no angn) */
} Mol;

static void mol_free(Mol *m) {
    free(m->sh); free(m->atom_sh0); free(m->ao_shell); free(m->con_coeff); free(m->angn_ao);
    free(m->coeff); free(m->Sao); free(m->Hao); free(m->cvec);
}

/* AO index of the k-th (1-based) OBS AO of orbital p: walk OBS atoms, their
 * type shells and CCA components in order (the scatter in ovint/int1e/int2e). */
static int obs_ao_global(const Mol *m, int p, int k1) {
    const orc_system *s = m->s;
    int k = k1 - 1;
    for (int ia = 0; ia < s->orb_natom[p]; ++ia) {
        int a = s->orb_atoms[p * s->mxctr + ia];
        int t = s->atom_type[a];
        int base = m->sh[m->atom_sh0[a]].ao;
        int n = 0;
        for (int ts = s->type_shell_start[t]; ts < s->type_shell_start[t + 1]; ++ts) n += NCART(s->shell_l[ts]);
        if (k < n) return base + k;
        k -= n;
    }
    return -1;
}

/* Effective raw weight vector of orbital p over global AOs (scatter incl. NDF
 * expansion through ndf2obs, valence.F90:2917-2945 and 2191-2247). */
static void build_cvec(Mol *m, int p) {
    const orc_system *s = m->s;
    double *c = m->cvec + (size_t)p * m->nbf;
    memset(c, 0, sizeof(double) * m->nbf);
    for (int i = s->orb_start[p]; i < s->orb_start[p + 1]; ++i) {
        int xp = s->xpset[i];
        if (xp < 1) {
            int indf = xp + s->norbs - 1; /* 0-based NDF orbital */
            for (int j = s->orb_start[indf]; j < s->orb_start[indf + 1]; ++j) {
                int g = obs_ao_global(m, indf, s->xpset[j]);
                if (g >= 0) c[g] += m->coeff[j] * m->coeff[i];
            }
        } else {
            int g = obs_ao_global(m, p, xp);
            if (g >= 0) c[g] += m->coeff[i];
        }
    }
}

static double bilinear(const Mol *m, const double *M, int p, int q) {
    const double *cp = m->cvec + (size_t)p * m->nbf, *cq = m->cvec + (size_t)q * m->nbf;
    double sum = 0.0;
    for (int mu = 0; mu < m->nbf; ++mu) {
        double a = cp[mu] * m->angn_ao[mu];
        if (a == 0.0) continue;
        const double *row = M + (size_t)mu * m->nbf;
        for (int nu = 0; nu < m->nbf; ++nu) sum += a * m->angn_ao[nu] * cq[nu] * row[nu];
    }
    return sum;
}

static int mol_init(Mol *m, const orc_system *s, double dtol_opt, double itol_opt) {
    memset(m, 0, sizeof(*m));
    init_cart();
    m->s = s;
    m->nsp = 2 * s->npair + s->nunpd + s->ndocc;
    m->dtol = dtol_opt >= 0 ? dtol_opt : pow(10.0, -(double)s->ntol_d);
    m->itol = itol_opt >= 0 ? itol_opt : pow(10.0, -(double)s->ntol_i);
    /* 1. norm_prim per type shell (calculate_vsvb_energy, valence.F90:71-79) */
    m->con_coeff = (double *)malloc(sizeof(double) * (s->nprim > 0 ? s->nprim : 1));
    for (int ts = 0; ts < s->nshell_type; ++ts) {
        int p0 = s->shell_prim_start[ts], p1 = s->shell_prim_start[ts + 1];
        if (s->shell_l[ts] > LMAX) return -1;
        orc_norm_prim(s->shell_l[ts], p1 - p0, s->prim_exp + p0, s->prim_coef + p0, m->con_coeff + p0);
    }
    /* 2. global shells */
    m->nsh = 0;
    for (int a = 0; a < s->natom; ++a) {
        int t = s->atom_type[a];
        m->nsh += s->type_shell_start[t + 1] - s->type_shell_start[t];
    }
    m->sh = (Shell *)calloc(m->nsh, sizeof(Shell));
    m->atom_sh0 = (int *)calloc(s->natom + 1, sizeof(int));
    int k = 0, ao = 0;
    for (int a = 0; a < s->natom; ++a) {
        int t = s->atom_type[a];
        m->atom_sh0[a] = k;
        for (int ts = s->type_shell_start[t]; ts < s->type_shell_start[t + 1]; ++ts) {
            Shell *sh = &m->sh[k++];
            sh->atom = a;
            sh->l = s->shell_l[ts];
            sh->tshell = ts;
            sh->nprim = s->shell_prim_start[ts + 1] - s->shell_prim_start[ts];
            sh->exps = s->prim_exp + s->shell_prim_start[ts];
            sh->coef = m->con_coeff + s->shell_prim_start[ts];
            sh->A = s->coords + 3 * a;
            sh->ao = ao;
            ao += NCART(sh->l);
        }
    }
    m->atom_sh0[s->natom] = k;
    m->nbf = ao;
    m->ao_shell = (int *)malloc(sizeof(int) * ao);
    m->angn_ao = (double *)malloc(sizeof(double) * ao);
    for (int i = 0; i < m->nsh; ++i)
        for (int c = 0; c < NCART(m->sh[i].l); ++c) {
            m->ao_shell[m->sh[i].ao + c] = i;
            m->angn_ao[m->sh[i].ao + c] = cart_angn[m->sh[i].l][c];
        }
    /* 3. AO one-electron matrices */
    size_t nn = (size_t)m->nbf * m->nbf;
    m->Sao = (double *)calloc(nn, sizeof(double));
    m->Hao = (double *)calloc(nn, sizeof(double));
    double *charges = (double *)malloc(sizeof(double) * s->natom);
    for (int a = 0; a < s->natom; ++a) charges[a] = s->nuc_charge[s->atom_type[a]];
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < m->nsh; ++i) {
        double S[36], T[36], V[36];
        for (int j = 0; j < m->nsh; ++j) {
            Shell *a = &m->sh[i], *b = &m->sh[j];
            shell_1e(a->A, a->l, a->nprim, a->exps, a->coef, b->A, b->l, b->nprim, b->exps, b->coef,
                     s->natom, s->coords, charges, S, T, V);
            int nb = NCART(b->l);
            for (int x = 0; x < NCART(a->l); ++x)
                for (int y = 0; y < nb; ++y) {
                    size_t idx = (size_t)(a->ao + x) * m->nbf + (b->ao + y);
                    m->Sao[idx] = S[x * nb + y];
                    m->Hao[idx] = T[x * nb + y] + V[x * nb + y];
                }
        }
    }
    free(charges);
    /* 4. orbitals and normalization: NDFs first, then the rest
     *    (calculate_vsvb_energy, valence.F90:144-145; normal 2157-2182). */
    m->coeff = (double *)malloc(sizeof(double) * (s->totlen > 0 ? s->totlen : 1));
    memcpy(m->coeff, s->coeff, sizeof(double) * s->totlen);
    m->cvec = (double *)calloc((size_t)(s->norbs > 0 ? s->norbs : 1) * m->nbf, sizeof(double));
    int first_ndf = s->norbs - s->ndf;
    for (int pass = 0; pass < 2; ++pass) {
        int p0 = pass == 0 ? first_ndf : 0, p1 = pass == 0 ? s->norbs : first_ndf;
        for (int p = p0; p < p1; ++p) {
            build_cvec(m, p);
            double sint = bilinear(m, m->Sao, p, p);
            sint = pow(sint, -0.5);
            for (int i = s->orb_start[p]; i < s->orb_start[p + 1]; ++i) m->coeff[i] *= sint;
        }
    }
    for (int p = 0; p < s->norbs; ++p) build_cvec(m, p);
    return 0;
}

int orc_normalized_coeff(const orc_system *s, double *out) {
    Mol m;
    if (mol_init(&m, s, -1, -1)) return -1;
    memcpy(out, m.coeff, sizeof(double) * s->totlen);
    mol_free(&m);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Two-electron orbital integrals                                       */
/* ------------------------------------------------------------------ */
/* Per-(orbital, shell) weight screen of int2e (valence.F90:3296-3300): a shell
 * is used only if the sum of its squared raw weights exceeds dtol. */
static void screened_cvec(const Mol *m, int p, double *out /* nbf, angn folded */) {
    const double *c = m->cvec + (size_t)p * m->nbf;
    for (int i = 0; i < m->nsh; ++i) {
        const Shell *sh = &m->sh[i];
        double sum = 0.0;
        for (int k = 0; k < NCART(sh->l); ++k) sum += c[sh->ao + k] * c[sh->ao + k];
        for (int k = 0; k < NCART(sh->l); ++k)
            out[sh->ao + k] = sum > m->dtol ? c[sh->ao + k] * m->angn_ao[sh->ao + k] : 0.0;
    }
}

/* int2e restated directly (valence.F90:3184-3438): loop the four OBSs' shells
 * with the per-shell screen, one shell-quartet block each, contract. */
static double int2e_direct(const Mol *m, int io, int jo, int ko, int lo, double *blk) {
    const orc_system *s = m->s;
    double *cv[4];
    int orbs[4] = {io, jo, ko, lo};
    for (int x = 0; x < 4; ++x) {
        cv[x] = (double *)malloc(sizeof(double) * m->nbf);
        screened_cvec(m, orbs[x], cv[x]);
    }
    double g = 0.0;
    int sh_of[4][64], nsh_of[4];
    for (int x = 0; x < 4; ++x) {
        nsh_of[x] = 0;
        int p = orbs[x];
        for (int ia = 0; ia < s->orb_natom[p]; ++ia) {
            int a = s->orb_atoms[p * s->mxctr + ia];
            for (int i = m->atom_sh0[a]; i < m->atom_sh0[a + 1]; ++i) {
                const Shell *sh = &m->sh[i];
                int nz = 0;
                for (int k = 0; k < NCART(sh->l); ++k) nz |= cv[x][sh->ao + k] != 0.0;
                if (nz) sh_of[x][nsh_of[x]++] = i;
            }
        }
    }
    for (int a = 0; a < nsh_of[0]; ++a)
        for (int b = 0; b < nsh_of[1]; ++b)
            for (int c = 0; c < nsh_of[2]; ++c)
                for (int d = 0; d < nsh_of[3]; ++d) {
                    const Shell *A = &m->sh[sh_of[0][a]], *B = &m->sh[sh_of[1][b]];
                    const Shell *C = &m->sh[sh_of[2][c]], *D = &m->sh[sh_of[3][d]];
                    orc_eri_block(A->A, A->l, A->nprim, A->exps, A->coef, B->A, B->l, B->nprim, B->exps, B->coef,
                                  C->A, C->l, C->nprim, C->exps, C->coef, D->A, D->l, D->nprim, D->exps, D->coef, blk);
                    int nA = NCART(A->l), nB = NCART(B->l), nC = NCART(C->l), nD = NCART(D->l);
                    int n = 0;
                    for (int i = 0; i < nA; ++i)
                        for (int j = 0; j < nB; ++j) {
                            double dij = cv[0][A->ao + i] * cv[1][B->ao + j];
                            for (int k = 0; k < nC; ++k)
                                for (int l = 0; l < nD; ++l, ++n)
                                    g += cv[2][C->ao + k] * cv[3][D->ao + l] * dij * blk[n];
                        }
                }
    for (int x = 0; x < 4; ++x) free(cv[x]);
    return g;
}

/* Full table G[p][q][r][s] = (pq|rs) over spatial orbitals, same screened
 * weights as int2e, by a half transformation of AO shell-quartet blocks. */
static double *eri_table(const Mol *m) {
    int n = m->nsp, nbf = m->nbf;
    size_t n2 = (size_t)n * n;
    double *cs = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * nbf);
    for (int p = 0; p < n; ++p) screened_cvec(m, p, cs + (size_t)p * nbf);
    /* orbitals touching each shell */
    int *touch = (int *)calloc((size_t)m->nsh * (n > 0 ? n : 1), sizeof(int));
    int *ntouch = (int *)calloc(m->nsh, sizeof(int));
    for (int i = 0; i < m->nsh; ++i)
        for (int p = 0; p < n; ++p) {
            int nz = 0;
            for (int k = 0; k < NCART(m->sh[i].l); ++k) nz |= cs[(size_t)p * nbf + m->sh[i].ao + k] != 0.0;
            if (nz) touch[(size_t)i * n + ntouch[i]++] = p;
        }
    /* shell pairs M >= N and their row offsets in H */
    int npair_sh = m->nsh * (m->nsh + 1) / 2;
    size_t *off = (size_t *)malloc(sizeof(size_t) * (npair_sh + 1));
    int *pm = (int *)malloc(sizeof(int) * npair_sh), *pn = (int *)malloc(sizeof(int) * npair_sh);
    size_t rows = 0;
    int idx = 0;
    for (int M = 0; M < m->nsh; ++M)
        for (int N = 0; N <= M; ++N) {
            pm[idx] = M; pn[idx] = N; off[idx] = rows;
            rows += (size_t)NCART(m->sh[M].l) * NCART(m->sh[N].l);
            ++idx;
        }
    off[npair_sh] = rows;
    double *H = (double *)calloc(rows * n2, sizeof(double));
    /* Each shell quartet block is evaluated once, for ket pair <= bra pair, and
     * added to both rows: (ab|cd) to the row of (ab) with ket (cd), and, by
     * (ab|cd) = (cd|ab), to the row of (cd) with ket (ab).  Rows receive adds
     * from several threads, so every add is atomic. */
#pragma omp parallel
    {
        double *blk = (double *)malloc(sizeof(double) * 1296);
#pragma omp for schedule(dynamic, 1)
        for (int bp = 0; bp < npair_sh; ++bp) {
            if (!ntouch[pm[bp]] || !ntouch[pn[bp]]) continue; /* no orbital uses this pair in either slot */
            for (int kp = 0; kp <= bp; ++kp) {
                if (!ntouch[pm[kp]] || !ntouch[pn[kp]]) continue;
                const Shell *A = &m->sh[pm[bp]], *B = &m->sh[pn[bp]];
                const Shell *C = &m->sh[pm[kp]], *D = &m->sh[pn[kp]];
                orc_eri_block(A->A, A->l, A->nprim, A->exps, A->coef, B->A, B->l, B->nprim, B->exps, B->coef,
                              C->A, C->l, C->nprim, C->exps, C->coef, D->A, D->l, D->nprim, D->exps, D->coef, blk);
                int nA = NCART(A->l), nB = NCART(B->l), nC = NCART(C->l), nD = NCART(D->l);
                int nab = nA * nB, ncd = nC * nD;
                for (int side = 0; side < (kp < bp ? 2 : 1); ++side) {
                    /* side 0: row pair (A,B), ket shells (C,D); side 1: row pair (C,D), ket (A,B) */
                    const int row_pair = side ? kp : bp, nrow = side ? ncd : nab;
                    const int L = side ? pm[bp] : pm[kp], S = side ? pn[bp] : pn[kp];
                    const Shell *X = side ? A : C, *Y = side ? B : D;
                    const int nX = side ? nA : nC, nY = side ? nB : nD;
                    double *Hrow = H + off[row_pair] * n2;
                    for (int ij = 0; ij < nrow; ++ij) {
                        double *h = Hrow + (size_t)ij * n2;
                        /* element (row ij, ket x y) of the block */
#define BLK(x_, y_) (side ? blk[((x_) * nY + (y_)) * ncd + ij] : blk[ij * ncd + (x_) * nY + (y_)])
                        for (int x = 0; x < ntouch[L]; ++x) {
                            int r = touch[(size_t)L * n + x];
                            const double *cr = cs + (size_t)r * nbf;
                            for (int y = 0; y < ntouch[S]; ++y) {
                                int s2 = touch[(size_t)S * n + y];
                                const double *cq = cs + (size_t)s2 * nbf;
                                double acc = 0.0;
                                for (int k = 0; k < nX; ++k)
                                    for (int l = 0; l < nY; ++l) acc += BLK(k, l) * cr[X->ao + k] * cq[Y->ao + l];
#pragma omp atomic
                                h[(size_t)r * n + s2] += acc;
                            }
                        }
                        if (L != S)
                            for (int x = 0; x < ntouch[S]; ++x) {
                                int r = touch[(size_t)S * n + x];
                                const double *cr = cs + (size_t)r * nbf;
                                for (int y = 0; y < ntouch[L]; ++y) {
                                    int s2 = touch[(size_t)L * n + y];
                                    const double *cq = cs + (size_t)s2 * nbf;
                                    double acc = 0.0;
                                    for (int k = 0; k < nX; ++k)
                                        for (int l = 0; l < nY; ++l) acc += BLK(k, l) * cr[Y->ao + l] * cq[X->ao + k];
#pragma omp atomic
                                    h[(size_t)r * n + s2] += acc;
                                }
                            }
#undef BLK
                    }
                }
            }
        }
        free(blk);
    }
    double *G = (double *)calloc(n2 * n2 > 0 ? n2 * n2 : 1, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int pq = 0; pq < (int)n2; ++pq) {
        int p = pq / n, q = pq % n;
        const double *cp = cs + (size_t)p * nbf, *cq = cs + (size_t)q * nbf;
        double *g = G + (size_t)pq * n2;
        for (int bp = 0; bp < npair_sh; ++bp) {
            const Shell *A = &m->sh[pm[bp]], *B = &m->sh[pn[bp]];
            int nA = NCART(A->l), nB = NCART(B->l);
            for (int i = 0; i < nA; ++i)
                for (int j = 0; j < nB; ++j) {
                    double w = cp[A->ao + i] * cq[B->ao + j];
                    if (pm[bp] != pn[bp]) w += cp[B->ao + j] * cq[A->ao + i];
                    if (w == 0.0) continue;
                    const double *h = H + (off[bp] + (size_t)i * nB + j) * n2;
                    for (size_t rs = 0; rs < n2; ++rs) g[rs] += w * h[rs];
                }
        }
    }
    free(H); free(off); free(pm); free(pn); free(touch); free(ntouch); free(cs);
    return G;
}

int orc_orbital_integrals(const orc_system *s, double dtol, double *S, double *h, double *g) {
    Mol m;
    if (mol_init(&m, s, dtol, -1)) return -1;
    int n = m.nsp;
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            if (S) S[p * n + q] = bilinear(&m, m.Sao, p, q);
            if (h) h[p * n + q] = bilinear(&m, m.Hao, p, q);
        }
    if (g) {
        double *G = eri_table(&m);
        memcpy(g, G, sizeof(double) * (size_t)n * n * n * n);
        free(G);
    }
    mol_free(&m);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Givens determinant kernel (givens_single, src/givens.F90:230-262)    */
/* column-major a(lda, n)                                               */
/* ------------------------------------------------------------------ */
#define A_(r, c) a[((c) - 1) * (size_t)lda + ((r) - 1)]
void orc_givens_single(double *a, int lda, int n, double tol) {
    for (int j = 1; j <= n - 1; ++j)
        for (int i = n; i >= j + 1; --i) {
            double cx = A_(j, i - 1), sx = A_(j, i);
            double r = cx * cx + sx * sx;
            if (fabs(r) > tol) {
                r = pow(r, -0.5);
                cx *= r;
                sx *= r;
                for (int k = 1; k <= n; ++k) {
                    double t1 = A_(k, i - 1), t2 = A_(k, i);
                    A_(k, i - 1) = t1 * cx + t2 * sx;
                    A_(k, i) = -t1 * sx + t2 * cx;
                }
            }
        }
}

double orc_givens_det(const double *a_in, int lda, int n, double tol) {
    double *a = (double *)malloc(sizeof(double) * (size_t)lda * (n > 0 ? n : 1));
    memcpy(a, a_in, sizeof(double) * (size_t)lda * n);
    orc_givens_single(a, lda, n, tol);
    double d = 1.0;
    for (int i = 1; i <= n; ++i) d *= A_(i, i);
    free(a);
    return d;
}
#undef A_

/* ------------------------------------------------------------------ */
/* Cofactor densities (valence.F90:1535-2144, 2437-2714)                */
/* ------------------------------------------------------------------ */
typedef struct {
    int npair, nunpd, ndocc, nalpha, nbeta, nelec, nspinc, lda;
    const int *pair_sc;    /* [nspinc][npair][2] */
    const double *coeff_sc;
    const double *wdet;    /* (nelec, nelec) column-major */
    double dtol;
    int *bra_a, *bra_b, *ket_a, *ket_b, *bexch, *kexch; /* 1-based indices, arrays 1..n */
    double *aket, *bket;
    double *abra_npair, *bbra_npair, *abra_docc_un, *bbra_docc_un, *aket_docc_un, *bket_docc_un;
    int dme_b[3], dme_k[3];
    long long ndet;
} DWork;

#define WDET(i, j) w->wdet[((j) - 1) * (size_t)w->nelec + ((i) - 1)]
#define AKET(r, c) w->aket[((c) - 1) * (size_t)w->lda + ((r) - 1)]
#define BKET(r, c) w->bket[((c) - 1) * (size_t)w->lda + ((r) - 1)]
#define ABRA_NPAIR(k, l) w->abra_npair[((l) - 1) * (size_t)(w->npair > 0 ? w->npair : 1) + ((k) - 1)]
#define BBRA_NPAIR(k, l) w->bbra_npair[((l) - 1) * (size_t)(w->npair > 0 ? w->npair : 1) + ((k) - 1)]
#define NDU (w->ndocc + w->nunpd > 0 ? w->ndocc + w->nunpd : 1)
#define ABRA_DU(i, j) w->abra_docc_un[((j) - 1) * (size_t)NDU + ((i) - 1)]
#define BBRA_DU(i, j) w->bbra_docc_un[((j) - 1) * (size_t)NDU + ((i) - 1)]
#define AKET_DU(i, j) w->aket_docc_un[((j) - 1) * (size_t)NDU + ((i) - 1)]
#define BKET_DU(i, j) w->bket_docc_un[((j) - 1) * (size_t)NDU + ((i) - 1)]
#define PAIR_SC(k, ab, isc) w->pair_sc[(((isc) - 1) * (w->npair) + ((k) - 1)) * 2 + ((ab) - 1)]

static void dwork_init(DWork *w, const orc_system *s, const double *wdet, double dtol) {
    memset(w, 0, sizeof(*w));
    w->npair = s->npair; w->nunpd = s->nunpd; w->ndocc = s->ndocc;
    w->nalpha = s->npair + s->nunpd + s->ndocc;
    w->nbeta = s->npair + s->ndocc;
    w->nelec = 2 * s->npair + 2 * s->ndocc + s->nunpd;
    w->nspinc = s->nspinc;
    w->lda = w->nalpha > 0 ? w->nalpha : 1; /* padded_size = nalpha (valence.F90:161) */
    w->pair_sc = s->pair_sc;
    w->coeff_sc = s->coeff_sc;
    w->wdet = wdet;
    w->dtol = dtol;
    int na = w->nalpha + 2, np = w->npair + 2;
    w->bra_a = (int *)calloc(na, sizeof(int)); w->ket_a = (int *)calloc(na, sizeof(int));
    w->bra_b = (int *)calloc(na, sizeof(int)); w->ket_b = (int *)calloc(na, sizeof(int));
    w->bexch = (int *)calloc(np, sizeof(int)); w->kexch = (int *)calloc(np, sizeof(int));
    size_t sq = (size_t)w->lda * w->lda;
    w->aket = (double *)calloc(sq, sizeof(double));
    w->bket = (double *)calloc(sq, sizeof(double));
    size_t npe = (size_t)(w->npair > 0 ? w->npair : 1) * (w->nelec > 0 ? w->nelec : 1);
    w->abra_npair = (double *)calloc(npe, sizeof(double));
    w->bbra_npair = (double *)calloc(npe, sizeof(double));
    size_t du = (size_t)NDU * (2 * w->npair > 0 ? 2 * w->npair : 1);
    w->abra_docc_un = (double *)calloc(du, sizeof(double));
    w->bbra_docc_un = (double *)calloc(du, sizeof(double));
    w->aket_docc_un = (double *)calloc((size_t)NDU * NDU, sizeof(double));
    w->bket_docc_un = (double *)calloc((size_t)NDU * NDU, sizeof(double));
}

static void dwork_free(DWork *w) {
    free(w->bra_a); free(w->ket_a); free(w->bra_b); free(w->ket_b); free(w->bexch); free(w->kexch);
    free(w->aket); free(w->bket); free(w->abra_npair); free(w->bbra_npair);
    free(w->abra_docc_un); free(w->bbra_docc_un); free(w->aket_docc_un); free(w->bket_docc_un);
}

/* set_up_unpaired_docc (valence.F90:2444-2517) */
static void set_up_unpaired_docc(DWork *w) {
    int dima = w->npair + w->ndocc + w->nunpd, dimb = w->npair + w->ndocc, nsc = 2 * w->npair;
    int j = 2 * w->npair + 1;
    for (int i = w->npair + 1; i <= w->npair + w->nunpd; ++i) { w->bra_a[i] = j; w->ket_a[i] = j; ++j; }
    j = 2 * w->npair + w->nunpd + 1;
    for (int i = w->npair + w->nunpd + 1; i <= w->npair + w->nunpd + w->ndocc; ++i) { w->bra_a[i] = j; w->ket_a[i] = j; j += 2; }
    j = 2 * w->npair + w->nunpd + 2;
    for (int i = w->npair + 1; i <= w->npair + w->ndocc; ++i) { w->bra_b[i] = j; w->ket_b[i] = j; j += 2; }
    for (int jj = 1; jj <= nsc; ++jj) {
        for (int i = 1; i <= dima - w->npair; ++i) ABRA_DU(i, jj) = WDET(w->bra_a[i + w->npair], jj);
        for (int i = 1; i <= dimb - w->npair; ++i) BBRA_DU(i, jj) = WDET(w->bra_b[i + w->npair], jj);
    }
    for (int i = 1; i <= dima - w->npair; ++i)
        for (int jj = 1; jj <= dima - w->npair; ++jj) AKET_DU(jj, i) = WDET(w->bra_a[jj + w->npair], w->ket_a[i + w->npair]);
    for (int i = 1; i <= dimb - w->npair; ++i)
        for (int jj = 1; jj <= dimb - w->npair; ++jj) BKET_DU(jj, i) = WDET(w->bra_b[jj + w->npair], w->ket_b[i + w->npair]);
}

/* build_abket (valence.F90:2524-2586) */
static void build_abket(DWork *w, int dima, int dimb) {
    int np = w->npair;
    if (np > 0) {
        for (int k = 1; k <= np; ++k) {
            for (int l = 1 + np; l <= dima; ++l) AKET(l, k) = ABRA_DU(l - np, w->ket_a[k]);
            for (int l = 1 + np; l <= dimb; ++l) BKET(l, k) = BBRA_DU(l - np, w->ket_b[k]);
        }
        for (int k = 1; k <= dima; ++k)
            for (int l = 1; l <= np; ++l) AKET(l, k) = ABRA_NPAIR(l, w->ket_a[k]);
        for (int k = 1; k <= dimb; ++k)
            for (int l = 1; l <= np; ++l) BKET(l, k) = BBRA_NPAIR(l, w->ket_b[k]);
    }
    for (int k = 1; k <= w->ndocc; ++k) {
        for (int l = 1 + np; l <= np + w->ndocc; ++l) {
            AKET(l, k + np) = AKET_DU(l - np, k);
            BKET(l, k + np) = BKET_DU(l - np, k);
        }
        for (int l = 1 + np + w->ndocc; l <= np + w->nunpd + w->ndocc; ++l) AKET(l, k + np) = AKET_DU(l - np, k);
    }
    for (int k = 1 + w->ndocc; k <= w->ndocc + w->nunpd; ++k)
        for (int l = 1 + np; l <= np + w->ndocc + w->nunpd; ++l) AKET(l, k + np) = AKET_DU(l - np, k);
}

/* check_spin_and_locate (valence.F90:2670-2714) */
static void check_spin(const DWork *w, int sb, int sk, int *lab, int *lak, int *lbb, int *lbk, int *both_a, int *both_b) {
    int ba = 0, ka = 0, bb = 0, kb = 0;
    for (int i = 1; i <= w->nalpha; ++i) {
        if (sb == w->bra_a[i]) { ba = 1; *lab = i; }
        if (sk == w->ket_a[i]) { ka = 1; *lak = i; }
    }
    for (int i = 1; i <= w->nbeta; ++i) {
        if (sb == w->bra_b[i]) { bb = 1; *lbb = i; }
        if (sk == w->ket_b[i]) { kb = 1; *lbk = i; }
    }
    *both_a = ba && ka;
    *both_b = bb && kb;
}

static double givdr(DWork *w, double *m, int n) {
    ++w->ndet;
    orc_givens_single(m, w->lda, n, w->dtol);
    double d = 1.0;
    for (int i = 1; i <= n; ++i) d *= m[(size_t)(i - 1) * w->lda + (i - 1)];
    return d;
}

/* det (valence.F90:1895-2056) */
static void det_(DWork *w, int dima, int dimb, int nord, double *dens, double *xdens, int calc_d, int calc_x) {
    int erep_nz = 0, exch_nz = 0, ba = 0, bb = 0;
    int lab = 0, lak = 0, lbb = 0, lbk = 0;
    double ad = 0.0, bd = 0.0;
    if (calc_d) {
        for (int iord = 1; iord <= nord; ++iord) {
            check_spin(w, w->dme_b[iord], w->dme_k[iord], &lab, &lak, &lbb, &lbk, &ba, &bb);
            if (ba) {
                for (int i = 1; i <= w->nalpha; ++i) { AKET(lab, i) = 0.0; AKET(i, lak) = 0.0; }
                AKET(lab, lak) = 1.0;
            } else if (bb) {
                for (int i = 1; i <= w->nbeta; ++i) { BKET(lbb, i) = 0.0; BKET(i, lbk) = 0.0; }
                BKET(lbb, lbk) = 1.0;
            } else
                break;
        }
        erep_nz = bb || ba;
    }
    int setup_e = calc_d && erep_nz;
    if (calc_x) {
        for (int iord = 1; iord <= nord; ++iord) {
            check_spin(w, w->dme_b[iord], w->dme_k[nord - iord + 1], &lab, &lak, &lbb, &lbk, &ba, &bb);
            if (ba) {
                if (!setup_e) {
                    for (int i = 1; i <= w->nalpha; ++i) { AKET(lab, i) = 0.0; AKET(i, lak) = 0.0; }
                    AKET(lab, lak) = 1.0;
                }
            } else if (bb) {
                if (!setup_e) {
                    for (int i = 1; i <= w->nbeta; ++i) { BKET(lbb, i) = 0.0; BKET(i, lbk) = 0.0; }
                    BKET(lbb, lbk) = 1.0;
                }
            } else
                break;
        }
        exch_nz = ba || bb;
    }
    int setup_x = calc_x && exch_nz;
    if (setup_e || setup_x) {
        if (dima == 0) ad = 1.0;
        else if (dima == 1) ad = AKET(1, 1);
        else ad = givdr(w, w->aket, w->nalpha);
        if (dimb == 0) bd = 1.0;
        else if (dimb == 1) bd = BKET(1, 1);
        else bd = givdr(w, w->bket, w->nbeta);
    }
    if (setup_e) *dens += ad * bd;
    if (setup_x && !setup_e) *xdens -= ad * bd;
    if (setup_x && setup_e) *xdens += ad * bd;
}

/* dket (valence.F90:1777-1870) */
static void dket(DWork *w, int nord, int nexk, int dima, int dimb, double *d, double *x, int jsc, int cd, int cx) {
    for (int k = 1; k <= w->npair; ++k) { w->ket_a[k] = PAIR_SC(k, 1, jsc); w->ket_b[k] = PAIR_SC(k, 2, jsc); }
    build_abket(w, dima, dimb);
    det_(w, dima, dimb, nord, d, x, cd, cx);
    for (int jo = 1; jo <= nexk; ++jo) {
        for (int k = 1; k <= jo; ++k) w->kexch[k] = k;
        for (int k = 1; k <= w->npair; ++k) { w->ket_a[k] = PAIR_SC(k, 1, jsc); w->ket_b[k] = PAIR_SC(k, 2, jsc); }
        for (int k = 1; k <= jo; ++k) { int t = w->ket_a[k]; w->ket_a[k] = w->ket_b[k]; w->ket_b[k] = t; }
        build_abket(w, dima, dimb);
        det_(w, dima, dimb, nord, d, x, cd, cx);
        int j = jo;
        while (w->kexch[1] < 1 + nexk - jo) {
            if (w->kexch[j] < j + nexk - jo) {
                w->kexch[j] += 1;
                for (int k = j + 1; k <= jo; ++k) w->kexch[k] = w->kexch[k - 1] + 1;
                for (int k = 1; k <= w->npair; ++k) { w->ket_a[k] = PAIR_SC(k, 1, jsc); w->ket_b[k] = PAIR_SC(k, 2, jsc); }
                for (int k = 1; k <= jo; ++k) {
                    int l = w->kexch[k], t = w->ket_a[l];
                    w->ket_a[l] = w->ket_b[l]; w->ket_b[l] = t;
                }
                build_abket(w, dima, dimb);
                det_(w, dima, dimb, nord, d, x, cd, cx);
                j = jo;
            } else
                j = j - 1;
        }
    }
}

static void load_bra_rows(DWork *w) {
    for (int l = 1; l <= w->nelec; ++l)
        for (int k = 1; k <= w->npair; ++k) {
            ABRA_NPAIR(k, l) = WDET(w->bra_a[k], l);
            BBRA_NPAIR(k, l) = WDET(w->bra_b[k], l);
        }
}

/* dbra (valence.F90:1648-1761) */
static void dbra(DWork *w, int nord, int nexb, int nexk, int dima, int dimb, double *d, double *x, int isc, int jsc, int cd, int cx) {
    for (int k = 1; k <= w->npair; ++k) { w->bra_a[k] = PAIR_SC(k, 1, isc); w->bra_b[k] = PAIR_SC(k, 2, isc); }
    load_bra_rows(w);
    dket(w, nord, nexk, dima, dimb, d, x, jsc, cd, cx);
    for (int io = 1; io <= nexb; ++io) {
        for (int k = 1; k <= io; ++k) w->bexch[k] = k;
        for (int k = 1; k <= w->npair; ++k) { w->bra_a[k] = PAIR_SC(k, 1, isc); w->bra_b[k] = PAIR_SC(k, 2, isc); }
        for (int k = 1; k <= io; ++k) { int t = w->bra_a[k]; w->bra_a[k] = w->bra_b[k]; w->bra_b[k] = t; }
        load_bra_rows(w);
        dket(w, nord, nexk, dima, dimb, d, x, jsc, cd, cx);
        int i = io;
        while (w->bexch[1] < 1 + nexb - io) {
            if (w->bexch[i] < i + nexb - io) {
                w->bexch[i] += 1;
                for (int k = i + 1; k <= io; ++k) w->bexch[k] = w->bexch[k - 1] + 1;
                for (int k = 1; k <= w->npair; ++k) { w->bra_a[k] = PAIR_SC(k, 1, isc); w->bra_b[k] = PAIR_SC(k, 2, isc); }
                for (int k = 1; k <= io; ++k) {
                    int l = w->bexch[k], t = w->bra_a[l];
                    w->bra_a[l] = w->bra_b[l]; w->bra_b[l] = t;
                }
                load_bra_rows(w);
                dket(w, nord, nexk, dima, dimb, d, x, jsc, cd, cx);
                i = io;
            } else
                i = i - 1;
        }
    }
}

/* density (valence.F90:1535-1600), spinopt = .false. */
static void density(DWork *w, int nord, double *d, double *x, int cd, int cx) {
    int dima = w->npair + w->ndocc + w->nunpd, dimb = w->npair + w->ndocc;
    if (w->nspinc > 0) {
        *d = 0.0;
        *x = 0.0;
        for (int isc = 1; isc <= w->nspinc; ++isc)
            for (int jsc = 1; jsc <= w->nspinc; ++jsc) {
                double ds = 0.0, xs = 0.0; /* density_sc zeroes its outputs */
                dbra(w, nord, w->npair, w->npair, dima, dimb, &ds, &xs, isc, jsc, cd, cx);
                *d += ds * w->coeff_sc[isc - 1] * w->coeff_sc[jsc - 1];
                *x += xs * w->coeff_sc[isc - 1] * w->coeff_sc[jsc - 1];
            }
    } else {
        *d = 0.0;
        *x = 0.0;
        build_abket(w, dima, dimb);
        det_(w, dima, dimb, nord, d, x, cd, cx);
    }
}

/* xm_dtriang / xm_dtriang8 (xm_module.F90:953-991) */
static void dtriang(long long ij, int *i, int *j) {
    long long ii = (long long)sqrt((double)(2 * ij));
    long long k = (ii * ii + ii) / 2;
    while (k < ij) { ++ii; k = (ii * ii + ii) / 2; }
    *i = (int)ii;
    *j = (int)(ij - (ii * ii - ii) / 2);
}

static int indx(int i, int j) {
    int ix = i > j ? i : j, mn = i < j ? i : j;
    return (ix * ix - ix) / 2 + mn;
}

/* ------------------------------------------------------------------ */
/* guess_energy -> wfndet, schwarz_ints, vsvb_energy (valence.F90)       */
/* ------------------------------------------------------------------ */
typedef struct {
    const Mol *m;
    double *G;    /* table or NULL */
    int n;
} Eri;

static double eri_get(const Eri *e, int p, int q, int r, int s) { /* 1-based spatial orbitals */
    if (e->G) {
        size_t n = e->n;
        return e->G[(((size_t)(p - 1) * n + (q - 1)) * n + (r - 1)) * n + (s - 1)];
    }
    double blk[1296];
    return int2e_direct(e->m, p - 1, q - 1, r - 1, s - 1, blk);
}

int orc_energy(const orc_system *s, const orc_options *opt, orc_result *res) {
    memset(res, 0, sizeof(*res));
#ifdef _OPENMP
    if (opt->nthreads > 0) omp_set_num_threads(opt->nthreads);
#endif
    double t0 = now_s();
    Mol m;
    if (mol_init(&m, s, opt->dtol, opt->itol)) return -1;
    /* nuclear repulsion (tools_module.F90:19-40) */
    double enuc = 0.0;
    for (int i = 1; i < s->natom; ++i)
        for (int j = 0; j < i; ++j) {
            double zij = s->nuc_charge[s->atom_type[i]] * s->nuc_charge[s->atom_type[j]];
            if (fabs(zij) > 1e-12) {
                const double *a = s->coords + 3 * i, *b = s->coords + 3 * j;
                double rsq = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
                if (rsq > 1e-5) enuc += zij * pow(rsq, -0.5);
            }
        }
    res->e_nuc = enuc;
    int nnd = 2 * s->npair + s->nunpd;
    int nsp = m.nsp;
    int nelec = 2 * s->npair + 2 * s->ndocc + s->nunpd;
    /* bra/ket electron slot lists (guess_energy, valence.F90:324-336) */
    int *bra = (int *)calloc(nelec + 1, sizeof(int));
    for (int i = 1; i <= nnd; ++i) bra[i] = i;
    for (int d = nnd + 1, i = nnd + 1; d <= nnd + s->ndocc; ++d, i += 2) { bra[i] = d; bra[i + 1] = d; }
    /* orbital-level one-electron integrals */
    double *So = (double *)malloc(sizeof(double) * (size_t)nsp * nsp + 1);
    double *Ho = (double *)malloc(sizeof(double) * (size_t)nsp * nsp + 1);
    for (int p = 0; p < nsp; ++p)
        for (int q = 0; q < nsp; ++q) {
            So[p * nsp + q] = bilinear(&m, m.Sao, p, q);
            Ho[p * nsp + q] = bilinear(&m, m.Hao, p, q);
        }
    /* wfndet (valence.F90:1440-1480): symmetric fill from the lower triangle */
    double *wdet = (double *)calloc((size_t)nelec * nelec + 1, sizeof(double));
    for (int i = 1; i <= nelec; ++i)
        for (int j = 1; j <= i; ++j) {
            double v = So[(bra[i] - 1) * nsp + (bra[j] - 1)];
            wdet[(size_t)(j - 1) * nelec + (i - 1)] = v;
            wdet[(size_t)(i - 1) * nelec + (j - 1)] = v;
        }
    Eri eri = {&m, NULL, nsp};
    if (opt->eri_table) eri.G = eri_table(&m);
    /* schwarz_ints (valence.F90:1489-1523).  In the literal sampling mode
     * (direct int2e, a task slice) values are computed on first use instead. */
    int P = nsp * (nsp + 1) / 2;
    double *schwarz = (double *)calloc(P + 2, sizeof(double));
    const int lazy = !opt->eri_table && opt->task_end >= 0;
    for (int i = 1; i <= nsp; ++i)
        for (int j = 1; j <= i; ++j) {
            if (lazy) { schwarz[indx(i, j)] = -1.0; continue; }
            int ie = i > nnd ? 2 * i - nnd : i, je = j > nnd ? 2 * j - nnd : j;
            double g = eri_get(&eri, bra[ie], bra[je], bra[ie], bra[je]);
            schwarz[indx(i, j)] = sqrt(g);
        }
    res->seconds_setup = now_s() - t0;
    t0 = now_s();
    double dtol = m.dtol, itol = m.itol;
    /* one-electron part (valence.F90:1060-1106) */
    double energy1 = 0.0, wfnorm = 0.0;
    long long ndet = 0;
    {
        DWork w;
        dwork_init(&w, s, wdet, dtol);
        set_up_unpaired_docc(&w);
        if (!opt->skip_1e)
            for (int i = 1; i <= nelec; ++i) {
                int idocc = i > nnd;
                for (int j = 1; j <= nelec; ++j) {
                    int jdocc = j > nnd;
                    if (idocc && jdocc && (i % 2) != (j % 2)) continue;
                    double sint = So[(bra[i] - 1) * nsp + (bra[j] - 1)];
                    double hint = Ho[(bra[i] - 1) * nsp + (bra[j] - 1)];
                    w.dme_b[1] = i;
                    w.dme_k[1] = j;
                    double d1, dummy;
                    density(&w, 1, &d1, &dummy, 1, 0);
                    wfnorm += sint * d1;
                    energy1 += hint * d1;
                }
            }
        ndet += w.ndet;
        dwork_free(&w);
    }
    wfnorm /= (double)nelec;
    res->seconds_1e = now_s() - t0;
    t0 = now_s();
    /* two-electron part (valence.F90:1153-1433), ijkl symmetry enabled */
    long long ntasks = (long long)P * (P + 1) / 2;
    long long tb = opt->task_begin > 0 ? opt->task_begin : 0;
    long long te = opt->task_end >= 0 && opt->task_end < ntasks ? opt->task_end : ntasks;
    double energy2 = 0.0;
    long long nsk = 0, npo = 0, ncalls = 0, nsch = 0;
#pragma omp parallel reduction(+ : energy2, nsk, npo, ndet, ncalls, nsch) if (!lazy)
    {
        DWork w;
        dwork_init(&w, s, wdet, dtol);
        set_up_unpaired_docc(&w);
#pragma omp for schedule(dynamic, 64)
        for (long long task = tb; task < te; ++task) {
            int ijorb, klorb, io, jo, ko, lo;
            dtriang(1 + task, &ijorb, &klorb);
            dtriang(ijorb, &io, &jo);
            dtriang(klorb, &ko, &lo);
            if ((io == jo && io <= nnd) || (ko == lo && ko <= nnd)) continue;
            if (lazy) {
                int pr[4][2] = {{io, ko}, {jo, lo}, {io, lo}, {jo, ko}};
                for (int x = 0; x < 4; ++x) {
                    int ix = indx(pr[x][0], pr[x][1]);
                    if (schwarz[ix] < 0.0) {
                        schwarz[ix] = sqrt(eri_get(&eri, pr[x][0], pr[x][1], pr[x][0], pr[x][1]));
                        ++nsch;
                    }
                }
            }
            int erep_sig = schwarz[indx(io, ko)] * schwarz[indx(jo, lo)] > itol;
            int exch_sig = schwarz[indx(io, lo)] * schwarz[indx(jo, ko)] > itol;
            double gint_e = 0.0, gint_x = 0.0;
            if (erep_sig || exch_sig) {
                if (io == jo && ko == lo) { /* nonsub shortcut, valence.F90:1213-1216 */
                    gint_e = schwarz[indx(io, ko)] * schwarz[indx(io, ko)];
                    gint_x = gint_e;
                } else {
                    int ie = io > nnd ? 2 * io - nnd - 1 : io, je = jo > nnd ? 2 * jo - nnd - 1 : jo;
                    int ke = ko > nnd ? 2 * ko - nnd - 1 : ko, le = lo > nnd ? 2 * lo - nnd - 1 : lo;
                    if (erep_sig) { gint_e = eri_get(&eri, bra[ie], bra[ke], bra[je], bra[le]); ++ncalls; }
                    if (exch_sig) {
                        if (erep_sig && bra[le] == bra[ke]) gint_x = gint_e;
                        else { gint_x = eri_get(&eri, bra[ie], bra[le], bra[je], bra[ke]); ++ncalls; }
                    }
                }
            } else
                ++nsk;
            erep_sig = fabs(gint_e) > itol;
            exch_sig = fabs(gint_x) > itol;
            double esum = 0.0, xsum = 0.0;
            if (erep_sig || exch_sig) {
                int idc = io > nnd, jdc = jo > nnd, kdc = ko > nnd, ldc = lo > nnd;
                int iso = idc ? 2 : 1, jso = jdc ? 2 : 1, kso = kdc ? 2 : 1, lso = ldc ? 2 : 1;
                for (int is = 1; is <= iso; ++is) {
                    int i = idc ? 2 * io - nnd + is - 2 : io;
                    for (int js = 1; js <= jso; ++js) {
                        int j = jdc ? 2 * jo - nnd + js - 2 : jo;
                        if (!(j < i)) continue;
                        for (int ks = 1; ks <= kso; ++ks) {
                            int k = kdc ? 2 * ko - nnd + ks - 2 : ko;
                            for (int ls = 1; ls <= lso; ++ls) {
                                int l = ldc ? 2 * lo - nnd + ls - 2 : lo;
                                if (!(l < k)) continue;
                                int ik0 = idc && kdc && (i % 2) != (k % 2);
                                int jl0 = jdc && ldc && (j % 2) != (l % 2);
                                int ce = !(ik0 || jl0);
                                int il0 = idc && ldc && (i % 2) != (l % 2);
                                int jk0 = jdc && kdc && (j % 2) != (k % 2);
                                int cx = !(il0 || jk0);
                                if ((ce && erep_sig) || (cx && exch_sig)) {
                                    w.dme_b[1] = i; w.dme_k[1] = k; w.dme_b[2] = j; w.dme_k[2] = l;
                                    double de, dx;
                                    density(&w, 2, &de, &dx, erep_sig, exch_sig);
                                    if (erep_sig) esum += de * gint_e;
                                    if (exch_sig) xsum += dx * gint_x;
                                }
                            }
                        }
                    }
                }
            } else
                ++npo;
            if (opt->plant == 1) xsum = 0.0;
            if (opt->plant == 2) xsum = -xsum;
            if (!(ko == io && lo == jo)) { esum *= 2.0; xsum *= 2.0; }
            energy2 += esum - xsum;
        }
        ndet += w.ndet;
        dwork_free(&w);
    }
    res->seconds_2e = now_s() - t0;
    res->e1 = energy1;
    res->e2 = energy2;
    res->wfnorm = wfnorm;
    res->energy = (energy1 + energy2) / wfnorm + enuc;
    res->ntasks = ntasks;
    res->tasks_done = te - tb;
    res->n_schwarz_skip = nsk;
    res->n_post_skip = npo;
    res->n_dets = ndet;
    res->n_int2e_task = ncalls;
    res->n_int2e_schwarz = nsch;
    free(eri.G); free(schwarz); free(wdet); free(So); free(Ho); free(bra);
    mol_free(&m);
    return 0;
}
