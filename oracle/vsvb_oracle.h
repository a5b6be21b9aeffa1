/*
 * CPU oracle for the VSVB energy path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference's plain-energy algorithm
 * (src/valence.F90 guess_energy -> vsvb_energy, its integral layer
 * ovint/int1e/int2e and the Givens determinant kernel src/givens.F90:230-262)
 * so that the HIP product path can be checked against it.  It is linked only
 * by the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark; nothing in the product calls it.
 *
 * Integrals are McMurchie-Davidson (Hermite) formulas, deliberately a
 * different algorithm from the Obara-Saika/Head-Gordon-Pople kernels of the
 * product, so agreement between the two pins both.
 */
#ifndef VSVB_ORACLE_H
#define VSVB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same content as the reference's module state after xm_input
 * (src/integrals_module.F90, src/density_module.F90), 0-based where noted. */
typedef struct {
    int natom, ntype, nshell_type, nprim, norbs, mxctr, totlen;
    int npair, nunpd, ndocc, ndf, nspinc;
    int ntol_d, ntol_i;
    const int *atom_type;        /* [natom] 0-based */
    const double *coords;        /* [natom][3] bohr */
    const double *nuc_charge;    /* [ntype] */
    const int *type_shell_start; /* [ntype+1] */
    const int *shell_l;          /* [nshell_type] */
    const int *shell_prim_start; /* [nshell_type+1] */
    const double *prim_exp;      /* [nprim] */
    const double *prim_coef;     /* [nprim] unnormalized */
    const int *orb_natom;        /* [norbs] */
    const int *orb_atoms;        /* [norbs][mxctr] 0-based */
    const int *orb_start;        /* [norbs+1] */
    const int *xpset;            /* [totlen] 1-based OBS AO index, <1 = NDF */
    const double *coeff;         /* [totlen] */
    const int *pair_sc;          /* [max(nspinc,1)][npair][2] 1-based slots */
    const double *coeff_sc;      /* [max(nspinc,1)] */
} orc_system;

typedef struct {
    double itol;          /* < 0: 10^-ntol_i as the reference */
    double dtol;          /* < 0: 10^-ntol_d as the reference */
    int nthreads;         /* OpenMP threads (0 = library default) */
    int eri_table;        /* 1: orbital ERI table by AO transform; 0: direct int2e per call */
    long long task_begin; /* 2e task slice [begin, end), 0-based; end < 0 = all */
    long long task_end;
    int skip_1e;          /* 1: skip the one-electron loop (task-slice timing) */
    int plant;            /* tests only: 1 = drop the exchange term, 2 = flip its sign */
} orc_options;

typedef struct {
    double energy;        /* E = (E1+E2)/N + E_nuc (guess_energy, valence.F90:344) */
    double e_nuc;
    double e1;            /* unnormalized one-electron numerator */
    double e2;            /* unnormalized two-electron numerator (the task slice) */
    double wfnorm;        /* N */
    long long ntasks;     /* tasks in the full list */
    long long tasks_done; /* tasks in the slice */
    long long n_schwarz_skip;  /* tasks with both ERIs Schwarz-screened */
    long long n_post_skip;     /* tasks with both ERIs below itol after evaluation */
    long long n_dets;     /* Givens determinants evaluated */
    long long n_int2e_task;     /* orbital ERIs evaluated inside the task slice */
    long long n_int2e_schwarz;  /* orbital ERIs evaluated for Schwarz values */
    double seconds_setup, seconds_1e, seconds_2e;
} orc_result;

int orc_energy(const orc_system *sys, const orc_options *opt, orc_result *res);

/* Normalized, in-place-updated orbital coefficients after `normal`
 * (valence.F90:2157-2182); out has totlen entries. */
int orc_normalized_coeff(const orc_system *sys, double *out);

/* Orbital-level integrals over spatial orbitals (0-based, after `normal`):
 * S, h (n x n row-major) and, if g != NULL and n^4 fits, the ERI table
 * g[((p*n+q)*n+r)*n+s] = (pq|rs) with the reference's per-shell weight screen. */
int orc_orbital_integrals(const orc_system *sys, double dtol, double *S, double *h, double *g);

/* Building blocks exposed for unit tests. */
void   orc_boys(int nmax, double T, double *F);           /* F[0..nmax] */
int    orc_eri_block(const double *A, int la, int na, const double *ea, const double *ca,
                     const double *B, int lb, int nb, const double *eb, const double *cb,
                     const double *C, int lc, int nc, const double *ec, const double *cc,
                     const double *D, int ld, int nd, const double *ed, const double *cd,
                     double *out);   /* raw (no angn), i slowest, CCA order */
void   orc_norm_prim(int l, int n, const double *exps, const double *unnorm, double *out);
void   orc_givens_single(double *a, int lda, int n, double tol);
double orc_givens_det(const double *a, int lda, int n, double tol);

#ifdef __cplusplus
}
#endif
#endif
