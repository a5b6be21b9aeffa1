/*
 * vsvb_hip.h -- C ABI of the MI355X VSVB energy path.
 *
 * The library replaces the body of VALENCE's energy kernel
 *   subroutine vsvb_energy(iorb, num_non_docc, num_spatial_orbs, energy, wfnorm,
 *                          spinav, ijkl_symmetry_is_enabled)      src/valence.F90:1010-1011
 * as called from guess_energy (src/valence.F90:341), so that the public entry
 * points valence_api_initialize / valence_api_calculate_energy /
 * valence_api_finalize (src/valence_api.F90:9, 37, 114) keep their signatures.
 * A Fortran host binds these functions with ISO_C_BINDING interfaces; every
 * argument is a plain pointer or scalar, arrays are host memory unless the name
 * says "device", and no GPU type appears in a signature except the opaque
 * stream handle of the *_device entry points.
 *
 * Error behaviour: the reference has no error codes (it calls xm_abort).  Every
 * function here returns 0 on success and a negative code otherwise;
 * vsvb_last_error() gives the message (thread-local).  A Fortran shim calls
 * xm_abort on a non-zero return.
 */
#ifndef VSVB_HIP_H
#define VSVB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSVB_OK 0
#define VSVB_ERR_ARG (-1)      /* invalid argument / unsupported input */
#define VSVB_ERR_DEVICE (-2)   /* HIP runtime or kernel error */
#define VSVB_ERR_NUMERIC (-3)  /* singular overlap block, non-finite result */
#define VSVB_ERR_UNSUPPORTED (-4)

/*
 * The wave function, molecule and basis as the reference holds them after
 * xm_input (src/xm_module.F90:76-335; module integrals, src/integrals_module.F90;
 * module densitywork, src/density_module.F90).  Indices are 0-based except
 * xpset and pair_sc, which keep the reference's 1-based meaning.  Coordinates
 * are in bohr (after angs2bohr, src/tools_module.F90:6-17).  Arrays are read
 * during the call only; the library keeps copies.
 */
typedef struct vsvb_system {
    int32_t natom, ntype, nshell_type, nprim, norbs, mxctr, totlen;
    int32_t npair, nunpd, ndocc, ndf, nspinc;
    int32_t ntol_d, ntol_i;             /* dtol = 10^-ntol_d, itol = 10^-ntol_i */
    const int32_t *atom_type;           /* [natom]            atom_t - 1          */
    const double *coords;               /* [natom][3] bohr    coords              */
    const double *nuc_charge;           /* [ntype]            nuc_charge          */
    const int32_t *type_shell_start;    /* [ntype+1]          map_atom2shell - 1  */
    const int32_t *shell_l;             /* [nshell_type]      ang_mom             */
    const int32_t *shell_prim_start;    /* [nshell_type+1]    map_shell2prim - 1  */
    const double *prim_exp;             /* [nprim]            exponent            */
    const double *prim_coef;            /* [nprim]            unnormalized_con_coeff */
    const int32_t *orb_natom;           /* [norbs]            orbas_atnum         */
    const int32_t *orb_atoms;           /* [norbs][mxctr]     orbas_atset - 1     */
    const int32_t *orb_start;           /* [norbs+1]          map_orbs - 1        */
    const int32_t *xpset;               /* [totlen]           xpset (1-based; <1 = NDF) */
    const double *coeff;                /* [totlen]           coeff               */
    const int32_t *pair_sc;             /* [max(nspinc,1)][2][npair] pair_sc(npair,2,nspinc) as stored (1-based slots) */
    const double *coeff_sc;             /* [max(nspinc,1)]    coeff_sc            */
} vsvb_system;

typedef struct vsvb_options {
    double quartet_tau;  /* AO shell-quartet Schwarz cut: skip if Q_ab*Q_cd < tau (<=0: default 1e-17) */
    int32_t device;      /* HIP device ordinal */
    int32_t pair_chunk;  /* pairs of the wave-uniform side per work-unit claim of the straight-line quartet
                          * kernels (ket pairs; bra pairs for transposed J/K units): 0 = chosen from the
                          * system size, else 1..64.  Results do not depend on it. */
    int32_t grid_inverse_min; /* orbital overlap blocks of at least this many rows are inverted by the
                               * grid-wide elimination (one pair of launches per pivot), smaller ones by one
                               * workgroup: 0 = default (129), else >= 1 (1 = always grid-wide).  Results
                               * do not depend on it beyond rounding. */
    int32_t beta_only_last;   /* 1: the last doubly occupied orbital holds only a beta electron, so that the
                               * alpha and beta orbital sets differ in one orbital -- the determinant of the
                               * first-order orbital step, whose alpha copy of a doubly occupied orbital is
                               * substituted (src/valence.F90:566-590): the substitute is the last unpaired
                               * orbital, the original the last doubly occupied one.  Needs nunpd >= 1
                               * and ndocc >= 1.  0 (default): the reference's wave function. */
} vsvb_options;

typedef struct vsvb_context vsvb_context;

/* Creates the device-resident state for one molecule/basis/orbital layout
 * (called where valence_initialize reads the input, src/valence_initialize_module.F90:4-30).
 * All device memory the energy calls need is allocated here. */
int vsvb_context_create(const vsvb_system *sys, const vsvb_options *opt, vsvb_context **ctx);

/* Frees everything (valence_finalize, src/valence_finalize_module.F90:3-21). */
void vsvb_context_destroy(vsvb_context *ctx);

/* Plain energy: E = (E1 + E2)/N + E_nuc as guess_energy returns it
 * (src/valence.F90:309-345).  coords (bohr, [natom][3]) and coeff ([totlen])
 * may be NULL to reuse the previous call's values (initially the ones in the
 * vsvb_system).  parts, if not NULL, receives {E_el, N, E1, E_nuc} where
 * E_el = (E1+E2)/N; the library returns N = 1 (numerator and norm are
 * defined up to a common factor and it reports them normalized). */
int vsvb_energy(vsvb_context *ctx, const double *coords, const double *coeff, double *energy, double *parts);

/* The internal seam itself: replaces vsvb_energy's output pair (energy, wfnorm)
 * (src/valence.F90:1010-1011, contract at :342-344: caller computes
 * energy/wfnorm + enucrep).  Under MPI, call it on one rank and pass zeros on
 * the others so that xm_equalize_scalar sums to the same values. */
int vsvb_energy_parts(vsvb_context *ctx, const double *coeff, double *energy_numerator, double *wfnorm);

/* Energy and its gradient with respect to every LCAO weight coeff[k]
 * (adjoint pass over the same screened quartets, integrals recomputed).
 * grad has totlen entries; d E / d coeff[k] at the given (unnormalized) weights. */
int vsvb_energy_grad(vsvb_context *ctx, const double *coords, const double *coeff, double *energy, double *grad);

/* Same as vsvb_energy_grad with device pointers (no host staging of the
 * weights or results): coeff_device [totlen], energy_device [1] and grad_device
 * [totlen] are device memory of the context's GPU; stream is a hipStream_t
 * passed as void* (NULL = the null stream).  The evaluation is ordered after the
 * work already queued on that stream, and work queued there afterwards waits for
 * the results; the call returns once they are written (so that numerical
 * errors can be reported).  Geometry: the context's current coordinates. */
int vsvb_energy_grad_device(vsvb_context *ctx, const double *coeff_device, double *energy_device,
                            double *grad_device, void *stream);

/* Allocation-free energy and gradient: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  The same evaluation as vsvb_energy_grad on a context
 * made from sys, but with no context: every device buffer is carved from the
 * library's static 32 MiB device scratch array (one per GPU, held until the
 * process ends); the call allocates and frees nothing, creates no stream or
 * event, and runs on the null stream of opt->device (opt may be NULL).
 * Systems whose buffers exceed the array (C8H18 needs about 11 MiB, C12H26
 * with 310 AOs about 23 MiB) and spin-coupled wave functions (npair > 0) are
 * refused with
 * VSVB_ERR_UNSUPPORTED.  Calls on one GPU share the array: they are
 * serialized by a lock inside the process.  energy: [1]; grad: NULL or
 * [totlen] (host memory). */
int vsvb_energy_grad_static(const vsvb_system *sys, const vsvb_options *opt, double *energy, double *grad);

/* The NITROGEN potential-surface plug-in (src/valence_api_nitrogen.F90):
 * calcsurface (:17-27), the energy at Cartesian coordinates x[natom][3] in
 * angstrom (converted with angs2bohr's 1.889725987722, src/tools_module.F90:11),
 * returned in cm^-1 (x 219474.631).  The wave function is the context's (the
 * reference runs its optimizer inside the call when the input asks for it;
 * this is the guess energy of the reference's output). */
int vsvb_surface(vsvb_context *ctx, const double *x_angstrom, double *v_cm1);

/* Spin-coupling matrices of a spin-coupled wave function (nspinc couplings),
 * replacing the ham/ovl accumulation of the reference's spin optimization
 * (spin_opt, src/valence.F90:850-936; density's spinopt branch, :1549-1569):
 * ham[I][J] = the energy numerator E1 + E2 over bra determinants of coupling I
 * and ket determinants of coupling J, without the coupling weights; ovl[I][J]
 * the same for the overlap (the reference's ovl after its division by nelec).
 * Both [nspinc][nspinc], symmetric, electronic (no E_nuc), for orbitals
 * normalized as the reference's normal does: for weights c,
 * E = c.ham.c / c.ovl.c + E_nuc.  coeff: NULL = the context's weights;
 * geometry: the context's current coordinates. */
int vsvb_spin_matrices(vsvb_context *ctx, const double *coeff, double *ham, double *ovl);
/* The reference's spin_opt step: the lowest root of ham c = E ovl c.  energy =
 * that root + E_nuc; coeff_sc[nspinc] = its eigenvector, normalized to
 * c.ovl.c = 1 with its largest component positive.  The context keeps its
 * weights (the reference overwrites coeff_sc, :931-934): pass them to
 * vsvb_set_spin_coupling to do the same.  VSVB_ERR_NUMERIC when ovl is not
 * positive definite (the reference's rsg failure, :913-927). */
int vsvb_spin_opt(vsvb_context *ctx, const double *coeff, double *energy, double *coeff_sc);
/* New spin-coupling weights coeff_sc[nspinc] for the later calls on this
 * context (the reference's coeff_sc, set by spin_opt, :931-934). */
int vsvb_set_spin_coupling(vsvb_context *ctx, const double *coeff_sc);

/* Batched evaluation of many wave functions of uneven size (displaced
 * geometries, trial orbital rotations, different molecules); no counterpart in
 * the reference, which evaluates one wave function per process call.  Wave
 * functions on one molecule share one pass over the shell quartets; the groups
 * run on several streams so that they overlap.  The first run is enqueued
 * directly and then captured into a HIP graph; every later run is one graph
 * launch (the systems are fixed at creation). */
typedef struct vsvb_batch vsvb_batch;

int vsvb_batch_create(int32_t nwfn, const vsvb_system *const *systems, const vsvb_options *opt, vsvb_batch **batch);
/* energies: [nwfn]; grads: NULL or [nwfn] host pointers of totlen doubles each (NULL entries skipped). */
int vsvb_batch_run(vsvb_batch *batch, int32_t want_grad, double *energies, double *const *grads);
void vsvb_batch_destroy(vsvb_batch *batch);
/* 1 when runs replay the captured graph, 0 when they are enqueued directly
 * (before the first run, or after a failed capture: vsvb_last_error says why). */
int vsvb_batch_uses_graph(vsvb_batch *batch);
/* on = 0 keeps every run direct (an existing graph is released); on = 1 (the
 * default) captures after the next direct run.  No counterpart in the reference.
 * For tools that time or trace dispatches one by one and do not follow graph
 * launches; results are the same either way. */
int vsvb_batch_set_graph(vsvb_batch *batch, int32_t on);
/* Groups of a batch run on n streams (1..4, default 4; more is clipped), largest
 * estimated work first, each to the least-loaded stream, so that independent
 * groups overlap on the GPU.  An
 * existing graph is released (the next run captures again unless graphs are
 * off).  n = 1 serializes the groups, for per-kernel timing without overlap.
 * No counterpart in the reference. */
int vsvb_batch_set_lanes(vsvb_batch *batch, int32_t n);
/* Class launches over many groups (no counterpart in the reference: the
 * reference evaluates one wave function per call).  chunks >= 1:
 * each stream's groups (closed- and high-spin; spin-coupled groups keep their
 * own pass) run in `chunks` consecutive chunks of about equal estimated work,
 * and each chunk launches every quartet class once over all its groups, with
 * persistent workgroups that drain one group's work units after another;
 * chunks = 0 (the default): every group launches its own classes, which
 * overlaps best with the groups' orbital and gradient work on several streams.
 * Results are the same up
 * to the order of floating-point atomic additions.  An existing graph is
 * released.  vsvb_batch_class_stats then reports, per class, the summed time
 * and the number of these launches. */
int vsvb_batch_set_group_launch(vsvb_batch *batch, int32_t chunks);
/* The last run's group launches (no counterpart in the reference): *n = their
 * number (0 without group launches); the first min(*n, cap) go to out as rows
 * of 6 doubles: quartet class, kernel (0 straight-line, 1 straight-line
 * transposed, 2 cooperative), groups, stream, device time in ms (from the
 * kernels' wall-clock stamps; 0 unless profiling is on), algorithmic FLOPs. */
int vsvb_batch_launches(vsvb_batch *batch, double *out, int32_t cap, int32_t *n);
/* New coordinates (bohr) for every wave function of the batch, concatenated
 * [nwfn][natom_i][3], for the next runs (no counterpart in the reference: a
 * batch of displaced geometries, e.g. a finite-difference stencil).  Wave
 * functions created on one geometry share one quartet pass and must receive
 * identical coordinates.  A captured graph is released; the next run is direct. */
int vsvb_batch_set_coords(vsvb_batch *batch, const double *coords);
/* calcsurface over the whole batch in one run: x_angstrom concatenated as in
 * vsvb_batch_set_coords, v_cm1[nwfn] energies in cm^-1. */
int vsvb_surface_batch(vsvb_batch *batch, const double *x_angstrom, double *v_cm1);
/* Launch choices of the batch tuner, which runs before the first run, for
 * groups of one wave function (bucket 0, also applied to groups of up to four)
 * and groups of more than four sharing one quartet pass (bucket 1): out[2][21][8] = per bucket and class (J/K mode)
 * {cooperative kernel 0/1, workgroups per CU (0 = occupancy), LDS exchange
 * staging 0/1, units per claim (0 = automatic), transposed units 0/1, best ms,
 * default-configuration ms, tuned 0/1}; zeros where nothing was tuned. */
int vsvb_batch_tuning(vsvb_batch *batch, double *out);

/* Diagnostics. */
const char *vsvb_last_error(void);
/* Counters of the last call: [0] shell quartets evaluated, [1] primitive quartets,
 * [2] shell quartets skipped by the Schwarz cut, [3] device milliseconds of the ERI stage. */
int vsvb_last_counters(vsvb_context *ctx, double *out4);
/* Enable per-class HIP-event timing of the shell-quartet launches (off by default). */
int vsvb_set_profiling(vsvb_context *ctx, int32_t on);
/* Per quartet class (21 classes, bra pair class major: ss|ss, ps|ss, ps|ps, pp|ss, ...):
 * out[21][6] = {device ms (profiling on), shell quartets, primitive quartets,
 * quartets skipped by the Schwarz cut, algorithmic FP64 operations, launches} of the last call. */
int vsvb_class_stats(vsvb_context *ctx, double *out);
/* Batch versions: profiling switch for every wave function, and the per-class
 * statistics summed over the wave functions of the last run. */
int vsvb_batch_set_profiling(vsvb_batch *batch, int32_t on);
int vsvb_batch_class_stats(vsvb_batch *batch, double *out);
/* Copy n doubles of a named internal device buffer to host memory (diagnostics and
 * tests): "S_ao", "h_ao", "C", "C_eri", "S_orb", "h_orb", "P", "R", "J", "K", "F",
 * "gradC", "scalars".  Layouts are row-major; see the implementation for shapes. */
int vsvb_debug_buffer(vsvb_context *ctx, const char *name, double *out, int64_t n);
/* Evaluate the device Boys function F_0..F_L at n arguments T (diagnostics/tests):
 * out[i*(L+1) + m] = F_m(T[i]), 0 <= L <= 8, from the full Taylor table (the
 * cooperative and one-electron kernels); L + 16 selects the compact per-L
 * table that the straight-line quartet kernels stage in LDS. */
int vsvb_debug_boys(int32_t L, const double *T, int32_t n, double *out);
/* The library's dense product on the device (diagnostics/tests), row-major:
 * C[M x N] = alpha op(A) op(B) + beta C, op(A) M x K (A stored K x M when ta),
 * op(B) K x N (B stored N x K when tb); C is read and overwritten. */
int vsvb_debug_gemm(int32_t M, int32_t N, int32_t K, int32_t ta, int32_t tb, double alpha, const double *A,
                    const double *B, double beta, double *C);
/* Library version string. */
const char *vsvb_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VSVB_HIP_H */
