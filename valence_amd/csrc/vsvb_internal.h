// Internal declarations shared by the translation units of libvsvb_hip.so:
// error reporting, device buffers, the per-molecule context and the host
// routines of the energy pipeline that more than one unit calls.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vsvb_hip.h"
#include "vsvb_eri_kernels.h"

namespace vsvb {

// Largest exchange LDS block of a J/K workgroup (nbf x ket columns doubles).
constexpr size_t kMaxLdsK = 60 * 1024;
// Device counter words of a quartet stage: per class {evaluated, primitive
// quartets, skipped, spare}, then per class {earliest start, latest end} stamps.
constexpr int kCounterWords = 6 * kNQuartetClass;

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

#define HIP_OK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess)                                                                              \
      return ::vsvb::fail(VSVB_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));         \
  } while (0)

struct PPSrc {
  int sA, sB, ia, ib;  // shell indices of A and B, global primitive indices of A and B
};

struct OrbEntry {
  int orb;   // owning orbital
  int kind;  // 0: AO, 1: NDF reference
  int tgt;   // AO index (kind 0) or NDF orbital index (kind 1)
  int pad;
};

// A bump region of device memory that buffers are carved from instead of
// allocated (the allocation-free entry point: the library's static scratch
// array).  need counts every request, including those that did not fit.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0, need = 0;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool own = true;
  hipError_t alloc(size_t count) {
    n = count;
    own = true;
    return hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  // Take the next 256-byte-aligned slice of the arena (nothing is allocated or
  // freed; hipErrorOutOfMemory when the arena is full).
  hipError_t carve(Arena& a, size_t count) {
    const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    a.need += bytes;
    if (a.used + bytes > a.cap) return hipErrorOutOfMemory;
    p = reinterpret_cast<T*>(a.base + a.used);
    a.used += bytes;
    n = count;
    own = false;
    return hipSuccess;
  }
  // Point at memory owned elsewhere (a batch group shares geometry data and
  // slices of group-wide density / Coulomb / exchange arrays).
  void alias(T* q, size_t count) {
    if (p && own) (void)hipFree(p);
    p = q;
    n = count;
    own = false;
  }
  ~DevBuf() {
    if (p && own) (void)hipFree(p);
  }
};


// Spin-coupled determinant expansion (density_sc / dbra / dket,
// src/valence.F90:1612-1870): determinant d = isc * 2^npair + mask has the
// alpha list (pair slots, swapped where mask has the bit; unpaired; docc) and
// the beta list (the partners; docc), as spatial-orbital indices, and weight
// coeff_sc[isc].  Work arrays are per ordered determinant pair.
struct SpinCoupled {
  int ndet = 0, na = 0, nb = 0, npq = 0;
  int nspinc = 0, npair = 0;             // determinant d belongs to spin coupling d >> npair
  bool want_mats = false;                // the next pass also accumulates the spin-coupling matrices
  long npairs = 0;
  DevBuf<double> d_ham, d_ovl;           // [nspinc][nspinc] lower triangles (bra coupling >= ket coupling)
  DevBuf<int> d_alist, d_blist;          // [ndet][na], [ndet][nb]
  DevBuf<int> d_pq;                      // [npq][2] unique orbital pairs (r >= s)
  DevBuf<double> d_cdet;                 // [ndet]
  DevBuf<double> d_pdens, d_pj;          // [npq][nbf][nbf] pair densities / their Coulomb matrices
  DevBuf<double> d_W;                    // [npq][nbf][nsp]  J^{rs} C^T
  DevBuf<double> d_eri, d_erik;          // [nsp^2][nsp^2]   (pq|rs) and its exchange ordering
  DevBuf<double> d_Ga, d_Gb;             // [npairs][na^2], [npairs][nb^2] inverses
  DevBuf<double> d_Xa, d_Xb, d_Xt;       // [npairs][nsp^2] transition densities
  DevBuf<double> d_JX, d_KXa, d_KXb;     // [npairs][nsp^2]
  DevBuf<double> d_w, d_num, d_phi, d_e1w, d_wv;  // [npairs] omega, omega*Phi, Phi, omega*e1, omega/N
  DevBuf<double> d_Shat, d_hhat, d_hsym;          // [nsp^2] dE/dS_orb, dE/dh_orb, symmetrized copy
  DevBuf<double> d_D2, d_Dk;                      // [nsp^2][nsp^2] dE/d(pq|rs), exchange outer product
  DevBuf<double> d_Dh, d_Xw;                      // [nsp^2][npq], [npairs][nsp^2] adjoint work
  DevBuf<double> d_Y, d_JY, d_KY;                 // [npairs][2][3][nsp^2] split cofactor matrices and images
  DevBuf<double> d_q4;                            // [npairs][2][4] sign, P_R, d1, d2
};

// Launch configuration of one quartet class in one mode (energy or J/K).
struct ClassLaunch {
  bool coop = false;  // cooperative kernel (one wavefront per contracted quartet)
  int units = 0;      // work units handed out by the class counter
  int blocks = 0;     // workgroups
  size_t lds = 0;     // straight-line J/K: exchange staging block (bytes, 0 = global atomics)
  int claim = 1;      // cooperative kernel: units per claim
  int per_cu = 0;     // workgroups per CU requested (0 = the kernel's occupancy)
  bool tr = false;    // straight-line J/K: transposed units (one bra pair x 256 ket pairs)
  int strips = 1;     // straight-line: energy segments, J/K bra strips (ket strips transposed), one work counter each
  int kchunk = 1;     // straight-line: pairs of the uniform side (ket; bra when transposed) per claim
  int seg_tiles = 1;  // straight-line energy: bra tiles per counter segment
  bool overflow = false;  // the class has more work units than 32-bit indices allow in this shape
};

}  // namespace vsvb

struct vsvb_context {
  using SpinCoupled = vsvb::SpinCoupled;
  template <typename T> using DevBuf = vsvb::DevBuf<T>;
  using Shell = vsvb::Shell;
  using Pair = vsvb::Pair;
  using PrimPair = vsvb::PrimPair;
  using PPSrc = vsvb::PPSrc;
  using OrbEntry = vsvb::OrbEntry;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ev_valid = false;  // ev0 / ev1 bracket the last quartet pass (a direct run_quartets)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against a caller's stream (device entry)
  double tau = 1e-17;
  int pair_chunk = 0;  // straight-line units: uniform-side pairs per claim (0: from the system size)
  int grid_inverse_min = 129;  // overlap blocks from this many rows up are inverted grid-wide
  int wall_khz = 0;            // device wall-clock rate (kHz) of the kernels' timing stamps
  vsvb::Arena* arena = nullptr;  // device buffers carved from this region instead of allocated (no events,
                                 // the null stream): the allocation-free entry point
  // sizes
  int natom = 0, nsh = 0, nbf = 0, npairs = 0, npp = 0, nsp = 0, norbs = 0, totlen = 0;
  int npair = 0, nunpd = 0, ndocc = 0, ndf = 0, first_ndf = 0;
  int nalpha_rows = 0;  // spatial orbitals with an alpha electron: rows [0, nalpha_rows) of C (nsp, or
                        // nsp - 1 when the last doubly occupied orbital holds only a beta electron)
  double dtol = 0.0;
  int pair_off[vsvb::kNPairClass + 1] = {0};
  // host copies
  std::vector<double> coords, charges, coeff;
  // host copies for the per-geometry pair schedule
  std::vector<Shell> h_shells;
  std::vector<Pair> h_pairs;
  std::vector<double> h_pexp, h_pcoef, sched_coords;
  std::vector<OrbEntry> ent;
  std::vector<int> orb_start;
  // device
  DevBuf<Shell> d_shells;
  DevBuf<Pair> d_pairs;
  DevBuf<PPSrc> d_ppsrc;
  DevBuf<PrimPair> d_pp;
  DevBuf<double> d_pexp, d_pcoef, d_xyz, d_charge, d_boys, d_angn;
  DevBuf<unsigned long long> d_stamps;  // per-class wall-clock slots of the last stage (profiling)
  DevBuf<int> d_ao_shell, d_orb_start, d_status;
  DevBuf<double> d_colk;  // pivot column of the large-matrix inverse
  DevBuf<OrbEntry> d_ent;
  DevBuf<double> d_ndf_d, d_ndf_g;  // NDF gradient scratch: [ndf][nbf] each
  DevBuf<double> d_coeff, d_ndf_scale, d_S, d_H, d_Ct, d_Cs, d_Sorb, d_Horb, d_P, d_R, d_tmp, d_tmp2;
  DevBuf<double> d_partial, d_scal, d_J, d_K, d_F, d_gradC, d_grad, d_Jsym, d_Rfull, d_X, d_Z;
  DevBuf<unsigned long long> d_counters;
  DevBuf<int> d_tilectr;  // work counters: class q owns [ctr_off[q], ctr_off[q] + ctr_len[q])
  int ctr_off[vsvb::kNQuartetClass] = {0}, ctr_len[vsvb::kNQuartetClass] = {0};
  int ctr_total = 0;

  size_t partial_per_class[vsvb::kNQuartetClass] = {0};
  // launch configuration per quartet class: energy mode and J/K mode
  vsvb::ClassLaunch le[vsvb::kNQuartetClass], ljk[vsvb::kNQuartetClass];
  int coop_idx[vsvb::kNQuartetClass] = {0};  // cooperative step tables of the class, or -1
  int ncu = 0;
  size_t partial_total = 0;
  bool geom_ready = false;
  double enuc = 0.0;
  double last_counters[4] = {0, 0, 0, 0};
  // per-class profiling (events around each class launch when enabled)
  bool profile = false;
  bool last_jk = false;
  int last_nwfn = 1, last_nspin = 1;  // densities and spin blocks of the last quartet launch
  bool ran_quartets = false;  // this context launched the quartet stage in the current call
  int class_rev = 0;          // quartet classes launched in reverse class order (odd batch lanes)
  bool capturing = false;     // enqueueing into a HIP graph capture: no host uploads
  double class_stats[vsvb::kNQuartetClass][6] = {};
  // spin-coupled wave functions (npair > 0): determinant lists and work arrays
  SpinCoupled sc;
};

namespace vsvb {

// host routines of the pipeline (vsvb_core.hip)
hipError_t gemm(hipStream_t st, int M, int N, int K, double alpha, const double* A, int lda, bool ta,
                const double* B, int ldb, bool tb, double beta, double* C, int ldc);
hipError_t launch_fill(hipStream_t st, double* p, size_t n, double v);
hipError_t launch_zero(hipStream_t st, void* p, size_t bytes);  // bytes: a multiple of 4
hipError_t launch_reduce(hipStream_t st, const double* partial, long n, int nw, double* out);
hipError_t launch_gather_grad(vsvb_context* c, const double* gradC, double* grad);
int prepare_geometry(vsvb_context* c, const double* coords);
int prepare_orbital_basis(vsvb_context* c, const double* coeff_host, const double* coeff_dev);
int run_quartets(vsvb_context* c, bool jk, double* e2_dev, int nwfn = 1, const double* dens = nullptr,
                 double* jacc = nullptr, double* kacc = nullptr, int nspin_override = -1);
int finish_counters(vsvb_context* c, const unsigned long long* host_counters = nullptr);
int check_status(vsvb_context* c, bool for_grad);

// cooperative kernel for the register-heavy classes (vsvb_eri_coop.hip)
int coop_class(int la, int lb, int lc, int ld);  // index of the class, or -1
int coop_occupancy(int idx, bool jk);
bool coop_default(int idx);
hipError_t coop_prepare(int device);              // step tables resident on the device
hipError_t launch_coop(int idx, const QuartetArgs& a, bool jk, int nblocks, int device, hipStream_t st);
// J/K mode over ng groups' arguments (device array gs) in one launch; stamp: the
// launch's wall-clock slots or null
hipError_t launch_coop_groups(int idx, const QuartetArgs* gs, int ng, int nblocks, int device,
                              unsigned long long* stamp, hipStream_t st);

// spin-coupled path (vsvb_sc.hip)
int sc_setup(vsvb_context* c, const vsvb_system* s);
int sc_enqueue(vsvb_context* c, bool want_grad, double* grad_dev);
int sc_finish(vsvb_context* c, double* energy, double* parts, double* grad_host);
// Spin-coupling weights coeff_sc[nspinc] into the per-determinant weights.
int sc_set_coupling(vsvb_context* c, const double* coeff_sc);
// Spin-coupling matrices (electronic Hamiltonian numerator and overlap,
// [nspinc][nspinc] symmetric) at the context's geometry and orbitals.
int sc_matrices(vsvb_context* c, double* ham, double* ovl);

}  // namespace vsvb
