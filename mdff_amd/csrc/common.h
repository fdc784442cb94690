// common.h -- shared device helpers and context layout of the gfx950 engforce library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mdff_gpu.h"

#define MDFF_PI 3.14159265358979323846264338327950

// ---------------------------------------------------------------------------
// scalar device math shared by every kernel
// ---------------------------------------------------------------------------

// Fortran NINT (round half away from zero).  x - trunc(x) is exact for
// |x| < 2^52, so the 0.5 comparisons are exact and the result equals the
// reference's nint for every double.
__host__ __device__ __forceinline__ double nint_d(double x) {
  const double t = trunc(x);
  const double r = x - t;                          // exact
  // +1 / -1 by selects (no branch): round half away from zero
  return t + ((r >= 0.5) ? 1.0 : 0.0) - ((r <= -0.5) ? 1.0 : 0.0);
}

// Numerical Recipes erfcc, as src/functions.f90:28-40 (x >= 0 only here:
// every caller passes alpha*d >= 0).
// The FP32 form takes t from the hardware reciprocal (within 1 ulp).
template <typename T>
__device__ __forceinline__ T errfc_pos(T z) {
  T t;
  if constexpr (sizeof(T) == 4) t = __builtin_amdgcn_rcpf(1.f + 0.5f * z);
  else t = T(1) / (T(1) + T(0.5) * z);
  T p = t * (T(1.48851587) + t * (T(-.82215223) + t * T(.17087277)));
  p = t * (T(-1.13520398) + p);
  p = t * (T(.27886807) + p);
  p = t * (T(-.18628806) + p);
  p = t * (T(.09678418) + p);
  p = t * (T(.37409196) + p);
  p = t * (T(1.00002368) + p);
  return t * exp(-z * z - T(1.26551223) + p);
}

// Tang-Toennies f_n and f'_n (src/tt_damp.f90:68-97); c_k = 1/k!.
template <typename T>
__device__ __forceinline__ void tt_damp(T b, T c, T r, int order, T& f, T& fd) {
  const T E[9] = {T(1), T(1), T(0.5), T(1.0 / 6.0), T(1.0 / 24.0), T(1.0 / 120.0), T(1.0 / 720.0),
                  T(1.0 / 5040.0), T(1.0 / 40320.0)};
  T br = b * r;
  T ex = exp(-br) * c;
  T ff = E[order];
  for (int k = order - 1; k >= 1; --k) ff = ff * br + E[k];
  ff = ff * br + E[0];
  f = T(1) - ff * ex;
  T brn = T(1), a = br;
  for (int k = order;;) {
    if (k & 1) brn *= a;
    k >>= 1;
    if (k == 0) break;
    a *= a;
  }
  fd = E[order] * brn * ex * b;
}

// ---------------------------------------------------------------------------
// per type-pair parameters (built on the host from the descriptor)
// ---------------------------------------------------------------------------
struct PairParam {
  double rcutsq;                          // short-range cutoff^2 (strict <)
  double A, B, C, D, BD;                  // BMHFT(D) (src/non_bonded.f90:103-107)
  double epsp, sigsq, pl, ql, uc, uc1, uc2, fc;  // LJ derived (src/non_bonded.f90:563-600)
  double epsmor, rhomor, rs;              // Morse
  double db_i, dc_i, db_j, dc_j;          // TT damping of the field at i (from j) / at j
  int32_t dk_i, dk_j, ldamp, pad;
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// kernel classes timed by the in-library event profiler
enum ProfKind { PK_PAIR = 0, PK_FIELD, PK_TABLES, PK_SK, PK_SKRED, PK_GATHER, PK_GFINAL, PK_NLIST, PK_AMAT, PK_INNER, PK_COUNT };
struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

// device sums of one force evaluation (one slot of mdff_ctx::raw)
constexpr int RAW_PA = 0;        // 15 per-atom column sums of the pair kernel
constexpr int RAW_EREC = 15;     // 7 reciprocal sums: energy, stress (xx yy zz xy xz yz)
constexpr int RAW_MOM = 22;      // musq, mutot(3), qtot(3)
constexpr int RAW_EKIN = 29;     // sum m v^2
constexpr int RAW_ENVT = 30;     // extended-system energy of a Nose-Hoover chain step
constexpr int RAW_SCRATCH = 32;  // 7 moments of the SCF field (slot 0 only)
constexpr int RAW_LJ = 40;       // 8 sums of the LJ pass beside BMHFT(D): u, vir, stress (6)
constexpr int RAW_N = 48;
constexpr int RAW_SLOTS = 1024;  // evaluations held before the sums go to the host

// A deferred fixed-order final sum: out[0..ncol) = sums over nb block
// partials part[b * ncol + c].  Queued finals run together in one launch
// (one block each) at flush_finals.
struct FinalSum {
  const double* part;
  double* out;
  int nb, ncol;
};
constexpr int kMaxFinals = 8;

// Queued final sums handed to a launch that runs them in extra blocks (one
// block each), plus the skin check's completion signal (seq_val written to
// seq_out by the first of them; the launch starts after the check kernel).
struct FinalBatch {
  FinalSum f[kMaxFinals];
  int nfin;
  double* seq_out;
  double seq_val;
};

// Workgroups of a launch are dispatched round-robin over the 8 XCDs by
// linear id.  xcd_order maps a linear id to a logical one so that each XCD
// runs one contiguous range of the logical order (the first nbk % 8 XCDs get
// one block more): neighbouring logical blocks, which read the same atoms'
// table rows, then share one XCD's L2.  A bijection of [0, nbk).
__device__ __forceinline__ int xcd_order(int bid, int nbk) {
  const int xcd = bid & 7, slot = bid >> 3, per = nbk >> 3, ext = nbk & 7;
  return (xcd < ext ? xcd * (per + 1) : ext * (per + 1) + (xcd - ext) * per) + slot;
}

// Last step of the reciprocal gather for one atom: the per-split partial
// moments part[(s * moments + c) * n + i] summed over the splits, then the
// Cartesian field and force (src/ewald_rec.f90:196-233, scaled as :300-358:
// x2 for the half mesh, x 4pi/V).  B is column-major: b[0] = B(1,1),
// b[1] = B(2,1), ...
struct RecFinal {
  int n, nsplit, moments, use_q;
  const double* part;
  const double* q;
  const double* mu;               // 3n AoS (moments == 9)
  double b[9];
  double fpi_V;
};

// v[0..moments) = sum over splits; g = 2pi B v[0..3) (the gradient of the
// charge phase sums)
__device__ __forceinline__ void rec_final_moments(const RecFinal& r, int i, double v[9], double g[3]) {
#pragma unroll
  for (int c = 0; c < 9; ++c) v[c] = 0.0;
  // splits in order, the loads of 4 splits issued before their adds
  for (int s0 = 0; s0 < r.nsplit; s0 += 4) {
    double t[4][9];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 9; ++c)
        t[u][c] = (s0 + u < r.nsplit && c < r.moments) ? r.part[((size_t)(s0 + u) * r.moments + c) * r.n + i] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 9; ++c)
        if (s0 + u < r.nsplit && c < r.moments) v[c] += t[u][c];
  }
  g[0] = r.b[0] * v[0] + r.b[3] * v[1] + r.b[6] * v[2];
  g[1] = r.b[1] * v[0] + r.b[4] * v[1] + r.b[7] * v[2];
  g[2] = r.b[2] * v[0] + r.b[5] * v[1] + r.b[8] * v[2];
}

__device__ __forceinline__ void rec_final_field(const RecFinal& r, const double g[3], double ef[3]) {
  const double tp = 2.0 * MDFF_PI, sc = 2.0 * r.fpi_V;
  for (int a = 0; a < 3; ++a) ef[a] = sc * tp * g[a];
}

__device__ __forceinline__ void rec_final_force(const RecFinal& r, int i, const double v[9], const double g[3],
                                                double f[3]) {
  const double tp = 2.0 * MDFF_PI, sc = 2.0 * r.fpi_V;
  const double* b = r.b;
  const double qi = r.use_q ? r.q[i] : 0.0;
  double fx = qi * tp * g[0], fy = qi * tp * g[1], fz = qi * tp * g[2];
  if (r.moments == 9) {
    const double u0 = r.mu[3 * i], u1 = r.mu[3 * i + 1], u2 = r.mu[3 * i + 2];
    // m = B^T mu; p = (symmetric second moments) m
    const double m0 = u0 * b[0] + u1 * b[1] + u2 * b[2];
    const double m1 = u0 * b[3] + u1 * b[4] + u2 * b[5];
    const double m2 = u0 * b[6] + u1 * b[7] + u2 * b[8];
    const double p0 = v[3] * m0 + v[6] * m1 + v[7] * m2;
    const double p1 = v[6] * m0 + v[4] * m1 + v[8] * m2;
    const double p2 = v[7] * m0 + v[8] * m1 + v[5] * m2;
    const double t2 = tp * tp;
    fx += t2 * (b[0] * p0 + b[3] * p1 + b[6] * p2);
    fy += t2 * (b[1] * p0 + b[4] * p1 + b[7] * p2);
    fz += t2 * (b[2] * p0 + b[5] * p1 + b[8] * p2);
  }
  f[0] = sc * fx;
  f[1] = sc * fy;
  f[2] = sc * fz;
}

// one atom of the pair kernels, in cell-sorted order: fractional coordinates
// and charge (two 16-byte loads), plus its type and original index
struct alignas(16) AtomRec {
  double x, y, z, q;
};

struct mdff_ctx {
  mdff_desc d;
  int n = 0, ntype = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[12] = {};
  // pair kernel overlapped with the reciprocal kernels (charges, valid
  // lists): it runs on stream2 after the phase-table launch (ov_tab) and the
  // force sums wait for it (ov_pair)
  int overlap = 0;
  hipStream_t stream2 = nullptr;
  hipEvent_t ov_tab = nullptr, ov_pair = nullptr;
  std::string err;

  // cell
  double A[9], B[9], omega, W[3];
  double rc_vdw, rc_coul, alpha;
  int task_qq = 0, task_qd = 0, task_dd = 0, linduced = 0, ldip_static = 0;
  int sr_model = 0;                       // 0 none, 1 LJ, 2 BMH, 3 Morse
  // reciprocal kernels in the FP32-class form (split-half / FP32 tables): the
  // mixed form.  A dipole system keeps them only where the FP64 inner k-box
  // of its force evaluation exists (else, or with MDFF_DIP_RECIP_FP64=1, its
  // reciprocal sums run in FP64): without that box the FP32-class dipole
  // structure factors set the force error above 1e-5
  int rec_f32 = 0;
  // LJ together with BMHFT(D) (src/engforce_driver.f90:130-141): the BMHFT(D)
  // pass gives the forces (each reference routine zeroes them first), an LJ
  // pass its energy, virial and stress into peratom_lj (forces into f_lj,
  // unused), summed into raw[RAW_LJ..]
  int sr_lj_extra = 0;
  double* peratom_lj = nullptr;
  double* f_lj = nullptr;
  double* lj_part = nullptr;
  double u_tail = 0, p_tail = 0;
  std::vector<PairParam> pp;              // ntype x ntype, (it + ntype*jt)

  // per-atom device arrays (SoA FP64 unless noted)
  double *rx = nullptr, *ry = nullptr, *rz = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *fsx = nullptr, *fsy = nullptr, *fsz = nullptr;   // saved forces (VV)
  double *sx = nullptr, *sy = nullptr, *sz = nullptr;      // fractional coordinates
  double *xs = nullptr, *ys = nullptr, *zs = nullptr;      // saved fractional (vnlistcheck)
  double *bx = nullptr, *by = nullptr, *bz = nullptr;      // fractional at the last build
  int32_t* itype = nullptr;               // 0-based
  double* q = nullptr;
  double* mass = nullptr;
  double* pol = nullptr;                  // 9 per atom (column-major 3x3), 0 if not polar
  double* invpol = nullptr;
  double* dipia = nullptr;                // static dipoles, 3 per atom
  double* mu = nullptr;                   // total dipoles (3n AoS)
  double* mu_ind = nullptr;
  double* efield = nullptr;               // output field ef_t (3n)
  double* e_stat = nullptr;               // SCF static field (3n)
  double* e_tot = nullptr;                // SCF total field (3n)
  double* e_ind = nullptr;                // SCF induced field (3n)
  double* hist = nullptr;                 // ASPC history 5 x 3n
  double* frec = nullptr;                 // reciprocal force scratch (3n)
  double* peratom = nullptr;              // per-atom energy/virial/stress partials (NPA x n)
  double* red = nullptr;                  // reduction outputs
  int mu_static = 0;                      // mu holds the static dipoles (no induced dipoles)
  double *cx = nullptr, *cy = nullptr, *cz = nullptr;   // fractional coordinates of a pending list check
  double* h_chk = nullptr;                // mapped pinned host memory: block maxima of the pending check
  double* d_chk = nullptr;                // the device address of h_chk
  hipEvent_t chk_ev = nullptr;            // the pending check's maxima are in h_chk
  int chk_pending = 0;
  // instead of the event, the next final-sum launch writes chk_seq into
  // h_chk[2 nb] once the check kernel has completed (1: to be attached to
  // that launch, 2: attached); the host waits for the value
  int chk_seq_wait = 0;
  double chk_seq = 0.0;
  double* red_part = nullptr;             // reduction block partials
  double* fin_part = nullptr;             // block partials of the pair sums [0, 1024*16) and kinetic energy
  double* nhc_part = nullptr;             // Nose-Hoover chain steps: two sets of kinetic block partials
  double* nhc_st = nullptr;               // Nose-Hoover chain state (xi, vxi, Q, scale)
  double* rec_blk = nullptr;              // k-block partials of the reciprocal energy and stress
  FinalSum fin[kMaxFinals] = {};          // queued final sums
  RecFinal rfin = {};                     // reciprocal forces left to the force kernel
  int rfin_pending = 0;
  int nfin = 0;
  int efield_zero = 0;                    // efield holds zeros (no field, dipoles or surface term)
  PairParam* d_pp = nullptr;
  int itime = 0;
  mdff_terms last_terms = {};             // terms of the last force evaluation (engforce or MD step)
  int32_t* polar = nullptr;               // 1 for polarisable ions
  int npol = 0;                           // number of polarisable ions
  double qsq = 0.0;                       // sum of q^2 (self term)
  double qmax = 0.0;                      // max |q|
  double* zeros = nullptr;                // n zeros (charges switched off)

  // neighbour list (full rows, cut = max(rc_vdw, rc_coul) + skin)
  int32_t* cell_of = nullptr;             // cell index per atom
  int32_t* cell_key = nullptr;            // sorted keys
  int32_t* cell_atoms = nullptr;          // atoms sorted by cell
  int32_t* cell_atoms_in = nullptr;
  int32_t* cell_start = nullptr;          // ncell + 1
  int32_t* nl_count = nullptr;            // n (+1): row lengths
  int32_t* nl_start = nullptr;            // n + 1: row starts
  int32_t* nl_j = nullptr;                // capacity entries
  uint8_t* nl_flag = nullptr;             // bit0: r2 <= (rc_vdw+skin)^2, bit1: <= (rc_coul+skin)^2,
                                          // bit2: r2 <= (rc_inner+skin)^2 (inner entry, mixed form)
  int64_t nl_cap = 0;
  // Mixed form: a row holds its inner entries (within rc_inner + skin at the
  // build) first, nl_count_in of them, then the rest.  The inner entries run
  // through the FP64 pair kernel, the rest through the FP32 one: the FP32
  // rounding of the large forces of close pairs would otherwise dominate the
  // force error.  The sweep writes the rows into nl_j_tmp / nl_flag_tmp and a
  // partition pass reorders them into nl_j / nl_flag.  rc_inner = 0: no split.
  double rc_inner = 0.0;
  int32_t* nl_count_in = nullptr;         // n
  int32_t* nl_j_tmp = nullptr;            // capacity entries (split only)
  uint8_t* nl_flag_tmp = nullptr;
  int nl_row_w = 0;                       // fixed row width of the one-pass build (0: rows from a count)
  int ncell[3] = {1, 1, 1};
  int nb_lo[3] = {0, 0, 0}, nb_hi[3] = {0, 0, 0}, nb_all[3] = {1, 1, 1};  // neighbour cell offsets
  float4* bf = nullptr;                   // wrapped single-precision copy of (bx, by, bz)
  int ncell_alloc = 0;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int lists_valid = 0;
  int32_t* d_flag = nullptr;              // device flags: [0] rebuild, [1] overflow
  int32_t* h_flag = nullptr;              // pinned host mirror
  double* h_red = nullptr;                // pinned host reduction mirror
  double* raw = nullptr;                  // device sums of evaluations: RAW_SLOTS x RAW_N
  double* h_raw = nullptr;                // pinned host mirror of raw
  int timings_pending = 0;

  // reciprocal space
  int K[3] = {0, 0, 0};
  int NXk = 0, NYk = 0, NZk = 0;          // box dims: Kx+1, 2Ky+1, 2Kz+1
  int nk_box = 0, nk = 0;
  double* kA = nullptr;                   // A(n) on the box, 0 where excluded (nk_box)
  double* kvec = nullptr;                 // kx,ky,kz,kcoe per box entry (4 x nk_box)
  void* tab = nullptr;                    // per-atom phase tables (precision dependent)
  void* sk_part = nullptr;                // structure-factor partials
  double* sk = nullptr;                   // S(n) (2 x nk_box) FP64
  double* wk = nullptr;                   // W(n) = A(n) conj(S(n)) (2 x nk_box)
  void* wk_t = nullptr;                   // W in the working precision
  double* rec_part = nullptr;             // per (atom, split) partial moments
  int rec_splits = 1, sk_chunks = 1;
  void* amat = nullptr;                   // matrix-core gather A operand
  int mma_nks = 0, mma_mtiles = 0, mma_ct = 0, mma_splits = 1;
  int use_mma = 1;                        // matrix-core reciprocal kernels
  int sk_ctm = 1;                         // m-tiles per wave of the structure factor
  int sk_ta = 0;                          // atoms per staged tile of the structure factor
  int sk_mma_ch = 1;                      // atom chunks of the matrix-core structure factor
  // mixed form: split-half (FP16 hi + lo) matrix-core reciprocal kernels
  int h3 = 0;                             // split-half gather (and its A operand from the reduction)
  int h3_sk = 0;                          // split-half structure factor as well
  int h3_wide = 0;                        // ... in its wide-mesh form (more than 64 staged X / Y columns)
  int h3_ch = 1, h3_nsh = 0, h3_ct = 1, h3_splits = 1, h3_mt8 = 0;
  float h3_pscale = 1.f;                  // charge scale of the structure factor's B operand
  float* mu_pscale = nullptr;             // device: scales of the two dipole passes (k_mu_pscale)
  // Mixed form, charge-only evaluations: the k-vectors of the inner box
  // |nx| <= kin[0], |ny| <= kin[1], |nz| <= kin[2] (every k with |k| <=
  // 2.5 alpha) carry most of the weight of the force sums and run in FP64
  // (k_in_phase, k_in_sk_s, k_in_reduce_r, k_in_gather_s, or the generic
  // k_in_sk, k_in_reduce, k_in_gather); the main kernels see A = 0 there
  // (kA_mask).  kin_on = 0: no inner pass.
  int kin_on = 0;
  int kin[3] = {0, 0, 0};
  int nb_in = 0, in_chunk = 0, in_nch = 0, kin_tile = 0, kin_msplit = 0, kin_splits = 0;
  int kin_reg = 0;                        // 1: the scalar-operand forms (Kz' small enough)
  int kin_ts = 0, kin_sg = 0;             // k_in_sk_s: threads per subgroup, subgroups
  int kin_jb = 1, kin_kb = 0;             // scalar-operand forms: blocks of j = |nz|, KZ (j per block - 1)
  int kin_zpad = 0;                       // j columns of the phase rows and the W table: kin_jb (kin_kb + 1)
  size_t kin_lds = 0;                     // LDS bytes of the structure-factor kernel
  void* tab_in = nullptr;                 // per atom X, Y, q Z rows (FP64 complex)
  double* wt_in = nullptr;                // per (m, j) sum / difference table of W
  double* kA_mask = nullptr;              // kA with the inner box zeroed (nk_box)
  double* kA_in = nullptr;                // A of the inner box, 0 outside the half space (nb_in)
  double* kvec_in = nullptr;              // kx, ky, kz, kcoe of the inner box (4 x nb_in)
  double* sk_in_part = nullptr;           // [chunk][2 nb_in] FP64 partials
  double* wk_in = nullptr;                // W = A conj(S) of the inner box (2 x nb_in)
  double* dip_part = nullptr;             // dipole systems: [q, mu_x, mu_y, mu_z][chunk][2 nb_in]
  void* amat_h3 = nullptr;                // gather A operand, split halves
  float* ascale_h3 = nullptr;             // inverse scale per gather m-tile
  double* kO_ws = nullptr;                // scf_kO_v2 work vectors (11 x 3n)
  AtomRec* arec = nullptr;                // atoms in cell-sorted order (refreshed every evaluation)
  int2* ameta = nullptr;                  // (type, original index) in the same order

  // electric field gradient efg_t (9 x n, column-major per atom) and its
  // work arrays: real-space part (6 x n), S(k) / reciprocal partials, W = A S
  double* efg = nullptr;
  double* efg_dir = nullptr;
  double* efg_part = nullptr;
  double* efg_w = nullptr;
  int efg_sk_nch = 1, efg_sk_chunk = 1, efg_rec_nch = 1, efg_rec_chunk = 1;

  double last_ms[8] = {0};
  int n_last = 0;

  // event profiler (mdff_profile): one event pair per timed launch
  bool prof_on = false;
  std::vector<ProfRec> prof;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
};

// Record a launch between two events on the context stream when profiling.
hipEvent_t prof_event(mdff_ctx* c);
#define PROF_LAUNCH(c, kind, ...)                                   \
  do {                                                              \
    if ((c)->prof_on) {                                             \
      ProfRec _r{(kind), prof_event(c), prof_event(c)};             \
      (void)hipEventRecord(_r.a, (c)->stream);                      \
      __VA_ARGS__;                                                  \
      (void)hipEventRecord(_r.b, (c)->stream);                      \
      (c)->prof.push_back(_r);                                      \
    } else {                                                        \
      __VA_ARGS__;                                                  \
    }                                                               \
  } while (0)

// ---------------------------------------------------------------------------
// error helpers
// ---------------------------------------------------------------------------
#define HIP_TRY(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
      return MDFF_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)

#define MDFF_TRY(call)                    \
  do {                                    \
    int _rc = (call);                     \
    if (_rc != MDFF_OK) return _rc;       \
  } while (0)

static inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// launches implemented in the .hip files
int nl_build(mdff_ctx* c);
int nl_check(mdff_ctx* c, int* rebuilt);
int nl_check_launch(mdff_ctx* c);
int nl_check_finish(mdff_ctx* c, int* rebuilt);
// velocity Verlet: position update; step end + list check (+ the next
// step's position update)
int vv_pos_launch(mdff_ctx* c, double dt);
int vv_end_launch(mdff_ctx* c, double dt, bool next, double* ke_part);
int reduce_cols_dev(mdff_ctx* c, const double* in, int ncol, double* dst);
int reduce_cols_part(mdff_ctx* c, const double* in, int ncol, double* part, int* nb);
int queue_final(mdff_ctx* c, const double* part, int nb, int ncol, double* out);
int flush_finals(mdff_ctx* c);
int claim_partials(mdff_ctx* c, const double* part);
// the queued finals (and a pending skin-check signal) as a batch for the next
// launch, which must run them; the queue is left empty
FinalBatch take_finals(mdff_ctx* c);
int nl_get_pairs(mdff_ctx* c, int which, int32_t* point, int32_t* list, int64_t cap, int64_t* n);
void launch_kardir(mdff_ctx* c);
void pack_atoms(mdff_ctx* c);
int realspace_field(mdff_ctx* c, int mode, const double* mu_in, double* ef_out, double alpha = 0.0);
int recip_setup(mdff_ctx* c);
int recip_tables(mdff_ctx* c);
// kardir + tables in one launch; pack: also pack_atoms (in the current cell
// order)
int recip_tables_from_r(mdff_ctx* c, bool pack);
// forces with a null f_out: the split partials stay in rec_part and c->rfin
// describes them, for a later kernel that finishes the forces itself
int recip_eval(mdff_ctx* c, const double* q, const double* mu, bool use_q, bool use_mu,
               bool forces, bool field, double* ef_out, double* f_out, double* e_out);
void recip_free(mdff_ctx* c);
// electric field gradient (efg.hip; its real-space part in realspace.hip)
int realspace_efg(mdff_ctx* c, double* out);
int efg_alloc(mdff_ctx* c);
int efg_eval(mdff_ctx* c);
int realspace_lj_extra(mdff_ctx* c, double* f_scratch, double* peratom);
void efg_free(mdff_ctx* c);
