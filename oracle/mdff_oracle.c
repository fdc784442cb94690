/*
 * mdff_oracle.c -- CPU restatement of MDFF's engforce path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker.
 * Nothing in the product library links or calls it.
 *
 * Reference: FilipeVasconcelos/mdff, Fortran 90 (all citations are paths in
 * that tree).  Arithmetic is FP64 in the reference's operation order
 * (Fortran evaluates a+b+c as (a+b)+c); build with -ffp-contract=off.
 *
 * Deliberate differences (each also stated where it applies):
 *  - positions are not round-tripped Cartesian->fractional->Cartesian in place
 *    (src/lattice.f90:183-264 called around every kernel); the round trip only
 *    perturbs coordinates at the last ULP between calls.
 *  - a Coulomb-only engforce starts from zero forces (the reference only zeroes
 *    in the short-range routines, src/non_bonded.f90:1104-1106, so Coulomb-only
 *    runs accumulate forces across calls).
 *  - Morse uses the analytic derivative of the reference energy
 *    (src/non_bonded.f90:1322-1325); the reference force (:1335) reads an
 *    uninitialised variable, and engforce never calls it.
 *  - LJ together with BMHFT(D) follows the reference: both energies, virials
 *    and stresses (tau = the sum, tau_nonb = the BMHFT(D) one), the forces of
 *    BMHFT(D) alone, since each routine zeroes the forces first
 *    (src/engforce_driver.f90:130-141, src/non_bonded.f90:699-701,1104-1106).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  errfc           - bit-exact against the reference's own errfc compiled from
 *                    src/functions.f90 into oracle/_ref, and against the
 *                    published NR bound |erfc-errfc| < 1.2e-7 (math.erfc).
 *  TT damping      - against the reference tt_damp.f90 compiled into oracle/_ref,
 *                    and against the closed form 1 - c e^{-br} sum (br)^k/k!.
 *  lattice         - B.A = I, omega = det A, widths = omega/|b x c|.
 *  vnlist_pbc      - brute-force pair enumeration in numpy with the ownership
 *                    rule; half list + reversed = full set within rc+skin.
 *  nmlj            - CCD LJ cluster energies (examples/ex02/config/REFERENCE)
 *                    and the 4000-atom FCC sample run (doc/mdff.tex:303-421).
 *  bmhftd          - one isolated pair against A e^{-Br} - f6 C/r^6 - f8 D/r^8
 *                    written from the TT series (energy, force, virial,
 *                    stress), the A = 0 skip and the strict cutoff; F = -dU/dr
 *                    and stress = dU/dstrain by central finite differences.
 *  morse           - harmonic period of a pair near its minimum; FD forces.
 *  kmesh           - closed-form nk (src/coulomb.f90:472); k = 2 pi B n.
 *  ewald q-q, mu-mu- Nymand & Linse values (examples/ex04/run_ex4.sh:145,179);
 *                    NaCl Madelung constant; alpha independence of the total;
 *                    FD forces and stress.
 *  efg             - Nymand & Linse V1xx, V1yy (examples/ex04/run_ex4.sh:145,
 *                    179); upper triangle = dE_a/dr_b of the (pinned) field by
 *                    central differences in a cubic cell; alpha independence;
 *                    zero trace for a neutral system.  The lower triangle
 *                    (real space + self only, src/ewald_rec.f90:209-216 fills
 *                    the upper one) and the real-space mu-mu term only between
 *                    two non-zero dipoles (src/ewald_dir.f90:557) follow the
 *                    reference lines; no data of the reference pins them.
 *  scf             - converged mu against a dense solve (alpha^-1 - T) mu = E;
 *                    F = -dU/dr of the converged PIM energy (damped terms).
 *  damped q-mu     - an isolated charge + polarisable ion: mu -> pol q f_k(br)/r^2
 *                    after extrapolating three box sizes (1/L^3 image trend).
 *  aspc            - coefficients equal Kolafa's closed form (J. Comput. Chem.
 *                    25, 335 (2004)), linear histories predicted exactly; the
 *                    MD history rules against a numpy restatement step by step.
 *  integrators     - NVE energy conservation; harmonic closed form; NHC of two
 *                    conserves the extended energy; longer chains equal the
 *                    MTK half step with the reference's zeroed middle forces;
 *                    Yoshida weights: sum w = 1, sum w^3 = 0.
 *  csvr            - mean and variance of the new kinetic energy over sampled
 *                    noise (Bussi et al. 2007, eq. A7); COM removal.
 */
#include "mdff_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define M3(m, i, j) (m)[(i) + 3 * (j)]     /* 0-based (i,j) of a column-major 3x3 */
#define T2I(it, jt) ((it) + 16 * (jt))     /* 0-based type-pair index             */
#define T3I(k, i, j) ((k) + 16 * (i) + 256 * (j))

static const double PI = 3.14159265358979323846264338327950;
static const double VNLMAX = 2000.0;       /* src/config.f90:43 */

struct orc_ctx {
  mdff_desc d;
  mdff_terms last_terms;
  int n;
  int32_t *itype;                /* 1-based */
  double A[9], B[9], omega, W[3];
  double rc_vdw, rc_coul, alpha;
  int task_qq, task_qd, task_dd; /* task_coul(1:3) */
  int linduced;
  /* per atom */
  double *q, *mass, *pol, *invpol, *dipia;
  double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz;
  double *xs, *ys, *zs;
  /* non-bonded derived tables (src/non_bonded.f90:455-644) */
  double rcutsq[256], epsp[256], sigsq[256], uc[256], uc1[256], uc2[256], fc[256], rs[256];
  double u_tail, p_tail;
  double E_TT[9];
  /* k mesh */
  int nk;
  double *kx, *ky, *kz, *kk, *Ak, *kcoe;
  /* lists (1-based content, reference layout) */
  int32_t *pv, *lv, *pc, *lc;
  int64_t nv, nc;
  int lists_valid;
  /* SCF state */
  double *hist;                  /* dipia_ind_t(5,3,natm): hist[k*3n + 3*ia + a] */
  int itime;
  double *mu_t, *ef_t;
  double *efg_t;                 /* efg(3,3,natm): efg_t[9*ia + a + 3*b] */
  double tau[9];
  int nthreads;
  char err[256];
};

/* ------------------------------------------------------------------------ */
/* scalar helpers                                                            */
/* ------------------------------------------------------------------------ */

/* Fortran NINT: round half away from zero.  x - trunc(x) is exact for
 * |x| < 2^52, so the comparison against 0.5 is exact. */
double orc_nint(double x) {
  double t = trunc(x);
  double r = x - t;
  if (r >= 0.5) t += 1.0;
  else if (r <= -0.5) t -= 1.0;
  return t;
}

/* src/functions.f90:28-40 (Numerical Recipes erfcc). */
double orc_errfc(double x) {
  double z = fabs(x);
  double t = 1.0 / (1.0 + 0.5 * z);
  double p = t * (1.48851587 + t * (-.82215223 + t * .17087277));
  p = t * (-1.13520398 + p);
  p = t * (.27886807 + p);
  p = t * (-.18628806 + p);
  p = t * (.09678418 + p);
  p = t * (.37409196 + p);
  p = t * (1.00002368 + p);
  double r = t * exp(-z * z - 1.26551223 + p);
  if (x < 0.0) r = 2.0 - r;
  return r;
}

/* E_TT(k) = 1/k! (src/tt_damp.f90:42-58) */
static void tt_coeffs(double E[9]) {
  E[0] = 1.0;
  for (int k = 1; k <= 8; ++k) E[k] = E[k - 1] / (double)k;
}

/* src/tt_damp.f90:68-97 */
void orc_tt_damp(double b, double c, double r, int order, double *f, double *fd) {
  double E[9];
  tt_coeffs(E);
  double br = b * r;
  double expbdr = exp(-br) * c;
  double ff = E[order];
  for (int k = order - 1; k >= 1; --k) ff = ff * br + E[k];
  ff = ff * br + E[0];
  *f = 1.0 - ff * expbdr;
  /* (br)**order by square-and-multiply, the integer-power algorithm of the
   * Fortran runtime (compiler-rt __powidf2) */
  double brn = 1.0, a = br;
  for (int k = order;;) {
    if (k & 1) brn *= a;
    k /= 2;
    if (k == 0) break;
    a *= a;
  }
  *fd = E[order] * brn * expbdr * b;
}

/* src/lattice.f90:60-112 (reciprocal basis B = (b x c, c x a, a x b)/omega),
 * src/tools.f90:780-795 (expro). */
void orc_lattice(const double A[9], double B[9], double *omega, double W[3]) {
  const double *a1 = &A[0], *a2 = &A[3], *a3 = &A[6];
  double *b1 = &B[0], *b2 = &B[3], *b3 = &B[6];
#define EXPRO(H, U1, U2)                                  \
  do {                                                    \
    H[0] = U1[1] * U2[2] - U1[2] * U2[1];                 \
    H[1] = U1[2] * U2[0] - U1[0] * U2[2];                 \
    H[2] = U1[0] * U2[1] - U1[1] * U2[0];                 \
  } while (0)
  EXPRO(b1, a2, a3);
  EXPRO(b2, a3, a1);
  EXPRO(b3, a1, a2);
#undef EXPRO
  double om = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
  *omega = om;
  for (int j = 0; j < 3; ++j) {
    double *bj = &B[3 * j];
    W[j] = om / sqrt(bj[0] * bj[0] + bj[1] * bj[1] + bj[2] * bj[2]);
  }
  for (int i = 0; i < 9; ++i) B[i] = B[i] / om;
}

/* kardir (src/lattice.f90:183-216): s = B^T r, in the reference order. */
static inline void kardir1(const double *Bm, double x, double y, double z, double *s) {
  s[0] = x * M3(Bm, 0, 0) + y * M3(Bm, 1, 0) + z * M3(Bm, 2, 0);
  s[1] = x * M3(Bm, 0, 1) + y * M3(Bm, 1, 1) + z * M3(Bm, 2, 1);
  s[2] = x * M3(Bm, 0, 2) + y * M3(Bm, 1, 2) + z * M3(Bm, 2, 2);
}

/* minimum image of a fractional difference, then r = A s
 * (src/non_bonded.f90:1141-1150, src/ewald_dir.f90:185-200). */
static inline double min_image(const double *Am, double dsx, double dsy, double dsz, double r[3]) {
  double sx = dsx - orc_nint(dsx);
  double sy = dsy - orc_nint(dsy);
  double sz = dsz - orc_nint(dsz);
  r[0] = sx * M3(Am, 0, 0) + sy * M3(Am, 0, 1) + sz * M3(Am, 0, 2);
  r[1] = sx * M3(Am, 1, 0) + sy * M3(Am, 1, 1) + sz * M3(Am, 1, 2);
  r[2] = sx * M3(Am, 2, 0) + sy * M3(Am, 2, 1) + sz * M3(Am, 2, 2);
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

/* nk with the half-space symmetry (src/coulomb.f90:472) */
int orc_kmesh_count(const int32_t kmax[3]) {
  return kmax[2] + kmax[1] * (2 * kmax[2] + 1) + kmax[0] * (2 * kmax[1] + 1) * (2 * kmax[2] + 1);
}

/* stable merge sort by key (src/tools.f90:734-777 with merge_1 :660-706):
 * ties keep the original order. */
static void stable_sort_idx(const double *key, int n, int *idx) {
  int *tmp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) tmp[k++] = (key[idx[i]] <= key[idx[j]]) ? idx[i++] : idx[j++];
      while (i < mid) tmp[k++] = idx[i++];
      while (j < hi) tmp[k++] = idx[j++];
    }
    memcpy(idx, tmp, sizeof(int) * (size_t)n);
  }
  free(tmp);
}

/* kpoint_sum_init + reorder_kpt (src/kspace.f90:143-207, 286-342). */
int orc_kmesh(const double B[9], double alpha, const int32_t kmax[3], double *okx, double *oky,
              double *okz, double *okk, double *oAk, double *okcoe) {
  int nk = orc_kmesh_count(kmax);
  double *kx = malloc(sizeof(double) * nk * 6);
  double *ky = kx + nk, *kz = ky + nk, *kk = kz + nk, *Ak = kk + nk, *kc = Ak + nk;
  double alpha2 = alpha * alpha;
  double tpi = 2.0 * PI;
  int n = 0;
  for (int nx = 0; nx <= kmax[0]; ++nx) {
    int nymin = (nx == 0) ? 0 : -kmax[1];
    for (int ny = nymin; ny <= kmax[1]; ++ny) {
      int nzmin = (nx == 0 && ny == 0) ? 1 : -kmax[2];
      for (int nz = nzmin; nz <= kmax[2]; ++nz) {
        double fx = (double)nx, fy = (double)ny, fz = (double)nz;
        double x = tpi * (fx * M3(B, 0, 0) + fy * M3(B, 0, 1) + fz * M3(B, 0, 2));
        double y = tpi * (fx * M3(B, 1, 0) + fy * M3(B, 1, 1) + fz * M3(B, 1, 2));
        double z = tpi * (fx * M3(B, 2, 0) + fy * M3(B, 2, 1) + fz * M3(B, 2, 2));
        double k2 = x * x + y * y + z * z;
        kx[n] = x; ky[n] = y; kz[n] = z; kk[n] = k2;
        Ak[n] = exp(-(k2 * 0.25) / alpha2) / k2;
        kc[n] = 2.0 * (1.0 / k2 + 1.0 / alpha2 / 4.0);
        ++n;
      }
    }
  }
  int *idx = malloc(sizeof(int) * (size_t)nk);
  stable_sort_idx(kk, nk, idx);
  for (int i = 0; i < nk; ++i) {
    int l = idx[i];
    if (okx) okx[i] = kx[l];
    if (oky) oky[i] = ky[l];
    if (okz) okz[i] = kz[l];
    if (okk) okk[i] = kk[l];
    if (oAk) oAk[i] = Ak[l];
    if (okcoe) okcoe[i] = kc[l];
  }
  free(idx);
  free(kx);
  return nk;
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */

static int fail(orc_ctx *c, int code, const char *msg) {
  if (c) snprintf(c->err, sizeof c->err, "%s", msg);
  return code;
}

const char *orc_last_error(const orc_ctx *c) { return c ? c->err : "null context"; }

void orc_set_threads(orc_ctx *c, int nthreads) { c->nthreads = nthreads > 0 ? nthreads : 1; }

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* 3x3 inverse by cofactors (replaces LAPACK DGETRF/DGETRI, src/read_pos.f90:203-212). */
static void inv3(const double *m, double *o) {
  double a = M3(m, 0, 0), b = M3(m, 0, 1), cc = M3(m, 0, 2);
  double d = M3(m, 1, 0), e = M3(m, 1, 1), f = M3(m, 1, 2);
  double g = M3(m, 2, 0), h = M3(m, 2, 1), k = M3(m, 2, 2);
  double det = a * (e * k - f * h) - b * (d * k - f * g) + cc * (d * h - e * g);
  M3(o, 0, 0) = (e * k - f * h) / det;  M3(o, 0, 1) = (cc * h - b * k) / det; M3(o, 0, 2) = (b * f - cc * e) / det;
  M3(o, 1, 0) = (f * g - d * k) / det;  M3(o, 1, 1) = (a * k - cc * g) / det; M3(o, 1, 2) = (cc * d - a * f) / det;
  M3(o, 2, 0) = (d * h - e * g) / det;  M3(o, 2, 1) = (b * g - a * h) / det;  M3(o, 2, 2) = (a * e - b * d) / det;
}

/* initialize_param_non_bonded (src/non_bonded.f90:455-644) */
static void init_nonbonded(orc_ctx *c) {
  const mdff_desc *d = &c->d;
  int nt = d->ntype;
  double two16 = pow(2.0, 1.0 / 6.0);
  double tpi = 2.0 * PI;
  int natmi[16] = {0};
  for (int i = 0; i < c->n; ++i) natmi[c->itype[i] - 1]++;
  c->u_tail = 0.0;
  double ptail = 0.0;
  for (int it = 0; it < nt; ++it)
    for (int jt = 0; jt < nt; ++jt) {
      int p = T2I(it, jt);
      double pp = d->plj[p], qq = d->qlj[p];
      double rcut = d->cutshortrange;
      double rcutsq = rcut * rcut;
      double rcut3 = rcut * rcutsq;
      double ppqq = pp * qq;
      c->rcutsq[p] = rcutsq;
      c->epsp[p] = d->epslj[p] / (qq - pp);
      c->sigsq[p] = two16 * two16 * d->sigmalj[p] * d->sigmalj[p];
      double sr2 = c->sigsq[p] / rcutsq;
      double sr = sqrt(sr2);
      double srp = pow(sr, pp), srq = pow(sr, qq);
      c->uc[p] = c->epsp[p] * (pp * srq - qq * srp);
      c->uc1[p] = c->epsp[p] * ppqq / (2.0 * rcutsq) * (srq - srp);
      c->uc2[p] = 0.5 * c->epsp[p] * ((2.0 * pp + ppqq) * srq - (2.0 * qq + ppqq) * srp);
      c->fc[p] = ppqq * c->epsp[p] / c->sigsq[p];
      c->rs[p] = exp(d->rhomor[p] * d->sigmamor[p]);
      double ut = c->epsp[p] * (pp * srq / (qq - 3.0) - qq * srp / (pp - 3.0));
      double pt = ppqq * c->epsp[p] * (srq / (qq - 3.0) - srp / (pp - 3.0));
      ut = ut * rcut3 * tpi;
      pt = pt * rcut3 * tpi;
      if (natmi[it] != 0 && natmi[jt] != 0) {
        c->u_tail += ut * natmi[it] * natmi[jt] / c->omega;
        ptail += pt * natmi[it] * natmi[jt] / c->omega;
      }
    }
  double pvtail = ptail / c->omega / 3.0;
  c->p_tail = pvtail / d->press_unit;
}

int orc_create(const mdff_desc *d, const int32_t *itype, orc_ctx **out) {
  *out = NULL;
  if (!d || !itype || d->natm <= 0 || d->ntype <= 0 || d->ntype > 16) return MDFF_ERR_BAD_ARG;
  if (d->lmorse && (d->lnmlj || d->lbmhft || d->lbmhftd)) return MDFF_ERR_UNSUPPORTED;
  orc_ctx *c = calloc(1, sizeof *c);
  c->d = *d;
  c->n = d->natm;
  c->nthreads = 1;
  int n = c->n;
  c->itype = malloc(sizeof(int32_t) * n);
  for (int i = 0; i < n; ++i) {
    if (itype[i] < 1 || itype[i] > d->ntype) { free(c->itype); free(c); return MDFF_ERR_BAD_ARG; }
    c->itype[i] = itype[i];
  }
  memcpy(c->A, d->cell, sizeof c->A);
  orc_lattice(c->A, c->B, &c->omega, c->W);
  /* ewald_param clamps cutlongrange to half the smallest width (src/coulomb.f90:377-383);
   * verlet_coul%cut is taken after that (src/config.f90:129-130). */
  double half = 0.5 * fmin(c->W[0], fmin(c->W[1], c->W[2]));
  c->rc_coul = fmin(d->cutlongrange, half);
  c->rc_vdw = d->cutshortrange;
  c->alpha = d->alphaES;
  tt_coeffs(c->E_TT);
  /* tasks: coulomb_check_tag (src/coulomb.f90:188-238), pim_check_tag (src/pim.f90:147-183) */
  int lqch = 0, ldip = 0;
  for (int i = 0; i < 16; ++i) if (d->qch[i] != 0.0) lqch = 1;
  for (int i = 0; i < 48; ++i) if (d->dip[i] != 0.0) ldip = 1;
  c->linduced = 0;
  for (int i = 0; i < d->ntype; ++i) if (d->ldip_polar[i]) c->linduced = 1;
  c->task_qq = lqch;
  c->task_qd = ldip && lqch;
  c->task_dd = ldip;
  if (c->linduced) { c->task_qd |= lqch; c->task_dd = 1; }
  /* per-atom type info (src/read_pos.f90:160-234) */
  c->q = dalloc(n); c->mass = dalloc(n); c->pol = dalloc(9 * n); c->invpol = dalloc(9 * n);
  c->dipia = dalloc(3 * n);
  for (int i = 0; i < n; ++i) {
    int it = c->itype[i] - 1;
    c->q[i] = d->qch[it];
    c->mass[i] = d->mass[it];
    for (int a = 0; a < 3; ++a) c->dipia[3 * i + a] = d->dip[it + 16 * a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) c->pol[9 * i + a + 3 * b] = d->poldip[it + 16 * a + 48 * b];
    if (d->ldip_polar[it]) inv3(&c->pol[9 * i], &c->invpol[9 * i]);
  }
  c->rx = dalloc(n); c->ry = dalloc(n); c->rz = dalloc(n);
  c->vx = dalloc(n); c->vy = dalloc(n); c->vz = dalloc(n);
  c->fx = dalloc(n); c->fy = dalloc(n); c->fz = dalloc(n);
  c->xs = dalloc(n); c->ys = dalloc(n); c->zs = dalloc(n);
  c->mu_t = dalloc(3 * n); c->ef_t = dalloc(3 * n); c->efg_t = dalloc(9 * (size_t)n);
  c->hist = dalloc(15 * (size_t)n);
  init_nonbonded(c);
  if (d->lcoulomb) {
    c->nk = orc_kmesh_count(d->kES);
    c->kx = dalloc(c->nk); c->ky = dalloc(c->nk); c->kz = dalloc(c->nk);
    c->kk = dalloc(c->nk); c->Ak = dalloc(c->nk); c->kcoe = dalloc(c->nk);
    orc_kmesh(c->B, c->alpha, d->kES, c->kx, c->ky, c->kz, c->kk, c->Ak, c->kcoe);
  }
  c->pv = malloc(sizeof(int32_t) * (n + 1));
  c->pc = malloc(sizeof(int32_t) * (n + 1));
  *out = c;
  return MDFF_OK;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  void *ps[] = {c->itype, c->q, c->mass, c->pol, c->invpol, c->dipia, c->rx, c->ry, c->rz,
                c->vx, c->vy, c->vz, c->fx, c->fy, c->fz, c->xs, c->ys, c->zs, c->mu_t,
                c->ef_t, c->efg_t, c->hist, c->kx, c->ky, c->kz, c->kk, c->Ak, c->kcoe, c->pv,
                c->lv, c->pc, c->lc};
  for (size_t i = 0; i < sizeof ps / sizeof ps[0]; ++i) free(ps[i]);
  free(c);
}

#define COPY3(dst1, dst2, dst3, s1, s2, s3)                        \
  do {                                                             \
    memcpy(dst1, s1, sizeof(double) * c->n);                       \
    memcpy(dst2, s2, sizeof(double) * c->n);                       \
    memcpy(dst3, s3, sizeof(double) * c->n);                       \
  } while (0)

int orc_set_positions(orc_ctx *c, const double *x, const double *y, const double *z) {
  COPY3(c->rx, c->ry, c->rz, x, y, z);
  return MDFF_OK;
}
int orc_get_positions(orc_ctx *c, double *x, double *y, double *z) {
  COPY3(x, y, z, c->rx, c->ry, c->rz);
  return MDFF_OK;
}
int orc_set_velocities(orc_ctx *c, const double *x, const double *y, const double *z) {
  COPY3(c->vx, c->vy, c->vz, x, y, z);
  return MDFF_OK;
}
int orc_get_velocities(orc_ctx *c, double *x, double *y, double *z) {
  COPY3(x, y, z, c->vx, c->vy, c->vz);
  return MDFF_OK;
}
int orc_get_forces(orc_ctx *c, double *x, double *y, double *z) {
  COPY3(x, y, z, c->fx, c->fy, c->fz);
  return MDFF_OK;
}
int orc_set_itime(orc_ctx *c, int32_t itime) { c->itime = itime; return MDFF_OK; }
int orc_get_cutoffs(orc_ctx *c, double *rv, double *rcl, double *alpha, int32_t *nk) {
  if (rv) *rv = c->rc_vdw;
  if (rcl) *rcl = c->rc_coul;
  if (alpha) *alpha = c->alpha;
  if (nk) *nk = c->nk;
  return MDFF_OK;
}

static double *frac_coords(orc_ctx *c) {
  double *s = dalloc(3 * (size_t)c->n);
  for (int i = 0; i < c->n; ++i) kardir1(c->B, c->rx[i], c->ry[i], c->rz[i], &s[3 * i]);
  return s;
}

/* ------------------------------------------------------------------------ */
/* Verlet lists                                                              */
/* ------------------------------------------------------------------------ */

/* vnlist_pbc (src/tools.f90:273-368): both half lists in one O(N^2) sweep.
 * Ownership (ia>ja and ia+ja even) or (ia<ja and ia+ja odd) with 1-based ia,ja;
 * acceptance r^2 <= (cut+skindiff)^2; j ascending per i. */
int orc_vnlist_build(orc_ctx *c) {
  int n = c->n;
  double *s = frac_coords(c);
  double rv = c->rc_vdw + c->d.skindiff, rcl = c->rc_coul + c->d.skindiff;
  double rsv = rv * rv, rsc = rcl * rcl;
  int64_t cap = (int64_t)(VNLMAX * n);
  /* count pass */
  int64_t cv = 0, cc = 0;
  for (int pass = 0; pass < 2; ++pass) {
    cv = 0; cc = 0;
    for (int i = 0; i < n; ++i) {
      if (pass) { c->pv[i] = (int32_t)(cv + 1); c->pc[i] = (int32_t)(cc + 1); }
      for (int j = 0; j < n; ++j) {
        int own = (i > j && ((i + j) % 2 == 0)) || (i < j && ((i + j) % 2 != 0));
        if (!own) continue;
        double r[3];
        double r2 = min_image(c->A, s[3 * i] - s[3 * j], s[3 * i + 1] - s[3 * j + 1],
                              s[3 * i + 2] - s[3 * j + 2], r);
        if (r2 <= rsv) { if (pass) c->lv[cv] = j + 1; cv++; }
        if (r2 <= rsc) { if (pass) c->lc[cc] = j + 1; cc++; }
      }
    }
    if (!pass) {
      if (cv > cap || cc > cap) { free(s); return fail(c, MDFF_ERR_LIST_OVERFLOW, "ERROR: out of bound list in vnlist_pbc"); }
      free(c->lv); free(c->lc);
      c->lv = malloc(sizeof(int32_t) * (cv ? cv : 1));
      c->lc = malloc(sizeof(int32_t) * (cc ? cc : 1));
    }
  }
  c->pv[n] = (int32_t)(cv + 1);
  c->pc[n] = (int32_t)(cc + 1);
  c->nv = cv; c->nc = cc;
  c->lists_valid = 1;
  free(s);
  return MDFF_OK;
}

/* vnlistcheck (src/tools.f90:444-537): rebuild when the sum of the two
 * largest displacements since the last save exceeds skindiff. */
int orc_vnlistcheck(orc_ctx *c, int32_t *rebuilt) {
  double *s = frac_coords(c);
  double m1 = 0.0, m2 = 0.0;
  for (int i = 0; i < c->n; ++i) {
    double r[3];
    double r2 = min_image(c->A, s[3 * i] - c->xs[i], s[3 * i + 1] - c->ys[i], s[3 * i + 2] - c->zs[i], r);
    double dr = sqrt(r2);
    if (dr > m1) { m2 = m1; m1 = dr; }
    else if (dr > m2) m2 = dr;
  }
  int rb = 0;
  if (m1 + m2 > c->d.skindiff) {
    int st = orc_vnlist_build(c);
    if (st) { free(s); return st; }
    for (int i = 0; i < c->n; ++i) { c->xs[i] = s[3 * i]; c->ys[i] = s[3 * i + 1]; c->zs[i] = s[3 * i + 2]; }
    rb = 1;
  }
  if (rebuilt) *rebuilt = rb;
  free(s);
  return MDFF_OK;
}

int orc_get_pairs(orc_ctx *c, int32_t which, int32_t *point, int32_t *list, int64_t cap, int64_t *n) {
  if (!c->lists_valid) return fail(c, MDFF_ERR_BAD_ARG, "lists not built");
  int64_t cnt = which ? c->nc : c->nv;
  if (n) *n = cnt;
  if (cap < cnt) return MDFF_ERR_BAD_ARG;
  memcpy(point, which ? c->pc : c->pv, sizeof(int32_t) * (c->n + 1));
  memcpy(list, which ? c->lc : c->lv, sizeof(int32_t) * cnt);
  return MDFF_OK;
}

/* iterate the half list of atom i (0-based) or, without lists, all j > i */
#define FOR_PAIRS(c, point, list, i, j, ...)                                         \
  do {                                                                                \
    if ((c)->d.lvnlist) {                                                             \
      for (int32_t _e = (point)[i] - 1; _e < (point)[(i) + 1] - 1; ++_e) {            \
        int j = (list)[_e] - 1;                                                       \
        if (j == (i)) continue;                                                       \
        __VA_ARGS__                                                                   \
      }                                                                               \
    } else {                                                                          \
      for (int j = (i) + 1; j < (c)->n; ++j) { __VA_ARGS__ }                          \
    }                                                                                 \
  } while (0)

/* ------------------------------------------------------------------------ */
/* short-range pair potentials                                               */
/* ------------------------------------------------------------------------ */

typedef struct { double u, vir, tau[9]; } sr_out;

/* engforce_bmhftd_pbc (src/non_bonded.f90:1076-1228), engforce_nmlj_pbc
 * (:660-827), Morse energy of engforce_morse_pbc (:1241-1390) with its
 * analytic derivative.  model: 1 = LJ, 2 = BMHFT(D), 3 = Morse. */
static void short_range(orc_ctx *c, const double *s, int model, sr_out *o) {
  const mdff_desc *d = &c->d;
  int n = c->n;
  memset(o, 0, sizeof *o);
  for (int i = 0; i < n; ++i) { c->fx[i] = 0.0; c->fy[i] = 0.0; c->fz[i] = 0.0; }
  double u = 0.0, vir = 0.0, tau[9] = {0};
  for (int i = 0; i < n; ++i) {
    double sxi = s[3 * i], syi = s[3 * i + 1], szi = s[3 * i + 2];
    int p1 = c->itype[i] - 1;
    FOR_PAIRS(c, c->pv, c->lv, i, j, {
      double r[3];
      double r2 = min_image(c->A, sxi - s[3 * j], syi - s[3 * j + 1], szi - s[3 * j + 2], r);
      int p = T2I(p1, c->itype[j] - 1);
      double wij;
      if (model == 2) {
        if (d->Abmhftd[p] == 0.0) continue;
        if (!(r2 < c->rcutsq[p])) continue;
        double ir2 = 1.0 / r2;
        double dd = sqrt(r2);
        double erh = d->Abmhftd[p] * exp(-d->Bbmhftd[p] * dd);
        double ir6 = ir2 * ir2 * ir2;
        double ir8 = ir6 * ir2;
        ir6 = ir6 * d->Cbmhftd[p];
        ir8 = ir8 * d->Dbmhftd[p];
        double f6 = 1.0, f8 = 1.0, fd6 = 0.0, fd8 = 0.0, ir6d = ir6, ir8d = ir8;
        if (d->lbmhftd) {
          orc_tt_damp(d->BDbmhftd[p], 1.0, dd, 6, &f6, &fd6);
          orc_tt_damp(d->BDbmhftd[p], 1.0, dd, 8, &f8, &fd8);
          ir6d = ir6 * f6;
          ir8d = ir8 * f8;
        }
        double ir7 = 6.0 * ir6d / dd;
        double ir9 = 8.0 * ir8d / dd;
        u = u + erh - ir6d - ir8d;
        wij = d->Bbmhftd[p] * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8;
        wij = wij / dd;
      } else if (model == 1) {
        if (!(r2 < c->rcutsq[p])) continue;
        double sr2 = c->sigsq[p] / r2;
        double srp = pow(sr2, d->plj[p] * 0.5);
        double srq = pow(sr2, d->qlj[p] * 0.5);
        double e = c->epsp[p] * (d->plj[p] * srq - d->qlj[p] * srp);
        if (d->trunc == 0) u = u + e;
        if (d->trunc == 1) u = u + e - c->uc[p];
        if (d->trunc == 2) u = u + e + c->uc1[p] * r2 - c->uc2[p];
        wij = c->fc[p] * (srq - srp) * sr2;
      } else {
        if (!(r2 < c->rcutsq[p])) continue;
        double dd = sqrt(r2);
        double erh = exp(-dd * d->rhomor[p]);
        double x = erh * c->rs[p] - 1.0;
        u = u + d->epsmor[p] * (x * x - 1.0);
        /* -dV/dr / r with V = eps ((e^{-rho(r-sigma)} - 1)^2 - 1) */
        wij = 2.0 * d->epsmor[p] * d->rhomor[p] * erh * c->rs[p] * x / dd;
      }
      double fxij = wij * r[0], fyij = wij * r[1], fzij = wij * r[2];
      if (model == 1) vir = vir + wij * r2;
      else vir = vir + (r[0] * fxij + r[1] * fyij + r[2] * fzij);
      c->fx[i] += fxij; c->fy[i] += fyij; c->fz[i] += fzij;
      c->fx[j] -= fxij; c->fy[j] -= fyij; c->fz[j] -= fzij;
      double f[3] = {fxij, fyij, fzij};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double t = r[a] * f[b] + r[b] * f[a];
          if (model == 1) t = t * 0.5;
          tau[a + 3 * b] += t;
        }
    });
  }
  o->u = u;
  o->vir = vir / 3.0;
  for (int k = 0; k < 9; ++k) {
    if (model == 1) o->tau[k] = tau[k] / c->omega / d->press_unit;
    else o->tau[k] = tau[k] / c->omega / d->press_unit * 0.5;
  }
}

/* ------------------------------------------------------------------------ */
/* Ewald                                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int qq, qd, dd;      /* task(1:3) */
  int damp;            /* damp_ind */
  int efield, forces, stress;
  int norec;           /* do_rec = .false. */
  double alpha;        /* alphaES for this call when > 0 (the scf_kO_v2 inner loop) */
  int efg;             /* do_efg */
} es_task;

/* multipole_ES_dir (src/ewald_dir.f90:40-798) restricted to charges and
 * dipoles (tasks 1-3; Thole off, src/coulomb.f90:173).  Outputs are
 * accumulated into ef (3n), efg (9n, only with tk.efg), f (3n, AoS), tau[9]
 * (already /omega*0.5). */
static void ewald_dir(orc_ctx *c, const double *s, const double *q, const double *mu, es_task tk,
                      double *u_out, double *ef, double *efg, double *f, double tau[9]) {
  const mdff_desc *d = &c->d;
  int n = c->n;
  double cutsq = c->rc_coul * c->rc_coul;
  double al = tk.alpha > 0.0 ? tk.alpha : c->alpha, a2 = al * al, a3 = a2 * al;
  double piroot = sqrt(PI);
  double u = 0.0, tt[9] = {0};
  for (int i = 0; i < n; ++i) {
    int ita = c->itype[i] - 1;
    double qi = q[i];
    const double *mui = &mu[3 * i];
    int dip_i = (mui[0] != 0.0 || mui[1] != 0.0 || mui[2] != 0.0);
    FOR_PAIRS(c, c->pc, c->lc, i, j, {
      int jta = c->itype[j] - 1;
      double qj = q[j];
      const double *muj = &mu[3 * j];
      int dip_j = (muj[0] != 0.0 || muj[1] != 0.0 || muj[2] != 0.0);
      double qij = qi * qj;
      double rij[3];
      double d2 = min_image(c->A, s[3 * i] - s[3 * j], s[3 * i + 1] - s[3 * j + 1],
                            s[3 * i + 2] - s[3 * j + 2], rij);
      if (d2 > cutsq) continue;
      double dd = sqrt(d2);
      double d3 = d2 * dd;
      double dm1 = 1.0 / dd, dm3 = dm1 / d2, dm5 = dm3 / d2, dm7 = dm5 / d2;
      int ldamp = (d->ldip_damping[T3I(ita, ita, jta)] || d->ldip_damping[T3I(jta, ita, jta)]) && tk.damp;
      double fdamp = 1.0, fdamp2 = 1.0, fdd = 0.0, fdd2 = 0.0;
      if (ldamp) {
        orc_tt_damp(d->pol_damp_b[T3I(ita, ita, jta)], d->pol_damp_c[T3I(ita, ita, jta)], dd,
                    d->pol_damp_k[T3I(ita, ita, jta)], &fdamp, &fdd);
        orc_tt_damp(d->pol_damp_b[T3I(jta, ita, jta)], d->pol_damp_c[T3I(jta, ita, jta)], dd,
                    d->pol_damp_k[T3I(jta, ita, jta)], &fdamp2, &fdd2);
      }
      double expon = exp(-a2 * d2) / piroot;
      double F0 = orc_errfc(al * dd);
      double F1 = F0 + 2.0 * al * dd * expon;
      double F2 = F1 + 4.0 * a3 * d3 * expon / 3.0;
      double d5 = d3 * d2;
      double a5 = a3 * a2;
      double F3 = F2 + 8.0 * a5 * d5 * expon / 15.0;
      double F1d = -fdamp + 1.0;
      double F2d = F1d + (dd / 3.0) * fdd;
      double F1d2 = -fdamp2 + 1.0;
      double F2d2 = F1d2 + (dd / 3.0) * fdd2;
      double T0 = dm1 * F0;
      double T1[3], T1d[3], T1d2[3];
      for (int a = 0; a < 3; ++a) {
        T1[a] = -rij[a] * dm3;
        T1d[a] = T1[a] * F1d;
        T1d2[a] = T1[a] * F1d2;
        T1[a] = T1[a] * F1;
      }
      double F2_dm5 = 3.0 * F2 * dm5, F2d_dm5 = 3.0 * F2d * dm5, F2d2_dm5 = 3.0 * F2d2 * dm5;
      double F1_dm3 = F1 * dm3, F1d_dm3 = F1d * dm3, F1d2_dm3 = F1d2 * dm3;
      double T2[9], T2d[9], T2d2[9];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          T2[a + 3 * b] = rij[a] * rij[b] * F2_dm5;
          T2d[a + 3 * b] = rij[a] * rij[b] * F2d_dm5;
          T2d2[a + 3 * b] = rij[a] * rij[b] * F2d2_dm5;
          if (a == b) {
            T2[a + 3 * b] -= F1_dm3;
            T2d[a + 3 * b] -= F1d_dm3;
            T2d2[a + 3 * b] -= F1d2_dm3;
          }
        }
      double F3_dm7 = F3 * dm7 * 15.0;
      double fij[3] = {0.0, 0.0, 0.0};
      if (tk.qq) {
        u = u + qij * T0;
        if (tk.efield) {
          for (int a = 0; a < 3; ++a) {
            ef[3 * i + a] = ef[3 * i + a] - qj * T1[a];
            ef[3 * j + a] = ef[3 * j + a] + qi * T1[a];
          }
          if (ldamp)
            for (int a = 0; a < 3; ++a) {
              ef[3 * i + a] = ef[3 * i + a] + qj * T1d[a];
              ef[3 * j + a] = ef[3 * j + a] - qi * T1d2[a];
            }
        }
        /* src/ewald_dir.f90:543-546 */
        if (tk.efg)
          for (int x = 0; x < 9; ++x) {
            efg[9 * i + x] = efg[9 * i + x] - qj * T2[x];
            efg[9 * j + x] = efg[9 * j + x] - qi * T2[x];
          }
        if (tk.forces)
          for (int a = 0; a < 3; ++a) fij[a] = fij[a] + qij * T1[a];
      }
      if (tk.dd && dip_i && dip_j) {
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) u = u - mui[a] * T2[a + 3 * b] * muj[b];
        if (tk.efield)
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
              ef[3 * i + a] = ef[3 * i + a] + T2[a + 3 * b] * muj[b];
              ef[3 * j + a] = ef[3 * j + a] + T2[a + 3 * b] * mui[b];
            }
        /* src/ewald_dir.f90:577-582: efg(:,:,ia) += T3(:,:,k) muj(k),
         * efg(:,:,ja) -= T3(:,:,k) mui(k), T3 as at :436-447 */
        if (tk.efg)
          for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 3; ++b)
              for (int a = 0; a < 3; ++a) {
                double t3 = -rij[a] * rij[b] * rij[k] * F3_dm7;
                if (a == b) t3 += rij[k] * F2_dm5;
                if (a == k) t3 += rij[b] * F2_dm5;
                if (b == k) t3 += rij[a] * F2_dm5;
                efg[9 * i + a + 3 * b] = efg[9 * i + a + 3 * b] + t3 * muj[k];
                efg[9 * j + a + 3 * b] = efg[9 * j + a + 3 * b] - t3 * mui[k];
              }
        if (tk.forces) {
          /* T3(a,b,k) for fixed pair (src/ewald_dir.f90:436-447) */
          for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) {
              for (int b = 0; b < 3; ++b) {
                double t3 = -rij[a] * rij[b] * rij[k] * F3_dm7;
                if (a == b) t3 += rij[k] * F2_dm5;
                if (a == k) t3 += rij[b] * F2_dm5;
                if (b == k) t3 += rij[a] * F2_dm5;
                fij[b] = fij[b] - mui[a] * t3 * muj[k];
              }
            }
        }
      }
      if (tk.qd && (dip_i || dip_j)) {
        for (int k = 0; k < 3; ++k) {
          u = u - qi * T1[k] * muj[k];
          u = u + qj * T1[k] * mui[k];
          if (ldamp) {
            u = u + qi * T1d2[k] * muj[k];
            u = u - qj * T1d[k] * mui[k];
          }
        }
        if (tk.forces) {
          for (int b = 0; b < 3; ++b)
            for (int k = 0; k < 3; ++k) {
              fij[k] = fij[k] - qi * T2[k + 3 * b] * muj[b];
              fij[k] = fij[k] + qj * T2[k + 3 * b] * mui[b];
            }
          if (ldamp)
            for (int b = 0; b < 3; ++b)
              for (int k = 0; k < 3; ++k) {
                fij[k] = fij[k] - qj * T2d[b + 3 * k] * mui[b];
                fij[k] = fij[k] + qi * T2d2[b + 3 * k] * muj[b];
              }
        }
      }
      if (tk.forces)
        for (int a = 0; a < 3; ++a) {
          f[3 * i + a] -= fij[a];
          f[3 * j + a] += fij[a];
        }
      if (tk.stress)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) tt[a + 3 * b] = tt[a + 3 * b] - (rij[a] * fij[b] + rij[b] * fij[a]);
    });
  }
  *u_out = u;
  if (tk.stress)
    for (int k = 0; k < 9; ++k) tau[k] = tt[k] / c->omega * 0.5;
}

/* multipole_ES_rec (src/ewald_rec.f90:38-365) for charges and dipoles.
 * Parallel over k-vectors when nthreads > 1 (private accumulators).  efg (9n)
 * receives the upper triangle and the diagonal only (:209-216). */
static void ewald_rec(orc_ctx *c, const double *q, const double *mu, es_task tk, double *u_out,
                      double *ef, double *efg, double *f, double tau[9]) {
  int n = c->n, nk = c->nk;
  int ldip = tk.qd || tk.dd;
  double tpi_V = 2.0 * PI / c->omega;
  double fpi_V = tpi_V * 2.0;
  int nth = c->nthreads;
  const size_t per = 12 * (size_t)n + 10;   /* ef 3n, f 3n, u, tau 9, efg 6n */
  double *acc = dalloc((size_t)nth * per);
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *my = acc + (size_t)tid * per;
    double *mef = my, *mf = my + 3 * n, *mu_rec = my + 6 * n, *mtau = my + 6 * n + 1;
    double *mefg = my + 6 * n + 10;
    double *ckr = dalloc(n), *skr = dalloc(n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int ik = 0; ik < nk; ++ik) {
      if (c->kk[ik] == 0.0) continue;
      double kx = c->kx[ik], ky = c->ky[ik], kz = c->kz[ik], Ak = c->Ak[ik], kcoe = c->kcoe[ik];
      double rR = 0.0, rI = 0.0;
      for (int ia = 0; ia < n; ++ia) {
        double qi = q[ia];
        double kr = (kx * c->rx[ia] + ky * c->ry[ia] + kz * c->rz[ia]);
        ckr[ia] = cos(kr);
        skr[ia] = sin(kr);
        rR = rR + qi * ckr[ia];
        rI = rI + qi * skr[ia];
        if (!ldip) continue;
        double kmu = (mu[3 * ia] * kx + mu[3 * ia + 1] * ky + mu[3 * ia + 2] * kz);
        rR = rR - kmu * skr[ia];
        rI = rI + kmu * ckr[ia];
      }
      double str = (rR * rR + rI * rI) * Ak;
      *mu_rec += str;
      for (int ia = 0; ia < n; ++ia) {
        double ca = ckr[ia], sa = skr[ia];
        double recarg2 = Ak * (rR * ca + rI * sa);
        double recarg = Ak * (rI * ca - rR * sa);
        double fxij = kx * recarg, fyij = ky * recarg, fzij = kz * recarg;
        if (tk.efield) {
          mef[3 * ia] -= fxij; mef[3 * ia + 1] -= fyij; mef[3 * ia + 2] -= fzij;
        }
        if (tk.efg) {
          /* (1,1) (2,2) (3,3) (1,2) (1,3) (2,3), src/ewald_rec.f90:209-216 */
          double *g = &mefg[6 * ia];
          g[0] = g[0] + kx * kx * recarg2;
          g[1] = g[1] + ky * ky * recarg2;
          g[2] = g[2] + kz * kz * recarg2;
          g[3] = g[3] + kx * ky * recarg2;
          g[4] = g[4] + kx * kz * recarg2;
          g[5] = g[5] + ky * kz * recarg2;
        }
        if (tk.forces) {
          double qi = q[ia];
          mf[3 * ia] -= qi * fxij; mf[3 * ia + 1] -= qi * fyij; mf[3 * ia + 2] -= qi * fzij;
          if (!ldip) continue;
          double kmu = (mu[3 * ia] * kx + mu[3 * ia + 1] * ky + mu[3 * ia + 2] * kz) * recarg2;
          mf[3 * ia] += kx * kmu; mf[3 * ia + 1] += ky * kmu; mf[3 * ia + 2] += kz * kmu;
        }
      }
      if (tk.stress) {
        double kv[3] = {kx, ky, kz};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            double t = -kcoe * kv[a] * kv[b] * str;
            if (a == b) t = (1.0 - kcoe * kv[a] * kv[a]) * str;
            mtau[a + 3 * b] += t;
          }
      }
    }
    free(ckr);
    free(skr);
  }
  double u = 0.0, tr[9] = {0};
  static const int up[6] = {0, 4, 8, 3, 6, 7};   /* (1,1) (2,2) (3,3) (1,2) (1,3) (2,3) */
  for (int t = 0; t < nth; ++t) {
    double *my = acc + (size_t)t * per;
    u += my[6 * n];
    for (int k = 0; k < 9; ++k) tr[k] += my[6 * n + 1 + k];
    for (size_t k = 0; k < 3 * (size_t)n; ++k) {
      ef[k] += my[k] * 2.0 * fpi_V;
      f[k] += my[3 * n + k] * 2.0 * fpi_V;
    }
    if (tk.efg)
      for (int ia = 0; ia < n; ++ia)
        for (int x = 0; x < 6; ++x) efg[9 * ia + up[x]] += my[6 * n + 10 + 6 * ia + x] * 2.0 * fpi_V;
  }
  free(acc);
  u = u * 2.0;
  for (int k = 0; k < 9; ++k) tr[k] *= 2.0;
  double sym[9];
  memcpy(sym, tr, sizeof sym);
  sym[3] = (tr[3] + tr[1]) * 0.5; sym[6] = (tr[6] + tr[2]) * 0.5; sym[7] = (tr[7] + tr[5]) * 0.5;
  sym[1] = sym[3]; sym[2] = sym[6]; sym[5] = sym[7];
  *u_out = u * tpi_V;
  if (tk.stress)
    for (int k = 0; k < 9; ++k) tau[k] = sym[k] * tpi_V / c->omega;
}

/* multipole_ES (src/ewald.f90:49-314) without quadrupoles.  ef receives
 * ef_dir+ef_rec+ef_self (+ef_surf) in internal units; forces are added to
 * c->fx.. scaled by coul_unit; returns u_coul WITHOUT u_pol. */
typedef struct { double u_dir, u_rec, u_self, u_surf, tau_coul[9]; } es_out;

static void multipole_es(orc_ctx *c, const double *s, const double *q, const double *mu, es_task tk,
                         double *ef, double *efg, es_out *o) {
  int n = c->n;
  const mdff_desc *d = &c->d;
  double *ef_dir = dalloc(3 * n), *ef_rec = dalloc(3 * n);
  double *efg_dir = dalloc(9 * (size_t)n), *efg_rec = dalloc(9 * (size_t)n);
  double *f_dir = dalloc(3 * n), *f_rec = dalloc(3 * n);
  double tau_dir[9] = {0}, tau_rec[9] = {0};
  double mutot[3] = {0}, qtot[3] = {0}, musq = 0.0, qsq = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) mutot[a] += mu[3 * i + a];
    musq = musq + (mu[3 * i] * mu[3 * i] + mu[3 * i + 1] * mu[3 * i + 1] + mu[3 * i + 2] * mu[3 * i + 2]);
    qtot[0] += q[i] * c->rx[i]; qtot[1] += q[i] * c->ry[i]; qtot[2] += q[i] * c->rz[i];
    qsq = qsq + q[i] * q[i];
  }
  double tpi_V = 2.0 * PI / c->omega, tpi_3V = tpi_V / 3.0, fpi_3V = tpi_3V * 2.0;
  const double al = tk.alpha > 0.0 ? tk.alpha : c->alpha;
  double alpha2 = al * al;
  double selfa = al / sqrt(PI);
  double selfa2 = 2.0 * selfa * alpha2 / 3.0;
  ewald_dir(c, s, q, mu, tk, &o->u_dir, ef_dir, efg_dir, f_dir, tau_dir);
  o->u_rec = 0.0;
  if (!tk.norec) ewald_rec(c, q, mu, tk, &o->u_rec, ef_rec, efg_rec, f_rec, tau_rec);
  double uq = qtot[0] * qtot[0] + qtot[1] * qtot[1] + qtot[2] * qtot[2];
  double ud = mutot[0] * mutot[0] + mutot[1] * mutot[1] + mutot[2] * mutot[2];
  double uqd = 2.0 * (qtot[0] * mutot[0] + qtot[1] * mutot[1] + qtot[2] * mutot[2]);
  o->u_surf = (uq + uqd + ud) * tpi_3V;
  o->u_self = -selfa * qsq + (-selfa2 * musq);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      double e = ef_dir[3 * i + a] + ef_rec[3 * i + a];
      if (d->lsurf) e += -(qtot[a] + mutot[a]) * fpi_3V;
      e += 2.0 * selfa2 * mu[3 * i + a];
      ef[3 * i + a] = e;
    }
  /* efg = (efg_dir + efg_rec + efg_self) * coul_unit, the self term
   * -2 selfa2 q on the diagonal whether or not do_efg (src/ewald.f90:214-216,231) */
  if (efg)
    for (int i = 0; i < n; ++i)
      for (int x = 0; x < 9; ++x) {
        double self = (x % 4 == 0) ? -2.0 * selfa2 * q[i] : 0.0;
        efg[9 * i + x] = (efg_dir[9 * i + x] + efg_rec[9 * i + x] + self) * d->coul_unit;
      }
  for (int k = 0; k < 9; ++k) o->tau_coul[k] = (tau_dir[k] + tau_rec[k]) * d->coul_unit / d->press_unit;
  if (tk.forces)
    for (int i = 0; i < n; ++i) {
      c->fx[i] = c->fx[i] + (f_rec[3 * i] + f_dir[3 * i]) * d->coul_unit;
      c->fy[i] = c->fy[i] + (f_rec[3 * i + 1] + f_dir[3 * i + 1]) * d->coul_unit;
      c->fz[i] = c->fz[i] + (f_rec[3 * i + 2] + f_dir[3 * i + 2]) * d->coul_unit;
    }
  free(ef_dir); free(ef_rec); free(f_dir); free(f_rec); free(efg_dir); free(efg_rec);
}

int orc_multipole_es(orc_ctx *c, const double *mu, const int32_t task[3], int32_t damp, double *ef,
                     double *fx, double *fy, double *fz, double tau_coul[9], double *u_dir,
                     double *u_rec, double *u_self, double *u_surf) {
  if (c->d.lvnlist && !c->lists_valid) {
    int st = orc_vnlist_build(c);
    if (st) return st;
  }
  double *s = frac_coords(c);
  es_task tk = {task[0], task[1], task[2], damp, 1, 1, 1};
  es_out o;
  for (int i = 0; i < c->n; ++i) { c->fx[i] = 0; c->fy[i] = 0; c->fz[i] = 0; }
  double *e = dalloc(3 * c->n);
  multipole_es(c, s, c->q, mu, tk, e, NULL, &o);
  if (ef) memcpy(ef, e, sizeof(double) * 3 * c->n);
  if (fx) orc_get_forces(c, fx, fy, fz);
  if (tau_coul) memcpy(tau_coul, o.tau_coul, sizeof o.tau_coul);
  if (u_dir) *u_dir = o.u_dir;
  if (u_rec) *u_rec = o.u_rec;
  if (u_self) *u_self = o.u_self;
  if (u_surf) *u_surf = o.u_surf;
  free(e);
  free(s);
  return MDFF_OK;
}

/* Field of a dipole set alone (charges zeroed, mu-mu task, undamped): the
 * induced-field evaluation of the SCF loop (src/moment_from_pola.f90:401-404). */
int orc_dipole_field(orc_ctx *c, const double *mu, double *ef) {
  if (c->d.lvnlist && !c->lists_valid) {
    int st = orc_vnlist_build(c);
    if (st) return st;
  }
  double *s = frac_coords(c), *zq = dalloc(c->n);
  es_task ti = {0, 0, 1, 0, 1, 0, 0};
  es_out o;
  multipole_es(c, s, zq, mu, ti, ef, NULL, &o);
  free(s);
  free(zq);
  return MDFF_OK;
}

/* ------------------------------------------------------------------------ */
/* induced dipoles                                                           */
/* ------------------------------------------------------------------------ */

/* induced_moment (src/moment_from_pola.f90:955-993) */
static void induced_moment(orc_ctx *c, const double *E, double *mu) {
  for (int i = 0; i < c->n; ++i) {
    int it = c->itype[i] - 1;
    for (int a = 0; a < 3; ++a) mu[3 * i + a] = 0.0;
    if (!c->d.ldip_polar[it]) continue;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) mu[3 * i + a] = mu[3 * i + a] + c->pol[9 * i + a + 3 * b] * E[3 * i + b];
  }
}

/* get_rmsd_mu (src/moment_from_pola.f90:998-1027) */
static double rmsd_mu(orc_ctx *c, const double *E, const double *mu) {
  double r = 0.0;
  int npol = 0;
  for (int i = 0; i < c->n; ++i) {
    int it = c->itype[i] - 1;
    if (!c->d.ldip_polar[it]) continue;
    npol++;
    for (int a = 0; a < 3; ++a) {
      double p = c->pol[9 * i + a + 3 * a];
      if (p == 0.0) continue;
      double x = mu[3 * i + a] / p - E[3 * i + a];
      r = r + x * x;
    }
  }
  return sqrt(r / (double)npol);
}

/* Predictor coefficients B[0..ext] (B[k] multiplies the dipoles of k steps
 * back) and corrector weight w of the ASPC order ext in use
 * (src/moment_from_pola.f90:526-562). */
void orc_aspc_coeffs(int ext, double B[6], double *w) {
  for (int k = 0; k < 6; ++k) B[k] = 0.0;
  switch (ext) {
    case 0: B[0] = 1.0; *w = 0.0; break;
    case 1: B[0] = 2.0; B[1] = -1.0; *w = 2.0 / 3.0; break;
    case 2: B[0] = 2.5; B[1] = -2.0; B[2] = 0.5; *w = 0.6; break;
    case 3: B[0] = 2.8; B[1] = -2.8; B[2] = 1.2; B[3] = -0.2; *w = 4.0 / 7.0; break;
    default: B[0] = 3.0; B[1] = -24.0 / 7.0; B[2] = 27.0 / 14.0; B[3] = -4.0 / 7.0; B[4] = 1.0 / 14.0; *w = 5.0 / 9.0; break;
  }
}

/* extrapolate_dipole_aspc (src/moment_from_pola.f90:493-598) */
static void aspc(orc_ctx *c, double *mu, const double *E, int key) {
  int n = c->n, ord = c->d.extrapolate_order, ext;
  if (c->itime >= ord + 1) ext = ord;
  else if (c->itime == 1 || c->itime == 0) ext = 0;
  else ext = c->itime - 1;
  double Bc[6], w;
  orc_aspc_coeffs(ext, Bc, &w);
  size_t N3 = 3 * (size_t)n;
  if (key == 1) {
    for (size_t x = 0; x < N3; ++x) {
      double m = 0.0;
      for (int k = 0; k <= ext; ++k) m = m + Bc[k] * c->hist[k * N3 + x];
      mu[x] = m;
    }
    for (int k = ext; k >= 1; --k) memcpy(&c->hist[k * N3], &c->hist[(k - 1) * N3], sizeof(double) * N3);
  } else {
    double *save = dalloc(N3);
    memcpy(save, mu, sizeof(double) * N3);
    induced_moment(c, E, mu);
    for (size_t x = 0; x < N3; ++x) mu[x] = w * mu[x] + (1.0 - w) * save[x];
    free(save);
  }
}

/* moment_from_pola_scf (src/moment_from_pola.f90:240-485). Returns mu_ind. */
static void scf(orc_ctx *c, const double *s, double *mu_ind, double *u_pol, int *n_scf, double *rmsd_out) {
  int n = c->n;
  size_t N3 = 3 * (size_t)n;
  const mdff_desc *d = &c->d;
  double *Es = dalloc(N3), *Ei = dalloc(N3), *E = dalloc(N3), *zq = dalloc(n);
  int ldip = 0;
  for (int i = 0; i < 48; ++i) if (d->dip[i] != 0.0) ldip = 1;
  es_task ts = {1, ldip, ldip, 1, 1, 0, 0};
  es_out o;
  multipole_es(c, s, c->q, c->dipia, ts, Es, NULL, &o);
  memcpy(E, Es, sizeof(double) * N3);
  es_task ti = {0, 0, 1, 0, 1, 0, 0};
  int iscf = 0;
  double rmsd = 1.79769313486231571e308;
  double upol = 0.0;
  while ((iscf < d->max_scf_pol_iter && rmsd > d->conv_tol_ind) || iscf < d->min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d->calc_md) induced_moment(c, E, mu_ind);
    else aspc(c, mu_ind, E, 1);
    upol = 0.0;
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) upol = upol + mu_ind[3 * i + a] * c->invpol[9 * i + a + 3 * b] * mu_ind[3 * i + b];
    upol = upol * 0.5;
    multipole_es(c, s, zq, mu_ind, ti, Ei, NULL, &o);
    for (size_t x = 0; x < N3; ++x) E[x] = Es[x] + Ei[x];
    if (iscf == 1 && d->calc_md) aspc(c, mu_ind, E, 2);
    rmsd = rmsd_mu(c, E, mu_ind);
  }
  memcpy(&c->hist[0], mu_ind, sizeof(double) * N3);
  *u_pol = upol;
  *n_scf = iscf;
  *rmsd_out = rmsd;
  free(Es); free(Ei); free(E); free(zq);
}

/* induced_moment_inner (src/moment_from_pola.f90:822-942): damped fixed-point
 * iteration mu <- (1-w) mu + w alpha (f_ext + E_dir(mu)) with a real-space
 * dipole field at alphaES = 0.001 (no reciprocal part), at most 10 steps,
 * until the residual falls below a tenth of |f_ext|. */
static void kO_inner(orc_ctx *c, const double *s, const double *fext, double *mu) {
  int n = c->n;
  size_t N3 = 3 * (size_t)n;
  double *zero = dalloc(N3), *f = dalloc(N3), *ft = dalloc(N3), *mprev = dalloc(N3), *zq = dalloc(n);
  double w = c->d.omegakO;
  double rmsd_ref = rmsd_mu(c, fext, zero);
  induced_moment(c, fext, mu);
  for (size_t x = 0; x < N3; ++x) mu[x] = w * mu[x];
  double rmsd_inner = 1.79769313486231571e308;
  es_task ti = {0, 0, 1, 0, 1, 0, 0, 1, 0.001};
  es_out o;
  for (int it = 1; it <= 10 && rmsd_inner > 0.1 * rmsd_ref; ++it) {
    memcpy(mprev, mu, sizeof(double) * N3);
    multipole_es(c, s, zq, mu, ti, f, NULL, &o);
    for (size_t x = 0; x < N3; ++x) ft[x] = fext[x] + f[x];
    rmsd_inner = rmsd_mu(c, ft, mu);
    induced_moment(c, ft, mu);
    for (size_t x = 0; x < N3; ++x) mu[x] = (1.0 - w) * mprev[x] + w * mu[x];
  }
  free(zero); free(f); free(ft); free(mprev); free(zq);
}

/* find_min (src/moment_from_pola.f90:2415-2455): line minimum between two
 * residuals g0, g1 along the step d1. */
static void kO_find_min(orc_ctx *c, const double *g0, const double *g1, const double *d1, double *gmin,
                        double *dmin) {
  int n = c->n;
  double g0sq = 0.0, g0g1mg0 = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!c->d.ldip_polar[c->itype[i] - 1]) continue;
    for (int k = 0; k < 3; ++k) {
      double a = g0[3 * i + k];
      g0sq = g0sq + a * a;
      g0g1mg0 = g0g1mg0 + a * (g1[3 * i + k] - a);
    }
  }
  double al = -g0sq / g0g1mg0;
  if (fabs(g0g1mg0) < 1e-10) al = 1.0;
  for (size_t x = 0; x < 3 * (size_t)n; ++x) {
    gmin[x] = g0[x] + al * (g1[x] - g0[x]);
    dmin[x] = al * d1[x];
  }
}

/* moment_from_pola_scf_kO_v2 (src/moment_from_pola.f90:1335-1639) */
static void scf_kO_v2(orc_ctx *c, const double *s, double *mu_ind, double *u_pol, int *n_scf, double *rmsd_out) {
  int n = c->n;
  size_t N3 = 3 * (size_t)n;
  const mdff_desc *d = &c->d;
  double *Es = dalloc(N3), *Ei = dalloc(N3), *E = dalloc(N3), *zq = dalloc(n);
  double *find = dalloc(N3), *dmu = dalloc(N3), *fprev = dalloc(N3), *gmin = dalloc(N3), *mprev = dalloc(N3),
         *dmin = dalloc(N3), *zero = dalloc(N3), *dd = dalloc(N3);
  int ldip = 0;
  for (int i = 0; i < 48; ++i) if (d->dip[i] != 0.0) ldip = 1;
  es_task ts = {1, ldip, ldip, 1, 1, 0, 0, 0, 0.0};
  es_out o;
  multipole_es(c, s, c->q, c->dipia, ts, Es, NULL, &o);
  memcpy(E, Es, sizeof(double) * N3);
  memcpy(find, E, sizeof(double) * N3);
  memset(mu_ind, 0, sizeof(double) * N3);
  es_task ti = {0, 0, 1, 0, 1, 0, 0, 0, 0.0};
  int iscf = 0;
  double rmsd = 1.79769313486231571e308, upol = 0.0;
  while ((iscf < d->max_scf_pol_iter && rmsd > d->conv_tol_ind) || iscf < d->min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d->calc_md) {
      kO_inner(c, s, iscf <= 2 ? find : gmin, dmu);
      for (size_t x = 0; x < N3; ++x) mu_ind[x] = mu_ind[x] + dmu[x];
    } else {
      aspc(c, mu_ind, E, 1);
    }
    upol = 0.0;
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) upol = upol + mu_ind[3 * i + a] * c->invpol[9 * i + a + 3 * b] * mu_ind[3 * i + b];
    upol = upol * 0.5;
    multipole_es(c, s, zq, mu_ind, ti, Ei, NULL, &o);
    for (size_t x = 0; x < N3; ++x) E[x] = Es[x] + Ei[x];
    if (iscf == 1 && d->calc_md) aspc(c, mu_ind, E, 2);
    for (int i = 0; i < n; ++i) {
      if (!d->ldip_polar[c->itype[i] - 1]) continue;
      for (int a = 0; a < 3; ++a) find[3 * i + a] = E[3 * i + a] - mu_ind[3 * i + a] / c->pol[9 * i + a + 3 * a];
    }
    if (iscf != 1) {
      for (size_t x = 0; x < N3; ++x) dd[x] = mu_ind[x] - mprev[x];
      kO_find_min(c, fprev, find, dd, gmin, dmin);
      for (size_t x = 0; x < N3; ++x) mu_ind[x] = mprev[x] + dmin[x];
      rmsd = rmsd_mu(c, gmin, zero);
    }
    memcpy(mprev, mu_ind, sizeof(double) * N3);
    memcpy(fprev, gmin, sizeof(double) * N3);
  }
  memcpy(&c->hist[0], mu_ind, sizeof(double) * N3);
  *u_pol = upol;
  *n_scf = iscf;
  *rmsd_out = rmsd;
  free(Es); free(Ei); free(E); free(zq); free(find); free(dmu); free(fprev); free(gmin); free(mprev);
  free(dmin); free(zero); free(dd);
}

/* moment_from_pola_scf_kO_v1 (src/moment_from_pola.f90:1042-1320): the inner
 * loop on the residual f = E - mu/alpha of the last iteration, no line
 * minimum; rmsd of (E, mu) every iteration (:1236). */
static void scf_kO_v1(orc_ctx *c, const double *s, double *mu_ind, double *u_pol, int *n_scf, double *rmsd_out) {
  int n = c->n;
  size_t N3 = 3 * (size_t)n;
  const mdff_desc *d = &c->d;
  double *Es = dalloc(N3), *Ei = dalloc(N3), *E = dalloc(N3), *zq = dalloc(n);
  double *find = dalloc(N3), *dmu = dalloc(N3);
  int ldip = 0;
  for (int i = 0; i < 48; ++i) if (d->dip[i] != 0.0) ldip = 1;
  es_task ts = {1, ldip, ldip, 1, 1, 0, 0, 0, 0.0};
  es_out o;
  multipole_es(c, s, c->q, c->dipia, ts, Es, NULL, &o);
  memcpy(E, Es, sizeof(double) * N3);
  memcpy(find, E, sizeof(double) * N3);
  memset(mu_ind, 0, sizeof(double) * N3);
  es_task ti = {0, 0, 1, 0, 1, 0, 0, 0, 0.0};
  int iscf = 0;
  double rmsd = 1.79769313486231571e308, upol = 0.0;
  while ((iscf < d->max_scf_pol_iter && rmsd > d->conv_tol_ind) || iscf < d->min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d->calc_md) {
      kO_inner(c, s, find, dmu);
      for (size_t x = 0; x < N3; ++x) mu_ind[x] = mu_ind[x] + dmu[x];
    } else {
      aspc(c, mu_ind, E, 1);
    }
    upol = 0.0;
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) upol = upol + mu_ind[3 * i + a] * c->invpol[9 * i + a + 3 * b] * mu_ind[3 * i + b];
    upol = upol * 0.5;
    multipole_es(c, s, zq, mu_ind, ti, Ei, NULL, &o);
    for (size_t x = 0; x < N3; ++x) E[x] = Es[x] + Ei[x];
    if (iscf == 1 && d->calc_md) aspc(c, mu_ind, E, 2);
    rmsd = rmsd_mu(c, E, mu_ind);
    for (int i = 0; i < n; ++i) {
      if (!d->ldip_polar[c->itype[i] - 1]) continue;
      for (int a = 0; a < 3; ++a) find[3 * i + a] = E[3 * i + a] - mu_ind[3 * i + a] / c->pol[9 * i + a + 3 * a];
    }
  }
  memcpy(&c->hist[0], mu_ind, sizeof(double) * N3);
  *u_pol = upol;
  *n_scf = iscf;
  *rmsd_out = rmsd;
  free(Es); free(Ei); free(E); free(zq); free(find); free(dmu);
}

/* ------------------------------------------------------------------------ */
/* engforce                                                                  */
/* ------------------------------------------------------------------------ */

/* SUBROUTINE engforce (src/engforce_driver.f90:102-190) */
int orc_engforce(orc_ctx *c, double *fx, double *fy, double *fz, mdff_terms *t) {
  const mdff_desc *d = &c->d;
  int n = c->n;
  mdff_terms tm;
  memset(&tm, 0, sizeof tm);
  if (d->lvnlist && !c->lists_valid) {
    int st = orc_vnlist_build(c);
    if (st) return st;
    tm.list_rebuilt = 1;
  }
  double *s = frac_coords(c);
  for (int i = 0; i < n; ++i) { c->fx[i] = 0.0; c->fy[i] = 0.0; c->fz[i] = 0.0; }
  sr_out so;
  if (d->lnmlj) {
    short_range(c, s, 1, &so);
    tm.u_lj = so.u; tm.vir_lj = so.vir;
    for (int k = 0; k < 9; ++k) { tm.tau_nonb[k] = so.tau[k]; tm.tau[k] += so.tau[k]; }
    tm.u_tail = c->u_tail; tm.p_tail = c->p_tail;
  }
  if (d->lbmhft || d->lbmhftd) {
    short_range(c, s, 2, &so);
    tm.u_bmhft = so.u; tm.vir_bmhft = so.vir;
    for (int k = 0; k < 9; ++k) { tm.tau_nonb[k] = so.tau[k]; tm.tau[k] += so.tau[k]; }
  }
  if (d->lmorse) {
    short_range(c, s, 3, &so);
    tm.u_morse = so.u; tm.vir_morse = so.vir;
    for (int k = 0; k < 9; ++k) { tm.tau_nonb[k] = so.tau[k]; tm.tau[k] += so.tau[k]; }
  }
  if (d->lcoulomb) {
    size_t N3 = 3 * (size_t)n;
    double *mu = dalloc(N3), *mu_ind = dalloc(N3), *ef = dalloc(N3);
    double upol = 0.0;
    if (c->linduced) {
      double *fsx = dalloc(n), *fsy = dalloc(n), *fsz = dalloc(n);
      memcpy(fsx, c->fx, sizeof(double) * n); memcpy(fsy, c->fy, sizeof(double) * n); memcpy(fsz, c->fz, sizeof(double) * n);
      if (d->algo_moment_from_pola == 1) scf_kO_v2(c, s, mu_ind, &upol, &tm.n_scf, &tm.rmsd);
      else if (d->algo_moment_from_pola == 2) scf_kO_v1(c, s, mu_ind, &upol, &tm.n_scf, &tm.rmsd);
      else scf(c, s, mu_ind, &upol, &tm.n_scf, &tm.rmsd);
      memcpy(c->fx, fsx, sizeof(double) * n); memcpy(c->fy, fsy, sizeof(double) * n); memcpy(c->fz, fsz, sizeof(double) * n);
      free(fsx); free(fsy); free(fsz);
    }
    for (size_t x = 0; x < N3; ++x) mu[x] = c->dipia[x] + mu_ind[x];
    /* polarisation catastrophe (src/moment_from_pola.f90:176-216): any component > 10 */
    for (size_t x = 0; x < N3; ++x)
      if (mu[x] > 10.0) {
        free(mu); free(mu_ind); free(ef); free(s);
        return fail(c, MDFF_ERR_POLAR_CATASTROPHE, "ERROR : POLARIZATION CATASTROPH");
      }
    es_task tk = {c->task_qq, c->task_qd, c->task_dd, 1, d->doefield, 1, 1};
    tk.efg = d->doefg;
    es_out o;
    multipole_es(c, s, c->q, mu, tk, ef, c->efg_t, &o);
    tm.u_dir = o.u_dir; tm.u_rec = o.u_rec; tm.u_self = o.u_self;
    tm.u_surf = d->lsurf ? o.u_surf : 0.0;
    tm.u_pol = upol;
    tm.u_coul = (o.u_dir + o.u_rec + tm.u_surf + o.u_self + upol) * d->coul_unit;
    for (int k = 0; k < 9; ++k) {
      tm.tau_coul[k] = o.tau_coul[k];
      tm.tau[k] += 2.0 * o.tau_coul[k];  /* added in multipole_ES and again in engforce */
    }
    tm.pvirial_coul = 1.0 / 3.0 * (o.tau_coul[0] + o.tau_coul[4] + o.tau_coul[8]) * d->press_unit;
    memcpy(c->mu_t, mu, sizeof(double) * N3);
    memcpy(c->ef_t, ef, sizeof(double) * N3);
    free(mu); free(mu_ind); free(ef);
  }
  memcpy(c->tau, tm.tau, sizeof c->tau);
  free(s);
  if (fx) orc_get_forces(c, fx, fy, fz);
  if (t) *t = tm;
  c->last_terms = tm;
  return MDFF_OK;
}

/* terms of the last force evaluation (engforce or an MD step) */
int orc_last_terms(orc_ctx *c, mdff_terms *t) {
  if (!c || !t) return MDFF_ERR_BAD_ARG;
  *t = c->last_terms;
  return MDFF_OK;
}

int orc_get_dipoles(orc_ctx *c, double *mu, double *ef) {
  if (mu) memcpy(mu, c->mu_t, sizeof(double) * 3 * c->n);
  if (ef) memcpy(ef, c->ef_t, sizeof(double) * 3 * c->n);
  return MDFF_OK;
}

/* efg_t (src/coulomb.f90:78) of the last engforce: zeros until a Coulomb
 * evaluation, then efg as multipole_ES leaves it */
int orc_get_efg(orc_ctx *c, double *efg) {
  if (!c || !efg) return MDFF_ERR_BAD_ARG;
  memcpy(efg, c->efg_t, sizeof(double) * 9 * (size_t)c->n);
  return MDFF_OK;
}

/* ------------------------------------------------------------------------ */
/* integrators                                                               */
/* ------------------------------------------------------------------------ */

static double upot_of(const mdff_terms *t) { return t->u_lj + t->u_bmhft + t->u_morse + t->u_coul; }

/* prop_velocity_verlet (src/integration.f90:132-194) followed by vnlistcheck
 * as md_run does (src/runmd.f90:294-304). */
int orc_md_vv(orc_ctx *c, int32_t nsteps, double dt, double *ekin, double *upot) {
  int n = c->n;
  double *fsx = dalloc(n), *fsy = dalloc(n), *fsz = dalloc(n);
  double dtsq2 = dt * dt * 0.5, dt2 = dt * 0.5;
  for (int st = 0; st < nsteps; ++st) {
    memcpy(fsx, c->fx, sizeof(double) * n); memcpy(fsy, c->fy, sizeof(double) * n); memcpy(fsz, c->fz, sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
      c->rx[i] = c->rx[i] + c->vx[i] * dt + (c->fx[i] * dtsq2) / c->mass[i];
      c->ry[i] = c->ry[i] + c->vy[i] * dt + (c->fy[i] * dtsq2) / c->mass[i];
      c->rz[i] = c->rz[i] + c->vz[i] * dt + (c->fz[i] * dtsq2) / c->mass[i];
    }
    mdff_terms t;
    c->itime++;   /* md%itime of this step (src/runmd.f90:257) */
    int rc = orc_engforce(c, NULL, NULL, NULL, &t);
    if (rc) { free(fsx); free(fsy); free(fsz); return rc; }
    double kin = 0.0;
    for (int i = 0; i < n; ++i) {
      c->vx[i] = c->vx[i] + (fsx[i] + c->fx[i]) * dt2 / c->mass[i];
      c->vy[i] = c->vy[i] + (fsy[i] + c->fy[i]) * dt2 / c->mass[i];
      c->vz[i] = c->vz[i] + (fsz[i] + c->fz[i]) * dt2 / c->mass[i];
      kin = kin + (c->vx[i] * c->vx[i] + c->vy[i] * c->vy[i] + c->vz[i] * c->vz[i]) * c->mass[i];
    }
    if (ekin) ekin[st] = kin * 0.5;
    if (upot) upot[st] = upot_of(&t);
    if (c->d.lvnlist) { rc = orc_vnlistcheck(c, NULL); if (rc) { free(fsx); free(fsy); free(fsz); return rc; } }
  }
  free(fsx); free(fsy); free(fsz);
  return MDFF_OK;
}

/* get_yoshida_weigth (src/integration.f90:1059-1104) */
void orc_yoshida(int order, double *w) {
  if (order == 3) {
    w[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0)); w[1] = 1.0 - 2.0 * w[0]; w[2] = w[0];
  } else if (order == 5) {
    /* the reference puts the middle weight last (not a symmetric composition) */
    w[0] = 1.0 / (4.0 - pow(4.0, 1.0 / 3.0)); w[1] = w[0]; w[2] = w[0]; w[3] = w[0]; w[4] = 1.0 - 4.0 * w[0];
  } else if (order == 7) {
    w[0] = .78451361047756; w[1] = .235573213359357; w[2] = -1.17767998417887;
    w[3] = 1.0 - 2.0 * (w[0] + w[1] + w[2]);
    w[4] = w[2]; w[5] = w[1]; w[6] = w[0];
  } else if (order == 9) {
    w[0] = 0.192; w[1] = 0.554910818409783619692725006662999;
    w[2] = 0.124659619941888644216504240951585; w[3] = -0.843182063596933505315033808282941;
    w[4] = 1.0 - 2.0 * (w[0] + w[1] + w[2] + w[3]);
    w[5] = w[3]; w[6] = w[2]; w[7] = w[1]; w[8] = w[0];
  } else {
    w[0] = 1.0;
  }
}

/* chain_nhcn (src/integration.f90:619-701), thermostat part: propagates
 * (xi, vxi) by one half step for kinetic energy kin and returns the velocity
 * scale s.  Like the reference, G(2..nhc-1) start at zero on every call (they
 * are set only by the forward sweep of the first Yoshida sub-step). */
double orc_nhc_chain(double kin, double *vxi, double *xi, const double *Q, double L, double temp,
                     double dt, int nhc, int yosh, int mults) {
  double w[9], G[16] = {0};
  orc_yoshida(yosh, w);
  double s = 1.0;
  G[0] = 2.0 * kin - L * temp;
  for (int k = 0; k < mults; ++k)
    for (int j = 0; j < yosh; ++j) {
      double dts = (w[j] * dt) / (double)mults;
      double dts2 = dts * 0.5, dts4 = dts2 * 0.5, dts8 = dts4 * 0.5;
      G[nhc - 1] = (vxi[nhc - 2] * vxi[nhc - 2] / Q[nhc - 2] - temp);
      vxi[nhc - 1] = vxi[nhc - 1] + G[nhc - 1] * dts4;
      for (int h = nhc - 2; h >= 0; --h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
      }
      s = s * exp(-vxi[0] * dts2 / Q[0]);
      for (int h = 0; h < nhc; ++h) xi[h] = xi[h] + vxi[h] * dts2 / Q[h];
      G[0] = 2.0 * kin * s * s - L * temp;
      for (int h = 0; h < nhc - 1; ++h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        G[h + 1] = (vxi[h] * vxi[h] / Q[h] - temp);
      }
      vxi[nhc - 1] = vxi[nhc - 1] + G[nhc - 1] * dts4;
    }
  return s;
}

/* chain_nhcn: the thermostat half step, then the velocities scaled by s */
static void chain_nhcn(orc_ctx *c, double *kin, double *vxi, double *xi, const double *Q, double L,
                       double temp, double dt, int nhc, int yosh, int mults) {
  double s = orc_nhc_chain(*kin, vxi, xi, Q, L, temp, dt, nhc, yosh, mults);
  double k2 = 0.0;
  for (int i = 0; i < c->n; ++i) {
    c->vx[i] = s * c->vx[i]; c->vy[i] = s * c->vy[i]; c->vz[i] = s * c->vz[i];
    k2 = k2 + (c->vx[i] * c->vx[i] + c->vy[i] * c->vy[i] + c->vz[i] * c->vz[i]) * c->mass[i];
  }
  *kin = k2 * 0.5;
}

static double kinetic(orc_ctx *c) {
  double e = 0.0;
  for (int i = 0; i < c->n; ++i)
    e = e + (c->vx[i] * c->vx[i] + c->vy[i] * c->vy[i] + c->vz[i] * c->vz[i]) * c->mass[i];
  return e * 0.5;
}

/* nose_hoover_chain_n (src/integration.f90:538-608) with prop_pos_vel_verlet (:396-447) */
int orc_md_nhcn(orc_ctx *c, int32_t nsteps, double dt, double temp, double timesca, int32_t nhc,
                int32_t yosh, int32_t mults, double *xi, double *vxi, double *ekin, double *upot,
                double *e_nvt) {
  int n = c->n;
  if (nhc < 2 || nhc > 16) return MDFF_ERR_BAD_ARG;
  double L = 3.0 * (double)n;
  double Q[16];
  for (int h = 0; h < nhc; ++h) Q[h] = pow(timesca, 2.0) * temp;
  Q[0] = Q[0] * L;
  double dt2 = dt * 0.5;
  for (int st = 0; st < nsteps; ++st) {
    double kin = kinetic(c);
    chain_nhcn(c, &kin, vxi, xi, Q, L, temp, dt, nhc, yosh, mults);
    for (int i = 0; i < n; ++i) {
      c->vx[i] = c->vx[i] + c->fx[i] * dt2 / c->mass[i];
      c->vy[i] = c->vy[i] + c->fy[i] * dt2 / c->mass[i];
      c->vz[i] = c->vz[i] + c->fz[i] * dt2 / c->mass[i];
      c->rx[i] = c->rx[i] + c->vx[i] * dt;
      c->ry[i] = c->ry[i] + c->vy[i] * dt;
      c->rz[i] = c->rz[i] + c->vz[i] * dt;
    }
    mdff_terms t;
    c->itime++;   /* md%itime of this step (src/runmd.f90:257) */
    int rc = orc_engforce(c, NULL, NULL, NULL, &t);
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
      c->vx[i] = c->vx[i] + c->fx[i] * dt2 / c->mass[i];
      c->vy[i] = c->vy[i] + c->fy[i] * dt2 / c->mass[i];
      c->vz[i] = c->vz[i] + c->fz[i] * dt2 / c->mass[i];
    }
    kin = kinetic(c);
    chain_nhcn(c, &kin, vxi, xi, Q, L, temp, dt, nhc, yosh, mults);
    kin = kinetic(c);
    if (ekin) ekin[st] = kin;
    if (upot) upot[st] = upot_of(&t);
    if (e_nvt) {
      double e3 = 0.0, e4 = 0.0;
      for (int h = 1; h < nhc; ++h) { e3 = e3 + vxi[h] * vxi[h] * 0.5 / Q[h]; e4 = e4 + xi[h]; }
      e_nvt[st] = L * temp * xi[0] + vxi[0] * vxi[0] * 0.5 / Q[0] + e3 + temp * e4;
    }
    if (c->d.lvnlist) { rc = orc_vnlistcheck(c, NULL); if (rc) return rc; }
  }
  return MDFF_OK;
}

/* rescale_velocities with lcsvr (src/kinetic.f90:122-198) and resamplekin
 * (:560-635); the Gaussian deviates are inputs. */
int orc_csvr_rescale(orc_ctx *c, double temp, double taut, double r1, double sumnoise, double *conint) {
  int n = c->n;
  double ekin = kinetic(c);
  int ndeg = 3 * n - 3;
  double sigma = ndeg * temp * 0.5;
  double factor = taut > 0.1 ? exp(-1.0 / taut) : 0.0;
  double knew = ekin + (1.0 - factor) * (sigma * (sumnoise + r1 * r1) / ndeg - ekin) +
                2.0 * r1 * sqrt(ekin * sigma / ndeg * (1.0 - factor) * factor);
  if (conint) *conint = *conint - (knew - ekin);
  double lambda = sqrt(knew / ekin);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int i = 0; i < n; ++i) {
    c->vx[i] *= lambda; c->vy[i] *= lambda; c->vz[i] *= lambda;
  }
  for (int i = 0; i < n; ++i) { sx += c->vx[i]; sy += c->vy[i]; sz += c->vz[i]; }
  sx /= (double)n; sy /= (double)n; sz /= (double)n;
  for (int i = 0; i < n; ++i) { c->vx[i] -= sx; c->vy[i] -= sy; c->vz[i] -= sz; }
  return MDFF_OK;
}

/* rescale_velocities with a given factor (src/kinetic.f90:155-180) */
int orc_rescale_velocities(orc_ctx *c, double lambda) {
  int n = c->n;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int i = 0; i < n; ++i) {
    c->vx[i] *= lambda; c->vy[i] *= lambda; c->vz[i] *= lambda;
  }
  for (int i = 0; i < n; ++i) { sx += c->vx[i]; sy += c->vy[i]; sz += c->vz[i]; }
  sx /= (double)n; sy /= (double)n; sz /= (double)n;
  for (int i = 0; i < n; ++i) { c->vx[i] -= sx; c->vy[i] -= sy; c->vz[i] -= sz; }
  return MDFF_OK;
}

/* calc_temp (src/kinetic.f90:527-552) */
int orc_calc_temp(orc_ctx *c, double boltz_unit, double *temp, double *ekin) {
  double e = kinetic(c);
  if (ekin) *ekin = e;
  if (temp) *temp = 2.0 * e / boltz_unit / (3.0 * c->n);
  return MDFF_OK;
}
