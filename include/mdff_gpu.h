/*
 * mdff_gpu.h -- C-ABI of the MI355X (gfx950) engforce library.
 *
 * Drop-in boundary for MDFF's force/energy/stress/induced-dipole evaluation
 * (reference: src/engforce_driver.f90:102-190, SUBROUTINE engforce).  The
 * reference routine takes no arguments and works on module globals; this C-ABI
 * replaces those globals by one context object (mdff_ctx) created from a plain
 * descriptor (mdff_desc) whose fields mirror the reference namelists:
 *   &controltag   src/control.f90:47-66        (lnmlj, lbmhft, lbmhftd, lcoulomb, lsurf, ...)
 *   &non_bondedtag src/non_bonded.f90:76-116    (LJ n-m, Morse, BMHFT(D) tables)
 *   &coulombtag   src/coulomb.f90:47-79        (alphaES, kES, doefield, doefg)
 *   &fieldtag     src/field.f90:113-188        (mass, qch, dip)
 *   &pimtag       src/pim.f90:63-96            (poldip, TT damping tables, SCF controls)
 *   &mdtag        src/md.f90:37                (itime for the ASPC dipole predictor)
 *
 * Conventions
 *  - Plain C types only: int32_t, int64_t, double, pointers and sizes.
 *  - 3x3 matrices are stored column-major exactly as Fortran stores them:
 *    M(i,j) (1-based) lives at m[(i-1) + 3*(j-1)].  cell[] is simu_cell%A:
 *    column j is lattice vector j (src/lattice.f90:33-45).
 *  - Type tables are ntypemax x ntypemax (ntypemax = 16, src/config.f90:42),
 *    column-major: T(it,jt) at t[(it-1) + 16*(jt-1)].  Three-index damping
 *    tables (src/pim.f90:78-81) use t[(k-1) + 16*(i-1) + 256*(j-1)].
 *  - Per-atom vectors: positions, velocities and forces are SoA (rx,ry,rz as in
 *    src/config.f90:61-63); dipoles and fields are AoS 3 x natm as in the
 *    reference arrays mu(3,natm), ef(3,natm) (src/coulomb.f90:75-78).
 *  - itype[] is 1-based (1..ntype), as in src/config.f90:49.
 *  - Units are the reference's internal units (src/constants.f90:20-50):
 *    Angstrom, eV, amu; stresses are returned in GPa (divided by press_unit).
 *  - Every entry point returns an mdff_status.  The reference prints and STOPs
 *    on these conditions (src/tools.f90:337-340, src/moment_from_pola.f90:176-216,
 *    src/pim.f90:201-204); the C side returns the code and keeps a message
 *    readable through mdff_last_error().
 *  - A context owns two HIP streams (default priority) on the device current
 *    at creation: its work stream and one on which the pair kernel runs beside
 *    the reciprocal kernels; every call returns with both joined.  It is not
 *    thread safe (the reference is single-threaded per rank).
 */
#ifndef MDFF_GPU_H
#define MDFF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDFF_NTYPEMAX 16
#define MDFF_ABI_VERSION 3

typedef enum mdff_status {
  MDFF_OK = 0,
  MDFF_ERR_BAD_ARG = 1,            /* invalid descriptor / argument                 */
  MDFF_ERR_LIST_OVERFLOW = 2,      /* Verlet list capacity exceeded (tools.f90:337) */
  MDFF_ERR_SCF_NOT_CONVERGED = 3,  /* reserved: like the reference, the SCF stops at
                                      max_scf_pol_iter and the call succeeds
                                      (src/moment_from_pola.f90:365) */
  MDFF_ERR_POLAR_CATASTROPHE = 4,  /* some mu component > 10 e.A (moment_from_pola.f90:176) */
  MDFF_ERR_HIP = 5,                /* HIP runtime error                             */
  MDFF_ERR_NO_DEVICE = 6,          /* no usable gfx950 device                       */
  MDFF_ERR_UNSUPPORTED = 7         /* option outside what this library implements  */
} mdff_status;

typedef enum mdff_precision {
  MDFF_PREC_FP64 = 0,   /* every operation in FP64 (parity form)                         */
  MDFF_PREC_MIXED = 1   /* FP32 pair / dipole-field arithmetic, FP64 force and field
                           accumulation (pair energies / stresses: FP32 within one
                           lane's row share, FP64 beyond), FP64 minimum image,
                           FP64 k.r phases, FP64 SCF residual, the reciprocal box
                           |k| <= 2.5 alpha of the force evaluation in FP64          */
} mdff_precision;

typedef enum mdff_trunc {  /* src/non_bonded.f90:54-56 ctrunc */
  MDFF_TRUNC_NONE = 0, MDFF_TRUNC_LINEAR = 1, MDFF_TRUNC_QUADRATIC = 2
} mdff_trunc;

/* Descriptor: everything the reference reads from its namelists and from POSFF
 * headers.  Zero-initialise it and fill what is used (defaults of the
 * reference are listed per field; mdff_desc_defaults() writes them). */
typedef struct mdff_desc {
  int32_t abi_version;             /* = MDFF_ABI_VERSION                              */
  int32_t natm;                    /* number of ions                                  */
  int32_t ntype;                   /* number of types (<= 16)                         */
  int32_t precision;               /* mdff_precision                                  */
  double  cell[9];                 /* simu_cell%A, column-major                       */

  /* &controltag (src/control.f90:201-245) */
  int32_t lnmlj, lbmhft, lbmhftd, lmorse, lcoulomb, lsurf, lvnlist;
  int32_t calc_md;                 /* 1: calc='md' (ASPC predictor active), 0: other   */
  double  cutshortrange;           /* vdW cutoff                                      */
  double  cutlongrange;            /* real-space Ewald cutoff (clamped, coulomb.f90:377-383) */
  double  skindiff;                /* Verlet skin, default 0.15                       */

  /* units (src/constants.f90:42-50; lreduced sets both to 1) */
  double  coul_unit;               /* 14.3996441494161                                */
  double  press_unit;              /* 0.00624150964712042                             */

  /* &non_bondedtag (src/non_bonded.f90:76-116), symmetric type tables */
  int32_t trunc;                   /* mdff_trunc, default NONE                        */
  int32_t pad0;
  double  epslj[256], sigmalj[256], plj[256], qlj[256];
  double  epsmor[256], sigmamor[256], rhomor[256];
  double  Abmhftd[256], Bbmhftd[256], Cbmhftd[256], Dbmhftd[256], BDbmhftd[256];

  /* &coulombtag (src/coulomb.f90:160-180) */
  double  alphaES;                 /* Ewald alpha, default 1.0                        */
  int32_t kES[3];                  /* kmax per reciprocal direction, default 10       */
  int32_t doefield;                /* return the electric field                       */

  /* &fieldtag (src/field.f90:113-188): per type */
  double  mass[16];
  double  qch[16];
  double  dip[48];                 /* dip(it,a) at [it-1 + 16*(a-1)] (static dipoles)  */

  /* &pimtag (src/pim.f90:107-140) */
  int32_t ldip_polar[16];
  double  poldip[144];             /* poldip(it,a,b) at [it-1 + 16*(a-1) + 48*(b-1)]  */
  int32_t ldip_damping[4096];
  double  pol_damp_b[4096];
  double  pol_damp_c[4096];
  int32_t pol_damp_k[4096];
  double  conv_tol_ind;            /* default 1e-6                                    */
  int32_t min_scf_pol_iter;        /* default 3                                       */
  int32_t max_scf_pol_iter;        /* default 100                                     */
  int32_t extrapolate_order;       /* ASPC order 0..4, default 0                      */
  int32_t algo_moment_from_pola;   /* 0: 'scf' (default), 1: 'scf_kO_v2', 2: 'scf_kO_v1' */
  double  omegakO;                 /* mixing of the scf_kO_v2 inner loop, default 0.7 */

  /* &coulombtag, continued */
  int32_t doefg;                   /* compute the electric field gradient (efg_t)     */
  int32_t pad1;
} mdff_desc;

/* Energies, virials and stress of one engforce call (src/thermo.f90:41-46,
 * src/config.f90:55-57, src/ewald.f90:237-246,310). */
typedef struct mdff_terms {
  double u_lj, u_bmhft, u_morse;   /* short-range energies (eV)                       */
  double vir_lj, vir_bmhft, vir_morse;
  double u_coul;                   /* (u_dir+u_rec+u_self[+u_surf]+u_pol)*coul_unit   */
  double u_pol;                    /* 1/2 mu.alpha^-1.mu, internal units (not scaled)  */
  double u_dir, u_rec, u_self, u_surf; /* Ewald parts, internal units (x coul_unit = eV) */
  double pvirial_coul;
  double u_tail, p_tail;           /* LJ long-range corrections (non_bonded.f90:610-633) */
  double tau[9];                   /* total stress as the reference accumulates it    */
  double tau_nonb[9];              /* short-range stress, GPa                         */
  double tau_coul[9];              /* Coulomb stress, GPa                             */
  int32_t n_scf;                   /* SCF iterations of the induced-dipole solve      */
  int32_t list_rebuilt;            /* 1 if this call rebuilt the Verlet lists         */
  double rmsd;                     /* final SCF residual                              */
} mdff_terms;

typedef struct mdff_ctx mdff_ctx;

/* Fill a descriptor with the reference defaults (src/control.f90:201-245,
 * src/non_bonded.f90:206-240, src/coulomb.f90:160-180, src/pim.f90:107-140). */
void mdff_desc_defaults(mdff_desc *d);

/* Create a context: replaces the setup chain read_pos -> field_init ->
 * initialize_param_non_bonded / ewald_param / initialize_coulomb /
 * kpoint_sum_init / pim_init / typeinfo_init (src/read_pos.f90:36-148,
 * src/non_bonded.f90:455-644, src/coulomb.f90:363-493, src/kspace.f90:143-342,
 * src/pim.f90:246-315, src/read_pos.f90:160-234).  itype: natm 1-based types. */
int mdff_ctx_create(const mdff_desc *d, const int32_t *itype, mdff_ctx **out);
int mdff_ctx_destroy(mdff_ctx *ctx);
const char *mdff_last_error(const mdff_ctx *ctx);
const char *mdff_version(void);

/* Copy host SoA positions in (Cartesian, Angstrom).  Replaces the module
 * arrays rx,ry,rz (src/config.f90:61). */
int mdff_set_positions(mdff_ctx *ctx, const double *rx, const double *ry, const double *rz);
int mdff_get_positions(mdff_ctx *ctx, double *rx, double *ry, double *rz);
int mdff_set_velocities(mdff_ctx *ctx, const double *vx, const double *vy, const double *vz);
int mdff_get_velocities(mdff_ctx *ctx, double *vx, double *vy, double *vz);
int mdff_get_forces(mdff_ctx *ctx, double *fx, double *fy, double *fz);

/* New cell (src/lattice.f90:60-168).  Like the reference, the k-mesh built at
 * creation is kept (src/coulomb.f90:485 builds it once). */
int mdff_set_cell(mdff_ctx *ctx, const double cell[9]);

/* Verlet half lists: vnlist_pbc (src/tools.f90:273-368) and vnlistcheck
 * (src/tools.f90:444-537).  *rebuilt (may be NULL) is set to 1 on rebuild. */
int mdff_vnlist_build(mdff_ctx *ctx);
int mdff_vnlistcheck(mdff_ctx *ctx, int32_t *rebuilt);

/* Read back a half list in the reference CSR layout: point[natm+1] and
 * list[*n] hold 1-based indices exactly as verlet_vdw / verlet_coul
 * (src/config.f90:88-95).  which: 0 = vdw, 1 = coul.  If cap is smaller than
 * the list, *n is set and MDFF_ERR_BAD_ARG returned without copying. */
int mdff_get_pairs(mdff_ctx *ctx, int32_t which, int32_t *point, int32_t *list,
                   int64_t cap, int64_t *n);

/* The drop-in for SUBROUTINE engforce (src/engforce_driver.f90:102-190).
 * Forces are left on the device; fetch them with mdff_get_forces or pass
 * non-NULL fx/fy/fz to have them copied.  terms may be NULL. */
int mdff_engforce(mdff_ctx *ctx, double *fx, double *fy, double *fz, mdff_terms *terms);

/* Energies, virials and stress of the last force evaluation, whether made by
 * mdff_engforce or by the last step of mdff_md_vv / mdff_md_nhcn: the module
 * globals the reference's calc_thermo reads after each step
 * (src/thermo.f90:122-211). */
int mdff_last_terms(mdff_ctx *ctx, mdff_terms *terms);

/* Total dipoles mu_t and field ef_t of the last engforce (src/coulomb.f90:75-78),
 * AoS 3 x natm.  Either pointer may be NULL. */
int mdff_get_dipoles(mdff_ctx *ctx, double *mu, double *ef);

/* Electric field gradient efg_t of the last force evaluation
 * (src/coulomb.f90:78, set at src/engforce_driver.f90:171), 9 x natm in the
 * reference layout efg(3,3,natm): component (a,b) of atom i at
 * efg[9i + (a-1) + 3(b-1)], in eV/(e A^2) (x coul_unit, src/ewald.f90:231).
 * As in the reference (src/ewald.f90:208-216,231; src/ewald_rec.f90:209-216):
 *  - with doefg, real space (-q_j T2 for the q-q task; T3.mu_j for the mu-mu
 *    task between two non-zero dipoles, src/ewald_dir.f90:543-546,577-582),
 *    reciprocal space (upper triangle and diagonal only) and the self term
 *    -2 selfa2 q_i on the diagonal;
 *  - without doefg, the self term alone;
 *  - without lcoulomb, zeros.
 * The lower triangle therefore lacks the reciprocal part; the EFGALL file
 * prints xx, yy, zz, xy, xz, yz (src/field.f90:423-425). */
int mdff_get_efg(mdff_ctx *ctx, double *efg);

/* MD step counter used by the ASPC predictor (md%itime, src/md.f90:37). */
int mdff_set_itime(mdff_ctx *ctx, int32_t itime);

/* Device-resident velocity-Verlet steps (src/integration.f90:132-194): for
 * each step r += v dt + f dt^2/2m, engforce, v += (f_old+f) dt/2m, then the
 * skin check (src/runmd.f90:294-304).  dt in internal time units.  ekin/upot
 * (may be NULL) receive per-step kinetic and potential energy (eV), nsteps
 * entries each.  Every step advances the step counter (md%itime,
 * mdff_set_itime) by one before its force call, as the reference's MD loop
 * does (src/runmd.f90:257); the ASPC order follows it. */
int mdff_md_vv(mdff_ctx *ctx, int32_t nsteps, double dt, double *ekin, double *upot);

/* Nose-Hoover chain NVT steps (src/integration.f90:538-701): chain_nhcn,
 * prop_pos_vel_verlet, chain_nhcn.  temp = kB*T (eV), Q from timesca_thermo
 * as in the reference.  xi/vxi: nhc_n thermostat positions/momenta, updated
 * in place.  e_nvt (may be NULL) receives the extended-system energy per step.
 * Each step advances md%itime by one, as mdff_md_vv. */
int mdff_md_nhcn(mdff_ctx *ctx, int32_t nsteps, double dt, double temp,
                 double timesca_thermo, int32_t nhc_n, int32_t yosh_order,
                 int32_t nhc_mults, double *xi, double *vxi,
                 double *ekin, double *upot, double *e_nvt);

/* Bussi CSVR rescale (src/kinetic.f90:122-198) with the random numbers of
 * resamplekin supplied by the caller: r1 ~ N(0,1), sumnoise = sum of
 * (3N-4) squared N(0,1) deviates.  Removes the COM velocity afterwards.
 * *conint accumulates the conserved-quantity integral (src/thermo.f90:190-193). */
int mdff_csvr_rescale(mdff_ctx *ctx, double temp, double taut, double r1,
                      double sumnoise, double *conint);

/* rescale_velocities with the factor given (src/kinetic.f90:122-198): v *=
 * lambda, then the mean velocity is removed.  For the Berendsen rescale the
 * caller forms lambda = sqrt(1 + dt/tauTberendsen (T0/T - 1)) (or the
 * annealing factor) from mdff_calc_temp. */
int mdff_rescale_velocities(mdff_ctx *ctx, double lambda);

/* Kinetic energy and temperature (src/kinetic.f90:527-552). */
int mdff_calc_temp(mdff_ctx *ctx, double boltz_unit, double *temp, double *ekin);

/* mdff_calc_temp for velocities already in device memory, with no allocation
 * of any kind.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops).  n atoms; mass, vx, vy, vz: device arrays of n doubles; out: device
 * memory for 2 doubles, receiving the kinetic energy (eV) and the
 * temperature (K, with boltz_unit); stream: a hipStream_t (NULL: the null
 * stream).  Asynchronous on that stream; the same sums, in the same order,
 * as mdff_calc_temp.  The block partials live in a fixed device array of
 * the library (1,026 doubles per device, kept until the process ends), so
 * any natm fits (at most 1,024 blocks, each thread striding over the atoms);
 * calls that overlap on one device race on that array.  Returns
 * MDFF_ERR_BAD_ARG for n <= 0, null pointers or boltz_unit <= 0. */
int mdff_calc_temp_noalloc(int32_t n, const double *mass, const double *vx, const double *vy,
                           const double *vz, double boltz_unit, void *stream, double *out);

/* Timing of the last engforce on the context stream (HIP events), ms:
 * [0] total, [1] list build/check, [2] short-range+real-space, [3] reciprocal,
 * [4] SCF loop.  n = number of entries written (<= 8); the events are
 * recorded only while mdff_profile is on, n = 0 for an evaluation without. */
int mdff_last_timings(mdff_ctx *ctx, double *ms, int32_t *n);

/* Kernel timing with HIP events on the context stream; replaces the
 * reference's MPI_WTIME accumulators (src/time.f90:36-127, src/symbol.h:1-21).
 * mdff_profile(ctx, 1) clears and starts recording one event pair around
 * each timed launch; mdff_profile_read returns, per kernel class k <
 * nkinds, the summed duration (ms) and launch count; mdff_profile_name(k)
 * names the class (9 classes). */
int mdff_profile(mdff_ctx *ctx, int32_t on);
int mdff_profile_read(mdff_ctx *ctx, double *ms_total, int32_t *count, int32_t nkinds);
const char *mdff_profile_name(int32_t kind);

/* ------------------------------------------------------------------------
 * Batched force-field fit.  No counterpart in the reference (it has no
 * gradient code); it serves fitting of the engforce parameters on many
 * fixed snapshots.  The forward of each snapshot is the engforce of this
 * library restricted to Buckingham (lbmhft) + Ewald + PIM with TT damping:
 * forces, stress (tau_nonb + tau_coul, GPa) and the SCF dipoles.  The loss is
 *   L = sum_s [ sum_i |F_i - F*_i|^2 + w sum_ab (tau_ab - tau*_ab)^2 ]
 * and its gradient with respect to the parameter vector theta is computed
 * through the converged SCF by one adjoint solve per snapshot.  All
 * snapshots run together in a sequence of launches, each over the blocks of
 * atoms or k-vectors of every snapshot.
 * ---------------------------------------------------------------------- */
typedef struct mdff_fit mdff_fit;

/* theta entry kinds: spec[3p] = kind, spec[3p+1] = t1, spec[3p+2] = t2 (1-based types) */
#define MDFF_FIT_A      0   /* Buckingham A of pair (t1,t2)                      */
#define MDFF_FIT_RHO    1   /* Buckingham rho = 1/B of pair (t1,t2)              */
#define MDFF_FIT_C      2   /* Buckingham C of pair (t1,t2)                      */
#define MDFF_FIT_ALPHA  3   /* isotropic polarisability of type t1               */
#define MDFF_FIT_TTB    4   /* TT b of the field at a dipole of type t1 from charges of type t2 */
#define MDFF_FIT_TTC    5   /* TT c, same entry                                   */
#define MDFF_FIT_MAXPARAM 32

/* d: the force field (types <= 4, charges, masses, TT orders, unit constants,
 * conv_tol_ind / min / max SCF iterations for both the SCF and the adjoint
 * solve, values of the parameters that are not fitted).  Per snapshot s:
 * natm[s], cells[9s..9s+8] (column-major), cut_alpha[3s..3s+2] =
 * (cutshortrange, cutlongrange, alphaES), kmax[3s..3s+2].  itype and pos
 * (AoS 3 x total atoms, Cartesian) are the snapshots concatenated. */
int mdff_fit_create(const mdff_desc *d, int32_t nsnap, const int32_t *natm,
                    const double *cells, const double *cut_alpha, const int32_t *kmax,
                    const int32_t *itype, const double *pos, int32_t nparam,
                    const int32_t *spec, mdff_fit **out);
/* targets: forces AoS 3 x total atoms (eV/A), stress 9 x nsnap (GPa), weight w */
int mdff_fit_set_targets(mdff_fit *f, const double *fref, const double *tauref, double w_stress);
/* block: threads per workgroup of the phase launches (a power of two, 8 to
 * 256); every snapshot's atoms and k-vectors are cut into blocks of that
 * many.  grid (>= 1) is not used: the workgroup counts follow the blocks. */
int mdff_fit_set_launch(mdff_fit *f, int32_t block, int32_t grid);
/* Evaluate at theta: loss; grad[nparam] if want_grad; optional forces
 * (3 x total atoms), stress (9 x nsnap) and iteration counts nscf[2s] (SCF),
 * nscf[2s+1] (adjoint).  precision: MDFF_PREC_FP64 or MDFF_PREC_MIXED. */
int mdff_fit_eval(mdff_fit *f, const double *theta, int32_t precision, int32_t want_grad,
                  double *loss, double *grad, double *forces, double *stress, int32_t *nscf);
int mdff_fit_destroy(mdff_fit *f);
const char *mdff_fit_last_error(const mdff_fit *f);

#ifdef __cplusplus
}
#endif
#endif /* MDFF_GPU_H */
