/*
 * mdff_oracle.h -- CPU restatement of MDFF's engforce path (test infrastructure).
 *
 * THIS IS THE ORACLE.  It exists only to check the GPU library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libmdff_gpu.so) never calls it.
 *
 * Every function restates the reference algorithm (Fortran, FilipeVasconcelos/mdff)
 * in plain C with FP64 arithmetic in the reference's operation order, compiled
 * with -ffp-contract=off so no FMA contraction changes a rounding.  The API
 * mirrors include/mdff_gpu.h (orc_* <-> mdff_*) so tests drive both the same way.
 *
 * Pinning (what checks each function against something other than itself) is
 * listed in the header comment of mdff_oracle.c.
 */
#ifndef MDFF_ORACLE_H
#define MDFF_ORACLE_H
#include <stdint.h>
#include "../include/mdff_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* scalar helpers (exported so tests can pin them one by one) */
double orc_nint(double x);                                   /* Fortran NINT    */
double orc_errfc(double x);                                  /* functions.f90   */
void   orc_tt_damp(double b, double c, double r, int order, double *f, double *fd);
void   orc_lattice(const double A[9], double B[9], double *omega, double W[3]);
int    orc_kmesh_count(const int32_t kmax[3]);
int    orc_kmesh(const double B[9], double alpha, const int32_t kmax[3],
                 double *kx, double *ky, double *kz, double *kk, double *Ak, double *kcoe);

int    orc_create(const mdff_desc *d, const int32_t *itype, orc_ctx **out);
void   orc_destroy(orc_ctx *c);
void   orc_set_threads(orc_ctx *c, int nthreads);
const char *orc_last_error(const orc_ctx *c);
int    orc_set_positions(orc_ctx *c, const double *rx, const double *ry, const double *rz);
int    orc_get_positions(orc_ctx *c, double *rx, double *ry, double *rz);
int    orc_set_velocities(orc_ctx *c, const double *vx, const double *vy, const double *vz);
int    orc_get_velocities(orc_ctx *c, double *vx, double *vy, double *vz);
int    orc_get_forces(orc_ctx *c, double *fx, double *fy, double *fz);
int    orc_set_itime(orc_ctx *c, int32_t itime);
int    orc_vnlist_build(orc_ctx *c);
int    orc_vnlistcheck(orc_ctx *c, int32_t *rebuilt);
int    orc_get_pairs(orc_ctx *c, int32_t which, int32_t *point, int32_t *list, int64_t cap, int64_t *n);
int    orc_engforce(orc_ctx *c, double *fx, double *fy, double *fz, mdff_terms *t);
int    orc_get_dipoles(orc_ctx *c, double *mu, double *ef);
int    orc_get_efg(orc_ctx *c, double *efg);
int    orc_last_terms(orc_ctx *c, mdff_terms *t);
int    orc_rescale_velocities(orc_ctx *c, double lambda);
int    orc_get_cutoffs(orc_ctx *c, double *rc_vdw, double *rc_coul, double *alpha, int32_t *nk);

/* Ewald parts on their own (multipole_ES with explicit mu and task), for the
 * Nymand-Linse and Madelung pins.  task[0..2] = q-q, q-mu, mu-mu. */
int    orc_multipole_es(orc_ctx *c, const double *mu, const int32_t task[3], int32_t damp,
                        double *ef, double *fx, double *fy, double *fz, double tau_coul[9],
                        double *u_dir, double *u_rec, double *u_self, double *u_surf);

/* ASPC predictor coefficients and corrector weight of order ext (0..4) */
void   orc_aspc_coeffs(int ext, double B[6], double *w);
/* Yoshida-Suzuki weights of order 1, 3, 5, 7 or 9 (w[0..order)) */
void   orc_yoshida(int order, double *w);
/* thermostat half step of chain_nhcn for kinetic energy kin; returns the
 * velocity scale */
double orc_nhc_chain(double kin, double *vxi, double *xi, const double *Q, double L, double temp,
                     double dt, int nhc, int yosh, int mults);

/* induced-dipole field operator E_ind(mu) (charges zero, undamped) */
int    orc_dipole_field(orc_ctx *c, const double *mu, double *ef);

/* Integrators and thermostat (src/integration.f90, src/kinetic.f90) */
int    orc_md_vv(orc_ctx *c, int32_t nsteps, double dt, double *ekin, double *upot);
int    orc_md_nhcn(orc_ctx *c, int32_t nsteps, double dt, double temp, double timesca_thermo,
                   int32_t nhc_n, int32_t yosh_order, int32_t nhc_mults, double *xi, double *vxi,
                   double *ekin, double *upot, double *e_nvt);
int    orc_csvr_rescale(orc_ctx *c, double temp, double taut, double r1, double sumnoise,
                        double *conint);
int    orc_calc_temp(orc_ctx *c, double boltz_unit, double *temp, double *ekin);

#ifdef __cplusplus
}
#endif
#endif
