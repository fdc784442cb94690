// mdff_gpu.hip -- context, parameter tables and the engforce driver (C-ABI).
//
// Drop-in for SUBROUTINE engforce (src/engforce_driver.f90:102-190) and the
// setup chain that feeds it (src/read_pos.f90:160-234,
// src/non_bonded.f90:455-644, src/coulomb.f90:363-493, src/pim.f90:147-240).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "common.h"
#include "reduce.h"

int realspace_force_impl(mdff_ctx* c, double* fsr, double* fcl, double* ef, bool with_field);

namespace {

constexpr int NPA = 15;  // per-atom columns of the pair kernel

// -------------------------------------------------------------- small kernels

__global__ void k_fill(size_t n, double v, double* p) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// column sums over atoms: in [col][n] (ncol <= 16) -> block partials
// [block][ncol], then one block adds them in block order -> out[ncol]

// fixed-order sum of a 256-thread block: wave shuffles, then the 4 wave sums
// in wave order (deterministic); result valid in thread 0
__device__ __forceinline__ double block_sum256(double v, double* sh4) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((sh4[0] + sh4[1]) + sh4[2]) + sh4[3];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_colsum(int n, int ncol, const double* __restrict__ in,
                                                double* __restrict__ part) {
  __shared__ double sh[4 * kRedMaxCols];
  colsum_partials(n, ncol, in, part, sh);
}

// out[0..ncol) = sums of the block partials in block order (one block)
__global__ void __launch_bounds__(256) k_colsum_final(int nb, int ncol, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  __shared__ double sh[4 * kRedMaxCols];
  sum_block_partials(nb, ncol, part, out, sh);
}

// the queued final sums, one block each
// (+ the skin check's completion signal: seq_val into mapped host memory;
// this launch starts after the check kernel has completed)
__global__ void __launch_bounds__(256) k_final_sums(FinalBatch a) {
  __shared__ double sh[4 * kRedMaxCols];
  run_final_batch(a, blockIdx.x, sh);
}

// Total force f = f_sr + (f_rec + f_dir) * coul_unit (src/ewald.f90:237-246),
// AoS inputs, SoA output; the pair kernel delivers f_sr + coul_unit f_dir
// (combined per pair) in fsr; f_rec from the array frec, or finished here
// from the reciprocal split partials (rf_on).  The same blocks then write the
// block partials of the pair kernel's per-atom columns (column sums over
// atoms, as k_colsum).
struct ForceSumArgs {
  int n, ncol;
  const double* fsr;
  const double* frec;
  RecFinal rf;
  int rf_on, has_cl;
  double coul;
  double *fx, *fy, *fz;
  const double* in;              // [ncol][n]
  double* part;                  // [block][ncol]
};
__global__ void __launch_bounds__(256) k_force_sums(ForceSumArgs a) {
  __shared__ double sh[4 * kRedMaxCols];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    double fr[3] = {0.0, 0.0, 0.0};
    if (a.has_cl) {
      if (a.rf_on) {
        double v[9], g[3];
        rec_final_moments(a.rf, i, v, g);
        rec_final_force(a.rf, i, v, g, fr);
      } else {
        fr[0] = a.frec[3 * i];
        fr[1] = a.frec[3 * i + 1];
        fr[2] = a.frec[3 * i + 2];
      }
    }
    double f[3];
    for (int d = 0; d < 3; ++d) {
      // fsr: the pair kernel's f_sr + coul_unit f_dir, combined per pair
      double v = a.fsr[3 * i + d];
      if (a.has_cl) v = v + fr[d] * a.coul;
      f[d] = v;
    }
    a.fx[i] = f[0];
    a.fy[i] = f[1];
    a.fz[i] = f[2];
  }
  colsum_partials(a.n, a.ncol, a.in, a.part, sh);
}

// ef = ef_dir + ef_rec + ef_surf + ef_self (src/ewald.f90:224-235);
// ef_surf = -(sum q r + sum mu) 4pi/3V when lsurf; ef_self = 4 alpha^3/(3 sqrt(pi)) mu
// total field of every atom: ef_dir + ef_rec (+ self term of the dipoles)
// (+ surface field -(sum q r | sum mu) 4pi/3V, from the device moments
// mom = [musq, mutot(3), qtot(3)]; which: 0 both, 1 charges, 2 dipoles);
// null edir / erec count as zero fields
__global__ void k_combine_field(int n, const double* __restrict__ edir, const double* __restrict__ erec,
                                const double* __restrict__ mu, double self2, int lsurf,
                                const double* __restrict__ mom, int which, double fpi_3V,
                                double* __restrict__ ef) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) {
    double e = (edir ? edir[3 * i + a] : 0.0) + (erec ? erec[3 * i + a] : 0.0);
    if (lsurf) {
      const double m = which == 0 ? mom[4 + a] + mom[1 + a] : (which == 1 ? mom[4 + a] : mom[1 + a]);
      e += -m * fpi_3V;
    }
    if (mu) e += self2 * mu[3 * i + a];
    ef[3 * i + a] = e;
  }
}

// per-atom columns for the self and surface terms: mu^2, mu, q r
__global__ void k_moments(int n, const double* __restrict__ q, const double* __restrict__ mu,
                          const double* __restrict__ rx, const double* __restrict__ ry,
                          const double* __restrict__ rz, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m0 = mu ? mu[3 * i] : 0.0, m1 = mu ? mu[3 * i + 1] : 0.0, m2 = mu ? mu[3 * i + 2] : 0.0;
  out[i] = m0 * m0 + m1 * m1 + m2 * m2;
  out[(size_t)n + i] = m0;
  out[(size_t)2 * n + i] = m1;
  out[(size_t)3 * n + i] = m2;
  out[(size_t)4 * n + i] = q[i] * rx[i];
  out[(size_t)5 * n + i] = q[i] * ry[i];
  out[(size_t)6 * n + i] = q[i] * rz[i];
}

// induced_moment (src/moment_from_pola.f90:955-993): mu = alpha E (alpha = 0
// for non-polarisable ions), plus the per-atom u_pol term mu.alpha^-1.mu
__global__ void k_induced(int n, const double* __restrict__ pol, const double* __restrict__ invpol,
                          const double* __restrict__ E, double* __restrict__ mu, double* __restrict__ upol) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = pol + 9 * (size_t)i;
  double m[3];
  for (int a = 0; a < 3; ++a) {
    double v = 0.0;
    for (int b = 0; b < 3; ++b) v = v + p[a + 3 * b] * E[3 * i + b];
    m[a] = v;
    mu[3 * i + a] = v;
  }
  if (upol) {
    const double* ip = invpol + 9 * (size_t)i;
    double u = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) u = u + m[a] * ip[a + 3 * b] * m[b];
    upol[i] = u;
  }
}

__global__ void k_upol(int n, const double* __restrict__ invpol, const double* __restrict__ mu,
                       double* __restrict__ upol) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* ip = invpol + 9 * (size_t)i;
  double u = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) u = u + mu[3 * i + a] * ip[a + 3 * b] * mu[3 * i + b];
  upol[i] = u;
}

// ASPC predictor (src/moment_from_pola.f90:566-575) and history shift
__global__ void k_aspc_predict(size_t n3, int ext, double b0, double b1, double b2, double b3, double b4,
                               double* __restrict__ hist, double* __restrict__ mu) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n3) return;
  double B[5] = {b0, b1, b2, b3, b4};
  double m = 0.0;
  for (int k = 0; k <= ext; ++k) m = m + B[k] * hist[k * n3 + x];
  mu[x] = m;
  for (int k = ext; k >= 1; --k) hist[k * n3 + x] = hist[(k - 1) * n3 + x];
}

// ASPC corrector (:579-583): mu = w alpha E + (1 - w) mu_p
__global__ void k_aspc_correct(int n, double w, const double* __restrict__ pol, const double* __restrict__ E,
                               double* __restrict__ mu) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = pol + 9 * (size_t)i;
  for (int a = 0; a < 3; ++a) {
    double v = 0.0;
    for (int b = 0; b < 3; ++b) v = v + p[a + 3 * b] * E[3 * i + b];
    mu[3 * i + a] = w * v + (1.0 - w) * mu[3 * i + a];
  }
}

// get_rmsd_mu (:998-1027) per-atom terms; columns: residual^2, polar flag,
// max mu component (catastrophe check, :176-190)
__global__ void k_rmsd_terms(int n, const int32_t* __restrict__ polar, const double* __restrict__ pol,
                             const double* __restrict__ E, const double* __restrict__ mu,
                             double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  if (polar[i]) {
    for (int a = 0; a < 3; ++a) {
      double p = pol[9 * (size_t)i + a + 3 * a];
      if (p == 0.0) continue;
      double x = mu[3 * i + a] / p - E[3 * i + a];
      r = r + x * x;
    }
  }
  out[i] = r;
}

// out = a x + b y
__global__ void k_axpby(size_t n3, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                        double* __restrict__ out) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) out[k] = a * x[k] + b * y[k];
}

// scf_kO_v2 residual f = E - mu / alpha on polarisable atoms (others keep f)
__global__ void k_kO_resid(int n, const int32_t* __restrict__ polar, const double* __restrict__ pol,
                           const double* __restrict__ E, const double* __restrict__ mu, double* __restrict__ f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !polar[i]) return;
  for (int a = 0; a < 3; ++a) f[3 * i + a] = E[3 * i + a] - mu[3 * i + a] / pol[9 * (size_t)i + a + 3 * a];
}

// find_min (src/moment_from_pola.f90:2415-2455) per-atom dot products over
// polarisable atoms: g0.g0 and g0.(g1 - g0)
__global__ void k_kO_dots(int n, const int32_t* __restrict__ polar, const double* __restrict__ g0,
                          const double* __restrict__ g1, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s0 = 0.0, s1 = 0.0;
  if (polar[i])
    for (int k = 0; k < 3; ++k) {
      const double a = g0[3 * i + k];
      s0 = s0 + a * a;
      s1 = s1 + a * (g1[3 * i + k] - a);
    }
  out[i] = s0;
  out[(size_t)n + i] = s1;
}

// gmin = g0 + al (g1 - g0), dmin = al d1
__global__ void k_kO_step(size_t n3, double al, const double* __restrict__ g0, const double* __restrict__ g1,
                          const double* __restrict__ d1, double* __restrict__ gmin, double* __restrict__ dmin) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  gmin[k] = g0[k] + al * (g1[k] - g0[k]);
  dmin[k] = al * d1[k];
}

__global__ void k_add3(size_t n3, const double* __restrict__ a, const double* __restrict__ b,
                       double* __restrict__ o) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n3) o[x] = a[x] + b[x];
}

__global__ void k_maxcomp(int n, const double* __restrict__ mu, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = mu[3 * i];
  m = fmax(m, mu[3 * i + 1]);
  m = fmax(m, mu[3 * i + 2]);
  out[i] = m > 10.0 ? 1.0 : 0.0;
}

// prop_pos_vel_verlet (:420-427): half kick + drift
__global__ void k_kick_drift(int n, double dt2, double dt, const double* __restrict__ m, double* rx, double* ry,
                             double* rz, double* vx, double* vy, double* vz, const double* fx, const double* fy,
                             const double* fz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] + fx[i] * dt2 / m[i];
  vy[i] = vy[i] + fy[i] * dt2 / m[i];
  vz[i] = vz[i] + fz[i] * dt2 / m[i];
  rx[i] = rx[i] + vx[i] * dt;
  ry[i] = ry[i] + vy[i] * dt;
  rz[i] = rz[i] + vz[i] * dt;
}

__global__ void k_kick(int n, double dt2, const double* __restrict__ m, double* vx, double* vy, double* vz,
                       const double* fx, const double* fy, const double* fz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] + fx[i] * dt2 / m[i];
  vy[i] = vy[i] + fy[i] * dt2 / m[i];
  vz[i] = vz[i] + fz[i] * dt2 / m[i];
}

__global__ void k_scale_v(int n, double s, double sx, double sy, double sz, double* vx, double* vy, double* vz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] * s - sx;
  vy[i] = vy[i] * s - sy;
  vz[i] = vz[i] * s - sz;
}

// kinetic energy column and velocity sums
__global__ void k_kin(int n, const double* __restrict__ m, const double* vx, const double* vy, const double* vz,
                      double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) * m[i];
  out[(size_t)n + i] = vx[i];
  out[(size_t)2 * n + i] = vy[i];
  out[(size_t)3 * n + i] = vz[i];
}

inline int nblk(int n) { return div_up(n, 256); }

// ---- Nose-Hoover chain steps on the device (nose_hoover_chain_n,
// src/integration.f90:538-701).  The chain state lives in device memory
// (NHC_XI, NHC_VXI, NHC_Q: nhc_n entries each; NHC_S: the last scale); the
// kinetic energy reaches the chain as block partials of sum m v^2.
constexpr int NHC_XI = 0, NHC_VXI = 16, NHC_Q = 32, NHC_S = 48, NHC_N = 64;

struct NhcArgs {
  int nhc, yosh, mults, nb;
  double L, temp, dt;
  double w[9];                   // Yoshida weights (get_yoshida_weigth)
  const double* part;            // nb block partials of sum m v^2
  double* st;                    // chain state (NHC_N doubles)
  double* envt;                  // e_nvt of the step (null: not wanted)
};

// chain_nhcn (src/integration.f90:619-701): the thermostat update from the
// kinetic energy, in one thread after the block's sum of the partials; the
// velocity scale s goes to st[NHC_S]
__global__ void __launch_bounds__(256) k_nhc_chain(NhcArgs a) {
  __shared__ double sh[4 * kRedMaxCols];
  __shared__ double tot[1];
  sum_block_partials(a.nb, 1, a.part, tot, sh);
  if (threadIdx.x != 0) return;
  const double kin = tot[0] * 0.5;
  double* xi = a.st + NHC_XI;
  double* vxi = a.st + NHC_VXI;
  const double* Q = a.st + NHC_Q;
  const int nhc = a.nhc;
  double G[16];
  for (int h = 0; h < 16; ++h) G[h] = 0.0;
  double s = 1.0;
  G[0] = 2.0 * kin - a.L * a.temp;
  for (int k = 0; k < a.mults; ++k)
    for (int j = 0; j < a.yosh; ++j) {
      const double dts = (a.w[j] * a.dt) / (double)a.mults, dts2 = dts * 0.5, dts4 = dts2 * 0.5, dts8 = dts4 * 0.5;
      G[nhc - 1] = (vxi[nhc - 2] * vxi[nhc - 2] / Q[nhc - 2] - a.temp);
      vxi[nhc - 1] = vxi[nhc - 1] + G[nhc - 1] * dts4;
      for (int h = nhc - 2; h >= 0; --h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
      }
      s = s * exp(-vxi[0] * dts2 / Q[0]);
      for (int h = 0; h < nhc; ++h) xi[h] = xi[h] + vxi[h] * dts2 / Q[h];
      G[0] = 2.0 * kin * s * s - a.L * a.temp;
      for (int h = 0; h < nhc - 1; ++h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        G[h + 1] = (vxi[h] * vxi[h] / Q[h] - a.temp);
      }
      vxi[nhc - 1] = vxi[nhc - 1] + G[nhc - 1] * dts4;
    }
  a.st[NHC_S] = s;
  if (a.envt) {
    // conserved quantity of the extended system (src/integration.f90:582-593)
    double e3 = 0.0, e4 = 0.0;
    for (int h = 1; h < nhc; ++h) {
      e3 = e3 + vxi[h] * vxi[h] * 0.5 / Q[h];
      e4 = e4 + xi[h];
    }
    a.envt[0] = a.L * a.temp * xi[0] + vxi[0] * vxi[0] * 0.5 / Q[0] + e3 + a.temp * e4;
  }
}

// v = s v with the chain's scale (scale != 0), then the block partials of
// sum m v^2
__global__ void __launch_bounds__(256) k_nhc_scale_kin(int n, int scale, const double* __restrict__ st,
                                                       const double* __restrict__ m, double* vx, double* vy,
                                                       double* vz, double* __restrict__ part) {
  __shared__ double sh[4 * kRedMaxCols];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double kv[kRedMaxCols] = {};
  if (i < n) {
    double a = vx[i], b = vy[i], cc = vz[i];
    if (scale) {
      const double s = st[NHC_S];
      a = s * a;
      b = s * b;
      cc = s * cc;
      vx[i] = a;
      vy[i] = b;
      vz[i] = cc;
    }
    kv[0] = (a * a + b * b + cc * cc) * m[i];
  }
  const double r = block_sum_cols(kv, 1, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// v = s v, then prop_pos_vel_verlet's half kick and drift (:420-427)
__global__ void k_nhc_scale_kick_drift(int n, const double* __restrict__ st, double dt2, double dt,
                                       const double* __restrict__ m, double* rx, double* ry, double* rz, double* vx,
                                       double* vy, double* vz, const double* fx, const double* fy, const double* fz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = st[NHC_S];
  const double a = s * vx[i], b = s * vy[i], cc = s * vz[i];
  const double ax = a + fx[i] * dt2 / m[i];
  const double by = b + fy[i] * dt2 / m[i];
  const double cz = cc + fz[i] * dt2 / m[i];
  vx[i] = ax;
  vy[i] = by;
  vz[i] = cz;
  rx[i] = rx[i] + ax * dt;
  ry[i] = ry[i] + by * dt;
  rz[i] = rz[i] + cz * dt;
}

// the second half kick of prop_pos_vel_verlet (:436-442), with the block
// partials of sum m v^2
__global__ void __launch_bounds__(256) k_nhc_kick_kin(int n, double dt2, const double* __restrict__ m, double* vx,
                                                      double* vy, double* vz, const double* fx, const double* fy,
                                                      const double* fz, double* __restrict__ part) {
  __shared__ double sh[4 * kRedMaxCols];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double kv[kRedMaxCols] = {};
  if (i < n) {
    const double a = vx[i] + fx[i] * dt2 / m[i];
    const double b = vy[i] + fy[i] * dt2 / m[i];
    const double cc = vz[i] + fz[i] * dt2 / m[i];
    vx[i] = a;
    vy[i] = b;
    vz[i] = cc;
    kv[0] = (a * a + b * b + cc * cc) * m[i];
  }
  const double r = block_sum_cols(kv, 1, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

}  // namespace

// ------------------------------------------------------------------ helpers

static int fail(mdff_ctx* c, int code, const char* msg) {
  c->err = msg;
  return code;
}

// deterministic column sums of a [ncol][n] device array; result on the host
int reduce_cols(mdff_ctx* c, const double* in, int ncol, double* host_out) {
  MDFF_TRY(reduce_cols_dev(c, in, ncol, c->red));
  HIP_TRY(c, hipMemcpyAsync(c->h_red, c->red, sizeof(double) * ncol, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < ncol; ++k) host_out[k] = c->h_red[k];
  return MDFF_OK;
}

// blocks of a column-sum launch over the atoms
static int colsum_blocks(int n) { return std::min(1024, std::max(1, div_up(n, 256))); }

// block partials of the column sums into part ([nb][ncol]); the final sum is
// queued or launched by the caller
int reduce_cols_part(mdff_ctx* c, const double* in, int ncol, double* part, int* nb_out) {
  if (ncol < 1 || ncol > kRedMaxCols) return fail(c, MDFF_ERR_BAD_ARG, "column sum over too many columns");
  MDFF_TRY(claim_partials(c, part));
  const int nb = colsum_blocks(c->n);
  hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, c->stream, c->n, ncol, in, part);
  *nb_out = nb;
  return MDFF_OK;
}

FinalBatch take_finals(mdff_ctx* c) {
  FinalBatch a{};
  for (int k = 0; k < c->nfin; ++k) a.f[k] = c->fin[k];
  a.nfin = c->nfin;
  if (c->chk_seq_wait == 1) {
    a.seq_out = c->d_chk + 2 * nblk(c->n);
    a.seq_val = c->chk_seq;
    c->chk_seq_wait = 2;
  }
  c->nfin = 0;
  return a;
}

int flush_finals(mdff_ctx* c) {
  if (c->nfin == 0 && c->chk_seq_wait != 1) return MDFF_OK;
  const FinalBatch a = take_finals(c);
  hipLaunchKernelGGL(k_final_sums, dim3(std::max(a.nfin, 1)), dim3(256), 0, c->stream, a);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int queue_final(mdff_ctx* c, const double* part, int nb, int ncol, double* out) {
  if (ncol < 1 || ncol > kRedMaxCols || nb < 1) return fail(c, MDFF_ERR_BAD_ARG, "bad final sum");
  if (c->nfin == kMaxFinals) MDFF_TRY(flush_finals(c));
  c->fin[c->nfin++] = FinalSum{part, out, nb, ncol};
  return MDFF_OK;
}

// before a kernel writes block partials into part: run the queued finals
// that still read it
int claim_partials(mdff_ctx* c, const double* part) {
  for (int k = 0; k < c->nfin; ++k)
    if (c->fin[k].part == part) return flush_finals(c);
  return MDFF_OK;
}

// the same sums left in device memory (dst[0..ncol)), no host round trip
int reduce_cols_dev(mdff_ctx* c, const double* in, int ncol, double* dst) {
  if (ncol < 1 || ncol > kRedMaxCols) return fail(c, MDFF_ERR_BAD_ARG, "column sum over too many columns");
  MDFF_TRY(claim_partials(c, c->red_part));
  const int nb = colsum_blocks(c->n);
  hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, c->stream, c->n, ncol, in, c->red_part);
  hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(256), 0, c->stream, nb, ncol, c->red_part, dst);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}


// 3x3 inverse (replaces LAPACK DGETRF/DGETRI of src/read_pos.f90:203-212)
static void inv3(const double* m, double* o) {
  double a = m[0], b = m[3], cc = m[6], d = m[1], e = m[4], f = m[7], g = m[2], h = m[5], k = m[8];
  double det = a * (e * k - f * h) - b * (d * k - f * g) + cc * (d * h - e * g);
  o[0] = (e * k - f * h) / det; o[3] = (cc * h - b * k) / det; o[6] = (b * f - cc * e) / det;
  o[1] = (f * g - d * k) / det; o[4] = (a * k - cc * g) / det; o[7] = (cc * d - a * f) / det;
  o[2] = (d * h - e * g) / det; o[5] = (b * g - a * h) / det; o[8] = (a * e - b * d) / det;
}

// lattice (src/lattice.f90:60-112)
static void lattice(const double* A, double* B, double* omega, double* W) {
  const double *a1 = A, *a2 = A + 3, *a3 = A + 6;
  double* b[3] = {B, B + 3, B + 6};
  const double* u[3][2] = {{a2, a3}, {a3, a1}, {a1, a2}};
  for (int j = 0; j < 3; ++j) {
    const double *U1 = u[j][0], *U2 = u[j][1];
    b[j][0] = U1[1] * U2[2] - U1[2] * U2[1];
    b[j][1] = U1[2] * U2[0] - U1[0] * U2[2];
    b[j][2] = U1[0] * U2[1] - U1[1] * U2[0];
  }
  double om = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
  *omega = om;
  for (int j = 0; j < 3; ++j) W[j] = om / sqrt(b[j][0] * b[j][0] + b[j][1] * b[j][1] + b[j][2] * b[j][2]);
  for (int k = 0; k < 9; ++k) B[k] = B[k] / om;
}

// pair tables (src/non_bonded.f90:455-644, src/pim.f90:78-81)
static void build_pair_params(mdff_ctx* c, const int32_t* natmi) {
  const mdff_desc& d = c->d;
  const int nt = c->ntype;
  const double two16 = pow(2.0, 1.0 / 6.0), tpi = 2.0 * MDFF_PI;
  c->pp.assign(nt * nt, PairParam{});
  c->u_tail = 0.0;
  double ptail = 0.0;
  for (int it = 0; it < nt; ++it)
    for (int jt = 0; jt < nt; ++jt) {
      const int p = it + 16 * jt;
      PairParam& P = c->pp[it + nt * jt];
      double pl = d.plj[p], ql = d.qlj[p];
      double rc = d.cutshortrange, rcsq = rc * rc, rc3 = rc * rcsq, ppqq = pl * ql;
      P.rcutsq = rcsq;
      P.A = d.Abmhftd[p]; P.B = d.Bbmhftd[p]; P.C = d.Cbmhftd[p]; P.D = d.Dbmhftd[p]; P.BD = d.BDbmhftd[p];
      P.pl = pl; P.ql = ql;
      P.epsp = d.epslj[p] / (ql - pl);
      P.sigsq = two16 * two16 * d.sigmalj[p] * d.sigmalj[p];
      double sr2 = P.sigsq / rcsq, sr = sqrt(sr2), srp = pow(sr, pl), srq = pow(sr, ql);
      P.uc = P.epsp * (pl * srq - ql * srp);
      P.uc1 = P.epsp * ppqq / (2.0 * rcsq) * (srq - srp);
      P.uc2 = 0.5 * P.epsp * ((2.0 * pl + ppqq) * srq - (2.0 * ql + ppqq) * srp);
      P.fc = ppqq * P.epsp / P.sigsq;
      P.epsmor = d.epsmor[p]; P.rhomor = d.rhomor[p];
      P.rs = exp(d.rhomor[p] * d.sigmamor[p]);
      double ut = P.epsp * (pl * srq / (ql - 3.0) - ql * srp / (pl - 3.0)) * rc3 * tpi;
      double pt = ppqq * P.epsp * (srq / (ql - 3.0) - srp / (pl - 3.0)) * rc3 * tpi;
      if (d.lnmlj && natmi[it] != 0 && natmi[jt] != 0) {
        c->u_tail += ut * natmi[it] * natmi[jt] / c->omega;
        ptail += pt * natmi[it] * natmi[jt] / c->omega;
      }
      // TT damping of the field at i (type it) from j (type jt) and the converse
      const int ki = it + 16 * it + 256 * jt, kj = jt + 16 * it + 256 * jt;
      P.ldamp = (d.ldip_damping[ki] || d.ldip_damping[kj]) ? 1 : 0;
      P.db_i = d.pol_damp_b[ki]; P.dc_i = d.pol_damp_c[ki]; P.dk_i = d.pol_damp_k[ki];
      P.db_j = d.pol_damp_b[kj]; P.dc_j = d.pol_damp_c[kj]; P.dk_j = d.pol_damp_k[kj];
    }
  c->p_tail = d.lnmlj ? ptail / c->omega / 3.0 / d.press_unit : 0.0;
  if (!d.lnmlj) c->u_tail = 0.0;
}

#define DMALLOC(ptr, count)                                                            \
  do {                                                                                 \
    HIP_TRY(c, hipMalloc((void**)&(ptr), sizeof(*(ptr)) * (size_t)(count)));           \
    HIP_TRY(c, hipMemset((ptr), 0, sizeof(*(ptr)) * (size_t)(count)));                 \
  } while (0)

static int ctx_init(mdff_ctx* c, const mdff_desc* d, const int32_t* itype) {
  c->d = *d;
  c->n = d->natm;
  c->ntype = d->ntype;
  const int n = c->n;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(c, MDFF_ERR_NO_DEVICE, "no HIP device");
  HIP_TRY(c, hipGetDevice(&c->device));
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) HIP_TRY(c, hipEventCreate(&e));
  memcpy(c->A, d->cell, sizeof c->A);
  lattice(c->A, c->B, &c->omega, c->W);
  const double half = 0.5 * std::min(c->W[0], std::min(c->W[1], c->W[2]));
  c->rc_coul = std::min(d->cutlongrange, half);  // ewald_param (src/coulomb.f90:377-383)
  c->rc_vdw = d->cutshortrange;
  c->alpha = d->alphaES;
  // tasks (src/coulomb.f90:188-238, src/pim.f90:147-183)
  int lqch = 0;
  for (int i = 0; i < 16; ++i) lqch |= d->qch[i] != 0.0;
  for (int i = 0; i < 48; ++i) c->ldip_static |= d->dip[i] != 0.0;
  for (int i = 0; i < d->ntype; ++i) c->linduced |= d->ldip_polar[i] != 0;
  c->task_qq = lqch;
  c->task_qd = c->ldip_static && lqch;
  c->task_dd = c->ldip_static;
  if (c->linduced) { c->task_qd |= lqch; c->task_dd = 1; }
  if (d->lnmlj) c->sr_model = 1;
  if (d->lbmhft || d->lbmhftd) c->sr_model = 2;
  c->sr_lj_extra = (d->lnmlj && (d->lbmhft || d->lbmhftd)) ? 1 : 0;
  if (d->lmorse) c->sr_model = 3;
  // per-atom type data (src/read_pos.f90:160-234)
  std::vector<int32_t> it0(n);
  int32_t natmi[16] = {0};
  std::vector<double> q(n), m(n), pol(9 * (size_t)n, 0.0), ipol(9 * (size_t)n, 0.0), dip(3 * (size_t)n);
  std::vector<int32_t> polar(n);
  for (int i = 0; i < n; ++i) {
    int t = itype[i] - 1;
    it0[i] = t;
    natmi[t]++;
    q[i] = d->qch[t];
    m[i] = d->mass[t];
    for (int a = 0; a < 3; ++a) dip[3 * i + a] = d->dip[t + 16 * a];
    polar[i] = d->ldip_polar[t];
    if (d->ldip_polar[t]) {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) pol[9 * (size_t)i + a + 3 * b] = d->poldip[t + 16 * a + 48 * b];
      inv3(&pol[9 * (size_t)i], &ipol[9 * (size_t)i]);
    }
  }
  build_pair_params(c, natmi);
  DMALLOC(c->rx, n); DMALLOC(c->ry, n); DMALLOC(c->rz, n);
  DMALLOC(c->vx, n); DMALLOC(c->vy, n); DMALLOC(c->vz, n);
  DMALLOC(c->fx, n); DMALLOC(c->fy, n); DMALLOC(c->fz, n);
  DMALLOC(c->fsx, n); DMALLOC(c->fsy, n); DMALLOC(c->fsz, n);
  DMALLOC(c->sx, n); DMALLOC(c->sy, n); DMALLOC(c->sz, n);
  DMALLOC(c->xs, n); DMALLOC(c->ys, n); DMALLOC(c->zs, n);
  DMALLOC(c->bx, n); DMALLOC(c->by, n); DMALLOC(c->bz, n);
  DMALLOC(c->bf, n);
  DMALLOC(c->itype, n);
  DMALLOC(c->q, n);
  DMALLOC(c->mass, n);
  DMALLOC(c->pol, 9 * (size_t)n);
  DMALLOC(c->invpol, 9 * (size_t)n);
  DMALLOC(c->dipia, 3 * (size_t)n);
  DMALLOC(c->mu, 3 * (size_t)n);
  DMALLOC(c->mu_ind, 3 * (size_t)n);
  DMALLOC(c->efield, 3 * (size_t)n);
  DMALLOC(c->e_stat, 3 * (size_t)n);
  DMALLOC(c->e_tot, 3 * (size_t)n);
  DMALLOC(c->e_ind, 3 * (size_t)n);
  DMALLOC(c->hist, 15 * (size_t)n);
  DMALLOC(c->frec, 3 * (size_t)n);
  DMALLOC(c->zeros, n);
  DMALLOC(c->arec, n);
  DMALLOC(c->ameta, n);
  // per-atom scratch: NPA columns + f_sr, f_dir, ef_dir, ef_rec (3n each) + 8 extra columns
  DMALLOC(c->peratom, (size_t)(NPA + 12 + 8) * n);
  if (c->sr_lj_extra) {
    DMALLOC(c->peratom_lj, (size_t)NPA * n);
    DMALLOC(c->f_lj, 3 * (size_t)n);
    DMALLOC(c->lj_part, (size_t)8 * colsum_blocks(n));
  }
  DMALLOC(c->d_pp, c->pp.size());
  HIP_TRY(c, hipMemcpy(c->d_pp, c->pp.data(), sizeof(PairParam) * c->pp.size(), hipMemcpyHostToDevice));
  DMALLOC(c->cell_of, n);
  DMALLOC(c->cell_key, n);
  DMALLOC(c->cell_atoms, n);
  DMALLOC(c->cell_atoms_in, n);
  DMALLOC(c->nl_count, n + 1);
  DMALLOC(c->nl_start, n + 1);
  DMALLOC(c->nl_count_in, n + 1);
  // mixed form: pairs within 0.4 of the largest cutoff (plus the skin, at
  // the build) run in FP64; MDFF_MIXED_INNER=<fraction> sets another
  // fraction, 0 none
  c->rc_inner = 0.0;
  if (d->precision == MDFF_PREC_MIXED) {
    double frac = 0.4;
    const char* ev = getenv("MDFF_MIXED_INNER");
    if (ev && ev[0]) frac = atof(ev);
    if (frac > 0.0) c->rc_inner = frac * std::max(c->rc_vdw, c->rc_coul);
  }
  DMALLOC(c->d_flag, 4);
  DMALLOC(c->cx, n); DMALLOC(c->cy, n); DMALLOC(c->cz, n);
  // written by the check kernels over the bus, read by the host after the
  // check's event: mapped, coherent (uncached on the device side)
  HIP_TRY(c, hipHostMalloc((void**)&c->h_chk, sizeof(double) * (2 * (size_t)nblk(n) + 2),
                           hipHostMallocMapped | hipHostMallocCoherent));
  HIP_TRY(c, hipHostGetDevicePointer((void**)&c->d_chk, c->h_chk, 0));
  memset(c->h_chk, 0, sizeof(double) * (2 * (size_t)nblk(n) + 2));
  HIP_TRY(c, hipEventCreateWithFlags(&c->chk_ev, hipEventDisableTiming));
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIP_TRY(c, hipEventCreateWithFlags(&c->ov_tab, hipEventDisableTiming));
  HIP_TRY(c, hipEventCreateWithFlags(&c->ov_pair, hipEventDisableTiming));
  {
    // MDFF_PAIR_OVERLAP=0 keeps every kernel on the one stream
    const char* ov = getenv("MDFF_PAIR_OVERLAP");
    c->overlap = (ov && ov[0] == '0') ? 0 : 1;
  }
  HIP_TRY(c, hipHostMalloc((void**)&c->h_flag, sizeof(int32_t) * 4));
  const size_t nred = std::max<size_t>(64 * 1024, 2 * (size_t)nblk(n) + 2);
  HIP_TRY(c, hipHostMalloc((void**)&c->h_red, sizeof(double) * nred));
  DMALLOC(c->red_part, nred);
  DMALLOC(c->red, 64);
  DMALLOC(c->fin_part, 1024 * 16 + (size_t)nblk(n) + 16);
  DMALLOC(c->nhc_part, 2 * (size_t)nblk(n));
  DMALLOC(c->nhc_st, NHC_N);
  DMALLOC(c->raw, (size_t)RAW_N * RAW_SLOTS);
  HIP_TRY(c, hipHostMalloc((void**)&c->h_raw, sizeof(double) * RAW_N * RAW_SLOTS));
  HIP_TRY(c, hipMemcpy(c->itype, it0.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->q, q.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->mass, m.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->pol, pol.data(), sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->invpol, ipol.data(), sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->dipia, dip.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMalloc((void**)&c->polar, sizeof(int32_t) * n));
  HIP_TRY(c, hipMemcpy(c->polar, polar.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  c->npol = 0;
  c->qsq = 0.0;
  for (int i = 0; i < n; ++i) {
    c->npol += polar[i] != 0;
    c->qsq = c->qsq + q[i] * q[i];
    c->qmax = std::max(c->qmax, std::fabs(q[i]));
  }
  // reciprocal kernels: matrix-core by default; MDFF_RECIP_VALU=1 selects the
  // vector-ALU kernels (same sums, kept for comparison and for shapes the
  // matrix-core kernels do not cover)
  const char* ev = getenv("MDFF_RECIP_VALU");
  c->use_mma = (ev && ev[0] == '1') ? 0 : 1;
  // m-tiles per wave of the matrix-core structure factor: one (two would
  // exceed 256 VGPRs with the unrolled atom tile)
  c->sk_ctm = 1;
  {
    const char* ev = getenv("MDFF_DIP_RECIP_FP64");
    const bool dip64 = ev && ev[0] == '1' && (c->task_qd || c->task_dd);
    c->rec_f32 = (d->precision == MDFF_PREC_MIXED && !dip64) ? 1 : 0;
  }
  if (d->lcoulomb) MDFF_TRY(recip_setup(c));
  return MDFF_OK;
}

// scratch views inside c->peratom
static double* col(mdff_ctx* c, int k) { return c->peratom + (size_t)k * c->n; }
static double* f_sr(mdff_ctx* c) { return col(c, NPA); }          // 3n
static double* f_dir(mdff_ctx* c) { return col(c, NPA + 3); }     // 3n
static double* ef_dir(mdff_ctx* c) { return col(c, NPA + 6); }    // 3n
static double* ef_rec(mdff_ctx* c) { return col(c, NPA + 9); }    // 3n
static double* extra(mdff_ctx* c) { return col(c, NPA + 12); }    // 8 columns

// ------------------------------------------------------------------ Ewald pieces

// self / surface sums (src/ewald.f90:112-143)
// musq, mutot(3), qtot(3) of (q, mu) left in device memory at dst[0..7)
static int moments_dev(mdff_ctx* c, const double* q, const double* mu, double* dst) {
  hipLaunchKernelGGL(k_moments, dim3(nblk(c->n)), dim3(256), 0, c->stream, c->n, q, mu, c->rx, c->ry, c->rz,
                     extra(c));
  return reduce_cols_dev(c, extra(c), 7, dst);
}

// full Ewald field of a charge set (damped, static SCF field) or a dipole set
// (undamped induced field): ef_dir + ef_rec + ef_self (+ ef_surf)
static int ewald_field(mdff_ctx* c, bool charges, const double* mu, double* out) {
  const double alpha2 = c->alpha * c->alpha;
  const double selfa = c->alpha / sqrt(MDFF_PI), selfa2 = 2.0 * selfa * alpha2 / 3.0;
  const double fpi_3V = 2.0 * (2.0 * MDFF_PI / c->omega / 3.0);
  MDFF_TRY(realspace_field(c, charges ? 0 : 1, mu, ef_dir(c)));
  MDFF_TRY(recip_eval(c, c->q, mu, charges, !charges, false, true, ef_rec(c), nullptr, nullptr));
  // the SCF zeroes the charges for the induced field (src/moment_from_pola.f90:347-351)
  double* mom = c->raw + RAW_SCRATCH;
  if (c->d.lsurf) MDFF_TRY(moments_dev(c, c->q, charges ? nullptr : mu, mom));
  hipLaunchKernelGGL(k_combine_field, dim3(nblk(c->n)), dim3(256), 0, c->stream, c->n, ef_dir(c), ef_rec(c),
                     charges ? nullptr : mu, 2.0 * selfa2, c->d.lsurf, mom, charges ? 1 : 2, fpi_3V, out);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

// static field of the SCF (src/moment_from_pola.f90:313-330): the charges'
// field (TT-damped) plus, with static dipoles, theirs (task_static = all:
// mu-mu real space between two non-zero static dipoles, reciprocal, self,
// surface), the two added as the sum is linear in the sources
static int static_field(mdff_ctx* c) {
  if (!c->ldip_static) return ewald_field(c, true, nullptr, c->e_stat);
  // e_tot and e_ind are free until the SCF loop starts
  MDFF_TRY(ewald_field(c, true, nullptr, c->e_tot));
  MDFF_TRY(ewald_field(c, false, c->dipia, c->e_ind));
  const size_t n3 = 3 * (size_t)c->n;
  hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->e_tot, c->e_ind, c->e_stat);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

// moment_from_pola_scf (src/moment_from_pola.f90:240-485)
static int scf(mdff_ctx* c, mdff_terms* t) {
  const int n = c->n;
  const size_t n3 = 3 * (size_t)n;
  const mdff_desc& d = c->d;
  MDFF_TRY(static_field(c));
  HIP_TRY(c, hipMemcpyAsync(c->e_tot, c->e_stat, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  int iscf = 0;
  double rmsd = 1.79769313486231571e308;
  const int npol = c->npol;
  int32_t* polar = c->polar;
  // ASPC order (src/moment_from_pola.f90:505-522)
  int ext;
  if (c->itime >= d.extrapolate_order + 1) ext = d.extrapolate_order;
  else if (c->itime == 1 || c->itime == 0) ext = 0;
  else ext = c->itime - 1;
  double Bc[5] = {0, 0, 0, 0, 0}, w = 0.0;
  switch (ext) {
    case 0: Bc[0] = 1.0; w = 0.0; break;
    case 1: Bc[0] = 2.0; Bc[1] = -1.0; w = 2.0 / 3.0; break;
    case 2: Bc[0] = 2.5; Bc[1] = -2.0; Bc[2] = 0.5; w = 0.6; break;
    case 3: Bc[0] = 2.8; Bc[1] = -2.8; Bc[2] = 1.2; Bc[3] = -0.2; w = 4.0 / 7.0; break;
    default: Bc[0] = 3.0; Bc[1] = -24.0 / 7.0; Bc[2] = 27.0 / 14.0; Bc[3] = -4.0 / 7.0; Bc[4] = 1.0 / 14.0; w = 5.0 / 9.0; break;
  }
  double upol = 0.0;
  while ((iscf < d.max_scf_pol_iter && rmsd > d.conv_tol_ind) || iscf < d.min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d.calc_md) {
      hipLaunchKernelGGL(k_induced, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->pol, c->invpol, c->e_tot,
                         c->mu_ind, extra(c));
    } else {
      hipLaunchKernelGGL(k_aspc_predict, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, ext, Bc[0], Bc[1],
                         Bc[2], Bc[3], Bc[4], c->hist, c->mu_ind);
      hipLaunchKernelGGL(k_upol, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->invpol, c->mu_ind, extra(c));
    }
    double h[1];
    MDFF_TRY(reduce_cols(c, extra(c), 1, h));
    upol = 0.5 * h[0];
    MDFF_TRY(ewald_field(c, false, c->mu_ind, c->e_ind));
    hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->e_stat, c->e_ind, c->e_tot);
    if (iscf == 1 && d.calc_md)
      hipLaunchKernelGGL(k_aspc_correct, dim3(nblk(n)), dim3(256), 0, c->stream, n, w, c->pol, c->e_tot, c->mu_ind);
    hipLaunchKernelGGL(k_rmsd_terms, dim3(nblk(n)), dim3(256), 0, c->stream, n, polar, c->pol, c->e_tot, c->mu_ind,
                       extra(c));
    MDFF_TRY(reduce_cols(c, extra(c), 1, h));
    rmsd = sqrt(h[0] / (double)npol);
  }
  HIP_TRY(c, hipMemcpyAsync(c->hist, c->mu_ind, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  t->u_pol = upol;
  t->n_scf = iscf;
  t->rmsd = rmsd;
  return MDFF_OK;
}

// rmsd of get_rmsd_mu (src/moment_from_pola.f90:998-1027) of (E, mu); mu may be
// the zero vector
static int rmsd_of(mdff_ctx* c, const double* E, const double* mu, double* out) {
  hipLaunchKernelGGL(k_rmsd_terms, dim3(nblk(c->n)), dim3(256), 0, c->stream, c->n, c->polar, c->pol, E, mu,
                     extra(c));
  double h[1];
  MDFF_TRY(reduce_cols(c, extra(c), 1, h));
  *out = sqrt(h[0] / (double)c->npol);
  return MDFF_OK;
}

// induced_moment_inner (src/moment_from_pola.f90:822-942): mu <- (1-w) mu +
// w alpha (f_ext + E(mu)) with E the real-space dipole field at alphaES =
// 0.001 plus its self (and surface) term, at most 10 steps, until the
// residual is below a tenth of |f_ext|.  Scratch: kw[0..3) 3n vectors.
static int kO_inner(mdff_ctx* c, const double* fext, double* mu, double* const kw[4]) {
  const int n = c->n;
  const size_t n3 = 3 * (size_t)n;
  const double w = c->d.omegakO, al = 0.001;
  double* ft = kw[0];
  double* mprev = kw[1];
  double* f = kw[2];
  double* zero = kw[3];
  double rmsd_ref;
  MDFF_TRY(rmsd_of(c, fext, zero, &rmsd_ref));
  hipLaunchKernelGGL(k_induced, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->pol, c->invpol, fext, mu, nullptr);
  hipLaunchKernelGGL(k_axpby, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, w, mu, 0.0, mu, mu);
  double rmsd_inner = 1.79769313486231571e308;
  const double self2 = 2.0 * (2.0 * (al / sqrt(MDFF_PI)) * al * al / 3.0);
  const double fpi_3V = 2.0 * (2.0 * MDFF_PI / c->omega / 3.0);
  double* mom = c->raw + RAW_SCRATCH;
  for (int it = 1; it <= 10 && rmsd_inner > 0.1 * rmsd_ref; ++it) {
    HIP_TRY(c, hipMemcpyAsync(mprev, mu, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
    MDFF_TRY(realspace_field(c, 1, mu, ef_dir(c), al));
    if (c->d.lsurf) MDFF_TRY(moments_dev(c, c->q, mu, mom));
    hipLaunchKernelGGL(k_combine_field, dim3(nblk(n)), dim3(256), 0, c->stream, n, ef_dir(c), zero, mu, self2,
                       c->d.lsurf, mom, 2, fpi_3V, f);
    hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, fext, f, ft);
    MDFF_TRY(rmsd_of(c, ft, mu, &rmsd_inner));
    hipLaunchKernelGGL(k_induced, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->pol, c->invpol, ft, mu, nullptr);
    hipLaunchKernelGGL(k_axpby, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, 1.0 - w, mprev, w, mu, mu);
  }
  return MDFF_OK;
}

// moment_from_pola_scf_kO_v2 (src/moment_from_pola.f90:1335-1639): inner
// fixed-point loop, full Ewald induced field, line minimum between the last
// two residuals (find_min); rmsd = |g_min|.
static int scf_kO_v2(mdff_ctx* c, mdff_terms* t) {
  const int n = c->n;
  const size_t n3 = 3 * (size_t)n;
  const mdff_desc& d = c->d;
  if (!c->kO_ws) HIP_TRY(c, hipMalloc((void**)&c->kO_ws, sizeof(double) * 11 * n3));
  double* W = c->kO_ws;
  double* find = W;
  double* dmu = W + n3;
  double* fprev = W + 2 * n3;
  double* gmin = W + 3 * n3;
  double* mprev = W + 4 * n3;
  double* dmin = W + 5 * n3;
  double* dd = W + 6 * n3;
  double* const kw[4] = {W + 7 * n3, W + 8 * n3, W + 9 * n3, W + 10 * n3};   // kw[3]: zeros
  hipLaunchKernelGGL(k_fill, dim3(div_up(11 * n3, 256)), dim3(256), 0, c->stream, 11 * n3, 0.0, W);
  MDFF_TRY(static_field(c));
  HIP_TRY(c, hipMemcpyAsync(c->e_tot, c->e_stat, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(find, c->e_stat, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_fill, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, 0.0, c->mu_ind);
  int ext;
  if (c->itime >= d.extrapolate_order + 1) ext = d.extrapolate_order;
  else if (c->itime == 1 || c->itime == 0) ext = 0;
  else ext = c->itime - 1;
  double Bc[5] = {0, 0, 0, 0, 0}, w = 0.0;
  switch (ext) {
    case 0: Bc[0] = 1.0; w = 0.0; break;
    case 1: Bc[0] = 2.0; Bc[1] = -1.0; w = 2.0 / 3.0; break;
    case 2: Bc[0] = 2.5; Bc[1] = -2.0; Bc[2] = 0.5; w = 0.6; break;
    case 3: Bc[0] = 2.8; Bc[1] = -2.8; Bc[2] = 1.2; Bc[3] = -0.2; w = 4.0 / 7.0; break;
    default: Bc[0] = 3.0; Bc[1] = -24.0 / 7.0; Bc[2] = 27.0 / 14.0; Bc[3] = -4.0 / 7.0; Bc[4] = 1.0 / 14.0; w = 5.0 / 9.0; break;
  }
  int iscf = 0;
  double rmsd = 1.79769313486231571e308, upol = 0.0;
  while ((iscf < d.max_scf_pol_iter && rmsd > d.conv_tol_ind) || iscf < d.min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d.calc_md) {
      MDFF_TRY(kO_inner(c, iscf <= 2 ? find : gmin, dmu, kw));
      hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->mu_ind, dmu, c->mu_ind);
    } else {
      hipLaunchKernelGGL(k_aspc_predict, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, ext, Bc[0], Bc[1],
                         Bc[2], Bc[3], Bc[4], c->hist, c->mu_ind);
    }
    hipLaunchKernelGGL(k_upol, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->invpol, c->mu_ind, extra(c));
    double h[2];
    MDFF_TRY(reduce_cols(c, extra(c), 1, h));
    upol = 0.5 * h[0];
    MDFF_TRY(ewald_field(c, false, c->mu_ind, c->e_ind));
    hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->e_stat, c->e_ind, c->e_tot);
    if (iscf == 1 && d.calc_md)
      hipLaunchKernelGGL(k_aspc_correct, dim3(nblk(n)), dim3(256), 0, c->stream, n, w, c->pol, c->e_tot, c->mu_ind);
    hipLaunchKernelGGL(k_kO_resid, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->polar, c->pol, c->e_tot,
                       c->mu_ind, find);
    if (iscf != 1) {
      hipLaunchKernelGGL(k_axpby, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, 1.0, c->mu_ind, -1.0, mprev,
                         dd);
      hipLaunchKernelGGL(k_kO_dots, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->polar, fprev, find, extra(c));
      MDFF_TRY(reduce_cols(c, extra(c), 2, h));
      double al = -h[0] / h[1];
      if (fabs(h[1]) < 1e-10) al = 1.0;
      hipLaunchKernelGGL(k_kO_step, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, al, fprev, find, dd, gmin,
                         dmin);
      hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, mprev, dmin, c->mu_ind);
      MDFF_TRY(rmsd_of(c, gmin, kw[3], &rmsd));
    }
    HIP_TRY(c, hipMemcpyAsync(mprev, c->mu_ind, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(fprev, gmin, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  }
  HIP_TRY(c, hipMemcpyAsync(c->hist, c->mu_ind, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  t->u_pol = upol;
  t->n_scf = iscf;
  t->rmsd = rmsd;
  return MDFF_OK;
}

// moment_from_pola_scf_kO_v1 (src/moment_from_pola.f90:1042-1320): the inner
// fixed-point loop on the residual f = E - mu/alpha of the previous
// iteration, mu += its step, no line minimum; rmsd of (E, mu) each iteration.
static int scf_kO_v1(mdff_ctx* c, mdff_terms* t) {
  const int n = c->n;
  const size_t n3 = 3 * (size_t)n;
  const mdff_desc& d = c->d;
  if (!c->kO_ws) HIP_TRY(c, hipMalloc((void**)&c->kO_ws, sizeof(double) * 11 * n3));
  double* W = c->kO_ws;
  double* find = W;
  double* dmu = W + n3;
  double* const kw[4] = {W + 7 * n3, W + 8 * n3, W + 9 * n3, W + 10 * n3};   // kw[3]: zeros
  hipLaunchKernelGGL(k_fill, dim3(div_up(11 * n3, 256)), dim3(256), 0, c->stream, 11 * n3, 0.0, W);
  MDFF_TRY(static_field(c));
  HIP_TRY(c, hipMemcpyAsync(c->e_tot, c->e_stat, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(find, c->e_stat, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_fill, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, 0.0, c->mu_ind);
  int ext;
  if (c->itime >= d.extrapolate_order + 1) ext = d.extrapolate_order;
  else if (c->itime == 1 || c->itime == 0) ext = 0;
  else ext = c->itime - 1;
  double Bc[5] = {0, 0, 0, 0, 0}, w = 0.0;
  switch (ext) {
    case 0: Bc[0] = 1.0; w = 0.0; break;
    case 1: Bc[0] = 2.0; Bc[1] = -1.0; w = 2.0 / 3.0; break;
    case 2: Bc[0] = 2.5; Bc[1] = -2.0; Bc[2] = 0.5; w = 0.6; break;
    case 3: Bc[0] = 2.8; Bc[1] = -2.8; Bc[2] = 1.2; Bc[3] = -0.2; w = 4.0 / 7.0; break;
    default: Bc[0] = 3.0; Bc[1] = -24.0 / 7.0; Bc[2] = 27.0 / 14.0; Bc[3] = -4.0 / 7.0; Bc[4] = 1.0 / 14.0; w = 5.0 / 9.0; break;
  }
  int iscf = 0;
  double rmsd = 1.79769313486231571e308, upol = 0.0;
  while ((iscf < d.max_scf_pol_iter && rmsd > d.conv_tol_ind) || iscf < d.min_scf_pol_iter) {
    iscf++;
    if (iscf != 1 || !d.calc_md) {
      MDFF_TRY(kO_inner(c, find, dmu, kw));
      hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->mu_ind, dmu, c->mu_ind);
    } else {
      hipLaunchKernelGGL(k_aspc_predict, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, ext, Bc[0], Bc[1],
                         Bc[2], Bc[3], Bc[4], c->hist, c->mu_ind);
    }
    hipLaunchKernelGGL(k_upol, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->invpol, c->mu_ind, extra(c));
    double h[1];
    MDFF_TRY(reduce_cols(c, extra(c), 1, h));
    upol = 0.5 * h[0];
    MDFF_TRY(ewald_field(c, false, c->mu_ind, c->e_ind));
    hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->e_stat, c->e_ind, c->e_tot);
    if (iscf == 1 && d.calc_md)
      hipLaunchKernelGGL(k_aspc_correct, dim3(nblk(n)), dim3(256), 0, c->stream, n, w, c->pol, c->e_tot, c->mu_ind);
    MDFF_TRY(rmsd_of(c, c->e_tot, c->mu_ind, &rmsd));
    hipLaunchKernelGGL(k_kO_resid, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->polar, c->pol, c->e_tot,
                       c->mu_ind, find);
  }
  HIP_TRY(c, hipMemcpyAsync(c->hist, c->mu_ind, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  t->u_pol = upol;
  t->n_scf = iscf;
  t->rmsd = rmsd;
  return MDFF_OK;
}

// stage events of one evaluation, recorded only while the event profiler is
// on (every event record costs a few microseconds of stream time)
static void ev_rec(mdff_ctx* c, int k) {
  if (c->prof_on) (void)hipEventRecord(c->ev[k], c->stream);
}

// whether an evaluation needs the moment sums (musq, sum mu, sum q r):
// only with dipoles (self energy) or the surface term
static bool has_moments(const mdff_ctx* c) { return c->task_dd || c->d.lsurf; }

// SUBROUTINE engforce (src/engforce_driver.f90:102-190), launch half: every
// kernel of the force evaluation is queued; the sums the energies, virials
// and stress are made of stay in device memory at raw[0..RAW_N) and the terms
// known on the host (list rebuild, SCF) go to t.  engforce_finish turns the
// sums into the terms once they are on the host.
// Stage events of the profiler: [0] total, [1] lists, [2] short range +
// real space, [3] reciprocal, [4] SCF; stage k between events 2k and 2k+1.
enum { EVS_TOTAL = 0, EVS_LIST = 1, EVS_REAL = 2, EVS_RECIP = 3, EVS_SCF = 4 };

static int engforce_launch(mdff_ctx* c, double* raw, mdff_terms* t, bool defer_finals) {
  const mdff_desc& d = c->d;
  const int n = c->n;
  memset(t, 0, sizeof *t);
  ev_rec(c, 2 * EVS_TOTAL);
  const bool coul = d.lcoulomb != 0;
  const size_t n3 = 3 * (size_t)n;
  const bool dip = c->task_qd || c->task_dd;
  const bool want_field = coul && d.doefield;
  // Without induced dipoles the reciprocal sums depend on the positions
  // only: they are queued first, so that a list check still pending from the
  // previous step (its flag on the way to the host) is finished while the
  // device works on them.
  const bool recip_first = coul && !c->linduced;
  auto recip_stage = [&]() -> int {
    ev_rec(c, 2 * EVS_RECIP);
    if (coul) MDFF_TRY(recip_eval(c, c->q, c->mu, true, dip, true, want_field, ef_rec(c), nullptr, raw + RAW_EREC));
    ev_rec(c, 2 * EVS_RECIP + 1);
    return MDFF_OK;
  };
  // atoms packed for the pair kernels by the table kernel (valid unless the
  // lists are rebuilt below)
  bool packed = false;
  bool ov = false;
  if (recip_first) {
    packed = d.lvnlist && c->lists_valid;
    MDFF_TRY(recip_tables_from_r(c, packed));  // kardir and the phase tables in one launch
    ov = c->overlap && packed && !dip && !want_field && !c->prof_on && !c->sr_lj_extra;
    if (ov) HIP_TRY(c, hipEventRecord(c->ov_tab, c->stream));
    if (!c->mu_static) {
      // no induced dipoles: mu is the static set for the life of the context
      HIP_TRY(c, hipMemcpyAsync(c->mu, c->dipia, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
      c->mu_static = 1;
    }
    MDFF_TRY(recip_stage());
  }
  ev_rec(c, 2 * EVS_LIST);
  int rb = 0;
  MDFF_TRY(nl_check_finish(c, &rb));
  if (!d.lvnlist || !c->lists_valid) {
    MDFF_TRY(nl_build(c));
    t->list_rebuilt = 1;
  } else if (!recip_first) {
    launch_kardir(c);
  }
  if (!packed || rb || t->list_rebuilt) pack_atoms(c);
  ev_rec(c, 2 * EVS_LIST + 1);
  if (!recip_first && coul) MDFF_TRY(recip_tables(c));
  // induced dipoles (get_dipole_moments, src/moment_from_pola.f90:84-224)
  ev_rec(c, 2 * EVS_SCF);
  if (coul && c->linduced) {
    if (d.algo_moment_from_pola == 1) MDFF_TRY(scf_kO_v2(c, t));
    else if (d.algo_moment_from_pola == 2) MDFF_TRY(scf_kO_v1(c, t));
    else MDFF_TRY(scf(c, t));
    hipLaunchKernelGGL(k_add3, dim3(div_up(n3, 256)), dim3(256), 0, c->stream, n3, c->dipia, c->mu_ind, c->mu);
  } else if (!c->mu_static) {
    HIP_TRY(c, hipMemcpyAsync(c->mu, c->dipia, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
    c->mu_static = 1;
  }
  ev_rec(c, 2 * EVS_SCF + 1);
  if (coul && c->linduced) {
    hipLaunchKernelGGL(k_maxcomp, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->mu, extra(c));
    double h[1];
    MDFF_TRY(reduce_cols(c, extra(c), 1, h));
    if (h[0] > 0.0) return fail(c, MDFF_ERR_POLAR_CATASTROPHE, "ERROR : POLARIZATION CATASTROPH");
  }
  // short range + real-space Ewald
  ev_rec(c, 2 * EVS_REAL);
  if (ov && !rb && !t->list_rebuilt) {
    // the atoms packed by the table launch and the unchanged lists are all
    // the pair kernel reads: it runs beside the reciprocal kernels
    HIP_TRY(c, hipStreamWaitEvent(c->stream2, c->ov_tab, 0));
    hipStream_t s0 = c->stream;
    c->stream = c->stream2;
    const int rc = realspace_force_impl(c, f_sr(c), f_dir(c), ef_dir(c), want_field);
    c->stream = s0;
    MDFF_TRY(rc);
    HIP_TRY(c, hipEventRecord(c->ov_pair, c->stream2));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ov_pair, 0));
  } else {
    MDFF_TRY(realspace_force_impl(c, f_sr(c), f_dir(c), ef_dir(c), want_field));
  }
  if (c->sr_lj_extra) {
    int nbl = 0;
    MDFF_TRY(realspace_lj_extra(c, c->f_lj, c->peratom_lj));
    MDFF_TRY(reduce_cols_part(c, c->peratom_lj, 8, c->lj_part, &nbl));
    MDFF_TRY(queue_final(c, c->lj_part, nbl, 8, raw + RAW_LJ));
  }
  ev_rec(c, 2 * EVS_REAL + 1);
  if (!recip_first) MDFF_TRY(recip_stage());
  {
    // total forces (the reciprocal ones finished from their split partials)
    // and the block partials of the per-atom columns, in one launch
    MDFF_TRY(claim_partials(c, c->fin_part));
    ForceSumArgs fa{};
    fa.n = n;
    fa.ncol = NPA;
    fa.fsr = f_sr(c);
    fa.frec = c->frec;
    fa.rf = c->rfin;
    fa.rf_on = (coul && c->rfin_pending) ? 1 : 0;
    fa.has_cl = coul ? 1 : 0;
    fa.coul = d.coul_unit;
    fa.fx = c->fx; fa.fy = c->fy; fa.fz = c->fz;
    fa.in = c->peratom;
    fa.part = c->fin_part;
    c->rfin_pending = 0;
    const int nbp = colsum_blocks(n);
    hipLaunchKernelGGL(k_force_sums, dim3(nbp), dim3(256), 0, c->stream, fa);
    MDFF_TRY(queue_final(c, c->fin_part, nbp, NPA, raw + RAW_PA));
  }
  if (coul) {
    // moments: sum mu^2 for the dipole self energy, sum mu and sum q r for the
    // surface terms; with neither dipoles nor lsurf they are all zero
    if (has_moments(c)) MDFF_TRY(moments_dev(c, c->q, c->mu, raw + RAW_MOM));
    // ef_t = (ef_dir + ef_rec if doefield) + ef_self (+ ef_surf); without a
    // field, dipoles or the surface term it is zero (written once)
    const bool zero_field = !want_field && !has_moments(c);
    if (!(zero_field && c->efield_zero)) {
      const double alpha2 = c->alpha * c->alpha;
      const double selfa = c->alpha / sqrt(MDFF_PI), selfa2 = 2.0 * selfa * alpha2 / 3.0;
      const double fpi_3V = 2.0 * (2.0 * MDFF_PI / c->omega / 3.0);
      hipLaunchKernelGGL(k_combine_field, dim3(nblk(n)), dim3(256), 0, c->stream, n,
                         want_field ? ef_dir(c) : nullptr, want_field ? ef_rec(c) : nullptr, c->mu, 2.0 * selfa2,
                         d.lsurf, raw + RAW_MOM, 0, fpi_3V, c->efield);
    }
    c->efield_zero = zero_field ? 1 : 0;
    // efg_t (src/engforce_driver.f90:171): with doefg every evaluation;
    // without it only the self term, the same every time
    if (d.doefg || !c->efg) MDFF_TRY(efg_eval(c));
  }
  if (!defer_finals) MDFF_TRY(flush_finals(c));
  ev_rec(c, 2 * EVS_TOTAL + 1);
  HIP_TRY(c, hipGetLastError());
  c->timings_pending = c->prof_on ? 1 : 0;
  if (!c->prof_on) c->n_last = 0;
  return MDFF_OK;
}

// engforce, finishing half: energies, virials and stress from the device sums
static void engforce_finish(const mdff_ctx* c, const double* raw, mdff_terms* t) {
  const mdff_desc& d = c->d;
  const double* pa = raw + RAW_PA;
  const double* erec = raw + RAW_EREC;
  // short range (full rows count every pair twice)
  const double inv_vp = 1.0 / c->omega / d.press_unit;
  double tnb[9];
  {
    const double* s = pa + 2;
    double v[9] = {s[0], s[3], s[4], s[3], s[1], s[5], s[4], s[5], s[2]};
    for (int k = 0; k < 9; ++k) tnb[k] = 0.5 * v[k] * inv_vp;
  }
  double usr = 0.5 * pa[0], vir = 0.5 * pa[1] / 3.0;
  if (c->sr_model == 1) { t->u_lj = usr; t->vir_lj = vir; t->u_tail = c->u_tail; t->p_tail = c->p_tail; }
  if (c->sr_model == 2) { t->u_bmhft = usr; t->vir_bmhft = vir; }
  if (c->sr_model == 3) { t->u_morse = usr; t->vir_morse = vir; }
  if (c->sr_model) for (int k = 0; k < 9; ++k) { t->tau_nonb[k] = tnb[k]; t->tau[k] += tnb[k]; }
  if (c->sr_lj_extra) {
    // LJ beside BMHFT(D): its energy, virial and stress; tau_nonb stays the
    // BMHFT(D) one, tau holds both (src/engforce_driver.f90:130-141)
    const double* s = raw + RAW_LJ;
    const double v[9] = {s[2], s[5], s[6], s[5], s[3], s[7], s[6], s[7], s[4]};
    for (int k = 0; k < 9; ++k) t->tau[k] += 0.5 * v[k] * inv_vp;
    t->u_lj = 0.5 * s[0];
    t->vir_lj = 0.5 * s[1] / 3.0;
    t->u_tail = c->u_tail;
    t->p_tail = c->p_tail;
  }
  if (d.lcoulomb) {
    static const double zero7[7] = {0, 0, 0, 0, 0, 0, 0};
    const double* mom = has_moments(c) ? raw + RAW_MOM : zero7;
    const double musq = mom[0];
    const double* mutot = mom + 1;
    const double* qtot = mom + 4;
    const double qsq = c->qsq;
    const double tpi_V = 2.0 * MDFF_PI / c->omega, tpi_3V = tpi_V / 3.0;
    const double alpha2 = c->alpha * c->alpha;
    const double selfa = c->alpha / sqrt(MDFF_PI), selfa2 = 2.0 * selfa * alpha2 / 3.0;
    t->u_dir = 0.5 * pa[8];
    t->u_rec = 2.0 * erec[0] * tpi_V;
    t->u_self = -selfa * qsq + (-selfa2 * musq);
    double uq = 0, ud = 0, uqd = 0;
    for (int a = 0; a < 3; ++a) {
      uq += qtot[a] * qtot[a];
      ud += mutot[a] * mutot[a];
      uqd += qtot[a] * mutot[a];
    }
    t->u_surf = d.lsurf ? (uq + 2.0 * uqd + ud) * tpi_3V : 0.0;
    t->u_coul = (t->u_dir + t->u_rec + t->u_surf + t->u_self + t->u_pol) * d.coul_unit;
    // tau_dir: per-atom -(r f + f r) summed over full rows: x 0.5 (double count) x 0.5 / omega
    double tdir[9], trec[9];
    {
      const double* s = pa + 9;
      double v[9] = {s[0], s[3], s[4], s[3], s[1], s[5], s[4], s[5], s[2]};
      for (int k = 0; k < 9; ++k) tdir[k] = 0.25 * v[k] / c->omega;
      double r[9] = {erec[1], erec[4], erec[5], erec[4], erec[2], erec[6], erec[5], erec[6], erec[3]};
      for (int k = 0; k < 9; ++k) trec[k] = 2.0 * r[k] * tpi_V / c->omega;
    }
    for (int k = 0; k < 9; ++k) {
      t->tau_coul[k] = (tdir[k] + trec[k]) * d.coul_unit / d.press_unit;
      t->tau[k] += 2.0 * t->tau_coul[k];
    }
    t->pvirial_coul = 1.0 / 3.0 * (t->tau_coul[0] + t->tau_coul[4] + t->tau_coul[8]) * d.press_unit;
  }
}

// sums of one or more evaluations to the host: raw slots [0, nslots)
static int fetch_raw(mdff_ctx* c, int nslots) {
  HIP_TRY(c, hipMemcpyAsync(c->h_raw, c->raw, sizeof(double) * RAW_N * nslots, hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

static int engforce(mdff_ctx* c, mdff_terms* t) {
  MDFF_TRY(engforce_launch(c, c->raw, t, false));
  MDFF_TRY(fetch_raw(c, 1));
  engforce_finish(c, c->h_raw, t);
  c->last_terms = *t;
  return MDFF_OK;
}

// stage timings of the last evaluation (needs its events complete)
static void read_timings(mdff_ctx* c) {
  if (!c->timings_pending) return;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return;
  // stage k of last_ms: events (2k, 2k+1); a stage that does not run
  // records both at the same point
  for (int k = 0; k < 5; ++k) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]);
    c->last_ms[k] = ms;
  }
  c->n_last = 5;
  c->timings_pending = 0;
}

hipEvent_t prof_event(mdff_ctx* c) {
  if (c->evused == c->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evpool.push_back(e);
  }
  return c->evpool[c->evused++];
}

static const char* kProfNames[PK_COUNT] = {"k_pair_force", "k_pair_field", "k_tables",       "k_sk",
                                           "k_sk_reduce",  "k_gather",     "k_gather_final", "k_nl_sweep",
                                           "k_amat",       "k_inner_box"};

// ====================================================================== C-ABI

extern "C" {

void mdff_desc_defaults(mdff_desc* d) {
  memset(d, 0, sizeof *d);
  d->abi_version = MDFF_ABI_VERSION;
  d->lvnlist = 1;
  d->calc_md = 1;
  d->skindiff = 0.15;
  d->coul_unit = 14.3996441494161;
  d->press_unit = 0.00624150964712042;
  for (int i = 0; i < 256; ++i) { d->plj[i] = 6.0; d->qlj[i] = 12.0; }
  d->alphaES = 1.0;
  d->kES[0] = d->kES[1] = d->kES[2] = 10;
  d->conv_tol_ind = 1e-6;
  d->min_scf_pol_iter = 3;
  d->max_scf_pol_iter = 100;
  d->algo_moment_from_pola = 0;
  d->omegakO = 0.7;
}

const char* mdff_version(void) { return "mdff_gpu 1 (gfx950)"; }

const char* mdff_last_error(const mdff_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mdff_ctx_create(const mdff_desc* d, const int32_t* itype, mdff_ctx** out) {
  if (!out) return MDFF_ERR_BAD_ARG;
  *out = nullptr;
  if (!d || !itype || d->abi_version != MDFF_ABI_VERSION || d->natm <= 0 || d->ntype <= 0 || d->ntype > 8)
    return MDFF_ERR_BAD_ARG;
  if (d->lmorse && (d->lnmlj || d->lbmhft || d->lbmhftd)) return MDFF_ERR_UNSUPPORTED;
  if (d->algo_moment_from_pola < 0 || d->algo_moment_from_pola > 2) return MDFF_ERR_UNSUPPORTED;
  for (int i = 0; i < d->natm; ++i)
    if (itype[i] < 1 || itype[i] > d->ntype) return MDFF_ERR_BAD_ARG;
  for (int i = 0; i < 4096; ++i)
    if (d->pol_damp_k[i] < 0 || d->pol_damp_k[i] > 8) return MDFF_ERR_BAD_ARG;  // src/pim.f90:201-204
  mdff_ctx* c = new mdff_ctx();
  int rc = ctx_init(c, d, itype);
  if (rc != MDFF_OK) {
    fprintf(stderr, "mdff_ctx_create: %s\n", c->err.c_str());
    mdff_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return MDFF_OK;
}

int mdff_ctx_destroy(mdff_ctx* c) {
  if (!c) return MDFF_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* ps[] = {c->rx, c->ry, c->rz, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz, c->fsx, c->fsy, c->fsz,
                c->sx, c->sy, c->sz, c->xs, c->ys, c->zs, c->bx, c->by, c->bz, c->bf, c->itype, c->q, c->mass,
                c->pol, c->invpol, c->dipia, c->mu, c->mu_ind, c->efield, c->e_stat, c->e_tot, c->e_ind,
                c->hist, c->frec, c->zeros, c->arec, c->ameta, c->peratom, c->red, c->red_part, c->d_pp, c->cell_of, c->cell_key,
                c->cell_atoms, c->cell_atoms_in, c->cell_start, c->nl_count, c->nl_start, c->nl_j,
                c->nl_flag, c->nl_count_in, c->nl_j_tmp, c->nl_flag_tmp, c->sort_tmp, c->d_flag, c->raw,
                c->cx, c->cy, c->cz, c->fin_part, c->nhc_part, c->nhc_st, c->peratom_lj, c->f_lj, c->lj_part};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (c->polar) (void)hipFree(c->polar);
  recip_free(c);
  efg_free(c);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  if (c->h_red) (void)hipHostFree(c->h_red);
  if (c->h_raw) (void)hipHostFree(c->h_raw);
  if (c->h_chk) (void)hipHostFree(c->h_chk);
  if (c->chk_ev) (void)hipEventDestroy(c->chk_ev);
  if (c->ov_tab) (void)hipEventDestroy(c->ov_tab);
  if (c->ov_pair) (void)hipEventDestroy(c->ov_pair);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->kO_ws) (void)hipFree(c->kO_ws);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->evpool)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MDFF_OK;
}

static int copy3_in(mdff_ctx* c, double* dx, double* dy, double* dz, const double* x, const double* y,
                    const double* z) {
  if (!c || !x || !y || !z) return MDFF_ERR_BAD_ARG;
  const size_t b = sizeof(double) * c->n;
  HIP_TRY(c, hipMemcpyAsync(dx, x, b, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dy, y, b, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(dz, z, b, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

static int copy3_out(mdff_ctx* c, const double* dx, const double* dy, const double* dz, double* x, double* y,
                     double* z) {
  if (!c || !x || !y || !z) return MDFF_ERR_BAD_ARG;
  const size_t b = sizeof(double) * c->n;
  HIP_TRY(c, hipMemcpyAsync(x, dx, b, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(y, dy, b, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(z, dz, b, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

int mdff_set_positions(mdff_ctx* c, const double* x, const double* y, const double* z) {
  return copy3_in(c, c->rx, c->ry, c->rz, x, y, z);
}
int mdff_get_positions(mdff_ctx* c, double* x, double* y, double* z) {
  return copy3_out(c, c->rx, c->ry, c->rz, x, y, z);
}
int mdff_set_velocities(mdff_ctx* c, const double* x, const double* y, const double* z) {
  return copy3_in(c, c->vx, c->vy, c->vz, x, y, z);
}
int mdff_get_velocities(mdff_ctx* c, double* x, double* y, double* z) {
  return copy3_out(c, c->vx, c->vy, c->vz, x, y, z);
}
int mdff_get_forces(mdff_ctx* c, double* x, double* y, double* z) {
  return copy3_out(c, c->fx, c->fy, c->fz, x, y, z);
}

int mdff_set_cell(mdff_ctx* c, const double cell[9]) {
  if (!c || !cell) return MDFF_ERR_BAD_ARG;
  memcpy(c->A, cell, sizeof c->A);
  lattice(c->A, c->B, &c->omega, c->W);
  c->lists_valid = 0;
  return MDFF_OK;
}

int mdff_set_itime(mdff_ctx* c, int32_t itime) {
  if (!c) return MDFF_ERR_BAD_ARG;
  c->itime = itime;
  return MDFF_OK;
}

int mdff_vnlist_build(mdff_ctx* c) {
  if (!c) return MDFF_ERR_BAD_ARG;
  MDFF_TRY(nl_build(c));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

int mdff_vnlistcheck(mdff_ctx* c, int32_t* rebuilt) {
  if (!c) return MDFF_ERR_BAD_ARG;
  int rb = 0;
  MDFF_TRY(nl_check(c, &rb));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (rebuilt) *rebuilt = rb;
  return MDFF_OK;
}

int mdff_get_pairs(mdff_ctx* c, int32_t which, int32_t* point, int32_t* list, int64_t cap, int64_t* n) {
  if (!c || (which != 0 && which != 1)) return MDFF_ERR_BAD_ARG;
  return nl_get_pairs(c, which, point, list, cap, n);
}

int mdff_engforce(mdff_ctx* c, double* fx, double* fy, double* fz, mdff_terms* terms) {
  if (!c) return MDFF_ERR_BAD_ARG;
  mdff_terms t;
  MDFF_TRY(engforce(c, &t));
  if (terms) *terms = t;
  if (fx && fy && fz) MDFF_TRY(mdff_get_forces(c, fx, fy, fz));
  return MDFF_OK;
}

int mdff_last_terms(mdff_ctx* c, mdff_terms* terms) {
  if (!c || !terms) return MDFF_ERR_BAD_ARG;
  *terms = c->last_terms;
  return MDFF_OK;
}

int mdff_get_dipoles(mdff_ctx* c, double* mu, double* ef) {
  if (!c) return MDFF_ERR_BAD_ARG;
  const size_t b = 3 * sizeof(double) * c->n;
  if (mu) HIP_TRY(c, hipMemcpyAsync(mu, c->mu, b, hipMemcpyDeviceToHost, c->stream));
  if (ef) HIP_TRY(c, hipMemcpyAsync(ef, c->efield, b, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

int mdff_get_efg(mdff_ctx* c, double* efg) {
  if (!c || !efg) return MDFF_ERR_BAD_ARG;
  const size_t b = 9 * sizeof(double) * c->n;
  if (c->efg) {
    HIP_TRY(c, hipMemcpyAsync(efg, c->efg, b, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  } else {
    // before the first Coulomb evaluation (or without lcoulomb): efg_t = 0
    // as allocated (src/coulomb.f90:434-439)
    memset(efg, 0, b);
  }
  return MDFF_OK;
}

static double upot_of(const mdff_terms& t) { return t.u_lj + t.u_bmhft + t.u_morse + t.u_coul; }

int mdff_md_vv(mdff_ctx* c, int32_t nsteps, double dt, double* ekin, double* upot) {
  if (!c || nsteps < 0) return MDFF_ERR_BAD_ARG;
  const int n = c->n;
  std::vector<mdff_terms> held(RAW_SLOTS);
  int st0 = 0;   // first step of the slots in flight
  for (int st = 0; st < nsteps; ++st) {
    const int slot = st - st0;
    double* raw = c->raw + (size_t)RAW_N * slot;
    // the first position update; later ones are made by the previous step's
    // vv_end_launch
    if (st == 0) MDFF_TRY(vv_pos_launch(c, dt));
    c->itime++;   // md%itime of this step (src/runmd.f90:257), read by the ASPC order
    MDFF_TRY(engforce_launch(c, raw, &held[slot], true));
    double* ke_part = c->fin_part + 1024 * 16;
    MDFF_TRY(claim_partials(c, ke_part));
    // velocity update, kinetic sum, vnlistcheck after the step (finished by
    // the next step's engforce_launch once that step's reciprocal work is
    // queued, or below) and the next step's position update
    MDFF_TRY(vv_end_launch(c, dt, st + 1 < nsteps, ke_part));
    if (ekin) MDFF_TRY(queue_final(c, ke_part, nblk(n), 1, raw + RAW_EKIN));
    // the step's final sums run in the next step's first launch (the phase
    // tables), or here before the sums go to the host
    if (slot == RAW_SLOTS - 1 || st == nsteps - 1) MDFF_TRY(flush_finals(c));
    if (slot == RAW_SLOTS - 1 || st == nsteps - 1) {
      MDFF_TRY(fetch_raw(c, slot + 1));
      for (int k = 0; k <= slot; ++k) {
        const double* h = c->h_raw + (size_t)RAW_N * k;
        engforce_finish(c, h, &held[k]);
        c->last_terms = held[k];
        if (ekin) ekin[st0 + k] = 0.5 * h[RAW_EKIN];
        if (upot) upot[st0 + k] = upot_of(held[k]);
      }
      st0 = st + 1;
    }
  }
  int rb = 0;
  MDFF_TRY(nl_check_finish(c, &rb));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

static int kinetic(mdff_ctx* c, double* kin, double* vsum) {
  hipLaunchKernelGGL(k_kin, dim3(nblk(c->n)), dim3(256), 0, c->stream, c->n, c->mass, c->vx, c->vy, c->vz,
                     extra(c));
  double h[4];
  MDFF_TRY(reduce_cols(c, extra(c), 4, h));
  *kin = h[0] * 0.5;
  if (vsum) { vsum[0] = h[1]; vsum[1] = h[2]; vsum[2] = h[3]; }
  return MDFF_OK;
}

// get_yoshida_weigth (src/integration.f90:1059-1104); order 5 keeps the
// reference's placement of the middle weight last
static void yoshida_weights(int order, double w[9]) {
  for (int k = 0; k < 9; ++k) w[k] = 0.0;
  w[0] = 1.0;
  if (order == 3) {
    w[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0)); w[1] = 1.0 - 2.0 * w[0]; w[2] = w[0];
  } else if (order == 5) {
    w[0] = 1.0 / (4.0 - pow(4.0, 1.0 / 3.0)); w[1] = w[2] = w[3] = w[0]; w[4] = 1.0 - 4.0 * w[0];
  } else if (order == 7) {
    w[0] = .78451361047756; w[1] = .235573213359357; w[2] = -1.17767998417887;
    w[3] = 1.0 - 2.0 * (w[0] + w[1] + w[2]);
    w[4] = w[2]; w[5] = w[1]; w[6] = w[0];
  } else if (order == 9) {
    w[0] = 0.192; w[1] = 0.554910818409783619692725006662999;
    w[2] = 0.124659619941888644216504240951585; w[3] = -0.843182063596933505315033808282941;
    w[4] = 1.0 - 2.0 * (w[0] + w[1] + w[2] + w[3]);
    w[5] = w[3]; w[6] = w[2]; w[7] = w[1]; w[8] = w[0];
  }
}

int mdff_md_nhcn(mdff_ctx* c, int32_t nsteps, double dt, double temp, double timesca, int32_t nhc, int32_t yosh,
                 int32_t mults, double* xi, double* vxi, double* ekin, double* upot, double* e_nvt) {
  if (!c || nhc < 2 || nhc > 16 || !xi || !vxi || (yosh != 1 && yosh != 3 && yosh != 5 && yosh != 7 && yosh != 9) ||
      mults < 1 || nsteps < 0)
    return MDFF_ERR_BAD_ARG;
  const int n = c->n, nb = nblk(n);
  // thermostat masses (src/integration.f90:563-566) and the chain state
  double st[NHC_N] = {0};
  const double L = 3.0 * (double)n;
  for (int h = 0; h < nhc; ++h) {
    st[NHC_XI + h] = xi[h];
    st[NHC_VXI + h] = vxi[h];
    st[NHC_Q + h] = pow(timesca, 2.0) * temp;
  }
  st[NHC_Q] = st[NHC_Q] * L;
  HIP_TRY(c, hipMemcpyAsync(c->nhc_st, st, sizeof st, hipMemcpyHostToDevice, c->stream));
  NhcArgs a{};
  a.nhc = nhc;
  a.yosh = yosh;
  a.mults = mults;
  a.nb = nb;
  a.L = L;
  a.temp = temp;
  a.dt = dt;
  yoshida_weights(yosh, a.w);
  a.st = c->nhc_st;
  double* pa = c->nhc_part;        // after the second half kick
  double* pb = c->nhc_part + nb;   // after the second chain's scale (the step's kinetic energy)
  // kinetic energy of the starting velocities (calc_temp, :569)
  MDFF_TRY(claim_partials(c, pb));
  hipLaunchKernelGGL(k_nhc_scale_kin, dim3(nb), dim3(256), 0, c->stream, n, 0, c->nhc_st, c->mass, c->vx, c->vy,
                     c->vz, pb);
  std::vector<mdff_terms> held(RAW_SLOTS);
  int st0 = 0;
  for (int step = 0; step < nsteps; ++step) {
    const int slot = step - st0;
    double* raw = c->raw + (size_t)RAW_N * slot;
    // chain_nhcn, then prop_pos_vel_verlet: v = s v, half kick, drift
    a.part = pb;
    a.envt = nullptr;
    hipLaunchKernelGGL(k_nhc_chain, dim3(1), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_nhc_scale_kick_drift, dim3(nb), dim3(256), 0, c->stream, n, c->nhc_st, dt * 0.5, dt,
                       c->mass, c->rx, c->ry, c->rz, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz);
    c->itime++;   // md%itime of this step (src/runmd.f90:257)
    MDFF_TRY(engforce_launch(c, raw, &held[slot], true));
    hipLaunchKernelGGL(k_nhc_kick_kin, dim3(nb), dim3(256), 0, c->stream, n, dt * 0.5, c->mass, c->vx, c->vy,
                       c->vz, c->fx, c->fy, c->fz, pa);
    // chain_nhcn again; its scale and the step's kinetic energy; e_nvt
    a.part = pa;
    a.envt = raw + RAW_ENVT;
    hipLaunchKernelGGL(k_nhc_chain, dim3(1), dim3(256), 0, c->stream, a);
    MDFF_TRY(claim_partials(c, pb));
    hipLaunchKernelGGL(k_nhc_scale_kin, dim3(nb), dim3(256), 0, c->stream, n, 1, c->nhc_st, c->mass, c->vx,
                       c->vy, c->vz, pb);
    MDFF_TRY(queue_final(c, pb, nb, 1, raw + RAW_EKIN));
    // vnlistcheck after the step (src/runmd.f90:294-304), finished by the
    // next step's force evaluation
    if (c->d.lvnlist) MDFF_TRY(nl_check_launch(c));
    if (slot == RAW_SLOTS - 1 || step == nsteps - 1) {
      MDFF_TRY(flush_finals(c));
      MDFF_TRY(fetch_raw(c, slot + 1));
      for (int k = 0; k <= slot; ++k) {
        const double* h = c->h_raw + (size_t)RAW_N * k;
        engforce_finish(c, h, &held[k]);
        c->last_terms = held[k];
        if (ekin) ekin[st0 + k] = 0.5 * h[RAW_EKIN];
        if (upot) upot[st0 + k] = upot_of(held[k]);
        if (e_nvt) e_nvt[st0 + k] = h[RAW_ENVT];
      }
      st0 = step + 1;
    }
  }
  int rb = 0;
  MDFF_TRY(nl_check_finish(c, &rb));
  HIP_TRY(c, hipMemcpyAsync(st, c->nhc_st, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int h = 0; h < nhc; ++h) {
    xi[h] = st[NHC_XI + h];
    vxi[h] = st[NHC_VXI + h];
  }
  return MDFF_OK;
}

int mdff_csvr_rescale(mdff_ctx* c, double temp, double taut, double r1, double sumnoise, double* conint) {
  if (!c) return MDFF_ERR_BAD_ARG;
  const int n = c->n;
  double ekin;
  MDFF_TRY(kinetic(c, &ekin, nullptr));
  const int ndeg = 3 * n - 3;
  const double sigma = ndeg * temp * 0.5;
  const double factor = taut > 0.1 ? exp(-1.0 / taut) : 0.0;
  const double knew = ekin + (1.0 - factor) * (sigma * (sumnoise + r1 * r1) / ndeg - ekin) +
                      2.0 * r1 * sqrt(ekin * sigma / ndeg * (1.0 - factor) * factor);
  if (conint) *conint = *conint - (knew - ekin);
  const double lambda = sqrt(knew / ekin);
  hipLaunchKernelGGL(k_scale_v, dim3(nblk(n)), dim3(256), 0, c->stream, n, lambda, 0.0, 0.0, 0.0, c->vx, c->vy,
                     c->vz);
  double k2, vs[3];
  MDFF_TRY(kinetic(c, &k2, vs));
  hipLaunchKernelGGL(k_scale_v, dim3(nblk(n)), dim3(256), 0, c->stream, n, 1.0, vs[0] / n, vs[1] / n, vs[2] / n,
                     c->vx, c->vy, c->vz);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

int mdff_rescale_velocities(mdff_ctx* c, double lambda) {
  if (!c || !(lambda >= 0.0)) return MDFF_ERR_BAD_ARG;
  const int n = c->n;
  hipLaunchKernelGGL(k_scale_v, dim3(nblk(n)), dim3(256), 0, c->stream, n, lambda, 0.0, 0.0, 0.0, c->vx, c->vy,
                     c->vz);
  double k2, vs[3];
  MDFF_TRY(kinetic(c, &k2, vs));
  hipLaunchKernelGGL(k_scale_v, dim3(nblk(n)), dim3(256), 0, c->stream, n, 1.0, vs[0] / n, vs[1] / n, vs[2] / n,
                     c->vx, c->vy, c->vz);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MDFF_OK;
}

int mdff_calc_temp(mdff_ctx* c, double boltz_unit, double* temp, double* ekin) {
  if (!c) return MDFF_ERR_BAD_ARG;
  double e;
  MDFF_TRY(kinetic(c, &e, nullptr));
  if (ekin) *ekin = e;
  if (temp) *temp = 2.0 * e / boltz_unit / (3.0 * c->n);
  return MDFF_OK;
}

int mdff_profile(mdff_ctx* c, int32_t on) {
  if (!c) return MDFF_ERR_BAD_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->prof.clear();
  c->evused = 0;
  c->prof_on = on != 0;
  return MDFF_OK;
}

int mdff_profile_read(mdff_ctx* c, double* ms_total, int32_t* count, int32_t nkinds) {
  if (!c || !ms_total || !count || nkinds <= 0) return MDFF_ERR_BAD_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < nkinds; ++k) { ms_total[k] = 0.0; count[k] = 0; }
  for (const ProfRec& r : c->prof) {
    if (r.kind >= nkinds || !r.a || !r.b) continue;
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, r.a, r.b));
    ms_total[r.kind] += ms;
    count[r.kind] += 1;
  }
  return MDFF_OK;
}

const char* mdff_profile_name(int32_t kind) { return (kind >= 0 && kind < PK_COUNT) ? kProfNames[kind] : ""; }

int mdff_last_timings(mdff_ctx* c, double* ms, int32_t* n) {
  if (!c || !ms) return MDFF_ERR_BAD_ARG;
  read_timings(c);
  for (int k = 0; k < c->n_last; ++k) ms[k] = c->last_ms[k];
  if (n) *n = c->n_last;
  return MDFF_OK;
}

}  // extern "C"
