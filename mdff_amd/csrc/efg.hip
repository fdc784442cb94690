// efg.hip -- electric field gradient at the ions (doefg) for gfx950.
//
// Reference: multipole_ES (src/ewald.f90:49-314) with do_efg, its parts
//   real space   multipole_ES_dir  src/ewald_dir.f90:543-546, 577-582
//   reciprocal   multipole_ES_rec  src/ewald_rec.f90:116-148 (S(k)),
//                                  :209-216 (k_a k_b A Re(S e^{-ik.r})), :302, :351
//   self         src/ewald.f90:214-216; total x coul_unit :231
// and its copy into efg_t (src/engforce_driver.f90:171).
//
// The gradient is an output option, not part of the force: every part runs
// in FP64 whatever the precision form, with the reference's direct cos/sin
// of k.r per (k, atom).
//  - k_efg_sk: S(k) over the k box (k_x >= 0 half, A = 0 outside the half
//    space), one thread per k, atoms staged through LDS, atom chunks over
//    blockIdx.y into partials;
//  - k_efg_w:  the partials summed in chunk order, W = A S;
//  - k_efg_rec: one thread per atom, k staged through LDS, k chunks over
//    blockIdx.y into partials of the six upper components;
//  - k_efg_combine: (dir + rec + self) coul_unit into the reference layout
//    efg(3,3,natm); the reciprocal part only on the upper triangle and the
//    diagonal, as the reference fills it.
#include <algorithm>

#include "common.h"

namespace {

constexpr int kTile = 256;

struct EfgArgs {
  int n, nk;                         // atoms, k box entries
  int chunk;                         // atoms (k_efg_sk) or k entries (k_efg_rec) per blockIdx.y
  int ldip;                          // dipoles enter S(k) (task q-mu or mu-mu)
  const double *rx, *ry, *rz, *q, *mu;
  const double* kA;                  // A(k) on the box, 0 where excluded
  const double* kvec;                // kx, ky, kz, kcoe per box entry
  double* part;                      // S partials [chunk][2 nk] / rec partials [chunk][6][n]
  double* w;                         // W = A S (2 nk)
};

__global__ void __launch_bounds__(kTile) k_efg_sk(EfgArgs a) {
  __shared__ double sr[kTile][7];
  const int k = blockIdx.x * kTile + threadIdx.x;
  const bool on = k < a.nk && a.kA[k] != 0.0;
  double kx = 0, ky = 0, kz = 0;
  if (on) {
    kx = a.kvec[4 * k];
    ky = a.kvec[4 * k + 1];
    kz = a.kvec[4 * k + 2];
  }
  const int i0 = blockIdx.y * a.chunk, i1 = min(a.n, i0 + a.chunk);
  double rR = 0.0, rI = 0.0;
  for (int t0 = i0; t0 < i1; t0 += kTile) {
    const int m = min(kTile, i1 - t0);
    __syncthreads();
    if (threadIdx.x < m) {
      const int i = t0 + threadIdx.x;
      sr[threadIdx.x][0] = a.rx[i];
      sr[threadIdx.x][1] = a.ry[i];
      sr[threadIdx.x][2] = a.rz[i];
      sr[threadIdx.x][3] = a.q[i];
      if (a.ldip) {
        sr[threadIdx.x][4] = a.mu[3 * i];
        sr[threadIdx.x][5] = a.mu[3 * i + 1];
        sr[threadIdx.x][6] = a.mu[3 * i + 2];
      }
    }
    __syncthreads();
    if (!on) continue;
    for (int l = 0; l < m; ++l) {
      const double kr = kx * sr[l][0] + ky * sr[l][1] + kz * sr[l][2];
      double s, c;
      sincos(kr, &s, &c);
      rR += sr[l][3] * c;
      rI += sr[l][3] * s;
      if (a.ldip) {
        const double kmu = sr[l][4] * kx + sr[l][5] * ky + sr[l][6] * kz;
        rR -= kmu * s;
        rI += kmu * c;
      }
    }
  }
  if (k < a.nk) {
    a.part[(size_t)blockIdx.y * 2 * a.nk + 2 * k] = rR;
    a.part[(size_t)blockIdx.y * 2 * a.nk + 2 * k + 1] = rI;
  }
}

__global__ void k_efg_w(EfgArgs a, int nch) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nk) return;
  double rR = 0.0, rI = 0.0;
  for (int c = 0; c < nch; ++c) {
    rR += a.part[(size_t)c * 2 * a.nk + 2 * k];
    rI += a.part[(size_t)c * 2 * a.nk + 2 * k + 1];
  }
  const double A = a.kA[k];
  a.w[2 * k] = A * rR;
  a.w[2 * k + 1] = A * rI;
}

__global__ void __launch_bounds__(kTile) k_efg_rec(EfgArgs a) {
  __shared__ double sk[kTile][5];
  const int i = blockIdx.x * kTile + threadIdx.x;
  double x = 0, y = 0, z = 0;
  if (i < a.n) {
    x = a.rx[i];
    y = a.ry[i];
    z = a.rz[i];
  }
  const int k0 = blockIdx.y * a.chunk, k1 = min(a.nk, k0 + a.chunk);
  double g[6] = {0, 0, 0, 0, 0, 0};
  for (int t0 = k0; t0 < k1; t0 += kTile) {
    const int m = min(kTile, k1 - t0);
    __syncthreads();
    if (threadIdx.x < m) {
      const int k = t0 + threadIdx.x;
      sk[threadIdx.x][0] = a.kvec[4 * k];
      sk[threadIdx.x][1] = a.kvec[4 * k + 1];
      sk[threadIdx.x][2] = a.kvec[4 * k + 2];
      sk[threadIdx.x][3] = a.w[2 * k];
      sk[threadIdx.x][4] = a.w[2 * k + 1];
    }
    __syncthreads();
    if (i >= a.n) continue;
    for (int l = 0; l < m; ++l) {
      const double wr = sk[l][3], wi = sk[l][4];
      if (wr == 0.0 && wi == 0.0) continue;   // excluded k (A = 0): adds exact zeros
      const double kx = sk[l][0], ky = sk[l][1], kz = sk[l][2];
      double s, c;
      sincos(kx * x + ky * y + kz * z, &s, &c);
      const double r2 = wr * c + wi * s;     // A (rho_R cos + rho_I sin)
      g[0] += kx * kx * r2;
      g[1] += ky * ky * r2;
      g[2] += kz * kz * r2;
      g[3] += kx * ky * r2;
      g[4] += kx * kz * r2;
      g[5] += ky * kz * r2;
    }
  }
  if (i < a.n)
    for (int v = 0; v < 6; ++v) a.part[((size_t)blockIdx.y * 6 + v) * a.n + i] = g[v];
}

// efg(3,3,n) column-major per atom.  dir: 6 per atom or null (no doefg: the
// self term alone); rec partials [nch][6][n] scaled by rscale = 2 (half mesh)
// x 4 pi / V.
__global__ void k_efg_combine(int n, const double* __restrict__ dir, const double* __restrict__ rec,
                              int nch, double rscale, const double* __restrict__ q, double self2,
                              double coul, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int up[6] = {0, 4, 8, 3, 6, 7};   // (1,1) (2,2) (3,3) (1,2) (1,3) (2,3)
  constexpr int lo[3] = {1, 2, 5};            // (2,1) (3,1) (3,2) of xy, xz, yz
  double* o = out + 9 * (size_t)i;
  for (int v = 0; v < 6; ++v) {
    const double d = dir ? dir[6 * (size_t)i + v] : 0.0;
    double r = 0.0;
    if (rec) {
      for (int c = 0; c < nch; ++c) r += rec[((size_t)c * 6 + v) * n + i];
      r *= rscale;
    }
    const double self = v < 3 ? -self2 * q[i] : 0.0;
    o[up[v]] = (d + r + self) * coul;
    if (v >= 3) o[lo[v - 3]] = d * coul;
  }
}

}  // namespace

int efg_alloc(mdff_ctx* c) {
  if (c->efg) return MDFF_OK;
  const int n = c->n;
  // every entry is written by each efg_eval (its only caller)
  HIP_TRY(c, hipMalloc((void**)&c->efg, sizeof(double) * 9 * (size_t)n));
  if (!c->d.doefg || !c->d.lcoulomb) return MDFF_OK;
  // atom chunks of S(k): about 2048 blocks; k chunks of the gather: about
  // 2048 blocks
  const int nk = c->nk_box;
  const int kb = div_up(nk, kTile);
  c->efg_sk_nch = std::max(1, std::min(div_up(n, kTile), div_up(2048, kb)));
  c->efg_sk_chunk = div_up(n, c->efg_sk_nch);
  c->efg_sk_nch = div_up(n, c->efg_sk_chunk);
  const int ab = div_up(n, kTile);
  c->efg_rec_nch = std::max(1, std::min(kb, div_up(2048, ab)));
  c->efg_rec_chunk = div_up(nk, c->efg_rec_nch);
  c->efg_rec_nch = div_up(nk, c->efg_rec_chunk);
  const size_t part = std::max((size_t)c->efg_sk_nch * 2 * nk, (size_t)c->efg_rec_nch * 6 * n);
  HIP_TRY(c, hipMalloc((void**)&c->efg_dir, sizeof(double) * 6 * (size_t)n));
  HIP_TRY(c, hipMalloc((void**)&c->efg_part, sizeof(double) * part));
  HIP_TRY(c, hipMalloc((void**)&c->efg_w, sizeof(double) * 2 * (size_t)nk));
  return MDFF_OK;
}

void efg_free(mdff_ctx* c) {
  for (double* p : {c->efg, c->efg_dir, c->efg_part, c->efg_w})
    if (p) (void)hipFree(p);
  c->efg = c->efg_dir = c->efg_part = c->efg_w = nullptr;
}

// efg_t of a Coulomb evaluation whose atoms, lists and dipoles (c->mu) are
// current: with doefg the three parts, otherwise the self term alone.
int efg_eval(mdff_ctx* c) {
  MDFF_TRY(efg_alloc(c));
  const int n = c->n;
  const double alpha2 = c->alpha * c->alpha;
  const double selfa = c->alpha / sqrt(MDFF_PI), selfa2 = 2.0 * selfa * alpha2 / 3.0;
  const double fpi_V = 4.0 * MDFF_PI / c->omega;
  if (!c->d.doefg) {
    hipLaunchKernelGGL(k_efg_combine, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, nullptr, nullptr, 0,
                       0.0, c->q, 2.0 * selfa2, c->d.coul_unit, c->efg);
    HIP_TRY(c, hipGetLastError());
    return MDFF_OK;
  }
  MDFF_TRY(realspace_efg(c, c->efg_dir));
  EfgArgs a{};
  a.n = n;
  a.nk = c->nk_box;
  a.ldip = (c->task_qd || c->task_dd) ? 1 : 0;
  a.rx = c->rx; a.ry = c->ry; a.rz = c->rz;
  a.q = c->q;
  a.mu = c->mu;
  a.kA = c->kA;
  a.kvec = c->kvec;
  a.part = c->efg_part;
  a.w = c->efg_w;
  a.chunk = c->efg_sk_chunk;
  hipLaunchKernelGGL(k_efg_sk, dim3(div_up(a.nk, kTile), c->efg_sk_nch), dim3(kTile), 0, c->stream, a);
  hipLaunchKernelGGL(k_efg_w, dim3(div_up(a.nk, 256)), dim3(256), 0, c->stream, a, c->efg_sk_nch);
  a.chunk = c->efg_rec_chunk;
  hipLaunchKernelGGL(k_efg_rec, dim3(div_up(n, kTile), c->efg_rec_nch), dim3(kTile), 0, c->stream, a);
  hipLaunchKernelGGL(k_efg_combine, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, c->efg_dir, c->efg_part,
                     c->efg_rec_nch, 2.0 * fpi_V, c->q, 2.0 * selfa2, c->d.coul_unit, c->efg);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}
