// fit.hip -- batched force-field fit for gfx950: many PIM snapshots of uneven
// size, forward (forces, stress, converged induced dipoles) and the gradient
// of a force/stress loss with respect to the potential parameters through the
// converged dipole SCF (implicit adjoint), in one persistent launch.
//
// Work layout: one launch per phase over all snapshots at once.  The atoms of
// every snapshot are cut into blocks of blockDim atoms and its k-vectors into
// blocks of blockDim k-vectors (host-made maps: block -> snapshot and range),
// so a launch spreads the large and the small snapshots alike over the GPU.
// Per snapshot sums (SCF residuals, stress, loss, gradients) are block
// partials added in block order by a one-block control kernel, which also
// advances the SCF / adjoint iterations of each snapshot and counts those
// still iterating; the host repeats an iteration's launches while any is.
// A snapshot that has converged skips the later iterations' work.
//
// Physics (the same terms as the single-configuration path, per snapshot):
//   Buckingham  V = A exp(-r/rho) - C/r^6       (src/non_bonded.f90:1156-1189)
//   Ewald q-q, q-mu (TT-damped), mu-mu, real + reciprocal + self
//                                               (src/ewald_dir.f90:230-644,
//                                                src/ewald_rec.f90:116-358)
//   SCF mu = alpha (E_stat(damped) + T mu)      (src/moment_from_pola.f90:240-485)
// Loss per snapshot: sum_i |F_i - F*_i|^2 + w sum_ab (tau_ab - tau*_ab)^2
// with tau = tau_nonb + tau_coul (GPa).
//
// Gradient: with g = dL/dF, G = dL/dtau and R(mu) = mu/alpha - E_stat - T mu,
//   dL/dtheta = dL/dtheta|_mu - lambda . dR/dtheta,  (1/alpha - T) lambda = h,
//   h = (dF/dmu)^T g + (dtau/dmu)^T G.
// The adjoint system is solved with the SCF's own iteration (same operator,
// T symmetric).  Real-space h, the explicit Buckingham and Tang-Toennies
// derivatives and the lambda terms are pair loops; the reciprocal h needs two
// extra structure factors' worth of gathers (see k_fit, phase 6).
//
// Precision: T = double (FP64 form, used for the finite-difference checks) or
// float (pair arithmetic and phase-table products in FP32); minimum image,
// phase angles, every accumulation, the SCF and adjoint residuals in FP64.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

template <typename T>
struct c2 {
  T re, im;
};
template <typename T>
__device__ __forceinline__ c2<T> cm(c2<T> a, c2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// parameter kinds of the fit vector
enum { FP_A = 0, FP_RHO = 1, FP_C = 2, FP_ALPHA = 3, FP_TTB = 4, FP_TTC = 5 };
constexpr int FIT_MAXP = 32;  // parameters
constexpr int FIT_NT = 4;     // types

struct FitSnap {
  int n, a0;                 // atoms, first atom
  int nk, k0;                // k-vectors, first k
  int Kx, Ky, Kz, ncol;
  long long tab0;            // first table element
  long long nl0;             // first neighbour entry
  int row0;                  // first row start (n + 1 entries)
  int npol;
  double A[9];               // cell, column-major
  double omega, rc2_sr, rc2_c, alpha;
};

// per type pair (it, jt): Buckingham and the damping of the field at a
// dipole of type it from a charge of type jt, with their gradient slots
struct FitPair {
  double A, B, C, rho;
  double db, dc;
  int dk, ldamp;
  int sA, sRho, sC, sB, sC2;   // slots (-1 none): A, rho, C, TT b, TT c
  int pad;
};

struct FitArgs {
  int nsnap, ntype, want_grad, min_iter, max_iter, ncol_max;
  double tol, tol_adj, coul_unit, press_unit, w_stress;
  const FitSnap* snaps;
  const FitPair* pairs;                 // ntype x ntype, [it + ntype jt]
  double alpha_t[FIT_NT];               // polarisability per type (0: none)
  int s_alpha[FIT_NT];                  // slot of alpha per type
  const int* itype;                     // global atoms, 0-based types
  const double* q;
  const double *sx, *sy, *sz;           // fractional coordinates
  const int* row;                       // row starts (local entry index), per snapshot n+1
  const int* nl;                        // local neighbour indices
  const double* kv;                     // per k: kx ky kz A kcoe
  const int* kn;                        // per k: nx, ny, nz
  const void* tab;                      // phase tables [snap][col][atom] complex T
  const double* fref;                   // 3 x atoms
  const double* tauref;                 // 9 x snaps
  double* ws;                           // per atom scratch, FIT_WS doubles
  double* kws;                          // per k scratch, FIT_KWS doubles
  double* F;                            // 3 x atoms out
  double* tau;                          // 9 x snaps out
  double* loss;                         // per snap
  double* grad;                         // FIT_MAXP x snaps
  int* nscf;                            // 2 x snaps: SCF and adjoint iterations
  // blocks: atom blocks (snapshot, first atom, end) and k blocks (snapshot,
  // first k, end); per snapshot the first atom / k block (nsnap + 1 each)
  const int4* ablk;
  const int4* kblk;
  const int* snap_ab0;
  const int* snap_kb0;
  double* pa;                           // atom-block partials [block][FIT_PA]
  double* pk;                           // k-block partials [block][FIT_PK]
  int* active;                          // per snap: still iterating
  int* iter;                            // per snap: iterations made
  double* hn;                           // per snap: RMS of h (adjoint tolerance)
  double* Gs;                           // per snap: symmetric stress weight (9)
  double* gsave;                        // per snap: gradient partial sums of the pair pass (FIT_MAXP)
  int* count;                           // snapshots still iterating (control kernel)
  const int4* gblk;                     // gather blocks: (snapshot, first atom, first k, k chunk)
  const int* snap_g0;                   // per snapshot: first gather record (atom x k chunk)
  const int* snap_nch;                  // per snapshot: k chunks
  double* gp;                           // gather records [record][FIT_GC]
  const int4* sblk;                     // structure-factor blocks: (snapshot, first k, first atom, chunk)
  const int* snap_s0;                   // per snapshot: first structure-factor record (k x atom chunk)
  const int* snap_sch;                  // per snapshot: atom chunks
  double* sp;                           // structure-factor records [record][2]
  int ws_dst, ws_src;                   // kf_update: ws slots (3 n each) dst = alpha src
};

constexpr int FIT_PA = 40;              // atom-block partial slots
constexpr int FIT_GC = 12;              // gather record: im3, re3, re6
constexpr int FIT_KC = 512;             // k-vectors per gather chunk
constexpr int FIT_AC = 512;             // atoms per structure-factor chunk
constexpr int FIT_PK = 8;               // k-block partial slots

// per-atom scratch: seven 3-vectors (static field, mu, total field, g, h,
// lambda, tmp), stored per snapshot as seven 3 x n blocks
constexpr int FIT_WS = 21;
// per-k scratch: S_q (2), S (2), C1 C2 C3 (2 each)
enum { K_SQ = 0, K_S = 2, K_C1 = 4, K_C2 = 6, K_C3 = 8, FIT_KWS = 10 };

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Tang-Toennies f, f' and their derivatives with respect to b and c
// (f = 1 - c e^{-x} sum_{k<=K} x^k/k!, x = b r; d f/d b = r f'/b,
// d f'/d b = f' (1 + K - x)/b, d f/d c = (f - 1)/c, d f'/d c = f'/c)
template <typename T>
__device__ __forceinline__ void tt_all(T b, T c, T r, int K, T& f, T& fd, T& fb, T& fdb, T& fc, T& fdc) {
  tt_damp<T>(b, c, r, K, f, fd);
  fb = r * fd / b;
  fdb = fd * (T(1) + (T)K - b * r) / b;
  fc = (f - T(1)) / c;
  fdc = fd / c;
}

struct Geo {
  double rx, ry, rz, r2;
};

__device__ __forceinline__ Geo min_image(const FitSnap& S, const double* sx, const double* sy, const double* sz,
                                         int i, int j) {
  double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
  dx -= nint_d(dx);
  dy -= nint_d(dy);
  dz -= nint_d(dz);
  Geo g;
  g.rx = dx * S.A[0] + dy * S.A[3] + dz * S.A[6];
  g.ry = dx * S.A[1] + dy * S.A[4] + dz * S.A[7];
  g.rz = dx * S.A[2] + dy * S.A[5] + dz * S.A[8];
  g.r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
  return g;
}

// Ewald radial functions F0..F3 of the real-space tensors (src/ewald_dir.f90:230-236)
template <typename T>
struct Erf {
  T dm1, dm2, dm3, dm5, F1, F2, F3;
};
template <typename T>
__device__ __forceinline__ Erf<T> erf_terms(T d, T r2, T alpha) {
  Erf<T> e;
  const T a2 = alpha * alpha;
  const T expon = exp(-a2 * r2) * T(0.56418958354775628695);
  const T F0 = errfc_pos<T>(alpha * d);
  e.dm1 = T(1) / d;
  e.dm2 = e.dm1 * e.dm1;
  e.dm3 = e.dm1 * e.dm2;
  e.dm5 = e.dm3 * e.dm2;
  e.F1 = F0 + T(2) * alpha * d * expon;
  e.F2 = e.F1 + T(4.0 / 3.0) * alpha * a2 * r2 * d * expon;
  e.F3 = e.F2 + T(8.0 / 15.0) * alpha * a2 * a2 * r2 * r2 * d * expon;
  return e;
}

template <typename T>
__device__ __forceinline__ c2<T> tabv(const c2<T>* tab, int n, int col, int j, bool conj) {
  c2<T> v = tab[(size_t)col * n + j];
  if (conj) v.im = -v.im;
  return v;
}

// exp(i k.r_j) from the snapshot's tables
template <typename T>
__device__ __forceinline__ c2<T> phase(const c2<T>* tab, const FitSnap& S, int j, int nx, int ny, int nz) {
  const c2<T> X = tabv(tab, S.n, nx, j, false);
  const c2<T> Y = tabv(tab, S.n, S.Kx + 1 + (ny < 0 ? -ny : ny), j, ny < 0);
  const c2<T> Z = tabv(tab, S.n, S.Kx + S.Ky + 2 + (nz < 0 ? -nz : nz), j, nz < 0);
  return cm(cm(X, Y), Z);
}

// ---------------------------------------------------------------------------
// Reciprocal building blocks (VALU, thread per k or per atom)
// ---------------------------------------------------------------------------

// S(k) = sum_j amp_j(k) exp(i k.r_j) with
// mode 0: q; 1: i k.v (v = 3-column block); 2: q + i k.mu; 3: (k.g)(q + i k.mu).
// One thread per k; the atoms' phase tables and amplitude data are staged in
// LDS in chunks of FIT_CA atoms (coalesced loads, then every thread reads its
// k's three factors from LDS).  Called by the whole workgroup (barriers).
constexpr int FIT_CA = 32;
template <typename T>
__device__ void struct_factor(const FitArgs& a, const FitSnap& S, int kb0, int kb1, int ja0, int ja1,
                              const c2<T>* tab, int mode, const double* v, const double* g,
                              double* out /* per k of the snapshot: 2 doubles at out + ostride k */, int ostride,
                              c2<T>* lt, double* la) {
  const double* q = a.q + S.a0;
  const int ncol = S.ncol;
  for (int k0 = kb0; k0 < kb1; k0 += blockDim.x) {
    const int k = k0 + threadIdx.x;
    const bool kok = k < kb1;
    const int kg = S.k0 + (kok ? k : kb0);
    const double kx = a.kv[5 * kg], ky = a.kv[5 * kg + 1], kz = a.kv[5 * kg + 2];
    const int nx = a.kn[3 * kg], ny = a.kn[3 * kg + 1], nz = a.kn[3 * kg + 2];
    const int cy = S.Kx + 1 + (ny < 0 ? -ny : ny), cz = S.Kx + S.Ky + 2 + (nz < 0 ? -nz : nz);
    const T sy = ny < 0 ? T(-1) : T(1), sz = nz < 0 ? T(-1) : T(1);
    double sr = 0.0, si = 0.0;
    for (int j0 = ja0; j0 < ja1; j0 += FIT_CA) {
      const int nc = min(FIT_CA, ja1 - j0);
      __syncthreads();
      for (int x = threadIdx.x; x < ncol * nc; x += blockDim.x) {
        const int col = x / nc, jj = x % nc;
        lt[jj * ncol + col] = tab[(size_t)col * S.n + j0 + jj];
      }
      for (int jj = threadIdx.x; jj < nc; jj += blockDim.x) {
        const int j = j0 + jj;
        la[7 * jj] = q[j];
        for (int c = 0; c < 3; ++c) {
          la[7 * jj + 1 + c] = v ? v[3 * j + c] : 0.0;
          la[7 * jj + 4 + c] = g ? g[3 * j + c] : 0.0;
        }
      }
      __syncthreads();
      if (kok) {
        // FP32 form: the tile's (at most FIT_CA) terms summed in T, the tile
        // sums in FP64; the iteration's k.v in T as well.  FP64 form: every
        // term added in FP64.
        T tr = T(0), ti = T(0);
        for (int jj = 0; jj < nc; ++jj) {
          const c2<T>* row = lt + jj * ncol;
          const c2<T> X = row[nx];
          c2<T> Y = row[cy];
          c2<T> Z = row[cz];
          Y.im *= sy;
          Z.im *= sz;
          const c2<T> e = cm(cm(X, Y), Z);
          const double* A = la + 7 * jj;
          T art = T(0), ait = T(0);
          if (mode == 1 && sizeof(T) == 4) {
            ait = (T)kx * (T)A[1] + (T)ky * (T)A[2] + (T)kz * (T)A[3];
          } else {
            double ar = 0.0, ai = 0.0;
            if (mode == 0) {
              ar = A[0];
            } else if (mode == 1) {
              ai = kx * A[1] + ky * A[2] + kz * A[3];
            } else {
              ar = A[0];
              ai = kx * A[1] + ky * A[2] + kz * A[3];
              if (mode == 3) {
                const double kg3 = kx * A[4] + ky * A[5] + kz * A[6];
                ar *= kg3;
                ai *= kg3;
              }
            }
            art = (T)ar;
            ait = (T)ai;
          }
          if constexpr (sizeof(T) == 4) {
            tr += art * e.re - ait * e.im;
            ti += art * e.im + ait * e.re;
          } else {
            sr += (double)(art * e.re - ait * e.im);
            si += (double)(art * e.im + ait * e.re);
          }
        }
        if constexpr (sizeof(T) == 4) {
          sr += (double)tr;
          si += (double)ti;
        }
      }
    }
    if (kok) {
      out[(size_t)ostride * k] = sr;
      out[(size_t)ostride * k + 1] = si;
    }
  }
  __syncthreads();
}

// the k data of one gather block staged in LDS (FIT_KC entries): k vector,
// packed (nx, ny + 512, nz + 512) and the C1 / C2 / C3 pairs the phase needs
struct GatherStage {
  double *kx, *ky, *kz, *c1, *c2, *c3;
  int* n;
};

__device__ __forceinline__ GatherStage gather_stage_of(double* lds) {
  GatherStage g;
  g.kx = lds;
  g.ky = lds + FIT_KC;
  g.kz = lds + 2 * FIT_KC;
  g.c1 = lds + 3 * FIT_KC;
  g.c2 = lds + 5 * FIT_KC;
  g.c3 = lds + 7 * FIT_KC;
  g.n = reinterpret_cast<int*>(lds + 9 * FIT_KC);
  return g;
}
// lanes per atom of the iteration's real-space dipole field (kf_dipreal)
constexpr int kDipLanes = 4;
constexpr size_t kGatherLds = sizeof(double) * 9 * FIT_KC + sizeof(int) * FIT_KC;

// the block's k-vectors kb0..kb1 (kb1 - kb0 <= FIT_KC) into LDS, by every
// thread of the block; c1 / c2i / c3: kws slots (-1: not needed)
__device__ void gather_stage(const FitArgs& a, const FitSnap& S, int kb0, int kb1, int c1, int c2i, int c3,
                             const GatherStage& g) {
  for (int k = kb0 + (int)threadIdx.x; k < kb1; k += blockDim.x) {
    const int kl = k - kb0, kg = S.k0 + k;
    g.kx[kl] = a.kv[5 * kg];
    g.ky[kl] = a.kv[5 * kg + 1];
    g.kz[kl] = a.kv[5 * kg + 2];
    g.n[kl] = a.kn[3 * kg] | ((a.kn[3 * kg + 1] + 512) << 10) | ((a.kn[3 * kg + 2] + 512) << 20);
    const double* K = a.kws + (size_t)FIT_KWS * kg;
    if (c1 >= 0) { g.c1[2 * kl] = K[c1]; g.c1[2 * kl + 1] = K[c1 + 1]; }
    if (c2i >= 0) { g.c2[2 * kl] = K[c2i]; g.c2[2 * kl + 1] = K[c2i + 1]; }
    if (c3 >= 0) { g.c3[2 * kl] = K[c3]; g.c3[2 * kl + 1] = K[c3 + 1]; }
  }
  __syncthreads();
}

// per atom j over the staged k-vectors: im3 = sum_k k Im(C1_k e^{-ik r_j}),
// re3 = sum_k k Re(C2_k e^{-ik r_j}), re6 = sum_k k k Re(C3_k e^{-ik r_j})
// (xx yy zz xy xz yz); flags pick the sums
template <typename T>
__device__ void gather(const FitSnap& S, const c2<T>* tab, int j, int nkl, bool do1, bool do2, bool do3,
                       const GatherStage& g, double im3[3], double re3[3], double re6[6]) {
  for (int c = 0; c < 3; ++c) im3[c] = re3[c] = 0.0;
  for (int c = 0; c < 6; ++c) re6[c] = 0.0;
  // the k list runs (nx, ny) rows with nz innermost: X Y of the atom is
  // formed once per row (the same products as phase())
  int prow = -1;
  c2<T> XY{T(1), T(0)};
  // Z of four k-vectors loaded ahead of their use (independent loads in
  // flight); the k-vectors themselves in order.  FP32 form: the terms of the
  // four summed in T (P), those sums in FP64; FP64 form: each term in FP64.
  constexpr int U = 4;
  constexpr bool F32 = sizeof(T) == 4;
  for (int kb = 0; kb < nkl; kb += U) {
    c2<T> Zs[U];
    int pks[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      pks[u] = g.n[min(kb + u, nkl - 1)];
      const int nz = ((pks[u] >> 20) & 1023) - 512;
      Zs[u] = tabv(tab, S.n, S.Kx + S.Ky + 2 + (nz < 0 ? -nz : nz), j, nz < 0);
    }
    T P[12] = {};
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kl = kb + u;
      if (kl >= nkl) break;
      const double kx = g.kx[kl], ky = g.ky[kl], kz = g.kz[kl];
      const int pk = pks[u];
      const int nx = pk & 1023, ny = ((pk >> 10) & 1023) - 512;
      const int row = pk & ((1 << 20) - 1);
      if (row != prow) {
        XY = cm(tabv(tab, S.n, nx, j, false), tabv(tab, S.n, S.Kx + 1 + (ny < 0 ? -ny : ny), j, ny < 0));
        prow = row;
      }
      const c2<T> e = cm(XY, Zs[u]);   // e^{+ik r_j}
      const T tx = (T)kx, ty = (T)ky, tz = (T)kz;
      if (do1) {   // Im(C e^-) = Ci er - Cr ei
        const T vt = (T)g.c1[2 * kl + 1] * e.re - (T)g.c1[2 * kl] * e.im;
        if constexpr (F32) {
          P[0] += tx * vt;
          P[1] += ty * vt;
          P[2] += tz * vt;
        } else {
          im3[0] += kx * vt;
          im3[1] += ky * vt;
          im3[2] += kz * vt;
        }
      }
      if (do2) {  // Re(C e^-) = Cr er + Ci ei
        const T vt = (T)g.c2[2 * kl] * e.re + (T)g.c2[2 * kl + 1] * e.im;
        if constexpr (F32) {
          P[3] += tx * vt;
          P[4] += ty * vt;
          P[5] += tz * vt;
        } else {
          re3[0] += kx * vt;
          re3[1] += ky * vt;
          re3[2] += kz * vt;
        }
      }
      if (do3) {
        const T vt = (T)g.c3[2 * kl] * e.re + (T)g.c3[2 * kl + 1] * e.im;
        if constexpr (F32) {
          P[6] += tx * tx * vt;
          P[7] += ty * ty * vt;
          P[8] += tz * tz * vt;
          P[9] += tx * ty * vt;
          P[10] += tx * tz * vt;
          P[11] += ty * tz * vt;
        } else {
          re6[0] += kx * kx * vt;
          re6[1] += ky * ky * vt;
          re6[2] += kz * kz * vt;
          re6[3] += kx * ky * vt;
          re6[4] += kx * kz * vt;
          re6[5] += ky * kz * vt;
        }
      }
    }
    if constexpr (F32) {
      for (int c = 0; c < 3; ++c) {
        im3[c] += (double)P[c];
        re3[c] += (double)P[3 + c];
      }
      for (int c = 0; c < 6; ++c) re6[c] += (double)P[6 + c];
    }
  }
}

// ---------------------------------------------------------------------------
// Real-space field passes
// ---------------------------------------------------------------------------

// static field of the charges at polarisable atoms, TT-damped (SCF), plus
// the lambda . dE_stat/d(b, c) gradient terms when lam != nullptr
template <typename T>
__device__ void field_static(const FitArgs& a, const FitSnap& S, int ib0, int ib1, double* Es, const double* lam,
                             double* gacc) {
  const double* sx = a.sx + S.a0;
  const double* sy = a.sy + S.a0;
  const double* sz = a.sz + S.a0;
  const int* it = a.itype + S.a0;
  const double* q = a.q + S.a0;
  const T alpha = (T)S.alpha;
  for (int i = ib0 + threadIdx.x; i < ib1; i += blockDim.x) {
    double e0 = 0, e1 = 0, e2 = 0;
    const int ti = it[i];
    if (a.alpha_t[ti] != 0.0) {
      for (int e = a.row[S.row0 + i]; e < a.row[S.row0 + i + 1]; ++e) {
        const int j = a.nl[S.nl0 + e];
        const Geo G = min_image(S, sx, sy, sz, i, j);
        if (G.r2 > S.rc2_c) continue;
        const FitPair& P = a.pairs[ti + a.ntype * it[j]];
        const T d = sqrt((T)G.r2);
        const Erf<T> E = erf_terms<T>(d, (T)G.r2, alpha);
        const T qj = (T)q[j];
        T gg = qj * E.F1 * E.dm3;
        if (P.ldamp) {
          T f, fd, fb, fdb, fc, fdc;
          tt_all<T>((T)P.db, (T)P.dc, d, P.dk, f, fd, fb, fdb, fc, fdc);
          gg -= qj * (T(1) - f) * E.dm3;
          if (lam) {
            const T lr = (T)(lam[3 * i] * G.rx + lam[3 * i + 1] * G.ry + lam[3 * i + 2] * G.rz);
            if (P.sB >= 0) gacc[P.sB] += (double)(qj * fb * E.dm3 * lr);
            if (P.sC2 >= 0) gacc[P.sC2] += (double)(qj * fc * E.dm3 * lr);
          }
        }
        e0 += (double)(gg * (T)G.rx);
        e1 += (double)(gg * (T)G.ry);
        e2 += (double)(gg * (T)G.rz);
      }
    }
    if (Es) {
      Es[3 * i] = e0;
      Es[3 * i + 1] = e1;
      Es[3 * i + 2] = e2;
    }
  }
}

// field of a dipole set v (undamped, real space only), at polarisable atoms,
// from polarisable atoms (the mu-mu pairs of src/ewald_dir.f90:557-593)
template <typename T>
__device__ void field_dip_real(const FitArgs& a, const FitSnap& S, int ib0, int ib1, const double* v, double* out,
                               double* red) {
  const double* sx = a.sx + S.a0;
  const double* sy = a.sy + S.a0;
  const double* sz = a.sz + S.a0;
  const int* it = a.itype + S.a0;
  const T alpha = (T)S.alpha;
  // gridDim.y lanes per atom stride its row; blockIdx.y picks the block's
  // share of atoms (blockDim / gridDim.y of them)
  const int L = (int)gridDim.y, sub = (int)threadIdx.x % L, per = (int)blockDim.x / L;
  const int i = ib0 + (int)blockIdx.y * per + (int)threadIdx.x / L;
  double e0 = 0, e1 = 0, e2 = 0;
  if (i < ib1 && a.alpha_t[it[i]] != 0.0) {
    for (int e = a.row[S.row0 + i] + sub; e < a.row[S.row0 + i + 1]; e += L) {
      const int j = a.nl[S.nl0 + e];
      if (a.alpha_t[it[j]] == 0.0) continue;
      const Geo G = min_image(S, sx, sy, sz, i, j);
      if (G.r2 > S.rc2_c) continue;
      const T d = sqrt((T)G.r2);
      const Erf<T> E = erf_terms<T>(d, (T)G.r2, alpha);
      const T rx = (T)G.rx, ry = (T)G.ry, rz = (T)G.rz;
      const T vx = (T)v[3 * j], vy = (T)v[3 * j + 1], vz = (T)v[3 * j + 2];
      const T rv = rx * vx + ry * vy + rz * vz;
      const T c3 = T(3) * E.F2 * E.dm5 * rv, c1 = E.F1 * E.dm3;
      e0 += (double)(c3 * rx - c1 * vx);
      e1 += (double)(c3 * ry - c1 * vy);
      e2 += (double)(c3 * rz - c1 * vz);
    }
  }
  // the L lane sums of an atom added in lane order
  red[3 * threadIdx.x] = e0;
  red[3 * threadIdx.x + 1] = e1;
  red[3 * threadIdx.x + 2] = e2;
  __syncthreads();
  if (sub == 0 && i < ib1) {
    double f0 = 0, f1 = 0, f2 = 0;
    for (int l = 0; l < L; ++l) {
      f0 += red[3 * (threadIdx.x + l)];
      f1 += red[3 * (threadIdx.x + l) + 1];
      f2 += red[3 * (threadIdx.x + l) + 2];
    }
    out[3 * i] = f0;
    out[3 * i + 1] = f1;
    out[3 * i + 2] = f2;
  }
}

// ---------------------------------------------------------------------------
// Forces, stress and the adjoint pair pass
// ---------------------------------------------------------------------------

// One pass over the rows.  FORCE: F (eV/A) and the real-space stress sums.
// ADJ: h_real (into h), explicit Buckingham and TT gradients (into gacc).
template <typename T, bool ADJ>
__device__ void pair_pass(const FitArgs& a, const FitSnap& S, int ib0, int ib1, const double* mu, const double* gv,
                          const double Gs[9], double* Fout, double* h, double* gacc, double st_nonb[6],
                          double st_dir[6]) {
  const double* sx = a.sx + S.a0;
  const double* sy = a.sy + S.a0;
  const double* sz = a.sz + S.a0;
  const int* it = a.itype + S.a0;
  const double* q = a.q + S.a0;
  const T alpha = (T)S.alpha;
  const double cu = a.coul_unit;
  const double sgs = 1.0 / (S.omega * a.press_unit);
  for (int i = ib0 + threadIdx.x; i < ib1; i += blockDim.x) {
    const int ti = it[i];
    const T qi = (T)q[i];
    const T mix = (T)mu[3 * i], miy = (T)mu[3 * i + 1], miz = (T)mu[3 * i + 2];
    const bool dip_i = a.alpha_t[ti] != 0.0;
    double fsr[3] = {0, 0, 0}, fcl[3] = {0, 0, 0}, hh[3] = {0, 0, 0};
    for (int e = a.row[S.row0 + i]; e < a.row[S.row0 + i + 1]; ++e) {
      const int j = a.nl[S.nl0 + e];
      const int tj = it[j];
      const Geo G = min_image(S, sx, sy, sz, i, j);
      const FitPair& P = a.pairs[ti + a.ntype * tj];
      const bool in_sr = G.r2 < S.rc2_sr && P.A != 0.0;
      const bool in_cl = !(G.r2 > S.rc2_c);
      if (!in_sr && !in_cl) continue;
      const T rx = (T)G.rx, ry = (T)G.ry, rz = (T)G.rz, r2 = (T)G.r2;
      const T d = sqrt(r2);
      const T dm1 = T(1) / d;
      // virtual displacement of the pair for the adjoint: (g_i - g_j) + Gs r / (omega press_unit)
      T dvx = T(0), dvy = T(0), dvz = T(0);
      if (ADJ) {
        const double gx = gv[3 * i] - gv[3 * j], gy = gv[3 * i + 1] - gv[3 * j + 1], gz = gv[3 * i + 2] - gv[3 * j + 2];
        dvx = (T)(gx + sgs * (Gs[0] * G.rx + Gs[3] * G.ry + Gs[6] * G.rz));
        dvy = (T)(gy + sgs * (Gs[1] * G.rx + Gs[4] * G.ry + Gs[7] * G.rz));
        dvz = (T)(gz + sgs * (Gs[2] * G.rx + Gs[5] * G.ry + Gs[8] * G.rz));
      }
      if (in_sr) {
        const T B = (T)P.B, Ab = (T)P.A, Cb = (T)P.C;
        const T erh = Ab * exp(-B * d);
        const T ir6 = Cb * dm1 * dm1 * dm1 * dm1 * dm1 * dm1;
        const T w = (B * erh - T(6) * ir6 * dm1) * dm1;
        if (!ADJ) {
          fsr[0] += (double)(w * rx);
          fsr[1] += (double)(w * ry);
          fsr[2] += (double)(w * rz);
          st_nonb[0] += (double)(w * rx * rx);
          st_nonb[1] += (double)(w * ry * ry);
          st_nonb[2] += (double)(w * rz * rz);
          st_nonb[3] += (double)(w * rx * ry);
          st_nonb[4] += (double)(w * rx * rz);
          st_nonb[5] += (double)(w * ry * rz);
        } else {
          // half (two rows per pair) of dw/dtheta [(g_i - g_j).r + r.Gs.r/(omega pu)]
          const T proj = T(0.5) * (dvx * rx + dvy * ry + dvz * rz);
          const T ex = exp(-B * d);
          if (P.sA >= 0) gacc[P.sA] += (double)(proj * B * ex * dm1);
          if (P.sC >= 0) gacc[P.sC] += (double)(proj * T(-6) * dm1 * dm1 * dm1 * dm1 * dm1 * dm1 * dm1 * dm1);
          if (P.sRho >= 0) {
            const T rho = (T)P.rho;
            gacc[P.sRho] += (double)(proj * (-(Ab * ex * (T(1) - B * d)) * dm1 / (rho * rho)));
          }
        }
      }
      if (in_cl) {
        const Erf<T> E = erf_terms<T>(d, r2, alpha);
        const T qj = (T)q[j];
        const bool dip_j = a.alpha_t[tj] != 0.0;
        const T mjx = (T)mu[3 * j], mjy = (T)mu[3 * j + 1], mjz = (T)mu[3 * j + 2];
        const T F1_dm3 = E.F1 * E.dm3, F2_dm5 = T(3) * E.F2 * E.dm5;
        const T rmi = rx * mix + ry * miy + rz * miz, rmj = rx * mjx + ry * mjy + rz * mjz;
        // damping of the field at i (from q_j) and at j (from q_i)
        const FitPair& Pj = a.pairs[tj + a.ntype * ti];
        T f = T(1), fd = T(0), fb = T(0), fdb = T(0), fc = T(0), fdc = T(0);
        T f2 = T(1), fd2 = T(0), fb2 = T(0), fdb2 = T(0), fc2 = T(0), fdc2 = T(0);
        const bool dmp_i = dip_i && P.ldamp, dmp_j = dip_j && Pj.ldamp;
        if (dmp_i) tt_all<T>((T)P.db, (T)P.dc, d, P.dk, f, fd, fb, fdb, fc, fdc);
        if (dmp_j) tt_all<T>((T)Pj.db, (T)Pj.dc, d, Pj.dk, f2, fd2, fb2, fdb2, fc2, fdc2);
        if (!ADJ) {
          T fij0 = T(0), fij1 = T(0), fij2 = T(0);
          const T qij = qi * qj;
          fij0 -= qij * F1_dm3 * rx;
          fij1 -= qij * F1_dm3 * ry;
          fij2 -= qij * F1_dm3 * rz;
          if (dip_i || dip_j) {
            const T vx = qj * mix - qi * mjx, vy = qj * miy - qi * mjy, vz = qj * miz - qi * mjz;
            const T rv = rx * vx + ry * vy + rz * vz;
            fij0 += F2_dm5 * rx * rv - F1_dm3 * vx;
            fij1 += F2_dm5 * ry * rv - F1_dm3 * vy;
            fij2 += F2_dm5 * rz * rv - F1_dm3 * vz;
            if (dmp_i || dmp_j) {
              const T F1d = T(1) - f, F2d = F1d + (d / T(3)) * fd;
              const T F1d2 = T(1) - f2, F2d2 = F1d2 + (d / T(3)) * fd2;
              const T ci = T(3) * F2d * E.dm5 * rmi, cj = T(3) * F2d2 * E.dm5 * rmj;
              const T di = F1d * E.dm3, dj = F1d2 * E.dm3;
              fij0 += -qj * (ci * rx - di * mix) + qi * (cj * rx - dj * mjx);
              fij1 += -qj * (ci * ry - di * miy) + qi * (cj * ry - dj * mjy);
              fij2 += -qj * (ci * rz - di * miz) + qi * (cj * rz - dj * mjz);
            }
          }
          if (dip_i && dip_j) {
            const T mm = mix * mjx + miy * mjy + miz * mjz;
            const T F3_dm7 = T(15) * E.F3 * E.dm5 * E.dm2;
            const T s3 = -F3_dm7 * rmi * rmj;
            fij0 -= s3 * rx + F2_dm5 * (mix * rmj + mm * rx + rmi * mjx);
            fij1 -= s3 * ry + F2_dm5 * (miy * rmj + mm * ry + rmi * mjy);
            fij2 -= s3 * rz + F2_dm5 * (miz * rmj + mm * rz + rmi * mjz);
          }
          fcl[0] -= (double)fij0;
          fcl[1] -= (double)fij1;
          fcl[2] -= (double)fij2;
          st_dir[0] -= (double)(T(2) * rx * fij0);
          st_dir[1] -= (double)(T(2) * ry * fij1);
          st_dir[2] -= (double)(T(2) * rz * fij2);
          st_dir[3] -= (double)(rx * fij1 + ry * fij0);
          st_dir[4] -= (double)(rx * fij2 + rz * fij0);
          st_dir[5] -= (double)(ry * fij2 + rz * fij1);
        } else {
          // Phi_pair = 1/2 sum_rows dvec . f_ij, dvec = -cu dv: d(dvec.f_ij)/d mu_i into h_i,
          // explicit TT derivatives of the damped q-mu terms (both sides, half per row)
          const T cuT = (T)cu;
          const T ex = -cuT * dvx, ey = -cuT * dvy, ez = -cuT * dvz;   // dvec
          const T dr = ex * rx + ey * ry + ez * rz;
          double h0 = 0, h1 = 0, h2 = 0;
          if (dip_i) {
            // undamped q-mu: f_ij += F2_dm5 r (r.v) - F1_dm3 v, v = qj mu_i - qi mu_j
            h0 += (double)(qj * (F2_dm5 * dr * rx - F1_dm3 * ex));
            h1 += (double)(qj * (F2_dm5 * dr * ry - F1_dm3 * ey));
            h2 += (double)(qj * (F2_dm5 * dr * rz - F1_dm3 * ez));
            if (dmp_i) {
              // f_ij += -qj (3 F2d dm5 rmi r - F1d dm3 mu_i)
              const T F1d = T(1) - f, F2d = F1d + (d / T(3)) * fd;
              h0 -= (double)(qj * (T(3) * F2d * E.dm5 * dr * rx - F1d * E.dm3 * ex));
              h1 -= (double)(qj * (T(3) * F2d * E.dm5 * dr * ry - F1d * E.dm3 * ey));
              h2 -= (double)(qj * (T(3) * F2d * E.dm5 * dr * rz - F1d * E.dm3 * ez));
            }
            if (dip_j) {
              // mu-mu: f_ij -= s3 r + F2_dm5 (mu_i rmj + mm r + rmi mu_j), s3 = -F3_dm7 rmi rmj
              const T F3_dm7 = T(15) * E.F3 * E.dm5 * E.dm2;
              const T edm = ex * mjx + ey * mjy + ez * mjz;
              h0 -= (double)(-F3_dm7 * rmj * dr * rx + F2_dm5 * (rmj * ex + dr * mjx + edm * rx));
              h1 -= (double)(-F3_dm7 * rmj * dr * ry + F2_dm5 * (rmj * ey + dr * mjy + edm * ry));
              h2 -= (double)(-F3_dm7 * rmj * dr * rz + F2_dm5 * (rmj * ez + dr * mjz + edm * rz));
            }
          }
          hh[0] += h0;
          hh[1] += h1;
          hh[2] += h2;
          if (dmp_i && (P.sB >= 0 || P.sC2 >= 0)) {
            // d f_ij / d theta (i side) = -qj (3 dF2d dm5 rmi r - dF1d dm3 mu_i),
            // dF1d = -df, dF2d = -df + d/3 dfd
            const T edmi = ex * mix + ey * miy + ez * miz;
            const T base3 = T(3) * E.dm5 * rmi * dr, base1 = E.dm3 * edmi;
            if (P.sB >= 0) {
              const T dF1 = -fb, dF2 = -fb + (d / T(3)) * fdb;
              gacc[P.sB] += (double)(T(0.5) * (-qj) * (dF2 * base3 - dF1 * base1));
            }
            if (P.sC2 >= 0) {
              const T dF1 = -fc, dF2 = -fc + (d / T(3)) * fdc;
              gacc[P.sC2] += (double)(T(0.5) * (-qj) * (dF2 * base3 - dF1 * base1));
            }
          }
          if (dmp_j && (Pj.sB >= 0 || Pj.sC2 >= 0)) {
            // d f_ij / d theta (j side) = qi (3 dF2d2 dm5 rmj r - dF1d2 dm3 mu_j)
            const T edmj = ex * mjx + ey * mjy + ez * mjz;
            const T base3 = T(3) * E.dm5 * rmj * dr, base1 = E.dm3 * edmj;
            if (Pj.sB >= 0) {
              const T dF1 = -fb2, dF2 = -fb2 + (d / T(3)) * fdb2;
              gacc[Pj.sB] += (double)(T(0.5) * qi * (dF2 * base3 - dF1 * base1));
            }
            if (Pj.sC2 >= 0) {
              const T dF1 = -fc2, dF2 = -fc2 + (d / T(3)) * fdc2;
              gacc[Pj.sC2] += (double)(T(0.5) * qi * (dF2 * base3 - dF1 * base1));
            }
          }
        }
      }
    }
    if (!ADJ) {
      for (int c = 0; c < 3; ++c) Fout[3 * i + c] = fsr[c] + cu * fcl[c];
    } else {
      for (int c = 0; c < 3; ++c) h[3 * i + c] = hh[c];
    }
  }
}

// ---------------------------------------------------------------------------
// Phase kernels.  Atom-block kernels: blockIdx.x -> ablk[b] = (snapshot,
// first atom, end atom, -); k-block kernels: kblk[b] = (snapshot, first k,
// end k, -).  The per-snapshot scratch: seven 3 x n blocks (ws) and ten
// doubles per k (kws), as listed at FIT_WS / FIT_KWS.
// ---------------------------------------------------------------------------
enum { W_ES = 0, W_MU = 1, W_ET = 2, W_G = 3, W_H = 4, W_LAM = 5, W_TMP = 6 };

__device__ __forceinline__ double* ws_of(const FitArgs& a, const FitSnap& S, int slot) {
  return a.ws + (size_t)FIT_WS * S.a0 + (size_t)slot * 3 * S.n;
}

// per-block sums of v[0..n) into the block's partial slots
__device__ __forceinline__ void block_partials(double* v, int n, double* red, double* out) {
  for (int c = 0; c < n; ++c) {
    const double r = block_sum(v[c], red);
    if (threadIdx.x == 0) out[c] = r;
  }
}

// static field of the charges (real space, damped) at the block's atoms
template <typename T>
__global__ void __launch_bounds__(256) kf_static(FitArgs a) {
  const int4 B = a.ablk[blockIdx.x];
  const FitSnap S = a.snaps[B.x];
  field_static<T>(a, S, B.y, B.z, ws_of(a, S, W_ES), nullptr, nullptr);
}

// dst = alpha src at the block's atoms, for snapshots still iterating
__global__ void __launch_bounds__(256) kf_update(FitArgs a) {
  const int4 B = a.ablk[blockIdx.x];
  if (!a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  const int* it = a.itype + S.a0;
  double* dst = ws_of(a, S, a.ws_dst);
  const double* src = ws_of(a, S, a.ws_src);
  for (int j = B.y + threadIdx.x; j < B.z; j += blockDim.x) {
    const double al = a.alpha_t[it[j]];
    for (int c = 0; c < 3; ++c) dst[3 * j + c] = al * src[3 * j + c];
  }
}

// real-space dipole field of v = ws[slot] into ws[W_TMP] (iterating snapshots)
template <typename T>
__global__ void __launch_bounds__(256) kf_dipreal(FitArgs a, int slot) {
  extern __shared__ double red[];
  const int4 B = a.ablk[blockIdx.x];
  if (!a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  field_dip_real<T>(a, S, B.y, B.z, ws_of(a, S, slot), ws_of(a, S, W_TMP), red);
}

// structure factors of the block's k-vectors and their per-k follow-ups.
// mode 0: q -> K_SQ, C1 = A S; 1: i k.v -> K_S, C1 = A S (iterating
// snapshots only); 2: q + i k.mu -> K_S, C1 = A S and the reciprocal stress
// partials; 3: (k.g)(q + i k.mu) -> K_C2 = A S3, C3 = (tr Gs - kcoe k.Gs.k)
// A S.  v / g: ws slots (-1: none).
// one block of k-vectors over one chunk of FIT_AC atoms: the chunk's
// structure-factor partial sums (sblk[b] = (snapshot, first k, first atom,
// chunk)); a large snapshot's sums are split over several workgroups and
// added in chunk order by kf_sf
template <typename T>
__global__ void __launch_bounds__(256) kf_sfpart(FitArgs a, int mode, int vslot, int gslot) {
  extern __shared__ double red[];                 // blockDim doubles, then staging
  c2<T>* lt = reinterpret_cast<c2<T>*>(red + blockDim.x);            // FIT_CA x ncol_max
  double* la = reinterpret_cast<double*>(lt + (size_t)FIT_CA * a.ncol_max);   // FIT_CA x 7
  const int4 B = a.sblk[blockIdx.x];
  if (mode == 1 && !a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  const c2<T>* tab = reinterpret_cast<const c2<T>*>(a.tab) + S.tab0;
  const double* v = vslot >= 0 ? ws_of(a, S, vslot) : nullptr;
  const double* g = gslot >= 0 ? ws_of(a, S, gslot) : nullptr;
  double* out = a.sp + 2 * ((size_t)a.snap_s0[B.x] + (size_t)B.w * S.nk);
  struct_factor<T>(a, S, B.y, min(S.nk, B.y + (int)blockDim.x), B.z, min(S.n, B.z + FIT_AC), tab, mode, v, g, out, 2,
                   lt, la);
}

// the block's structure factors (chunk partials in chunk order) and their
// per-k follow-ups
template <typename T>
__global__ void __launch_bounds__(256) kf_sf(FitArgs a, int mode) {
  extern __shared__ double red[];
  const int4 B = a.kblk[blockIdx.x];
  if (mode == 1 && !a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  const int slot = mode == 0 ? K_SQ : (mode == 3 ? K_C2 : K_S);
  const int nch = a.snap_sch[B.x];
  for (int k = B.y + threadIdx.x; k < B.z; k += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
      const double* p = a.sp + 2 * ((size_t)a.snap_s0[B.x] + (size_t)ch * S.nk + k);
      sr += p[0];
      si += p[1];
    }
    double* K = a.kws + (size_t)FIT_KWS * (S.k0 + k);
    K[slot] = sr;
    K[slot + 1] = si;
  }
  double strec[6] = {0, 0, 0, 0, 0, 0};
  const double* Gs = a.Gs + 9 * (size_t)B.x;
  for (int k = B.y + threadIdx.x; k < B.z; k += blockDim.x) {
    const int kg = S.k0 + k;
    double* K = a.kws + (size_t)FIT_KWS * kg;
    const double kx = a.kv[5 * kg], ky = a.kv[5 * kg + 1], kz = a.kv[5 * kg + 2], Ak = a.kv[5 * kg + 3],
                 kc = a.kv[5 * kg + 4];
    if (mode == 0) {
      K[K_C1] = Ak * K[K_SQ];
      K[K_C1 + 1] = Ak * K[K_SQ + 1];
    } else if (mode == 1 || mode == 2) {
      K[K_C1] = Ak * K[K_S];
      K[K_C1 + 1] = Ak * K[K_S + 1];
      if (mode == 2) {
        const double str = (K[K_S] * K[K_S] + K[K_S + 1] * K[K_S + 1]) * Ak;
        strec[0] += (1.0 - kc * kx * kx) * str;
        strec[1] += (1.0 - kc * ky * ky) * str;
        strec[2] += (1.0 - kc * kz * kz) * str;
        strec[3] -= kc * kx * ky * str;
        strec[4] -= kc * kx * kz * str;
        strec[5] -= kc * ky * kz * str;
      }
    } else {
      // K_C1 already holds A S (forces)
      K[K_C2] *= Ak;
      K[K_C2 + 1] *= Ak;
      const double kGk = kx * (Gs[0] * kx + Gs[3] * ky + Gs[6] * kz) + ky * (Gs[1] * kx + Gs[4] * ky + Gs[7] * kz) +
                         kz * (Gs[2] * kx + Gs[5] * ky + Gs[8] * kz);
      const double wk = (Gs[0] + Gs[4] + Gs[8] - kc * kGk) * Ak;
      K[K_C3] = wk * K[K_S];
      K[K_C3 + 1] = wk * K[K_S + 1];
    }
  }
  if (mode == 2) block_partials(strec, 6, red, a.pk + (size_t)FIT_PK * blockIdx.x);
}

// gather sums of one chunk of k-vectors at the block's atoms into the
// (atom, chunk) records: the sums of a large snapshot are split over several
// workgroups, and the records are added in chunk order by kf_gather.
// PH_STATIC / PH_ITER: im3 of C1 at polarisable atoms; PH_FORCE: im3 and
// re6 of C1 at every atom; PH_H: im3 of C3, re3 of C2, re6 of C1 at
// polarisable atoms.
template <typename T, int PH>
__global__ void __launch_bounds__(256) kf_gpart(FitArgs a) {
  extern __shared__ double red[];
  const int4 B = a.gblk[blockIdx.x];
  if (PH == 1 && !a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  const c2<T>* tab = reinterpret_cast<const c2<T>*>(a.tab) + S.tab0;
  const int* it = a.itype + S.a0;
  const int k1 = min(S.nk, B.z + FIT_KC);
  // PH 0, 1: im3 of C1; 2: + re6 of C1; 3: im3 of C3, re3 of C2, re6 of C1
  const int c1 = PH == 3 ? K_C3 : K_C1, c2i = PH == 3 ? K_C2 : -1, c3 = PH >= 2 ? K_C1 : -1;
  const GatherStage g = gather_stage_of(red);
  gather_stage(a, S, B.z, k1, c1, c2i, c3, g);
  const int j = B.y + threadIdx.x;
  if (j >= min(S.n, B.y + (int)blockDim.x)) return;
  if (PH != 2 && a.alpha_t[it[j]] == 0.0) return;
  double im3[3], re3[3], re6[6];
  gather<T>(S, tab, j, k1 - B.z, true, c2i >= 0, c3 >= 0, g, im3, re3, re6);
  double* r = a.gp + ((size_t)a.snap_g0[B.x] + (size_t)B.w * S.n + j) * FIT_GC;
  for (int c = 0; c < 3; ++c) r[c] = im3[c];
  if (PH == 3)
    for (int c = 0; c < 3; ++c) r[3 + c] = re3[c];
  if (PH >= 2)
    for (int c = 0; c < 6; ++c) r[6 + c] = re6[c];
}

// the gather sums of atom j: its chunk records added in chunk order
__device__ __forceinline__ void gather_sum(const FitArgs& a, int s, const FitSnap& S, int j, int ncol,
                                           double im3[3], double re3[3], double re6[6]) {
  for (int c = 0; c < 3; ++c) im3[c] = re3[c] = 0.0;
  for (int c = 0; c < 6; ++c) re6[c] = 0.0;
  const int nch = a.snap_nch[s];
  for (int ch = 0; ch < nch; ++ch) {
    const double* r = a.gp + ((size_t)a.snap_g0[s] + (size_t)ch * S.n + j) * FIT_GC;
    for (int c = 0; c < 3; ++c) im3[c] += r[c];
    if (ncol == 12)
      for (int c = 0; c < 3; ++c) re3[c] += r[3 + c];
    if (ncol >= 9)
      for (int c = 0; c < 6; ++c) re6[c] += r[6 + c];
  }
}

// gathers at the block's atoms and their per-atom follow-ups.
//   PH_STATIC: E_s += reciprocal field of the charges; E_t = E_s, mu = 0
//   PH_ITER:   tmp += reciprocal + self field of v = ws[vslot]; E_t = base +
//              tmp; residual |v / alpha - E_t|^2 partial (iterating only);
//              base = ws[bslot] (E_s for the SCF, h for the adjoint)
//   PH_FORCE:  F += reciprocal force; g = 2 (F - F*), loss partial
//   PH_H:      h += reciprocal h; E_t = h; |h|^2 partial
enum { PH_STATIC = 0, PH_ITER = 1, PH_FORCE = 2, PH_H = 3 };
template <typename T, int PH>
__global__ void __launch_bounds__(256) kf_gather(FitArgs a, int vslot, int bslot) {
  extern __shared__ double red[];
  const int4 B = a.ablk[blockIdx.x];
  if (PH == PH_ITER && !a.active[B.x]) return;
  const FitSnap S = a.snaps[B.x];
  const c2<T>* tab = reinterpret_cast<const c2<T>*>(a.tab) + S.tab0;
  const int* it = a.itype + S.a0;
  const double* q = a.q + S.a0;
  const double pre = -8.0 * MDFF_PI / S.omega;
  double part = 0.0;
  for (int j = B.y + threadIdx.x; j < B.z; j += blockDim.x) {
    const double al = a.alpha_t[it[j]];
    double im3[3], re3[3], re6[6];
    if (PH == PH_STATIC) {
      double* vES = ws_of(a, S, W_ES);
      if (al != 0.0) {
        gather_sum(a, B.x, S, j, 3, im3, re3, re6);
        for (int c = 0; c < 3; ++c) vES[3 * j + c] += pre * im3[c];
      }
      double* vET = ws_of(a, S, W_ET);
      double* vMU = ws_of(a, S, W_MU);
      for (int c = 0; c < 3; ++c) {
        vET[3 * j + c] = vES[3 * j + c];
        vMU[3 * j + c] = 0.0;
      }
    } else if (PH == PH_ITER) {
      const double* v = ws_of(a, S, vslot);
      const double* base = ws_of(a, S, bslot);
      double* tmp = ws_of(a, S, W_TMP);
      double* vET = ws_of(a, S, W_ET);
      const double self = 4.0 * S.alpha * S.alpha * S.alpha / (3.0 * sqrt(MDFF_PI));
      if (al == 0.0) {
        tmp[3 * j] = tmp[3 * j + 1] = tmp[3 * j + 2] = 0.0;
      } else {
        gather_sum(a, B.x, S, j, 3, im3, re3, re6);
        for (int c = 0; c < 3; ++c) tmp[3 * j + c] += pre * im3[c] + self * v[3 * j + c];
      }
      for (int c = 0; c < 3; ++c) {
        const double et = base[3 * j + c] + tmp[3 * j + c];
        vET[3 * j + c] = et;
        if (al != 0.0) {
          const double rr = v[3 * j + c] / al - et;
          part += rr * rr;
        }
      }
    } else if (PH == PH_FORCE) {
      const double* vMU = ws_of(a, S, W_MU);
      double* vG = ws_of(a, S, W_G);
      double* Fs = a.F + 3 * (size_t)S.a0;
      gather_sum(a, B.x, S, j, 9, im3, re3, re6);
      const double m0 = vMU[3 * j], m1 = vMU[3 * j + 1], m2 = vMU[3 * j + 2];
      // F_rec = -(8 pi / V) [q im3 - re6 mu]
      const double r0 = re6[0] * m0 + re6[3] * m1 + re6[4] * m2;
      const double r1 = re6[3] * m0 + re6[1] * m1 + re6[5] * m2;
      const double r2 = re6[4] * m0 + re6[5] * m1 + re6[2] * m2;
      Fs[3 * j] += a.coul_unit * pre * (q[j] * im3[0] - r0);
      Fs[3 * j + 1] += a.coul_unit * pre * (q[j] * im3[1] - r1);
      Fs[3 * j + 2] += a.coul_unit * pre * (q[j] * im3[2] - r2);
      for (int c = 0; c < 3; ++c) {
        const double dF = Fs[3 * j + c] - a.fref[3 * ((size_t)S.a0 + j) + c];
        part += dF * dF;
        vG[3 * j + c] = 2.0 * dF;
      }
    } else {
      //   h_j += -cu (8pi/V) [re3(A P) - re6(A S) g_j] + (cu/pu)(8pi/V^2) im3(w S)
      double* vH = ws_of(a, S, W_H);
      const double* vG = ws_of(a, S, W_G);
      double* vET = ws_of(a, S, W_ET);
      if (al == 0.0) {
        vH[3 * j] = vH[3 * j + 1] = vH[3 * j + 2] = 0.0;
      } else {
        const double cu = a.coul_unit;
        const double pv2 = cu / a.press_unit * 8.0 * MDFF_PI / (S.omega * S.omega);
        gather_sum(a, B.x, S, j, 12, im3, re3, re6);
        const double g0 = vG[3 * j], g1 = vG[3 * j + 1], g2 = vG[3 * j + 2];
        const double e0 = re6[0] * g0 + re6[3] * g1 + re6[4] * g2;
        const double e1 = re6[3] * g0 + re6[1] * g1 + re6[5] * g2;
        const double e2 = re6[4] * g0 + re6[5] * g1 + re6[2] * g2;
        vH[3 * j] += cu * pre * (re3[0] - e0) + pv2 * im3[0];
        vH[3 * j + 1] += cu * pre * (re3[1] - e1) + pv2 * im3[1];
        vH[3 * j + 2] += cu * pre * (re3[2] - e2) + pv2 * im3[2];
      }
      for (int c = 0; c < 3; ++c) {
        vET[3 * j + c] = vH[3 * j + c];
        part += vH[3 * j + c] * vH[3 * j + c];
      }
    }
  }
  // slot 12 for the force loss (the pair pass holds 0..11), else slot 0
  if (PH != PH_STATIC) block_partials(&part, 1, red, a.pa + (size_t)FIT_PA * blockIdx.x + (PH == PH_FORCE ? 12 : 0));
}

// pair pass at the block's atoms.  !ADJ: forces (ws-independent) and the
// real-space stress partials (slots 0..11); ADJ: h (real space, overwritten)
// and the explicit gradient partials (slots 0..FIT_MAXP-1)
template <typename T, bool ADJ>
__global__ void __launch_bounds__(256) kf_pair(FitArgs a) {
  extern __shared__ double red[];
  const int4 B = a.ablk[blockIdx.x];
  const FitSnap S = a.snaps[B.x];
  double stn[6] = {0, 0, 0, 0, 0, 0}, std_[6] = {0, 0, 0, 0, 0, 0};
  double* out = a.pa + (size_t)FIT_PA * blockIdx.x;
  if (!ADJ) {
    const double Gdummy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    pair_pass<T, false>(a, S, B.y, B.z, ws_of(a, S, W_MU), nullptr, Gdummy, a.F + 3 * (size_t)S.a0, nullptr,
                        nullptr, stn, std_);
    double v[12];
    for (int c = 0; c < 6; ++c) {
      v[c] = stn[c];
      v[6 + c] = std_[c];
    }
    block_partials(v, 12, red, out);
  } else {
    double gacc[FIT_MAXP];
    for (int p = 0; p < FIT_MAXP; ++p) gacc[p] = 0.0;
    double Gs[9];
    for (int c = 0; c < 9; ++c) Gs[c] = a.Gs[9 * (size_t)B.x + c];
    // the non-bonded stress carries 1/press_unit, the Coulomb one coul_unit/press_unit:
    // the pair pass scales its virtual displacement by 1/(omega pu) and the
    // Coulomb terms by coul_unit itself
    pair_pass<T, true>(a, S, B.y, B.z, ws_of(a, S, W_MU), ws_of(a, S, W_G), Gs, nullptr, ws_of(a, S, W_H), gacc,
                       stn, std_);
    block_partials(gacc, FIT_MAXP, red, out);
  }
}

// - lambda . dR/dtheta at the block's atoms: alpha (lambda.mu / alpha^2) and
// TT b, c (lambda . dE_s/dtheta); gradient partials (slots 0..FIT_MAXP-1)
template <typename T>
__global__ void __launch_bounds__(256) kf_lambda(FitArgs a) {
  extern __shared__ double red[];
  const int4 B = a.ablk[blockIdx.x];
  const FitSnap S = a.snaps[B.x];
  const int* it = a.itype + S.a0;
  const double* vLAM = ws_of(a, S, W_LAM);
  const double* vMU = ws_of(a, S, W_MU);
  double gacc[FIT_MAXP];
  for (int p = 0; p < FIT_MAXP; ++p) gacc[p] = 0.0;
  for (int j = B.y + threadIdx.x; j < B.z; j += blockDim.x) {
    const int tj = it[j];
    const double al = a.alpha_t[tj];
    if (al == 0.0 || a.s_alpha[tj] < 0) continue;
    const double lm = vLAM[3 * j] * vMU[3 * j] + vLAM[3 * j + 1] * vMU[3 * j + 1] + vLAM[3 * j + 2] * vMU[3 * j + 2];
    gacc[a.s_alpha[tj]] += lm / (al * al);
  }
  field_static<T>(a, S, B.y, B.z, nullptr, vLAM, gacc);
  block_partials(gacc, FIT_MAXP, red, a.pa + (size_t)FIT_PA * blockIdx.x);
}

// Control of every snapshot (one block, thread-strided over snapshots; each
// snapshot's block partials added in block order):
//   C_START:  iterations to 0, every snapshot iterating
//   C_SCF:    SCF residual of the iteration; rmsd, iteration count, continue?
//             (src/moment_from_pola.f90:365-439)
//   C_FORCE:  stress (GPa), loss, Gs = sym(2 w (tau - tau*)); SCF count out
//   C_PAIRG:  keep the pair pass's gradient partial sums
//   C_HN:     RMS of h, adjoint iterations to 0, every snapshot iterating
//   C_ADJ:    adjoint residual; continue while above tol_adj |h|
//   C_GRAD:   gradient = pair-pass sums + lambda-pass sums; adjoint count out
// count = snapshots still iterating after C_START / C_SCF / C_HN / C_ADJ.
enum { C_START = 0, C_SCF = 1, C_FORCE = 2, C_PAIRG = 3, C_HN = 4, C_ADJ = 5, C_GRAD = 6 };
__global__ void __launch_bounds__(64) kf_ctl(FitArgs a, int what) {
  __shared__ int cnt[64];
  int mine = 0;
  for (int s = threadIdx.x; s < a.nsnap; s += blockDim.x) {
    const FitSnap S = a.snaps[s];
    const int b0 = a.snap_ab0[s], b1 = a.snap_ab0[s + 1];
    auto asum = [&](int c) {
      double t = 0.0;
      for (int b = b0; b < b1; ++b) t += a.pa[(size_t)FIT_PA * b + c];
      return t;
    };
    if (what == C_START || what == C_HN) {
      if (what == C_HN) a.hn[s] = sqrt(asum(0) / (double)max(S.npol, 1));
      a.iter[s] = 0;
      a.active[s] = 1;
    } else if (what == C_SCF || what == C_ADJ) {
      if (a.active[s]) {
        const double rms = sqrt(asum(0) / (double)max(S.npol, 1));
        const int n = ++a.iter[s];
        const double tol = what == C_SCF ? a.tol : a.tol_adj * a.hn[s];
        a.active[s] = (S.npol > 0 && ((n < a.max_iter && rms > tol) || n < a.min_iter)) ? 1 : 0;
      }
    } else if (what == C_FORCE) {
      double sn[6], sd[6], sr[6];
      for (int c = 0; c < 6; ++c) {
        sn[c] = asum(c);
        sd[c] = asum(6 + c);
        double t = 0.0;
        for (int b = a.snap_kb0[s]; b < a.snap_kb0[s + 1]; ++b) t += a.pk[(size_t)FIT_PK * b + c];
        sr[c] = t;
      }
      // the force loss: the gather's partials (slot 12)
      double l = 0.0;
      for (int b = b0; b < b1; ++b) l += a.pa[(size_t)FIT_PA * b + 12];
      const double om = S.omega;
      const double fn = 0.5 / om / a.press_unit;
      const double fc = a.coul_unit / a.press_unit;
      const double tpi_V = 2.0 * MDFF_PI / om;
      const int ix[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
      double tau[9], G[9];
      double lt = 0.0;
      for (int c = 0; c < 9; ++c) {
        const double nb = fn * sn[ix[c]];
        const double dir = 0.25 * sd[ix[c]] / om;
        const double rec = 2.0 * sr[ix[c]] * tpi_V / om;
        tau[c] = nb + (dir + rec) * fc;
        const double dt = tau[c] - a.tauref[9 * s + c];
        lt += dt * dt;
        G[c] = 2.0 * a.w_stress * dt;
      }
      l += a.w_stress * lt;
      // d(G:tau)/d tau_coul-like symmetric contraction uses Gs = (G + G^T)/2
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a.Gs[9 * (size_t)s + r + 3 * c] = 0.5 * (G[r + 3 * c] + G[c + 3 * r]);
      a.loss[s] = l;
      for (int c = 0; c < 9; ++c) a.tau[9 * s + c] = tau[c];
      a.nscf[2 * s] = a.iter[s];
    } else if (what == C_PAIRG) {
      for (int p = 0; p < FIT_MAXP; ++p) a.gsave[(size_t)FIT_MAXP * s + p] = asum(p);
    } else if (what == C_GRAD) {
      for (int p = 0; p < FIT_MAXP; ++p) a.grad[(size_t)FIT_MAXP * s + p] = a.gsave[(size_t)FIT_MAXP * s + p] + asum(p);
      a.nscf[2 * s + 1] = a.iter[s];
    }
    mine += a.active[s];
  }
  cnt[threadIdx.x] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int x = 0; x < (int)blockDim.x; ++x) t += cnt[x];
    a.count[0] = t;
  }
}

// phase tables of all snapshots: one workgroup per snapshot
template <typename T>
__global__ void k_fit_tables(int nsnap, const FitSnap* snaps, const double* sx, const double* sy, const double* sz,
                             void* tab_all) {
  for (int s = blockIdx.x; s < nsnap; s += gridDim.x) {
    const FitSnap S = snaps[s];
    c2<T>* tab = reinterpret_cast<c2<T>*>(tab_all) + S.tab0;
    for (int x = threadIdx.x; x < S.ncol * S.n; x += blockDim.x) {
      const int col = x / S.n, j = x % S.n;
      int m;
      double sv;
      if (col <= S.Kx) {
        m = col;
        sv = sx[S.a0 + j];
      } else if (col <= S.Kx + S.Ky + 1) {
        m = col - S.Kx - 1;
        sv = sy[S.a0 + j];
      } else {
        m = col - S.Kx - S.Ky - 2;
        sv = sz[S.a0 + j];
      }
      // angle 2 pi m s reduced exactly: m s in FP64, then sincospi
      double sn, cs;
      sincospi(2.0 * ((double)m * sv), &sn, &cs);
      tab[x] = c2<T>{(T)cs, (T)sn};
    }
  }
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================

struct mdff_fit {
  std::string err;
  int nsnap = 0, ntype = 0, natm = 0, nk = 0, nparam = 0;
  std::vector<int> spec;          // 3 x nparam
  std::vector<FitSnap> snaps;
  std::vector<FitPair> pairs0;    // base pair table (desc values)
  double alpha0[FIT_NT] = {0, 0, 0, 0};
  mdff_desc base;
  int block = 256, ncol_max = 0;
  hipStream_t stream = nullptr;
  // device
  FitSnap* d_snaps = nullptr;
  // block maps for the current block size
  int nab = 0, nkb = 0;
  int4 *d_ablk = nullptr, *d_kblk = nullptr;
  int *d_ab0 = nullptr, *d_kb0 = nullptr, *d_active = nullptr, *d_iter = nullptr, *d_count = nullptr;
  int ngb = 0;
  int4* d_gblk = nullptr;
  int nsb = 0;
  int4* d_sblk = nullptr;
  int *d_s0 = nullptr, *d_sch = nullptr;
  double* d_sp = nullptr;
  int *d_g0 = nullptr, *d_nch = nullptr;
  double* d_gp = nullptr;
  double *d_pa = nullptr, *d_pk = nullptr, *d_hn = nullptr, *d_Gs = nullptr, *d_gsave = nullptr;
  FitPair* d_pairs = nullptr;
  int *d_itype = nullptr, *d_row = nullptr, *d_nl = nullptr, *d_kn = nullptr, *d_nscf = nullptr;
  double *d_q = nullptr, *d_sx = nullptr, *d_sy = nullptr, *d_sz = nullptr, *d_kv = nullptr;
  void* d_tab32 = nullptr;
  void* d_tab64 = nullptr;
  double *d_fref = nullptr, *d_tauref = nullptr, *d_ws = nullptr, *d_kws = nullptr, *d_F = nullptr,
         *d_tau = nullptr, *d_loss = nullptr, *d_grad = nullptr;
  double w_stress = 0.0;
  long long tab_elems = 0;
};

#define FIT_TRY(f, expr)                                                       \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      (f)->err = std::string("HIP: ") + hipGetErrorString(_e) + " at " #expr; \
      return MDFF_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define FIT_RC(call)                   \
  do {                                 \
    const int _rc = (call);            \
    if (_rc != MDFF_OK) return _rc;    \
  } while (0)

static void fit_lattice(const double* A, double* B, double* omega, double* W) {
  // B = A^{-T}, column-major, omega = det A, widths omega / |a_j x a_k|
  const double a11 = A[0], a21 = A[1], a31 = A[2], a12 = A[3], a22 = A[4], a32 = A[5], a13 = A[6], a23 = A[7],
               a33 = A[8];
  const double det = a11 * (a22 * a33 - a32 * a23) - a12 * (a21 * a33 - a31 * a23) + a13 * (a21 * a32 - a31 * a22);
  double inv[9];
  inv[0] = (a22 * a33 - a23 * a32) / det;
  inv[3] = (a13 * a32 - a12 * a33) / det;
  inv[6] = (a12 * a23 - a13 * a22) / det;
  inv[1] = (a23 * a31 - a21 * a33) / det;
  inv[4] = (a11 * a33 - a13 * a31) / det;
  inv[7] = (a13 * a21 - a11 * a23) / det;
  inv[2] = (a21 * a32 - a22 * a31) / det;
  inv[5] = (a12 * a31 - a11 * a32) / det;
  inv[8] = (a11 * a22 - a12 * a21) / det;
  // B(i,j) = inv(j,i)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i + 3 * j] = inv[j + 3 * i];
  *omega = fabs(det);
  for (int k = 0; k < 3; ++k) {
    const int b = (k + 1) % 3, c = (k + 2) % 3;
    const double cx = A[1 + 3 * b] * A[2 + 3 * c] - A[2 + 3 * b] * A[1 + 3 * c];
    const double cy = A[2 + 3 * b] * A[0 + 3 * c] - A[0 + 3 * b] * A[2 + 3 * c];
    const double cz = A[0 + 3 * b] * A[1 + 3 * c] - A[1 + 3 * b] * A[0 + 3 * c];
    W[k] = *omega / sqrt(cx * cx + cy * cy + cz * cz);
  }
}

static int fit_fail(mdff_fit* f, int code, const std::string& m) {
  f->err = m;
  return code;
}

// Block maps for blocks of `block` threads: every snapshot's atoms in blocks
// of `block` atoms and its k-vectors in blocks of `block` k-vectors, in
// snapshot order; per snapshot the first block of each kind.
static int fit_blocks(mdff_fit* f, int block) {
  std::vector<int4> ab, kb, gb, sb;
  std::vector<int> ab0(1, 0), kb0(1, 0), g0, nch, s0, sch;
  size_t nrec = 0, nsrec = 0;
  for (int s = 0; s < f->nsnap; ++s) {
    const FitSnap& S = f->snaps[s];
    for (int i = 0; i < S.n; i += block) ab.push_back(make_int4(s, i, std::min(S.n, i + block), 0));
    for (int k = 0; k < S.nk; k += block) kb.push_back(make_int4(s, k, std::min(S.nk, k + block), 0));
    ab0.push_back((int)ab.size());
    kb0.push_back((int)kb.size());
    // gather blocks: atom blocks x chunks of FIT_KC k-vectors
    const int nc = std::max(1, (S.nk + FIT_KC - 1) / FIT_KC);
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < S.n; i += block) gb.push_back(make_int4(s, i, c * FIT_KC, c));
    g0.push_back((int)nrec);
    nch.push_back(nc);
    nrec += (size_t)nc * S.n;
    // structure-factor blocks: k blocks x chunks of FIT_AC atoms
    const int na = std::max(1, (S.n + FIT_AC - 1) / FIT_AC);
    for (int c = 0; c < na; ++c)
      for (int k = 0; k < S.nk; k += block) sb.push_back(make_int4(s, k, c * FIT_AC, c));
    s0.push_back((int)nsrec);
    sch.push_back(na);
    nsrec += (size_t)na * S.nk;
  }
  if (sb.empty()) sb.push_back(make_int4(0, 0, 0, 0));
  if (kb.empty()) kb.push_back(make_int4(0, 0, 0, 0));
  void* old[] = {f->d_ablk, f->d_kblk, f->d_ab0, f->d_kb0, f->d_pa, f->d_pk, f->d_gblk, f->d_g0, f->d_nch, f->d_gp,
                 f->d_sblk, f->d_s0, f->d_sch, f->d_sp};
  for (void* p : old)
    if (p) (void)hipFree(p);
  f->nab = (int)ab.size();
  f->nkb = (int)kb.size();
  f->block = block;
  FIT_TRY(f, hipMalloc((void**)&f->d_ablk, sizeof(int4) * ab.size()));
  FIT_TRY(f, hipMemcpy(f->d_ablk, ab.data(), sizeof(int4) * ab.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_kblk, sizeof(int4) * kb.size()));
  FIT_TRY(f, hipMemcpy(f->d_kblk, kb.data(), sizeof(int4) * kb.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_ab0, sizeof(int) * ab0.size()));
  FIT_TRY(f, hipMemcpy(f->d_ab0, ab0.data(), sizeof(int) * ab0.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_kb0, sizeof(int) * kb0.size()));
  FIT_TRY(f, hipMemcpy(f->d_kb0, kb0.data(), sizeof(int) * kb0.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_pa, sizeof(double) * FIT_PA * ab.size()));
  FIT_TRY(f, hipMalloc((void**)&f->d_pk, sizeof(double) * FIT_PK * kb.size()));
  f->ngb = (int)gb.size();
  FIT_TRY(f, hipMalloc((void**)&f->d_gblk, sizeof(int4) * gb.size()));
  FIT_TRY(f, hipMemcpy(f->d_gblk, gb.data(), sizeof(int4) * gb.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_g0, sizeof(int) * g0.size()));
  FIT_TRY(f, hipMemcpy(f->d_g0, g0.data(), sizeof(int) * g0.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_nch, sizeof(int) * nch.size()));
  FIT_TRY(f, hipMemcpy(f->d_nch, nch.data(), sizeof(int) * nch.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_gp, sizeof(double) * FIT_GC * std::max<size_t>(nrec, 1)));
  f->nsb = (int)sb.size();
  FIT_TRY(f, hipMalloc((void**)&f->d_sblk, sizeof(int4) * sb.size()));
  FIT_TRY(f, hipMemcpy(f->d_sblk, sb.data(), sizeof(int4) * sb.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_s0, sizeof(int) * s0.size()));
  FIT_TRY(f, hipMemcpy(f->d_s0, s0.data(), sizeof(int) * s0.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_sch, sizeof(int) * sch.size()));
  FIT_TRY(f, hipMemcpy(f->d_sch, sch.data(), sizeof(int) * sch.size(), hipMemcpyHostToDevice));
  FIT_TRY(f, hipMalloc((void**)&f->d_sp, sizeof(double) * 2 * std::max<size_t>(nsrec, 1)));
  return MDFF_OK;
}

// One forward (+ adjoint gradient) evaluation: the phase kernels in order;
// the SCF and the adjoint solve repeat their iteration's launches while the
// control kernel counts a snapshot still iterating (read back per iteration).
template <typename T>
static int fit_run(mdff_fit* f, FitArgs a, size_t tsz) {
  hipStream_t st = f->stream;
  const dim3 ga(f->nab), gk(f->nkb), gg(f->ngb), gs(f->nsb), b(f->block);
  const size_t red = sizeof(double) * f->block;
  const size_t shm_sf = red + tsz * FIT_CA * f->ncol_max + sizeof(double) * 7 * FIT_CA;
  int cnt = 0;
  auto active_count = [&]() -> int {
    FIT_TRY(f, hipMemcpyAsync(&cnt, f->d_count, sizeof(int), hipMemcpyDeviceToHost, st));
    FIT_TRY(f, hipStreamSynchronize(st));
    return MDFF_OK;
  };
  auto iterate = [&](int vslot, int bslot, int what) -> int {
    for (int it = 0; it < a.max_iter; ++it) {
      a.ws_dst = vslot;
      a.ws_src = W_ET;
      hipLaunchKernelGGL(kf_update, ga, b, 0, st, a);
      hipLaunchKernelGGL((kf_dipreal<T>), dim3(f->nab, kDipLanes), b, 3 * red, st, a, vslot);
      hipLaunchKernelGGL((kf_sfpart<T>), gs, b, shm_sf, st, a, 1, vslot, -1);
      hipLaunchKernelGGL((kf_sf<T>), gk, b, red, st, a, 1);
      hipLaunchKernelGGL((kf_gpart<T, PH_ITER>), gg, b, kGatherLds, st, a);
      hipLaunchKernelGGL((kf_gather<T, PH_ITER>), ga, b, red, st, a, vslot, bslot);
      hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, what);
      FIT_TRY(f, hipGetLastError());
      const int rc = active_count();
      if (rc != MDFF_OK) return rc;
      if (cnt == 0) break;
    }
    return MDFF_OK;
  };
  // static field, SCF
  hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, (int)C_START);
  hipLaunchKernelGGL((kf_static<T>), ga, b, 0, st, a);
  hipLaunchKernelGGL((kf_sfpart<T>), gs, b, shm_sf, st, a, 0, -1, -1);
  hipLaunchKernelGGL((kf_sf<T>), gk, b, red, st, a, 0);
  hipLaunchKernelGGL((kf_gpart<T, PH_STATIC>), gg, b, kGatherLds, st, a);
  hipLaunchKernelGGL((kf_gather<T, PH_STATIC>), ga, b, red, st, a, -1, -1);
  FIT_TRY(f, hipGetLastError());
  FIT_RC(iterate(W_MU, W_ES, C_SCF));
  // forces, stress, loss
  hipLaunchKernelGGL((kf_pair<T, false>), ga, b, red, st, a);
  hipLaunchKernelGGL((kf_sfpart<T>), gs, b, shm_sf, st, a, 2, (int)W_MU, -1);
  hipLaunchKernelGGL((kf_sf<T>), gk, b, red, st, a, 2);
  hipLaunchKernelGGL((kf_gpart<T, PH_FORCE>), gg, b, kGatherLds, st, a);
  hipLaunchKernelGGL((kf_gather<T, PH_FORCE>), ga, b, red, st, a, -1, -1);
  hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, (int)C_FORCE);
  FIT_TRY(f, hipGetLastError());
  if (!a.want_grad) return MDFF_OK;
  // explicit pair gradients and h; adjoint solve; lambda terms
  hipLaunchKernelGGL((kf_pair<T, true>), ga, b, red, st, a);
  hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, (int)C_PAIRG);
  hipLaunchKernelGGL((kf_sfpart<T>), gs, b, shm_sf, st, a, 3, (int)W_MU, (int)W_G);
  hipLaunchKernelGGL((kf_sf<T>), gk, b, red, st, a, 3);
  hipLaunchKernelGGL((kf_gpart<T, PH_H>), gg, b, kGatherLds, st, a);
  hipLaunchKernelGGL((kf_gather<T, PH_H>), ga, b, red, st, a, -1, -1);
  hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, (int)C_HN);
  FIT_TRY(f, hipGetLastError());
  FIT_RC(iterate(W_LAM, W_H, C_ADJ));
  hipLaunchKernelGGL((kf_lambda<T>), ga, b, red, st, a);
  hipLaunchKernelGGL(kf_ctl, dim3(1), dim3(64), 0, st, a, (int)C_GRAD);
  FIT_TRY(f, hipGetLastError());
  return MDFF_OK;
}

extern "C" {

const char* mdff_fit_last_error(const mdff_fit* f) { return f ? f->err.c_str() : "null fit context"; }

int mdff_fit_create(const mdff_desc* d, int32_t nsnap, const int32_t* natm, const double* cells,
                    const double* cut_alpha, const int32_t* kmax, const int32_t* itype, const double* pos,
                    int32_t nparam, const int32_t* spec, mdff_fit** out) {
  if (!out || !d || !natm || !cells || !cut_alpha || !kmax || !itype || !pos || (nparam > 0 && !spec))
    return MDFF_ERR_BAD_ARG;
  *out = nullptr;
  if (nsnap <= 0 || nparam < 0 || nparam > FIT_MAXP || d->ntype < 1 || d->ntype > FIT_NT) return MDFF_ERR_BAD_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MDFF_ERR_NO_DEVICE;
  mdff_fit* f = new mdff_fit();
  *out = f;
  f->base = *d;
  f->nsnap = nsnap;
  f->ntype = d->ntype;
  f->nparam = nparam;
  f->spec.assign(spec, spec + 3 * nparam);
  if (!(d->lbmhft && !d->lbmhftd && !d->lnmlj && !d->lmorse && d->lcoulomb))
    return fit_fail(f, MDFF_ERR_UNSUPPORTED, "the fit supports Buckingham (lbmhft) + Coulomb only");
  for (int i = 0; i < 48; ++i)
    if (d->dip[i] != 0.0) return fit_fail(f, MDFF_ERR_UNSUPPORTED, "static dipoles are not supported by the fit");
  const int nt = f->ntype;
  // base pair table and slots
  f->pairs0.assign(nt * nt, FitPair{});
  for (int it = 0; it < nt; ++it)
    for (int jt = 0; jt < nt; ++jt) {
      FitPair& P = f->pairs0[it + nt * jt];
      const int p = it + 16 * jt;
      P.A = d->Abmhftd[p];
      P.B = d->Bbmhftd[p];
      P.rho = P.B != 0.0 ? 1.0 / P.B : 0.0;
      P.C = d->Cbmhftd[p];
      const int k3 = it + 16 * it + 256 * jt;
      P.ldamp = d->ldip_damping[k3] ? 1 : 0;
      P.db = d->pol_damp_b[k3];
      P.dc = d->pol_damp_c[k3];
      P.dk = d->pol_damp_k[k3];
      P.sA = P.sRho = P.sC = P.sB = P.sC2 = -1;
    }
  for (int t = 0; t < nt; ++t) f->alpha0[t] = d->ldip_polar[t] ? d->poldip[t] : 0.0;
  for (int p = 0; p < nparam; ++p) {
    const int kind = spec[3 * p], t1 = spec[3 * p + 1] - 1, t2 = spec[3 * p + 2] - 1;
    if (t1 < 0 || t1 >= nt || ((kind != FP_ALPHA) && (t2 < 0 || t2 >= nt)))
      return fit_fail(f, MDFF_ERR_BAD_ARG, "parameter type index out of range");
    switch (kind) {
      case FP_A: f->pairs0[t1 + nt * t2].sA = p; f->pairs0[t2 + nt * t1].sA = p; break;
      case FP_RHO: f->pairs0[t1 + nt * t2].sRho = p; f->pairs0[t2 + nt * t1].sRho = p; break;
      case FP_C: f->pairs0[t1 + nt * t2].sC = p; f->pairs0[t2 + nt * t1].sC = p; break;
      case FP_ALPHA:
        if (!d->ldip_polar[t1]) return fit_fail(f, MDFF_ERR_BAD_ARG, "alpha parameter on a non-polarisable type");
        break;
      case FP_TTB: f->pairs0[t1 + nt * t2].sB = p; break;
      case FP_TTC: f->pairs0[t1 + nt * t2].sC2 = p; break;
      default: return fit_fail(f, MDFF_ERR_BAD_ARG, "unknown parameter kind");
    }
  }
  // snapshots
  std::vector<int> itv;
  std::vector<double> q, sx, sy, sz, kv;
  std::vector<int> kn, row, nl;
  int a0 = 0, k0 = 0, row0 = 0;
  long long tab0 = 0, nl0 = 0;
  f->snaps.resize(nsnap);
  for (int s = 0; s < nsnap; ++s) {
    FitSnap& S = f->snaps[s];
    const int n = natm[s];
    if (n <= 0) return fit_fail(f, MDFF_ERR_BAD_ARG, "empty snapshot");
    S.n = n;
    S.a0 = a0;
    memcpy(S.A, cells + 9 * (size_t)s, sizeof S.A);
    double B[9], W[3];
    fit_lattice(S.A, B, &S.omega, W);
    const double half = 0.5 * std::min(W[0], std::min(W[1], W[2]));
    const double rc_sr = cut_alpha[3 * s], rc_c = std::min(cut_alpha[3 * s + 1], half);
    S.rc2_sr = rc_sr * rc_sr;
    S.rc2_c = rc_c * rc_c;
    S.alpha = cut_alpha[3 * s + 2];
    S.Kx = kmax[3 * s];
    S.Ky = kmax[3 * s + 1];
    S.Kz = kmax[3 * s + 2];
    S.ncol = S.Kx + S.Ky + S.Kz + 3;
    f->ncol_max = std::max(f->ncol_max, S.ncol);
    S.tab0 = tab0;
    tab0 += (long long)S.ncol * n;
    S.k0 = k0;
    S.npol = 0;
    // per atom
    std::vector<double> fx(n), fy(n), fz(n);
    for (int i = 0; i < n; ++i) {
      const int t = itype[a0 + i] - 1;
      if (t < 0 || t >= nt) return fit_fail(f, MDFF_ERR_BAD_ARG, "atom type out of range");
      itv.push_back(t);
      q.push_back(d->qch[t]);
      S.npol += d->ldip_polar[t] ? 1 : 0;
      const double* r = pos + 3 * ((size_t)a0 + i);
      // s = B^T r
      fx[i] = B[0] * r[0] + B[1] * r[1] + B[2] * r[2];
      fy[i] = B[3] * r[0] + B[4] * r[1] + B[5] * r[2];
      fz[i] = B[6] * r[0] + B[7] * r[1] + B[8] * r[2];
      sx.push_back(fx[i]);
      sy.push_back(fy[i]);
      sz.push_back(fz[i]);
    }
    // full rows within max cutoff (positions are fixed for the fit)
    const double rmax2 = std::max(S.rc2_sr, S.rc2_c) * (1.0 + 1e-9);
    S.row0 = row0;
    S.nl0 = nl0;
    int cnt = 0;
    row.push_back(0);
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double dx = fx[i] - fx[j], dy = fy[i] - fy[j], dz = fz[i] - fz[j];
        dx -= nearbyint(dx);
        dy -= nearbyint(dy);
        dz -= nearbyint(dz);
        const double rx = dx * S.A[0] + dy * S.A[3] + dz * S.A[6];
        const double ry = dx * S.A[1] + dy * S.A[4] + dz * S.A[7];
        const double rz = dx * S.A[2] + dy * S.A[5] + dz * S.A[8];
        if (rx * rx + ry * ry + rz * rz <= rmax2) {
          nl.push_back(j);
          ++cnt;
        }
      }
      row.push_back(cnt);
    }
    row0 += n + 1;
    nl0 += cnt;
    // k-vectors of the half space (src/kspace.f90:162-191)
    const double tpi = 2.0 * MDFF_PI, a2 = S.alpha * S.alpha;
    int nks = 0;
    for (int nx = 0; nx <= S.Kx; ++nx)
      for (int ny = -S.Ky; ny <= S.Ky; ++ny)
        for (int nz = -S.Kz; nz <= S.Kz; ++nz) {
          const bool halfk = (nx > 0) || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
          if (!halfk) continue;
          const double kx = tpi * (nx * B[0] + ny * B[3] + nz * B[6]);
          const double ky = tpi * (nx * B[1] + ny * B[4] + nz * B[7]);
          const double kz = tpi * (nx * B[2] + ny * B[5] + nz * B[8]);
          const double kk = kx * kx + ky * ky + kz * kz;
          kv.push_back(kx);
          kv.push_back(ky);
          kv.push_back(kz);
          kv.push_back(exp(-(kk * 0.25) / a2) / kk);
          kv.push_back(2.0 * (1.0 / kk + 1.0 / a2 / 4.0));
          kn.push_back(nx);
          kn.push_back(ny);
          kn.push_back(nz);
          ++nks;
        }
    S.nk = nks;
    k0 += nks;
    a0 += n;
  }
  f->natm = a0;
  f->nk = k0;
  f->tab_elems = tab0;
  FIT_TRY(f, hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
#define FUP(dst, vec)                                                                                   \
  do {                                                                                                  \
    FIT_TRY(f, hipMalloc((void**)&(dst), sizeof((vec)[0]) * std::max<size_t>(1, (vec).size())));       \
    if (!(vec).empty())                                                                                 \
      FIT_TRY(f, hipMemcpy((dst), (vec).data(), sizeof((vec)[0]) * (vec).size(), hipMemcpyHostToDevice)); \
  } while (0)
  FUP(f->d_snaps, f->snaps);
  FUP(f->d_itype, itv);
  FUP(f->d_q, q);
  FUP(f->d_sx, sx);
  FUP(f->d_sy, sy);
  FUP(f->d_sz, sz);
  FUP(f->d_row, row);
  FUP(f->d_nl, nl);
  FUP(f->d_kv, kv);
  FUP(f->d_kn, kn);
#undef FUP
  FIT_TRY(f, hipMalloc((void**)&f->d_pairs, sizeof(FitPair) * nt * nt));
  FIT_TRY(f, hipMalloc(&f->d_tab32, sizeof(c2<float>) * (size_t)tab0));
  FIT_TRY(f, hipMalloc(&f->d_tab64, sizeof(c2<double>) * (size_t)tab0));
  FIT_TRY(f, hipMalloc((void**)&f->d_fref, sizeof(double) * 3 * (size_t)a0));
  FIT_TRY(f, hipMemset(f->d_fref, 0, sizeof(double) * 3 * (size_t)a0));
  FIT_TRY(f, hipMalloc((void**)&f->d_tauref, sizeof(double) * 9 * (size_t)nsnap));
  FIT_TRY(f, hipMemset(f->d_tauref, 0, sizeof(double) * 9 * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_ws, sizeof(double) * FIT_WS * (size_t)a0));
  FIT_TRY(f, hipMalloc((void**)&f->d_kws, sizeof(double) * FIT_KWS * (size_t)std::max(k0, 1)));
  FIT_TRY(f, hipMalloc((void**)&f->d_F, sizeof(double) * 3 * (size_t)a0));
  FIT_TRY(f, hipMalloc((void**)&f->d_tau, sizeof(double) * 9 * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_loss, sizeof(double) * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_grad, sizeof(double) * FIT_MAXP * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_nscf, sizeof(int) * 2 * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_active, sizeof(int) * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_iter, sizeof(int) * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_count, sizeof(int)));
  FIT_TRY(f, hipMalloc((void**)&f->d_hn, sizeof(double) * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_Gs, sizeof(double) * 9 * (size_t)nsnap));
  FIT_TRY(f, hipMalloc((void**)&f->d_gsave, sizeof(double) * FIT_MAXP * (size_t)nsnap));
  const int tg = std::min(nsnap, 1024);
  hipLaunchKernelGGL((k_fit_tables<float>), dim3(tg), dim3(256), 0, f->stream, nsnap, f->d_snaps, f->d_sx,
                     f->d_sy, f->d_sz, f->d_tab32);
  hipLaunchKernelGGL((k_fit_tables<double>), dim3(tg), dim3(256), 0, f->stream, nsnap, f->d_snaps, f->d_sx,
                     f->d_sy, f->d_sz, f->d_tab64);
  FIT_TRY(f, hipGetLastError());
  FIT_TRY(f, hipStreamSynchronize(f->stream));
  return fit_blocks(f, 256);
}

int mdff_fit_set_targets(mdff_fit* f, const double* fref, const double* tauref, double w_stress) {
  if (!f || !fref || !tauref) return MDFF_ERR_BAD_ARG;
  FIT_TRY(f, hipMemcpy(f->d_fref, fref, sizeof(double) * 3 * (size_t)f->natm, hipMemcpyHostToDevice));
  FIT_TRY(f, hipMemcpy(f->d_tauref, tauref, sizeof(double) * 9 * (size_t)f->nsnap, hipMemcpyHostToDevice));
  f->w_stress = w_stress;
  return MDFF_OK;
}

int mdff_fit_set_launch(mdff_fit* f, int32_t block, int32_t grid) {
  if (!f || block < 8 || block > 256 || (block & (block - 1)) || grid < 1) return MDFF_ERR_BAD_ARG;
  return fit_blocks(f, block);
}

int mdff_fit_eval(mdff_fit* f, const double* theta, int32_t precision, int32_t want_grad, double* loss,
                  double* grad, double* forces, double* tau, int32_t* nscf) {
  if (!f || (f->nparam > 0 && !theta) || !loss) return MDFF_ERR_BAD_ARG;
  if (want_grad && f->nparam > 0 && !grad) return MDFF_ERR_BAD_ARG;
  const int nt = f->ntype;
  std::vector<FitPair> pairs = f->pairs0;
  double alpha_t[FIT_NT];
  for (int t = 0; t < FIT_NT; ++t) alpha_t[t] = t < nt ? f->alpha0[t] : 0.0;
  FitArgs a{};
  for (int t = 0; t < FIT_NT; ++t) a.s_alpha[t] = -1;
  for (int p = 0; p < f->nparam; ++p) {
    const int kind = f->spec[3 * p], t1 = f->spec[3 * p + 1] - 1, t2 = f->spec[3 * p + 2] - 1;
    const double v = theta[p];
    switch (kind) {
      case FP_A: pairs[t1 + nt * t2].A = v; pairs[t2 + nt * t1].A = v; break;
      case FP_RHO:
        if (!(v > 0.0)) return fit_fail(f, MDFF_ERR_BAD_ARG, "rho must be positive");
        pairs[t1 + nt * t2].rho = v; pairs[t2 + nt * t1].rho = v;
        pairs[t1 + nt * t2].B = 1.0 / v; pairs[t2 + nt * t1].B = 1.0 / v;
        break;
      case FP_C: pairs[t1 + nt * t2].C = v; pairs[t2 + nt * t1].C = v; break;
      case FP_ALPHA:
        if (!(v > 0.0)) return fit_fail(f, MDFF_ERR_BAD_ARG, "polarisability must be positive");
        alpha_t[t1] = v; a.s_alpha[t1] = p;
        break;
      case FP_TTB:
        if (!(v > 0.0)) return fit_fail(f, MDFF_ERR_BAD_ARG, "TT b must be positive");
        pairs[t1 + nt * t2].db = v;
        break;
      case FP_TTC:
        if (v == 0.0) return fit_fail(f, MDFF_ERR_BAD_ARG, "TT c must be non-zero");
        pairs[t1 + nt * t2].dc = v;
        break;
    }
  }
  FIT_TRY(f, hipMemcpyAsync(f->d_pairs, pairs.data(), sizeof(FitPair) * nt * nt, hipMemcpyHostToDevice, f->stream));
  a.nsnap = f->nsnap;
  a.ntype = nt;
  a.want_grad = want_grad ? 1 : 0;
  a.min_iter = std::max(1, f->base.min_scf_pol_iter);
  a.max_iter = std::max(a.min_iter, f->base.max_scf_pol_iter);
  a.tol = f->base.conv_tol_ind;
  a.coul_unit = f->base.coul_unit;
  a.press_unit = f->base.press_unit;
  a.w_stress = f->w_stress;
  a.snaps = f->d_snaps;
  a.ablk = f->d_ablk;
  a.kblk = f->d_kblk;
  a.snap_ab0 = f->d_ab0;
  a.snap_kb0 = f->d_kb0;
  a.pa = f->d_pa;
  a.pk = f->d_pk;
  a.active = f->d_active;
  a.iter = f->d_iter;
  a.hn = f->d_hn;
  a.Gs = f->d_Gs;
  a.gsave = f->d_gsave;
  a.count = f->d_count;
  a.gblk = f->d_gblk;
  a.snap_g0 = f->d_g0;
  a.snap_nch = f->d_nch;
  a.gp = f->d_gp;
  a.sblk = f->d_sblk;
  a.snap_s0 = f->d_s0;
  a.snap_sch = f->d_sch;
  a.sp = f->d_sp;
  a.pairs = f->d_pairs;
  for (int t = 0; t < FIT_NT; ++t) a.alpha_t[t] = alpha_t[t];
  a.itype = f->d_itype;
  a.q = f->d_q;
  a.sx = f->d_sx;
  a.sy = f->d_sy;
  a.sz = f->d_sz;
  a.row = f->d_row;
  a.nl = f->d_nl;
  a.kv = f->d_kv;
  a.kn = f->d_kn;
  a.fref = f->d_fref;
  a.tauref = f->d_tauref;
  a.ws = f->d_ws;
  a.kws = f->d_kws;
  a.F = f->d_F;
  a.tau = f->d_tau;
  a.loss = f->d_loss;
  a.grad = f->d_grad;
  a.nscf = f->d_nscf;
  a.ncol_max = f->ncol_max;
  // relative tolerance of the adjoint solve: |residual| <= tol |h| (FP32: not
  // below 2e-6, the resolution of its field sums)
  a.tol_adj = precision == MDFF_PREC_MIXED ? std::max(a.tol, 2e-6) : a.tol;
  if (precision == MDFF_PREC_MIXED) {
    a.tab = f->d_tab32;
    FIT_RC(fit_run<float>(f, a, sizeof(c2<float>)));
  } else {
    a.tab = f->d_tab64;
    FIT_RC(fit_run<double>(f, a, sizeof(c2<double>)));
  }
  FIT_TRY(f, hipGetLastError());
  std::vector<double> ls(f->nsnap), gr((size_t)FIT_MAXP * f->nsnap);
  FIT_TRY(f, hipMemcpyAsync(ls.data(), f->d_loss, sizeof(double) * f->nsnap, hipMemcpyDeviceToHost, f->stream));
  if (want_grad)
    FIT_TRY(f, hipMemcpyAsync(gr.data(), f->d_grad, sizeof(double) * FIT_MAXP * f->nsnap, hipMemcpyDeviceToHost,
                              f->stream));
  if (forces)
    FIT_TRY(f, hipMemcpyAsync(forces, f->d_F, sizeof(double) * 3 * (size_t)f->natm, hipMemcpyDeviceToHost, f->stream));
  if (tau) FIT_TRY(f, hipMemcpyAsync(tau, f->d_tau, sizeof(double) * 9 * f->nsnap, hipMemcpyDeviceToHost, f->stream));
  if (nscf)
    FIT_TRY(f, hipMemcpyAsync(nscf, f->d_nscf, sizeof(int) * 2 * f->nsnap, hipMemcpyDeviceToHost, f->stream));
  FIT_TRY(f, hipStreamSynchronize(f->stream));
  // deterministic sums in snapshot order
  double L = 0.0;
  for (int s = 0; s < f->nsnap; ++s) L += ls[s];
  *loss = L;
  if (want_grad)
    for (int p = 0; p < f->nparam; ++p) {
      double g = 0.0;
      for (int s = 0; s < f->nsnap; ++s) g += gr[(size_t)FIT_MAXP * s + p];
      grad[p] = g;
    }
  return MDFF_OK;
}

int mdff_fit_destroy(mdff_fit* f) {
  if (!f) return MDFF_ERR_BAD_ARG;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  void* ptrs[] = {f->d_snaps, f->d_sblk, f->d_s0, f->d_sch, f->d_sp, f->d_gblk, f->d_g0, f->d_nch, f->d_gp, f->d_ablk, f->d_kblk, f->d_ab0, f->d_kb0, f->d_pa, f->d_pk, f->d_active, f->d_iter,
                  f->d_count, f->d_hn, f->d_Gs, f->d_gsave, f->d_pairs, f->d_itype, f->d_row, f->d_nl, f->d_kn, f->d_nscf,
                  f->d_q, f->d_sx, f->d_sy, f->d_sz, f->d_kv, f->d_tab32, f->d_tab64, f->d_fref,
                  f->d_tauref, f->d_ws, f->d_kws, f->d_F, f->d_tau, f->d_loss, f->d_grad};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
  return MDFF_OK;
}

}  // extern "C"
