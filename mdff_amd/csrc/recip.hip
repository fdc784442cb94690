// recip.hip -- reciprocal-space Ewald sums for gfx950.
//
// Reference: multipole_ES_rec (src/ewald_rec.f90:38-365) with the half-space
// k mesh of kpoint_sum_init (src/kspace.f90:143-207).  The reference loops
// over (k, atom) pairs and calls FP64 COS/SIN of k.r for each.  Here the
// mesh is treated as the box nx in [0,Kx], ny in [-Ky,Ky], nz in [-Kz,Kz]
// (entries outside the half space, and k = 0, carry A(n) = 0), and
// e^{i k.r} = X[nx] Y[ny] Z[nz] with per-dimension phase tables
// X[n] = e^{2 pi i n s_x} (s = B^T r, k.r = 2 pi n.s).  Every phase angle is
// formed in FP64; the tables are stored in the working precision T.
//
//   phase 1  S(n) = sum_i a_i(n) X_i[nx] Y_i[ny] Z_i[nz],
//            a_i(n) = q_i + i 2 pi n.m_i,  m_i = B^T mu_i   (src/ewald_rec.f90:116-148)
//   reduce   S in FP64; W(n) = A(n) conj(S(n)); energy sum A|S|^2 and the
//            stress sum (delta - kcoe k k) A|S|^2              (:173-174, :252-263)
//   phase 2  per atom: U1_d = sum_n n_d Im(W e_i), M_de = sum_n n_d n_e Re(W e_i)
//            field = 2 (4 pi/V) 2 pi B U1; force = 2 (4 pi/V) [q_i 2 pi B U1
//            + (2 pi)^2 B M m_i]                               (:196-233, :300-358)
#include <cstdlib>
#include <type_traits>
#include <utility>

#include "common.h"
#include "reduce.h"

namespace {

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Phase tables, layout [col][atom], col = 0..Kx (X), Kx+1..Kx+Ky+1 (Y), then Z.
// A thread writes a run of up to kTabRun consecutive columns of one
// direction: exp(2 pi i k0 s) by sincospi of the reduced angle k0 s - rint(k0 s),
// then E(k + 1) = E(k) E(1) in FP64 (at most kTabRun - 1 products, each
// within a few units of 2^-53).
constexpr int kTabRun = 8;

// runs of the three directions: blockIdx.y -> (first column, k0, count, which s)
struct TabRun {
  int col, k0, cnt, dir;
};
__device__ __forceinline__ TabRun tab_run(int y, int Kx, int Ky, int Kz) {
  const int rx = (Kx + kTabRun) / kTabRun, ry = (Ky + kTabRun) / kTabRun;
  TabRun r;
  int K, base;
  if (y < rx) { r.dir = 0; K = Kx; base = 0; }
  else if (y < rx + ry) { r.dir = 1; K = Ky; base = Kx + 1; y -= rx; }
  else { r.dir = 2; K = Kz; base = Kx + Ky + 2; y -= rx + ry; }
  r.k0 = y * kTabRun;
  r.cnt = min(kTabRun, K + 1 - r.k0);
  r.col = base + r.k0;
  return r;
}
static int tab_runs(const int* K) {
  return (K[0] + kTabRun) / kTabRun + (K[1] + kTabRun) / kTabRun + (K[2] + kTabRun) / kTabRun;
}

template <typename T>
__device__ __forceinline__ void table_run(double s, const TabRun& r, int n, int i, cplx<T>* __restrict__ tab) {
  double t = (double)r.k0 * s;
  t -= rint(t);
  double sn, cs;
  sincospi(2.0 * t, &sn, &cs);
  double t1 = s - rint(s);
  double s1, c1;
  sincospi(2.0 * t1, &s1, &c1);
  cplx<T>* dst = tab + (size_t)r.col * n + i;
  dst[0] = {(T)cs, (T)sn};
  for (int u = 1; u < r.cnt; ++u) {
    const double c2 = cs * c1 - sn * s1, s2 = cs * s1 + sn * c1;
    cs = c2;
    sn = s2;
    dst[(size_t)u * n] = {(T)cs, (T)sn};
  }
}

template <typename T>
__global__ void k_tables(int n, int Kx, int Ky, int Kz, const double* __restrict__ sx,
                         const double* __restrict__ sy, const double* __restrict__ sz,
                         cplx<T>* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const TabRun r = tab_run(blockIdx.y, Kx, Ky, Kz);
  const double s = r.dir == 0 ? sx[i] : (r.dir == 1 ? sy[i] : sz[i]);
  table_run<T>(s, r, n, i, tab);
}

// The same tables straight from the Cartesian positions: s = B^T r formed
// as kardir does (src/lattice.f90:183-216, products and sums rounded
// separately, no fused multiply-add), so s is bit-identical to the list
// code's; the blocks of the first run (X columns from 0) also store s (sx,
// sy, sz).
struct TabRArgs {
  int n, Kx, Ky, Kz;
  const double *rx, *ry, *rz;
  double b[9];                   // B column-major
  double *sx, *sy, *sz;
  void* tab;
  // pack != 0: also the atoms in the cell order of the last list build
  // (rec[p], meta[p] of atom order[p], as pack_atoms)
  int pack;
  const int32_t* order;
  const double* q;
  const int32_t* itype;
  AtomRec* rec;
  int2* meta;
};

// fin_row >= 0: the blocks of row blockIdx.y == fin_row run the queued final
// sums of the previous MD step instead (block x: sum x).
template <typename T>
__global__ void __launch_bounds__(256) k_tables_r(TabRArgs a, FinalBatch fb, int fin_row) {
  if ((int)blockIdx.y == fin_row) {
    __shared__ double sh[4 * kRedMaxCols];
    run_final_batch(fb, blockIdx.x, sh);
    return;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  auto dot = [](double x, double y, double z, double u, double v, double w) {
    return __dadd_rn(__dadd_rn(__dmul_rn(x, u), __dmul_rn(y, v)), __dmul_rn(z, w));
  };
  const double* b = a.b;
  const double x = a.rx[i], y = a.ry[i], z = a.rz[i];
  const TabRun r = tab_run(blockIdx.y, a.Kx, a.Ky, a.Kz);
  if (blockIdx.y == 0) {
    a.sx[i] = dot(x, y, z, b[0], b[1], b[2]);
    a.sy[i] = dot(x, y, z, b[3], b[4], b[5]);
    a.sz[i] = dot(x, y, z, b[6], b[7], b[8]);
    if (a.pack) {
      const int j = a.order[i];
      const double xj = a.rx[j], yj = a.ry[j], zj = a.rz[j];
      a.rec[i] = AtomRec{dot(xj, yj, zj, b[0], b[1], b[2]), dot(xj, yj, zj, b[3], b[4], b[5]),
                         dot(xj, yj, zj, b[6], b[7], b[8]), a.q[j]};
      a.meta[i] = make_int2(a.itype[j], j);
    }
  }
  const double s = r.dir == 0 ? dot(x, y, z, b[0], b[1], b[2])
                              : (r.dir == 1 ? dot(x, y, z, b[3], b[4], b[5]) : dot(x, y, z, b[6], b[7], b[8]));
  table_run<T>(s, r, a.n, i, reinterpret_cast<cplx<T>*>(a.tab));
}

struct SkArgs {
  int n, Kx, Ky, Kz, NY, NZ, M, NG, GZ;
  int use_q, use_mu;
  int a0_step;                   // atoms per chunk
  int TA;                        // atoms per LDS tile
  const double* q;
  const double* mu;              // 3n AoS
  double b[9];                   // B column-major
  const void* tab;
  void* part;                    // [chunk][2*nk_box] in T
  int nk_box;
};

// Phase 1.  Work item = (g, m): m = nx*NY + (ny+Ky), g = group of GZ
// consecutive nz.  Consecutive lanes take consecutive m, so the Y loads are
// consecutive and the Z loads are wave-uniform (LDS broadcast).
template <typename T, int GZ>
__global__ void __launch_bounds__(256) k_sk(SkArgs a) {
  const int TA = a.TA;
  extern __shared__ unsigned char smem_raw[];
  const int ncol = a.Kx + 1 + a.Ky + 1 + a.Kz + 1;
  cplx<T>* st = reinterpret_cast<cplx<T>*>(smem_raw);          // [ncol][TA]
  T* sq = reinterpret_cast<T*>(st + (size_t)ncol * TA);         // q, mx, my, mz per atom
  const cplx<T>* tab = reinterpret_cast<const cplx<T>*>(a.tab);
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = item % a.M;
  const int g = item / a.M;
  const bool active = g < a.NG;
  const int nx = m / a.NY, ny = m % a.NY - a.Ky;
  const int nz0 = g * GZ - a.Kz;
  const T tp = T(2.0 * MDFF_PI);
  cplx<T> acc[GZ];
#pragma unroll
  for (int k = 0; k < GZ; ++k) acc[k] = {T(0), T(0)};
  const int abeg = blockIdx.y * a.a0_step;
  const int aend = min(a.n, abeg + a.a0_step);
  for (int t0 = abeg; t0 < aend; t0 += TA) {
    const int nt = min(TA, aend - t0);
    __syncthreads();
    for (int x = threadIdx.x; x < ncol * TA; x += blockDim.x) {
      int col = x / TA, ia = x % TA;
      st[x] = (ia < nt) ? tab[(size_t)col * a.n + t0 + ia] : cplx<T>{T(0), T(0)};
    }
    for (int ia = threadIdx.x; ia < TA; ia += blockDim.x) {
      T qq = T(0), mx = T(0), my = T(0), mz = T(0);
      if (ia < nt) {
        int i = t0 + ia;
        if (a.use_q) qq = (T)a.q[i];
        if (a.use_mu) {
          double u0 = a.mu[3 * i], u1 = a.mu[3 * i + 1], u2 = a.mu[3 * i + 2];
          // m = B^T mu
          mx = (T)(u0 * a.b[0] + u1 * a.b[1] + u2 * a.b[2]);
          my = (T)(u0 * a.b[3] + u1 * a.b[4] + u2 * a.b[5]);
          mz = (T)(u0 * a.b[6] + u1 * a.b[7] + u2 * a.b[8]);
        }
      }
      sq[4 * ia] = qq;
      sq[4 * ia + 1] = mx;
      sq[4 * ia + 2] = my;
      sq[4 * ia + 3] = mz;
    }
    __syncthreads();
    if (active) {
      for (int ia = 0; ia < nt; ++ia) {
        const cplx<T> X = st[nx * TA + ia];
        cplx<T> Y = st[(a.Kx + 1 + (ny >= 0 ? ny : -ny)) * TA + ia];
        if (ny < 0) Y.im = -Y.im;
        const cplx<T> XY = cmul(X, Y);
        const T qq = sq[4 * ia];
        cplx<T> P, Pz{T(0), T(0)};
        if (a.use_mu) {
          // a = q + i 2pi (nx mx + ny my) ; the nz part is added per nz
          const T b = tp * ((T)nx * sq[4 * ia + 1] + (T)ny * sq[4 * ia + 2]);
          P = {qq * XY.re - b * XY.im, qq * XY.im + b * XY.re};
          const T c = tp * sq[4 * ia + 3];
          Pz = {-c * XY.im, c * XY.re};
        } else {
          P = {qq * XY.re, qq * XY.im};
        }
        const int zb = (a.Kx + 1 + a.Ky + 1) * TA + ia;
#pragma unroll
        for (int k = 0; k < GZ; ++k) {
          const int nz = nz0 + k;
          const int az = nz >= 0 ? nz : -nz;
          if (az > a.Kz) continue;
          cplx<T> Z = st[zb + az * TA];
          if (nz < 0) Z.im = -Z.im;
          cplx<T> Pk = P;
          if (a.use_mu) {
            Pk.re += (T)nz * Pz.re;
            Pk.im += (T)nz * Pz.im;
          }
          acc[k].re += Pk.re * Z.re - Pk.im * Z.im;
          acc[k].im += Pk.re * Z.im + Pk.im * Z.re;
        }
      }
    }
  }
  if (!active) return;
  T* out = reinterpret_cast<T*>(a.part) + (size_t)blockIdx.y * 2 * a.nk_box;
#pragma unroll
  for (int k = 0; k < GZ; ++k) {
    const int nz = nz0 + k;
    if (nz > a.Kz) continue;
    const int idx = m * a.NZ + (nz + a.Kz);
    out[2 * idx] = acc[k].re;
    out[2 * idx + 1] = acc[k].im;
  }
}

template <typename T>
struct Mma;

// S(n) = sum of chunk partials (FP64, fixed order); W(n) = A conj(S);
// per-block sums of A|S|^2 and (delta_ab - kcoe k_a k_b) A|S|^2.  With amat,
// also the A operand of the matrix-core gather: amat[(mtile * nks + s) * 64
// + lane] holds, for m = 4 mtile + ml and K index kk = 4 s + lane / 16 =
// (nz index t, Re/Im cc), row R = lane % 16 = (ml, component comp):
// (Re, Im) of W Z as a real dot product with (Re Z, Im Z), times nz for
// comp >= 2.  Entries outside the mesh are never written (zero from setup).
template <typename T, typename TP>
__global__ void __launch_bounds__(256) k_sk_reduce(int nk_box, int nchunk, const TP* __restrict__ part,
                                                   const double* __restrict__ kA,
                                                   const double* __restrict__ kvec,
                                                   double* __restrict__ sk, double* __restrict__ wk,
                                                   T* __restrict__ wk_t, double* __restrict__ blk,
                                                   T* __restrict__ amat, int NZ, int Kz, int nks) {
  __shared__ double sh[4 * kRedMaxCols];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  double v[kRedMaxCols];
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) v[c] = 0.0;
  if (k < nk_box) {
    // chunks in order; the loads of 8 chunks are issued before their adds
    double re = 0.0, im = 0.0;
    int c = 0;
    for (; c + 8 <= nchunk; c += 8) {
      TP vr[8], vi[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        vr[u] = part[(size_t)(c + u) * 2 * nk_box + 2 * k];
        vi[u] = part[(size_t)(c + u) * 2 * nk_box + 2 * k + 1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        re += (double)vr[u];
        im += (double)vi[u];
      }
    }
    for (; c < nchunk; ++c) {
      re += (double)part[(size_t)c * 2 * nk_box + 2 * k];
      im += (double)part[(size_t)c * 2 * nk_box + 2 * k + 1];
    }
    sk[2 * k] = re;
    sk[2 * k + 1] = im;
    const double Ak = kA[k];
    wk[2 * k] = Ak * re;
    wk[2 * k + 1] = -Ak * im;
    wk_t[2 * k] = (T)(Ak * re);
    wk_t[2 * k + 1] = (T)(-Ak * im);
    if (amat) {
      const double wr = Ak * re, wi = -Ak * im;
      const int m = k / NZ, t = k - (k / NZ) * NZ;
      const int mt = m >> 2, ml = m & 3;
      const double fz = (double)(t - Kz);
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const int kk = 2 * t + cc;
        T* dst = amat + ((size_t)mt * nks + (kk >> 2)) * 64 + 16 * (kk & 3);
#pragma unroll
        for (int comp = 0; comp < 4; ++comp) {
          double v = ((comp & 1) == 0) ? (cc == 0 ? wr : -wi) : (cc == 0 ? wi : wr);
          if (comp >= 2) v *= fz;
          dst[Mma<T>::row_of(ml, comp)] = (T)v;
        }
      }
    }
    const double str = (re * re + im * im) * Ak;
    const double kx = kvec[4 * k], ky = kvec[4 * k + 1], kz = kvec[4 * k + 2], kc = kvec[4 * k + 3];
    v[0] = str;
    v[1] = (1.0 - kc * kx * kx) * str;
    v[2] = (1.0 - kc * ky * ky) * str;
    v[3] = (1.0 - kc * kz * kz) * str;
    v[4] = -kc * kx * ky * str;
    v[5] = -kc * kx * kz * str;
    v[6] = -kc * ky * kz * str;
  }
  const double r = block_sum_cols(v, 7, sh);
  if (threadIdx.x < 7) blk[blockIdx.x * 7 + threadIdx.x] = r;
}

struct GaArgs {
  int n, Kx, Ky, Kz, NY, NZ, M, m_per_split;
  int moments;                   // 3 (field / charge forces) or 9 (dipole forces)
  const void* tab;
  const void* wk;                // W(n) in T, 2 x nk_box
  double* out;                   // [split][moments][n]
};

// Phase 2: one thread per atom, W streamed through LDS (wave-uniform reads),
// the atom's own Z row held in LDS as [nz][thread] (conflict-free).
template <typename T, int MOM>
__global__ void __launch_bounds__(256) k_gather(GaArgs a) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int MT = 8;                                     // m values per W tile
  const int nthr = blockDim.x;
  cplx<T>* zt = reinterpret_cast<cplx<T>*>(smem_raw);       // [Kz+1][nthr]
  cplx<T>* wt = zt + (size_t)(a.Kz + 1) * nthr;             // [MT][NZ]
  const cplx<T>* tab = reinterpret_cast<const cplx<T>*>(a.tab);
  const cplx<T>* W = reinterpret_cast<const cplx<T>*>(a.wk);
  const int i = blockIdx.x * nthr + threadIdx.x;
  const bool act = i < a.n;
  const int zcol = a.Kx + 1 + a.Ky + 1;
  for (int k = 0; k <= a.Kz; ++k)
    zt[k * nthr + threadIdx.x] = act ? tab[(size_t)(zcol + k) * a.n + i] : cplx<T>{T(0), T(0)};
  double u0 = 0, u1 = 0, u2 = 0, mxx = 0, myy = 0, mzz = 0, mxy = 0, mxz = 0, myz = 0;
  const int mbeg = blockIdx.y * a.m_per_split;
  const int mend = min(a.M, mbeg + a.m_per_split);
  for (int mt = mbeg; mt < mend; mt += MT) {
    const int nm = min(MT, mend - mt);
    __syncthreads();
    for (int x = threadIdx.x; x < nm * a.NZ; x += nthr) wt[x] = W[(size_t)mt * a.NZ + x];
    __syncthreads();
    if (!act) continue;
    for (int mm = 0; mm < nm; ++mm) {
      const int m = mt + mm;
      const int nx = m / a.NY, ny = m % a.NY - a.Ky;
      cplx<T> X = tab[(size_t)nx * a.n + i];
      cplx<T> Y = tab[(size_t)(a.Kx + 1 + (ny >= 0 ? ny : -ny)) * a.n + i];
      if (ny < 0) Y.im = -Y.im;
      const cplx<T> XY = cmul(X, Y);
      T z0r = 0, z0i = 0, z1r = 0, z1i = 0, z2r = 0, z2i = 0;
      const cplx<T>* wrow = wt + mm * a.NZ;
      // negative nz: Z = conj(Z[|nz|])
      for (int az = a.Kz; az >= 1; --az) {
        const cplx<T> w = wrow[a.Kz - az];
        const cplx<T> z = zt[az * nthr + threadIdx.x];
        const T tr = w.re * z.re + w.im * z.im;   // w * conj(z)
        const T ti = w.im * z.re - w.re * z.im;
        const T fz = -(T)az;
        z0r += tr; z0i += ti;
        z1r += fz * tr; z1i += fz * ti;
        if (MOM == 9) { z2r += fz * fz * tr; z2i += fz * fz * ti; }
      }
      for (int az = 0; az <= a.Kz; ++az) {
        const cplx<T> w = wrow[a.Kz + az];
        const cplx<T> z = zt[az * nthr + threadIdx.x];
        const T tr = w.re * z.re - w.im * z.im;
        const T ti = w.re * z.im + w.im * z.re;
        const T fz = (T)az;
        z0r += tr; z0i += ti;
        z1r += fz * tr; z1i += fz * ti;
        if (MOM == 9) { z2r += fz * fz * tr; z2i += fz * fz * ti; }
      }
      // Im and Re of XY * Z_p
      const T i0 = XY.re * z0i + XY.im * z0r, r0 = XY.re * z0r - XY.im * z0i;
      const T i1 = XY.re * z1i + XY.im * z1r, r1 = XY.re * z1r - XY.im * z1i;
      const double fx = (double)nx, fy = (double)ny;
      u0 += fx * (double)i0;
      u1 += fy * (double)i0;
      u2 += (double)i1;
      if (MOM == 9) {
        const T r2 = XY.re * z2r - XY.im * z2i;
        mxx += fx * fx * (double)r0;
        myy += fy * fy * (double)r0;
        mxy += fx * fy * (double)r0;
        mxz += fx * (double)r1;
        myz += fy * (double)r1;
        mzz += (double)r2;
      }
    }
  }
  if (!act) return;
  double* o = a.out + (size_t)blockIdx.y * MOM * a.n;
  o[i] = u0;
  o[(size_t)a.n + i] = u1;
  o[(size_t)2 * a.n + i] = u2;
  if (MOM == 9) {
    o[(size_t)3 * a.n + i] = mxx;
    o[(size_t)4 * a.n + i] = myy;
    o[(size_t)5 * a.n + i] = mzz;
    o[(size_t)6 * a.n + i] = mxy;
    o[(size_t)7 * a.n + i] = mxz;
    o[(size_t)8 * a.n + i] = myz;
  }
}

// sum over splits, then field / force in Cartesian (src/ewald_rec.f90:196-233,
// scaled as :300-358: x2 for the half mesh, x 4pi/V)
__global__ void k_gather_final(RecFinal r, int field, int forces, double* __restrict__ ef,
                               double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.n) return;
  double v[9], g[3];
  rec_final_moments(r, i, v, g);
  if (field) {
    double e[3];
    rec_final_field(r, g, e);
    ef[3 * i] = e[0];
    ef[3 * i + 1] = e[1];
    ef[3 * i + 2] = e[2];
  }
  if (forces) {
    double fv[3];
    rec_final_force(r, i, v, g, fv);
    f[3 * i] = fv[0];
    f[3 * i + 1] = fv[1];
    f[3 * i + 2] = fv[2];
  }
}


// ---------------------------------------------------------------------------
// Matrix-core form of phase 2 (field / charge forces).
//
// For one 16-atom tile the contraction over nz is a real GEMM on MFMA
// 16x16x4: rows R = (m_local, comp) of 4 values of m = (nx, ny) and four
// components comp = (Re U0, Im U0, Re U1, Im U1), U_p(i, m) = sum_nz nz^p
// W(m, nz) Z_i(nz); K = (nz, re/im of Z) padded to a multiple of 4; columns =
// the 16 atoms.  The A operand (W, arranged per lane) is written once per
// evaluation by k_sk_reduce; the B operand (the atom's Z row) stays in
// registers for the whole kernel.  With the row order below, lane l holds the
// four components of atom (l % 16) for m_local = l / 16 in its four
// accumulator registers, so the (nx, ny) epilogue (multiply by
// X_i[nx] Y_i[ny], weight by nx, ny) is lane-local; the four m_local lane
// groups are summed with two cross-lane shuffles at the very end.
// Lane layouts (checked on gfx950): A(row = l%16, k = l/16), B(k = l/16,
// col = l%16); D f32: row = 4*(l/16) + r; D f64: row = 4*r + l/16.
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;
template <>
struct Mma<float> {
  typedef f32x4 acc_t;
  __device__ static __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // A row R -> (m_local, comp) so that D register r = comp for m_local = l/16
  __host__ __device__ static __forceinline__ void row(int R, int& ml, int& comp) { ml = R >> 2; comp = R & 3; }
  __host__ __device__ static __forceinline__ int row_of(int ml, int comp) { return 4 * ml + comp; }
};
template <>
struct Mma<double> {
  typedef f64x4 acc_t;
  __device__ static __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __host__ __device__ static __forceinline__ void row(int R, int& ml, int& comp) { comp = R >> 2; ml = R & 3; }
  __host__ __device__ static __forceinline__ int row_of(int ml, int comp) { return 4 * comp + ml; }
};

struct GmArgs {
  int n, Kx, Ky, Kz, NY, NZ, M, nks, mt_per_split, mtiles;
  const void* tab;
  const void* amat;
  double* out;                   // [split][3][n]
};

// Two waves per SIMD where the B operand, the double-buffered A operand and
// the accumulators fit in 256 registers without spills (the instances
// listed), one otherwise.
template <typename T, int CT, int NKS>
constexpr int gather_waves() {
  return sizeof(T) == 4 ? ((CT == 4 && NKS <= 25) ? 2 : 1)
                        : (((CT == 2 && NKS <= 25) || (CT == 1 && NKS <= 33)) ? 2 : 1);
}

template <typename T, int CT, int NKS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gather_waves<T, CT, NKS>())))
k_gather_mma(GmArgs a) {
  typedef typename Mma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const int atom0 = (blockIdx.x * (blockDim.x >> 6) + wave) * (16 * CT);
  const cplx<T>* tab = reinterpret_cast<const cplx<T>*>(a.tab);
  const T* amat = reinterpret_cast<const T*>(a.amat);
  const int zcol = a.Kx + 1 + a.Ky + 1;
  // B operand: the Z row of this lane's atoms, zero beyond NZ.  Every load
  // reads a clamped, valid address and the zero is selected afterwards, so
  // all CT * NKS loads are in flight together.
  T breg[CT][NKS];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int ia = atom0 + ct * 16 + col;
    const bool ok = ia < a.n;
    const int iac = min(ia, a.n - 1);
#pragma unroll
    for (int s = 0; s < NKS; ++s) {
      const int k = 4 * s + g;
      const int t = k >> 1;
      const int nz = t - a.Kz;
      const int az = min(nz < 0 ? -nz : nz, a.Kz);
      // Re (k even) or Im (k odd) of Z, conjugated for nz < 0
      const T v = reinterpret_cast<const T*>(tab + (size_t)(zcol + az) * a.n + iac)[k & 1];
      const T w = (ok && t < a.NZ) ? (((k & 1) && nz < 0) ? T(-1) : T(1)) : T(0);
      breg[ct][s] = w * v;
    }
  }
  double ax[CT], ay[CT], az3[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) { ax[ct] = 0.0; ay[ct] = 0.0; az3[ct] = 0.0; }
  const int mt0 = blockIdx.y * a.mt_per_split;
  const int mt1 = min(a.mtiles, mt0 + a.mt_per_split);
  // The A operand of tile mt+1 is loaded at the top of tile mt; the
  // epilogue's X[nx] Y[ny] rows of tile mt at its top, used after its MFMAs
  // (a scheduling barrier keeps their use, and so the wait for them, behind
  // the MFMAs).  Addresses are clamped to valid rows, lanes beyond M get a
  // zero weight, so nothing here branches.  The lane's m advances by 4 per
  // tile; (nx, ny + Ky) are stepped, not divided.
  int m = 4 * mt0 + g;
  int nx = min(m, a.M - 1) / a.NY;
  int nyo = min(m, a.M - 1) - nx * a.NY;                          // ny + Ky
  int iac[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) iac[ct] = min(atom0 + ct * 16 + col, a.n - 1);
  T abuf[NKS];
  auto load_a = [&](int mtl) {
    const T* ap = amat + (size_t)mtl * NKS * 64 + lane;
#pragma unroll
    for (int s = 0; s < NKS; ++s) abuf[s] = ap[(size_t)s * 64];
  };
  if (mt0 < mt1) load_a(mt0);
  for (int mt = mt0; mt < mt1; ++mt) {
    T cur[NKS];
#pragma unroll
    for (int s = 0; s < NKS; ++s) cur[s] = abuf[s];
    cplx<T> Xv[CT], Yv[CT];
    {
      const int nyy = nyo - a.Ky;
      const size_t xrow = (size_t)min(nx, a.Kx) * a.n, yrow = (size_t)(a.Kx + 1 + (nyy >= 0 ? nyy : -nyy)) * a.n;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        Xv[ct] = tab[xrow + iac[ct]];
        Yv[ct] = tab[yrow + iac[ct]];
      }
    }
    const T mw = m < a.M ? T(1) : T(0);
    const int ny = nyo - a.Ky;
    const double fx = (double)nx, fy = (double)ny;
    const T ysg = ny < 0 ? T(-1) : T(1);
    // the next tile's m
    m += 4;
    nyo += 4;
    while (nyo >= a.NY) { nyo -= a.NY; ++nx; }          // NY = 2 Ky + 1 may be below 4
    if (mt + 1 < mt1) load_a(mt + 1);
    acc_t acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[ct] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int s = 0; s < NKS; ++s) {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[ct] = Mma<T>::op(cur[s], breg[ct][s], acc[ct]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      cplx<T> Y = Yv[ct];
      Y.im *= ysg;
      cplx<T> XY = cmul(Xv[ct], Y);
      XY.re *= mw;
      XY.im *= mw;
      const double im0 = (double)(XY.re * acc[ct][1] + XY.im * acc[ct][0]);
      const double im1 = (double)(XY.re * acc[ct][3] + XY.im * acc[ct][2]);
      ax[ct] += fx * im0;
      ay[ct] += fy * im0;
      az3[ct] += im1;
    }
  }
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    ax[ct] += __shfl_xor(ax[ct], 16);
    ax[ct] += __shfl_xor(ax[ct], 32);
    ay[ct] += __shfl_xor(ay[ct], 16);
    ay[ct] += __shfl_xor(ay[ct], 32);
    az3[ct] += __shfl_xor(az3[ct], 16);
    az3[ct] += __shfl_xor(az3[ct], 32);
  }
  if (g == 0) {
    double* o = a.out + (size_t)blockIdx.y * 3 * a.n;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int ia = atom0 + ct * 16 + col;
      if (ia < a.n) {
        o[ia] = ax[ct];
        o[(size_t)a.n + ia] = ay[ct];
        o[(size_t)2 * a.n + ia] = az3[ct];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Matrix-core structure factor, transposed form: S^T(nz, m) = sum_i
// Z_i(nz) P_i(m).  Rows R = (nz, Re/Im) in 7 tiles of 16 (8 nz each; 56 nz
// per block, blockIdx.z picks the group of 56), taken straight from the Z table; columns = 16 values of m = (nx, ny); K = (atom,
// Re/Im of P).  The generated operand P_i(m) = a_i X_i[nx] Y_i[ny] is made
// once per lane and K step and feeds the 7 row tiles.
//   mode 0: a = q            (charges)
//   mode 1: a = q + i 2pi (nx m_x + ny m_y)   (charges and dipoles, first part)
//   mode 2: a = m_z          (second part: S = S1 + i 2pi nz S2)
// FP32/FP64 MFMA accumulation within one staged atom tile, FP64 across tiles;
// per-chunk FP64 partials.
// ---------------------------------------------------------------------------
struct Sm2Args {
  int n, Kx, Ky, Kz, NY, NZ, M, mtiles, chunk, TA, mode, nk_box;
  const double* q;
  const double* mu;
  double b[9];
  const void* tab;
  double* part;                  // [chunk][2*nk_box] FP64
};

template <typename T, int CTM, int TA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_sk_mma2(Sm2Args a) {
  typedef typename Mma<T>::acc_t acc_t;
  constexpr int RT = 7;                                          // row tiles: 56 nz values max
  constexpr int TAP = TA + 2;                                    // padded atom stride of a staged column
  static_assert(256 % TA == 0 && TA % 2 == 0, "atom tile must divide the block");
  extern __shared__ unsigned char smem_raw[];
  const int ncol = a.Kx + a.Ky + a.Kz + 3;
  // staged tile, column-major: st[col * TAP + atom]; then the a-coefficients
  // sq[c * TAP + atom], c = 0..2.  A K step advances the atom by 2, a fixed
  // byte offset, so the unrolled step loop addresses LDS with immediates.
  cplx<T>* st = reinterpret_cast<cplx<T>*>(smem_raw);
  T* sq = reinterpret_cast<T*>(st + (size_t)ncol * TAP);
  const cplx<T>* tab = reinterpret_cast<const cplx<T>*>(a.tab);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const int kk = g & 1, ja = g >> 1;                             // K: atom 2s+ja, Re/Im kk
  const int zcol = a.Kx + 1 + a.Ky + 1;
  // A rows of this lane (R = col): nz index and part per row tile.  A =
  // part 0: (Zr, -Zi); part 1: (Zi, Zr), with Z(-n) = conj Z(n).  Rows
  // beyond NZ read a valid column; their D rows are never stored.
  const int zoff = blockIdx.z * (RT * 8);                       // nz rows of this block
  const bool take_im = (col & 1) ^ kk;                           // which Z component feeds A
  const T* stT = reinterpret_cast<const T*>(st);
  const T* zp[RT];
  T zs[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const int nzi = zoff + rt * 8 + (col >> 1);
    const int part = col & 1;
    const int nz = nzi - a.Kz;
    const int zc = nzi < a.NZ ? zcol + (nz < 0 ? -nz : nz) : zcol;
    const T sgn_im = nz < 0 ? T(-1) : T(1);
    zs[rt] = part == 0 ? (kk == 0 ? T(1) : -sgn_im) : (kk == 0 ? sgn_im : T(1));
    zp[rt] = stT + 2 * ((size_t)zc * TAP + ja) + (take_im ? 1 : 0);
  }
  // B columns of this lane: m per m-tile
  const int mt0 = (blockIdx.x * (blockDim.x >> 6) + wave) * CTM;
  const cplx<T>* xp[CTM];
  const cplx<T>* yp[CTM];
  T bnyf[CTM], bnxf[CTM], bw[CTM];
  bool bneg[CTM], bok[CTM];
#pragma unroll
  for (int j = 0; j < CTM; ++j) {
    const int m = (mt0 + j) * 16 + col;
    bok[j] = m < a.M;
    const int nx = bok[j] ? m / a.NY : 0, ny = bok[j] ? m % a.NY - a.Ky : 0;
    xp[j] = st + (size_t)nx * TAP + ja;
    yp[j] = st + (size_t)(a.Kx + 1 + (ny >= 0 ? ny : -ny)) * TAP + ja;
    bneg[j] = ny < 0;
    bnxf[j] = (T)nx;
    bnyf[j] = (T)ny;
    bw[j] = bok[j] ? T(1) : T(0);      // columns beyond M: zero B (a multiply, no branch)
  }
  const T* qp = sq + ja;
  acc_t acc[RT][CTM];
  double dacc[RT][CTM][4];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int j = 0; j < CTM; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) dacc[rt][j][r] = 0.0;
  const int abeg = blockIdx.y * a.chunk;
  const int aend = min(a.n, abeg + a.chunk);
  // staging: thread -> (atom sia, columns scg, scg + CG, ...); loads from
  // clamped addresses in batches of 8, so they are in flight together
  constexpr int CG = 256 / TA;
  const int sia = threadIdx.x % TA, scg = threadIdx.x / TA;
  for (int t0 = abeg; t0 < aend; t0 += TA) {
    const int nt = min(TA, aend - t0);
    const bool sok = sia < nt;
    const size_t src = (size_t)t0 + min(sia, nt - 1);
    __syncthreads();
    for (int c0 = scg; c0 < ncol; c0 += 8 * CG) {
      cplx<T> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = tab[(size_t)min(c0 + u * CG, ncol - 1) * a.n + src];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (c0 + u * CG < ncol) st[(size_t)(c0 + u * CG) * TAP + sia] = sok ? v[u] : cplx<T>{T(0), T(0)};
    }
    if (threadIdx.x < TA) {
      const int ia = threadIdx.x;
      T c0 = T(0), c1 = T(0), c2 = T(0);
      if (ia < nt) {
        const int i = t0 + ia;
        if (a.mode <= 1) c0 = (T)a.q[i];
        if (a.mode >= 1) {
          const double u0 = a.mu[3 * i], u1 = a.mu[3 * i + 1], u2 = a.mu[3 * i + 2];
          if (a.mode == 1) {
            c1 = (T)(2.0 * MDFF_PI * (u0 * a.b[0] + u1 * a.b[1] + u2 * a.b[2]));
            c2 = (T)(2.0 * MDFF_PI * (u0 * a.b[3] + u1 * a.b[4] + u2 * a.b[5]));
          } else {
            c0 = (T)(u0 * a.b[6] + u1 * a.b[7] + u2 * a.b[8]);
          }
        }
      }
      sq[ia] = c0;
      sq[TAP + ia] = c1;
      sq[2 * TAP + ia] = c2;
    }
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int j = 0; j < CTM; ++j) acc[rt][j] = acc_t{T(0), T(0), T(0), T(0)};
    // K steps of 4: atoms 2s and 2s+1 (Re/Im of P).  Rows beyond nt are the
    // zero padding of the staged tile, so every tile runs all TA/2 steps and
    // no step needs a mask.
#pragma unroll
    for (int s = 0; s < TA / 2; ++s) {
      T av[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) av[rt] = zs[rt] * zp[rt][4 * s];
      const T c0 = qp[2 * s], c1 = qp[TAP + 2 * s], c2 = qp[2 * TAP + 2 * s];
      T bv[CTM];
#pragma unroll
      for (int j = 0; j < CTM; ++j) {
        const cplx<T> X = xp[j][2 * s];
        const cplx<T> Y = yp[j][2 * s];
        const T yim = bneg[j] ? -Y.im : Y.im;
        const T xyr = X.re * Y.re - X.im * yim;
        const T xyi = X.re * yim + X.im * Y.re;
        // (c0 + i bb) XY, bb = 2pi (nx mx + ny my) in mode 1 (c1, c2 carry 2pi)
        const T bb = bnxf[j] * c1 + bnyf[j] * c2;
        const T pre = c0 * xyr - bb * xyi;
        const T pim = c0 * xyi + bb * xyr;
        bv[j] = bw[j] * (kk == 0 ? pre : pim);
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int j = 0; j < CTM; ++j) acc[rt][j] = Mma<T>::op(av[rt], bv[j], acc[rt][j]);
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int j = 0; j < CTM; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) dacc[rt][j][r] += (double)acc[rt][j][r];
  }
  // D: lane holds rows Rd(r) of column m; row Rd -> nz = rt*8 + Rd/2, part Rd%2
  double* out = a.part + (size_t)blockIdx.y * 2 * a.nk_box;
#pragma unroll
  for (int j = 0; j < CTM; ++j) {
    if (!bok[j]) continue;
    const int m = (mt0 + j) * 16 + col;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int Rd = (sizeof(T) == 4) ? (4 * g + r) : (4 * r + g);
        const int nzi = zoff + rt * 8 + (Rd >> 1);
        if (nzi >= a.NZ) continue;
        out[2 * ((size_t)m * a.NZ + nzi) + (Rd & 1)] = dacc[rt][j][r];
      }
  }
}

// S = S1 + i 2pi nz S2 from two partial sets (dipole amplitude)
template <typename TP>
__global__ void k_sk_combine(int nk_box, int NZ, int Kz, int nchunk, TP* __restrict__ p1,
                             const TP* __restrict__ p2) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)nchunk * nk_box) return;
  const int k = (int)(x % nk_box);
  const double fz = 2.0 * MDFF_PI * (double)(k % NZ - Kz);
  const double s2r = p2[2 * x], s2i = p2[2 * x + 1];
  p1[2 * x] = (TP)((double)p1[2 * x] - fz * s2i);
  p1[2 * x + 1] = (TP)((double)p1[2 * x + 1] + fz * s2r);
}

// ---------------------------------------------------------------------------
// Split-half matrix-core kernels of the mixed form.
//
// The FP32 operands of the two GEMMs (structure factor, gather) are each
// split into two FP16 halves, v = hi + lo (both rounded to nearest: |v - hi
// - lo| <= 2^-22 |v|), and a product is formed from three FP16 MFMAs with
// FP32 accumulation: a b ~ ah bh + ah bl + al bh (the dropped al bl is below
// 2^-22 |a b|).  That is FP32-class accuracy at the FP16 matrix rate of
// v_mfma_f32_32x32x16_f16.  Lane layouts (checked on gfx950):
// A(row = l%32, k = 8*(l/32) + j), B(k = 8*(l/32) + j, col = l%32),
// D register r: row = 8*(r/4) + 4*(l/32) + r%4, col = l%32.
// Operands whose range is not bounded (the gather's W) are scaled per tile
// by a power of two before the split and unscaled after the product.
// ---------------------------------------------------------------------------
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes: one half8

__device__ __forceinline__ f32x16 mma_h(half8 a, half8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// (m / d, m % d) for 0 <= m < 2^20 from a float reciprocal inv = 1/d and
// one correction: branch-free, so loads that use the result issue together
__device__ __forceinline__ void divmod_f(int m, int d, float inv, int& q, int& r) {
  q = (int)((float)m * inv);
  r = m - q * d;
  q = r < 0 ? q - 1 : (r >= d ? q + 1 : q);
  r = m - q * d;
}

// v ~ hi + lo
__device__ __forceinline__ void split_h(float v, _Float16& hi, _Float16& lo) {
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

__device__ __forceinline__ uint32_t pack_h(_Float16 a, _Float16 b) {
  return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// Power-of-two scales of the dipole passes of the split-half structure
// factor, from the current dipoles: the largest |P| an atom can give is
// brought into [256, 512), so that the FP16 halves of the products stay
// normal numbers (a lo half below 2^-14 loses bits).  Mode 1:
// |q| + 2 pi (Kx |m.b1| + Ky |m.b2|); mode 2: |m.b3|.  One block.
__global__ void __launch_bounds__(256) k_mu_pscale(int n, int Kx, int Ky, const double* __restrict__ q,
                                                   const double* __restrict__ mu, double b0, double b1, double b2,
                                                   double b3, double b4, double b5, double b6, double b7,
                                                   double b8, float* __restrict__ out) {
  __shared__ double s1[256], s2[256];
  double m1 = 0.0, m2 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double u0 = mu[3 * i], u1 = mu[3 * i + 1], u2 = mu[3 * i + 2];
    const double c1 = fabs(u0 * b0 + u1 * b1 + u2 * b2), c2 = fabs(u0 * b3 + u1 * b4 + u2 * b5);
    const double c3 = fabs(u0 * b6 + u1 * b7 + u2 * b8);
    m1 = fmax(m1, fabs(q[i]) + 2.0 * MDFF_PI * (Kx * c1 + Ky * c2));
    m2 = fmax(m2, c3);
  }
  s1[threadIdx.x] = m1;
  s2[threadIdx.x] = m2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 0;
    out[0] = s1[0] > 0.0 ? (float)ldexp(1.0, (frexp(s1[0], &e), 9 - e)) : 1.f;
    out[1] = s2[0] > 0.0 ? (float)ldexp(1.0, (frexp(s2[0], &e), 9 - e)) : 1.f;
  }
}

// ---- structure factor: S^T(nz, m) = sum_i Z_i(nz) P_i(m), split halves ----
// Rows R = (nz, Re/Im) in 4 tiles of 32 (16 nz each; 64 nz per block,
// blockIdx.z picks the group); columns = 32 values of m per wave; K = (atom,
// Re/Im of P), 8 atoms per MFMA.  The A operand (Z, signed per row part) is
// split once per staged tile into LDS; the B operand P_i(m) = a_i X_i[nx]
// Y_i[ny] is formed and split per lane.  Modes as k_sk_mma2; the charges
// are scaled by pscale (a power of two) and the sums written times
// 1/pscale, FP32 per atom chunk.  With dipoles the scale of the pass comes
// from device memory (pscale_dev[0]: mode 1, [1]: mode 2), set by
// k_mu_pscale from the current dipoles.
struct Sh3Args {
  int n, Kx, Ky, Kz, NY, NZ, M, chunk, mode, nk_box;
  float pscale;
  const float* pscale_dev;
  const double* q;
  const double* mu;
  double b[9];
  const void* tab;               // FP32 phase tables
  float* part;                   // [chunk][2*nk_box]
};

// WIDE: up to 128 staged X / Y columns (16 per staging thread) and Y staged
// for ny >= 0 only, conjugated on read for ny < 0 (the LDS of a wide mesh,
// e.g. Kx = Ky = 56, stays within two blocks per CU)
template <int TA, int MODE, bool WIDE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_sk_h3(Sh3Args a) {
  constexpr bool CQ = MODE != 1;
  const float psc = a.pscale_dev ? a.pscale_dev[MODE == 2 ? 1 : 0] : a.pscale;
  // CQ: P = c X Y with one coefficient per atom (modes 0 and 2): c X is
  // formed once per tile when X is staged; otherwise (mode 1) P =
  // (c0 - i (nx c1 + ny c2)) X Y is formed per column
  constexpr int RT = 4;                                   // row tiles of 32 rows
  constexpr int RW = TA + 4;                              // 32-bit words per staged Z row
  constexpr int TAP = TA + 4;                             // atoms per staged X / Y column
  static_assert(256 % TA == 0 && TA % 8 == 0, "atom tile");
  extern __shared__ unsigned char smem_raw[];
  uint32_t* zs = reinterpret_cast<uint32_t*>(smem_raw);                 // [plane hi/lo][128 rows][RW]
  const int nxy = a.Kx + a.Ky + 2;                                       // X and Y columns
  // staged columns: X(nx) (times c when CQ), nx = 0..Kx, then Y(ny),
  // ny = -Ky..Ky (Y(-n) = conj Y(n)); WIDE: ny = 0..Ky
  const int nst = WIDE ? nxy : a.Kx + 1 + 2 * a.Ky + 1;
  cplx<float>* st = reinterpret_cast<cplx<float>*>(zs + 2 * 128 * RW);  // [nst][TAP]
  float* sq = reinterpret_cast<float*>(st + (size_t)nst * TAP);         // [3][TAP]
  const cplx<float>* tab = reinterpret_cast<const cplx<float>*>(a.tab);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hl = lane >> 5, col = lane & 31;
  // logical block (m block bx fastest, then chunk by, then z block bz) in
  // XCD order: the m blocks of one atom chunk run on one XCD
  const int gx = gridDim.x, gy = gridDim.y;
  const int lin = xcd_order(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * (int)gridDim.z);
  const int bx = lin % gx, by = (lin / gx) % gy, bz = lin / (gx * gy);
  const int zoff = bz * (RT * 16);
  const int zcol = a.Kx + 1 + a.Ky + 1;
  // B column of this lane
  const int m = (bx * 4 + wave) * 32 + col;
  const bool bok = m < a.M;
  const int nx = bok ? m / a.NY : 0, ny = bok ? m % a.NY - a.Ky : 0;
  const float4* xp = reinterpret_cast<const float4*>(st + (size_t)nx * TAP + 4 * hl);
  const int ycol = WIDE ? a.Kx + 1 + (ny < 0 ? -ny : ny) : a.Kx + 1 + a.Ky + ny;
  const float4* yp = reinterpret_cast<const float4*>(st + (size_t)ycol * TAP + 4 * hl);
  const float ysg = (WIDE && ny < 0) ? -1.f : 1.f;          // conj Y(|ny|) for ny < 0
  // columns m >= M (bok false) read valid columns; their D columns are not
  // stored, and every D column depends on its own B column only
  const float fnx = (float)nx, fny = (float)ny;
  // per-tile sums (MFMA accumulators, restarted every tile) and the chunk
  // totals they are added to: the FP32 rounding of each accumulation is
  // relative to a tile's sum, not to the growing chunk sum
  f32x16 acc[RT], tot[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int r = 0; r < 16; ++r) tot[rt][r] = 0.f;
  const int abeg = by * a.chunk;
  const int aend = min(a.n, abeg + a.chunk);
  // staging: thread -> atom sia of the tile; Z rows scg + CG u (u < ZU), X/Y
  // columns scg + CG u (u < XU, nxy <= CG XU); the loads of tile t+1 are
  // issued before the MFMAs of tile t and written to LDS after them
  constexpr int CG = 256 / TA;
  constexpr int ZU = RT * 16 / CG;
  constexpr int XU = WIDE ? 16 : 8;
  const int sia = threadIdx.x % TA, scg = threadIdx.x / TA;
  cplx<float> zn[ZU], xn[XU];
  double qd = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;        // raw a-coefficient inputs (used at store time)
#define SKH3_FETCH(T0)                                                                        \
  do {                                                                                        \
    const int nt_ = min(TA, aend - (T0));                                                     \
    const size_t src_ = (size_t)(T0) + min(sia, nt_ - 1);                                      \
    _Pragma("unroll") for (int u = 0; u < ZU; ++u) {                                          \
      const int nz_ = zoff + scg + CG * u - a.Kz;                                             \
      zn[u] = tab[(size_t)(zcol + min(nz_ < 0 ? -nz_ : nz_, a.Kz)) * a.n + src_];             \
    }                                                                                         \
    _Pragma("unroll") for (int u = 0; u < XU; ++u)                                            \
      xn[u] = tab[(size_t)min(scg + CG * u, nxy - 1) * a.n + src_];                           \
    if (CQ || threadIdx.x < TA) {                                                             \
      const int i_ = (int)src_;                                                               \
      if (MODE <= 1) qd = a.q[i_];                                                          \
      if (MODE >= 1) {                                                                      \
        m0 = a.mu[3 * i_];                                                                    \
        m1 = a.mu[3 * i_ + 1];                                                                \
        m2 = a.mu[3 * i_ + 2];                                                                \
      }                                                                                       \
    }                                                                                         \
  } while (0)
  if (abeg < aend) SKH3_FETCH(abeg);
  for (int t0 = abeg; t0 < aend; t0 += TA) {
    const int nt = min(TA, aend - t0);
    const bool sok = sia < nt;
    // every other tile stages -Z and its sums are subtracted: an offset that
    // the matrix cores' FP32 accumulation adds whatever the sign of the terms
    // (it grew with the atom count and acted as a charge at the phase origin:
    // 3.3e-5 force error on a relaxed 200k LiF-like melt, 2.1e-6 with the
    // alternation) cancels between tiles
    const float tsg = (((t0 - abeg) / TA) & 1) ? -1.f : 1.f;
    __syncthreads();
    // Z rows: part 0 (Zr, -Zi), part 1 (Zi, Zr), Z(-n) = conj Z(n); atoms
    // beyond the chunk get zeros (their B is zero too)
#pragma unroll
    for (int u = 0; u < ZU; ++u) {
      const int nzl = scg + CG * u;
      const int nzi = zoff + nzl;
      cplx<float> z{tsg * zn[u].re, tsg * zn[u].im};
      if (nzi - a.Kz < 0) z.im = -z.im;
      if (!sok || nzi >= a.NZ) z = cplx<float>{0.f, 0.f};
      _Float16 rh, rl, ih, il;
      split_h(z.re, rh, rl);
      split_h(z.im, ih, il);
      uint32_t* r0 = zs + (size_t)(2 * nzl) * RW + sia;
      r0[0] = pack_h(rh, -ih);
      r0[RW] = pack_h(ih, rh);
      r0[128 * RW] = pack_h(rl, -il);
      r0[129 * RW] = pack_h(il, rl);
    }
    float cs = 1.f;                                       // c of atom sia (CQ)
    if (CQ) cs = MODE == 0 ? (float)(qd * psc) : (float)(psc * (m0 * a.b[6] + m1 * a.b[7] + m2 * a.b[8]));
#pragma unroll
    for (int u = 0; u < XU; ++u) {
      const int cidx = scg + CG * u;
      if (cidx < nxy) {
        cplx<float> v = sok ? xn[u] : cplx<float>{0.f, 0.f};
        if (cidx <= a.Kx) {
          if (CQ) v = cplx<float>{cs * v.re, cs * v.im};
          st[(size_t)cidx * TAP + sia] = v;
        } else if (WIDE) {
          st[(size_t)cidx * TAP + sia] = v;
        } else {
          const int nyc = cidx - a.Kx - 1;
          st[(size_t)(a.Kx + 1 + a.Ky + nyc) * TAP + sia] = v;
          if (nyc > 0) st[(size_t)(a.Kx + 1 + a.Ky - nyc) * TAP + sia] = cplx<float>{v.re, -v.im};
        }
      }
    }
    if (!CQ && threadIdx.x < TA) {
      float c0 = 0.f, c1 = 0.f, c2 = 0.f;
      if ((int)threadIdx.x < nt) {
        if (MODE <= 1) c0 = (float)(qd * psc);
        if (MODE == 1) {
          c1 = (float)(2.0 * MDFF_PI * psc * (m0 * a.b[0] + m1 * a.b[1] + m2 * a.b[2]));
          c2 = (float)(2.0 * MDFF_PI * psc * (m0 * a.b[3] + m1 * a.b[4] + m2 * a.b[5]));
        } else if (MODE == 2) {
          c0 = (float)(psc * (m0 * a.b[6] + m1 * a.b[7] + m2 * a.b[8]));
        }
      }
      sq[threadIdx.x] = c0;
      sq[TAP + threadIdx.x] = c1;
      sq[2 * TAP + threadIdx.x] = c2;
    }
    __syncthreads();
    if (t0 + TA < aend) SKH3_FETCH(t0 + TA);
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[rt][r] = 0.f;
    const float4* qp = reinterpret_cast<const float4*>(sq + 4 * hl);
#pragma unroll
    for (int s = 0; s < TA / 8; ++s) {
      // B: P of atoms 8s + 4hl + u (u < 4) at this lane's m, split
      const float4 x01 = xp[4 * s], x23 = xp[4 * s + 1];
      const float4 y01 = yp[4 * s], y23 = yp[4 * s + 1];
      const float xr[4] = {x01.x, x01.z, x23.x, x23.z}, xi[4] = {x01.y, x01.w, x23.y, x23.w};
      const float yr[4] = {y01.x, y01.z, y23.x, y23.z};
      const float yi[4] = {ysg * y01.y, ysg * y01.w, ysg * y23.y, ysg * y23.w};
      float cq[4] = {0.f, 0.f, 0.f, 0.f}, cx[4] = {0.f, 0.f, 0.f, 0.f}, cy[4] = {0.f, 0.f, 0.f, 0.f};
      if (!CQ) {
        const float4 c0v = qp[2 * s], c1v = qp[TAP / 4 + 2 * s], c2v = qp[TAP / 2 + 2 * s];
        cq[0] = c0v.x; cq[1] = c0v.y; cq[2] = c0v.z; cq[3] = c0v.w;
        cx[0] = c1v.x; cx[1] = c1v.y; cx[2] = c1v.z; cx[3] = c1v.w;
        cy[0] = c2v.x; cy[1] = c2v.y; cy[2] = c2v.z; cy[3] = c2v.w;
      }
      half8 bh, bl;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float xyr = xr[u] * yr[u] - xi[u] * yi[u];
        const float xyi = xr[u] * yi[u] + xi[u] * yr[u];
        float pre = xyr, pim = xyi;                         // CQ: c is in X
        if (!CQ) {
          const float bb = fnx * cx[u] + fny * cy[u];
          pre = cq[u] * xyr - bb * xyi;
          pim = cq[u] * xyi + bb * xyr;
        }
        _Float16 h, l;
        split_h(pre, h, l);
        bh[2 * u] = h;
        bl[2 * u] = l;
        split_h(pim, h, l);
        bh[2 * u + 1] = h;
        bl[2 * u + 1] = l;
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        const uint32_t* row = zs + (size_t)(32 * rt + col) * RW + 8 * s + 4 * hl;
        const half8 ah = __builtin_bit_cast(half8, *reinterpret_cast<const u32x4*>(row));
        const half8 al = __builtin_bit_cast(half8, *reinterpret_cast<const u32x4*>(row + 128 * RW));
        acc[rt] = mma_h(ah, bh, acc[rt]);
        acc[rt] = mma_h(ah, bl, acc[rt]);
        acc[rt] = mma_h(al, bh, acc[rt]);
      }
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) tot[rt][r] = fmaf(tsg, acc[rt][r], tot[rt][r]);
  }
  // D: lane holds rows 8*(r/4) + 4*hl + r%4 of its column m; row R of tile
  // rt -> nz index zoff + (32 rt + R) / 2, part (32 rt + R) % 2
  if (!bok) return;
  float* out = a.part + (size_t)by * 2 * a.nk_box + 2 * (size_t)m * a.NZ;
  const float inv = 1.0f / psc;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int R = 32 * rt + 8 * (r / 4) + 4 * hl + (r % 4);
      const int nzi = zoff + (R >> 1);
      if (nzi < a.NZ) out[2 * nzi + (R & 1)] = tot[rt][r] * inv;
    }
}
#undef SKH3_FETCH

// ---- chunk sums of S and the gather A operand, per m-tile of 8 m values ----
// One block per m-tile (k = m NZ + t for its 8 values of m, every nz): S(k)
// as k_sk_reduce (chunks in order), W = A conj(S), the energy / stress block
// partials, and the split A operand of the gather:
// amat[(mt * NSH + s) * 128 + {0: hi, 1: lo} * 64 + lane] = 8 halves (hi and
// lo planes apart, so a wave's 16-byte reads are consecutive in LDS): row
// l%32 = (m_local = row / 4, comp = row % 4), K = 16 s + 8 (l/32) + j =
// (nz index t, Re/Im cc): comp 0 (Re U0): (wr, -wi), comp 1 (Im U0):
// (wi, wr), comp 2/3 the same times nz.  Values times 2^(14 - e), e the
// exponent of the tile's largest value; ascale[mt] = 2^(e - 14).  The tile is
// assembled in LDS (zeros where t >= NZ or m >= M) and stored coalesced.
template <int NSH, typename TP>
__global__ void __launch_bounds__(256) k_sk_reduce_h3(int nk_box, int nchunk, const TP* __restrict__ part,
                                                      const double* __restrict__ kA,
                                                      const double* __restrict__ kvec, double* __restrict__ sk,
                                                      double* __restrict__ wk, float* __restrict__ wk_t,
                                                      double* __restrict__ blk, int NZ, int Kz,
                                                      u32x4* __restrict__ amat, float* __restrict__ ascale) {
  constexpr int NU = (NSH * 16 * 4 + 255) / 256;           // k entries per thread (8 NZ <= 64 NSH)
  constexpr int NT = NSH * 128 * 4;                        // 32-bit words of the tile
  __shared__ double sh[4 * kRedMaxCols];
  __shared__ uint32_t tile[NT];
  __shared__ float wmax[4];
  const int mt = blockIdx.x;
  const int k0 = 8 * mt * NZ;
  const int kn = min(nk_box, k0 + 8 * NZ) - k0;
  for (int x = threadIdx.x; x < NT; x += 256) tile[x] = 0u;
  double v[kRedMaxCols];
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) v[c] = 0.0;
  // chunk sums in chunk order for the NU entries of the thread; the loads of
  // 16 chunks of every entry are issued before their adds (clamped indices)
  constexpr int CQ = 16;
  double sre[NU], sim[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) { sre[u] = 0.0; sim[u] = 0.0; }
  for (int c = 0; c < nchunk; c += CQ) {
    using TP2 = typename std::conditional<sizeof(TP) == 4, float2, double2>::type;
    TP2 pv[NU][CQ];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int k = k0 + min((int)threadIdx.x + 256 * u, kn - 1);
#pragma unroll
      for (int q = 0; q < CQ; ++q)
        pv[u][q] = *reinterpret_cast<const TP2*>(part + (size_t)min(c + q, nchunk - 1) * 2 * nk_box + 2 * k);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int q = 0; q < CQ; ++q)
        if (c + q < nchunk) {
          sre[u] += (double)pv[u][q].x;
          sim[u] += (double)pv[u][q].y;
        }
  }
  float val[NU][8];
  float vmax = 0.f;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
#pragma unroll
    for (int j = 0; j < 8; ++j) val[u][j] = 0.f;
    const int kl = threadIdx.x + 256 * u;
    if (kl >= kn) continue;
    const int k = k0 + kl;
    const double re = sre[u], im = sim[u];
    sk[2 * k] = re;
    sk[2 * k + 1] = im;
    const double Ak = kA[k];
    const double wr = Ak * re, wi = -Ak * im;
    wk[2 * k] = wr;
    wk[2 * k + 1] = wi;
    wk_t[2 * k] = (float)wr;
    wk_t[2 * k + 1] = (float)wi;
    const double str = (re * re + im * im) * Ak;
    const double kx = kvec[4 * k], ky = kvec[4 * k + 1], kz = kvec[4 * k + 2], kc = kvec[4 * k + 3];
    v[0] += str;
    v[1] += (1.0 - kc * kx * kx) * str;
    v[2] += (1.0 - kc * ky * ky) * str;
    v[3] += (1.0 - kc * kz * kz) * str;
    v[4] += -kc * kx * ky * str;
    v[5] += -kc * kx * kz * str;
    v[6] += -kc * ky * kz * str;
    const double fz = (double)(kl % NZ - Kz);
#pragma unroll
    for (int comp = 0; comp < 4; ++comp)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        double x = ((comp & 1) == 0) ? (cc == 0 ? wr : -wi) : (cc == 0 ? wi : wr);
        if (comp >= 2) x *= fz;
        val[u][2 * comp + cc] = (float)x;
        vmax = fmaxf(vmax, fabsf((float)x));
      }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
  if (lane == 0) wmax[w] = vmax;
  __syncthreads();                                       // also orders the tile zeroing
  vmax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  int e = 0;
  if (vmax > 0.f) (void)frexpf(vmax, &e);                  // vmax in [2^(e-1), 2^e)
  // tiles of vanishing W (A_k far below FP32's normal range at large k):
  // the scale stays finite, their halves flush to zero
  e = max(e, -100);
  const float sc = ldexpf(1.f, 14 - e);
  if (threadIdx.x == 0) ascale[mt] = ldexpf(1.f, e - 14);
  uint16_t* th = reinterpret_cast<uint16_t*>(tile);
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int kl = threadIdx.x + 256 * u;
    if (kl >= kn) continue;
    const int ml = kl / NZ, t = kl - ml * NZ;
#pragma unroll
    for (int comp = 0; comp < 4; ++comp)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        _Float16 h, l;
        split_h(val[u][2 * comp + cc] * sc, h, l);
        const int K = 2 * t + cc, row = 4 * ml + comp;
        const int ln = 32 * ((K >> 3) & 1) + row;
        const int at = ((K >> 4) * 128 + ln) * 8 + (K & 7);
        th[at] = __builtin_bit_cast(uint16_t, h);
        th[at + 64 * 8] = __builtin_bit_cast(uint16_t, l);
      }
  }
  __syncthreads();
  const u32x4* t4 = reinterpret_cast<const u32x4*>(tile);
  for (int x = threadIdx.x; x < NSH * 128; x += 256) amat[(size_t)mt * NSH * 128 + x] = t4[x];
  const double r = block_sum_cols(v, 7, sh);
  if (threadIdx.x < 7) blk[blockIdx.x * 7 + threadIdx.x] = r;
}

// ---- gather: U(m, i) per m-tile of 8 m values and atom tiles of 32 ----
// D(rows (m_local, comp), atoms) = sum_K A(rows, K) B(K, atoms), B = the
// atom's split Z row (registers, whole kernel), A staged per m-tile in LDS
// (double buffered, shared by the block's 4 waves, which run the same
// m-tiles).  Epilogue as k_gather_mma: Im(X Y U0) weighted by nx, ny and
// Im(X Y U1), accumulated in FP64.
struct Gh3Args {
  int n, Kx, Ky, Kz, NY, NZ, M, mt_per_split, mtiles;
  const void* tab;               // FP32 phase tables
  const u32x4* amat;
  const float* ascale;
  double* out;                   // [split][3][n]
};

// word 3 of a raw buffer resource on gfx950 (32-bit data, no swizzle)
constexpr int kBufWord3 = 0x00020000;

template <int CT, int NSH>
constexpr int gather_h3_waves() {
  return CT * NSH <= 14 ? 2 : 1;        // two waves per SIMD where the split B rows fit
}

template <int CT, int NSH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gather_h3_waves<CT, NSH>())))
k_gather_h3(Gh3Args a) {
  __shared__ u32x4 as[2][NSH * 128];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hl = lane >> 5, col = lane & 31;
  const int atom0 = (blockIdx.x * 4 + wave) * 32 * CT;
  const cplx<float>* tab = reinterpret_cast<const cplx<float>*>(a.tab);
  const int zcol = a.Kx + 1 + a.Ky + 1;
  int iac[CT];
  bool aok[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    aok[ct] = atom0 + 32 * ct + col < a.n;
    iac[ct] = min(atom0 + 32 * ct + col, a.n - 1);
  }
  // B: split Z rows (K = 16 s + 8 hl + j = (t, cc): Re / Im of Z(t - Kz))
  half8 bh[CT][NSH], bl[CT][NSH];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int s = 0; s < NSH; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int kk = 16 * s + 8 * hl + j;
        const int t = kk >> 1, cc = kk & 1;
        const int nz = t - a.Kz;
        const int az = min(nz < 0 ? -nz : nz, a.Kz);
        const float v = reinterpret_cast<const float*>(tab + (size_t)(zcol + az) * a.n + iac[ct])[cc];
        const float w = (aok[ct] && t < a.NZ) ? ((cc && nz < 0) ? -1.f : 1.f) : 0.f;
        _Float16 h, l;
        split_h(w * v, h, l);
        bh[ct][s][j] = h;
        bl[ct][s][j] = l;
      }
  double ax[CT], ay[CT], az3[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) { ax[ct] = 0.0; ay[ct] = 0.0; az3[ct] = 0.0; }
  const int mt0 = blockIdx.y * a.mt_per_split;
  const int mt1 = min(a.mtiles, mt0 + a.mt_per_split);
  constexpr int NA = NSH * 128;                       // 16-byte units per staged m-tile
  constexpr int PER = (NA + 255) / 256;
  // staging of the A operand: thread -> unit x = threadIdx.x + 256 u (loads
  // from clamped indices, stores only below NA)
  u32x4 pre[PER];
  if (mt0 < mt1) {
#pragma unroll
    for (int u = 0; u < PER; ++u) pre[u] = a.amat[(size_t)mt0 * NA + min((int)threadIdx.x + 256 * u, NA - 1)];
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if ((int)threadIdx.x + 256 * u < NA) as[0][threadIdx.x + 256 * u] = pre[u];
  }
  __syncthreads();
  const float inv_ny = 1.0f / (float)a.NY;
  // the phase tables as a buffer resource: 32-bit byte offsets, no 64-bit
  // address arithmetic per load (tables below 2 GiB: checked at set-up)
  const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void*>(a.tab), (short)0, (int)(8 * (size_t)(a.Kx + a.Ky + a.Kz + 3) * a.n), kBufWord3);
  const int rowb = 8 * a.n;                           // bytes per table column
  int ioff[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) ioff[ct] = 8 * iac[ct];
  for (int mt = mt0; mt < mt1; ++mt) {
    const int buf = (mt - mt0) & 1;
    if (mt + 1 < mt1) {
#pragma unroll
      for (int u = 0; u < PER; ++u)
        pre[u] = a.amat[(size_t)(mt + 1) * NA + min((int)threadIdx.x + 256 * u, NA - 1)];
    }
    // epilogue operands of this tile (used after the MFMAs): the lane's m
    // values 8 mt + 2 q + hl, q < 4, clamped below M (the A rows of m >= M
    // are zero, so those D rows are zero)
    cplx<float> Xv[CT][4], Yv[CT][4];
    float fx[4], fy[4], ysg[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int mq = 8 * mt + 2 * q + hl;
      int nxx, nyo;
      divmod_f(min(mq, a.M - 1), a.NY, inv_ny, nxx, nyo);
      const int ny = nyo - a.Ky;
      fx[q] = (float)nxx;
      fy[q] = (float)ny;
      ysg[q] = ny < 0 ? -1.f : 1.f;
      const int xrow = nxx * rowb;
      const int yrow = (a.Kx + 1 + (ny >= 0 ? ny : -ny)) * rowb;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        Xv[ct][q] = __builtin_bit_cast(cplx<float>, __builtin_amdgcn_raw_buffer_load_b64(trs, xrow + ioff[ct], 0, 0));
        Yv[ct][q] = __builtin_bit_cast(cplx<float>, __builtin_amdgcn_raw_buffer_load_b64(trs, yrow + ioff[ct], 0, 0));
      }
    }
    f32x16 acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[ct][r] = 0.f;
#pragma unroll
    for (int s = 0; s < NSH; ++s) {
      const half8 ah = __builtin_bit_cast(half8, as[buf][s * 128 + lane]);
      const half8 al = __builtin_bit_cast(half8, as[buf][s * 128 + 64 + lane]);
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        acc[ct] = mma_h(ah, bh[ct][s], acc[ct]);
        acc[ct] = mma_h(ah, bl[ct][s], acc[ct]);
        acc[ct] = mma_h(al, bh[ct][s], acc[ct]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // Im(X Y U0) weighted by nx, ny and Im(X Y U1) over the lane's 4 m
    // values in FP32, times the tile's inverse scale, then added in FP64
    const float inv = a.ascale[mt];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cplx<float> Y = Yv[ct][q];
        Y.im *= ysg[q];
        const cplx<float> XY = cmul(Xv[ct][q], Y);
        const float im0 = XY.re * acc[ct][4 * q + 1] + XY.im * acc[ct][4 * q];
        const float im1 = XY.re * acc[ct][4 * q + 3] + XY.im * acc[ct][4 * q + 2];
        sx += fx[q] * im0;
        sy += fy[q] * im0;
        sz += im1;
      }
      ax[ct] += (double)(sx * inv);
      ay[ct] += (double)(sy * inv);
      az3[ct] += (double)(sz * inv);
    }
    if (mt + 1 < mt1) {
#pragma unroll
      for (int u = 0; u < PER; ++u)
        if ((int)threadIdx.x + 256 * u < NA) as[buf ^ 1][threadIdx.x + 256 * u] = pre[u];
    }
    __syncthreads();
  }
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    ax[ct] += __shfl_xor(ax[ct], 32);
    ay[ct] += __shfl_xor(ay[ct], 32);
    az3[ct] += __shfl_xor(az3[ct], 32);
  }
  if (hl == 0) {
    double* o = a.out + (size_t)blockIdx.y * 3 * a.n;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
      if (aok[ct]) {
        const int ia = atom0 + 32 * ct + col;
        o[ia] = ax[ct];
        o[(size_t)a.n + ia] = ay[ct];
        o[(size_t)2 * a.n + ia] = az3[ct];
      }
  }
}

// ---------------------------------------------------------------------------
// Inner k-box in FP64 (mixed form, charge-only evaluations).
//
// The weight of a k-vector in the force sums, |k| A(k), falls off as
// exp(-k^2 / 4 alpha^2) / |k|, so the few k-vectors of small |k| carry the
// large, mutually cancelling contributions; their FP32-class structure
// factors (24-bit accumulation over thousands of atoms) set the mixed form's
// force error.  The box |nx| <= Kx', |ny| <= Ky', |nz| <= Kz' around k = 0
// (nx >= 0, entries outside the half space and k = 0 carry A = 0) is summed
// here in FP64 and masked out of the main kernels (A = 0 there): per k the
// same terms as the main path, so the split changes which precision sums
// a k-vector, not the sum.
// Generic forms (any Kz'; the scalar-operand forms below replace them when
// Kz' <= kInKZmax and the (nx, ny) columns fit one block):
//   k_in_sk      S(k) partials per chunk of atoms, one thread per k, FP64
//                phase tables of the chunk's atoms staged in LDS
//   k_in_reduce  S(k) over the chunks in order; W = A conj(S); the energy
//                and stress block partials (appended to the main blocks)
//   k_in_gather  one thread per atom: the moments sum_k n_d Im(W e^{ik.r}),
//                written as one more split of the gather's partials
// ---------------------------------------------------------------------------
struct InArgs {
  int n, Kx, Ky, Kz, NY, NZ, nb, chunk, tile, msplit, ts;
  int zpad;                       // scalar-operand forms: Z columns of a phase row, j entries of a W table row
  const double *sx, *sy, *sz, *q;
  int qstride;                    // the per-atom weight is q[i * qstride] (3: a dipole component)
  const double* kA;               // A of the inner box (nb)
  const double* kvec;             // kx, ky, kz, kcoe (4 nb)
  double* part;                   // [chunk][2 nb]
  double* wk;                     // 2 nb
  double* blk;                    // energy / stress block partials [block][7]
  double* out;                    // gather moments [split][3][n], one split per m range
  cplx<double>* tab;              // per atom: X(0..Kx'), Y(0..Ky'), q Z(0..Kz')
  double* wt;                     // per (m, j = 0..Kz'): sums / differences of W(+-j)
};

// e^{2 pi i k s} for k = 0..K into dst[0..K] (stride st): sincospi of the
// reduced angle for k = 1, then products (each within a few units of 2^-53)
// (times w: the products are formed unscaled, each entry scaled once)
__device__ __forceinline__ void phase_row(double s, int K, cplx<double>* dst, int st, double w = 1.0) {
  double t1 = s - rint(s);
  double s1, c1;
  sincospi(2.0 * t1, &s1, &c1);
  double cs = 1.0, sn = 0.0;
  dst[0] = {w, 0.0};
  for (int k = 1; k <= K; ++k) {
    const double c2 = cs * c1 - sn * s1, s2 = cs * s1 + sn * c1;
    cs = c2;
    sn = s2;
    dst[(size_t)k * st] = {w * cs, w * sn};
  }
}

// grid (k blocks of 256, atom chunks); the chunk's atoms pass through LDS
// in tiles of a.tile atoms (FP64 tables and charges)
__global__ void __launch_bounds__(256) k_in_sk(InArgs a) {
  extern __shared__ unsigned char smem_raw[];
  const int ncol = a.Kx + a.Ky + a.Kz + 3;
  cplx<double>* tb = reinterpret_cast<cplx<double>*>(smem_raw);          // [atom][ncol]
  double* qs = reinterpret_cast<double*>(tb + (size_t)a.tile * ncol);    // [atom]
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = k < a.nb;
  const int kk = act ? k : 0;
  const int nx = kk / (a.NY * a.NZ);
  const int ny = (kk / a.NZ) % a.NY - a.Ky;
  const int nz = kk % a.NZ - a.Kz;
  const int cx = nx, cy = a.Kx + 1 + (ny < 0 ? -ny : ny), cz = a.Kx + a.Ky + 2 + (nz < 0 ? -nz : nz);
  const double sy = ny < 0 ? -1.0 : 1.0, sz = nz < 0 ? -1.0 : 1.0;   // negative n: conjugate
  double re = 0.0, im = 0.0;
  const int c0 = blockIdx.y * a.chunk, c1 = min(a.n, c0 + a.chunk);
  for (int t0 = c0; t0 < c1; t0 += a.tile) {
    const int na = min(a.tile, c1 - t0);
    __syncthreads();
    // one (atom, direction) row per thread
    for (int r = threadIdx.x; r < 3 * na; r += blockDim.x) {
      const int at = r / 3, d = r - 3 * (r / 3);
      const int i = t0 + at;
      const double sv = d == 0 ? a.sx[i] : (d == 1 ? a.sy[i] : a.sz[i]);
      const int K = d == 0 ? a.Kx : (d == 1 ? a.Ky : a.Kz);
      const int col = d == 0 ? 0 : (d == 1 ? a.Kx + 1 : a.Kx + a.Ky + 2);
      phase_row(sv, K, tb + (size_t)at * ncol + col, 1);
      if (d == 0) qs[at] = a.q[i];
    }
    __syncthreads();
    if (act) {
      for (int at = 0; at < na; ++at) {
        const cplx<double>* row = tb + (size_t)at * ncol;
        const cplx<double> X = row[cx];
        cplx<double> Y = row[cy], Z = row[cz];
        Y.im *= sy;
        Z.im *= sz;
        const cplx<double> e = cmul(cmul(X, Y), Z);
        const double qa = qs[at];
        re += qa * e.re;
        im += qa * e.im;
      }
    }
  }
  if (!act) return;
  a.part[(size_t)blockIdx.y * 2 * a.nb + 2 * k] = re;
  a.part[(size_t)blockIdx.y * 2 * a.nb + 2 * k + 1] = im;
}

// 64 k-vectors per block: wave w sums the chunks [w q, (w + 1) q) in order,
// the four wave sums are added in wave order; then W = A conj(S) and the
// energy / stress block partials
__global__ void __launch_bounds__(256) k_in_reduce(InArgs a, int nchunk) {
  __shared__ double sh[4 * kRedMaxCols];
  __shared__ double ws[4][64][2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + lane;
  const int q = (nchunk + 3) / 4;
  const int c0 = w * q, c1 = min(nchunk, c0 + q);
  double re = 0.0, im = 0.0;
  if (k < a.nb) {
    int c = c0;
    for (; c + 4 <= c1; c += 4) {
      double vr[4], vi[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        vr[u] = a.part[(size_t)(c + u) * 2 * a.nb + 2 * k];
        vi[u] = a.part[(size_t)(c + u) * 2 * a.nb + 2 * k + 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        re += vr[u];
        im += vi[u];
      }
    }
    for (; c < c1; ++c) {
      re += a.part[(size_t)c * 2 * a.nb + 2 * k];
      im += a.part[(size_t)c * 2 * a.nb + 2 * k + 1];
    }
  }
  ws[w][lane][0] = re;
  ws[w][lane][1] = im;
  __syncthreads();
  double v[kRedMaxCols];
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) v[c] = 0.0;
  if (w == 0 && k < a.nb) {
    re = ((ws[0][lane][0] + ws[1][lane][0]) + ws[2][lane][0]) + ws[3][lane][0];
    im = ((ws[0][lane][1] + ws[1][lane][1]) + ws[2][lane][1]) + ws[3][lane][1];
    const double Ak = a.kA[k];
    a.wk[2 * k] = Ak * re;
    a.wk[2 * k + 1] = -Ak * im;
    const double str = (re * re + im * im) * Ak;
    const double kx = a.kvec[4 * k], ky = a.kvec[4 * k + 1], kz = a.kvec[4 * k + 2], kc = a.kvec[4 * k + 3];
    v[0] = str;
    v[1] = (1.0 - kc * kx * kx) * str;
    v[2] = (1.0 - kc * ky * ky) * str;
    v[3] = (1.0 - kc * kz * kz) * str;
    v[4] = -kc * kx * ky * str;
    v[5] = -kc * kx * kz * str;
    v[6] = -kc * ky * kz * str;
  }
  const double r = block_sum_cols(v, 7, sh);
  if (threadIdx.x < 7) a.blk[blockIdx.x * 7 + threadIdx.x] = r;
}

// Forms for boxes with Kz' <= kInKZmax, where the operands shared by a wave
// come through scalar loads rather than LDS.  The nz range is taken in
// blocks of KZ + 1 values of j = |nz| (KZ <= kInKZ a template argument;
// block b holds j = b (KZ + 1) + 0..KZ, the last block's surplus j masked),
// one grid row per block (k_in_sk_s: blockIdx.y, k_in_gather_s:
// blockIdx.z); boxes with Kz' <= kInKZ are one block with KZ = Kz':
//   k_in_phase     per atom the FP64 rows X(0..Kx'), Y(0..Ky'), q Z(0..Kz')
//   k_in_sk_s      one thread per m = (nx, ny) and subgroup of atoms, the
//                  S(m, +-nz) sums in registers; q Z of the wave's atom is
//                  wave-uniform (scalar loads), X and Y from LDS; the
//                  subgroups' sums added in subgroup order through LDS
//   k_in_reduce_r  whole rows m per block: S over the chunks, W, the
//                  energy / stress partials, and per (m, j) the sums and
//                  differences of W(m, +j) and W(m, -j) that the gather reads
//   k_in_gather_s  one thread per atom, Z(0..Kz') in registers, the (m, j)
//                  table through scalar loads
constexpr int kInKZ = 8;       // largest KZ: the two operand sets' q Z in scalar registers
constexpr int kInKZmax = 31;   // NZ = 2 Kz' + 1 <= 64: one row m per k_in_reduce_r wave

__global__ void __launch_bounds__(256) k_in_phase(InArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * a.n) return;
  const int i = r / 3, d = r - 3 * (r / 3);
  const int ncol = a.Kx + a.Ky + 2 + a.zpad;
  const double sv = d == 0 ? a.sx[i] : (d == 1 ? a.sy[i] : a.sz[i]);
  const int K = d == 0 ? a.Kx : (d == 1 ? a.Ky : a.Kz);
  const int col = d == 0 ? 0 : (d == 1 ? a.Kx + 1 : a.Kx + a.Ky + 2);
  phase_row(sv, K, a.tab + (size_t)i * ncol + col, 1, d == 2 ? a.q[(size_t)i * a.qstride] : 1.0);
}

// grid: one block per chunk of atoms; block: SG subgroups of a.ts threads
// (a.ts a multiple of 64, >= M); subgroup g takes atoms g, g + SG, ... of
// each tile
// one atom's operands of k_in_sk_s: q Z(0..KZ) (wave-uniform: scalar loads)
// and X(nx), Y(|ny|) from the LDS tile
template <int KZ>
__device__ __forceinline__ void in_sk_load(const cplx<double>* __restrict__ tab, size_t zrow,
                                           const cplx<double>* xy, int xrow, int nx, int cy,
                                           cplx<double> (&z)[KZ + 1], cplx<double>& X, cplx<double>& Y) {
#pragma unroll
  for (int j = 0; j <= KZ; ++j) z[j] = tab[zrow + j];
  X = xy[xrow + nx];
  Y = xy[xrow + cy];
}

// one atom's terms: P = X Y (Y conjugated for ny < 0) and the block's first
// j: S(m, +j0) += P q Z(j0), and S(m, -j0) += P q conj Z(j0) unless j0 = 0
template <int KZ>
__device__ __forceinline__ cplx<double> in_sk_first(const cplx<double>& X, cplx<double> Y, double sy,
                                                    const cplx<double> (&z)[KZ + 1], bool j0pos,
                                                    double (&pr)[KZ + 1], double (&pi)[KZ + 1],
                                                    double (&qr)[KZ + 1], double (&qi)[KZ + 1]) {
  Y.im *= sy;
  const cplx<double> P = cmul(X, Y);
  pr[0] = fma(P.re, z[0].re, pr[0]);
  pr[0] = fma(-P.im, z[0].im, pr[0]);
  pi[0] = fma(P.re, z[0].im, pi[0]);
  pi[0] = fma(P.im, z[0].re, pi[0]);
  if (j0pos) {
    qr[0] = fma(P.re, z[0].re, qr[0]);
    qr[0] = fma(P.im, z[0].im, qr[0]);
    qi[0] = fma(P.im, z[0].re, qi[0]);
    qi[0] = fma(-P.re, z[0].im, qi[0]);
  }
  return P;
}

// then S(m, +j) and S(m, -j) += P q Z(+-j), j = 1..KZ
template <int KZ>
__device__ __forceinline__ void in_sk_rest(const cplx<double>& P, const cplx<double> (&z)[KZ + 1],
                                           double (&pr)[KZ + 1], double (&pi)[KZ + 1], double (&qr)[KZ + 1],
                                           double (&qi)[KZ + 1]) {
#pragma unroll
  for (int j = 1; j <= KZ; ++j) {
    const double zr = z[j].re, zi = z[j].im;
    pr[j] = fma(P.re, zr, pr[j]);
    pr[j] = fma(-P.im, zi, pr[j]);
    pi[j] = fma(P.re, zi, pi[j]);
    pi[j] = fma(P.im, zr, pi[j]);
    qr[j] = fma(P.re, zr, qr[j]);
    qr[j] = fma(P.im, zi, qr[j]);
    qi[j] = fma(P.im, zr, qi[j]);
    qi[j] = fma(-P.re, zi, qi[j]);
  }
}

template <int KZ, bool MB>
__global__ void __launch_bounds__(1024) k_in_sk_s(InArgs a) {
  extern __shared__ unsigned char smem_raw[];
  const int nxy = a.Kx + a.Ky + 2;
  const int ncol = nxy + a.zpad;                                 // Z(0..Kz') and zero padding
  const int j0 = MB ? blockIdx.y * (KZ + 1) : 0;                 // this block's first j
  const bool j0pos = MB && j0 > 0;
  cplx<double>* xy = reinterpret_cast<cplx<double>*>(smem_raw);   // [tile][nxy]
  const int M = (a.Kx + 1) * a.NY;
  const int SG = blockDim.x / a.ts;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / a.ts);
  const int m = threadIdx.x - g * a.ts;
  const int mm = m < M ? m : 0;
  const int nx = mm / a.NY, ny = mm - (mm / a.NY) * a.NY - a.Ky;
  const int cy = a.Kx + 1 + (ny < 0 ? -ny : ny);
  const double sy = ny < 0 ? -1.0 : 1.0;
  const cplx<double>* __restrict__ tab = a.tab;
  double pr[KZ + 1], pi[KZ + 1], qr[KZ + 1], qi[KZ + 1];   // nz = +j, nz = -j
#pragma unroll
  for (int j = 0; j <= KZ; ++j) pr[j] = pi[j] = qr[j] = qi[j] = 0.0;
  const int c0 = blockIdx.x * a.chunk, c1 = min(a.n, c0 + a.chunk);
  for (int t0 = c0; t0 < c1; t0 += a.tile) {
    const int na = min(a.tile, c1 - t0);
    __syncthreads();
    for (int x = threadIdx.x; x < na * nxy; x += blockDim.x) {
      const int at = x / nxy;
      xy[x] = tab[(size_t)(t0 + at) * ncol + (x - at * nxy)];
    }
    __syncthreads();
    // two operand sets, A and B, alternate between atoms.  Each atom's
    // first terms (P = X Y and nz = 0) come before the next atom's loads
    // (q Z by scalar loads, X and Y from LDS) issue, so the wait for its own
    // operands, which were issued one atom earlier, falls before them; the
    // next atom's operands then land while the rest of its terms are summed.
    // (Scalar loads return out of order: a wait on them is a full wait.)  An
    // atom index past the tile loads the last atom's row again (in range,
    // unused).
    cplx<double> zA[KZ + 1], zB[KZ + 1];
    cplx<double> XA{0.0, 0.0}, YA{0.0, 0.0}, XB{0.0, 0.0}, YB{0.0, 0.0};
    if (g < na) in_sk_load<KZ>(tab, (size_t)(t0 + g) * ncol + nxy + j0, xy, g * nxy, nx, cy, zA, XA, YA);
    for (int at = g; at < na; at += 2 * SG) {
      const int b = at + SG;
      {
        const cplx<double> P = in_sk_first<KZ>(XA, YA, sy, zA, j0pos, pr, pi, qr, qi);
        __builtin_amdgcn_sched_barrier(0);
        const int an = __builtin_amdgcn_readfirstlane(min(b, na - 1));
        in_sk_load<KZ>(tab, (size_t)(t0 + an) * ncol + nxy + j0, xy, an * nxy, nx, cy, zB, XB, YB);
        __builtin_amdgcn_sched_barrier(0);
        in_sk_rest<KZ>(P, zA, pr, pi, qr, qi);
      }
      if (b < na) {
        const cplx<double> P = in_sk_first<KZ>(XB, YB, sy, zB, j0pos, pr, pi, qr, qi);
        __builtin_amdgcn_sched_barrier(0);
        const int an = __builtin_amdgcn_readfirstlane(min(b + SG, na - 1));
        in_sk_load<KZ>(tab, (size_t)(t0 + an) * ncol + nxy + j0, xy, an * nxy, nx, cy, zA, XA, YA);
        __builtin_amdgcn_sched_barrier(0);
        in_sk_rest<KZ>(P, zB, pr, pi, qr, qi);
      }
    }
  }
  // subgroups 1..SG-1 into subgroup 0, in order, one j at a time
  double* red = reinterpret_cast<double*>(smem_raw);            // [SG - 1][ts][4]
#pragma unroll
  for (int j = 0; j <= KZ; ++j) {
    __syncthreads();
    if (g > 0) {
      double* r = red + ((size_t)(g - 1) * a.ts + m) * 4;
      r[0] = pr[j]; r[1] = pi[j]; r[2] = qr[j]; r[3] = qi[j];
    }
    __syncthreads();
    if (g == 0) {
      for (int h = 1; h < SG; ++h) {
        const double* r = red + ((size_t)(h - 1) * a.ts + m) * 4;
        pr[j] += r[0]; pi[j] += r[1]; qr[j] += r[2]; qi[j] += r[3];
      }
    }
  }
  if (g != 0 || m >= M) return;
  double* out = a.part + (size_t)blockIdx.x * 2 * a.nb + 2 * (size_t)m * a.NZ;
#pragma unroll
  for (int jj = 0; jj <= KZ; ++jj) {
    const int j = j0 + jj;
    if (j > a.Kz) break;
    out[2 * (a.Kz + j)] = pr[jj];
    out[2 * (a.Kz + j) + 1] = pi[jj];
    if (j > 0) {
      out[2 * (a.Kz - j)] = qr[jj];
      out[2 * (a.Kz - j) + 1] = qi[jj];
    }
  }
}

// R = floor(64 / NZ) rows m per block, lane l the k = m0 NZ + l; wave w of
// 16 sums the chunks [w q, (w + 1) q) in order, the 16 sums are added in
// wave order
__global__ void __launch_bounds__(1024) k_in_reduce_r(InArgs a, int nchunk) {
  __shared__ double ws[16][64][2];
  __shared__ double wl[64][2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R = 64 / a.NZ;
  const int M = (a.Kx + 1) * a.NY;
  const int m0 = blockIdx.x * R;
  const int rows = min(R, M - m0);
  const bool act = lane < rows * a.NZ;
  const int k = m0 * a.NZ + (act ? lane : 0);
  const int q = (nchunk + 15) / 16;
  const int c0 = w * q, c1 = min(nchunk, c0 + q);
  double re = 0.0, im = 0.0;
  if (act) {
    int c = c0;
    for (; c + 4 <= c1; c += 4) {
      double vr[4], vi[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        vr[u] = a.part[(size_t)(c + u) * 2 * a.nb + 2 * k];
        vi[u] = a.part[(size_t)(c + u) * 2 * a.nb + 2 * k + 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        re += vr[u];
        im += vi[u];
      }
    }
    for (; c < c1; ++c) {
      re += a.part[(size_t)c * 2 * a.nb + 2 * k];
      im += a.part[(size_t)c * 2 * a.nb + 2 * k + 1];
    }
  }
  ws[w][lane][0] = re;
  ws[w][lane][1] = im;
  __syncthreads();
  if (w == 0) {
    re = 0.0;
    im = 0.0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      re += ws[v][lane][0];
      im += ws[v][lane][1];
    }
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double wr = 0.0, wi = 0.0;
    if (act) {
      const double Ak = a.kA[k];
      wr = Ak * re;
      wi = -Ak * im;
      a.wk[2 * k] = wr;
      a.wk[2 * k + 1] = wi;
      const double str = (re * re + im * im) * Ak;
      const double kx = a.kvec[4 * k], ky = a.kvec[4 * k + 1], kz = a.kvec[4 * k + 2], kc = a.kvec[4 * k + 3];
      v[0] = str;
      v[1] = (1.0 - kc * kx * kx) * str;
      v[2] = (1.0 - kc * ky * ky) * str;
      v[3] = (1.0 - kc * kz * kz) * str;
      v[4] = -kc * kx * ky * str;
      v[5] = -kc * kx * kz * str;
      v[6] = -kc * ky * kz * str;
    }
    wl[lane][0] = wr;
    wl[lane][1] = wi;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      double x = v[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
      if (lane == 0) a.blk[blockIdx.x * 7 + c] = x;
    }
  }
  __syncthreads();
  // (m, j): W(+j) + W(-j) and W(+j) - W(-j); j = 0: W(0)
  const int nj = a.Kz + 1;
  const int t = threadIdx.x;
  if (t < rows * nj) {
    const int r = t / nj, j = t - (t / nj) * nj;
    double* o = a.wt + ((size_t)(m0 + r) * a.zpad + j) * 4;
    const double* p = wl[r * a.NZ + a.Kz + j];
    const double* n = wl[r * a.NZ + a.Kz - j];
    if (j == 0) {
      o[0] = p[0]; o[1] = p[1]; o[2] = 0.0; o[3] = 0.0;
    } else {
      o[0] = p[0] + n[0]; o[1] = p[1] + n[1]; o[2] = p[0] - n[0]; o[3] = p[1] - n[1];
    }
  }
}

// one thread per atom and one range of m = (nx, ny) values per blockIdx.y
// (a.msplit values each; the ranges' moments are separate splits of the
// output): W of the range through LDS in tiles of kInMT values of m
// (wave-uniform reads); X, Y and Z of the atom by recurrence from
// e^{2 pi i s} (each product within a few units of 2^-53)
constexpr int kInMT = 8;
__global__ void __launch_bounds__(256) k_in_gather(InArgs a) {
  extern __shared__ unsigned char smem_raw[];
  cplx<double>* ws = reinterpret_cast<cplx<double>*>(smem_raw);   // [kInMT][NZ]
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < a.n;
  const int ii = act ? i : 0;
  double x1r, x1i, y1r, y1i, z1r, z1i;
  {
    double t = a.sx[ii] - rint(a.sx[ii]);
    sincospi(2.0 * t, &x1i, &x1r);
    t = a.sy[ii] - rint(a.sy[ii]);
    sincospi(2.0 * t, &y1i, &y1r);
    t = a.sz[ii] - rint(a.sz[ii]);
    sincospi(2.0 * t, &z1i, &z1r);
  }
  const cplx<double> x1{x1r, x1i}, y1{y1r, y1i}, zz1{z1r, z1i};
  // Y(-Ky) and Z(-Kz): conjugate powers
  cplx<double> y0{1.0, 0.0}, z0{1.0, 0.0};
  for (int k = 0; k < a.Ky; ++k) y0 = cmul(y0, cplx<double>{y1r, -y1i});
  for (int k = 0; k < a.Kz; ++k) z0 = cmul(z0, cplx<double>{z1r, -z1i});
  const int M = (a.Kx + 1) * a.NY;
  const int mb = blockIdx.y * a.msplit, me = min(M, mb + a.msplit);
  // X and Y at the first m of the range
  cplx<double> X{1.0, 0.0}, Y = y0;
  {
    const int nx0 = mb / a.NY, iy0 = mb - (mb / a.NY) * a.NY;
    for (int k = 0; k < nx0; ++k) X = cmul(X, x1);
    for (int k = 0; k < iy0; ++k) Y = cmul(Y, y1);
  }
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int mt = mb; mt < me; mt += kInMT) {
    const int nm = min(kInMT, me - mt);
    __syncthreads();
    for (int x = threadIdx.x; x < nm * a.NZ; x += blockDim.x)
      ws[x] = {a.wk[2 * ((size_t)mt * a.NZ + x)], a.wk[2 * ((size_t)mt * a.NZ + x) + 1]};
    __syncthreads();
    for (int mm = 0; mm < nm; ++mm) {
      const int m = mt + mm;
      const int nx = m / a.NY, ny = m - (m / a.NY) * a.NY - a.Ky;
      const cplx<double> XY = cmul(X, Y);
      const cplx<double>* wrow = ws + (size_t)mm * a.NZ;
      double s0r = 0.0, s0i = 0.0, s1r = 0.0, s1i = 0.0;
      cplx<double> Z = z0;
      for (int t = 0; t < a.NZ; ++t) {
        const cplx<double> w = wrow[t];
        const double tr = w.re * Z.re - w.im * Z.im;
        const double ti = w.re * Z.im + w.im * Z.re;
        const double fz = (double)(t - a.Kz);
        s0r += tr; s0i += ti;
        s1r += fz * tr; s1i += fz * ti;
        Z = cmul(Z, zz1);
      }
      const double i0 = XY.re * s0i + XY.im * s0r;
      const double i1 = XY.re * s1i + XY.im * s1r;
      u0 += (double)nx * i0;
      u1 += (double)ny * i0;
      u2 += i1;
      // next m: ny + 1, or the next nx with ny = -Ky
      if (ny == a.Ky) {
        X = cmul(X, x1);
        Y = y0;
      } else {
        Y = cmul(Y, y1);
      }
    }
  }
  if (!act) return;
  double* o = a.out + (size_t)blockIdx.y * 3 * a.n;
  o[i] = u0;
  o[(size_t)a.n + i] = u1;
  o[(size_t)2 * a.n + i] = u2;
}

// one thread per atom, one range of a.msplit values of m per blockIdx.y;
// per m: s0 = sum_nz W Z, s1 = sum_nz nz W Z from the (m, j) table,
// Z(+j) = z_j, Z(-j) = conj(z_j):
//   s0 = W(0) + sum_j [(Sr c - Di s) + i (Dr s + Si c)]
//   s1 = sum_j [(Dr jc - Si js) + i (Sr js + Di jc)]
// with S = W(+j) + W(-j), D = W(+j) - W(-j), z_j = c + i s, (jc, js) = j z_j
template <int KZ, int MOM, bool MB>
__global__ void __launch_bounds__(256) k_in_gather_s(InArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < a.n;
  const int ii = act ? i : 0;
  // this block's j = j0 + 0..KZ (blockIdx.z); the W table's entries of the
  // surplus j beyond Kz' are zero padding
  const int j0 = MB ? blockIdx.z * (KZ + 1) : 0;
  double x1r, x1i, y1r, y1i;
  double zc[KZ + 1], zs[KZ + 1];
  {
    double t = a.sx[ii] - rint(a.sx[ii]);
    sincospi(2.0 * t, &x1i, &x1r);
    t = a.sy[ii] - rint(a.sy[ii]);
    sincospi(2.0 * t, &y1i, &y1r);
    double s1, c1;
    t = a.sz[ii] - rint(a.sz[ii]);
    sincospi(2.0 * t, &s1, &c1);
    if (j0 == 0) {
      zc[0] = 1.0;
      zs[0] = 0.0;
    } else {
      sincospi(2.0 * (double)j0 * t, &zs[0], &zc[0]);
    }
#pragma unroll
    for (int k = 1; k <= KZ; ++k) {
      zc[k] = zc[k - 1] * c1 - zs[k - 1] * s1;
      zs[k] = zc[k - 1] * s1 + zs[k - 1] * c1;
    }
  }
  double jc[KZ + 1], js[KZ + 1];          // j z_j
#pragma unroll
  for (int k = 0; k <= KZ; ++k) {
    jc[k] = (double)(j0 + k) * zc[k];
    js[k] = (double)(j0 + k) * zs[k];
  }
  const cplx<double> x1{x1r, x1i}, y1{y1r, y1i};
  cplx<double> y0{1.0, 0.0};
  for (int k = 0; k < a.Ky; ++k) y0 = cmul(y0, cplx<double>{y1r, -y1i});
  const int M = (a.Kx + 1) * a.NY;
  const int mb = blockIdx.y * a.msplit, me = min(M, mb + a.msplit);
  cplx<double> X{1.0, 0.0}, Y = y0;
  {
    const int nx0 = mb / a.NY, iy0 = mb - (mb / a.NY) * a.NY;
    for (int k = 0; k < nx0; ++k) X = cmul(X, x1);
    for (int k = 0; k < iy0; ++k) Y = cmul(Y, y1);
  }
  const double* __restrict__ wt = a.wt;
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  // MOM 9: the second moments n_a n_b Re(W e^{ik.r}) as well (xx yy zz xy
  // xz yz, the dipole force terms), with the nz^2-weighted sum s2
  double mxx = 0.0, myy = 0.0, mzz = 0.0, mxy = 0.0, mxz = 0.0, myz = 0.0;
  for (int m = mb; m < me; ++m) {
    const int nx = m / a.NY, ny = m - (m / a.NY) * a.NY - a.Ky;
    const double* e = wt + ((size_t)m * a.zpad + j0) * 4;
    double s0r = 0.0, s0i = 0.0, s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0;
    if (j0 == 0) {
      // W(0)
      s0r = e[0];
      s0i = e[1];
    }
#pragma unroll
    for (int j = 0; j <= KZ; ++j) {
      if (j == 0 && j0 == 0) continue;
      const double sr = e[4 * j], si = e[4 * j + 1], dr = e[4 * j + 2], di = e[4 * j + 3];
      s0r = fma(sr, zc[j], s0r);
      s0r = fma(-di, zs[j], s0r);
      s0i = fma(dr, zs[j], s0i);
      s0i = fma(si, zc[j], s0i);
      s1r = fma(dr, jc[j], s1r);
      s1r = fma(-si, js[j], s1r);
      s1i = fma(sr, js[j], s1i);
      s1i = fma(di, jc[j], s1i);
      if (MOM == 9) {
        // j^2 (W(+j) z + W(-j) conj z) = j (sigma j z terms)
        const double fj = (double)(j0 + j);
        s2r = fma(fj, fma(sr, jc[j], -di * js[j]), s2r);
        s2i = fma(fj, fma(dr, js[j], si * jc[j]), s2i);
      }
    }
    const cplx<double> XY = cmul(X, Y);
    const double i0 = XY.re * s0i + XY.im * s0r;
    const double i1 = XY.re * s1i + XY.im * s1r;
    u0 += (double)nx * i0;
    u1 += (double)ny * i0;
    u2 += i1;
    if (MOM == 9) {
      const double r0 = XY.re * s0r - XY.im * s0i;
      const double r1 = XY.re * s1r - XY.im * s1i;
      const double r2 = XY.re * s2r - XY.im * s2i;
      const double fx = (double)nx, fy = (double)ny;
      mxx += fx * fx * r0;
      myy += fy * fy * r0;
      mxy += fx * fy * r0;
      mxz += fx * r1;
      myz += fy * r1;
      mzz += r2;
    }
    if (ny == a.Ky) {
      X = cmul(X, x1);
      Y = y0;
    } else {
      Y = cmul(Y, y1);
    }
  }
  if (!act) return;
  double* o = a.out + ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * MOM * a.n;
  o[i] = u0;
  o[(size_t)a.n + i] = u1;
  o[(size_t)2 * a.n + i] = u2;
  if (MOM == 9) {
    o[(size_t)3 * a.n + i] = mxx;
    o[(size_t)4 * a.n + i] = myy;
    o[(size_t)5 * a.n + i] = mzz;
    o[(size_t)6 * a.n + i] = mxy;
    o[(size_t)7 * a.n + i] = mxz;
    o[(size_t)8 * a.n + i] = myz;
  }
}

// The inner box's structure factor of a charge + dipole set from the chunk
// partials of its four weighted sums (weights q, mu_x, mu_y, mu_z, each
// through k_in_phase / k_in_sk_s): S_c(k) = Sq + i (kx Sx + ky Sy + kz Sz)
// per chunk c, as (q + i k.mu) e^{ik.r} summed over the chunk's atoms
// (src/ewald_rec.f90:116-148); written where k_in_reduce_r reads it.
__global__ void k_in_dip_combine(InArgs a, int nchunk, int use_q, const double* __restrict__ pq,
                                 const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)nchunk * a.nb) return;
  const int k = (int)(x % a.nb);
  const double kx = a.kvec[4 * k], ky = a.kvec[4 * k + 1], kz = a.kvec[4 * k + 2];
  const size_t o = 2 * x;
  const double tr = kx * px[o] + ky * py[o] + kz * pz[o];
  const double ti = kx * px[o + 1] + ky * py[o + 1] + kz * pz[o + 1];
  const double qr = use_q ? pq[o] : 0.0, qi = use_q ? pq[o + 1] : 0.0;
  a.part[o] = qr - ti;
  a.part[o + 1] = qi + tr;
}

using InKernel = void (*)(InArgs);
template <int... I>
InKernel pick_in_sk(int kz, bool mb, std::integer_sequence<int, I...>) {
  static const InKernel t[] = {k_in_sk_s<I + 1, false>...};
  static const InKernel tb[] = {k_in_sk_s<I + 1, true>...};
  return mb ? tb[kz - 1] : t[kz - 1];
}
template <int... I>
InKernel pick_in_gather(int kz, int mom, bool mb, std::integer_sequence<int, I...>) {
  static const InKernel t3[] = {k_in_gather_s<I + 1, 3, false>...};
  static const InKernel t9[] = {k_in_gather_s<I + 1, 9, false>...};
  static const InKernel t3b[] = {k_in_gather_s<I + 1, 3, true>...};
  static const InKernel t9b[] = {k_in_gather_s<I + 1, 9, true>...};
  return mom == 9 ? (mb ? t9b[kz - 1] : t9[kz - 1]) : (mb ? t3b[kz - 1] : t3[kz - 1]);
}
// the register-blocked kernels for KZ = kz (1 <= kz <= kInKZ); mb: more
// than one block of j (the block offset taken from the grid)
InKernel in_sk_kernel(int kz, bool mb) {
  return pick_in_sk(kz, mb, std::make_integer_sequence<int, kInKZ>{});
}
InKernel in_gather_kernel(int kz, int mom, bool mb) {
  return pick_in_gather(kz, mom, mb, std::make_integer_sequence<int, kInKZ>{});
}

}  // namespace

// K-step counts compiled for the matrix-core gather (Kz + 1 rounded up)
static const int kMmaBuckets[] = {9, 13, 17, 19, 25, 33, 41, 57};
static int mma_bucket(int nks) {
  for (int b : kMmaBuckets)
    if (nks <= b) return b;
  return 0;
}
// atom tiles of 16 per wave: register budget of the B operand
static int mma_ct(bool fp32, int nks) {
  if (fp32) return nks <= 33 ? 4 : 2;
  return nks <= 25 ? 2 : 1;
}

// the compiled matrix-core gather for (CT, NKS), or null
template <typename T>
static const void* gather_kernel(int ct, int nks) {
#define GM_CASE(CTV, NKSV) \
  case (CTV) * 1000 + (NKSV): return reinterpret_cast<const void*>(&k_gather_mma<T, CTV, NKSV>);
  if constexpr (sizeof(T) == 4) {
    switch (ct * 1000 + nks) {
      GM_CASE(4, 9) GM_CASE(4, 13) GM_CASE(4, 17) GM_CASE(4, 19) GM_CASE(4, 25) GM_CASE(4, 33)
      GM_CASE(2, 41) GM_CASE(2, 57)
      default: return nullptr;
    }
  } else {
    switch (ct * 1000 + nks) {
      GM_CASE(2, 9) GM_CASE(2, 13) GM_CASE(2, 17) GM_CASE(2, 19) GM_CASE(2, 25)
      GM_CASE(1, 33) GM_CASE(1, 41) GM_CASE(1, 57)
      default: return nullptr;
    }
  }
#undef GM_CASE
}

// the compiled matrix-core structure factor for (CTM, TA), or null
template <typename T>
static const void* sk_kernel(int ctm, int ta) {
#define SK_CASE(C, A) \
  case (C) * 100 + (A): return reinterpret_cast<const void*>(&k_sk_mma2<T, C, A>);
  switch (ctm * 100 + ta) {
    SK_CASE(1, 64) SK_CASE(1, 32) SK_CASE(1, 16) SK_CASE(1, 8)
    default: return nullptr;
  }
#undef SK_CASE
}

// Split count of a gather over mtiles m-tiles for atom_waves waves of atoms:
// every wave first loads its atoms' B operand (a prologue worth about
// kGatherPrologue m-tiles), then runs its share of the m-tiles; the time is
// taken as rounds of resident waves x (prologue + m-tiles per wave).  At
// least 4 m-tiles per wave; fewer splits win ties.
constexpr double kGatherPrologue = 4.0;
static int gather_splits(int mtiles, int64_t atom_waves, int64_t conc) {
  int best_s = 1;
  double best = -1.0;
  for (int sp = 1; sp <= std::min(mtiles, 64); ++sp) {
    if (sp > 1 && mtiles / sp < 4) break;
    const int mps = div_up(mtiles, sp);
    const int nsp = div_up(mtiles, mps);
    const int64_t rounds = (atom_waves * nsp + conc - 1) / conc;
    const double cost = (double)rounds * (kGatherPrologue + mps);
    if (best < 0.0 || cost < best - 1e-9) {
      best = cost;
      best_s = nsp;
    }
  }
  return best_s;
}

// resident 256-thread blocks per CU of a kernel with `lds` bytes of LDS
static int blocks_per_cu(const void* fn, size_t lds) {
  int nb = 0;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, lds) != hipSuccess || nb <= 0) nb = 1;
  return nb;
}

// LDS bytes of the matrix-core structure factor for tile TA
static size_t sk_lds(size_t tsz, int ncol, int ta) {
  return (tsz * (size_t)ncol + tsz / 2 * 3) * (size_t)(ta + 2);
}

// split-half kernels: staged atoms per tile, K steps (of 16) compiled for
// the gather, LDS bytes of the structure factor
constexpr int kH3TA = 32;
constexpr int kH3MaxChunk = 2048;   // atoms per structure-factor chunk (FP32 chunk totals)
static const int kH3Nsh[] = {4, 7, 10, 15};
static int h3_bucket(int nsh) {
  for (int b : kH3Nsh)
    if (nsh <= b) return b;
  return 0;
}
static size_t sk_h3_lds(int kx, int ky, int ta, bool wide = false) {
  const int nst = wide ? kx + ky + 2 : kx + 1 + 2 * ky + 1;
  return (size_t)2 * 128 * (ta + 4) * 4 + (size_t)nst * (ta + 4) * 8 + (size_t)3 * (ta + 4) * 4;
}
static const void* gather_h3_kernel(int ct, int nsh) {
  switch (ct * 100 + nsh) {
    case 204: return reinterpret_cast<const void*>(&k_gather_h3<2, 4>);
    case 207: return reinterpret_cast<const void*>(&k_gather_h3<2, 7>);
    case 110: return reinterpret_cast<const void*>(&k_gather_h3<1, 10>);
    case 115: return reinterpret_cast<const void*>(&k_gather_h3<1, 15>);
    default: return nullptr;
  }
}
template <typename TP>
static const void* sk_reduce_h3_kernel(int nsh) {
  switch (nsh) {
    case 4: return reinterpret_cast<const void*>(&k_sk_reduce_h3<4, TP>);
    case 7: return reinterpret_cast<const void*>(&k_sk_reduce_h3<7, TP>);
    case 10: return reinterpret_cast<const void*>(&k_sk_reduce_h3<10, TP>);
    case 15: return reinterpret_cast<const void*>(&k_sk_reduce_h3<15, TP>);
    default: return nullptr;
  }
}

// k-mesh on the box with the reference's k vectors and A(n), kcoe
// (src/kspace.f90:162-191); allocation of the reciprocal buffers.
// The inner k-box of the mixed form (charges only): half-widths from the
// cutoff |k| <= 2.5 alpha (|n_d| <= |k| |a_d| / 2 pi, a_d lattice vector d),
// within the main box; A and k of its entries as the main mesh's; the main
// mesh's A copied with the box zeroed; chunking of k_in_sk.
// a positive number from the environment, or def
static double env_pos(const char* name, double def) {
  const char* ev = getenv(name);
  if (!ev || !*ev) return def;
  const double v = atof(ev);
  return v > 0.0 ? v : def;
}

// Tuning knobs read here (context creation): MDFF_KIN_ALPHA, the box's |k|
// bound in units of alpha (2.5); MDFF_KIN_SG, subgroups of the structure
// factor's blocks; MDFF_KIN_CHUNKS, its atom chunks (256); MDFF_KIN_GBLOCKS,
// the gather's target block count (768).
static int recip_inner_setup(mdff_ctx* c, const std::vector<double>& Amain) {
  c->kin_on = 0;
  const bool fp32 = c->rec_f32 != 0;
  if (!fp32 || !c->task_qq) return MDFF_OK;
  const double kcut = env_pos("MDFF_KIN_ALPHA", 2.5) * c->alpha;
  int kin[3];
  for (int d = 0; d < 3; ++d) {
    const double* a = c->A + 3 * d;
    const double len = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    kin[d] = std::min(c->K[d], (int)ceil(kcut * len / (2.0 * MDFF_PI)));
    if (kin[d] < 1) return MDFF_OK;
  }
  for (int d = 0; d < 3; ++d) c->kin[d] = kin[d];
  const int NY = 2 * kin[1] + 1, NZ = 2 * kin[2] + 1;
  const int nbi = (kin[0] + 1) * NY * NZ;
  c->nb_in = nbi;
  std::vector<double> Ai(nbi, 0.0), kv(4 * (size_t)nbi, 0.0), Am(Amain);
  const int Ky = c->K[1], Kz = c->K[2];
  for (int nx = 0; nx <= kin[0]; ++nx)
    for (int ny = -kin[1]; ny <= kin[1]; ++ny)
      for (int nz = -kin[2]; nz <= kin[2]; ++nz) {
        const int im = (nx * c->NYk + (ny + Ky)) * c->NZk + (nz + Kz);
        const int ii = (nx * NY + (ny + kin[1])) * NZ + (nz + kin[2]);
        Ai[ii] = Amain[im];
        Am[im] = 0.0;
      }
  // k vectors and kcoe of the inner entries: the main mesh's values
  std::vector<double> kvm(4 * (size_t)c->nk_box);
  HIP_TRY(c, hipMemcpy(kvm.data(), c->kvec, sizeof(double) * 4 * c->nk_box, hipMemcpyDeviceToHost));
  for (int nx = 0; nx <= kin[0]; ++nx)
    for (int ny = -kin[1]; ny <= kin[1]; ++ny)
      for (int nz = -kin[2]; nz <= kin[2]; ++nz) {
        const int im = (nx * c->NYk + (ny + Ky)) * c->NZk + (nz + Kz);
        const int ii = (nx * NY + (ny + kin[1])) * NZ + (nz + kin[2]);
        for (int u = 0; u < 4; ++u) kv[4 * (size_t)ii + u] = kvm[4 * (size_t)im + u];
      }
  HIP_TRY(c, hipMalloc((void**)&c->kA_in, sizeof(double) * nbi));
  HIP_TRY(c, hipMalloc((void**)&c->kvec_in, sizeof(double) * 4 * nbi));
  HIP_TRY(c, hipMalloc((void**)&c->kA_mask, sizeof(double) * c->nk_box));
  HIP_TRY(c, hipMalloc((void**)&c->wk_in, sizeof(double) * 2 * nbi));
  HIP_TRY(c, hipMemcpy(c->kA_in, Ai.data(), sizeof(double) * nbi, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->kvec_in, kv.data(), sizeof(double) * 4 * nbi, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->kA_mask, Am.data(), sizeof(double) * c->nk_box, hipMemcpyHostToDevice));
  const int ncol = kin[0] + kin[1] + kin[2] + 3;
  const int M = (kin[0] + 1) * NY;
  const int ts = div_up(M, 64) * 64;
  c->kin_reg = (kin[2] <= kInKZmax && ts <= 1024) ? 1 : 0;
  // blocks of at most kInKZ + 1 values of j = 0..Kz'
  c->kin_jb = c->kin_reg ? div_up(kin[2] + 1, kInKZ + 1) : 1;
  c->kin_kb = c->kin_reg ? div_up(kin[2] + 1, c->kin_jb) - 1 : 0;
  if (c->kin_reg) {
    // k_in_sk_s: SG subgroups of ts threads (at most 1024 threads), one
    // block per chunk; the chunk's X / Y rows through LDS in tiles within
    // 48 KiB.  Charge-only systems, whose pair kernel runs beside the
    // reciprocal kernels on the second stream in MD steps that reuse the
    // lists: one subgroup and about 384 chunks (1.5 blocks per CU), which
    // leaves room on the SIMDs for the pair kernel's waves (24k a-SiO2:
    // 0.298 against 0.307 ms per MD step, although the kernel alone takes
    // longer); dipole systems: 4 subgroups and about 256 chunks (one per
    // CU).  The shape depends on the system only, so that a run with every
    // kernel on one stream sums in the same order.
    // k_in_gather_s: m ranges for about 768 blocks of 256 atoms.
    const bool beside = !(c->task_qd || c->task_dd);
    const int sg = std::min((int)env_pos("MDFF_KIN_SG", beside ? 1 : 4), 1024 / ts);
    const int nxy = kin[0] + kin[1] + 2;
    int chunk = std::max(div_up(c->n, (int)env_pos("MDFF_KIN_CHUNKS", beside ? 384 : 256)), 4 * sg);
    const int tile = std::min(chunk, 49152 / (16 * nxy));
    chunk = div_up(chunk, tile) * tile;
    c->kin_ts = ts;
    c->kin_sg = sg;
    c->in_chunk = chunk;
    c->in_nch = div_up(c->n, chunk);
    c->kin_tile = tile;
    c->kin_lds = std::max<size_t>((size_t)tile * nxy * 16, (size_t)(sg - 1) * ts * 32);
    const int sp = std::max(1, std::min(M, div_up((int)env_pos("MDFF_KIN_GBLOCKS", 768), div_up(c->n, 256))));
    c->kin_msplit = div_up(M, sp);
    c->kin_splits = div_up(M, c->kin_msplit);
    // phase rows and W table rows padded to whole blocks of j; the padding
    // is zero from here on (no kernel writes it)
    c->kin_zpad = c->kin_jb * (c->kin_kb + 1);
    const size_t ncolp = (size_t)nxy + c->kin_zpad;
    HIP_TRY(c, hipMalloc((void**)&c->tab_in, 16 * ncolp * c->n));
    HIP_TRY(c, hipMalloc((void**)&c->wt_in, sizeof(double) * 4 * (size_t)M * c->kin_zpad));
    HIP_TRY(c, hipMemset(c->tab_in, 0, 16 * ncolp * c->n));
    HIP_TRY(c, hipMemset(c->wt_in, 0, sizeof(double) * 4 * (size_t)M * c->kin_zpad));
  } else {
    // k_in_sk: about 2048 blocks; tiles of atoms within 48 KiB of LDS.
    // k_in_gather: m ranges of whole kInMT tiles, about 2048 blocks
    const int tile = (int)std::min<size_t>(64, 49152 / (16 * (size_t)ncol + 8));
    if (tile < 1) return MDFF_OK;
    const int kblocks = div_up(nbi, 256);
    const int want = std::max(1, 2048 / kblocks);
    int chunk = div_up(c->n, want);
    chunk = div_up(chunk, tile) * tile;
    c->in_chunk = chunk;
    c->in_nch = div_up(c->n, chunk);
    c->kin_tile = tile;
    c->kin_lds = (size_t)tile * (16 * (size_t)ncol + 8);
    const int tiles = div_up(M, kInMT);
    const int sp = std::max(1, std::min(tiles, div_up(2048, div_up(c->n, 256))));
    c->kin_msplit = div_up(tiles, sp) * kInMT;
    c->kin_splits = div_up(M, c->kin_msplit);
  }
  HIP_TRY(c, hipMalloc((void**)&c->sk_in_part, sizeof(double) * 2 * (size_t)nbi * c->in_nch));
  // dipole systems: the chunk partials of the four weighted structure
  // factors (q, mu_x, mu_y, mu_z) of the force evaluation
  if ((c->task_qd || c->task_dd) && c->kin_reg)
    HIP_TRY(c, hipMalloc((void**)&c->dip_part, sizeof(double) * 8 * (size_t)nbi * c->in_nch));
  c->kin_on = 1;
  return MDFF_OK;
}

// whether the register-blocked inner box exists for this cell and kmax (the
// condition recip_inner_setup applies): dipole systems need it to keep the
// FP32-class reciprocal kernels
static bool inner_reg_possible(const mdff_ctx* c) {
  if (!c->task_qq) return false;
  const double kcut = env_pos("MDFF_KIN_ALPHA", 2.5) * c->alpha;
  int kin[3];
  for (int d = 0; d < 3; ++d) {
    const double* a = c->A + 3 * d;
    const double len = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    kin[d] = std::min(c->d.kES[d], (int)ceil(kcut * len / (2.0 * MDFF_PI)));
    if (kin[d] < 1) return false;
  }
  const int M = (kin[0] + 1) * (2 * kin[1] + 1);
  return kin[2] <= kInKZmax && div_up(M, 64) * 64 <= 1024;
}

int recip_setup(mdff_ctx* c) {
  const int Kx = c->d.kES[0], Ky = c->d.kES[1], Kz = c->d.kES[2];
  // a dipole system keeps the FP32-class reciprocal kernels only with the
  // FP64 inner box of its force evaluation; otherwise its sums run in FP64
  if (c->rec_f32 && (c->task_qd || c->task_dd) && !inner_reg_possible(c)) c->rec_f32 = 0;
  c->K[0] = Kx; c->K[1] = Ky; c->K[2] = Kz;
  c->NXk = Kx + 1; c->NYk = 2 * Ky + 1; c->NZk = 2 * Kz + 1;
  c->nk_box = c->NXk * c->NYk * c->NZk;
  c->nk = Kz + Ky * (2 * Kz + 1) + Kx * (2 * Ky + 1) * (2 * Kz + 1);
  const int nb = c->nk_box;
  std::vector<double> A(nb, 0.0), kv(4 * (size_t)nb, 0.0);
  const double tpi = 2.0 * MDFF_PI, alpha2 = c->alpha * c->alpha;
  const double* B = c->B;
  for (int nx = 0; nx <= Kx; ++nx)
    for (int ny = -Ky; ny <= Ky; ++ny)
      for (int nz = -Kz; nz <= Kz; ++nz) {
        bool half = (nx > 0) || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
        int idx = (nx * c->NYk + (ny + Ky)) * c->NZk + (nz + Kz);
        double fx = nx, fy = ny, fz = nz;
        double kx = tpi * (fx * B[0] + fy * B[3] + fz * B[6]);
        double ky = tpi * (fx * B[1] + fy * B[4] + fz * B[7]);
        double kz = tpi * (fx * B[2] + fy * B[5] + fz * B[8]);
        double kk = kx * kx + ky * ky + kz * kz;
        kv[4 * idx] = kx; kv[4 * idx + 1] = ky; kv[4 * idx + 2] = kz;
        if (half && kk != 0.0) {
          A[idx] = exp(-(kk * 0.25) / alpha2) / kk;
          kv[4 * idx + 3] = 2.0 * (1.0 / kk + 1.0 / alpha2 / 4.0);
        }
      }
  HIP_TRY(c, hipMalloc((void**)&c->kA, sizeof(double) * nb));
  HIP_TRY(c, hipMalloc((void**)&c->kvec, sizeof(double) * 4 * nb));
  HIP_TRY(c, hipMemcpy(c->kA, A.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->kvec, kv.data(), sizeof(double) * 4 * nb, hipMemcpyHostToDevice));
  const size_t tsz = c->rec_f32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const int ncol = Kx + Ky + Kz + 3;
  HIP_TRY(c, hipMalloc(&c->tab, tsz * ncol * (size_t)c->n));
  // structure-factor chunks: enough work items to fill the GPU
  const int GZ = 8;
  const int NG = div_up(c->NZk, GZ);
  const int64_t items = (int64_t)c->NXk * c->NYk * NG;
  int chunks = (int)std::max<int64_t>(1, (int64_t)256 * 1024 * 8 / std::max<int64_t>(items, 1));
  chunks = std::min(chunks, std::max(1, c->n / 64));
  chunks = std::min(chunks, 64);
  c->sk_chunks = chunks;
  HIP_TRY(c, hipMalloc((void**)&c->sk, sizeof(double) * 2 * nb));
  HIP_TRY(c, hipMalloc((void**)&c->wk, sizeof(double) * 2 * nb));
  HIP_TRY(c, hipMalloc(&c->wk_t, tsz * (size_t)nb));
  // gather splits over (nx, ny): about 256k threads in flight
  const int M = c->NXk * c->NYk;
  int splits = std::max(1, std::min(M, div_up((int64_t)262144, c->n)));
  // matrix-core gather: A operand and its own split count
  const bool fp32 = c->rec_f32 != 0;
  c->mma_nks = mma_bucket(Kz + 1);
  c->mma_mtiles = div_up(M, 4);
  c->mma_ct = mma_ct(fp32, c->mma_nks);
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0)
    cus = 256;
  // splits over m-tiles (gather_splits)
  const int atom_waves = div_up(c->n, 16 * c->mma_ct);
  const void* gfn = fp32 ? gather_kernel<float>(c->mma_ct, c->mma_nks) : gather_kernel<double>(c->mma_ct, c->mma_nks);
  const int64_t conc = 4 * (int64_t)cus * blocks_per_cu(gfn, 0);
  const int best_s = gather_splits(c->mma_mtiles, atom_waves, conc);
  c->mma_splits = best_s;
  // matrix-core structure factor: atoms per staged tile within 64 KiB of
  // LDS; atom chunks chosen from the kernel's occupancy so that the blocks
  // fill whole rounds of resident blocks with the fewest tiles per block
  {
    int TA = 64;
    while (TA > 8 && sk_lds(tsz, ncol, TA) > 65536) TA /= 2;
    c->sk_ta = sk_lds(tsz, ncol, TA) <= 65536 ? TA : 0;
    c->sk_mma_ch = 1;
    if (c->sk_ta) {
      const void* sfn = fp32 ? sk_kernel<float>(c->sk_ctm, TA) : sk_kernel<double>(c->sk_ctm, TA);
      const int64_t cap = (int64_t)cus * blocks_per_cu(sfn, sk_lds(tsz, ncol, TA));
      const int64_t mblocks = div_up(div_up(M, 16), 4 * c->sk_ctm) * (int64_t)div_up(c->NZk, 56);
      int64_t best = -1;
      for (int ch = 1; ch <= 128 && (ch == 1 || div_up(c->n, ch) >= TA); ++ch) {
        const int per = div_up(c->n, ch);
        const int nch = div_up(c->n, per);
        const int64_t rounds = (mblocks * nch + cap - 1) / cap;
        const int64_t cost = rounds * div_up(per, TA);
        if (best < 0 || cost < best) {
          best = cost;
          c->sk_mma_ch = nch;
        }
      }
    }
  }
  HIP_TRY(c, hipMalloc(&c->amat, tsz / 2 * (size_t)c->mma_mtiles * c->mma_nks * 64));
  HIP_TRY(c, hipMemset(c->amat, 0, tsz / 2 * (size_t)c->mma_mtiles * c->mma_nks * 64));
  // block partials of the energy / stress: k-blocks of 256, or m-tiles of 8,
  // then those of the inner box (k-blocks of 64, or of whole rows m, at
  // least 46 k each)
  const size_t nblk_rec = std::max<size_t>(div_up(nb, 256), div_up((int64_t)c->NXk * c->NYk, 8)) +
                          div_up((int64_t)(Kx + 1) * (2 * Ky + 1) * (2 * Kz + 1), 32);
  HIP_TRY(c, hipMalloc((void**)&c->rec_blk, sizeof(double) * 7 * nblk_rec));
  // mixed form: split-half kernels when their shapes fit (MDFF_RECIP_F32MMA=1
  // keeps the FP32 matrix-core kernels)
  c->h3 = 0;
  c->h3_sk = 0;
  {
    const char* ev = getenv("MDFF_RECIP_F32MMA");
    const bool off = ev && ev[0] == '1';
    const bool sk_off = ev && ev[0] == '2';         // FP32 matrix-core structure factor only
    const int nxy = Kx + Ky + 2;
    const int nsh = h3_bucket(div_up(2 * c->NZk, 16));
    // the gather addresses the FP32 phase tables with 32-bit byte offsets
    const bool tab_ok = 8.0 * (Kx + Ky + Kz + 3) * (double)c->n < 2147483648.0;
    if (fp32 && c->use_mma && !off && nsh > 0 && tab_ok) {
      // the gather for any kz its K steps cover; the structure factor where
      // its staged columns fit (else the FP32 matrix-core one feeds it)
      c->h3 = 1;
      c->h3_sk = !sk_off && nxy <= 8 * (256 / kH3TA) && sk_h3_lds(c->K[0], c->K[1], kH3TA) <= 65536;
      if (!sk_off && !c->h3_sk && nxy <= 16 * (256 / kH3TA)) {
        // wide mesh: the wide form where its LDS fits a block on this device
        int maxlds = 0;
        const size_t lw = sk_h3_lds(c->K[0], c->K[1], kH3TA, true);
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess &&
            lw <= (size_t)maxlds) {
          bool ok = true;
          for (const void* fn : {reinterpret_cast<const void*>(&k_sk_h3<kH3TA, 0, true>),
                                 reinterpret_cast<const void*>(&k_sk_h3<kH3TA, 1, true>),
                                 reinterpret_cast<const void*>(&k_sk_h3<kH3TA, 2, true>)})
            ok = ok && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw) == hipSuccess;
          (void)hipGetLastError();
          c->h3_sk = c->h3_wide = ok ? 1 : 0;
        }
      }
      c->h3_nsh = nsh;
      c->h3_ct = nsh <= 7 ? 2 : 1;
      c->h3_mt8 = div_up(M, 8);
      // B operand of the structure factor: max |q| scaled into [256, 512)
      int e = 0;
      if (c->qmax > 0.0) (void)frexp(c->qmax, &e);
      c->h3_pscale = (float)ldexp(1.0, 9 - e);
      // gather splits over m-tiles (gather_splits)
      const int aw = div_up(c->n, 32 * c->h3_ct);
      const int64_t gconc = 4 * (int64_t)cus * blocks_per_cu(gather_h3_kernel(c->h3_ct, nsh), 0);
      const int bs = gather_splits(c->h3_mt8, aw, gconc);
      c->h3_splits = bs;
      // structure-factor chunks from the kernel's occupancy
      if (c->h3_sk) {
        const size_t lds = sk_h3_lds(c->K[0], c->K[1], kH3TA, c->h3_wide);
        const void* fn = c->h3_wide ? reinterpret_cast<const void*>(&k_sk_h3<kH3TA, 0, true>)
                                    : reinterpret_cast<const void*>(&k_sk_h3<kH3TA, 0, false>);
        const int64_t cap = (int64_t)cus * blocks_per_cu(fn, lds);
        const int64_t mblocks = div_up(div_up(M, 32), 4) * (int64_t)div_up(c->NZk, 64);
        // at most kH3MaxChunk atoms per chunk: a chunk's FP32 total adds one
        // tile sum per 32 atoms, so its rounding grows with the chunk (the
        // chunks are summed in FP64)
        const int ch0 = div_up(c->n, kH3MaxChunk);
        int64_t best = -1;
        for (int ch = ch0; ch <= std::max(128, ch0) && (ch == ch0 || div_up(c->n, ch) >= kH3TA); ++ch) {
          const int per = div_up(c->n, ch);
          const int nch = div_up(c->n, per);
          const int64_t cost = ((mblocks * nch + cap - 1) / cap) * div_up(per, kH3TA);
          if (best < 0 || cost < best) { best = cost; c->h3_ch = nch; }
        }
      }
      HIP_TRY(c, hipMalloc(&c->amat_h3, 16 * 128 * (size_t)nsh * c->h3_mt8));
      HIP_TRY(c, hipMalloc((void**)&c->ascale_h3, sizeof(float) * (size_t)c->h3_mt8));
      HIP_TRY(c, hipMalloc((void**)&c->mu_pscale, sizeof(float) * 2));
    }
  }
  HIP_TRY(c, hipMalloc(&c->sk_part, sizeof(double) * 4 * (size_t)std::max({c->sk_chunks, c->sk_mma_ch, c->h3_ch}) * nb));
  MDFF_TRY(recip_inner_setup(c, A));
  splits = std::max(splits, c->mma_splits);
  if (c->h3) splits = std::max(splits, c->h3_splits);
  c->rec_splits = splits;
  // moments of every split, plus (3 moments each) the inner box's
  HIP_TRY(c, hipMalloc((void**)&c->rec_part,
                       sizeof(double) * (9 * (size_t)splits + 9 * (size_t)c->kin_splits * c->kin_jb) * c->n));
  return MDFF_OK;
}

void recip_free(mdff_ctx* c) {
  void* ps[] = {c->kA, c->kvec, c->tab, c->sk_part, c->sk, c->wk, c->wk_t, c->rec_part, c->amat, c->rec_blk,
                c->amat_h3, c->ascale_h3, c->kA_mask, c->kA_in, c->kvec_in, c->sk_in_part, c->wk_in, c->tab_in, c->wt_in,
                c->mu_pscale, c->dip_part};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

template <typename T>
static void tables_t(mdff_ctx* c) {
  PROF_LAUNCH(c, PK_TABLES,
              hipLaunchKernelGGL((k_tables<T>), dim3(div_up(c->n, 256), tab_runs(c->K)), dim3(256), 0, c->stream, c->n,
                                 c->K[0], c->K[1], c->K[2], c->sx, c->sy, c->sz,
                                 reinterpret_cast<cplx<T>*>(c->tab)));
}

template <typename T>
static void tables_r_t(mdff_ctx* c, bool pack) {
  TabRArgs a{};
  a.n = c->n;
  a.Kx = c->K[0]; a.Ky = c->K[1]; a.Kz = c->K[2];
  a.rx = c->rx; a.ry = c->ry; a.rz = c->rz;
  for (int k = 0; k < 9; ++k) a.b[k] = c->B[k];
  a.sx = c->sx; a.sy = c->sy; a.sz = c->sz;
  a.tab = c->tab;
  a.pack = pack ? 1 : 0;
  a.order = c->cell_atoms;
  a.q = c->q;
  a.itype = c->itype;
  a.rec = c->arec;
  a.meta = c->ameta;
  // the previous MD step's final sums ride along as one more block row
  const int nbx = div_up(c->n, 256);
  if (c->nfin > nbx) (void)flush_finals(c);
  const FinalBatch fb = take_finals(c);
  const int runs = tab_runs(c->K);
  const bool extra = fb.nfin > 0 || fb.seq_out;
  PROF_LAUNCH(c, PK_TABLES,
              hipLaunchKernelGGL((k_tables_r<T>), dim3(nbx, runs + (extra ? 1 : 0)), dim3(256), 0, c->stream, a,
                                 fb, extra ? runs : -1));
}

int recip_tables_from_r(mdff_ctx* c, bool pack) {
  if (c->rec_f32) tables_r_t<float>(c, pack);
  else tables_r_t<double>(c, pack);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int recip_tables(mdff_ctx* c) {
  if (c->rec_f32) tables_t<float>(c);
  else tables_t<double>(c);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

template <typename T>
static int launch_gather_mma(mdff_ctx* c, const GmArgs& gm, int ct, dim3 grid) {
  const void* fn = gather_kernel<T>(ct, gm.nks);
  if (!fn) {
    c->err = "no matrix-core gather kernel for this kmax";
    return MDFF_ERR_UNSUPPORTED;
  }
  void* args[] = {const_cast<GmArgs*>(&gm)};
  PROF_LAUNCH(c, PK_GATHER, (void)hipLaunchKernel(fn, grid, dim3(256), args, 0, c->stream));
  return MDFF_OK;
}

template <typename T>
static void launch_sk_mma2(int ctm, dim3 grid, size_t lds, hipStream_t st, const Sm2Args& sm) {
  void* args[] = {const_cast<Sm2Args*>(&sm)};
  (void)hipLaunchKernel(sk_kernel<T>(ctm, sm.TA), grid, dim3(256), args, lds, st);
}

template <typename T>
static int eval_t(mdff_ctx* c, const double* q, const double* mu, bool use_q, bool use_mu, bool forces,
                  bool field, double* ef_out, double* f_out, double* e_out) {
  constexpr int GZ = 8;
  SkArgs s{};
  s.n = c->n;
  s.Kx = c->K[0]; s.Ky = c->K[1]; s.Kz = c->K[2];
  s.NY = c->NYk; s.NZ = c->NZk;
  s.M = c->NXk * c->NYk;
  s.NG = div_up(c->NZk, GZ);
  s.GZ = GZ;
  s.use_q = use_q; s.use_mu = use_mu;
  s.a0_step = div_up(c->n, c->sk_chunks);
  s.q = q; s.mu = mu;
  for (int k = 0; k < 9; ++k) s.b[k] = c->B[k];
  s.tab = c->tab;
  s.part = c->sk_part;
  s.nk_box = c->nk_box;
  const int ncol = c->K[0] + c->K[1] + c->K[2] + 3;
  // atoms per LDS tile within a 64 KiB budget
  int TA = 32;
  while (TA > 1 && sizeof(cplx<T>) * (size_t)ncol * TA + sizeof(T) * 4 * TA > 65536) TA /= 2;
  s.TA = TA;
  const size_t shm = sizeof(cplx<T>) * (size_t)ncol * TA + sizeof(T) * 4 * TA;
  if (shm > 65536) {
    c->err = "kES too large for the structure-factor kernel";
    return MDFF_ERR_UNSUPPORTED;
  }
  int chunks = div_up(c->n, s.a0_step);
  bool sk_fp64 = false;
  const bool h3 = sizeof(T) == 4 && c->h3;           // split-half gather
  if (h3 && c->h3_sk) {
    // split-half structure factor, FP32 partials per chunk
    Sh3Args sh{};
    sh.n = c->n;
    sh.Kx = c->K[0]; sh.Ky = c->K[1]; sh.Kz = c->K[2];
    sh.NY = c->NYk; sh.NZ = c->NZk; sh.M = s.M;
    int ch = c->h3_ch;
    sh.chunk = div_up(c->n, ch);
    ch = div_up(c->n, sh.chunk);
    chunks = ch;
    sh.nk_box = c->nk_box;
    sh.q = use_q ? q : c->zeros;
    sh.mu = mu;
    for (int k = 0; k < 9; ++k) sh.b[k] = c->B[k];
    sh.tab = c->tab;
    float* p1 = reinterpret_cast<float*>(c->sk_part);
    sh.part = p1;
    sh.mode = use_mu ? 1 : 0;
    sh.pscale = c->h3_pscale;
    sh.pscale_dev = nullptr;
    if (use_mu) {
      const double* b = c->B;
      hipLaunchKernelGGL(k_mu_pscale, dim3(1), dim3(256), 0, c->stream, c->n, c->K[0], c->K[1], sh.q, mu, b[0],
                         b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], c->mu_pscale);
      sh.pscale_dev = c->mu_pscale;
    }
    const bool wide = c->h3_wide != 0;
    const size_t lds = sk_h3_lds(c->K[0], c->K[1], kH3TA, wide);
    dim3 gs(div_up(div_up(s.M, 32), 4), ch, div_up(c->NZk, 64));
    if (use_mu && wide)
      PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 1, true>), gs, dim3(256), lds, c->stream, sh));
    else if (use_mu)
      PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 1, false>), gs, dim3(256), lds, c->stream, sh));
    else if (wide)
      PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 0, true>), gs, dim3(256), lds, c->stream, sh));
    else
      PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 0, false>), gs, dim3(256), lds, c->stream, sh));
    if (use_mu) {
      Sh3Args s2 = sh;
      s2.mode = 2;
      s2.part = p1 + (size_t)2 * ch * c->nk_box;
      const size_t tot = (size_t)ch * c->nk_box;
      if (wide)
        PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 2, true>), gs, dim3(256), lds, c->stream, s2));
      else
        PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk_h3<kH3TA, 2, false>), gs, dim3(256), lds, c->stream, s2));
      PROF_LAUNCH(c, PK_SK,
                  hipLaunchKernelGGL(k_sk_combine<float>, dim3(div_up(tot, 256)), dim3(256), 0, c->stream,
                                     c->nk_box, c->NZk, c->K[2], ch, p1, s2.part));
    }
  } else if (c->use_mma) {
    sk_fp64 = true;
    Sm2Args sm{};
    sm.n = c->n;
    sm.Kx = c->K[0]; sm.Ky = c->K[1]; sm.Kz = c->K[2];
    sm.NY = c->NYk; sm.NZ = c->NZk; sm.M = s.M;
    sm.mtiles = div_up(s.M, 16);
    const int CTM = c->sk_ctm;
    const int mblocks = div_up(sm.mtiles, 4 * CTM);
    int ch = c->sk_mma_ch;
    sm.chunk = div_up(c->n, ch);
    ch = div_up(c->n, sm.chunk);
    chunks = ch;
    sm.nk_box = c->nk_box;
    sm.q = q; sm.mu = mu;
    for (int k = 0; k < 9; ++k) sm.b[k] = c->B[k];
    sm.tab = c->tab;
    const int TA = c->sk_ta;
    auto lds = [&](int ta) { return sk_lds(sizeof(cplx<T>), ncol, ta); };
    if (TA == 0) {
      c->err = "kES too large for the structure-factor kernel";
      return MDFF_ERR_UNSUPPORTED;
    }
    sm.TA = TA;
    dim3 gs(mblocks, ch, div_up(c->NZk, 56));
    double* p1 = reinterpret_cast<double*>(c->sk_part);
    sm.part = p1;
    // first pass: mode 0 (charges) or mode 1 (charges, zero when switched
    // off, plus the x/y dipole phase); second pass for dipoles: mode 2
    sm.mode = use_mu ? 1 : 0;
    if (!use_q) sm.q = c->zeros;
    PROF_LAUNCH(c, PK_SK, launch_sk_mma2<T>(CTM, gs, lds(TA), c->stream, sm));
    if (use_mu) {
      Sm2Args s2 = sm;
      s2.mode = 2;
      s2.part = p1 + (size_t)2 * ch * c->nk_box;
      const size_t tot = (size_t)ch * c->nk_box;
      PROF_LAUNCH(c, PK_SK, launch_sk_mma2<T>(CTM, gs, lds(TA), c->stream, s2);
                  hipLaunchKernelGGL(k_sk_combine<double>, dim3(div_up(tot, 256)), dim3(256), 0, c->stream, c->nk_box,
                                     c->NZk, c->K[2], ch, p1, s2.part));
    }
  } else {
    dim3 g1(div_up((int64_t)s.M * s.NG, 256), chunks);
    PROF_LAUNCH(c, PK_SK, hipLaunchKernelGGL((k_sk<T, GZ>), g1, dim3(256), shm, c->stream, s));
  }
  // inner k-box in FP64 (mixed form, charges only): its structure factor
  // here; the main kernels below see A = 0 on the box
  // charges: every evaluation; dipoles: the force evaluation (its error is
  // the mixed form's; the SCF fields stay on the FP32-class kernels)
  const bool inner_q = sizeof(T) == 4 && c->kin_on && use_q && !use_mu;
  const bool inner_d = sizeof(T) == 4 && c->kin_on && c->kin_reg && c->dip_part && forces && use_mu;
  const bool inner = inner_q || inner_d;
  const double* kA = inner ? c->kA_mask : c->kA;
  InArgs ia{};
  if (inner) {
    ia.n = c->n;
    ia.Kx = c->kin[0]; ia.Ky = c->kin[1]; ia.Kz = c->kin[2];
    ia.NY = 2 * c->kin[1] + 1; ia.NZ = 2 * c->kin[2] + 1;
    ia.nb = c->nb_in;
    ia.zpad = c->kin_zpad;
    ia.chunk = c->in_chunk;
    ia.tile = c->kin_tile;
    ia.sx = c->sx; ia.sy = c->sy; ia.sz = c->sz;
    ia.q = q;
    ia.qstride = 1;
    ia.kA = c->kA_in;
    ia.kvec = c->kvec_in;
    ia.part = c->sk_in_part;
    ia.wk = c->wk_in;
    const size_t lds = c->kin_lds;
    if (inner_d) {
      // four weighted structure factors (q, mu_x, mu_y, mu_z), then
      // Sq + i k.(Sx, Sy, Sz) per chunk into sk_in_part
      ia.ts = c->kin_ts;
      ia.tab = reinterpret_cast<cplx<double>*>(c->tab_in);
      ia.wt = c->wt_in;
      const size_t pstride = 2 * (size_t)ia.nb * c->in_nch;
      for (int w = use_q ? 0 : 1; w < 4; ++w) {
        InArgs iw = ia;
        iw.q = w == 0 ? q : mu + (w - 1);
        iw.qstride = w == 0 ? 1 : 3;
        iw.part = c->dip_part + (size_t)w * pstride;
        PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_phase, dim3(div_up(3 * c->n, 256)), dim3(256), 0, c->stream, iw));
        PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(in_sk_kernel(c->kin_kb, c->kin_jb > 1), dim3(c->in_nch, c->kin_jb),
                                                 dim3(c->kin_ts * c->kin_sg), lds, c->stream, iw));
      }
      const size_t tot = (size_t)c->in_nch * ia.nb;
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_dip_combine, dim3(div_up(tot, 256)), dim3(256), 0, c->stream, ia,
                                               c->in_nch, use_q ? 1 : 0, c->dip_part, c->dip_part + pstride,
                                               c->dip_part + 2 * pstride, c->dip_part + 3 * pstride));
    } else if (c->kin_reg) {
      ia.ts = c->kin_ts;
      ia.tab = reinterpret_cast<cplx<double>*>(c->tab_in);
      ia.wt = c->wt_in;
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_phase, dim3(div_up(3 * c->n, 256)), dim3(256), 0, c->stream, ia));
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(in_sk_kernel(c->kin_kb, c->kin_jb > 1), dim3(c->in_nch, c->kin_jb),
                                               dim3(c->kin_ts * c->kin_sg), lds, c->stream, ia));
    } else {
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_sk, dim3(div_up(ia.nb, 256), c->in_nch), dim3(256), lds,
                                               c->stream, ia));
    }
  }
  // the 7 energy / stress sums over the k-blocks: block partials here, the
  // final sum queued into e_out (device memory)
  const bool dipf = forces && use_mu;
  const bool mma_gather = (forces || field) && !dipf && c->mma_nks > 0 && c->use_mma && !h3;
  // split-half gather: its A operand is written by the reduction (per m-tile)
  const bool h3_gather = h3 && !dipf && (forces || field);
  T* amat = mma_gather ? reinterpret_cast<T*>(c->amat) : nullptr;
  const int nb = h3_gather ? c->h3_mt8 : div_up(c->nk_box, 256);
  MDFF_TRY(claim_partials(c, c->rec_blk));
  const int rows_in = inner && c->kin_reg ? 64 / ia.NZ : 0;
  const int nbi = !inner ? 0 : c->kin_reg ? div_up((ia.Kx + 1) * ia.NY, rows_in) : div_up(c->nb_in, 64);
  if (inner) {
    ia.blk = c->rec_blk + (size_t)7 * nb;     // after the main blocks
    if (c->kin_reg) {
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_reduce_r, dim3(nbi), dim3(1024), 0, c->stream, ia, c->in_nch));
    } else {
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_reduce, dim3(nbi), dim3(256), 0, c->stream, ia, c->in_nch));
    }
  }
  PROF_LAUNCH(c, PK_SKRED,
              if (h3_gather) {
                const void* prt = c->sk_part;        // FP32 (split-half) or FP64 (FP32 matrix-core) partials
                float* wkt = reinterpret_cast<float*>(c->wk_t);
                u32x4* am = reinterpret_cast<u32x4*>(c->amat_h3);
                void* rargs[] = {(void*)&c->nk_box, (void*)&chunks, (void*)&prt, (void*)&kA, (void*)&c->kvec,
                                 (void*)&c->sk, (void*)&c->wk, (void*)&wkt, (void*)&c->rec_blk, (void*)&c->NZk,
                                 (void*)&c->K[2], (void*)&am, (void*)&c->ascale_h3};
                const void* fn = sk_fp64 ? sk_reduce_h3_kernel<double>(c->h3_nsh) : sk_reduce_h3_kernel<float>(c->h3_nsh);
                (void)hipLaunchKernel(fn, dim3(nb), dim3(256), rargs, 0, c->stream);
              } else if (sk_fp64) {
                hipLaunchKernelGGL((k_sk_reduce<T, double>), dim3(nb), dim3(256), 0, c->stream, c->nk_box, chunks,
                                   reinterpret_cast<const double*>(c->sk_part), kA, c->kvec, c->sk, c->wk,
                                   reinterpret_cast<T*>(c->wk_t), c->rec_blk, amat, c->NZk, c->K[2],
                                   c->mma_nks);
              } else {
                hipLaunchKernelGGL((k_sk_reduce<T, T>), dim3(nb), dim3(256), 0, c->stream, c->nk_box, chunks,
                                   reinterpret_cast<const T*>(c->sk_part), kA, c->kvec, c->sk, c->wk,
                                   reinterpret_cast<T*>(c->wk_t), c->rec_blk, amat, c->NZk, c->K[2],
                                   c->mma_nks);
              });
  if (e_out) MDFF_TRY(queue_final(c, c->rec_blk, nb + nbi, 7, e_out));
  if (!forces && !field) return MDFF_OK;
  int nsplit = 0, moments = dipf ? 9 : 3;
  if (h3_gather) {
    // split-half gather (A operand per m-tile of 8 m values from the reduction)
    const int nsh = c->h3_nsh, mt8 = c->h3_mt8;
    Gh3Args gh{};
    gh.n = c->n;
    gh.Kx = c->K[0]; gh.Ky = c->K[1]; gh.Kz = c->K[2];
    gh.NY = c->NYk; gh.NZ = c->NZk; gh.M = s.M;
    gh.mtiles = mt8;
    gh.mt_per_split = div_up(mt8, c->h3_splits);
    nsplit = div_up(mt8, gh.mt_per_split);
    gh.tab = c->tab;
    gh.amat = reinterpret_cast<const u32x4*>(c->amat_h3);
    gh.ascale = c->ascale_h3;
    gh.out = c->rec_part;
    void* gargs[] = {(void*)&gh};
    dim3 gg(div_up(c->n, 4 * 32 * c->h3_ct), nsplit);
    PROF_LAUNCH(c, PK_GATHER, (void)hipLaunchKernel(gather_h3_kernel(c->h3_ct, nsh), gg, dim3(256), gargs, 0,
                                                    c->stream));
  } else if (mma_gather) {
    // matrix-core path (A operand written by k_sk_reduce)
    GmArgs gm{};
    gm.n = c->n;
    gm.Kx = c->K[0]; gm.Ky = c->K[1]; gm.Kz = c->K[2];
    gm.NY = c->NYk; gm.NZ = c->NZk; gm.M = s.M;
    gm.nks = c->mma_nks;
    gm.mtiles = c->mma_mtiles;
    gm.mt_per_split = div_up(c->mma_mtiles, c->mma_splits);
    nsplit = div_up(c->mma_mtiles, gm.mt_per_split);
    gm.tab = c->tab;
    gm.amat = c->amat;
    gm.out = c->rec_part;
    const int CT = c->mma_ct;
    dim3 gg(div_up(c->n, 4 * 16 * CT), nsplit);
    MDFF_TRY(launch_gather_mma<T>(c, gm, CT, gg));
  } else {
    GaArgs g{};
    g.n = c->n;
    g.Kx = c->K[0]; g.Ky = c->K[1]; g.Kz = c->K[2];
    g.NY = c->NYk; g.NZ = c->NZk;
    g.M = s.M;
    g.m_per_split = div_up(s.M, c->rec_splits);
    nsplit = div_up(s.M, g.m_per_split);
    g.moments = moments;
    g.tab = c->tab;
    g.wk = c->wk_t;
    g.out = c->rec_part;
    // threads per block from the LDS budget (64 KiB)
    int nthr = 256;
    while (nthr > 64 && sizeof(cplx<T>) * ((size_t)(c->K[2] + 1) * nthr + 8 * (size_t)c->NZk) > 65536) nthr -= 64;
    const size_t shm2 = sizeof(cplx<T>) * ((size_t)(c->K[2] + 1) * nthr + 8 * (size_t)c->NZk);
    if (shm2 > 65536) {
      c->err = "kES(3) too large for the reciprocal gather kernel";
      return MDFF_ERR_UNSUPPORTED;
    }
    dim3 g2(div_up(c->n, nthr), nsplit);
    if (dipf) PROF_LAUNCH(c, PK_GATHER, hipLaunchKernelGGL((k_gather<T, 9>), g2, dim3(nthr), shm2, c->stream, g));
    else PROF_LAUNCH(c, PK_GATHER, hipLaunchKernelGGL((k_gather<T, 3>), g2, dim3(nthr), shm2, c->stream, g));
  }
  if (inner) {
    // the inner box's moments: kin_splits more splits after the main ones
    ia.out = c->rec_part + (size_t)nsplit * moments * c->n;
    ia.msplit = c->kin_msplit;
    const size_t lds = (size_t)kInMT * ia.NZ * 16;
    const dim3 gg(div_up(c->n, 256), c->kin_splits, c->kin_jb);
    if (c->kin_reg) {
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(in_gather_kernel(c->kin_kb, moments, c->kin_jb > 1), gg, dim3(256),
                                               0, c->stream, ia));
    } else {
      PROF_LAUNCH(c, PK_INNER, hipLaunchKernelGGL(k_in_gather, gg, dim3(256), lds, c->stream, ia));
    }
    nsplit += c->kin_splits * c->kin_jb;
  }
  RecFinal rf{};
  rf.n = c->n;
  rf.nsplit = nsplit;
  rf.moments = moments;
  rf.use_q = use_q ? 1 : 0;
  rf.part = c->rec_part;
  rf.q = q;
  rf.mu = mu;
  for (int k = 0; k < 9; ++k) rf.b[k] = c->B[k];
  rf.fpi_V = 2.0 * (2.0 * MDFF_PI / c->omega);
  // forces wanted with no output array: the caller's force kernel finishes
  // them from the split partials (c->rfin)
  const bool defer = forces && !f_out;
  if (defer) {
    c->rfin = rf;
    c->rfin_pending = 1;
  }
  if (field || !defer)
    PROF_LAUNCH(c, PK_GFINAL,
                hipLaunchKernelGGL(k_gather_final, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, rf,
                                   field ? 1 : 0, (forces && !defer) ? 1 : 0, ef_out, f_out));
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int recip_eval(mdff_ctx* c, const double* q, const double* mu, bool use_q, bool use_mu, bool forces,
               bool field, double* ef_out, double* f_out, double* e_out) {
  if (c->rec_f32) return eval_t<float>(c, q, mu, use_q, use_mu, forces, field, ef_out, f_out, e_out);
  return eval_t<double>(c, q, mu, use_q, use_mu, forces, field, ef_out, f_out, e_out);
}
