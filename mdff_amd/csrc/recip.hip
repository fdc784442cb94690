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
#include "common.h"

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
template <typename T>
__global__ void k_tables(int n, int Kx, int Ky, int Kz, const double* __restrict__ sx,
                         const double* __restrict__ sy, const double* __restrict__ sz,
                         cplx<T>* __restrict__ tab) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int col = blockIdx.y;
  if (i >= n) return;
  double s;
  int k;
  if (col <= Kx) { s = sx[i]; k = col; }
  else if (col <= Kx + Ky + 1) { s = sy[i]; k = col - Kx - 1; }
  else { s = sz[i]; k = col - Kx - Ky - 2; }
  double t = (double)k * s;
  t -= rint(t);
  double sn, cs;
  sincospi(2.0 * t, &sn, &cs);
  tab[(size_t)col * n + i] = {(T)cs, (T)sn};
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

// S(n) = sum of chunk partials (FP64, fixed order); W(n) = A conj(S);
// per-block sums of A|S|^2 and (delta_ab - kcoe k_a k_b) A|S|^2.
template <typename T>
__global__ void __launch_bounds__(256) k_sk_reduce(int nk_box, int nchunk, const T* __restrict__ part,
                                                   const double* __restrict__ kA,
                                                   const double* __restrict__ kvec,
                                                   double* __restrict__ sk, double* __restrict__ wk,
                                                   T* __restrict__ wk_t, double* __restrict__ blk) {
  __shared__ double sh[7][256];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (k < nk_box) {
    double re = 0.0, im = 0.0;
    for (int c = 0; c < nchunk; ++c) {
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
  for (int c = 0; c < 7; ++c) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 7; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 7) blk[blockIdx.x * 7 + threadIdx.x] = sh[threadIdx.x][0];
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
__global__ void k_gather_final(int n, int nsplit, int moments, const double* __restrict__ part,
                               const double* __restrict__ q, const double* __restrict__ mu, double b11,
                               double b21, double b31, double b12, double b22, double b32, double b13,
                               double b23, double b33, double fpi_V, int use_q, int field, int forces,
                               double* __restrict__ ef, double* __restrict__ f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = 0; s < nsplit; ++s)
    for (int c = 0; c < moments; ++c) v[c] += part[((size_t)s * moments + c) * n + i];
  const double tp = 2.0 * MDFF_PI;
  // B U1 (Cartesian), B column-major: (B u)_a = sum_d B(a,d) u_d
  const double gx = b11 * v[0] + b12 * v[1] + b13 * v[2];
  const double gy = b21 * v[0] + b22 * v[1] + b23 * v[2];
  const double gz = b31 * v[0] + b32 * v[1] + b33 * v[2];
  const double sc = 2.0 * fpi_V;
  if (field) {
    ef[3 * i] = sc * tp * gx;
    ef[3 * i + 1] = sc * tp * gy;
    ef[3 * i + 2] = sc * tp * gz;
  }
  if (forces) {
    double qi = use_q ? q[i] : 0.0;
    double fxv = qi * tp * gx, fyv = qi * tp * gy, fzv = qi * tp * gz;
    if (moments == 9) {
      const double u0 = mu[3 * i], u1 = mu[3 * i + 1], u2 = mu[3 * i + 2];
      // m = B^T mu
      const double m0 = u0 * b11 + u1 * b21 + u2 * b31;
      const double m1 = u0 * b12 + u1 * b22 + u2 * b32;
      const double m2 = u0 * b13 + u1 * b23 + u2 * b33;
      const double p0 = v[3] * m0 + v[6] * m1 + v[7] * m2;
      const double p1 = v[6] * m0 + v[4] * m1 + v[8] * m2;
      const double p2 = v[7] * m0 + v[8] * m1 + v[5] * m2;
      const double t2 = tp * tp;
      fxv += t2 * (b11 * p0 + b12 * p1 + b13 * p2);
      fyv += t2 * (b21 * p0 + b22 * p1 + b23 * p2);
      fzv += t2 * (b31 * p0 + b32 * p1 + b33 * p2);
    }
    f[3 * i] = sc * fxv;
    f[3 * i + 1] = sc * fyv;
    f[3 * i + 2] = sc * fzv;
  }
}

}  // namespace

// k-mesh on the box with the reference's k vectors and A(n), kcoe
// (src/kspace.f90:162-191); allocation of the reciprocal buffers.
int recip_setup(mdff_ctx* c) {
  const int Kx = c->d.kES[0], Ky = c->d.kES[1], Kz = c->d.kES[2];
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
  const size_t tsz = c->d.precision == MDFF_PREC_MIXED ? sizeof(cplx<float>) : sizeof(cplx<double>);
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
  HIP_TRY(c, hipMalloc(&c->sk_part, tsz * (size_t)chunks * nb));
  HIP_TRY(c, hipMalloc((void**)&c->sk, sizeof(double) * 2 * nb));
  HIP_TRY(c, hipMalloc((void**)&c->wk, sizeof(double) * 2 * nb));
  HIP_TRY(c, hipMalloc(&c->wk_t, tsz * (size_t)nb));
  // gather splits over (nx, ny): about 256k threads in flight
  const int M = c->NXk * c->NYk;
  int splits = std::max(1, std::min(M, div_up((int64_t)262144, c->n)));
  c->rec_splits = splits;
  HIP_TRY(c, hipMalloc((void**)&c->rec_part, sizeof(double) * 9 * (size_t)splits * c->n));
  return MDFF_OK;
}

void recip_free(mdff_ctx* c) {
  void* ps[] = {c->kA, c->kvec, c->tab, c->sk_part, c->sk, c->wk, c->wk_t, c->rec_part};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

template <typename T>
static void tables_t(mdff_ctx* c) {
  const int ncol = c->K[0] + c->K[1] + c->K[2] + 3;
  hipLaunchKernelGGL((k_tables<T>), dim3(div_up(c->n, 256), ncol), dim3(256), 0, c->stream, c->n, c->K[0],
                     c->K[1], c->K[2], c->sx, c->sy, c->sz, reinterpret_cast<cplx<T>*>(c->tab));
}

int recip_tables(mdff_ctx* c) {
  if (c->d.precision == MDFF_PREC_MIXED) tables_t<float>(c);
  else tables_t<double>(c);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
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
  const int chunks = div_up(c->n, s.a0_step);
  dim3 g1(div_up((int64_t)s.M * s.NG, 256), chunks);
  hipLaunchKernelGGL((k_sk<T, GZ>), g1, dim3(256), shm, c->stream, s);
  const int nb = div_up(c->nk_box, 256);
  hipLaunchKernelGGL((k_sk_reduce<T>), dim3(nb), dim3(256), 0, c->stream, c->nk_box, chunks,
                     reinterpret_cast<const T*>(c->sk_part), c->kA, c->kvec, c->sk, c->wk,
                     reinterpret_cast<T*>(c->wk_t), c->red_part);
  if (e_out) {
    // 7 sums of the k-blocks, reduced on the host in block order
    std::vector<double> blk(7 * (size_t)nb);
    HIP_TRY(c, hipMemcpyAsync(blk.data(), c->red_part, sizeof(double) * 7 * nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 7; ++k) e_out[k] = 0.0;
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < 7; ++k) e_out[k] += blk[7 * (size_t)b + k];
  }
  if (!forces && !field) return MDFF_OK;
  GaArgs g{};
  g.n = c->n;
  g.Kx = c->K[0]; g.Ky = c->K[1]; g.Kz = c->K[2];
  g.NY = c->NYk; g.NZ = c->NZk;
  g.M = s.M;
  g.m_per_split = div_up(s.M, c->rec_splits);
  const int nsplit = div_up(s.M, g.m_per_split);
  const bool dipf = forces && use_mu;
  g.moments = dipf ? 9 : 3;
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
  if (dipf) hipLaunchKernelGGL((k_gather<T, 9>), g2, dim3(nthr), shm2, c->stream, g);
  else hipLaunchKernelGGL((k_gather<T, 3>), g2, dim3(nthr), shm2, c->stream, g);
  const double* B = c->B;
  const double fpi_V = 2.0 * (2.0 * MDFF_PI / c->omega);
  hipLaunchKernelGGL(k_gather_final, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, c->n, nsplit, g.moments,
                     c->rec_part, q, mu, B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], fpi_V,
                     use_q ? 1 : 0, field ? 1 : 0, forces ? 1 : 0, ef_out, f_out);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int recip_eval(mdff_ctx* c, const double* q, const double* mu, bool use_q, bool use_mu, bool forces,
               bool field, double* ef_out, double* f_out, double* e_out) {
  if (c->d.precision == MDFF_PREC_MIXED)
    return eval_t<float>(c, q, mu, use_q, use_mu, forces, field, ef_out, f_out, e_out);
  return eval_t<double>(c, q, mu, use_q, use_mu, forces, field, ef_out, f_out, e_out);
}
