// nlist.hip -- cell-list Verlet build and skin check for gfx950.
//
// Replaces vnlist_pbc (src/tools.f90:273-368) and vnlistcheck
// (src/tools.f90:444-537).  The reference tests every ordered pair of the
// O(N^2) sweep with the ownership rule; here atoms are binned into cells in
// fractional space (cell width >= cutoff / perpendicular width, so every
// partner lies in the 27 neighbouring cells) and one wavefront per atom
// sweeps those cells.  Each device row holds ALL partners j != i with
// r^2 <= (max cut + skin)^2, plus a flag byte saying which of the two
// reference lists (vdw, coul) accepts the pair.  The reference half lists are
// the ownership-filtered rows; r^2(i,j) and r^2(j,i) are bitwise equal (the
// fractional difference only changes sign), so the filtered set is exactly
// the reference pair set.
//
// This file is compiled with -ffp-contract=off: r^2 must be computed with the
// reference's rounding (s = B^T r, d - nint(d), A.s, x*x+y*y+z*z) so that
// the <= decisions match the reference bit for bit.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.h"
#include "reduce.h"

// kardir (src/lattice.f90:183-216): s = B^T r in the reference's order.
__global__ void k_kardir(int n, const double* __restrict__ rx, const double* __restrict__ ry,
                         const double* __restrict__ rz, double b11, double b21, double b31,
                         double b12, double b22, double b32, double b13, double b23, double b33,
                         double* __restrict__ sx, double* __restrict__ sy, double* __restrict__ sz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = rx[i], y = ry[i], z = rz[i];
  sx[i] = x * b11 + y * b21 + z * b31;
  sy[i] = x * b12 + y * b22 + z * b32;
  sz[i] = x * b13 + y * b23 + z * b33;
}

static void kardir_to(mdff_ctx* c, double* ox, double* oy, double* oz) {
  const double* B = c->B;
  hipLaunchKernelGGL(k_kardir, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, c->n, c->rx, c->ry,
                     c->rz, B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], ox, oy, oz);
}

void launch_kardir(mdff_ctx* c) { kardir_to(c, c->sx, c->sy, c->sz); }

// atoms in the cell order of the last list build: rec[p], meta[p] of atom
// order[p] (the rows store sorted positions p, so neighbours that are close
// in space are close in memory)
__global__ void k_pack_atoms(int n, const int32_t* __restrict__ order, const double* __restrict__ sx,
                             const double* __restrict__ sy, const double* __restrict__ sz,
                             const double* __restrict__ q, const int32_t* __restrict__ itype,
                             AtomRec* __restrict__ rec, int2* __restrict__ meta) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = order[p];
  rec[p] = AtomRec{sx[i], sy[i], sz[i], q[i]};
  meta[p] = make_int2(itype[i], i);
}

void pack_atoms(mdff_ctx* c) {
  hipLaunchKernelGGL(k_pack_atoms, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cell_atoms, c->sx,
                     c->sy, c->sz, c->q, c->itype, c->arec, c->ameta);
}

__device__ __forceinline__ int wrap_cell(double s, int nc) {
  double w = s - floor(s);
  int k = (int)(w * (double)nc);
  return k < 0 ? 0 : (k >= nc ? nc - 1 : k);
}

__global__ void k_cell_assign(int n, const double* __restrict__ sx, const double* __restrict__ sy,
                              const double* __restrict__ sz, int nx, int ny, int nz,
                              int32_t* __restrict__ cell_of, int32_t* __restrict__ iota) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx = wrap_cell(sx[i], nx), cy = wrap_cell(sy[i], ny), cz = wrap_cell(sz[i], nz);
  cell_of[i] = (cx * ny + cy) * nz + cz;
  iota[i] = i;
}

// start offset of every cell from the sorted keys (empty cells included)
__global__ void k_cell_start(int n, int ncell, const int32_t* __restrict__ key,
                             int32_t* __restrict__ start) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int k = key[p];
  int prev = (p == 0) ? -1 : key[p - 1];
  for (int c = prev + 1; c <= k; ++c) start[c] = p;
  if (p == n - 1)
    for (int c = k + 1; c <= ncell; ++c) start[c] = n;
}

__global__ void k_gather_sorted(int n, const int32_t* __restrict__ order, const double* __restrict__ sx,
                                const double* __restrict__ sy, const double* __restrict__ sz,
                                double* __restrict__ ox, double* __restrict__ oy, double* __restrict__ oz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = order[p];
  ox[p] = sx[i];
  oy[p] = sy[i];
  oz[p] = sz[i];
}

struct NLArgs {
  int n, nx, ny, nz;
  double a11, a21, a31, a12, a22, a32, a13, a23, a33;  // simu_cell%A, A(i,j)
  double cut2_max, cut2_v, cut2_c;
  const double *sx, *sy, *sz;             // fractional coords, original order
  const double *px, *py, *pz;             // fractional coords, cell order
  const int32_t* cell_atoms;
  const int32_t* cell_start;
  const int32_t* cell_of;
  int32_t* count;
  const int32_t* row_start;
  int32_t* nl_j;
  uint8_t* nl_flag;
  int64_t cap;
  int32_t* overflow;
};

// r^2 of the minimum image of a fractional difference, in the reference order
// (src/tools.f90:322-333): ds - nint(ds), then A.s row by row, then x*x+y*y+z*z.
__device__ __forceinline__ double min_image_r2(const NLArgs& a, double dx, double dy, double dz) {
  double sxx = dx - nint_d(dx), syy = dy - nint_d(dy), szz = dz - nint_d(dz);
  double rx = sxx * a.a11 + syy * a.a12 + szz * a.a13;
  double ry = sxx * a.a21 + syy * a.a22 + szz * a.a23;
  double rz = sxx * a.a31 + syy * a.a32 + szz * a.a33;
  return rx * rx + ry * ry + rz * rz;
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_nl_sweep(NLArgs a) {
  // The neighbour cells' candidate ranges go to LDS once; the candidates are
  // then taken in chunks of 64 (cell by cell, in cell order, as before) and
  // the loads of the next chunk are issued before the current chunk is
  // tested, so the wave does not wait on memory at every chunk.
  __shared__ int crange[4][2][27];
  const int lane = threadIdx.x & 63;
  const int wl = threadIdx.x >> 6;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= a.n) return;                 // whole waves: no block barrier below
  const int i = a.cell_atoms[wave];        // atoms visited in cell order (L2 locality)
  const double si = a.sx[i], ti = a.sy[i], ui = a.sz[i];
  const int ci = a.cell_of[i];
  const int cz = ci % a.nz, cy = (ci / a.nz) % a.ny, cx = ci / (a.nz * a.ny);
  int64_t base = FILL ? (int64_t)a.row_start[i] : 0;
  const int ox0 = a.nx >= 3 ? -1 : 0, ox1 = a.nx >= 2 ? 1 : 0;
  const int oy0 = a.ny >= 3 ? -1 : 0, oy1 = a.ny >= 2 ? 1 : 0;
  const int oz0 = a.nz >= 3 ? -1 : 0, oz1 = a.nz >= 2 ? 1 : 0;
  const int ncx = ox1 - ox0 + 1, ncy = oy1 - oy0 + 1, ncz = oz1 - oz0 + 1;
  const int ncell = ncx * ncy * ncz;
  if (lane < ncell) {
    // cell k in the order of the nested dx, dy, dz loops
    const int dz = oz0 + lane % ncz, dy = oy0 + (lane / ncz) % ncy, dx = ox0 + lane / (ncz * ncy);
    const int xx = (cx + dx + a.nx) % a.nx, yy = (cy + dy + a.ny) % a.ny, zz = (cz + dz + a.nz) % a.nz;
    const int c = (xx * a.ny + yy) * a.nz + zz;
    crange[wl][0][lane] = a.cell_start[c];
    crange[wl][1][lane] = a.cell_start[c + 1];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // first non-empty chunk
  int k = 0, p = 0;
  while (k < ncell && crange[wl][0][k] >= crange[wl][1][k]) ++k;
  if (k < ncell) p = crange[wl][0][k];
  int found = 0;
  // chunk loads: candidate j and its cell-ordered coordinates (clamped)
  int jn = -1;
  double xn = 0.0, yn = 0.0, zn = 0.0;
  int pend = k < ncell ? crange[wl][1][k] : 0;
  if (k < ncell && p + lane < pend) {
    const int q = p + lane;
    jn = a.cell_atoms[q];
    xn = a.px[q];
    yn = a.py[q];
    zn = a.pz[q];
  }
  while (k < ncell) {
    const int pc = p, pe = pend;
    const int j = jn;
    const double xj = xn, yj = yn, zj = zn;
    // next chunk: the rest of this cell, else the next non-empty cell
    p += 64;
    if (p >= pend) {
      ++k;
      while (k < ncell && crange[wl][0][k] >= crange[wl][1][k]) ++k;
      if (k < ncell) {
        p = crange[wl][0][k];
        pend = crange[wl][1][k];
      }
    }
    jn = -1;
    if (k < ncell && p + lane < pend) {
      const int q = p + lane;
      jn = a.cell_atoms[q];
      xn = a.px[q];
      yn = a.py[q];
      zn = a.pz[q];
    }
    // test the current chunk
    bool ok = false;
    uint8_t flag = 0;
    const int pp = pc + lane;
    if (pp < pe && j != i) {
      const double r2 = min_image_r2(a, si - xj, ti - yj, ui - zj);
      ok = (r2 <= a.cut2_max);
      flag = (uint8_t)((r2 <= a.cut2_v ? 1 : 0) | (r2 <= a.cut2_c ? 2 : 0));
    }
    const uint64_t m = __ballot(ok);
    if (FILL && ok) {
      const int off = found + __popcll(m & ((1ull << lane) - 1ull));
      const int64_t e = base + off;
      if (e < a.cap) {
        a.nl_j[e] = pp;                  // sorted position of j
        a.nl_flag[e] = flag;
      }
    }
    found += __popcll(m);
  }
  if (!FILL && lane == 0) a.count[i] = found;
  if (FILL && lane == 0 && base + found > a.cap) atomicExch(a.overflow, 1);
}

static int ensure(mdff_ctx* c, void** p, size_t& have, size_t need) {
  if (have >= need) return MDFF_OK;
  if (*p) HIP_TRY(c, hipFree(*p));
  *p = nullptr;
  HIP_TRY(c, hipMalloc(p, need));
  have = need;
  return MDFF_OK;
}

static NLArgs make_args(mdff_ctx* c) {
  NLArgs a{};
  a.n = c->n;
  a.nx = c->ncell[0];
  a.ny = c->ncell[1];
  a.nz = c->ncell[2];
  const double* A = c->A;
  a.a11 = A[0]; a.a21 = A[1]; a.a31 = A[2];
  a.a12 = A[3]; a.a22 = A[4]; a.a32 = A[5];
  a.a13 = A[6]; a.a23 = A[7]; a.a33 = A[8];
  // (cut + skindiff)^2 exactly as src/tools.f90:298-301
  double rv = c->rc_vdw + c->d.skindiff, rcl = c->rc_coul + c->d.skindiff;
  a.cut2_v = rv * rv;
  a.cut2_c = rcl * rcl;
  a.cut2_max = a.cut2_v > a.cut2_c ? a.cut2_v : a.cut2_c;
  a.sx = c->sx; a.sy = c->sy; a.sz = c->sz;
  a.cell_atoms = c->cell_atoms;
  a.cell_start = c->cell_start;
  a.cell_of = c->cell_of;
  a.count = c->nl_count;
  a.row_start = c->nl_start;
  a.nl_j = c->nl_j;
  a.nl_flag = c->nl_flag;
  a.cap = c->nl_cap;
  a.overflow = c->d_flag + 1;
  return a;
}

// vnlist_pbc: rebuild the rows from the current positions.
// list build from the fractional coordinates (px, py, pz) of the build
static int nl_build_from(mdff_ctx* c, const double* px, const double* py, const double* pz) {
  const int n = c->n;
  // cells: width >= cutoff in every reciprocal direction (1e-9 margin against
  // rounding of the wrapped coordinates)
  double rmax = std::max(c->rc_vdw, c->rc_coul) + c->d.skindiff;
  for (int d = 0; d < 3; ++d) {
    int k = (int)floor(c->W[d] / (rmax * (1.0 + 1e-9)));
    c->ncell[d] = k < 1 ? 1 : k;
  }
  int ncell = c->ncell[0] * c->ncell[1] * c->ncell[2];
  if (ncell + 1 > c->ncell_alloc) {
    if (c->cell_start) HIP_TRY(c, hipFree(c->cell_start));
    HIP_TRY(c, hipMalloc((void**)&c->cell_start, sizeof(int32_t) * (ncell + 1)));
    c->ncell_alloc = ncell + 1;
  }
  hipLaunchKernelGGL(k_cell_assign, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, px, py, pz,
                     c->ncell[0], c->ncell[1], c->ncell[2], c->cell_of, c->cell_atoms_in);
  int bits = 1;
  while ((1 << bits) <= ncell) ++bits;
  size_t tmp = 0;
  HIP_TRY(c, rocprim::radix_sort_pairs(nullptr, tmp, c->cell_of, c->cell_key, c->cell_atoms_in,
                                       c->cell_atoms, (unsigned)n, 0u, (unsigned)bits, c->stream));
  size_t tmp2 = 0;
  HIP_TRY(c, rocprim::exclusive_scan(nullptr, tmp2, c->nl_count, c->nl_start, 0, (size_t)n + 1,
                                     rocprim::plus<int32_t>(), c->stream));
  MDFF_TRY(ensure(c, &c->sort_tmp, c->sort_tmp_bytes, std::max(tmp, tmp2)));
  HIP_TRY(c, rocprim::radix_sort_pairs(c->sort_tmp, tmp, c->cell_of, c->cell_key, c->cell_atoms_in,
                                       c->cell_atoms, (unsigned)n, 0u, (unsigned)bits, c->stream));
  hipLaunchKernelGGL(k_cell_start, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, ncell, c->cell_key,
                     c->cell_start);
  // fractional coordinates in cell order (coalesced candidate loads); they
  // are kept as the positions of this build
  hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, c->cell_atoms,
                     px, py, pz, c->bx, c->by, c->bz);
  NLArgs a = make_args(c);
  a.sx = px; a.sy = py; a.sz = pz;
  a.px = c->bx; a.py = c->by; a.pz = c->bz;
  dim3 grid(div_up((int64_t)n * 64, 256));
  HIP_TRY(c, hipMemsetAsync(c->nl_count + n, 0, sizeof(int32_t), c->stream));
  PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_sweep<false>, grid, dim3(256), 0, c->stream, a));
  HIP_TRY(c, rocprim::exclusive_scan(c->sort_tmp, tmp2, c->nl_count, c->nl_start, 0, (size_t)n + 1,
                                     rocprim::plus<int32_t>(), c->stream));
  int32_t total = 0;
  HIP_TRY(c, hipMemcpyAsync(&total, c->nl_start + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (total < 0) { c->err = "neighbour list exceeds 2^31 entries"; return MDFF_ERR_LIST_OVERFLOW; }
  // reference capacity: vnlmax*natm entries per half list (src/config.f90:43,
  // src/tools.f90:337-340); the full rows hold twice the half-list entries
  if ((double)total > 2.0 * 2000.0 * (double)n) {
    c->err = "ERROR: out of bound list in vnlist_pbc";
    return MDFF_ERR_LIST_OVERFLOW;
  }
  if (total > c->nl_cap) {
    int64_t cap = (int64_t)(total * 1.25) + 1024;
    if (c->nl_j) HIP_TRY(c, hipFree(c->nl_j));
    if (c->nl_flag) HIP_TRY(c, hipFree(c->nl_flag));
    HIP_TRY(c, hipMalloc((void**)&c->nl_j, sizeof(int32_t) * cap));
    HIP_TRY(c, hipMalloc((void**)&c->nl_flag, cap));
    c->nl_cap = cap;
    a.nl_j = c->nl_j;
    a.nl_flag = c->nl_flag;
    a.cap = cap;
  }
  HIP_TRY(c, hipMemsetAsync(c->d_flag + 1, 0, sizeof(int32_t), c->stream));
  PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_sweep<true>, grid, dim3(256), 0, c->stream, a));
  HIP_TRY(c, hipGetLastError());
  c->nl_total = total;
  c->lists_valid = 1;
  return MDFF_OK;
}

int nl_build(mdff_ctx* c) {
  launch_kardir(c);
  return nl_build_from(c, c->sx, c->sy, c->sz);
}

// the two largest of the block's values m1 (m2 = 0 on entry) -> out[2 * block]
// (all threads of the block take part)
__device__ inline void block_top2(double m1, double m2, double* s1, double* s2, double* __restrict__ out) {
  s1[threadIdx.x] = m1;
  s2[threadIdx.x] = m2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double a1 = s1[threadIdx.x], a2 = s2[threadIdx.x];
      double b1 = s1[threadIdx.x + w], b2 = s2[threadIdx.x + w];
      double hi = a1 > b1 ? a1 : b1;
      double lo = a1 > b1 ? b1 : a1;
      double sec = a2 > b2 ? a2 : b2;
      s1[threadIdx.x] = hi;
      s2[threadIdx.x] = lo > sec ? lo : sec;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = s1[0];
    out[2 * blockIdx.x + 1] = s2[0];
  }
}

// displacement of atom i since the saved fractional coordinates
// (src/tools.f90:470-500); its fractional coordinates (kardir's arithmetic)
// go to (cx, cy, cz) for a rebuild at these positions
__device__ __forceinline__ double check_disp(const NLArgs& a, int i, double x, double y, double z, const double* b,
                                             double* __restrict__ cx, double* __restrict__ cy,
                                             double* __restrict__ cz, const double* __restrict__ xs,
                                             const double* __restrict__ ys, const double* __restrict__ zs) {
  const double u = x * b[0] + y * b[1] + z * b[2];
  const double v = x * b[3] + y * b[4] + z * b[5];
  const double w = x * b[6] + y * b[7] + z * b[8];
  cx[i] = u;
  cy[i] = v;
  cz[i] = w;
  return sqrt(min_image_r2(a, u - xs[i], v - ys[i], w - zs[i]));
}

struct CheckArgs {
  NLArgs a;
  double b[9];                    // B column-major
  double *cx, *cy, *cz;
  const double *xs, *ys, *zs;
  double* out;                    // [block][2] (mapped host memory)
};

// per-atom displacement, reduced to the two largest values per block
__global__ void __launch_bounds__(256) k_disp_top2(CheckArgs k, const double* __restrict__ rx,
                                                   const double* __restrict__ ry,
                                                   const double* __restrict__ rz) {
  __shared__ double s1[256], s2[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m1 = 0.0;
  if (i < k.a.n) m1 = check_disp(k.a, i, rx[i], ry[i], rz[i], k.b, k.cx, k.cy, k.cz, k.xs, k.ys, k.zs);
  block_top2(m1, 0.0, s1, s2, k.out);
}

// prop_velocity_verlet position update (src/integration.f90:165-169):
// f_saved = f, r += v dt + f dt^2/2 / m
__global__ void k_vv_pos(int n, double dt, double dtsq2, const double* __restrict__ m, double* rx, double* ry,
                         double* rz, const double* vx, const double* vy, const double* vz, const double* fx,
                         const double* fy, const double* fz, double* fsx, double* fsy, double* fsz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fsx[i] = fx[i];
  fsy[i] = fy[i];
  fsz[i] = fz[i];
  rx[i] = rx[i] + vx[i] * dt + (fx[i] * dtsq2) / m[i];
  ry[i] = ry[i] + vy[i] * dt + (fy[i] * dtsq2) / m[i];
  rz[i] = rz[i] + vz[i] * dt + (fz[i] * dtsq2) / m[i];
}

// End of a velocity Verlet step (src/integration.f90:178-185), the list
// check after it (vnlistcheck, src/tools.f90:444-537) and, when another step
// follows, that step's position update (:165-169), one thread per atom:
//   v += (f_saved + f) dt/2 / m; ke[block] = sum m v^2 over the block;
//   the check's displacement of the positions of this step (block top-2);
//   then f_saved = f and r += v dt + f dt^2/2 / m.
struct VvEndArgs {
  CheckArgs k;
  int n, check, next;
  double dt2, dt, dtsq2;
  const double* m;
  double *rx, *ry, *rz, *vx, *vy, *vz;
  double *fsx, *fsy, *fsz;
  const double *fx, *fy, *fz;
  double* ke;                     // [block]
};
__global__ void __launch_bounds__(256) k_vv_end(VvEndArgs e) {
  __shared__ double sh[4 * kRedMaxCols];
  __shared__ double s1[256], s2[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double kv[kRedMaxCols] = {};
  double m1 = 0.0;
  if (i < e.n) {
    const double m = e.m[i];
    const double fx = e.fx[i], fy = e.fy[i], fz = e.fz[i];
    const double a = e.vx[i] + (e.fsx[i] + fx) * e.dt2 / m;
    const double b = e.vy[i] + (e.fsy[i] + fy) * e.dt2 / m;
    const double cc = e.vz[i] + (e.fsz[i] + fz) * e.dt2 / m;
    e.vx[i] = a;
    e.vy[i] = b;
    e.vz[i] = cc;
    kv[0] = (a * a + b * b + cc * cc) * m;
    const double x = e.rx[i], y = e.ry[i], z = e.rz[i];
    if (e.check) m1 = check_disp(e.k.a, i, x, y, z, e.k.b, e.k.cx, e.k.cy, e.k.cz, e.k.xs, e.k.ys, e.k.zs);
    if (e.next) {
      e.fsx[i] = fx;
      e.fsy[i] = fy;
      e.fsz[i] = fz;
      e.rx[i] = x + a * e.dt + (fx * e.dtsq2) / m;
      e.ry[i] = y + b * e.dt + (fy * e.dtsq2) / m;
      e.rz[i] = z + cc * e.dt + (fz * e.dtsq2) / m;
    }
  }
  const double r = block_sum_cols(kv, 1, sh);
  if (threadIdx.x == 0) e.ke[blockIdx.x] = r;
  if (e.check) block_top2(m1, 0.0, s1, s2, e.k.out);
}

__global__ void k_copy3(int n, const double* a, const double* b, const double* cc, double* x, double* y,
                        double* z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = a[i];
  y[i] = b[i];
  z[i] = cc[i];
}

// vnlistcheck: rebuild when (largest + second largest displacement) > skindiff,
// then save the fractional coordinates (src/tools.f90:513-526).  Split in two
// halves so that a caller can queue more work between them: the launch half
// (nl_check_launch, or vv_end_launch at the end of an MD step) queues the
// displacement check of the current positions (their fractional coordinates
// go to cx, cy, cz; its block maxima to mapped host memory); the finishing
// half waits for it and, when the skin is exceeded, builds the lists from
// those same coordinates (the positions of the check, whatever the positions
// are by then).
static CheckArgs check_args(mdff_ctx* c) {
  CheckArgs k{};
  k.a = make_args(c);
  for (int j = 0; j < 9; ++j) k.b[j] = c->B[j];
  k.cx = c->cx; k.cy = c->cy; k.cz = c->cz;
  k.xs = c->xs; k.ys = c->ys; k.zs = c->zs;
  k.out = c->d_chk;
  return k;
}

// the check's block maxima are written straight into mapped host memory
// (h_chk); the event says when they are there
static int check_posted(mdff_ctx* c) {
  HIP_TRY(c, hipEventRecord(c->chk_ev, c->stream));
  c->chk_pending = 1;
  return MDFF_OK;
}

int nl_check_launch(mdff_ctx* c) {
  const int nb = div_up(c->n, 256);
  hipLaunchKernelGGL(k_disp_top2, dim3(nb), dim3(256), 0, c->stream, check_args(c), c->rx, c->ry, c->rz);
  HIP_TRY(c, hipGetLastError());
  return check_posted(c);
}

int vv_pos_launch(mdff_ctx* c, double dt) {
  hipLaunchKernelGGL(k_vv_pos, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, c->n, dt, dt * dt * 0.5, c->mass,
                     c->rx, c->ry, c->rz, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz, c->fsx, c->fsy, c->fsz);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int vv_end_launch(mdff_ctx* c, double dt, bool next, double* ke_part) {
  VvEndArgs e{};
  e.k = check_args(c);
  e.n = c->n;
  e.check = c->d.lvnlist ? 1 : 0;
  e.next = next ? 1 : 0;
  e.dt2 = dt * 0.5;
  e.dt = dt;
  e.dtsq2 = dt * dt * 0.5;
  e.m = c->mass;
  e.rx = c->rx; e.ry = c->ry; e.rz = c->rz;
  e.vx = c->vx; e.vy = c->vy; e.vz = c->vz;
  e.fsx = c->fsx; e.fsy = c->fsy; e.fsz = c->fsz;
  e.fx = c->fx; e.fy = c->fy; e.fz = c->fz;
  e.ke = ke_part;
  hipLaunchKernelGGL(k_vv_end, dim3(div_up(c->n, 256)), dim3(256), 0, c->stream, e);
  HIP_TRY(c, hipGetLastError());
  return e.check ? check_posted(c) : MDFF_OK;
}

int nl_check_finish(mdff_ctx* c, int* rebuilt) {
  *rebuilt = 0;
  if (!c->chk_pending) return MDFF_OK;
  c->chk_pending = 0;
  const int n = c->n;
  const int nb = div_up(n, 256);
  HIP_TRY(c, hipEventSynchronize(c->chk_ev));
  double m1 = 0.0, m2 = 0.0;
  for (int b = 0; b < nb; ++b) {
    double v[2] = {c->h_chk[2 * b], c->h_chk[2 * b + 1]};
    for (double d : v) {
      if (d > m1) { m2 = m1; m1 = d; }
      else if (d > m2) m2 = d;
    }
  }
  if (m1 + m2 > c->d.skindiff) {
    MDFF_TRY(nl_build_from(c, c->cx, c->cy, c->cz));
    hipLaunchKernelGGL(k_copy3, dim3(nb), dim3(256), 0, c->stream, n, c->cx, c->cy, c->cz, c->xs, c->ys,
                       c->zs);
    *rebuilt = 1;
  }
  return MDFF_OK;
}

int nl_check(mdff_ctx* c, int* rebuilt) {
  MDFF_TRY(nl_check_launch(c));
  return nl_check_finish(c, rebuilt);
}

// Reference CSR half lists from the device rows: ownership rule
// (src/tools.f90:318-319, 1-based ia+ja parity) and the list's own flag bit,
// j ascending within each row, 1-based content.
int nl_get_pairs(mdff_ctx* c, int which, int32_t* point, int32_t* list, int64_t cap, int64_t* nout) {
  if (!c->lists_valid) { c->err = "lists not built"; return MDFF_ERR_BAD_ARG; }
  const int n = c->n;
  std::vector<int32_t> start(n + 1), rows(c->nl_total > 0 ? c->nl_total : 1), order(n);
  std::vector<uint8_t> flags(c->nl_total > 0 ? c->nl_total : 1);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(start.data(), c->nl_start, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(order.data(), c->cell_atoms, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  if (c->nl_total > 0) {
    HIP_TRY(c, hipMemcpy(rows.data(), c->nl_j, sizeof(int32_t) * c->nl_total, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(flags.data(), c->nl_flag, c->nl_total, hipMemcpyDeviceToHost));
  }
  const uint8_t bit = which ? 2 : 1;
  int64_t cnt = 0;
  std::vector<int32_t> tmp;
  std::vector<int32_t> out;
  out.reserve(c->nl_total / 2 + 1);
  for (int i = 0; i < n; ++i) {
    if (point) point[i] = (int32_t)(cnt + 1);
    tmp.clear();
    for (int e = start[i]; e < start[i + 1]; ++e) {
      int j = order[rows[e]];             // rows hold cell-sorted positions
      bool own = (i > j && ((i + j) % 2 == 0)) || (i < j && ((i + j) % 2 != 0));
      if (own && (flags[e] & bit)) tmp.push_back(j + 1);
    }
    std::sort(tmp.begin(), tmp.end());
    out.insert(out.end(), tmp.begin(), tmp.end());
    cnt += (int64_t)tmp.size();
  }
  if (point) point[n] = (int32_t)(cnt + 1);
  if (nout) *nout = cnt;
  if (cap < cnt) return MDFF_ERR_BAD_ARG;
  if (list && cnt) std::copy(out.begin(), out.end(), list);
  return MDFF_OK;
}
