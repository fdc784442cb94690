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
#include <rocprim/device/device_reduce.hpp>
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

// fractional coordinates in cell order: the doubles, and a wrapped
// single-precision copy (s - floor(s), exact in double, then rounded) for
// the sweep's prefilter
__global__ void k_gather_sorted(int n, const int32_t* __restrict__ order, const double* __restrict__ sx,
                                const double* __restrict__ sy, const double* __restrict__ sz,
                                double* __restrict__ ox, double* __restrict__ oy, double* __restrict__ oz,
                                float4* __restrict__ of) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = order[p];
  const double x = sx[i], y = sy[i], z = sz[i];
  ox[p] = x;
  oy[p] = y;
  oz[p] = z;
  of[p] = make_float4((float)(x - floor(x)), (float)(y - floor(y)), (float)(z - floor(z)), 0.f);
}

struct NLArgs {
  int n, nx, ny, nz;
  double a11, a21, a31, a12, a22, a32, a13, a23, a33;  // simu_cell%A, A(i,j)
  double cut2_max, cut2_v, cut2_c;
  // inner entries (flag bit 2): r^2 <= cut2_in at the build; counted per row
  // into count_in when it is set (cut2_in < 0: none)
  double cut2_in;
  const double *sx, *sy, *sz;             // fractional coords, original order
  const double *px, *py, *pz;             // fractional coords, cell order
  const float4* pf;                       // wrapped single-precision copy, cell order
  // neighbour cells: offsets ox0..ox1 (x), oy0..oy1 (y) around the atom's
  // cell; in z either oz0..oz1 around it or, zall != 0, every z cell
  int ox0, ox1, oy0, oy1, oz0, oz1, zall;
  // prefilter: r^2 in single precision; decided exactly (FP64, reference
  // rounding) when within marg of a cutoff or when prefilter == 0
  int prefilter;
  float a11f, a21f, a31f, a12f, a22f, a32f, a13f, a23f, a33f;
  float c2f_max, c2f_v, c2f_c, c2f_in, marg;
  const int32_t* cell_atoms;
  const int32_t* cell_start;
  const int32_t* cell_of;
  int32_t* count;
  int32_t* count_in;
  const int32_t* row_start;
  // row_w > 0: rows of fixed width row_w (row i at i * row_w), written in one
  // pass that also stores the row's start (row_out) and length (count);
  // row_w == 0: rows at the exclusive scan of a counting pass (row_start)
  int row_w;
  int32_t* row_out;
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

// Candidate index v of the wave's concatenated ranges -> sorted position:
// range k holds v in [pref[k], pref[k + 1]) starting at position st[k];
// unused entries hold pref = total, so the branch-free search over 64
// entries ends on the range that holds v.
__device__ __forceinline__ int range_pos(const int* pref, const int* st, int v) {
  int k = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) k += (pref[k + step] <= v) ? step : 0;
  return st[k] + (v - pref[k]);
}

// One wave per atom (atoms in cell order).  The neighbour cells form, per
// (x, y) column, one or two contiguous ranges of the cell-sorted atoms (the
// z cells of a column are consecutive cells); the ranges are concatenated
// and swept 64 candidates at a time, the loads of the next 64 issued before
// the current ones are tested.  A candidate is tested in single precision
// (wrapped coordinates, rint minimum image); one whose r^2 lies within marg
// of a cutoff is tested again in FP64 exactly as the reference does, so the
// accepted set and the flags equal the reference's bit for bit.
template <bool FILL>
__global__ void __launch_bounds__(256) k_nl_sweep(NLArgs a) {
  __shared__ int rpref[4][66], rst[4][65];
  const int lane = threadIdx.x & 63;
  const int wl = threadIdx.x >> 6;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave >= a.n) return;                 // whole waves: no block barrier below
  const int i = a.cell_atoms[wave];
  const double si = a.sx[i], ti = a.sy[i], ui = a.sz[i];
  const float sif = (float)(si - floor(si)), tif = (float)(ti - floor(ti)), uif = (float)(ui - floor(ui));
  const int ci = a.cell_of[i];
  const int cz = ci % a.nz, cy = (ci / a.nz) % a.ny, cx = ci / (a.nz * a.ny);
  const int64_t base = FILL ? (a.row_w ? (int64_t)i * a.row_w : (int64_t)a.row_start[i]) : 0;
  const int64_t lim = a.row_w ? base + a.row_w : a.cap;
  const int ncx = a.ox1 - a.ox0 + 1, ncy = a.oy1 - a.oy0 + 1;
  const int ncol = ncx * ncy;              // <= 25 columns, <= 2 ranges each
  int b0 = 0, e0 = 0, b1 = 0, e1 = 0;
  if (lane < ncol) {
    const int dy = a.oy0 + lane % ncy, dx = a.ox0 + lane / ncy;
    const int xx = (cx + dx + a.nx) % a.nx, yy = (cy + dy + a.ny) % a.ny;
    const int c0 = (xx * a.ny + yy) * a.nz;  // cell (xx, yy, 0)
    int z0 = 0, z1 = a.nz - 1, w0 = 0, w1 = -1;   // [z0, z1] and the wrapped [w0, w1]
    if (!a.zall) {
      z0 = cz + a.oz0;
      z1 = cz + a.oz1;
      if (z0 < 0) { w0 = z0 + a.nz; w1 = a.nz - 1; z0 = 0; }
      else if (z1 >= a.nz) { w0 = 0; w1 = z1 - a.nz; z1 = a.nz - 1; }
    }
    b0 = a.cell_start[c0 + z0];
    e0 = a.cell_start[c0 + z1 + 1];
    if (w1 >= w0) {
      b1 = a.cell_start[c0 + w0];
      e1 = a.cell_start[c0 + w1 + 1];
    }
  }
  // prefix of the range lengths over the lanes (2 ranges per lane)
  const int len = (e0 - b0) + (e1 - b1);
  int incl = len;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const int total = __shfl(incl, 63, 64);
  const int excl = incl - len;
  // range 2l: [b0, e0), range 2l+1: [b1, e1); 64 entries (2 * 25 + pad)
  if (lane < 32) {
    rpref[wl][2 * lane] = lane < ncol ? excl : total;
    rpref[wl][2 * lane + 1] = lane < ncol ? excl + (e0 - b0) : total;
    rst[wl][2 * lane] = b0;
    rst[wl][2 * lane + 1] = b1;
  }
  if (lane < 2) rpref[wl][64 + lane] = total;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int* pref = rpref[wl];
  const int* st = rst[wl];
  int found = 0, found_in = 0;
  // chunk loads: candidate sorted position, its atom and coordinates (clamped)
  int qn = 0, jn = -1;
  float4 fn = make_float4(0.f, 0.f, 0.f, 0.f);
  if (lane < total) {
    qn = range_pos(pref, st, lane);
    jn = a.cell_atoms[qn];
    fn = a.pf[qn];
  }
  for (int v0 = 0; v0 < total; v0 += 64) {
    const int q = qn, j = jn;
    const float4 f = fn;
    const bool live = v0 + lane < total;
    if (v0 + 64 + lane < total) {
      qn = range_pos(pref, st, v0 + 64 + lane);
      jn = a.cell_atoms[qn];
      fn = a.pf[qn];
    }
    bool ok = false, inner = false;
    uint8_t flag = 0;
    if (live && j != i) {
      float dx = sif - f.x, dy = tif - f.y, dz = uif - f.z;
      dx -= rintf(dx);
      dy -= rintf(dy);
      dz -= rintf(dz);
      const float rx = fmaf(dx, a.a11f, fmaf(dy, a.a12f, dz * a.a13f));
      const float ry = fmaf(dx, a.a21f, fmaf(dy, a.a22f, dz * a.a23f));
      const float rz = fmaf(dx, a.a31f, fmaf(dy, a.a32f, dz * a.a33f));
      const float r2f = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
      const bool near = fabsf(r2f - a.c2f_max) <= a.marg || fabsf(r2f - a.c2f_v) <= a.marg ||
                        fabsf(r2f - a.c2f_c) <= a.marg;
      if (!a.prefilter || near) {
        const double r2 = min_image_r2(a, si - a.px[q], ti - a.py[q], ui - a.pz[q]);
        ok = (r2 <= a.cut2_max);
        inner = r2 <= a.cut2_in;
        flag = (uint8_t)((r2 <= a.cut2_v ? 1 : 0) | (r2 <= a.cut2_c ? 2 : 0));
      } else {
        ok = (r2f <= a.c2f_max);
        inner = r2f <= a.c2f_in;
        flag = (uint8_t)((r2f <= a.c2f_v ? 1 : 0) | (r2f <= a.c2f_c ? 2 : 0));
      }
      if (inner) flag |= 4;
    }
    const uint64_t m = __ballot(ok);
    found_in += __popcll(__ballot(ok && inner));
    if (FILL && ok) {
      const int off = found + __popcll(m & ((1ull << lane) - 1ull));
      const int64_t e = base + off;
      if (e < lim) {
        a.nl_j[e] = q;                   // sorted position of j
        a.nl_flag[e] = flag;
      }
    }
    found += __popcll(m);
  }
  if (lane == 0) {
    if (!FILL || a.row_w) a.count[i] = found;
    if (FILL && a.count_in) a.count_in[i] = found_in;
    if (FILL && a.row_w) a.row_out[i] = (int32_t)base;
    if (FILL && base + found > lim) atomicExch(a.overflow, 1);
  }
}

// Mixed form: reorder each row so that its inner entries (flag bit 2) come
// first, count_in of them, both parts in their sweep order.  One wave per
// row; the rows are read from (sj, sf) and written to (dj, df) at the same
// row starts.
__global__ void __launch_bounds__(256) k_nl_partition(int n, const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ count,
                                                      const int32_t* __restrict__ count_in,
                                                      const int32_t* __restrict__ sj, const uint8_t* __restrict__ sf,
                                                      int32_t* __restrict__ dj, uint8_t* __restrict__ df) {
  const int lane = threadIdx.x & 63;
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (i >= n) return;                      // whole waves
  const int64_t base = start[i];
  const int cnt = count[i], nin = count_in[i];
  const uint64_t below = (1ull << lane) - 1ull;
  int oin = 0, oout = 0;
  for (int v0 = 0; v0 < cnt; v0 += 64) {
    const int e = v0 + lane;
    const bool live = e < cnt;
    int32_t j = 0;
    uint8_t f = 0;
    if (live) {
      j = sj[base + e];
      f = sf[base + e];
    }
    const bool in = live && (f & 4);
    const uint64_t mi = __ballot(in), mo = __ballot(live && !in);
    if (live) {
      const int64_t d = in ? base + oin + __popcll(mi & below) : base + nin + oout + __popcll(mo & below);
      dj[d] = j;
      df[d] = f;
    }
    oin += __popcll(mi);
    oout += __popcll(mo);
  }
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
  a.cut2_in = -1.0;
  if (c->rc_inner > 0.0) {
    const double rin = c->rc_inner + c->d.skindiff;
    a.cut2_in = rin * rin;
  }
  a.sx = c->sx; a.sy = c->sy; a.sz = c->sz;
  a.pf = c->bf;
  a.ox0 = c->nb_lo[0]; a.ox1 = c->nb_hi[0];
  a.oy0 = c->nb_lo[1]; a.oy1 = c->nb_hi[1];
  a.oz0 = c->nb_lo[2]; a.oz1 = c->nb_hi[2];
  a.zall = c->nb_all[2];
  // Single-precision prefilter.  The wrapped coordinates carry an error of
  // at most 2^-24, the difference, the minimum image and each Cartesian
  // component add a few units of 2^-24 relative to |A| (row sums of |A|),
  // so |r^2(single) - r^2| <= (2 sqrt(3) rmax max_row |A| + r^2) 2^-21
  // near the cutoffs; marg is 16 times that.  It is valid when every
  // perpendicular width exceeds 2 rmax (with room): a pair within range is
  // then far from a half-cell fractional offset, and the two nearest
  // integers cannot differ between the single and the double difference.
  {
    double amax = 0.0;
    for (int r = 0; r < 3; ++r)
      amax = std::max(amax, std::fabs(A[r]) + std::fabs(A[r + 3]) + std::fabs(A[r + 6]));
    const double rmax = std::sqrt(a.cut2_max);
    a.marg = (float)(16.0 * (2.0 * std::sqrt(3.0) * rmax * amax + a.cut2_max) * std::ldexp(1.0, -21));
    a.prefilter = 1;
    for (int d = 0; d < 3; ++d)
      if (!(c->W[d] > 2.0 * rmax * (1.0 + 1e-3))) a.prefilter = 0;
    a.a11f = (float)A[0]; a.a21f = (float)A[1]; a.a31f = (float)A[2];
    a.a12f = (float)A[3]; a.a22f = (float)A[4]; a.a32f = (float)A[5];
    a.a13f = (float)A[6]; a.a23f = (float)A[7]; a.a33f = (float)A[8];
    a.c2f_max = (float)a.cut2_max;
    a.c2f_v = (float)a.cut2_v;
    a.c2f_c = (float)a.cut2_c;
    a.c2f_in = (float)a.cut2_in;
  }
  a.cell_atoms = c->cell_atoms;
  a.cell_start = c->cell_start;
  a.cell_of = c->cell_of;
  a.count = c->nl_count;
  a.count_in = c->rc_inner > 0.0 ? c->nl_count_in : nullptr;
  a.row_start = c->nl_start;
  a.row_out = c->nl_start;
  a.row_w = 0;
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
  // cells: width >= cutoff / 2 in a reciprocal direction with room for at
  // least 5 of them (partners within +-2 cells), else width >= cutoff
  // (partners within +-1 cell); 1e-9 margin against rounding of the wrapped
  // coordinates.  With fewer cells than the reach covers, every cell of that
  // direction is visited once.
  double rmax = std::max(c->rc_vdw, c->rc_coul) + c->d.skindiff;
  for (int d = 0; d < 3; ++d) {
    const int k2 = (int)floor(c->W[d] / (0.5 * rmax * (1.0 + 1e-9)));
    int reach = 2, k = k2;
    if (k2 < 5) {
      reach = 1;
      k = (int)floor(c->W[d] / (rmax * (1.0 + 1e-9)));
    }
    c->ncell[d] = k < 1 ? 1 : k;
    const bool all = c->ncell[d] < 2 * reach + 1;
    c->nb_lo[d] = all ? 0 : -reach;
    c->nb_hi[d] = all ? c->ncell[d] - 1 : reach;
    c->nb_all[d] = all ? 1 : 0;
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
  size_t tmp2 = 0, tmp3 = 0;
  HIP_TRY(c, rocprim::exclusive_scan(nullptr, tmp2, c->nl_count, c->nl_start, 0, (size_t)n + 1,
                                     rocprim::plus<int32_t>(), c->stream));
  HIP_TRY(c, rocprim::reduce(nullptr, tmp3, c->nl_count, c->d_flag + 2, 0, (size_t)n,
                             rocprim::maximum<int32_t>(), c->stream));
  MDFF_TRY(ensure(c, &c->sort_tmp, c->sort_tmp_bytes, std::max(tmp, std::max(tmp2, tmp3))));
  HIP_TRY(c, rocprim::radix_sort_pairs(c->sort_tmp, tmp, c->cell_of, c->cell_key, c->cell_atoms_in,
                                       c->cell_atoms, (unsigned)n, 0u, (unsigned)bits, c->stream));
  hipLaunchKernelGGL(k_cell_start, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, ncell, c->cell_key,
                     c->cell_start);
  // fractional coordinates in cell order (coalesced candidate loads); they
  // are kept as the positions of this build
  hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, c->cell_atoms,
                     px, py, pz, c->bx, c->by, c->bz, c->bf);
  NLArgs a = make_args(c);
  a.sx = px; a.sy = py; a.sz = pz;
  a.px = c->bx; a.py = c->by; a.pz = c->bz;
  dim3 grid(div_up((int64_t)n * 64, 256));
  // with the inner split the sweep fills the staging rows and the partition
  // pass writes the final ones
  const bool split = c->rc_inner > 0.0;
  auto partition = [&]() {
    if (split)
      PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_partition, grid, dim3(256), 0, c->stream, n, c->nl_start,
                                                  c->nl_count, c->nl_count_in, c->nl_j_tmp, c->nl_flag_tmp,
                                                  c->nl_j, c->nl_flag));
  };
  if (split) {
    a.nl_j = c->nl_j_tmp;
    a.nl_flag = c->nl_flag_tmp;
  }
  if (c->nl_row_w > 0) {
    // one pass into rows of the width sized at an earlier build; a row that
    // does not fit sends the build through the counting pass below
    a.row_w = c->nl_row_w;
    HIP_TRY(c, hipMemsetAsync(c->d_flag + 1, 0, sizeof(int32_t), c->stream));
    PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_sweep<true>, grid, dim3(256), 0, c->stream, a));
    HIP_TRY(c, hipMemcpyAsync(c->h_flag + 1, c->d_flag + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->h_flag[1] == 0) {
      partition();
      HIP_TRY(c, hipGetLastError());
      c->lists_valid = 1;
      return MDFF_OK;
    }
    a.row_w = 0;
  }
  HIP_TRY(c, hipMemsetAsync(c->nl_count + n, 0, sizeof(int32_t), c->stream));
  PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_sweep<false>, grid, dim3(256), 0, c->stream, a));
  HIP_TRY(c, rocprim::exclusive_scan(c->sort_tmp, tmp2, c->nl_count, c->nl_start, 0, (size_t)n + 1,
                                     rocprim::plus<int32_t>(), c->stream));
  HIP_TRY(c, rocprim::reduce(c->sort_tmp, tmp3, c->nl_count, c->d_flag + 2, 0, (size_t)n,
                             rocprim::maximum<int32_t>(), c->stream));
  int32_t total = 0;
  HIP_TRY(c, hipMemcpyAsync(&total, c->nl_start + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(c->h_flag + 2, c->d_flag + 2, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (total < 0) { c->err = "neighbour list exceeds 2^31 entries"; return MDFF_ERR_LIST_OVERFLOW; }
  // reference capacity: vnlmax*natm entries per half list (src/config.f90:43,
  // src/tools.f90:337-340); the full rows hold twice the half-list entries
  if ((double)total > 2.0 * 2000.0 * (double)n) {
    c->err = "ERROR: out of bound list in vnlist_pbc";
    return MDFF_ERR_LIST_OVERFLOW;
  }
  // fixed-width rows for this and later builds: the widest row with room for
  // growth, while the rows stay within the reference capacity (a one-pass
  // build cannot exceed it then) and int32 positions
  const int64_t widest = c->h_flag[2];
  int64_t w = ((widest + widest / 4 + 16) + 31) / 32 * 32;
  if (w > 2 * 2000 || w * n >= ((int64_t)1 << 31)) w = 0;
  const int64_t need = w > 0 ? w * n : (int64_t)(total * 1.25) + 1024;
  if (need > c->nl_cap) {
    if (c->nl_j) HIP_TRY(c, hipFree(c->nl_j));
    if (c->nl_flag) HIP_TRY(c, hipFree(c->nl_flag));
    if (c->nl_j_tmp) HIP_TRY(c, hipFree(c->nl_j_tmp));
    if (c->nl_flag_tmp) HIP_TRY(c, hipFree(c->nl_flag_tmp));
    c->nl_j_tmp = nullptr;
    c->nl_flag_tmp = nullptr;
    HIP_TRY(c, hipMalloc((void**)&c->nl_j, sizeof(int32_t) * need));
    HIP_TRY(c, hipMalloc((void**)&c->nl_flag, need));
    if (split) {
      HIP_TRY(c, hipMalloc((void**)&c->nl_j_tmp, sizeof(int32_t) * need));
      HIP_TRY(c, hipMalloc((void**)&c->nl_flag_tmp, need));
    }
    c->nl_cap = need;
  }
  a.nl_j = split ? c->nl_j_tmp : c->nl_j;
  a.nl_flag = split ? c->nl_flag_tmp : c->nl_flag;
  a.cap = c->nl_cap;
  a.row_w = (int)w;
  c->nl_row_w = (int)w;
  HIP_TRY(c, hipMemsetAsync(c->d_flag + 1, 0, sizeof(int32_t), c->stream));
  PROF_LAUNCH(c, PK_NLIST, hipLaunchKernelGGL(k_nl_sweep<true>, grid, dim3(256), 0, c->stream, a));
  partition();
  HIP_TRY(c, hipGetLastError());
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
  if (!e.check) return MDFF_OK;
  // no event: the final-sum launch that follows in the step signals the host
  c->chk_seq += 1.0;
  c->chk_seq_wait = 1;
  c->chk_pending = 1;
  return MDFF_OK;
}

int nl_check_finish(mdff_ctx* c, int* rebuilt) {
  *rebuilt = 0;
  if (!c->chk_pending) return MDFF_OK;
  c->chk_pending = 0;
  const int n = c->n;
  const int nb = div_up(n, 256);
  if (c->chk_seq_wait) {
    // the check's maxima are in h_chk once the sequence value is: spin on the
    // mapped word, then (if the signal was never attached or is slow) wait
    // for the stream and look again
    volatile const double* sig = c->h_chk + 2 * nb;
    bool seen = false;
    if (c->chk_seq_wait == 2)
      for (long k = 0; k < 20000000L && !seen; ++k) seen = (*sig == c->chk_seq);
    if (!seen) {
      if (c->chk_seq_wait == 1) MDFF_TRY(flush_finals(c));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      if (*sig != c->chk_seq) { c->err = "skin check signal missing"; return MDFF_ERR_HIP; }
    }
    c->chk_seq_wait = 0;
  } else {
    HIP_TRY(c, hipEventSynchronize(c->chk_ev));
  }
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
  std::vector<int32_t> start(n), count(n), order(n);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(start.data(), c->nl_start, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(count.data(), c->nl_count, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(order.data(), c->cell_atoms, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  int64_t span = 0, total = 0;
  for (int i = 0; i < n; ++i) {
    span = std::max<int64_t>(span, (int64_t)start[i] + count[i]);
    total += count[i];
  }
  std::vector<int32_t> rows(span > 0 ? span : 1);
  std::vector<uint8_t> flags(span > 0 ? span : 1);
  if (span > 0) {
    HIP_TRY(c, hipMemcpy(rows.data(), c->nl_j, sizeof(int32_t) * span, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(flags.data(), c->nl_flag, span, hipMemcpyDeviceToHost));
  }
  const uint8_t bit = which ? 2 : 1;
  int64_t cnt = 0;
  std::vector<int32_t> tmp;
  std::vector<int32_t> out;
  out.reserve(total / 2 + 1);
  for (int i = 0; i < n; ++i) {
    if (point) point[i] = (int32_t)(cnt + 1);
    tmp.clear();
    for (int e = start[i]; e < start[i] + count[i]; ++e) {
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
