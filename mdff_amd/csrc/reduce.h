// reduce.h -- deterministic block and grid reductions for 256-thread blocks.
//
// Every sum here has a fixed association order for a given launch shape, so
// repeated runs give bit-identical results: wave shuffles, then the four wave
// sums in wave order; across blocks (a second, one-block launch), per-thread
// strided partial sums over the block index followed by the same block
// reduction.  No grid-wide completion counter: on a GPU whose L2 is split
// over several dies, the device-scope fences such a counter needs write the
// L2 back and cost more than the second launch.
#pragma once
#include <hip/hip_runtime.h>

constexpr int kRedMaxCols = 16;   // columns one reduction can carry

// Column sums of v[0..ncol) over the 256 threads of the block.  The result
// for column c is returned in thread c (c < ncol); sh holds 4 * kRedMaxCols
// doubles.
__device__ inline double block_sum_cols(const double* v, int ncol, double* sh) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) {
    if (c < ncol) {
      double x = v[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
      if ((threadIdx.x & 63) == 0) sh[w * kRedMaxCols + c] = x;
    }
  }
  __syncthreads();
  double r = 0.0;
  const int c = threadIdx.x;
  if (c < ncol) r = ((sh[c] + sh[kRedMaxCols + c]) + sh[2 * kRedMaxCols + c]) + sh[3 * kRedMaxCols + c];
  __syncthreads();
  return r;
}

// Column sums of a [ncol][n] array (ncol <= kRedMaxCols): block partials
// part[block][ncol].  Every column is loaded from a clamped, valid address and
// masked afterwards, so one thread's loads are in flight together.
__device__ inline void colsum_partials(int n, int ncol, const double* __restrict__ in,
                                       double* __restrict__ part, double* sh) {
  double v[kRedMaxCols];
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) v[c] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double x[kRedMaxCols];
#pragma unroll
    for (int c = 0; c < kRedMaxCols; ++c) x[c] = in[(size_t)(c < ncol ? c : ncol - 1) * n + i];
#pragma unroll
    for (int c = 0; c < kRedMaxCols; ++c) v[c] += c < ncol ? x[c] : 0.0;
  }
  const double r = block_sum_cols(v, ncol, sh);
  if ((int)threadIdx.x < ncol) part[(size_t)blockIdx.x * ncol + threadIdx.x] = r;
}

// One block: out[c] = sum over b < nb of part[b * ncol + c], per-thread
// strided sums over b, then the block reduction.
__device__ inline void sum_block_partials(int nb, int ncol, const double* __restrict__ part,
                                          double* __restrict__ out, double* sh) {
  double v[kRedMaxCols];
#pragma unroll
  for (int c = 0; c < kRedMaxCols; ++c) v[c] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    double x[kRedMaxCols];
#pragma unroll
    for (int c = 0; c < kRedMaxCols; ++c) x[c] = part[(size_t)b * ncol + (c < ncol ? c : ncol - 1)];
#pragma unroll
    for (int c = 0; c < kRedMaxCols; ++c) v[c] += c < ncol ? x[c] : 0.0;
  }
  const double r = block_sum_cols(v, ncol, sh);
  if ((int)threadIdx.x < ncol) out[threadIdx.x] = r;
}

// (FinalBatch and FinalSum: common.h, included first.)
// Block b of a launch that carries a FinalBatch: final sum b (b < nfin) and,
// in block 0, the completion signal.  Whole blocks (256 threads).
__device__ inline void run_final_batch(const FinalBatch& fb, int b, double* sh) {
  if (fb.seq_out && b == 0 && threadIdx.x == 0) fb.seq_out[0] = fb.seq_val;
  if (b >= fb.nfin) return;
  const FinalSum f = fb.f[b];
  sum_block_partials(f.nb, f.ncol, f.part, f.out, sh);
}
