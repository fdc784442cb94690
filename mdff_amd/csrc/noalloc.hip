// noalloc.hip -- kinetic energy and temperature of device-resident velocities
// with no allocation: the variant of mdff_calc_temp for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).
//
// The sums are those of mdff_calc_temp (calc_temp, src/kinetic.f90:527-552):
// per-thread strided sums of m v^2 over a grid of min(1024, ceil(n/256))
// blocks of 256 threads, the fixed-order block reduction of reduce.h, and one
// block adding the block partials in block order, so both entry points give
// the same bits for the same velocities and masses.
//
// Scratch: the block partials and the two results live in `scratch`, one
// file-scope device array of kScratchDoubles doubles (8,208 bytes) that every
// device the library runs on holds until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) at the first call on the device; host code passes that address,
// never the symbol itself, to the kernels.  Every call on one device shares
// its array: calls that overlap on one device race.
#include <mutex>

#include "common.h"
#include "reduce.h"

namespace {

constexpr int kMaxBlocks = 1024;                      // block partials of one call
constexpr int kScratchDoubles = kMaxBlocks + 2;       // partials, then ekin and temperature
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

// per-atom m v^2 summed per block: the arithmetic of k_kin and the column
// sums of reduce.h, fused (no per-atom column)
__global__ void __launch_bounds__(256) k_kin_partials(int n, const double* __restrict__ m,
                                                      const double* __restrict__ vx,
                                                      const double* __restrict__ vy,
                                                      const double* __restrict__ vz, double* __restrict__ part) {
  __shared__ double sh[4 * kRedMaxCols];
  double v = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double x = (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) * m[i];
    {
      // an add of its own, as when x is stored and summed by another kernel
#pragma clang fp contract(off)
      v += x;
    }
  }
  const double r = block_sum_cols(&v, 1, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// one block: the partials in block order -> 2 ekin, then ekin and the
// temperature into out
__global__ void __launch_bounds__(256) k_kin_final(int nb, int n, double boltz_unit,
                                                   const double* __restrict__ part, double* __restrict__ res,
                                                   double* __restrict__ out) {
  __shared__ double sh[4 * kRedMaxCols];
  sum_block_partials(nb, 1, part, res, sh);
  if (threadIdx.x == 0) {
    const double e = res[0] * 0.5;
    out[0] = e;
    out[1] = 2.0 * e / boltz_unit / (3.0 * n);
  }
}

std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};

// the device address of this device's scratch array (looked up once)
int scratch_of(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return MDFF_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> g(g_lock);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MDFF_ERR_HIP;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return MDFF_OK;
}

}  // namespace

int mdff_calc_temp_noalloc(int32_t n, const double* mass, const double* vx, const double* vy, const double* vz,
                           double boltz_unit, void* stream, double* out) {
  if (n <= 0 || !mass || !vx || !vy || !vz || !out || !(boltz_unit > 0.0)) return MDFF_ERR_BAD_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MDFF_ERR_NO_DEVICE;
  double* s = nullptr;
  const int rc = scratch_of(dev, &s);
  if (rc != MDFF_OK) return rc;
  const int nb = std::min(kMaxBlocks, std::max(1, div_up(n, 256)));
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_kin_partials, dim3(nb), dim3(256), 0, st, n, mass, vx, vy, vz, s);
  hipLaunchKernelGGL(k_kin_final, dim3(1), dim3(256), 0, st, nb, n, boltz_unit, s, s + kMaxBlocks, out);
  return hipGetLastError() == hipSuccess ? MDFF_OK : MDFF_ERR_HIP;
}
