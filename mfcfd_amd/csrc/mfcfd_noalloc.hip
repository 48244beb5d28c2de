// Force coefficients without allocation: mf_compute_cl_cd_cm_noalloc (include/mfcfd.h).
//
// compute_force_coeffs.F90:11-102 on caller-owned device arrays, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops).  The
// per-block partial sums live in `scratch`, a file-scope device array of fixed size
// that each GPU holds until the process ends; its device address is looked up once
// per device with hipGetSymbolAddress, and only that address is handed to kernels.
// Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../include/mfcfd.h"
#include "lskum_kernels.h"

using namespace mf;

// defined in mfcfd.hip: sets the message mf_last_error returns
int mf_noalloc_fail(int code, const char* msg);

namespace {

constexpr int NA_BLOCKS = 256;                    // partial-sum blocks at most
constexpr int NA_SHAPES = MF_NOALLOC_MAX_SHAPES;  // shapes at most
constexpr int NA_SLOTS = NA_BLOCKS * NA_SHAPES * 3;

// (H, V, pitching moment) partials: [block][shape][3].  Every launch writes every slot
// it later reads (all blocks of the grid, all `shapes` of each), so nothing needs
// clearing between calls.
static __device__ __align__(256) double scratch[NA_SLOTS];

// compute_force_coeffs.F90:37-77 per shape point, summed per shape in this block.
__global__ __launch_bounds__(BLOCK) void k_forces_partial(
    int n, int n_shape, int shapes, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ nx, const double* __restrict__ ny, const double* __restrict__ prim,
    const int32_t* __restrict__ left, const int32_t* __restrict__ right, const int32_t* __restrict__ flag_2,
    const int32_t* __restrict__ shape_idx, double cp_scale, double* __restrict__ cp,
    double* __restrict__ part) {
  const double pr_inf = 1.0 / 1.4;
  double acc[NA_SHAPES][3];
#pragma unroll
  for (int s = 0; s < NA_SHAPES; ++s) acc[s][0] = acc[s][1] = acc[s][2] = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n_shape; j += gridDim.x * blockDim.x) {
    const int m = shape_idx[j];
    if (m < 0 || m >= n) {
      cp[j] = __builtin_nan("");
      continue;
    }
    const int l = left[m], r = right[m], sh = flag_2[m] - 1;
    if (l < 0 || l >= n || r < 0 || r >= n || sh < 0 || sh >= shapes) {
      cp[j] = __builtin_nan("");
      continue;
    }
    const double lx = x[l], ly = y[l], mx = x[m], my = y[m], rx = x[r], ry = y[r];
    double ds1 = (mx - lx) * (mx - lx) + (my - ly) * (my - ly);
    ds1 = sqrt(ds1);
    double ds2 = (rx - mx) * (rx - mx) + (ry - my) * (ry - my);
    ds2 = sqrt(ds2);
    const double ds = 0.5 * (ds1 + ds2);
    const double nxm = nx[m], nym = ny[m];
    double c = prim[4 * (size_t)m + 3] - pr_inf;
    c = -c / cp_scale;
    cp[j] = c;
    const double h = c * nxm * ds, v = c * nym * ds;
    const double pm = (-c * nym * ds * (mx - 0.25) + c * nxm * ds * (my));
#pragma unroll
    for (int s = 0; s < NA_SHAPES; ++s)
      if (s == sh) {
        acc[s][0] += h;
        acc[s][1] += v;
        acc[s][2] += pm;
      }
  }
  for (int s = 0; s < shapes; ++s)
    for (int k = 0; k < 3; ++k) {
      const double t = block_sum(acc[s][k]);
      if (threadIdx.x == 0) part[((size_t)blockIdx.x * NA_SHAPES + s) * 3 + k] = t;
      __syncthreads();
    }
}

// compute_force_coeffs.F90:79-81: one block per shape sums the blocks' partials in
// block order, then Cl = V cos - H sin, Cd = H cos + V sin, Cm.
__global__ __launch_bounds__(BLOCK) void k_forces_total(int nblocks, double ct, double st,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ clcdcm) {
  const int s = blockIdx.x;
  double t[3];
  for (int k = 0; k < 3; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[((size_t)b * NA_SHAPES + s) * 3 + k];
    t[k] = block_sum(v);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double H = t[0], V = t[1];
    clcdcm[3 * s] = V * ct - H * st;
    clcdcm[3 * s + 1] = H * ct + V * st;
    clcdcm[3 * s + 2] = t[2];
  }
}

std::mutex g_mu;
double* g_scratch[64] = {};  // device address of `scratch` on each device

int scratch_of_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return mf_noalloc_fail(MF_ERR_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= 64) return mf_noalloc_fail(MF_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return mf_noalloc_fail(MF_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return MF_OK;
}

}  // namespace

extern "C" int mf_compute_cl_cd_cm_noalloc(int32_t n_points, int32_t n_shape, int32_t shapes, double mach,
                                           double aoa, const double* x, const double* y, const double* nx,
                                           const double* ny, const double* prim, const int32_t* left,
                                           const int32_t* right, const int32_t* flag_2,
                                           const int32_t* shape_idx, double* clcdcm, double* cp,
                                           void* stream) {
  if (n_points <= 0 || n_shape <= 0 || n_shape > n_points)
    return mf_noalloc_fail(MF_ERR_INVALID, "mf_compute_cl_cd_cm_noalloc: bad point counts");
  if (shapes < 1 || shapes > NA_SHAPES)
    return mf_noalloc_fail(MF_ERR_INVALID, "mf_compute_cl_cd_cm_noalloc: shapes must be 1..MF_NOALLOC_MAX_SHAPES");
  if (!x || !y || !nx || !ny || !prim || !left || !right || !flag_2 || !shape_idx || !clcdcm || !cp)
    return mf_noalloc_fail(MF_ERR_INVALID, "mf_compute_cl_cd_cm_noalloc: null array");
  double* part = nullptr;
  int rc = scratch_of_device(&part);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = std::min(NA_BLOCKS, (n_shape + BLOCK - 1) / BLOCK);
  // parameter.F90:16-34: rho_inf = 1, theta = aoa in radians
  const double pi = 4.0 * atan(1.0);
  const double theta = aoa * pi / 180.0;
  const double cp_scale = 0.5 * 1.0 * mach * mach;
  hipLaunchKernelGGL(k_forces_partial, dim3(nb), dim3(BLOCK), 0, st, (int)n_points, (int)n_shape, (int)shapes, x, y,
                     nx, ny, prim, left, right, flag_2, shape_idx, cp_scale, cp, part);
  hipLaunchKernelGGL(k_forces_total, dim3(shapes), dim3(BLOCK), 0, st, nb, cos(theta), sin(theta),
                     (const double*)part, clcdcm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mf_noalloc_fail(MF_ERR_HIP, hipGetErrorString(e));
  return MF_OK;
}
