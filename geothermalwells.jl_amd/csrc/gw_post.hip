// gw_post.hip -- post-processing reductions over a resident field (the
// reference's validation helpers, src/util.jl:274-356, which read a saved
// host snapshot).
#include "gw_kernels.hpp"

namespace gw {

namespace {

constexpr int kMeanLevels = 64;  // levels per block
constexpr int kMeanSlices = 4;   // column slices per level (one wave each)

// Per-level means over a list of pipe columns (get_temperatures_along_z_single_well,
// src/util.jl:330-356): block (level chunk, list) with blockIdx.y = 0 for the
// inner-pipe list, 1 for the annulus list; a wave per column slice, a lane per
// level (the loads of one column are 64 consecutive levels), the slices added
// in a fixed order.
template <class TS>
__global__ __launch_bounds__(kMeanLevels * kMeanSlices) void k_level_means(const TS* __restrict__ phi,
                                                                           const int64_t* __restrict__ off,
                                                                           int n_in, int n_an, int nlev,
                                                                           double* __restrict__ means) {
  __shared__ double part[kMeanSlices][kMeanLevels];
  const int list = blockIdx.y;
  const int n = list == 0 ? n_in : n_an;
  const int64_t* o = off + (list == 0 ? 0 : n_in);
  const int l = threadIdx.x % kMeanLevels;
  const int sl = threadIdx.x / kMeanLevels;
  const int lev = blockIdx.x * kMeanLevels + l;
  double s = 0.0;
  if (lev < nlev)
    for (int c = sl; c < n; c += kMeanSlices) s += (double)phi[o[c] + lev];
  part[sl][l] = s;
  __syncthreads();
  if (sl == 0 && lev < nlev) {
    double t = part[0][l];
    for (int q = 1; q < kMeanSlices; ++q) t += part[q][l];
    means[list * nlev + lev] = t / (double)n;
  }
}

// One x-line of the field at grid row j, interpolated between levels k_lo and
// k_hi with the reference's weights, two products and a sum each rounded (no
// contraction: the reference's broadcast has no @muladd), extract_x_profile,
// src/util.jl:274-318.
template <class TS>
__global__ __launch_bounds__(256) void k_x_profile(const TS* __restrict__ phi, int nx, int ny, int nz, int j,
                                                   int k_lo, int k_hi, double w_lo, double w_hi,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const int64_t col = ((int64_t)j + (int64_t)ny * i) * nz;
  const double a = (double)phi[col + k_lo];
  out[i] = k_lo == k_hi ? a : w_lo * a + w_hi * (double)phi[col + k_hi];
}

}  // namespace

hipError_t launch_level_means(int prec, const void* phi, const int64_t* off, int n_in, int n_an, int nlev,
                              double* means, hipStream_t st) {
  if (nlev <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nlev + kMeanLevels - 1) / kMeanLevels), 2);
  if (prec == 0)
    hipLaunchKernelGGL(k_level_means<double>, grid, dim3(kMeanLevels * kMeanSlices), 0, st,
                       static_cast<const double*>(phi), off, n_in, n_an, nlev, means);
  else
    hipLaunchKernelGGL(k_level_means<float>, grid, dim3(kMeanLevels * kMeanSlices), 0, st,
                       static_cast<const float*>(phi), off, n_in, n_an, nlev, means);
  return hipGetLastError();
}

hipError_t launch_x_profile(int prec, const void* phi, int nx, int ny, int nz, int j, int k_lo, int k_hi,
                            double w_lo, double w_hi, double* out, hipStream_t st) {
  const dim3 grid((unsigned)((nx + 255) / 256));
  if (prec == 0)
    hipLaunchKernelGGL(k_x_profile<double>, grid, dim3(256), 0, st, static_cast<const double*>(phi), nx, ny, nz, j,
                       k_lo, k_hi, w_lo, w_hi, out);
  else
    hipLaunchKernelGGL(k_x_profile<float>, grid, dim3(256), 0, st, static_cast<const float*>(phi), nx, ny, nz, j,
                       k_lo, k_hi, w_lo, w_hi, out);
  return hipGetLastError();
}

}  // namespace gw
