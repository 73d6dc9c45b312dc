// Stand-alone launches of the two-body RPY move (the production step calls the
// same __device__ functions from k_prod_post).
#include <hip/hip_runtime.h>
#include "pst_kernels.hpp"
#include "pst_rpy.hpp"

namespace pst {

// Columns of the lab-frame mobility M (f = e_j, no noise, Df dt = 1) and of
// the noise factor G (xi = e_j, no load, 2 Df dt = 1), both 12x12 row-major.
__global__ void k_rpy_eval(int n, const double* pos, double a0, double a1, double* mob,
                           double* fac) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = pos + 3 * (size_t)i;
  for (int j = 0; j < 12; ++j) {
    double f[12], xi[12], dX[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) { f[k] = k == j ? 1.0 : 0.0; xi[k] = 0.0; }
    if (mob) {
      rpy_displacement(a0, a1, 1.0, 1.0, p, f, f + 3, f + 6, f + 9, xi, dX);
      for (int k = 0; k < 12; ++k) mob[(size_t)i * 144 + k * 12 + j] = dX[k];
    }
    if (fac) {
#pragma unroll
      for (int k = 0; k < 12; ++k) { xi[k] = f[k]; f[k] = 0.0; }
      rpy_displacement(a0, a1, 1.0, 0.5, p, f, f + 3, f + 6, f + 9, xi, dX);
      for (int k = 0; k < 12; ++k) fac[(size_t)i * 144 + k * 12 + j] = dX[k];
    }
  }
}

__global__ void k_rpy_move(int n, double a0, double a1, double Df, int torque_mode,
                           const double* dt, double* pos, double* quat, const double* F,
                           const double* T, const double* xi, double* dX_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* p = pos + 3 * (size_t)i;
  double FR[3], TR[3], dX[12];
  rpy_receptor_load(torque_mode, p, F + 3 * (size_t)i, T + 3 * (size_t)i, FR, TR);
  rpy_displacement(a0, a1, Df, dt[i], p, FR, F + 3 * (size_t)i, TR, T + 3 * (size_t)i,
                   xi + 12 * (size_t)i, dX);
  if (dX_out)
    for (int k = 0; k < 12; ++k) dX_out[12 * (size_t)i + k] = dX[k];
  rpy_apply(dX, p, quat + 4 * (size_t)i);
}

hipError_t launch_rpy_eval(int n, const double* pos, double a0, double a1, double* mob,
                           double* fac, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rpy_eval, dim3((n + 127) / 128), dim3(128), 0, st, n, pos, a0, a1, mob,
                     fac);
  return hipGetLastError();
}

hipError_t launch_rpy_move(int n, double a0, double a1, double Df, int torque_mode,
                           const double* dt, double* pos, double* quat, const double* F,
                           const double* T, const double* xi, double* dX_out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rpy_move, dim3((n + 127) / 128), dim3(128), 0, st, n, a0, a1, Df,
                     torque_mode, dt, pos, quat, F, T, xi, dX_out);
  return hipGetLastError();
}

}  // namespace pst
