// Synthetic APBS-style grids evaluated on the device, one thread per grid
// node, atoms streamed through LDS in tiles.  Used to build the benchmark
// systems (APBS itself is not part of this library).
//
// kind 0, screened Coulomb / Debye-Hueckel potential in kBT/e:
//   phi(x) = sum_j q_j exp(-(d - a_j)/lambda) / ((1 + a_j/lambda) 4 pi eps0 eps_out d),
//   d = max(|x - r_j|, a_j, 0.5 A); with a_j = 0 this is the reference's
//   Yukawa monopole V_fac e^{-r/lambda}/r (gpu_batch_engine.py:407-410).
// kind 1, analytic Born desolvation field (image-charge strength of a low-
// dielectric atom of radius a_j in solvent):
//   phi_B(x) = (eps_out - eps_in)/(eps_out + eps_in) / (8 pi eps0 eps_out)
//              * sum_j a_j^3 / max(|x - r_j|, a_j, 0.5)^4,
//   multiplied by alpha q^2 in the force kernel (gpu_batch_engine.py:116-121).
#include <hip/hip_runtime.h>
#include "pst_kernels.hpp"

namespace pst {

constexpr double kEps0 = 0.000142;  // e^2/(kBT A), pystarc/global_defs/constants.py:47-49

__global__ void __launch_bounds__(256) k_grid(float* out32, double* out64, int nx, int ny, int nz,
                                              double ox, double oy, double oz, double sx,
                                              double sy, double sz, const double* atoms, int n,
                                              double lam, double eps_out, double eps_in, int kind) {
  __shared__ double tile[256 * 5];
  const long long total = (long long)nx * ny * nz;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = idx < total;
  const int k = valid ? (int)(idx % nz) : 0;
  const int j = valid ? (int)((idx / nz) % ny) : 0;
  const int i = valid ? (int)(idx / ((long long)ny * nz)) : 0;
  const double x = ox + i * sx, y = oy + j * sy, z = oz + k * sz;
  double acc = 0.0;
  const double pref0 = 1.0 / (4.0 * M_PI * kEps0 * eps_out);
  const double pref1 = (eps_out - eps_in) / (eps_out + eps_in) / (8.0 * M_PI * kEps0 * eps_out);
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    const int a = base + threadIdx.x;
    if (a < n) {
      for (int c = 0; c < 5; ++c) tile[threadIdx.x * 5 + c] = atoms[5 * a + c];
    }
    __syncthreads();
    const int cnt = min(256, n - base);
    for (int t = 0; t < cnt; ++t) {
      const double dx = x - tile[5 * t], dy = y - tile[5 * t + 1], dz = z - tile[5 * t + 2];
      const double q = tile[5 * t + 3], rad = tile[5 * t + 4];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (kind == 0) {
        d = fmax(d, fmax(rad, 0.5));
        acc += q * exp(-(d - rad) / lam) / ((1.0 + rad / lam) * d) * pref0;
      } else {
        d = fmax(d, fmax(rad, 0.5));
        const double d2 = d * d;
        acc += pref1 * rad * rad * rad / (d2 * d2);
      }
    }
  }
  if (valid) {
    if (out32) out32[idx] = (float)acc;
    if (out64) out64[idx] = acc;
  }
}

__global__ void k_interleave(const float* a, const float* b, float2* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(a[i], b[i]);
}

// C-order grid (z fastest) -> the bricked layout of the force kernels
// (bricks of 4x4x4 nodes in C order, z fastest inside a brick).
template <typename T>
__global__ void k_to_bricks(const T* src, T* dst, int nx, int ny, int nz) {
  const long n = (long)nx * ny * nz;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int k = (int)(idx % nz);
  const long r = idx / nz;
  const int j = (int)(r % ny), i = (int)(r / ny);
  const long nbz = (nz + 3) >> 2, nby = (ny + 3) >> 2;
  const long o = (long)(i >> 2) * nby * nbz * 64 + (long)(j >> 2) * nbz * 64 + ((long)(k >> 2) << 6) +
                 ((i & 3) << 4) + ((j & 3) << 2) + (k & 3);
  dst[o] = src[idx];
}

hipError_t launch_to_bricks(const void* src, void* dst, int elem_bytes, int nx, int ny, int nz,
                            hipStream_t st) {
  const long n = (long)nx * ny * nz;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (elem_bytes == 8)
    hipLaunchKernelGGL(k_to_bricks<unsigned long long>, grid, block, 0, st,
                       (const unsigned long long*)src, (unsigned long long*)dst, nx, ny, nz);
  else
    hipLaunchKernelGGL(k_to_bricks<unsigned int>, grid, block, 0, st, (const unsigned int*)src,
                       (unsigned int*)dst, nx, ny, nz);
  return hipGetLastError();
}

hipError_t launch_interleave(const float* a, const float* b, float2* out, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_interleave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, out,
                     n);
  return hipGetLastError();
}

hipError_t launch_grid_dh(float* out, double* out64, int nx, int ny, int nz, const double* origin,
                          const double* sp, const double* atoms, int n_atoms, double lam,
                          double eps_out, double eps_in, int kind, hipStream_t st) {
  const long long total = (long long)nx * ny * nz;
  const long long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_grid, dim3((unsigned)blocks), dim3(256), 0, st, out, out64, nx, ny, nz,
                     origin[0], origin[1], origin[2], sp[0], sp[1], sp[2], atoms, n_atoms, lam,
                     eps_out, eps_in, kind);
  return hipGetLastError();
}

}  // namespace pst
