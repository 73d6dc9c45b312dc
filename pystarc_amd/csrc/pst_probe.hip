// Gather ceiling of the receptor grids (measurement only, no reference
// counterpart).  Each thread gathers the 20-node potential + Born stencil of
// the FP32 force path (grid_phi_grad_pair: float2 nodes in 4x4x4 bricks,
// 160 algorithmic bytes per point) at one point and writes one float, so the
// kernel does the force kernel's gathers with none of its arithmetic:
//   mode 0: points uniform over the grid's valid box (every gather far from
//           the last: the line traffic is served by the Infinity Cache / HBM);
//   mode k >= 1: the points of one workgroup inside a cube of edge 12 / k A
//           at a random place (mode 1: the gathers of a ligand's atoms,
//           neighbouring stencils sharing bricks and lines in L2; higher k:
//           tighter clusters, the locality a better ordering could reach).
// Its rate is the ceiling a gather-bound force kernel can reach on the same
// grids.  Fetch variants: 0 the plain stencil (20 single-node loads); 1 the
// z-pair fetches of the force kernel's z-pair mode (16 aligned pair loads);
// 2 the same z-pair fetches issued cooperatively by teams of 8 lanes
// (stencil_team_load), each load instruction covering the columns of 8
// points instead of one column of 64.
#include <hip/hip_runtime.h>
#include "pst_device.hpp"
#include "pst_kernels.hpp"

namespace pst {

__device__ inline double hash_unit(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

template <int VAR>
__global__ void __launch_bounds__(256) k_gather_probe(GridDev g, long long n, int mode,
                                                      unsigned long long seed, float* out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const double ext[3] = {g.hi[0] - g.lo[0], g.hi[1] - g.lo[1], g.hi[2] - g.lo[2]};
  __shared__ Pair<float> team_buf[VAR == 2 ? 4 : 1][VAR == 2 ? kTeamWords : 1];
  // every lane of a wave takes the same number of trips (team loads need the whole wave)
  const long long first = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
  for (long long base = first; base < n; base += stride) {
    const long long p = base + (threadIdx.x & 63);
    const bool live = p < n;
    if (VAR != 2 && !live) break;
    double x[3];
    if (mode == 0) {
      for (int d = 0; d < 3; ++d)
        x[d] = g.lo[0 + d] + ext[d] * hash_unit(seed ^ (unsigned long long)(3 * p + d));
    } else {
      // the workgroup's cluster (one per 256 consecutive points)
      const long long c = p >> 8;
      for (int d = 0; d < 3; ++d) {
        const double h0 = 6.0 / mode;
        const double half = h0 < 0.5 * ext[d] ? h0 : 0.5 * ext[d];
        const double cen = g.lo[d] + half + (ext[d] - 2.0 * half) *
                                                hash_unit(seed ^ (0x5555ull + 3 * (unsigned long long)c + d));
        x[d] = cen + half * (2.0 * hash_unit(seed ^ (unsigned long long)(3 * p + d + 0x777)) - 1.0);
      }
    }
    Pair<float> phi, gx, gy, gz;
    if constexpr (VAR == 2) {
      const StencilGeo<float> sg = stencil_geo<float>(g, x[0], x[1], x[2]);
      StencilNodes<Pair<float>> nodes;
      stencil_team_load<float2, Pair<float>>((const float2*)g.data, global_acc<true>(g), sg, live,
                                              team_buf[threadIdx.x >> 6], nodes);
      if (!live) continue;
      stencil_eval<float, Pair<float>>(g, sg, nodes, phi, gx, gy, gz);
    } else {
      grid_phi_grad_pair<float, VAR == 1>(g, x[0], x[1], x[2], phi, gx, gy, gz);
    }
    out[p] = phi.a + phi.b + gx.a + gx.b + gy.a + gy.b + gz.a + gz.b;
  }
}

hipError_t launch_gather_probe(const GridDev& g, long long n, int mode, int variant,
                               unsigned long long seed, float* out, int blocks, hipStream_t st) {
  if (variant == 2)
    hipLaunchKernelGGL(k_gather_probe<2>, dim3(blocks), dim3(256), 0, st, g, n, mode, seed, out);
  else if (variant == 1)
    hipLaunchKernelGGL(k_gather_probe<1>, dim3(blocks), dim3(256), 0, st, g, n, mode, seed, out);
  else
    hipLaunchKernelGGL(k_gather_probe<0>, dim3(blocks), dim3(256), 0, st, g, n, mode, seed, out);
  return hipGetLastError();
}

}  // namespace pst
