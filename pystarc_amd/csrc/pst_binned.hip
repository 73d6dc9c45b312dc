// Brick-binned force evaluation for gfx950: the same forces as k_forces, with
// the grid gathers regrouped by the region of the grid they read.
//
// k_forces walks trajectories; each in-grid atom pulls its 20 stencil nodes
// from L2 / Infinity Cache, about 8 scattered cache lines per atom, and the
// lines are rarely reused before eviction.  Here one force evaluation is five
// kernels:
//   count    every (trajectory, atom) gather request is assigned to the
//            brick (8^3 cells) of the grid cell it interpolates in; requests
//            per brick are counted (wave-aggregated atomics over runs of
//            equal bricks in consecutive lanes)
//   scan     exclusive scan of the counts (one workgroup), work items of at
//            most kItem requests per brick, counts reset for the next call
//   scatter  the same requests written into brick order
//   gather   one workgroup per work item stages the brick and its halo
//            (11^3 nodes: float2 potential + Born pair, or one FP32 field)
//            in LDS once and evaluates all of the brick's requests from LDS;
//            items are cut into eight contiguous ranges, one per XCD, so
//            neighbouring bricks share an L2
//   reduce   per trajectory, the per-atom results are summed in the lane
//            order of k_forces, with the Yukawa far-field for atoms outside
//            the grid, reverse Born and torques as k_forces does
// Per-atom arithmetic (positions, assignment, interpolation, charges) is
// k_forces' own, through an LDS accessor, so the sums are bit-identical to
// the direct kernel with the same G.  Restates gpu_batch_engine.py:289-391
// (forward gathers) and :607-657 (reverse Born) like k_forces.
#include <hip/hip_runtime.h>
#include "pst_device.hpp"
#include "pst_kernels.hpp"

namespace pst {

constexpr int kBrick = 8;                        // cells per brick edge
constexpr int kTile = kBrick + 3;                // tile nodes per edge: -1 .. kBrick + 1
constexpr int kTileNodes = kTile * kTile * kTile;
constexpr int kItem = 1024;                      // requests per gather work item

template <typename TC> struct Out4 { typedef double4 type; };
template <> struct Out4<float> { typedef float4 type; };

// Brick of the cell that an interpolation at p starts from (i0 = floor).
__device__ inline int brick_of(const GridDev& g, const int* nb, double x, double y, double z) {
  const int i0 = (int)floor((x - g.o[0]) * g.inv[0]);
  const int j0 = (int)floor((y - g.o[1]) * g.inv[1]);
  const int k0 = (int)floor((z - g.o[2]) * g.inv[2]);
  return ((i0 / kBrick) * nb[1] + j0 / kBrick) * nb[2] + k0 / kBrick;
}

// Ligand body coordinate as k_forces forms it (TC-rounded when staged in LDS).
template <typename TC>
__device__ inline double body_coord(const ForceSys& S, int idx) {
  return S.M <= kLdsAtomsMax ? (double)(TC)S.lig_body[idx] : S.lig_body[idx];
}

__device__ inline void pose_of(const ForceWork& W, int t, double* R, double& cx, double& cy,
                               double& cz) {
  quat_to_rot(W.quat + 4 * (size_t)t, R);
  cx = W.pos[3 * t]; cy = W.pos[3 * t + 1]; cz = W.pos[3 * t + 2];
}

// Lab position of ligand atom i: R b + c (gpu_batch_simulator.py:935-937).
template <typename TC>
__device__ inline void atom_lab(const ForceSys& S, const double* R, double cx, double cy,
                                double cz, int i, double& px, double& py, double& pz) {
  const double bx = body_coord<TC>(S, 3 * i), by = body_coord<TC>(S, 3 * i + 1),
               bz = body_coord<TC>(S, 3 * i + 2);
  px = R[0] * bx + R[1] * by + R[2] * bz + cx;
  py = R[3] * bx + R[4] * by + R[5] * bz + cy;
  pz = R[6] * bx + R[7] * by + R[8] * bz + cz;
}

// Receptor atom j in the ligand frame: R^T (r - c) (:641-644).
__device__ inline void rec_in_lig(const ForceSys& S, const double* R, double cx, double cy,
                                  double cz, int j, double& x, double& y, double& z) {
  const double dx = S.rec_pos[3 * j] - cx, dy = S.rec_pos[3 * j + 1] - cy,
               dz = S.rec_pos[3 * j + 2] - cz;
  x = R[0] * dx + R[3] * dy + R[6] * dz;
  y = R[1] * dx + R[4] * dy + R[7] * dz;
  z = R[2] * dx + R[5] * dy + R[8] * dz;
}

// Wave-aggregated emission of one request per lane (key < 0: none).  Runs of
// equal keys in consecutive lanes share one atomic.  PASS 0 counts, PASS 1
// writes the request id at its position in brick order.
template <int PASS>
__device__ inline void bin_emit(const BinBufs& B, int key, int rid) {
  const int l = __lane_id();
  const int kp = __shfl_up(key, 1);
  const bool head = (l == 0) || (key != kp);
  const unsigned long long hm = __ballot(head);
  const unsigned long long above = (l == 63) ? 0ull : (hm >> (l + 1));
  const int len = above ? __ffsll((long long)above) : (64 - l);
  int pos = 0;
  if (head && key >= 0) {
    if (PASS == 0)
      atomicAdd(&B.hist[key], len);
    else
      pos = atomicAdd(&B.cursor[key], len);
  }
  if (PASS == 1) {
    const unsigned long long le = (l == 63) ? hm : (hm & ((2ull << l) - 1ull));
    const int h = 63 - __clzll((long long)le);
    const int base = __shfl(pos, h);
    if (key >= 0) B.sorted[base + (l - h)] = rid;
  }
}

template <typename TC, int G, int PASS>
__global__ void __launch_bounds__(kForceBlock) k_bin_pass(ForceSys S, ForceWork W, BinBufs B) {
  const int n = W.n_dev ? *W.n_dev : W.n;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = kForceBlock / G;
  const int M = S.M;
  const bool do_rev = S.has_rev && S.alpha > 1e-12;
  for (int base = blockIdx.x * gpb; base < n; base += gridDim.x * gpb) {
    const int w = base + grp;
    const bool valid = w < n;
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, cx = 0, cy = 0, cz = 0;
    if (valid) pose_of(W, W.list ? W.list[w] : w, R, cx, cy, cz);
    const double rc = norm3(cx, cy, cz);
    const bool grids = valid && rc - S.lig_rmax <= S.grid_reach;
    for (int i0 = 0; i0 < M; i0 += G) {
      const int i = i0 + lane;
      int key = -1;
      if (grids && i < M && fabs(S.lig_q[i]) >= 1e-9) {
        double px, py, pz;
        atom_lab<TC>(S, R, cx, cy, cz, i, px, py, pz);
        if (grid_inside(S.pair, px, py, pz)) key = brick_of(S.pair, B.rb, px, py, pz);
      }
      bin_emit<PASS>(B, key, w * M + i);
    }
    if (do_rev) {
      const bool live = valid && !(rc - S.rec_max_r > S.rev_cull);
      // runs of receptor atoms out of the ligand grid's reach emit nothing;
      // a run is skipped when no lane of the wave needs it (bin_emit is a
      // wave-wide operation)
      for (int k = 0; k < S.n_rec_chunk; ++k) {
        const bool near = live && rec_chunk_near(S, k, cx, cy, cz, S.rev_cull);
        if (__ballot(near) == 0ull) continue;
        const int j1 = min((k + 1) * PST_REC_CHUNK, S.n_rec);
        for (int j0 = k * PST_REC_CHUNK; j0 < j1; j0 += G) {
          const int j = j0 + lane;
          int key = -1;
          if (near && j < j1 && fabs(S.rec_q[j]) >= 1e-9) {
            double x, y, z;
            rec_in_lig(S, R, cx, cy, cz, j, x, y, z);
            if (grid_inside(S.lig_born, x, y, z)) key = B.nb_rec + brick_of(S.lig_born, B.lb, x, y, z);
          }
          bin_emit<PASS>(B, key, B.rev_base + w * S.n_rec + j);
        }
      }
    }
  }
}

// Exclusive scan of the brick counts, work items, counts reset.
__global__ void __launch_bounds__(1024) k_bin_scan(BinBufs B) {
  __shared__ int s_req[1024], s_it[1024];
  const int nb = B.nb_rec + B.nb_lig;
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per, b1 = min(nb, b0 + per);
  int req = 0, it = 0;
  for (int b = b0; b < b1; ++b) {
    const int c = B.hist[b];
    req += c;
    it += (c + kItem - 1) / kItem;
  }
  s_req[t] = req;
  s_it[t] = it;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int ar = t >= off ? s_req[t - off] : 0, ai = t >= off ? s_it[t - off] : 0;
    __syncthreads();
    s_req[t] += ar;
    s_it[t] += ai;
    __syncthreads();
  }
  int rpos = s_req[t] - req, ipos = s_it[t] - it;
  for (int b = b0; b < b1; ++b) {
    const int c = B.hist[b];
    B.start[b] = rpos;
    B.cursor[b] = rpos;
    for (int k = 0; k * kItem < c && ipos < B.max_items; ++k, ++ipos) B.items[ipos] = make_int2(b, k * kItem);
    rpos += c;
    B.hist[b] = 0;
  }
  if (t == 1023) {
    B.start[nb] = s_req[1023];
    *B.n_items = min(s_it[1023], B.max_items);
  }
}

template <typename TC>
__global__ void __launch_bounds__(256) k_bin_gather(ForceSys S, ForceWork W, BinBufs B) {
  typedef typename Out4<TC>::type O4;
  __shared__ float2 tile[kTileNodes];
  const int n = W.n_dev ? *W.n_dev : W.n;
  const int n_items = *B.n_items;
  const int M = S.M;
  O4* out = (O4*)B.out;
  O4* out_rev = (O4*)B.out_rev;
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
  const int lo = (int)(((long)n_items * xcd) >> 3), hi = (int)(((long)n_items * (xcd + 1)) >> 3);
  for (int it = lo + (blockIdx.x >> 3); it < hi; it += per) {
    const int2 item = B.items[it];
    const int bin = item.x;
    const bool rev = bin >= B.nb_rec;
    const GridDev& g = rev ? S.lig_born : S.pair;
    const int* nb = rev ? B.lb : B.rb;
    const int local = rev ? bin - B.nb_rec : bin;
    const int bk = local % nb[2], bj = (local / nb[2]) % nb[1], bi = local / (nb[1] * nb[2]);
    const int ti = bi * kBrick - 1, tj = bj * kBrick - 1, tk = bk * kBrick - 1;
    __syncthreads();
    for (int e = threadIdx.x; e < kTileNodes; e += blockDim.x) {
      const int a = e / (kTile * kTile), b = (e / kTile) % kTile, c = e % kTile;
      const int i = ti + a, j = tj + b, k = tk + c;
      if (i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz) {
        const long gi = global_acc(g).at(i, j, k);
        if (rev)
          ((float*)tile)[e] = ((const float*)g.data)[gi];
        else
          tile[e] = ((const float2*)g.data)[gi];
      }
    }
    __syncthreads();
    const TileAcc acc{(long)kTile * kTile, (long)kTile, ti, tj, tk};
    const int first = B.start[bin] + item.y;
    const int last = min(B.start[bin + 1], first + kItem);
    for (int r = first + (int)threadIdx.x; r < last; r += blockDim.x) {
      const int rid = B.sorted[r];
      if (!rev) {
        if (rid < 0 || rid >= n * M) continue;
        const int w = rid / M, i = rid - w * M;
        double R[9], cx, cy, cz, px, py, pz;
        pose_of(W, W.list ? W.list[w] : w, R, cx, cy, cz);
        atom_lab<TC>(S, R, cx, cy, cz, i, px, py, pz);
        const double q = S.lig_q[i];
        Pair<TC> phi, gx, gy, gz;
        grid_phi_grad_pair_a<TC>(S.pair, tile, acc, px, py, pz, phi, gx, gy, gz);
        const TC qq = (TC)q, c = (TC)(S.alpha * q * q);
        TC fx = 0, fy = 0, fz = 0, e = 0;
        fx -= qq * gx.a + c * gx.b;
        fy -= qq * gy.a + c * gy.b;
        fz -= qq * gz.a + c * gz.b;
        e += qq * phi.a + c * phi.b;
        O4 v;
        v.x = fx; v.y = fy; v.z = fz; v.w = e;
        out[rid] = v;
      } else {
        const int rr = rid - B.rev_base;
        if (rr < 0 || rr >= n * S.n_rec) continue;
        const int w = rr / S.n_rec, j = rr - w * S.n_rec;
        double R[9], cx, cy, cz, x, y, z;
        pose_of(W, W.list ? W.list[w] : w, R, cx, cy, cz);
        rec_in_lig(S, R, cx, cy, cz, j, x, y, z);
        const double q = S.rec_q[j];
        TC phi, gx, gy, gz;
        grid_phi_grad_f32_a<TC>(S.lig_born, (const float*)tile, acc, x, y, z, phi, gx, gy, gz);
        const TC c = (TC)(S.alpha * q * q);
        O4 v;
        v.x = -c * gx; v.y = -c * gy; v.z = -c * gz; v.w = 0;
        out_rev[rr] = v;
      }
    }
  }
}

// Per-trajectory sums in k_forces' order, writes F / T / E like k_forces.
template <typename TC, int G>
__global__ void __launch_bounds__(kForceBlock) k_bin_reduce(ForceSys S, ForceWork W, BinBufs B) {
  typedef typename Out4<TC>::type O4;
  const int n = W.n_dev ? *W.n_dev : W.n;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = kForceBlock / G;
  const int M = S.M;
  const O4* out = (const O4*)B.out;
  const O4* out_rev = (const O4*)B.out_rev;
  for (int base = blockIdx.x * gpb; base < n; base += gridDim.x * gpb) {
    const int w = base + grp;
    const bool valid = w < n;
    const int t = valid ? (W.list ? W.list[w] : w) : 0;
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, cx = 0, cy = 0, cz = 0;
    if (valid) pose_of(W, t, R, cx, cy, cz);
    TC fx = 0, fy = 0, fz = 0, e = 0, tx = 0, ty = 0, tz = 0;
    double qx = 0, qy = 0, qz = 0;
    if (S.torque_mode == 1) {
      qx = cx + R[0] * S.cq_body[0] + R[1] * S.cq_body[1] + R[2] * S.cq_body[2];
      qy = cy + R[3] * S.cq_body[0] + R[4] * S.cq_body[1] + R[5] * S.cq_body[2];
      qz = cz + R[6] * S.cq_body[0] + R[7] * S.cq_body[1] + R[8] * S.cq_body[2];
    }
    const double rc = norm3(cx, cy, cz);
    const bool grids = rc - S.lig_rmax <= S.grid_reach;
    if (valid) {
      for (int i = lane; i < M; i += G) {
        double px, py, pz;
        atom_lab<TC>(S, R, cx, cy, cz, i, px, py, pz);
        const double q = S.lig_q[i];
        TC ax = 0, ay = 0, az = 0, ae = 0;
        if (grids && grid_inside(S.pair, px, py, pz)) {
          if (fabs(q) >= 1e-9) {
            const O4 v = out[(size_t)w * M + i];
            ax = v.x; ay = v.y; az = v.z; ae = v.w;
          }
        } else if (S.has_yukawa && q != 0.0) {
          yukawa_atom<TC>(S, (TC)px, (TC)py, (TC)pz, (TC)q, ax, ay, az, ae);
        }
        fx += ax; fy += ay; fz += az; e += ae;
        if (S.torque_mode == 1) {
          const TC lx = (TC)(px - qx), ly = (TC)(py - qy), lz = (TC)(pz - qz);
          tx += ly * az - lz * ay; ty += lz * ax - lx * az; tz += lx * ay - ly * ax;
        }
      }
    }
    TC rx = 0, ry = 0, rz = 0, rtx = 0, rty = 0, rtz = 0;
    if (S.has_rev && valid && S.alpha > 1e-12 && !(rc - S.rec_max_r > S.rev_cull)) {
      for (int k = 0; k < S.n_rec_chunk; ++k) {
        if (!rec_chunk_near(S, k, cx, cy, cz, S.rev_cull)) continue;
        const int j1 = min((k + 1) * PST_REC_CHUNK, S.n_rec);
        for (int j = k * PST_REC_CHUNK + lane; j < j1; j += G) {
          double x, y, z;
          rec_in_lig(S, R, cx, cy, cz, j, x, y, z);
          if (grid_inside(S.lig_born, x, y, z) && fabs(S.rec_q[j]) >= 1e-9) {
            const O4 v = out_rev[(size_t)w * S.n_rec + j];
            const TC ax = v.x, ay = v.y, az = v.z;
            rx += ax; ry += ay; rz += az;
            if (S.torque_mode == 1) {
              const TC lx = (TC)(x - S.cq_body[0]), ly = (TC)(y - S.cq_body[1]),
                       lz = (TC)(z - S.cq_body[2]);
              rtx += ly * az - lz * ay; rty += lz * ax - lx * az; rtz += lx * ay - ly * ax;
            }
          }
        }
      }
    }
    double Fx = group_sum<G>((double)fx), Fy = group_sum<G>((double)fy), Fz = group_sum<G>((double)fz);
    double Ee = group_sum<G>((double)e);
    double Rx = group_sum<G>((double)rx), Ry = group_sum<G>((double)ry), Rz = group_sum<G>((double)rz);
    double Tx = 0, Ty = 0, Tz = 0;
    if (S.torque_mode == 1) {
      Tx = group_sum<G>((double)tx); Ty = group_sum<G>((double)ty); Tz = group_sum<G>((double)tz);
      double a = group_sum<G>((double)rtx), b = group_sum<G>((double)rty), c = group_sum<G>((double)rtz);
      Tx -= R[0] * a + R[1] * b + R[2] * c;
      Ty -= R[3] * a + R[4] * b + R[5] * c;
      Tz -= R[6] * a + R[7] * b + R[8] * c;
    }
    if (valid && lane == 0) {
      Fx -= R[0] * Rx + R[1] * Ry + R[2] * Rz;
      Fy -= R[3] * Rx + R[4] * Ry + R[5] * Rz;
      Fz -= R[6] * Rx + R[7] * Ry + R[8] * Rz;
      W.F[3 * (size_t)t] = Fx; W.F[3 * (size_t)t + 1] = Fy; W.F[3 * (size_t)t + 2] = Fz;
      if (W.T) { W.T[3 * (size_t)t] = Tx; W.T[3 * (size_t)t + 1] = Ty; W.T[3 * (size_t)t + 2] = Tz; }
      if (W.E) W.E[t] = Ee;
    }
  }
}

template <typename TC, int G>
static hipError_t launch_binned_g(const ForceSys& S, const ForceWork& W, const BinBufs& B,
                                  int max_n, int blocks_cap, hipStream_t st) {
  const int gpb = kForceBlock / G;
  int blocks = (max_n + gpb - 1) / gpb;
  if (blocks > blocks_cap) blocks = blocks_cap;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_bin_pass<TC, G, 0>), dim3(blocks), dim3(kForceBlock), 0, st, S, W, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_bin_pass<TC, G, 1>), dim3(blocks), dim3(kForceBlock), 0, st, S, W, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_bin_gather<TC>), dim3(B.gather_blocks), dim3(256), 0, st, S, W, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_bin_reduce<TC, G>), dim3(blocks), dim3(kForceBlock), 0, st, S, W, B);
  return hipGetLastError();
}

bool binned_eligible(const ForceSys& S, const ForceWork& W, const BinBufs& B, int max_n) {
  if (!S.paired || !W.quat || W.lig_pos || !W.pos || W.out_compact || W.n_cap > 0) return false;
  if (W.rev_skip_mode != 0) return false;
  if (S.has_rev && S.lig_born.dtype != 0) return false;
  if (!B.hist || max_n > B.max_traj || S.M > B.max_m || S.n_rec > B.max_rec) return false;
  return true;
}

hipError_t launch_forces_binned(const ForceSys& S, const ForceWork& W, const BinBufs& B,
                                int precision, int G, int max_n, int blocks_cap, hipStream_t st) {
  if (max_n <= 0) return hipSuccess;
  if (precision == 0) {
    switch (G) {
      case 8: return launch_binned_g<double, 8>(S, W, B, max_n, blocks_cap, st);
      case 16: return launch_binned_g<double, 16>(S, W, B, max_n, blocks_cap, st);
      case 32: return launch_binned_g<double, 32>(S, W, B, max_n, blocks_cap, st);
      default: return launch_binned_g<double, 64>(S, W, B, max_n, blocks_cap, st);
    }
  }
  switch (G) {
    case 8: return launch_binned_g<float, 8>(S, W, B, max_n, blocks_cap, st);
    case 16: return launch_binned_g<float, 16>(S, W, B, max_n, blocks_cap, st);
    case 32: return launch_binned_g<float, 32>(S, W, B, max_n, blocks_cap, st);
    default: return launch_binned_g<float, 64>(S, W, B, max_n, blocks_cap, st);
  }
}

int bricks_per_axis(int n_nodes) { return (n_nodes - 1 + kBrick - 1) / kBrick; }
int bin_items_bound(long max_req, int n_bins) { return n_bins + (int)(max_req / kItem) + 1; }

}  // namespace pst
