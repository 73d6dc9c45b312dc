// Per-trajectory BD step kernels (FP64 state) for gfx950.
//
// Restates the flattened-pairs step of GPUBatchSimulator.run
// (pystarc/simulation/gpu_batch_simulator.py:783-1356):
//   reaction check :855-871, escape / return :879-911, adaptive dt :961-1067,
//   HI D_parallel :1074-1084, Ermak-McCammon translation :1085-1094,
//   rotation :1095-1101, overlap rejection :1106-1128, Brownian bridge
//   :1142-1252, bookkeeping :1253-1254, early stop :1347-1348.
//
// Two drivers share the same __device__ physics:
//  * replay kernels take the reference's host NumPy draws, indexed by the
//    rank of a trajectory in the reference's ordered index lists, so a run
//    reproduces the reference's reaction / escape flags bit for bit;
//  * production kernels draw Philox4x32-10 numbers keyed by (trajectory id,
//    trajectory step, stream), keep a fixed pool of slots busy by refilling a
//    finished slot with the next trajectory, and never synchronise with the host.
#include <hip/hip_runtime.h>
#include "pst_device.hpp"
#include "pst_kernels.hpp"
#include "pst_rpy.hpp"
#include "pst_outer.hpp"

namespace pst {

constexpr int kStepBlock = 256;

// Random-stream id of the run's trajectory t: trajectory ids of a shard run
// are offset by its first id, so a shard draws exactly what the same
// trajectories draw in the whole run.
__device__ inline uint64_t rng_key(const StepParams& P, long long t) {
  return (uint64_t)(P.traj_base + t);
}

// ----------------------------------------------------------------------------
// shared step physics
// ----------------------------------------------------------------------------
// Pair distance |R b_p + c - r_p| (gpu_batch_simulator.py:1737-1751).
__device__ inline double pair_dist(const StepParams& P, const double* R, double cx, double cy,
                                   double cz, int p) {
  const double bx = P.lig_gho[3 * p], by = P.lig_gho[3 * p + 1], bz = P.lig_gho[3 * p + 2];
  const double gx = R[0] * bx + R[1] * by + R[2] * bz + cx - P.rec_gho[3 * p];
  const double gy = R[3] * bx + R[4] * by + R[5] * bz + cy - P.rec_gho[3 * p + 1];
  const double gz = R[6] * bx + R[7] * by + R[8] * bz + cz - P.rec_gho[3 * p + 2];
  return sqrt(gx * gx + gy * gy + gz * gz);
}

// Pairs within their cutoff (the checks below use each distance as it is
// formed: no per-lane array, which would live in scratch memory).
__device__ inline int pairs_within(const StepParams& P, const double* pos, const double* q) {
  double R[9];
  quat_to_rot(q, R);
  int n = 0;
  for (int p = 0; p < P.n_pairs; ++p)
    n += pair_dist(P, R, pos[0], pos[1], pos[2], p) < P.cutoff[p] ? 1 : 0;
  return n;
}

__device__ inline bool reacted_at(const StepParams& P, const double* pos, const double* q) {
  if (P.n_pairs == 0) return false;
  return pairs_within(P, pos, q) >= P.n_needed;
}

// dt per trajectory before the cap and the floor, flattened mode (:975-1004).
__device__ inline double adaptive_dt_raw(const StepParams& P, double r) {
  const double D_t = P.D_t;
  const bool in_outer = r > 1.1 * P.r_b;
  const double dt_pair = (0.1 * 0.1 / 2.0) * r * r / D_t;
  double dt_force = dt_pair;
  if (P.adt_has_force) {
    const double lam = P.adt_debye;
    const double e = exp(-r / lam);
    const double Fr0 = P.adt_V0 * e * (1.0 / (r * r) + 1.0 / (r * lam));
    const double Fr1 = -P.adt_V0 * e / (lam * lam) - 2.0 * Fr0 / r;
    const double DFr1 = fabs(D_t * Fr1);
    if (DFr1 > 1e-15 && r < 3.0 * lam) dt_force = 0.01 / DFr1;
  }
  const double dist_b = fmax(r - P.r_b, 1e-3);
  const double dist_e = fmax(P.r_esc - r, 1e-3);
  const double de = fmin(dist_b, dist_e);
  const double dt_edge = de * de / (18.0 * D_t);
  return in_outer ? fmin(fmin(dt_force, dt_edge), dt_pair) : fmin(dt_force, dt_pair);
}

// dt per trajectory, flattened mode (:975-1004, :1061-1067): cap, then floor.
__device__ inline double adaptive_dt(const StepParams& P, double r) {
  double dt = adaptive_dt_raw(P, r);
  if (P.max_dt > 0) dt = fmin(dt, P.max_dt);
  if (P.min_core_dt > 0) dt = fmax(dt, P.min_core_dt);
  return dt;
}

// State-machine mode: the same dt, refined near a reachable reaction surface
// before the cap and floor (:1014-1067): rc = min over reactions reachable
// from the current state of max(d_r - cutoff_r, 0); when 12 sqrt(2 D_t dt)
// exceeds rc, dt <- min(dt, max(rc^2 / (288 D_t), min_core_rxn_dt)).
__device__ inline double adaptive_dt_sm(const StepParams& P, double r, int state,
                                        const double* R, const double* pos) {
  double dt = adaptive_dt_raw(P, r);
  double rc = 1.0e30;
  for (int p = 0; p < P.n_pairs; ++p)
    if (P.sm_before[p] == state)
      rc = fmin(rc, fmax(pair_dist(P, R, pos[0], pos[1], pos[2], p) - P.cutoff[p], 0.0));
  const double sig = sqrt(2.0 * P.D_t * dt);
  if (12.0 * sig > rc) {
    double crit = rc * rc / (288.0 * P.D_t);
    if (P.min_core_rxn_dt > 0) crit = fmax(crit, P.min_core_rxn_dt);
    dt = fmin(dt, crit);
  }
  if (P.max_dt > 0) dt = fmin(dt, P.max_dt);
  if (P.min_core_dt > 0) dt = fmax(dt, P.min_core_dt);
  return dt;
}

// State-machine reaction check (:1683-1735): the lowest-index reaction whose
// source state is the trajectory's and whose pair is within its cutoff.
__device__ inline int sm_check(const StepParams& P, const double* pos, const double* q,
                               int state) {
  double R[9];
  quat_to_rot(q, R);
  for (int p = 0; p < P.n_pairs; ++p)
    if (P.sm_before[p] == state && pair_dist(P, R, pos[0], pos[1], pos[2], p) < P.cutoff[p])
      return p;
  return -1;
}

// Scalar relative diffusion, isotropic D_parallel with HI (:1074-1084).
__device__ inline double diffusion(const StepParams& P, double r) {
  if (!P.use_hi) return P.D_t;
  const double rm1 = 1.0 / fmax(r, 0.01);
  const double rm3 = rm1 * rm1 * rm1;
  const double Di = -2.0 * P.hi_Df * (2.0 * rm1 - (4.0 / 3.0) * P.hi_a2 * rm3) / (8.0 * M_PI);
  return fmax(P.hi_Df * P.hi_ainv + Di, 1e-6);
}

// Overlap: min_j |c - r_j| < threshold (:1106-1128).  The test is skipped
// when |c| - max|r_j| already exceeds the threshold; otherwise only the
// receptor cells that intersect the cube of half-edge threshold (+1e-9)
// about c are visited.  Every atom closer than the threshold lies in one of
// them, so the minimum compared with the threshold, and the outcome, are
// those of the loop over all atoms.
__device__ inline bool overlaps(const StepParams& P, const ForceSys& S, double cx, double cy,
                                double cz) {
  const double rc = norm3(cx, cy, cz);
  const double reach = P.overlap_thresh + 1e-9;
  if (rc - S.rec_max_r > reach) return false;
  const double c[3] = {cx, cy, cz};
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = max((int)floor((c[d] - reach - S.cell_o[d]) * S.cell_inv), 0);
    hi[d] = min((int)floor((c[d] + reach - S.cell_o[d]) * S.cell_inv), S.cell_n[d] - 1);
    if (lo[d] > hi[d]) return false;
  }
  double m2 = 1e300;
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j) {
      const int row = (i * S.cell_n[1] + j) * S.cell_n[2];
      const int a0 = S.cell_start[row + lo[2]], a1 = S.cell_start[row + hi[2] + 1];
      for (int a = a0; a < a1; ++a) {
        const double dx = cx - S.cell_pos[3 * a], dy = cy - S.cell_pos[3 * a + 1],
                     dz = cz - S.cell_pos[3 * a + 2];
        m2 = fmin(m2, dx * dx + dy * dy + dz * dz);
      }
    }
  return sqrt(m2) < P.overlap_thresh;
}

// Pair distances before a move; keeps the pose for an overlap rejection.
struct MoveSave { double p[3], q[4]; };

__device__ inline void move_begin(const StepParams& P, const double* pos, const double* q,
                                  double* x0, int xs, MoveSave& sv) {
  double R[9];
  quat_to_rot(q, R);
  for (int p = 0; p < P.n_pairs; ++p)
    x0[p * xs] = pair_dist(P, R, pos[0], pos[1], pos[2], p) - P.cutoff[p];
  sv.p[0] = pos[0]; sv.p[1] = pos[1]; sv.p[2] = pos[2];
  sv.q[0] = q[0]; sv.q[1] = q[1]; sv.q[2] = q[2]; sv.q[3] = q[3];
}

// Overlap rejection (:1106-1128) and new pair distances.  Returns whether any
// pair is above its cutoff both before and after the move.
__device__ inline bool move_end(const StepParams& P, const ForceSys& S, double* pos, double* q,
                                const MoveSave& sv, const double* x0, double* x1, int xs,
                                bool* rejected) {
  *rejected = false;
  if (P.use_overlap && S.n_rec > 0 && overlaps(P, S, pos[0], pos[1], pos[2])) {
    pos[0] = sv.p[0]; pos[1] = sv.p[1]; pos[2] = sv.p[2];
    q[0] = sv.q[0]; q[1] = sv.q[1]; q[2] = sv.q[2]; q[3] = sv.q[3];
    *rejected = true;
  }
  bool any = false;
  if (P.n_pairs) {
    double R[9];
    quat_to_rot(q, R);
    for (int p = 0; p < P.n_pairs; ++p) {
      const double a = x0[p * xs], b = pair_dist(P, R, pos[0], pos[1], pos[2], p) - P.cutoff[p];
      x1[p * xs] = b;
      any = any || (a > 0 && b > 0);
    }
  }
  return any;
}

// Scalar-diffusion move: Ermak-McCammon translation (:1085-1094) and the
// right-multiplied rotation (:1095-1101).
__device__ inline bool move_one(const StepParams& P, const ForceSys& S, double* pos, double* q,
                                const double* Fv, const double* Tv, const double* nt,
                                const double* nr, double dt, double D, double* x0, double* x1,
                                int xs, bool* rejected) {
  MoveSave sv;
  move_begin(P, pos, q, x0, xs, sv);
  const double sig = sqrt(2.0 * D * dt);
  pos[0] += D * dt * Fv[0] + sig * nt[0];
  pos[1] += D * dt * Fv[1] + sig * nt[1];
  pos[2] += D * dt * Fv[2] + sig * nt[2];
  const double sr = sqrt(2.0 * P.D_r * dt);
  apply_rotation(q, P.D_r * dt * Tv[0] + sr * nr[0], P.D_r * dt * Tv[1] + sr * nr[1],
                 P.D_r * dt * Tv[2] + sr * nr[2]);
  return move_end(P, S, pos, q, sv, x0, x1, xs, rejected);
}

// Two-body RPY move: both bodies translate and turn with the 12x12 mobility
// and its Cholesky noise (xi: 12 standard normals), then the pose is
// re-expressed in the turned receptor frame.
__device__ inline bool move_one_rpy(const StepParams& P, const ForceSys& S, double* pos,
                                    double* q, const double* FL, const double* TL,
                                    const double* xi, double dt, double* x0, double* x1, int xs,
                                    bool* rejected) {
  MoveSave sv;
  move_begin(P, pos, q, x0, xs, sv);
  double FR[3], TR[3], dX[12];
  rpy_receptor_load(S.torque_mode, pos, FL, TL, FR, TR);
  rpy_displacement(P.a_rec, P.a_lig, P.kT_over_eta, dt, pos, FR, FL, TR, TL, xi, dX);
  rpy_apply(dX, pos, q);
  return move_end(P, S, pos, q, sv, x0, x1, xs, rejected);
}

// Brownian-bridge pair count (:1157-1184, :1236-1239).
__device__ inline int bridge_count(const StepParams& P, const double* x0, const double* x1,
                                   int xs, double dt, const double* u) {
  int n = 0;
  for (int p = 0; p < P.n_pairs; ++p) {
    const double a = x0[p * xs], b = x1[p * xs];
    double pc = 0.0;
    if (a > 0 && b > 0) {
      const double Ddt = fmax((P.has_bb_deff ? P.bb_deff[p] : P.D_t) * dt, 1e-30);
      pc = exp(-a * b / Ddt);
    }
    const bool fired = (u[p] < pc) || (b < 0.0);
    n += fired ? 1 : 0;
  }
  return n;
}

// ----------------------------------------------------------------------------
// ordered compaction (one workgroup; replay mode and small lists)
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_compact_ordered(const int* in_list, int n_in,
                                                          const int* n_in_dev,
                                                          const unsigned char* flag, int* out,
                                                          int* n_out) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int n = n_in_dev ? *n_in_dev : n_in;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + tid;
    int s = -1;
    bool f = false;
    if (i < n) {
      s = in_list ? in_list[i] : i;
      f = flag[s] != 0;
    }
    const unsigned long long m = __ballot(f);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int wp = 0, tot = 0;
    for (int k = 0; k < 16; ++k) {
      if (k < wave) wp += wsum[k];
      tot += wsum[k];
    }
    if (f) out[carry + wp + before] = s;
    __syncthreads();
    if (tid == 0) carry += tot;
    __syncthreads();
  }
  if (tid == 0) *n_out = carry;
}

hipError_t launch_compact_ordered(const int* in_list, int n_in, const int* n_in_dev,
                                  const unsigned char* flag, int* out, int* n_out, hipStream_t st) {
  hipLaunchKernelGGL(k_compact_ordered, dim3(1), dim3(1024), 0, st, in_list, n_in, n_in_dev, flag,
                     out, n_out);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// replay kernels
// ----------------------------------------------------------------------------
// Phase A: reaction check on running trajectories; flag escape candidates.
__global__ void k_replay_begin(StepParams P, ReplayBufs B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n_slots) return;
  B.f_esc[s] = 0;
  B.f_ret[s] = 0;
  B.f_sr[s] = 0;
  if (B.status[s] != PST_RUNNING) { B.f_run[s] = 0; return; }
  B.f_run[s] = 1;
  atomicAdd(&B.counts[0], 1);
  const double* pos = B.pos + 3 * (size_t)s;
  const bool react = reacted_at(P, pos, B.quat + 4 * (size_t)s);
  if (react) {
    B.status[s] = PST_REACTED;
    atomicAdd(&B.counts[7], 1);
  }
  const double r = norm3(pos[0], pos[1], pos[2]);
  if (r >= P.r_esc && !react) B.f_esc[s] = 1;
}

hipError_t replay_begin(const StepParams& P, const ReplayBufs& B, hipStream_t st) {
  hipError_t e = hipMemsetAsync(B.counts, 0, 8 * sizeof(int), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_replay_begin, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = launch_compact_ordered(nullptr, B.n_slots, nullptr, B.f_run, B.run_list, &B.counts[0], st);
  if (e != hipSuccess) return e;
  return launch_compact_ordered(B.run_list, 0, &B.counts[0], B.f_esc, B.esc_list, &B.counts[1],
                                st);
}

// Phase B1: u_outer[e] < return_prob decides return vs escape (:884-911).
__global__ void k_replay_esc_decide(StepParams P, ReplayBufs B, const double* u) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B.counts[1]) return;
  const int s = B.esc_list[e];
  if (u[e] < P.return_prob) {
    B.f_ret[s] = 1;
  } else {
    B.status[s] = PST_ESCAPED;
  }
}

hipError_t replay_escape_decide(const StepParams& P, const ReplayBufs& B, const double* u_outer,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_replay_esc_decide, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B, u_outer);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_compact_ordered(B.esc_list, 0, &B.counts[1], B.f_ret, B.ret_list, &B.counts[2],
                                st);
}

// Phase B2: snap returners to the b-sphere with a fresh Shoemake rotation.
__global__ void k_replay_esc_apply(StepParams P, ReplayBufs B, const double* uq) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B.counts[2]) return;
  const int s = B.ret_list[j];
  double* pos = B.pos + 3 * (size_t)s;
  const double r = norm3(pos[0], pos[1], pos[2]);
  pos[0] = pos[0] / r * P.r_b;
  pos[1] = pos[1] / r * P.r_b;
  pos[2] = pos[2] / r * P.r_b;
  shoemake(uq[3 * j], uq[3 * j + 1], uq[3 * j + 2], B.quat + 4 * (size_t)s);
}

hipError_t replay_escape_apply(const StepParams& P, const ReplayBufs& B, const double* u_quat,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_replay_esc_apply, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B, u_quat);
  return hipGetLastError();
}

__global__ void k_flag_sr(ReplayBufs B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n_slots) return;
  B.f_sr[s] = (B.f_run[s] && B.status[s] == PST_RUNNING) ? 1 : 0;
}

hipError_t replay_build_sr(const ReplayBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_flag_sr, dim3((B.n_slots + kStepBlock - 1) / kStepBlock), dim3(kStepBlock),
                     0, st, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_compact_ordered(B.run_list, 0, &B.counts[0], B.f_sr, B.sr_list, &B.counts[3], st);
}

// Phase C: dt, D, Ermak-McCammon move, overlap, new pair distances.
__global__ void k_replay_move(StepParams P, ForceSys S, ReplayBufs B, const double* nt,
                              const double* nr, int n_sr) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_sr) return;
  const int s = B.sr_list[w];
  double* pos = B.pos + 3 * (size_t)s;
  double* q = B.quat + 4 * (size_t)s;
  const double r = norm3(pos[0], pos[1], pos[2]);
  const double dt = adaptive_dt(P, r);
  const double D = diffusion(P, r);
  B.dt[w] = dt;
  bool rej;
  const bool any = move_one(P, S, pos, q, B.F + 3 * (size_t)s, B.T + 3 * (size_t)s, nt + 3 * (size_t)w,
                            nr + 3 * (size_t)w, dt, D, B.x0 + (size_t)w * PST_MAX_PAIRS,
                            B.x1 + (size_t)w * PST_MAX_PAIRS, 1, &rej);
  if (rej) atomicAdd(&B.counts[6], 1);
  if (any) atomicOr(&B.counts[4], 1);
}

hipError_t replay_move(const StepParams& P, const ForceSys& S, const ReplayBufs& B,
                       const double* noise_t, const double* noise_r, int n_sr, hipStream_t st) {
  if (n_sr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_replay_move, dim3((n_sr + kStepBlock - 1) / kStepBlock), dim3(kStepBlock),
                     0, st, P, S, B, noise_t, noise_r, n_sr);
  return hipGetLastError();
}

// Phase D: bridge with host uniforms (only drawn when any pair is above on
// both ends), then n_steps += 1 for every still-running trajectory.
__global__ void k_replay_bridge(StepParams P, ReplayBufs B, const double* u_bb, int n_sr) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_sr) return;
  const int s = B.sr_list[w];
  if (u_bb && P.n_pairs) {
    const int n = bridge_count(P, B.x0 + (size_t)w * PST_MAX_PAIRS,
                               B.x1 + (size_t)w * PST_MAX_PAIRS, 1, B.dt[w],
                               u_bb + (size_t)w * P.n_pairs);
    if (n >= P.n_needed && B.status[s] == PST_RUNNING) {
      B.status[s] = PST_REACTED;
      atomicAdd(&B.counts[5], 1);
    }
  }
  B.n_steps[s] += 1;
}

hipError_t replay_bridge(const StepParams& P, const ReplayBufs& B, const double* u_bb, int n_sr,
                         hipStream_t st) {
  if (n_sr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_replay_bridge, dim3((n_sr + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B, u_bb, n_sr);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------
// production kernels
// ----------------------------------------------------------------------------
// Start trajectory t in slot s: centroid uniformly on the b-sphere, uniform
// orientation (gpu_batch_simulator.py:509-513, Philox stream RNG_INIT).
__device__ inline void start_traj(const StepParams& P, const ProdBufs& B, int s, long long t) {
  double z0, z1, z2, z3, u0, u1, u2, u3;
  philox_normal2(P.seed, rng_key(P, t), 0xFFFFFFFFu, RNG_INIT, 0, z0, z1);
  philox_normal2(P.seed, rng_key(P, t), 0xFFFFFFFFu, RNG_INIT, 1, z2, z3);
  philox_uniform2(P.seed, rng_key(P, t), 0xFFFFFFFFu, RNG_INIT, 2, u0, u1);
  philox_uniform2(P.seed, rng_key(P, t), 0xFFFFFFFFu, RNG_INIT, 3, u2, u3);
  const double n = norm3(z0, z1, z2);
  double* pos = B.pos + 3 * (size_t)s;
  pos[0] = z0 / n * P.r_b;
  pos[1] = z1 / n * P.r_b;
  pos[2] = z2 / n * P.r_b;
  shoemake(u0, u1, u2, B.quat + 4 * (size_t)s);
  B.traj[s] = t;
  B.n_steps[s] = 0;
  B.status[s] = PST_RUNNING;
  if (P.sm_active) {
    B.sm_state[s] = 0;
    B.sm_mask[s] = 0u;
  }
  if (B.obs.traj) {
    const ObsBufs& O = B.obs;
    double* rec = O.traj + (size_t)t * kObsTraj;
    const double* q = B.quat + 4 * (size_t)s;
    for (int k = 0; k < 3; ++k) { rec[k] = pos[k]; O.s_nm[7 * (size_t)s + k] = pos[k]; }
    for (int k = 0; k < 4; ++k) { rec[3 + k] = q[k]; O.s_nm[7 * (size_t)s + 3 + k] = q[k]; }
    O.s_min[s] = 1e30;
    O.s_step[s] = 0;
    O.s_time[s] = 0.0;
    O.s_ret[s] = 0;
    O.s_prev_r[s] = P.r_b;   // on the b-sphere by construction (no rounding of |pos|)
  }
}

// Copy a slot's accumulated observables into its trajectory record.
__device__ inline void obs_flush(const StepParams& P, const ProdBufs& B, int s, int bridge) {
  const ObsBufs& O = B.obs;
  double* rec = O.traj + (size_t)B.traj[s] * kObsTraj;
  const double* pos = B.pos + 3 * (size_t)s;
  const double* q = B.quat + 4 * (size_t)s;
  rec[7] = O.s_min[s];
  rec[8] = O.s_step[s];
  for (int k = 0; k < 7; ++k) rec[9 + k] = O.s_nm[7 * (size_t)s + k];
  rec[16] = O.s_time[s];
  rec[17] = O.s_ret[s];
  rec[18] = bridge;
  for (int k = 0; k < 3; ++k) rec[19 + k] = pos[k];
  for (int k = 0; k < 4; ++k) rec[22 + k] = q[k];
  rec[26] = P.n_pairs ? pairs_within(P, pos, q) : 0;
}

__device__ inline void finish_traj(const StepParams& P, const ProdBufs& B, int s, int outcome,
                                   int bridge = 0) {
  const long long t = B.traj[s];
  if (B.obs.traj) obs_flush(P, B, s, bridge);
  if (outcome == PST_REACTED) atomicAdd(&B.ctr->n_reacted, 1ull);
  if (outcome == PST_ESCAPED) atomicAdd(&B.ctr->n_escaped, 1ull);
  if (outcome == PST_MAXSTEPS) atomicAdd(&B.ctr->n_maxsteps, 1ull);
  if (B.out_status) B.out_status[t] = (signed char)outcome;
  if (B.out_steps) B.out_steps[t] = (int)B.n_steps[s];
  const unsigned long long nt = atomicAdd(&B.ctr->next_traj, 1ull);
  if ((long long)nt < P.n_total) {
    start_traj(P, B, s, (long long)nt);
  } else {
    B.status[s] = PST_IDLE;
  }
}

// Fire reaction r on slot s: new state, counts; true when the new state is terminal.
__device__ inline bool sm_fire(const StepParams& P, const ProdBufs& B, int s, int r, bool bridge) {
  const int after = P.sm_after[r];
  B.sm_state[s] = after;
  if (bridge)
    atomicAdd(&B.ctr->sm_fired_bridge[r], 1ull);
  else
    atomicAdd(&B.ctr->sm_fired[r], 1ull);
  const unsigned int bit = 1u << r;
  if (!(B.sm_mask[s] & bit)) {
    B.sm_mask[s] |= bit;
    atomicAdd(&B.ctr->sm_unique[r], 1ull);
  }
  return after >= 0 && ((P.sm_terminal >> after) & 1);
}

__global__ void k_prod_init(StepParams P, ProdBufs B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n_slots) return;
  if (s < P.n_total) {
    start_traj(P, B, s, s);
  } else {
    B.status[s] = PST_IDLE;
  }
  if (s == 0) {
    RunCounters z = {};
    z.next_traj = (unsigned long long)min((long long)B.n_slots, P.n_total);
    *B.ctr = z;
  }
}

hipError_t prod_init(const StepParams& P, const ProdBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_prod_init, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B);
  return hipGetLastError();
}

// Workgroup-aggregated append to an unordered list: the waves reserve their
// ranges in LDS and the workgroup takes one global range (one atomic on the
// counter per workgroup instead of one per wave; every wave of the launch
// appends to the same counter).  All threads of the workgroup call it.
__device__ inline void block_append(int* list, int* counter, int v, bool pred) {
  __shared__ int wg_count, wg_base;
  if (threadIdx.x == 0) wg_count = 0;
  __syncthreads();
  const unsigned long long m = __ballot(pred);
  const int lane = __lane_id();
  const int leader = m ? __ffsll((long long)m) - 1 : 0;
  int off = 0;
  if (m && lane == leader) off = atomicAdd(&wg_count, __popcll(m));
  off = __shfl(off, leader);
  __syncthreads();
  if (threadIdx.x == 0) wg_base = wg_count ? atomicAdd(counter, wg_count) : 0;
  __syncthreads();
  if (pred) list[wg_base + off + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

// Start of step: stop test, reaction check, escape / return, refill,
// partition the still-running slots into the grid list and the surrogate list.
// Workgroup `blk` of this system's launch.
__device__ inline void prod_pre_body(const StepParams& P, const ProdBufs& B, int par, int blk) {
  const int s = blk * blockDim.x + threadIdx.x;
  const RunCounters* C = B.ctr;
  const bool stop = C->stop != 0;
  bool live = false, far = false;
  if (s < B.n_slots && !stop && B.status[s] == PST_RUNNING) {
    // a refilled slot is checked again as its new trajectory (guarded); the
    // outcome of a check ends the trajectory at one call site, so the refill
    // code is inlined once
    for (int guard = 0; guard < 8; ++guard) {
      double* pos = B.pos + 3 * (size_t)s;
      double* q = B.quat + 4 * (size_t)s;
      const long long t = B.traj[s];
      const unsigned int ns = B.n_steps[s];
      int outcome = 0;
      if ((long long)ns >= P.max_steps) {
        outcome = PST_MAXSTEPS;
      } else {
        bool fired = false;
        if (P.sm_active) {
          const int rx = sm_check(P, pos, q, B.sm_state[s]);
          if (rx >= 0) {
            fired = true;
            if (sm_fire(P, B, s, rx, false)) outcome = PST_REACTED;
          }
        } else if (reacted_at(P, pos, q)) {
          outcome = PST_REACTED;
        }
        if (!outcome) {
          const double r = norm3(pos[0], pos[1], pos[2]);
          if (P.lmz_on) {
            // the outer excursion ran in k_prod_outer this step (returned: the
            // pose is already back on the b-sphere)
            if (B.lmz_escaped[s]) {
              B.lmz_escaped[s] = 0;
              outcome = PST_ESCAPED;
            }
          } else if (r >= P.r_esc && !fired) {
            double u, unused;
            philox_uniform2(P.seed, rng_key(P, t), ns, RNG_ESC, 0, u, unused);
            if (u < P.return_prob) {
              pos[0] = pos[0] / r * P.r_b;
              pos[1] = pos[1] / r * P.r_b;
              pos[2] = pos[2] / r * P.r_b;
              double u0, u1, u2, u3;
              philox_uniform2(P.seed, rng_key(P, t), ns, RNG_RETQ, 0, u0, u1);
              philox_uniform2(P.seed, rng_key(P, t), ns, RNG_RETQ, 1, u2, u3);
              shoemake(u0, u1, u2, q);
              atomicAdd(&B.ctr->n_returns, 1ull);
              if (B.obs.traj) B.obs.s_ret[s] += 1;
            } else {
              outcome = PST_ESCAPED;
            }
          }
        }
      }
      if (outcome) {
        finish_traj(P, B, s, outcome);
        if (B.status[s] != PST_RUNNING) break;
        continue;
      }
      live = true;
      far = P.surrogate_r > 0 && norm3(pos[0], pos[1], pos[2]) > P.surrogate_r;
      break;
    }
  }
  RunCounters* Cw = B.ctr;
  block_append(B.mlp_list, &Cw->n_mlp, s, live && far);
  // Grid-path trajectories get a direction key for the locality sort.  For
  // the fused force kernel (no reverse list): lab octant of the centroid
  // (bits 5-7, the receptor-grid region it gathers from) and the octant +
  // dominant axis of the receptor direction in the ligand frame (bits 0-4,
  // the ligand-grid region of its reverse Born).  For the split evaluation
  // the two passes get keys of their own: the cube-map cell of the centroid
  // direction in the lab frame (bits 0-7) orders the ligand-atom pass, whose
  // gathers depend only on where the ligand sits in the receptor grids, and
  // the cube-map cell of the receptor direction in the ligand frame (bits
  // 8-15) the reverse pass.  Surrogate slots get -2, the rest -1.
  __shared__ int hist[256];
  __shared__ int hist2[256];
  hist[threadIdx.x] = 0;
  hist2[threadIdx.x] = 0;
  __syncthreads();
  if (s < B.n_slots) {
    int k = live && far ? -2 : -1;
    if (live && !far) {
      const double* pos = B.pos + 3 * (size_t)s;
      double R[9];
      quat_to_rot(B.quat + 4 * (size_t)s, R);
      const double cx = pos[0], cy = pos[1], cz = pos[2];
      // d = R^T (-c)
      const double dx = -(R[0] * cx + R[3] * cy + R[6] * cz);
      const double dy = -(R[1] * cx + R[4] * cy + R[7] * cz);
      const double dz = -(R[2] * cx + R[5] * cy + R[8] * cz);
      if (B.rev_list) {
        k = dir_cell(cx, cy, cz);
        const int c = dir_cell(dx, dy, dz);
        atomicAdd(&hist2[c], 1);
        k |= c << 8;
      } else {
        const int lab = (cx > 0) | ((cy > 0) << 1) | ((cz > 0) << 2);
        const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
        const int dom = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        const int lig = ((dx > 0) | ((dy > 0) << 1) | ((dz > 0) << 2)) * 3 + dom;
        k = lab * 32 + lig;
      }
      atomicAdd(&hist[k & 255], 1);
    }
    B.key[s] = k;
  }
  __syncthreads();
  if (hist[threadIdx.x]) atomicAdd(&B.bins[threadIdx.x], hist[threadIdx.x]);
  if (hist2[threadIdx.x]) atomicAdd(&B.bins[512 + threadIdx.x], hist2[threadIdx.x]);
}

// Outer excursions of the step (LMZ mode), before k_prod_pre: a running
// trajectory that k_prod_pre would take to its escape test (no max-steps
// end, no reaction at the check) and that sits at r >= 1.1 b is carried
// through the outer region (nam_simulator.py:354-373).  Back on the b-sphere
// its pose is updated; an escape is left for k_prod_pre to finish (and refill)
// in the same step.  Its own kernel, so that the propagator's registers and
// stack stay out of k_prod_pre.
__device__ inline void prod_outer_body(const StepParams& P, const ProdBufs& B, int blk) {
  const int s = blk * blockDim.x + threadIdx.x;
  if (!P.lmz_on || s >= B.n_slots || B.ctr->stop != 0 || B.status[s] != PST_RUNNING) return;
  const unsigned int ns = B.n_steps[s];
  if ((long long)ns >= P.max_steps) return;
  double* pos = B.pos + 3 * (size_t)s;
  double* q = B.quat + 4 * (size_t)s;
  if (P.sm_active) {
    if (sm_check(P, pos, q, B.sm_state[s]) >= 0) return;
  } else if (reacted_at(P, pos, q)) {
    return;
  }
  if (norm3(pos[0], pos[1], pos[2]) < P.lmz_trigger) return;
  if (outer_propagate(P, rng_key(P, B.traj[s]), ns, pos, q)) {
    atomicAdd(&B.ctr->n_returns, 1ull);
    if (B.obs.traj) B.obs.s_ret[s] += 1;
  } else {
    B.lmz_escaped[s] = 1;
  }
}

__global__ void __launch_bounds__(kStepBlock) k_prod_outer(StepParams P, ProdBufs B) {
  prod_outer_body(P, B, blockIdx.x);
}

// Several systems in one launch: workgroup b belongs to system l with
// start[l] <= b < start[l + 1].
__device__ inline int slice_of(const int* start, int L, int b) {
  int l = 0;
  while (l + 1 < L && b >= start[l + 1]) ++l;
  return l;
}

__global__ void __launch_bounds__(kStepBlock) k_prod_pre(StepParams P, ProdBufs B, int par) {
  prod_pre_body(P, B, par, blockIdx.x);
}

__global__ void __launch_bounds__(kStepBlock) k_prod_outer_group(const StepParams* Pg,
                                                                 const ProdBufs* Bg,
                                                                 const int* start, int L) {
  const int l = slice_of(start, L, blockIdx.x);
  prod_outer_body(Pg[l], Bg[l], blockIdx.x - start[l]);
}

__global__ void __launch_bounds__(kStepBlock) k_prod_pre_group(const StepParams* Pg,
                                                               const ProdBufs* Bg,
                                                               const int* start, int L, int par) {
  const int l = slice_of(start, L, blockIdx.x);
  prod_pre_body(Pg[l], Bg[l], par, blockIdx.x - start[l]);
}

// Exclusive scan of 256 histogram bins (h) into scatter cursors (h + 256);
// clears the histogram for the next step and returns the total.
__device__ inline int scan256(int* h) {
  __shared__ int sc[256];
  const int t = threadIdx.x;
  const int v = h[t];
  sc[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = t >= off ? sc[t - off] : 0;
    __syncthreads();
    sc[t] += add;
    __syncthreads();
  }
  h[256 + t] = sc[t] - v;
  h[t] = 0;
  const int total = sc[255];
  __syncthreads();
  return total;
}

// Scatter cursors of the grid-path list (and of its reverse-pass order);
// sets the length of the grid-path list.
__device__ inline void bins_scan_body(const ProdBufs& B) {
  const int n = scan256(B.bins);
  if (B.rev_list) scan256(B.bins + 512);
  if (threadIdx.x == 255) B.ctr->n_sr = n;
}

__global__ void __launch_bounds__(256) k_bins_scan(ProdBufs B) { bins_scan_body(B); }

// one workgroup per system
__global__ void __launch_bounds__(256) k_bins_scan_group(const ProdBufs* Bg) {
  bins_scan_body(Bg[blockIdx.x]);
}

// The scatter of the counting sort.  Workgroup `blk` takes kSPT slots per
// thread (blockDim.x * kSPT consecutive slots): it ranks its slots inside
// each bin with LDS atomics, reserves one range per non-empty bin with one
// global atomic on the bin's cursor, then writes.  Every slot of the launch
// bumps one of ~216 shared cursors otherwise, and those atomics serialise
// per address; with 4,096 slots per workgroup there is one per bin and
// workgroup.  The order inside a bin is arbitrary either way.
template <int kSPT>
__device__ inline void bins_scatter_body(const ProdBufs& B, int blk) {
  __shared__ int cnt[256], cnt2[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) { cnt[i] = 0; cnt2[i] = 0; }
  __syncthreads();
  const int s0 = blk * blockDim.x * kSPT;
  const bool rev = B.rev_list != nullptr;
  int key[kSPT], loc[kSPT], loc2[kSPT];
#pragma unroll
  for (int j = 0; j < kSPT; ++j) {
    const int s = s0 + j * blockDim.x + threadIdx.x;
    const int k = s < B.n_slots ? B.key[s] : -1;
    key[j] = k;
    loc[j] = loc2[j] = 0;
    if (k >= 0) {
      loc[j] = atomicAdd(&cnt[k & 255], 1);
      if (rev) loc2[j] = atomicAdd(&cnt2[k >> 8], 1);
    }
  }
  __syncthreads();
  // the bins' base offsets replace their local counts
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    cnt[i] = cnt[i] ? atomicAdd(&B.bins[256 + i], cnt[i]) : 0;
    if (rev) cnt2[i] = cnt2[i] ? atomicAdd(&B.bins[768 + i], cnt2[i]) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSPT; ++j) {
    const int k = key[j];
    if (k < 0) continue;
    const int s = s0 + j * blockDim.x + threadIdx.x;
    B.plain_list[cnt[k & 255] + loc[j]] = s;
    if (rev) B.rev_list[cnt2[k >> 8] + loc2[j]] = s;
  }
}

constexpr int kScatterSPT = 16;   // slots per thread of k_bins_scatter (4,096 per workgroup)

__global__ void __launch_bounds__(256) k_bins_scatter(ProdBufs B) {
  bins_scatter_body<kScatterSPT>(B, blockIdx.x);
}

// grouped: the step launches' partition (kStepBlock slots per workgroup)
__global__ void k_bins_scatter_group(const ProdBufs* Bg, const int* start, int L) {
  const int l = slice_of(start, L, blockIdx.x);
  bins_scatter_body<1>(Bg[l], blockIdx.x - start[l]);
}

hipError_t prod_step_sort(const ProdBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_bins_scan, dim3(1), dim3(256), 0, st, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int per = 256 * kScatterSPT;
  hipLaunchKernelGGL(k_bins_scatter, dim3((B.n_slots + per - 1) / per), dim3(256), 0, st, B);
  return hipGetLastError();
}

// Histogram bins of the observables (the reference's linspace edges).
__device__ inline int obs_bin(double r, double rmax, int nb) {
  const int b = (int)floor(r / (rmax / nb));
  return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

// Move every still-running slot (both lists).  With OBS the step also
// accumulates the observables: per slot closest approach, step at it and the
// pose there, simulated time; per workgroup in LDS the radial histogram, the
// angular occupancy of the centroid direction, milestone crossings, the
// radial transition matrix and pair contact counts (flushed with one global
// atomic per non-zero bin).
template <bool OBS>
__device__ inline void prod_post_body(const StepParams& P, const ForceSys& S, const ProdBufs& B,
                                      int par, int blk) {
  __shared__ int h[OBS ? kObsHist : 1];
  if (OBS) {
    for (int i = threadIdx.x; i < kObsHist; i += blockDim.x) h[i] = 0;
    __syncthreads();
  }
  // the slots of this step's grid-path and surrogate lists, in slot order
  // (coalesced per-slot reads; the lists are in direction order)
  const int s = blk * blockDim.x + threadIdx.x;
  const bool active = s < B.n_slots && B.key[s] != -1;
  bool any = false;
  if (active) {
    double* pos = B.pos + 3 * (size_t)s;
    double* q = B.quat + 4 * (size_t)s;
    const long long t = B.traj[s];
    const unsigned int ns = B.n_steps[s];
    const double r = norm3(pos[0], pos[1], pos[2]);
    double dt;
    if (P.sm_active) {
      double R[9];
      quat_to_rot(q, R);
      dt = adaptive_dt_sm(P, r, B.sm_state[s], R, pos);
    } else {
      dt = adaptive_dt(P, r);
    }
    const double D = diffusion(P, r);
    B.dt[s] = dt;
    bool rej;
    // pair distances [pair][slot]: one coalesced row per pair
    const int xs = B.n_slots;
    double* x0 = B.x0 + s;
    double* x1 = B.x1 + s;
    if (P.hydro_mode == 1) {
      double xi[12];
#pragma unroll
      for (int k = 0; k < 6; ++k)
        philox_normal2(P.seed, rng_key(P, t), ns, RNG_NOISE, (uint32_t)k, xi[2 * k], xi[2 * k + 1]);
      any = move_one_rpy(P, S, pos, q, B.F + 3 * (size_t)s, B.T + 3 * (size_t)s, xi, dt, x0, x1,
                         xs, &rej);
    } else {
      double nt[3], nr[3];
      philox_normal2(P.seed, rng_key(P, t), ns, RNG_NOISE, 0, nt[0], nt[1]);
      philox_normal2(P.seed, rng_key(P, t), ns, RNG_NOISE, 1, nt[2], nr[0]);
      philox_normal2(P.seed, rng_key(P, t), ns, RNG_NOISE, 2, nr[1], nr[2]);
      any = move_one(P, S, pos, q, B.F + 3 * (size_t)s, B.T + 3 * (size_t)s, nt, nr, dt, D, x0, x1,
                     xs, &rej);
    }
    if (rej) atomicAdd(&B.ctr->n_overlap, 1ull);
    if (OBS) {
      const ObsBufs& O = B.obs;
      const double rn = norm3(pos[0], pos[1], pos[2]);
      O.s_time[s] += dt;
      if (rn < O.s_min[s]) {
        O.s_min[s] = rn;
        O.s_step[s] = (int)ns;
        for (int k = 0; k < 3; ++k) O.s_nm[7 * (size_t)s + k] = pos[k];
        for (int k = 0; k < 4; ++k) O.s_nm[7 * (size_t)s + 3 + k] = q[k];
      }
      const double rmax = 1.2 * P.r_esc;
      atomicAdd(&h[obs_bin(rn, rmax, kObsRad)], 1);
      const double rs = fmax(rn, 1e-30);
      const double th = acos(fmin(fmax(pos[2] / rs, -1.0), 1.0));
      const double ph = atan2(pos[1], pos[0]) + M_PI;
      int ti = (int)(th / M_PI * kObsTheta), pj = (int)(ph / (2.0 * M_PI) * kObsPhi);
      ti = ti < 0 ? 0 : (ti >= kObsTheta ? kObsTheta - 1 : ti);
      pj = pj < 0 ? 0 : (pj >= kObsPhi ? kObsPhi - 1 : pj);
      atomicAdd(&h[kObsRad + ti * kObsPhi + pj], 1);
      const double rp = O.s_prev_r[s];
      const int ms0 = kObsRad + kObsTheta * kObsPhi;
      for (int m = 0; m < kObsMs; ++m) {
        const double mr = P.r_b + (P.r_esc - P.r_b) * m / (kObsMs - 1);
        if (rp < mr && rn >= mr) atomicAdd(&h[ms0 + m], 1);
        if (rp >= mr && rn < mr) atomicAdd(&h[ms0 + kObsMs + m], 1);
      }
      const int tr0 = ms0 + 2 * kObsMs;
      atomicAdd(&h[tr0 + obs_bin(rp, rmax, kObsTrans) * kObsTrans + obs_bin(rn, rmax, kObsTrans)], 1);
      const int c0 = tr0 + kObsTrans * kObsTrans;
      // contacts at the pose the step started from (gpu_batch_simulator.py:1300-1311)
      for (int p = 0; p < P.n_pairs; ++p)
        if (x0[p * xs] < 0.0) atomicAdd(&h[c0 + p], 1);
      atomicAdd(&h[c0 + PST_MAX_PAIRS], 1);
      O.s_prev_r[s] = rn;
    }
  }
  const unsigned long long m = __ballot(any);
  // one flag for the whole launch: set it only while it is still clear (once
  // set, every later wave would serialise on the same address for nothing)
  if (m && __lane_id() == __ffsll((long long)m) - 1 &&
      __hip_atomic_load(&B.ctr->any_above[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    atomicOr(&B.ctr->any_above[par], 1);
  if (OBS) {
    __syncthreads();
    for (int i = threadIdx.x; i < kObsHist; i += blockDim.x)
      if (h[i]) atomicAdd(&B.obs.hist[i], (unsigned long long)h[i]);
  }
}

template <bool OBS>
__global__ void __launch_bounds__(kStepBlock) k_prod_post(StepParams P, ForceSys S, ProdBufs B,
                                                          int par) {
  prod_post_body<OBS>(P, S, B, par, blockIdx.x);
}

template <bool OBS>
__global__ void __launch_bounds__(kStepBlock) k_prod_post_group(const StepParams* Pg,
                                                                const ForceSys* Sg,
                                                                const ProdBufs* Bg,
                                                                const int* start, int L, int par) {
  const int l = slice_of(start, L, blockIdx.x);
  prod_post_body<OBS>(Pg[l], Sg[l], Bg[l], par, blockIdx.x - start[l]);
}

// Observables of the slots still running (after an early stop): their
// accumulators go into their trajectory records as they stand.
__global__ void k_prod_flush_obs(StepParams P, ProdBufs B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n_slots || B.status[s] != PST_RUNNING) return;
  obs_flush(P, B, s, 0);
}

hipError_t prod_flush_observables(const StepParams& P, const ProdBufs& B, hipStream_t st) {
  if (!B.obs.traj) return hipSuccess;
  hipLaunchKernelGGL(k_prod_flush_obs, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B);
  return hipGetLastError();
}

// Bridge decision, n_steps += 1, finish + refill of bridged trajectories.
__device__ inline void prod_bridge_body(const StepParams& P, const ProdBufs& B, int par, int blk) {
  // the slots of this step's lists, in slot order (as k_prod_post)
  const int s = blk * blockDim.x + threadIdx.x;
  RunCounters* C = B.ctr;
  if (s == 0) atomicAdd(&C->total_steps, (unsigned long long)(C->n_sr + C->n_mlp));
  if (s >= B.n_slots || B.key[s] == -1) return;
  const long long t = B.traj[s];
  const unsigned int ns = B.n_steps[s];
  bool react = false;
  if (P.n_pairs && C->any_above[par]) {
    double u[PST_MAX_PAIRS];
    for (int p = 0; p < P.n_pairs; p += 2) {
      double a, b;
      philox_uniform2(P.seed, rng_key(P, t), ns, RNG_BRIDGE, (uint32_t)(p >> 1), a, b);
      u[p] = a;
      if (p + 1 < PST_MAX_PAIRS) u[p + 1] = b;
    }
    // pair distances [pair][slot] (written by the post phase)
    const int xs = B.n_slots;
    const double* x0 = B.x0 + s;
    const double* x1 = B.x1 + s;
    if (P.sm_active) {
      // fired pairs gated by the current state, lowest index (:1185-1234)
      const int state = B.sm_state[s];
      for (int p = 0; p < P.n_pairs; ++p) {
        if (P.sm_before[p] != state) continue;
        const double a = x0[p * xs], b = x1[p * xs];
        double pc = 0.0;
        if (a > 0 && b > 0) {
          const double Ddt = fmax((P.has_bb_deff ? P.bb_deff[p] : P.D_t) * B.dt[s], 1e-30);
          pc = exp(-a * b / Ddt);
        }
        if (u[p] < pc || b < 0.0) {
          react = sm_fire(P, B, s, p, true);
          break;
        }
      }
    } else {
      const int n = bridge_count(P, x0, x1, xs, B.dt[s], u);
      react = n >= P.n_needed;
    }
  }
  B.n_steps[s] = ns + 1;
  if (react) {
    atomicAdd(&C->n_bridge, 1ull);
    finish_traj(P, B, s, PST_REACTED, 1);
  }
}

__global__ void k_prod_bridge(StepParams P, ProdBufs B, int par) {
  prod_bridge_body(P, B, par, blockIdx.x);
}

__global__ void k_prod_bridge_group(const StepParams* Pg, const ProdBufs* Bg, const int* start,
                                    int L, int par) {
  const int l = slice_of(start, L, blockIdx.x);
  prod_bridge_body(Pg[l], Bg[l], par, blockIdx.x - start[l]);
}

// Between steps: decide the early stop (:1347-1348) and reset the per-step
// counters for the next step.
__device__ inline void prod_stop_body(const StepParams& P, RunCounters* C, int par) {
  const unsigned long long done = C->n_reacted + C->n_escaped;
  if ((long long)done >= P.stop_done) C->stop = 1;
  C->n_sr = 0;
  C->n_mlp = 0;
  C->any_above[par] = 0;
}

__global__ void k_prod_stop(StepParams P, RunCounters* C, int par) { prod_stop_body(P, C, par); }

// one thread per system
__global__ void k_prod_stop_group(const StepParams* Pg, const ProdBufs* Bg, int L, int par) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < L) prod_stop_body(Pg[l], Bg[l].ctr, par);
}

hipError_t prod_step_pre(const StepParams& P, const ProdBufs& B, int par, hipStream_t st) {
  hipLaunchKernelGGL(k_prod_stop, dim3(1), dim3(1), 0, st, P, B.ctr, par);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (P.lmz_on) {
    hipLaunchKernelGGL(k_prod_outer, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                       dim3(kStepBlock), 0, st, P, B);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_prod_pre, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B, par);
  return hipGetLastError();
}

hipError_t prod_step_post(const StepParams& P, const ForceSys& S, const ProdBufs& B, int par,
                          hipStream_t st) {
  if (B.obs.traj)
    hipLaunchKernelGGL(k_prod_post<true>, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                       dim3(kStepBlock), 0, st, P, S, B, par);
  else
    hipLaunchKernelGGL(k_prod_post<false>, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                       dim3(kStepBlock), 0, st, P, S, B, par);
  return hipGetLastError();
}

hipError_t prod_step_bridge(const StepParams& P, const ProdBufs& B, int par, hipStream_t st) {
  hipLaunchKernelGGL(k_prod_bridge, dim3((B.n_slots + kStepBlock - 1) / kStepBlock),
                     dim3(kStepBlock), 0, st, P, B, par);
  return hipGetLastError();
}

int prod_step_block() { return kStepBlock; }

// The phases of one step for L systems at once (device arrays of their
// parameters and buffers; system l owns workgroups start[l] .. start[l+1]-1,
// ceil(n_slots / prod_step_block()) of them).  Each system's arithmetic is
// that of its own single-system launches.
hipError_t prod_group_pre(const StepParams* Pg, const ProdBufs* Bg, const int* start, int L,
                          int blocks, int par, bool outer, hipStream_t st) {
  hipLaunchKernelGGL(k_prod_stop_group, dim3((L + 63) / 64), dim3(64), 0, st, Pg, Bg, L, par);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (outer) {
    hipLaunchKernelGGL(k_prod_outer_group, dim3(blocks), dim3(kStepBlock), 0, st, Pg, Bg, start,
                       L);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_prod_pre_group, dim3(blocks), dim3(kStepBlock), 0, st, Pg, Bg, start, L, par);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bins_scan_group, dim3(L), dim3(256), 0, st, Bg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bins_scatter_group, dim3(blocks), dim3(kStepBlock), 0, st, Bg, start, L);
  return hipGetLastError();
}

hipError_t prod_group_post(const StepParams* Pg, const ForceSys* Sg, const ProdBufs* Bg,
                           const int* start, int L, int blocks, bool obs, int par,
                           hipStream_t st) {
  if (obs)
    hipLaunchKernelGGL(k_prod_post_group<true>, dim3(blocks), dim3(kStepBlock), 0, st, Pg, Sg, Bg,
                       start, L, par);
  else
    hipLaunchKernelGGL(k_prod_post_group<false>, dim3(blocks), dim3(kStepBlock), 0, st, Pg, Sg,
                       Bg, start, L, par);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prod_bridge_group, dim3(blocks), dim3(kStepBlock), 0, st, Pg, Bg, start, L,
                     par);
  return hipGetLastError();
}

}  // namespace pst
