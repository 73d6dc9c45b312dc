// Surrogate branch of the production step.  Trajectories whose centre
// distance exceeds r_on are listed by k_prod_pre in the surrogate list; for
// them the MLP replaces the grid + Born + far-field force evaluation.
//
// Features (9, FP32): centroid in the receptor frame divided by b, and the
// first two columns of R(q) (a 6-D rotation representation).  Outputs (13):
// ligand force and torque, receptor force and torque, energy, each divided
// by a per-group scale and by the envelope
//   env(r) = exp(-(r - r0)/lam) (r0/r)^2 (1 + r/lam) / (1 + r0/lam),
// the shape of a screened-Coulomb force normalised to 1 at r0, so the
// network learns O(1) targets over the whole shell.  An audit evaluates the
// plain grid path on `audit_n` surrogate trajectories of every step (a stride
// through the list from a moving offset) and accumulates the squared force
// and torque errors and the squared reference force and torque.
#include <hip/hip_runtime.h>
#include "pst_device.hpp"
#include "pst_kernels.hpp"
#include "pst_mlp.hpp"
#include "pst_surrogate.hpp"

namespace pst {

__device__ inline double envelope(const SurrParams& S, double r) {
  const double x = S.r0 / r;
  return exp(-(r - S.r0) / S.lam) * x * x * (1.0 + r / S.lam) / (1.0 + S.r0 / S.lam);
}

// FP32 features of list entry w into row w of x0f (the first layer runs in FP32).
__device__ inline void features_row(const SurrParams& S, const double* pos, const double* quat,
                                    const int* list, float* x0f, int w) {
  const int s = list[w];
  double R[9];
  quat_to_rot(quat + 4 * (size_t)s, R);
  float* row = x0f + (size_t)w * kMlpInLd;
  row[0] = (float)(pos[3 * s] * S.inv_b);
  row[1] = (float)(pos[3 * s + 1] * S.inv_b);
  row[2] = (float)(pos[3 * s + 2] * S.inv_b);
  row[3] = (float)R[0]; row[4] = (float)R[3]; row[5] = (float)R[6];
  row[6] = (float)R[1]; row[7] = (float)R[4]; row[8] = (float)R[7];
}

// Force (and torque) of list entry w from row w of the network's output.
__device__ inline void apply_row(const SurrParams& S, const double* pos, const int* list,
                                 const float* y, double* F, double* T, int w) {
  const int s = list[w];
  const double r = norm3(pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]);
  const double env = envelope(S, r);
  const float* o = y + (size_t)w * 64;
  const double fs = S.scale_F * env, ts = S.scale_T * env;
  F[3 * s] = o[0] * fs; F[3 * s + 1] = o[1] * fs; F[3 * s + 2] = o[2] * fs;
  if (S.use_torque) {
    T[3 * s] = o[3] * ts; T[3 * s + 1] = o[4] * ts; T[3 * s + 2] = o[5] * ts;
  } else {
    T[3 * s] = 0.0; T[3 * s + 1] = 0.0; T[3 * s + 2] = 0.0;
  }
}

__global__ void k_surr_features(SurrParams S, const double* pos, const double* quat,
                                 const int* list, const int* n_dev, float* x0f) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= *n_dev) return;
  features_row(S, pos, quat, list, x0f, w);
}

__global__ void k_surr_apply(SurrParams S, const double* pos, const int* list, const int* n_dev,
                             const float* y, double* F, double* T) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= *n_dev) return;
  apply_row(S, pos, list, y, F, T, w);
}

// Audit sample of one step: k = min(n, cap) rows of the surrogate list at a
// stride of n / k from an offset that a hash of the run's trajectory-step
// counter moves every step (stable within the step: only the pre and post
// phases change it), so the audit sees every part of the list over a run.
// Entry i of the sample (i == 0 also stores k).
__device__ inline void audit_pick_entry(const int* list, int n, int cap, const RunCounters* C,
                                        int* out, int* n_out, int i) {
  const int k = min(n, cap);
  if (i == 0) *n_out = k;
  if (i >= k) return;
  unsigned long long h = (C->total_steps + 1) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 31;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 27;
  const long long off = (long long)(h % (unsigned long long)n);
  const long long stride = n / k;
  out[i] = list[(off + (long long)i * stride) % n];
}

__global__ void k_surr_audit_pick(const int* list, const int* n_dev, int cap,
                                  const RunCounters* C, int* out, int* n_out) {
  audit_pick_entry(list, *n_dev, cap, C, out, n_out, blockIdx.x * blockDim.x + threadIdx.x);
}

// Squared errors of the surrogate's forces (and, with Tref, torques) against
// the grid path on audited rows w = w0, w0 + dw, ... < n, summed into the run
// counters (one atomic per wave and sum).
__device__ inline void audit_rows(const int* list, int n, int w0, int dw, const double* F,
                                  const double* T, const double* Fref, const double* Tref,
                                  RunCounters* C) {
  double e2 = 0, r2 = 0, te2 = 0, tr2 = 0;
  for (int w = w0; w < n; w += dw) {
    const int s = list[w];
    for (int d = 0; d < 3; ++d) {
      const double a = F[3 * s + d], b = Fref[3 * w + d];
      e2 += (a - b) * (a - b);
      r2 += b * b;
    }
    if (Tref) {
      for (int d = 0; d < 3; ++d) {
        const double a = T[3 * s + d], b = Tref[3 * w + d];
        te2 += (a - b) * (a - b);
        tr2 += b * b;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    e2 += __shfl_xor(e2, off);
    r2 += __shfl_xor(r2, off);
    te2 += __shfl_xor(te2, off);
    tr2 += __shfl_xor(tr2, off);
  }
  if (__lane_id() == 0 && (e2 > 0 || r2 > 0)) {
    atomicAdd(&C->audit_err2, e2);
    atomicAdd(&C->audit_ref2, r2);
  }
  if (__lane_id() == 0 && (te2 > 0 || tr2 > 0)) {
    atomicAdd(&C->audit_terr2, te2);
    atomicAdd(&C->audit_tref2, tr2);
  }
  if (w0 == 0) atomicAdd(&C->audit_n, (unsigned long long)n);
}

__global__ void k_surr_audit(const int* list, const int* n_dev, int cap, const double* F,
                             const double* T, const double* Fref, const double* Tref,
                             RunCounters* C) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  audit_rows(list, min(*n_dev, cap), w, 1 << 30, F, T, Fref, Tref, C);
}

// Grouped forms: task l of L for workgroup b (start[l] <= b < start[l + 1]).
__device__ inline int task_of(const int* start, int L, int b) {
  int l = 0;
  while (l + 1 < L && b >= start[l + 1]) ++l;
  return l;
}

__global__ void k_surr_group_features(const SurrTask* __restrict__ Tk, const int* start, int L) {
  const int l = task_of(start, L, blockIdx.x);
  const SurrTask& t = Tk[l];
  const int w = (blockIdx.x - start[l]) * blockDim.x + threadIdx.x;
  if (w >= *t.n_dev) return;
  features_row(t.sp, t.pos, t.quat, t.list, t.x0f, w);
}

__global__ void k_surr_group_apply(const SurrTask* __restrict__ Tk, const int* start, int L) {
  const int l = task_of(start, L, blockIdx.x);
  const SurrTask& t = Tk[l];
  const int w = (blockIdx.x - start[l]) * blockDim.x + threadIdx.x;
  if (w >= *t.n_dev) return;
  apply_row(t.sp, t.pos, t.list, t.y, t.F, t.T, w);
}

// one workgroup per task
__global__ void k_surr_group_audit_pick(const SurrTask* __restrict__ Tk) {
  const SurrTask& t = Tk[blockIdx.x];
  const int n = *t.n_dev;
  for (int i = threadIdx.x; i < max(t.sp.audit_n, 1); i += blockDim.x)
    audit_pick_entry(t.list, n, t.sp.audit_n, t.ctr, t.audit_list, t.audit_count, i);
}

__global__ void k_surr_group_audit(const SurrTask* __restrict__ Tk) {
  const SurrTask& t = Tk[blockIdx.x];
  if (t.sp.audit_n <= 0) return;
  audit_rows(t.audit_list, min(*t.n_dev, t.sp.audit_n), threadIdx.x, blockDim.x, t.F, t.T,
             t.audit_F, t.audit_T, t.ctr);
}

hipError_t surr_features(const SurrParams& S, const MlpBufs& m, const double* pos,
                         const double* quat, const int* list, const int* n_dev, int n_slots,
                         hipStream_t st) {
  const int blocks = (n_slots + 255) / 256;
  hipLaunchKernelGGL(k_surr_features, dim3(blocks), dim3(256), 0, st, S, pos, quat, list, n_dev,
                     m.x0f);
  return hipGetLastError();
}

hipError_t surr_apply(const SurrParams& S, const MlpBufs& m, const double* pos, const int* list,
                      const int* n_dev, int n_slots, double* F, double* T, hipStream_t st) {
  const int blocks = (n_slots + 255) / 256;
  hipLaunchKernelGGL(k_surr_apply, dim3(blocks), dim3(256), 0, st, S, pos, list, n_dev, m.y, F, T);
  return hipGetLastError();
}

hipError_t surr_step(const SurrParams& S, const MlpBufs& m, const double* pos, const double* quat,
                     const int* list, const int* n_dev, int n_slots, double* F, double* T,
                     hipStream_t st) {
  hipError_t e = surr_features(S, m, pos, quat, list, n_dev, n_slots, st);
  if (e != hipSuccess) return e;
  const int Sp = (n_slots + 63) / 64 * 64;
  e = mlp_forward_packed(m, Sp, n_dev, false, m.y, st);
  if (e != hipSuccess) return e;
  return surr_apply(S, m, pos, list, n_dev, n_slots, F, T, st);
}

hipError_t surr_audit_pick(const int* list, const int* n_dev, int cap, const RunCounters* C,
                           int* out, int* n_out, hipStream_t st) {
  hipLaunchKernelGGL(k_surr_audit_pick, dim3((cap + 255) / 256), dim3(256), 0, st, list, n_dev,
                     cap, C, out, n_out);
  return hipGetLastError();
}

hipError_t surr_audit(const int* list, const int* n_dev, int cap, const double* F, const double* T,
                      const double* Fref, const double* Tref, RunCounters* C, hipStream_t st) {
  hipLaunchKernelGGL(k_surr_audit, dim3((cap + 255) / 256), dim3(256), 0, st, list, n_dev, cap, F,
                     T, Fref, Tref, C);
  return hipGetLastError();
}

hipError_t surr_group_features(const SurrTask* T, const int* start, int L, int blocks,
                               hipStream_t st) {
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_surr_group_features, dim3(blocks), dim3(256), 0, st, T, start, L);
  return hipGetLastError();
}

hipError_t surr_group_apply(const SurrTask* T, const int* start, int L, int blocks, hipStream_t st) {
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_surr_group_apply, dim3(blocks), dim3(256), 0, st, T, start, L);
  return hipGetLastError();
}

hipError_t surr_group_audit_pick(const SurrTask* T, int L, hipStream_t st) {
  if (L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_surr_group_audit_pick, dim3(L), dim3(256), 0, st, T);
  return hipGetLastError();
}

hipError_t surr_group_audit(const SurrTask* T, int L, hipStream_t st) {
  if (L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_surr_group_audit, dim3(L), dim3(256), 0, st, T);
  return hipGetLastError();
}

}  // namespace pst
