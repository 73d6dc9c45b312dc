// Surrogate branch of the production step (see pst_surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "pst_mlp.hpp"
#include "pst_types.hpp"

namespace pst {

struct SurrParams {
  double r_on;      // centre distance above which the surrogate replaces the grid path
  double r0;        // envelope reference distance
  double lam;       // envelope screening length
  double inv_b;     // feature scale 1/b
  double scale_F;   // force scale (kBT/A) at r0
  double scale_T;   // torque scale
  double scale_E;   // energy scale
  int use_torque;   // apply the predicted ligand torque (physical torque mode)
  int audit_n;      // surrogate trajectories re-evaluated on the grid path per step
};

// The step's surrogate forces in three parts (features, then the network's
// forward, then forces / torques applied to the surrogate list's slots).
hipError_t surr_features(const SurrParams& S, const MlpBufs& m, const double* pos,
                         const double* quat, const int* list, const int* n_dev, int n_slots,
                         hipStream_t st);
hipError_t surr_apply(const SurrParams& S, const MlpBufs& m, const double* pos, const int* list,
                      const int* n_dev, int n_slots, double* F, double* T, hipStream_t st);
hipError_t surr_step(const SurrParams& S, const MlpBufs& m, const double* pos, const double* quat,
                     const int* list, const int* n_dev, int n_slots, double* F, double* T,
                     hipStream_t st);
hipError_t surr_audit_pick(const int* list, const int* n_dev, int cap, const RunCounters* C,
                           int* out, int* n_out, hipStream_t st);
hipError_t surr_audit(const int* list, const int* n_dev, int cap, const double* F, const double* T,
                      const double* Fref, const double* Tref, RunCounters* C, hipStream_t st);

// One surrogate-carrying system of a grouped step (an element of a device
// array): the same buffers the single-system calls above take.
struct SurrTask {
  SurrParams sp;
  const double* pos;
  const double* quat;
  const int* list;          // surrogate list and its length
  const int* n_dev;
  float* x0f;               // network input rows
  const float* y;           // network output rows
  double* F;
  double* T;
  RunCounters* ctr;
  int* audit_list;          // audit sample (slots) and its length
  int* audit_count;
  const double* audit_F;    // grid-path forces / torques of the sample
  const double* audit_T;    // null unless the torques are audited
};
// Features and applied forces of every task, one launch each: task l owns
// workgroups start[l] .. start[l + 1] - 1 (256 threads) of `blocks`.
hipError_t surr_group_features(const SurrTask* T, const int* start, int L, int blocks,
                               hipStream_t st);
hipError_t surr_group_apply(const SurrTask* T, const int* start, int L, int blocks, hipStream_t st);
// Audit sample and audit of every task, one workgroup per task.
hipError_t surr_group_audit_pick(const SurrTask* T, int L, hipStream_t st);
hipError_t surr_group_audit(const SurrTask* T, int L, hipStream_t st);

}  // namespace pst
