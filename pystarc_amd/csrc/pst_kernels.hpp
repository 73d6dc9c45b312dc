// Kernel launchers shared between the translation units of libpystarc_hip.
#pragma once
#include <hip/hip_runtime.h>
#include "pst_types.hpp"

namespace pst {

hipError_t launch_forces(const ForceSys& S, const ForceWork& W, int precision, int G, int max_n,
                         int blocks_cap, hipStream_t st, bool zp = false);
// Split evaluation over a device-counted pose-mode list (G in {16, 32, 64};
// W.rev_order = the same trajectories ordered for the reverse-Born pass):
// the ligand-atom and reverse-Born items in one launch (k_forces_split,
// corrections into W.F_rev / W.T_rev), then F -= F_rev, T -= T_rev
// (k_rev_apply).  Bit-identical to launch_forces with the same G.
hipError_t launch_forces_split(const ForceSys& S, const ForceWork& W, int precision, int G,
                               int max_n, int blocks_cap, hipStream_t st, bool zp);
hipError_t launch_chunk_min_r(const double* pos, const int* list, int n, int fb, int cs, int cpf,
                              double* out, int n_chunks, hipStream_t st);
int pick_group(int M, int n_rec_active);
// One launch over L systems (device arrays of ForceSys / ForceWork, pose
// mode); workgroups [start[i], start[i+1]) evaluate system order[i] (system
// i when order is null); xcd_major: start[] indexes workgroups XCD by XCD
// (blocks a multiple of 8); zp: z-pair / z-row fetches (the same node values).
hipError_t launch_forces_group(const ForceSys* Sg, const ForceWork* Wg, const int* start, int L,
                               int blocks, size_t lds, int precision, int G, bool zp,
                               hipStream_t st, bool xcd_major = false,
                               const int* order = nullptr);
// Gathers of one force launch over W's list (pose source): out4 += {trajectories,
// elec-grid atoms, born-grid atoms, reverse-grid atoms}.
hipError_t launch_census(const ForceSys& S, const ForceWork& W, unsigned long long* out4,
                         hipStream_t st);
// Empty marker kernel bracketing a timing window in a kernel trace.
hipError_t launch_window_mark(hipStream_t st);
// Gather ceiling: n points, the FP32 pair-grid stencil at each (pst_probe.hip).
hipError_t launch_gather_probe(const GridDev& g, long long n, int mode, int variant,
                               unsigned long long seed, float* out, int blocks, hipStream_t st);

// Brick-binned force evaluation: request counts / order per brick of the
// receptor float2 pair grid and of the ligand Born grid, and per-request results.
struct BinBufs {
  int* hist;            // [nb_rec + nb_lig] request counts (reset by the scan)
  int* start;           // [nb_rec + nb_lig + 1] first request of each brick
  int* cursor;          // [nb_rec + nb_lig] scatter cursors
  int2* items;          // [max_items] (brick, offset) work items of the gather kernel
  int* n_items;         // [1]
  int* sorted;          // request ids in brick order: w*M + i, rev_base + w*n_rec + j
  void* out;            // [max_traj * max_m] per-atom {F, E} (double4 / float4)
  void* out_rev;        // [max_traj * max_rec] per-receptor-atom reverse force
  int nb_rec, nb_lig;   // bricks of each grid
  int rb[3], lb[3];     // bricks per axis
  int rev_base;         // max_traj * max_m
  int max_items, max_traj, max_m, max_rec;
  int gather_blocks;    // multiple of 8
};
bool binned_eligible(const ForceSys& S, const ForceWork& W, const BinBufs& B, int max_n);
hipError_t launch_forces_binned(const ForceSys& S, const ForceWork& W, const BinBufs& B,
                                int precision, int G, int max_n, int blocks_cap, hipStream_t st);
int bricks_per_axis(int n_nodes);
int bin_items_bound(long max_req, int n_bins);

// Ordered (stable) compaction of in_list (or 0..n-1) by flag[slot].
hipError_t launch_compact_ordered(const int* in_list, int n_in, const int* n_in_dev,
                                  const unsigned char* flag, int* out, int* n_out, hipStream_t st);

// Replay-mode step phases (host-supplied noise, reference draw order).
struct ReplayBufs {
  double* pos; double* quat; int* status; long long* n_steps;
  unsigned char* f_run; unsigned char* f_esc; unsigned char* f_ret; unsigned char* f_sr;
  int* run_list; int* esc_list; int* ret_list; int* sr_list;
  int* counts;           // [0]=n_run [1]=n_esc [2]=n_ret [3]=n_sr [4]=any_above [5]=n_bridge [6]=n_overlap [7]=n_react_start
  double* F; double* T; double* dt; double* x0; double* x1;
  double* old_pos; double* old_quat;
  int n_slots;
};
hipError_t replay_begin(const StepParams& P, const ReplayBufs& B, hipStream_t st);
hipError_t replay_escape_decide(const StepParams& P, const ReplayBufs& B, const double* u_outer,
                                hipStream_t st);
hipError_t replay_escape_apply(const StepParams& P, const ReplayBufs& B, const double* u_quat,
                               hipStream_t st);
hipError_t replay_build_sr(const ReplayBufs& B, hipStream_t st);
hipError_t replay_move(const StepParams& P, const ForceSys& S, const ReplayBufs& B,
                       const double* noise_t, const double* noise_r, int n_sr, hipStream_t st);
hipError_t replay_bridge(const StepParams& P, const ReplayBufs& B, const double* u_bb, int n_sr,
                         hipStream_t st);

// Per-trajectory and run-wide observables of the production step (the
// reference's sim_data, gpu_batch_simulator.py:576-620, :1255-1340).
constexpr int kObsRad = 200;     // radial histogram bins over [0, 1.2 r_esc)
constexpr int kObsTheta = 36;    // angular occupancy (theta, phi)
constexpr int kObsPhi = 72;
constexpr int kObsMs = 11;       // milestone radii linspace(b, r_esc, 11)
constexpr int kObsTrans = 50;    // radial transition matrix bins over [0, 1.2 r_esc)
constexpr int kObsTraj = 27;     // doubles per trajectory record
// histogram block: rad | ang | ms_out | ms_in | trans | contact[16] | contact_steps
constexpr int kObsHist = kObsRad + kObsTheta * kObsPhi + 2 * kObsMs + kObsTrans * kObsTrans +
                         PST_MAX_PAIRS + 1;
struct ObsBufs {
  // per slot (accumulated while the slot holds a trajectory)
  double* s_min; int* s_step; double* s_time; int* s_ret; double* s_prev_r; double* s_nm;  // s_nm (S,7)
  // per trajectory id: (N, kObsTraj) = start pos 0-2, start q 3-6, min dist 7, step at min 8,
  // near-miss pos 9-11, near-miss q 12-15, total time 16, returns 17, bridge 18,
  // end pos 19-21, end q 22-25, pairs within cutoff at the end 26
  double* traj;
  unsigned long long* hist;   // [kObsHist]
};

// Production step (device Philox, refilling slots).
struct ProdBufs {
  double* pos; double* quat; int* status; unsigned int* n_steps; long long* traj;
  int* plain_list; int* mlp_list;
  int* rev_list;               // the grid-path list ordered for the reverse-Born pass
  // per-slot step role: -1 not stepped, -2 surrogate list, else the grid-path
  // direction key (bits 0-7) and the reverse-pass cell (bits 8-15)
  int* key;
  int* bins;                   // [256] key histogram, [256..511] scatter cursors,
                               // [512..767] / [768..1023] the same for the reverse cells
  double* F; double* T; double* dt;
  double* x0; double* x1;      // pair distance minus cutoff before / after the move,
                               // [PST_MAX_PAIRS][n_slots] (pair-major)
  RunCounters* ctr;
  signed char* out_status;     // per trajectory id (nullable)
  int* out_steps;              // per trajectory id (nullable)
  ObsBufs obs;                 // observables (obs.traj null: off)
  int* sm_state;               // per-slot reaction state (state-machine mode)
  unsigned int* sm_mask;       // per-slot reactions fired so far (bit r)
  int* lmz_escaped;            // per slot: 1 when this step's outer excursion escaped
  int n_slots;
};
hipError_t prod_flush_observables(const StepParams& P, const ProdBufs& B, hipStream_t st);
hipError_t prod_init(const StepParams& P, const ProdBufs& B, hipStream_t st);
hipError_t prod_step_pre(const StepParams& P, const ProdBufs& B, int par, hipStream_t st);
hipError_t prod_step_sort(const ProdBufs& B, hipStream_t st);
hipError_t prod_step_post(const StepParams& P, const ForceSys& S, const ProdBufs& B, int par,
                          hipStream_t st);
hipError_t prod_step_bridge(const StepParams& P, const ProdBufs& B, int par, hipStream_t st);
int prod_step_block();
hipError_t prod_group_pre(const StepParams* Pg, const ProdBufs* Bg, const int* start, int L,
                          int blocks, int par, bool outer, hipStream_t st);
hipError_t prod_group_post(const StepParams* Pg, const ForceSys* Sg, const ProdBufs* Bg,
                           const int* start, int L, int blocks, bool obs, int par,
                           hipStream_t st);

// Two-body RPY: 12x12 mobility / noise factor columns, and the move itself.
hipError_t launch_rpy_eval(int n, const double* pos, double a0, double a1, double* mob,
                           double* fac, hipStream_t st);
hipError_t launch_rpy_move(int n, double a0, double a1, double Df, int torque_mode,
                           const double* dt, double* pos, double* quat, const double* F,
                           const double* T, const double* xi, double* dX_out, hipStream_t st);

// Interleave two FP32 grids of identical geometry into float2 nodes.
hipError_t launch_interleave(const float* a, const float* b, float2* out, long n, hipStream_t st);
hipError_t launch_to_bricks(const void* src, void* dst, int elem_bytes, int nx, int ny, int nz,
                            hipStream_t st);
// Nodes of a grid stored in 4x4x4 bricks, padding included.
inline long brick_nodes(int nx, int ny, int nz) {
  return (long)((nx + 3) >> 2) * ((ny + 3) >> 2) * ((nz + 3) >> 2) * 64;
}

// Synthetic grids.
hipError_t launch_grid_dh(float* out, double* out64, int nx, int ny, int nz, const double* origin,
                          const double* sp, const double* atoms, int n_atoms, double kappa,
                          double eps_out, double eps_in, int kind, hipStream_t st);

}  // namespace pst
