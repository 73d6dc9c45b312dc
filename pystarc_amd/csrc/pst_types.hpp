// Plain-old-data structures shared by the host C-ABI layer and the gfx950
// kernels.  Everything here is passed to kernels by value.
#pragma once
#include <stdint.h>

#define PST_MAX_GRIDS 4
#define PST_MAX_PAIRS 16
#define PST_REC_CHUNK 64

// Trajectory status codes (pystarc/simulation/gpu_batch_simulator.py:514).
enum { PST_RUNNING = 0, PST_REACTED = 1, PST_ESCAPED = 2, PST_MAXSTEPS = 3, PST_IDLE = 4 };

// One APBS-style grid (node (i, j, k) of an nx x ny x nz grid, DX order
// (i*ny + j)*nz + k on upload, pystarc/forces/electrostatic/grid_force.py:144),
// held on the device in bricks of 4x4x4 nodes (see GlobalAcc).  lo/hi are
// the 3-spacing BC-safe box of gpu_batch_engine.py:267-273.
struct GridDev {
  const void* data;
  int dtype;  // 0 = float32, 1 = float64
  int nx, ny, nz;
  double o[3];
  double sp[3];
  double inv[3];
  double lo[3];
  double hi[3];
};

// Everything the force kernels need about the system.
struct ForceSys {
  GridDev elec[PST_MAX_GRIDS];
  GridDev born[PST_MAX_GRIDS];
  GridDev lig_born;
  // one elec and one Born grid of identical geometry, interleaved float2
  // (potential, Born) per node; the kernels then fetch both with one stencil
  GridDev pair;
  int paired;
  int n_elec, n_born, has_rev;
  double alpha;
  // Yukawa multipole far-field (gpu_batch_engine.py:161-175, :393-472)
  int has_yukawa, has_dip, has_quad;
  double V_fac, debye, fpe;
  double dip[3];
  double quad[9];
  // the same constants pre-divided on the host: 1/debye, 1/debye^2,
  // 1/(3 debye^2), 1/(3 debye^3), dipole/fpe, quadrupole/fpe
  double inv_lam, inv_lam2, c3, c3l;
  double dip_s[3];
  double quad_s[9];
  // whole-ligand cull: no ligand atom can be inside any receptor grid box when
  // |c| - lig_rmax > grid_reach
  double lig_rmax, grid_reach;
  // ligand, body frame about the centroid
  int M;
  const double* lig_body;   // (M,3)
  const double* lig_q;      // (M,)
  const float* lig_f4;      // (M,4) x,y,z,q in float
  double cq_body[3];        // centroid of charged atoms, body frame
  // receptor atoms for reverse Born (and overlap)
  int n_rec;
  const double* rec_pos;    // (Nr,3)
  const double* rec_q;      // (Nr,)
  const float* rec_f4;      // (Nr,4)
  double rec_max_r;         // max |r_rec|
  // receptor atoms are stored in spatially compact chunks of PST_REC_CHUNK
  // consecutive atoms; chunk k covers atoms [k*PST_REC_CHUNK, (k+1)*PST_REC_CHUNK)
  // and rec_chunk[4k..4k+3] = (centre x, y, z, radius) bounds all of them
  int n_rec_chunk;
  const double* rec_chunk;  // (n_rec_chunk,4)
  // the same atoms binned on a uniform grid of cubic cells (edge 1/cell_inv)
  // for short-range queries: cell (i,j,k) holds cell_pos rows
  // cell_start[c] .. cell_start[c+1]-1, c = (i*cell_n[1] + j)*cell_n[2] + k
  int cell_n[3];
  double cell_o[3];
  double cell_inv;
  const int* cell_start;    // (n_cells+1,)
  const double* cell_pos;   // (Nr,3)
  double rev_radius;        // max(|lo_x|,|hi_x|) of the ligand Born grid (:623)
  double rev_cull;          // exact cull: largest distance from the ligand centre to the valid box
  int torque_mode;          // 0 = reference (zero), 1 = lever arm about charged centroid
};

// Work description of one force launch.
struct ForceWork {
  const int* list;          // trajectory indices, or null for 0..n-1
  const int* n_dev;         // device count, or null
  int n;                    // host count (used when n_dev is null)
  // pose source: pose mode reads pos/quat; positions mode reads lig_pos (N,M,3)
  const double* pos;        // (N,3) centroids
  const double* quat;       // (N,4) or null
  const double* rot;        // (N,9) or null
  const double* lig_pos;    // (N,M,3) or null
  const double* lig_q;      // (M,) charges for positions mode (else sys.lig_q)
  double* F;                // (N,3)
  double* T;                // (N,3)
  double* E;                // (N,)
  // reverse-Born skip: 0 = exact geometric cull per trajectory,
  // 1 = the reference's chunked skip using chunk_min_r
  int rev_skip_mode;
  const double* chunk_min_r;
  int chunk_fb;             // force-batch size
  int chunk_cs;             // chunk size within a force batch
  int chunks_per_fb;
  int out_compact;          // 1: write outputs at list position, 0: at trajectory index
  int n_cap;                // > 0: process at most n_cap list entries
  int xcd_chunks;           // 1: XCD x (= blockIdx % 8) takes the x-th eighth of the list
  // split evaluation: room for the reverse pass's order (list capacity) and
  // the reverse-Born corrections per trajectory (indexed like F / T)
  int* rev_order;
  double* F_rev;
  double* T_rev;
};

// Step parameters (gpu_batch_simulator.py:480-781).
struct StepParams {
  int64_t n_total;       // trajectories in the run
  int64_t n_slots;       // concurrent slots
  uint64_t seed;
  int64_t traj_base;     // random-stream id of the run's trajectory 0 (a shard's first id)
  double r_b, r_esc, return_prob;
  double D_t, D_r;
  int use_hi;
  double hi_Df, hi_ainv, hi_a2;
  double adt_debye, adt_V0;
  int adt_has_force;
  double min_core_dt, max_dt;
  int use_overlap;
  double overlap_thresh;
  int n_pairs, n_needed, has_bb_deff;
  double rec_gho[PST_MAX_PAIRS * 3];
  double lig_gho[PST_MAX_PAIRS * 3];
  double cutoff[PST_MAX_PAIRS];
  double bb_deff[PST_MAX_PAIRS];
  int64_t max_steps;
  int hydro_mode;        // 0 = scalar D (reference), 1 = two-body 12x12 RPY
  double a_rec, a_lig, kT_over_eta;
  double surrogate_r;    // centre distance above which the surrogate is used (<=0: off)
  int64_t stop_done;     // early stop when done >= this (int(0.995 N))
  // state-machine reactions: reaction r is pair r (its pair0), fires only from
  // state sm_before[r], moves the trajectory to sm_after[r]; a state whose bit
  // is set in sm_terminal ends the trajectory as reacted
  int sm_active;
  int sm_before[PST_MAX_PAIRS];
  int sm_after[PST_MAX_PAIRS];
  int sm_terminal;       // bit mask over state indices (< 32 states)
  double min_core_rxn_dt;
  // LMZ outer propagator (outer_propagator.py:259-401): a trajectory at
  // r >= lmz_trigger is carried through b < r < q analytically, returning
  // to the b-sphere or escaping (instead of the return probability at r_esc)
  int lmz_on, lmz_hi;
  double lmz_trigger, lmz_b, lmz_q, lmz_b_cover, lmz_q_cover;
  double lmz_V, lmz_L;          // U/kT = lmz_V e^{-r/lmz_L} / r
  double lmz_return_q;          // rate(b) / rate(q)
  double lmz_Dt, lmz_dpre, lmz_ainv, lmz_a2;
  double lmz_drot0, lmz_drot1;  // rotational diffusion of receptor and ligand
};

// Device-side run counters (production mode).
struct RunCounters {
  unsigned long long n_reacted;
  unsigned long long n_escaped;
  unsigned long long n_maxsteps;
  unsigned long long total_steps;
  unsigned long long next_traj;
  unsigned long long n_returns;
  unsigned long long n_overlap;
  unsigned long long n_bridge;
  unsigned long long audit_n;
  int n_sr;
  int n_mlp;
  int any_above[2];
  int stop;
  int pad;
  double audit_err2;
  double audit_ref2;
  // state-machine fire counts per reaction: at the reaction check, in the
  // Brownian bridge, and distinct trajectories that fired it
  unsigned long long sm_fired[PST_MAX_PAIRS];
  unsigned long long sm_fired_bridge[PST_MAX_PAIRS];
  unsigned long long sm_unique[PST_MAX_PAIRS];
  // surrogate audit, torques (physical torque mode): squared error and
  // squared grid-path torque summed over the audited trajectories
  double audit_terr2;
  double audit_tref2;
};
