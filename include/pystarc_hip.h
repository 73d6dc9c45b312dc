/*
 * pystarc_hip.h - C ABI of libpystarc_hip.so, the MI355X (gfx950) kernels
 * behind the pystarc batch Brownian-dynamics force operator and step.
 *
 * Plain C types only: device buffers are raw pointers to device memory
 * (e.g. torch.Tensor.data_ptr() of a CUDA/HIP tensor), host buffers are raw
 * host pointers, sizes are ints.  Every entry point returns 0 on success and a
 * non-zero code on failure; pst_last_error() then holds the message.  One
 * context belongs to one host thread; all work of a context is enqueued on
 * its stream.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the PySTARC source tree).
 */
#ifndef PYSTARC_HIP_H
#define PYSTARC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pst_ctx pst_ctx;
typedef struct pst_sim pst_sim;

/* Grid kinds for pst_add_grid. */
#define PST_GRID_REC_ELEC 0 /* receptor potential grids  (receptor[0-9].dx)      */
#define PST_GRID_REC_BORN 1 /* receptor Born grids       (receptor[0-9]_born.dx) */
#define PST_GRID_LIG_BORN 2 /* ligand Born grids         (ligand[0-9]_born.dx)   */

/* Data types. */
#define PST_F32 0
#define PST_F64 1

/* ---- context -------------------------------------------------------------
 * Replaces the CuPy device set-up of GPUBatchForceEngine.__init__
 * (pystarc/forces/gpu_batch_engine.py:155-160).  With use_stream != 0 all
 * work goes to `stream` (NULL = the legacy default stream, which orders the
 * library with a caller such as torch); otherwise the library creates its own
 * non-blocking stream at default priority.  */
int pst_ctx_create(pst_ctx** out, int device, void* stream, int use_stream);
int pst_ctx_destroy(pst_ctx* ctx);
const char* pst_last_error(const pst_ctx* ctx);
int pst_abi_version(void);
/* sizeof of the structs a binding mirrors: 0 = StepParams, 1 = RunCounters
 * (no reference counterpart; lets a binding verify its struct layout). */
int pst_struct_size(int which);

/* ---- system set-up -------------------------------------------------------
 * Replaces GPUBatchForceEngine.__init__ / _upload_grids
 * (gpu_batch_engine.py:139-246, :248-287).  Grids are C-order (x slowest),
 * copied from host memory once and kept resident.  The caller passes the grids
 * already selected as the constructor selects them (finest first, :186-211). */
int pst_add_grid(pst_ctx* ctx, int kind, const void* host_data, int dtype, int nx, int ny, int nz,
                 const double* origin3, const double* spacing3);
int pst_clear_grids(pst_ctx* ctx);
/* dst uses src's receptor potential / Born grids (and their float2 pair)
 * without a copy: one resident set for many ligand contexts.  src must
 * outlive dst.  No reference counterpart (the reference builds one engine per
 * run, gpu_batch_engine.py:139-246). */
int pst_share_receptor_grids(pst_ctx* dst, const pst_ctx* src);
/* Yukawa far-field (gpu_batch_engine.py:161-175): dipole3 / quad9 may be NULL
 * (no multipole expansion, monopole only). */
int pst_set_farfield(pst_ctx* ctx, double receptor_charge, double debye_length, double sdie,
                     const double* dipole3, const double* quad9);
/* Born desolvation alpha (gpu_batch_engine.py:143). */
int pst_set_alpha(pst_ctx* ctx, double alpha);
/* Receptor atoms for reverse Born and overlap (gpu_batch_engine.py:207-217,
 * gpu_batch_simulator.py:775-777); host arrays (n,3) and (n,). */
int pst_set_receptor(pst_ctx* ctx, int n, const double* pos, const double* charges);
/* Ligand body-frame coordinates about the centroid and charges
 * (gpu_batch_simulator.py:171-176); host arrays (m,3) and (m,).  A ligand of
 * more than 64 atoms is held on the device in a spatial order (compact runs
 * of 32 and 64 atoms); forces, torques and energies are sums over atoms, so
 * only their summation order depends on it. */
int pst_set_ligand(pst_ctx* ctx, int m, const double* body, const double* charges);
/* 0 = reference torque (zero, gpu_batch_engine.py:597-599);
 * 1 = lever arm about the centroid of charged atoms (pystarc/forces/engine.py:144-183). */
int pst_set_torque_mode(pst_ctx* ctx, int mode);
/* Lanes per trajectory for the force kernel (1..64, power of two; 0 = automatic). */
int pst_set_force_group(pst_ctx* ctx, int lanes);
/* Force evaluation of the pose entry points and of new simulators: 0 = the
 * direct kernel (one trajectory per lane group, gathers from the grids), 1 =
 * brick-binned (gather requests sorted by 8^3-cell brick, each brick staged
 * once in LDS; same results bit for bit; needs the float2 receptor pair grid,
 * else the direct kernel runs), 2 = the direct kernel with z-pair fetches
 * (FP32), 3 / 4 = split (production steps: the ligand-atom terms over the
 * grid-path list, then the reverse Born over the list re-sorted by the
 * ligand-frame direction of the receptor; same results bit for bit; 4 with
 * z-pair fetches; elsewhere modes 0 / 2 run).  Mode 1 sizes its buffers for
 * max_traj trajectories (~32 B x (ligand + receptor atoms) each).  No
 * reference counterpart (the reference has one kernel,
 * gpu_batch_engine.py:65-123). */
int pst_set_gather_mode(pst_ctx* ctx, int mode, int max_traj);
/* What the context's last force evaluation launched: its mode (0 to 4 as
 * above, -1: none yet) and its lanes per trajectory. */
int pst_last_force_eval(const pst_ctx* ctx, int* mode, int* G);

/* ---- force operator ------------------------------------------------------
 * Replaces GPUBatchForceEngine.__call__(lig_positions, lig_charges,
 * R_matrices, centroids) (gpu_batch_engine.py:531-605).  All pointers are
 * device pointers.  lig_pos (n,m,3) f64, lig_q (m,) f64, rot (n,3,3) and
 * cen (n,3) f64 (both NULL disables reverse Born, as in the reference);
 * outputs F (n,3), T (n,3), E (n,) f64.  precision 0 = FP64 gather and sums,
 * 1 = FP32 gather with FP64 cross-lane reduction.  rev_skip 1 reproduces the
 * reference's 500 MB chunk skip (:626-639), 0 culls per trajectory exactly. */
int pst_forces_positions(pst_ctx* ctx, int n, int m, const double* lig_pos, const double* lig_q,
                         const double* rot, const double* cen, double* F, double* T, double* E,
                         int precision, int rev_skip);
/* pst_forces_positions without allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as pst_forces_positions with the direct
 * kernel (the context's gather mode is not used).  The reverse-Born chunk
 * skip (rev_skip 1) keeps its per-chunk minimum centroid distances in a
 * static device array of 65,536 doubles (512 KiB per GPU, held until the
 * process ends): a call needing more chunks than that (more than 65,536 x
 * floor(500 MiB / (24 N_rec)) trajectories) fails with an error.  Every call
 * on one GPU shares that array, so calls that overlap on one GPU race. */
int pst_forces_positions_noalloc(pst_ctx* ctx, int n, int m, const double* lig_pos,
                                 const double* lig_q, const double* rot, const double* cen,
                                 double* F, double* T, double* E, int precision, int rev_skip);
/* Fused variant: ligand atoms placed in registers from the pose (centroid
 * pos (n,3), quaternion quat (n,4) w,x,y,z) as gpu_batch_simulator.py:935-943
 * does, without materialising (n,m,3). */
int pst_forces_pose(pst_ctx* ctx, int n, const double* pos, const double* quat, double* F,
                    double* T, double* E, int precision);
/* Several ligand contexts (receptor grids shared with pst_share_receptor_grids,
 * each with its own ligand) evaluated in ONE launch of the grouped force
 * kernel, as pst_group_run's step does: per context l, n[l] poses pos[l]
 * (n,3) / quat[l] (n,4) and outputs F[l], T[l], E[l] (T / E arrays or their
 * entries may be NULL).  Pointer arrays are host arrays of device pointers.
 * Replaces one GPUBatchForceEngine.__call__ per ligand
 * (gpu_batch_engine.py:531-605, one process per ligand in multi_GPU_runs.py). */
int pst_forces_pose_group(pst_ctx* const* ctxs, int L, const int* n, const double* const* pos,
                          const double* const* quat, double* const* F, double* const* T,
                          double* const* E, int precision);
/* Times the fused force kernel: `reps` launches on the context stream,
 * bracketed by HIP events; writes the mean launch time in ms and, when
 * census4 is non-NULL, {trajectories, elec-grid atoms, born-grid atoms,
 * reverse-grid atoms} for one launch. */
int pst_time_forces_pose(pst_ctx* ctx, int n, const double* pos, const double* quat, double* F,
                         double* T, double* E, int precision, int reps, double* ms_per_launch,
                         double* census4);

/* Gather ceiling of the context's FP32 receptor potential + Born pair grid
 * (measurement only; no reference counterpart): n_points 20-node stencil
 * gathers (160 algorithmic bytes each, the FP32 force path's gather) at
 * points uniform over the grid (mode 0) or clustered 256 to a cube of edge
 * 12 / mode A (mode 1..16), `reps` timed launches; plus 32 x the fetch
 * variant (0 the plain stencil, 1 the z-pair fetches of the force kernel's
 * z-pair mode, 2 the same fetches issued by teams of 8 lanes).  out4 = {ms per launch, algorithmic GB/s, points, bytes per
 * point}. */
int pst_gather_ceiling(pst_ctx* ctx, int mode, long long n_points, int reps, double* out4);

/* ---- two-body RPY hydrodynamics ------------------------------------------
 * Extends MobilityTensor / rpy_offdiagonal
 * (pystarc/hydrodynamics/rotne_prager.py:67-124, :127-215), whose
 * translational coupling feeds the scalar D_parallel of
 * gpu_batch_simulator.py:1074-1084, to the full 12x12 grand mobility of the
 * two spheres (Zuk et al. 2014 three regimes for tt, rr and rt couplings),
 * order [x_rec, x_lig, theta_rec, theta_lig], 1/eta units, receptor at the
 * origin.  The production step uses it when StepParams.hydro_mode == 1.
 * pst_rpy12_eval writes, per pose (device pos (n,3)), the mobility M and the
 * noise factor G (G G^T = M), each (n,12,12) row-major; either may be NULL.
 * pst_rpy12_move applies one move with the step's own code: receptor load
 * from F / T of the ligand (torque_mode 0: zero torques as the reference,
 * 1: angular-momentum balance), dX = kT_over_eta dt M f
 * + sqrt(2 kT_over_eta dt) G xi (xi (n,12) standard normals), then pos /
 * quat (device, in-out) re-expressed in the turned receptor frame; dX (n,12)
 * may be NULL. */
int pst_rpy12_eval(pst_ctx* ctx, int n, const double* pos, double a_rec, double a_lig,
                   double* mobility, double* factor);
int pst_rpy12_move(pst_ctx* ctx, int n, double a_rec, double a_lig, double kT_over_eta,
                   int torque_mode, const double* dt, double* pos, double* quat, const double* F,
                   const double* T, const double* xi, double* dX);

/* ---- simulator -----------------------------------------------------------
 * Replaces GPUBatchSimulator (pystarc/simulation/gpu_batch_simulator.py:143-1751).
 * `params` points to a pst_step_params struct whose layout is StepParams
 * in pystarc_amd/csrc/pst_types.hpp (mirrored by the ctypes wrapper). */
int pst_sim_create(pst_ctx* ctx, const void* params, int precision, pst_sim** out);
int pst_sim_destroy(pst_sim* sim);
/* Host state in / out, (n,3) positions, (n,4) quaternions. */
int pst_sim_set_state(pst_sim* sim, int n, const double* pos, const double* quat);
int pst_sim_get_state(pst_sim* sim, double* pos, double* quat, int* status, long long* n_steps);

/* Replay mode: one reference loop iteration (gpu_batch_simulator.py:783-1254)
 * in four phases, the host supplying NumPy draws in the reference's order.
 * begin:  reaction check + escape candidates (:855-884).
 * escape: u_outer (n_esc) and u_quat (n_ret,3) host arrays (:884-911).
 * move:   forces, dt, D, noise_t / noise_r (n_sr,3) host arrays (:931-1128).
 * bridge: u_bb (n_sr, n_pairs) host array or NULL (:1142-1254).        */
int pst_replay_begin(pst_sim* sim, int* n_run, int* n_esc, int* n_react);
int pst_replay_escape(pst_sim* sim, const double* u_outer, int n_esc, const double* u_quat,
                      int n_ret, int* n_sr);
int pst_replay_move(pst_sim* sim, const double* noise_t, const double* noise_r, int n_sr,
                    int* any_above, int* n_overlap);
int pst_replay_bridge(pst_sim* sim, const double* u_bb, int n_sr, int* n_bridge);
/* Forces of the last move phase for the ordered still-running list (n_sr,3). */
int pst_replay_forces(pst_sim* sim, double* F_host, int n_sr);

/* Production mode: Philox noise on the device, a fixed pool of slots refilled
 * from the trajectory counter, no host synchronisation inside a launch.
 * record != 0 keeps per-trajectory outcome / step count (n_total each);
 * record = 2 also collects the observables (pst_prod_observables). */
int pst_prod_init(pst_sim* sim, int record);
/* Runs up to max_steps device steps, `graph_steps` per captured hipGraph
 * launch (0 = plain launches); stops early at the reference's 99.5 % rule
 * (gpu_batch_simulator.py:1347-1348) or when every slot is idle.
 * counters16 receives {n_reacted, n_escaped, n_maxsteps, total_steps,
 * next_traj, n_returns, n_overlap, n_bridge, steps_run, stop, n_sr, n_mlp,
 * 0, 0, 0, 0} as doubles. */
int pst_prod_run(pst_sim* sim, long long max_steps, int graph_steps, double* counters16);
int pst_prod_outcomes(pst_sim* sim, signed char* status, int* steps);
/* Observables of a run initialised with record = 2, collected on the device
 * instead of the reference's per-step host loops (sim_data of
 * gpu_batch_simulator.py:576-620, :1255-1340, :1524-1557).  traj: host
 * (n_total, traj_doubles) = start pos 0-2, start quaternion 3-6, closest
 * centroid distance 7, the trajectory's own step at it 8, pose there 9-15,
 * simulated time (ps) 16, returns 17, reacted in a Brownian bridge 18, last
 * pose 19-25, pairs within cutoff at the last pose 26 (slots still running
 * report their state so far).  hist: host (hist_len) counts = radial
 * histogram (200 bins over [0, 1.2 r_esc)), angular occupancy of the
 * centroid direction (36 theta x 72 phi), milestone crossings out / in at
 * linspace(b, r_esc, 11), radial transition matrix (50 x 50 bins over
 * [0, 1.2 r_esc)), per-pair contact counts at the pose a step starts from
 * (16 slots), trajectory-steps counted.  Either pointer may be NULL. */
int pst_prod_obs_sizes(int* traj_doubles, int* hist_len);
/* Several simulators (one ligand each; receptor grids shared with
 * pst_share_receptor_grids; same precision) stepped in lock-step: per step
 * each simulator's pre phases, ONE force launch over all grid-path lists
 * (workgroups partitioned by simulator; with L >= 8 each simulator's
 * workgroups on one XCD, re-balanced between graph launches), the surrogate
 * phase of every simulator with a surrogate beside it on a second stream,
 * each simulator's post phases; graph_steps per captured hipGraph.  c16 = (L,16) counters as
 * pst_prod_run; ms_per_step = device time per group step.  No reference
 * counterpart (the reference runs one ligand per process, multi_GPU_runs.py). */
int pst_group_run(pst_sim** sims, int L, long long max_steps, int graph_steps, double* c16,
                  double* ms_per_step);
/* State-machine reaction counts (StepParams.sm_active, reactions r = pair r;
 * gpu_batch_simulator.py:216-318, :818-854, :1185-1234): out48 = fires at the
 * reaction check [16], fires in the Brownian bridge [16], distinct
 * trajectories that fired each reaction [16]. */
int pst_prod_sm_counts(pst_sim* sim, unsigned long long* out48);
int pst_prod_observables(pst_sim* sim, double* traj, unsigned long long* hist);
/* Mean device time per step (ms) of the last pst_prod_run, HIP events. */
int pst_prod_last_ms_per_step(pst_sim* sim, double* ms);
/* The run counters as they stand, in pst_prod_run's 16-value layout (steps
 * field 0), without stepping (the counters of gpu_batch_simulator.py:
 * 1347-1348, 1524-1557 read between launches). */
int pst_prod_counters(pst_sim* sim, double* c16);
/* Autotuner of the production force launch (no reference counterpart: the
 * reference launches batch_force_kernel with a fixed (16,16) block,
 * gpu_batch_engine.py:328-332).  Candidates: the direct kernel with lanes per
 * trajectory G in {g0/4, g0/2, g0} (g0 = smallest power of two >= max(ligand
 * atoms, reverse-Born receptor atoms), at most 64), resident-workgroup cap in
 * {1024 ... 8192} and the XCD-chunked list mapping on/off; and, when the
 * receptor grids form a float2 pair, the brick-binned evaluation with G >= 8
 * and caps {1024, 4096}; in FP32, the direct kernel with z-pair fetches at
 * caps {2048, 8192}.  Each candidate runs `steps_per_cand` real production
 * steps per round, `rounds` round-robin passes, its force launch timed by HIP
 * events; the 4 fastest then run 4 more interleaved passes and the fastest of
 * that runoff becomes the launch of later steps.
 * table (max_rows,5) = {mode, G, cap, xcd, mean ms of the round-robin passes}
 * per candidate (mode 0 direct, 1 binned, 2 z-pair, 3 split, 4 split with
 * z-pair fetches);
 * best5 = {mode, G, cap, xcd, steps run}.
 * Both may be NULL. */
int pst_prod_autotune(pst_sim* sim, int rounds, int steps_per_cand, double* table, int max_rows,
                      int* n_rows, int* best5);
/* The same launch configuration set by hand (mode -1: the context's gather
 * mode, else 0 to 4 as pst_set_gather_mode; G = 0: heuristic). */
int pst_prod_set_launch(pst_sim* sim, int mode, int G, int blocks_cap, int xcd);
/* Times the force kernel inside the production step (no reference
 * counterpart): `steps` real steps with plain launches, each force launch
 * bracketed by HIP events on the context stream and preceded by a gather
 * census of its trajectory list; empty k_timing_window_mark dispatches bracket
 * the window in a kernel trace.  out6 = {mean ms per force launch, steps run,
 * and per launch the mean trajectories, receptor-potential-grid atoms,
 * receptor-Born-grid atoms, ligand-Born-grid (reverse) atoms}. */
int pst_prod_time_forces(pst_sim* sim, int steps, double* out6);

/* Checkpoint / resume of a production run (the reference saves pos, q,
 * status, n_steps and counters to checkpoint.npz every checkpoint_interval
 * completed trajectories and resumes from it, gpu_batch_simulator.py:548-575,
 * :1400-1420).  The saved state is every device buffer that carries from one
 * step to the next (slots, counters, outcomes, observables) behind a header
 * and the run's StepParams; noise is keyed by (trajectory, own step), so a run
 * resumed from it continues exactly as the uninterrupted run.  load requires
 * a sim created with the same StepParams and initialised with the same record
 * level.  bytes = pst_prod_state_bytes. */
int pst_prod_state_bytes(pst_sim* sim, long long* bytes);
int pst_prod_save_state(pst_sim* sim, void* host, long long bytes);
int pst_prod_load_state(pst_sim* sim, const void* host, long long bytes);
/* Path and energetics snapshots of the running trajectories that have moved
 * (gpu_batch_simulator.py:1318-1340): path8 (n_slots,8) rows = trajectory,
 * step, centroid x y z, q0 q1 q2; energy6 (n_slots,6) rows = trajectory, step,
 * force x y z, dt; rows in trajectory order; host arrays (either may be
 * NULL), n_rows written. */
int pst_prod_snapshot(pst_sim* sim, long long step, double* path8, double* energy6, int* n_rows);
/* Surrogate audit sums since the last pst_sim_reset_audit: out6 = {force
 * squared error, squared grid-path force, audited rows, torque squared error,
 * squared grid-path torque (torques audited when the surrogate applies its
 * torque), 0}.  No reference counterpart. */
int pst_prod_audit(pst_sim* sim, double* out6);

/* ---- learned surrogate ---------------------------------------------------
 * No reference counterpart: a 9 -> 4x256 (SiLU) -> 13 MLP standing in for
 * the grid + Born + far-field interaction outside a contact shell, with
 * forward and backward on BF16 MFMA (FP32 accumulate).  Parameters are one
 * device FP32 vector of pst_mlp_param_count() floats, [out][in] per layer
 * (W0 256x9, b0, W1..W3 256x256 + biases, W4 13x256, b4).                */
typedef struct pst_mlp pst_mlp;
int pst_mlp_param_count(void);
int pst_mlp_create(pst_ctx* ctx, int max_rows, pst_mlp** out);
int pst_mlp_destroy(pst_mlp* mlp);
/* Casts device FP32 parameters to the kernels' padded BF16 weights. */
int pst_mlp_load(pst_mlp* mlp, const float* params);
/* x (B,9) f32 -> y (B,13) f32, device pointers; train != 0 keeps the
 * activations for pst_mlp_backward. */
int pst_mlp_forward(pst_mlp* mlp, int B, const float* x, float* y, int train);
/* Average duration (HIP events on the context stream) of the fused
 * inference kernel alone over `reps` launches on the first B rows of the
 * features left by the last pst_mlp_forward of at least B rows; the outputs
 * stay in the library's buffer.  Measurement only; no reference counterpart. */
int pst_mlp_time_infer(pst_mlp* mlp, int B, int reps, double* ms_per_launch);
/* dL/dy (B,13) f32 -> dL/dparams (flat f32), for the last train forward. */
int pst_mlp_backward(pst_mlp* mlp, int B, const float* gy, float* gparams);
/* Attaches (mlp != NULL) or detaches the surrogate branch of the production
 * step.  p9 = {r_on, r0, lambda_env, 1/b, scale_F, scale_T, scale_E,
 * use_torque, audit_n}: trajectories with centre distance > r_on take the
 * MLP force; the grid path re-evaluates audit_n of them per step (a stride
 * through the surrogate list from an offset that changes every step) and the
 * squared error / reference sums come back in counters16[12..14] (forces) and
 * pst_prod_audit (forces and torques). */
int pst_sim_set_surrogate(pst_sim* sim, pst_mlp* mlp, const double* p9);
int pst_sim_reset_audit(pst_sim* sim);

/* ---- synthetic grids -----------------------------------------------------
 * Fills a host array (nx,ny,nz) C-order with a Debye-Hueckel potential
 * (kind 0) or an analytic Born desolvation field (kind 1) of `atoms`
 * (n,5: x, y, z, charge, radius), evaluated on the device. dtype PST_F32/PST_F64. */
int pst_make_grid(pst_ctx* ctx, int kind, int nx, int ny, int nz, const double* origin3,
                  const double* spacing3, const double* atoms, int n_atoms, double debye_length,
                  double eps_out, double eps_in, int dtype, void* host_out);

#ifdef __cplusplus
}
#endif
#endif /* PYSTARC_HIP_H */
