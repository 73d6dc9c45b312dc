// C ABI of libpystarc_hip.so (declared in include/pystarc_hip.h): contexts
// owning resident grids, ligand / receptor tables and simulator state, and
// the entry points that enqueue the gfx950 kernels on the context stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <string>
#include <vector>
#include <algorithm>
#include <mutex>

#include "../../include/pystarc_hip.h"
#include "pst_kernels.hpp"
#include "pst_mlp.hpp"
#include "pst_order.hpp"
#include "pst_surrogate.hpp"

using namespace pst;

namespace {

struct DevGrid {
  GridDev g{};
  void* mem = nullptr;
  bool owned = true;   // false: device memory of another context's grid
};

static void free_grid(DevGrid& g) {
  if (g.owned && g.mem) hipFree(g.mem);
  g.mem = nullptr;
  g.g.data = nullptr;
  g.owned = true;
}

}  // namespace

struct pst_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  std::vector<DevGrid> elec, born, lig_born;
  DevGrid pair;  // interleaved elec + Born copy when both share a geometry
  double receptor_charge = 0, debye = 7.858, sdie = 78.0;
  bool has_mp = false;
  double dip[3] = {0, 0, 0}, quad[9] = {0};
  double alpha = 0.07957747;
  int M = 0;
  double* d_body = nullptr;
  double* d_q = nullptr;
  double cq_body[3] = {0, 0, 0};
  double lig_rmax = 1e300;
  int n_rec = 0;
  double* d_rec = nullptr;
  double* d_rec_q = nullptr;
  double* d_rec_chunk = nullptr;
  int n_rec_chunk = 0;
  int cell_n[3] = {0, 0, 0};
  double cell_o[3] = {0, 0, 0};
  int* d_cell_start = nullptr;
  double* d_cell_pos = nullptr;
  double rec_max_r = 0;
  int torque_mode = 0;
  int force_group = 0;
  // scratch for the reference chunk skip
  double* d_chunk = nullptr;
  int chunk_cap = 0;
  // brick-binned force evaluation (0 = direct kernel, 1 = binned when eligible)
  int gather_mode = 0;
  int last_mode = -1, last_G = 0;  // what the last force evaluation launched
  BinBufs bins{};
  std::vector<void*> bin_allocs;
};

struct pst_mlp {
  pst_ctx* ctx = nullptr;
  MlpBufs m;
  std::vector<void*> allocs;
};

struct pst_sim {
  pst_ctx* ctx = nullptr;
  StepParams P;
  int precision = 0;
  int S = 0;
  // shared state
  double* pos = nullptr;
  double* quat = nullptr;
  int* status = nullptr;
  double* F = nullptr;
  double* T = nullptr;
  double* dt = nullptr;
  double* x0 = nullptr;
  double* x1 = nullptr;
  // replay
  long long* n_steps64 = nullptr;
  unsigned char* flags = nullptr;  // 4 * S
  int* lists = nullptr;            // 4 * S
  int* counts = nullptr;           // 8
  double* h_buf = nullptr;         // device staging for host noise
  size_t h_buf_cap = 0;
  // production
  unsigned int* n_steps32 = nullptr;
  long long* traj = nullptr;
  int* plain_list = nullptr;
  int* rev_list = nullptr;           // split evaluation: reverse-Born order,
  double* F_rev = nullptr;           // and the reverse-Born corrections
  double* T_rev = nullptr;
  int* mlp_list = nullptr;
  int* key = nullptr;
  int* bins = nullptr;
  RunCounters* ctr = nullptr;
  signed char* out_status = nullptr;
  int* out_steps = nullptr;
  ObsBufs obs{};                     // observables (record level 2)
  int* sm_state = nullptr;           // state-machine reactions
  unsigned int* sm_mask = nullptr;
  int* lmz_escaped = nullptr;        // outer-propagator escapes of the step
  std::vector<void*> obs_allocs;
  double last_ms_per_step = 0;
  bool prod_ready = false;
  // surrogate branch
  pst_mlp* mlp = nullptr;
  SurrParams sp;
  double* audit_F = nullptr;
  double* audit_T = nullptr;
  int* audit_list = nullptr;    // slot ids of the audited surrogate trajectories
  int* audit_count = nullptr;   // their number (device)
  int audit_cap = 0;
  // the captured step graph of pst_prod_run, re-launched while everything it
  // captured (launch configuration, parameters, buffers, system) is unchanged
  hipGraphExec_t g_exec = nullptr;
  hipGraph_t g_graph = nullptr;
  hipStream_t g_cap = nullptr;
  std::vector<unsigned char> g_key;
  // force-launch configuration of the production step (autotuned or default)
  int tune_G = 0;          // lanes per trajectory; 0 = pick_group heuristic
  int tune_cap = 2048;     // cap on resident workgroups (grid-stride beyond)
  int tune_xcd = 1;        // XCD-chunked mapping of the sorted grid-path list
  int tune_mode = -1;      // 0 direct, 1 brick-binned, -1 the context's gather mode
  hipEvent_t tev0 = nullptr, tev1 = nullptr;  // brackets the force launch while tuning
  // the surrogate audit's grid-path launch runs beside the main force launch
  // on a second stream (fork after the sort, join before the comparison)
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  unsigned long long* census = nullptr;       // gather census of timed force launches
  bool census_on = false;
};

#define CK(expr)                                                                      \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)

#define FAIL(msg)          \
  do {                     \
    ctx->err = (msg);      \
    return 1;              \
  } while (0)

namespace {
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
  }
};
}  // namespace

static hipError_t set_dev(pst_ctx* ctx) { return hipSetDevice(ctx->device); }

static GridDev grid_record(const DevGrid& d) { return d.g; }

// Cubic cells of kRecCell A over the receptor's bounding box plus one cell on
// each side; atom a goes to cell floor((x_a - o) / kRecCell) per axis (the
// division written as a multiplication by the reciprocal, as the kernels
// form their query ranges).
constexpr double kRecCell = 2.0;

static void rec_cells(int n, const double* pos, int* cn, double* co, std::vector<int>& cstart,
                      std::vector<double>& cpos) {
  const double inv = 1.0 / kRecCell;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int j = 0; j < n; ++j)
    for (int d = 0; d < 3; ++d) {
      mn[d] = fmin(mn[d], pos[3 * j + d]);
      mx[d] = fmax(mx[d], pos[3 * j + d]);
    }
  for (int d = 0; d < 3; ++d) {
    co[d] = mn[d] - kRecCell;
    cn[d] = (int)floor((mx[d] - co[d]) * inv) + 2;
  }
  const long nc = (long)cn[0] * cn[1] * cn[2];
  std::vector<int> cell(n);
  cstart.assign(nc + 1, 0);
  for (int j = 0; j < n; ++j) {
    int ijk[3];
    for (int d = 0; d < 3; ++d) {
      ijk[d] = (int)floor((pos[3 * j + d] - co[d]) * inv);
      ijk[d] = ijk[d] < 0 ? 0 : (ijk[d] >= cn[d] ? cn[d] - 1 : ijk[d]);
    }
    cell[j] = (ijk[0] * cn[1] + ijk[1]) * cn[2] + ijk[2];
    cstart[cell[j] + 1] += 1;
  }
  for (long c = 0; c < nc; ++c) cstart[c + 1] += cstart[c];
  std::vector<int> cur(cstart.begin(), cstart.end() - 1);
  cpos.assign(3 * (size_t)n, 0.0);
  for (int j = 0; j < n; ++j) {
    const int p = cur[cell[j]]++;
    for (int d = 0; d < 3; ++d) cpos[3 * (size_t)p + d] = pos[3 * j + d];
  }
}

static void fill_sys(pst_ctx* ctx, ForceSys& S) {
  memset(&S, 0, sizeof(S));
  S.n_elec = (int)ctx->elec.size();
  for (int i = 0; i < S.n_elec; ++i) S.elec[i] = grid_record(ctx->elec[i]);
  S.n_born = (int)ctx->born.size();
  for (int i = 0; i < S.n_born; ++i) S.born[i] = grid_record(ctx->born[i]);
  S.paired = ctx->pair.mem != nullptr ? 1 : 0;
  if (S.paired) S.pair = ctx->pair.g;
  S.has_rev = (!ctx->lig_born.empty() && ctx->n_rec > 0) ? 1 : 0;
  if (S.has_rev) {
    S.lig_born = ctx->lig_born[0].g;
    // skip radius of the reference (:620-623) and exact cull radius
    S.rev_radius = fmax(fabs(S.lig_born.lo[0]), fabs(S.lig_born.hi[0]));
    double m = 0;
    for (int c = 0; c < 8; ++c) {
      double x = (c & 1) ? S.lig_born.hi[0] : S.lig_born.lo[0];
      double y = (c & 2) ? S.lig_born.hi[1] : S.lig_born.lo[1];
      double z = (c & 4) ? S.lig_born.hi[2] : S.lig_born.lo[2];
      m = fmax(m, sqrt(x * x + y * y + z * z));
    }
    S.rev_cull = m * (1.0 + 1e-12) + 1e-9;
  }
  S.alpha = ctx->alpha;
  const double eps_s = ctx->sdie * 0.000142;
  S.V_fac = eps_s > 0 ? ctx->receptor_charge / (4.0 * M_PI * eps_s) : 0.0;
  S.has_yukawa = (fabs(ctx->receptor_charge) > 1e-9 && ctx->debye > 0) ? 1 : 0;
  S.debye = ctx->debye;
  S.fpe = 4.0 * M_PI * eps_s;
  if (ctx->has_mp) {
    double pm = sqrt(ctx->dip[0] * ctx->dip[0] + ctx->dip[1] * ctx->dip[1] + ctx->dip[2] * ctx->dip[2]);
    double qm = 0;
    for (int i = 0; i < 9; ++i) qm += ctx->quad[i] * ctx->quad[i];
    qm = sqrt(qm);
    S.has_dip = pm > 1e-9;
    S.has_quad = qm > 1e-9;
    for (int i = 0; i < 3; ++i) S.dip[i] = ctx->dip[i];
    for (int i = 0; i < 9; ++i) S.quad[i] = ctx->quad[i];
  }
  const double lam = ctx->debye > 0 ? ctx->debye : 1.0;
  S.inv_lam = 1.0 / lam;
  S.inv_lam2 = 1.0 / (lam * lam);
  S.c3 = 1.0 / (3.0 * lam * lam);
  S.c3l = 1.0 / (3.0 * lam * lam * lam);
  for (int i = 0; i < 3; ++i) S.dip_s[i] = S.dip[i] / S.fpe;
  for (int i = 0; i < 9; ++i) S.quad_s[i] = S.quad[i] / S.fpe;
  S.lig_rmax = ctx->lig_rmax;
  double reach = 0;
  auto corner_reach = [&](const GridDev& g) {
    for (int c = 0; c < 8; ++c) {
      const double x = (c & 1) ? g.hi[0] : g.lo[0];
      const double y = (c & 2) ? g.hi[1] : g.lo[1];
      const double z = (c & 4) ? g.hi[2] : g.lo[2];
      reach = fmax(reach, sqrt(x * x + y * y + z * z));
    }
  };
  for (int i = 0; i < S.n_elec; ++i) corner_reach(S.elec[i]);
  for (int i = 0; i < S.n_born; ++i) corner_reach(S.born[i]);
  S.grid_reach = reach * (1.0 + 1e-12) + 1e-9;
  S.M = ctx->M;
  S.lig_body = ctx->d_body;
  S.lig_q = ctx->d_q;
  for (int i = 0; i < 3; ++i) S.cq_body[i] = ctx->cq_body[i];
  S.n_rec = ctx->n_rec;
  S.rec_pos = ctx->d_rec;
  S.rec_q = ctx->d_rec_q;
  S.rec_max_r = ctx->rec_max_r;
  S.n_rec_chunk = ctx->n_rec_chunk;
  S.rec_chunk = ctx->d_rec_chunk;
  for (int d = 0; d < 3; ++d) { S.cell_n[d] = ctx->cell_n[d]; S.cell_o[d] = ctx->cell_o[d]; }
  S.cell_inv = 1.0 / kRecCell;
  S.cell_start = ctx->d_cell_start;
  S.cell_pos = ctx->d_cell_pos;
  S.torque_mode = ctx->torque_mode;
}

static int group_for(pst_ctx* ctx, const ForceSys& S) {
  if (ctx->force_group > 0) return ctx->force_group;
  return pick_group(S.M, S.has_rev ? S.n_rec : 0);
}

extern "C" {

int pst_abi_version(void) { return 1; }

int pst_struct_size(int which) {
  if (which == 0) return (int)sizeof(StepParams);
  if (which == 1) return (int)sizeof(RunCounters);
  return -1;
}

int pst_ctx_create(pst_ctx** out, int device, void* stream, int use_stream) {
  if (!out) return 1;
  pst_ctx* ctx = new pst_ctx();
  ctx->device = device;
  *out = ctx;
  CK(set_dev(ctx));
  if (use_stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    CK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  return 0;
}

static void free_grids(std::vector<DevGrid>& v) {
  for (auto& g : v) free_grid(g);
  v.clear();
}

int pst_ctx_destroy(pst_ctx* ctx) {
  if (!ctx) return 0;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_grids(ctx->elec);
  free_grids(ctx->born);
  free_grids(ctx->lig_born);
  free_grid(ctx->pair);
  if (ctx->d_body) hipFree(ctx->d_body);
  if (ctx->d_q) hipFree(ctx->d_q);
  if (ctx->d_rec) hipFree(ctx->d_rec);
  if (ctx->d_rec_q) hipFree(ctx->d_rec_q);
  if (ctx->d_rec_chunk) hipFree(ctx->d_rec_chunk);
  if (ctx->d_cell_start) hipFree(ctx->d_cell_start);
  if (ctx->d_cell_pos) hipFree(ctx->d_cell_pos);
  if (ctx->d_chunk) hipFree(ctx->d_chunk);
  for (void* p : ctx->bin_allocs) hipFree(p);
  if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char* pst_last_error(const pst_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Keep an interleaved float2 copy of the receptor potential and Born grids
// when there is exactly one of each, both FP32, with identical geometry.
static int update_pair(pst_ctx* ctx) {
  free_grid(ctx->pair);
  if (ctx->elec.size() != 1 || ctx->born.size() != 1) return 0;
  const GridDev& a = ctx->elec[0].g;
  const GridDev& b = ctx->born[0].g;
  if (a.dtype != 0 || b.dtype != 0 || a.nx != b.nx || a.ny != b.ny || a.nz != b.nz) return 0;
  for (int i = 0; i < 3; ++i)
    if (a.o[i] != b.o[i] || a.sp[i] != b.sp[i]) return 0;
  const long n = brick_nodes(a.nx, a.ny, a.nz);  // bricked layout, padding included
  CK(hipMalloc(&ctx->pair.mem, (size_t)n * sizeof(float2)));
  CK(launch_interleave((const float*)a.data, (const float*)b.data, (float2*)ctx->pair.mem, n,
                       ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  ctx->pair.g = a;
  ctx->pair.g.data = ctx->pair.mem;
  ctx->pair.g.dtype = 2;
  return 0;
}

int pst_add_grid(pst_ctx* ctx, int kind, const void* host_data, int dtype, int nx, int ny, int nz,
                 const double* origin3, const double* spacing3) {
  if (!ctx) return 1;
  if (kind < 0 || kind > 2) FAIL("pst_add_grid: unknown grid kind");
  if (dtype != PST_F32 && dtype != PST_F64) FAIL("pst_add_grid: dtype must be PST_F32 or PST_F64");
  if (nx < 2 || ny < 2 || nz < 2) FAIL("pst_add_grid: every grid dimension must be >= 2");
  if (!host_data || !origin3 || !spacing3) FAIL("pst_add_grid: null pointer");
  std::vector<DevGrid>& dst = kind == 0 ? ctx->elec : (kind == 1 ? ctx->born : ctx->lig_born);
  if ((int)dst.size() >= PST_MAX_GRIDS) FAIL("pst_add_grid: too many grids of this kind");
  for (int i = 0; i < 3; ++i)
    if (!(spacing3[i] > 0)) FAIL("pst_add_grid: spacing must be positive");
  CK(set_dev(ctx));
  DevGrid d;
  // uploaded in C order, stored in the kernels' bricked layout (zero padding)
  const int elem = dtype == PST_F64 ? 8 : 4;
  const size_t count = (size_t)nx * ny * nz;
  const size_t bytes = (size_t)brick_nodes(nx, ny, nz) * elem;
  void* tmp = nullptr;
  CK(hipMalloc(&d.mem, bytes));
  hipError_t e = hipMemset(d.mem, 0, bytes);
  if (e == hipSuccess) e = hipMalloc(&tmp, count * elem);
  if (e == hipSuccess) e = hipMemcpy(tmp, host_data, count * elem, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_to_bricks(tmp, d.mem, elem, nx, ny, nz, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (tmp) hipFree(tmp);
  if (e != hipSuccess) {
    hipFree(d.mem);
    ctx->err = std::string("pst_add_grid upload: ") + hipGetErrorString(e);
    return 2;
  }
  d.g.data = d.mem;
  d.g.dtype = dtype;
  d.g.nx = nx; d.g.ny = ny; d.g.nz = nz;
  const int n[3] = {nx, ny, nz};
  for (int i = 0; i < 3; ++i) {
    d.g.o[i] = origin3[i];
    d.g.sp[i] = spacing3[i];
    d.g.inv[i] = 1.0 / spacing3[i];
    d.g.lo[i] = origin3[i] + 3.0 * spacing3[i];
    d.g.hi[i] = origin3[i] + (n[i] - 1 - 3.0) * spacing3[i];
  }
  dst.push_back(d);
  // the float2 pair depends on the receptor grids only
  return kind == 2 ? 0 : update_pair(ctx);
}

int pst_share_receptor_grids(pst_ctx* dst, const pst_ctx* src) {
  if (!dst) return 1;
  pst_ctx* ctx = dst;
  if (!src) FAIL("pst_share_receptor_grids: null source context");
  if (src->device != dst->device) FAIL("pst_share_receptor_grids: contexts on different devices");
  if (!dst->elec.empty() || !dst->born.empty())
    FAIL("pst_share_receptor_grids: the destination already has receptor grids");
  CK(set_dev(ctx));
  for (const auto& g : src->elec) { DevGrid d = g; d.owned = false; dst->elec.push_back(d); }
  for (const auto& g : src->born) { DevGrid d = g; d.owned = false; dst->born.push_back(d); }
  free_grid(dst->pair);
  dst->pair = src->pair;
  dst->pair.owned = false;
  return 0;
}

int pst_clear_grids(pst_ctx* ctx) {
  if (!ctx) return 1;
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  free_grids(ctx->elec);
  free_grids(ctx->born);
  free_grids(ctx->lig_born);
  return update_pair(ctx);
}

int pst_set_farfield(pst_ctx* ctx, double receptor_charge, double debye_length, double sdie,
                     const double* dipole3, const double* quad9) {
  if (!ctx) return 1;
  ctx->receptor_charge = receptor_charge;
  ctx->debye = debye_length;
  ctx->sdie = sdie;
  ctx->has_mp = dipole3 != nullptr && quad9 != nullptr;
  for (int i = 0; i < 3; ++i) ctx->dip[i] = dipole3 ? dipole3[i] : 0.0;
  for (int i = 0; i < 9; ++i) ctx->quad[i] = quad9 ? quad9[i] : 0.0;
  return 0;
}

int pst_set_alpha(pst_ctx* ctx, double alpha) {
  if (!ctx) return 1;
  ctx->alpha = alpha;
  return 0;
}

int pst_set_torque_mode(pst_ctx* ctx, int mode) {
  if (!ctx) return 1;
  if (mode != 0 && mode != 1) FAIL("pst_set_torque_mode: mode must be 0 or 1");
  ctx->torque_mode = mode;
  return 0;
}

int pst_set_force_group(pst_ctx* ctx, int lanes) {
  if (!ctx) return 1;
  if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 &&
      lanes != 32 && lanes != 64)
    FAIL("pst_set_force_group: lanes must be 0 or a power of two <= 64");
  ctx->force_group = lanes;
  return 0;
}

int pst_set_receptor(pst_ctx* ctx, int n, const double* pos, const double* charges) {
  if (!ctx) return 1;
  if (n < 0 || (n > 0 && (!pos || !charges))) FAIL("pst_set_receptor: bad arguments");
  CK(set_dev(ctx));
  if (ctx->d_rec) { hipFree(ctx->d_rec); ctx->d_rec = nullptr; }
  if (ctx->d_rec_q) { hipFree(ctx->d_rec_q); ctx->d_rec_q = nullptr; }
  if (ctx->d_rec_chunk) { hipFree(ctx->d_rec_chunk); ctx->d_rec_chunk = nullptr; }
  if (ctx->d_cell_start) { hipFree(ctx->d_cell_start); ctx->d_cell_start = nullptr; }
  if (ctx->d_cell_pos) { hipFree(ctx->d_cell_pos); ctx->d_cell_pos = nullptr; }
  ctx->n_rec = n;
  ctx->n_rec_chunk = 0;
  for (int d = 0; d < 3; ++d) ctx->cell_n[d] = 0;
  ctx->rec_max_r = 0;
  if (n == 0) return 0;
  // Device copy in chunk order: spatially compact runs of PST_REC_CHUNK atoms
  // with a bounding sphere each, so the reverse-Born and overlap loops skip a
  // whole run that cannot reach the ligand.  Both loops only add (reverse
  // Born: per lane, in atom order) or take a minimum over the atoms, so the
  // order the atoms are stored in is the only thing this changes.
  std::vector<int> order;
  std::vector<double> spheres;
  rec_chunk_order(n, pos, order, spheres);
  std::vector<double> sp(3 * (size_t)n), sq(n);
  for (int j = 0; j < n; ++j) {
    const int a = order[j];
    sp[3 * j] = pos[3 * a]; sp[3 * j + 1] = pos[3 * a + 1]; sp[3 * j + 2] = pos[3 * a + 2];
    sq[j] = charges[a];
  }
  ctx->n_rec_chunk = (int)(spheres.size() / 4);
  CK(hipMalloc(&ctx->d_rec, sizeof(double) * 3 * n));
  CK(hipMalloc(&ctx->d_rec_q, sizeof(double) * n));
  CK(hipMalloc(&ctx->d_rec_chunk, sizeof(double) * spheres.size()));
  CK(hipMemcpy(ctx->d_rec, sp.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->d_rec_q, sq.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->d_rec_chunk, spheres.data(), sizeof(double) * spheres.size(),
               hipMemcpyHostToDevice));
  // uniform cell grid (counting sort by cell) for the overlap test
  std::vector<int> cstart;
  std::vector<double> cpos;
  rec_cells(n, pos, ctx->cell_n, ctx->cell_o, cstart, cpos);
  CK(hipMalloc(&ctx->d_cell_start, sizeof(int) * cstart.size()));
  CK(hipMalloc(&ctx->d_cell_pos, sizeof(double) * cpos.size()));
  CK(hipMemcpy(ctx->d_cell_start, cstart.data(), sizeof(int) * cstart.size(),
               hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->d_cell_pos, cpos.data(), sizeof(double) * cpos.size(), hipMemcpyHostToDevice));
  for (int j = 0; j < n; ++j) {
    const double r = sqrt(pos[3 * j] * pos[3 * j] + pos[3 * j + 1] * pos[3 * j + 1] +
                          pos[3 * j + 2] * pos[3 * j + 2]);
    if (r > ctx->rec_max_r) ctx->rec_max_r = r;
  }
  return 0;
}

int pst_set_ligand(pst_ctx* ctx, int m, const double* body, const double* charges) {
  if (!ctx) return 1;
  if (m <= 0 || !body || !charges) FAIL("pst_set_ligand: bad arguments");
  CK(set_dev(ctx));
  if (ctx->d_body) { hipFree(ctx->d_body); ctx->d_body = nullptr; }
  if (ctx->d_q) { hipFree(ctx->d_q); ctx->d_q = nullptr; }
  ctx->M = m;
  CK(hipMalloc(&ctx->d_body, sizeof(double) * 3 * m));
  CK(hipMalloc(&ctx->d_q, sizeof(double) * m));
  // device order of the atoms (spatial for ligands of more than 64 atoms)
  std::vector<int> order;
  ligand_order(m, body, order);
  std::vector<double> hb(3 * (size_t)m), hq(m);
  for (int i = 0; i < m; ++i) {
    const int a = order[i];
    for (int d = 0; d < 3; ++d) hb[3 * (size_t)i + d] = body[3 * (size_t)a + d];
    hq[i] = charges[a];
  }
  CK(hipMemcpy(ctx->d_body, hb.data(), sizeof(double) * 3 * m, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->d_q, hq.data(), sizeof(double) * m, hipMemcpyHostToDevice));
  double c[3] = {0, 0, 0};
  int nc = 0;
  for (int i = 0; i < m; ++i)
    if (fabs(charges[i]) > 1e-9) {
      for (int d = 0; d < 3; ++d) c[d] += body[3 * i + d];
      ++nc;
    }
  for (int d = 0; d < 3; ++d) ctx->cq_body[d] = nc ? c[d] / nc : 0.0;
  double rmax = 0;
  for (int i = 0; i < m; ++i)
    rmax = fmax(rmax, sqrt(body[3 * i] * body[3 * i] + body[3 * i + 1] * body[3 * i + 1] +
                           body[3 * i + 2] * body[3 * i + 2]));
  ctx->lig_rmax = rmax * (1.0 + 1e-12) + 1e-9;
  return 0;
}

static int ensure_chunk(pst_ctx* ctx, int n) {
  if (n <= ctx->chunk_cap) return 0;
  if (ctx->d_chunk) hipFree(ctx->d_chunk);
  ctx->d_chunk = nullptr;
  ctx->chunk_cap = 0;
  CK(hipMalloc(&ctx->d_chunk, sizeof(double) * n));
  ctx->chunk_cap = n;
  return 0;
}

// Reference chunking of a force batch: 500 MiB / (N_rec * 24) trajectories
// (gpu_batch_engine.py:627-629) within force batches of `fb` trajectories.
static void chunk_layout(const pst_ctx* ctx, int n, int fb, int* cs, int* cpf, int* n_chunks) {
  int c = (int)fmax(1.0, floor(500.0 * 1024 * 1024 / ((double)ctx->n_rec * 3 * 8)));
  if (c > fb) c = fb;
  *cs = c;
  *cpf = (fb + c - 1) / c;
  const int nfb = (n + fb - 1) / fb;
  *n_chunks = nfb * (*cpf);
}

// Bin buffers for up to max_traj trajectories of the current system (grown,
// never shrunk; not called while a graph is being captured).  Left empty when
// the system has no float2 pair grid or the request ids would not fit int32.
static void free_bins(pst_ctx* ctx) {
  for (void* p : ctx->bin_allocs) hipFree(p);
  ctx->bin_allocs.clear();
  memset(&ctx->bins, 0, sizeof(BinBufs));
}

static int ensure_bins(pst_ctx* ctx, int max_traj) {
  if (!ctx->pair.mem || max_traj <= 0 || ctx->M <= 0) return 0;
  const GridDev& pg = ctx->pair.g;
  const bool rev = !ctx->lig_born.empty() && ctx->n_rec > 0;
  int rb[3] = {bricks_per_axis(pg.nx), bricks_per_axis(pg.ny), bricks_per_axis(pg.nz)};
  int lb[3] = {0, 0, 0};
  if (rev) {
    const GridDev& lg = ctx->lig_born[0].g;
    lb[0] = bricks_per_axis(lg.nx); lb[1] = bricks_per_axis(lg.ny); lb[2] = bricks_per_axis(lg.nz);
  }
  BinBufs& B = ctx->bins;
  const int need_rec = rev ? ctx->n_rec : 0;
  if (B.hist && B.max_traj >= max_traj && B.max_m >= ctx->M && B.max_rec >= need_rec &&
      !memcmp(B.rb, rb, sizeof(rb)) && !memcmp(B.lb, lb, sizeof(lb)))
    return 0;
  const int mt = B.hist && B.max_traj > max_traj ? B.max_traj : max_traj;
  const long max_req = (long)mt * ((long)ctx->M + need_rec);
  free_bins(ctx);
  if (max_req >= 2147483647L) return 0;
  const int nb_rec = rb[0] * rb[1] * rb[2], nb_lig = lb[0] * lb[1] * lb[2], nb = nb_rec + nb_lig;
  const int max_items = bin_items_bound(max_req, nb);
  auto get = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    ctx->bin_allocs.push_back(p);
    return p;
  };
  B.hist = (int*)get((size_t)nb * 4);
  B.start = (int*)get((size_t)(nb + 1) * 4);
  B.cursor = (int*)get((size_t)nb * 4);
  B.items = (int2*)get((size_t)max_items * 8);
  B.n_items = (int*)get(4);
  B.sorted = (int*)get((size_t)max_req * 4);
  B.out = get((size_t)mt * ctx->M * 32);
  B.out_rev = get((size_t)mt * need_rec * 32);
  if (!B.hist || !B.start || !B.cursor || !B.items || !B.n_items || !B.sorted || !B.out ||
      !B.out_rev) {
    free_bins(ctx);
    ctx->err = "ensure_bins: out of device memory";
    return 2;
  }
  CK(hipMemsetAsync(B.hist, 0, (size_t)nb * 4, ctx->stream));
  CK(hipMemsetAsync(B.n_items, 0, 4, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  B.nb_rec = nb_rec; B.nb_lig = nb_lig;
  memcpy(B.rb, rb, sizeof(rb));
  memcpy(B.lb, lb, sizeof(lb));
  B.max_traj = mt; B.max_m = ctx->M; B.max_rec = need_rec;
  B.rev_base = mt * ctx->M;
  B.max_items = max_items;
  B.gather_blocks = 2048;
  return 0;
}

// The binned buffers must describe the current grids.
static bool bins_match(pst_ctx* ctx, const ForceSys& S) {
  const BinBufs& B = ctx->bins;
  if (!B.hist || !S.paired) return false;
  if (B.rb[0] != bricks_per_axis(S.pair.nx) || B.rb[1] != bricks_per_axis(S.pair.ny) ||
      B.rb[2] != bricks_per_axis(S.pair.nz))
    return false;
  if (S.has_rev) {
    if (B.lb[0] != bricks_per_axis(S.lig_born.nx) || B.lb[1] != bricks_per_axis(S.lig_born.ny) ||
        B.lb[2] != bricks_per_axis(S.lig_born.nz))
      return false;
  }
  return true;
}

// Whether the split evaluation (modes 3 / 4) applies: pose mode over a
// device-counted list with per-trajectory outputs, a buffer for the reverse
// order, the exact cull, G of a part kernel.
static bool split_eligible(const ForceSys& S, const ForceWork& W, int G) {
  return S.has_rev && W.rev_order && W.F_rev && W.list && W.n_dev && W.quat && !W.lig_pos &&
         !W.out_compact && W.rev_skip_mode == 0 && (G == 16 || G == 32 || G == 64);
}

// One force evaluation: the brick-binned pipeline when asked for and
// eligible (same results as the direct kernel with the same G); the split
// evaluation (the ligand-atom terms over the list and the reverse Born over
// the list re-sorted by the ligand-frame direction of the receptor, in one
// launch; the same results as the direct kernel with the same G); else
// k_forces.  Modes 2 and 4 use z-pair fetches.
static hipError_t launch_force_eval(pst_ctx* ctx, const ForceSys& S, const ForceWork& W,
                                    int precision, int G, int max_n, int cap, int mode) {
  ctx->last_G = G;
  if ((mode == 3 || mode == 4) && split_eligible(S, W, G)) {
    ctx->last_mode = mode;
    return launch_forces_split(S, W, precision, G, max_n, cap, ctx->stream, mode == 4);
  }
  if (mode == 3) mode = 0;
  if (mode == 4) mode = 2;
  if (mode == 1 && G >= 8 && bins_match(ctx, S) && binned_eligible(S, W, ctx->bins, max_n)) {
    ctx->last_mode = 1;
    return launch_forces_binned(S, W, ctx->bins, precision, G, max_n, cap, ctx->stream);
  }
  if (mode == 2) {
    ctx->last_mode = 2;
    return launch_forces(S, W, precision, G, max_n, cap, ctx->stream, true);
  }
  ctx->last_mode = 0;
  return launch_forces(S, W, precision, G, max_n, cap, ctx->stream);
}

// fixed_chunk != nullptr: the chunk-skip scratch is that buffer of
// fixed_cap doubles (nothing is allocated; too many chunks is an error) and
// the direct kernel runs whatever the context's gather mode.
static int forces_common(pst_ctx* ctx, ForceWork& W, int n, int precision, int fb,
                         double* fixed_chunk = nullptr, int fixed_cap = 0) {
  ForceSys S;
  fill_sys(ctx, S);
  if (W.rev_skip_mode == 1 && S.has_rev && n > 0) {
    int cs, cpf, nch;
    chunk_layout(ctx, n, fb, &cs, &cpf, &nch);
    double* buf = fixed_chunk;
    if (fixed_chunk) {
      if (nch > fixed_cap) {
        ctx->err = "pst_forces_positions_noalloc: " + std::to_string(n) + " trajectories need " +
                   std::to_string(nch) + " reverse-Born chunks, more than the " +
                   std::to_string(fixed_cap) + " the static scratch holds";
        return 1;
      }
    } else {
      if (ensure_chunk(ctx, nch)) return 2;
      buf = ctx->d_chunk;
    }
    CK(launch_chunk_min_r(W.pos, W.list, n, fb, cs, cpf, buf, nch, ctx->stream));
    W.chunk_min_r = buf;
    W.chunk_fb = fb;
    W.chunk_cs = cs;
    W.chunks_per_fb = cpf;
  }
  CK(launch_force_eval(ctx, S, W, precision, group_for(ctx, S), n, 2048,
                       fixed_chunk ? 0 : ctx->gather_mode));
  return 0;
}

int pst_last_force_eval(const pst_ctx* ctx, int* mode, int* G) {
  if (!ctx) return 1;
  if (mode) *mode = ctx->last_mode;
  if (G) *G = ctx->last_G;
  return 0;
}

int pst_set_gather_mode(pst_ctx* ctx, int mode, int max_traj) {
  if (!ctx) return 1;
  if (mode < 0 || mode > 4) FAIL("pst_set_gather_mode: mode must be 0 to 4");
  if (mode == 1 && max_traj <= 0) FAIL("pst_set_gather_mode: max_traj must be positive");
  CK(set_dev(ctx));
  ctx->gather_mode = mode;
  if (mode == 1) return ensure_bins(ctx, max_traj);
  return 0;
}

int pst_forces_positions(pst_ctx* ctx, int n, int m, const double* lig_pos, const double* lig_q,
                         const double* rot, const double* cen, double* F, double* T, double* E,
                         int precision, int rev_skip) {
  if (!ctx) return 1;
  if (n < 0 || m <= 0) FAIL("pst_forces_positions: bad sizes");
  if (n == 0) return 0;
  if (!lig_pos || !lig_q || !F) FAIL("pst_forces_positions: null pointer");
  if ((rot == nullptr) != (cen == nullptr)) FAIL("pst_forces_positions: rot and cen go together");
  if (precision != 0 && precision != 1) FAIL("pst_forces_positions: precision must be 0 or 1");
  CK(set_dev(ctx));
  const int saved_M = ctx->M;
  ForceWork W;
  memset(&W, 0, sizeof(W));
  W.n = n;
  W.pos = cen;
  W.rot = rot;
  W.lig_pos = lig_pos;
  W.lig_q = lig_q;
  W.F = F; W.T = T; W.E = E;
  W.rev_skip_mode = rev_skip;
  ctx->M = m;
  int rc = forces_common(ctx, W, n, precision, n);
  ctx->M = saved_M;
  return rc;
}

// ---- force operator without allocation ---------------------------------------
// The reverse-Born chunk skip's per-chunk minimum centroid distance lives in a
// file-scope device array (one copy per GPU, held until the process ends).
// Host code uses the per-device address from hipGetSymbolAddress, never the
// array's name.  Every call on one GPU shares the array: calls that overlap
// on one GPU race.  k_chunk_min_r writes every chunk it later reads, so the
// array needs no clearing between calls.
constexpr int kNoallocChunks = 65536;  // 512 KiB
static __device__ __align__(256) double scratch[kNoallocChunks];
constexpr int kNoallocMaxDevices = 64;
static double* g_scratch_dev[kNoallocMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

// Address of the device's copy of `scratch`, looked up once per device with
// that device current (the caller has made it current).
static int noalloc_scratch(pst_ctx* ctx, double** out) {
  if (ctx->device < 0 || ctx->device >= kNoallocMaxDevices)
    FAIL("pst_forces_positions_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ctx->err = std::string("pst_forces_positions_noalloc: hipGetSymbolAddress(scratch): ") +
                 hipGetErrorString(e);
      return 2;
    }
    g_scratch_dev[ctx->device] = (double*)p;
  }
  *out = g_scratch_dev[ctx->device];
  return 0;
}

int pst_forces_positions_noalloc(pst_ctx* ctx, int n, int m, const double* lig_pos,
                                 const double* lig_q, const double* rot, const double* cen,
                                 double* F, double* T, double* E, int precision, int rev_skip) {
  if (!ctx) return 1;
  if (n < 0 || m <= 0) FAIL("pst_forces_positions_noalloc: bad sizes");
  if (n == 0) return 0;
  if (!lig_pos || !lig_q || !F) FAIL("pst_forces_positions_noalloc: null pointer");
  if ((rot == nullptr) != (cen == nullptr))
    FAIL("pst_forces_positions_noalloc: rot and cen go together");
  if (precision != 0 && precision != 1)
    FAIL("pst_forces_positions_noalloc: precision must be 0 or 1");
  CK(set_dev(ctx));
  double* chunk = nullptr;
  if (noalloc_scratch(ctx, &chunk)) return 2;
  const int saved_M = ctx->M;
  ForceWork W;
  memset(&W, 0, sizeof(W));
  W.n = n;
  W.pos = cen;
  W.rot = rot;
  W.lig_pos = lig_pos;
  W.lig_q = lig_q;
  W.F = F; W.T = T; W.E = E;
  W.rev_skip_mode = rev_skip;
  ctx->M = m;
  int rc = forces_common(ctx, W, n, precision, n, chunk, kNoallocChunks);
  ctx->M = saved_M;
  return rc;
}

int pst_forces_pose(pst_ctx* ctx, int n, const double* pos, const double* quat, double* F,
                    double* T, double* E, int precision) {
  if (!ctx) return 1;
  if (n < 0) FAIL("pst_forces_pose: bad size");
  if (n == 0) return 0;
  if (!pos || !quat || !F) FAIL("pst_forces_pose: null pointer");
  if (ctx->M <= 0) FAIL("pst_forces_pose: ligand not set");
  CK(set_dev(ctx));
  ForceWork W;
  memset(&W, 0, sizeof(W));
  W.n = n;
  W.pos = pos;
  W.quat = quat;
  W.F = F; W.T = T; W.E = E;
  return forces_common(ctx, W, n, precision, n);
}

// Whether a force evaluation fetches z-pairs / z-rows (modes 2 / 4).
static bool group_zp(const pst_ctx* ctx) { return ctx->gather_mode == 2 || ctx->gather_mode == 4; }
static bool sim_zp(const pst_sim* s) {
  const int mode = s->tune_mode >= 0 ? s->tune_mode : s->ctx->gather_mode;
  return mode == 2 || mode == 4;
}

int pst_forces_pose_group(pst_ctx* const* ctxs, int L, const int* n, const double* const* pos,
                          const double* const* quat, double* const* F, double* const* T,
                          double* const* E, int precision) {
  if (!ctxs || L <= 0 || !ctxs[0]) return 1;
  pst_ctx* ctx = ctxs[0];
  if (L > 64) FAIL("pst_forces_pose_group: at most 64 systems");
  if (!n || !pos || !quat || !F) FAIL("pst_forces_pose_group: null pointer");
  if (precision != 0 && precision != 1) FAIL("pst_forces_pose_group: precision must be 0 or 1");
  for (int l = 0; l < L; ++l) {
    if (!ctxs[l] || ctxs[l]->device != ctx->device)
      FAIL("pst_forces_pose_group: every context on the same device");
    if (ctxs[l]->M <= 0) FAIL("pst_forces_pose_group: ligand not set");
    if (n[l] < 0) FAIL("pst_forces_pose_group: bad size");
    if (n[l] > 0 && (!pos[l] || !quat[l] || !F[l])) FAIL("pst_forces_pose_group: null pointer");
  }
  CK(set_dev(ctx));
  for (int l = 1; l < L; ++l) CK(hipStreamSynchronize(ctxs[l]->stream));
  // descriptors and the workgroup partition as the grouped step builds them
  std::vector<ForceSys> S(L);
  std::vector<ForceWork> W(L);
  std::vector<int> start(L + 1, 0);
  int G = 16;
  size_t lds = 0;
  for (int l = 0; l < L; ++l) {
    fill_sys(ctxs[l], S[l]);
    memset(&W[l], 0, sizeof(ForceWork));
    W[l].n = n[l];
    W[l].pos = pos[l];
    W[l].quat = quat[l];
    W[l].F = F[l];
    W[l].T = T ? T[l] : nullptr;
    W[l].E = E ? E[l] : nullptr;
    W[l].xcd_chunks = 1;
    G = std::max(G, group_for(ctxs[l], S[l]));
    if (S[l].M <= 1536) lds = std::max(lds, (size_t)S[l].M * (precision == 0 ? 32 : 16));
  }
  G = G >= 64 ? 64 : (G >= 32 ? 32 : 16);
  const int gpb = 256 / G;
  for (int l = 0; l < L; ++l) {
    int b = (n[l] + gpb - 1) / gpb;
    if (b > 2048) b = 2048;
    b = (b + 7) / 8 * 8;
    start[l + 1] = start[l] + b;
  }
  ForceSys* dS = nullptr;
  ForceWork* dW = nullptr;
  int* dstart = nullptr;
  auto release = [&]() {
    if (dS) hipFree(dS);
    if (dW) hipFree(dW);
    if (dstart) hipFree(dstart);
  };
  hipError_t e = hipMalloc(&dS, sizeof(ForceSys) * L);
  if (e == hipSuccess) e = hipMalloc(&dW, sizeof(ForceWork) * L);
  if (e == hipSuccess) e = hipMalloc(&dstart, sizeof(int) * (L + 1));
  if (e == hipSuccess) e = hipMemcpy(dS, S.data(), sizeof(ForceSys) * L, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dW, W.data(), sizeof(ForceWork) * L, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dstart, start.data(), sizeof(int) * (L + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_forces_group(dS, dW, dstart, L, start[L], lds, precision, G, group_zp(ctx),
                            ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release();
  if (e != hipSuccess) {
    ctx->err = std::string("pst_forces_pose_group: ") + hipGetErrorString(e);
    return 2;
  }
  for (int l = 0; l < L; ++l) {
    ctxs[l]->last_mode = 0;
    ctxs[l]->last_G = G;
  }
  return 0;
}

int pst_time_forces_pose(pst_ctx* ctx, int n, const double* pos, const double* quat, double* F,
                         double* T, double* E, int precision, int reps, double* ms_per_launch,
                         double* census4) {
  if (!ctx) return 1;
  if (n <= 0 || reps <= 0 || !ms_per_launch) FAIL("pst_time_forces_pose: bad arguments");
  CK(set_dev(ctx));
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  int rc = pst_forces_pose(ctx, n, pos, quat, F, T, E, precision);  // warm-up
  if (rc) return rc;
  CK(hipEventRecord(a, ctx->stream));
  for (int i = 0; i < reps; ++i) {
    rc = pst_forces_pose(ctx, n, pos, quat, F, T, E, precision);
    if (rc) return rc;
  }
  CK(hipEventRecord(b, ctx->stream));
  CK(hipEventSynchronize(b));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, a, b));
  *ms_per_launch = ms / reps;
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (census4) {
    // atom census on the host from the current poses (one launch's work)
    std::vector<double> hp(3 * (size_t)n), hq(4 * (size_t)n), body(3 * (size_t)ctx->M),
        lq(ctx->M), rp(3 * (size_t)ctx->n_rec), rq(ctx->n_rec);
    CK(hipMemcpy(hp.data(), pos, hp.size() * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(hq.data(), quat, hq.size() * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(body.data(), ctx->d_body, body.size() * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(lq.data(), ctx->d_q, lq.size() * 8, hipMemcpyDeviceToHost));
    if (ctx->n_rec) {
      CK(hipMemcpy(rp.data(), ctx->d_rec, rp.size() * 8, hipMemcpyDeviceToHost));
      CK(hipMemcpy(rq.data(), ctx->d_rec_q, rq.size() * 8, hipMemcpyDeviceToHost));
    }
    ForceSys S;
    fill_sys(ctx, S);
    double ne = 0, nb = 0, nr = 0;
    auto inside = [](const GridDev& g, double x, double y, double z) {
      return x > g.lo[0] && y > g.lo[1] && z > g.lo[2] && x < g.hi[0] && y < g.hi[1] && z < g.hi[2];
    };
    for (int t = 0; t < n; ++t) {
      const double* q = &hq[4 * (size_t)t];
      const double w = q[0], x = q[1], y = q[2], z = q[3];
      const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                           2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                           2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
      const double cx = hp[3 * t], cy = hp[3 * t + 1], cz = hp[3 * t + 2];
      for (int i = 0; i < ctx->M; ++i) {
        if (fabs(lq[i]) < 1e-9) continue;
        const double bx = body[3 * i], by = body[3 * i + 1], bz = body[3 * i + 2];
        const double px = R[0] * bx + R[1] * by + R[2] * bz + cx;
        const double py = R[3] * bx + R[4] * by + R[5] * bz + cy;
        const double pz = R[6] * bx + R[7] * by + R[8] * bz + cz;
        for (int g = 0; g < S.n_elec; ++g)
          if (inside(S.elec[g], px, py, pz)) { ne += 1; break; }
        for (int g = 0; g < S.n_born; ++g)
          if (inside(S.born[g], px, py, pz)) { nb += 1; break; }
      }
      if (S.has_rev) {
        const double rc = sqrt(cx * cx + cy * cy + cz * cz);
        if (rc - S.rec_max_r > S.rev_cull) continue;
        for (int j = 0; j < ctx->n_rec; ++j) {
          if (fabs(rq[j]) < 1e-9) continue;
          const double dx = rp[3 * j] - cx, dy = rp[3 * j + 1] - cy, dz = rp[3 * j + 2] - cz;
          const double xx = R[0] * dx + R[3] * dy + R[6] * dz;
          const double yy = R[1] * dx + R[4] * dy + R[7] * dz;
          const double zz = R[2] * dx + R[5] * dy + R[8] * dz;
          if (inside(S.lig_born, xx, yy, zz)) nr += 1;
        }
      }
    }
    census4[0] = n; census4[1] = ne; census4[2] = nb; census4[3] = nr;
  }
  return 0;
}

int pst_gather_ceiling(pst_ctx* ctx, int mode, long long n_points, int reps, double* out4) {
  if (!ctx) return 1;
  if (!ctx->pair.mem) FAIL("pst_gather_ceiling: needs the FP32 potential + Born pair grid");
  const int variant = mode >> 5;
  mode &= 31;
  if (mode > 16 || variant > 2) FAIL("pst_gather_ceiling: mode must be 0..16 plus 32 x variant 0..2");
  if (n_points <= 0 || n_points > (1ll << 30) || reps <= 0 || !out4)
    FAIL("pst_gather_ceiling: bad arguments");
  CK(set_dev(ctx));
  float* out = nullptr;
  CK(hipMalloc(&out, (size_t)n_points * 4));
  EventPair ev;
  hipError_t e = hipEventCreate(&ev.a);
  if (e == hipSuccess) e = hipEventCreate(&ev.b);
  const long long nb = (n_points + 255) / 256;
  const int blocks = (int)(nb < 65536 ? nb : 65536);
  if (e == hipSuccess) e = launch_gather_probe(ctx->pair.g, n_points, mode, variant, 12345, out,
                                               blocks, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev.a, ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = launch_gather_probe(ctx->pair.g, n_points, mode, variant, 777 + r, out, blocks,
                            ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev.b, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev.b);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev.a, ev.b);
  hipFree(out);
  if (e != hipSuccess) {
    ctx->err = std::string("pst_gather_ceiling: ") + hipGetErrorString(e);
    return 2;
  }
  const double per = ms / reps;
  out4[0] = per;
  out4[1] = 160.0 * (double)n_points / (per * 1e-3) / 1e9;
  out4[2] = (double)n_points;
  out4[3] = 160.0;
  return 0;
}

// ----------------------------------------------------------------------------
// two-body RPY mobility
// ----------------------------------------------------------------------------
int pst_rpy12_eval(pst_ctx* ctx, int n, const double* pos, double a_rec, double a_lig,
                   double* mobility, double* factor) {
  if (!ctx) return 1;
  if (n < 0 || !(a_rec > 0) || !(a_lig > 0)) FAIL("pst_rpy12_eval: bad arguments");
  if (n == 0) return 0;
  if (!pos || (!mobility && !factor)) FAIL("pst_rpy12_eval: null pointer");
  CK(set_dev(ctx));
  CK(launch_rpy_eval(n, pos, a_rec, a_lig, mobility, factor, ctx->stream));
  return 0;
}

int pst_rpy12_move(pst_ctx* ctx, int n, double a_rec, double a_lig, double kT_over_eta,
                   int torque_mode, const double* dt, double* pos, double* quat, const double* F,
                   const double* T, const double* xi, double* dX) {
  if (!ctx) return 1;
  if (n < 0 || !(a_rec > 0) || !(a_lig > 0) || !(kT_over_eta > 0))
    FAIL("pst_rpy12_move: bad arguments");
  if (torque_mode != 0 && torque_mode != 1) FAIL("pst_rpy12_move: torque_mode must be 0 or 1");
  if (n == 0) return 0;
  if (!dt || !pos || !quat || !F || !T || !xi) FAIL("pst_rpy12_move: null pointer");
  CK(set_dev(ctx));
  CK(launch_rpy_move(n, a_rec, a_lig, kT_over_eta, torque_mode, dt, pos, quat, F, T, xi, dX,
                     ctx->stream));
  return 0;
}

// ----------------------------------------------------------------------------
// simulator
// ----------------------------------------------------------------------------
int pst_sim_create(pst_ctx* ctx, const void* params, int precision, pst_sim** out) {
  if (!ctx) return 1;
  if (!params || !out) FAIL("pst_sim_create: null pointer");
  if (precision != 0 && precision != 1) FAIL("pst_sim_create: precision must be 0 or 1");
  const StepParams* P = (const StepParams*)params;
  if (P->n_slots <= 0 || P->n_slots > (1ll << 30)) FAIL("pst_sim_create: n_slots out of range");
  if (P->n_pairs < 0 || P->n_pairs > PST_MAX_PAIRS) FAIL("pst_sim_create: too many reaction pairs");
  if (ctx->M <= 0) FAIL("pst_sim_create: ligand not set");
  if (P->hydro_mode != 0 && P->hydro_mode != 1) FAIL("pst_sim_create: hydro_mode must be 0 or 1");
  if (P->hydro_mode == 1 && !(P->a_rec > 0 && P->a_lig > 0 && P->kT_over_eta > 0))
    FAIL("pst_sim_create: hydro_mode 1 needs a_rec, a_lig and kT_over_eta > 0");
  if (P->lmz_on &&
      !(P->lmz_b > 0 && P->lmz_q > P->lmz_trigger && P->lmz_trigger >= P->lmz_b &&
        P->lmz_b_cover > P->lmz_b && P->lmz_q_cover < P->lmz_q && P->lmz_L > 0 && P->lmz_Dt > 0 &&
        (!P->lmz_hi || (P->lmz_dpre > 0 && P->lmz_ainv > 0))))
    FAIL("pst_sim_create: inconsistent outer-propagator parameters (need 0 < b <= trigger < q, "
         "covers inside (b, q), Debye length and D_t > 0)");
  if (P->sm_active) {
    for (int r = 0; r < P->n_pairs; ++r)
      if (P->sm_before[r] < 0 || P->sm_before[r] >= 32 || P->sm_after[r] < -1 || P->sm_after[r] >= 32)
        FAIL("pst_sim_create: state-machine state indices must be in [0, 32)");
  }
  CK(set_dev(ctx));
  pst_sim* s = new pst_sim();
  s->ctx = ctx;
  s->P = *P;
  s->precision = precision;
  s->S = (int)P->n_slots;
  const size_t S = s->S;
  *out = s;
  CK(hipMalloc(&s->pos, S * 3 * 8));
  CK(hipMalloc(&s->quat, S * 4 * 8));
  CK(hipMalloc(&s->status, S * 4));
  CK(hipMalloc(&s->F, S * 3 * 8));
  CK(hipMalloc(&s->T, S * 3 * 8));
  CK(hipMalloc(&s->dt, S * 8));
  CK(hipMalloc(&s->x0, S * PST_MAX_PAIRS * 8));
  CK(hipMalloc(&s->x1, S * PST_MAX_PAIRS * 8));
  CK(hipMemsetAsync(s->F, 0, S * 3 * 8, ctx->stream));
  CK(hipMemsetAsync(s->T, 0, S * 3 * 8, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pst_sim_destroy(pst_sim* s) {
  if (!s) return 0;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  void* ptrs[] = {s->pos, s->quat, s->status, s->F, s->T, s->dt, s->x0, s->x1,
                  s->n_steps64, s->flags, s->lists, s->counts, s->h_buf, s->n_steps32,
                  s->traj, s->plain_list, s->mlp_list, s->ctr, s->out_status, s->out_steps,
                  s->audit_F, s->audit_T, s->audit_list, s->audit_count, s->key, s->bins,
                  s->census, s->sm_state, s->sm_mask, s->lmz_escaped, s->rev_list,
                  s->F_rev, s->T_rev};
  if (s->g_exec) hipGraphExecDestroy(s->g_exec);
  if (s->g_graph) hipGraphDestroy(s->g_graph);
  if (s->g_cap) hipStreamDestroy(s->g_cap);
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (void* p : s->obs_allocs) hipFree(p);
  if (s->aux) { hipStreamSynchronize(s->aux); hipStreamDestroy(s->aux); }
  if (s->fork_ev) hipEventDestroy(s->fork_ev);
  if (s->join_ev) hipEventDestroy(s->join_ev);
  delete s;
  return 0;
}

static int ensure_replay(pst_sim* s) {
  pst_ctx* ctx = s->ctx;
  if (s->flags) return 0;
  const size_t S = s->S;
  CK(hipMalloc(&s->n_steps64, S * 8));
  CK(hipMalloc(&s->flags, S * 4));
  CK(hipMalloc(&s->lists, S * 4 * 4));
  CK(hipMalloc(&s->counts, 8 * 4));
  CK(hipMemsetAsync(s->n_steps64, 0, S * 8, ctx->stream));
  CK(hipMemsetAsync(s->counts, 0, 8 * 4, ctx->stream));
  return 0;
}

static ReplayBufs replay_bufs(pst_sim* s) {
  ReplayBufs B;
  const size_t S = s->S;
  B.pos = s->pos; B.quat = s->quat; B.status = s->status; B.n_steps = s->n_steps64;
  B.f_run = s->flags; B.f_esc = s->flags + S; B.f_ret = s->flags + 2 * S; B.f_sr = s->flags + 3 * S;
  B.run_list = s->lists; B.esc_list = s->lists + S; B.ret_list = s->lists + 2 * S;
  B.sr_list = s->lists + 3 * S;
  B.counts = s->counts;
  B.F = s->F; B.T = s->T; B.dt = s->dt; B.x0 = s->x0; B.x1 = s->x1;
  B.old_pos = nullptr; B.old_quat = nullptr;
  B.n_slots = s->S;
  return B;
}

static int stage(pst_sim* s, const double* host, size_t count, double** dev) {
  pst_ctx* ctx = s->ctx;
  if (count == 0) { *dev = s->h_buf; return 0; }
  if (count > s->h_buf_cap) {
    if (s->h_buf) hipFree(s->h_buf);
    s->h_buf = nullptr;
    s->h_buf_cap = 0;
    CK(hipMalloc(&s->h_buf, count * 8));
    s->h_buf_cap = count;
  }
  CK(hipMemcpyAsync(s->h_buf, host, count * 8, hipMemcpyHostToDevice, ctx->stream));
  *dev = s->h_buf;
  return 0;
}

int pst_sim_set_state(pst_sim* s, int n, const double* pos, const double* quat) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (n != s->S || !pos || !quat) FAIL("pst_sim_set_state: n must equal n_slots");
  CK(set_dev(ctx));
  if (ensure_replay(s)) return 2;
  CK(hipMemcpyAsync(s->pos, pos, (size_t)n * 24, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(s->quat, quat, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemsetAsync(s->status, 0, (size_t)n * 4, ctx->stream));
  CK(hipMemsetAsync(s->n_steps64, 0, (size_t)n * 8, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pst_sim_get_state(pst_sim* s, double* pos, double* quat, int* status, long long* n_steps) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  const size_t S = s->S;
  if (pos) CK(hipMemcpy(pos, s->pos, S * 24, hipMemcpyDeviceToHost));
  if (quat) CK(hipMemcpy(quat, s->quat, S * 32, hipMemcpyDeviceToHost));
  if (status) CK(hipMemcpy(status, s->status, S * 4, hipMemcpyDeviceToHost));
  if (n_steps) {
    if (s->n_steps64) {
      CK(hipMemcpy(n_steps, s->n_steps64, S * 8, hipMemcpyDeviceToHost));
    } else if (s->n_steps32) {
      std::vector<unsigned int> tmp(S);
      CK(hipMemcpy(tmp.data(), s->n_steps32, S * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < S; ++i) n_steps[i] = tmp[i];
    }
  }
  return 0;
}

int pst_replay_begin(pst_sim* s, int* n_run, int* n_esc, int* n_react) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (s->P.hydro_mode != 0) FAIL("pst_replay_begin: replay follows the scalar diffusion (hydro_mode 0)");
  if (s->P.sm_active) FAIL("pst_replay_begin: state-machine reactions run in production mode");
  if (s->P.lmz_on) FAIL("pst_replay_begin: the outer propagator runs in production mode");
  CK(set_dev(ctx));
  if (ensure_replay(s)) return 2;
  ReplayBufs B = replay_bufs(s);
  CK(replay_begin(s->P, B, ctx->stream));
  int h[8];
  CK(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (n_run) *n_run = h[0];
  if (n_esc) *n_esc = h[1];
  if (n_react) *n_react = h[7];
  return 0;
}

int pst_replay_escape(pst_sim* s, const double* u_outer, int n_esc, const double* u_quat, int n_ret,
                      int* n_sr) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  ReplayBufs B = replay_bufs(s);
  int h[8];
  CK(hipMemcpy(h, s->counts, sizeof(h), hipMemcpyDeviceToHost));
  if (h[1] != n_esc) FAIL("pst_replay_escape: n_esc does not match the device escape list");
  if (n_esc > 0) {
    if (!u_outer) FAIL("pst_replay_escape: u_outer is null");
    double* du;
    if (stage(s, u_outer, (size_t)n_esc, &du)) return 2;
    CK(replay_escape_decide(s->P, B, du, ctx->stream));
    CK(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (h[2] != n_ret) FAIL("pst_replay_escape: n_ret does not match u_outer < return_prob");
    if (n_ret > 0) {
      if (!u_quat) FAIL("pst_replay_escape: u_quat is null");
      double* dq;
      if (stage(s, u_quat, (size_t)n_ret * 3, &dq)) return 2;
      CK(replay_escape_apply(s->P, B, dq, ctx->stream));
    }
  } else if (n_ret != 0) {
    FAIL("pst_replay_escape: n_ret must be 0 when n_esc is 0");
  } else {
    CK(hipMemsetAsync(s->counts + 2, 0, 4, ctx->stream));
  }
  CK(replay_build_sr(B, ctx->stream));
  CK(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (n_sr) *n_sr = h[3];
  return 0;
}

int pst_replay_move(pst_sim* s, const double* noise_t, const double* noise_r, int n_sr,
                    int* any_above, int* n_overlap) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  ReplayBufs B = replay_bufs(s);
  int h[8];
  CK(hipMemcpy(h, s->counts, sizeof(h), hipMemcpyDeviceToHost));
  if (h[3] != n_sr) FAIL("pst_replay_move: n_sr does not match the device list");
  if (n_sr > 0) {
    if (!noise_t || !noise_r) FAIL("pst_replay_move: null noise");
    // forces on the ordered still-running list, reference batching
    ForceWork W;
    memset(&W, 0, sizeof(W));
    W.list = B.sr_list;
    W.n = n_sr;
    W.pos = s->pos;
    W.quat = s->quat;
    W.F = s->F; W.T = s->T; W.E = nullptr;
    W.rev_skip_mode = 1;
    const int M = ctx->M;
    int fb = (int)fmax(1000.0, floor(4.0 * 1024 * 1024 * 1024 / fmax(1.0, (double)M * 150)));
    if (fb > n_sr) fb = n_sr;
    int rc = forces_common(ctx, W, n_sr, s->precision, fb);
    if (rc) return rc;
    // host noise staged after the forces, in one buffer
    std::vector<double> buf((size_t)n_sr * 6);
    memcpy(buf.data(), noise_t, (size_t)n_sr * 24);
    memcpy(buf.data() + (size_t)n_sr * 3, noise_r, (size_t)n_sr * 24);
    double* dn;
    if (stage(s, buf.data(), buf.size(), &dn)) return 2;
    ForceSys S;
    fill_sys(ctx, S);
    CK(hipMemsetAsync(s->counts + 4, 0, 4, ctx->stream));
    CK(hipMemsetAsync(s->counts + 6, 0, 4, ctx->stream));
    CK(replay_move(s->P, S, B, dn, dn + (size_t)n_sr * 3, n_sr, ctx->stream));
  }
  CK(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (any_above) *any_above = n_sr > 0 ? h[4] : 0;
  if (n_overlap) *n_overlap = n_sr > 0 ? h[6] : 0;
  return 0;
}

int pst_replay_bridge(pst_sim* s, const double* u_bb, int n_sr, int* n_bridge) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  ReplayBufs B = replay_bufs(s);
  CK(hipMemsetAsync(s->counts + 5, 0, 4, ctx->stream));
  double* du = nullptr;
  if (u_bb && n_sr > 0 && s->P.n_pairs > 0) {
    if (stage(s, u_bb, (size_t)n_sr * s->P.n_pairs, &du)) return 2;
  }
  CK(replay_bridge(s->P, B, du, n_sr, ctx->stream));
  int h[8];
  CK(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (n_bridge) *n_bridge = h[5];
  return 0;
}

int pst_replay_forces(pst_sim* s, double* F_host, int n_sr) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  if (!s->lists) FAIL("pst_replay_forces: no replay state");
  std::vector<int> idx(n_sr);
  std::vector<double> all((size_t)s->S * 3);
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(idx.data(), s->lists + 3 * (size_t)s->S, (size_t)n_sr * 4, hipMemcpyDeviceToHost));
  CK(hipMemcpy(all.data(), s->F, all.size() * 8, hipMemcpyDeviceToHost));
  for (int w = 0; w < n_sr; ++w)
    for (int d = 0; d < 3; ++d) F_host[3 * w + d] = all[3 * (size_t)idx[w] + d];
  return 0;
}

// ---- production ------------------------------------------------------------
// Whether the sim's force evaluation is the split one (modes 3 / 4): its
// pre-phase then sorts the ligand-atom list and the reverse-Born order by
// keys of their own.
static bool sim_split(const pst_sim* s) {
  const int mode = s->tune_mode >= 0 ? s->tune_mode : s->ctx->gather_mode;
  return mode == 3 || mode == 4;
}

static ProdBufs prod_bufs(pst_sim* s) {
  ProdBufs B;
  B.pos = s->pos; B.quat = s->quat; B.status = s->status; B.n_steps = s->n_steps32;
  B.traj = s->traj; B.plain_list = s->plain_list; B.mlp_list = s->mlp_list;
  B.key = s->key; B.bins = s->bins; B.rev_list = sim_split(s) ? s->rev_list : nullptr;
  B.F = s->F; B.T = s->T; B.dt = s->dt; B.x0 = s->x0; B.x1 = s->x1;
  B.ctr = s->ctr; B.out_status = s->out_status; B.out_steps = s->out_steps;
  B.obs = s->obs;
  B.sm_state = s->sm_state;
  B.sm_mask = s->sm_mask;
  B.lmz_escaped = s->lmz_escaped;
  B.n_slots = s->S;
  return B;
}

// Bin buffers for the sim's slots when its force evaluation is binned.
static int sim_ensure_bins(pst_sim* s) {
  const int mode = s->tune_mode >= 0 ? s->tune_mode : s->ctx->gather_mode;
  return mode == 1 ? ensure_bins(s->ctx, s->S) : 0;
}

int pst_prod_init(pst_sim* s, int record) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  const size_t S = s->S;
  if (!s->n_steps32) {
    CK(hipMalloc(&s->n_steps32, S * 4));
    CK(hipMalloc(&s->traj, S * 8));
    CK(hipMalloc(&s->plain_list, S * 4));
    CK(hipMalloc(&s->rev_list, S * 4));
    CK(hipMalloc(&s->F_rev, S * 3 * 8));
    CK(hipMalloc(&s->T_rev, S * 3 * 8));
    CK(hipMalloc(&s->mlp_list, S * 4));
    CK(hipMalloc(&s->ctr, sizeof(RunCounters)));
    CK(hipMalloc(&s->key, S * 4));
    CK(hipMalloc(&s->bins, 1024 * 4));
    CK(hipMalloc(&s->sm_state, S * 4));
    CK(hipMalloc(&s->sm_mask, S * 4));
    CK(hipMalloc(&s->lmz_escaped, S * 4));
  }
  CK(hipMemsetAsync(s->bins, 0, 1024 * 4, ctx->stream));
  CK(hipMemsetAsync(s->lmz_escaped, 0, S * 4, ctx->stream));
  if (s->out_status) { hipFree(s->out_status); s->out_status = nullptr; }
  if (s->out_steps) { hipFree(s->out_steps); s->out_steps = nullptr; }
  for (void* p : s->obs_allocs) hipFree(p);
  s->obs_allocs.clear();
  memset(&s->obs, 0, sizeof(ObsBufs));
  if (record >= 2) {
    const size_t N = (size_t)s->P.n_total;
    auto get = [&](size_t bytes) -> void* {
      void* p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
      s->obs_allocs.push_back(p);
      return p;
    };
    ObsBufs& O = s->obs;
    O.s_min = (double*)get(S * 8);
    O.s_step = (int*)get(S * 4);
    O.s_time = (double*)get(S * 8);
    O.s_ret = (int*)get(S * 4);
    O.s_prev_r = (double*)get(S * 8);
    O.s_nm = (double*)get(S * 7 * 8);
    O.traj = (double*)get(N * kObsTraj * 8);
    O.hist = (unsigned long long*)get(kObsHist * 8);
    if (!O.s_min || !O.s_step || !O.s_time || !O.s_ret || !O.s_prev_r || !O.s_nm || !O.traj ||
        !O.hist) {
      for (void* p : s->obs_allocs) hipFree(p);
      s->obs_allocs.clear();
      memset(&s->obs, 0, sizeof(ObsBufs));
      FAIL("pst_prod_init: out of device memory for observables");
    }
    CK(hipMemsetAsync(O.traj, 0, N * kObsTraj * 8, ctx->stream));
    CK(hipMemsetAsync(O.hist, 0, kObsHist * 8, ctx->stream));
  }
  if (record) {
    const size_t N = (size_t)s->P.n_total;
    CK(hipMalloc(&s->out_status, N));
    CK(hipMalloc(&s->out_steps, N * 4));
    CK(hipMemsetAsync(s->out_status, 0, N, ctx->stream));
    CK(hipMemsetAsync(s->out_steps, 0, N * 4, ctx->stream));
  }
  CK(prod_init(s->P, prod_bufs(s), ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  s->prod_ready = true;
  return sim_ensure_bins(s);
}

// The main force list of a production step: still-running grid-path slots.
static ForceWork main_work(pst_sim* s) {
  ForceWork W;
  memset(&W, 0, sizeof(W));
  W.list = s->plain_list;
  W.rev_order = s->rev_list;
  W.F_rev = s->F_rev;
  W.T_rev = s->T_rev;
  W.n_dev = &s->ctr->n_sr;
  W.xcd_chunks = s->tune_xcd;
  W.pos = s->pos;
  W.quat = s->quat;
  W.F = s->F; W.T = s->T; W.E = nullptr;
  W.rev_skip_mode = 0;
  return W;
}

// Step phases before the force evaluation: reaction / escape / refill and
// the grid-path list sorted by direction.
static int enqueue_pre(pst_sim* s, int par) {
  pst_ctx* ctx = s->ctx;
  ProdBufs B = prod_bufs(s);
  CK(prod_step_pre(s->P, B, par, ctx->stream));
  CK(prod_step_sort(B, ctx->stream));
  return 0;
}

// Surrogate forces of the step and their audit (after the grid-path forces).
static bool audits(const pst_sim* s) {
  return s->mlp && s->P.surrogate_r > 0 && s->sp.audit_n > 0;
}

// The plain path checks the surrogate on audit_n surrogate trajectories per
// step, picked by surr_audit_pick (a stride through the surrogate list from an
// offset that moves with the step counter): their grid-path forces and
// torques into audit_F / audit_T.
static hipError_t launch_audit_forces(pst_sim* s, const ForceSys& S, const ForceWork& W,
                                      hipStream_t st) {
  hipError_t e = surr_audit_pick(s->mlp_list, &s->ctr->n_mlp, s->sp.audit_n, s->ctr,
                                 s->audit_list, s->audit_count, st);
  if (e != hipSuccess) return e;
  ForceWork A = W;
  A.list = s->audit_list;
  A.n_dev = s->audit_count;
  A.n_cap = s->sp.audit_n;
  A.out_compact = 1;
  A.xcd_chunks = 0;
  A.F = s->audit_F;
  A.T = s->audit_T;
  return launch_forces(S, A, s->precision, group_for(s->ctx, S), s->sp.audit_n, 2048, st);
}

// Whether the step's surrogate work (the network's forward for the far
// trajectories and the audit's grid-path launch) runs on the second stream.
static bool forks(const pst_sim* s) {
  return s->mlp && s->P.surrogate_r > 0;
}

// Fork: the surrogate trajectories' work on the second stream, ordered after
// everything enqueued so far on the context stream (captured graphs take it
// as a parallel branch beside the grid-path force launch).  It touches only
// the surrogate list's slots of F / T, the network's buffers and the audit
// buffers, none of which the grid-path launch reads or writes.
static int fork_surr(pst_sim* s, const ForceSys& S, const ForceWork& W) {
  pst_ctx* ctx = s->ctx;
  if (!s->aux) {
    CK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
    CK(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
  }
  CK(hipEventRecord(s->fork_ev, ctx->stream));
  CK(hipStreamWaitEvent(s->aux, s->fork_ev, 0));
  if (audits(s)) CK(launch_audit_forces(s, S, W, s->aux));
  CK(surr_step(s->sp, s->mlp->m, s->pos, s->quat, s->mlp_list, &s->ctr->n_mlp, s->S, s->F, s->T,
               s->aux));
  // the audit compares the network's forces with the audit's grid-path
  // forces, both produced on this stream
  if (audits(s))
    CK(surr_audit(s->audit_list, s->audit_count, s->sp.audit_n, s->F, s->T, s->audit_F,
                  s->sp.use_torque ? s->audit_T : nullptr, s->ctr, s->aux));
  CK(hipEventRecord(s->join_ev, s->aux));
  return 0;
}

// Surrogate forces of the step and their audit (after the grid-path forces).
// forked: fork_surr issued the forward, the audit's grid-path launch and the
// audit comparison.
static int enqueue_surr(pst_sim* s, const ForceSys& S, const ForceWork& W, bool forked = false) {
  pst_ctx* ctx = s->ctx;
  if (s->mlp && s->P.surrogate_r > 0) {
    if (forked) {
      CK(hipStreamWaitEvent(ctx->stream, s->join_ev, 0));
    } else {
      CK(surr_step(s->sp, s->mlp->m, s->pos, s->quat, s->mlp_list, &s->ctr->n_mlp, s->S, s->F,
                   s->T, ctx->stream));
      if (s->sp.audit_n > 0) {
        CK(launch_audit_forces(s, S, W, ctx->stream));
        CK(surr_audit(s->audit_list, s->audit_count, s->sp.audit_n, s->F, s->T, s->audit_F,
                      s->sp.use_torque ? s->audit_T : nullptr, s->ctr, ctx->stream));
      }
    }
  }
  return 0;
}

// Step phases after the grid-path forces: surrogate forces and their audit,
// the move, the Brownian bridge.
static int enqueue_post(pst_sim* s, const ForceSys& S, const ForceWork& W, int par,
                        bool forked = false) {
  pst_ctx* ctx = s->ctx;
  ProdBufs B = prod_bufs(s);
  const int r = enqueue_surr(s, S, W, forked);
  if (r) return r;
  CK(prod_step_post(s->P, S, B, par, ctx->stream));
  CK(prod_step_bridge(s->P, B, par, ctx->stream));
  return 0;
}

static int enqueue_step(pst_sim* s, int par) {
  pst_ctx* ctx = s->ctx;
  int rc = enqueue_pre(s, par);
  if (rc) return rc;
  ForceWork W = main_work(s);
  ForceSys S;
  fill_sys(ctx, S);
  const int G = s->tune_G > 0 ? s->tune_G : group_for(ctx, S);
  const bool forked = forks(s);
  if (forked) {
    const int r = fork_surr(s, S, W);
    if (r) return r;
  }
  if (s->census_on) CK(launch_census(S, W, s->census, ctx->stream));
  if (s->tev0) CK(hipEventRecord(s->tev0, ctx->stream));
  const int mode = s->tune_mode >= 0 ? s->tune_mode : ctx->gather_mode;
  CK(launch_force_eval(ctx, S, W, s->precision, G, s->S, s->tune_cap, mode));
  if (s->tev1) CK(hipEventRecord(s->tev1, ctx->stream));
  return enqueue_post(s, S, W, par, forked);
}

// Everything a captured step graph depends on: the launch configuration,
// the step and surrogate parameters (kernel arguments by value), the system
// record of the context and every buffer address a kernel argument holds.
static void graph_key(pst_sim* s, int gs, std::vector<unsigned char>& key) {
  pst_ctx* ctx = s->ctx;
  key.clear();
  auto put = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    key.insert(key.end(), b, b + n);
  };
  const int cfg[9] = {gs, s->tune_G, s->tune_cap, s->tune_xcd, s->tune_mode, ctx->gather_mode,
                      (int)s->census_on, s->precision, s->audit_cap};
  put(cfg, sizeof(cfg));
  put(&s->P, sizeof(StepParams));
  ForceSys S;
  fill_sys(ctx, S);
  put(&S, sizeof(S));
  const void* ptrs[] = {s->mlp, s->mlp ? (const void*)s->mlp->m.x0 : nullptr, s->audit_F,
                        s->audit_T, s->audit_list, s->audit_count, s->out_status, s->out_steps,
                        s->mlp ? (const void*)s->mlp->m.y : nullptr, s->obs.traj, s->obs.hist,
                        s->obs.s_min, s->obs.s_step, s->obs.s_time, s->obs.s_ret,
                        s->obs.s_prev_r, s->obs.s_nm, s->aux, ctx->stream,
                        ctx->bins.hist, ctx->bins.items, ctx->bins.sorted, ctx->bins.out,
                        ctx->bins.out_rev, s->tev0, s->tev1, s->rev_list, s->F_rev, s->T_rev};
  put(ptrs, sizeof(ptrs));
  if (s->mlp) {
    put(&s->sp, sizeof(SurrParams));
    put(&s->mlp->m, sizeof(MlpBufs));
  }
  const int binsz[3] = {ctx->bins.max_traj, ctx->bins.max_m, ctx->bins.max_rec};
  put(binsz, sizeof(binsz));
}

static void drop_graph(pst_sim* s) {
  if (s->g_exec) hipGraphExecDestroy(s->g_exec);
  if (s->g_graph) hipGraphDestroy(s->g_graph);
  s->g_exec = nullptr;
  s->g_graph = nullptr;
  s->g_key.clear();
}

// The step graph of gs steps: re-used while its key is unchanged, captured
// again (on the sim's private capture stream) otherwise.
static int step_graph(pst_sim* s, int gs) {
  pst_ctx* ctx = s->ctx;
  std::vector<unsigned char> key;
  graph_key(s, gs, key);
  if (s->g_exec && key == s->g_key) return 0;
  drop_graph(s);
  if (!s->g_cap) CK(hipStreamCreateWithFlags(&s->g_cap, hipStreamNonBlocking));
  CK(hipStreamBeginCapture(s->g_cap, hipStreamCaptureModeThreadLocal));
  hipStream_t saved = ctx->stream;
  ctx->stream = s->g_cap;
  int rc = 0;
  for (int i = 0; i < gs && !rc; ++i) rc = enqueue_step(s, i & 1);
  ctx->stream = saved;
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(s->g_cap, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) {
    if (g) hipGraphDestroy(g);
    ctx->err = std::string("pst_prod_run: hipStreamEndCapture: ") + hipGetErrorString(ee);
    return 2;
  }
  s->g_graph = g;
  const hipError_t ei = hipGraphInstantiate(&s->g_exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    s->g_exec = nullptr;
    drop_graph(s);
    ctx->err = std::string("pst_prod_run: hipGraphInstantiate: ") + hipGetErrorString(ei);
    return 2;
  }
  s->g_key.swap(key);
  return 0;
}

int pst_prod_run(pst_sim* s, long long max_steps, int graph_steps, double* c16) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_run: call pst_prod_init first");
  if (max_steps < 0 || graph_steps < 0) FAIL("pst_prod_run: bad arguments");
  CK(set_dev(ctx));
  if (sim_ensure_bins(s)) return 2;
  EventPair ev;
  CK(hipEventCreate(&ev.a));
  CK(hipEventCreate(&ev.b));
  long long steps = 0;
  RunCounters h;
  memset(&h, 0, sizeof(h));
  const int gs = graph_steps;
  if (gs > 0 && max_steps >= gs) {
    const int rc = step_graph(s, gs);
    if (rc) return rc;
  }
  CK(hipEventRecord(ev.a, ctx->stream));
  bool stop = false;
  while (steps < max_steps && !stop) {
    if (gs > 0 && steps + gs <= max_steps) {
      CK(hipGraphLaunch(s->g_exec, ctx->stream));
      steps += gs;
    } else {
      int rc = enqueue_step(s, (int)(steps & 1));
      if (rc) return rc;
      steps += 1;
    }
    CK(hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (h.stop || (h.n_sr + h.n_mlp) == 0) stop = true;
  }
  CK(hipEventRecord(ev.b, ctx->stream));
  CK(hipEventSynchronize(ev.b));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, ev.a, ev.b));
  s->last_ms_per_step = steps ? ms / steps : 0.0;
  if (c16) {
    double v[16] = {(double)h.n_reacted, (double)h.n_escaped, (double)h.n_maxsteps,
                    (double)h.total_steps, (double)h.next_traj, (double)h.n_returns,
                    (double)h.n_overlap, (double)h.n_bridge, (double)steps, (double)h.stop,
                    (double)h.n_sr, (double)h.n_mlp, h.audit_err2, h.audit_ref2,
                    (double)h.audit_n, 0};
    memcpy(c16, v, sizeof(v));
  }
  return 0;
}

// Autotuner of the production force launch: lanes per trajectory G (atoms
// per wave slice), the resident-workgroup cap (tile count of the persistent
// grid-stride launch) and the XCD-chunked list mapping.  Every candidate runs
// real production steps (plain launches, round-robin over `rounds` so that a
// slowly drifting population is seen by all), its force kernel bracketed by
// HIP events; the fastest total becomes the configuration of later steps.
int pst_prod_autotune(pst_sim* s, int rounds, int steps_per_cand, double* table, int max_rows,
                      int* n_rows, int* best5) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_autotune: call pst_prod_init first");
  if (rounds <= 0 || steps_per_cand <= 0) FAIL("pst_prod_autotune: bad arguments");
  CK(set_dev(ctx));
  ForceSys S;
  fill_sys(ctx, S);
  const int g0 = group_for(ctx, S);
  std::vector<int> Gs;
  for (int g = g0 / 4; g <= g0; g *= 2)
    if (g >= 1) Gs.push_back(g);
  if (Gs.empty()) Gs.push_back(g0);
  // brick-binned candidates when the system has the float2 pair grid
  bool binned = false;
  if (S.paired) {
    if (ensure_bins(ctx, s->S)) return 2;
    fill_sys(ctx, S);
    binned = bins_match(ctx, S);
  }
  struct Cand { int mode, G, cap, xcd; double ms; int n; double ms2; int n2; };
  std::vector<Cand> cands;
  const int caps[] = {1024, 2048, 4096, 8192};
  for (int g : Gs)
    for (int c : caps)
      for (int x = 0; x < 2; ++x) cands.push_back({0, g, c, x, 0.0, 0, 0.0, 0});
  if (binned)
    for (int g : Gs)
      if (g >= 8)
        for (int c : {1024, 4096}) cands.push_back({1, g, c, 1, 0.0, 0, 0.0, 0});
  // the direct kernel with z-pair fetches
  for (int g : Gs)
    for (int c : {2048, 8192})
      for (int x = 0; x < 2; ++x) cands.push_back({2, g, c, x, 0.0, 0, 0.0, 0});
  // split evaluation (reverse-Born pass in its own order), XCD-chunked
  if (S.has_rev)
    for (int g : Gs)
      if (g >= 16)
        for (int c : {2048, 8192}) {
          cands.push_back({3, g, c, 1, 0.0, 0, 0.0, 0});
          cands.push_back({4, g, c, 1, 0.0, 0, 0.0, 0});
        }
  CK(hipEventCreate(&s->tev0));
  CK(hipEventCreate(&s->tev1));
  RunCounters h;
  memset(&h, 0, sizeof(h));
  bool stop = false;
  long long k = 0;
  int rc = 0;
  // one production step under candidate c; its force launch time is added to *acc
  auto timed_step = [&](const Cand& c, double* acc, int* cnt) {
    s->tune_mode = c.mode; s->tune_G = c.G; s->tune_cap = c.cap; s->tune_xcd = c.xcd;
    rc = enqueue_step(s, (int)(k++ & 1));
    if (rc) return;
    hipError_t e = hipEventSynchronize(s->tev1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->tev0, s->tev1);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = 2; return; }
    *acc += ms;
    *cnt += 1;
    if (h.stop || (h.n_sr + h.n_mlp) == 0) stop = true;
  };
  for (int r = 0; r < rounds && !stop && !rc; ++r)
    for (auto& c : cands)
      for (int i = 0; i < steps_per_cand && !stop && !rc; ++i) timed_step(c, &c.ms, &c.n);
  // runoff: the kRunoff fastest candidates of the round-robin passes (those
  // measured in every pass) are timed again over kRunoffRounds interleaved
  // passes, and the choice rests on the runoff alone; the trajectory
  // population drifts from step to step, so single-step timings of launches
  // a few per cent apart can swap order
  constexpr int kRunoff = 4, kRunoffRounds = 4;
  int nmax = 0;
  for (auto& c : cands) nmax = c.n > nmax ? c.n : nmax;
  std::vector<Cand*> fin;
  for (auto& c : cands)
    if (c.n == nmax && nmax > 0) fin.push_back(&c);
  std::stable_sort(fin.begin(), fin.end(), [](const Cand* a, const Cand* b) { return a->ms < b->ms; });
  if (fin.size() > (size_t)kRunoff) fin.resize(kRunoff);
  if (fin.size() > 1)
    for (int r = 0; r < kRunoffRounds && !stop && !rc; ++r)
      for (Cand* c : fin)
        for (int i = 0; i < steps_per_cand && !stop && !rc; ++i) timed_step(*c, &c->ms2, &c->n2);
  hipEventDestroy(s->tev0);
  hipEventDestroy(s->tev1);
  s->tev0 = s->tev1 = nullptr;
  int n2max = 0;
  for (Cand* c : fin) n2max = c->n2 > n2max ? c->n2 : n2max;
  const Cand* best = nullptr;
  for (Cand* c : fin) {
    const double a = n2max > 0 ? c->ms2 : c->ms, b = best ? (n2max > 0 ? best->ms2 : best->ms) : 0.0;
    if ((n2max == 0 || c->n2 == n2max) && (!best || a < b)) best = c;
  }
  if (best) {
    s->tune_mode = best->mode; s->tune_G = best->G; s->tune_cap = best->cap; s->tune_xcd = best->xcd;
  } else {
    s->tune_mode = -1; s->tune_G = 0; s->tune_cap = 2048; s->tune_xcd = 1;
  }
  if (rc) return rc;
  if (table && n_rows) {
    int m = 0;
    for (auto& c : cands) {
      if (m >= max_rows) break;
      table[5 * m] = c.mode; table[5 * m + 1] = c.G; table[5 * m + 2] = c.cap;
      table[5 * m + 3] = c.xcd; table[5 * m + 4] = c.n ? c.ms / c.n : -1.0;
      ++m;
    }
    *n_rows = m;
  }
  if (best5) {
    best5[0] = s->tune_mode >= 0 ? s->tune_mode : ctx->gather_mode;
    best5[1] = s->tune_G > 0 ? s->tune_G : g0; best5[2] = s->tune_cap; best5[3] = s->tune_xcd;
    best5[4] = (int)k;
  }
  return 0;
}

// Times the production force kernel in its own step: `steps` production steps
// (plain launches, current launch configuration), each force launch bracketed
// by HIP events on the context stream and preceded by a gather census of its
// list; marker kernels bracket the window in a kernel trace.
// out6 = {mean ms per force launch, steps run, mean trajectories, mean elec-grid
// atoms, mean born-grid atoms, mean reverse-grid atoms} per launch.
int pst_prod_time_forces(pst_sim* s, int steps, double* out6) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_time_forces: call pst_prod_init first");
  if (steps <= 0 || !out6) FAIL("pst_prod_time_forces: bad arguments");
  CK(set_dev(ctx));
  if (sim_ensure_bins(s)) return 2;
  if (!s->census) CK(hipMalloc(&s->census, 4 * sizeof(unsigned long long)));
  CK(hipMemsetAsync(s->census, 0, 4 * sizeof(unsigned long long), ctx->stream));
  CK(hipEventCreate(&s->tev0));
  CK(hipEventCreate(&s->tev1));
  s->census_on = true;
  CK(launch_window_mark(ctx->stream));
  double total = 0;
  int done = 0, rc = 0;
  RunCounters h;
  memset(&h, 0, sizeof(h));
  for (int k = 0; k < steps; ++k) {
    rc = enqueue_step(s, k & 1);
    if (rc) break;
    hipError_t e = hipEventSynchronize(s->tev1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->tev0, s->tev1);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = 2; break; }
    total += ms;
    ++done;
    if (h.stop || (h.n_sr + h.n_mlp) == 0) break;
  }
  s->census_on = false;
  hipEventDestroy(s->tev0);
  hipEventDestroy(s->tev1);
  s->tev0 = s->tev1 = nullptr;
  if (rc) return rc;
  CK(launch_window_mark(ctx->stream));
  unsigned long long c[4];
  CK(hipMemcpyAsync(c, s->census, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  const double d = done > 0 ? (double)done : 1.0;
  out6[0] = total / d; out6[1] = done;
  out6[2] = c[0] / d; out6[3] = c[1] / d; out6[4] = c[2] / d; out6[5] = c[3] / d;
  return 0;
}

// Force-launch configuration of the production step, set by hand.
int pst_prod_set_launch(pst_sim* s, int mode, int G, int blocks_cap, int xcd) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (mode < -1 || mode > 4) FAIL("pst_prod_set_launch: mode must be -1 to 4");
  if (!(G == 0 || G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64))
    FAIL("pst_prod_set_launch: G must be 0 or a power of two up to 64");
  if (blocks_cap < 1) FAIL("pst_prod_set_launch: blocks_cap must be positive");
  CK(set_dev(ctx));
  s->tune_mode = mode; s->tune_G = G; s->tune_cap = blocks_cap; s->tune_xcd = xcd ? 1 : 0;
  return s->prod_ready ? sim_ensure_bins(s) : 0;
}

// Workgroup partition of a grouped force launch, XCD by XCD (per workgroups
// each, per >= L): systems are placed longest-first (weights w) on the XCD
// with the least weight so far, and each XCD's workgroups are split among its
// systems in proportion to their weights, one at least each.  Slice i =
// [start[i], start[i + 1]) of the XCD-major workgroup index belongs to
// system order[i]; the slices of XCD x lie in [x per, (x + 1) per).
static void xcd_partition(const std::vector<double>& w, int per, std::vector<int>& start,
                          std::vector<int>& order) {
  const int L = (int)w.size();
  std::vector<int> idx(L), xcd(L);
  for (int l = 0; l < L; ++l) idx[l] = l;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] > w[b]; });
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int l : idx) {
    int x = 0;
    for (int y = 1; y < 8; ++y)
      if (load[y] < load[x]) x = y;
    xcd[l] = x;
    load[x] += w[l];
  }
  start.assign(L + 1, 0);
  order.assign(L, 0);
  int i = 0;
  for (int x = 0; x < 8; ++x) {
    int k = 0;
    for (int l = 0; l < L; ++l) k += xcd[l] == x;
    const int extra = per - k;   // every system keeps one workgroup
    double cum = 0;
    int j = 0;
    for (int l = 0; l < L; ++l) {
      if (xcd[l] != x) continue;
      start[i] = x * per + j + (int)(extra * (load[x] > 0 ? cum / load[x] : 0.0));
      order[i] = l;
      cum += w[l];
      ++i;
      ++j;
    }
  }
  start[L] = 8 * per;
}

// Several simulators (one ligand each, receptor grids shared) stepped in
// lock-step: per step every simulator's pre phases, ONE force launch over
// all of their grid-path lists, then every simulator's post phases; captured
// graph_steps at a time in a hipGraph.  Stops when every simulator has
// stopped or has no live trajectory.  c16 = (L,16) counters as pst_prod_run.
int pst_group_run(pst_sim** sims, int L, long long max_steps, int graph_steps, double* c16,
                  double* ms_per_step) {
  if (!sims || L <= 0 || !sims[0]) return 1;
  pst_ctx* ctx = sims[0]->ctx;
  if (L > 64) FAIL("pst_group_run: at most 64 simulators");
  if (max_steps < 0 || graph_steps < 0) FAIL("pst_group_run: bad arguments");
  const bool obs = prod_bufs(sims[0]).obs.traj != nullptr;
  for (int l = 0; l < L; ++l) {
    if (!sims[l] || !sims[l]->prod_ready) FAIL("pst_group_run: every simulator needs pst_prod_init");
    if (sims[l]->precision != sims[0]->precision) FAIL("pst_group_run: mixed precisions");
    if (sims[l]->ctx->device != ctx->device) FAIL("pst_group_run: simulators on different devices");
    if ((prod_bufs(sims[l]).obs.traj != nullptr) != obs)
      FAIL("pst_group_run: every simulator needs the same record mode (observables or not)");
  }
  CK(set_dev(ctx));
  const hipStream_t st = ctx->stream;
  std::vector<hipStream_t> saved(L);
  for (int l = 0; l < L; ++l) {
    saved[l] = sims[l]->ctx->stream;
    CK(hipStreamSynchronize(saved[l]));
  }
  // per-simulator force descriptors, step parameters and buffers, and the
  // workgroup partitions of the force launch and of the step-phase launches
  std::vector<ForceSys> S(L);
  std::vector<ForceWork> W(L);
  std::vector<StepParams> P(L);
  std::vector<ProdBufs> B(L);
  std::vector<int> start(L + 1, 0), sstart(L + 1, 0);
  int G = 16;
  size_t lds = 0;
  const int sb = prod_step_block();
  bool any_outer = false;
  for (int l = 0; l < L; ++l) {
    any_outer = any_outer || sims[l]->P.lmz_on;
    fill_sys(sims[l]->ctx, S[l]);
    P[l] = sims[l]->P;
    B[l] = prod_bufs(sims[l]);
    B[l].rev_list = nullptr;   // the grouped force launch has no split evaluation
    sstart[l + 1] = sstart[l] + (B[l].n_slots + sb - 1) / sb;
    W[l] = main_work(sims[l]);
    W[l].xcd_chunks = 1;
    G = std::max(G, group_for(sims[l]->ctx, S[l]));
    if (S[l].M <= 1536)
      lds = std::max(lds, (size_t)S[l].M * (sims[0]->precision == 0 ? 32 : 16));
  }
  G = G >= 64 ? 64 : (G >= 32 ? 32 : 16);
  const bool zp = sim_zp(sims[0]);
  // workgroups per system: the grid-stride loop covers any count, so the
  // launch holds about 4096 resident workgroups in all (a launch of mostly
  // idle workgroups costs dispatch time when few trajectories are left)
  const int gpb = 256 / G;
  const int wg_cap = std::max(8, (4096 / L + 7) / 8 * 8);
  for (int l = 0; l < L; ++l) {
    int b = (sims[l]->S + gpb - 1) / gpb;
    if (b > wg_cap) b = wg_cap;
    b = (b + 7) / 8 * 8;
    start[l + 1] = start[l] + b;
  }
  // With 8 or more systems each system's workgroups run on one XCD, so an
  // XCD's L2 holds the ligand grids of about L / 8 systems instead of all L
  // (xcd_partition; re-balanced between graph launches from the systems'
  // grid-path counts of the last step).
  const bool xcd_major = L >= 8;
  const int n_wg = start[L];
  std::vector<int> order(L);
  for (int l = 0; l < L; ++l) order[l] = l;
  if (xcd_major) {
    std::vector<double> w(L);
    for (int l = 0; l < L; ++l) w[l] = start[l + 1] - start[l];
    xcd_partition(w, n_wg / 8, start, order);
    for (int l = 0; l < L; ++l) W[l].xcd_chunks = 0;
  }
  // the simulators whose surrogate is attached: one grouped network forward
  // per step over all of them
  std::vector<int> surr_l;
  std::vector<MlpInfer> MI;
  std::vector<int> mstart(1, 0);
  for (int l = 0; l < L; ++l) {
    pst_sim* sl = sims[l];
    if (!sl->mlp || !(sl->P.surrogate_r > 0)) continue;
    const int Sp = (sl->S + 63) / 64 * 64;
    surr_l.push_back(l);
    MI.push_back(mlp_infer_args(sl->mlp->m, Sp, &sl->ctr->n_mlp, sl->mlp->m.y));
    mstart.push_back(mstart.back() + (Sp + kInferRows - 1) / kInferRows);
  }
  const int n_surr = (int)surr_l.size();
  // their surrogate phase: features, the grouped forward and the applied
  // forces (workgroup partition fstart), the audit sample and the audit's
  // grid-path launch (one system per surrogate, partition astart)
  std::vector<SurrTask> tasks(n_surr);
  std::vector<ForceSys> Sa(n_surr);
  std::vector<ForceWork> Wa(n_surr);
  std::vector<int> fstart(n_surr + 1, 0), astart(n_surr + 1, 0);
  for (int i = 0; i < n_surr; ++i) {
    pst_sim* sl = sims[surr_l[i]];
    SurrTask& t = tasks[i];
    t.sp = sl->sp;
    t.pos = sl->pos;
    t.quat = sl->quat;
    t.list = sl->mlp_list;
    t.n_dev = &sl->ctr->n_mlp;
    t.x0f = sl->mlp->m.x0f;
    t.y = sl->mlp->m.y;
    t.F = sl->F;
    t.T = sl->T;
    t.ctr = sl->ctr;
    t.audit_list = sl->audit_list;
    t.audit_count = sl->audit_count;
    t.audit_F = sl->audit_F;
    t.audit_T = sl->sp.use_torque ? sl->audit_T : nullptr;
    fstart[i + 1] = fstart[i] + (sl->S + 255) / 256;
    Sa[i] = S[surr_l[i]];
    ForceWork A = W[surr_l[i]];
    A.list = sl->audit_list;
    A.n_dev = sl->audit_count;
    A.n_cap = sl->sp.audit_n;
    A.out_compact = 1;
    A.xcd_chunks = 0;
    A.F = sl->audit_F;
    A.T = sl->audit_T;
    Wa[i] = A;
    astart[i + 1] = astart[i] + (audits(sl) ? (sl->sp.audit_n + gpb - 1) / gpb : 0);
  }
  ForceSys* dS = nullptr;
  ForceWork* dW = nullptr;
  StepParams* dP = nullptr;
  ProdBufs* dB = nullptr;
  int* dstart = nullptr;
  int* dsstart = nullptr;
  MlpInfer* dMI = nullptr;
  int* dmstart = nullptr;
  int* dorder = nullptr;
  SurrTask* dT = nullptr;
  ForceSys* dSa = nullptr;
  ForceWork* dWa = nullptr;
  int* dfstart = nullptr;
  int* dastart = nullptr;
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  auto cleanup = [&]() {
    for (int l = 0; l < L; ++l) sims[l]->ctx->stream = saved[l];
    if (dS) hipFree(dS);
    if (dW) hipFree(dW);
    if (dP) hipFree(dP);
    if (dB) hipFree(dB);
    if (dstart) hipFree(dstart);
    if (dsstart) hipFree(dsstart);
    if (dMI) hipFree(dMI);
    if (dmstart) hipFree(dmstart);
    if (dorder) hipFree(dorder);
    if (dT) hipFree(dT);
    if (dSa) hipFree(dSa);
    if (dWa) hipFree(dWa);
    if (dfstart) hipFree(dfstart);
    if (dastart) hipFree(dastart);
    if (fork_ev) hipEventDestroy(fork_ev);
    if (join_ev) hipEventDestroy(join_ev);
    if (aux) hipStreamDestroy(aux);
  };
  if (hipMalloc(&dS, sizeof(ForceSys) * L) != hipSuccess ||
      hipMalloc(&dW, sizeof(ForceWork) * L) != hipSuccess ||
      hipMalloc(&dP, sizeof(StepParams) * L) != hipSuccess ||
      hipMalloc(&dB, sizeof(ProdBufs) * L) != hipSuccess ||
      hipMalloc(&dstart, sizeof(int) * (L + 1)) != hipSuccess ||
      hipMalloc(&dorder, sizeof(int) * L) != hipSuccess ||
      hipMalloc(&dsstart, sizeof(int) * (L + 1)) != hipSuccess ||
      (n_surr > 0 && (hipMalloc(&dMI, sizeof(MlpInfer) * n_surr) != hipSuccess ||
                      hipMalloc(&dmstart, sizeof(int) * (n_surr + 1)) != hipSuccess ||
                      hipMalloc(&dT, sizeof(SurrTask) * n_surr) != hipSuccess ||
                      hipMalloc(&dSa, sizeof(ForceSys) * n_surr) != hipSuccess ||
                      hipMalloc(&dWa, sizeof(ForceWork) * n_surr) != hipSuccess ||
                      hipMalloc(&dfstart, sizeof(int) * (n_surr + 1)) != hipSuccess ||
                      hipMalloc(&dastart, sizeof(int) * (n_surr + 1)) != hipSuccess))) {
    cleanup();
    FAIL("pst_group_run: out of device memory");
  }
  if (n_surr > 0 &&
      (hipMemcpy(dMI, MI.data(), sizeof(MlpInfer) * n_surr, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dmstart, mstart.data(), sizeof(int) * (n_surr + 1), hipMemcpyHostToDevice) !=
           hipSuccess ||
       hipMemcpy(dT, tasks.data(), sizeof(SurrTask) * n_surr, hipMemcpyHostToDevice) !=
           hipSuccess ||
       hipMemcpy(dSa, Sa.data(), sizeof(ForceSys) * n_surr, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dWa, Wa.data(), sizeof(ForceWork) * n_surr, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dfstart, fstart.data(), sizeof(int) * (n_surr + 1), hipMemcpyHostToDevice) !=
           hipSuccess ||
       hipMemcpy(dastart, astart.data(), sizeof(int) * (n_surr + 1), hipMemcpyHostToDevice) !=
           hipSuccess)) {
    cleanup();
    FAIL("pst_group_run: upload failed");
  }
  if (n_surr > 0 &&
      (hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&join_ev, hipEventDisableTiming) != hipSuccess)) {
    cleanup();
    FAIL("pst_group_run: stream / event creation failed");
  }
  if (hipMemcpy(dS, S.data(), sizeof(ForceSys) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dW, W.data(), sizeof(ForceWork) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dP, P.data(), sizeof(StepParams) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dB, B.data(), sizeof(ProdBufs) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dstart, start.data(), sizeof(int) * (L + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dorder, order.data(), sizeof(int) * L, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dsstart, sstart.data(), sizeof(int) * (L + 1), hipMemcpyHostToDevice) !=
          hipSuccess) {
    cleanup();
    FAIL("pst_group_run: upload failed");
  }
  // one step of every system: each phase is one launch over all systems; the
  // surrogate phase (features, every network's forward, applied forces, the
  // audit sample, its grid-path launch and the audit comparison) runs on the
  // second stream beside the grouped force launch, touching only the
  // surrogate lists' slots and the networks' and audits' buffers
  auto step = [&](int par) -> int {
    const hipStream_t cs = sims[0]->ctx->stream;
    const int prec = sims[0]->precision;
    hipError_t e = prod_group_pre(dP, dB, dsstart, L, sstart[L], par, any_outer, cs);
    if (n_surr > 0) {
      if (e == hipSuccess) e = hipEventRecord(fork_ev, cs);
      if (e == hipSuccess) e = hipStreamWaitEvent(aux, fork_ev, 0);
      if (e == hipSuccess) e = surr_group_audit_pick(dT, n_surr, aux);
      if (e == hipSuccess)
        e = launch_forces_group(dSa, dWa, dastart, n_surr, astart[n_surr], lds, prec, G, false,
                                aux);
      if (e == hipSuccess) e = surr_group_features(dT, dfstart, n_surr, fstart[n_surr], aux);
      if (e == hipSuccess) e = launch_mlp_infer_group(dMI, dmstart, n_surr, mstart[n_surr], aux);
      if (e == hipSuccess) e = surr_group_apply(dT, dfstart, n_surr, fstart[n_surr], aux);
      if (e == hipSuccess) e = surr_group_audit(dT, n_surr, aux);
      if (e == hipSuccess) e = hipEventRecord(join_ev, aux);
    }
    if (e == hipSuccess)
      e = launch_forces_group(dS, dW, dstart, L, n_wg, lds, prec, G, zp, cs, xcd_major, dorder);
    if (n_surr > 0 && e == hipSuccess) e = hipStreamWaitEvent(cs, join_ev, 0);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return 2; }
    e = prod_group_post(dP, dS, dB, dsstart, L, sstart[L], obs, par, cs);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return 2; }
    return 0;
  };
  for (int l = 0; l < L; ++l) sims[l]->ctx->stream = st;
  int rc = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  const int gs = graph_steps;
  if (gs > 0) {
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { cleanup(); ctx->err = hipGetErrorString(e); return 2; }
    for (int l = 0; l < L; ++l) sims[l]->ctx->stream = cap;
    for (int i = 0; i < gs && !rc; ++i) rc = step(i & 1);
    for (int l = 0; l < L; ++l) sims[l]->ctx->stream = st;
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cap, &g);
    if (rc || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      hipStreamDestroy(cap);
      cleanup();
      if (!rc) { ctx->err = hipGetErrorString(e); rc = 2; }
      return rc;
    }
    graph = g;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(graph);
      hipStreamDestroy(cap);
      cleanup();
      ctx->err = hipGetErrorString(e);
      return 2;
    }
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, st);
  long long steps = 0;
  std::vector<RunCounters> h(L);
  bool stop = false;
  while (steps < max_steps && !stop && !rc) {
    if (gs > 0 && steps + gs <= max_steps) {
      hipError_t e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = 2; break; }
      steps += gs;
    } else {
      rc = step((int)(steps & 1));
      if (rc) break;
      steps += 1;
    }
    bool all_done = true;
    for (int l = 0; l < L; ++l)
      hipMemcpyAsync(&h[l], sims[l]->ctr, sizeof(RunCounters), hipMemcpyDeviceToHost, st);
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = 2; break; }
    for (int l = 0; l < L; ++l)
      if (!(h[l].stop || (h[l].n_sr + h[l].n_mlp) == 0)) all_done = false;
    stop = all_done;
    if (xcd_major && !stop) {
      // weights: each system's grid-path trajectories of the last step, plus a
      // floor so that a system whose count grows within the next graph
      // launch keeps a share
      std::vector<double> w(L);
      double mean = 0;
      for (int l = 0; l < L; ++l) mean += (double)h[l].n_sr / L;
      for (int l = 0; l < L; ++l) w[l] = (double)h[l].n_sr + 0.05 * mean + 1.0;
      std::vector<int> st2(L + 1), or2(L);
      xcd_partition(w, n_wg / 8, st2, or2);
      if (st2 != start || or2 != order) {
        start = st2;
        order = or2;
        hipError_t e2 = hipMemcpyAsync(dstart, start.data(), sizeof(int) * (L + 1),
                                       hipMemcpyHostToDevice, st);
        if (e2 == hipSuccess)
          e2 = hipMemcpyAsync(dorder, order.data(), sizeof(int) * L, hipMemcpyHostToDevice, st);
        if (e2 != hipSuccess) { ctx->err = hipGetErrorString(e2); rc = 2; break; }
      }
    }
  }
  hipEventRecord(e1, st);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (cap) hipStreamDestroy(cap);
  cleanup();
  if (rc) return rc;
  if (ms_per_step) *ms_per_step = steps ? ms / steps : 0.0;
  if (c16) {
    for (int l = 0; l < L; ++l) {
      const RunCounters& x = h[l];
      double v[16] = {(double)x.n_reacted, (double)x.n_escaped, (double)x.n_maxsteps,
                      (double)x.total_steps, (double)x.next_traj, (double)x.n_returns,
                      (double)x.n_overlap, (double)x.n_bridge, (double)steps, (double)x.stop,
                      (double)x.n_sr, (double)x.n_mlp, x.audit_err2, x.audit_ref2,
                      (double)x.audit_n, 0};
      memcpy(c16 + 16 * l, v, sizeof(v));
    }
  }
  return 0;
}

int pst_prod_sm_counts(pst_sim* s, unsigned long long* out48) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready || !out48) FAIL("pst_prod_sm_counts: call pst_prod_init first");
  CK(set_dev(ctx));
  RunCounters h;
  CK(hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < PST_MAX_PAIRS; ++r) {
    out48[r] = h.sm_fired[r];
    out48[PST_MAX_PAIRS + r] = h.sm_fired_bridge[r];
    out48[2 * PST_MAX_PAIRS + r] = h.sm_unique[r];
  }
  return 0;
}

int pst_prod_obs_sizes(int* traj_doubles, int* hist_len) {
  if (traj_doubles) *traj_doubles = kObsTraj;
  if (hist_len) *hist_len = kObsHist;
  return 0;
}

int pst_prod_observables(pst_sim* s, double* traj, unsigned long long* hist) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->obs.traj) FAIL("pst_prod_observables: run was not recorded at level 2");
  CK(set_dev(ctx));
  CK(prod_flush_observables(s->P, prod_bufs(s), ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (traj)
    CK(hipMemcpy(traj, s->obs.traj, (size_t)s->P.n_total * kObsTraj * 8, hipMemcpyDeviceToHost));
  if (hist) CK(hipMemcpy(hist, s->obs.hist, kObsHist * 8, hipMemcpyDeviceToHost));
  return 0;
}

int pst_prod_outcomes(pst_sim* s, signed char* status, int* steps) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->out_status) FAIL("pst_prod_outcomes: run was not recorded (pst_prod_init record=0)");
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  const size_t N = (size_t)s->P.n_total;
  if (status) CK(hipMemcpy(status, s->out_status, N, hipMemcpyDeviceToHost));
  if (steps) CK(hipMemcpy(steps, s->out_steps, N * 4, hipMemcpyDeviceToHost));
  return 0;
}

int pst_prod_counters(pst_sim* s, double* c16) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready || !c16) FAIL("pst_prod_counters: call pst_prod_init first");
  CK(set_dev(ctx));
  RunCounters h;
  CK(hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  const double v[16] = {(double)h.n_reacted, (double)h.n_escaped, (double)h.n_maxsteps,
                        (double)h.total_steps, (double)h.next_traj, (double)h.n_returns,
                        (double)h.n_overlap, (double)h.n_bridge, 0.0, (double)h.stop,
                        (double)h.n_sr, (double)h.n_mlp, h.audit_err2, h.audit_ref2,
                        (double)h.audit_n, 0};
  memcpy(c16, v, sizeof(v));
  return 0;
}

// ---- checkpoint / resume -------------------------------------------------------
// Every device buffer whose contents carry from one production step to the
// next, in a fixed order: slot state, run counters, per-trajectory outcomes
// (record >= 1) and observables (record 2).  Noise is keyed by (trajectory,
// own step), so restoring these restores the run exactly.
namespace {
struct Region {
  void* dev;
  size_t bytes;
};
constexpr long long kStateMagic = 0x3153544154535350ll;  // "PSSTATS1"
constexpr int kStateHeader = 8;                           // int64 words
}  // namespace

static std::vector<Region> state_regions(pst_sim* s) {
  const size_t S = s->S, N = (size_t)s->P.n_total;
  std::vector<Region> r = {
      {s->pos, S * 3 * 8},       {s->quat, S * 4 * 8},        {s->status, S * 4},
      {s->n_steps32, S * 4},     {s->traj, S * 8},            {s->F, S * 3 * 8},
      {s->T, S * 3 * 8},         {s->dt, S * 8},              {s->x0, S * PST_MAX_PAIRS * 8},
      {s->x1, S * PST_MAX_PAIRS * 8}, {s->sm_state, S * 4},   {s->sm_mask, S * 4},
      {s->lmz_escaped, S * 4},   {s->key, S * 4},             {s->bins, 1024 * 4},
      {s->ctr, sizeof(RunCounters)}};
  if (s->out_status) {
    r.push_back({s->out_status, N});
    r.push_back({s->out_steps, N * 4});
  }
  if (s->obs.traj) {
    const ObsBufs& O = s->obs;
    r.push_back({O.s_min, S * 8});
    r.push_back({O.s_step, S * 4});
    r.push_back({O.s_time, S * 8});
    r.push_back({O.s_ret, S * 4});
    r.push_back({O.s_prev_r, S * 8});
    r.push_back({O.s_nm, S * 7 * 8});
    r.push_back({O.traj, N * kObsTraj * 8});
    r.push_back({O.hist, (size_t)kObsHist * 8});
  }
  return r;
}

static long long state_bytes(pst_sim* s) {
  long long b = kStateHeader * 8 + (long long)sizeof(StepParams);
  for (const Region& r : state_regions(s)) b += (long long)r.bytes;
  return b;
}

static int record_level(const pst_sim* s) { return s->obs.traj ? 2 : (s->out_status ? 1 : 0); }

int pst_prod_state_bytes(pst_sim* s, long long* bytes) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_state_bytes: call pst_prod_init first");
  if (!bytes) FAIL("pst_prod_state_bytes: null pointer");
  *bytes = state_bytes(s);
  return 0;
}

int pst_prod_save_state(pst_sim* s, void* host, long long bytes) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_save_state: call pst_prod_init first");
  if (!host) FAIL("pst_prod_save_state: null pointer");
  if (bytes != state_bytes(s)) FAIL("pst_prod_save_state: buffer size differs from pst_prod_state_bytes");
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  long long* hdr = (long long*)host;
  const long long h[kStateHeader] = {kStateMagic, 1, (long long)s->S, (long long)s->P.n_total,
                                     record_level(s), (long long)sizeof(RunCounters),
                                     (long long)sizeof(StepParams), bytes};
  memcpy(hdr, h, sizeof(h));
  unsigned char* p = (unsigned char*)host + kStateHeader * 8;
  memcpy(p, &s->P, sizeof(StepParams));
  p += sizeof(StepParams);
  for (const Region& r : state_regions(s)) {
    CK(hipMemcpy(p, r.dev, r.bytes, hipMemcpyDeviceToHost));
    p += r.bytes;
  }
  return 0;
}

int pst_prod_load_state(pst_sim* s, const void* host, long long bytes) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_load_state: call pst_prod_init (same record level) first");
  if (!host) FAIL("pst_prod_load_state: null pointer");
  if (bytes < kStateHeader * 8 + (long long)sizeof(StepParams))
    FAIL("pst_prod_load_state: buffer too small for a saved state");
  long long h[kStateHeader];
  memcpy(h, host, sizeof(h));
  if (h[0] != kStateMagic || h[1] != 1) FAIL("pst_prod_load_state: not a saved production state");
  if (h[2] != s->S || h[3] != s->P.n_total || h[4] != record_level(s))
    FAIL("pst_prod_load_state: saved slots / trajectories / record level differ from this run");
  if (h[5] != (long long)sizeof(RunCounters) || h[6] != (long long)sizeof(StepParams) ||
      h[7] != bytes || bytes != state_bytes(s))
    FAIL("pst_prod_load_state: saved layout differs from this library's");
  const unsigned char* p = (const unsigned char*)host + kStateHeader * 8;
  if (memcmp(p, &s->P, sizeof(StepParams)) != 0)
    FAIL("pst_prod_load_state: saved step parameters differ from this run's");
  p += sizeof(StepParams);
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  for (const Region& r : state_regions(s)) {
    CK(hipMemcpy(r.dev, p, r.bytes, hipMemcpyHostToDevice));
    p += r.bytes;
  }
  return 0;
}

// Audit sums of the run counters (since the last pst_sim_reset_audit).
int pst_prod_audit(pst_sim* s, double* out6) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->ctr) FAIL("pst_prod_audit: call pst_prod_init first");
  if (!out6) FAIL("pst_prod_audit: null pointer");
  CK(set_dev(ctx));
  RunCounters h;
  CK(hipMemcpyAsync(&h, s->ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  out6[0] = h.audit_err2; out6[1] = h.audit_ref2; out6[2] = (double)h.audit_n;
  out6[3] = h.audit_terr2; out6[4] = h.audit_tref2; out6[5] = 0.0;
  return 0;
}

// Path and energetics rows of the running trajectories that have moved
// (gpu_batch_simulator.py:1318-1340): path8 (S,8) = trajectory, step,
// centroid x y z, q0 q1 q2; energy6 (S,6) = trajectory, step, force x y z,
// dt.  Rows in trajectory order.
int pst_prod_snapshot(pst_sim* s, long long step, double* path8, double* energy6, int* n_rows) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->prod_ready) FAIL("pst_prod_snapshot: call pst_prod_init first");
  if (!n_rows) FAIL("pst_prod_snapshot: null pointer");
  CK(set_dev(ctx));
  CK(hipStreamSynchronize(ctx->stream));
  const size_t S = s->S;
  std::vector<double> pos(S * 3), q(S * 4), F(S * 3), dt(S);
  std::vector<int> st(S);
  std::vector<unsigned int> ns(S);
  std::vector<long long> tr(S);
  CK(hipMemcpy(st.data(), s->status, S * 4, hipMemcpyDeviceToHost));
  CK(hipMemcpy(ns.data(), s->n_steps32, S * 4, hipMemcpyDeviceToHost));
  CK(hipMemcpy(tr.data(), s->traj, S * 8, hipMemcpyDeviceToHost));
  CK(hipMemcpy(pos.data(), s->pos, S * 24, hipMemcpyDeviceToHost));
  CK(hipMemcpy(q.data(), s->quat, S * 32, hipMemcpyDeviceToHost));
  CK(hipMemcpy(F.data(), s->F, S * 24, hipMemcpyDeviceToHost));
  CK(hipMemcpy(dt.data(), s->dt, S * 8, hipMemcpyDeviceToHost));
  // running trajectories that have moved (a slot refilled at the end of the
  // step holds a new trajectory that has not), in trajectory order as the
  // reference lists them (which slot a trajectory occupies depends on the
  // order in which slots were refilled)
  std::vector<size_t> order;
  for (size_t i = 0; i < S; ++i)
    if (st[i] == PST_RUNNING && ns[i] != 0) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return tr[a] < tr[b]; });
  int n = 0;
  for (size_t i : order) {
    const double t = (double)tr[i];
    if (path8) {
      double* r = path8 + 8 * (size_t)n;
      r[0] = t; r[1] = (double)step;
      r[2] = pos[3 * i]; r[3] = pos[3 * i + 1]; r[4] = pos[3 * i + 2];
      r[5] = q[4 * i]; r[6] = q[4 * i + 1]; r[7] = q[4 * i + 2];
    }
    if (energy6) {
      double* r = energy6 + 6 * (size_t)n;
      r[0] = t; r[1] = (double)step;
      r[2] = F[3 * i]; r[3] = F[3 * i + 1]; r[4] = F[3 * i + 2];
      r[5] = dt[i];
    }
    ++n;
  }
  *n_rows = n;
  return 0;
}

int pst_prod_last_ms_per_step(pst_sim* s, double* ms) {
  if (!s || !ms) return 1;
  *ms = s->last_ms_per_step;
  return 0;
}

// ---- synthetic grids ---------------------------------------------------------
int pst_make_grid(pst_ctx* ctx, int kind, int nx, int ny, int nz, const double* origin3,
                  const double* spacing3, const double* atoms, int n_atoms, double debye_length,
                  double eps_out, double eps_in, int dtype, void* host_out) {
  if (!ctx) return 1;
  if (kind != 0 && kind != 1) FAIL("pst_make_grid: kind must be 0 or 1");
  if (nx < 2 || ny < 2 || nz < 2 || n_atoms <= 0 || !atoms || !host_out || !origin3 || !spacing3)
    FAIL("pst_make_grid: bad arguments");
  CK(set_dev(ctx));
  const size_t count = (size_t)nx * ny * nz;
  double* d_atoms = nullptr;
  void* d_out = nullptr;
  CK(hipMalloc(&d_atoms, (size_t)n_atoms * 5 * 8));
  CK(hipMalloc(&d_out, count * (dtype == PST_F64 ? 8 : 4)));
  CK(hipMemcpy(d_atoms, atoms, (size_t)n_atoms * 40, hipMemcpyHostToDevice));
  CK(launch_grid_dh(dtype == PST_F64 ? nullptr : (float*)d_out,
                    dtype == PST_F64 ? (double*)d_out : nullptr, nx, ny, nz, origin3, spacing3,
                    d_atoms, n_atoms, debye_length, eps_out, eps_in, kind, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(host_out, d_out, count * (dtype == PST_F64 ? 8 : 4), hipMemcpyDeviceToHost));
  hipFree(d_atoms);
  hipFree(d_out);
  return 0;
}

// ---- surrogate MLP ------------------------------------------------------------
int pst_mlp_param_count(void) { return kMlpParams; }

int pst_mlp_create(pst_ctx* ctx, int max_rows, pst_mlp** out) {
  if (!ctx) return 1;
  if (max_rows <= 0 || !out) FAIL("pst_mlp_create: bad arguments");
  CK(set_dev(ctx));
  pst_mlp* p = new pst_mlp();
  p->ctx = ctx;
  *out = p;
  MlpBufs& m = p->m;
  const size_t R = (size_t)(max_rows + kInferRows - 1) / kInferRows * kInferRows;
  m.max_rows = (int)R;
  const size_t H = kMlpWidth;
  auto al = [&](void** dst, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, bytes);
    if (e == hipSuccess) p->allocs.push_back(*dst);
    return e;
  };
  CK(al((void**)&m.w0, H * 64 * 2));
  CK(al((void**)&m.w1, H * H * 2));
  CK(al((void**)&m.w2, H * H * 2));
  CK(al((void**)&m.w3, H * H * 2));
  CK(al((void**)&m.w4, 64 * H * 2));
  CK(al((void**)&m.b0, H * 4));
  CK(al((void**)&m.b1, H * 4));
  CK(al((void**)&m.b2, H * 4));
  CK(al((void**)&m.b3, H * 4));
  CK(al((void**)&m.b4, 64 * 4));
  CK(al((void**)&m.x0, R * 64 * 2));
  CK(al((void**)&m.w0f, H * kMlpIn * 4));
  CK(al((void**)&m.x0f, R * kMlpInLd * 4));
  __bf16** hs[4] = {&m.h1, &m.h2, &m.h3, &m.h4};
  float** zs[4] = {&m.z1, &m.z2, &m.z3, &m.z4};
  __bf16** gs[4] = {&m.g1, &m.g2, &m.g3, &m.g4};
  for (int l = 0; l < 4; ++l) {
    CK(al((void**)hs[l], R * H * 2));
    CK(al((void**)zs[l], R * H * 4));
    CK(al((void**)gs[l], R * H * 2));
  }
  CK(al((void**)&m.y, R * 64 * 4));
  CK(al((void**)&m.gy, R * 64 * 2));
  CK(al((void**)&m.gw, (H * 64 * 2 + H * H * 3) * 4));
  CK(al((void**)&m.gcol, (R + kColRows - 1) / kColRows * 256 * 4));
  CK(hipMemsetAsync(m.x0, 0, R * 64 * 2, ctx->stream));
  CK(hipMemsetAsync(m.x0f, 0, R * kMlpInLd * 4, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pst_mlp_destroy(pst_mlp* p) {
  if (!p) return 0;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  for (void* a : p->allocs) hipFree(a);
  delete p;
  return 0;
}

int pst_mlp_load(pst_mlp* p, const float* params) {
  if (!p) return 1;
  pst_ctx* ctx = p->ctx;
  if (!params) FAIL("pst_mlp_load: null parameters");
  CK(set_dev(ctx));
  CK(mlp_cast(p->m, params, ctx->stream));
  return 0;
}

int pst_mlp_forward(pst_mlp* p, int B, const float* x, float* y, int train) {
  if (!p) return 1;
  pst_ctx* ctx = p->ctx;
  if (B <= 0 || B > p->m.max_rows || !x || !y) FAIL("pst_mlp_forward: bad batch or pointers");
  CK(set_dev(ctx));
  CK(mlp_forward(p->m, B, x, y, train != 0, ctx->stream));
  return 0;
}

int pst_mlp_time_infer(pst_mlp* p, int B, int reps, double* ms_per_launch) {
  if (!p) return 1;
  pst_ctx* ctx = p->ctx;
  if (B <= 0 || B > p->m.max_rows || reps <= 0 || !ms_per_launch)
    FAIL("pst_mlp_time_infer: bad arguments");
  CK(set_dev(ctx));
  const int Bp = (B + 63) / 64 * 64;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) e = mlp_forward_packed(p->m, Bp, nullptr, false, p->m.y, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
  for (int i = 0; i < reps && e == hipSuccess; ++i)
    e = mlp_forward_packed(p->m, Bp, nullptr, false, p->m.y, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (e != hipSuccess) {
    ctx->err = std::string("pst_mlp_time_infer: ") + hipGetErrorString(e);
    return 2;
  }
  *ms_per_launch = ms / reps;
  return 0;
}

int pst_mlp_backward(pst_mlp* p, int B, const float* gy, float* gparams) {
  if (!p) return 1;
  pst_ctx* ctx = p->ctx;
  if (B <= 0 || B > p->m.max_rows || !gy || !gparams) FAIL("pst_mlp_backward: bad batch or pointers");
  CK(set_dev(ctx));
  CK(mlp_backward(p->m, B, gy, gparams, ctx->stream));
  return 0;
}

int pst_sim_set_surrogate(pst_sim* s, pst_mlp* mlp, const double* p9) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  CK(set_dev(ctx));
  if (!mlp) {
    s->mlp = nullptr;
    s->P.surrogate_r = -1.0;
    return 0;
  }
  if (!p9) FAIL("pst_sim_set_surrogate: null parameters");
  if (mlp->m.max_rows < s->S) FAIL("pst_sim_set_surrogate: MLP buffers smaller than the slot pool");
  SurrParams sp;
  sp.r_on = p9[0]; sp.r0 = p9[1]; sp.lam = p9[2]; sp.inv_b = p9[3];
  sp.scale_F = p9[4]; sp.scale_T = p9[5]; sp.scale_E = p9[6];
  sp.use_torque = p9[7] != 0.0; sp.audit_n = (int)p9[8];
  if (!(sp.r_on > 0) || !(sp.r0 > 0) || !(sp.lam > 0)) FAIL("pst_sim_set_surrogate: bad parameters");
  if (sp.audit_n < 0 || sp.audit_n > 65536) FAIL("pst_sim_set_surrogate: audit_n out of range");
  if (sp.audit_n > s->audit_cap || !s->audit_count) {
    if (s->audit_F) hipFree(s->audit_F);
    if (s->audit_T) hipFree(s->audit_T);
    if (s->audit_list) hipFree(s->audit_list);
    s->audit_F = s->audit_T = nullptr;
    s->audit_list = nullptr;
    s->audit_cap = 0;
    if (!s->audit_count) CK(hipMalloc(&s->audit_count, sizeof(int)));
    const size_t n = sp.audit_n > 0 ? (size_t)sp.audit_n : 1;
    CK(hipMalloc(&s->audit_F, n * 3 * 8));
    CK(hipMalloc(&s->audit_T, n * 3 * 8));
    CK(hipMalloc(&s->audit_list, n * 4));
    s->audit_cap = (int)n;
  }
  s->sp = sp;
  s->mlp = mlp;
  s->P.surrogate_r = sp.r_on;
  return 0;
}

int pst_sim_reset_audit(pst_sim* s) {
  if (!s) return 1;
  pst_ctx* ctx = s->ctx;
  if (!s->ctr) FAIL("pst_sim_reset_audit: call pst_prod_init first");
  CK(set_dev(ctx));
  CK(hipMemsetAsync(&s->ctr->audit_n, 0, sizeof(unsigned long long), ctx->stream));
  CK(hipMemsetAsync(&s->ctr->audit_err2, 0, 2 * sizeof(double), ctx->stream));
  CK(hipMemsetAsync(&s->ctr->audit_terr2, 0, 2 * sizeof(double), ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // extern "C"
