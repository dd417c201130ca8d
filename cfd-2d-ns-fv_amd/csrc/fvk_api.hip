// fvk_api.hip — the C ABI (include/fvk.h): contexts, device memory, and the
// SSP-RK3 driver (ssp_rk3.cpp:18-77 as a sequence of kernel launches on one
// stream).
//
// One engine serves both entry-point families.  A single solve (fvk_ctx) is an
// engine with one instance; a batch (fvk_batch) concatenates its instances'
// device layouts: each instance's padded cell range, block list, slot table,
// halo lists and block face lists are laid end to end, with the global cell
// indices in them (neighbour codes, boundary ghost sources, halo cells)
// offset by the instance's first cell.  Block-local data need no change.  The
// kernels find a block's instance in a per-block table and take that
// instance's constants (freestream, CFL, limiter switch) and dt maximum from
// per-instance arrays, so one launch advances every instance by one stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "fv_kernels.hpp"
#include "fvk.h"
#include "mesh_prep.hpp"

// Scratch of fvk_residual_l2_noalloc, one copy per GPU for the life of the
// process: [4][kScratchBlocks] per-block partial sums, then the 4 norms
// (2 MiB).  Host code uses only its device address (hipGetSymbolAddress).
constexpr int kScratchBlocks = 65536;   // most mesh blocks a call may have
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[4 * kScratchBlocks + 4];

namespace {

thread_local std::string g_create_error;

struct Instance {
  fvk::HostMesh hm;             // the instance's own layout (cells numbered from 0)
  fvk_options opt{};
  double qinf[4] = {0, 0, 0, 0};
  int32_t cell_off = 0;         // first padded cell in the engine's arrays
  int32_t block_off = 0;        // first mesh block
  int32_t slot_off = 0;         // first face slot
  int32_t halo_off = 0;         // first halo-list entry
  int32_t bface_off = 0;        // first block-face entry
  bool active = true;
  int64_t steps = 0;            // steps run since the last (re)start
  int32_t why = 0;              // FVK_RETIRED_*: why the instance is not active
  double best = 0.0;            // run_to_drop: lowest density residual norm that counted
  int32_t since_best = 0;       // run_to_drop: checks since `best` was set
  double ref0 = 0.0;            // run_to_drop: caller's reference norm (> 0), else the first step's
};

struct Engine {
  std::string err;
  fvk_options opt{};            // instance 0's options (the kernel-shaping ones are common)
  std::vector<Instance> in;
  bool batched = false;
  hipStream_t stream = nullptr;
  std::vector<void*> bufs;      // device allocations, freed by engine_free
  fvk::DevMesh dm{};
  bool fp32 = false;            // FP32-mixed: state = float(Q - Q_inf)
  bool lsq32 = false;           // FP32-mixed with FP32 LSQ inverses
  size_t tsize = 8;             // bytes per state value
  void* flows = nullptr;        // FlowT<T>[n]
  int32_t* blk_inst = nullptr;  // batch: per-block instance
  int32_t* active = nullptr;    // batch: active block list
  int32_t nactive = 0;
  // state arrays of the arithmetic type (double, or float perturbations)
  void* q[3] = {nullptr, nullptr, nullptr};   // Qn, Q1, Q2 (4*ncp each)
  void* g[3] = {nullptr, nullptr, nullptr};   // stage 1..3 limited gradients (8*ncp each)
  void* sigma = nullptr;                      // ncp
  void* res = nullptr;                        // 4*ncp
  unsigned long long* smax = nullptr;         // [2][n], FP64 bits
  double* time = nullptr;                     // [n] physical time of the uniform dt
  double* partial = nullptr;                  // [10][nblocks]: residual and conservation sums
  int32_t* inst_blocks = nullptr;             // [n+1]: first mesh block of each instance
  double* norms = nullptr;                    // [4][n]: residual norms
  int device = 0;                             // the HIP device the context was created on
  int block = 256;
  int variant = fvk::kVariantLds;
  int smax_cur = 0;
  bool have_state = false;
  bool sentinel = true;
  bool fs_prims = false;
  bool res_valid = false;
  int32_t stall_checks = 0;                   // run_to_drop stall retirement (0: off)
  double stall_factor = 0.5;
  bool use_graph = true;
  int graph_steps = 2;                        // steps per captured graph (even)
  hipGraphExec_t graph = nullptr;             // graph_steps steps, captured for `graph_nactive`
  int32_t graph_nactive = -1;
};

}  // namespace

struct fvk_ctx {
  Engine e;
};

struct fvk_batch {
  Engine e;
};

namespace {

// automatic launch options (measured on MI355X: 2k-200k-cell single meshes,
// a 7.8M-cell mesh and an 11.3M-cell batch of 256 meshes)
constexpr int64_t kAutoLargeCells = 2000000;    // FP32-mixed: 128-cell blocks above this
constexpr int32_t kAutoContiguousBlocks = 4096; // contiguous XCD ranges up to this many blocks
// steps per captured graph by padded cells in the solve: 2 from this size up,
constexpr int64_t kGraphTwoStepCells = 262144;
// 8 from this size up, 32 below
constexpr int64_t kGraphEightStepCells = 16384;
// FP32-mixed: FP32 LSQ inverses when no cell's normal matrix has a condition
// number above this (the gradient's relative error then stays below ~6e-6)
constexpr double kLsq32MaxCond = 100.0;

// host threads of the mesh preparation: FVK_PREP_THREADS (1..64) when set,
// else the hardware threads, at most 16
int prep_threads() {
  if (const char* v = std::getenv("FVK_PREP_THREADS")) {
    const int t = std::atoi(v);
    if (t >= 1 && t <= 64) return t;
  }
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(16, hw));
}

// instance i's options with the engine-wide launch choices resolved
fvk_options resolved_options(const Engine& c, int32_t i) {
  fvk_options o = c.in[i].opt;
  o.block_size = c.dm.blk;
  o.xcd_remap = c.opt.xcd_remap;
  o.kernel_variant = c.variant;
  return o;
}

int set_err(Engine* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_err(Engine* c, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_err(c, FVK_ERR_HIP, m);
}

#define FVK_HIP(ctx, call)                                \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_err(ctx, e_, #call); \
  } while (0)

template <typename T>
int dev_alloc(Engine* c, T** out, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
  if (e != hipSuccess) return hip_err(c, e, "hipMalloc");
  c->bufs.push_back(p);
  *out = static_cast<T*>(p);
  return FVK_OK;
}

// Concatenate one HostMesh field over the instances into a new device array.
// `adjust(instance, x)` rewrites a value (global indices get the instance's
// offsets); prefix arrays (`prefix`) drop every instance's last entry except
// the last instance's.
template <typename T, typename Vec, typename Adjust>
int upload_cat(Engine* c, Vec fvk::HostMesh::*field, const T** out, Adjust adjust,
               bool prefix = false) {
  const size_t n = c->in.size();
  auto count = [&](size_t i) {
    const size_t sz = (c->in[i].hm.*field).size();
    return (prefix && i + 1 < n) ? sz - 1 : sz;
  };
  size_t total = 0;
  for (size_t i = 0; i < n; ++i) total += count(i);
  T* d = nullptr;
  int rc = dev_alloc(c, &d, total);
  if (rc) return rc;
  size_t off = 0;
  std::vector<T> tmp;
  for (size_t i = 0; i < n; ++i) {
    const Vec& v = c->in[i].hm.*field;
    const size_t k = count(i);
    const T* src = v.data();
    if (!std::is_same<Adjust, std::nullptr_t>::value) {   // copy, rewriting the values
      tmp.assign(v.begin(), v.begin() + k);
      if constexpr (!std::is_same<Adjust, std::nullptr_t>::value)
        for (T& x : tmp) x = adjust(c->in[i], x);
      src = tmp.data();
    }
    if (k) {
      hipError_t e = hipMemcpy(d + off, src, k * sizeof(T), hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_err(c, e, "hipMemcpy H2D");
    }
    off += k;
  }
  *out = d;
  return FVK_OK;
}

// values unchanged: straight from the host array
template <typename T, typename Vec>
int upload_cat(Engine* c, Vec fvk::HostMesh::*field, const T** out) {
  return upload_cat(c, field, out, nullptr);
}

// boundary code ~((src << 2) | kind) with the ghost source cell moved by `off`
inline int32_t shift_boundary(int32_t code, int32_t off) { return ~((~code) + (off << 2)); }

// stages per step: 3 (SSP_RK3, and SSP_RK2, which the reference runs as
// SSP-RK3, main.cpp:123-126), or 2 for the two-stage SSP-RK2 extension
inline int stages_of(const fvk_options& o) { return o.time_scheme == 3 ? 2 : 3; }

// the LDS-staged shapes: one workgroup per block, or a persistent grid
inline bool lds_staged(int variant) {
  return variant == fvk::kVariantLds || variant == fvk::kVariantPersist;
}

int check_options(const fvk_options& o, std::string* msg) {
  if (o.order != 1 && o.order != 2) { *msg = "order must be 1 or 2 (solution_poly_order 0 or 1)"; return FVK_ERR_OPTION; }
  if (o.limiter < 0 || o.limiter > 3) { *msg = "limiter must be NONE, MINMOD, SQUEEZE or 3 (Barth-Jespersen)"; return FVK_ERR_OPTION; }
  if (o.eqn != 1 && o.eqn != 2) { *msg = "eqn must be EULER or NS"; return FVK_ERR_OPTION; }
  if (o.flux_solver != 1 && o.flux_solver != 2) { *msg = "flux_solver must be RUSANOV or ROE"; return FVK_ERR_OPTION; }
  if (o.flux_solver == 2 && ((o.kernel_variant != 0 && !lds_staged(o.kernel_variant)) || o.accumulate)) {
    *msg = "flux_solver ROE runs in the LDS-staged gather pass B only (kernel_variant 0, 3 or 4, accumulate 0)";
    return FVK_ERR_OPTION;
  }
  if (o.time_scheme < 1 || o.time_scheme > 3) { *msg = "time_scheme must be SSP_RK2, SSP_RK3 or 3 (two-stage SSP-RK2)"; return FVK_ERR_OPTION; }
  if (o.time_step_type < 1 || o.time_step_type > 3) { *msg = "time_step_type must be GLOBAL, LOCAL or 3 (uniform)"; return FVK_ERR_OPTION; }
  if (o.bc1 < -4 || o.bc1 > -1 || o.bc2 < -4 || o.bc2 > -1) { *msg = "BC1/BC2 must be FREESTREAM, SLIP_WALL, EXTRAPOLATION or NO_SLIP_WALL"; return FVK_ERR_OPTION; }
  if (o.precision != FVK_FP64 && o.precision != FVK_FP32_MIXED) { *msg = "precision must be FVK_FP64 or FVK_FP32_MIXED"; return FVK_ERR_OPTION; }
  if (o.kernel_variant < 0 || o.kernel_variant > 4) { *msg = "kernel_variant must be 0, 1, 2, 3 or 4"; return FVK_ERR_OPTION; }
  if (o.renumber < 0 || o.renumber > 2) { *msg = "renumber must be 0, 1 or 2"; return FVK_ERR_OPTION; }
  if (o.accumulate != 0 && o.accumulate != 1) { *msg = "accumulate must be 0 (gather) or 1 (atomic scatter)"; return FVK_ERR_OPTION; }
  if (o.accumulate == 1 && o.kernel_variant != 0 && !lds_staged(o.kernel_variant)) {
    *msg = "accumulate = 1 needs the LDS-staged kernels (kernel_variant 0, 3 or 4)";
    return FVK_ERR_OPTION;
  }
  if (o.xcd_remap < -1 || o.xcd_remap > 4096) { *msg = "xcd_remap must be -1 (auto), 0, 1 or a run length 2..4096"; return FVK_ERR_OPTION; }
  if (o.block_size != 0 && (o.block_size < 64 || o.block_size > 256 || o.block_size % 64)) {
    *msg = "block_size must be 0 or a multiple of 64 in [64, 256]";
    return FVK_ERR_OPTION;
  }
  if (!(o.rho > 0) || !(o.P > 0) || !(o.gamma > 1)) { *msg = "freestream needs rho > 0, P > 0, gamma > 1"; return FVK_ERR_OPTION; }
  return FVK_OK;
}

// The options that select kernels or the launch shape must agree across the
// instances of a batch; the flow conditions, gas constants, CFL, limiter
// switch, BC kinds and renumbering may differ.
int check_common(const fvk_options& a, const fvk_options& b, std::string* msg) {
  const bool same = a.order == b.order && a.eqn == b.eqn && a.time_step_type == b.time_step_type &&
                    a.precision == b.precision && a.block_size == b.block_size &&
                    a.kernel_variant == b.kernel_variant && a.xcd_remap == b.xcd_remap &&
                    a.accumulate == b.accumulate && a.flux_solver == b.flux_solver &&
                    stages_of(a) == stages_of(b);
  if (!same) {
    *msg = "batch instances must agree on order, eqn, time_step_type, precision, block_size, "
           "kernel_variant, xcd_remap, accumulate, flux_solver and the stage count of time_scheme";
    return FVK_ERR_OPTION;
  }
  if (a.kernel_variant == fvk::kVariantCell || a.kernel_variant == fvk::kVariantSlot) {
    *msg = "batches run the LDS-staged kernels only (kernel_variant 0, 3 or 4)";
    return FVK_ERR_OPTION;
  }
  return FVK_OK;
}

// class_flow (class_flow.hpp:28-46) and class_Q constants in type T
template <typename T>
fvk::FlowT<T> make_flow(const fvk_options& o, bool delta, int32_t first_block) {
  fvk::FlowT<T> fl{};
  fl.gamma = T(o.gamma);
  fl.gm1 = T(o.gamma - 1);
  fl.Rgas = T(o.R);
  fl.mu = T(o.mu);
  const double Cp = o.R * o.gamma / (o.gamma - 1);
  fl.k = T((o.mu * Cp) / o.Pr);
  fl.cfl2 = T(2 * o.cfl);
  const double qinf[4] = {o.rho, o.rho * o.u, o.rho * o.v,
                          (o.P / (o.gamma - 1)) + 0.5 * o.rho * (o.u * o.u + o.v * o.v)};
  for (int k = 0; k < 4; ++k) {
    fl.qinf[k] = T(qinf[k]);
    // the reference's -101 initial face states (class_q.hpp:59-60)
    fl.sent[k] = T(delta ? -101.0 - qinf[k] : -101.0);
  }
  fl.rho = T(o.rho);
  fl.u = T(o.u);
  fl.v = T(o.v);
  fl.P = T(o.P);
  fl.vn0 = T(std::sqrt(o.v * o.v + o.u * o.u));
  fl.c0 = T(std::sqrt(o.gamma * (o.P / o.rho)));
  for (int k = 0; k < 4; ++k) fl.qinf64[k] = qinf[k];
  fl.gm1_64 = o.gamma - 1;
  fl.Rgas64 = o.R;
  fl.mu64 = o.mu;
  fl.k64 = (o.mu * Cp) / o.Pr;
  fl.squeeze = o.limiter == 2 || o.limiter == 3;
  fl.bj = o.limiter == 3;
  fl.first_block = first_block;
  return fl;
}

template <typename T>
int upload_flows(Engine* c) {
  std::vector<fvk::FlowT<T>> h;
  for (const Instance& i : c->in) h.push_back(make_flow<T>(i.opt, c->fp32, i.block_off));
  fvk::FlowT<T>* d = nullptr;
  int rc = dev_alloc(c, &d, h.size());
  if (rc) return rc;
  FVK_HIP(c, hipMemcpy(d, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
  c->flows = d;
  return FVK_OK;
}

template <typename T>
fvk::StageArgsT<T> base_args(Engine* c) {
  fvk::StageArgsT<T> a{};
  a.m = c->dm;
  a.flows = static_cast<const fvk::FlowT<T>*>(c->flows);
  a.blk_inst = c->blk_inst;
  a.active = c->active;
  a.nactive = c->nactive;
  a.dt_global = c->opt.time_step_type == 1 || c->opt.time_step_type == 3;
  a.dt_uniform = c->opt.time_step_type == 3;
  a.time = c->time;
  a.xcd_remap = c->opt.xcd_remap;
  a.variant = c->variant;
  a.roe = c->opt.flux_solver == 2;
  a.sigma = static_cast<T*>(c->sigma);
  a.res = static_cast<T*>(c->res);
  a.accumulate = c->opt.accumulate;
  const size_t n = c->in.size();
  a.smax_cur = c->smax + c->smax_cur * n;
  a.smax_next = c->smax + (1 - c->smax_cur) * n;
  return a;
}

// stage roles of the three state and gradient buffers (ssp_rk3.cpp:38-76):
// stage s reads Q_{s-1} and the previous stage's Q and limited gradient (the
// lagged limiter), writes G_s and Q_s; stage 3 writes Q_n in place.  The
// two-stage SSP-RK2 (time_scheme 3) writes Q* to buffer 1 and Q_n+1 in place;
// its stage 1 takes the lagged data of the previous step's stage 2 (Q*, G_2).
template <typename T>
void set_roles(Engine* c, int stage, bool rk2, fvk::StageArgsT<T>& a) {
  T* const* q = reinterpret_cast<T* const*>(c->q);
  T* const* g = reinterpret_cast<T* const*>(c->g);
  a.qn = q[0];
  if (rk2) {
    if (stage == 1) { a.qc = q[0]; a.qp = q[1]; a.gp = g[1]; a.g = g[0]; a.qout = q[1]; }
    if (stage == 2) { a.qc = q[1]; a.qp = q[0]; a.gp = g[0]; a.g = g[1]; a.qout = q[0]; }
    return;
  }
  if (stage == 1) { a.qc = q[0]; a.qp = q[2]; a.gp = g[2]; a.g = g[0]; a.qout = q[1]; }
  if (stage == 2) { a.qc = q[1]; a.qp = q[0]; a.gp = g[0]; a.g = g[1]; a.qout = q[2]; }
  if (stage == 3) { a.qc = q[2]; a.qp = q[1]; a.gp = g[1]; a.g = g[2]; a.qout = q[0]; }
}



// the stage combination Q_out = ca Q_n + cb (Q + dt R) of stages 2 and 3
template <typename T>
void stage_coefficients(int nst, int stage, fvk::StageArgsT<T>& a) {
  if (nst == 2) { a.ca = T(0.5); a.cb = T(0.5); }
  else if (stage == 2) { a.ca = T(0.75); a.cb = T(0.25); }
  else { a.ca = T(1.0 / 3.0); a.cb = T(2.0 / 3.0); }
}

template <typename T>
int fill_t(Engine* c, void* p, int64_t n, double v) {
  hipError_t e = fvk::launch_fill(static_cast<T*>(p), n, T(v), c->stream);
  return e == hipSuccess ? FVK_OK : hip_err(c, e, "fill");
}

int fill(Engine* c, void* p, int64_t n, double v) {
  return c->fp32 ? fill_t<float>(c, p, n, v) : fill_t<double>(c, p, n, v);
}

void* elem(Engine* c, void* base, size_t index) {
  return static_cast<unsigned char*>(base) + index * c->tsize;
}

bool any_squeeze(const Engine* c) {
  for (const Instance& i : c->in)
    if (i.opt.limiter == 2 || i.opt.limiter == 3) return true;
  return false;
}

// Spectral sums |V_inf| + c_inf on every face and their max: the lagged face
// data of a fresh start (class_q.hpp:61-62, main.cpp:117-118).
int arm_start(Engine* c) {
  c->smax_cur = 0;
  FVK_HIP(c, fvk::launch_fill(reinterpret_cast<double*>(c->smax), 2 * (int64_t)c->in.size(), 0.0,
                              c->stream));
  FVK_HIP(c, fvk::launch_fill(c->time, (int64_t)c->in.size(), 0.0, c->stream));
  if (c->fp32) FVK_HIP(c, fvk::launch_init_sigma(base_args<float>(c), c->stream));
  else FVK_HIP(c, fvk::launch_init_sigma(base_args<double>(c), c->stream));
  c->sentinel = true;
  c->res_valid = false;
  c->have_state = true;
  for (Instance& i : c->in) i.steps = 0;
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  return FVK_OK;
}

enum { kKindA = 1, kKindB = 2 };

// One ssp_rk3 call.  launch(kind, fn) lets fvk_time_steps bracket launches.
template <typename T, typename Launch>
int run_step_t(Engine* c, bool store_res, Launch&& launch) {
  const fvk_options& o = c->opt;
  // the SQUEEZE pass A when any instance limits; per-instance switch inside
  const bool squeeze = any_squeeze(c);
  const int nst = stages_of(o);
  for (int stage = 1; stage <= nst; ++stage) {
    fvk::StageArgsT<T> a = base_args<T>(c);
    a.stage = stage;
    a.last = stage == nst ? 1 : 0;
    stage_coefficients(nst, stage, a);
    a.sentinel = (stage == 1 && c->sentinel) ? 1 : 0;
    a.fs_prims = (stage == 1 && c->fs_prims) ? 1 : 0;
    a.store_res = (a.last && store_res) ? 1 : 0;
    set_roles(c, stage, nst == 2, a);
    if (o.order == 2) {
      int rc = launch(kKindA, [&]() { return fvk::launch_pass_a(a, squeeze, c->block, c->stream); });
      if (rc) return rc;
    }
    int rc = launch(kKindB, [&]() { return fvk::launch_pass_b(a, o.order, o.eqn == 2, c->block, c->stream); });
    if (rc) return rc;
  }
  c->sentinel = false;
  c->fs_prims = false;
  c->smax_cur = 1 - c->smax_cur;
  if (store_res) c->res_valid = true;
  return FVK_OK;
}

template <typename Launch>
int run_step(Engine* c, bool store_res, Launch&& launch) {
  return c->fp32 ? run_step_t<float>(c, store_res, launch) : run_step_t<double>(c, store_res, launch);
}

int plain_step(Engine* c, bool store_res) {
  auto plain = [c](int, auto&& fn) -> int {
    hipError_t e = fn();
    return e == hipSuccess ? FVK_OK : hip_err(c, e, "kernel launch");
  };
  int rc = run_step(c, store_res, plain);
  if (rc) return rc;
  for (Instance& i : c->in)
    if (i.active) ++i.steps;
  return FVK_OK;
}

void drop_graph(Engine* c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_nactive = -1;
}

// Capture graph_steps steps (an even count, so the graph starts and ends on
// dt-maximum parity 0) into a graph.  The launches read everything that
// changes from device memory (states, spectral sums, dt maxima, the active
// block list), so the graph stays valid across restarts and checkpoint loads
// until the number of active blocks changes.
int capture_graph(Engine* c) {
  drop_graph(c);
  hipGraph_t g = nullptr;
  FVK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  auto plain = [](int, auto&& fn) -> int { return fn() == hipSuccess ? FVK_OK : FVK_ERR_HIP; };
  int rc = FVK_OK;
  for (int k = 0; k < c->graph_steps && !rc; ++k) rc = run_step(c, false, plain);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return set_err(c, FVK_ERR_HIP, "kernel launch during graph capture failed");
  }
  if (e != hipSuccess) return hip_err(c, e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return hip_err(c, e, "hipGraphInstantiate");
  }
  c->graph_nactive = c->nactive;
  return FVK_OK;
}

// nsteps steps of every active instance.  Steps are replayed from the
// captured graph wherever a step is not the first after a start, starts on
// parity 0 and leaves at least one step for the end of the call.
int engine_step(Engine* c, int32_t nsteps) {
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "step: no state (initialise it first)");
  if (nsteps < 0) return set_err(c, FVK_ERR_ARG, "step: negative step count");
  if (c->batched && c->nactive == 0) return FVK_OK;   // every instance retired
  int32_t left = nsteps;
  while (left > 0) {
    // the graph's steps keep no residual: the call's last step (eager) does
    const bool graphable = c->use_graph && !c->sentinel && !c->fs_prims && c->smax_cur == 0 &&
                           left > c->graph_steps;
    if (!graphable) {
      int rc = plain_step(c, left == 1);
      if (rc) return rc;
      --left;
      continue;
    }
    if (!c->graph || c->graph_nactive != c->nactive) {
      int rc = capture_graph(c);
      if (rc) return rc;
    }
    FVK_HIP(c, hipGraphLaunch(c->graph, c->stream));
    for (Instance& i : c->in)
      if (i.active) i.steps += c->graph_steps;
    left -= c->graph_steps;
  }
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  return FVK_OK;
}

// instance i's values of a cell-major [ncp][nrows] device array of the
// arithmetic type, as doubles [nrows][ncp_i] in the instance's padded numbering
int download_rows(Engine* c, const void* src, int nrows, size_t i, std::vector<double>* out) {
  const Instance& in = c->in[i];
  const size_t n = in.hm.ncp, count = n * nrows;
  const void* p = static_cast<const unsigned char*>(src) + (size_t)in.cell_off * nrows * c->tsize;
  out->resize(count);
  if (c->fp32) {
    std::vector<float> f(count);
    FVK_HIP(c, hipMemcpy(f.data(), p, count * 4, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < n; ++j)
      for (int k = 0; k < nrows; ++k) (*out)[k * n + j] = f[j * nrows + k];
  } else {
    std::vector<double> d(count);
    FVK_HIP(c, hipMemcpy(d.data(), p, count * 8, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < n; ++j)
      for (int k = 0; k < nrows; ++k) (*out)[k * n + j] = d[j * nrows + k];
  }
  return FVK_OK;
}

// unpermute [nrows][ncp_i] into the caller's [nrows][nc] (grid-file order),
// adding add[k] to row k when given
void unpermute(const fvk::HostMesh& h, const std::vector<double>& buf, int nrows, double* out,
               const double* add = nullptr) {
  const size_t n = h.ncp;
  for (int k = 0; k < nrows; ++k)
    for (int32_t j = 0; j < h.ncp; ++j)
      if (h.perm[j] >= 0)
        out[k * (size_t)h.nc + h.perm[j]] = add ? add[k] + buf[k * n + j] : buf[k * n + j];
}

int get_state(Engine* c, size_t i, double* q, double* grad) {
  if (!q && !grad) return set_err(c, FVK_ERR_ARG, "get_state: nothing requested");
  const Instance& in = c->in[i];
  std::vector<double> buf;
  if (q) {
    int rc = download_rows(c, c->q[0], 4, i, &buf);
    if (rc) return rc;
    unpermute(in.hm, buf, 4, q, c->fp32 ? in.qinf : nullptr);
  }
  if (grad) {
    int rc = download_rows(c, c->g[0], 8, i, &buf);
    if (rc) return rc;
    unpermute(in.hm, buf, 8, grad);
  }
  return FVK_OK;
}

// store instance i's conserved state q[4*nc] (grid-file order), or its
// freestream when q is NULL
int put_state(Engine* c, size_t i, const double* q) {
  const Instance& in = c->in[i];
  const fvk::HostMesh& h = in.hm;
  const size_t n = h.ncp;
  void* dst = elem(c, c->q[0], 4 * (size_t)in.cell_off);
  if (c->fp32) {
    // the perturbation from freestream (0 for the freestream start)
    std::vector<float> buf(4 * n, 0.0f);
    if (q)
      for (int32_t j = 0; j < h.ncp; ++j)
        if (h.perm[j] >= 0)
          for (int k = 0; k < 4; ++k)
            buf[4 * (size_t)j + k] = float(q[k * (size_t)h.nc + h.perm[j]] - in.qinf[k]);
    FVK_HIP(c, hipMemcpy(dst, buf.data(), buf.size() * 4, hipMemcpyHostToDevice));
  } else {
    // padding cells hold Q_inf (or 1.0 after a restart); they are never read
    std::vector<double> buf(4 * n);
    for (int32_t j = 0; j < h.ncp; ++j)
      for (int k = 0; k < 4; ++k)
        buf[4 * (size_t)j + k] = !q ? in.qinf[k] : h.perm[j] >= 0 ? q[k * (size_t)h.nc + h.perm[j]] : 1.0;
    FVK_HIP(c, hipMemcpy(dst, buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
  }
  return FVK_OK;
}

// Neumaier-compensated sum of n values
struct Neumaier {
  double s = 0.0, comp = 0.0;
  void add(double x) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) comp += (s - t) + x;
    else comp += (x - t) + s;
    s = t;
  }
  double value() const { return s + comp; }
};

// sqrt(sum R_k^2) of every instance's last stored residual: out[4*i + k].
// Per-block partial sums, then per instance a compensated sum of its block
// partials (k_res_norm), both on the device; only the 4 norms per instance
// cross to the host.
int residual_l2(Engine* c, double* out) {
  const int32_t nb = c->dm.nblocks, n = (int32_t)c->in.size();
  if (c->fp32)
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const float*>(c->res), c->dm.ncp, c->dm.blk, c->partial, c->stream));
  else
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const double*>(c->res), c->dm.ncp, c->dm.blk, c->partial, c->stream));
  FVK_HIP(c, fvk::launch_res_norm_inst(c->partial, nb, c->inst_blocks, n, c->norms, c->stream));
  FVK_HIP(c, hipMemcpyAsync(out, c->norms, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  for (int32_t i = 0; i < 4 * n; ++i)
    if (!std::isfinite(out[i])) return set_err(c, FVK_ERR_NONFINITE, "residual norm is not finite");
  return FVK_OK;
}

// Global conservation sums of every instance's current state:
// out[4*i + k] = sum over its cells of V * Q_k, compensated (per-block
// TwoSum pairs on the device, a Neumaier sum of the pairs here).  FP32-mixed:
// Q_inf,k * sum(V) + sum(V * dQ_k), the product and the final sum in
// extended precision.
int conservation_sums(Engine* c, double* out) {
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "conservation sums: no state");
  const int32_t nb = c->dm.nblocks;
  if (c->fp32)
    FVK_HIP(c, fvk::launch_cons_sum(static_cast<const float*>(c->q[0]), c->dm, c->partial, c->stream));
  else
    FVK_HIP(c, fvk::launch_cons_sum(static_cast<const double*>(c->q[0]), c->dm, c->partial, c->stream));
  std::vector<double> part(10 * (size_t)nb);
  FVK_HIP(c, hipMemcpyAsync(part.data(), c->partial, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < c->in.size(); ++i) {
    const Instance& in = c->in[i];
    double sum[5];
    for (int j = 0; j < 5; ++j) {
      Neumaier acc;
      for (int32_t b = in.block_off; b < in.block_off + in.hm.nblocks; ++b) {
        acc.add(part[(2 * j) * (size_t)nb + b]);
        acc.add(part[(2 * j + 1) * (size_t)nb + b]);
      }
      sum[j] = acc.value();
    }
    for (int k = 0; k < 4; ++k)
      out[4 * i + k] = c->fp32 ? double((long double)in.qinf[k] * sum[4] + (long double)sum[k]) : sum[k];
  }
  return FVK_OK;
}

// the active block list from the instances' active flags
int upload_active(Engine* c) {
  std::vector<int32_t> list;
  for (const Instance& in : c->in)
    if (in.active)
      for (int32_t b = 0; b < in.hm.nblocks; ++b) list.push_back(in.block_off + b);
  c->nactive = (int32_t)list.size();
  if (!list.empty())
    FVK_HIP(c, hipMemcpy(c->active, list.data(), list.size() * 4, hipMemcpyHostToDevice));
  return FVK_OK;
}

// `reps` back-to-back launches of the stage-2 pass A and pass B kernels, each
// kind between one event pair.  Stage 2 reads Q1, Qn, G1 and writes G2, Q2, so
// the repeats leave the state unchanged.
template <typename T>
int time_kernels_t(Engine* c, int32_t reps, double avg_us[2]) {
  const fvk_options& o = c->opt;
  fvk::StageArgsT<T> a = base_args<T>(c);
  a.stage = 2;
  a.last = 0;
  stage_coefficients(3, 2, a);
  set_roles(c, 2, false, a);
  if (stages_of(o) == 2) {
    // the two-stage scheme's lagged data lives in buffers 0 and 1: the timed
    // launches write buffers 2 only, which it does not use
    T* const* g = reinterpret_cast<T* const*>(c->g);
    a.g = g[2];
  }
  const bool squeeze = any_squeeze(c);
  hipEvent_t e0, e1;
  FVK_HIP(c, hipEventCreate(&e0));
  FVK_HIP(c, hipEventCreate(&e1));
  avg_us[0] = avg_us[1] = 0.0;
  int rc = FVK_OK;
  for (int kind = 0; kind < 2 && !rc; ++kind) {
    if (kind == 0 && o.order != 2) continue;
    hipError_t e = hipEventRecord(e0, c->stream);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r)
      e = kind == 0 ? fvk::launch_pass_a(a, squeeze, c->block, c->stream)
                    : fvk::launch_pass_b(a, o.order, o.eqn == 2, c->block, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = hip_err(c, e, "time_kernels");
    else avg_us[kind] = ms * 1e3 / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int time_kernels(Engine* c, int32_t reps, double avg_us[2]) {
  if (!avg_us || reps <= 0) return set_err(c, FVK_ERR_ARG, "time_kernels: bad arguments");
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "time_kernels: no state");
  return c->fp32 ? time_kernels_t<float>(c, reps, avg_us) : time_kernels_t<double>(c, reps, avg_us);
}

void engine_free(Engine* c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (void* p : c->bufs) (void)hipFree(p);
  c->bufs.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

// Host preparation of every instance (in parallel), then the concatenated
// device layout, state buffers and per-instance constants.
int engine_init(Engine* c, int32_t n, const fvk_mesh* meshes, const fvk_options* opts, bool batched) {
  std::string msg;
  for (int32_t i = 0; i < n; ++i) {
    int rc = check_options(opts[i], &msg);
    if (!rc && batched) rc = check_common(opts[0], opts[i], &msg);
    if (rc) return set_err(nullptr, rc, batched ? "instance " + std::to_string(i) + ": " + msg : msg);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(nullptr, FVK_ERR_HIP, "no HIP device");
  // automatic launch options, resolved once for the whole solve (a batch
  // counts all its instances): 128-cell blocks for FP64 and for large solves,
  // where the smaller LDS footprint keeps more blocks resident per CU
  std::vector<fvk_options> ro(opts, opts + n);
  int64_t total_cells = 0;
  for (int32_t i = 0; i < n; ++i) total_cells += meshes[i].num_cells;
  for (fvk_options& o : ro)
    if (o.block_size == 0)
      o.block_size = (o.precision == FVK_FP64 || total_cells > kAutoLargeCells) ? 128 : 256;
  c->opt = ro[0];
  c->batched = batched;
  c->in.resize(n);
  {
    std::vector<int> rcs(n, FVK_OK);
    std::vector<std::string> errs(n);
    std::atomic<int32_t> next{0};
    const int pool_size = prep_threads();
    const int nthreads = std::max(1, std::min<int>(n, pool_size));
    // threads inside each instance's preparation: all of them for one mesh
    const int inner = std::max(1, pool_size / nthreads);
    auto work = [&]() {
      for (int32_t i = next++; i < n; i = next++) {
        c->in[i].opt = ro[i];
        rcs[i] = fvk::prepare_mesh(meshes[i], ro[i], &c->in[i].hm, &errs[i], inner);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (int32_t i = 0; i < n; ++i)
      if (rcs[i])
        return set_err(nullptr, rcs[i], batched ? "instance " + std::to_string(i) + ": " + errs[i] : errs[i]);
  }
  int64_t ncells = 0;
  for (Instance& in : c->in) {
    if (!in.hm.warnings.empty()) fputs(in.hm.warnings.c_str(), stderr);
    const fvk_options& o = in.opt;
    in.qinf[0] = o.rho;
    in.qinf[1] = o.rho * o.u;
    in.qinf[2] = o.rho * o.v;
    in.qinf[3] = (o.P / (o.gamma - 1)) + 0.5 * o.rho * (o.u * o.u + o.v * o.v);
    ncells += in.hm.ncp;
  }
  if (ncells >= (int64_t(1) << 29)) return set_err(nullptr, FVK_ERR_MESH, "more than 2^29 cells in total");
  {
    // the concatenated layout's offsets are 32-bit on the device
    int64_t cell = 0, block = 0, slot = 0, halo = 0, bface = 0;
    for (Instance& in : c->in) {
      in.cell_off = (int32_t)cell; in.block_off = (int32_t)block; in.slot_off = (int32_t)slot;
      in.halo_off = (int32_t)halo; in.bface_off = (int32_t)bface;
      cell += in.hm.ncp; block += in.hm.nblocks; slot += in.hm.chunk_off.back();
      halo += (int64_t)in.hm.halo_idx.size(); bface += (int64_t)in.hm.bf_a.size();
      if (std::max({slot, halo, bface}) >= (int64_t(1) << 31))
        return set_err(nullptr, FVK_ERR_MESH, "more than 2^31 face slots, halo entries or block "
                                              "faces in total");
    }
  }
  const fvk_options& o0 = opts[0];
  c->fp32 = o0.precision == FVK_FP32_MIXED;
  c->tsize = c->fp32 ? sizeof(float) : sizeof(double);
  c->block = o0.block_size ? o0.block_size : 256;   // thread block of the cell-parallel shape
  c->variant = o0.kernel_variant ? o0.kernel_variant : fvk::kVariantLds;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return hip_err(nullptr, e, "hipGetDevice");
  e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    c->stream = nullptr;
    return hip_err(nullptr, e, "hipStreamCreate");
  }
  fvk::DevMesh& d = c->dm;
  d.nc = 0;
  d.ncp = (int32_t)ncells;
  d.blk = c->in[0].hm.blk;
  d.nblocks = 0;
  d.max_halo = 0;
  d.max_bfaces = 0;
  d.max_deg = 0;
  for (const Instance& in : c->in) {
    d.nc += in.hm.nc;
    d.nblocks += in.hm.nblocks;
    d.max_halo = std::max(d.max_halo, in.hm.max_halo);
    d.max_bfaces = std::max(d.max_bfaces, in.hm.max_bfaces);
    d.max_deg = std::max(d.max_deg, in.hm.max_deg);
  }
  // Steps per captured graph: a graph launch costs the host a few
  // microseconds, which a step of a small mesh (a few blocks per kernel, a few
  // microseconds of device time) does not hide, so small solves capture more
  // steps per graph.
  c->graph_steps = d.ncp >= kGraphTwoStepCells ? 2 : d.ncp >= kGraphEightStepCells ? 8 : 32;
  // XCD order: contiguous ranges while a launch is a few blocks per CU (every
  // block is resident at once), runs of 8 blocks dealt in turn beyond that
  // (ranges of unequal work would leave XCDs idle at the end)
  if (c->opt.xcd_remap < 0) c->opt.xcd_remap = d.nblocks <= kAutoContiguousBlocks ? 1 : 8;
  using HM = fvk::HostMesh;
  int rc;
  if ((rc = upload_cat<int32_t>(c, &HM::chunk_off, &d.chunk_off,
                                [](const Instance& in, int32_t x) { return x + in.slot_off; }, true))) return rc;
  if ((rc = upload_cat(c, &HM::deg, &d.deg))) return rc;
  if ((rc = upload_cat<int32_t>(c, &HM::code, &d.code, [](const Instance& in, int32_t x) {
         return x >= 0 ? x + (in.cell_off << 1) : shift_boundary(x, in.cell_off);
       }))) return rc;
  if ((rc = upload_cat(c, &HM::lcode, &d.lcode))) return rc;    // block-local: no offsets
  if ((rc = upload_cat(c, &HM::sface, &d.sface))) return rc;
  if ((rc = upload_cat<int32_t>(c, &HM::halo_ptr, &d.halo_ptr,
                                [](const Instance& in, int32_t x) { return x + in.halo_off; }, true))) return rc;
  if ((rc = upload_cat<int32_t>(c, &HM::halo_idx, &d.halo_idx,
                                [](const Instance& in, int32_t x) { return x + in.cell_off; }))) return rc;
  if ((rc = upload_cat<int32_t>(c, &HM::bf_ptr, &d.bf_ptr,
                                [](const Instance& in, int32_t x) { return x + in.bface_off; }, true))) return rc;
  if ((rc = upload_cat(c, &HM::bf_a, &d.bf_a))) return rc;
  if ((rc = upload_cat(c, &HM::bf_b, &d.bf_b))) return rc;
  if ((rc = upload_cat(c, &HM::bf_nx, &d.bf_nx))) return rc;
  if ((rc = upload_cat(c, &HM::bf_ny, &d.bf_ny))) return rc;
  if ((rc = upload_cat(c, &HM::bf_area, &d.bf_area))) return rc;
  if ((rc = upload_cat(c, &HM::bf_fcx, &d.bf_fcx))) return rc;
  if ((rc = upload_cat(c, &HM::bf_fcy, &d.bf_fcy))) return rc;
  if ((rc = upload_cat(c, &HM::s_nx, &d.s_nx))) return rc;
  if ((rc = upload_cat(c, &HM::s_ny, &d.s_ny))) return rc;
  if ((rc = upload_cat(c, &HM::s_area, &d.s_area))) return rc;
  if ((rc = upload_cat(c, &HM::s_fcx, &d.s_fcx))) return rc;
  if ((rc = upload_cat(c, &HM::s_fcy, &d.s_fcy))) return rc;
  if ((rc = upload_cat(c, &HM::c_ccx, &d.ccx))) return rc;
  if ((rc = upload_cat(c, &HM::c_ccy, &d.ccy))) return rc;
  if ((rc = upload_cat(c, &HM::c_vol, &d.vol))) return rc;
  // LSQ inverses: FP64, or in the FP32-mixed form the same values rounded
  // to FP32 when every cell's normal matrix is well conditioned
  double cond = 0;
  for (const Instance& in : c->in) cond = std::max(cond, in.hm.max_lsq_cond);
  const char* force64 = std::getenv("FVK_LSQ_FP64");
  c->lsq32 = c->fp32 && cond <= kLsq32MaxCond && !(force64 && std::atoi(force64) == 1);
  d.ixx = d.iyy = d.ixy = nullptr;
  d.ixx32 = d.iyy32 = d.ixy32 = nullptr;
  if (c->lsq32) {
    if ((rc = upload_cat(c, &HM::c_ixx32, &d.ixx32))) return rc;
    if ((rc = upload_cat(c, &HM::c_iyy32, &d.iyy32))) return rc;
    if ((rc = upload_cat(c, &HM::c_ixy32, &d.ixy32))) return rc;
  } else {
    if ((rc = upload_cat(c, &HM::c_ixx, &d.ixx))) return rc;
    if ((rc = upload_cat(c, &HM::c_iyy, &d.iyy))) return rc;
    if ((rc = upload_cat(c, &HM::c_ixy, &d.ixy))) return rc;
  }
  if (c->fp32) {
    // the FP32 kernels' single-precision geometry (mesh_prep.hpp)
    if ((rc = upload_cat(c, &HM::s_dfx32, &d.s_dfx32))) return rc;
    if ((rc = upload_cat(c, &HM::s_dfy32, &d.s_dfy32))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_nx, &d.bf32_nx))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_ny, &d.bf32_ny))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_area, &d.bf32_area))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_d1x, &d.bf32_d1x))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_d1y, &d.bf32_d1y))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_d2x, &d.bf32_d2x))) return rc;
    if ((rc = upload_cat(c, &HM::bf32_d2y, &d.bf32_d2y))) return rc;
    if (o0.eqn == 2) {
      if ((rc = upload_cat(c, &HM::bf32_lx, &d.bf32_lx))) return rc;
      if ((rc = upload_cat(c, &HM::bf32_ly, &d.bf32_ly))) return rc;
      if ((rc = upload_cat(c, &HM::bf32_dist, &d.bf32_dist))) return rc;
    }
  }
  if (lds_staged(c->variant)) {
    // the LDS-staged kernels need a block plus its halo (the largest) in LDS
    int dev = 0;
    hipDeviceProp_t prop;
    if ((e = hipGetDevice(&dev)) != hipSuccess || (e = hipGetDeviceProperties(&prop, dev)) != hipSuccess)
      return hip_err(c, e, "hipGetDeviceProperties");
    const size_t need = std::max(fvk::lds_bytes_pass_a(d, c->tsize),
                                 fvk::lds_bytes_pass_b(d, o0.order, true, c->tsize));
    // (16-bit local codes: blk + max_halo and max_bfaces below 2^14, far
    // above what fits in LDS)
    const bool codes_fit = d.blk + d.max_halo < (1 << 14) && d.max_bfaces < (1 << 14);
    if (need > prop.sharedMemPerBlock || !codes_fit) {
      if (lds_staged(o0.kernel_variant) || batched || o0.accumulate || o0.flux_solver == 2)
        return set_err(c, FVK_ERR_MESH, "mesh blocks need " + std::to_string(need) +
                                            " bytes of LDS, more than the device allows");
      c->variant = fvk::kVariantCell;
    }
  }
  if (batched) {
    std::vector<int32_t> bi;
    for (size_t i = 0; i < c->in.size(); ++i) bi.insert(bi.end(), c->in[i].hm.nblocks, (int32_t)i);
    if ((rc = dev_alloc(c, &c->blk_inst, bi.size()))) return rc;
    FVK_HIP(c, hipMemcpy(c->blk_inst, bi.data(), bi.size() * 4, hipMemcpyHostToDevice));
    if ((rc = dev_alloc(c, &c->active, bi.size()))) return rc;
    if ((rc = upload_active(c))) return rc;
  }
  if ((rc = c->fp32 ? upload_flows<float>(c) : upload_flows<double>(c))) return rc;
  const size_t ncp = d.ncp;
  auto alloc = [&](void** p, size_t count) {
    unsigned char* b = nullptr;
    const int r = dev_alloc(c, &b, count * c->tsize);
    *p = b;
    return r;
  };
  for (int i = 0; i < 3; ++i) {
    if ((rc = alloc(&c->q[i], 4 * ncp))) return rc;
    if ((rc = alloc(&c->g[i], 8 * ncp))) return rc;
  }
  if ((rc = alloc(&c->sigma, ncp))) return rc;
  if ((rc = alloc(&c->res, 4 * ncp))) return rc;
  if ((rc = dev_alloc(c, &c->smax, 2 * c->in.size()))) return rc;
  if ((rc = dev_alloc(c, &c->time, c->in.size()))) return rc;
  if ((rc = fill(c, c->time, (int64_t)c->in.size(), 0.0))) return rc;
  if ((rc = dev_alloc(c, &c->partial, 10 * (size_t)d.nblocks))) return rc;
  if ((rc = dev_alloc(c, &c->norms, 4 * c->in.size()))) return rc;
  {
    std::vector<int32_t> ib;
    for (const Instance& in : c->in) ib.push_back(in.block_off);
    ib.push_back(d.nblocks);
    if ((rc = dev_alloc(c, &c->inst_blocks, ib.size()))) return rc;
    FVK_HIP(c, hipMemcpy(c->inst_blocks, ib.data(), ib.size() * 4, hipMemcpyHostToDevice));
  }
  for (int i = 0; i < 3; ++i) {
    if ((rc = fill(c, c->q[i], 4 * (int64_t)ncp, 0.0))) return rc;
    if ((rc = fill(c, c->g[i], 8 * (int64_t)ncp, 0.0))) return rc;
  }
  if ((rc = fill(c, c->res, 4 * (int64_t)ncp, 0.0))) return rc;
  if ((rc = fill(c, c->sigma, (int64_t)ncp, 0.0))) return rc;
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  return FVK_OK;
}

// ---------------------------------------------------------------------------
// checkpoints: the complete device state, so that a run resumes exactly
// ---------------------------------------------------------------------------
struct CheckpointHeader {
  char magic[8];               // "FVKCKPT2"
  int32_t precision, instances;
  int64_t ncp, nblocks, bytes;
  int32_t flags[6];            // have_state, sentinel, fs_prims, res_valid, smax_cur, pad
  int64_t steps_sum;
  uint64_t flow_hash;          // the instances' flow options (options_hash)
};

// FNV-1a hash of every instance's options that the stored state depends on:
// the freestream (the FP32-mixed state is stored relative to it), the gas
// constants, CFL and the scheme switches (flux, and the stage count, which
// fixes the roles of the stored buffers).  A checkpoint is refused by a
// context whose hash differs.
uint64_t options_hash(const Engine* c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  };
  for (const Instance& in : c->in) {
    const fvk_options& o = in.opt;
    const double d[10] = {in.qinf[0], in.qinf[1], in.qinf[2], in.qinf[3], o.cfl, o.gamma,
                          o.R, o.mu, o.Pr, o.P};
    const int32_t k[9] = {o.order, o.limiter, o.eqn, o.time_step_type, o.bc1, o.bc2, o.precision,
                          o.flux_solver, stages_of(o)};
    mix(d, sizeof(d));
    mix(k, sizeof(k));
  }
  return h;
}

// the device arrays a checkpoint holds, with their byte sizes
std::vector<std::pair<void*, size_t>> checkpoint_arrays(Engine* c) {
  const size_t ncp = c->dm.ncp, t = c->tsize;
  std::vector<std::pair<void*, size_t>> v;
  for (int i = 0; i < 3; ++i) v.push_back({c->q[i], 4 * ncp * t});
  for (int i = 0; i < 3; ++i) v.push_back({c->g[i], 8 * ncp * t});
  v.push_back({c->sigma, ncp * t});
  v.push_back({c->res, 4 * ncp * t});
  v.push_back({c->smax, 2 * c->in.size() * sizeof(unsigned long long)});
  v.push_back({c->time, c->in.size() * sizeof(double)});
  return v;
}

int64_t checkpoint_bytes(Engine* c) {
  int64_t n = sizeof(CheckpointHeader);
  for (auto& a : checkpoint_arrays(c)) n += (int64_t)a.second;
  n += (int64_t)c->in.size() * (1 + 8);   // active flags, step counters
  return n;
}

int checkpoint_save(Engine* c, void* buf, int64_t bytes) {
  const int64_t need = checkpoint_bytes(c);
  if (!buf || bytes < need) return set_err(c, FVK_ERR_ARG, "checkpoint: buffer smaller than the checkpoint size");
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  CheckpointHeader h{};
  std::memcpy(h.magic, "FVKCKPT2", 8);
  h.precision = c->fp32 ? FVK_FP32_MIXED : FVK_FP64;
  h.instances = (int32_t)c->in.size();
  h.ncp = c->dm.ncp;
  h.nblocks = c->dm.nblocks;
  h.bytes = need;
  h.flags[0] = c->have_state;
  h.flags[1] = c->sentinel;
  h.flags[2] = c->fs_prims;
  h.flags[3] = c->res_valid;
  h.flags[4] = c->smax_cur;
  h.flow_hash = options_hash(c);
  unsigned char* p = static_cast<unsigned char*>(buf);
  std::memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  for (auto& a : checkpoint_arrays(c)) {
    FVK_HIP(c, hipMemcpy(p, a.first, a.second, hipMemcpyDeviceToHost));
    p += a.second;
  }
  for (const Instance& in : c->in) *p++ = in.active ? 1 : 0;
  for (const Instance& in : c->in) {
    std::memcpy(p, &in.steps, 8);
    p += 8;
  }
  return FVK_OK;
}

int checkpoint_load(Engine* c, const void* buf, int64_t bytes) {
  if (!buf || bytes < (int64_t)sizeof(CheckpointHeader)) return set_err(c, FVK_ERR_ARG, "checkpoint: buffer too small");
  CheckpointHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, "FVKCKPT2", 8) != 0) return set_err(c, FVK_ERR_ARG, "checkpoint: not a checkpoint");
  if (h.precision != (c->fp32 ? FVK_FP32_MIXED : FVK_FP64) || h.instances != (int32_t)c->in.size() ||
      h.ncp != c->dm.ncp || h.nblocks != c->dm.nblocks || h.bytes != checkpoint_bytes(c) || bytes < h.bytes)
    return set_err(c, FVK_ERR_ARG, "checkpoint: written for another mesh layout, instance count or precision");
  if (h.flow_hash != options_hash(c))
    return set_err(c, FVK_ERR_ARG, "checkpoint: written under other flow options (freestream, gas "
                                   "constants, CFL or scheme switches)");
  const unsigned char* p = static_cast<const unsigned char*>(buf) + sizeof(h);
  for (auto& a : checkpoint_arrays(c)) {
    FVK_HIP(c, hipMemcpy(a.first, p, a.second, hipMemcpyHostToDevice));
    p += a.second;
  }
  for (Instance& in : c->in) in.active = *p++ != 0;
  for (Instance& in : c->in) {
    std::memcpy(&in.steps, p, 8);
    p += 8;
    in.why = in.active ? FVK_RETIRED_NONE : FVK_RETIRED_CALLER;
  }
  c->have_state = h.flags[0] != 0;
  c->sentinel = h.flags[1] != 0;
  c->fs_prims = h.flags[2] != 0;
  c->res_valid = h.flags[3] != 0;
  c->smax_cur = h.flags[4];
  if (c->batched) return upload_active(c);
  return FVK_OK;
}

// ---------------------------------------------------------------------------
// fvk_residual_l2_noalloc: the residual norm with a static device scratch
// array in place of the context's partial-sum and norm buffers.
// Every call on a GPU shares that GPU's copy of the array (calls that
// overlap on one GPU race); each call overwrites the slots it reads.
// ---------------------------------------------------------------------------
std::mutex g_scratch_mu;
double* g_scratch[kScratchDevices] = {};   // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_on_device(Engine* c, double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(c, e, "hipGetDevice");
  if (dev != c->device)
    return set_err(c, FVK_ERR_ARG, "fvk_residual_l2_noalloc: the context's device is not current");
  if (dev < 0 || dev >= kScratchDevices) return set_err(c, FVK_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_err(c, e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return FVK_OK;
}
}  // namespace

extern "C" {

void fvk_default_options(fvk_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->order = 1;
  o->limiter = 0;
  o->eqn = 1;
  o->flux_solver = 1;
  o->time_scheme = 2;
  o->time_step_type = 1;
  o->bc1 = FVK_BC_SLIP_WALL;
  o->bc2 = FVK_BC_FREESTREAM;
  o->cfl = 0.5;
  o->gamma = 1.4;
  o->R = 287.0;
  o->mu = 0.0;
  o->Pr = 0.72;
  o->P = 1e5;
  o->rho = 1.225;
  o->u = 0.0;
  o->v = 0.0;
  o->precision = FVK_FP64;
  o->renumber = 1;
  o->block_size = 0;
  o->xcd_remap = -1;
  o->kernel_variant = 0;
}

const char* fvk_last_error(const fvk_ctx* c) { return c ? c->e.err.c_str() : g_create_error.c_str(); }

int fvk_create(const fvk_mesh* mesh, const fvk_options* opt, fvk_ctx** out) {
  g_create_error.clear();
  if (!mesh || !opt || !out) return set_err(nullptr, FVK_ERR_ARG, "fvk_create: NULL argument");
  *out = nullptr;
  fvk_ctx* c = new fvk_ctx();
  int rc = engine_init(&c->e, 1, mesh, opt, false);
  if (rc) {
    if (!c->e.err.empty()) g_create_error = c->e.err;
    engine_free(&c->e);
    delete c;
    return rc;
  }
  *out = c;
  return FVK_OK;
}

int fvk_destroy(fvk_ctx* c) {
  if (!c) return FVK_OK;
  engine_free(&c->e);
  delete c;
  return FVK_OK;
}

int fvk_init_freestream(fvk_ctx* c) {
  if (!c) return FVK_ERR_ARG;
  int rc = put_state(&c->e, 0, nullptr);
  if (!rc) rc = arm_start(&c->e);
  c->e.fs_prims = true;
  return rc;
}

int fvk_set_state(fvk_ctx* c, const double* q) {
  if (!c) return FVK_ERR_ARG;
  if (!q) return set_err(&c->e, FVK_ERR_ARG, "fvk_set_state: NULL state");
  int rc = put_state(&c->e, 0, q);
  if (!rc) rc = arm_start(&c->e);
  c->e.fs_prims = false;
  return rc;
}

int fvk_step(fvk_ctx* c, int32_t nsteps) {
  if (!c) return FVK_ERR_ARG;
  return engine_step(&c->e, nsteps);
}

int fvk_get_state(fvk_ctx* c, double* q, double* grad) {
  if (!c) return FVK_ERR_ARG;
  return get_state(&c->e, 0, q, grad);
}

int fvk_get_residual(fvk_ctx* c, double* r) {
  if (!c) return FVK_ERR_ARG;
  if (!r) return set_err(&c->e, FVK_ERR_ARG, "fvk_get_residual: NULL output");
  std::vector<double> buf;
  int rc = download_rows(&c->e, c->e.res, 4, 0, &buf);
  if (rc) return rc;
  unpermute(c->e.in[0].hm, buf, 4, r);
  return FVK_OK;
}

int fvk_residual_l2(fvk_ctx* c, double out[4]) {
  if (!c || !out) return FVK_ERR_ARG;
  return residual_l2(&c->e, out);
}

int fvk_residual_l2_noalloc(fvk_ctx* cx, double out[4]) {
  if (!cx || !out) return FVK_ERR_ARG;
  Engine* c = &cx->e;
  const int32_t nb = c->dm.nblocks;
  if (nb > kScratchBlocks)
    return set_err(c, FVK_ERR_ARG, "fvk_residual_l2_noalloc: the mesh has more than 65536 blocks");
  double* ptr = nullptr;
  int rc = scratch_on_device(c, &ptr);
  if (rc) return rc;
  double* norms = ptr + 4 * (size_t)kScratchBlocks;
  if (c->fp32)
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const float*>(c->res), c->dm.ncp, c->dm.blk, ptr, c->stream));
  else
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const double*>(c->res), c->dm.ncp, c->dm.blk, ptr, c->stream));
  FVK_HIP(c, fvk::launch_res_norm(ptr, nb, norms, c->stream));
  FVK_HIP(c, hipMemcpyAsync(out, norms, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(out[k])) return set_err(c, FVK_ERR_NONFINITE, "residual norm is not finite");
  return FVK_OK;
}

int fvk_conservation_sums(fvk_ctx* c, double out[4]) {
  if (!c || !out) return FVK_ERR_ARG;
  return conservation_sums(&c->e, out);
}

int fvk_use_graph(fvk_ctx* c, int32_t on) {
  if (!c) return FVK_ERR_ARG;
  c->e.use_graph = on != 0;
  if (!on) drop_graph(&c->e);
  return FVK_OK;
}

int fvk_get_geometry(fvk_ctx* c, const char* name, double* out) {
  if (!c || !name || !out) return FVK_ERR_ARG;
  const fvk::HostMesh& h = c->e.in[0].hm;
  const fvk::hvec<double>* v = nullptr;
  const std::string n = name;
  if (n == "face_area") v = &h.area;
  else if (n == "face_nx") v = &h.nx;
  else if (n == "face_ny") v = &h.ny;
  else if (n == "face_centerx") v = &h.fcx;
  else if (n == "face_centery") v = &h.fcy;
  else if (n == "cell_vol") v = &h.vol;
  else if (n == "cell_centerx") v = &h.ccx;
  else if (n == "cell_centery") v = &h.ccy;
  else if (n == "Ixx") v = &h.ixx;
  else if (n == "Iyy") v = &h.iyy;
  else if (n == "Ixy") v = &h.ixy;
  if (!v) return set_err(&c->e, FVK_ERR_ARG, "fvk_get_geometry: unknown array " + n);
  std::memcpy(out, v->data(), v->size() * sizeof(double));
  return FVK_OK;
}

int fvk_geometry(const fvk_mesh* mesh, const fvk_options* opt, const char* name, double* out) {
  g_create_error.clear();
  if (!mesh || !opt || !name || !out) return set_err(nullptr, FVK_ERR_ARG, "fvk_geometry: NULL argument");
  fvk_ctx tmp;
  tmp.e.in.resize(1);
  std::string msg;
  int rc = fvk::prepare_mesh(*mesh, *opt, &tmp.e.in[0].hm, &msg, prep_threads());
  if (rc) return set_err(nullptr, rc, msg);
  rc = fvk_get_geometry(&tmp, name, out);
  if (rc) g_create_error = tmp.e.err;
  return rc;
}

int fvk_get_info(fvk_ctx* c, int64_t info[6]) {
  if (!c || !info) return FVK_ERR_ARG;
  const fvk::HostMesh& h = c->e.in[0].hm;
  info[0] = h.nc;
  info[1] = h.nf;
  info[2] = h.nb;
  info[3] = h.chunk_off.empty() ? 0 : h.chunk_off.back();
  info[4] = h.max_deg;
  info[5] = h.num_wall;
  return FVK_OK;
}

// physical time of every instance (uniform dt only)
static int get_time(Engine* c, double* out) {
  if (c->opt.time_step_type != 3)
    return set_err(c, FVK_ERR_OPTION, "physical time is defined only for the uniform time step "
                                      "(time_step_type 3): GLOBAL and LOCAL steps differ per cell");
  FVK_HIP(c, hipMemcpy(out, c->time, c->in.size() * sizeof(double), hipMemcpyDeviceToHost));
  return FVK_OK;
}

int fvk_get_time(fvk_ctx* c, double* t) {
  if (!c || !t) return FVK_ERR_ARG;
  return get_time(&c->e, t);
}

int fvk_get_options(fvk_ctx* c, fvk_options* out) {
  if (!c || !out) return FVK_ERR_ARG;
  *out = resolved_options(c->e, 0);
  return FVK_OK;
}

int fvk_time_steps(fvk_ctx* cx, int32_t nsteps, double out_ms[4], int64_t counts[4]) {
  if (!cx || !out_ms || !counts) return FVK_ERR_ARG;
  Engine* c = &cx->e;
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "fvk_time_steps: no state");
  for (int i = 0; i < 4; ++i) { out_ms[i] = 0; counts[i] = 0; }
  const int nev = 2 * 6 * (nsteps > 0 ? nsteps : 1) + 2;
  std::vector<hipEvent_t> ev(nev);
  for (auto& e : ev) FVK_HIP(c, hipEventCreate(&e));
  std::vector<int> kinds;
  int used = 2;
  FVK_HIP(c, hipEventRecord(ev[0], c->stream));
  auto timed = [&](int kind, auto&& fn) -> int {
    hipError_t e = hipEventRecord(ev[used], c->stream);
    if (e == hipSuccess) e = fn();
    if (e == hipSuccess) e = hipEventRecord(ev[used + 1], c->stream);
    if (e != hipSuccess) return hip_err(c, e, "timed launch");
    kinds.push_back(kind);
    used += 2;
    return FVK_OK;
  };
  int rc = FVK_OK;
  for (int32_t n = 0; n < nsteps && !rc; ++n) rc = run_step(c, n == nsteps - 1, timed);
  if (!rc) {
    FVK_HIP(c, hipEventRecord(ev[1], c->stream));
    FVK_HIP(c, hipEventSynchronize(ev[1]));
    float ms = 0;
    FVK_HIP(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    out_ms[0] = ms;
    for (size_t i = 0; i < kinds.size(); ++i) {
      float t = 0;
      FVK_HIP(c, hipEventElapsedTime(&t, ev[2 + 2 * i], ev[3 + 2 * i]));
      out_ms[kinds[i]] += t;
      counts[kinds[i]] += 1;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int fvk_checkpoint_size(fvk_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return FVK_ERR_ARG;
  *bytes = checkpoint_bytes(&c->e);
  return FVK_OK;
}

int fvk_checkpoint_save(fvk_ctx* c, void* buf, int64_t bytes) {
  if (!c) return FVK_ERR_ARG;
  return checkpoint_save(&c->e, buf, bytes);
}

int fvk_checkpoint_load(fvk_ctx* c, const void* buf, int64_t bytes) {
  if (!c) return FVK_ERR_ARG;
  return checkpoint_load(&c->e, buf, bytes);
}

int fvk_time_kernels(fvk_ctx* c, int32_t reps, double avg_us[2]) {
  if (!c) return FVK_ERR_ARG;
  return time_kernels(&c->e, reps, avg_us);
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

const char* fvk_batch_last_error(const fvk_batch* b) {
  return b ? b->e.err.c_str() : g_create_error.c_str();
}

int fvk_batch_create(int32_t n, const fvk_mesh* meshes, const fvk_options* opts, fvk_batch** out) {
  g_create_error.clear();
  if (n <= 0 || !meshes || !opts || !out) return set_err(nullptr, FVK_ERR_ARG, "fvk_batch_create: bad arguments");
  *out = nullptr;
  fvk_batch* b = new fvk_batch();
  int rc = engine_init(&b->e, n, meshes, opts, true);
  if (rc) {
    if (!b->e.err.empty()) g_create_error = b->e.err;
    engine_free(&b->e);
    delete b;
    return rc;
  }
  *out = b;
  return FVK_OK;
}

int fvk_batch_destroy(fvk_batch* b) {
  if (!b) return FVK_OK;
  engine_free(&b->e);
  delete b;
  return FVK_OK;
}

int fvk_batch_init(fvk_batch* b, const double* const* q) {
  if (!b) return FVK_ERR_ARG;
  Engine* c = &b->e;
  size_t given = 0;
  for (size_t i = 0; q && i < c->in.size(); ++i) given += q[i] != nullptr;
  for (size_t i = 0; i < c->in.size(); ++i) {
    // an instance without a state starts from its freestream
    int rc = put_state(c, i, given ? q[i] : nullptr);
    if (rc) return rc;
    c->in[i].active = true;
    c->in[i].why = FVK_RETIRED_NONE;
  }
  int rc = upload_active(c);
  if (!rc) rc = arm_start(c);
  c->fs_prims = given == 0;
  return rc;
}

int fvk_batch_step(fvk_batch* b, int32_t nsteps) {
  if (!b) return FVK_ERR_ARG;
  return engine_step(&b->e, nsteps);
}

int fvk_batch_set_active(fvk_batch* b, const uint8_t* active) {
  if (!b || !active) return FVK_ERR_ARG;
  Engine* c = &b->e;
  for (size_t i = 0; i < c->in.size(); ++i) {
    c->in[i].active = active[i] != 0;
    c->in[i].why = c->in[i].active ? FVK_RETIRED_NONE : FVK_RETIRED_CALLER;
  }
  return upload_active(c);
}

int fvk_batch_set_stall(fvk_batch* b, int32_t checks, double factor) {
  if (!b) return FVK_ERR_ARG;
  if (checks < 0 || !(factor > 0.0 && factor < 1.0))
    return set_err(&b->e, FVK_ERR_ARG, "fvk_batch_set_stall: checks >= 0 and 0 < factor < 1");
  b->e.stall_checks = checks;
  b->e.stall_factor = factor;
  return FVK_OK;
}

int fvk_batch_set_reference(fvk_batch* b, const double* res0) {
  if (!b) return FVK_ERR_ARG;
  for (size_t i = 0; i < b->e.in.size(); ++i) {
    const double r = res0 ? res0[i] : 0.0;
    b->e.in[i].ref0 = (std::isfinite(r) && r > 0.0) ? r : 0.0;
  }
  return FVK_OK;
}

int fvk_batch_retired(fvk_batch* b, int32_t* why) {
  if (!b || !why) return FVK_ERR_ARG;
  for (size_t i = 0; i < b->e.in.size(); ++i) why[i] = b->e.in[i].why;
  return FVK_OK;
}

int fvk_batch_residual_l2(fvk_batch* b, double* out) {
  if (!b || !out) return FVK_ERR_ARG;
  if (!b->e.res_valid) return set_err(&b->e, FVK_ERR_ARG, "fvk_batch_residual_l2: no step run yet");
  return residual_l2(&b->e, out);
}

int fvk_batch_run_to_drop(fvk_batch* b, double drop, int32_t max_steps, int32_t check_every,
                          int32_t* steps, double* res0, double* res_end) {
  if (!b || !steps || check_every <= 0 || max_steps < 0) return FVK_ERR_ARG;
  Engine* c = &b->e;
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "fvk_batch_run_to_drop: no state");
  const size_t n = c->in.size();
  std::vector<double> r(4 * n), r0(n, -1.0);
  std::vector<int64_t> start(n);
  for (size_t i = 0; i < n; ++i) {
    start[i] = c->in[i].steps;
    c->in[i].since_best = 0;
  }
  int32_t done = 0;
  bool any = c->nactive > 0;
  while (any && done < max_steps) {
    // the first check follows the first step: it fixes each instance's reference norm
    const int32_t k = done == 0 ? 1 : std::min(check_every, max_steps - done);
    int rc = engine_step(c, k);
    if (rc) return rc;
    done += k;
    rc = residual_l2(c, r.data());
    if (rc && rc != FVK_ERR_NONFINITE) return rc;
    bool changed = false;
    any = false;
    for (size_t i = 0; i < n; ++i) {
      Instance& in = c->in[i];
      if (!in.active) continue;
      const double rho_res = r[4 * i];
      if (r0[i] < 0) {
        r0[i] = in.ref0 > 0.0 ? in.ref0 : rho_res;
        in.best = rho_res;
      }
      if (res_end) res_end[i] = rho_res;
      // retire on a non-finite residual (state frozen), on convergence and,
      // when enabled, on a stall: stall_checks checks in a row without the
      // norm falling below stall_factor times the lowest norm that counted
      int why = FVK_RETIRED_NONE;
      if (!std::isfinite(rho_res)) {
        why = FVK_RETIRED_NONFINITE;
      } else if (!(rho_res > drop * r0[i])) {
        why = FVK_RETIRED_DROP;
      } else if (c->stall_checks > 0) {
        if (rho_res < c->stall_factor * in.best) {
          in.best = rho_res;
          in.since_best = 0;
        } else if (++in.since_best >= c->stall_checks) {
          why = FVK_RETIRED_STALL;
        }
      }
      if (why != FVK_RETIRED_NONE) {
        in.active = false;
        in.why = why;
        changed = true;
      } else {
        any = true;
      }
    }
    if (changed && (rc = upload_active(c))) return rc;
  }
  for (size_t i = 0; i < n; ++i) {
    steps[i] = (int32_t)(c->in[i].steps - start[i]);
    if (res0) res0[i] = r0[i];
  }
  return FVK_OK;
}

int fvk_batch_conservation_sums(fvk_batch* b, double* out) {
  if (!b || !out) return FVK_ERR_ARG;
  return conservation_sums(&b->e, out);
}

int fvk_batch_get_state(fvk_batch* b, int32_t i, double* q, double* grad) {
  if (!b || i < 0 || i >= (int32_t)b->e.in.size()) return FVK_ERR_ARG;
  return get_state(&b->e, i, q, grad);
}

int fvk_batch_get_info(fvk_batch* b, int64_t info[6]) {
  if (!b || !info) return FVK_ERR_ARG;
  const Engine& c = b->e;
  int64_t nf = 0, active = 0;
  for (const Instance& in : c.in) {
    nf += in.hm.nf;
    active += in.active;
  }
  info[0] = (int64_t)c.in.size();
  info[1] = c.dm.nc;
  info[2] = nf;
  info[3] = c.dm.nblocks;
  info[4] = c.nactive;
  info[5] = active;
  return FVK_OK;
}

int fvk_batch_get_time(fvk_batch* b, double* t) {
  if (!b || !t) return FVK_ERR_ARG;
  return get_time(&b->e, t);
}

int fvk_batch_get_options(fvk_batch* b, int32_t i, fvk_options* out) {
  if (!b || !out) return FVK_ERR_ARG;
  if (i < 0 || i >= (int32_t)b->e.in.size()) return set_err(&b->e, FVK_ERR_ARG, "instance out of range");
  *out = resolved_options(b->e, i);
  return FVK_OK;
}

int fvk_batch_use_graph(fvk_batch* b, int32_t on) {
  if (!b) return FVK_ERR_ARG;
  b->e.use_graph = on != 0;
  if (!on) drop_graph(&b->e);
  return FVK_OK;
}

int fvk_batch_checkpoint_size(fvk_batch* b, int64_t* bytes) {
  if (!b || !bytes) return FVK_ERR_ARG;
  *bytes = checkpoint_bytes(&b->e);
  return FVK_OK;
}

int fvk_batch_checkpoint_save(fvk_batch* b, void* buf, int64_t bytes) {
  if (!b) return FVK_ERR_ARG;
  return checkpoint_save(&b->e, buf, bytes);
}

int fvk_batch_checkpoint_load(fvk_batch* b, const void* buf, int64_t bytes) {
  if (!b) return FVK_ERR_ARG;
  return checkpoint_load(&b->e, buf, bytes);
}

int fvk_batch_time_kernels(fvk_batch* b, int32_t reps, double avg_us[2]) {
  if (!b) return FVK_ERR_ARG;
  return time_kernels(&b->e, reps, avg_us);
}

}  // extern "C"
