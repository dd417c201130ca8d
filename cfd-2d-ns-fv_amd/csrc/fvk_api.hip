// fvk_api.hip — the C ABI (include/fvk.h): context, device memory, the SSP-RK3
// driver (ssp_rk3.cpp:18-77 as a sequence of kernel launches on one stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fv_kernels.hpp"
#include "fvk.h"
#include "mesh_prep.hpp"

namespace {

thread_local std::string g_create_error;

}  // namespace

struct fvk_ctx {
  std::string err;
  fvk_options opt{};
  fvk::HostMesh hm;
  hipStream_t stream = nullptr;
  std::vector<void*> bufs;   // device allocations, freed by fvk_destroy
  fvk::DevMesh dm{};
  fvk::FlowT<double> fl64{};
  fvk::FlowT<float> fl32{};
  double qinf[4] = {0, 0, 0, 0};                // freestream conserved state (FP64)
  bool fp32 = false;                            // FP32-mixed: state = float(Q - Q_inf)
  size_t tsize = 8;                             // bytes per state value
  // state arrays of the arithmetic type (double, or float perturbations)
  void* q[3] = {nullptr, nullptr, nullptr};     // Qn, Q1, Q2 (4*ncp each)
  void* g[3] = {nullptr, nullptr, nullptr};     // stage 1..3 limited gradients (8*ncp each)
  void* sigma = nullptr;                        // ncp
  void* res = nullptr;                          // 4*ncp
  unsigned long long* smax = nullptr;           // [2], FP64 bits
  double* partial = nullptr;
  int block = 256;
  int variant = fvk::kVariantLds;
  int res_blocks = 0;
  int smax_cur = 0;
  bool have_state = false;
  bool sentinel = true;
  bool fs_prims = false;
  bool res_valid = false;
};

namespace {

int set_err(fvk_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_err(fvk_ctx* c, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_err(c, FVK_ERR_HIP, m);
}

#define FVK_HIP(ctx, call)                         \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return hip_err(ctx, e_, #call); \
  } while (0)

template <typename T>
int dev_alloc(fvk_ctx* c, T** out, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
  if (e != hipSuccess) return hip_err(c, e, "hipMalloc");
  c->bufs.push_back(p);
  *out = static_cast<T*>(p);
  return FVK_OK;
}

template <typename T>
int upload(fvk_ctx* c, const std::vector<T>& v, const T** out) {
  T* d = nullptr;
  int rc = dev_alloc(c, &d, v.size());
  if (rc) return rc;
  if (!v.empty()) {
    hipError_t e = hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err(c, e, "hipMemcpy H2D");
  }
  *out = d;
  return FVK_OK;
}

int check_options(const fvk_options& o, std::string* msg) {
  if (o.order != 1 && o.order != 2) { *msg = "order must be 1 or 2 (solution_poly_order 0 or 1)"; return FVK_ERR_OPTION; }
  if (o.limiter < 0 || o.limiter > 2) { *msg = "limiter must be NONE, MINMOD or SQUEEZE"; return FVK_ERR_OPTION; }
  if (o.eqn != 1 && o.eqn != 2) { *msg = "eqn must be EULER or NS"; return FVK_ERR_OPTION; }
  if (o.flux_solver == 2) { *msg = "flux_solver ROE is not implemented by the reference solver library"; return FVK_ERR_UNSUPPORTED; }
  if (o.flux_solver != 1) { *msg = "flux_solver must be RUSANOV"; return FVK_ERR_OPTION; }
  if (o.time_scheme != 1 && o.time_scheme != 2) { *msg = "time_scheme must be SSP_RK2 or SSP_RK3"; return FVK_ERR_OPTION; }
  if (o.time_step_type != 1 && o.time_step_type != 2) { *msg = "time_step_type must be GLOBAL or LOCAL"; return FVK_ERR_OPTION; }
  if (o.bc1 < -4 || o.bc1 > -1 || o.bc2 < -4 || o.bc2 > -1) { *msg = "BC1/BC2 must be FREESTREAM, SLIP_WALL, EXTRAPOLATION or NO_SLIP_WALL"; return FVK_ERR_OPTION; }
  if (o.precision != FVK_FP64 && o.precision != FVK_FP32_MIXED) { *msg = "precision must be FVK_FP64 or FVK_FP32_MIXED"; return FVK_ERR_OPTION; }
  if (o.kernel_variant < 0 || o.kernel_variant > 3) { *msg = "kernel_variant must be 0, 1, 2 or 3"; return FVK_ERR_OPTION; }
  if (o.renumber < 0 || o.renumber > 2) { *msg = "renumber must be 0, 1 or 2"; return FVK_ERR_OPTION; }
  if (o.block_size != 0 && (o.block_size < 64 || o.block_size > 256 || o.block_size % 64)) {
    *msg = "block_size must be 0 or a multiple of 64 in [64, 256]";
    return FVK_ERR_OPTION;
  }
  if (!(o.rho > 0) || !(o.P > 0) || !(o.gamma > 1)) { *msg = "freestream needs rho > 0, P > 0, gamma > 1"; return FVK_ERR_OPTION; }
  return FVK_OK;
}

template <typename T> const fvk::FlowT<T>& flow(const fvk_ctx* c);
template <> const fvk::FlowT<double>& flow<double>(const fvk_ctx* c) { return c->fl64; }
template <> const fvk::FlowT<float>& flow<float>(const fvk_ctx* c) { return c->fl32; }

template <typename T>
fvk::StageArgsT<T> base_args(fvk_ctx* c) {
  fvk::StageArgsT<T> a{};
  a.m = c->dm;
  a.fl = flow<T>(c);
  a.dt_global = c->opt.time_step_type == 1;
  a.xcd_remap = c->opt.xcd_remap;
  a.variant = c->variant;
  a.sigma = static_cast<T*>(c->sigma);
  a.res = static_cast<T*>(c->res);
  a.smax_cur = c->smax + c->smax_cur;
  a.smax_next = c->smax + (1 - c->smax_cur);
  return a;
}

// stage roles of the three state and gradient buffers (ssp_rk3.cpp:38-76):
// stage s reads Q_{s-1} and the previous stage's Q and limited gradient (the
// lagged limiter), writes G_s and Q_s; stage 3 writes Q_n in place
template <typename T>
void set_roles(fvk_ctx* c, int stage, fvk::StageArgsT<T>& a) {
  T* const* q = reinterpret_cast<T* const*>(c->q);
  T* const* g = reinterpret_cast<T* const*>(c->g);
  a.qn = q[0];
  if (stage == 1) { a.qc = q[0]; a.qp = q[2]; a.gp = g[2]; a.g = g[0]; a.qout = q[1]; }
  if (stage == 2) { a.qc = q[1]; a.qp = q[0]; a.gp = g[0]; a.g = g[1]; a.qout = q[2]; }
  if (stage == 3) { a.qc = q[2]; a.qp = q[1]; a.gp = g[1]; a.g = g[2]; a.qout = q[0]; }
}

// Spectral sums |V_inf| + c_inf on every face and their max: the lagged face
// data of a fresh start (class_q.hpp:61-62, main.cpp:117-118).
int arm_start(fvk_ctx* c) {
  c->smax_cur = 0;
  FVK_HIP(c, fvk::launch_fill(reinterpret_cast<double*>(c->smax), 2, 0.0, c->stream));
  if (c->fp32) FVK_HIP(c, fvk::launch_init_sigma(base_args<float>(c), c->block, c->stream));
  else FVK_HIP(c, fvk::launch_init_sigma(base_args<double>(c), c->block, c->stream));
  c->sentinel = true;
  c->res_valid = false;
  c->have_state = true;
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  return FVK_OK;
}

enum { kKindA = 1, kKindB = 2 };

// One ssp_rk3 call.  launch(kind, fn) lets fvk_time_steps bracket launches.
template <typename T, typename Launch>
int run_step_t(fvk_ctx* c, bool store_res, Launch&& launch) {
  const fvk_options& o = c->opt;
  for (int stage = 1; stage <= 3; ++stage) {
    fvk::StageArgsT<T> a = base_args<T>(c);
    a.stage = stage;
    a.sentinel = (stage == 1 && c->sentinel) ? 1 : 0;
    a.fs_prims = (stage == 1 && c->fs_prims) ? 1 : 0;
    a.store_res = (stage == 3 && store_res) ? 1 : 0;
    set_roles(c, stage, a);
    if (o.order == 2) {
      int rc = launch(kKindA, [&]() { return fvk::launch_pass_a(a, o.limiter == 2, c->block, c->stream); });
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
int run_step(fvk_ctx* c, bool store_res, Launch&& launch) {
  return c->fp32 ? run_step_t<float>(c, store_res, launch) : run_step_t<double>(c, store_res, launch);
}

// device array of n values of the arithmetic type -> doubles
int download(fvk_ctx* c, const void* src, size_t n, std::vector<double>* out) {
  out->resize(n);
  if (!c->fp32) {
    FVK_HIP(c, hipMemcpy(out->data(), src, n * 8, hipMemcpyDeviceToHost));
    return FVK_OK;
  }
  std::vector<float> f(n);
  FVK_HIP(c, hipMemcpy(f.data(), src, n * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) (*out)[i] = f[i];
  return FVK_OK;
}

template <typename T>
int fill_t(fvk_ctx* c, void* p, int64_t n, double v) {
  hipError_t e = fvk::launch_fill(static_cast<T*>(p), n, T(v), c->stream);
  return e == hipSuccess ? FVK_OK : hip_err(c, e, "fill");
}

int fill(fvk_ctx* c, void* p, int64_t n, double v) {
  return c->fp32 ? fill_t<float>(c, p, n, v) : fill_t<double>(c, p, n, v);
}

// class_flow (class_flow.hpp:28-46) and class_Q constants in type T
template <typename T>
void make_flow(const fvk_options& o, fvk::FlowT<T>* fl, bool delta) {
  fl->gamma = T(o.gamma);
  fl->gm1 = T(o.gamma - 1);
  fl->Rgas = T(o.R);
  fl->mu = T(o.mu);
  const double Cp = o.R * o.gamma / (o.gamma - 1);
  fl->k = T((o.mu * Cp) / o.Pr);
  fl->cfl2 = T(2 * o.cfl);
  const double qinf[4] = {o.rho, o.rho * o.u, o.rho * o.v,
                          (o.P / (o.gamma - 1)) + 0.5 * o.rho * (o.u * o.u + o.v * o.v)};
  for (int k = 0; k < 4; ++k) {
    fl->qinf[k] = T(qinf[k]);
    // the reference's -101 initial face states (class_q.hpp:59-60)
    fl->sent[k] = T(delta ? -101.0 - qinf[k] : -101.0);
  }
  fl->rho = T(o.rho);
  fl->u = T(o.u);
  fl->v = T(o.v);
  fl->P = T(o.P);
  fl->vn0 = T(std::sqrt(o.v * o.v + o.u * o.u));
  fl->c0 = T(std::sqrt(o.gamma * (o.P / o.rho)));
}

// `reps` back-to-back launches of the stage-2 pass A and pass B kernels, each
// kind between one event pair.  Stage 2 reads Q1, Qn, G1 and writes G2, Q2, so
// the repeats leave the state unchanged.
template <typename T>
int time_kernels_t(fvk_ctx* c, int32_t reps, double avg_us[2]) {
  const fvk_options& o = c->opt;
  fvk::StageArgsT<T> a = base_args<T>(c);
  a.stage = 2;
  set_roles(c, 2, a);
  hipEvent_t e0, e1;
  FVK_HIP(c, hipEventCreate(&e0));
  FVK_HIP(c, hipEventCreate(&e1));
  avg_us[0] = avg_us[1] = 0.0;
  int rc = FVK_OK;
  for (int kind = 0; kind < 2 && !rc; ++kind) {
    if (kind == 0 && o.order != 2) continue;
    hipError_t e = hipEventRecord(e0, c->stream);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r)
      e = kind == 0 ? fvk::launch_pass_a(a, o.limiter == 2, c->block, c->stream)
                    : fvk::launch_pass_b(a, o.order, o.eqn == 2, c->block, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = hip_err(c, e, "fvk_time_kernels");
    else avg_us[kind] = ms * 1e3 / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
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
  o->xcd_remap = 1;
  o->kernel_variant = 0;
}

const char* fvk_last_error(const fvk_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int fvk_create(const fvk_mesh* mesh, const fvk_options* opt, fvk_ctx** out) {
  g_create_error.clear();
  if (!mesh || !opt || !out) return set_err(nullptr, FVK_ERR_ARG, "fvk_create: NULL argument");
  *out = nullptr;
  std::string msg;
  int rc = check_options(*opt, &msg);
  if (rc) return set_err(nullptr, rc, msg);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(nullptr, FVK_ERR_HIP, "fvk_create: no HIP device");
  fvk_ctx* c = new fvk_ctx();
  c->opt = *opt;
  rc = fvk::prepare_mesh(*mesh, *opt, &c->hm, &msg);
  if (rc) {
    delete c;
    return set_err(nullptr, rc, msg);
  }
  if (!c->hm.warnings.empty()) fputs(c->hm.warnings.c_str(), stderr);
  const fvk::HostMesh& h = c->hm;
  c->block = opt->block_size ? opt->block_size : 256;
  c->variant = opt->kernel_variant ? opt->kernel_variant : fvk::kVariantLds;
  c->res_blocks = (h.ncp + 255) / 256;
  c->fp32 = opt->precision == FVK_FP32_MIXED;
  c->tsize = c->fp32 ? sizeof(float) : sizeof(double);
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    hip_err(nullptr, e, "hipStreamCreate");
    delete c;
    return FVK_ERR_HIP;
  }
  auto fail = [&](int code) {
    g_create_error = c->err;
    fvk_destroy(c);
    return code;
  };
  fvk::DevMesh& d = c->dm;
  d.nc = h.nc;
  d.ncp = h.ncp;
  d.blk = h.blk;
  d.nblocks = h.nblocks;
  d.max_halo = h.max_halo;
  d.max_bfaces = h.max_bfaces;
  if ((rc = upload(c, h.chunk_off, &d.chunk_off))) return fail(rc);
  if ((rc = upload(c, h.deg, &d.deg))) return fail(rc);
  if ((rc = upload(c, h.code, &d.code))) return fail(rc);
  if ((rc = upload(c, h.lcode, &d.lcode))) return fail(rc);
  if ((rc = upload(c, h.sface, &d.sface))) return fail(rc);
  if ((rc = upload(c, h.halo_ptr, &d.halo_ptr))) return fail(rc);
  if ((rc = upload(c, h.halo_idx, &d.halo_idx))) return fail(rc);
  if ((rc = upload(c, h.bf_ptr, &d.bf_ptr))) return fail(rc);
  if ((rc = upload(c, h.bf_a, &d.bf_a))) return fail(rc);
  if ((rc = upload(c, h.bf_b, &d.bf_b))) return fail(rc);
  if ((rc = upload(c, h.bf_nx, &d.bf_nx))) return fail(rc);
  if ((rc = upload(c, h.bf_ny, &d.bf_ny))) return fail(rc);
  if ((rc = upload(c, h.bf_area, &d.bf_area))) return fail(rc);
  if ((rc = upload(c, h.bf_fcx, &d.bf_fcx))) return fail(rc);
  if ((rc = upload(c, h.bf_fcy, &d.bf_fcy))) return fail(rc);
  if ((rc = upload(c, h.s_nx, &d.s_nx))) return fail(rc);
  if ((rc = upload(c, h.s_ny, &d.s_ny))) return fail(rc);
  if ((rc = upload(c, h.s_area, &d.s_area))) return fail(rc);
  if ((rc = upload(c, h.s_fcx, &d.s_fcx))) return fail(rc);
  if ((rc = upload(c, h.s_fcy, &d.s_fcy))) return fail(rc);
  if ((rc = upload(c, h.c_ccx, &d.ccx))) return fail(rc);
  if ((rc = upload(c, h.c_ccy, &d.ccy))) return fail(rc);
  if ((rc = upload(c, h.c_vol, &d.vol))) return fail(rc);
  if ((rc = upload(c, h.c_ixx, &d.ixx))) return fail(rc);
  if ((rc = upload(c, h.c_iyy, &d.iyy))) return fail(rc);
  if ((rc = upload(c, h.c_ixy, &d.ixy))) return fail(rc);
  if (c->variant == fvk::kVariantLds) {
    // the LDS-staged kernels need the block plus its largest halo in LDS
    int dev = 0;
    hipDeviceProp_t prop;
    if ((e = hipGetDevice(&dev)) != hipSuccess || (e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) {
      hip_err(c, e, "hipGetDeviceProperties");
      return fail(FVK_ERR_HIP);
    }
    const size_t need = std::max(fvk::lds_bytes_pass_a(d, c->tsize),
                                 fvk::lds_bytes_pass_b(d, opt->order, 3, c->tsize));
    if (need > prop.sharedMemPerBlock) {
      if (opt->kernel_variant == fvk::kVariantLds) {
        set_err(c, FVK_ERR_MESH, "mesh blocks need " + std::to_string(need) +
                                     " bytes of LDS, more than the device allows");
        return fail(FVK_ERR_MESH);
      }
      c->variant = fvk::kVariantCell;
    }
  }
  const size_t ncp = h.ncp;
  auto alloc = [&](void** p, size_t n) {
    unsigned char* d = nullptr;
    const int r = dev_alloc(c, &d, n * c->tsize);
    *p = d;
    return r;
  };
  for (int i = 0; i < 3; ++i) {
    if ((rc = alloc(&c->q[i], 4 * ncp))) return fail(rc);
    if ((rc = alloc(&c->g[i], 8 * ncp))) return fail(rc);
  }
  if ((rc = alloc(&c->sigma, ncp))) return fail(rc);
  if ((rc = alloc(&c->res, 4 * ncp))) return fail(rc);
  if ((rc = dev_alloc(c, &c->smax, 2))) return fail(rc);
  if ((rc = dev_alloc(c, &c->partial, 4 * (size_t)c->res_blocks))) return fail(rc);
  for (int i = 0; i < 3; ++i) {
    if ((rc = fill(c, c->q[i], 4 * (int64_t)ncp, 0.0))) return fail(rc);
    if ((rc = fill(c, c->g[i], 8 * (int64_t)ncp, 0.0))) return fail(rc);
  }
  if ((rc = fill(c, c->res, 4 * (int64_t)ncp, 0.0))) return fail(rc);
  if ((rc = fill(c, c->sigma, (int64_t)ncp, 0.0))) return fail(rc);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) { hip_err(c, e, "sync"); return fail(FVK_ERR_HIP); }

  const fvk_options& o = *opt;
  make_flow(o, &c->fl64, false);
  make_flow(o, &c->fl32, true);
  c->qinf[0] = o.rho;
  c->qinf[1] = o.rho * o.u;
  c->qinf[2] = o.rho * o.v;
  c->qinf[3] = (o.P / (o.gamma - 1)) + 0.5 * o.rho * (o.u * o.u + o.v * o.v);
  *out = c;
  return FVK_OK;
}

int fvk_destroy(fvk_ctx* c) {
  if (!c) return FVK_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->bufs) (void)hipFree(p);
  c->bufs.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FVK_OK;
}

int fvk_init_freestream(fvk_ctx* c) {
  if (!c) return FVK_ERR_ARG;
  const size_t ncp = c->hm.ncp;
  for (int k = 0; k < 4; ++k) {
    // FP64: Q_inf; FP32-mixed: the perturbation 0
    void* blk = static_cast<unsigned char*>(c->q[0]) + k * ncp * c->tsize;
    const int rc = fill(c, blk, (int64_t)ncp, c->fp32 ? 0.0 : c->qinf[k]);
    if (rc) return rc;
  }
  int rc = arm_start(c);
  c->fs_prims = true;
  return rc;
}

int fvk_set_state(fvk_ctx* c, const double* q) {
  if (!c) return FVK_ERR_ARG;
  if (!q) return set_err(c, FVK_ERR_ARG, "fvk_set_state: NULL state");
  const fvk::HostMesh& h = c->hm;
  const size_t n = 4 * (size_t)h.ncp;
  if (c->fp32) {
    std::vector<float> buf(n, 0.0f);
    for (int k = 0; k < 4; ++k)
      for (int32_t i = 0; i < h.ncp; ++i)
        if (h.perm[i] >= 0)
          buf[k * (size_t)h.ncp + i] = float(q[k * (size_t)h.nc + h.perm[i]] - c->qinf[k]);
    FVK_HIP(c, hipMemcpy(c->q[0], buf.data(), n * 4, hipMemcpyHostToDevice));
  } else {
    std::vector<double> buf(n, 1.0);
    for (int k = 0; k < 4; ++k)
      for (int32_t i = 0; i < h.ncp; ++i)
        if (h.perm[i] >= 0) buf[k * (size_t)h.ncp + i] = q[k * (size_t)h.nc + h.perm[i]];
    FVK_HIP(c, hipMemcpy(c->q[0], buf.data(), n * 8, hipMemcpyHostToDevice));
  }
  int rc = arm_start(c);
  c->fs_prims = false;
  return rc;
}

int fvk_step(fvk_ctx* c, int32_t nsteps) {
  if (!c) return FVK_ERR_ARG;
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "fvk_step: no state (call fvk_init_freestream or fvk_set_state)");
  if (nsteps < 0) return set_err(c, FVK_ERR_ARG, "fvk_step: negative step count");
  auto plain = [c](int, auto&& fn) -> int {
    hipError_t e = fn();
    return e == hipSuccess ? FVK_OK : hip_err(c, e, "kernel launch");
  };
  for (int32_t n = 0; n < nsteps; ++n) {
    int rc = run_step(c, n == nsteps - 1, plain);
    if (rc) return rc;
  }
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  return FVK_OK;
}

int fvk_get_state(fvk_ctx* c, double* q, double* grad) {
  if (!c) return FVK_ERR_ARG;
  if (!q && !grad) return set_err(c, FVK_ERR_ARG, "fvk_get_state: nothing requested");
  const fvk::HostMesh& h = c->hm;
  const size_t ncp = h.ncp;
  if (q) {
    std::vector<double> buf;
    int rc = download(c, c->q[0], 4 * ncp, &buf);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) {
      const double base = c->fp32 ? c->qinf[k] : 0.0;
      for (int32_t i = 0; i < h.ncp; ++i)
        if (h.perm[i] >= 0)
          q[k * (size_t)h.nc + h.perm[i]] = c->fp32 ? base + buf[k * ncp + i] : buf[k * ncp + i];
    }
  }
  if (grad) {
    std::vector<double> buf;
    int rc = download(c, c->g[0], 8 * ncp, &buf);
    if (rc) return rc;
    for (int k = 0; k < 8; ++k)
      for (int32_t i = 0; i < h.ncp; ++i)
        if (h.perm[i] >= 0) grad[k * (size_t)h.nc + h.perm[i]] = buf[k * ncp + i];
  }
  return FVK_OK;
}

int fvk_get_residual(fvk_ctx* c, double* r) {
  if (!c) return FVK_ERR_ARG;
  if (!r) return set_err(c, FVK_ERR_ARG, "fvk_get_residual: NULL output");
  const fvk::HostMesh& h = c->hm;
  std::vector<double> buf;
  int rc = download(c, c->res, 4 * (size_t)h.ncp, &buf);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k)
    for (int32_t i = 0; i < h.ncp; ++i)
      if (h.perm[i] >= 0) r[k * (size_t)h.nc + h.perm[i]] = buf[k * (size_t)h.ncp + i];
  return FVK_OK;
}

int fvk_residual_l2(fvk_ctx* c, double out[4]) {
  if (!c || !out) return FVK_ERR_ARG;
  const fvk::HostMesh& h = c->hm;
  if (c->fp32)
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const float*>(c->res), h.ncp, c->partial, c->res_blocks, 256, c->stream));
  else
    FVK_HIP(c, fvk::launch_res_sq(static_cast<const double*>(c->res), h.ncp, c->partial, c->res_blocks, 256, c->stream));
  std::vector<double> part(4 * (size_t)c->res_blocks);
  FVK_HIP(c, hipMemcpyAsync(part.data(), c->partial, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
  FVK_HIP(c, hipStreamSynchronize(c->stream));
  bool finite = true;
  for (int k = 0; k < 4; ++k) {
    // Neumaier-compensated sum of the block partials
    double s = 0.0, comp = 0.0;
    for (int b = 0; b < c->res_blocks; ++b) {
      const double x = part[k * (size_t)c->res_blocks + b];
      const double t = s + x;
      if (std::fabs(s) >= std::fabs(x)) comp += (s - t) + x;
      else comp += (x - t) + s;
      s = t;
    }
    out[k] = std::sqrt(s + comp);
    if (!std::isfinite(out[k])) finite = false;
  }
  if (!finite) return set_err(c, FVK_ERR_NONFINITE, "residual norm is not finite");
  return FVK_OK;
}

int fvk_get_geometry(fvk_ctx* c, const char* name, double* out) {
  if (!c || !name || !out) return FVK_ERR_ARG;
  const fvk::HostMesh& h = c->hm;
  const std::vector<double>* v = nullptr;
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
  if (!v) return set_err(c, FVK_ERR_ARG, "fvk_get_geometry: unknown array " + n);
  std::memcpy(out, v->data(), v->size() * sizeof(double));
  return FVK_OK;
}

int fvk_geometry(const fvk_mesh* mesh, const fvk_options* opt, const char* name, double* out) {
  g_create_error.clear();
  if (!mesh || !opt || !name || !out) return set_err(nullptr, FVK_ERR_ARG, "fvk_geometry: NULL argument");
  fvk_ctx tmp;
  std::string msg;
  int rc = fvk::prepare_mesh(*mesh, *opt, &tmp.hm, &msg);
  if (rc) return set_err(nullptr, rc, msg);
  rc = fvk_get_geometry(&tmp, name, out);
  if (rc) g_create_error = tmp.err;
  return rc;
}

int fvk_get_info(fvk_ctx* c, int64_t info[6]) {
  if (!c || !info) return FVK_ERR_ARG;
  const fvk::HostMesh& h = c->hm;
  info[0] = h.nc;
  info[1] = h.nf;
  info[2] = h.nb;
  info[3] = h.chunk_off.empty() ? 0 : h.chunk_off.back();
  info[4] = h.max_deg;
  info[5] = h.num_wall;
  return FVK_OK;
}

int fvk_time_steps(fvk_ctx* c, int32_t nsteps, double out_ms[4], int64_t counts[4]) {
  if (!c || !out_ms || !counts) return FVK_ERR_ARG;
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

int fvk_time_kernels(fvk_ctx* c, int32_t reps, double avg_us[2]) {
  if (!c || !avg_us || reps <= 0) return FVK_ERR_ARG;
  if (!c->have_state) return set_err(c, FVK_ERR_ARG, "fvk_time_kernels: no state");
  return c->fp32 ? time_kernels_t<float>(c, reps, avg_us) : time_kernels_t<double>(c, reps, avg_us);
}

}  // extern "C"
