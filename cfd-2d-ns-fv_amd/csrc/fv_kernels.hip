// fv_kernels.hip — explicit FV kernels for gfx950 (MI355X), FP64 and FP32-mixed.
//
// Layout (mesh_prep.hpp): cells in blocks (recursive coordinate bisection),
// numbered block-major; per block a halo list and a face list; per cell a
// sliced-ELL list of face slots in increasing ORIGINAL face index.  The
// reference's face loops add into a cell in increasing face index, so adding
// a cell's face terms in slot order reproduces its sums term for term.  Built
// with -ffp-contract=off: every expression is evaluated exactly as written,
// in the reference's association order, so the FP64 instantiation is
// bit-identical to the reference.
//
// Every kernel is a template on the state type T (fv_kernels.hpp): double is
// the FP64 path (state = conserved variables); float is the FP32-mixed path
// (state = perturbation from freestream, kDelta below).  The FP64 expressions
// are the T = double case of the same code: the perturbation-only steps are
// compiled out for it.
//
// Kernels per RK stage:
//   pass A (order 2): LSQ gradient (cell_gradient.cpp:9-130) + squeeze limiter
//          (squeeze.cpp:9-154) with the lagged face states rebuilt from the
//          previous stage's state and limited gradient of the same cell.
//   pass B: face reconstruction with BC ghosts (reconstruction.cpp:15-80),
//          NS face gradients and stresses (reconstruction.cpp:83-164,
//          class_q.hpp:163-169, class_f.hpp:52-59), Rusanov flux
//          (rusanov.cpp:12-36, class_f.hpp:41-50), face-to-cell accumulation
//          (res.cpp:74-109), the SSP-RK3 stage update (ssp_rk3.cpp:51-76) and,
//          at stage 3, the next step's spectral sums (ssp_rk3.cpp:28-42).
//
// Three launch shapes share the per-face device functions:
//   cell-parallel  one thread per cell walks its slots, neighbours from HBM;
//   slot-parallel  one thread per (slot row, cell), slot terms summed in slot
//                  order through LDS;
//   LDS-staged     one block per mesh block: the block's cells and its halo
//                  are staged in LDS with coalesced loads; pass B computes
//                  each face of the block once (one thread per face, two
//                  threads per cell of the block) into LDS and each cell then
//                  adds its faces' terms in slot order;
//   persistent     the LDS-staged blocks walked by as many workgroups as are
//                  resident at once, each over its XCD's share of the blocks.
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "fv_kernels.hpp"

// FP32-mixed arithmetic: 1 = reciprocal and hardware square root (1 ulp) in
// the perturbation flux, fused multiply-adds in the FP32 face states and
// flux; 0 = the FP64 path's operation sequence in FP32.  The FP64 path never
// changes.
#ifndef FVK_FAST32
#define FVK_FAST32 1
#endif
// FP32-mixed NS: 1 = the viscous face flux in FP64, 0 = in FP32.
// FP32-mixed squeeze factors with one hardware reciprocal instead of a
// correctly rounded quotient
#ifndef FVK_A_RCP
#define FVK_A_RCP 1
#endif
#ifndef FVK_NS64
#define FVK_NS64 0
#endif

namespace fvk {
namespace {

constexpr int kMaxBlock = 256;
constexpr int kPassBThreadsPerCell = 2;  // LDS-staged pass B: blockDim = 2 * blk
constexpr int kRows = 4;  // slot rows per block in the slot-parallel shape

// the FP32-mixed state is a perturbation from freestream
template <typename T>
constexpr bool kDelta = std::is_same<T, float>::value;

template <typename T>
__device__ __forceinline__ T min2(T a, T b) { return (b < a) ? b : a; }  // std::min({a,b})
template <typename T>
__device__ __forceinline__ T max2(T a, T b) { return (a < b) ? b : a; }  // std::max({a,b})

// XCD-aware block order: blocks are dealt to the 8 XCDs round-robin, so block
// b runs on XCD b%8; give each XCD one contiguous range of the logical blocks
// so that the neighbour gathers of its blocks hit its own L2.
// A negative argument -G asks for runs of G consecutive logical blocks per
// XCD, the runs dealt to the XCDs in turn: neighbour blocks still share an L2
// while every XCD walks the whole block range at the same pace.
__device__ __forceinline__ int logical_block_of(int b, int per_xcd) {
  if (per_xcd > 0) return (b & 7) * per_xcd + (b >> 3);
  if (per_xcd < 0) {
    const int g = -per_xcd, j = b >> 3;
    return ((j / g) * 8 + (b & 7)) * g + j % g;
  }
  return b;
}

__device__ __forceinline__ int logical_block(int per_xcd) {
  return logical_block_of(blockIdx.x, per_xcd);
}

// Mesh block and instance of logical block lb: through the active-block list
// of a batch (retired instances' blocks are not in it), else the identity.
// False for the padding blocks of the launch.
template <typename T>
__device__ __forceinline__ bool block_of(const StageArgsT<T>& a, int lb, int& b, int& inst) {
  if (a.active) {
    if (lb >= a.nactive) return false;
    b = a.active[lb];
    inst = a.blk_inst[b];
    return true;
  }
  b = lb;
  inst = 0;
  return lb < a.m.nblocks;
}

template <typename T>
struct Prim {
  T rho, u, v, E, P, c;
};

// stored state, limited gradient and centroid of one cell
template <typename T>
struct CellQG {
  T q[4], gx[4], gy[4];
  double x, y;
};

// conserved state from the stored state
template <typename T>
__device__ __forceinline__ T absq(const FlowT<T>& fl, int k, T d) {
  if constexpr (kDelta<T>) return fl.qinf[k] + d;
  else return d;
}

// class_Q::updateflow_int (class_q.hpp:143-151), without T (unused downstream)
template <typename T>
__device__ __forceinline__ void to_prim(const T q[4], T gamma, T gm1, Prim<T>& p) {
  p.rho = q[0];
  p.E = q[3];
  p.u = q[1] / p.rho;
  p.v = q[2] / p.rho;
  p.P = (p.E - (T(0.5) * p.rho * (p.u * p.u + p.v * p.v))) * gm1;
  p.c = sqrt(gamma * (p.P / p.rho));
}

// class_flow::updateQ of a ghost state (class_flow.hpp:28-34)
template <typename T>
__device__ __forceinline__ void ghost_q(T gm1, T rho, T u, T v, T P, T w[4]) {
  w[0] = rho;
  w[1] = rho * u;
  w[2] = rho * v;
  w[3] = (P / gm1) + T(0.5) * rho * (u * u + v * v);
}

// primitives rho, u, v, P of a cell (class_Q::updateflow) from its stored
// state, or the exact freestream values the very first stage sees
// (class_q.hpp:55-58)
template <typename T>
__device__ __forceinline__ void cell_prim(const T qs[4], const FlowT<T>& fl, int fs, T& rho, T& u,
                                          T& v, T& P) {
  if (fs) {
    rho = fl.rho; u = fl.u; v = fl.v; P = fl.P;
    return;
  }
  rho = absq(fl, 0, qs[0]);
  const T E = absq(fl, 3, qs[3]);
  u = absq(fl, 1, qs[1]) / rho;
  v = absq(fl, 2, qs[2]) / rho;
  P = (E - (T(0.5) * rho * (u * u + v * v))) * fl.gm1;
}

// Cell-major state: the 4 values of a cell's state (or residual) are
// contiguous, and so are the 8 of its gradient (gx[4], gy[4]), so a cell is
// read and written with 16-byte vector accesses (one per 4 floats, two per 4
// doubles) -- coalesced across a wavefront's consecutive cells, and a single
// access per halo or neighbour cell where a component-major layout needs four.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void ld4(const float* p, float v[4]) {
  const f4v x = *reinterpret_cast<const f4v*>(p);
  v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
}
__device__ __forceinline__ void ld4(const double* p, double v[4]) {
  const d2v x = reinterpret_cast<const d2v*>(p)[0];
  const d2v y = reinterpret_cast<const d2v*>(p)[1];
  v[0] = x[0]; v[1] = x[1]; v[2] = y[0]; v[3] = y[1];
}
__device__ __forceinline__ void st4(float* p, const float v[4]) {
  f4v x;
  x[0] = v[0]; x[1] = v[1]; x[2] = v[2]; x[3] = v[3];
  *reinterpret_cast<f4v*>(p) = x;
}
__device__ __forceinline__ void st4(double* p, const double v[4]) {
  d2v x, y;
  x[0] = v[0]; x[1] = v[1]; y[0] = v[2]; y[1] = v[3];
  reinterpret_cast<d2v*>(p)[0] = x;
  reinterpret_cast<d2v*>(p)[1] = y;
}

// Non-temporal access for the streams that a launch reads or writes once, so
// that they do not displace from the L2 the cell data that neighbouring
// blocks read again as their halo.  Build option FVK_NT, a bit mask, measured
// per launch on MI355X (256-instance batch, ramp, NS cylinder):
//   1  pass-B loads of the face records, slot map and stage-update data:
//      FP32-mixed pass B 3-14% slower;
//   2  FP32-mixed stores of the new state, spectral sums, residual and
//      gradients: pass A 3.6% and pass B 2.6% faster on the batch;
//   4  pass-A loads of the slot data, LSQ inverse and lagged state: FP64
//      pass A 16-25% slower;
//   8  FP64 stores of the new state, spectral sums and residual: pass B
//      1-2% faster;
//  16  FP64 gradient stores of whole 1 KB ranges per wave (FVK_A_GW): FP64
//      pass A 5% faster on the batch;
//  32  pass-B loads of the stage-update data alone: FP32-mixed pass B 1%
//      faster on the batch, 2-4% slower on the ramp and the cylinder;
//  64  pass-A loads of the lagged state alone: pass A 2-14% slower;
// 128  FP64 gradient stores of one record per lane (four 16-byte stores 64
//      bytes apart: the cell- and slot-parallel shapes, and pass A with
//      FVK_A_GW = 0): FP64 pass A 27% slower on the batch.
#ifndef FVK_NT
#define FVK_NT 26
#endif
template <int BIT, typename P>
__device__ __forceinline__ P sld(const P* p) {
  if constexpr ((FVK_NT & BIT) != 0) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int BIT, typename P>
__device__ __forceinline__ void sst(P* p, P v) {
  if constexpr ((FVK_NT & BIT) != 0) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <int BIT>
__device__ __forceinline__ void ld4s(const float* p, float v[4]) {
  const f4v x = sld<BIT>(reinterpret_cast<const f4v*>(p));
  v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
}
template <int BIT>
__device__ __forceinline__ void ld4s(const double* p, double v[4]) {
  const d2v x = sld<BIT>(reinterpret_cast<const d2v*>(p));
  const d2v y = sld<BIT>(reinterpret_cast<const d2v*>(p) + 1);
  v[0] = x[0]; v[1] = x[1]; v[2] = y[0]; v[3] = y[1];
}
template <int BIT>
__device__ __forceinline__ void st4s(float* p, const float v[4]) {
  f4v x;
  x[0] = v[0]; x[1] = v[1]; x[2] = v[2]; x[3] = v[3];
  sst<BIT>(reinterpret_cast<f4v*>(p), x);
}
template <int BIT>
__device__ __forceinline__ void st4s(double* p, const double v[4]) {
  d2v x, y;
  x[0] = v[0]; x[1] = v[1]; y[0] = v[2]; y[1] = v[3];
  sst<BIT>(reinterpret_cast<d2v*>(p), x);
  sst<BIT>(reinterpret_cast<d2v*>(p) + 1, y);
}

template <typename T>
__device__ __forceinline__ void load_q(const T* __restrict__ qa, int c, T q[4]) {
  ld4(qa + 4 * (size_t)c, q);
}

template <typename T>
__device__ __forceinline__ void load_g(const T* __restrict__ ga, int c, T gx[4], T gy[4]) {
  ld4(ga + 8 * (size_t)c, gx);
  ld4(ga + 8 * (size_t)c + 4, gy);
}

template <typename T>
__device__ __forceinline__ void load_cell(const StageArgsT<T>& a, int c, bool order2,
                                          CellQG<T>& d) {
  load_q(a.qc, c, d.q);
  if (order2) {
    load_g(a.g, c, d.gx, d.gy);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) d.gx[k] = d.gy[k] = T(0);
  }
  d.x = a.m.ccx[c];
  d.y = a.m.ccy[c];
}

// Block-wide max of non-negative values, then one atomicMax on the FP64 bit
// pattern (the order of non-negative IEEE doubles is that of their bits).
__device__ __forceinline__ void block_max_to(double val, unsigned long long* dst) {
  __shared__ double wmax[16];
  for (int off = 32; off > 0; off >>= 1) val = fmax(val, __shfl_xor(val, off, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) wmax[w] = val;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, wmax[i]);
    atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// ---------------------------------------------------------------------------
// per-face terms
// ---------------------------------------------------------------------------

// Pass A, interior face: the LSQ right-hand-side terms (Qj - Qi)*dxj,
// (Qj - Qi)*dyj with i = face_cell1 (cell_gradient.cpp:74-91), and the face's
// min/max candidates min/max(Q1, Q2) (squeeze.cpp:22-40).  `own`: this cell is
// face_cell1.  Centroid differences are formed in FP64.
template <typename T>
__device__ __forceinline__ void lsq_interior(bool own, const T q[4], double xc, double yc,
                                             const T qn[4], double xn, double yn, T px[4],
                                             T py[4], T tmin[4], T tmax[4]) {
  T d[4];
  double dxj, dyj;
  if (own) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      d[k] = qn[k] - q[k];
      tmin[k] = min2(q[k], qn[k]);
      tmax[k] = max2(q[k], qn[k]);
    }
    dxj = xn - xc;
    dyj = yn - yc;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      d[k] = q[k] - qn[k];
      tmin[k] = min2(qn[k], q[k]);
      tmax[k] = max2(qn[k], q[k]);
    }
    dxj = xc - xn;
    dyj = yc - yn;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    px[k] = d[k] * T(dxj);
    py[k] = d[k] * T(dyj);
  }
}

// Pass A, boundary face: ghost-neighbour terms (cell_gradient.cpp:37-73,
// ghost centroid readmesh.cpp:238-243); the face offers only Q_i to min/max.
template <typename T>
__device__ __forceinline__ void lsq_boundary(const StageArgsT<T>& a, const FlowT<T>& fl, int code,
                                             const T q[4],
                                             double xc, double yc, double nx, double ny,
                                             double fx, double fy, T px[4], T py[4], T tmin[4],
                                             T tmax[4]) {
  const int v = ~code;
  const int kind = v & 3;
  const double rx = fx - xc, ry = fy - yc;
  const double t = 2.0 * (rx * nx + ry * ny);
  const double gx = xc + t * nx, gy = yc + t * ny;
  const double dxj = gx - xc, dyj = gy - yc;
  T d[4];
  if (kind == 0) {            // FREESTREAM: Q_inf - Q_i
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if constexpr (kDelta<T>) d[k] = -q[k];
      else d[k] = fl.qinf[k] - q[k];
    }
  } else if (kind == 1) {     // SLIP_WALL: reflect the cell velocity
    const T tnx = T(nx), tny = T(ny);
    if constexpr (kDelta<T>) {  // W - Q_i = (0, -2(m.n)n, 0): same rho and E
      const T mn = (fl.qinf[1] + q[1]) * tnx + (fl.qinf[2] + q[2]) * tny;
      d[0] = T(0);
      d[1] = -T(2.0) * mn * tnx;
      d[2] = -T(2.0) * mn * tny;
      d[3] = T(0);
    } else {
      T rho, u, vv, P;
      cell_prim(q, fl, a.fs_prims, rho, u, vv, P);
      const T Vn = u * tnx + vv * tny;
      const T wu = u - T(2.0) * Vn * tnx;
      const T wv = vv - T(2.0) * Vn * tny;
      T w[4];
      ghost_q(fl.gm1, rho, wu, wv, P, w);
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = w[k] - q[k];
    }
  } else if (kind == 2) {     // EXTRAPOLATION
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = T(0);
  } else {                    // NO_SLIP_WALL: state of the cell whose index is the face index
    T qs[4];
    load_q(a.qc, v >> 2, qs);
    if constexpr (kDelta<T>) {  // W = (rho_s, -m_s, E_s)
      d[0] = qs[0] - q[0];
      d[1] = (-qs[1] - T(2.0) * fl.qinf[1]) - q[1];
      d[2] = (-qs[2] - T(2.0) * fl.qinf[2]) - q[2];
      d[3] = qs[3] - q[3];
    } else {
      T rho, u, vv, P;
      cell_prim(qs, fl, a.fs_prims, rho, u, vv, P);
      T w[4];
      ghost_q(fl.gm1, rho, -u, -vv, P, w);
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = w[k] - q[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    px[k] = d[k] * T(dxj);
    py[k] = d[k] * T(dyj);
    tmin[k] = q[k];
    tmax[k] = q[k];
  }
}

// Pass A, second loop: the squeeze factor of one face for this cell from the
// face state the previous res() call left on this cell's side
// (squeeze.cpp:68-121): Q_prev + grad_prev . (x_f - x_i), or -101 at first.
// The Barth-Jespersen limiter (limiter 3, an extension) passes the current
// state and unlimited gradient instead, and no sentinel.
template <typename T>
__device__ __forceinline__ void face_phi(bool sentinel, const FlowT<T>& fl, T dx, T dy,
                                         const T q[4], const T qmin[4], const T qmax[4],
                                         const T qp[4], const T gpx[4], const T gpy[4], T ph[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T qf = sentinel ? fl.sent[k] : qp[k] + gpx[k] * dx + gpy[k] * dy;
    T p = T(1.0);
#if FVK_FAST32 && FVK_A_RCP
    if constexpr (kDelta<T>) {
      // FP32-mixed: one hardware reciprocal (1 ulp); the factor stays within
      // [0, 1 + ulp] and the minimum with 1 below caps it.  A cell at its
      // own bound gets 0 outright: the quotient's 0 / (qf - q) would be
      // 0 * inf when qf - q is too small for the reciprocal.
      const T bound = qf < qmin[k] ? qmin[k] : qmax[k];
      const T num = bound - q[k];
      if (qf < qmin[k] || qf > qmax[k])
        p = num == T(0) ? T(0) : num * __builtin_amdgcn_rcpf(qf - q[k]);
      ph[k] = p;
      continue;
    }
#endif
    if (qf < qmin[k]) p = (qmin[k] - q[k]) / (qf - q[k]);
    else if (qf > qmax[k]) p = (qmax[k] - q[k]) / (qf - q[k]);
    ph[k] = p;
  }
}

// Face state q + gx*dx + gy*dy (reconstruction.cpp:26-29,73-76): the FP64
// path keeps the reference's two roundings per term; the FP32 path fuses
// them where FVK_FAST32 allows.
template <typename T>
__device__ __forceinline__ T recon(T q, T gx, T gy, T dx, T dy) {
#if FVK_FAST32
  if constexpr (kDelta<T>) return __builtin_fmaf(gy, dy, __builtin_fmaf(gx, dx, q));
#endif
  return q + gx * dx + gy * dy;
}

// Geometry of one face as the flux uses it, in the arithmetic type: normal,
// length, face centre minus owner / neighbour centroid and, for NS, the unit
// owner->neighbour vector and centroid distance (interior) or the owner's
// wall distance (no-slip wall).  Every value is formed in FP64 and rounded
// once, so the FP64 path keeps the reference's expressions and the FP32 path
// may read the same values precomputed.
template <typename T>
struct FaceGeo {
  T nx, ny, A;
  T d1x, d1y, d2x, d2y;
  T lx, ly, dist;
};

template <typename T, bool NS>
__device__ __forceinline__ FaceGeo<T> geo_from(double nx, double ny, double A, double fx, double fy,
                                               const CellQG<T>& c1, const CellQG<T>* c2, int kind) {
  FaceGeo<T> g;
  g.nx = T(nx);
  g.ny = T(ny);
  g.A = T(A);
  g.d1x = T(fx - c1.x);   // reconstruction.cpp:26-29 (face_cell1_dx)
  g.d1y = T(fy - c1.y);
  g.d2x = g.d2y = g.lx = g.ly = g.dist = T(0);
  if (c2) {
    g.d2x = T(fx - c2->x);
    g.d2y = T(fy - c2->y);
    if (NS) {             // face_dist and l (readmesh.cpp:394-399)
      const double tx = c2->x - c1.x, ty = c2->y - c1.y;
      const double dist = sqrt(tx * tx + ty * ty);
      g.lx = T(tx / dist);
      g.ly = T(ty / dist);
      g.dist = T(dist);
    }
  } else if (NS && kind == 3) {   // reconstruction.cpp:112-113
    const double dx1 = fx - c1.x, dy1 = fy - c1.y;
    g.dist = T(fabs(sqrt(dx1 * dx1 + dy1 * dy1)));
  }
  return g;
}

// the geometry of face slot s, from the per-slot FP64 arrays
template <typename T, bool NS>
__device__ __forceinline__ FaceGeo<T> slot_geo(const DevMesh& m, int s, const CellQG<T>& c1,
                                               const CellQG<T>* c2, int kind) {
  return geo_from<T, NS>(m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], c1, c2, kind);
}

// Inviscid face terms of one side in perturbation form (FP32-mixed path):
// from d = Q - Q_inf, the primitives and F(Q).n - F(Q_inf).n computed so that
// every term that cancels between Q and Q_inf cancels exactly:
//   dVn = (dm.n - Vn_inf drho) / rho
//   dK  = dm.(u + u_inf) - drho (u.u_inf)      (= |m|^2/rho - |m_inf|^2/rho_inf)
//   dP  = (gamma - 1) (dE - dK/2)
//   dF  = (dm.n,  dm_x Vn + m_inf,x dVn + dP nx,  dm_y Vn + m_inf,y dVn + dP ny,
//          dVn H_inf + Vn (dE + dP)),  H_inf = E_inf + P_inf.
template <typename T>
struct SideFlux {
  T rho, u, v, Vn, P, c;
  T dF[4];
};

template <typename T>
__device__ __forceinline__ void side_flux_delta(const FlowT<T>& fl, const T d[4], T nx, T ny,
                                                SideFlux<T>& s) {
#if FVK_FAST32
  // one reciprocal of rho and the hardware square root (1 ulp each) and
  // explicit fused multiply-adds (explicit, so that every launch shape
  // evaluates the same operations)
  s.rho = fl.qinf[0] + d[0];
  const T inv = __builtin_amdgcn_rcpf(s.rho);
  const T mx = fl.qinf[1] + d[1], my = fl.qinf[2] + d[2];
  s.u = mx * inv;
  s.v = my * inv;
  s.Vn = __builtin_fmaf(s.v, ny, s.u * nx);
  const T vn_inf = __builtin_fmaf(fl.v, ny, fl.u * nx);
  const T dmn = __builtin_fmaf(d[2], ny, d[1] * nx);
  const T dVn = __builtin_fmaf(-vn_inf, d[0], dmn) * inv;
  const T uu = __builtin_fmaf(s.v, fl.v, s.u * fl.u);
  const T dK = __builtin_fmaf(d[1], s.u + fl.u, __builtin_fmaf(d[2], s.v + fl.v, -d[0] * uu));
  const T dP = fl.gm1 * __builtin_fmaf(T(-0.5), dK, d[3]);
  s.P = fl.P + dP;
  s.c = __builtin_amdgcn_sqrtf(fl.gamma * (s.P * inv));
  s.dF[0] = dmn;
  s.dF[1] = __builtin_fmaf(dP, nx, __builtin_fmaf(fl.qinf[1], dVn, d[1] * s.Vn));
  s.dF[2] = __builtin_fmaf(dP, ny, __builtin_fmaf(fl.qinf[2], dVn, d[2] * s.Vn));
  s.dF[3] = __builtin_fmaf(s.Vn, d[3] + dP, dVn * (fl.qinf[3] + fl.P));
#else
  s.rho = fl.qinf[0] + d[0];
  const T mx = fl.qinf[1] + d[1], my = fl.qinf[2] + d[2];
  s.u = mx / s.rho;
  s.v = my / s.rho;
  s.Vn = s.u * nx + s.v * ny;
  const T vn_inf = fl.u * nx + fl.v * ny;
  const T dmn = d[1] * nx + d[2] * ny;
  const T dVn = (dmn - vn_inf * d[0]) / s.rho;
  const T dK = d[1] * (s.u + fl.u) + d[2] * (s.v + fl.v) - d[0] * (s.u * fl.u + s.v * fl.v);
  const T dP = fl.gm1 * (d[3] - T(0.5) * dK);
  s.P = fl.P + dP;
  s.c = sqrt(fl.gamma * (s.P / s.rho));
  s.dF[0] = dmn;
  s.dF[1] = d[1] * s.Vn + fl.qinf[1] * dVn + dP * nx;
  s.dF[2] = d[2] * s.Vn + fl.qinf[2] * dVn + dP * ny;
  s.dF[3] = dVn * (fl.qinf[3] + fl.P) + s.Vn * (d[3] + dP);
#endif
}

// conserved component k, in type U, of a stored state value d
template <typename U, typename T>
__device__ __forceinline__ U conserved(const FlowT<T>& fl, int k, T d) {
  if constexpr (!kDelta<T>) return d;
  else if constexpr (std::is_same<U, T>::value) return fl.qinf[k] + d;
  else return fl.qinf64[k] + U(d);
}

// Viscous face flux Fv (reconstruction.cpp:83-164, class_q.hpp:163-169,
// class_f.hpp:52-59) in the arithmetic type U from the face states L, R
// (stored-state form) and the two cells: U = T is the path's own arithmetic
// (for FP64, the reference's expressions in its order); the FP32-mixed path
// may take it in FP64 (FVK_NS64), where the temperature gradient's two
// products cancel most of each other.
template <typename U, typename T>
__device__ __forceinline__ void viscous_flux(const StageArgsT<T>& a, const FlowT<T>& fl,
                                             const CellQG<T>& c1, const CellQG<T>* c2,
                                             const FaceGeo<T>& g, const T L[4], const T R[4],
                                             U Fv[4]) {
  constexpr bool kSame = std::is_same<U, T>::value;
  const U gm1 = kSame ? U(fl.gm1) : U(fl.gm1_64);
  const U nx = U(g.nx), ny = U(g.ny);
  // face-averaged state, reconstruction.cpp:86-92
  U fa[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) fa[k] = U(0.5) * (conserved<U>(fl, k, L[k]) + conserved<U>(fl, k, R[k]));
  const U frho = fa[0], fE = fa[3];
  const U fu = fa[1] / frho, fv = fa[2] / frho;
  const U fP = (fE - (U(0.5) * frho * (fu * fu + fv * fv))) * gm1;
  U ux, uy, vx, vy, Px = 0, Py = 0, rhox = 0, rhoy = 0;
  if (!c2) {              // no-slip wall, reconstruction.cpp:110-123
    U cu, cv;
    if (a.fs_prims) {
      cu = U(fl.u);
      cv = U(fl.v);
    } else {
      const U crho = conserved<U>(fl, 0, c1.q[0]);
      cu = conserved<U>(fl, 1, c1.q[1]) / crho;
      cv = conserved<U>(fl, 2, c1.q[2]) / crho;
    }
    const U dn = U(g.dist);
    const U dudn = -cu / dn;
    const U dvdn = -cv / dn;
    ux = dudn * nx;
    uy = dudn * ny;
    vx = dvdn * nx;
    vy = dvdn * ny;
  } else {                // interior, reconstruction.cpp:124-161
    const U dist = U(g.dist);                        // face_dist
    const U lx = U(g.lx), ly = U(g.ly);
    const U mx = -ny, my = nx;                       // unit p1->p2 (readmesh.cpp:400-406)
    U ddx[4], ddy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const U dl = (U(c2->q[k]) - U(c1.q[k])) / fabs(dist);
      const U dm = U(0.5) * (U(c1.gx[k]) * mx + U(c1.gy[k]) * my + U(c2->gx[k]) * mx +
                             U(c2->gy[k]) * my);
      ddy[k] = dl * ly + dm * ly;
      ddx[k] = dl * lx + dm * lx;
    }
    rhox = ddx[0];
    ux = ddx[1] - rhox * fu / frho;
    vx = ddx[2] - rhox * fv / frho;
    rhoy = ddy[0];
    uy = ddy[1] - rhoy * fu / frho;
    vy = ddy[2] - rhoy * fu / frho;   // the reference uses u here
    Px = gm1 * (ddx[3] - U(0.5) * rhox * (fu * fu + fv * fv) - frho * (fu * ux + fv * vx));
    Py = gm1 * (ddy[3] - U(0.5) * rhoy * (fu * fu + fv * fv) - frho * (fu * uy + fv * vy));
  }
  const U mu = kSame ? U(fl.mu) : U(fl.mu64);
  const U Rg = kSame ? U(fl.Rgas) : U(fl.Rgas64);
  const U kc = kSame ? U(fl.k) : U(fl.k64);
  const U txx = 2 * mu * (ux - (ux + vy) / 3);
  const U txy = mu * (uy + vx);
  const U tyy = 2 * mu * (vy - (ux + vy) / 3);
  const U dTdx = (Px * frho - rhox * fP) / (Rg * frho * frho);
  const U dTdy = (Py * frho - rhoy * fP) / (Rg * frho * frho);
  Fv[0] = 0;
  Fv[1] = txx * nx + txy * ny;
  Fv[2] = txy * nx + tyy * ny;
  Fv[3] = nx * (fu * txx + fv * txy + kc * dTdx) + ny * (fu * txy + fv * tyy + kc * dTdy);
}

#if FVK_FAST32
// The FP32-mixed viscous flux: viscous_flux's expressions with one hardware
// reciprocal (1 ulp) per divisor, 1/rho_f, 1/|d| (interior) or 1/d_wall and
// 1/(R rho_f^2), in place of its fourteen correctly rounded quotients, and the
// face-averaged state formed from the perturbations, Q_inf + (dL + dR)/2.
template <typename T>
__device__ __forceinline__ void viscous_flux_fast(const StageArgsT<T>& a, const FlowT<T>& fl,
                                                  const CellQG<T>& c1, const CellQG<T>* c2,
                                                  const FaceGeo<T>& g, const T L[4], const T R[4],
                                                  T Fv[4]) {
  static_assert(kDelta<T>, "the FP32-mixed path only");
  const T gm1 = fl.gm1;
  const T nx = g.nx, ny = g.ny;
  T fa[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) fa[k] = __builtin_fmaf(T(0.5), L[k] + R[k], fl.qinf[k]);
  const T frho = fa[0], fE = fa[3];
  const T irho = __builtin_amdgcn_rcpf(frho);
  const T fu = fa[1] * irho, fv = fa[2] * irho;
  const T ke = __builtin_fmaf(fv, fv, fu * fu);
  const T fP = __builtin_fmaf(T(-0.5) * frho, ke, fE) * gm1;
  T ux, uy, vx, vy, Px = 0, Py = 0, rhox = 0, rhoy = 0;
  if (!c2) {              // no-slip wall
    T cu, cv;
    if (a.fs_prims) {
      cu = fl.u;
      cv = fl.v;
    } else {
      const T ic = __builtin_amdgcn_rcpf(fl.qinf[0] + c1.q[0]);
      cu = (fl.qinf[1] + c1.q[1]) * ic;
      cv = (fl.qinf[2] + c1.q[2]) * ic;
    }
    const T idn = __builtin_amdgcn_rcpf(g.dist);
    const T dudn = -cu * idn;
    const T dvdn = -cv * idn;
    ux = dudn * nx;
    uy = dudn * ny;
    vx = dvdn * nx;
    vy = dvdn * ny;
  } else {                // interior
    const T idist = __builtin_amdgcn_rcpf(fabsf(g.dist));
    const T mx = -ny, my = nx;
    T ddx[4], ddy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T dl = (c2->q[k] - c1.q[k]) * idist;
      const T dm = T(0.5) * __builtin_fmaf(c2->gy[k], my, __builtin_fmaf(c2->gx[k], mx,
                                           __builtin_fmaf(c1.gy[k], my, c1.gx[k] * mx)));
      const T s = dl + dm;
      ddy[k] = s * g.ly;
      ddx[k] = s * g.lx;
    }
    rhox = ddx[0];
    rhoy = ddy[0];
    const T rx = rhox * irho, ry = rhoy * irho;
    ux = __builtin_fmaf(-rx, fu, ddx[1]);
    vx = __builtin_fmaf(-rx, fv, ddx[2]);
    uy = __builtin_fmaf(-ry, fu, ddy[1]);
    vy = __builtin_fmaf(-ry, fu, ddy[2]);   // the reference uses u here
    Px = gm1 * __builtin_fmaf(-frho, __builtin_fmaf(fv, vx, fu * ux),
                              __builtin_fmaf(T(-0.5) * rhox, ke, ddx[3]));
    Py = gm1 * __builtin_fmaf(-frho, __builtin_fmaf(fv, vy, fu * uy),
                              __builtin_fmaf(T(-0.5) * rhoy, ke, ddy[3]));
  }
  const T mu = fl.mu, kc = fl.k;
  const T dv3 = (ux + vy) * T(1.0 / 3.0);
  const T txx = T(2) * mu * (ux - dv3);
  const T txy = mu * (uy + vx);
  const T tyy = T(2) * mu * (vy - dv3);
  const T iRT = __builtin_amdgcn_rcpf(fl.Rgas * frho * frho);
  const T dTdx = __builtin_fmaf(Px, frho, -rhox * fP) * iRT;
  const T dTdy = __builtin_fmaf(Py, frho, -rhoy * fP) * iRT;
  Fv[0] = 0;
  Fv[1] = __builtin_fmaf(txy, ny, txx * nx);
  Fv[2] = __builtin_fmaf(tyy, ny, txy * nx);
  Fv[3] = __builtin_fmaf(ny, __builtin_fmaf(kc, dTdy, __builtin_fmaf(fv, tyy, fu * txy)),
                         nx * __builtin_fmaf(kc, dTdx, __builtin_fmaf(fv, txy, fu * txx)));
}
#endif

// F(Q).n of one face side from its primitives (rusanov.cpp:20-31): the
// normal velocity Vn and the four flux components
template <typename T>
__device__ __forceinline__ void side_flux(const Prim<T>& p, T nx, T ny, T F[4], T& Vn) {
  Vn = p.u * nx + p.v * ny;
  F[0] = p.rho * Vn;
  F[1] = p.rho * p.u * Vn + p.P * nx;
  F[2] = p.rho * p.v * Vn + p.P * ny;
  F[3] = Vn * (p.E + p.P);
}

// The Rusanov flux of the two sides and the face's residual terms
// (rusanov.cpp:32-37, res.cpp:81-101, ssp_rk3.cpp:32-34): x = (Fv - F)*A,
// the owner's term, and sterm = (|Vn_avg| + c_avg)*A.
template <typename T>
__device__ __forceinline__ void rusanov_combine(const T L[4], const T R[4], const T FL[4],
                                                const T FR[4], T VnL, T VnR, T cL, T cR,
                                                const T Fv[4], T A, T x[4], T& sterm) {
  const T vn_avg = (VnL + VnR) / 2;
  const T c_avg = (cL + cR) / 2;
  const T lam = (fabs(vn_avg) + c_avg) / 2;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T F = (FL[k] + FR[k]) / 2 - lam * (R[k] - L[k]);
#if FVK_FAST32
    if constexpr (kDelta<T>) {
      x[k] = (Fv[k] - F) * A;
      continue;
    }
#endif
    x[k] = (Fv[k] * A) - (F * A);
  }
  sterm = (fabs(vn_avg) + c_avg) * A;
}

// Roe's flux difference |A~|(Q_R - Q_L) across a face of unit normal n (P. L.
// Roe, J. Comput. Phys. 43, 1981: density-weighted averages of u, v and H;
// split into the two acoustic waves, the entropy wave and the shear wave,
// as in Blazek's textbook treatment of the scheme), from the two face
// states' primitives and total energies and their differences (dr, du, dv,
// dp).  The acoustic eigenvalues take Harten's entropy fix with delta =
// c~/10.  The reference reads ROE (read_inputs.cpp:89-91) but its
// roe_flux.cpp is an unfinished draft outside its solver library
// (res.cpp:66-68 leaves the flux at -101), so this is an extension.
template <typename T>
__device__ __forceinline__ void roe_dissipation(T rL, T uL, T vL, T PL, T EL, T rR, T uR, T vR,
                                                T PR, T ER, T dr, T du, T dv, T dp, T nx, T ny,
                                                T gm1, T D[4]) {
  const T sL = sqrt(rL), sR = sqrt(rR);
  const T w = sL + sR;
  const T HL = (EL + PL) / rL;
  const T HR = (ER + PR) / rR;
  const T rt = sL * sR;                          // Roe averages
  const T ut = (sL * uL + sR * uR) / w;
  const T vt = (sL * vL + sR * vR) / w;
  const T Ht = (sL * HL + sR * HR) / w;
  const T kt = (ut * ut + vt * vt) / 2;
  const T ct = sqrt(gm1 * (Ht - kt));
  const T Vt = ut * nx + vt * ny;
  const T dV = du * nx + dv * ny;
  const T del = ct / 10;
  T l1 = fabs(Vt - ct), l5 = fabs(Vt + ct);
  const T l2 = fabs(Vt);
  if (l1 < del) l1 = (l1 * l1 + del * del) / (2 * del);
  if (l5 < del) l5 = (l5 * l5 + del * del) / (2 * del);
  const T c2 = ct * ct;
  const T a1 = l1 * (dp - rt * ct * dV) / (2 * c2);   // acoustic wave Vn - c
  const T a2 = l2 * (dr - dp / c2);                   // entropy wave
  const T a3 = l2 * rt;                               // shear wave
  const T a5 = l5 * (dp + rt * ct * dV) / (2 * c2);   // acoustic wave Vn + c
  D[0] = a1 + a2 + a5;
  D[1] = a1 * (ut - ct * nx) + a2 * ut + a3 * (du - dV * nx) + a5 * (ut + ct * nx);
  D[2] = a1 * (vt - ct * ny) + a2 * vt + a3 * (dv - dV * ny) + a5 * (vt + ct * ny);
  D[3] = a1 * (Ht - ct * Vt) + a2 * kt + a3 * (ut * du + vt * dv - Vt * dV) + a5 * (Ht + ct * Vt);
}

// The Roe flux F = (F_L + F_R)/2 - D/2 and the face's residual terms, as
// rusanov_combine (sterm keeps the side-averaged |Vn| + c of the time step)
template <typename T>
__device__ __forceinline__ void roe_combine(const T FL[4], const T FR[4], const T D[4], T VnL,
                                            T VnR, T cL, T cR, const T Fv[4], T A, T x[4],
                                            T& sterm) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T F = (FL[k] + FR[k]) / 2 - D[k] / 2;
#if FVK_FAST32
    if constexpr (kDelta<T>) {
      x[k] = (Fv[k] - F) * A;
      continue;
    }
#endif
    x[k] = (Fv[k] * A) - (F * A);
  }
  const T vn_avg = (VnL + VnR) / 2;
  const T c_avg = (cL + cR) / 2;
  sterm = (fabs(vn_avg) + c_avg) * A;
}

// The viscous face flux Fv of face_flux (zero unless NS at second order on an
// interior or no-slip face), in the path's arithmetic.
template <int ORDER, bool NS, typename T>
__device__ __forceinline__ void face_viscous(const StageArgsT<T>& a, const FlowT<T>& fl,
                                             const CellQG<T>& c1, const CellQG<T>* c2, int kind,
                                             const FaceGeo<T>& g, const T L[4], const T R[4],
                                             T Fv[4]) {
  if (NS && ORDER == 2 && (c2 || kind == 3)) {
#if FVK_NS64
    if constexpr (kDelta<T>) {
      double Fd[4];
      viscous_flux<double>(a, fl, c1, c2, g, L, R, Fd);
#pragma unroll
      for (int k = 0; k < 4; ++k) Fv[k] = T(Fd[k]);
    } else
#elif FVK_FAST32
    if constexpr (kDelta<T>) {
      viscous_flux_fast(a, fl, c1, c2, g, L, R, Fv);
    } else
#endif
    {
      viscous_flux<T>(a, fl, c1, c2, g, L, R, Fv);
    }
  }
}

// Pass B: one face with owner c1 and neighbour c2 (interior) or a BC ghost of
// the given kind.  x = (Fv - F)*A, the owner's residual term (the neighbour's
// is -x, res.cpp:81-101); sterm = (|Vn_avg| + c_avg) * A (ssp_rk3.cpp:32-34).
//
// FP32-mixed path: the ghost states are formed as perturbations (a slip wall
// reflects the momentum, m - 2(m.n)n, a no-slip wall reverses it, both keep
// rho and E, which is what the reference's reflected primitives give), and the
// central flux is (dF_L + dF_R)/2: the freestream flux F(Q_inf).n it leaves
// out adds F(Q_inf) . sum(n A) over a closed cell, zero up to the rounding of
// the face normals.
template <int ORDER, bool NS, bool ROE = false, typename T>
__device__ __forceinline__ void face_flux(const StageArgsT<T>& a, const FlowT<T>& fl,
                                          const CellQG<T>& c1,
                                          const CellQG<T>* c2, int kind, const FaceGeo<T>& g,
                                          T x[4], T& sterm) {
  const T gamma = fl.gamma, gm1 = fl.gm1;
  const T nx = g.nx, ny = g.ny, A = g.A;
  T L[4], R[4];               // face states (conserved, or perturbations)
  if (ORDER == 2) {           // reconstruction.cpp:26-29,73-76
    const T dx = g.d1x, dy = g.d1y;
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] = recon(c1.q[k], c1.gx[k], c1.gy[k], dx, dy);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] = c1.q[k];
  }
  T FL[4], FR[4], VnL, VnR, cL, cR;
  // the viscous flux is taken as soon as both face states exist, so that the
  // cells' gradients are no longer live while the inviscid terms are formed
  T Fv[4] = {0, 0, 0, 0};
  if constexpr (kDelta<T>) {
    SideFlux<T> sl, sr;
    side_flux_delta(fl, L, nx, ny, sl);
    if (c2) {
      if (ORDER == 2) {
        const T dx = g.d2x, dy = g.d2y;
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = recon(c2->q[k], c2->gx[k], c2->gy[k], dx, dy);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = c2->q[k];
      }
    } else if (kind == 0) {   // FREESTREAM
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = T(0);
    } else if (kind == 1) {   // SLIP_WALL
      const T mn = sl.rho * sl.Vn;
      R[0] = L[0];
      R[1] = L[1] - T(2.0) * mn * nx;
      R[2] = L[2] - T(2.0) * mn * ny;
      R[3] = L[3];
    } else if (kind == 2) {   // EXTRAPOLATION
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = L[k];
    } else {                  // NO_SLIP_WALL
      R[0] = L[0];
      R[1] = -L[1] - T(2.0) * fl.qinf[1];
      R[2] = -L[2] - T(2.0) * fl.qinf[2];
      R[3] = L[3];
    }
    face_viscous<ORDER, NS>(a, fl, c1, c2, kind, g, L, R, Fv);
    side_flux_delta(fl, R, nx, ny, sr);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      FL[k] = sl.dF[k];
      FR[k] = sr.dF[k];
    }
    VnL = sl.Vn; VnR = sr.Vn; cL = sl.c; cR = sr.c;
    if constexpr (ROE) {
      T D[4];
      roe_dissipation(sl.rho, sl.u, sl.v, sl.P, fl.qinf[3] + L[3], sr.rho, sr.u, sr.v, sr.P,
                      fl.qinf[3] + R[3], R[0] - L[0], sr.u - sl.u, sr.v - sl.v, sr.P - sl.P, nx,
                      ny, gm1, D);
      roe_combine(FL, FR, D, VnL, VnR, cL, cR, Fv, A, x, sterm);
      return;
    }
  } else {
    Prim<T> pl;
    to_prim(L, gamma, gm1, pl);
    if (c2) {
      if (ORDER == 2) {
        const T dx = g.d2x, dy = g.d2y;
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = recon(c2->q[k], c2->gx[k], c2->gy[k], dx, dy);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = c2->q[k];
      }
    } else if (kind == 0) {   // FREESTREAM
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = fl.qinf[k];
    } else if (kind == 1) {   // SLIP_WALL: reflect the reconstructed velocity
      const T Vn = pl.u * nx + pl.v * ny;
      const T wu = pl.u - T(2.0) * Vn * nx;
      const T wv = pl.v - T(2.0) * Vn * ny;
      ghost_q(gm1, pl.rho, wu, wv, pl.P, R);
    } else if (kind == 2) {   // EXTRAPOLATION
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = L[k];
    } else {                  // NO_SLIP_WALL
      ghost_q(gm1, pl.rho, -pl.u, -pl.v, pl.P, R);
    }
    face_viscous<ORDER, NS>(a, fl, c1, c2, kind, g, L, R, Fv);
    Prim<T> pr;
    to_prim(R, gamma, gm1, pr);
    side_flux(pl, nx, ny, FL, VnL);
    side_flux(pr, nx, ny, FR, VnR);
    cL = pl.c;
    cR = pr.c;
    if constexpr (ROE) {
      T D[4];
      roe_dissipation(pl.rho, pl.u, pl.v, pl.P, pl.E, pr.rho, pr.u, pr.v, pr.P, pr.E,
                      pr.rho - pl.rho, pr.u - pl.u, pr.v - pl.v, pr.P - pl.P, nx, ny, gm1, D);
      roe_combine(FL, FR, D, VnL, VnR, cL, cR, Fv, A, x, sterm);
      return;
    }
  }
  rusanov_combine(L, R, FL, FR, VnL, VnR, cL, cR, Fv, A, x, sterm);
}

// Residual / volume, dt, SSP-RK3 (or two-stage SSP-RK2) stage combination and
// stores of one cell.  The combinations' coefficients sum to one, so they
// apply unchanged to a perturbation state.  The step's last stage stores the
// spectral sums of the next step's dt.
// Uniform time step (time_step_type 3, an extension: one dt for every cell,
// dt = 2*CFL / max_i(Sigma_i / V_i), so the run is time-accurate): the dt of
// the step whose maximum has FP64 bits `bits`.
template <typename T>
__device__ __forceinline__ double uniform_dt(const FlowT<T>& fl, unsigned long long bits) {
  return double(fl.cfl2) / __longlong_as_double((long long)bits);
}

// the value stage 3 maximises for the next step: Sigma_i (GLOBAL) or
// Sigma_i / V_i (uniform)
template <typename T>
__device__ __forceinline__ double dt_key(const StageArgsT<T>& a, T sg, double V) {
  return a.dt_uniform ? double(sg) / V : double(sg);
}

// stage 1, by the instance's first block: clear the maximum stage 3 will
// build and, with the uniform dt, advance the instance's physical time
template <typename T>
__device__ __forceinline__ void step_start(const StageArgsT<T>& a, const FlowT<T>& fl, int inst) {
  if (a.dt_uniform) a.time[inst] += uniform_dt(fl, a.smax_cur[inst]);
  a.smax_next[inst] = 0ull;
}

template <typename T>
__device__ __forceinline__ void cell_update(const StageArgsT<T>& a, const FlowT<T>& fl, int inst,
                                            int c, const T q[4], const T qn[4], T r[4], T sg,
                                            double V_d, T sig_old) {
  const T V = T(V_d);
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = r[k] / V;
  T dt;
  if (a.dt_uniform) {
    dt = T(uniform_dt(fl, a.smax_cur[inst]));
  } else {
    const T denom = a.dt_global ? T(__longlong_as_double((long long)a.smax_cur[inst])) : sig_old;
    dt = (fl.cfl2 * V) / denom;
  }
  T out[4];
  if (a.stage == 1) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = q[k] + dt * r[k];
  } else {
    // ca Q_n + cb (Q + dt R): SSP-RK3 stage 2 (3/4, 1/4) and 3 (1/3, 2/3)
    // (ssp_rk3.cpp:58-59,68-69), the two-stage SSP-RK2's stage 2 (1/2, 1/2)
    const T ca = a.ca, cb = a.cb;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = ca * qn[k] + cb * q[k] + cb * dt * r[k];
  }
  constexpr int kNT = kDelta<T> ? 2 : 8;
  st4s<kNT>(a.qout + 4 * (size_t)c, out);
  if (a.last) {
    sst<kNT>(a.sigma + c, sg);
    if (a.store_res) st4s<kNT>(a.res + 4 * (size_t)c, r);
  }
}

// cell_update with its own loads (cell-parallel and slot-parallel shapes)
template <typename T>
__device__ __forceinline__ void cell_update_ld(const StageArgsT<T>& a, const FlowT<T>& fl, int c,
                                               const T q[4], T r[4], T sg) {
  T qn[4] = {0, 0, 0, 0};
  if (a.stage > 1) load_q(a.qn, c, qn);
  const T sig_old = a.dt_global ? T(1) : a.sigma[c];
  cell_update(a, fl, 0, c, q, qn, r, sg, a.m.vol[c], sig_old);
}

// LSQ inverse of cell c: FP64, or the same values rounded once to FP32 where
// the engine provides them (FP32-mixed on meshes whose every normal matrix is
// well conditioned: the products below stay FP64, so the gradient's relative
// error stays within about 6e-8 times the condition number)
__device__ __forceinline__ void load_lsq_inv(const DevMesh& m, int c, double& ixx, double& iyy,
                                             double& ixy) {
  if (m.ixx32) {
    ixx = sld<4>(m.ixx32 + c);
    iyy = sld<4>(m.iyy32 + c);
    ixy = sld<4>(m.ixy32 + c);
  } else {
    ixx = sld<4>(m.ixx + c);
    iyy = sld<4>(m.iyy + c);
    ixy = sld<4>(m.ixy + c);
  }
}

// gradient from the LSQ sums (cell_gradient.cpp:120-130); the FP32-mixed
// path multiplies by the inverse in FP64
template <typename T>
__device__ __forceinline__ void lsq_solve(double ixx, double iyy, double ixy, const T sx[4],
                                          const T sy[4], T gx[4], T gy[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (kDelta<T>) {
      gx[k] = T(double(sx[k]) * iyy - double(sy[k]) * ixy);
      gy[k] = T(double(sy[k]) * ixx - double(sx[k]) * ixy);
    } else {
      gx[k] = (sx[k] * iyy - sy[k] * ixy);
      gy[k] = (sy[k] * ixx - sx[k] * ixy);
    }
  }
}

template <typename T>
__device__ __forceinline__ void write_grad(const StageArgsT<T>& a, int c, const T gx[4],
                                           const T gy[4]) {
  constexpr int kNT = kDelta<T> ? 2 : 128;
  st4s<kNT>(a.g + 8 * (size_t)c, gx);
  st4s<kNT>(a.g + 8 * (size_t)c + 4, gy);
}

template <typename T>
__device__ __forceinline__ void load_lagged(const StageArgsT<T>& a, int c, T qp[4], T gpx[4],
                                            T gpy[4]) {
  // the buffers always hold finite values (zero-filled at creation), and
  // face_phi ignores them while the sentinel is armed
  ld4s<64>(a.qp + 4 * (size_t)c, qp);
  ld4s<64>(a.gp + 8 * (size_t)c, gpx);
  ld4s<64>(a.gp + 8 * (size_t)c + 4, gpy);
}

// The Barth-Jespersen limiter (limiter 3, an extension: T. J. Barth and D. C.
// Jespersen, AIAA 89-0366, the multidimensional form of the minmod-type
// limiter min(1, r)) bounds each face value of the CURRENT reconstruction,
// Q + grad Q . (x_f - x_i), by the min/max of the cell and its face
// neighbours: the squeeze's test with the current state and unlimited
// gradient in place of the lagged ones.
template <typename T>
__device__ __forceinline__ void current_as_lagged(const T q[4], const T gx[4], const T gy[4],
                                                  T qp[4], T gpx[4], T gpy[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qp[k] = q[k];
    gpx[k] = gx[k];
    gpy[k] = gy[k];
  }
}

// the reference's min/max start values 1e11 / -1e11 (squeeze.cpp:10-11)
template <typename T>
__device__ __forceinline__ void init_minmax(T qmin[4], T qmax[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qmin[k] = T(100000000000.0);
    qmax[k] = T(-100000000000.0);
  }
}

template <typename T>
__device__ __forceinline__ void accumulate_lsq(T sx[4], T sy[4], T qmin[4], T qmax[4],
                                               const T px[4], const T py[4], const T tmin[4],
                                               const T tmax[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sx[k] += px[k];
    sy[k] += py[k];
    qmin[k] = min2(qmin[k], tmin[k]);
    qmax[k] = max2(qmax[k], tmax[k]);
  }
}

// ---------------------------------------------------------------------------
// cell-parallel shape
// ---------------------------------------------------------------------------
template <typename T, bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_cell(StageArgsT<T> a) {
  const FlowT<T>& fl = a.flows[0];   // single-instance shape
  const int c = logical_block(a.xcd_blocks) * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  if (c >= m.ncp) return;
  const int deg = m.deg[c];
  if (deg == 0) return;
  T q[4];
  load_q(a.qc, c, q);
  const double xc = m.ccx[c], yc = m.ccy[c];
  const int base = m.chunk_off[c >> 6] + (c & 63);
  T sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
  init_minmax(qmin, qmax);
  for (int j = 0; j < deg; ++j) {
    const int s = base + j * 64;
    const int code = m.code[s];
    T px[4], py[4], tmin[4], tmax[4];
    if (code >= 0) {
      const int nb = code >> 1;
      T qn[4];
      load_q(a.qc, nb, qn);
      lsq_interior(code & 1, q, xc, yc, qn, m.ccx[nb], m.ccy[nb], px, py, tmin, tmax);
    } else {
      lsq_boundary(a, fl, code, q, xc, yc, m.s_nx[s], m.s_ny[s], m.s_fcx[s], m.s_fcy[s], px, py,
                   tmin, tmax);
    }
    accumulate_lsq(sx, sy, qmin, qmax, px, py, tmin, tmax);
  }
  T gx[4], gy[4];
  double ixx, iyy, ixy;
  load_lsq_inv(m, c, ixx, iyy, ixy);
  lsq_solve(ixx, iyy, ixy, sx, sy, gx, gy);
  if (SQUEEZE) {
    T phi[4] = {1, 1, 1, 1}, qp[4], gpx[4], gpy[4];
    if (fl.bj) current_as_lagged(q, gx, gy, qp, gpx, gpy);
    else load_lagged(a, c, qp, gpx, gpy);
    const bool sent = a.sentinel && !fl.bj;
    for (int j = 0; j < deg; ++j) {
      const int s = base + j * 64;
      T ph[4];
      face_phi(sent, fl, T(m.s_fcx[s] - xc), T(m.s_fcy[s] - yc), q, qmin, qmax, qp, gpx, gpy, ph);
#pragma unroll
      for (int k = 0; k < 4; ++k) phi[k] = min2(ph[k], phi[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gx[k] = gx[k] * phi[k];
      gy[k] = gy[k] * phi[k];
    }
  }
  write_grad(a, c, gx, gy);
}

template <typename T, int ORDER, bool NS>
__global__ __launch_bounds__(kMaxBlock) void k_pass_b_cell(StageArgsT<T> a) {
  const FlowT<T>& fl = a.flows[0];   // single-instance shape
  const int c = logical_block(a.xcd_blocks) * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  double sig_new = 0.0;
  if (c < m.ncp && m.deg[c] > 0) {
    const int deg = m.deg[c];
    CellQG<T> me;
    load_cell(a, c, ORDER == 2, me);
    const int base = m.chunk_off[c >> 6] + (c & 63);
    T r[4] = {0, 0, 0, 0}, sg = 0;
    for (int j = 0; j < deg; ++j) {
      const int s = base + j * 64;
      const int code = m.code[s];
      T x[4], st;
      bool own = true;
      if (code >= 0) {
        CellQG<T> nb;
        load_cell(a, code >> 1, ORDER == 2, nb);
        own = (code & 1) != 0;
        if (own) face_flux<ORDER, NS>(a, fl, me, &nb, -1, slot_geo<T, NS>(m, s, me, &nb, -1), x, st);
        else face_flux<ORDER, NS>(a, fl, nb, &me, -1, slot_geo<T, NS>(m, s, nb, &me, -1), x, st);
      } else {
        const int kind = (~code) & 3;
        face_flux<ORDER, NS>(a, fl, me, (const CellQG<T>*)nullptr, kind,
                             slot_geo<T, NS>(m, s, me, nullptr, kind), x, st);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = r[k] + (own ? x[k] : -x[k]);
      sg += st;
    }
    cell_update_ld(a, fl, c, me.q, r, sg);
    sig_new = dt_key(a, sg, m.vol[c]);
  }
  if (a.last && a.dt_global) block_max_to(sig_new, a.smax_next);
  if (a.stage == 1 && blockIdx.x == 0 && threadIdx.x == 0) step_start(a, fl, 0);
}

// ---------------------------------------------------------------------------
// slot-parallel shape: block = one 64-cell chunk, kRows slot rows at a time
// ---------------------------------------------------------------------------
template <typename T, bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_slot(StageArgsT<T> a) {
  const FlowT<T>& fl = a.flows[0];   // single-instance shape
  __shared__ T t1[kRows][16][64];        // per row: px[4] py[4] tmin[4] tmax[4]
  __shared__ T cv[16][64];               // qmin[4] qmax[4] (and gx[4] gy[4]) of each cell
  const DevMesh& m = a.m;
  const int chunk = logical_block(a.xcd_blocks);
  if (chunk * 64 >= m.ncp) return;         // whole block idle (uniform)
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int c = chunk * 64 + lane;
  const int deg = m.deg[c];
  const bool active = deg > 0;
  const int cbase = m.chunk_off[chunk];
  const int nrows = (m.chunk_off[chunk + 1] - cbase) >> 6;
  const int base = cbase + lane;
  T q[4] = {0, 0, 0, 0};
  double xc = 0, yc = 0;
  if (active) {
    load_q(a.qc, c, q);
    xc = m.ccx[c];
    yc = m.ccy[c];
  }
  T sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
  init_minmax(qmin, qmax);
  for (int j0 = 0; j0 < nrows; j0 += kRows) {
    const int j = j0 + row;
    if (j < deg) {
      const int s = base + j * 64;
      const int code = m.code[s];
      T px[4], py[4], tmin[4], tmax[4];
      if (code >= 0) {
        const int nb = code >> 1;
        T qn[4];
        load_q(a.qc, nb, qn);
        lsq_interior(code & 1, q, xc, yc, qn, m.ccx[nb], m.ccy[nb], px, py, tmin, tmax);
      } else {
        lsq_boundary(a, fl, code, q, xc, yc, m.s_nx[s], m.s_ny[s], m.s_fcx[s], m.s_fcy[s], px, py,
                     tmin, tmax);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        t1[row][k][lane] = px[k];
        t1[row][4 + k][lane] = py[k];
        t1[row][8 + k][lane] = tmin[k];
        t1[row][12 + k][lane] = tmax[k];
      }
    }
    __syncthreads();
    if (row == 0) {
      const int jn = min(deg, j0 + kRows);
      for (int jj = j0; jj < jn; ++jj) {
        const int rr = jj - j0;
        T px[4], py[4], tmin[4], tmax[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          px[k] = t1[rr][k][lane];
          py[k] = t1[rr][4 + k][lane];
          tmin[k] = t1[rr][8 + k][lane];
          tmax[k] = t1[rr][12 + k][lane];
        }
        accumulate_lsq(sx, sy, qmin, qmax, px, py, tmin, tmax);
      }
    }
    __syncthreads();
  }
  T gx[4], gy[4];
  if (row == 0 && active) {
    double ixx, iyy, ixy;
    load_lsq_inv(m, c, ixx, iyy, ixy);
    lsq_solve(ixx, iyy, ixy, sx, sy, gx, gy);
  }
  if (SQUEEZE) {
    if (row == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        cv[k][lane] = qmin[k];
        cv[4 + k][lane] = qmax[k];
        if (fl.bj && active) {
          cv[8 + k][lane] = gx[k];
          cv[12 + k][lane] = gy[k];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      qmin[k] = cv[k][lane];
      qmax[k] = cv[4 + k][lane];
    }
    T qp[4] = {0, 0, 0, 0}, gpx[4] = {0, 0, 0, 0}, gpy[4] = {0, 0, 0, 0};
    if (active) {
      if (fl.bj) {
        T ux[4], uy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          ux[k] = cv[8 + k][lane];
          uy[k] = cv[12 + k][lane];
        }
        current_as_lagged(q, ux, uy, qp, gpx, gpy);
      } else {
        load_lagged(a, c, qp, gpx, gpy);
      }
    }
    const bool sent = a.sentinel && !fl.bj;
    T phi[4] = {1, 1, 1, 1};
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        const int s = base + j * 64;
        T ph[4];
        face_phi(sent, fl, T(m.s_fcx[s] - xc), T(m.s_fcy[s] - yc), q, qmin, qmax, qp, gpx, gpy, ph);
#pragma unroll
        for (int k = 0; k < 4; ++k) t1[row][k][lane] = ph[k];
      }
      __syncthreads();
      if (row == 0) {
        const int jn = min(deg, j0 + kRows);
        for (int jj = j0; jj < jn; ++jj)
#pragma unroll
          for (int k = 0; k < 4; ++k) phi[k] = min2(t1[jj - j0][k][lane], phi[k]);
      }
      __syncthreads();
    }
    if (row == 0 && active) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gx[k] = gx[k] * phi[k];
        gy[k] = gy[k] * phi[k];
      }
    }
  }
  if (row == 0 && active) write_grad(a, c, gx, gy);
}

template <typename T, int ORDER, bool NS>
__global__ __launch_bounds__(kMaxBlock) void k_pass_b_slot(StageArgsT<T> a) {
  const FlowT<T>& fl = a.flows[0];   // single-instance shape
  __shared__ T t1[kRows][5][64];         // per row: signed residual terms[4], sigma term
  const DevMesh& m = a.m;
  const int chunk = logical_block(a.xcd_blocks);
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int c = chunk * 64 + lane;
  double sig_new = 0.0;
  if (chunk * 64 < m.ncp) {                // uniform over the block
    const int deg = m.deg[c];
    const bool active = deg > 0;
    const int cbase = m.chunk_off[chunk];
    const int nrows = (m.chunk_off[chunk + 1] - cbase) >> 6;
    const int base = cbase + lane;
    CellQG<T> me;
    load_cell(a, c, ORDER == 2, me);   // padding cells too: c < ncp
    T r[4] = {0, 0, 0, 0}, sg = 0;
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        const int s = base + j * 64;
        const int code = m.code[s];
        T x[4], st;
        bool own = true;
        if (code >= 0) {
          CellQG<T> nb;
          load_cell(a, code >> 1, ORDER == 2, nb);
          own = (code & 1) != 0;
          if (own) face_flux<ORDER, NS>(a, fl, me, &nb, -1, slot_geo<T, NS>(m, s, me, &nb, -1), x, st);
          else face_flux<ORDER, NS>(a, fl, nb, &me, -1, slot_geo<T, NS>(m, s, nb, &me, -1), x, st);
        } else {
          const int kind = (~code) & 3;
          face_flux<ORDER, NS>(a, fl, me, (const CellQG<T>*)nullptr, kind,
                               slot_geo<T, NS>(m, s, me, nullptr, kind), x, st);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) t1[row][k][lane] = own ? x[k] : -x[k];
        t1[row][4][lane] = st;
      }
      __syncthreads();
      if (row == 0) {
        const int jn = min(deg, j0 + kRows);
        for (int jj = j0; jj < jn; ++jj) {
#pragma unroll
          for (int k = 0; k < 4; ++k) r[k] = r[k] + t1[jj - j0][k][lane];
          sg += t1[jj - j0][4][lane];
        }
      }
      __syncthreads();
    }
    if (row == 0 && active) {
      cell_update_ld(a, fl, c, me.q, r, sg);
      sig_new = dt_key(a, sg, m.vol[c]);
    }
  }
  if (a.last && a.dt_global) block_max_to(sig_new, a.smax_next);
  if (a.stage == 1 && blockIdx.x == 0 && threadIdx.x == 0) step_start(a, fl, 0);
}

// ---------------------------------------------------------------------------
// LDS-staged shape: one block per mesh block (blockDim == blk, or 2*blk for
// pass B).  LDS holds NQ rows of T per local cell (state, and for pass B the
// gradient), then, where the kernel needs them, two rows of FP64 centroids,
// then (pass B) the face terms.  The FP32 pass B reads per-face offsets
// instead of centroids (FaceGeo from the bf32 arrays).  Every global load a
// thread needs is issued before the block's first barrier, so a block pays
// one round trip to memory (two for halo cells, whose index is read first)
// instead of one per face slot.
// ---------------------------------------------------------------------------
// Tuning constants of the LDS-staged kernels (measured on MI355X):
//   PF / kFPF   face slots per cell / block faces per thread prefetched
//               before the staging barrier.  PF is a template parameter:
//               kPF for meshes whose cells have at most kPF faces, kPFDeep
//               for the rest (polygonal meshes: 12% less time per launch
//               there, while the deep form costs FP64 pass B 8% on
//               triangle/quad meshes);
//   kMinBlocksA / kMinBlocksB  minimum resident blocks per CU the register
//               allocation must allow, per state type (FP64, FP32);
//   FVK_NS_LATE the FP64 NS pass B prefetches one face (not kFPF) and loads
//               the stage-update data after its face phase: 124 VGPRs
//               instead of 154, 4 waves per SIMD instead of 3;
//   FVK_A_GW    pass A writes each wave's gradient records through LDS, with
//               stores that each cover a contiguous 1 KB: on the 256-instance
//               batch FP32-mixed 327 -> 318 us per launch, FP64 637 -> 590 us
//               (560 us with the non-temporal stores);
//   FVK_A_LATE  pass A loads the limiter's data (lagged state and gradient,
//               face positions) after the gradient loop (1), the LSQ
//               inverse too (2), or all before the staging barrier (0).  1
//               takes the SQUEEZE pass A from 85 to 66 VGPRs (FP32, 7 waves
//               per SIMD) and from 145 to 110 (FP64, 4 waves): 8% (FP32) and
//               5% (FP64) less time per launch on the 256-instance batch.
#ifndef FVK_PF
#define FVK_PF 4
#endif
#ifndef FVK_PF_DEEP
#define FVK_PF_DEEP 8
#endif
#ifndef FVK_FPF
#define FVK_FPF 2
#endif
#ifndef FVK_NS_LATE
#define FVK_NS_LATE 1
#endif
#ifndef FVK_A_LATE
#define FVK_A_LATE 1
#endif
#ifndef FVK_A_GW
#define FVK_A_GW 1
#endif
#ifndef FVK_MINBLK_A64
#define FVK_MINBLK_A64 1
#endif
#ifndef FVK_MINBLK_A32
#define FVK_MINBLK_A32 1
#endif
#ifndef FVK_MINBLK_B64
#define FVK_MINBLK_B64 1
#endif
#ifndef FVK_MINBLK_B32
#define FVK_MINBLK_B32 1
#endif
constexpr int kPF = FVK_PF;
constexpr int kPFDeep = FVK_PF_DEEP;
constexpr int kFPF = FVK_FPF;
template <typename T>
constexpr int kMinBlocksA = kDelta<T> ? FVK_MINBLK_A32 : FVK_MINBLK_A64;
template <typename T>
constexpr int kMinBlocksB = kDelta<T> ? FVK_MINBLK_B32 : FVK_MINBLK_B64;

// A block's staged cells in LDS: one record of NQ values per cell (state,
// then the limited gradient's x and y parts), and the centroids as (x, y)
// pairs.  A record is a whole number of 16-byte words, so a cell's state or
// gradient part moves with one 128-bit LDS access (FP32) or two (FP64)
// instead of four 32- or 64-bit ones at four addresses.
template <typename T, int NQ, bool XY>
struct LdsCells {
  static_assert((NQ * sizeof(T)) % 16 == 0, "records of whole 16-byte words");
  T* q;        // [stride][NQ]
  double* xy;  // [stride][2] (XY only)
  int stride;
  __device__ LdsCells(unsigned char* base, int s)
      : q(reinterpret_cast<T*>(base)),
        xy(reinterpret_cast<double*>(base + NQ * s * (int)sizeof(T))),
        stride(s) {}
  __device__ size_t bytes() const {
    return NQ * stride * sizeof(T) + (XY ? 2 * stride * sizeof(double) : 0);
  }
  __device__ T* rec(int l) const { return q + NQ * l; }
};

// four consecutive values of a 16-byte aligned LDS record
template <typename T>
__device__ __forceinline__ void lds_ld4(const T* p, T v[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  } else {
    const double2 t0 = *reinterpret_cast<const double2*>(p);
    const double2 t1 = *reinterpret_cast<const double2*>(p + 2);
    v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
  }
}

template <typename T>
__device__ __forceinline__ void lds_st4(T* p, const T v[4]) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
  }
}

// centroid of staged cell l
// The gradient records of a wave's 64 consecutive cells, first cell c0
// (FVK_A_GW): each lane puts its cell's record (32 or 64 bytes) in the wave's
// own LDS area, then writes the 16-byte words 64k + l of the wave's range, so
// that each store covers a contiguous 1 KB where per-lane stores of the
// records would leave partial lines.  The area is the wave's alone, and a
// wave's LDS accesses complete in order.  All 64 lanes take part.
template <typename T>
__device__ __forceinline__ void write_grad_lds(T* g, int c0, const T gx[4], const T gy[4],
                                               T* lw) {
  const int l = threadIdx.x & 63;
  lds_st4(lw + 8 * l, gx);
  lds_st4(lw + 8 * l + 4, gy);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  constexpr int kNT = kDelta<T> ? 2 : 16;
  constexpr int kWords = 8 * (int)sizeof(T) / 16;   // 16-byte words per record
  constexpr int kPer = 16 / (int)sizeof(T);         // values per word
  T* base = g + 8 * (size_t)c0;
#pragma unroll
  for (int k = 0; k < kWords; ++k) {
    const int w = 64 * k + l;
    const f4v x = *reinterpret_cast<const f4v*>(lw + kPer * w);
    sst<kNT>(reinterpret_cast<f4v*>(base + kPer * w), x);
  }
}

template <typename T, int NQ, bool XY>
__device__ __forceinline__ void lds_xy(const LdsCells<T, NQ, XY>& L, int l, double& x, double& y) {
  const double2 c = *reinterpret_cast<const double2*>(L.xy + 2 * l);
  x = c.x;
  y = c.y;
}

template <typename T, int NQ, bool XY>
__device__ __forceinline__ void stage_block(const StageArgsT<T>& a, int b, LdsCells<T, NQ, XY>& L) {
  const DevMesh& m = a.m;
  const int hb = m.halo_ptr[b], nh = m.halo_ptr[b + 1] - hb;
  const int n = m.blk + nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int gc = i < m.blk ? b * m.blk + i : m.halo_idx[hb + i - m.blk];
    T v[4];
    load_q(a.qc, gc, v);
    T* r = L.rec(i);
    lds_st4(r, v);
    if (NQ == 12) {
      T gx[4], gy[4];
      load_g(a.g, gc, gx, gy);
      lds_st4(r + 4, gx);
      lds_st4(r + 8, gy);
    }
    if (XY) *reinterpret_cast<double2*>(L.xy + 2 * i) = make_double2(m.ccx[gc], m.ccy[gc]);
  }
}

// Face data of slot s that the limiter needs: the FP64 face centre (FP64
// path; offset from the centroid taken after staging), or the precomputed
// FP32 offset from the centroid (FP32 path).
template <typename T>
using SlotPos = typename std::conditional<kDelta<T>, float, double>::type;

template <typename T>
__device__ __forceinline__ void slot_pos(const DevMesh& m, int s, SlotPos<T>& px, SlotPos<T>& py) {
  if constexpr (kDelta<T>) {
    px = sld<4>(m.s_dfx32 + s);
    py = sld<4>(m.s_dfy32 + s);
  } else {
    px = sld<4>(m.s_fcx + s);
    py = sld<4>(m.s_fcy + s);
  }
}

// ---------------------------------------------------------------------------
// Launch over the mesh blocks (tiles).  PERSIST = false: one workgroup per
// block, the kernel body runs once.  PERSIST = true (kernel_variant 4): a
// persistent grid of as many workgroups as the device holds at once, a
// multiple of 8; workgroup w runs on XCD w % 8 and walks the tiles that the
// one-workgroup-per-block launch deals to that XCD, in the same order,
// striding by the XCD's workgroup count, so the blocks that share an L2 are
// the same.  The assignment is static: no counters, no atomics, no workgroup
// waits for another (a kernel boundary still separates the passes and the
// stages), and each tile's work is that of the per-block launch, bit for bit.
// The body is written once, in the kernel, for both: moved into a device
// function it is optimised apart from the kernel and comes back with more
// registers.
// ---------------------------------------------------------------------------
template <bool PERSIST>
struct TileWalk;

template <>
struct TileWalk<false> {
  template <typename T>
  __device__ explicit TileWalk(const StageArgsT<T>&) {}
  template <typename T>
  __device__ int block(const StageArgsT<T>& a) const { return logical_block(a.xcd_blocks); }
  template <typename T>
  __device__ bool next(const StageArgsT<T>&) { return false; }
};

template <>
struct TileWalk<true> {
  int k;
  template <typename T>
  __device__ explicit TileWalk(const StageArgsT<T>&) : k(blockIdx.x >> 3) {}
  // the tile's logical block, or one that block_of() turns down
  template <typename T>
  __device__ int block(const StageArgsT<T>& a) const {
    const int v = k * 8 + (blockIdx.x & 7);   // its workgroup in the per-block launch
    return k < a.tiles_per_xcd && v < a.tiles_total ? logical_block_of(v, a.xcd_blocks)
                                                    : 0x7fffffff;
  }
  template <typename T>
  __device__ bool next(const StageArgsT<T>& a) {
    __syncthreads();   // the tile's LDS reads are done before the next staging
    k += gridDim.x >> 3;
    return k < a.tiles_per_xcd;
  }
};

// The argument block a tile reads.  Per block: the kernel's.  Persistent:
// read again from the kernarg segment for every tile -- left loop-invariant,
// the compiler hoists the argument loads and everything derived from them out
// of the tile loop, and the spilled scalars cost registers and occupancy.
template <bool PERSIST, typename T>
__device__ __forceinline__ const StageArgsT<T>& tile_args(const StageArgsT<T>& args) {
  if constexpr (!PERSIST) {
    return args;
  } else {
    using CP = const __attribute__((address_space(4))) StageArgsT<T>*;
    CP p = (CP)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const StageArgsT<T>*)p;
  }
}

// Pass A, one thread per cell of the block
template <typename T, bool SQUEEZE, int PF, bool PERSIST = false>
__global__ __launch_bounds__(kMaxBlock, kMinBlocksA<T>) void k_pass_a_lds(StageArgsT<T> args) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TileWalk<PERSIST> walk(args);
  do {
    const StageArgsT<T>& a = tile_args<PERSIST>(args);
    do {   // one tile; `break` ends it
      const DevMesh& m = a.m;
      int b, inst;
      if (!block_of(a, walk.block(a), b, inst)) break;   // uniform over the block
      const FlowT<T>& fl = a.flows[inst];
      const bool squeeze = SQUEEZE && fl.squeeze;
      LdsCells<T, 4, true> L(lds_raw, m.blk + m.max_halo);
      const int i = threadIdx.x;
      const int c = b * m.blk + i;
      const int deg = m.deg[c];
      const int base = m.chunk_off[c >> 6] + (c & 63);
      // prefetch: slot codes and face positions, lagged state, LSQ inverse (the
      // limiter's data after the gradient loop with FVK_A_LATE >= 1, the LSQ
      // inverse too with 2, so that fewer values are live through the loop)
      int lc[PF];
      SlotPos<T> pfx[PF], pfy[PF];
      auto load_pf = [&]() {
#pragma unroll
        for (int j = 0; j < PF; ++j)
          if (j < deg) slot_pos<T>(m, base + j * 64, pfx[j], pfy[j]);
      };
#pragma unroll
      for (int j = 0; j < PF; ++j)
        if (j < deg) lc[j] = sld<4>(m.lcode + base + j * 64);
      T qp[4], gpx[4], gpy[4];
      if (FVK_A_LATE < 1) {
        load_pf();
        if (squeeze && !fl.bj && deg > 0) load_lagged(a, c, qp, gpx, gpy);
      }
      double ixx = 0, iyy = 0, ixy = 0;
      if (FVK_A_LATE < 2 && deg > 0) load_lsq_inv(m, c, ixx, iyy, ixy);
      stage_block(a, b, L);
      __syncthreads();
      // FVK_A_GW: each wave writes its cells' gradient records together, so
      // the padding cells (no faces) stay to the store and write zeros
      constexpr bool kGw = FVK_A_GW != 0;
      if (!kGw && deg == 0) break;
      T gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
      if (deg > 0) {
        T q[4];
        lds_ld4(L.rec(i), q);
        double xc, yc;
        lds_xy(L, i, xc, yc);
        T sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
        init_minmax(qmin, qmax);
        auto slot = [&](int j, int code) {
          T px[4], py[4], tmin[4], tmax[4];
          if (code >= 0) {
            const int nb = code >> 1;
            T qn[4];
            lds_ld4(L.rec(nb), qn);
            double xn, yn;
            lds_xy(L, nb, xn, yn);
            lsq_interior(code & 1, q, xc, yc, qn, xn, yn, px, py, tmin, tmax);
          } else {
            // the 16-bit local code holds the kind only; the global code also
            // names the no-slip ghost's source cell
            const int s = base + j * 64;
            lsq_boundary(a, fl, m.code[s], q, xc, yc, m.s_nx[s], m.s_ny[s], m.s_fcx[s], m.s_fcy[s],
                         px, py, tmin, tmax);
          }
          accumulate_lsq(sx, sy, qmin, qmax, px, py, tmin, tmax);
        };
#pragma unroll
        for (int j = 0; j < PF; ++j)
          if (j < deg) slot(j, lc[j]);
        for (int j = PF; j < deg; ++j) slot(j, sld<4>(m.lcode + base + j * 64));
        if (FVK_A_LATE >= 1) {
          load_pf();
          if (squeeze && !fl.bj) load_lagged(a, c, qp, gpx, gpy);
        }
        if (FVK_A_LATE >= 2) load_lsq_inv(m, c, ixx, iyy, ixy);
        lsq_solve(ixx, iyy, ixy, sx, sy, gx, gy);
        if (squeeze) {
          T phi[4] = {1, 1, 1, 1};
          if (fl.bj) current_as_lagged(q, gx, gy, qp, gpx, gpy);
          const bool sent = a.sentinel && !fl.bj;
          auto lim = [&](SlotPos<T> px, SlotPos<T> py) {
            T ph[4];
            if constexpr (kDelta<T>) face_phi(sent, fl, px, py, q, qmin, qmax, qp, gpx, gpy, ph);
            else face_phi(sent, fl, T(px - xc), T(py - yc), q, qmin, qmax, qp, gpx, gpy, ph);
#pragma unroll
            for (int k = 0; k < 4; ++k) phi[k] = min2(ph[k], phi[k]);
          };
#pragma unroll
          for (int j = 0; j < PF; ++j)
            if (j < deg) lim(pfx[j], pfy[j]);
          for (int j = PF; j < deg; ++j) {
            SlotPos<T> px, py;
            slot_pos<T>(m, base + j * 64, px, py);
            lim(px, py);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            gx[k] = gx[k] * phi[k];
            gy[k] = gy[k] * phi[k];
          }
        }
      }
      if constexpr (kGw)
        write_grad_lds(a.g, c - (i & 63), gx, gy,
                       reinterpret_cast<T*>(lds_raw + L.bytes()) + 8 * (i & ~63));
      else
        write_grad(a, c, gx, gy);
    } while (false);
  } while (walk.next(args));
}

// One block face: owner and neighbour (local indices, or the boundary code)
// and its geometry -- FP64 for the FP64 path (FaceGeo built once the
// centroids are in LDS), precomputed FP32 for the FP32 path.
template <typename T, bool NS>
struct FaceRec {
  int ia, ib;
  double nx, ny, A, fx, fy;
};

template <bool NS>
struct FaceRec<float, NS> {
  int ia, ib;
  FaceGeo<float> g;
};

template <typename T, bool NS>
__device__ __forceinline__ void load_face(const DevMesh& m, int e, FaceRec<T, NS>& f) {
  f.ia = sld<1>(m.bf_a + e);
  f.ib = sld<1>(m.bf_b + e);
  if constexpr (kDelta<T>) {
    f.g.nx = sld<1>(m.bf32_nx + e);
    f.g.ny = sld<1>(m.bf32_ny + e);
    f.g.A = sld<1>(m.bf32_area + e);
    f.g.d1x = sld<1>(m.bf32_d1x + e);
    f.g.d1y = sld<1>(m.bf32_d1y + e);
    f.g.d2x = sld<1>(m.bf32_d2x + e);
    f.g.d2y = sld<1>(m.bf32_d2y + e);
    if (NS) {
      f.g.lx = sld<1>(m.bf32_lx + e);
      f.g.ly = sld<1>(m.bf32_ly + e);
      f.g.dist = sld<1>(m.bf32_dist + e);
    } else {
      f.g.lx = f.g.ly = f.g.dist = 0.0f;
    }
  } else {
    f.nx = sld<1>(m.bf_nx + e);
    f.ny = sld<1>(m.bf_ny + e);
    f.A = sld<1>(m.bf_area + e);
    f.fx = sld<1>(m.bf_fcx + e);
    f.fy = sld<1>(m.bf_fcy + e);
  }
}

template <typename T, int ORDER, int NQ, bool XY>
__device__ __forceinline__ void lds_cell(const LdsCells<T, NQ, XY>& L, int l, CellQG<T>& d) {
  const T* r = L.rec(l);
  lds_ld4(r, d.q);
  if constexpr (ORDER == 2) {
    lds_ld4(r + 4, d.gx);
    lds_ld4(r + 8, d.gy);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) d.gx[k] = d.gy[k] = T(0);
  }
  if constexpr (XY) {
    lds_xy(L, l, d.x, d.y);
  } else {
    d.x = d.y = 0.0;
  }
}

// Pass B, two threads per cell of the block.  LDS: cell rows, then [4 or
// 5][max_bfaces] face terms (the spectral-sum term only at stage 3, where the
// next step's dt needs it)
template <typename T, int ORDER, bool NS, int PF, bool ROE = false, bool PERSIST = false>
__global__ __launch_bounds__(kMaxBlock * kPassBThreadsPerCell, kMinBlocksB<T>)
void k_pass_b_lds(StageArgsT<T> args) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TileWalk<PERSIST> walk(args);
  do {
    const StageArgsT<T>& a = tile_args<PERSIST>(args);
    constexpr int NQ = ORDER == 2 ? 12 : 4;
    constexpr bool XY = !kDelta<T>;
    constexpr bool kLate = FVK_NS_LATE && NS && !kDelta<T>;   // FP64 NS: the leaner prefetch
    constexpr int kFPFB = kLate ? 1 : kFPF;
    const DevMesh& m = a.m;
    int b = 0, inst = 0;
    const bool live = block_of(a, walk.block(a), b, inst);   // uniform
    const FlowT<T>& fl = a.flows[inst];
    double sig_new = 0.0;
    if (live) {
      LdsCells<T, NQ, XY> L(lds_raw, m.blk + m.max_halo);
      T* lf = reinterpret_cast<T*>(lds_raw + L.bytes());
      const int fstride = m.max_bfaces;
      const int fb = m.bf_ptr[b], nfb = m.bf_ptr[b + 1] - fb;
      // threads [0, blk) own the block's cells; all threads share the faces
      const int i = threadIdx.x;
      const bool cell_thread = i < m.blk;
      const int c = b * m.blk + (cell_thread ? i : 0);
      const int deg = cell_thread ? m.deg[c] : 0;
      const int base = m.chunk_off[c >> 6] + (c & 63);
      // prefetch: this thread's faces, its cell's slot-to-face map and the
      // cell data of the stage update.  The FP64 NS kernel, whose face terms
      // need the most registers, prefetches one face and loads the cell data
      // after the face phase, so that fewer values stay live through it.
      FaceRec<T, NS> fr[kFPFB];
#pragma unroll
      for (int t = 0; t < kFPFB; ++t) {
        const int e = i + t * blockDim.x;
        if (e < nfb) load_face(m, fb + e, fr[t]);
      }
      int sf[PF];
#pragma unroll
      for (int j = 0; j < PF; ++j)
        if (j < deg) sf[j] = sld<1>(m.sface + base + j * 64);
      T qn[4] = {0, 0, 0, 0}, sig_old = T(1);
      double V = 1.0;
      if (!kLate && deg > 0) {
        if (a.stage > 1) ld4s<32>(a.qn + 4 * (size_t)c, qn);
        V = sld<32>(m.vol + c);
        if (!a.dt_global) sig_old = sld<32>(a.sigma + c);
      }
      stage_block(a, b, L);
      __syncthreads();
      // one thread per face of the block
      auto face = [&](int e, const FaceRec<T, NS>& f) {
        T x[4], st;
        CellQG<T> c1, c2;
        lds_cell<T, ORDER>(L, f.ia, c1);
        if (f.ib >= 0) {
          lds_cell<T, ORDER>(L, f.ib, c2);
          FaceGeo<T> g;
          if constexpr (kDelta<T>) g = f.g;
          else g = geo_from<T, NS>(f.nx, f.ny, f.A, f.fx, f.fy, c1, &c2, -1);
          face_flux<ORDER, NS, ROE>(a, fl, c1, &c2, -1, g, x, st);
        } else {
          const int kind = (~f.ib) & 3;
          FaceGeo<T> g;
          if constexpr (kDelta<T>) g = f.g;
          else g = geo_from<T, NS>(f.nx, f.ny, f.A, f.fx, f.fy, c1, nullptr, kind);
          face_flux<ORDER, NS, ROE>(a, fl, c1, (const CellQG<T>*)nullptr, kind, g, x, st);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) lf[k * fstride + e] = x[k];
        if (a.last) lf[4 * fstride + e] = st;
      };
#pragma unroll
      for (int t = 0; t < kFPFB; ++t) {
        const int e = i + t * blockDim.x;
        if (e < nfb) face(e, fr[t]);
      }
      for (int e = i + kFPFB * blockDim.x; e < nfb; e += blockDim.x) {
        FaceRec<T, NS> f;
        load_face(m, fb + e, f);
        face(e, f);
      }
      if (kLate && deg > 0) {
        if (a.stage > 1) ld4s<32>(a.qn + 4 * (size_t)c, qn);
        V = sld<32>(m.vol + c);
        if (!a.dt_global) sig_old = sld<32>(a.sigma + c);
      }
      __syncthreads();
      // one thread per cell: add the face terms in slot order (res.cpp:74-102)
      if (deg > 0) {
        T r[4] = {0, 0, 0, 0}, sg = 0;
        auto add = [&](int code) {
          const int e = code >> 1;
          const bool own = (code & 1) != 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const T x = lf[k * fstride + e];
            r[k] = r[k] + (own ? x : -x);
          }
          if (a.last) sg += lf[4 * fstride + e];
        };
#pragma unroll
        for (int j = 0; j < PF; ++j)
          if (j < deg) add(sf[j]);
        for (int j = PF; j < deg; ++j) add(sld<1>(m.sface + base + j * 64));
        T q[4];
        lds_ld4(L.rec(i), q);
        cell_update(a, fl, inst, c, q, qn, r, sg, V, sig_old);
        sig_new = dt_key(a, sg, V);
      }
    }
    if (live && a.last && a.dt_global) block_max_to(sig_new, a.smax_next + inst);
    if (live && a.stage == 1 && threadIdx.x == 0 && b == fl.first_block) step_start(a, fl, inst);
  } while (walk.next(args));
}

// ---------------------------------------------------------------------------
// Atomic-scatter accumulation (accumulate = 1), the alternative to the owner
// gather above.  Each block computes the faces of its cells, as the gather
// does, and scatters each face's terms with LDS atomics (ds_add) into the
// accumulators of whichever of the face's two cells are its own; the halo
// side of a face on a block boundary is skipped, since the neighbouring block
// computes that face too.  The stage update then reads each cell's sums from
// LDS.  No device-scope atomics: on gfx950 those are performed beyond the
// per-XCD L2 caches, at a small fraction of the memory bandwidth.  The
// accumulators are FP64 in both precisions: ds_add_f64 keeps this pass within
// 1.1x of the gather on the 256-instance batch, where ds_add_f32 made it 2.8x
// (and the FP32-mixed face terms are summed in FP64).  The sums take their
// terms in run-time order, so results differ from the gather (and the
// reference) by summation order only.
// ---------------------------------------------------------------------------
template <typename T, int ORDER, bool NS, bool PERSIST = false>
__global__ __launch_bounds__(kMaxBlock * kPassBThreadsPerCell, kMinBlocksB<T>)
void k_pass_b_atomic(StageArgsT<T> args) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TileWalk<PERSIST> walk(args);
  do {
    const StageArgsT<T>& a = tile_args<PERSIST>(args);
    constexpr int NQ = ORDER == 2 ? 12 : 4;
    constexpr bool XY = !kDelta<T>;
    const DevMesh& m = a.m;
    int b = 0, inst = 0;
    const bool live = block_of(a, walk.block(a), b, inst);   // uniform
    const FlowT<T>& fl = a.flows[inst];
    double sig_new = 0.0;
    if (live) {
      LdsCells<T, NQ, XY> L(lds_raw, m.blk + m.max_halo);
      // [5][blk]: the face sums and the spectral sum of each cell of the block
      double* acl = reinterpret_cast<double*>(lds_raw + ((L.bytes() + 7) & ~size_t(7)));
      const int fb = m.bf_ptr[b], nfb = m.bf_ptr[b + 1] - fb;
      const int i = threadIdx.x;
      if (i < m.blk) {
#pragma unroll
        for (int k = 0; k < 5; ++k) acl[k * m.blk + i] = 0.0;
      }
      stage_block(a, b, L);
      __syncthreads();
      for (int e = i; e < nfb; e += blockDim.x) {
        FaceRec<T, NS> f;
        load_face(m, fb + e, f);
        CellQG<T> c1, c2;
        lds_cell<T, ORDER>(L, f.ia, c1);
        T x[4], st;
        if (f.ib >= 0) {
          lds_cell<T, ORDER>(L, f.ib, c2);
          FaceGeo<T> g;
          if constexpr (kDelta<T>) g = f.g;
          else g = geo_from<T, NS>(f.nx, f.ny, f.A, f.fx, f.fy, c1, &c2, -1);
          face_flux<ORDER, NS>(a, fl, c1, &c2, -1, g, x, st);
          if (f.ib < m.blk) {
#pragma unroll
            for (int k = 0; k < 4; ++k) atomicAdd(acl + k * m.blk + f.ib, -double(x[k]));
            if (a.last) atomicAdd(acl + 4 * m.blk + f.ib, double(st));
          }
        } else {
          const int kind = (~f.ib) & 3;
          FaceGeo<T> g;
          if constexpr (kDelta<T>) g = f.g;
          else g = geo_from<T, NS>(f.nx, f.ny, f.A, f.fx, f.fy, c1, nullptr, kind);
          face_flux<ORDER, NS>(a, fl, c1, (const CellQG<T>*)nullptr, kind, g, x, st);
        }
        if (f.ia < m.blk) {
#pragma unroll
          for (int k = 0; k < 4; ++k) atomicAdd(acl + k * m.blk + f.ia, double(x[k]));
          if (a.last) atomicAdd(acl + 4 * m.blk + f.ia, double(st));
        }
      }
      __syncthreads();
      if (i < m.blk) {
        const int c = b * m.blk + i;
        if (m.deg[c] > 0) {
          T r[4], q[4], qn[4] = {0, 0, 0, 0};
#pragma unroll
          for (int k = 0; k < 4; ++k) r[k] = T(acl[k * m.blk + i]);
          const T sg = T(acl[4 * m.blk + i]);
          lds_ld4(L.rec(i), q);
          if (a.stage > 1) load_q(a.qn, c, qn);
          const T sig_old = a.dt_global ? T(1) : a.sigma[c];
          cell_update(a, fl, inst, c, q, qn, r, sg, m.vol[c], sig_old);
          sig_new = dt_key(a, sg, m.vol[c]);
        }
      }
    }
    if (live && a.last && a.dt_global) block_max_to(sig_new, a.smax_next + inst);
    if (live && a.stage == 1 && threadIdx.x == 0 && b == fl.first_block) step_start(a, fl, inst);
  } while (walk.next(args));
}

// Spectral sums of the first step: every face carries |V_inf| and c_inf
// (class_q.hpp:61-62), accumulated as ssp_rk3.cpp:28-36 does.
// One thread block per mesh block (blockDim = blk).
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_init_sigma(StageArgsT<T> a) {
  const DevMesh& m = a.m;
  const int b = blockIdx.x;
  const int inst = a.blk_inst ? a.blk_inst[b] : 0;
  const FlowT<T>& fl = a.flows[inst];
  const int c = b * m.blk + threadIdx.x;
  const int deg = m.deg[c];
  const int base = m.chunk_off[c >> 6] + (c & 63);
  T sg = 0;
  for (int j = 0; j < deg; ++j) sg += (fabs(fl.vn0) + fl.c0) * T(m.s_area[base + j * 64]);
  a.sigma[c] = sg;
  block_max_to(dt_key(a, sg, m.vol[c]), a.smax_cur + inst);
}

// One thread block per mesh block (blockDim = blk, grid = ncp / blk, so every
// thread has a cell): partial[k*nblocks + b].
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_res_sq(const T* __restrict__ res,
                                                      double* __restrict__ partial) {
  __shared__ double ws[4][kMaxBlock / 64];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double r = double(res[4 * (size_t)c + k]);
    s[k] = r * r;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) ws[k][threadIdx.x >> 6] = s[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += ws[threadIdx.x][i];
    partial[threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// Residual norms from the per-block partial sums partial[k*nblocks + b]:
// one wavefront (blockDim = 64) per instance i < n and component k < 4
// (grid = 4n, block 4i + k), over the instance's blocks [inst_off[i],
// inst_off[i+1]) (inst_off null: one instance, all blocks).  Lane l keeps a
// Neumaier-compensated sum of the blocks b = b0 + l, b0 + l + 64, ...; the
// first lane then adds the 64 lane sums in lane order, with their
// compensations, so the result does not depend on anything but the partials;
// out[4*i + k] = its square root.
__global__ __launch_bounds__(64) void k_res_norm(const double* __restrict__ partial, int nblocks,
                                                 const int32_t* __restrict__ inst_off,
                                                 double* __restrict__ out) {
  __shared__ double ls[64], lc[64];
  const int i = blockIdx.x >> 2, k = blockIdx.x & 3, lane = threadIdx.x;
  const int b0 = inst_off ? inst_off[i] : 0, b1 = inst_off ? inst_off[i + 1] : nblocks;
  auto add = [](double& s, double& comp, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) comp += (s - t) + x;
    else comp += (x - t) + s;
    s = t;
  };
  double s = 0.0, comp = 0.0;
  for (int b = b0 + lane; b < b1; b += 64) add(s, comp, partial[k * (size_t)nblocks + b]);
  ls[lane] = s;
  lc[lane] = comp;
  __syncthreads();
  if (lane == 0) {
    double S = 0.0, C = 0.0;
    for (int j = 0; j < 64; ++j) {
      add(S, C, ls[j]);
      C += lc[j];
    }
    out[4 * i + k] = sqrt(S + C);
  }
}

// Error-free transformations (Knuth's TwoSum; the exact product error by one
// fused multiply-add): s + e == a + b and p + e == a * b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Global conservation sums: per mesh block, the compensated sums over its
// cells of V_i * q_k,i (k = 0..3, the stored state: conserved variables, or
// the FP32-mixed perturbation) and of V_i, as (hi, lo) pairs whose sum is the
// block's sum to about twice the working precision:
// partial[(2j)*nblocks + b] = hi_j, partial[(2j+1)*nblocks + b] = lo_j, j = 0..4.
// One thread block per mesh block (blockDim = blk, grid = ncp / blk).
template <typename T>
__global__ __launch_bounds__(kMaxBlock) void k_cons_sum(const T* __restrict__ q,
                                                        const double* __restrict__ vol,
                                                        const uint8_t* __restrict__ deg,
                                                        double* __restrict__ partial) {
  __shared__ double ws[10][kMaxBlock / 64];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double hi[5], lo[5];
  const bool live = deg[c] > 0;
  const double V = live ? vol[c] : 0.0;
  T qc[4];
  ld4(q + 4 * (size_t)c, qc);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double x = live ? double(qc[k]) : 0.0;
    hi[k] = V * x;
    lo[k] = __builtin_fma(V, x, -hi[k]);
  }
  hi[4] = V;
  lo[4] = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    for (int off = 32; off > 0; off >>= 1) {
      const double oh = __shfl_xor(hi[j], off, 64), ol = __shfl_xor(lo[j], off, 64);
      double s, e;
      two_sum(hi[j], oh, s, e);
      hi[j] = s;
      lo[j] = (lo[j] + ol) + e;
    }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      ws[2 * j][threadIdx.x >> 6] = hi[j];
      ws[2 * j + 1][threadIdx.x >> 6] = lo[j];
    }
  __syncthreads();
  if (threadIdx.x < 5) {
    const int j = threadIdx.x;
    double h = ws[2 * j][0], l = ws[2 * j + 1][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      double s, e;
      two_sum(h, ws[2 * j][w], s, e);
      h = s;
      l = (l + ws[2 * j + 1][w]) + e;
    }
    partial[(2 * j) * gridDim.x + blockIdx.x] = h;
    partial[(2 * j + 1) * gridDim.x + blockIdx.x] = l;
  }
}

template <typename T>
__global__ void k_fill(T* p, long long n, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// grid size and XCD range length of a launch over `nlogical` blocks
template <typename T>
void shape(const StageArgsT<T>& a, int nlogical, int* grid, int* per_xcd) {
  if (a.xcd_remap == 1) {
    *per_xcd = (nlogical + 7) / 8;
    *grid = 8 * *per_xcd;
  } else if (a.xcd_remap > 1) {
    const int g = a.xcd_remap;
    *per_xcd = -g;
    *grid = (nlogical + 8 * g - 1) / (8 * g) * (8 * g);
  } else {
    *per_xcd = 0;
    *grid = nlogical;
  }
}

size_t lds_cells_bytes(int stride, int nq, size_t tsize, bool xy = true) {
  return nq * (size_t)stride * tsize + (xy ? 2 * (size_t)stride * 8 : 0);
}

}  // namespace

size_t lds_bytes_pass_a(const DevMesh& m, size_t tsize) {
  return lds_cells_bytes(m.blk + m.max_halo, 4, tsize) +
         (FVK_A_GW ? 8 * (size_t)m.blk * tsize : 0);   // the waves' record areas
}

size_t lds_bytes_pass_b(const DevMesh& m, int order, bool last, size_t tsize) {
  return lds_cells_bytes(m.blk + m.max_halo, order == 2 ? 12 : 4, tsize, tsize == 8) +
         (last ? 5 : 4) * (size_t)m.max_bfaces * tsize;
}

// the atomic-scatter pass B: cell rows, then (8-byte aligned) the block's
// cells' five FP64 sums (no face terms)
static size_t lds_bytes_atomic(const DevMesh& m, int order, size_t tsize) {
  const size_t cells = lds_cells_bytes(m.blk + m.max_halo, order == 2 ? 12 : 4, tsize, tsize == 8);
  return ((cells + 7) & ~size_t(7)) + 5 * (size_t)m.blk * sizeof(double);
}

template <typename T>
using StageKernel = void (*)(StageArgsT<T>);

// Workgroups of a persistent launch: as many as are resident at once (the
// kernel's occupancy at this LDS size times the CUs), a multiple of 8 so that
// every XCD has as many, and no more than the tiles.  The occupancy
// is looked up once per kernel, LDS size and device.
template <typename T>
hipError_t persist_grid(StageKernel<T> k, int threads, size_t lds, int tiles, int* grid) {
  struct Key {
    const void* k;
    int threads, dev;
    size_t lds;
  };
  static std::mutex mu;
  static std::vector<std::pair<Key, int>> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  int resident = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& kv : cache)
      if (kv.first.k == (const void*)k && kv.first.threads == threads && kv.first.dev == dev &&
          kv.first.lds == lds)
        resident = kv.second;
  }
  if (resident == 0) {
    int cus = 0, per_cu = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, threads, lds)) !=
        hipSuccess)
      return e;
    resident = std::max(8, per_cu * cus / 8 * 8);
    std::lock_guard<std::mutex> lock(mu);
    if (cache.size() >= 256) cache.clear();   // many meshes in one process: start again
    cache.push_back({Key{(const void*)k, threads, dev, lds}, resident});
  }
  *grid = std::min(resident, (tiles + 7) / 8 * 8);
  return hipSuccess;
}

// an LDS-staged pass over `tiles` logical blocks: one workgroup per block
// (k1), or the persistent grid (kp) walking them
template <typename T>
hipError_t launch_tiles(StageArgsT<T> a, StageKernel<T> k1, StageKernel<T> kp, int tiles,
                        int threads, size_t lds, hipStream_t s) {
  StageKernel<T> k = k1;
  int grid = tiles;
  if (a.variant == kVariantPersist) {
    const hipError_t e = persist_grid(kp, threads, lds, tiles, &grid);
    if (e != hipSuccess) return e;
    a.tiles_total = tiles;
    a.tiles_per_xcd = (tiles + 7) / 8;
    k = kp;
  }
  void* args[] = {&a};
  return hipLaunchKernel((const void*)k, dim3(grid), dim3(threads), args, lds, s);
}

template <typename T, bool SQUEEZE, int PF>
void pick_a(StageKernel<T>* k1, StageKernel<T>* kp) {
  *k1 = k_pass_a_lds<T, SQUEEZE, PF>;
  *kp = k_pass_a_lds<T, SQUEEZE, PF, true>;
}

template <typename T, int PF, bool ROE>
void pick_b(int order, int ns, StageKernel<T>* k1, StageKernel<T>* kp) {
  if (order == 1) {
    *k1 = k_pass_b_lds<T, 1, false, PF, ROE>;
    *kp = k_pass_b_lds<T, 1, false, PF, ROE, true>;
  } else if (ns) {
    *k1 = k_pass_b_lds<T, 2, true, PF, ROE>;
    *kp = k_pass_b_lds<T, 2, true, PF, ROE, true>;
  } else {
    *k1 = k_pass_b_lds<T, 2, false, PF, ROE>;
    *kp = k_pass_b_lds<T, 2, false, PF, ROE, true>;
  }
}

template <typename T>
hipError_t launch_pass_a(const StageArgsT<T>& a0, int squeeze, int block, hipStream_t s) {
  StageArgsT<T> a = a0;
  int grid;
  if (a.variant == kVariantLds || a.variant == kVariantPersist) {
    shape(a, a.active ? a.nactive : a.m.nblocks, &grid, &a.xcd_blocks);
    if (grid == 0) return hipSuccess;
    StageKernel<T> k1, kp;
    if (a.m.max_deg > kPF) {
      if (squeeze) pick_a<T, true, kPFDeep>(&k1, &kp);
      else pick_a<T, false, kPFDeep>(&k1, &kp);
    } else {
      if (squeeze) pick_a<T, true, kPF>(&k1, &kp);
      else pick_a<T, false, kPF>(&k1, &kp);
    }
    return launch_tiles(a, k1, kp, grid, a.m.blk, lds_bytes_pass_a(a.m, sizeof(T)), s);
  } else if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL((k_pass_a_slot<T, true>), dim3(grid), dim3(64 * kRows), 0, s, a);
    else hipLaunchKernelGGL((k_pass_a_slot<T, false>), dim3(grid), dim3(64 * kRows), 0, s, a);
  } else {
    shape(a, (a.m.ncp + block - 1) / block, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL((k_pass_a_cell<T, true>), dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL((k_pass_a_cell<T, false>), dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pass_b(const StageArgsT<T>& a0, int order, int ns, int block, hipStream_t s) {
  StageArgsT<T> a = a0;
  int grid;
  if (a.variant == kVariantLds || a.variant == kVariantPersist) {
    shape(a, a.active ? a.nactive : a.m.nblocks, &grid, &a.xcd_blocks);
    if (grid == 0) return hipSuccess;
    const int threads = a.m.blk * kPassBThreadsPerCell;
    StageKernel<T> k1, kp;
    if (a.accumulate) {
      if (order == 1) {
        k1 = k_pass_b_atomic<T, 1, false>;
        kp = k_pass_b_atomic<T, 1, false, true>;
      } else if (ns) {
        k1 = k_pass_b_atomic<T, 2, true>;
        kp = k_pass_b_atomic<T, 2, true, true>;
      } else {
        k1 = k_pass_b_atomic<T, 2, false>;
        kp = k_pass_b_atomic<T, 2, false, true>;
      }
      return launch_tiles(a, k1, kp, grid, threads, lds_bytes_atomic(a.m, order, sizeof(T)), s);
    }
    const bool deep = a.m.max_deg > kPF;
    if (a.roe) {   // the Roe-flux pass B (an extension)
      if (deep) pick_b<T, kPFDeep, true>(order, ns, &k1, &kp);
      else pick_b<T, kPF, true>(order, ns, &k1, &kp);
    } else {
      if (deep) pick_b<T, kPFDeep, false>(order, ns, &k1, &kp);
      else pick_b<T, kPF, false>(order, ns, &k1, &kp);
    }
    return launch_tiles(a, k1, kp, grid, threads, lds_bytes_pass_b(a.m, order, a.last != 0, sizeof(T)), s);
  } else if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    const dim3 blk(64 * kRows);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_slot<T, 1, false>), dim3(grid), blk, 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_slot<T, 2, true>), dim3(grid), blk, 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_slot<T, 2, false>), dim3(grid), blk, 0, s, a);
  } else {
    shape(a, (a.m.ncp + block - 1) / block, &grid, &a.xcd_blocks);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_cell<T, 1, false>), dim3(grid), dim3(block), 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_cell<T, 2, true>), dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_cell<T, 2, false>), dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_init_sigma(const StageArgsT<T>& a, hipStream_t s) {
  hipLaunchKernelGGL(k_init_sigma<T>, dim3(a.m.nblocks), dim3(a.m.blk), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_res_sq(const T* res, int32_t ncp, int32_t blk, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_res_sq<T>, dim3(ncp / blk), dim3(blk), 0, s, res, partial);
  return hipGetLastError();
}

hipError_t launch_res_norm(const double* partial, int32_t nblocks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_res_norm, dim3(4), dim3(64), 0, s, partial, (int)nblocks,
                     (const int32_t*)nullptr, out);
  return hipGetLastError();
}

hipError_t launch_res_norm_inst(const double* partial, int32_t nblocks, const int32_t* inst_off,
                               int32_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_res_norm, dim3(4 * n), dim3(64), 0, s, partial, (int)nblocks, inst_off,
                     out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cons_sum(const T* q, const DevMesh& m, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_cons_sum<T>, dim3(m.ncp / m.blk), dim3(m.blk), 0, s, q, m.vol, m.deg, partial);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int block = 256;
  const long long grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)grid), dim3(block), 0, s, p, (long long)n, v);
  return hipGetLastError();
}

#define FVK_INSTANTIATE(T)                                                                      \
  template hipError_t launch_pass_a<T>(const StageArgsT<T>&, int, int, hipStream_t);            \
  template hipError_t launch_pass_b<T>(const StageArgsT<T>&, int, int, int, hipStream_t);       \
  template hipError_t launch_init_sigma<T>(const StageArgsT<T>&, hipStream_t);                  \
  template hipError_t launch_res_sq<T>(const T*, int32_t, int32_t, double*, hipStream_t);       \
  template hipError_t launch_cons_sum<T>(const T*, const DevMesh&, double*, hipStream_t);        \
  template hipError_t launch_fill<T>(T*, int64_t, T, hipStream_t);
FVK_INSTANTIATE(double)
FVK_INSTANTIATE(float)
#undef FVK_INSTANTIATE

}  // namespace fvk
