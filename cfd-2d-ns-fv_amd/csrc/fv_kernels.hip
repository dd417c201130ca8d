// fv_kernels.hip — FP64 explicit FV kernels for gfx950 (MI355X).
//
// Layout (mesh_prep.hpp): cells in blocks (recursive coordinate bisection),
// numbered block-major; per block a halo list and a face list; per cell a
// sliced-ELL list of face slots in increasing ORIGINAL face index.  The
// reference's face loops add into a cell in increasing face index, so adding
// a cell's face terms in slot order reproduces its sums term for term.  Built
// with -ffp-contract=off: every expression is evaluated exactly as written,
// in the reference's association order.
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
//                  adds its faces' terms in slot order.
#include "fv_kernels.hpp"

namespace fvk {
namespace {

constexpr int kMaxBlock = 256;
constexpr int kPassBThreadsPerCell = 2;  // LDS-staged pass B: blockDim = 2 * blk
constexpr int kRows = 4;  // slot rows per block in the slot-parallel shape

__device__ __forceinline__ double min2(double a, double b) { return (b < a) ? b : a; }  // std::min({a,b})
__device__ __forceinline__ double max2(double a, double b) { return (a < b) ? b : a; }  // std::max({a,b})

// XCD-aware block order: blocks are dealt to the 8 XCDs round-robin, so block
// b runs on XCD b%8; give each XCD one contiguous range of the logical blocks
// so that the neighbour gathers of its blocks hit its own L2.
__device__ __forceinline__ int logical_block(int per_xcd) {
  const int b = blockIdx.x;
  return per_xcd > 0 ? (b & 7) * per_xcd + (b >> 3) : b;
}

struct Prim {
  double rho, u, v, E, P, c;
};

// state, limited gradient and centroid of one cell
struct CellQG {
  double q[4], gx[4], gy[4];
  double x, y;
};

// class_Q::updateflow_int (class_q.hpp:143-151), without T (unused downstream)
__device__ __forceinline__ void to_prim(const double q[4], double gamma, double gm1, Prim& p) {
  p.rho = q[0];
  p.E = q[3];
  p.u = q[1] / p.rho;
  p.v = q[2] / p.rho;
  p.P = (p.E - (0.5 * p.rho * (p.u * p.u + p.v * p.v))) * gm1;
  p.c = sqrt(gamma * (p.P / p.rho));
}

// class_flow::updateQ of a ghost state (class_flow.hpp:28-34)
__device__ __forceinline__ void ghost_q(double gm1, double rho, double u, double v, double P,
                                        double w[4]) {
  w[0] = rho;
  w[1] = rho * u;
  w[2] = rho * v;
  w[3] = (P / gm1) + 0.5 * rho * (u * u + v * v);
}

// primitives rho, u, v, P of a cell state (class_Q::updateflow), or the exact
// freestream values the very first stage sees (class_q.hpp:55-58)
__device__ __forceinline__ void cell_prim(const double q[4], const Flow& fl, int fs, double& rho,
                                          double& u, double& v, double& P) {
  if (fs) {
    rho = fl.rho; u = fl.u; v = fl.v; P = fl.P;
    return;
  }
  rho = q[0];
  const double E = q[3];
  u = q[1] / rho;
  v = q[2] / rho;
  P = (E - (0.5 * rho * (u * u + v * v))) * fl.gm1;
}

__device__ __forceinline__ void load_q(const double* __restrict__ qa, int ncp, int c, double q[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = qa[k * ncp + c];
}

__device__ __forceinline__ void load_cell(const StageArgs& a, int c, bool order2, CellQG& d) {
  const int ncp = a.m.ncp;
  load_q(a.qc, ncp, c, d.q);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    d.gx[k] = order2 ? a.g[k * ncp + c] : 0.0;
    d.gy[k] = order2 ? a.g[(4 + k) * ncp + c] : 0.0;
  }
  d.x = a.m.ccx[c];
  d.y = a.m.ccy[c];
}

// Block-wide max of non-negative doubles, then one atomicMax on the bit
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
// face_cell1.
__device__ __forceinline__ void lsq_interior(bool own, const double q[4], double xc, double yc,
                                             const double qn[4], double xn, double yn,
                                             double px[4], double py[4], double tmin[4],
                                             double tmax[4]) {
  double d[4], dxj, dyj;
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
    px[k] = d[k] * dxj;
    py[k] = d[k] * dyj;
  }
}

// Pass A, boundary face: ghost-neighbour terms (cell_gradient.cpp:37-73,
// ghost centroid readmesh.cpp:238-243); the face offers only Q_i to min/max.
__device__ __forceinline__ void lsq_boundary(const StageArgs& a, int code, const double q[4],
                                             double xc, double yc, double nx, double ny,
                                             double fx, double fy, double px[4], double py[4],
                                             double tmin[4], double tmax[4]) {
  const int v = ~code;
  const int kind = v & 3;
  const double rx = fx - xc, ry = fy - yc;
  const double t = 2.0 * (rx * nx + ry * ny);
  const double gx = xc + t * nx, gy = yc + t * ny;
  const double dxj = gx - xc, dyj = gy - yc;
  const double gm1 = a.fl.gm1;
  double d[4];
  if (kind == 0) {            // FREESTREAM
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = a.fl.qinf[k] - q[k];
  } else if (kind == 1) {     // SLIP_WALL: reflect the cell velocity
    double rho, u, vv, P;
    cell_prim(q, a.fl, a.fs_prims, rho, u, vv, P);
    const double Vn = u * nx + vv * ny;
    const double wu = u - 2.0 * Vn * nx;
    const double wv = vv - 2.0 * Vn * ny;
    double w[4];
    ghost_q(gm1, rho, wu, wv, P, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = w[k] - q[k];
  } else if (kind == 2) {     // EXTRAPOLATION
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = 0;
  } else {                    // NO_SLIP_WALL: state of the cell whose index is the face index
    double qs[4];
    load_q(a.qc, a.m.ncp, v >> 2, qs);
    double rho, u, vv, P;
    cell_prim(qs, a.fl, a.fs_prims, rho, u, vv, P);
    double w[4];
    ghost_q(gm1, rho, -u, -vv, P, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = w[k] - q[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    px[k] = d[k] * dxj;
    py[k] = d[k] * dyj;
    tmin[k] = q[k];
    tmax[k] = q[k];
  }
}

// Pass A, second loop: the squeeze factor of one face for this cell from the
// face state the previous res() call left on this cell's side
// (squeeze.cpp:68-121): Q_prev + grad_prev . (x_f - x_i), or -101 at first.
__device__ __forceinline__ void face_phi(int sentinel, double fx, double fy, double xc, double yc,
                                         const double q[4], const double qmin[4],
                                         const double qmax[4], const double qp[4],
                                         const double gpx[4], const double gpy[4], double ph[4]) {
  const double dx = fx - xc, dy = fy - yc;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double qf = sentinel ? -101.0 : qp[k] + gpx[k] * dx + gpy[k] * dy;
    double p = 1.0;
    if (qf < qmin[k]) p = (qmin[k] - q[k]) / (qf - q[k]);
    else if (qf > qmax[k]) p = (qmax[k] - q[k]) / (qf - q[k]);
    ph[k] = p;
  }
}

// Pass B: one face with owner c1 and neighbour c2 (interior) or a BC ghost of
// the given kind.  x = (Fv - F)*A, the owner's residual term (the neighbour's
// is -x, res.cpp:81-101); sterm = (|Vn_avg| + c_avg) * A (ssp_rk3.cpp:32-34).
template <int ORDER, bool NS>
__device__ __forceinline__ void face_flux(const StageArgs& a, const CellQG& c1, const CellQG* c2,
                                          int kind, double nx, double ny, double A, double fx,
                                          double fy, double x[4], double& sterm) {
  const Flow& fl = a.fl;
  const double gamma = fl.gamma, gm1 = fl.gm1;
  double L[4], R[4];
  if (ORDER == 2) {           // reconstruction.cpp:26-29,73-76
    const double dx = fx - c1.x, dy = fy - c1.y;
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] = c1.q[k] + c1.gx[k] * dx + c1.gy[k] * dy;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] = c1.q[k];
  }
  Prim pl;
  to_prim(L, gamma, gm1, pl);
  if (c2) {
    if (ORDER == 2) {
      const double dx = fx - c2->x, dy = fy - c2->y;
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = c2->q[k] + c2->gx[k] * dx + c2->gy[k] * dy;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = c2->q[k];
    }
  } else if (kind == 0) {     // FREESTREAM
#pragma unroll
    for (int k = 0; k < 4; ++k) R[k] = fl.qinf[k];
  } else if (kind == 1) {     // SLIP_WALL: reflect the reconstructed velocity
    const double Vn = pl.u * nx + pl.v * ny;
    const double wu = pl.u - 2.0 * Vn * nx;
    const double wv = pl.v - 2.0 * Vn * ny;
    ghost_q(gm1, pl.rho, wu, wv, pl.P, R);
  } else if (kind == 2) {     // EXTRAPOLATION
#pragma unroll
    for (int k = 0; k < 4; ++k) R[k] = L[k];
  } else {                    // NO_SLIP_WALL
    ghost_q(gm1, pl.rho, -pl.u, -pl.v, pl.P, R);
  }
  Prim pr;
  to_prim(R, gamma, gm1, pr);
  const double VnL = pl.u * nx + pl.v * ny;
  const double VnR = pr.u * nx + pr.v * ny;
  double FL[4], FR[4];
  FL[0] = pl.rho * VnL;
  FL[1] = pl.rho * pl.u * VnL + pl.P * nx;
  FL[2] = pl.rho * pl.v * VnL + pl.P * ny;
  FL[3] = VnL * (pl.E + pl.P);
  FR[0] = pr.rho * VnR;
  FR[1] = pr.rho * pr.u * VnR + pr.P * nx;
  FR[2] = pr.rho * pr.v * VnR + pr.P * ny;
  FR[3] = VnR * (pr.E + pr.P);
  const double vn_avg = (VnL + VnR) / 2;
  const double c_avg = (pl.c + pr.c) / 2;
  double Fv[4] = {0, 0, 0, 0};
  if (NS && ORDER == 2 && (c2 || kind == 3)) {
    // face-averaged state, reconstruction.cpp:86-92
    double fa[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) fa[k] = 0.5 * (L[k] + R[k]);
    const double frho = fa[0], fE = fa[3];
    const double fu = fa[1] / frho, fv = fa[2] / frho;
    const double fP = (fE - (0.5 * frho * (fu * fu + fv * fv))) * gm1;
    double ux, uy, vx, vy, Px = 0, Py = 0, rhox = 0, rhoy = 0;
    if (!c2) {              // no-slip wall, reconstruction.cpp:110-123
      double crho, cu, cv, cP;
      cell_prim(c1.q, fl, a.fs_prims, crho, cu, cv, cP);
      const double dx1 = fx - c1.x, dy1 = fy - c1.y;
      const double dn = fabs(sqrt(dx1 * dx1 + dy1 * dy1));
      const double dudn = -cu / dn;
      const double dvdn = -cv / dn;
      ux = dudn * nx;
      uy = dudn * ny;
      vx = dvdn * nx;
      vy = dvdn * ny;
    } else {                // interior, reconstruction.cpp:124-161
      const double tx = c2->x - c1.x, ty = c2->y - c1.y;
      const double dist = sqrt(tx * tx + ty * ty);   // face_dist, and the norm of l
      const double lx = tx / dist, ly = ty / dist;
      const double mx = -ny, my = nx;                 // unit p1->p2 (readmesh.cpp:400-406)
      double ddx[4], ddy[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dl = (c2->q[k] - c1.q[k]) / fabs(dist);
        const double dm =
            0.5 * (c1.gx[k] * mx + c1.gy[k] * my + c2->gx[k] * mx + c2->gy[k] * my);
        ddy[k] = dl * ly + dm * ly;
        ddx[k] = dl * lx + dm * lx;
      }
      rhox = ddx[0];
      ux = ddx[1] - rhox * fu / frho;
      vx = ddx[2] - rhox * fv / frho;
      rhoy = ddy[0];
      uy = ddy[1] - rhoy * fu / frho;
      vy = ddy[2] - rhoy * fu / frho;   // the reference uses u here
      Px = gm1 * (ddx[3] - 0.5 * rhox * (fu * fu + fv * fv) - frho * (fu * ux + fv * vx));
      Py = gm1 * (ddy[3] - 0.5 * rhoy * (fu * fu + fv * fv) - frho * (fu * uy + fv * vy));
    }
    const double mu = fl.mu;
    const double txx = 2 * mu * (ux - (ux + vy) / 3);
    const double txy = mu * (uy + vx);
    const double tyy = 2 * mu * (vy - (ux + vy) / 3);
    const double dTdx = (Px * frho - rhox * fP) / (fl.Rgas * frho * frho);
    const double dTdy = (Py * frho - rhoy * fP) / (fl.Rgas * frho * frho);
    Fv[1] = txx * nx + txy * ny;
    Fv[2] = txy * nx + tyy * ny;
    Fv[3] = nx * (fu * txx + fv * txy + fl.k * dTdx) + ny * (fu * txy + fv * tyy + fl.k * dTdy);
  }
  const double lam = (fabs(vn_avg) + c_avg) / 2;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double F = (FL[k] + FR[k]) / 2 - lam * (R[k] - L[k]);
    x[k] = (Fv[k] * A) - (F * A);
  }
  sterm = (fabs(vn_avg) + c_avg) * A;
}

// Residual / volume, dt, SSP-RK3 stage combination and stores of one cell.
__device__ __forceinline__ void cell_update(const StageArgs& a, int c, const double q[4],
                                            double r[4], double sg) {
  const int ncp = a.m.ncp;
  const double V = a.m.vol[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = r[k] / V;
  const double denom = a.dt_global ? __longlong_as_double((long long)*a.smax_cur) : a.sigma[c];
  const double dt = (a.fl.cfl2 * V) / denom;
  double out[4];
  if (a.stage == 1) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = q[k] + dt * r[k];
  } else if (a.stage == 2) {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = 0.75 * a.qn[k * ncp + c] + 0.25 * q[k] + 0.25 * dt * r[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      out[k] = (1.0 / 3.0) * a.qn[k * ncp + c] + (2.0 / 3.0) * q[k] + (2.0 / 3.0) * dt * r[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) a.qout[k * ncp + c] = out[k];
  if (a.stage == 3) {
    a.sigma[c] = sg;
    if (a.store_res) {
#pragma unroll
      for (int k = 0; k < 4; ++k) a.res[k * ncp + c] = r[k];
    }
  }
}

// gradient from the LSQ sums (cell_gradient.cpp:120-130)
__device__ __forceinline__ void lsq_solve(const StageArgs& a, int c, const double sx[4],
                                          const double sy[4], double gx[4], double gy[4]) {
  const double ixx = a.m.ixx[c], iyy = a.m.iyy[c], ixy = a.m.ixy[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    gx[k] = (sx[k] * iyy - sy[k] * ixy);
    gy[k] = (sy[k] * ixx - sx[k] * ixy);
  }
}

__device__ __forceinline__ void write_grad(const StageArgs& a, int c, const double gx[4],
                                           const double gy[4]) {
  const int ncp = a.m.ncp;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a.g[k * ncp + c] = gx[k];
    a.g[(4 + k) * ncp + c] = gy[k];
  }
}

__device__ __forceinline__ void load_lagged(const StageArgs& a, int c, double qp[4], double gpx[4],
                                            double gpy[4]) {
  const int ncp = a.m.ncp;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qp[k] = a.sentinel ? 0.0 : a.qp[k * ncp + c];
    gpx[k] = a.sentinel ? 0.0 : a.gp[k * ncp + c];
    gpy[k] = a.sentinel ? 0.0 : a.gp[(4 + k) * ncp + c];
  }
}

__device__ __forceinline__ void init_minmax(double qmin[4], double qmax[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qmin[k] = 100000000000.0;
    qmax[k] = -100000000000.0;
  }
}

// ---------------------------------------------------------------------------
// cell-parallel shape
// ---------------------------------------------------------------------------
template <bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_cell(StageArgs a) {
  const int c = logical_block(a.xcd_blocks) * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  if (c >= m.ncp) return;
  const int deg = m.deg[c];
  if (deg == 0) return;
  double q[4];
  load_q(a.qc, m.ncp, c, q);
  const double xc = m.ccx[c], yc = m.ccy[c];
  const int base = m.chunk_off[c >> 6] + (c & 63);
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
  init_minmax(qmin, qmax);
  for (int j = 0; j < deg; ++j) {
    const int s = base + j * 64;
    const int code = m.code[s];
    double px[4], py[4], tmin[4], tmax[4];
    if (code >= 0) {
      const int nb = code >> 1;
      double qn[4];
      load_q(a.qc, m.ncp, nb, qn);
      lsq_interior(code & 1, q, xc, yc, qn, m.ccx[nb], m.ccy[nb], px, py, tmin, tmax);
    } else {
      lsq_boundary(a, code, q, xc, yc, m.s_nx[s], m.s_ny[s], m.s_fcx[s], m.s_fcy[s], px, py,
                   tmin, tmax);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sx[k] += px[k];
      sy[k] += py[k];
      qmin[k] = min2(qmin[k], tmin[k]);
      qmax[k] = max2(qmax[k], tmax[k]);
    }
  }
  double gx[4], gy[4];
  lsq_solve(a, c, sx, sy, gx, gy);
  if (SQUEEZE) {
    double phi[4] = {1, 1, 1, 1}, qp[4], gpx[4], gpy[4];
    load_lagged(a, c, qp, gpx, gpy);
    for (int j = 0; j < deg; ++j) {
      const int s = base + j * 64;
      double ph[4];
      face_phi(a.sentinel, m.s_fcx[s], m.s_fcy[s], xc, yc, q, qmin, qmax, qp, gpx, gpy, ph);
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

template <int ORDER, bool NS>
__global__ __launch_bounds__(kMaxBlock) void k_pass_b_cell(StageArgs a) {
  const int c = logical_block(a.xcd_blocks) * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  double sig_new = 0.0;
  if (c < m.ncp && m.deg[c] > 0) {
    const int deg = m.deg[c];
    CellQG me;
    load_cell(a, c, ORDER == 2, me);
    const int base = m.chunk_off[c >> 6] + (c & 63);
    double r[4] = {0, 0, 0, 0}, sg = 0.0;
    for (int j = 0; j < deg; ++j) {
      const int s = base + j * 64;
      const int code = m.code[s];
      double x[4], st;
      bool own = true;
      if (code >= 0) {
        CellQG nb;
        load_cell(a, code >> 1, ORDER == 2, nb);
        own = (code & 1) != 0;
        if (own) face_flux<ORDER, NS>(a, me, &nb, -1, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
        else face_flux<ORDER, NS>(a, nb, &me, -1, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
      } else {
        face_flux<ORDER, NS>(a, me, nullptr, (~code) & 3, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = r[k] + (own ? x[k] : -x[k]);
      sg += st;
    }
    cell_update(a, c, me.q, r, sg);
    sig_new = sg;
  }
  if (a.stage == 3 && a.dt_global) block_max_to(sig_new, a.smax_next);
  if (a.stage == 1 && blockIdx.x == 0 && threadIdx.x == 0) *a.smax_next = 0ull;
}

// ---------------------------------------------------------------------------
// slot-parallel shape: block = one 64-cell chunk, kRows slot rows at a time
// ---------------------------------------------------------------------------
template <bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_slot(StageArgs a) {
  __shared__ double t1[kRows][16][64];   // per row: px[4] py[4] tmin[4] tmax[4]
  __shared__ double cv[8][64];           // qmin[4] qmax[4] of each cell
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
  double q[4] = {0, 0, 0, 0}, xc = 0, yc = 0;
  if (active) {
    load_q(a.qc, m.ncp, c, q);
    xc = m.ccx[c];
    yc = m.ccy[c];
  }
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
  init_minmax(qmin, qmax);
  for (int j0 = 0; j0 < nrows; j0 += kRows) {
    const int j = j0 + row;
    if (j < deg) {
      const int s = base + j * 64;
      const int code = m.code[s];
      double px[4], py[4], tmin[4], tmax[4];
      if (code >= 0) {
        const int nb = code >> 1;
        double qn[4];
        load_q(a.qc, m.ncp, nb, qn);
        lsq_interior(code & 1, q, xc, yc, qn, m.ccx[nb], m.ccy[nb], px, py, tmin, tmax);
      } else {
        lsq_boundary(a, code, q, xc, yc, m.s_nx[s], m.s_ny[s], m.s_fcx[s], m.s_fcy[s], px, py,
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
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          sx[k] += t1[rr][k][lane];
          sy[k] += t1[rr][4 + k][lane];
          qmin[k] = min2(qmin[k], t1[rr][8 + k][lane]);
          qmax[k] = max2(qmax[k], t1[rr][12 + k][lane]);
        }
      }
    }
    __syncthreads();
  }
  double gx[4], gy[4];
  if (row == 0 && active) lsq_solve(a, c, sx, sy, gx, gy);
  if (SQUEEZE) {
    if (row == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        cv[k][lane] = qmin[k];
        cv[4 + k][lane] = qmax[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      qmin[k] = cv[k][lane];
      qmax[k] = cv[4 + k][lane];
    }
    double qp[4] = {0, 0, 0, 0}, gpx[4] = {0, 0, 0, 0}, gpy[4] = {0, 0, 0, 0};
    if (active) load_lagged(a, c, qp, gpx, gpy);
    double phi[4] = {1, 1, 1, 1};
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        const int s = base + j * 64;
        double ph[4];
        face_phi(a.sentinel, m.s_fcx[s], m.s_fcy[s], xc, yc, q, qmin, qmax, qp, gpx, gpy, ph);
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

template <int ORDER, bool NS>
__global__ __launch_bounds__(kMaxBlock) void k_pass_b_slot(StageArgs a) {
  __shared__ double t1[kRows][5][64];    // per row: signed residual terms[4], sigma term
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
    CellQG me;
    if (active) load_cell(a, c, ORDER == 2, me);
    double r[4] = {0, 0, 0, 0}, sg = 0.0;
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        const int s = base + j * 64;
        const int code = m.code[s];
        double x[4], st;
        bool own = true;
        if (code >= 0) {
          CellQG nb;
          load_cell(a, code >> 1, ORDER == 2, nb);
          own = (code & 1) != 0;
          if (own) face_flux<ORDER, NS>(a, me, &nb, -1, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
          else face_flux<ORDER, NS>(a, nb, &me, -1, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
        } else {
          face_flux<ORDER, NS>(a, me, nullptr, (~code) & 3, m.s_nx[s], m.s_ny[s], m.s_area[s], m.s_fcx[s], m.s_fcy[s], x, st);
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
      cell_update(a, c, me.q, r, sg);
      sig_new = sg;
    }
  }
  if (a.stage == 3 && a.dt_global) block_max_to(sig_new, a.smax_next);
  if (a.stage == 1 && blockIdx.x == 0 && threadIdx.x == 0) *a.smax_next = 0ull;
}

// ---------------------------------------------------------------------------
// LDS-staged shape: one block per mesh block (blockDim == blk)
// ---------------------------------------------------------------------------

// Stage the block's cells and its halo: rows [first_row, first_row + nrow) of
// the per-cell fields (q0..3, gx0..3, gy0..3, x, y) into lds[row][local].
template <int NROW>
__device__ __forceinline__ void stage_block(const StageArgs& a, int b, int stride, double* lds,
                                            bool with_grad) {
  const DevMesh& m = a.m;
  const int ncp = m.ncp;
  const int hb = m.halo_ptr[b], nh = m.halo_ptr[b + 1] - hb;
  const int n = m.blk + nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int gc = i < m.blk ? b * m.blk + i : m.halo_idx[hb + i - m.blk];
#pragma unroll
    for (int k = 0; k < 4; ++k) lds[k * stride + i] = a.qc[k * ncp + gc];
    if (with_grad) {
#pragma unroll
      for (int k = 0; k < 8; ++k) lds[(4 + k) * stride + i] = a.g[k * ncp + gc];
    }
    lds[(NROW - 2) * stride + i] = m.ccx[gc];
    lds[(NROW - 1) * stride + i] = m.ccy[gc];
  }
}

// Every global load a thread needs is issued before the block's first
// barrier, so a block pays one round trip to memory (two for halo cells,
// whose index is read first) instead of one per face slot.
constexpr int kPF = 4;   // face slots per cell whose data are prefetched
constexpr int kFPF = 2;  // block faces per thread whose records are prefetched

template <bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_lds(StageArgs a) {
  extern __shared__ double lds[];         // [6][stride]: q0..3, x, y
  const DevMesh& m = a.m;
  const int b = logical_block(a.xcd_blocks);
  if (b >= m.nblocks) return;             // uniform over the block
  const int stride = m.blk + m.max_halo;
  const int i = threadIdx.x;
  const int c = b * m.blk + i;
  const int deg = m.deg[c];
  const int base = m.chunk_off[c >> 6] + (c & 63);
  // prefetch: slot codes and face centres, lagged state, LSQ inverse
  int lc[kPF];
  double pfx[kPF], pfy[kPF];
#pragma unroll
  for (int j = 0; j < kPF; ++j) {
    if (j < deg) {
      const int s = base + j * 64;
      lc[j] = m.lcode[s];
      pfx[j] = m.s_fcx[s];
      pfy[j] = m.s_fcy[s];
    }
  }
  double qp[4], gpx[4], gpy[4];
  if (SQUEEZE && deg > 0) load_lagged(a, c, qp, gpx, gpy);
  double ixx = 0, iyy = 0, ixy = 0;
  if (deg > 0) { ixx = m.ixx[c]; iyy = m.iyy[c]; ixy = m.ixy[c]; }
  stage_block<6>(a, b, stride, lds, false);
  __syncthreads();
  if (deg == 0) return;
  double q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = lds[k * stride + i];
  const double xc = lds[4 * stride + i], yc = lds[5 * stride + i];
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
  init_minmax(qmin, qmax);
  auto slot = [&](int j, int code, double fx, double fy) {
    double px[4], py[4], tmin[4], tmax[4];
    if (code >= 0) {
      const int nb = code >> 1;
      double qn[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) qn[k] = lds[k * stride + nb];
      lsq_interior(code & 1, q, xc, yc, qn, lds[4 * stride + nb], lds[5 * stride + nb], px, py,
                   tmin, tmax);
    } else {
      const int s = base + j * 64;
      lsq_boundary(a, code, q, xc, yc, m.s_nx[s], m.s_ny[s], fx, fy, px, py, tmin, tmax);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sx[k] += px[k];
      sy[k] += py[k];
      qmin[k] = min2(qmin[k], tmin[k]);
      qmax[k] = max2(qmax[k], tmax[k]);
    }
  };
#pragma unroll
  for (int j = 0; j < kPF; ++j)
    if (j < deg) slot(j, lc[j], pfx[j], pfy[j]);
  for (int j = kPF; j < deg; ++j) {
    const int s = base + j * 64;
    slot(j, m.lcode[s], m.s_fcx[s], m.s_fcy[s]);
  }
  double gx[4], gy[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {   // cell_gradient.cpp:120-130
    gx[k] = (sx[k] * iyy - sy[k] * ixy);
    gy[k] = (sy[k] * ixx - sx[k] * ixy);
  }
  if (SQUEEZE) {
    double phi[4] = {1, 1, 1, 1};
    auto lim = [&](double fx, double fy) {
      double ph[4];
      face_phi(a.sentinel, fx, fy, xc, yc, q, qmin, qmax, qp, gpx, gpy, ph);
#pragma unroll
      for (int k = 0; k < 4; ++k) phi[k] = min2(ph[k], phi[k]);
    };
#pragma unroll
    for (int j = 0; j < kPF; ++j)
      if (j < deg) lim(pfx[j], pfy[j]);
    for (int j = kPF; j < deg; ++j) lim(m.s_fcx[base + j * 64], m.s_fcy[base + j * 64]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gx[k] = gx[k] * phi[k];
      gy[k] = gy[k] * phi[k];
    }
  }
  write_grad(a, c, gx, gy);
}

// one block face: owner/neighbour codes and geometry
struct FaceRec {
  int ia, ib;
  double nx, ny, A, fx, fy;
};

__device__ __forceinline__ void load_face(const DevMesh& m, int e, FaceRec& f) {
  f.ia = m.bf_a[e];
  f.ib = m.bf_b[e];
  f.nx = m.bf_nx[e];
  f.ny = m.bf_ny[e];
  f.A = m.bf_area[e];
  f.fx = m.bf_fcx[e];
  f.fy = m.bf_fcy[e];
}

template <int ORDER, int NROW>
__device__ __forceinline__ void lds_cell(const double* lds, int stride, int l, CellQG& d) {
#pragma unroll
  for (int k = 0; k < 4; ++k) d.q[k] = lds[k * stride + l];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    d.gx[k] = ORDER == 2 ? lds[(4 + k) * stride + l] : 0.0;
    d.gy[k] = ORDER == 2 ? lds[(8 + k) * stride + l] : 0.0;
  }
  d.x = lds[(NROW - 2) * stride + l];
  d.y = lds[(NROW - 1) * stride + l];
}

template <int ORDER, bool NS>
__global__ __launch_bounds__(kMaxBlock * kPassBThreadsPerCell) void k_pass_b_lds(StageArgs a) {
  // [NROW][stride] cell fields, then [4 or 5][max_bfaces] face terms (the
  // spectral-sum term only at stage 3, where the next step's dt needs it)
  extern __shared__ double lds[];
  constexpr int NROW = ORDER == 2 ? 14 : 6;
  const DevMesh& m = a.m;
  const int b = logical_block(a.xcd_blocks);
  double sig_new = 0.0;
  if (b < m.nblocks) {                    // uniform over the block
    const int stride = m.blk + m.max_halo;
    double* lf = lds + NROW * stride;
    const int fstride = m.max_bfaces;
    const int fb = m.bf_ptr[b], nfb = m.bf_ptr[b + 1] - fb;
    // threads [0, blk) own the block's cells; all threads share the faces
    const int i = threadIdx.x;
    const bool cell_thread = i < m.blk;
    const int c = b * m.blk + (cell_thread ? i : 0);
    const int deg = cell_thread ? m.deg[c] : 0;
    const int base = m.chunk_off[c >> 6] + (c & 63);
    const int ncp = m.ncp;
    // prefetch: this thread's faces, its cell's slot-to-face map and the
    // cell data of the stage update
    FaceRec fr[kFPF];
#pragma unroll
    for (int t = 0; t < kFPF; ++t) {
      const int e = i + t * blockDim.x;
      if (e < nfb) load_face(m, fb + e, fr[t]);
    }
    int sf[kPF];
#pragma unroll
    for (int j = 0; j < kPF; ++j)
      if (j < deg) sf[j] = m.sface[base + j * 64];
    double qn[4] = {0, 0, 0, 0}, V = 1.0, sig_old = 1.0;
    if (deg > 0) {
      if (a.stage > 1) load_q(a.qn, ncp, c, qn);
      V = m.vol[c];
      if (!a.dt_global) sig_old = a.sigma[c];
    }
    stage_block<NROW>(a, b, stride, lds, ORDER == 2);
    __syncthreads();
    // one thread per face of the block
    auto face = [&](int e, const FaceRec& f) {
      CellQG c1, c2;
      lds_cell<ORDER, NROW>(lds, stride, f.ia, c1);
      double x[4], st;
      if (f.ib >= 0) {
        lds_cell<ORDER, NROW>(lds, stride, f.ib, c2);
        face_flux<ORDER, NS>(a, c1, &c2, -1, f.nx, f.ny, f.A, f.fx, f.fy, x, st);
      } else {
        face_flux<ORDER, NS>(a, c1, nullptr, (~f.ib) & 3, f.nx, f.ny, f.A, f.fx, f.fy, x, st);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) lf[k * fstride + e] = x[k];
      if (a.stage == 3) lf[4 * fstride + e] = st;
    };
#pragma unroll
    for (int t = 0; t < kFPF; ++t) {
      const int e = i + t * blockDim.x;
      if (e < nfb) face(e, fr[t]);
    }
    for (int e = i + kFPF * blockDim.x; e < nfb; e += blockDim.x) {
      FaceRec f;
      load_face(m, fb + e, f);
      face(e, f);
    }
    __syncthreads();
    // one thread per cell: add the face terms in slot order (res.cpp:74-102)
    if (deg > 0) {
      double r[4] = {0, 0, 0, 0}, sg = 0.0;
      auto add = [&](int code) {
        const int e = code >> 1;
        const bool own = (code & 1) != 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double x = lf[k * fstride + e];
          r[k] = r[k] + (own ? x : -x);
        }
        if (a.stage == 3) sg += lf[4 * fstride + e];
      };
#pragma unroll
      for (int j = 0; j < kPF; ++j)
        if (j < deg) add(sf[j]);
      for (int j = kPF; j < deg; ++j) add(m.sface[base + j * 64]);
      double q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = lds[k * stride + i];
      // stage update (res.cpp:104-109, ssp_rk3.cpp:38-76)
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = r[k] / V;
      const double denom = a.dt_global ? __longlong_as_double((long long)*a.smax_cur) : sig_old;
      const double dt = (a.fl.cfl2 * V) / denom;
      double out[4];
      if (a.stage == 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = q[k] + dt * r[k];
      } else if (a.stage == 2) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = 0.75 * qn[k] + 0.25 * q[k] + 0.25 * dt * r[k];
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          out[k] = (1.0 / 3.0) * qn[k] + (2.0 / 3.0) * q[k] + (2.0 / 3.0) * dt * r[k];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) a.qout[k * ncp + c] = out[k];
      if (a.stage == 3) {
        a.sigma[c] = sg;
        if (a.store_res) {
#pragma unroll
          for (int k = 0; k < 4; ++k) a.res[k * ncp + c] = r[k];
        }
      }
      sig_new = sg;
    }
  }
  if (a.stage == 3 && a.dt_global) block_max_to(sig_new, a.smax_next);
  if (a.stage == 1 && blockIdx.x == 0 && threadIdx.x == 0) *a.smax_next = 0ull;
}

// Spectral sums of the first step: every face carries |V_inf| and c_inf
// (class_q.hpp:61-62), accumulated as ssp_rk3.cpp:28-36 does.
__global__ __launch_bounds__(kMaxBlock) void k_init_sigma(StageArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  double sg = 0.0;
  if (c < m.ncp) {
    const int deg = m.deg[c];
    const int base = m.chunk_off[c >> 6] + (c & 63);
    for (int j = 0; j < deg; ++j) sg += (fabs(a.fl.vn0) + a.fl.c0) * m.s_area[base + j * 64];
    a.sigma[c] = sg;
  }
  block_max_to(sg, a.smax_cur);
}

__global__ __launch_bounds__(kMaxBlock) void k_res_sq(const double* __restrict__ res, int ncp,
                                                      double* __restrict__ partial) {
  __shared__ double ws[4][kMaxBlock / 64];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double r = (c < ncp) ? res[k * ncp + c] : 0.0;
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

__global__ void k_fill(double* p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// grid size and XCD range length of a launch over `nlogical` blocks
void shape(const StageArgs& a, int nlogical, int* grid, int* per_xcd) {
  if (a.xcd_remap) {
    *per_xcd = (nlogical + 7) / 8;
    *grid = 8 * *per_xcd;
  } else {
    *per_xcd = 0;
    *grid = nlogical;
  }
}

}  // namespace

size_t lds_bytes_pass_a(const DevMesh& m) { return 6 * (size_t)(m.blk + m.max_halo) * 8; }

size_t lds_bytes_pass_b(const DevMesh& m, int order, int stage) {
  const size_t nrow = order == 2 ? 14 : 6;
  return (nrow * (size_t)(m.blk + m.max_halo) + (stage == 3 ? 5 : 4) * (size_t)m.max_bfaces) * 8;
}

hipError_t launch_pass_a(const StageArgs& a0, int squeeze, int block, hipStream_t s) {
  StageArgs a = a0;
  int grid;
  if (a.variant == kVariantLds) {
    shape(a, a.m.nblocks, &grid, &a.xcd_blocks);
    const size_t lds = lds_bytes_pass_a(a.m);
    if (squeeze) hipLaunchKernelGGL(k_pass_a_lds<true>, dim3(grid), dim3(a.m.blk), lds, s, a);
    else hipLaunchKernelGGL(k_pass_a_lds<false>, dim3(grid), dim3(a.m.blk), lds, s, a);
  } else if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL(k_pass_a_slot<true>, dim3(grid), dim3(64 * kRows), 0, s, a);
    else hipLaunchKernelGGL(k_pass_a_slot<false>, dim3(grid), dim3(64 * kRows), 0, s, a);
  } else {
    shape(a, (a.m.ncp + block - 1) / block, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL(k_pass_a_cell<true>, dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL(k_pass_a_cell<false>, dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_pass_b(const StageArgs& a0, int order, int ns, int block, hipStream_t s) {
  StageArgs a = a0;
  int grid;
  if (a.variant == kVariantLds) {
    shape(a, a.m.nblocks, &grid, &a.xcd_blocks);
    const size_t lds = lds_bytes_pass_b(a.m, order, a.stage);
    const dim3 blk(a.m.blk * kPassBThreadsPerCell);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_lds<1, false>), dim3(grid), blk, lds, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_lds<2, true>), dim3(grid), blk, lds, s, a);
    else hipLaunchKernelGGL((k_pass_b_lds<2, false>), dim3(grid), blk, lds, s, a);
  } else if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    const dim3 blk(64 * kRows);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_slot<1, false>), dim3(grid), blk, 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_slot<2, true>), dim3(grid), blk, 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_slot<2, false>), dim3(grid), blk, 0, s, a);
  } else {
    shape(a, (a.m.ncp + block - 1) / block, &grid, &a.xcd_blocks);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_cell<1, false>), dim3(grid), dim3(block), 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_cell<2, true>), dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_cell<2, false>), dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_init_sigma(const StageArgs& a, int block, hipStream_t s) {
  hipLaunchKernelGGL(k_init_sigma, dim3((a.m.ncp + block - 1) / block), dim3(block), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_res_sq(const double* res, int32_t ncp, double* partial, int nblocks, int block,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_res_sq, dim3(nblocks), dim3(block), 0, s, res, ncp, partial);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int block = 256;
  const long long grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)grid), dim3(block), 0, s, p, (long long)n, v);
  return hipGetLastError();
}

}  // namespace fvk
