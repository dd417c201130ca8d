// fv_kernels.hip — FP64 explicit FV kernels for gfx950 (MI355X).
//
// Layout (mesh_prep.hpp): cells in reverse Cuthill-McKee order, grouped in
// chunks of 64 (one wavefront); each cell's faces ("slots") in increasing
// ORIGINAL face index, stored slot-major within the chunk (slot j of lane l at
// chunk_off[chunk] + 64*j + l) so that the slot loads of a wavefront are
// contiguous.  The reference's face loops add into a cell in increasing face
// index, so accumulating a cell's slot terms in slot order reproduces its sums
// term for term.  Each interior face is evaluated by both of its cells with
// identical inputs and expressions, so both see the same bits.  Built with
// -ffp-contract=off: every expression is evaluated exactly as written, in the
// reference's association order.
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
// Two launch shapes share the per-slot device functions:
//   cell-parallel  one thread per cell walks its slots;
//   slot-parallel  one block per chunk, one thread per (slot row, cell); the
//                  slot terms go through LDS and the cell's wavefront adds
//                  them in slot order.  4x the threads and memory-level
//                  parallelism of the cell-parallel shape for the same sums.
#include "fv_kernels.hpp"

namespace fvk {
namespace {

constexpr int kMaxBlock = 256;
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
__device__ __forceinline__ void cell_prim(const double* qc, int ncp, int cell, const Flow& fl,
                                          int fs, double& rho, double& u, double& v, double& P) {
  if (fs) {
    rho = fl.rho; u = fl.u; v = fl.v; P = fl.P;
    return;
  }
  rho = qc[cell];
  const double E = qc[3 * ncp + cell];
  u = qc[ncp + cell] / rho;
  v = qc[2 * ncp + cell] / rho;
  P = (E - (0.5 * rho * (u * u + v * v))) * fl.gm1;
}

// Block-wide max of non-negative doubles, then one atomicMax on the bit
// pattern (the order of non-negative IEEE doubles is that of their bits).
__device__ __forceinline__ void block_max_to(double val, unsigned long long* dst) {
  __shared__ double wmax[kMaxBlock / 64];
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
// per-slot terms
// ---------------------------------------------------------------------------

// Pass A, first loop: one face's LSQ right-hand-side terms (Qj - Qi)*dxj,
// (Qj - Qi)*dyj with i = face_cell1 (cell_gradient.cpp:30-91), and the
// face's min/max candidates for this cell (squeeze.cpp:17-62).
__device__ __forceinline__ void slot_lsq(const StageArgs& a, int c, int s, const double q[4],
                                         double xc, double yc, double px[4], double py[4],
                                         double tmin[4], double tmax[4]) {
  const DevMesh& m = a.m;
  const int ncp = m.ncp;
  const double* __restrict__ qc = a.qc;
  const int code = m.code[s];
  double d[4], dxj, dyj;
  if (code >= 0) {
    const int nb = code >> 1;
    double qn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) qn[k] = qc[k * ncp + nb];
    const double xn = m.ccx[nb], yn = m.ccy[nb];
    if (code & 1) {  // this cell is face_cell1
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        d[k] = qn[k] - q[k];
        tmin[k] = min2(q[k], qn[k]);
        tmax[k] = max2(q[k], qn[k]);
      }
      dxj = xn - xc;
      dyj = yn - yc;
    } else {         // this cell is face_cell2: same face terms, owner-relative
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        d[k] = q[k] - qn[k];
        tmin[k] = min2(qn[k], q[k]);
        tmax[k] = max2(qn[k], q[k]);
      }
      dxj = xc - xn;
      dyj = yc - yn;
    }
  } else {
    const int v = ~code;
    const int kind = v & 3;
    const double nx = m.s_nx[s], ny = m.s_ny[s];
    const double rx = m.s_fcx[s] - xc, ry = m.s_fcy[s] - yc;
    const double t = 2.0 * (rx * nx + ry * ny);       // ghost centroid, readmesh.cpp:238-243
    const double gx = xc + t * nx, gy = yc + t * ny;
    dxj = gx - xc;
    dyj = gy - yc;
    const double gm1 = a.fl.gm1;
    if (kind == 0) {            // FREESTREAM
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = a.fl.qinf[k] - q[k];
    } else if (kind == 1) {     // SLIP_WALL: reflect the cell velocity
      double rho, u, vv, P;
      cell_prim(qc, ncp, c, a.fl, a.fs_prims, rho, u, vv, P);
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
      const int src = v >> 2;
      double rho, u, vv, P;
      cell_prim(qc, ncp, src, a.fl, a.fs_prims, rho, u, vv, P);
      double w[4];
      ghost_q(gm1, rho, -u, -vv, P, w);
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = w[k] - q[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tmin[k] = q[k];
      tmax[k] = q[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    px[k] = d[k] * dxj;
    py[k] = d[k] * dyj;
  }
}

// Pass A, second loop: the squeeze factor of one face for this cell from the
// face state the previous res() call left on this cell's side
// (squeeze.cpp:68-121): Q_prev + grad_prev . (x_f - x_i), or -101 at first.
__device__ __forceinline__ void slot_phi(const StageArgs& a, int s, double xc, double yc,
                                         const double q[4], const double qmin[4],
                                         const double qmax[4], const double qp[4],
                                         const double gpx[4], const double gpy[4], double ph[4]) {
  const double dx = a.m.s_fcx[s] - xc, dy = a.m.s_fcy[s] - yc;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double qf = a.sentinel ? -101.0 : qp[k] + gpx[k] * dx + gpy[k] * dy;
    double p = 1.0;
    if (qf < qmin[k]) p = (qmin[k] - q[k]) / (qf - q[k]);
    else if (qf > qmax[k]) p = (qmax[k] - q[k]) / (qf - q[k]);
    ph[k] = p;
  }
}

// Pass B: one face's contribution to this cell's residual, already signed
// (+ for face_cell1, - for face_cell2, res.cpp:81-101; a - b == a + (-b)),
// and its spectral-sum term (|Vn_avg| + c_avg) * A (ssp_rk3.cpp:32-34).
template <int ORDER, bool NS>
__device__ __forceinline__ void slot_flux(const StageArgs& a, int c, int s, const double q[4],
                                          const double gx[4], const double gy[4], double xc,
                                          double yc, double xo[4], double& sterm) {
  const DevMesh& m = a.m;
  const int ncp = m.ncp;
  const Flow& fl = a.fl;
  const double gamma = fl.gamma, gm1 = fl.gm1;
  const double* __restrict__ qc = a.qc;
  const double* __restrict__ g = a.g;
  const int code = m.code[s];
  const double nx = m.s_nx[s], ny = m.s_ny[s], A = m.s_area[s];
  double fx = 0, fy = 0;
  double o[4];  // this cell's side of the face
  if (ORDER == 2) {
    fx = m.s_fcx[s];
    fy = m.s_fcy[s];
    const double dx = fx - xc, dy = fy - yc;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = q[k] + gx[k] * dx + gy[k] * dy;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = q[k];
  }
  double L[4], R[4];
  double qn[4], gnx[4] = {0, 0, 0, 0}, gny[4] = {0, 0, 0, 0}, xn = 0, yn = 0;
  bool own = true;
  int kind = -1;
  if (code >= 0) {
    const int nb = code >> 1;
    own = (code & 1) != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) qn[k] = qc[k * ncp + nb];
    double t[4];
    if (ORDER == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gnx[k] = g[k * ncp + nb];
        gny[k] = g[(4 + k) * ncp + nb];
      }
      xn = m.ccx[nb];
      yn = m.ccy[nb];
      const double dx = fx - xn, dy = fy - yn;
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = qn[k] + gnx[k] * dx + gny[k] * dy;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = qn[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      L[k] = own ? o[k] : t[k];
      R[k] = own ? t[k] : o[k];
    }
  } else {
    kind = (~code) & 3;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      L[k] = o[k];
      qn[k] = 0;
    }
  }
  Prim pl;
  to_prim(L, gamma, gm1, pl);
  if (kind == 0) {          // FREESTREAM
#pragma unroll
    for (int k = 0; k < 4; ++k) R[k] = fl.qinf[k];
  } else if (kind == 1) {   // SLIP_WALL: reflect the reconstructed velocity
    const double Vn = pl.u * nx + pl.v * ny;
    const double wu = pl.u - 2.0 * Vn * nx;
    const double wv = pl.v - 2.0 * Vn * ny;
    ghost_q(gm1, pl.rho, wu, wv, pl.P, R);
  } else if (kind == 2) {   // EXTRAPOLATION
#pragma unroll
    for (int k = 0; k < 4; ++k) R[k] = L[k];
  } else if (kind == 3) {   // NO_SLIP_WALL
    ghost_q(gm1, pl.rho, -pl.u, -pl.v, pl.P, R);
  }
  Prim pr;
  to_prim(R, gamma, gm1, pr);
  // Rusanov flux
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
  if (NS && ORDER == 2 && (kind == -1 || kind == 3)) {
    // face-averaged state, reconstruction.cpp:86-92
    double fa[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) fa[k] = 0.5 * (L[k] + R[k]);
    const double frho = fa[0], fE = fa[3];
    const double fu = fa[1] / frho, fv = fa[2] / frho;
    const double fP = (fE - (0.5 * frho * (fu * fu + fv * fv))) * gm1;
    double ux, uy, vx, vy, Px = 0, Py = 0, rhox = 0, rhoy = 0;
    if (kind == 3) {        // no-slip wall, reconstruction.cpp:110-123
      double crho, cu, cv, cP;
      cell_prim(qc, ncp, c, fl, a.fs_prims, crho, cu, cv, cP);
      const double dx1 = fx - xc, dy1 = fy - yc;
      const double dn = fabs(sqrt(dx1 * dx1 + dy1 * dy1));
      const double dudn = -cu / dn;
      const double dvdn = -cv / dn;
      ux = dudn * nx;
      uy = dudn * ny;
      vx = dvdn * nx;
      vy = dvdn * ny;
    } else {                // interior, reconstruction.cpp:124-161
      const double* q1 = own ? q : qn;
      const double* q2 = own ? qn : q;
      const double* g1x = own ? gx : gnx;
      const double* g1y = own ? gy : gny;
      const double* g2x = own ? gnx : gx;
      const double* g2y = own ? gny : gy;
      const double x1 = own ? xc : xn, y1 = own ? yc : yn;
      const double x2 = own ? xn : xc, y2 = own ? yn : yc;
      const double tx = x2 - x1, ty = y2 - y1;
      const double dist = sqrt(tx * tx + ty * ty);   // face_dist, and the norm of l
      const double lx = tx / dist, ly = ty / dist;
      const double mx = -ny, my = nx;                 // unit p1->p2 (readmesh.cpp:400-406)
      double ddx[4], ddy[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dl = (q2[k] - q1[k]) / fabs(dist);
        const double dm = 0.5 * (g1x[k] * mx + g1y[k] * my + g2x[k] * mx + g2y[k] * my);
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
    const double x = (Fv[k] * A) - (F * A);
    xo[k] = own ? x : -x;
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

__device__ __forceinline__ void write_grad(const StageArgs& a, int c, const double gx[4],
                                           const double gy[4]) {
  const int ncp = a.m.ncp;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a.g[k * ncp + c] = gx[k];
    a.g[(4 + k) * ncp + c] = gy[k];
  }
}

// ---------------------------------------------------------------------------
// cell-parallel shape
// ---------------------------------------------------------------------------
template <bool SQUEEZE>
__global__ __launch_bounds__(kMaxBlock) void k_pass_a_cell(StageArgs a) {
  const int c = logical_block(a.xcd_blocks) * blockDim.x + threadIdx.x;
  const DevMesh& m = a.m;
  if (c >= m.nc) return;
  const int ncp = m.ncp;
  double q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = a.qc[k * ncp + c];
  const double xc = m.ccx[c], yc = m.ccy[c];
  const int deg = m.deg[c];
  const int base = m.chunk_off[c >> 6] + (c & 63);
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qmin[k] = 100000000000.0;
    qmax[k] = -100000000000.0;
  }
  for (int j = 0; j < deg; ++j) {
    double px[4], py[4], tmin[4], tmax[4];
    slot_lsq(a, c, base + j * 64, q, xc, yc, px, py, tmin, tmax);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sx[k] += px[k];
      sy[k] += py[k];
      qmin[k] = min2(qmin[k], tmin[k]);
      qmax[k] = max2(qmax[k], tmax[k]);
    }
  }
  const double ixx = m.ixx[c], iyy = m.iyy[c], ixy = m.ixy[c];
  double gx[4], gy[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    gx[k] = (sx[k] * iyy - sy[k] * ixy);
    gy[k] = (sy[k] * ixx - sx[k] * ixy);
  }
  if (SQUEEZE) {
    double phi[4] = {1, 1, 1, 1};
    double qp[4] = {0, 0, 0, 0}, gpx[4] = {0, 0, 0, 0}, gpy[4] = {0, 0, 0, 0};
    if (!a.sentinel) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        qp[k] = a.qp[k * ncp + c];
        gpx[k] = a.gp[k * ncp + c];
        gpy[k] = a.gp[(4 + k) * ncp + c];
      }
    }
    for (int j = 0; j < deg; ++j) {
      double ph[4];
      slot_phi(a, base + j * 64, xc, yc, q, qmin, qmax, qp, gpx, gpy, ph);
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
  const int ncp = m.ncp;
  double sig_new = 0.0;
  if (c < m.nc) {
    double q[4], gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = a.qc[k * ncp + c];
    double xc = 0, yc = 0;
    if (ORDER == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gx[k] = a.g[k * ncp + c];
        gy[k] = a.g[(4 + k) * ncp + c];
      }
      xc = m.ccx[c];
      yc = m.ccy[c];
    }
    const int deg = m.deg[c];
    const int base = m.chunk_off[c >> 6] + (c & 63);
    double r[4] = {0, 0, 0, 0}, sg = 0.0;
    for (int j = 0; j < deg; ++j) {
      double x[4], st;
      slot_flux<ORDER, NS>(a, c, base + j * 64, q, gx, gy, xc, yc, x, st);
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = r[k] + x[k];
      sg += st;
    }
    cell_update(a, c, q, r, sg);
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
  const int ncp = m.ncp;
  const int chunk = logical_block(a.xcd_blocks);
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int c = chunk * 64 + lane;
  if (chunk * 64 >= m.nc) return;          // whole block idle (uniform)
  const bool active = c < m.nc;
  const int deg = active ? m.deg[c] : 0;
  const int cbase = m.chunk_off[chunk];
  const int nrows = (m.chunk_off[chunk + 1] - cbase) >> 6;
  const int base = cbase + lane;
  double q[4] = {0, 0, 0, 0}, xc = 0, yc = 0;
  if (active) {
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = a.qc[k * ncp + c];
    xc = m.ccx[c];
    yc = m.ccy[c];
  }
  double sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, qmin[4], qmax[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qmin[k] = 100000000000.0;
    qmax[k] = -100000000000.0;
  }
  for (int j0 = 0; j0 < nrows; j0 += kRows) {
    const int j = j0 + row;
    if (j < deg) {
      double px[4], py[4], tmin[4], tmax[4];
      slot_lsq(a, c, base + j * 64, q, xc, yc, px, py, tmin, tmax);
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
  if (row == 0 && active) {
    const double ixx = m.ixx[c], iyy = m.iyy[c], ixy = m.ixy[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gx[k] = (sx[k] * iyy - sy[k] * ixy);
      gy[k] = (sy[k] * ixx - sx[k] * ixy);
    }
  }
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
    if (!a.sentinel && active) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        qp[k] = a.qp[k * ncp + c];
        gpx[k] = a.gp[k * ncp + c];
        gpy[k] = a.gp[(4 + k) * ncp + c];
      }
    }
    double phi[4] = {1, 1, 1, 1};
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        double ph[4];
        slot_phi(a, base + j * 64, xc, yc, q, qmin, qmax, qp, gpx, gpy, ph);
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
  const int ncp = m.ncp;
  const int chunk = logical_block(a.xcd_blocks);
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int c = chunk * 64 + lane;
  double sig_new = 0.0;
  if (chunk * 64 < m.nc) {                 // uniform over the block
    const bool active = c < m.nc;
    const int deg = active ? m.deg[c] : 0;
    const int cbase = m.chunk_off[chunk];
    const int nrows = (m.chunk_off[chunk + 1] - cbase) >> 6;
    const int base = cbase + lane;
    double q[4] = {0, 0, 0, 0}, gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0}, xc = 0, yc = 0;
    if (active) {
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = a.qc[k * ncp + c];
      if (ORDER == 2) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          gx[k] = a.g[k * ncp + c];
          gy[k] = a.g[(4 + k) * ncp + c];
        }
        xc = m.ccx[c];
        yc = m.ccy[c];
      }
    }
    double r[4] = {0, 0, 0, 0}, sg = 0.0;
    for (int j0 = 0; j0 < nrows; j0 += kRows) {
      const int j = j0 + row;
      if (j < deg) {
        double x[4], st;
        slot_flux<ORDER, NS>(a, c, base + j * 64, q, gx, gy, xc, yc, x, st);
#pragma unroll
        for (int k = 0; k < 4; ++k) t1[row][k][lane] = x[k];
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
      cell_update(a, c, q, r, sg);
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
  if (c < m.nc) {
    const int deg = m.deg[c];
    const int base = m.chunk_off[c >> 6] + (c & 63);
    for (int j = 0; j < deg; ++j) sg += (fabs(a.fl.vn0) + a.fl.c0) * m.s_area[base + j * 64];
    a.sigma[c] = sg;
  }
  block_max_to(sg, a.smax_cur);
}

__global__ __launch_bounds__(kMaxBlock) void k_res_sq(const double* __restrict__ res, int nc,
                                                      int ncp, double* __restrict__ partial) {
  __shared__ double ws[4][kMaxBlock / 64];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double r = (c < nc) ? res[k * ncp + c] : 0.0;
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

hipError_t launch_pass_a(const StageArgs& a0, int squeeze, int block, hipStream_t s) {
  StageArgs a = a0;
  int grid;
  if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL(k_pass_a_slot<true>, dim3(grid), dim3(64 * kRows), 0, s, a);
    else hipLaunchKernelGGL(k_pass_a_slot<false>, dim3(grid), dim3(64 * kRows), 0, s, a);
  } else {
    shape(a, (a.m.nc + block - 1) / block, &grid, &a.xcd_blocks);
    if (squeeze) hipLaunchKernelGGL(k_pass_a_cell<true>, dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL(k_pass_a_cell<false>, dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_pass_b(const StageArgs& a0, int order, int ns, int block, hipStream_t s) {
  StageArgs a = a0;
  int grid;
  if (a.variant == kVariantSlot) {
    shape(a, a.m.ncp / 64, &grid, &a.xcd_blocks);
    const dim3 blk(64 * kRows);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_slot<1, false>), dim3(grid), blk, 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_slot<2, true>), dim3(grid), blk, 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_slot<2, false>), dim3(grid), blk, 0, s, a);
  } else {
    shape(a, (a.m.nc + block - 1) / block, &grid, &a.xcd_blocks);
    if (order == 1) hipLaunchKernelGGL((k_pass_b_cell<1, false>), dim3(grid), dim3(block), 0, s, a);
    else if (ns) hipLaunchKernelGGL((k_pass_b_cell<2, true>), dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL((k_pass_b_cell<2, false>), dim3(grid), dim3(block), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_init_sigma(const StageArgs& a, int block, hipStream_t s) {
  hipLaunchKernelGGL(k_init_sigma, dim3((a.m.nc + block - 1) / block), dim3(block), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_res_sq(const double* res, int32_t nc, int32_t ncp, double* partial, int nblocks,
                         int block, hipStream_t s) {
  hipLaunchKernelGGL(k_res_sq, dim3(nblocks), dim3(block), 0, s, res, nc, ncp, partial);
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
