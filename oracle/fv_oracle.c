/* fv_oracle.c — CPU restatement of the reference FV solver. TEST INFRASTRUCTURE:
 * only the tests, the smoke check and the CPU-baseline timing load this file's
 * library; the product never does.  Every function names the reference lines it
 * restates.  Expressions keep the reference's evaluation order (C evaluates
 * a*b + c*d - e left to right exactly as C++ does) and the library is built
 * with -ffp-contract=off, so results are bit-identical to the reference's
 * x86-64 build, which has no FMA contraction either.
 */
#include "fv_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* --- one set of conserved/primitive arrays: class_Q (class_q.hpp:7-48) --- */
typedef struct {
  int n;
  double gamma, R, mu, k;
  double *p[4];
  double *rho, *u, *v, *E, *P, *c, *T;
  double *vn_avg, *c_avg;
  double *qx[4], *qy[4];
  double *txx, *txy, *tyy, *dtdx, *dtdy;
} qset;

struct or_case {
  int np, nf, nc, nb;
  or_options o;
  /* class_flow (class_flow.hpp:3-47) */
  double fQ[4], fE, fCp, fT, fk, flowx, flowy;
  /* class_mesh (class_mesh.hpp:5-68) */
  double *x, *y;
  int *fp1, *fp2, *fc1, *fc2;   /* fc2 after BC-kind substitution */
  double *area, *nx, *ny, *fcx, *fcy;
  double *vol, *ccx, *ccy;
  int *cf_ptr, *cf_idx;          /* cell -> faces, increasing face index */
  int *bc_faces, *face2bc;
  double *bccx, *bccy;
  double *c1dx, *c1dy, *c2dx, *c2dy, *dxj, *dyj;
  double *delta, *Ixx, *Iyy, *Ixy;
  double *fdist, *lx, *ly, *mx, *my;
  int num_bc1, num_bc2, num_wall;
  /* solver state */
  qset Qbar, Qface, Qc1, Qc2;
  double *res[4];
  /* scratch */
  double *dt;
  double time;                     /* physical time of the uniform dt (extension) */
  double fixed_dt;                 /* > 0: every step takes this dt (test hook) */
};

static char g_err[512];
const char* or_error(void) { return g_err; }

static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
static double* dfill(int n, double v) {
  double* a = (double*)xcalloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; ++i) a[i] = v;
  return a;
}

static void qset_alloc(qset* q, int n) {
  memset(q, 0, sizeof(*q));
  q->n = n;
  for (int k = 0; k < 4; ++k) {
    q->p[k] = dfill(n, 0); q->qx[k] = dfill(n, 0); q->qy[k] = dfill(n, 0);
  }
  q->rho = dfill(n, 0); q->u = dfill(n, 0); q->v = dfill(n, 0); q->E = dfill(n, 0);
  q->P = dfill(n, 0); q->c = dfill(n, 0); q->T = dfill(n, 0);
  q->vn_avg = dfill(n, 0); q->c_avg = dfill(n, 0);
  q->txx = dfill(n, 0); q->txy = dfill(n, 0); q->tyy = dfill(n, 0);
  q->dtdx = dfill(n, 0); q->dtdy = dfill(n, 0);
}
static void qset_free(qset* q) {
  for (int k = 0; k < 4; ++k) { free(q->p[k]); free(q->qx[k]); free(q->qy[k]); }
  free(q->rho); free(q->u); free(q->v); free(q->E); free(q->P); free(q->c); free(q->T);
  free(q->vn_avg); free(q->c_avg);
  free(q->txx); free(q->txy); free(q->tyy); free(q->dtdx); free(q->dtdy);
}

/* class_Q::init(size, class_flow*)  class_q.hpp:49-94 */
static void qset_init_flow(qset* q, const or_case* C) {
  const or_options* f = &C->o;
  q->gamma = f->gamma; q->R = f->R; q->mu = f->mu; q->k = C->fk;
  for (int i = 0; i < q->n; ++i) {
    q->P[i] = f->P; q->rho[i] = f->rho; q->u[i] = f->u; q->v[i] = f->v;
    q->vn_avg[i] = sqrt(f->v * f->v + f->u * f->u);
    q->c_avg[i] = sqrt(q->gamma * (f->P / f->rho));
    q->E[i] = (q->P[i] / (q->gamma - 1)) + 0.5 * q->rho[i] * (q->u[i] * q->u[i] + q->v[i] * q->v[i]);
    q->c[i] = sqrt(q->gamma * (q->P[i] / q->rho[i]));
    q->T[i] = q->P[i] / (q->rho[i] * q->R);
    for (int k = 0; k < 4; ++k) { q->p[k][i] = -101; q->qx[k][i] = 0; q->qy[k][i] = 0; }
    q->txx[i] = 0; q->txy[i] = 0; q->tyy[i] = 0; q->dtdx[i] = 0; q->dtdy[i] = 0;
  }
}

/* class_Q::init(size, gamma, R)  class_q.hpp:96-129 */
static void qset_init_plain(qset* q, double gamma, double R) {
  q->gamma = gamma; q->R = R;
  for (int i = 0; i < q->n; ++i) {
    q->P[i] = -101; q->rho[i] = -101; q->u[i] = -101; q->v[i] = -101;
    q->E[i] = -101; q->c[i] = -101; q->T[i] = -101;
    q->vn_avg[i] = -101; q->c_avg[i] = -101;
    for (int k = 0; k < 4; ++k) { q->p[k][i] = -101; q->qx[k][i] = 0; q->qy[k][i] = 0; }
  }
}

/* class_Q::updateflow_int  class_q.hpp:143-151 */
static void updateflow_int(qset* q, int i) {
  q->rho[i] = q->p[0][i];
  q->E[i] = q->p[3][i];
  q->u[i] = q->p[1][i] / q->rho[i];
  q->v[i] = q->p[2][i] / q->rho[i];
  q->P[i] = (q->E[i] - (0.5 * q->rho[i] * (q->u[i] * q->u[i] + q->v[i] * q->v[i]))) * (q->gamma - 1);
  q->c[i] = sqrt(q->gamma * (q->P[i] / q->rho[i]));
  q->T[i] = q->P[i] / (q->rho[i] * q->R);
}
/* class_Q::updateflow  class_q.hpp:131-141 */
static void updateflow(qset* q) {
  for (int i = 0; i < q->n; ++i) updateflow_int(q, i);
}
/* class_Q::updateQ  class_q.hpp:153-161 */
static void updateQ(qset* q) {
  for (int i = 0; i < q->n; ++i) {
    q->p[0][i] = q->rho[i];
    q->p[1][i] = q->rho[i] * q->u[i];
    q->p[2][i] = q->rho[i] * q->v[i];
    q->E[i] = (q->P[i] / (q->gamma - 1)) + 0.5 * q->rho[i] * (q->u[i] * q->u[i] + q->v[i] * q->v[i]);
    q->p[3][i] = q->E[i];
  }
}
/* class_Q::comp_vis  class_q.hpp:163-169 */
static void comp_vis(qset* q, int i, double ux, double uy, double vx, double vy, double Px,
                     double Py, double rhox, double rhoy) {
  q->txx[i] = 2 * q->mu * (ux - (ux + vy) / 3);
  q->txy[i] = q->mu * (uy + vx);
  q->tyy[i] = 2 * q->mu * (vy - (ux + vy) / 3);
  q->dtdx[i] = (Px * q->rho[i] - rhox * q->P[i]) / (q->R * q->rho[i] * q->rho[i]);
  q->dtdy[i] = (Py * q->rho[i] - rhoy * q->P[i]) / (q->R * q->rho[i] * q->rho[i]);
}

/* class_flow::updateQ for a ghost state  class_flow.hpp:28-34 */
static void ghost_Q(double gamma, double rho, double u, double v, double P, double Q[4]) {
  Q[0] = rho;
  Q[1] = rho * u;
  Q[2] = rho * v;
  double E = (P / (gamma - 1)) + 0.5 * rho * (u * u + v * v);
  Q[3] = E;
}

static int bc_kind(const or_options* o, int file_c2) {
  if (file_c2 == -1) return o->bc1;   /* readmesh.cpp:61-63 */
  if (file_c2 == -2) return o->bc2;   /* readmesh.cpp:64-66 */
  return file_c2;                     /* readmesh.cpp:67-69: raw -3/-4 pass through */
}

/* ------------------------------------------------------------------ */
/* geometry: readmesh.cpp:113-411                                      */
/* ------------------------------------------------------------------ */
static int build_geometry(or_case* C) {
  int nf = C->nf, nc = C->nc;
  /* face normals, areas, centres  readmesh.cpp:117-139 */
  for (int f = 0; f < nf; ++f) {
    double x1 = C->x[C->fp1[f]], y1 = C->y[C->fp1[f]];
    double x2 = C->x[C->fp2[f]], y2 = C->y[C->fp2[f]];
    double dx = x2 - x1, dy = y2 - y1;
    double area = sqrt(dx * dx + dy * dy);
    C->area[f] = area;
    C->nx[f] = dy / area;
    C->ny[f] = -dx / area;
    C->fcx[f] = dx / 2 + x1;
    C->fcy[f] = dy / 2 + y1;
  }
  /* centroids and volumes by fan triangles about the vertex mean,
   * readmesh.cpp:147-228.  The unique-vertex search compares against every
   * one of the vertex_count zero-initialised slots (readmesh.cpp:149-150,163-184). */
  int maxv = 0;
  for (int c = 0; c < nc; ++c) {
    int n = C->cf_ptr[c + 1] - C->cf_ptr[c];
    if (n > maxv) maxv = n;
  }
  double* px = (double*)xcalloc((size_t)maxv + 1, sizeof(double));
  double* py = (double*)xcalloc((size_t)maxv + 1, sizeof(double));
  for (int c = 0; c < nc; ++c) {
    int b = C->cf_ptr[c], n = C->cf_ptr[c + 1] - b, found = 0;
    for (int z = 0; z < n; ++z) { px[z] = 0; py[z] = 0; }
    for (int j = 0; j < n; ++j) {
      int f = C->cf_idx[b + j];
      double t1x = C->x[C->fp1[f]], t1y = C->y[C->fp1[f]];
      double t2x = C->x[C->fp2[f]], t2y = C->y[C->fp2[f]];
      int hit = 0;
      for (int z = 0; z < n; ++z)
        if (px[z] == t1x && py[z] == t1y) { hit = 1; break; }
      if (!hit && found < n) { px[found] = t1x; py[found] = t1y; found++; }
      hit = 0;
      for (int z = 0; z < n; ++z)
        if (px[z] == t2x && py[z] == t2y) { hit = 1; break; }
      if (!hit && found < n) { px[found] = t2x; py[found] = t2y; found++; }
    }
    double sx = 0, sy = 0;
    for (int j = 0; j < n; ++j) { sx += px[j]; sy += py[j]; }
    double cx = sx / n, cy = sy / n;
    double sum_vol = 0, scx = 0, scy = 0;
    for (int j = 0; j < n; ++j) {
      int f = C->cf_idx[b + j];
      double x1 = C->x[C->fp1[f]], y1 = C->y[C->fp1[f]];
      double x2 = C->x[C->fp2[f]], y2 = C->y[C->fp2[f]];
      double x3 = cx, y3 = cy;
      double vs = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
      sum_vol += vs;
      double csx = (x1 + x2 + x3) / 3, csy = (y1 + y2 + y3) / 3;
      scx += csx * vs;
      scy += csy * vs;
    }
    C->vol[c] = sum_vol;
    C->ccx[c] = scx / sum_vol;
    C->ccy[c] = scy / sum_vol;
  }
  free(px); free(py);
  /* ghost centroids: mirror of the owner centroid about the face,
   * readmesh.cpp:232-244 */
  for (int i = 0; i < C->nb; ++i) {
    int f = C->bc_faces[i];
    int c1 = C->fc1[f];
    double rx = C->fcx[f] - C->ccx[c1];
    double ry = C->fcy[f] - C->ccy[c1];
    double nt = 2.0 * (rx * C->nx[f] + ry * C->ny[f]);
    C->bccx[i] = C->ccx[c1] + nt * C->nx[f];
    C->bccy[i] = C->ccy[c1] + nt * C->ny[f];
  }
  /* face-centre offsets  readmesh.cpp:247-265 */
  for (int f = 0; f < nf; ++f) {
    C->c1dx[f] = C->fcx[f] - C->ccx[C->fc1[f]];
    C->c1dy[f] = C->fcy[f] - C->ccy[C->fc1[f]];
    if (C->fc2[f] >= 0) {
      C->c2dx[f] = C->fcx[f] - C->ccx[C->fc2[f]];
      C->c2dy[f] = C->fcy[f] - C->ccy[C->fc2[f]];
    } else {
      int b = C->face2bc[f];
      C->c2dx[f] = C->fcx[f] - C->bccx[b];
      C->c2dy[f] = C->fcy[f] - C->bccy[b];
    }
  }
  /* LSQ normal matrix incl. ghosts  readmesh.cpp:268-320 */
  double* d1 = dfill(nc, 0);
  double* d2 = dfill(nc, 0);
  double* d3 = dfill(nc, 0);
  for (int f = 0; f < nf; ++f) {
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double xi = C->ccx[c1], yi = C->ccy[c1], xj, yj;
    if (c2 >= 0) { xj = C->ccx[c2]; yj = C->ccy[c2]; }
    else { int b = C->face2bc[f]; xj = C->bccx[b]; yj = C->bccy[b]; }
    d1[c1] += (xi - xj) * (xi - xj);
    d2[c1] += (yi - yj) * (yi - yj);
    d3[c1] += ((xi - xj) * (yi - yj));
    if (c2 >= 0) {
      d1[c2] += (xj - xi) * (xj - xi);
      d2[c2] += (yj - yi) * (yj - yi);
      d3[c2] += ((xj - xi) * (yj - yi));
    }
    C->dxj[f] = xj - xi;
    C->dyj[f] = yj - yi;
  }
  for (int c = 0; c < nc; ++c) C->delta[c] = d1[c] * d2[c] - (d3[c] * d3[c]);
  for (int c = 0; c < nc; ++c) {
    C->Ixx[c] = d1[c] / C->delta[c];
    C->Iyy[c] = d2[c] / C->delta[c];
    C->Ixy[c] = d3[c] / C->delta[c];
  }
  free(d1); free(d2); free(d3);
  /* wall count  readmesh.cpp:356-360 (0 when neither group is a wall) */
  const or_options* o = &C->o;
  C->num_wall = 0;
  if (o->bc1 == -2 || o->bc1 == -4) C->num_wall = C->num_bc1;
  else if (o->bc2 == -2 || o->bc2 == -4) C->num_wall = C->num_bc2;
  /* NS metrics  readmesh.cpp:363-411.  On boundary faces the reference looks
   * the ghost up as BC_cell_center[BC_faces[idx]] (readmesh.cpp:373,394), which
   * is this face's own ghost when boundary faces come first; those values are
   * never read by the solver (reconstruction.cpp:100-124 uses only the interior
   * ones), so the face's own ghost is used here. */
  for (int f = 0; f < nf; ++f) {
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double tx, ty;
    if (c2 >= 0) { tx = C->ccx[c2] - C->ccx[c1]; ty = C->ccy[c2] - C->ccy[c1]; }
    else { int b = C->face2bc[f]; tx = C->bccx[b] - C->ccx[c1]; ty = C->bccy[b] - C->ccy[c1]; }
    C->fdist[f] = sqrt(tx * tx + ty * ty);
    C->lx[f] = tx;
    C->ly[f] = ty;
    C->mx[f] = (C->x[C->fp2[f]] - C->x[C->fp1[f]]);
    C->my[f] = (C->y[C->fp2[f]] - C->y[C->fp1[f]]);
    double t = sqrt(C->mx[f] * C->mx[f] + C->my[f] * C->my[f]);
    C->mx[f] = C->mx[f] / t;
    C->my[f] = C->my[f] / t;
    double t2 = sqrt(C->lx[f] * C->lx[f] + C->ly[f] * C->ly[f]);
    C->lx[f] = C->lx[f] / t2;
    C->ly[f] = C->ly[f] / t2;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* solver                                                              */
/* ------------------------------------------------------------------ */

/* compute_gradient  cell_gradient.cpp:9-130 */
static void compute_gradient(or_case* C, qset* Q) {
  int nf = C->nf, nc = C->nc;
  double g = C->o.gamma;
  double* sx[4];
  double* sy[4];
  double* fsx[4];
  double* fsy[4];
  for (int k = 0; k < 4; ++k) {
    sx[k] = dfill(nc, 0); sy[k] = dfill(nc, 0);
    fsx[k] = dfill(nf, 0); fsy[k] = dfill(nf, 0);
  }
  for (int f = 0; f < nf; ++f) {
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double d[4], W[4];
    switch (c2) {
      case -1: /* freestream  :37-42 */
        for (int k = 0; k < 4; ++k) d[k] = C->fQ[k] - Q->p[k][c1];
        break;
      case -2: { /* slip wall: reflect the cell velocity  :43-55 */
        double Vn = Q->u[c1] * C->nx[f] + Q->v[c1] * C->ny[f];
        double wu = Q->u[c1] - 2.0 * Vn * C->nx[f];
        double wv = Q->v[c1] - 2.0 * Vn * C->ny[f];
        ghost_Q(g, Q->rho[c1], wu, wv, Q->P[c1], W);
        for (int k = 0; k < 4; ++k) d[k] = W[k] - Q->p[k][c1];
        break;
      }
      case -3: /* extrapolation  :56-61 */
        for (int k = 0; k < 4; ++k) d[k] = 0;
        break;
      case -4: /* no-slip: state of the cell whose index equals the face index  :62-73 */
        ghost_Q(g, Q->rho[f], -Q->u[f], -Q->v[f], Q->P[f], W);
        for (int k = 0; k < 4; ++k) d[k] = W[k] - Q->p[k][c1];
        break;
      default: /* :74-79 */
        for (int k = 0; k < 4; ++k) d[k] = Q->p[k][c2] - Q->p[k][c1];
        break;
    }
    for (int k = 0; k < 4; ++k) {
      fsx[k][f] = d[k] * C->dxj[f];
      fsy[k][f] = d[k] * C->dyj[f];
    }
  }
  for (int f = 0; f < nf; ++f) {   /* :93-118 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    for (int k = 0; k < 4; ++k) { sx[k][c1] += fsx[k][f]; sy[k][c1] += fsy[k][f]; }
    if (c2 >= 0)
      for (int k = 0; k < 4; ++k) { sx[k][c2] += fsx[k][f]; sy[k][c2] += fsy[k][f]; }
  }
  for (int c = 0; c < nc; ++c) {   /* :120-130 */
    for (int k = 0; k < 4; ++k) {
      Q->qx[k][c] = (sx[k][c] * C->Iyy[c] - sy[k][c] * C->Ixy[c]);
      Q->qy[k][c] = (sy[k][c] * C->Ixx[c] - sx[k][c] * C->Ixy[c]);
    }
  }
  for (int k = 0; k < 4; ++k) { free(sx[k]); free(sy[k]); free(fsx[k]); free(fsy[k]); }
}

/* std::min({a, b}) / std::max({a, b}): the first of equal elements wins */
static double min2(double a, double b) { return (b < a) ? b : a; }
static double max2(double a, double b) { return (a < b) ? b : a; }

/* squeeze  squeeze.cpp:9-154; the face states are the previous res() call's */
static void squeeze(or_case* C, qset* Q) {
  int nf = C->nf, nc = C->nc;
  double *qmin[4], *qmax[4], *phc1[4], *phc2[4], *ph[4];
  for (int k = 0; k < 4; ++k) {
    qmin[k] = dfill(nc, 100000000000.0);
    qmax[k] = dfill(nc, -100000000000.0);
    phc1[k] = dfill(nf, 1);
    phc2[k] = dfill(nf, 1);
    ph[k] = dfill(nc, 1);
  }
  for (int f = 0; f < nf; ++f) {   /* :17-62 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double tmin[4], tmax[4];
    if (c2 >= 0) {
      for (int k = 0; k < 4; ++k) {
        tmin[k] = min2(Q->p[k][c1], Q->p[k][c2]);
        tmax[k] = max2(Q->p[k][c1], Q->p[k][c2]);
      }
      for (int k = 0; k < 4; ++k) {
        qmin[k][c2] = min2(qmin[k][c2], tmin[k]);
        qmax[k][c2] = max2(qmax[k][c2], tmax[k]);
      }
    } else {
      for (int k = 0; k < 4; ++k) { tmin[k] = Q->p[k][c1]; tmax[k] = Q->p[k][c1]; }
    }
    for (int k = 0; k < 4; ++k) {
      qmin[k][c1] = min2(qmin[k][c1], tmin[k]);
      qmax[k][c1] = max2(qmax[k][c1], tmax[k]);
    }
  }
  for (int f = 0; f < nf; ++f) {   /* :68-121 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    for (int k = 0; k < 4; ++k) {
      double q = C->Qc1.p[k][f];
      if (q < qmin[k][c1]) phc1[k][f] = (qmin[k][c1] - Q->p[k][c1]) / (q - Q->p[k][c1]);
      else if (q > qmax[k][c1]) phc1[k][f] = (qmax[k][c1] - Q->p[k][c1]) / (q - Q->p[k][c1]);
    }
    if (c2 >= 0) {   /* the reference also evaluates this side on boundary faces
                        (negative index, squeeze.cpp:98-120) and discards it */
      for (int k = 0; k < 4; ++k) {
        double q = C->Qc2.p[k][f];
        if (q < qmin[k][c2]) phc2[k][f] = (qmin[k][c2] - Q->p[k][c2]) / (q - Q->p[k][c2]);
        else if (q > qmax[k][c2]) phc2[k][f] = (qmax[k][c2] - Q->p[k][c2]) / (q - Q->p[k][c2]);
      }
    }
  }
  for (int f = 0; f < nf; ++f) {   /* :126-141 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    for (int k = 0; k < 4; ++k) ph[k][c1] = min2(phc1[k][f], ph[k][c1]);
    if (c2 >= 0)
      for (int k = 0; k < 4; ++k) ph[k][c2] = min2(phc2[k][f], ph[k][c2]);
  }
  for (int c = 0; c < nc; ++c)     /* :144-154 */
    for (int k = 0; k < 4; ++k) {
      Q->qx[k][c] = Q->qx[k][c] * ph[k][c];
      Q->qy[k][c] = Q->qy[k][c] * ph[k][c];
    }
  for (int k = 0; k < 4; ++k) { free(qmin[k]); free(qmax[k]); free(phc1[k]); free(phc2[k]); free(ph[k]); }
}

/* Barth-Jespersen limiter (limiter 3): an EXTENSION.  The reference maps
 * MINMOD to a no-op (res.cpp:36-38) and never builds its minmod.cpp draft;
 * limiter 1 keeps that.  This is the multidimensional minmod-type limiter of
 * T. J. Barth and D. C. Jespersen (AIAA 89-0366): squeeze's bounds and test
 * (the min/max of each cell and its face neighbours, phi = min over the
 * cell's faces of the factor that brings the face value inside them) applied
 * to the CURRENT reconstruction Q + grad Q . (x_f - x_i) instead of the
 * previous res() call's face states.  Pinned by tests/test_oracle_bj.py. */
static void barth_jespersen(or_case* C, qset* Q) {
  int nf = C->nf, nc = C->nc;
  double *qmin[4], *qmax[4], *ph[4];
  for (int k = 0; k < 4; ++k) {
    qmin[k] = dfill(nc, 100000000000.0);
    qmax[k] = dfill(nc, -100000000000.0);
    ph[k] = dfill(nc, 1);
  }
  for (int f = 0; f < nf; ++f) {   /* the bounds, as squeeze */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double tmin[4], tmax[4];
    if (c2 >= 0) {
      for (int k = 0; k < 4; ++k) {
        tmin[k] = min2(Q->p[k][c1], Q->p[k][c2]);
        tmax[k] = max2(Q->p[k][c1], Q->p[k][c2]);
      }
      for (int k = 0; k < 4; ++k) {
        qmin[k][c2] = min2(qmin[k][c2], tmin[k]);
        qmax[k][c2] = max2(qmax[k][c2], tmax[k]);
      }
    } else {
      for (int k = 0; k < 4; ++k) { tmin[k] = Q->p[k][c1]; tmax[k] = Q->p[k][c1]; }
    }
    for (int k = 0; k < 4; ++k) {
      qmin[k][c1] = min2(qmin[k][c1], tmin[k]);
      qmax[k][c1] = max2(qmax[k][c1], tmax[k]);
    }
  }
  for (int f = 0; f < nf; ++f) {
    int cs[2] = {C->fc1[f], C->fc2[f]};
    const double* dxs[2] = {C->c1dx, C->c2dx};
    const double* dys[2] = {C->c1dy, C->c2dy};
    for (int side = 0; side < 2; ++side) {
      int c = cs[side];
      if (c < 0) continue;
      for (int k = 0; k < 4; ++k) {
        double q = Q->p[k][c] + Q->qx[k][c] * dxs[side][f] + Q->qy[k][c] * dys[side][f];
        double p = 1.0;
        if (q < qmin[k][c]) p = (qmin[k][c] - Q->p[k][c]) / (q - Q->p[k][c]);
        else if (q > qmax[k][c]) p = (qmax[k][c] - Q->p[k][c]) / (q - Q->p[k][c]);
        ph[k][c] = min2(p, ph[k][c]);
      }
    }
  }
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k) {
      Q->qx[k][c] = Q->qx[k][c] * ph[k][c];
      Q->qy[k][c] = Q->qy[k][c] * ph[k][c];
    }
  for (int k = 0; k < 4; ++k) { free(qmin[k]); free(qmax[k]); free(ph[k]); }
}

/* reconstruction  reconstruction.cpp:15-167 */
static void reconstruction(or_case* C, qset* Q) {
  int nf = C->nf;
  double g = C->o.gamma;
  qset* L = &C->Qc1;
  qset* R = &C->Qc2;
  for (int f = 0; f < nf; ++f) {   /* :22-79 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    for (int k = 0; k < 4; ++k)
      L->p[k][f] = Q->p[k][c1] + Q->qx[k][c1] * C->c1dx[f] + Q->qy[k][c1] * C->c1dy[f];
    updateflow_int(L, f);
    double W[4];
    switch (c2) {
      case -1:
        for (int k = 0; k < 4; ++k) R->p[k][f] = C->fQ[k];
        break;
      case -2: {
        double Vn = L->u[f] * C->nx[f] + L->v[f] * C->ny[f];
        double wu = L->u[f] - 2.0 * Vn * C->nx[f];
        double wv = L->v[f] - 2.0 * Vn * C->ny[f];
        ghost_Q(g, L->rho[f], wu, wv, L->P[f], W);
        for (int k = 0; k < 4; ++k) R->p[k][f] = W[k];
        break;
      }
      case -3:
        for (int k = 0; k < 4; ++k) R->p[k][f] = L->p[k][f];
        break;
      case -4:
        ghost_Q(g, L->rho[f], -L->u[f], -L->v[f], L->P[f], W);
        for (int k = 0; k < 4; ++k) R->p[k][f] = W[k];
        break;
      default:
        for (int k = 0; k < 4; ++k)
          R->p[k][f] = Q->p[k][c2] + Q->qx[k][c2] * C->c2dx[f] + Q->qy[k][c2] * C->c2dy[f];
        break;
    }
  }
  updateflow(R);                    /* :80 */
  if (C->o.eqn != 2) return;
  qset* F = &C->Qface;              /* NS viscous face terms  :83-164 */
  for (int f = 0; f < nf; ++f)
    for (int k = 0; k < 4; ++k) F->p[k][f] = 0.5 * (L->p[k][f] + R->p[k][f]);
  updateflow(F);
  for (int f = 0; f < nf; ++f) {
    int c1 = C->fc1[f], c2 = C->fc2[f];
    switch (c2) {
      case -1:
      case -3:
        comp_vis(F, f, 0, 0, 0, 0, 0, 0, 0, 0);
        break;
      case -2: /* "SLIP_WALL CANNOT BE USED WITH NS": stresses left untouched */
        break;
      case -4: {
        double dn = fabs(sqrt(C->c1dx[f] * C->c1dx[f] + C->c1dy[f] * C->c1dy[f]));
        double dudn = -Q->u[c1] / dn;
        double dvdn = -Q->v[c1] / dn;
        double dudx = dudn * C->nx[f], dudy = dudn * C->ny[f];
        double dvdx = dvdn * C->nx[f], dvdy = dvdn * C->ny[f];
        comp_vis(F, f, dudx, dudy, dvdx, dvdy, 0, 0, 0, 0);
        break;
      }
      default: {
        double dl[4], dm[4], ddx[4], ddy[4];
        for (int k = 0; k < 4; ++k)
          dl[k] = (Q->p[k][c2] - Q->p[k][c1]) / fabs(C->fdist[f]);
        for (int k = 0; k < 4; ++k)
          dm[k] = 0.5 * (Q->qx[k][c1] * C->mx[f] + Q->qy[k][c1] * C->my[f] +
                         Q->qx[k][c2] * C->mx[f] + Q->qy[k][c2] * C->my[f]);
        for (int k = 0; k < 4; ++k) {
          ddy[k] = dl[k] * C->ly[f] + dm[k] * C->ly[f];
          ddx[k] = dl[k] * C->lx[f] + dm[k] * C->lx[f];
        }
        double drhodx = ddx[0];
        double dudx = ddx[1] - drhodx * F->u[f] / F->rho[f];
        double dvdx = ddx[2] - drhodx * F->v[f] / F->rho[f];
        double drhody = ddy[0];
        double dudy = ddy[1] - drhody * F->u[f] / F->rho[f];
        double dvdy = ddy[2] - drhody * F->u[f] / F->rho[f];   /* uses u (reference) */
        double dPdx = (g - 1) * (ddx[3] - 0.5 * drhodx * (F->u[f] * F->u[f] + F->v[f] * F->v[f]) -
                                 F->rho[f] * (F->u[f] * dudx + F->v[f] * dvdx));
        double dPdy = (g - 1) * (ddy[3] - 0.5 * drhody * (F->u[f] * F->u[f] + F->v[f] * F->v[f]) -
                                 F->rho[f] * (F->u[f] * dudy + F->v[f] * dvdy));
        comp_vis(F, f, dudx, dudy, dvdx, dvdy, dPdx, dPdy, drhodx, drhody);
        break;
      }
    }
  }
}

/* Roe's flux difference |A~|(Q_R - Q_L) for the face normal n: an
 * EXTENSION, not a restatement.  The reference reads ROE
 * (read_inputs.cpp:89-91) but its roe_flux.cpp is an unfinished draft that
 * its solver library does not build, and res.cpp:66-68 leaves the flux at
 * -101.  Roe's averages (P. L. Roe, J. Comput. Phys. 43, 1981), the
 * decomposition into two acoustic waves, the entropy wave and the shear wave,
 * and Harten's entropy fix (delta = c~/10) on the acoustic eigenvalues; the
 * device path (fv_kernels.hip roe_dissipation) evaluates the same expressions.
 * Pinned by tests/test_oracle_roe.py: consistency, upwinding of supersonic
 * states (Roe's property A~ dQ = dF), symmetry, rotation, closed-box
 * conservation and freestream preservation. */
static void roe_dissipation(double rL, double uL, double vL, double PL, double EL, double rR,
                            double uR, double vR, double PR, double ER, double dr, double du,
                            double dv, double dp, double nx, double ny, double gm1, double D[4]) {
  double sL = sqrt(rL), sR = sqrt(rR);
  double w = sL + sR;
  double HL = (EL + PL) / rL;
  double HR = (ER + PR) / rR;
  double rt = sL * sR;
  double ut = (sL * uL + sR * uR) / w;
  double vt = (sL * vL + sR * vR) / w;
  double Ht = (sL * HL + sR * HR) / w;
  double kt = (ut * ut + vt * vt) / 2;
  double ct = sqrt(gm1 * (Ht - kt));
  double Vt = ut * nx + vt * ny;
  double dV = du * nx + dv * ny;
  double del = ct / 10;
  double l1 = fabs(Vt - ct), l5 = fabs(Vt + ct);
  double l2 = fabs(Vt);
  if (l1 < del) l1 = (l1 * l1 + del * del) / (2 * del);
  if (l5 < del) l5 = (l5 * l5 + del * del) / (2 * del);
  double c2 = ct * ct;
  double a1 = l1 * (dp - rt * ct * dV) / (2 * c2);
  double a2 = l2 * (dr - dp / c2);
  double a3 = l2 * rt;
  double a5 = l5 * (dp + rt * ct * dV) / (2 * c2);
  D[0] = a1 + a2 + a5;
  D[1] = a1 * (ut - ct * nx) + a2 * ut + a3 * (du - dV * nx) + a5 * (ut + ct * nx);
  D[2] = a1 * (vt - ct * ny) + a2 * vt + a3 * (dv - dV * ny) + a5 * (vt + ct * ny);
  D[3] = a1 * (Ht - ct * Vt) + a2 * kt + a3 * (ut * du + vt * dv - Vt * dV) + a5 * (Ht + ct * Vt);
}

/* the primitives of a conserved state, as updateflow (class_q.hpp:35-48) */
static void prims_of(const double q[4], double gamma, double* rho, double* u, double* v,
                     double* E, double* P) {
  *rho = q[0];
  *E = q[3];
  *u = q[1] / *rho;
  *v = q[2] / *rho;
  *P = (*E - (0.5 * *rho * (*u * *u + *v * *v))) * (gamma - 1);
}

void or_roe_flux(const double qL[4], const double qR[4], double nx, double ny, double gamma,
                 double F[4]) {
  double rL, uL, vL, EL, PL, rR, uR, vR, ER, PR, D[4];
  prims_of(qL, gamma, &rL, &uL, &vL, &EL, &PL);
  prims_of(qR, gamma, &rR, &uR, &vR, &ER, &PR);
  double VnL = uL * nx + vL * ny, VnR = uR * nx + vR * ny;
  double FL[4] = {rL * VnL, rL * uL * VnL + PL * nx, rL * vL * VnL + PL * ny, VnL * (EL + PL)};
  double FR[4] = {rR * VnR, rR * uR * VnR + PR * nx, rR * vR * VnR + PR * ny, VnR * (ER + PR)};
  roe_dissipation(rL, uL, vL, PL, EL, rR, uR, vR, PR, ER, rR - rL, uR - uL, vR - vL, PR - PL,
                  nx, ny, gamma - 1, D);
  for (int k = 0; k < 4; ++k) F[k] = (FL[k] + FR[k]) / 2 - D[k] / 2;
}

/* res  res.cpp:21-109 (with rusanov.cpp:12-36 and class_f.hpp:26-59) */
static void res(or_case* C, qset* Q) {
  int nf = C->nf, nc = C->nc;
  double* Fp[4];
  double* Fv[4];
  for (int k = 0; k < 4; ++k) { Fp[k] = dfill(nf, -101); Fv[k] = dfill(nf, 0); }
  if (C->o.order == 1) {
    reconstruction(C, Q);
  } else {
    compute_gradient(C, Q);
    if (C->o.limiter == 2) squeeze(C, Q);
    else if (C->o.limiter == 3) barth_jespersen(C, Q);
    reconstruction(C, Q);
    if (C->o.eqn == 2) {           /* class_F::update_vis  class_f.hpp:52-59 */
      qset* F = &C->Qface;
      for (int f = 0; f < nf; ++f) {
        double nx = C->nx[f], ny = C->ny[f];
        Fv[0][f] = 0;
        Fv[1][f] = F->txx[f] * nx + F->txy[f] * ny;
        Fv[2][f] = F->txy[f] * nx + F->tyy[f] * ny;
        Fv[3][f] = nx * (F->u[f] * F->txx[f] + F->v[f] * F->txy[f] + F->k * F->dtdx[f]) +
                   ny * (F->u[f] * F->txy[f] + F->v[f] * F->tyy[f] + F->k * F->dtdy[f]);
      }
    }
  }
  if (C->o.flux == 1) {            /* rusanov.cpp:12-36 */
    qset* L = &C->Qc1;
    qset* R = &C->Qc2;
    for (int f = 0; f < nf; ++f) {
      double nx = C->nx[f], ny = C->ny[f];
      double VnL = L->u[f] * nx + L->v[f] * ny;      /* class_f.hpp:42-49 */
      double VnR = R->u[f] * nx + R->v[f] * ny;
      double FL[4], FR[4];
      FL[0] = L->rho[f] * VnL;
      FL[1] = L->rho[f] * L->u[f] * VnL + L->P[f] * nx;
      FL[2] = L->rho[f] * L->v[f] * VnL + L->P[f] * ny;
      FL[3] = VnL * (L->E[f] + L->P[f]);
      FR[0] = R->rho[f] * VnR;
      FR[1] = R->rho[f] * R->u[f] * VnR + R->P[f] * nx;
      FR[2] = R->rho[f] * R->v[f] * VnR + R->P[f] * ny;
      FR[3] = VnR * (R->E[f] + R->P[f]);
      double vn_avg = (VnL + VnR) / 2;
      double c_avg = (L->c[f] + R->c[f]) / 2;
      L->vn_avg[f] = vn_avg; L->c_avg[f] = c_avg;
      R->vn_avg[f] = vn_avg; R->c_avg[f] = c_avg;
      for (int k = 0; k < 4; ++k)
        Fp[k][f] = (FL[k] + FR[k]) / 2 - (fabs(vn_avg) + c_avg) / 2 * (R->p[k][f] - L->p[k][f]);
    }
  } else if (C->o.flux == 2) {     /* Roe (an extension, roe_dissipation above) */
    qset* L = &C->Qc1;
    qset* R = &C->Qc2;
    double gm1 = C->o.gamma - 1;
    for (int f = 0; f < nf; ++f) {
      double nx = C->nx[f], ny = C->ny[f];
      double VnL = L->u[f] * nx + L->v[f] * ny;
      double VnR = R->u[f] * nx + R->v[f] * ny;
      double FL[4], FR[4], D[4];
      FL[0] = L->rho[f] * VnL;
      FL[1] = L->rho[f] * L->u[f] * VnL + L->P[f] * nx;
      FL[2] = L->rho[f] * L->v[f] * VnL + L->P[f] * ny;
      FL[3] = VnL * (L->E[f] + L->P[f]);
      FR[0] = R->rho[f] * VnR;
      FR[1] = R->rho[f] * R->u[f] * VnR + R->P[f] * nx;
      FR[2] = R->rho[f] * R->v[f] * VnR + R->P[f] * ny;
      FR[3] = VnR * (R->E[f] + R->P[f]);
      double vn_avg = (VnL + VnR) / 2;
      double c_avg = (L->c[f] + R->c[f]) / 2;
      L->vn_avg[f] = vn_avg; L->c_avg[f] = c_avg;
      R->vn_avg[f] = vn_avg; R->c_avg[f] = c_avg;
      roe_dissipation(L->rho[f], L->u[f], L->v[f], L->P[f], L->E[f], R->rho[f], R->u[f],
                      R->v[f], R->P[f], R->E[f], R->rho[f] - L->rho[f], R->u[f] - L->u[f],
                      R->v[f] - L->v[f], R->P[f] - L->P[f], nx, ny, gm1, D);
      for (int k = 0; k < 4; ++k) Fp[k][f] = (FL[k] + FR[k]) / 2 - D[k] / 2;
    }
  }
  for (int k = 0; k < 4; ++k)      /* residual->reset()  res.cpp:73 */
    for (int c = 0; c < nc; ++c) C->res[k][c] = 0;
  for (int f = 0; f < nf; ++f) {   /* :74-102 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    double a = C->area[f];
    for (int k = 0; k < 4; ++k) {
      double st = (Fp[k][f] * a);
      double sv = (Fv[k][f] * a);
      C->res[k][c1] += sv - st;
      if (c2 >= 0) C->res[k][c2] -= sv - st;
    }
  }
  for (int c = 0; c < nc; ++c)     /* :104-109 */
    for (int k = 0; k < 4; ++k) C->res[k][c] = C->res[k][c] / C->vol[c];
  for (int k = 0; k < 4; ++k) { free(Fp[k]); free(Fv[k]); }
}

/* the step's dt from the last res() call's face wave speeds (ssp_rk3.cpp:18-47);
 * a fixed dt for every cell instead when C->fixed_dt > 0 (a test hook) */
static void step_dt(or_case* C) {
  int nf = C->nf, nc = C->nc;
  double* dt = C->dt;
  if (C->fixed_dt > 0) {
    for (int c = 0; c < nc; ++c) dt[c] = C->fixed_dt;
    C->time += C->fixed_dt;
    return;
  }
  for (int c = 0; c < nc; ++c) dt[c] = 0;
  for (int f = 0; f < nf; ++f) {   /* :28-36 */
    int c1 = C->fc1[f], c2 = C->fc2[f];
    dt[c1] += (fabs(C->Qc1.vn_avg[f]) + C->Qc1.c_avg[f]) * C->area[f];
    if (c2 >= 0) dt[c2] += (fabs(C->Qc2.vn_avg[f]) + C->Qc2.c_avg[f]) * C->area[f];
  }
  if (C->o.dt_type == 1) {         /* GLOBAL: max of the spectral sums  :38-42 */
    double m = dt[0];
    for (int c = 1; c < nc; ++c)
      if (m < dt[c]) m = dt[c];
    for (int c = 0; c < nc; ++c) dt[c] = (2 * C->o.cfl * C->vol[c]) / m;
  } else if (C->o.dt_type == 2) {  /* LOCAL  :43-47 */
    for (int c = 0; c < nc; ++c) dt[c] = (2 * C->o.cfl * C->vol[c]) / dt[c];
  } else if (C->o.dt_type == 3) {
    /* UNIFORM: an extension with no reference counterpart.  One dt for every
     * cell, 2*CFL / max_c(Sigma_c / V_c), the smallest LOCAL dt, so that the
     * run is time-accurate. */
    double m = dt[0] / C->vol[0];
    for (int c = 1; c < nc; ++c) {
      const double r = dt[c] / C->vol[c];
      if (m < r) m = r;
    }
    const double d = (2 * C->o.cfl) / m;
    for (int c = 0; c < nc; ++c) dt[c] = d;
    C->time += d;
  }
}

/* ssp_rk3  ssp_rk3.cpp:18-77 */
static void ssp_rk3(or_case* C) {
  int nc = C->nc;
  qset S;
  qset_alloc(&S, nc);
  qset_init_plain(&S, C->o.gamma, C->o.R);
  double* dt = C->dt;
  step_dt(C);
  qset* Q = &C->Qbar;
  res(C, Q);                       /* :50-57 */
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k) S.p[k][c] = Q->p[k][c] + dt[c] * C->res[k][c];
  updateflow(&S);
  res(C, &S);                      /* :60-67 */
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k)
      S.p[k][c] = 0.75 * Q->p[k][c] + 0.25 * S.p[k][c] + 0.25 * dt[c] * C->res[k][c];
  updateflow(&S);
  res(C, &S);                      /* :70-77 */
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k)
      Q->p[k][c] = (1.0 / 3.0) * Q->p[k][c] + (2.0 / 3.0) * S.p[k][c] + (2.0 / 3.0) * dt[c] * C->res[k][c];
  updateflow(Q);
  qset_free(&S);
}

/* Two-stage SSP-RK2 (scheme = 3): an EXTENSION.  The reference's
 * ssp_rk2.cpp is a draft its solver library never builds (main.cpp:123-126
 * runs ssp_rk3 for SSP_RK2); this follows that draft's stage structure
 * (ssp_rk2.cpp:44-60: Q* = Q + dt R(Q), Q_n+1 = (Q + Q* + dt R(Q*))/2, the
 * lagged limiter fed by the previous res() call), the last stage written in
 * ssp_rk3's Shu-Osher form (the three terms each halved), with the dt of the
 * scheme's time_step_type, as ssp_rk3.  Pinned by tests/test_oracle_rk2.py (second
 * order in time, conservation, freestream preservation). */
static void ssp_rk2(or_case* C) {
  int nc = C->nc;
  qset S;
  qset_alloc(&S, nc);
  qset_init_plain(&S, C->o.gamma, C->o.R);
  double* dt = C->dt;
  step_dt(C);
  qset* Q = &C->Qbar;
  res(C, Q);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k) S.p[k][c] = Q->p[k][c] + dt[c] * C->res[k][c];
  updateflow(&S);
  res(C, &S);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 4; ++k) Q->p[k][c] = 0.5 * Q->p[k][c] + 0.5 * S.p[k][c] + 0.5 * dt[c] * C->res[k][c];
  updateflow(Q);
  qset_free(&S);
}

/* ------------------------------------------------------------------ */
/* public API                                                          */
/* ------------------------------------------------------------------ */
or_case* or_create(int32_t np, int32_t nf, int32_t nc, const double* x, const double* y,
                   const int32_t* fp1, const int32_t* fp2, const int32_t* fc1,
                   const int32_t* fc2_file, const or_options* opt) {
  g_err[0] = 0;
  if (np <= 0 || nf <= 0 || nc <= 0) { snprintf(g_err, sizeof g_err, "empty mesh"); return NULL; }
  for (int f = 0; f < nf; ++f) {
    if (fp1[f] < 0 || fp1[f] >= np || fp2[f] < 0 || fp2[f] >= np || fc1[f] < 0 || fc1[f] >= nc ||
        fc2_file[f] >= nc || fc2_file[f] < -4) {
      snprintf(g_err, sizeof g_err, "face %d: index out of range", f);
      return NULL;
    }
  }
  or_case* C = (or_case*)xcalloc(1, sizeof(or_case));
  C->np = np; C->nf = nf; C->nc = nc;
  C->o = *opt;
  C->x = (double*)xcalloc(np, 8); memcpy(C->x, x, 8 * (size_t)np);
  C->y = (double*)xcalloc(np, 8); memcpy(C->y, y, 8 * (size_t)np);
  C->fp1 = (int*)xcalloc(nf, 4); memcpy(C->fp1, fp1, 4 * (size_t)nf);
  C->fp2 = (int*)xcalloc(nf, 4); memcpy(C->fp2, fp2, 4 * (size_t)nf);
  C->fc1 = (int*)xcalloc(nf, 4); memcpy(C->fc1, fc1, 4 * (size_t)nf);
  C->fc2 = (int*)xcalloc(nf, 4);
  C->face2bc = (int*)xcalloc(nf, 4);
  C->bc_faces = (int*)xcalloc(nf, 4);
  int nb = 0;
  for (int f = 0; f < nf; ++f) {   /* readmesh.cpp:58-85 */
    int c2 = fc2_file[f];
    C->fc2[f] = c2 >= 0 ? c2 : bc_kind(opt, c2);
    if (c2 == -1) C->num_bc1++;
    if (c2 == -2) C->num_bc2++;
    C->face2bc[f] = -101;
    if (c2 < 0) { C->bc_faces[nb] = f; C->face2bc[f] = nb; nb++; }
  }
  C->nb = nb;
  /* cell -> face lists in increasing face index (readmesh.cpp:72-78) */
  C->cf_ptr = (int*)xcalloc((size_t)nc + 1, 4);
  for (int f = 0; f < nf; ++f) {
    C->cf_ptr[fc1[f] + 1]++;
    if (fc2_file[f] >= 0) C->cf_ptr[fc2_file[f] + 1]++;
  }
  for (int c = 0; c < nc; ++c) C->cf_ptr[c + 1] += C->cf_ptr[c];
  C->cf_idx = (int*)xcalloc((size_t)C->cf_ptr[nc], 4);
  int* fill = (int*)xcalloc(nc, 4);
  for (int f = 0; f < nf; ++f) {
    C->cf_idx[C->cf_ptr[fc1[f]] + fill[fc1[f]]++] = f;
    if (fc2_file[f] >= 0) C->cf_idx[C->cf_ptr[fc2_file[f]] + fill[fc2_file[f]]++] = f;
  }
  free(fill);
  for (int c = 0; c < nc; ++c)
    if (C->cf_ptr[c + 1] - C->cf_ptr[c] < 3) {
      snprintf(g_err, sizeof g_err, "cell %d has fewer than 3 faces", c);
      or_destroy(C);
      return NULL;
    }
  C->area = dfill(nf, 0); C->nx = dfill(nf, 0); C->ny = dfill(nf, 0);
  C->fcx = dfill(nf, 0); C->fcy = dfill(nf, 0);
  C->vol = dfill(nc, 0); C->ccx = dfill(nc, 0); C->ccy = dfill(nc, 0);
  C->bccx = dfill(nb, -101); C->bccy = dfill(nb, -101);
  C->c1dx = dfill(nf, -101); C->c1dy = dfill(nf, -101);
  C->c2dx = dfill(nf, -101); C->c2dy = dfill(nf, -101);
  C->dxj = dfill(nf, -101); C->dyj = dfill(nf, -101);
  C->delta = dfill(nc, -101);
  C->Ixx = dfill(nc, -101); C->Iyy = dfill(nc, -101); C->Ixy = dfill(nc, -101);
  C->fdist = dfill(nf, -101); C->lx = dfill(nf, -101); C->ly = dfill(nf, -101);
  C->mx = dfill(nf, -101); C->my = dfill(nf, -101);
  build_geometry(C);
  /* the no-slip gradient ghost reads cell [face index]  cell_gradient.cpp:63-66 */
  if (C->o.order == 2)
    for (int f = 0; f < nf; ++f)
      if (C->fc2[f] == -4 && f >= nc) {
        snprintf(g_err, sizeof g_err, "no-slip face %d has no cell of the same index", f);
        or_destroy(C);
        return NULL;
      }
  /* freestream  main.cpp:77-79, class_flow.hpp:28-46 */
  const or_options* o = &C->o;
  C->fCp = o->R * o->gamma / (o->gamma - 1);
  C->fT = o->P / (o->rho * o->R);
  C->fk = (o->mu * C->fCp) / o->Pr;
  C->fQ[0] = o->rho;
  C->fQ[1] = o->rho * o->u;
  C->fQ[2] = o->rho * o->v;
  C->fE = (o->P / (o->gamma - 1)) + 0.5 * o->rho * (o->u * o->u + o->v * o->v);
  C->fQ[3] = C->fE;
  double nrm = sqrt(o->u * o->u + o->v * o->v);
  C->flowx = o->u / nrm;
  C->flowy = o->v / nrm;
  for (int k = 0; k < 4; ++k) C->res[k] = dfill(nc, 0);    /* residual.init  main.cpp:80 */
  C->dt = dfill(nc, 0);
  qset_alloc(&C->Qbar, nc);
  qset_init_flow(&C->Qbar, C);     /* main.cpp:108-109 */
  updateQ(&C->Qbar);
  qset_alloc(&C->Qface, nf);       /* main.cpp:114-118 */
  qset_alloc(&C->Qc1, nf);
  qset_alloc(&C->Qc2, nf);
  qset_init_flow(&C->Qface, C);
  qset_init_flow(&C->Qc1, C);
  qset_init_flow(&C->Qc2, C);
  return C;
}

void or_destroy(or_case* C) {
  if (!C) return;
  free(C->x); free(C->y); free(C->fp1); free(C->fp2); free(C->fc1); free(C->fc2);
  free(C->area); free(C->nx); free(C->ny); free(C->fcx); free(C->fcy);
  free(C->vol); free(C->ccx); free(C->ccy); free(C->cf_ptr); free(C->cf_idx);
  free(C->bc_faces); free(C->face2bc); free(C->bccx); free(C->bccy);
  free(C->c1dx); free(C->c1dy); free(C->c2dx); free(C->c2dy); free(C->dxj); free(C->dyj);
  free(C->delta); free(C->Ixx); free(C->Iyy); free(C->Ixy);
  free(C->fdist); free(C->lx); free(C->ly); free(C->mx); free(C->my);
  for (int k = 0; k < 4; ++k) free(C->res[k]);
  free(C->dt);
  if (C->Qbar.n) qset_free(&C->Qbar);
  if (C->Qface.n) qset_free(&C->Qface);
  if (C->Qc1.n) qset_free(&C->Qc1);
  if (C->Qc2.n) qset_free(&C->Qc2);
  free(C);
}

int or_set_state(or_case* C, const double* p1, const double* p2, const double* p3,
                 const double* p4) {
  /* main.cpp:94-106 then the face re-initialisation of main.cpp:114-118 */
  C->time = 0.0;
  qset_init_plain(&C->Qbar, C->o.gamma, C->o.R);
  const double* src[4] = {p1, p2, p3, p4};
  for (int k = 0; k < 4; ++k) memcpy(C->Qbar.p[k], src[k], 8 * (size_t)C->nc);
  updateflow(&C->Qbar);
  qset_init_flow(&C->Qface, C);
  qset_init_flow(&C->Qc1, C);
  qset_init_flow(&C->Qc2, C);
  return 0;
}

int or_step(or_case* C, int32_t n) {
  for (int i = 0; i < n; ++i) {
    if (C->o.scheme == 3) ssp_rk2(C);
    else ssp_rk3(C);
  }
  return 0;
}

void or_set_fixed_dt(or_case* C, double dt) { C->fixed_dt = dt; }

int or_res(or_case* C) {
  res(C, &C->Qbar);
  return 0;
}

void or_residual_l2(const or_case* C, double out[4]) {   /* save_sol.cpp:79-89 */
  for (int k = 0; k < 4; ++k) {
    double s = 0;
    for (int c = 0; c < C->nc; ++c) s += C->res[k][c] * C->res[k][c];
    out[k] = sqrt(s);
  }
}

void or_wall_forces(const or_case* C, double length, double out[2]) {
  /* save_sol.cpp:20-75: wall faces are faces 0..numWALL-1; the centroid and
   * face-centre coordinates are truncated to int (save_sol.cpp:28-32) */
  const qset* Q = &C->Qbar;
  const or_options* o = &C->o;
  double Fx = 0, Fy = 0;
  for (int face = 0; face < C->num_wall; ++face) {
    int c1 = C->fc1[face];
    int cellx = (int)C->ccx[c1], celly = (int)C->ccy[c1];
    int facex = (int)C->fcx[face], facey = (int)C->fcy[face];
    double fr[4];
    for (int k = 0; k < 4; ++k)
      fr[k] = Q->p[k][c1] + ((facex - cellx) * Q->qx[k][c1] + (facey - celly) * Q->qy[k][c1]);
    double rho = fr[0];
    double u = fr[1] / rho;
    double v = fr[2] / rho;
    double E = fr[3];
    double P = (E - (0.5 * rho * (u * u + v * v))) * (Q->gamma - 1);
    if (o->eqn == 2) {
      double dn = fabs(sqrt(C->c1dx[face] * C->c1dx[face] + C->c1dy[face] * C->c1dy[face]));
      double dudn = -Q->u[c1] / dn;
      double dvdn = -Q->v[c1] / dn;
      double tx = Q->mu * dudn;
      double ty = Q->mu * dvdn;
      double tw = sqrt(tx * tx + ty * ty);
      if (C->ny[face] < 0) {
        Fx += tw * C->area[face] * (-C->ny[face]);
        Fy += tw * C->area[face] * (C->nx[face]);
      } else {
        Fx += tw * C->area[face] * (C->ny[face]);
        Fy += tw * C->area[face] * (-C->nx[face]);
      }
    }
    double force = P * C->area[face];
    Fy += force * C->ny[face];
    Fx += force * C->nx[face];
  }
  double L = Fx * (-C->flowy) + Fy * C->flowx;
  double D = Fy * C->flowy + Fx * C->flowx;
  double Vmag = sqrt(o->u * o->u + o->v * o->v);
  out[0] = L / (0.5 * o->rho * Vmag * Vmag * length);
  out[1] = D / (0.5 * o->rho * Vmag * Vmag * length);
}

const void* or_array(const or_case* C, const char* name, int64_t* n, char* kind) {
  *kind = 'd';
  int nc = C->nc, nf = C->nf;
  const qset* Q = &C->Qbar;
  struct { const char* nm; const double* p; int len; } tab[] = {
    {"p1", Q->p[0], nc}, {"p2", Q->p[1], nc}, {"p3", Q->p[2], nc}, {"p4", Q->p[3], nc},
    {"rho", Q->rho, nc}, {"u", Q->u, nc}, {"v", Q->v, nc}, {"P", Q->P, nc},
    {"Qxp1", Q->qx[0], nc}, {"Qxp2", Q->qx[1], nc}, {"Qxp3", Q->qx[2], nc}, {"Qxp4", Q->qx[3], nc},
    {"Qyp1", Q->qy[0], nc}, {"Qyp2", Q->qy[1], nc}, {"Qyp3", Q->qy[2], nc}, {"Qyp4", Q->qy[3], nc},
    {"R1", C->res[0], nc}, {"R2", C->res[1], nc}, {"R3", C->res[2], nc}, {"R4", C->res[3], nc},
    {"Vn_avg", C->Qc1.vn_avg, nf}, {"c_avg", C->Qc1.c_avg, nf},
    {"fL1", C->Qc1.p[0], nf}, {"fL2", C->Qc1.p[1], nf}, {"fL3", C->Qc1.p[2], nf}, {"fL4", C->Qc1.p[3], nf},
    {"fR1", C->Qc2.p[0], nf}, {"fR2", C->Qc2.p[1], nf}, {"fR3", C->Qc2.p[2], nf}, {"fR4", C->Qc2.p[3], nf},
    {"tau_xx", C->Qface.txx, nf}, {"tau_xy", C->Qface.txy, nf}, {"tau_yy", C->Qface.tyy, nf},
    {"dTdx", C->Qface.dtdx, nf}, {"dTdy", C->Qface.dtdy, nf},
    {"dt", C->dt, nc}, {"time", &C->time, 1},
    {"geom/face_area", C->area, nf}, {"geom/face_nx", C->nx, nf}, {"geom/face_ny", C->ny, nf},
    {"geom/face_centerx", C->fcx, nf}, {"geom/face_centery", C->fcy, nf},
    {"geom/cell_vol", C->vol, nc}, {"geom/cell_centerx", C->ccx, nc}, {"geom/cell_centery", C->ccy, nc},
    {"geom/face_cell1_dx", C->c1dx, nf}, {"geom/face_cell1_dy", C->c1dy, nf},
    {"geom/face_cell2_dx", C->c2dx, nf}, {"geom/face_cell2_dy", C->c2dy, nf},
    {"geom/face_dxj", C->dxj, nf}, {"geom/face_dyj", C->dyj, nf},
    {"geom/Ixx", C->Ixx, nc}, {"geom/Iyy", C->Iyy, nc}, {"geom/Ixy", C->Ixy, nc},
    {"geom/face_dist", C->fdist, nf}, {"geom/face_lx", C->lx, nf}, {"geom/face_ly", C->ly, nf},
    {"geom/face_mx", C->mx, nf}, {"geom/face_my", C->my, nf},
    {"geom/BC_cell_centerx", C->bccx, C->nb}, {"geom/BC_cell_centery", C->bccy, C->nb},
  };
  for (size_t i = 0; i < sizeof tab / sizeof tab[0]; ++i)
    if (strcmp(tab[i].nm, name) == 0) { *n = tab[i].len; return tab[i].p; }
  if (strcmp(name, "geom/face_cell2") == 0) { *kind = 'i'; *n = nf; return C->fc2; }
  *n = 0;
  return NULL;
}
