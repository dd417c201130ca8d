/* fv_oracle.h — CPU restatement of the reference FV solver (TEST INFRASTRUCTURE).
 *
 * This library is the test oracle: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product (libfvk.so and the fv2d
 * CLI) never links or calls it.
 *
 * It restates, in plain C and in the reference's own loop order (face loops
 * that scatter into cells in increasing face index), the serial FP64 solver of
 * Avery-Hantla/CFD-2D-NS-FV:
 *   geometry        src/base/readmesh.cpp:113-411
 *   initialisation  src/base/main.cpp:77-118, src/support/class_q.hpp:19-61
 *   time step       src/solver/ssp_rk3.cpp:18-77 (and a two-stage SSP-RK2, an extension)
 *   residual        src/solver/res.cpp:21-109
 *   LSQ gradient    src/solver/cell_gradient.cpp:9-130
 *   squeeze limiter src/solver/squeeze.cpp:9-154
 *   reconstruction  src/solver/reconstruction.cpp:15-167
 *   Rusanov flux    src/solver/rusanov.cpp:12-36, src/support/class_f.hpp:41-59
 *   Roe flux        an extension (the reference's draft is not built; res.cpp:66-68)
 *   residual norms  src/base/save_sol.cpp:79-89
 *   wall CL/CD      src/base/save_sol.cpp:20-75
 * Unlike the reference it keeps each cell's faces in a CSR list, so cells with
 * more than six sides work (the reference's stride is six, readmesh.cpp:38).
 *
 * Pinning (see tests/test_oracle_*.py): bit-for-bit against the reference
 * compiled from its own sources (oracle/_ref) on the committed golden fixtures,
 * and against closed forms (areas/centroids/normals, LSQ exactness on linear
 * fields, Rusanov consistency, freestream preservation, conservation).
 */
#ifndef FV_ORACLE_H
#define FV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_case or_case;

typedef struct {
  int32_t order;          /* solution_poly_order + 1 (1 or 2) */
  int32_t limiter;        /* 0 NONE, 1 MINMOD (no-op), 2 SQUEEZE, 3 Barth-Jespersen (an
                             extension) */
  int32_t eqn;            /* 1 EULER, 2 NS */
  int32_t flux;           /* 1 RUSANOV, 2 ROE: an extension (the reference leaves the
                             flux at -101, res.cpp:66-68), see or_roe_flux */
  int32_t dt_type;        /* 1 GLOBAL, 2 LOCAL, 3 UNIFORM (extension) */
  int32_t bc1, bc2;       /* BC kinds of file markers 0 and -1: -1..-4 */
  int32_t scheme;         /* time_scheme: 1 SSP_RK2 and 2 SSP_RK3 both run SSP-RK3
                             (main.cpp:123-126); 3 the two-stage SSP-RK2, an extension */
  double cfl;
  double gamma, R, mu, Pr;
  double P, rho, u, v;    /* [init] freestream */
} or_options;

/* face_c2 holds what readmesh.cpp:59-69 reads *before* the BC-kind
 * substitution: >=0 neighbour, -1 BC1 group, -2 BC2 group, -3 / -4 raw kinds.
 * Returns NULL on error (message via or_error). */
or_case* or_create(int32_t np, int32_t nf, int32_t nc, const double* x, const double* y,
                   const int32_t* fp1, const int32_t* fp2, const int32_t* fc1,
                   const int32_t* fc2_file, const or_options* opt);
const char* or_error(void);
void or_destroy(or_case* c);

/* restart semantics of main.cpp:82-106 at full precision */
int or_set_state(or_case* c, const double* p1, const double* p2, const double* p3,
                 const double* p4);
/* n calls of ssp_rk3 */
int or_step(or_case* c, int32_t n);
/* dt > 0: every following step takes this dt in every cell (tests of the
 * time integrators' order); 0: the scheme's own dt */
void or_set_fixed_dt(or_case* c, double dt);
/* one res() call on the current state (tests) */
int or_res(or_case* c);
void or_residual_l2(const or_case* c, double out[4]);
/* CL, CD of the monitor block, save_sol.cpp:20-75 (numWALL faces first) */
void or_wall_forces(const or_case* c, double length, double out[2]);
/* the Roe flux F.n of the extension (flux = 2) between conserved states qL
 * and qR across the unit normal (nx, ny) (tests) */
void or_roe_flux(const double qL[4], const double qR[4], double nx, double ny, double gamma,
                 double F[4]);

/* named array access: returns pointer and sets *n (elements), *kind 'd' or 'i' */
const void* or_array(const or_case* c, const char* name, int64_t* n, char* kind);

#ifdef __cplusplus
}
#endif
#endif
