// fv_kernels.hpp — argument blocks and launchers of the FV kernels.
//
// The kernels are templates on the arithmetic type T of the state:
//   double  the FP64 path: state = conserved variables, arithmetic in FP64,
//           bit-identical to the reference;
//   float   the FP32-mixed path: state = FP32 perturbation from freestream
//           (dQ = Q - Q_inf), reconstruction, limiter, flux and update in
//           FP32; geometry and LSQ inverses stay FP64 (centroid and face-centre
//           offsets are formed in FP64 before rounding) and the LSQ solve
//           multiplies in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fvk {

// Read-only mesh data on the device (new numbering; see mesh_prep.hpp).
struct DevMesh {
  int32_t nc;                  // real cells
  int32_t ncp;                 // padded cells = nblocks * blk
  int32_t blk;                 // cells per block
  int32_t nblocks;
  int32_t max_halo;            // LDS row length is blk + max_halo
  int32_t max_bfaces;          // LDS face-result row length
  int32_t max_deg;             // most faces of any cell (the prefetch depth's choice)
  const int32_t* chunk_off;    // slot offset of each 64-cell chunk
  const uint8_t* deg;          // faces per cell (0 = padding)
  const int32_t* code;         // per slot: global neighbour/owner or boundary code
  const int16_t* lcode;        // per slot: (block-local neighbour << 1) | owner, or ~kind
  const int16_t* sface;        // per slot: (block face index << 1) | owner
  const double* s_nx;          // per slot: face unit normal (out of the owner)
  const double* s_ny;
  const double* s_area;        // per slot: face length
  const double* s_fcx;         // per slot: face centre
  const double* s_fcy;
  const double* ccx;           // per cell: centroid
  const double* ccy;
  const double* vol;           // per cell: area
  const double* ixx;           // per cell: LSQ inverse (readmesh.cpp:316-320)
  const double* iyy;
  const double* ixy;
  const float* ixx32;          // the same rounded to FP32 (FP32-mixed on well-conditioned
  const float* iyy32;          // meshes; null otherwise, and then ixx.. are read)
  const float* ixy32;
  const int32_t* halo_ptr;     // per block: halo cells
  const int32_t* halo_idx;
  const int32_t* bf_ptr;       // per block: face list
  const int16_t* bf_a;         // local owner index
  const int16_t* bf_b;         // local neighbour index or ~kind
  const double* bf_nx;
  const double* bf_ny;
  const double* bf_area;
  const double* bf_fcx;
  const double* bf_fcy;
  // FP32-mixed only (null otherwise): single-precision geometry, each value
  // computed in FP64 and rounded once
  const float* s_dfx32;        // per slot: face centre - cell centroid
  const float* s_dfy32;
  const float* bf32_nx;        // per block face: normal, length
  const float* bf32_ny;
  const float* bf32_area;
  const float* bf32_d1x;       // face centre - owner centroid
  const float* bf32_d1y;
  const float* bf32_d2x;       // face centre - neighbour centroid (interior)
  const float* bf32_d2y;
  const float* bf32_lx;        // NS: unit owner->neighbour, centroid distance
  const float* bf32_ly;        //     (interior); owner's wall distance (no-slip)
  const float* bf32_dist;
};

// Per-instance constants: gas and freestream in the arithmetic type
// (class_flow, class_Q), the limiter switch and the instance's first block.
// A single solve is a batch of one instance.
template <typename T>
struct FlowT {
  T gamma, gm1, Rgas, mu, k;
  T cfl2;                      // 2*CFL
  T qinf[4];                   // freestream conserved state (class_flow::updateQ)
  T rho, u, v, P;              // freestream primitives
  T vn0, c0;                   // initial face |Vn_avg| and c_avg (class_q.hpp:61-62)
  T sent[4];                   // the -101 initial face state, in stored-state units
  double qinf64[4];            // FP64 copies for steps an FP32 path takes in FP64
  double gm1_64, Rgas64, mu64, k64;
  int32_t squeeze;             // limiter SQUEEZE or Barth-Jespersen (1) or none (0)
  int32_t bj;                  // limiter 3, Barth-Jespersen: the current reconstruction, unlagged
  int32_t first_block;         // first mesh block of the instance
};

// One RK stage.  State arrays are cell-major: q[4*cell + k], g[8*cell + k]
// (gx[0..3], gy[0..3]), res[4*cell + k].
// A batch concatenates its instances' padded cell ranges and block lists;
// per-block instance ids select the instance's constants and dt maximum.
template <typename T>
struct StageArgsT {
  DevMesh m;
  const FlowT<T>* flows;       // per instance
  const int32_t* blk_inst;     // per mesh block: instance (batch only, else null)
  const int32_t* active;       // logical block -> mesh block (batch only, else null)
  int32_t nactive;             // active blocks (batch only)
  int32_t accumulate;          // 1: pass B scatters each face's terms with LDS atomics
  int32_t stage;               // 1, 2, 3
  int32_t last;                // the step's last stage (3, or 2 for the two-stage SSP-RK2):
                               // spectral sums for the next dt, GLOBAL maximum, residual
  T ca, cb;                    // stage > 1: Q_out = ca Q_n + cb (Q + dt R)
  int32_t dt_global;           // one maximum per step: time_step_type GLOBAL or uniform
  int32_t dt_uniform;          // time_step_type 3: one dt = 2 CFL / max(Sigma_i / V_i)
  int32_t sentinel;            // lagged face states are the -101 initial values
  int32_t fs_prims;            // stage input's primitives are the freestream ones
  int32_t store_res;           // write the residual (stage 3 of a reported step)
  int32_t xcd_remap;           // XCD block order: 0 off, 1 contiguous ranges, G>1 runs of G
  int32_t xcd_blocks;          // blocks per XCD range, or -G (set by the launcher)
  int32_t variant;             // kVariantCell, kVariantSlot, kVariantLds or kVariantPersist
  int32_t roe;                 // flux_solver ROE: the Roe-flux pass B (LDS-staged gather only)
  const T* qc;                 // stage input state (4 per cell)
  const T* qp;                 // previous stage's input state (lagged limiter)
  const T* gp;                 // previous stage's limited gradient (8 per cell)
  T* g;                        // this stage's limited gradient (8 per cell)
  const T* qn;                 // state at the start of the step
  T* qout;                     // stage output state
  T* sigma;                    // per-cell spectral sum of the last stage 3
  unsigned long long* smax_cur;   // per instance: max spectral sum used by this step
                                  // (GLOBAL dt), FP64 bits
  unsigned long long* smax_next;  // per instance: max spectral sum written by stage 3
  T* res;                      // residual (4 per cell)
  double* time;                // per instance: physical time of the uniform dt
  int32_t tiles_total;         // kVariantPersist (set by the launcher): logical blocks
  int32_t tiles_per_xcd;       //   of the one-workgroup-per-block launch, and per XCD
};

// launch shapes: one thread per cell; one thread per (cell, face slot);
// LDS-staged blocks (block + halo in LDS, one thread per face for the flux);
// or the LDS-staged blocks walked by a persistent grid, per XCD
enum { kVariantCell = 1, kVariantSlot = 2, kVariantLds = 3, kVariantPersist = 4 };

// LDS bytes the LDS-staged kernels need for this mesh and state type size
// (pass B stages centroids only for FP64: the FP32 kernels read per-face
// offsets instead)
size_t lds_bytes_pass_a(const DevMesh& m, size_t tsize);
size_t lds_bytes_pass_b(const DevMesh& m, int order, bool last, size_t tsize);

template <typename T>
hipError_t launch_pass_a(const StageArgsT<T>& a, int squeeze, int block, hipStream_t s);
template <typename T>
hipError_t launch_pass_b(const StageArgsT<T>& a, int order, int ns, int block, hipStream_t s);
// first-step spectral sums and their per-instance maxima
template <typename T>
hipError_t launch_init_sigma(const StageArgsT<T>& a, hipStream_t s);
// per-mesh-block partial sums of R_k^2: partial[k*nblocks + b]
template <typename T>
hipError_t launch_res_sq(const T* res, int32_t ncp, int32_t blk, double* partial, hipStream_t s);
// out[k] = sqrt of the compensated sum over b of partial[k*nblocks + b],
// k < 4: 64 strided Neumaier sums, then those added in a fixed order
hipError_t launch_res_norm(const double* partial, int32_t nblocks, double* out, hipStream_t s);
// out[4*i + k] = the same per instance i < n over its blocks
// [inst_off[i], inst_off[i+1])
hipError_t launch_res_norm_inst(const double* partial, int32_t nblocks, const int32_t* inst_off,
                               int32_t n, double* out, hipStream_t s);
// per-mesh-block compensated sums of V*q_k (k = 0..3) and of V:
// partial[(2j)*nblocks + b] = hi, partial[(2j+1)*nblocks + b] = lo, j = 0..4
template <typename T>
hipError_t launch_cons_sum(const T* q, const DevMesh& m, double* partial, hipStream_t s);
template <typename T>
hipError_t launch_fill(T* p, int64_t n, T v, hipStream_t s);

}  // namespace fvk
