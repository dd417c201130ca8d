// mesh_prep.hpp — host-side mesh preparation for the device solver.
//
// 1. Geometry of the reference reader (src/base/readmesh.cpp:113-320), computed
//    in the original numbering with the reference's expression order, so every
//    value is bit-identical to what readmesh stores in class_mesh.
// 2. Device layout.
//    * Cells are grouped into blocks of `blk` cells (a multiple of 64) and
//      numbered block-major, a block's cells first and its padding after.
//      Blocks come from recursive coordinate bisection of the centroids (so a
//      block is a compact patch whose neighbours are mostly inside it), or
//      from consecutive ranges of a reverse Cuthill-McKee or file order.
//    * Every block lists its halo: the cells outside it that share a face
//      with one of its cells.  Inside a block, cell i has local index i and
//      halo cell h has local index blk + h; kernels stage the block and its
//      halo in LDS and address neighbours by local index.
//    * Every block lists its faces (each face touching one of its cells,
//      once), with the owner and neighbour as local indices.
//    * A sliced-ELL "face slot" table: chunk of 64 consecutive cells (one
//      wavefront), slot j of lane l at chunk_off[chunk] + 64*j + l.  A cell's
//      slots list its faces in increasing ORIGINAL face index, which is the
//      order the reference's face loops add into each cell (res.cpp:74-102,
//      cell_gradient.cpp:93-118, ssp_rk3.cpp:28-36), so per-cell sums taken in
//      slot order reproduce its sums bit for bit.
#pragma once
#include <cstdint>
#include <string>
#include <memory>
#include <utility>
#include <vector>

#include "fvk.h"

namespace fvk {

constexpr int kChunk = 64;

// slot / face codes:
//   interior  = (neighbour_index << 1) | (this cell owns the face)
//               (global index in `code`, block-local index in `lcode`)
//   boundary  = ~((ghost_source_cell << 2) | kind), kind 0..3 = FREESTREAM,
//               SLIP_WALL, EXTRAPOLATION, NO_SLIP_WALL; the source cell (global
//               new index) is the cell whose original index equals the face's
//               original index, which the no-slip gradient ghost reads
//               (cell_gradient.cpp:63-66)
inline int32_t slot_code_interior(int32_t nb, bool owner) { return (nb << 1) | (owner ? 1 : 0); }
inline int32_t slot_code_boundary(int32_t src, int kind) { return ~((src << 2) | kind); }

// Allocator whose resize() leaves new elements default-initialised (for the
// arithmetic types here: not written), so that multi-gigabyte arrays are
// not zero-filled by one thread before the parallel loops that fill every
// element.
template <typename T>
struct UninitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind { using other = UninitAlloc<U>; };
  UninitAlloc() = default;
  template <typename U>
  UninitAlloc(const UninitAlloc<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) { ::new (static_cast<void*>(p)) U(std::forward<A>(a)...); }
};
template <typename T>
using hvec = std::vector<T, UninitAlloc<T>>;

struct HostMesh {
  int32_t np = 0, nf = 0, nc = 0, nb = 0;
  int32_t blk = 256;        // cells per block
  int32_t nblocks = 0;
  int32_t ncp = 0;          // nblocks * blk
  int32_t max_deg = 0;
  int32_t max_halo = 0;     // largest halo of a block
  int32_t max_bfaces = 0;   // largest face list of a block
  int32_t num_bc1 = 0, num_bc2 = 0, num_wall = 0;
  // original numbering (class_mesh fields)
  std::vector<int32_t> fc1, fc2;      // fc2 after BC-kind substitution
  std::vector<int32_t> cf_ptr, cf_idx;
  hvec<double> area, nx, ny, fcx, fcy;
  hvec<double> vol, ccx, ccy, ixx, iyy, ixy;
  double max_lsq_cond = 0;  // largest condition number of a cell's LSQ normal matrix
  // renumbering
  std::vector<int32_t> perm;   // ncp: new -> old, -1 for padding
  std::vector<int32_t> iperm;  // nc: old -> new
  // slot table (new numbering, padded)
  std::vector<int32_t> chunk_off;     // nchunks + 1, in slots
  std::vector<uint8_t> deg;           // ncp (0 for padding)
  hvec<int32_t> code;                 // per slot, global neighbour index
  // block-local codes are 16-bit: local indices stay below blk + max_halo and
  // block face indices below max_bfaces, both bounded far below 2^14 by LDS
  hvec<int16_t> lcode;                // per slot: (block-local neighbour << 1) | owner, or ~kind
  hvec<int16_t> sface;                // per slot, (index in the block's face list << 1) | owner
  hvec<double> s_nx, s_ny, s_area, s_fcx, s_fcy;   // per slot (0 in padding slots)
  // per cell (new numbering, padded)
  hvec<double> c_ccx, c_ccy, c_vol, c_ixx, c_iyy, c_ixy;
  // FP32-mixed only: the LSQ inverse entries rounded once to FP32 (used when
  // every cell's normal matrix is well conditioned)
  hvec<float> c_ixx32, c_iyy32, c_ixy32;
  // per block
  std::vector<int32_t> halo_ptr, halo_idx;            // halo cells (global new index)
  std::vector<int32_t> bf_ptr;                        // block face lists
  hvec<int16_t> bf_a, bf_b;                           // local owner; local neighbour or ~kind
  hvec<double> bf_nx, bf_ny, bf_area, bf_fcx, bf_fcy;
  // FP32-mixed only: the single-precision geometry the FP32 kernels use,
  // each value computed in FP64 from the arrays above and rounded once
  // (exactly what an FP32 kernel would compute from them):
  //   per slot    face centre - cell centroid;
  //   per block face  normal, length, face centre - owner / neighbour
  //   centroid; NS: unit owner->neighbour vector and centroid distance
  //   (interior), wall distance of the owner centroid (no-slip walls)
  hvec<float> s_dfx32, s_dfy32;
  hvec<float> bf32_nx, bf32_ny, bf32_area, bf32_d1x, bf32_d1y, bf32_d2x, bf32_d2y;
  hvec<float> bf32_lx, bf32_ly, bf32_dist;
  std::string warnings;
};

// Returns FVK_OK or an FVK_ERR_* code with *err set.  `threads` host threads
// share the work (1: serial); the result does not depend on it, bit for bit.
int prepare_mesh(const fvk_mesh& m, const fvk_options& o, HostMesh* hm, std::string* err,
                 int threads = 1);

}  // namespace fvk
