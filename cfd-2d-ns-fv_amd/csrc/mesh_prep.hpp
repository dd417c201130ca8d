// mesh_prep.hpp — host-side mesh preparation for the device solver.
//
// 1. Geometry of the reference reader (src/base/readmesh.cpp:113-320), computed
//    in the original numbering with the reference's expression order, so every
//    value is bit-identical to what readmesh stores in class_mesh.
// 2. Device layout: cells renumbered by reverse Cuthill-McKee so neighbour
//    gathers stay within a narrow index band, and a sliced-ELL "face slot"
//    table: chunk of 64 consecutive cells (one wavefront), slot j of lane l at
//    chunk_off[chunk] + 64*j + l.  A cell's slots list its faces in increasing
//    ORIGINAL face index, which is the order the reference's face loops add
//    into each cell (res.cpp:74-102, cell_gradient.cpp:93-118,
//    ssp_rk3.cpp:28-36), so per-cell gathers reproduce its sums bit for bit.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fvk.h"

namespace fvk {

constexpr int kChunk = 64;

// slot code: interior  = (neighbour_new_index << 1) | (this cell owns the face)
//            boundary  = ~((ghost_source_cell << 2) | kind), kind 0..3 =
//                        FREESTREAM, SLIP_WALL, EXTRAPOLATION, NO_SLIP_WALL
inline int32_t slot_code_interior(int32_t nb, bool owner) { return (nb << 1) | (owner ? 1 : 0); }
inline int32_t slot_code_boundary(int32_t src, int kind) { return ~((src << 2) | kind); }

struct HostMesh {
  int32_t np = 0, nf = 0, nc = 0, nb = 0;
  int32_t ncp = 0;          // cells padded to a multiple of kChunk
  int32_t max_deg = 0;
  int32_t num_bc1 = 0, num_bc2 = 0, num_wall = 0;
  // original numbering (class_mesh fields)
  std::vector<int32_t> fc1, fc2;      // fc2 after BC-kind substitution
  std::vector<int32_t> cf_ptr, cf_idx;
  std::vector<double> area, nx, ny, fcx, fcy;
  std::vector<double> vol, ccx, ccy, ixx, iyy, ixy;
  // renumbering
  std::vector<int32_t> perm;   // new -> old
  std::vector<int32_t> iperm;  // old -> new
  // device layout (new numbering, padded)
  std::vector<int32_t> chunk_off;     // nchunks + 1, in slots
  std::vector<uint8_t> deg;           // ncp
  std::vector<int32_t> code;          // per slot
  std::vector<double> s_nx, s_ny, s_area, s_fcx, s_fcy;   // per slot
  std::vector<double> c_ccx, c_ccy, c_vol, c_ixx, c_iyy, c_ixy;  // ncp
  std::string warnings;
};

// Returns FVK_OK or an FVK_ERR_* code with *err set.
int prepare_mesh(const fvk_mesh& m, const fvk_options& o, HostMesh* hm, std::string* err);

}  // namespace fvk
