// case_io.hpp — the drop-in CLI's file formats: input.in, the native grid
// file, and the monitor/solution/restart writers of the reference program.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fvk.h"

namespace fv2d {

// struct_inputs + struct_time + struct_BC + class_flow + struct_report
// (src/support/*.hpp) as read_inputs.cpp fills them.
struct Inputs {
  int order = -1, nmax = 0, monitor_step = 1, output_step = 1, restart = 0;
  std::string grid_file;
  int eqn = -1, flux_solver = -1, limiter = -1;
  double gamma = 1.4, R = 287.0, mu = 0.0, Pr = 0.72;
  double P = 0, rho = 0, u = 0, v = 0;
  int scheme = -1, dt_type = -1;
  double CFL = 0;
  double length = 1.0;
  int BC1 = 0, BC2 = 0;
  // optional trailing [gpu] section (read only after the nine standard ones)
  int renumber = 1, block_size = 0, xcd_remap = -1, kernel_variant = 0;
  int grid_cache = 0;          // [gpu] grid_cache = 1: binary grid cache beside the grid file
  int accumulate = 0;          // [gpu] accumulate = 1: atomic-scatter residual accumulation
  int precision = FVK_FP64;    // [gpu] precision = FP64 | FP32_MIXED
};

// read_inputs.cpp:10-219; returns false with *err on a missing/invalid file or
// an option the solver needs that was never set.
bool read_inputs(const std::string& path, Inputs* in, std::string* log, std::string* err);

struct Grid {
  int32_t np = 0, nf = 0, nc = 0;
  std::vector<double> x, y;
  std::vector<int32_t> p1, p2, c1, c2;  // 0-based; c2 = file value - 1 (readmesh.cpp:59-69)
};

// readmesh.cpp:31-85 token stream, parsed by several threads.  With `cache`,
// a binary copy `<path>.fvkgrid` is read instead while the text file keeps its
// size and modification time, and written after a text parse.
bool read_grid(const std::string& path, Grid* g, std::string* log, std::string* err,
               bool cache = false);

fvk_mesh as_mesh(const Grid& g);
fvk_options as_options(const Inputs& in);

// Everything save()/post() need on the host, filled once after fvk_create.
struct Geometry {
  std::vector<double> area, nx, ny, fcx, fcy, ccx, ccy, vol;
  std::vector<int32_t> fc2kind;               // face_cell2 after BC-kind substitution
  std::vector<int32_t> cf_ptr, cf_idx;        // faces of each cell, increasing index
  std::vector<int32_t> bc_faces;              // boundary faces in file order
  int num_wall = 0;
};
bool load_geometry(fvk_ctx* ctx, const Grid& g, const Inputs& in, Geometry* geo, std::string* err);

// mesh/cell_data.dat and mesh/face_data.dat (readmesh.cpp:416-445)
void write_mesh_dumps(const Grid& g, const Geometry& geo);

// save_sol.cpp:16-103 monitor block: CL, CD and the residual norms, appended
// to sol/output.dat and sol/res_history.dat and echoed to stdout; then this
// build's own lift and drag coefficients (wall_forces), appended to
// sol/forces.dat with 10 significant digits, followed there by the physical
// time t when t >= 0 (uniform time step).
void monitor(int ndx, const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
             const double* grad, const double res[4], double t = -1.0);
// Lift and drag coefficients of the wall group (BC1 if it is a wall, else
// BC2), per unit span referred to `length`.  Unlike save_sol.cpp:20-75,
// which truncates centroids and face centres to int, the wall state is the
// owner's state reconstructed to the true face centre with its limited
// gradient (grad may be null: first order), pressure acts along the outward
// normal, and no-slip walls add the shear mu * |u_c| / d_n along the
// centroid velocity (the wall gradient of reconstruction.cpp:110-123).
// q, grad: component-major as fvk_get_state returns them.
void wall_forces(const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
                 const double* grad, double* cl, double* cd);
// save_sol.cpp:105-234 Tecplot volume and surface files
void write_tecplot(int ndx, const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
                   const double* grad);
// post_process.cpp:13-28
void write_restart(int ndx, int nc, const double* q);
// main.cpp:82-104; returns false when the file is missing
bool read_restart(int nc, int* ndx, std::vector<double>* q, std::string* log);

}  // namespace fv2d
