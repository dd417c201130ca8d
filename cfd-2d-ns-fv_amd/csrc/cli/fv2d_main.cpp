// fv2d — drop-in command line program: `fv2d [input.in]`.
//
// `fv2d --cache-grid <grid file>` only parses the grid file and writes its
// binary cache `<grid file>.fvkgrid` (used by runs whose input.in sets
// `grid_cache = 1` in the [gpu] section); it needs no GPU.
//
// Same flow as the reference program (src/base/main.cpp:33-140): read
// input.in and the grid, write the output headers, start from freestream or
// sol/restart_sol.dat, run nmax SSP-RK3 iterations (here on the GPU through
// libfvk), monitor every monitor_step, write Tecplot files every output_step
// and the restart file at the end.  Output files have the reference's names
// and formats.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "case_io.hpp"
#include "fvk.h"

int main(int argc, char** argv) {
  if (argc == 3 && std::string(argv[1]) == "--cache-grid") {
    fv2d::Grid grid;
    std::string log, err;
    if (!fv2d::read_grid(argv[2], &grid, &log, &err, true)) {
      std::cout << err << std::endl;
      return 1;
    }
    std::cout << log;
    return 0;
  }
  ::mkdir("sol", 0755);
  ::mkdir("mesh", 0755);
  fv2d::Inputs in;
  std::string log, err;
  const std::string input = (argc == 2) ? argv[1] : "input.in";
  if (!fv2d::read_inputs(input, &in, &log, &err)) {
    std::cout << log << err << std::endl;
    return 1;
  }
  std::cout << log;
  fv2d::Grid grid;
  if (!fv2d::read_grid(in.grid_file, &grid, &log, &err, in.grid_cache != 0)) {
    std::cout << err << std::endl;
    return 1;
  }
  std::cout << log;
  if (in.restart == 0) {
    std::ofstream rh("sol/res_history.dat");
    if (rh.is_open()) rh << "#iter, res1, res2, res3, res4" << std::endl;
    std::ofstream out("sol/output.dat");
    if (out.is_open()) out << "#iter, CL, CD" << std::endl;
    std::ofstream fo("sol/forces.dat");
    if (fo.is_open()) fo << "#iter, CL, CD (wall integral at the true face centres)" << std::endl;
  }
  const fvk_mesh mesh = fv2d::as_mesh(grid);
  const fvk_options opt = fv2d::as_options(in);
  fvk_ctx* ctx = nullptr;
  int rc = fvk_create(&mesh, &opt, &ctx);
  if (rc != FVK_OK) {
    std::cout << "ERROR: " << fvk_last_error(nullptr) << std::endl;
    return 1;
  }
  fv2d::Geometry geo;
  if (!fv2d::load_geometry(ctx, grid, in, &geo, &err)) {
    std::cout << "ERROR: " << err << std::endl;
    fvk_destroy(ctx);
    return 1;
  }
  fv2d::write_mesh_dumps(grid, geo);

  int ndx = 1;
  if (in.restart == 1) {
    std::vector<double> q;
    std::string rlog;
    if (fv2d::read_restart(grid.nc, &ndx, &q, &rlog)) {
      std::cout << rlog;
      rc = fvk_set_state(ctx, q.data());
      ++ndx;
    } else {
      std::cout << "ERROR: Cannot Open Restart File\n";
      rc = fvk_init_freestream(ctx);
    }
  } else {
    rc = fvk_init_freestream(ctx);
  }
  if (rc != FVK_OK) {
    std::cout << "ERROR: " << fvk_last_error(ctx) << std::endl;
    fvk_destroy(ctx);
    return 1;
  }

  const size_t nc = grid.nc;
  std::vector<double> q(4 * nc), grad(8 * nc), r(4 * nc);
  while (ndx <= in.nmax) {
    // advance to the next iteration that monitors or writes output
    int stop = in.nmax;
    const int mnext = (ndx + in.monitor_step - 1) / in.monitor_step * in.monitor_step;
    const int onext = (ndx + in.output_step - 1) / in.output_step * in.output_step;
    if (mnext < stop) stop = mnext;
    if (onext < stop) stop = onext;
    rc = fvk_step(ctx, stop - ndx + 1);
    if (rc != FVK_OK) {
      std::cout << "ERROR: " << fvk_last_error(ctx) << std::endl;
      fvk_destroy(ctx);
      return 1;
    }
    ndx = stop;
    const bool mon = (ndx % in.monitor_step) == 0, outp = (ndx % in.output_step) == 0;
    if (mon || outp) fvk_get_state(ctx, q.data(), grad.data());
    if (mon) {
      // residual norms summed in cell order on the host, as save_sol.cpp:79-89
      fvk_get_residual(ctx, r.data());
      double res[4];
      for (int k = 0; k < 4; ++k) {
        double s = 0;
        for (size_t i = 0; i < nc; ++i) s += r[k * nc + i] * r[k * nc + i];
        res[k] = std::sqrt(s);
      }
      double t = -1.0;   // physical time: uniform time step only
      if (in.dt_type == 3) fvk_get_time(ctx, &t);
      fv2d::monitor(ndx, grid, geo, in, q.data(), grad.data(), res, t);
      // the reference carries on through NaNs; the drop-in stops after the
      // monitor line that shows them
      if (!(std::isfinite(res[0]) && std::isfinite(res[1]) && std::isfinite(res[2]) &&
            std::isfinite(res[3]))) {
        std::cout << "ERROR: non-finite residual norm at iteration " << ndx << std::endl;
        fvk_destroy(ctx);
        return 3;
      }
    }
    if (outp) fv2d::write_tecplot(ndx, grid, geo, in, q.data(), grad.data());
    ++ndx;
  }
  fvk_get_state(ctx, q.data(), grad.data());
  fv2d::write_restart(in.nmax, grid.nc, q.data());
  fvk_destroy(ctx);
  return 0;
}
