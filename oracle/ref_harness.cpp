// ref_harness.cpp — test-only driver around the UNMODIFIED reference solver.
//
// This file is test infrastructure (the oracle side), never part of the product.
// It is compiled by oracle/Makefile together with the reference's own sources,
// read in place from the reference tree (nothing is copied into this repository),
// into oracle/_ref/ref_harness.  It replaces only the reference's main()
// (src/base/main.cpp:33-140) so that the state can be dumped at full precision:
// the reference's own writers print 6 significant digits
// (src/base/post_process.cpp:22, src/base/save_sol.cpp:198).
//
// Usage:
//   ref_harness <input.in> <out.bin> <nsteps> [--state S.bin] [--dump a,b,c]
//               [--res-only] [--time]
//   --state S.bin  start from the 4*N_c doubles in S.bin (p1..p4 blocks), with
//                  the reference's restart semantics (main.cpp:82-106) but at
//                  full precision: Qbar.init(N,gamma,R), p set, updateflow().
//   --dump list    steps after which the state is written (default: nsteps).
//   --res-only     call res() once on the initial state and dump its outputs.
//   --time         print "REF_STEP_SECONDS <s>" for the step loop only.
//
// Output: "FVDUMP01" then records {u32 name_len, name, u8 kind ('d'|'i'),
// u64 count, payload}.
#include <chrono>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "src/support/class_mesh.hpp"
#include "src/support/class_q.hpp"
#include "src/support/class_f.hpp"
#include "src/support/class_flow.hpp"
#include "src/support/class_residual.hpp"
#include "src/support/struct_BC.hpp"
#include "src/support/struct_size.hpp"
#include "src/support/struct_inputs.hpp"
#include "src/support/struct_time.hpp"
#include "src/support/struct_report.hpp"

#include "src/base/read_inputs.cpp"
#include "src/base/readmesh.cpp"

#include "src/include/ssp_rk3.hpp"
#include "src/include/res.hpp"

static FILE* g_out = nullptr;

static void put(const std::string& name, const std::vector<double>& v) {
  uint32_t n = (uint32_t)name.size();
  fwrite(&n, 4, 1, g_out);
  fwrite(name.data(), 1, n, g_out);
  uint8_t k = 'd';
  fwrite(&k, 1, 1, g_out);
  uint64_t c = v.size();
  fwrite(&c, 8, 1, g_out);
  if (c) fwrite(v.data(), 8, c, g_out);
}

static void put_i(const std::string& name, const std::vector<int>& v) {
  uint32_t n = (uint32_t)name.size();
  fwrite(&n, 4, 1, g_out);
  fwrite(name.data(), 1, n, g_out);
  uint8_t k = 'i';
  fwrite(&k, 1, 1, g_out);
  uint64_t c = v.size();
  fwrite(&c, 8, 1, g_out);
  if (c) fwrite(v.data(), 4, c, g_out);
}

static void dump_state(const std::string& pre, class_Q& Qbar, class_residual& residual,
                       class_Q& Qface, class_Q& Qc1, class_Q& Qc2, int eqn) {
  put(pre + "p1", Qbar.p1); put(pre + "p2", Qbar.p2);
  put(pre + "p3", Qbar.p3); put(pre + "p4", Qbar.p4);
  put(pre + "rho", Qbar.rho); put(pre + "u", Qbar.u);
  put(pre + "v", Qbar.v); put(pre + "P", Qbar.P);
  put(pre + "Qxp1", Qbar.Qxp1); put(pre + "Qxp2", Qbar.Qxp2);
  put(pre + "Qxp3", Qbar.Qxp3); put(pre + "Qxp4", Qbar.Qxp4);
  put(pre + "Qyp1", Qbar.Qyp1); put(pre + "Qyp2", Qbar.Qyp2);
  put(pre + "Qyp3", Qbar.Qyp3); put(pre + "Qyp4", Qbar.Qyp4);
  put(pre + "R1", residual.p1); put(pre + "R2", residual.p2);
  put(pre + "R3", residual.p3); put(pre + "R4", residual.p4);
  put(pre + "Vn_avg", Qc1.Vn_avg); put(pre + "c_avg", Qc1.c_avg);
  put(pre + "fL1", Qc1.p1); put(pre + "fL2", Qc1.p2);
  put(pre + "fL3", Qc1.p3); put(pre + "fL4", Qc1.p4);
  put(pre + "fR1", Qc2.p1); put(pre + "fR2", Qc2.p2);
  put(pre + "fR3", Qc2.p3); put(pre + "fR4", Qc2.p4);
  if (eqn == 2) {
    put(pre + "tau_xx", Qface.tau_xx); put(pre + "tau_xy", Qface.tau_xy);
    put(pre + "tau_yy", Qface.tau_yy); put(pre + "dTdx", Qface.dTdx);
    put(pre + "dTdy", Qface.dTdy);
  }
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s input.in out.bin nsteps [--state S.bin] [--dump a,b] [--res-only] [--time]\n", argv[0]);
    return 2;
  }
  std::string input = argv[1], outpath = argv[2];
  int nsteps = atoi(argv[3]);
  std::string state_path;
  std::set<int> dumps;
  bool res_only = false, timeit = false;
  for (int i = 4; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--state" && i + 1 < argc) state_path = argv[++i];
    else if (a == "--dump" && i + 1 < argc) {
      std::string s = argv[++i];
      size_t p = 0;
      while (p < s.size()) {
        size_t q = s.find(',', p);
        if (q == std::string::npos) q = s.size();
        dumps.insert(atoi(s.substr(p, q - p).c_str()));
        p = q + 1;
      }
    } else if (a == "--res-only") res_only = true;
    else if (a == "--time") timeit = true;
  }
  if (dumps.empty()) dumps.insert(nsteps);

  class_mesh mesh;
  class_Q Qbar, Qface, Qface_c1, Qface_c2;
  class_flow freestream;
  class_residual residual;
  struct_size size;
  struct_inputs inputs;
  struct_BC BC;
  struct_time time;
  struct_report report;

  char* fake_argv[2] = {argv[0], (char*)input.c_str()};
  read_inputs(&freestream, &inputs, &BC, &time, &report, 2, fake_argv);
  readmesh(&mesh, inputs.grid_file, &size, &BC, &inputs);

  // Initialisation sequence of src/base/main.cpp:77-118.
  freestream.update_vis();
  freestream.updateQ();
  freestream.free_comp();
  residual.init(size.num_cells);
  if (!state_path.empty()) {
    std::vector<double> s(4 * (size_t)size.num_cells);
    FILE* f = fopen(state_path.c_str(), "rb");
    if (!f || fread(s.data(), 8, s.size(), f) != s.size()) {
      fprintf(stderr, "cannot read state %s\n", state_path.c_str());
      return 3;
    }
    fclose(f);
    Qbar.init(size.num_cells, freestream.gamma, freestream.R);
    for (int i = 0; i < size.num_cells; ++i) {
      Qbar.p1[i] = s[i];
      Qbar.p2[i] = s[size.num_cells + i];
      Qbar.p3[i] = s[2 * (size_t)size.num_cells + i];
      Qbar.p4[i] = s[3 * (size_t)size.num_cells + i];
    }
    Qbar.updateflow();
  } else {
    Qbar.init(size.num_cells, &freestream);
    Qbar.updateQ();
  }
  if (inputs.eqn == 2) Qface.init(size.num_faces, &freestream);
  Qface_c1.init(size.num_faces, &freestream);
  Qface_c2.init(size.num_faces, &freestream);

  g_out = fopen(outpath.c_str(), "wb");
  if (!g_out) { fprintf(stderr, "cannot open %s\n", outpath.c_str()); return 3; }
  fwrite("FVDUMP01", 1, 8, g_out);

  put("geom/face_area", mesh.face_area);
  put("geom/face_nx", mesh.face_nx);
  put("geom/face_ny", mesh.face_ny);
  put("geom/face_centerx", mesh.face_centerx);
  put("geom/face_centery", mesh.face_centery);
  put("geom/cell_vol", mesh.cell_vol);
  put("geom/cell_centerx", mesh.cell_centerx);
  put("geom/cell_centery", mesh.cell_centery);
  put("geom/face_cell1_dx", mesh.face_cell1_dx);
  put("geom/face_cell1_dy", mesh.face_cell1_dy);
  put("geom/face_cell2_dx", mesh.face_cell2_dx);
  put("geom/face_cell2_dy", mesh.face_cell2_dy);
  put("geom/face_dxj", mesh.face_dxj);
  put("geom/face_dyj", mesh.face_dyj);
  put("geom/Ixx", mesh.Ixx);
  put("geom/Iyy", mesh.Iyy);
  put("geom/Ixy", mesh.Ixy);
  put("geom/face_dist", mesh.face_dist);
  put("geom/face_lx", mesh.face_lx);
  put("geom/face_ly", mesh.face_ly);
  put("geom/face_mx", mesh.face_mx);
  put("geom/face_my", mesh.face_my);
  put("geom/BC_cell_centerx", mesh.BC_cell_centerx);
  put("geom/BC_cell_centery", mesh.BC_cell_centery);
  put_i("geom/face_cell2", mesh.face_cell2);

  if (res_only) {
    res(&residual, &mesh, &size, &inputs, &Qbar, &Qface, &Qface_c1, &Qface_c2, &freestream, &BC);
    dump_state("res/", Qbar, residual, Qface, Qface_c1, Qface_c2, inputs.eqn);
    fclose(g_out);
    return 0;
  }

  double secs = 0.0;
  for (int ndx = 1; ndx <= nsteps; ++ndx) {
    auto t0 = std::chrono::steady_clock::now();
    ssp_rk3(&mesh, &Qbar, &Qface, &Qface_c1, &Qface_c2, &residual, &size, &inputs, &freestream, &BC, &time);
    auto t1 = std::chrono::steady_clock::now();
    secs += std::chrono::duration<double>(t1 - t0).count();
    if (dumps.count(ndx))
      dump_state("step" + std::to_string(ndx) + "/", Qbar, residual, Qface, Qface_c1, Qface_c2, inputs.eqn);
  }
  fclose(g_out);
  if (timeit) printf("REF_STEP_SECONDS %.9f\n", secs);
  return 0;
}
