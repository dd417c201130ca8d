// case_io.cpp — file formats of the drop-in CLI.  Outputs use the default
// ostream formatting (6 significant digits) exactly like the reference's
// writers, so the files compare line for line.
#include "case_io.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <thread>

namespace fv2d {

namespace {

bool next_tok(std::istream& is, std::string* t) { return static_cast<bool>(is >> *t); }

int bc_code(const std::string& s) {
  if (s == "FREESTREAM") return FVK_BC_FREESTREAM;
  if (s == "SLIP_WALL") return FVK_BC_SLIP_WALL;
  if (s == "EXTRAPOLATION") return FVK_BC_EXTRAPOLATION;
  if (s == "NO_SLIP_WALL") return FVK_BC_NO_SLIP_WALL;
  return 0;
}

}  // namespace

bool read_inputs(const std::string& path, Inputs* in, std::string* log, std::string* err) {
  std::ifstream f(path);
  if (!f.is_open()) {
    *err = "ERROR: Cannot Open Input File";
    return false;
  }
  std::ostringstream o;
  o << "\nReading input.in \n";
  std::string head, key, eq, val;
  auto keyval = [&]() {
    key.clear(); val.clear();
    return next_tok(f, &key) && next_tok(f, &eq) && next_tok(f, &val);
  };
  // The reference reads exactly nine section headers with a fixed number of
  // "key = value" lines each (read_inputs.cpp:20-209); a trailing optional
  // [gpu] section may follow, which the reference never reads.
  for (int sec = 0; sec < 10 && next_tok(f, &head); ++sec) {
    if (head == "[case]") {
      for (int i = 0; i < 5 && keyval(); ++i) {
        if (key == "solution_poly_order") in->order = std::stoi(val) + 1;
        if (key == "nmax") in->nmax = std::stoi(val);
        if (key == "monitor_step") in->monitor_step = std::stoi(val);
        if (key == "output_step") in->output_step = std::stoi(val);
        if (key == "restart") in->restart = std::stoi(val);
      }
      o << "[case] \n  nmax = " << in->nmax << "\n  monitor_step = " << in->monitor_step
        << "\n  output_step = " << in->output_step << "\n  solution_poly_order = " << in->order - 1
        << "\n  restart = " << in->restart << "\n";
    } else if (head == "[limiter]") {
      o << "[limiter] \n";
      if (keyval() && key == "limiter") {
        if (val == "NONE") in->limiter = 0;
        if (val == "MINMOD") in->limiter = 1;
        if (val == "SQUEEZE") in->limiter = 2;
        o << "  limiter = " << val << " \n";
      }
    } else if (head == "[grid]") {
      if (keyval() && key == "grid_file") in->grid_file = val;
      o << "[grid] \n  grid_file = " << in->grid_file << "\n";
    } else if (head == "[equations]") {
      o << "[equations] \n";
      for (int i = 0; i < 2 && keyval(); ++i) {
        if (key == "eqn") {
          if (val == "EULER") in->eqn = 1;
          if (val == "NS") in->eqn = 2;
          o << "  eqn = " << val << " \n";
        }
        if (key == "flux_solver") {
          if (val == "RUSANOV") in->flux_solver = 1;
          if (val == "ROE") in->flux_solver = 2;
          o << "  flux_solver = " << val << " \n";
        }
      }
    } else if (head == "[fluid]") {
      for (int i = 0; i < 4 && keyval(); ++i) {
        if (key == "gamma") in->gamma = std::stod(val);
        if (key == "R") in->R = std::stod(val);
        if (key == "mu") in->mu = std::stod(val);
        if (key == "Pr") in->Pr = std::stod(val);
      }
      o << "[fluid] \n  gamma = " << in->gamma << "\n  R = " << in->R << "\n  mu = " << in->mu
        << "\n  Pr = " << in->Pr << "\n";
    } else if (head == "[init]") {
      for (int i = 0; i < 4 && keyval(); ++i) {
        if (key == "P") in->P = std::stod(val);
        if (key == "rho") in->rho = std::stod(val);
        if (key == "u") in->u = std::stod(val);
        if (key == "v") in->v = std::stod(val);
      }
      o << "[init] \n  P = " << in->P << "\n  rho = " << in->rho << "\n  u = " << in->u
        << "\n  v = " << in->v << "\n";
    } else if (head == "[time]") {
      o << "[time] \n";
      for (int i = 0; i < 3 && keyval(); ++i) {
        if (key == "CFL") in->CFL = std::stod(val);
        if (key == "time_step_type") {
          if (val == "GLOBAL") in->dt_type = 1;
          if (val == "LOCAL") in->dt_type = 2;
          if (val == "UNIFORM") in->dt_type = 3;   // extension: one dt for every cell
          o << "  time_step_type = " << val << " \n";
        }
        if (key == "time_scheme") {
          if (val == "SSP_RK2") in->scheme = 1;
          if (val == "SSP_RK3") in->scheme = 2;
          o << "  scheme = " << val << " \n";
        }
      }
      o << "  CFL = " << in->CFL << "\n";
    } else if (head == "[report]") {
      if (keyval() && key == "length") in->length = std::stod(val);
      o << "[report] \n  length = " << in->length << "\n";
    } else if (head == "[BC]") {
      o << "[BC] \n";
      for (int i = 0; i < 2 && keyval(); ++i) {
        if (key == "BC1") in->BC1 = bc_code(val);
        if (key == "BC2") in->BC2 = bc_code(val);
        o << "  " << key << " = " << val << " \n";
      }
    } else if (head == "[gpu]" && sec == 9) {
      while (keyval()) {
        if (key == "renumber") in->renumber = std::stoi(val);
        if (key == "block_size") in->block_size = std::stoi(val);
        if (key == "xcd_remap") in->xcd_remap = std::stoi(val);
        if (key == "kernel_variant") in->kernel_variant = std::stoi(val);
        if (key == "grid_cache") in->grid_cache = std::stoi(val);
        if (key == "accumulate") in->accumulate = std::stoi(val);
        if (key == "precision") in->precision = (val == "FP32_MIXED") ? FVK_FP32_MIXED : (val == "FP64" ? FVK_FP64 : -1);
      }
    }
  }
  *log = o.str();
  if (in->order < 1 || in->eqn < 0 || in->flux_solver < 0 || in->limiter < 0 || in->scheme < 0 ||
      in->dt_type < 0 || in->BC1 == 0 || in->BC2 == 0 || in->grid_file.empty()) {
    *err = "ERROR: input.in is missing one of solution_poly_order, eqn, flux_solver, limiter, "
           "time_scheme, time_step_type, BC1, BC2 or grid_file";
    return false;
  }
  if (in->monitor_step <= 0 || in->output_step <= 0) {
    *err = "ERROR: monitor_step and output_step must be positive";
    return false;
  }
  return true;
}

namespace {

inline bool is_space(char c) {   // the characters strtod/strtol skip (isspace in the C locale)
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

int reader_threads() {
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(16, hw));
}

// The whitespace-separated token stream of readmesh.cpp:31-85 after the
// header: 2*np coordinates, then 2*nf point ids, then 2*nf cell ids, then
// anything (ignored, as the reference ignores trailing tokens).  The text is
// cut into one range per thread at token boundaries; a first pass counts each
// range's tokens, so every thread knows the index of its first token and
// parses its range straight into the arrays.  Each value is the one
// strtod/strtol gives for its token, as the serial reader's.
bool parse_body(const char* p, const char* end, Grid* g, std::string* err) {
  const int nt = (end - p > (1 << 22)) ? reader_threads() : 1;
  std::vector<const char*> cut(nt + 1);
  cut[0] = p;
  cut[nt] = end;
  for (int t = 1; t < nt; ++t) {
    const char* q = p + (end - p) * t / nt;
    while (q < end && !is_space(*q)) ++q;   // move the cut to the end of a token
    cut[t] = std::max(q, cut[t - 1]);
  }
  std::vector<int64_t> count(nt + 1, 0);
  auto count_range = [&](int t) {
    int64_t n = 0;
    bool in_tok = false;
    for (const char* q = cut[t]; q < cut[t + 1]; ++q) {
      const bool sp = is_space(*q);
      if (!sp && !in_tok) ++n;
      in_tok = !sp;
    }
    count[t + 1] = n;
  };
  const int64_t need_pts = 2 * (int64_t)g->np, need = need_pts + 4 * (int64_t)g->nf;
  std::vector<int> bad(nt, 0);
  auto parse_range = [&](int t, int64_t k) {
    const char* q = cut[t];
    const char* e = cut[t + 1];
    while (k < need) {
      while (q < e && is_space(*q)) ++q;
      if (q >= e) break;
      char* stop = nullptr;
      if (k < need_pts) {
        const double v = std::strtod(q, &stop);
        (k & 1 ? g->y : g->x)[k >> 1] = v;
      } else {
        const long v = std::strtol(q, &stop, 10);
        const int64_t j = k - need_pts;
        const int64_t i = (j % (2 * (int64_t)g->nf)) >> 1;
        if (j < 2 * (int64_t)g->nf) (j & 1 ? g->p2 : g->p1)[i] = (int32_t)v - 1;
        else (j & 1 ? g->c2 : g->c1)[i] = (int32_t)v - 1;
      }
      if (stop == q || (stop < e && !is_space(*stop))) { bad[t] = 1; return; }
      q = stop;
      ++k;
    }
  };
  {
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(count_range, t);
    count_range(0);
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < nt; ++t) count[t + 1] += count[t];
  if (count[nt] < need) {
    *err = "ERROR: grid file ended early";
    return false;
  }
  {
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(parse_range, t, count[t]);
    parse_range(0, 0);
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < nt; ++t)
    if (bad[t]) {
      *err = "ERROR: grid file holds a token that is not a number";
      return false;
    }
  return true;
}

bool read_grid_text(const std::string& path, Grid* g, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *err = "ERROR: Cannot Open Grid File";
    return false;
  }
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf((size_t)len + 1, '\0');
  const size_t got = std::fread(&buf[0], 1, (size_t)len, f);
  std::fclose(f);
  buf[got] = '\0';
  char* p = &buf[0];
  bool ok = true;
  auto rd_l = [&]() -> long {
    char* e = nullptr;
    long v = std::strtol(p, &e, 10);
    if (e == p) ok = false;
    p = e;
    return v;
  };
  g->np = (int32_t)rd_l();
  g->nf = (int32_t)rd_l();
  g->nc = (int32_t)rd_l();
  if (!ok || g->np <= 0 || g->nf <= 0 || g->nc <= 0) {
    *err = "ERROR: bad grid header";
    return false;
  }
  g->x.resize(g->np); g->y.resize(g->np);
  g->p1.resize(g->nf); g->p2.resize(g->nf); g->c1.resize(g->nf); g->c2.resize(g->nf);
  return parse_body(p, &buf[0] + got, g, err);
}

// Binary cache of a parsed grid: header, then x, y, p1, p2, c1, c2 as read.
// It is used only while the text file keeps the size and modification time
// recorded in it.
struct CacheHeader {
  char magic[8];
  int64_t text_size;
  int64_t text_mtime_ns;
  int32_t np, nf, nc, pad;
};

bool text_stamp(const std::string& path, int64_t* size, int64_t* mtime_ns) {
  struct stat st;
  if (::stat(path.c_str(), &st) != 0) return false;
  *size = (int64_t)st.st_size;
  *mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000 + st.st_mtim.tv_nsec;
  return true;
}

bool load_cache(const std::string& path, const std::string& cpath, Grid* g) {
  int64_t size, mtime;
  if (!text_stamp(path, &size, &mtime)) return false;
  FILE* f = std::fopen(cpath.c_str(), "rb");
  if (!f) return false;
  CacheHeader h;
  bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, "FVKGRID1", 8) == 0 &&
            h.text_size == size && h.text_mtime_ns == mtime && h.np > 0 && h.nf > 0 && h.nc > 0;
  if (ok) {
    g->np = h.np; g->nf = h.nf; g->nc = h.nc;
    g->x.resize(h.np); g->y.resize(h.np);
    g->p1.resize(h.nf); g->p2.resize(h.nf); g->c1.resize(h.nf); g->c2.resize(h.nf);
    ok = std::fread(g->x.data(), sizeof(double), h.np, f) == (size_t)h.np &&
         std::fread(g->y.data(), sizeof(double), h.np, f) == (size_t)h.np;
    for (std::vector<int32_t>* v : {&g->p1, &g->p2, &g->c1, &g->c2})
      ok = ok && std::fread(v->data(), sizeof(int32_t), h.nf, f) == (size_t)h.nf;
  }
  std::fclose(f);
  return ok;
}

void save_cache(const std::string& path, const std::string& cpath, const Grid& g) {
  CacheHeader h{};
  std::memcpy(h.magic, "FVKGRID1", 8);
  if (!text_stamp(path, &h.text_size, &h.text_mtime_ns)) return;
  h.np = g.np; h.nf = g.nf; h.nc = g.nc;
  const std::string tmp = cpath + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
            std::fwrite(g.x.data(), sizeof(double), g.np, f) == (size_t)g.np &&
            std::fwrite(g.y.data(), sizeof(double), g.np, f) == (size_t)g.np;
  for (const std::vector<int32_t>* v : {&g.p1, &g.p2, &g.c1, &g.c2})
    ok = ok && std::fwrite(v->data(), sizeof(int32_t), g.nf, f) == (size_t)g.nf;
  ok = (std::fclose(f) == 0) && ok;
  if (ok) ok = std::rename(tmp.c_str(), cpath.c_str()) == 0;
  if (!ok) std::remove(tmp.c_str());
}

}  // namespace

bool read_grid(const std::string& path, Grid* g, std::string* log, std::string* err, bool cache) {
  const std::string cpath = path + ".fvkgrid";
  const bool cached = cache && load_cache(path, cpath, g);
  if (!cached) {
    if (!read_grid_text(path, g, err)) return false;
    if (cache) save_cache(path, cpath, *g);
  }
  std::vector<int32_t> cnt(g->nc, 0);
  for (int32_t i = 0; i < g->nf; ++i) {
    if (g->c1[i] >= 0 && g->c1[i] < g->nc) cnt[g->c1[i]]++;
    if (g->c2[i] >= 0 && g->c2[i] < g->nc) cnt[g->c2[i]]++;
  }
  long tri = 0, quad = 0, pent = 0, hex = 0, other = 0;
  for (int32_t c = 0; c < g->nc; ++c) {
    if (cnt[c] == 3) ++tri;
    else if (cnt[c] == 4) ++quad;
    else if (cnt[c] == 5) ++pent;
    else if (cnt[c] == 6) ++hex;
    else ++other;
  }
  std::ostringstream o;
  o << "Reading " << path << "\n" << g->np << " Points \n" << g->nf << " Faces \n" << g->nc
    << " Cells \n";
  o << "Elements: \n  Triangles: " << tri << "\n  Quads: " << quad << "\n  Pentagons: " << pent
    << "\n  Hexagons:" << hex << "\n";
  if (other) o << "  Other polygons: " << other << "\n";
  *log = o.str();
  return true;
}

fvk_mesh as_mesh(const Grid& g) {
  fvk_mesh m{};
  m.num_points = g.np;
  m.num_faces = g.nf;
  m.num_cells = g.nc;
  m.x = g.x.data();
  m.y = g.y.data();
  m.face_point1 = g.p1.data();
  m.face_point2 = g.p2.data();
  m.face_cell1 = g.c1.data();
  m.face_cell2 = g.c2.data();
  return m;
}

fvk_options as_options(const Inputs& in) {
  fvk_options o;
  fvk_default_options(&o);
  o.order = in.order;
  o.limiter = in.limiter;
  o.eqn = in.eqn;
  o.flux_solver = in.flux_solver;
  o.time_scheme = in.scheme;
  o.time_step_type = in.dt_type;
  o.bc1 = in.BC1;
  o.bc2 = in.BC2;
  o.cfl = in.CFL;
  o.gamma = in.gamma;
  o.R = in.R;
  o.mu = in.mu;
  o.Pr = in.Pr;
  o.P = in.P;
  o.rho = in.rho;
  o.u = in.u;
  o.v = in.v;
  o.renumber = in.renumber;
  o.block_size = in.block_size;
  o.xcd_remap = in.xcd_remap;
  o.accumulate = in.accumulate;
  o.kernel_variant = in.kernel_variant;
  o.precision = in.precision;
  return o;
}

bool load_geometry(fvk_ctx* ctx, const Grid& g, const Inputs& in, Geometry* geo, std::string* err) {
  auto get = [&](const char* name, std::vector<double>& v, size_t n) {
    v.resize(n);
    return fvk_get_geometry(ctx, name, v.data()) == FVK_OK;
  };
  if (!get("face_area", geo->area, g.nf) || !get("face_nx", geo->nx, g.nf) ||
      !get("face_ny", geo->ny, g.nf) || !get("face_centerx", geo->fcx, g.nf) ||
      !get("face_centery", geo->fcy, g.nf) || !get("cell_centerx", geo->ccx, g.nc) ||
      !get("cell_centery", geo->ccy, g.nc) || !get("cell_vol", geo->vol, g.nc)) {
    *err = fvk_last_error(ctx);
    return false;
  }
  geo->fc2kind.resize(g.nf);
  geo->cf_ptr.assign(g.nc + 1, 0);
  int nbc1 = 0, nbc2 = 0;
  for (int32_t f = 0; f < g.nf; ++f) {
    const int32_t c2 = g.c2[f];
    geo->fc2kind[f] = c2 >= 0 ? c2 : (c2 == -1 ? in.BC1 : (c2 == -2 ? in.BC2 : c2));
    if (c2 == -1) ++nbc1;
    if (c2 == -2) ++nbc2;
    if (c2 < 0) geo->bc_faces.push_back(f);
    geo->cf_ptr[g.c1[f] + 1]++;
    if (c2 >= 0) geo->cf_ptr[c2 + 1]++;
  }
  for (int32_t c = 0; c < g.nc; ++c) geo->cf_ptr[c + 1] += geo->cf_ptr[c];
  geo->cf_idx.resize(geo->cf_ptr[g.nc]);
  std::vector<int32_t> fill(g.nc, 0);
  for (int32_t f = 0; f < g.nf; ++f) {
    geo->cf_idx[geo->cf_ptr[g.c1[f]] + fill[g.c1[f]]++] = f;
    if (g.c2[f] >= 0) geo->cf_idx[geo->cf_ptr[g.c2[f]] + fill[g.c2[f]]++] = f;
  }
  geo->num_wall = 0;  // readmesh.cpp:356-360
  if (in.BC1 == FVK_BC_SLIP_WALL || in.BC1 == FVK_BC_NO_SLIP_WALL) geo->num_wall = nbc1;
  else if (in.BC2 == FVK_BC_SLIP_WALL || in.BC2 == FVK_BC_NO_SLIP_WALL) geo->num_wall = nbc2;
  return true;
}

void write_mesh_dumps(const Grid& g, const Geometry& geo) {
  std::ofstream out("mesh/cell_data.dat");
  if (out.is_open()) {
    out << "# cell_vol, cell_centerx, cell_centery, x points for cell, y points for cell" << std::endl;
    std::vector<double> px, py;
    for (int32_t c = 0; c < g.nc; ++c) {
      // distinct vertices in the order readmesh.cpp:147-185 finds them
      const int32_t b = geo.cf_ptr[c], n = geo.cf_ptr[c + 1] - b;
      px.assign(n, 0.0);
      py.assign(n, 0.0);
      int32_t found = 0;
      auto add = [&](double x, double y) {
        for (int32_t z = 0; z < n; ++z)
          if (px[z] == x && py[z] == y) return;
        if (found < n) { px[found] = x; py[found] = y; ++found; }
      };
      for (int32_t j = 0; j < n; ++j) {
        const int32_t f = geo.cf_idx[b + j];
        add(g.x[g.p1[f]], g.y[g.p1[f]]);
        add(g.x[g.p2[f]], g.y[g.p2[f]]);
      }
      out << geo.vol[c] << " " << geo.ccx[c] << " " << geo.ccy[c] << " ";
      for (int j = 0; j < 6; ++j) out << (j < n ? px[j] : -101.0) << (j < 5 ? " " : " ");
      for (int j = 0; j < 6; ++j) out << (j < n ? py[j] : -101.0) << (j < 5 ? " " : "");
      out << std::endl;
    }
  } else {
    std::cout << "ERROR: Cannot Save File \n";
  }
  out.close();
  out.open("mesh/face_data.dat");
  if (out.is_open()) {
    out << "# face point 1 x, face point 1 y, face point 2 x, face point 2y, face nx, face ny, "
           "face center x, face center y"
        << std::endl;
    for (int32_t f = 0; f < g.nf; ++f) {
      out << g.x[g.p1[f]] << " " << g.y[g.p1[f]] << " ";
      out << g.x[g.p2[f]] << " " << g.y[g.p2[f]] << " ";
      out << geo.nx[f] << " " << geo.ny[f] << " ";
      out << geo.fcx[f] << " " << geo.fcy[f] << std::endl;
    }
  } else {
    std::cout << "ERROR: Cannot Save File \n";
  }
}

void monitor(int ndx, const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
             const double* grad, const double res[4], double t) {
  const int32_t nc = g.nc;
  const double* p[4] = {q, q + nc, q + 2 * (size_t)nc, q + 3 * (size_t)nc};
  const double* qx[4] = {grad, grad + nc, grad + 2 * (size_t)nc, grad + 3 * (size_t)nc};
  const double* qy[4] = {grad + 4 * (size_t)nc, grad + 5 * (size_t)nc, grad + 6 * (size_t)nc,
                         grad + 7 * (size_t)nc};
  // wall forces, save_sol.cpp:20-75: wall faces are taken to be faces
  // 0..numWALL-1 and the coordinates are truncated to int (save_sol.cpp:24-32)
  double Fx = 0, Fy = 0;
  for (int face = 0; face < geo.num_wall; ++face) {
    const int c1 = g.c1[face];
    const int cellx = (int)geo.ccx[c1], celly = (int)geo.ccy[c1];
    const int facex = (int)geo.fcx[face], facey = (int)geo.fcy[face];
    double fr[4];
    for (int k = 0; k < 4; ++k)
      fr[k] = p[k][c1] + ((facex - cellx) * qx[k][c1] + (facey - celly) * qy[k][c1]);
    const double rho = fr[0];
    const double u = fr[1] / rho;
    const double v = fr[2] / rho;
    const double E = fr[3];
    const double P = (E - (0.5 * rho * (u * u + v * v))) * (in.gamma - 1);
    if (in.eqn == 2) {
      const double dx = geo.fcx[face] - geo.ccx[c1], dy = geo.fcy[face] - geo.ccy[c1];
      const double dn = std::fabs(std::sqrt(dx * dx + dy * dy));
      const double cu = p[1][c1] / p[0][c1], cv = p[2][c1] / p[0][c1];
      const double dudn = -cu / dn;
      const double dvdn = -cv / dn;
      const double tx = in.mu * dudn;
      const double ty = in.mu * dvdn;
      const double tw = std::sqrt(tx * tx + ty * ty);
      if (geo.ny[face] < 0) {
        Fx += tw * geo.area[face] * (-geo.ny[face]);
        Fy += tw * geo.area[face] * (geo.nx[face]);
      } else {
        Fx += tw * geo.area[face] * (geo.ny[face]);
        Fy += tw * geo.area[face] * (-geo.nx[face]);
      }
    }
    const double force = P * geo.area[face];
    Fy += force * geo.ny[face];
    Fx += force * geo.nx[face];
  }
  const double nrm = std::sqrt(in.u * in.u + in.v * in.v);
  const double flowx = in.u / nrm, flowy = in.v / nrm;
  const double L = Fx * (-flowy) + Fy * flowx;
  const double D = Fy * flowy + Fx * flowx;
  const double Vmag = std::sqrt(in.u * in.u + in.v * in.v);
  const double CL = L / (0.5 * in.rho * Vmag * Vmag * in.length);
  const double CD = D / (0.5 * in.rho * Vmag * Vmag * in.length);
  std::cout << "iter: " << ndx << ", CL: " << CL << ", CD: " << CD << ", res: " << res[0] << ", "
            << res[1] << ", " << res[2] << ", " << res[3] << std::endl;
  std::ofstream out("sol/output.dat", std::ios_base::app | std::ios_base::out);
  out << ndx << " " << CL << " " << CD << std::endl;
  out.close();
  std::ofstream rh("sol/res_history.dat", std::ios_base::app | std::ios_base::out);
  rh << ndx << " " << res[0] << " " << res[1] << " " << res[2] << " " << res[3] << std::endl;
  rh.close();
  double cl = 0, cd = 0;
  wall_forces(g, geo, in, q, grad, &cl, &cd);
  std::ofstream fo("sol/forces.dat", std::ios_base::app | std::ios_base::out);
  fo << std::setprecision(10) << ndx << " " << cl << " " << cd;
  if (t >= 0) fo << " " << t;
  fo << std::endl;
}

void wall_forces(const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
                 const double* grad, double* cl, double* cd) {
  const int32_t nc = g.nc;
  // the wall group of readmesh.cpp:356-360: BC1 when its kind is a wall, else BC2
  int32_t group = 0;
  if (in.BC1 == FVK_BC_SLIP_WALL || in.BC1 == FVK_BC_NO_SLIP_WALL) group = -1;
  else if (in.BC2 == FVK_BC_SLIP_WALL || in.BC2 == FVK_BC_NO_SLIP_WALL) group = -2;
  const bool noslip = (group == -1 ? in.BC1 : in.BC2) == FVK_BC_NO_SLIP_WALL;
  double fx = 0, fy = 0;
  for (int32_t f : geo.bc_faces) {
    if (group == 0 || g.c2[f] != group) continue;
    const int32_t c = g.c1[f];
    const double dx = geo.fcx[f] - geo.ccx[c], dy = geo.fcy[f] - geo.ccy[c];
    double qf[4];
    for (int k = 0; k < 4; ++k) {
      qf[k] = q[(size_t)k * nc + c];
      if (grad)
        qf[k] += grad[(size_t)k * nc + c] * dx + grad[(size_t)(4 + k) * nc + c] * dy;
    }
    const double P = (qf[3] - 0.5 * (qf[1] * qf[1] + qf[2] * qf[2]) / qf[0]) * (in.gamma - 1);
    fx += P * geo.area[f] * geo.nx[f];
    fy += P * geo.area[f] * geo.ny[f];
    if (in.eqn == 2 && noslip) {   // mu * (centroid velocity) / (normal distance), along it
      const double dn = std::sqrt(dx * dx + dy * dy);
      const double rho = q[c];
      fx += in.mu * (q[(size_t)nc + c] / rho) / dn * geo.area[f];
      fy += in.mu * (q[2 * (size_t)nc + c] / rho) / dn * geo.area[f];
    }
  }
  const double vinf2 = in.u * in.u + in.v * in.v, vinf = std::sqrt(vinf2);
  const double ex = in.u / vinf, ey = in.v / vinf;
  const double qdyn = 0.5 * in.rho * vinf2 * in.length;
  *cd = (fx * ex + fy * ey) / qdyn;
  *cl = (-fx * ey + fy * ex) / qdyn;
}

void write_tecplot(int ndx, const Grid& g, const Geometry& geo, const Inputs& in, const double* q,
                   const double* grad) {
  std::cout << "Saving Solution \n";
  const int32_t nc = g.nc, np = g.np;
  const double* p[4] = {q, q + nc, q + 2 * (size_t)nc, q + 3 * (size_t)nc};
  const double* qx[4] = {grad, grad + nc, grad + 2 * (size_t)nc, grad + 3 * (size_t)nc};
  const double* qy[4] = {grad + 4 * (size_t)nc, grad + 5 * (size_t)nc, grad + 6 * (size_t)nc,
                         grad + 7 * (size_t)nc};
  // node values: average of the cell reconstructions at each point,
  // save_sol.cpp:110-162
  std::vector<int> cnt(np, 0);
  std::vector<double> avg[4];
  for (int k = 0; k < 4; ++k) avg[k].assign(np, 0.0);
  for (int32_t f = 0; f < g.nf; ++f) {
    const int c1 = g.c1[f], c2 = geo.fc2kind[f];
    const int a = g.p1[f], b = g.p2[f];
    cnt[a]++;
    cnt[b]++;
    for (int k = 0; k < 4; ++k)
      avg[k][a] += p[k][c1] + ((g.x[a] - geo.ccx[c1]) * qx[k][c1] + (g.y[a] - geo.ccy[c1]) * qy[k][c1]);
    for (int k = 0; k < 4; ++k)
      avg[k][b] += p[k][c1] + ((g.x[b] - geo.ccx[c1]) * qx[k][c1] + (g.y[b] - geo.ccy[c1]) * qy[k][c1]);
    if (c2 >= 0) {
      for (int k = 0; k < 4; ++k)
        avg[k][a] += p[k][c2] + ((g.x[a] - geo.ccx[c2]) * qx[k][c2] + (g.y[a] - geo.ccy[c2]) * qy[k][c2]);
      for (int k = 0; k < 4; ++k)
        avg[k][b] += p[k][c2] + ((g.x[b] - geo.ccx[c2]) * qx[k][c2] + (g.y[b] - geo.ccy[c2]) * qy[k][c2]);
      cnt[a]++;
      cnt[b]++;
    }
  }
  std::vector<double> rho(np), u(np), v(np), P(np);
  for (int32_t i = 0; i < np; ++i) {
    double a[4];
    for (int k = 0; k < 4; ++k) a[k] = avg[k][i] / cnt[i];
    rho[i] = a[0];
    const double E = a[3];
    u[i] = a[1] / rho[i];
    v[i] = a[2] / rho[i];
    P[i] = (E - (0.5 * rho[i] * (u[i] * u[i] + v[i] * v[i]))) * (in.gamma - 1);
  }
  const int nw = geo.num_wall;
  std::vector<double> tau_x(nw > 0 ? nw : 0, 0.0), tau_y(nw > 0 ? nw : 0, 0.0);
  if (in.eqn == 2) {
    for (int i = 0; i < nw; ++i) {
      const int face = geo.bc_faces[i];
      const int c1 = g.c1[face];
      const double dx = geo.fcx[face] - geo.ccx[c1], dy = geo.fcy[face] - geo.ccy[c1];
      const double dn = std::fabs(std::sqrt(dx * dx + dy * dy));
      const double cu = p[1][c1] / p[0][c1], cv = p[2][c1] / p[0][c1];
      tau_x[i] = in.mu * (-cu / dn);
      tau_y[i] = in.mu * (-cv / dn);
    }
  }
  const bool last = (in.nmax == ndx);
  std::ofstream out(last ? std::string("sol/tec_sol.dat") : "sol/tec_sol_" + std::to_string(ndx) + ".dat");
  if (out.is_open()) {
    out << "VARIABLES = \"X\", \"Y\", \"density\", \"u\", \"v\", \"P\" \n";
    out << "Zone STRANDID=1, ZoneType=FETRIANGLE, SOLUTIONTIME=" << ndx << std::endl;
    out << "n=" << g.nf << " e=" << g.nf * 2 - (int)geo.bc_faces.size() << " DATAPACKING=POINT" << std::endl;
    out << "DT=(DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE )" << std::endl;
    for (int32_t i = 0; i < np; ++i)
      out << g.x[i] << " " << g.y[i] << " " << rho[i] << " " << u[i] << " " << v[i] << " " << P[i] << std::endl;
    for (int32_t c = 0; c < nc; ++c) {
      const double r = p[0][c], E = p[3][c];
      const double cu = p[1][c] / r, cv = p[2][c] / r;
      const double cP = (E - (0.5 * r * (cu * cu + cv * cv))) * (in.gamma - 1);
      out << geo.ccx[c] << " " << geo.ccy[c] << " " << r << " " << cu << " " << cv << " " << cP << std::endl;
    }
    for (int32_t c = 0; c < nc; ++c)
      for (int32_t j = geo.cf_ptr[c]; j < geo.cf_ptr[c + 1]; ++j) {
        const int f = geo.cf_idx[j];
        out << (double)(g.p1[f] + 1) << " " << (double)(g.p2[f] + 1) << " " << (double)(np + c + 1) << std::endl;
      }
  } else {
    std::cout << "ERROR: Cannot Save Tecplot File \n";
  }
  out.close();
  out.open(last ? std::string("sol/tec_surf.dat") : "sol/tec_surf_" + std::to_string(ndx) + ".dat");
  if (out.is_open()) {
    out << "TITLE     = \"Translation of CGNS file merged_sol_aver.cgns\"" << std::endl;
    out << "VARIABLES = \"CoordinateX\" \n \"CoordinateY\" \n \"rho\" \n \"u\" \n \"v\" \n \"P\" \n \"tau_x\" \n \"tau_y\" " << std::endl;
    out << "ZONE T=\"ZONE1\"\nSTRANDID=0, SOLUTIONTIME=" << ndx << "\nI=" << nw + 1 << ", J=1, K=1, ZONETYPE=Ordered\n";
    out << "DATAPACKING=POINT\nDT=(DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE )\n";
    for (int i = 0; i < nw + 1 && i < np; ++i) {
      // the reference reads one past its wall-stress arrays on the last line
      const double tx = i < nw ? tau_x[i] : 0.0, ty = i < nw ? tau_y[i] : 0.0;
      out << g.x[i] << " " << g.y[i] << " " << rho[i] << " " << u[i] << " " << v[i] << " " << P[i]
          << " " << tx << " " << ty << std::endl;
    }
  } else {
    std::cout << "ERROR: Cannot Save Tecplot Surface File \n";
  }
}

void write_restart(int ndx, int nc, const double* q) {
  std::cout << "Writing Restart File \n";
  std::ofstream out("sol/restart_sol.dat");
  if (!out.is_open()) {
    std::cout << "ERROR: Cannot Write Restart File \n";
    return;
  }
  out << ndx << " " << nc << std::endl;
  for (int i = 0; i < nc; ++i)
    out << q[i] << " " << q[nc + i] << " " << q[2 * (size_t)nc + i] << " " << q[3 * (size_t)nc + i] << std::endl;
}

bool read_restart(int nc, int* ndx, std::vector<double>* q, std::string* log) {
  std::ifstream in("sol/restart_sol.dat");
  if (!in.is_open()) return false;
  int n = 0;
  in >> *ndx >> n;
  if (n != nc) *log += "ERROR: Restart File Does Not Match Grid Input";
  q->assign(4 * (size_t)nc, -101.0);
  for (int i = 0; i < n && i < nc; ++i)
    in >> (*q)[i] >> (*q)[nc + i] >> (*q)[2 * (size_t)nc + i] >> (*q)[3 * (size_t)nc + i];
  return true;
}

}  // namespace fv2d
