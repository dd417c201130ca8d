// mesh_prep.cpp — geometry (bit-identical to src/base/readmesh.cpp) and the
// RCM + sliced-ELL device layout.  Built with -ffp-contract=off: every FP64
// expression below is evaluated exactly in the order written.
#include "mesh_prep.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <thread>

namespace fvk {
namespace {

// fn(begin, end) over [0, n) split into `threads` contiguous ranges, one
// thread each (the calling thread takes the first).  Every loop run through
// it writes only the elements of its own range, so the result does not
// depend on the split.
template <typename Fn>
void parallel_for(int64_t n, int threads, Fn fn) {
  const int64_t t = std::max<int64_t>(1, std::min<int64_t>(threads, n / 4096));
  if (t <= 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  for (int64_t k = 1; k < t; ++k) pool.emplace_back(fn, n * k / t, n * (k + 1) / t);
  fn(int64_t(0), n / t);
  for (auto& th : pool) th.join();
}

int bc_kind_of(const fvk_options& o, int32_t c2) {
  if (c2 == -1) return o.bc1;  // readmesh.cpp:61-63
  if (c2 == -2) return o.bc2;  // readmesh.cpp:64-66
  return c2;                   // readmesh.cpp:67-69: raw -3 / -4
}

// Face normals/areas/centres (readmesh.cpp:117-139).
void face_geometry(const fvk_mesh& m, HostMesh& h, int threads) {
  h.area.resize(h.nf); h.nx.resize(h.nf); h.ny.resize(h.nf);
  h.fcx.resize(h.nf); h.fcy.resize(h.nf);
  parallel_for(h.nf, threads, [&](int64_t f0, int64_t f1) {
  for (int32_t f = (int32_t)f0; f < (int32_t)f1; ++f) {
    const double xa = m.x[m.face_point1[f]], ya = m.y[m.face_point1[f]];
    const double xb = m.x[m.face_point2[f]], yb = m.y[m.face_point2[f]];
    const double dx = xb - xa, dy = yb - ya;
    const double len = std::sqrt(dx * dx + dy * dy);
    h.area[f] = len;
    h.nx[f] = dy / len;
    h.ny[f] = -dx / len;
    h.fcx[f] = dx / 2 + xa;
    h.fcy[f] = dy / 2 + ya;
  }
  });
}

// Cell volume and centroid (readmesh.cpp:147-228): triangles fanned about the
// mean of the cell's distinct vertices.  The distinct-vertex search scans all
// face-count slots, which start at (0,0), exactly as the reference does.
void cell_geometry(const fvk_mesh& m, HostMesh& h, int threads) {
  h.vol.resize(h.nc); h.ccx.resize(h.nc); h.ccy.resize(h.nc);
  parallel_for(h.nc, threads, [&](int64_t c0, int64_t c1) {
  std::vector<double> vx(h.max_deg), vy(h.max_deg);
  for (int32_t c = (int32_t)c0; c < (int32_t)c1; ++c) {
    const int32_t b = h.cf_ptr[c], n = h.cf_ptr[c + 1] - b;
    std::fill(vx.begin(), vx.begin() + n, 0.0);
    std::fill(vy.begin(), vy.begin() + n, 0.0);
    int32_t found = 0;
    auto add_vertex = [&](double px, double py) {
      for (int32_t z = 0; z < n; ++z)
        if (vx[z] == px && vy[z] == py) return;
      if (found < n) { vx[found] = px; vy[found] = py; ++found; }
    };
    for (int32_t j = 0; j < n; ++j) {
      const int32_t f = h.cf_idx[b + j];
      add_vertex(m.x[m.face_point1[f]], m.y[m.face_point1[f]]);
      add_vertex(m.x[m.face_point2[f]], m.y[m.face_point2[f]]);
    }
    double sx = 0, sy = 0;
    for (int32_t j = 0; j < n; ++j) { sx += vx[j]; sy += vy[j]; }
    const double mx = sx / n, my = sy / n;
    double vsum = 0, wx = 0, wy = 0;
    for (int32_t j = 0; j < n; ++j) {
      const int32_t f = h.cf_idx[b + j];
      const double x1 = m.x[m.face_point1[f]], y1 = m.y[m.face_point1[f]];
      const double x2 = m.x[m.face_point2[f]], y2 = m.y[m.face_point2[f]];
      const double tri = 0.5 * std::fabs(x1 * (y2 - my) + x2 * (my - y1) + mx * (y1 - y2));
      vsum += tri;
      const double tx = (x1 + x2 + mx) / 3, ty = (y1 + y2 + my) / 3;
      wx += tx * tri;
      wy += ty * tri;
    }
    h.vol[c] = vsum;
    h.ccx[c] = wx / vsum;
    h.ccy[c] = wy / vsum;
  }
  });
}

// Mirror image of the owner centroid about a boundary face (readmesh.cpp:238-243).
inline void ghost_centroid(const HostMesh& h, int32_t f, double* gx, double* gy) {
  const int32_t c = h.fc1[f];
  const double rx = h.fcx[f] - h.ccx[c];
  const double ry = h.fcy[f] - h.ccy[c];
  const double t = 2.0 * (rx * h.nx[f] + ry * h.ny[f]);
  *gx = h.ccx[c] + t * h.nx[f];
  *gy = h.ccy[c] + t * h.ny[f];
}

// LSQ normal-matrix inverse including ghost neighbours (readmesh.cpp:268-320),
// accumulated per cell over its faces in increasing face index.
// The reference's face loop (readmesh.cpp:246-285) adds each face's terms to
// its owner and neighbour in increasing face index.  Gathered per cell over
// its faces in that same order, each sum takes the same additions in the same
// order: the neighbour's terms (xj-xi)^2 ... are the cell's own
// (x_self - x_other)^2 ..., as the owner's are.
void lsq_geometry(HostMesh& h, int threads) {
  h.ixx.resize(h.nc); h.iyy.resize(h.nc); h.ixy.resize(h.nc);
  std::vector<double> cond(h.nc);
  parallel_for(h.nc, threads, [&](int64_t c0, int64_t c1) {
  for (int32_t c = (int32_t)c0; c < (int32_t)c1; ++c) {
    double sxx = 0.0, syy = 0.0, sxy = 0.0;
    const double xs = h.ccx[c], ys = h.ccy[c];
    for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) {
      const int32_t f = h.cf_idx[j];
      const int32_t a = h.fc1[f], b = h.fc2[f];
      double xo, yo;
      if (b < 0) ghost_centroid(h, f, &xo, &yo);
      else if (a == c) { xo = h.ccx[b]; yo = h.ccy[b]; }
      else { xo = h.ccx[a]; yo = h.ccy[a]; }
      sxx += (xs - xo) * (xs - xo);
      syy += (ys - yo) * (ys - yo);
      sxy += ((xs - xo) * (ys - yo));
    }
    const double det = sxx * syy - (sxy * sxy);
    h.ixx[c] = sxx / det;
    h.iyy[c] = syy / det;
    h.ixy[c] = sxy / det;
    // condition number of [[sxx, sxy], [sxy, syy]] from its eigenvalues
    const double tr = sxx + syy, disc = std::sqrt(std::max(0.0, tr * tr - 4 * det));
    const double lo = 0.5 * (tr - disc);
    cond[c] = lo > 0 ? 0.5 * (tr + disc) / lo : HUGE_VAL;
  }
  });
  h.max_lsq_cond = *std::max_element(cond.begin(), cond.end());
}

// Reverse Cuthill-McKee over the interior-face graph.  Each connected part
// starts from a pseudo-peripheral cell (two BFS sweeps from a min-degree cell).
std::vector<int32_t> rcm_order(const HostMesh& h) {
  const int32_t n = h.nc;
  std::vector<int32_t> aptr(n + 1, 0), adj;
  for (int32_t c = 0; c < n; ++c) {
    int32_t d = 0;
    for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j)
      if (h.fc2[h.cf_idx[j]] >= 0) ++d;
    aptr[c + 1] = aptr[c] + d;
  }
  adj.resize(aptr[n]);
  for (int32_t c = 0; c < n; ++c) {
    int32_t k = aptr[c];
    for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) {
      const int32_t f = h.cf_idx[j];
      if (h.fc2[f] < 0) continue;
      adj[k++] = (h.fc1[f] == c) ? h.fc2[f] : h.fc1[f];
    }
  }
  auto degree = [&](int32_t c) { return aptr[c + 1] - aptr[c]; };
  std::vector<int32_t> level(n, -1), order;
  order.reserve(n);
  std::vector<char> placed(n, 0);
  std::vector<int32_t> queue;
  queue.reserve(n);
  // BFS from s; returns the last level's min-degree cell and the depth
  auto bfs_far = [&](int32_t s, std::vector<int32_t>& touched) {
    queue.clear();
    queue.push_back(s);
    level[s] = 0;
    touched.push_back(s);
    size_t head = 0;
    int32_t far = s;
    while (head < queue.size()) {
      const int32_t c = queue[head++];
      if (level[c] > level[far] || (level[c] == level[far] && degree(c) < degree(far))) far = c;
      for (int32_t k = aptr[c]; k < aptr[c + 1]; ++k)
        if (level[adj[k]] < 0) {
          level[adj[k]] = level[c] + 1;
          queue.push_back(adj[k]);
          touched.push_back(adj[k]);
        }
    }
    return std::make_pair(far, level[far]);
  };
  std::vector<int32_t> by_degree(n);
  std::iota(by_degree.begin(), by_degree.end(), 0);
  std::stable_sort(by_degree.begin(), by_degree.end(),
                   [&](int32_t a, int32_t b) { return degree(a) < degree(b); });
  std::vector<int32_t> touched, nbrs;
  for (int32_t start : by_degree) {
    if (placed[start]) continue;
    int32_t s = start, depth = -1;
    for (int sweep = 0; sweep < 3; ++sweep) {
      touched.clear();
      auto r = bfs_far(s, touched);
      for (int32_t c : touched) level[c] = -1;
      if (r.second <= depth) break;
      depth = r.second;
      s = r.first;
    }
    // Cuthill-McKee BFS from s
    size_t first = order.size();
    order.push_back(s);
    placed[s] = 1;
    for (size_t head = first; head < order.size(); ++head) {
      const int32_t c = order[head];
      nbrs.clear();
      for (int32_t k = aptr[c]; k < aptr[c + 1]; ++k)
        if (!placed[adj[k]]) { nbrs.push_back(adj[k]); placed[adj[k]] = 1; }
      std::stable_sort(nbrs.begin(), nbrs.end(),
                       [&](int32_t a, int32_t b) { return degree(a) < degree(b); });
      order.insert(order.end(), nbrs.begin(), nbrs.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Recursive coordinate bisection: split ids[0..n) into k parts of at most
// cap cells, cutting each range across the longer extent of its centroids.
// In place: the parts end up contiguous, part after part.  The two halves of
// a cut are independent, so `threads` threads share them; the order of each
// half depends only on its own cells, so the result is the serial one.
int64_t rcb_left(int64_t n, int64_t k, int64_t cap) {
  const int64_t kl = k / 2, kr = k - kl;
  int64_t nl = (n * kl + k / 2) / k;
  nl = std::max(nl, n - cap * kr);
  return std::min(nl, cap * kl);
}

void rcb(int32_t* ids, int64_t n, int64_t k, int64_t cap, const double* cx, const double* cy,
         int threads) {
  if (k <= 1 || n <= 1) return;
  const int64_t kl = k / 2, kr = k - kl;
  const int64_t nl = rcb_left(n, k, cap);
  double x0 = cx[ids[0]], x1 = x0, y0 = cy[ids[0]], y1 = y0;
  for (int64_t i = 1; i < n; ++i) {
    x0 = std::min(x0, cx[ids[i]]); x1 = std::max(x1, cx[ids[i]]);
    y0 = std::min(y0, cy[ids[i]]); y1 = std::max(y1, cy[ids[i]]);
  }
  const double* key = (x1 - x0) >= (y1 - y0) ? cx : cy;
  std::nth_element(ids, ids + nl, ids + n, [key](int32_t a, int32_t b) {
    return key[a] < key[b] || (key[a] == key[b] && a < b);
  });
  if (threads > 1 && n > 65536) {
    std::thread left(rcb, ids, nl, kl, cap, cx, cy, threads / 2);
    rcb(ids + nl, n - nl, kr, cap, cx, cy, threads - threads / 2);
    left.join();
  } else {
    rcb(ids, nl, kl, cap, cx, cy, 1);
    rcb(ids + nl, n - nl, kr, cap, cx, cy, 1);
  }
}

// part sizes of rcb(n, k, cap), in part order
void rcb_sizes(int64_t n, int64_t k, int64_t cap, std::vector<int32_t>& sizes) {
  if (k <= 1 || n <= 1) {
    sizes.push_back((int32_t)n);
    for (int64_t i = 1; i < k; ++i) sizes.push_back(0);
    return;
  }
  const int64_t kl = k / 2, nl = rcb_left(n, k, cap);
  rcb_sizes(nl, kl, cap, sizes);
  rcb_sizes(n - nl, k - kl, cap, sizes);
}

}  // namespace

int prepare_mesh(const fvk_mesh& m, const fvk_options& o, HostMesh* out, std::string* err,
                 int threads) {
  threads = std::max(1, threads);
  HostMesh& h = *out;
  h = HostMesh();
  char buf[256];
  if (m.num_points <= 0 || m.num_faces <= 0 || m.num_cells <= 0 || !m.x || !m.y ||
      !m.face_point1 || !m.face_point2 || !m.face_cell1 || !m.face_cell2) {
    *err = "mesh: empty or NULL arrays";
    return FVK_ERR_ARG;
  }
  if (m.num_cells >= (1 << 29)) {
    *err = "mesh: more than 2^29 cells";
    return FVK_ERR_MESH;
  }
  h.np = m.num_points; h.nf = m.num_faces; h.nc = m.num_cells;
  h.fc1.resize(h.nf); h.fc2.resize(h.nf);
  std::vector<int32_t> cnt(h.nc, 0);
  for (int32_t f = 0; f < h.nf; ++f) {
    const int32_t p1 = m.face_point1[f], p2 = m.face_point2[f];
    const int32_t c1 = m.face_cell1[f], c2 = m.face_cell2[f];
    if (p1 < 0 || p1 >= h.np || p2 < 0 || p2 >= h.np || p1 == p2 || c1 < 0 || c1 >= h.nc ||
        c2 >= h.nc || c2 < -4 || c2 == c1) {
      snprintf(buf, sizeof buf, "mesh: face %d has an out-of-range or degenerate entry", f);
      *err = buf;
      return FVK_ERR_MESH;
    }
    h.fc1[f] = c1;
    h.fc2[f] = (c2 >= 0) ? c2 : bc_kind_of(o, c2);
    if (c2 == -1) ++h.num_bc1;
    if (c2 == -2) ++h.num_bc2;
    if (c2 < 0) ++h.nb;
    cnt[c1]++;
    if (c2 >= 0) cnt[c2]++;
  }
  for (int32_t f = 0; f < h.nf; ++f)
    if (h.fc2[f] < 0 && (h.fc2[f] < -4 || h.fc2[f] > -1)) {
      *err = "options: BC kind must be -1..-4";
      return FVK_ERR_OPTION;
    }
  h.cf_ptr.assign(h.nc + 1, 0);
  for (int32_t c = 0; c < h.nc; ++c) {
    if (cnt[c] < 3) {
      snprintf(buf, sizeof buf, "mesh: cell %d has %d faces (needs at least 3)", c, cnt[c]);
      *err = buf;
      return FVK_ERR_MESH;
    }
    if (cnt[c] > 255) {
      snprintf(buf, sizeof buf, "mesh: cell %d has %d faces (at most 255)", c, cnt[c]);
      *err = buf;
      return FVK_ERR_MESH;
    }
    h.cf_ptr[c + 1] = h.cf_ptr[c] + cnt[c];
    h.max_deg = std::max(h.max_deg, cnt[c]);
  }
  h.cf_idx.resize(h.cf_ptr[h.nc]);
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int32_t f = 0; f < h.nf; ++f) {
    const int32_t c1 = h.fc1[f], c2 = h.fc2[f];
    h.cf_idx[h.cf_ptr[c1] + cnt[c1]++] = f;
    if (c2 >= 0) h.cf_idx[h.cf_ptr[c2] + cnt[c2]++] = f;
  }
  // wall count, readmesh.cpp:356-360
  if (o.bc1 == FVK_BC_SLIP_WALL || o.bc1 == FVK_BC_NO_SLIP_WALL) h.num_wall = h.num_bc1;
  else if (o.bc2 == FVK_BC_SLIP_WALL || o.bc2 == FVK_BC_NO_SLIP_WALL) h.num_wall = h.num_bc2;

  face_geometry(m, h, threads);
  cell_geometry(m, h, threads);
  lsq_geometry(h, threads);

  // The reference's no-slip gradient ghost reads the state of the cell whose
  // index equals the face index (cell_gradient.cpp:63-66).
  int32_t noslip_mismatch = 0;
  for (int32_t f = 0; f < h.nf; ++f) {
    if (h.fc2[f] != FVK_BC_NO_SLIP_WALL) continue;
    if (f >= h.nc) {
      if (o.order == 2) {
        snprintf(buf, sizeof buf,
                 "mesh: no-slip face %d has no cell of the same index; second-order no-slip "
                 "walls need boundary faces first with wall face k owned by cell k",
                 f);
        *err = buf;
        return FVK_ERR_MESH;
      }
    } else if (h.fc1[f] != f) {
      ++noslip_mismatch;
    }
  }
  if (noslip_mismatch && o.order == 2) {
    snprintf(buf, sizeof buf,
             "warning: %d no-slip faces are not owned by the cell of the same index; the "
             "gradient ghost uses that cell's state, as the reference does\n",
             noslip_mismatch);
    h.warnings += buf;
  }
  int32_t first_interior = -1;
  for (int32_t f = 0; f < h.nf; ++f)
    if (h.fc2[f] >= 0) { first_interior = f; break; }
  for (int32_t f = first_interior >= 0 ? first_interior : h.nf; f < h.nf; ++f)
    if (h.fc2[f] < 0) {
      h.warnings += "warning: boundary faces are not all first (the reference reader reads "
                    "past its boundary arrays on such grids)\n";
      break;
    }

  // renumbering: blocks of blk cells, numbered block-major
  h.blk = o.block_size > 0 ? o.block_size : 256;
  h.nblocks = (h.nc + h.blk - 1) / h.blk;
  h.ncp = h.nblocks * h.blk;
  std::vector<int32_t> order, sizes;
  if (o.renumber == 1) {
    order.resize(h.nc);
    std::iota(order.begin(), order.end(), 0);
    rcb(order.data(), h.nc, h.nblocks, h.blk, h.ccx.data(), h.ccy.data(), threads);
    rcb_sizes(h.nc, h.nblocks, h.blk, sizes);
    // order each block's cells by a further bisection into 64-cell groups
    std::vector<int64_t> start(h.nblocks + 1, 0);
    for (int32_t b = 0; b < h.nblocks; ++b) start[b + 1] = start[b] + sizes[b];
    // then by face count, most first, keeping that order among equals: a
    // 64-cell chunk (one wavefront) then holds cells of one face count
    // wherever the block allows, so its slot rows carry little padding
    // (a mesh of mixed triangles and quads loses most of its padding)
    parallel_for(h.nblocks, threads, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; ++b) {
        const int32_t n = sizes[b];
        int32_t* first = order.data() + start[b];
        rcb(first, n, (n + kChunk - 1) / kChunk, kChunk, h.ccx.data(), h.ccy.data(), 1);
        std::stable_sort(first, first + n, [&](int32_t x, int32_t y) {
          return h.cf_ptr[x + 1] - h.cf_ptr[x] > h.cf_ptr[y + 1] - h.cf_ptr[y];
        });
      }
    });
  } else {
    if (o.renumber == 2) order = rcm_order(h);
    else { order.resize(h.nc); std::iota(order.begin(), order.end(), 0); }
    for (int32_t b = 0; b < h.nblocks; ++b) sizes.push_back(std::min(h.blk, h.nc - b * h.blk));
  }
  h.perm.assign(h.ncp, -1);
  {
    size_t pos = 0;
    for (int32_t b = 0; b < h.nblocks; ++b)
      for (int32_t i = 0; i < sizes[b]; ++i) h.perm[(size_t)b * h.blk + i] = order[pos++];
  }
  h.iperm.assign(h.nc, 0);
  for (int32_t i = 0; i < h.ncp; ++i)
    if (h.perm[i] >= 0) h.iperm[h.perm[i]] = i;

  // halos and local indices
  h.halo_ptr.assign(h.nblocks + 1, 0);
  std::vector<std::vector<int32_t>> halos(h.nblocks);
  parallel_for(h.nblocks, threads, [&](int64_t b0, int64_t b1) {
  for (int32_t b = (int32_t)b0; b < (int32_t)b1; ++b) {
    std::vector<int32_t>& hl = halos[b];
    for (int32_t i = 0; i < h.blk; ++i) {
      const int32_t c = h.perm[(size_t)b * h.blk + i];
      if (c < 0) continue;
      for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) {
        const int32_t f = h.cf_idx[j];
        if (h.fc2[f] < 0) continue;
        const int32_t other = h.iperm[(h.fc1[f] == c) ? h.fc2[f] : h.fc1[f]];
        if (other / h.blk != b) hl.push_back(other);
      }
    }
    std::sort(hl.begin(), hl.end());
    hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
  }
  });
  for (int32_t b = 0; b < h.nblocks; ++b) {
    h.halo_ptr[b + 1] = h.halo_ptr[b] + (int32_t)halos[b].size();
    h.max_halo = std::max<int32_t>(h.max_halo, (int32_t)halos[b].size());
  }
  h.halo_idx.reserve(h.halo_ptr[h.nblocks]);
  for (auto& hl : halos) h.halo_idx.insert(h.halo_idx.end(), hl.begin(), hl.end());
  auto local_of = [&](int32_t b, int32_t g) -> int32_t {
    if (g / h.blk == b) return g % h.blk;
    const std::vector<int32_t>& hl = halos[b];
    return h.blk + (int32_t)(std::lower_bound(hl.begin(), hl.end(), g) - hl.begin());
  };

  // block face lists, in increasing original face index
  h.bf_ptr.assign(h.nblocks + 1, 0);
  std::vector<std::vector<int32_t>> bfaces(h.nblocks);
  parallel_for(h.nblocks, threads, [&](int64_t b0, int64_t b1) {
  for (int32_t b = (int32_t)b0; b < (int32_t)b1; ++b) {
    std::vector<int32_t>& fl = bfaces[b];
    for (int32_t i = 0; i < h.blk; ++i) {
      const int32_t c = h.perm[(size_t)b * h.blk + i];
      if (c < 0) continue;
      for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) fl.push_back(h.cf_idx[j]);
    }
    std::sort(fl.begin(), fl.end());
    fl.erase(std::unique(fl.begin(), fl.end()), fl.end());
  }
  });
  for (int32_t b = 0; b < h.nblocks; ++b) {
    h.bf_ptr[b + 1] = h.bf_ptr[b] + (int32_t)bfaces[b].size();
    h.max_bfaces = std::max<int32_t>(h.max_bfaces, (int32_t)bfaces[b].size());
  }
  const size_t nbf = h.bf_ptr[h.nblocks];
  h.bf_a.resize(nbf); h.bf_b.resize(nbf);
  h.bf_nx.resize(nbf); h.bf_ny.resize(nbf); h.bf_area.resize(nbf);
  h.bf_fcx.resize(nbf); h.bf_fcy.resize(nbf);
  auto boundary_code = [&](int32_t f) {
    const int32_t c2 = h.fc2[f];
    const int32_t src = (c2 == FVK_BC_NO_SLIP_WALL && f < h.nc) ? h.iperm[f] : 0;
    return slot_code_boundary(src, -c2 - 1);
  };
  parallel_for(h.nblocks, threads, [&](int64_t b0, int64_t b1) {
  for (int32_t b = (int32_t)b0; b < (int32_t)b1; ++b) {
    size_t e = h.bf_ptr[b];
    for (int32_t f : bfaces[b]) {
      h.bf_a[e] = (int16_t)local_of(b, h.iperm[h.fc1[f]]);
      h.bf_b[e] = (int16_t)((h.fc2[f] >= 0) ? local_of(b, h.iperm[h.fc2[f]]) : ~(-h.fc2[f] - 1));
      h.bf_nx[e] = h.nx[f]; h.bf_ny[e] = h.ny[f]; h.bf_area[e] = h.area[f];
      h.bf_fcx[e] = h.fcx[f]; h.bf_fcy[e] = h.fcy[f];
      ++e;
    }
  }
  });

  // sliced-ELL slot table
  const int32_t nchunks = h.ncp / kChunk;
  h.deg.assign(h.ncp, 0);
  for (int32_t i = 0; i < h.ncp; ++i) {
    const int32_t c = h.perm[i];
    if (c >= 0) h.deg[i] = (uint8_t)(h.cf_ptr[c + 1] - h.cf_ptr[c]);
  }
  h.chunk_off.assign(nchunks + 1, 0);
  int64_t slots = 0;   // slot offsets are 32-bit on the device
  for (int32_t k = 0; k < nchunks; ++k) {
    int32_t mx = 0;
    for (int32_t l = 0; l < kChunk; ++l) mx = std::max<int32_t>(mx, h.deg[k * kChunk + l]);
    slots += (int64_t)mx * kChunk;
    if (slots >= (int64_t(1) << 31)) {
      *err = "mesh: more than 2^31 face slots";
      return FVK_ERR_MESH;
    }
    h.chunk_off[k + 1] = (int32_t)slots;
  }
  const int64_t nslots = h.chunk_off[nchunks];
  // every slot is written below: a cell's own slots, then zeros in the
  // padding slots of its lane (and in every slot of a padding cell's lane)
  h.code.resize(nslots); h.lcode.resize(nslots); h.sface.resize(nslots);
  h.s_nx.resize(nslots); h.s_ny.resize(nslots); h.s_area.resize(nslots);
  h.s_fcx.resize(nslots); h.s_fcy.resize(nslots);
  parallel_for(h.ncp, threads, [&](int64_t i0, int64_t i1) {
  for (int32_t i = (int32_t)i0; i < (int32_t)i1; ++i) {
    const int32_t c = h.perm[i];
    const int32_t k = i / kChunk;
    const int32_t base = h.chunk_off[k] + (i % kChunk);
    const int32_t width = (h.chunk_off[k + 1] - h.chunk_off[k]) / kChunk;
    for (int32_t j = c >= 0 ? h.cf_ptr[c + 1] - h.cf_ptr[c] : 0; j < width; ++j) {
      const int64_t s = base + (int64_t)j * kChunk;
      h.code[s] = 0; h.lcode[s] = 0; h.sface[s] = 0;
      h.s_nx[s] = 0.0; h.s_ny[s] = 0.0; h.s_area[s] = 0.0; h.s_fcx[s] = 0.0; h.s_fcy[s] = 0.0;
    }
    if (c < 0) continue;
    const int32_t b = i / h.blk;
    const std::vector<int32_t>& fl = bfaces[b];
    for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) {
      const int32_t f = h.cf_idx[j];
      const int64_t s = base + (int64_t)(j - h.cf_ptr[c]) * kChunk;
      const int32_t c2 = h.fc2[f];
      bool owner = true;
      if (c2 >= 0) {
        owner = (h.fc1[f] == c);
        const int32_t other = h.iperm[owner ? c2 : h.fc1[f]];
        h.code[s] = slot_code_interior(other, owner);
        h.lcode[s] = (int16_t)slot_code_interior(local_of(b, other), owner);
      } else {
        h.code[s] = boundary_code(f);
        h.lcode[s] = (int16_t)~(-c2 - 1);   // the kind; the ghost source is in code[s]
      }
      const int32_t fi = (int32_t)(std::lower_bound(fl.begin(), fl.end(), f) - fl.begin());
      h.sface[s] = (int16_t)((fi << 1) | (owner ? 1 : 0));
      h.s_nx[s] = h.nx[f]; h.s_ny[s] = h.ny[f]; h.s_area[s] = h.area[f];
      h.s_fcx[s] = h.fcx[f]; h.s_fcy[s] = h.fcy[f];
    }
  }
  });
  if (o.precision == FVK_FP32_MIXED) {
    // single-precision geometry of the FP32-mixed kernels
    h.s_dfx32.resize(nslots);
    h.s_dfy32.resize(nslots);
    parallel_for(h.ncp, threads, [&](int64_t i0, int64_t i1) {
    for (int32_t i = (int32_t)i0; i < (int32_t)i1; ++i) {
      const int32_t c = h.perm[i];
      const int32_t k = i / kChunk;
      const int32_t base = h.chunk_off[k] + (i % kChunk);
      const int32_t width = (h.chunk_off[k + 1] - h.chunk_off[k]) / kChunk;
      for (int32_t j = c >= 0 ? h.cf_ptr[c + 1] - h.cf_ptr[c] : 0; j < width; ++j) {
        h.s_dfx32[base + (int64_t)j * kChunk] = 0.0f;
        h.s_dfy32[base + (int64_t)j * kChunk] = 0.0f;
      }
      if (c < 0) continue;
      for (int32_t j = h.cf_ptr[c]; j < h.cf_ptr[c + 1]; ++j) {
        const int32_t f = h.cf_idx[j];
        const int64_t s = base + (int64_t)(j - h.cf_ptr[c]) * kChunk;
        h.s_dfx32[s] = (float)(h.fcx[f] - h.ccx[c]);
        h.s_dfy32[s] = (float)(h.fcy[f] - h.ccy[c]);
      }
    }
    });
    const bool ns = o.eqn == 2;
    h.bf32_nx.resize(nbf); h.bf32_ny.resize(nbf); h.bf32_area.resize(nbf);
    h.bf32_d1x.resize(nbf); h.bf32_d1y.resize(nbf);
    h.bf32_d2x.resize(nbf); h.bf32_d2y.resize(nbf);
    if (ns) {
      h.bf32_lx.resize(nbf); h.bf32_ly.resize(nbf); h.bf32_dist.resize(nbf);
    }
    parallel_for(h.nblocks, threads, [&](int64_t b0, int64_t b1) {
    for (int32_t b = (int32_t)b0; b < (int32_t)b1; ++b) {
      size_t e = h.bf_ptr[b];
      for (int32_t f : bfaces[b]) {
        const int32_t c1 = h.fc1[f], c2 = h.fc2[f];
        h.bf32_nx[e] = (float)h.nx[f];
        h.bf32_ny[e] = (float)h.ny[f];
        h.bf32_area[e] = (float)h.area[f];
        const double dx1 = h.fcx[f] - h.ccx[c1], dy1 = h.fcy[f] - h.ccy[c1];
        h.bf32_d1x[e] = (float)dx1;
        h.bf32_d1y[e] = (float)dy1;
        h.bf32_d2x[e] = 0.0f;
        h.bf32_d2y[e] = 0.0f;
        if (ns) {
          h.bf32_lx[e] = 0.0f;
          h.bf32_ly[e] = 0.0f;
          h.bf32_dist[e] = 0.0f;
        }
        if (c2 >= 0) {
          h.bf32_d2x[e] = (float)(h.fcx[f] - h.ccx[c2]);
          h.bf32_d2y[e] = (float)(h.fcy[f] - h.ccy[c2]);
          if (ns) {   // reconstruction.cpp:124-161 (face_dist, l)
            const double tx = h.ccx[c2] - h.ccx[c1], ty = h.ccy[c2] - h.ccy[c1];
            const double dist = std::sqrt(tx * tx + ty * ty);
            h.bf32_lx[e] = (float)(tx / dist);
            h.bf32_ly[e] = (float)(ty / dist);
            h.bf32_dist[e] = (float)dist;
          }
        } else if (ns && c2 == FVK_BC_NO_SLIP_WALL) {   // reconstruction.cpp:110-123
          h.bf32_dist[e] = (float)std::fabs(std::sqrt(dx1 * dx1 + dy1 * dy1));
        }
        ++e;
      }
    }
    });
  }
  auto permute = [&](const hvec<double>& src, hvec<double>& dst, double pad) {
    dst.resize(h.ncp);
    parallel_for(h.ncp, threads, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) dst[i] = h.perm[i] >= 0 ? src[h.perm[i]] : pad;
    });
  };
  permute(h.ccx, h.c_ccx, 0.0);
  permute(h.ccy, h.c_ccy, 0.0);
  permute(h.vol, h.c_vol, 1.0);
  permute(h.ixx, h.c_ixx, 0.0);
  permute(h.iyy, h.c_iyy, 0.0);
  permute(h.ixy, h.c_ixy, 0.0);
  if (o.precision == FVK_FP32_MIXED) {
    auto round32 = [&](const hvec<double>& src, hvec<float>& dst) {
      dst.resize(h.ncp);
      parallel_for(h.ncp, threads, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) dst[i] = (float)src[i];
      });
    };
    round32(h.c_ixx, h.c_ixx32);
    round32(h.c_iyy, h.c_iyy32);
    round32(h.c_ixy, h.c_ixy32);
  }
  return FVK_OK;
}

}  // namespace fvk
