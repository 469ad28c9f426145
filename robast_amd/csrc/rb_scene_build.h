// rb_scene_build.h -- host side of rbgpu_scene_upload: validation of a
// flattened scene, the per-node BVH build and packing into one blob whose
// layout the device reads through DScene. Plain C++ (no HIP calls).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rbgpu.h"
#include "rb_device.h"

namespace rbd {

inline int fail_msg(std::string& err, int code, const std::string& msg) {
  err = msg;
  return code;
}

// ---------------------------------------------------------------- BVH build
struct Box {
  double lo[3], hi[3];
};

Box node_box(const rb_scene_desc* d, int n) {
  const rb_node_desc& nd = d->nodes[n];
  const rb_shape_desc& sh = d->shapes[d->volumes[nd.volume].shape];
  Box b;
  for (int k = 0; k < 3; k++) {
    b.lo[k] = 1e300;
    b.hi[k] = -1e300;
  }
  for (int c = 0; c < 8; c++) {
    double l[3];
    for (int k = 0; k < 3; k++) l[k] = sh.bbox_origin[k] + ((c >> k) & 1 ? 1 : -1) * sh.bbox_half[k];
    for (int r = 0; r < 3; r++) {
      double m = nd.global.tr[r] + nd.global.rot[3 * r] * l[0] + nd.global.rot[3 * r + 1] * l[1] +
                 nd.global.rot[3 * r + 2] * l[2];
      b.lo[r] = std::min(b.lo[r], m);
      b.hi[r] = std::max(b.hi[r], m);
    }
  }
  // conservative expansion so FP32 slab tests never cull a true hit
  for (int k = 0; k < 3; k++) {
    b.lo[k] -= 1e-3 + 1e-6 * std::fabs(b.lo[k]);
    b.hi[k] += 1e-3 + 1e-6 * std::fabs(b.hi[k]);
  }
  return b;
}

float down(double v) {
  float f = float(v);
  return double(f) > v ? std::nextafter(f, -INFINITY) : f;
}
float up(double v) {
  float f = float(v);
  return double(f) < v ? std::nextafter(f, INFINITY) : f;
}

// Depth-first BVH with escape links (stackless traversal): node i's first
// child is i + 1; `escape` is the index just past its subtree. Leaves hold
// one scene node. Median split on the longest centroid axis.
struct BvhBuilder {
  std::vector<DBvh> nodes;

  int build(std::vector<std::pair<int, Box>>& items) {
    int root = int(nodes.size());
    fill(items, 0, int(items.size()));
    return root;
  }
  void fill(std::vector<std::pair<int, Box>>& items, int b, int e) {
    Box bb;
    for (int k = 0; k < 3; k++) {
      bb.lo[k] = 1e300;
      bb.hi[k] = -1e300;
    }
    for (int i = b; i < e; i++)
      for (int k = 0; k < 3; k++) {
        bb.lo[k] = std::min(bb.lo[k], items[i].second.lo[k]);
        bb.hi[k] = std::max(bb.hi[k], items[i].second.hi[k]);
      }
    int idx = int(nodes.size());
    nodes.push_back(DBvh{});
    DBvh nd{};
    for (int k = 0; k < 3; k++) {
      nd.lo[k] = down(bb.lo[k]);
      nd.hi[k] = up(bb.hi[k]);
    }
    if (e - b == 1) {
      nd.prim = items[b].first;
      nd.escape = idx + 1;
      nodes[idx] = nd;
      return;
    }
    double cl[3], ch[3];
    for (int k = 0; k < 3; k++) {
      cl[k] = 1e300;
      ch[k] = -1e300;
    }
    for (int i = b; i < e; i++)
      for (int k = 0; k < 3; k++) {
        double c = 0.5 * (items[i].second.lo[k] + items[i].second.hi[k]);
        cl[k] = std::min(cl[k], c);
        ch[k] = std::max(ch[k], c);
      }
    int ax = 0;
    for (int k = 1; k < 3; k++)
      if (ch[k] - cl[k] > ch[ax] - cl[ax]) ax = k;
    int mid = (b + e) / 2;
    std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e,
                     [ax](const std::pair<int, Box>& x, const std::pair<int, Box>& y) {
                       double cx = x.second.lo[ax] + x.second.hi[ax], cy = y.second.lo[ax] + y.second.hi[ax];
                       return cx < cy || (cx == cy && x.first < y.first);
                     });
    fill(items, b, mid);
    fill(items, mid, e);
    nd.prim = -1;
    nd.escape = int(nodes.size());
    nodes[idx] = nd;
  }
};

// Neighbour cells of node n's daughters. A daughter c's cell box C(c) is
// its FP32 box grown by kCellGrow of its extent on each axis; its neighbours
// are the daughters whose FP32 boxes meet C(c), c included. Any daughter
// that contains a point of C(c), or is crossed by a ray inside C(c), has a
// box meeting C(c) (boxes enclose their shapes), so the neighbours are a
// complete candidate set there, whatever the BVH would have visited.
// A cell with more than kCellMaxNbrs neighbours is marked unused.
constexpr int kCellMinChildren = 256;
constexpr int kCellMaxChildren = 1 << 20;
constexpr int kCellMaxNbrs = 48;
// Measured on the Winston cone array (1e7 rays): growth 0.03 / 0.05 / 0.1 /
// 0.2 / 0.3 -> 78.0 / 78.0 / 78.9 / 78.9 / 80.9 ms per trace kernel.
constexpr double kCellGrow = 0.05;

inline void build_cells(const rb_scene_desc* d, const rb_node_desc& n, std::vector<DNode>& nodes,
                        std::vector<DBvh>& cells, std::vector<DBvh>& nbrs) {
  const int m = n.n_children;
  std::vector<DBvh> box(m), cbox(m);
  for (int k = 0; k < m; k++) {
    const int c = d->children[n.first_child + k];
    const Box b = node_box(d, c);
    DBvh e{};
    for (int a = 0; a < 3; a++) {
      e.lo[a] = down(b.lo[a]);
      e.hi[a] = up(b.hi[a]);
    }
    e.prim = c;
    e.escape = 0;
    box[k] = e;
    DBvh g = e;
    for (int a = 0; a < 3; a++) {
      const double grow = kCellGrow * (double(e.hi[a]) - double(e.lo[a]));
      g.lo[a] = down(double(e.lo[a]) - grow);
      g.hi[a] = up(double(e.hi[a]) + grow);
    }
    cbox[k] = g;
  }
  // sweep on x: daughters ordered by box lo.x; the widest box bounds the scan
  std::vector<int> order(m);
  for (int k = 0; k < m; k++) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return box[a].lo[0] < box[b].lo[0] || (box[a].lo[0] == box[b].lo[0] && a < b);
  });
  std::vector<float> lox(m);
  float widest = 0.0f;
  for (int k = 0; k < m; k++) {
    lox[k] = box[order[k]].lo[0];
    widest = std::max(widest, box[order[k]].hi[0] - box[order[k]].lo[0]);
  }
  std::vector<int> found;
  for (int k = 0; k < m; k++) {
    const DBvh& g = cbox[k];
    found.clear();
    const float from = std::nextafter(g.lo[0] - widest, -INFINITY);
    auto it = std::lower_bound(lox.begin(), lox.end(), from);
    for (size_t j = size_t(it - lox.begin()); j < size_t(m) && lox[j] <= g.hi[0]; j++) {
      const DBvh& e = box[order[j]];
      bool meet = true;
      for (int a = 0; a < 3; a++) meet = meet && e.lo[a] <= g.hi[a] && g.lo[a] <= e.hi[a];
      if (meet) found.push_back(order[j]);
    }
    DBvh cell = g;
    const int c = box[k].prim;
    if (int(found.size()) > kCellMaxNbrs) {
      cell.escape = 0;
      cell.prim = -1;
    } else {
      std::sort(found.begin(), found.end(), [&](int a, int b) { return box[a].prim < box[b].prim; });
      cell.escape = int(nbrs.size());
      cell.prim = int(found.size());
      for (int j : found) nbrs.push_back(box[j]);
    }
    nodes[c].cell = int(cells.size());
    cells.push_back(cell);
  }
}

// A TGeoPgon the convex-frustum code takes: full phi and rmin = 0 throughout.
inline bool pgon_is_simple(const double* p) {
  if (p[1] < 360.0) return false;
  const int nz = int(p[3]);
  for (int k = 0; k < nz; k++)
    if (p[4 + 3 * k + 1] != 0.0) return false;
  return true;
}

int csg_depth(const rb_scene_desc* d, int shape, int guard) {
  if (guard > 16) return 99;
  const rb_shape_desc& s = d->shapes[shape];
  if (s.type != RB_SHAPE_COMPOSITE) return 0;
  const double* p = d->params + s.param_offset;
  int l = int(p[1]), r = int(p[2]);
  if (l < 0 || l >= d->n_shapes || r < 0 || r >= d->n_shapes) return 99;
  return 1 + std::max(csg_depth(d, l, guard + 1), csg_depth(d, r, guard + 1));
}

int validate(std::string& err, const rb_scene_desc* d) {
  if (!d || d->n_nodes < 1 || !d->nodes || !d->volumes || !d->shapes || !d->params)
    return fail_msg(err, -1, "scene: empty or null arrays");
  for (int i = 0; i < d->n_shapes; i++) {
    const rb_shape_desc& s = d->shapes[i];
    if (s.param_offset < 0 || s.param_offset + s.n_params > d->n_params)
      return fail_msg(err, -1, "scene: shape parameters out of range");
    switch (s.type) {
      case RB_SHAPE_BOX:
        if (s.n_params != 6) return fail_msg(err, -1, "scene: box needs 6 parameters");
        break;
      case RB_SHAPE_TUBE:
        if (s.n_params != 3) return fail_msg(err, -1, "scene: tube needs 3 parameters");
        break;
      case RB_SHAPE_SPHERE: {
        const double* p = d->params + s.param_offset;
        if (s.n_params != 14) return fail_msg(err, -1, "scene: sphere needs 14 parameters");
        if (!(p[9] > 0.0)) return fail_msg(err, -1, "scene: sphere needs dphi > 0");
        break;
      }
      case RB_SHAPE_PARABOLOID:
        if (s.n_params != 5) return fail_msg(err, -1, "scene: paraboloid needs 5 parameters");
        break;
      case RB_SHAPE_PGON:
      case RB_SHAPE_PCON: {
        const bool pgon = s.type == RB_SHAPE_PGON;
        const double* p = d->params + s.param_offset;
        const int nz = int(p[pgon ? 3 : 2]), ne = pgon ? int(p[2]) : 0;
        if (nz < 2 || (pgon && ne < 1) || s.n_params != (pgon ? 4 : 3) + 3 * nz + 2 * ne + 4)
          return fail_msg(err, -1, pgon ? "scene: TGeoPgon parameter count" : "scene: TGeoPcon parameter count");
        if (!(p[1] > 0.0)) return fail_msg(err, -1, "scene: TGeoPcon/TGeoPgon needs dphi > 0");
        const double* z = p + (pgon ? 4 : 3);
        for (int k = 0; k < nz; k++) {
          if (k > 0 && z[3 * k] < z[3 * (k - 1)]) return fail_msg(err, -1, "scene: TGeoPcon/TGeoPgon sections must not decrease in z");
          if (!(z[3 * k + 1] >= 0.0) || !(z[3 * k + 2] >= z[3 * k + 1]))
            return fail_msg(err, -1, "scene: TGeoPcon/TGeoPgon needs 0 <= rmin <= rmax");
        }
        if (pgon && p[1] >= 360.0 && ne < 3) return fail_msg(err, -1, "scene: a full TGeoPgon needs 3 or more sides");
        break;
      }
      case RB_SHAPE_ASPHERIC: {
        const double* p = d->params + s.param_offset;
        if (s.n_params < 10 || int(p[8]) < 0 || int(p[9]) < 0 || s.n_params != 10 + 2 * (int(p[8]) + int(p[9])))
          return fail_msg(err, -1, "scene: aspheric disk parameter count");
        break;
      }
      case RB_SHAPE_WINSTON: {
        const double* p = d->params + s.param_offset;
        if (s.n_params < 9 || int(p[2]) < 3 || s.n_params != 9 + 2 * int(p[2]))
          return fail_msg(err, -1, "scene: Winston cone parameter count");
        break;
      }
      case RB_SHAPE_WINSTON2D: {
        const double* p = d->params + s.param_offset;
        if (s.n_params != 14 || int(p[2]) != 2 || !(p[13] > 0.0) || !(p[0] > p[1]) || !(p[1] > 0.0))
          return fail_msg(err, -1, "scene: AGeoWinstonCone2D needs 14 parameters, r1 > r2 > 0 and dy > 0");
        break;
      }
      case RB_SHAPE_COMPOSITE:
        if (s.n_params != 27) return fail_msg(err, -1, "scene: composite needs 27 parameters");
        if (csg_depth(d, i, 0) > kMaxCsgDepth)
          return fail_msg(err, -3, "scene: composite nesting deeper than " + std::to_string(kMaxCsgDepth) + " boolean levels");
        break;
      default:
        return fail_msg(err, -3, "scene: unsupported shape type " + std::to_string(s.type));
    }
  }
  for (int i = 0; i < d->n_volumes; i++) {
    const rb_volume_desc& v = d->volumes[i];
    if (v.shape < 0 || v.shape >= d->n_shapes) return fail_msg(err, -1, "scene: volume shape out of range");
    if (v.kind < RB_KLENS || v.kind > RB_KOTHER) return fail_msg(err, -1, "scene: bad volume kind");
    int tabs[5] = {v.refl_table, v.k_table, v.qe_lambda_table, v.qe_angle_table,
                   v.index_model == RB_INDEX_TABLE ? v.index_ref : -1};
    for (int t : tabs)
      if (t >= d->n_tables) return fail_msg(err, -1, "scene: table index out of range");
    if (v.index_model == RB_INDEX_TABLE && v.index_ref < 0) return fail_msg(err, -1, "scene: missing index table");
    if ((v.index_model == RB_INDEX_SELLMEIER || v.index_model == RB_INDEX_SCHOTT) &&
        (v.index_ref < 0 || v.index_ref + 6 > d->n_params))
      return fail_msg(err, -1, "scene: index formula parameters out of range");
    if (v.refl_grid < -1 || v.refl_grid >= std::max(d->n_grids, 0))
      return fail_msg(err, -1, "scene: mirror reflectance table out of range");
    if (v.border_count < 0 || v.border_first < 0 ||
        int64_t(v.border_first) + v.border_count > std::max(d->n_borders, 0) || (v.border_count > 0 && !d->borders))
      return fail_msg(err, -1, "scene: border-condition range out of bounds");
  }
  for (int i = 0; i < d->n_borders; i++) {
    const rb_border_desc& b = d->borders[i];
    if (b.volume2 < -2 || b.volume2 >= d->n_volumes) return fail_msg(err, -1, "scene: border volume out of range");
    if (!(b.sigma >= 0.0) || !std::isfinite(b.sigma)) return fail_msg(err, -1, "scene: border roughness must be >= 0");
    if (b.multilayer < -1 || b.multilayer >= std::max(d->n_grids, 0))
      return fail_msg(err, -1, "scene: border multilayer table out of range");
  }
  if (d->n_grids > 0 && !d->grids) return fail_msg(err, -1, "scene: missing multilayer tables");
  for (int i = 0; i < d->n_grids; i++) {
    const rb_grid2d_desc& g = d->grids[i];
    const int64_t cells = int64_t(g.nx) * g.ny;
    if (g.nx < 1 || g.ny < 1 || cells > (int64_t(1) << 26))
      return fail_msg(err, -1, "scene: multilayer table needs 1 <= nx, ny and nx * ny <= 2^26");
    if (!(g.xhi > g.xlo) || !(g.yhi > g.ylo) || !std::isfinite(g.xhi - g.xlo) || !std::isfinite(g.yhi - g.ylo))
      return fail_msg(err, -1, "scene: multilayer table axes must be finite with hi > lo");
    if (g.r_offset < 0 || g.t_offset < 0 || g.r_offset + cells > d->n_params || g.t_offset + cells > d->n_params)
      return fail_msg(err, -1, "scene: multilayer table values out of range");
  }
  for (int i = 0; i < d->n_tables; i++) {
    const rb_table_desc& t = d->tables[i];
    if (t.offset < 0 || t.n < 0 || 2 * (int64_t(t.offset) + t.n + 4) > d->n_table_xy)
      return fail_msg(err, -1, "scene: table out of range");
    const double* xy = d->table_xy + 2 * t.offset;
    for (int k = 1; k < t.n; k++)
      if (xy[2 * k] < xy[2 * (k - 1)]) return fail_msg(err, -1, "scene: table x must be ascending");
  }
  if (d->nodes[0].parent != -1) return fail_msg(err, -1, "scene: node 0 must be the top node");
  for (int i = 0; i < d->n_nodes; i++) {
    const rb_node_desc& n = d->nodes[i];
    if (n.volume < 0 || n.volume >= d->n_volumes) return fail_msg(err, -1, "scene: node volume out of range");
    if (n.first_child < 0 || n.n_children < 0 || n.first_child + n.n_children > std::max(d->n_children, 0))
      return fail_msg(err, -1, "scene: child range out of bounds");
    for (int k = 0; k < n.n_children; k++) {
      int c = d->children[n.first_child + k];
      if (c <= 0 || c >= d->n_nodes || d->nodes[c].parent != i)
        return fail_msg(err, -1, "scene: inconsistent child list");
      if (k > 0 && c <= d->children[n.first_child + k - 1])
        return fail_msg(err, -1, "scene: child ids must ascend");
    }
  }
  return 0;
}

template <typename X>
size_t put(std::vector<char>& blob, const X* src, size_t count) {
  size_t off = (blob.size() + 255) & ~size_t(255);
  blob.resize(off + sizeof(X) * std::max<size_t>(count, 1));
  if (count) std::memcpy(blob.data() + off, src, sizeof(X) * count);
  return off;
}


// Offsets of the DScene arrays inside the packed blob.
struct SceneLayout {
  size_t nodes, children, bvh, shapes, params, volumes, tables, table_xy, borders, grids, cells, cell_nbrs, sbox;
  int32_t n_nodes;
  bool has_top_box = false;  // the top node has daughters: their bounding box follows
  float top_lo[3] = {0, 0, 0}, top_hi[3] = {0, 0, 0};
  int32_t top_bvh_lo = 0, top_bvh_n = 0;  // the top node's BVH entries
  int32_t max_depth = 0;                  // deepest boolean nesting of any shape
  bool has_lens = false;                  // a placed node is an ALens
  bool has_cells = false;                 // some node's daughters have neighbour cells
  bool has_borders = false;               // border-surface conditions are defined
  bool has_ext = false;                   // a TGeoPcon, a general TGeoPgon, an AGeoWinstonCone2D or a
                                          // TGeoSphere with a phi range (code compiled into the deep
                                          // kernel forms only)
  bool has_solver = false;                // an AGeoAsphericDisk or an AGeoWinstonConePoly
};

// Validate d, build BVHs and pack everything into blob. 0 = ok.
inline int pack_scene(const rb_scene_desc* d, std::vector<char>& blob, SceneLayout& lay, std::string& err) {
  int rc = validate(err, d);
  if (rc) return rc;
  std::vector<DNode> nodes(d->n_nodes);
  BvhBuilder bvh;
  for (int i = 0; i < d->n_nodes; i++) {
    const rb_node_desc& n = d->nodes[i];
    DNode& o = nodes[i];
    std::memcpy(o.rot, n.global.rot, sizeof(o.rot));
    std::memcpy(o.tr, n.global.tr, sizeof(o.tr));
    o.volume = n.volume;
    o.shape = d->volumes[n.volume].shape;
    o.kind = d->volumes[n.volume].kind;
    o.parent = n.parent;
    o.first_child = n.first_child;
    o.n_children = n.n_children;
    o.bvh_root = -1;
    o.cell = -1;
    if (o.kind == RB_KLENS) lay.has_lens = true;
    if (n.n_children > 0) {
      std::vector<std::pair<int, Box>> items;
      for (int k = 0; k < n.n_children; k++) {
        int c = d->children[n.first_child + k];
        items.emplace_back(c, node_box(d, c));
      }
      o.bvh_root = bvh.build(items);
    }
  }
  // neighbour cells (exact candidate sets for a point or a ray segment
  // inside a daughter's cell box): built for nodes with many daughters
  std::vector<DBvh> cells, nbrs;
  for (int i = 0; i < d->n_nodes; i++) {
    const rb_node_desc& n = d->nodes[i];
    if (n.n_children < kCellMinChildren || n.n_children > kCellMaxChildren) continue;
    build_cells(d, n, nodes, cells, nbrs);
  }
  std::vector<DShape> shapes(d->n_shapes);
  // each shape's local box, grown as node boxes are, for the boolean nodes'
  // child tests
  std::vector<double> sbox(size_t(std::max(d->n_shapes, 0)) * 6);
  for (int i = 0; i < d->n_shapes; i++)
    for (int k = 0; k < 3; k++) {
      const double lo = d->shapes[i].bbox_origin[k] - d->shapes[i].bbox_half[k];
      const double hi = d->shapes[i].bbox_origin[k] + d->shapes[i].bbox_half[k];
      sbox[size_t(i) * 6 + k] = lo - (1e-3 + 1e-6 * std::fabs(lo));
      sbox[size_t(i) * 6 + 3 + k] = hi + (1e-3 + 1e-6 * std::fabs(hi));
    }
  lay.max_depth = 0;
  for (int i = 0; i < d->n_shapes; i++) {
    shapes[i] = DShape{d->shapes[i].type, d->shapes[i].param_offset, csg_depth(d, i, 0), 0};
    if (shapes[i].type == RB_SHAPE_PGON && !pgon_is_simple(d->params + d->shapes[i].param_offset))
      shapes[i].type = kShapePgonGeneral;  // rmin > 0 or a phi range: the general polygon code
    if (shapes[i].type == RB_SHAPE_SPHERE && d->params[d->shapes[i].param_offset + 9] < 360.0)
      shapes[i].type = kShapeSphereWedge;  // a phi range: the wedge-constrained sphere code
    if (shapes[i].type == kShapePgonGeneral || shapes[i].type == RB_SHAPE_PCON ||
        shapes[i].type == RB_SHAPE_WINSTON2D || shapes[i].type == kShapeSphereWedge)
      lay.has_ext = true;
    if (shapes[i].type == RB_SHAPE_ASPHERIC || shapes[i].type == RB_SHAPE_WINSTON) lay.has_solver = true;
    lay.max_depth = std::max(lay.max_depth, shapes[i].depth);
  }
  blob.clear();
  lay.nodes = put(blob, nodes.data(), nodes.size());
  lay.children = put(blob, d->children, size_t(std::max(d->n_children, 0)));
  lay.bvh = put(blob, bvh.nodes.data(), bvh.nodes.size());
  lay.shapes = put(blob, shapes.data(), shapes.size());
  lay.params = put(blob, d->params, size_t(d->n_params));
  lay.volumes = put(blob, d->volumes, size_t(d->n_volumes));
  lay.tables = put(blob, d->tables, size_t(std::max(d->n_tables, 0)));
  lay.table_xy = put(blob, d->table_xy, size_t(std::max(d->n_table_xy, 0)));
  lay.borders = put(blob, d->borders, size_t(std::max(d->n_borders, 0)));
  lay.grids = put(blob, d->grids, size_t(std::max(d->n_grids, 0)));
  lay.cells = put(blob, cells.data(), cells.size());
  lay.has_cells = !cells.empty();
  lay.cell_nbrs = put(blob, nbrs.data(), nbrs.size());
  lay.sbox = put(blob, sbox.data(), sbox.size());
  // the surface-extras kernel form also reads TH2 mirror reflectance
  lay.has_borders = d->n_borders > 0;
  for (int i = 0; i < d->n_volumes; i++)
    if (d->volumes[i].refl_grid >= 0) lay.has_borders = true;
  lay.n_nodes = d->n_nodes;
  lay.has_top_box = d->n_nodes > 0 && nodes[0].bvh_root >= 0;
  if (lay.has_top_box) {
    const DBvh& r = bvh.nodes[size_t(nodes[0].bvh_root)];
    for (int k = 0; k < 3; k++) {
      lay.top_lo[k] = r.lo[k];
      lay.top_hi[k] = r.hi[k];
    }
    lay.top_bvh_lo = nodes[0].bvh_root;
    lay.top_bvh_n = r.escape - nodes[0].bvh_root;
  }
  return 0;
}

// DScene view of a packed blob at `base` (host or device address).
inline DScene scene_view(const char* base, const SceneLayout& lay) {
  DScene s;
  s.nodes = reinterpret_cast<const DNode*>(base + lay.nodes);
  s.children = reinterpret_cast<const int32_t*>(base + lay.children);
  s.bvh = reinterpret_cast<const DBvh*>(base + lay.bvh);
  s.shapes = reinterpret_cast<const DShape*>(base + lay.shapes);
  s.params = reinterpret_cast<const double*>(base + lay.params);
  s.volumes = reinterpret_cast<const rb_volume_desc*>(base + lay.volumes);
  s.tables = reinterpret_cast<const rb_table_desc*>(base + lay.tables);
  s.table_xy = reinterpret_cast<const double*>(base + lay.table_xy);
  s.borders = reinterpret_cast<const rb_border_desc*>(base + lay.borders);
  s.grids = reinterpret_cast<const rb_grid2d_desc*>(base + lay.grids);
  s.cells = reinterpret_cast<const DBvh*>(base + lay.cells);
  s.cell_nbrs = reinterpret_cast<const DBvh*>(base + lay.cell_nbrs);
  s.sbox = reinterpret_cast<const double*>(base + lay.sbox);
  s.n_nodes = lay.n_nodes;
  s.stage_lo = 0;
  s.stage_n = 0;
  return s;
}

}  // namespace rbd
