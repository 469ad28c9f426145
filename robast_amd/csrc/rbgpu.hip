// rbgpu.hip -- librbgpu: C ABI (include/rbgpu.h), scene upload with per-node
// BVH build, ray sources and the trace kernel launch for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rbgpu.h"
#include "rb_trace.h"

using namespace rbd;

struct rb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  int launches = 0;
  std::string err;
};

struct rb_scene {
  rb_ctx* ctx = nullptr;
  void* dmem = nullptr;
  DScene s{};
};

namespace {

int fail(rb_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, -2, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

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

struct BvhBuilder {
  std::vector<DBvh> nodes;
  std::vector<int32_t> prims;

  // Median split on the longest centroid axis; leaves hold <= 2 prims.
  // Returns the index of the subtree root; inner nodes keep both children
  // adjacent (right = left + 1).
  int build(std::vector<std::pair<int, Box>>& items, int b, int e) {
    int idx = int(nodes.size());
    nodes.push_back(DBvh{});
    fill(idx, items, b, e);
    return idx;
  }
  void fill(int idx, std::vector<std::pair<int, Box>>& items, int b, int e) {
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
    DBvh nd{};
    for (int k = 0; k < 3; k++) {
      nd.lo[k] = down(bb.lo[k]);
      nd.hi[k] = up(bb.hi[k]);
    }
    if (e - b <= 2) {
      nd.left = int(prims.size());
      nd.count = e - b;
      for (int i = b; i < e; i++) prims.push_back(items[i].first);
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
                       return x.second.lo[ax] + x.second.hi[ax] < y.second.lo[ax] + y.second.hi[ax];
                     });
    int left = int(nodes.size());
    nodes.push_back(DBvh{});
    nodes.push_back(DBvh{});
    nd.left = left;
    nd.count = 0;
    nodes[idx] = nd;
    fill(left, items, b, mid);
    fill(left + 1, items, mid, e);
  }
};

int csg_depth(const rb_scene_desc* d, int shape, int guard) {
  if (guard > 16) return 99;
  const rb_shape_desc& s = d->shapes[shape];
  if (s.type != RB_SHAPE_COMPOSITE) return 0;
  const double* p = d->params + s.param_offset;
  int l = int(p[1]), r = int(p[2]);
  if (l < 0 || l >= d->n_shapes || r < 0 || r >= d->n_shapes) return 99;
  return 1 + std::max(csg_depth(d, l, guard + 1), csg_depth(d, r, guard + 1));
}

int validate(rb_ctx* ctx, const rb_scene_desc* d) {
  if (!d || d->n_nodes < 1 || !d->nodes || !d->volumes || !d->shapes || !d->params)
    return fail(ctx, -1, "scene: empty or null arrays");
  for (int i = 0; i < d->n_shapes; i++) {
    const rb_shape_desc& s = d->shapes[i];
    if (s.param_offset < 0 || s.param_offset + s.n_params > d->n_params)
      return fail(ctx, -1, "scene: shape parameters out of range");
    switch (s.type) {
      case RB_SHAPE_BOX:
      case RB_SHAPE_TUBE:
      case RB_SHAPE_SPHERE:
      case RB_SHAPE_PARABOLOID:
        break;
      case RB_SHAPE_PGON: {
        const double* p = d->params + s.param_offset;
        int nz = int(p[3]);
        if (nz < 2 || int(p[2]) < 3 || s.n_params != 4 + 3 * nz || p[1] != 360.0)
          return fail(ctx, -3, "scene: unsupported TGeoPgon parameters");
        for (int k = 0; k < nz; k++)
          if (p[4 + 3 * k + 1] != 0.0) return fail(ctx, -3, "scene: TGeoPgon with rmin > 0 is unsupported");
        break;
      }
      case RB_SHAPE_COMPOSITE:
        if (s.n_params != 27) return fail(ctx, -1, "scene: composite needs 27 parameters");
        if (csg_depth(d, i, 0) > kMaxCsgDepth)
          return fail(ctx, -3, "scene: composite nesting deeper than 3 boolean levels");
        break;
      default:
        return fail(ctx, -3, "scene: unsupported shape type " + std::to_string(s.type));
    }
  }
  for (int i = 0; i < d->n_volumes; i++) {
    const rb_volume_desc& v = d->volumes[i];
    if (v.shape < 0 || v.shape >= d->n_shapes) return fail(ctx, -1, "scene: volume shape out of range");
    if (v.kind < RB_KLENS || v.kind > RB_KOTHER) return fail(ctx, -1, "scene: bad volume kind");
    int tabs[5] = {v.refl_table, v.k_table, v.qe_lambda_table, v.qe_angle_table,
                   v.index_model == RB_INDEX_TABLE ? v.index_ref : -1};
    for (int t : tabs)
      if (t >= d->n_tables) return fail(ctx, -1, "scene: table index out of range");
    if (v.index_model == RB_INDEX_TABLE && v.index_ref < 0) return fail(ctx, -1, "scene: missing index table");
    if ((v.index_model == RB_INDEX_SELLMEIER || v.index_model == RB_INDEX_SCHOTT) &&
        (v.index_ref < 0 || v.index_ref + 6 > d->n_params))
      return fail(ctx, -1, "scene: index formula parameters out of range");
  }
  for (int i = 0; i < d->n_tables; i++) {
    const rb_table_desc& t = d->tables[i];
    if (t.offset < 0 || t.n < 0 || 2 * (t.offset + t.n) > d->n_table_xy)
      return fail(ctx, -1, "scene: table out of range");
    const double* xy = d->table_xy + 2 * t.offset;
    for (int k = 1; k < t.n; k++)
      if (xy[2 * k] < xy[2 * (k - 1)]) return fail(ctx, -1, "scene: table x must be ascending");
  }
  if (d->nodes[0].parent != -1) return fail(ctx, -1, "scene: node 0 must be the top node");
  for (int i = 0; i < d->n_nodes; i++) {
    const rb_node_desc& n = d->nodes[i];
    if (n.volume < 0 || n.volume >= d->n_volumes) return fail(ctx, -1, "scene: node volume out of range");
    if (n.first_child < 0 || n.n_children < 0 || n.first_child + n.n_children > std::max(d->n_children, 0))
      return fail(ctx, -1, "scene: child range out of bounds");
    for (int k = 0; k < n.n_children; k++) {
      int c = d->children[n.first_child + k];
      if (c <= 0 || c >= d->n_nodes || d->nodes[c].parent != i)
        return fail(ctx, -1, "scene: inconsistent child list");
      if (k > 0 && c <= d->children[n.first_child + k - 1])
        return fail(ctx, -1, "scene: child ids must ascend");
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

}  // namespace

namespace {

struct DevRays {
  std::vector<void*> allocs;
  ~DevRays() {
    for (void* p : allocs) (void)hipFree(p);
  }
  template <typename X>
  hipError_t alloc(X** p, int64_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(X) * size_t(std::max<int64_t>(n, 1)));
    if (e == hipSuccess) allocs.push_back(q);
    *p = static_cast<X*>(q);
    return e;
  }
};

}  // namespace

// ---------------------------------------------------------------- kernels
__global__ void __launch_bounds__(256) k_trace(DScene S, RayIO io, int64_t n, TraceParams P) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (io.status[i] != RB_RUN) return;
  Lane L;
  L.p = mk(io.x[i], io.y[i], io.z[i]);
  L.d = mk(io.dx[i], io.dy[i], io.dz[i]);
  L.t = io.t[i];
  L.lambda = io.lambda[i];
  L.status = RB_RUN;
  L.npoints = io.npoints[i];
  L.last_node = io.last_node ? io.last_node[i] : -1;
  L.draw = 0;
  L.step = 0;
  L.hit_node = -1;
  L.hit_leaf = 0;
  trace_lane(S, P, io, i, L);
  io.x[i] = L.p.x;
  io.y[i] = L.p.y;
  io.z[i] = L.p.z;
  io.t[i] = L.t;
  io.dx[i] = L.d.x;
  io.dy[i] = L.d.y;
  io.dz[i] = L.d.z;
  io.status[i] = L.status;
  io.npoints[i] = L.npoints;
  if (io.last_node) io.last_node[i] = L.last_node;
}

__global__ void __launch_bounds__(256) k_generate(rb_raygen_desc g, RayIO io, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double th, lam, px, py;
  if (g.kind == RB_RAYGEN_GRID) {
    int ns = g.n_side;
    double delta = ns == 1 ? g.size / 2 : g.size / (ns - 1);
    int64_t a = i / ns, b = i % ns;
    px = a * delta - g.size / 2;
    py = b * delta - g.size / 2;
    th = g.theta_min;
    lam = g.lambda_min;
  } else {
    double u0 = uniform01(g.seed, uint64_t(i), 0u, 1u), u1 = uniform01(g.seed, uint64_t(i), 1u, 1u);
    double u2 = uniform01(g.seed, uint64_t(i), 2u, 1u), u3 = uniform01(g.seed, uint64_t(i), 3u, 1u);
    th = g.theta_min + (g.theta_max - g.theta_min) * u2;
    lam = g.lambda_min + (g.lambda_max - g.lambda_min) * u3;
    double rr = 0.5 * g.size * sqrt(u0), ph = 2 * kPi * u1;
    px = rr * cos(ph);
    py = rr * sin(ph);
  }
  double st = sin(th), ct = cos(th), t2 = kPi - th;
  double cx = -g.distance * st, cz = g.distance * ct;
  if (g.kind == RB_RAYGEN_GRID) {
    io.x[i] = px + cx;
    io.y[i] = py;
    io.z[i] = cz;
  } else {
    io.x[i] = cx + px * ct;
    io.y[i] = py;
    io.z[i] = cz + px * st;
  }
  io.t[i] = 0;
  io.dx[i] = sin(t2);
  io.dy[i] = 0;
  io.dz[i] = cos(t2);
  const_cast<double*>(io.lambda)[i] = lam;
  io.status[i] = RB_RUN;
  io.npoints[i] = 1;
  if (io.last_node) io.last_node[i] = -1;
}

// ---------------------------------------------------------------- C ABI
extern "C" {

int rbgpu_ctx_create(int device, rb_ctx** out) {
  if (!out) return -1;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return -2;
  if (device < 0 || device >= count) return -1;
  rb_ctx* c = new rb_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return -2;
  }
  *out = c;
  return 0;
}

void rbgpu_ctx_destroy(rb_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* rbgpu_last_error(const rb_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rbgpu_abi_sizes(int32_t* out) {
  out[0] = sizeof(rb_matrix);
  out[1] = sizeof(rb_shape_desc);
  out[2] = sizeof(rb_volume_desc);
  out[3] = sizeof(rb_node_desc);
  out[4] = sizeof(rb_table_desc);
  out[5] = sizeof(rb_scene_desc);
  out[6] = sizeof(rb_rays);
  out[7] = sizeof(rb_trace_opts);
  out[8] = sizeof(rb_raygen_desc);
  return 9;
}

int rbgpu_scene_upload(rb_ctx* ctx, const rb_scene_desc* d, rb_scene** out) {
  if (!ctx || !out) return -1;
  *out = nullptr;
  int rc = validate(ctx, d);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));

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
    o.pad = 0;
    if (n.n_children > 0) {
      std::vector<std::pair<int, Box>> items;
      for (int k = 0; k < n.n_children; k++) {
        int c = d->children[n.first_child + k];
        items.emplace_back(c, node_box(d, c));
      }
      o.bvh_root = bvh.build(items, 0, int(items.size()));
    }
  }
  std::vector<DShape> shapes(d->n_shapes);
  for (int i = 0; i < d->n_shapes; i++) shapes[i] = DShape{d->shapes[i].type, d->shapes[i].param_offset};

  std::vector<char> blob;
  size_t o_nodes = put(blob, nodes.data(), nodes.size());
  size_t o_children = put(blob, d->children, size_t(std::max(d->n_children, 0)));
  size_t o_bvh = put(blob, bvh.nodes.data(), bvh.nodes.size());
  size_t o_prims = put(blob, bvh.prims.data(), bvh.prims.size());
  size_t o_shapes = put(blob, shapes.data(), shapes.size());
  size_t o_params = put(blob, d->params, size_t(d->n_params));
  size_t o_vol = put(blob, d->volumes, size_t(d->n_volumes));
  size_t o_tab = put(blob, d->tables, size_t(std::max(d->n_tables, 0)));
  size_t o_txy = put(blob, d->table_xy, size_t(std::max(d->n_table_xy, 0)));

  rb_scene* sc = new rb_scene();
  sc->ctx = ctx;
  hipError_t e = hipMalloc(&sc->dmem, blob.size());
  if (e != hipSuccess) {
    delete sc;
    return fail(ctx, -2, std::string("hipMalloc scene: ") + hipGetErrorString(e));
  }
  e = hipMemcpy(sc->dmem, blob.data(), blob.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(sc->dmem);
    delete sc;
    return fail(ctx, -2, std::string("hipMemcpy scene: ") + hipGetErrorString(e));
  }
  char* base = static_cast<char*>(sc->dmem);
  sc->s.nodes = reinterpret_cast<const DNode*>(base + o_nodes);
  sc->s.children = reinterpret_cast<const int32_t*>(base + o_children);
  sc->s.bvh = reinterpret_cast<const DBvh*>(base + o_bvh);
  sc->s.bvh_prims = reinterpret_cast<const int32_t*>(base + o_prims);
  sc->s.shapes = reinterpret_cast<const DShape*>(base + o_shapes);
  sc->s.params = reinterpret_cast<const double*>(base + o_params);
  sc->s.volumes = reinterpret_cast<const rb_volume_desc*>(base + o_vol);
  sc->s.tables = reinterpret_cast<const rb_table_desc*>(base + o_tab);
  sc->s.table_xy = reinterpret_cast<const double*>(base + o_txy);
  sc->s.n_nodes = d->n_nodes;
  *out = sc;
  return 0;
}

void rbgpu_scene_free(rb_scene* s) {
  if (!s) return;
  if (s->dmem) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipFree(s->dmem);
  }
  delete s;
}

int rbgpu_trace(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, rb_rays* rays, int64_t n) {
  if (!ctx) return -1;
  if (!scene || !opts || !rays || n < 0) return fail(ctx, -1, "rbgpu_trace: null argument or negative n");
  if (opts->precision != RB_PRECISION_FP64 && opts->precision != RB_PRECISION_MIXED)
    return fail(ctx, -1, "rbgpu_trace: unknown precision mode");
  if (!rays->x || !rays->y || !rays->z || !rays->t || !rays->dx || !rays->dy || !rays->dz || !rays->lambda ||
      !rays->status || !rays->npoints)
    return fail(ctx, -1, "rbgpu_trace: missing ray array");
  if (rays->path_capacity < 0 || (rays->path_capacity > 0 && (!rays->path_xyzt || !rays->path_node)))
    return fail(ctx, -1, "rbgpu_trace: bad path record");
  ctx->launches = 0;
  ctx->last_ms = 0;
  if (n == 0) return 0;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TraceParams P;
  P.seed = opts->seed;
  P.id_base = opts->ray_id_base;
  P.limit = opts->limit > 0 ? opts->limit : 100;
  P.fresnel = opts->fresnel ? 1 : 0;

  RayIO io;
  DevRays dev;
  const bool host = !opts->device_ptrs;
  const size_t nd = size_t(n) * sizeof(double), ni = size_t(n) * sizeof(int32_t);
  const int64_t cap = rays->path_capacity;
  if (host) {
    double* f[8];
    for (int k = 0; k < 8; k++) HIPCHK(ctx, dev.alloc(&f[k], n));
    int32_t* g[3];
    for (int k = 0; k < 3; k++) HIPCHK(ctx, dev.alloc(&g[k], n));
    io.x = f[0]; io.y = f[1]; io.z = f[2]; io.t = f[3];
    io.dx = f[4]; io.dy = f[5]; io.dz = f[6]; io.lambda = f[7];
    io.status = g[0]; io.npoints = g[1]; io.last_node = g[2];
    const double* hs[8] = {rays->x, rays->y, rays->z, rays->t, rays->dx, rays->dy, rays->dz, rays->lambda};
    for (int k = 0; k < 8; k++) HIPCHK(ctx, hipMemcpyAsync(f[k], hs[k], nd, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(g[0], rays->status, ni, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(g[1], rays->npoints, ni, hipMemcpyHostToDevice, ctx->stream));
    if (rays->last_node)
      HIPCHK(ctx, hipMemcpyAsync(g[2], rays->last_node, ni, hipMemcpyHostToDevice, ctx->stream));
    else
      io.last_node = nullptr;
    io.path_cap = int32_t(cap);
    io.path_xyzt = nullptr;
    io.path_node = nullptr;
    if (cap > 0) {
      HIPCHK(ctx, dev.alloc(&io.path_xyzt, n * cap * 4));
      HIPCHK(ctx, dev.alloc(&io.path_node, n * cap));
      HIPCHK(ctx, hipMemcpyAsync(io.path_xyzt, rays->path_xyzt, nd * cap * 4, hipMemcpyHostToDevice, ctx->stream));
      HIPCHK(ctx, hipMemcpyAsync(io.path_node, rays->path_node, ni * cap, hipMemcpyHostToDevice, ctx->stream));
    }
  } else {
    io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
    io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
    io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
    io.path_cap = int32_t(cap);
    io.path_xyzt = rays->path_xyzt;
    io.path_node = rays->path_node;
  }

  const int block = 256;
  const int64_t grid = (n + block - 1) / block;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_trace, dim3(unsigned(grid)), dim3(block), 0, ctx->stream, scene->s, io, n, P);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->launches = 1;

  if (host) {
    double* hs[7] = {rays->x, rays->y, rays->z, rays->t, rays->dx, rays->dy, rays->dz};
    double* ds[7] = {io.x, io.y, io.z, io.t, io.dx, io.dy, io.dz};
    for (int k = 0; k < 7; k++) HIPCHK(ctx, hipMemcpyAsync(hs[k], ds[k], nd, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(rays->status, io.status, ni, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(rays->npoints, io.npoints, ni, hipMemcpyDeviceToHost, ctx->stream));
    if (rays->last_node)
      HIPCHK(ctx, hipMemcpyAsync(rays->last_node, io.last_node, ni, hipMemcpyDeviceToHost, ctx->stream));
    if (cap > 0) {
      HIPCHK(ctx, hipMemcpyAsync(rays->path_xyzt, io.path_xyzt, nd * cap * 4, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipMemcpyAsync(rays->path_node, io.path_node, ni * cap, hipMemcpyDeviceToHost, ctx->stream));
    }
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return 0;
}

int rbgpu_generate_rays(rb_ctx* ctx, const rb_raygen_desc* g, rb_rays* rays, int64_t n) {
  if (!ctx) return -1;
  if (!g || !rays || n < 0) return fail(ctx, -1, "rbgpu_generate_rays: null argument");
  if (g->kind != RB_RAYGEN_GRID && g->kind != RB_RAYGEN_DISK) return fail(ctx, -1, "rbgpu_generate_rays: bad kind");
  if (g->kind == RB_RAYGEN_GRID && (g->n_side < 1 || int64_t(g->n_side) * g->n_side < n))
    return fail(ctx, -1, "rbgpu_generate_rays: grid smaller than n");
  if (n == 0) return 0;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  RayIO io{};
  io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
  io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
  io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
  const int block = 256;
  const int64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_generate, dim3(unsigned(grid)), dim3(block), 0, ctx->stream, *g, io, n);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

double rbgpu_last_trace_ms(const rb_ctx* c) { return c ? c->last_ms : 0.0; }
int rbgpu_last_trace_launches(const rb_ctx* c) { return c ? c->launches : 0; }

}  // extern "C"
