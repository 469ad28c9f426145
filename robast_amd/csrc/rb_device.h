// rb_device.h -- device-side tracer for librbgpu (gfx950).
//
// One lane owns one ray. Per segment the lane (1) finds the next boundary in
// its current node: exit distance of the node's shape plus entry distances of
// the daughters that a BVH over their global AABBs cannot cull, (2) crosses
// and relocates the way TGeoNavigator::FindNextBoundaryAndStep /
// CrossBoundaryAndLocate do, (3) applies the optics of
// AOpticsManager::TraceNonSequential (src/AOpticsManager.cxx:335-520).
// Shape code is templated on the arithmetic type T (double for the plain
// form); the scene itself is always FP64.
#pragma once

#ifndef RB_HOST_EMULATION
#include <hip/hip_runtime.h>
#else
#include "rb_host_emulation.h"  // CPU build of the device code, for debugging and tests
#endif
#include <stdint.h>

#include "../../include/rbgpu.h"

// The aspheric and Winston-cone solvers (Newton iterations, per-side
// quadratics) are the register-heaviest primitives; kept as calls
// (RB_HEAVY_FN) they do not size the frame of every primitive dispatch.
#ifndef RB_HEAVY_FN
#define RB_HEAVY_FN
#endif

namespace rbd {

constexpr double kBig = 1e30;        // TGeoShape::Big()
constexpr double kTol = 1e-10;       // TGeoShape::Tolerance()
constexpr double kEps = 1e-6;        // AOpticsManager kEpsilon (src/AOpticsManager.cxx:20)
constexpr double kC = 2.99792458e10; // speed of light, cm/s
constexpr double kBoolPush = 1e-8;   // inside-test offset in boolean-shape iteration
constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxCsgDepth = 3;   // boolean nesting levels unrolled on the device

// ------------------------------------------------------------ scene on HBM
struct DNode {                 // 128 B
  double rot[9];
  double tr[3];
  int32_t shape, kind, volume, parent;
  int32_t first_child, n_children, bvh_root;
  int32_t cell;                // this node's neighbour cell among its parent's daughters, or -1
};
struct DBvh {                  // 32 B, FP32 boxes expanded outward, depth-first order
  float lo[3];
  int32_t escape;              // next node when this box is missed (end of its subtree)
  float hi[3];
  int32_t prim;                // leaf: scene node id; inner: -1 (first child is the next entry)
};
struct DShape {
  int32_t type, off;
  int32_t depth;  // boolean nesting levels below this shape (0 = primitive)
  int32_t pad;
};
// Shape sets a kernel form is compiled for (the primitive dispatch leaves the
// others out): kShapesExt the TGeoPcon, general TGeoPgon, AGeoWinstonCone2D
// and phi-limited TGeoSphere code; kShapesSolver the AGeoAsphericDisk and
// AGeoWinstonConePoly code.
constexpr unsigned kShapesExt = 1u, kShapesSolver = 2u;
struct DScene {
  const DNode* nodes;
  const int32_t* children;
  const DBvh* bvh;
  const DShape* shapes;
  const double* params;
  const rb_volume_desc* volumes;
  const rb_table_desc* tables;
  const double* table_xy;
  const rb_border_desc* borders;
  const rb_grid2d_desc* grids;  // multilayer (wavelength, angle) tables
  // Neighbour cells of the daughters of nodes with many daughters. cells[k]
  // (DBvh layout): the cell box C (the daughter's box grown by a tenth of
  // its extent per axis), escape = first neighbour record, prim = neighbour
  // count (-1: too many, not used). cell_nbrs: one record per daughter whose
  // box meets C (the daughter itself included): its box and prim = node id.
  const DBvh* cells;
  const DBvh* cell_nbrs;
  // per shape, its local bounding box grown by 1e-3 cm (+1e-6 relative):
  // lo[3], hi[3]. A boolean node tests a child's box before the child.
  const double* sbox;
  int32_t n_nodes;
  // BVH entries [stage_lo, stage_lo + stage_n) -- the top node's BVH, at
  // most kBvhStage entries -- are read from an LDS copy that each tracing
  // kernel makes at entry (stage_n = 0: none).
  int32_t stage_lo, stage_n;
  // deepest boolean nesting the navigation code is compiled for, and
  // whether it handles border-surface conditions (roughness, Lambertian)
  static constexpr int max_depth = kMaxCsgDepth;
  static constexpr bool borders_on = true;
  static constexpr bool ext_on = true;
  static constexpr bool lens_on = true;
  static constexpr unsigned shape_set = kShapesExt | kShapesSolver;
};
constexpr int kBvhStage = 256;  // 8 KB of DBvh per block
// The same scene, for kernels compiled for scenes whose composites nest at
// most MAXD levels: the deeper composite call sites, and the registers they
// make the kernel spill around them, are left out; likewise the border-surface
// code unless BORDERS, and the TGeoPcon / general TGeoPgon /
// AGeoWinstonCone2D code unless EXT (compiled into the deep forms, with or
// without border surfaces; a scene holding such a shape is traced with a
// deep form).
// BOX: boolean nodes test a child's bounding box before evaluating the child
// (compiled into the work-queue kernels: in the lane kernel the extra code
// cost more in register spills than it saved).
// LENS: the refracting-volume physics (absorption in the volume, Fresnel
// reflection and refraction at its faces); the lane kernels for scenes with
// no ALens are compiled without it (a smaller frame: 288 bytes against 336).
// SOLVER: the AGeoAsphericDisk and AGeoWinstonConePoly code (iterative
// roots); the lane kernels for scenes with neither (and no ALens) are
// compiled without it (160 bytes against 288, no VGPR spills).
template <int MAXD, bool BORDERS = false, bool EXT = (MAXD > 1), bool BOX = false, bool LENS = true,
          bool SOLVER = true>
struct DSceneOf : DScene {
  static constexpr int max_depth = MAXD;
  static constexpr bool borders_on = BORDERS;
  static constexpr bool ext_on = EXT;
  static constexpr bool box_on = BOX;
  static constexpr bool lens_on = LENS;
  static constexpr unsigned shape_set = (EXT ? kShapesExt : 0u) | (SOLVER ? kShapesSolver : 0u);
};

template <typename T>
struct V3 {
  T x, y, z;
};
template <typename T> __device__ __forceinline__ V3<T> mk(T x, T y, T z) { return {x, y, z}; }
template <typename T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> axpy(V3<T> p, T s, V3<T> d) {
  return {p.x + s * d.x, p.y + s * d.y, p.z + s * d.z};
}
template <typename T> __device__ __forceinline__ V3<T> scale(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }

// master -> local of a 4x3 matrix stored as rot[9], tr[3] (master = R local + T)
template <typename T>
__device__ __forceinline__ V3<T> m2l(const double* rot, const double* tr, V3<T> p) {
  T qx = p.x - T(tr[0]), qy = p.y - T(tr[1]), qz = p.z - T(tr[2]);
  return {T(rot[0]) * qx + T(rot[3]) * qy + T(rot[6]) * qz, T(rot[1]) * qx + T(rot[4]) * qy + T(rot[7]) * qz,
          T(rot[2]) * qx + T(rot[5]) * qy + T(rot[8]) * qz};
}
template <typename T>
__device__ __forceinline__ V3<T> m2lv(const double* rot, V3<T> q) {
  return {T(rot[0]) * q.x + T(rot[3]) * q.y + T(rot[6]) * q.z, T(rot[1]) * q.x + T(rot[4]) * q.y + T(rot[7]) * q.z,
          T(rot[2]) * q.x + T(rot[5]) * q.y + T(rot[8]) * q.z};
}
template <typename T>
__device__ __forceinline__ V3<T> l2mv(const double* rot, V3<T> q) {
  return {T(rot[0]) * q.x + T(rot[1]) * q.y + T(rot[2]) * q.z, T(rot[3]) * q.x + T(rot[4]) * q.y + T(rot[5]) * q.z,
          T(rot[6]) * q.x + T(rot[7]) * q.y + T(rot[8]) * q.z};
}

struct Hit {
  double s;
  uint32_t leaf;     // composite path: bit k = right child at depth k; depth in bits 24..31
  uint32_t exiting;  // 1 if the crossing leaves the leaf primitive, 0 if it enters it
};
__device__ __forceinline__ uint32_t leaf_push(uint32_t leaf, int side) {
  uint32_t depth = leaf >> 24;
  return ((depth + 1u) << 24) | (((leaf & 0xffffffu) << 1) | uint32_t(side));
}

// ------------------------------------------------------------ box
// params: dx dy dz ox oy oz
template <typename T> __device__ bool box_contains(const double* p, V3<T> q) {
  return fabs(q.x - T(p[3])) <= T(p[0]) && fabs(q.y - T(p[4])) <= T(p[1]) && fabs(q.z - T(p[5])) <= T(p[2]);
}
template <typename T> __device__ T box_dist_in(const double* p, V3<T> q, V3<T> d) {
  T n[3] = {q.x - T(p[3]), q.y - T(p[4]), q.z - T(p[5])};
  T dd[3] = {d.x, d.y, d.z};
  T smin = T(kBig);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (dd[i] != T(0)) {
      T s = (dd[i] > T(0) ? T(p[i]) - n[i] : T(p[i]) + n[i]) / fabs(dd[i]);
      smin = s < smin ? s : smin;
    }
  }
  return smin < T(0) ? T(0) : smin;
}
template <typename T> __device__ T box_dist_out(const double* p, V3<T> q, V3<T> d) {
  T n[3] = {q.x - T(p[3]), q.y - T(p[4]), q.z - T(p[5])};
  T dd[3] = {d.x, d.y, d.z};
  T h[3] = {T(p[0]), T(p[1]), T(p[2])};
  T saf[3];
  bool in = true;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    saf[i] = fabs(n[i]) - h[i];
    in = in && !(saf[i] > T(0));
  }
  if (in) return T(0);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (saf[i] < T(0) || n[i] * dd[i] >= T(0)) continue;
    T s = saf[i] / fabs(dd[i]);
    int j = (i + 1) % 3, k = (i + 2) % 3;
    if (fabs(n[j] + s * dd[j]) > h[j] || fabs(n[k] + s * dd[k]) > h[k]) continue;
    return s;
  }
  return T(kBig);
}
template <typename T> __device__ V3<T> box_normal(const double* p, V3<T> q, V3<T> d) {
  T s0 = fabs(fabs(q.x - T(p[3])) - T(p[0])), s1 = fabs(fabs(q.y - T(p[4])) - T(p[1])),
    s2 = fabs(fabs(q.z - T(p[5])) - T(p[2]));
  int i = 0;
  T m = s0;
  if (s1 < m) { i = 1; m = s1; }
  if (s2 < m) { i = 2; }
  if (i == 0) return mk<T>(d.x > T(0) ? T(1) : T(-1), T(0), T(0));
  if (i == 1) return mk<T>(T(0), d.y > T(0) ? T(1) : T(-1), T(0));
  return mk<T>(T(0), T(0), d.z > T(0) ? T(1) : T(-1));
}

// ------------------------------------------------------------ tube
// params: rmin rmax dz
template <typename T> __device__ bool cyl_roots(V3<T> q, V3<T> d, T R, T& s1, T& s2) {
  T nsq = d.x * d.x + d.y * d.y;
  if (nsq < T(1e-20)) return false;
  T b = (q.x * d.x + q.y * d.y) / nsq;
  T c = (q.x * q.x + q.y * q.y - R * R) / nsq;
  T disc = b * b - c;
  if (disc <= T(0)) return false;
  T sq = sqrt(disc);
  s1 = -b - sq;
  s2 = -b + sq;
  return true;
}
template <typename T> __device__ bool tube_contains(const double* p, V3<T> q) {
  if (fabs(q.z) > T(p[2])) return false;
  T r2 = q.x * q.x + q.y * q.y;
  return r2 >= T(p[0]) * T(p[0]) && r2 <= T(p[1]) * T(p[1]);
}
template <typename T> __device__ T tube_dist_in(const double* p, V3<T> q, V3<T> d) {
  T rmin = T(p[0]), rmax = T(p[1]), dz = T(p[2]);
  T s = T(kBig);
  if (d.z != T(0)) {
    s = ((d.z > T(0) ? dz : -dz) - q.z) / d.z;
    if (s <= T(0)) return T(0);
  }
  T s1, s2;
  if (rmin > T(0) && cyl_roots(q, d, rmin, s1, s2) && s1 > T(0)) s = s1 < s ? s1 : s;
  if (cyl_roots(q, d, rmax, s1, s2)) {
    if (s2 > T(0)) s = s2 < s ? s2 : s;
    else s = T(0);
  } else if (d.x * d.x + d.y * d.y >= T(1e-20)) {
    s = T(0);
  }
  return s < T(0) ? T(0) : s;
}
template <typename T> __device__ T tube_dist_out(const double* p, V3<T> q, V3<T> d) {
  T rmin = T(p[0]), rmax = T(p[1]), dz = T(p[2]);
  T best = T(kBig);
  if (fabs(q.z) >= dz && q.z * d.z < T(0)) {
    T s = (fabs(q.z) - dz) / fabs(d.z);
    T x = q.x + s * d.x, y = q.y + s * d.y, r2 = x * x + y * y;
    if (r2 >= rmin * rmin && r2 <= rmax * rmax) best = s;
  }
  // outer wall entry (first root) or bore exit into the solid (second root)
  T s1, s2;
  if (cyl_roots(q, d, rmax, s1, s2) && s1 >= T(0) && s1 < best) {
    if (fabs(q.z + s1 * d.z) <= dz) best = s1;
  }
  if (rmin > T(0) && cyl_roots(q, d, rmin, s1, s2) && s2 >= T(0) && s2 < best) {
    if (fabs(q.z + s2 * d.z) <= dz) best = s2;
  }
  return best;
}
template <typename T> __device__ V3<T> tube_normal(const double* p, V3<T> q, V3<T> d) {
  T r = sqrt(q.x * q.x + q.y * q.y);
  T a = fabs(T(p[2]) - fabs(q.z));
  T b = T(p[0]) > T(1e-10) ? fabs(r - T(p[0])) : T(kBig);
  T c = fabs(T(p[1]) - r);
  int i = 0;
  T m = a;
  if (b < m) { i = 1; m = b; }
  if (c < m) { i = 2; }
  if (i == 0) return mk<T>(T(0), T(0), d.z >= T(0) ? T(1) : T(-1));
  if (r <= T(0)) return mk<T>(T(1), T(0), T(0));
  V3<T> n = mk<T>(q.x / r, q.y / r, T(0));
  if (n.x * d.x + n.y * d.y < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ sphere
// params: rmin rmax theta1 theta2 (deg) cos1 cos2 sin1 sin2 phi1 dphi (deg)
// {cos sin}(phi1) {cos sin}(phi1 + dphi); constraint g <= 0 inside:
//   0: r - rmax   1: rmin - r   2: z - c1 r (theta >= theta1)   3: c2 r - z (theta <= theta2)
//   4, 5: -n.x for the inward normals n of the phi half-planes (W: a phi range)
// The phi range is a wedge: both of 4 and 5 hold for dphi <= 180 deg, either
// of them for a wider one. A sphere with a phi range has DShape.type
// kShapeSphereWedge (set at upload) and is compiled into the EXT forms only.
constexpr int kShapeSphereWedge = 65;
template <typename T> struct SphP {
  T rmin, rmax, c1, c2;
  bool t1, t2;
  T n1x, n1y, n2x, n2y;  // W: inward normals of the phi half-planes
  bool reflex;           // W: dphi > 180 deg (the union of the two half-spaces)
};
template <typename T, bool W = false> __device__ SphP<T> sph_p(const double* p) {
  SphP<T> s;
  s.rmin = T(p[0]);
  s.rmax = T(p[1]);
  s.t1 = p[2] > 0.0;
  s.t2 = p[3] < 180.0;
  s.c1 = T(p[4]);
  s.c2 = T(p[5]);
  if (W) {
    s.n1x = T(-p[11]);
    s.n1y = T(p[10]);
    s.n2x = T(p[13]);
    s.n2y = T(-p[12]);
    s.reflex = p[9] > 180.0;
  }
  return s;
}
template <typename T> __device__ __forceinline__ bool sph_in_wedge(const SphP<T>& s, V3<T> q, T lim) {
  const bool h1 = s.n1x * q.x + s.n1y * q.y >= -lim, h2 = s.n2x * q.x + s.n2y * q.y >= -lim;
  return s.reflex ? (h1 || h2) : (h1 && h2);
}
template <typename T, bool W = false> __device__ bool sphere_contains(const double* p, V3<T> q) {
  SphP<T> s = sph_p<T, W>(p);
  T r2 = dot(q, q);
  if (r2 < s.rmin * s.rmin || r2 > s.rmax * s.rmax) return false;
  if (W && !sph_in_wedge(s, q, T(0))) return false;
  if (r2 < T(1e-20)) return true;
  T r = sqrt(r2);
  if (s.t1 && q.z > s.c1 * r) return false;
  if (s.t2 && q.z < s.c2 * r) return false;
  return true;
}
// evaluate constraint k and its rate along d at x
template <typename T> __device__ __forceinline__ void sph_g(const SphP<T>& s, V3<T> x, V3<T> d, int k, T& g, T& dg) {
  if (k >= 4) {
    const T nx = k == 4 ? s.n1x : s.n2x, ny = k == 4 ? s.n1y : s.n2y;
    g = -(nx * x.x + ny * x.y);
    dg = -(nx * d.x + ny * d.y);
    return;
  }
  T r = sqrt(dot(x, x));
  T rd = r > T(0) ? dot(x, d) / r : T(0);
  if (k == 0) { g = r - s.rmax; dg = rd; }
  else if (k == 1) { g = s.rmin - r; dg = -rd; }
  else if (k == 2) { g = x.z - s.c1 * r; dg = d.z - s.c1 * rd; }
  else { g = s.c2 * r - x.z; dg = s.c2 * rd - d.z; }
}
// A candidate root t on surface w: accept for dist_out if entering through w
// and every other active constraint holds; for dist_in if leaving through w.
// With a reflex wedge, crossing one phi plane enters or leaves the solid only
// while the point is outside the other half-space.
template <typename T, bool OUT, bool W = false>
__device__ __forceinline__ void sph_try(const SphP<T>& s, V3<T> q, V3<T> d, T t, int w, T& best) {
  if (!(t >= T(0)) || t >= best) return;
  V3<T> x = axpy(q, t, d);
  T g, dg;
  sph_g(s, x, d, w, g, dg);
  T lim = T(1e-9) * (T(1) + s.rmax);
  if (OUT) {
    if (dg >= T(0)) return;
    for (int k = 0; k < 4; k++) {
      if (k == w) continue;
      if (k == 1 && !(s.rmin > T(0))) continue;
      if (k == 2 && !s.t1) continue;
      if (k == 3 && !s.t2) continue;
      T gk, dgk;
      sph_g(s, x, d, k, gk, dgk);
      if (gk > lim) return;
    }
    if (W) {
      if (w >= 4) {
        T go, dgo;
        sph_g(s, x, d, 9 - w, go, dgo);
        if (s.reflex ? !(go > T(0)) : go > lim) return;
      } else if (!sph_in_wedge(s, x, lim)) {
        return;
      }
    }
  } else {
    if (dg <= T(0)) return;
    if (W && w >= 4 && s.reflex) {
      T go, dgo;
      sph_g(s, x, d, 9 - w, go, dgo);
      if (!(go > T(0))) return;
    }
  }
  best = t;
}
template <typename T, bool OUT, bool W = false>
__device__ void sph_cone(const SphP<T>& s, V3<T> q, V3<T> d, T c, int w, T& best) {
  if (fabs(c) < T(1e-15)) {
    if (d.z != T(0)) sph_try<T, OUT, W>(s, q, d, -q.z / d.z, w, best);
    return;
  }
  T c2 = c * c;
  T A = d.z * d.z - c2 * dot(d, d);
  T B = T(2) * (q.z * d.z - c2 * dot(q, d));
  T C = q.z * q.z - c2 * dot(q, q);
  T r0 = T(-1), r1 = T(-1);
  int m = 0;
  if (fabs(A) < T(1e-15)) {
    if (B != T(0)) { r0 = -C / B; m = 1; }
  } else {
    T disc = B * B - T(4) * A * C;
    if (disc < T(0)) return;
    T sq = sqrt(disc);
    T qq = T(-0.5) * (B + (B >= T(0) ? sq : -sq));
    r0 = qq / A;
    m = 1;
    if (qq != T(0)) { r1 = C / qq; m = 2; }
  }
  if (m == 2 && r0 > r1) { T t = r0; r0 = r1; r1 = t; }
  for (int i = 0; i < m; i++) {
    T r = i == 0 ? r0 : r1;
    T z = q.z + r * d.z;
    if ((c > T(0) && z >= T(0)) || (c < T(0) && z <= T(0))) sph_try<T, OUT, W>(s, q, d, r, w, best);
  }
}
template <typename T, bool OUT, bool W = false> __device__ T sphere_dist(const double* p, V3<T> q, V3<T> d) {
  SphP<T> s = sph_p<T, W>(p);
  T best = T(kBig);
  T b = dot(q, d), qq = dot(q, q);
  T disc = b * b - (qq - s.rmax * s.rmax);
  if (disc > T(0)) {
    T sq = sqrt(disc);
    sph_try<T, OUT, W>(s, q, d, -b - sq, 0, best);
    sph_try<T, OUT, W>(s, q, d, -b + sq, 0, best);
  }
  if (s.rmin > T(0)) {
    disc = b * b - (qq - s.rmin * s.rmin);
    if (disc > T(0)) {
      T sq = sqrt(disc);
      sph_try<T, OUT, W>(s, q, d, -b - sq, 1, best);
      sph_try<T, OUT, W>(s, q, d, -b + sq, 1, best);
    }
  }
  if (s.t1) sph_cone<T, OUT, W>(s, q, d, s.c1, 2, best);
  if (s.t2) sph_cone<T, OUT, W>(s, q, d, s.c2, 3, best);
  if (W) {
    // the phi planes through the axis
    for (int k = 4; k < 6; k++) {
      const T nx = k == 4 ? s.n1x : s.n2x, ny = k == 4 ? s.n1y : s.n2y;
      const T fd = nx * d.x + ny * d.y;
      if (fd != T(0)) sph_try<T, OUT, W>(s, q, d, -(nx * q.x + ny * q.y) / fd, k, best);
    }
  }
  if (!OUT && best >= T(kBig)) best = T(0);
  return best;
}
template <typename T, bool W = false> __device__ V3<T> sphere_normal(const double* p, V3<T> q, V3<T> d) {
  SphP<T> s = sph_p<T, W>(p);
  T r = sqrt(dot(q, q));
  T rho = sqrt(q.x * q.x + q.y * q.y);
  T st1 = T(p[6]), st2 = T(p[7]);
  T dist[4] = {fabs(r - s.rmax), s.rmin > T(0) ? fabs(r - s.rmin) : T(kBig),
               s.t1 ? fabs(rho * s.c1 - q.z * st1) : T(kBig), s.t2 ? fabs(rho * s.c2 - q.z * st2) : T(kBig)};
  int k = 0;
  for (int i = 1; i < 4; i++)
    if (dist[i] < dist[k]) k = i;
  V3<T> n;
  if (k <= 1) {
    n = r > T(0) ? scale(q, T(1) / r) : mk<T>(T(0), T(0), T(1));
  } else {
    T c = k == 2 ? s.c1 : s.c2, st = k == 2 ? st1 : st2;
    T cp = rho > T(0) ? q.x / rho : T(1), sp = rho > T(0) ? q.y / rho : T(0);
    n = mk<T>(-c * cp, -c * sp, st);
  }
  if (W) {
    // a phi half-plane, where the point projects onto it
    T best = dist[k];
    for (int j = 0; j < 2; j++) {
      const T ux = T(p[10 + 2 * j]), uy = T(p[11 + 2 * j]);
      if (ux * q.x + uy * q.y < T(0)) continue;
      const T nx = j == 0 ? s.n1x : s.n2x, ny = j == 0 ? s.n1y : s.n2y;
      const T dj = fabs(nx * q.x + ny * q.y);
      if (dj < best) {
        best = dj;
        n = mk<T>(nx, ny, T(0));
      }
    }
  }
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ paraboloid
// params: rlo rhi dz a b ; z = a r^2 + b, solid: |z| <= dz and a (z - b) >= a^2 r^2
template <typename T> __device__ bool para_contains(const double* p, V3<T> q) {
  if (fabs(q.z) > T(p[2])) return false;
  T a = T(p[3]), b = T(p[4]);
  T aa = a * (q.z - b);
  if (aa < T(0)) return false;
  return aa >= a * a * (q.x * q.x + q.y * q.y);
}
template <typename T> __device__ int para_roots(const double* p, V3<T> q, V3<T> d, T& r0, T& r1) {
  T a = T(p[3]), b = T(p[4]);
  T A = a * (d.x * d.x + d.y * d.y);
  T B = T(2) * a * (q.x * d.x + q.y * d.y) - d.z;
  T C = a * (q.x * q.x + q.y * q.y) + b - q.z;
  if (fabs(A) < T(1e-30)) {
    if (B == T(0)) return 0;
    r0 = -C / B;
    return 1;
  }
  T disc = B * B - T(4) * A * C;
  if (disc < T(0)) return 0;
  T sq = sqrt(disc);
  T qq = T(-0.5) * (B + (B >= T(0) ? sq : -sq));
  r0 = qq / A;
  if (qq == T(0)) return 1;
  r1 = C / qq;
  if (r0 > r1) { T t = r0; r0 = r1; r1 = t; }
  return 2;
}
template <typename T> __device__ T para_dist_out(const double* p, V3<T> q, V3<T> d) {
  T rlo = T(p[0]), rhi = T(p[1]), dz = T(p[2]), a = T(p[3]);
  T best = T(kBig);
  if (q.z <= -dz && d.z > T(0)) {
    T s = -(dz + q.z) / d.z;
    T x = q.x + s * d.x, y = q.y + s * d.y;
    if (x * x + y * y <= rlo * rlo) best = s;
  } else if (q.z >= dz && d.z < T(0)) {
    T s = (dz - q.z) / d.z;
    T x = q.x + s * d.x, y = q.y + s * d.y;
    if (x * x + y * y <= rhi * rhi) best = s;
  }
  T r[2];
  int n = para_roots(p, q, d, r[0], r[1]);
  for (int i = 0; i < n; i++) {
    if (r[i] < T(0) || r[i] >= best) continue;
    V3<T> x = axpy(q, r[i], d);
    if (fabs(x.z) > dz) continue;
    T rate = a * d.z - T(2) * a * a * (x.x * d.x + x.y * d.y);
    if (rate <= T(0)) continue;
    best = r[i];
    break;
  }
  return best;
}
template <typename T> __device__ T para_dist_in(const double* p, V3<T> q, V3<T> d) {
  T dz = T(p[2]), a = T(p[3]);
  T best = T(kBig);
  if (d.z > T(0)) best = (dz - q.z) / d.z;
  else if (d.z < T(0)) best = (-dz - q.z) / d.z;
  if (best < T(0)) best = T(0);
  T r[2];
  int n = para_roots(p, q, d, r[0], r[1]);
  for (int i = 0; i < n; i++) {
    if (r[i] < T(0) || r[i] >= best) continue;
    V3<T> x = axpy(q, r[i], d);
    T rate = a * d.z - T(2) * a * a * (x.x * d.x + x.y * d.y);
    if (rate >= T(0)) continue;
    best = r[i];
    break;
  }
  return best;
}
template <typename T> __device__ V3<T> para_normal(const double* p, V3<T> q, V3<T> d) {
  T dz = T(p[2]), a = T(p[3]);
  T r2 = q.x * q.x + q.y * q.y;
  T zs = a * r2 + T(p[4]);
  T gl = sqrt(T(1) + T(4) * a * a * r2);
  T dpar = fabs(q.z - zs) / gl;
  T dtop = fabs(dz - q.z), dbot = fabs(q.z + dz);
  V3<T> n;
  if (dpar <= dtop && dpar <= dbot) n = mk<T>(T(2) * a * q.x / gl, T(2) * a * q.y / gl, T(-1) / gl);
  else n = mk<T>(T(0), T(0), T(1));
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ pgon (rmin = 0, full phi)
// params: phi1 dphi nedges nz {z rmin rmax}*nz. Each z segment is a convex
// frustum: half-spaces n_k.(x,y) - slope (z - z0) <= r0 inside the z slab.
template <typename T> __device__ __forceinline__ void pgon_face(const double* p, int k, T& cx, T& sy) {
  const double* f = p + 4 + 3 * int(p[3]) + 2 * k;   // precomputed outward side normals
  cx = T(f[0]);
  sy = T(f[1]);
}
template <typename T> __device__ bool pgon_clip(const double* p, int i, V3<T> q, V3<T> d, T& t0, T& t1) {
  const double* z = p + 4;
  T z0 = T(z[3 * i]), z1 = T(z[3 * (i + 1)]);
  if (!(z1 - z0 > T(0))) return false;
  T r0 = T(z[3 * i + 2]);
  T slope = (T(z[3 * (i + 1) + 2]) - r0) / (z1 - z0);
  t0 = T(-kBig);
  t1 = T(kBig);
  if (d.z != T(0)) {
    T a = (z0 - q.z) / d.z, b = (z1 - q.z) / d.z;
    if (a > b) { T t = a; a = b; b = t; }
    t0 = a > t0 ? a : t0;
    t1 = b < t1 ? b : t1;
  } else if (q.z < z0 || q.z > z1) {
    return false;
  }
  int ne = int(p[2]);
  for (int k = 0; k < ne; k++) {
    T cx, sy;
    pgon_face(p, k, cx, sy);
    T f0 = cx * q.x + sy * q.y - slope * (q.z - z0) - r0;
    T fd = cx * d.x + sy * d.y - slope * d.z;
    if (fd == T(0)) {
      if (f0 > T(0)) return false;
      continue;
    }
    T t = -f0 / fd;
    if (fd > T(0)) t1 = t < t1 ? t : t1;
    else t0 = t > t0 ? t : t0;
  }
  return t0 <= t1;
}
// Inside a regular full-phi polygon iff the projection on every side normal
// is within the side's apothem; the largest projection is the one on the
// side of the point's own phi sector (the test ROOT's TGeoPgon::Contains makes).
template <typename T> __device__ bool pgon_contains(const double* p, V3<T> q) {
  int nz = int(p[3]), ne = int(p[2]);
  const double* z = p + 4;
  if (q.z < T(z[0]) || q.z > T(z[3 * (nz - 1)])) return false;
  T r = T(-kBig);
  for (int k = 0; k < ne; k++) {
    T cx, sy;
    pgon_face(p, k, cx, sy);
    T v = q.x * cx + q.y * sy;
    r = v > r ? v : r;
  }
  for (int i = 0; i + 1 < nz; i++) {
    T za = T(z[3 * i]), zb = T(z[3 * (i + 1)]);
    if (q.z < za || q.z > zb) continue;
    T rmax;
    if (zb - za < T(1e-8)) {
      rmax = T(fmax(z[3 * i + 2], z[3 * (i + 1) + 2]));
    } else {
      T f = (q.z - za) / (zb - za);
      rmax = T(z[3 * i + 2]) + f * (T(z[3 * (i + 1) + 2]) - T(z[3 * i + 2]));
    }
    if (r <= rmax) return true;
  }
  return false;
}
template <typename T> __device__ T pgon_dist_out(const double* p, V3<T> q, V3<T> d) {
  int nz = int(p[3]);
  T best = T(kBig);
  for (int i = 0; i + 1 < nz; i++) {
    T t0, t1;
    if (!pgon_clip(p, i, q, d, t0, t1)) continue;
    if (t1 <= T(0) || t0 >= t1) continue;
    T t = t0 < T(0) ? T(0) : t0;
    best = t < best ? t : best;
  }
  return best;
}
template <typename T> __device__ T pgon_dist_in(const double* p, V3<T> q, V3<T> d) {
  int nz = int(p[3]);
  T cur = T(0);
  for (int guard = 0; guard < 2 * nz + 2; guard++) {
    bool moved = false;
    for (int i = 0; i + 1 < nz; i++) {
      T t0, t1;
      if (!pgon_clip(p, i, q, d, t0, t1)) continue;
      if (t0 <= cur + T(1e-9) && t1 > cur + T(1e-9)) {
        cur = t1;
        moved = true;
        break;
      }
    }
    if (!moved) break;
  }
  return cur;
}
template <typename T> __device__ V3<T> pgon_normal(const double* p, V3<T> q, V3<T> d) {
  int nz = int(p[3]), ne = int(p[2]);
  const double* z = p + 4;
  T best = T(kBig);
  V3<T> n = mk<T>(T(0), T(0), T(1));
  for (int i = 0; i < nz; i++) {
    T dzp = fabs(q.z - T(z[3 * i]));
    if (dzp < best) { best = dzp; n = mk<T>(T(0), T(0), T(1)); }
  }
  // the side of the point's phi sector: largest projection
  int ip = 0;
  T cx = T(0), sy = T(0), rmx = T(-kBig);
  for (int k = 0; k < ne; k++) {
    T a, b;
    pgon_face(p, k, a, b);
    T v = q.x * a + q.y * b;
    if (v > rmx) { rmx = v; ip = k; cx = a; sy = b; }
  }
  (void)ip;
  for (int i = 0; i + 1 < nz; i++) {
    T z0 = T(z[3 * i]), z1 = T(z[3 * (i + 1)]);
    if (!(z1 - z0 > T(0))) continue;
    if (q.z < z0 - T(1e-9) || q.z > z1 + T(1e-9)) continue;
    T r0 = T(z[3 * i + 2]);
    T slope = (T(z[3 * (i + 1) + 2]) - r0) / (z1 - z0);
    T len = sqrt(T(1) + slope * slope);
    T f = (cx * q.x + sy * q.y - slope * (q.z - z0) - r0) / len;
    if (fabs(f) < best) {
      best = fabs(f);
      n = mk<T>(cx / len, sy / len, -slope / len);
    }
  }
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ TGeoPcon, and TGeoPgon with rmin > 0 or a phi range
// Pcon params:  phi1 dphi nz {z rmin rmax}*nz {cos sin}(phi1) {cos sin}(phi1 + dphi)
// Pgon params:  phi1 dphi nedges nz {z rmin rmax}*nz {cos sin}*nedges {cos sin}(phi1) {cos sin}(phi1 + dphi)
// (z non-decreasing). Per z segment the solid is the region between the
// inner and the outer surface (cone frustums; for a Pgon pyramid frustums
// whose sides are the polygon's), inside the phi wedge. Along a ray the
// slab, the outer region and a wedge of at most 180 deg are convex (one
// interval each); the inner region and, for a wedge wider than 180 deg, the
// missing wedge are removed from it, so a segment holds at most three
// intervals of the ray and the solid holds their union over the segments.
#ifndef RB_POLY_FN
#define RB_POLY_FN __noinline__
#endif
constexpr int kShapePgonGeneral = 64;  // DShape.type of a TGeoPgon with rmin > 0 or a phi range (set at upload)

struct PolyView {
  const double* z;      // {z rmin rmax} * nz
  const double* faces;  // Pgon: outward side normals {cos sin} * ne
  const double* wedge;  // {cos sin} of phi1, then of phi1 + dphi
  int nz, ne;
  bool full;            // dphi >= 360
  bool reflex;          // dphi > 180: the wedge is the union of its two half-planes
};
template <bool PGON> __device__ __forceinline__ PolyView poly_view(const double* p) {
  PolyView v;
  if (PGON) {
    v.ne = int(p[2]);
    v.nz = int(p[3]);
    v.z = p + 4;
    v.faces = v.z + 3 * v.nz;
    v.wedge = v.faces + 2 * v.ne;
  } else {
    v.ne = 0;
    v.nz = int(p[2]);
    v.z = p + 3;
    v.faces = nullptr;
    v.wedge = v.z + 3 * v.nz;
  }
  v.full = p[1] >= 360.0;
  v.reflex = p[1] > 180.0;
  return v;
}
// inward normals (into the wedge) of the two phi half-planes
template <typename T> __device__ __forceinline__ void poly_wedge(const PolyView& v, T& n1x, T& n1y, T& n2x, T& n2y) {
  n1x = T(-v.wedge[1]);
  n1y = T(v.wedge[0]);
  n2x = T(v.wedge[3]);
  n2y = T(-v.wedge[2]);
}
// [a, b] clipped to {t : n . (q + t d)_xy >= 0}; false if it becomes empty
template <typename T> __device__ __forceinline__ bool half_clip(T nx, T ny, V3<T> q, V3<T> d, T& a, T& b) {
  const T f0 = nx * q.x + ny * q.y, fd = nx * d.x + ny * d.y;
  if (fd == T(0)) return f0 >= T(0) && a <= b;
  const T t = -f0 / fd;
  if (fd > T(0)) a = t > a ? t : a;
  else b = t < b ? t : b;
  return a <= b;
}
// [a, b] clipped to the slab of segment i; false if empty or a zero-length segment
template <typename T> __device__ __forceinline__ bool poly_slab(const PolyView& v, int i, V3<T> q, V3<T> d, T& a, T& b) {
  const T z0 = T(v.z[3 * i]), z1 = T(v.z[3 * (i + 1)]);
  if (!(z1 - z0 > T(0))) return false;
  a = T(-kBig);
  b = T(kBig);
  if (d.z != T(0)) {
    T u = (z0 - q.z) / d.z, w = (z1 - q.z) / d.z;
    if (u > w) { T x = u; u = w; w = x; }
    a = u;
    b = w;
    return true;
  }
  return q.z >= z0 && q.z <= z1;
}
// [a, b] clipped to the convex region radial(p) <= R(z) of segment i, where R
// interpolates column col (1 = rmin, 2 = rmax) linearly between its sections
// and radial is the distance from the axis (Pcon) or the projection on the
// side normals (Pgon); false if it becomes empty.
template <typename T, bool PGON>
__device__ bool poly_radial(const PolyView& v, int i, int col, V3<T> q, V3<T> d, T& a, T& b) {
  const T z0 = T(v.z[3 * i]), z1 = T(v.z[3 * (i + 1)]);
  const T r0 = T(v.z[3 * i + col]);
  const T s = (T(v.z[3 * (i + 1) + col]) - r0) / (z1 - z0);
  if (PGON) {
    for (int k = 0; k < v.ne; k++) {
      const T cx = T(v.faces[2 * k]), sy = T(v.faces[2 * k + 1]);
      const T f0 = cx * q.x + sy * q.y - s * (q.z - z0) - r0;
      const T fd = cx * d.x + sy * d.y - s * d.z;
      if (fd == T(0)) {
        if (f0 > T(0)) return false;
        continue;
      }
      const T t = -f0 / fd;
      if (fd > T(0)) b = t < b ? t : b;
      else a = t > a ? t : a;
    }
    return a <= b;
  }
  // cone frustum x^2 + y^2 <= R(z)^2 with R(z) = R0 + s dz t >= 0 inside the slab:
  // g(t) = A t^2 + 2 B t + C <= 0
  const T R0 = r0 + s * (q.z - z0);
  const T A = d.x * d.x + d.y * d.y - s * s * d.z * d.z;
  const T B = q.x * d.x + q.y * d.y - s * d.z * R0;
  const T C = q.x * q.x + q.y * q.y - R0 * R0;
  // the real nappe: R(t) >= 0
  const T rd = s * d.z;
  if (rd != T(0)) {
    const T t = -R0 / rd;
    if (rd > T(0)) a = t > a ? t : a;
    else b = t < b ? t : b;
  } else if (R0 < T(0)) {
    return false;
  }
  if (!(a <= b)) return false;
  const T eps = T(1e-12);
  if (fabs(A) <= eps * (d.x * d.x + d.y * d.y + s * s * d.z * d.z)) {
    // parallel to a generator: 2 B t + C <= 0
    if (B == T(0)) return C <= T(0);
    const T t = -C / (T(2) * B);
    if (B > T(0)) b = t < b ? t : b;
    else a = t > a ? t : a;
    return a <= b;
  }
  const T disc = B * B - A * C;
  if (A > T(0)) {
    if (disc < T(0)) return false;
    const T h = sqrt(disc);
    // roots without cancellation
    const T qq = B >= T(0) ? -(B + h) : -(B - h);
    T u = qq / A, w = qq != T(0) ? C / qq : T(0);
    if (u > w) { T x = u; u = w; w = x; }
    a = u > a ? u : a;
    b = w < b ? w : b;
    return a <= b;
  }
  // A < 0: g <= 0 outside the roots; the part on the real nappe is one half-line
  if (disc < T(0)) return true;
  const T h = sqrt(disc);
  const T qq = B >= T(0) ? -(B + h) : -(B - h);
  T u = qq / A, w = qq != T(0) ? C / qq : T(0);
  if (u > w) { T x = u; u = w; w = x; }
  if (R0 + rd * w >= T(0)) a = w > a ? w : a;
  else b = u < b ? u : b;
  return a <= b;
}
// True if segment i has an inner surface (rmin > 0 at either section)
__device__ __forceinline__ bool poly_has_inner(const PolyView& v, int i) {
  return v.z[3 * i + 1] > 0.0 || v.z[3 * (i + 1) + 1] > 0.0;
}
// Calls f(lo, hi) for each interval of the ray inside segment i.
template <typename T, bool PGON, typename F>
__device__ __forceinline__ void poly_pieces(const PolyView& v, int i, V3<T> q, V3<T> d, F& f) {
  T a, b;
  if (!poly_slab(v, i, q, d, a, b)) return;
  if (!poly_radial<T, PGON>(v, i, 2, q, d, a, b)) return;
  T n1x = T(0), n1y = T(0), n2x = T(0), n2y = T(0);
  if (!v.full) {
    poly_wedge(v, n1x, n1y, n2x, n2y);
    if (!v.reflex && (!half_clip(n1x, n1y, q, d, a, b) || !half_clip(n2x, n2y, q, d, a, b))) return;
  }
  // removed: the inner region [c, e], and for a reflex wedge the missing wedge [m0, m1]
  T c = T(kBig), e = T(-kBig), m0 = T(kBig), m1 = T(-kBig);
  if (poly_has_inner(v, i)) {
    T u = T(-kBig), w = T(kBig);
    if (poly_radial<T, PGON>(v, i, 1, q, d, u, w)) { c = u; e = w; }
  }
  if (!v.full && v.reflex) {
    T u = T(-kBig), w = T(kBig);
    if (half_clip(-n1x, -n1y, q, d, u, w) && half_clip(-n2x, -n2y, q, d, u, w)) { m0 = u; m1 = w; }
  }
  // [a, b] minus [c, e] minus [m0, m1]
  T lo[2], hi[2];
  int np = 0;
  if (c <= e && c < b && e > a) {
    if (a < c) { lo[np] = a; hi[np] = c; np++; }
    if (e < b) { lo[np] = e; hi[np] = b; np++; }
  } else {
    lo[0] = a; hi[0] = b; np = 1;
  }
  for (int k = 0; k < np; k++) {
    if (m0 <= m1 && m0 < hi[k] && m1 > lo[k]) {
      if (lo[k] < m0) f(lo[k], m0);
      if (m1 < hi[k]) f(m1, hi[k]);
    } else {
      f(lo[k], hi[k]);
    }
  }
}
template <typename T> struct PolyFirstEntry {
  T best;
  __device__ __forceinline__ void operator()(T lo, T hi) {
    if (hi <= T(0)) return;
    const T t = lo < T(0) ? T(0) : lo;
    best = t < best ? t : best;
  }
};
template <typename T> struct PolyAdvance {
  T cur;
  bool moved;
  __device__ __forceinline__ void operator()(T lo, T hi) {
    if (!moved && lo <= cur + T(1e-9) && hi > cur + T(1e-9)) {
      cur = hi;
      moved = true;
    }
  }
};
template <typename T, bool PGON> __device__ RB_POLY_FN T poly_dist_out(const double* p, V3<T> q, V3<T> d) {
  const PolyView v = poly_view<PGON>(p);
  PolyFirstEntry<T> f{T(kBig)};
  for (int i = 0; i + 1 < v.nz; i++) poly_pieces<T, PGON>(v, i, q, d, f);
  return f.best;
}
// the end of the run of touching intervals that holds the start point
template <typename T, bool PGON> __device__ RB_POLY_FN T poly_dist_in(const double* p, V3<T> q, V3<T> d) {
  const PolyView v = poly_view<PGON>(p);
  T cur = T(0);
  for (int guard = 0; guard < 3 * v.nz + 2; guard++) {
    PolyAdvance<T> f{cur, false};
    for (int i = 0; i + 1 < v.nz && !f.moved; i++) poly_pieces<T, PGON>(v, i, q, d, f);
    if (!f.moved) break;
    cur = f.cur;
  }
  return cur;
}
// radial measure: distance from the axis (Pcon); for a Pgon the projection on
// the side of the point's own sector, the largest over all sides
template <typename T, bool PGON> __device__ __forceinline__ T poly_rho(const PolyView& v, V3<T> q, int& face) {
  face = -1;
  if (!PGON) return sqrt(q.x * q.x + q.y * q.y);
  T r = T(-kBig);
  for (int k = 0; k < v.ne; k++) {
    const T w = q.x * T(v.faces[2 * k]) + q.y * T(v.faces[2 * k + 1]);
    if (w > r) { r = w; face = k; }
  }
  return r;
}
template <typename T> __device__ __forceinline__ bool poly_in_wedge(const PolyView& v, V3<T> q) {
  if (v.full) return true;
  T n1x, n1y, n2x, n2y;
  poly_wedge(v, n1x, n1y, n2x, n2y);
  const bool h1 = n1x * q.x + n1y * q.y >= T(0), h2 = n2x * q.x + n2y * q.y >= T(0);
  return v.reflex ? (h1 || h2) : (h1 && h2);
}
template <typename T, bool PGON> __device__ RB_POLY_FN bool poly_contains(const double* p, V3<T> q) {
  const PolyView v = poly_view<PGON>(p);
  if (q.z < T(v.z[0]) || q.z > T(v.z[3 * (v.nz - 1)])) return false;
  if (!poly_in_wedge(v, q)) return false;
  int face;
  const T rho = poly_rho<T, PGON>(v, q, face);
  for (int i = 0; i + 1 < v.nz; i++) {
    const T za = T(v.z[3 * i]), zb = T(v.z[3 * (i + 1)]);
    if (q.z < za || q.z > zb) continue;
    T rmin, rmax;
    if (zb - za < T(1e-8)) {
      rmin = T(fmin(v.z[3 * i + 1], v.z[3 * (i + 1) + 1]));
      rmax = T(fmax(v.z[3 * i + 2], v.z[3 * (i + 1) + 2]));
    } else {
      const T f = (q.z - za) / (zb - za);
      rmin = T(v.z[3 * i + 1]) + f * (T(v.z[3 * (i + 1) + 1]) - T(v.z[3 * i + 1]));
      rmax = T(v.z[3 * i + 2]) + f * (T(v.z[3 * (i + 1) + 2]) - T(v.z[3 * i + 2]));
    }
    if (rho >= rmin && rho <= rmax) return true;
  }
  return false;
}
// normal of the nearest surface: a section plane, the inner or outer surface
// of the point's segment, or a phi half-plane
template <typename T, bool PGON> __device__ RB_POLY_FN V3<T> poly_normal(const double* p, V3<T> q, V3<T> d) {
  const PolyView v = poly_view<PGON>(p);
  T best = T(kBig);
  V3<T> n = mk<T>(T(0), T(0), T(1));
  for (int k = 0; k < v.nz; k++) {
    const T dz = fabs(q.z - T(v.z[3 * k]));
    if (dz < best) best = dz;
  }
  int face;
  const T rho = poly_rho<T, PGON>(v, q, face);
  T cx, sy;
  if (PGON) {
    cx = T(v.faces[2 * face]);
    sy = T(v.faces[2 * face + 1]);
  } else {
    cx = rho > T(0) ? q.x / rho : T(1);
    sy = rho > T(0) ? q.y / rho : T(0);
  }
  for (int i = 0; i + 1 < v.nz; i++) {
    const T z0 = T(v.z[3 * i]), z1 = T(v.z[3 * (i + 1)]);
    if (!(z1 - z0 > T(0)) || q.z < z0 - T(1e-9) || q.z > z1 + T(1e-9)) continue;
    for (int col = 1; col <= 2; col++) {
      if (col == 1 && !poly_has_inner(v, i)) continue;
      const T r0 = T(v.z[3 * i + col]);
      const T s = (T(v.z[3 * (i + 1) + col]) - r0) / (z1 - z0);
      const T len = sqrt(T(1) + s * s);
      const T dist = fabs(rho - r0 - s * (q.z - z0)) / len;
      if (dist < best) {
        best = dist;
        n = mk<T>(cx / len, sy / len, -s / len);
      }
    }
  }
  if (!v.full) {
    T nx[2], ny[2];
    poly_wedge(v, nx[0], ny[0], nx[1], ny[1]);
    for (int k = 0; k < 2; k++) {
      // the half-plane reaches out from the axis along (cos phi, sin phi)
      const T ux = T(v.wedge[2 * k]), uy = T(v.wedge[2 * k + 1]);
      if (ux * q.x + uy * q.y < T(0)) continue;
      const T dist = fabs(nx[k] * q.x + ny[k] * q.y);
      if (dist < best) {
        best = dist;
        n = mk<T>(nx[k], ny[k], T(0));
      }
    }
  }
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ aspheric disk
// AGeoAsphericDisk (src/AGeoAsphericDisk.cxx). params: z1 c1 z2 c2 rmax rmin
// kappa1 kappa2 n1 n2 K1[n1] K2[n2]. Both distances are the minimum over the
// two aspheres and the two walls (:363-412); an asphere crossing is the
// Matsui spherical seed refined by Newton steps along the ray (:415-526).
template <typename T> __device__ __forceinline__ T newton_eps() { return T(1e-10); }
template <> __device__ __forceinline__ float newton_eps<float>() { return 2e-4f; }

// polynomial part with normalised coefficients: sum Kn_k h^k, h = r^2 / rmax^2
template <typename T>
__device__ __forceinline__ const double* asph_kn(const double* p, int s) {
  return p + 10 + int(p[8]) + int(p[9]) + (s ? int(p[8]) : 0);
}
template <typename T>
__device__ bool asph_sag(const double* p, int s, T r, T& f) {
  const T c = T(p[1 + 2 * s]), kap = T(p[6 + s]);
  const T pp = r * r * c * c * kap;
  if (T(1) - pp < T(0)) return false;
  const int n = int(p[8 + s]);
  const double* K = asph_kn<T>(p, s);
  const T r2 = r * r, h = r2 / T(p[4] * p[4]);
  T acc = r2 * c / (T(1) + sqrt(T(1) - pp)), hk = h;
  for (int i = 0; i < n; i++) {
    acc += T(K[i]) * hk;
    hk *= h;
  }
  f = T(p[2 * s]) + acc;
  return true;
}
template <typename T>
__device__ bool asph_dsag(const double* p, int s, T r, T& df) {
  const T c = T(p[1 + 2 * s]), kap = T(p[6 + s]);
  const T pp = r * r * c * c * kap;
  if (T(1) - pp <= T(0)) return false;
  const int n = int(p[8 + s]);
  const double* K = asph_kn<T>(p, s);
  const T R2inv = T(1.0 / (p[4] * p[4])), h = r * r * R2inv;
  T acc = r * c / sqrt(T(1) - pp), poly = T(0), hk = T(1);
  for (int i = 0; i < n; i++) {
    poly += T(2 * (i + 1)) * T(K[i]) * hk;
    hk *= h;
  }
  df = acc + poly * r * R2inv;
  return true;
}
template <typename T>
__device__ T asph_surface(const double* p, int s, V3<T> P, V3<T> D) {
  const T c = T(p[1 + 2 * s]), kap = T(p[6 + s]), dz = T(p[2 * s]);
  const int n = int(p[8 + s]);
  const double* K = asph_kn<T>(p, s);
  const T R2inv = T(1.0 / (p[4] * p[4]));
  T H2 = P.x * P.x + P.y * P.y;
  T pz = P.z - dz;
  T pr = -(pz * D.z + P.x * D.x + P.y * D.y);
  T M = pr * D.z + pz;
  T M2 = pz * pz + H2 - pr * pr;
  T chk = T(1) - (M2 * c - T(2) * M) * c / D.z / D.z;
  if (!(chk >= T(0))) return T(kBig);
  T q = pr + (M2 * c - T(2) * M) / (D.z * (T(1) + sqrt(chk)));
  T nx = P.x + q * D.x, ny = P.y + q * D.y, nz = pz + q * D.z;
  const T eps = newton_eps<T>();
  T e = T(kBig);
  for (int i = 0;; i++) {
    if (i > 100) {
      // FP32 may stall at its noise floor: keep a close iterate for the FP64 polish
      if (sizeof(T) == 4 && fabs(e) < T(1e-2)) break;
      return T(kBig);
    }
    H2 = nx * nx + ny * ny;
    chk = T(1) - kap * H2 * c * c;
    if (chk < T(0)) return T(kBig);
    T l = sqrt(chk);
    // conic sag c r^2 / (1 + l): the same value as the reference's
    // (1 - l) / (c kappa) without its cancellation
    T x = H2 * c / (T(1) + l);
    const T h = H2 * R2inv;
    T v = T(0), hk = T(1);
    for (int j = 0; j < n; j++) {
      T kj = T(K[j]);
      v += T(2 * (j + 1)) * kj * hk;
      hk *= h;
      x += kj * hk;
    }
    v = c * kap + l * v * R2inv;
    T m = -nx * v, mm = -ny * v;
    T inv = T(1) / sqrt(l * l + m * m + mm * mm);
    l *= inv; m *= inv; mm *= inv;
    chk = D.z * l + D.x * m + D.y * mm;
    if (chk == T(0)) return T(kBig);
    e = l * (x - nz) / chk;
    nx += e * D.x; ny += e * D.y; nz += e * D.z;
    if (fabs(e) < eps) break;
  }
  nz += dz;
  T ex = nx - P.x, ey = ny - P.y, ez = nz - P.z;
  if (D.x * ex + D.y * ey + D.z * ez < T(0)) return T(kBig);
  T rz = sqrt(nx * nx + ny * ny);
  if (rz < T(p[5]) || rz > T(p[4])) return T(kBig);
  return sqrt(ex * ex + ey * ey + ez * ez);
}
template <typename T>
__device__ T asph_wall(const double* p, bool outer, V3<T> P, V3<T> D) {
  const T R = outer ? T(p[4]) : T(p[5]);
  if (!outer && !(R > T(0))) return T(kBig);
  T rsq = P.x * P.x + P.y * P.y, nsq = D.x * D.x + D.y * D.y;
  if (sqrt(nsq) < T(kTol)) return T(kBig);
  T rdotn = P.x * D.x + P.y * D.y;
  T t1i = T(1) / nsq, b = t1i * rdotn, cc = t1i * (rsq - R * R);
  T delta = b * b - cc;
  if (!(delta > T(0))) return T(kBig);
  delta = sqrt(delta);
  T t1 = -b + delta, t2 = -b - delta;
  if (t1 < T(0) && t2 < T(0)) return T(kBig);
  T zmin, zmax;
  if (!asph_sag<T>(p, 0, R, zmin) || !asph_sag<T>(p, 1, R, zmax)) return T(kBig);
  if (t2 > T(0)) {
    T z1 = t1 * D.z + P.z, z2 = t2 * D.z + P.z;
    if (z1 < zmin || zmax < z1) t1 = T(kBig);
    if (z2 < zmin || zmax < z2) t2 = T(kBig);
    return t1 < t2 ? t1 : t2;
  } else if (t2 == T(0)) {
    if (t1 > T(0)) {
      if (zmin <= P.z && P.z <= zmax) return T(0);
      T z1 = t1 * D.z + P.z;
      if (zmin <= z1 && z1 <= zmax) return t1;
    } else if (t1 == T(0)) {
      if (zmin <= P.z && P.z <= zmax) return T(0);
    }
  } else {
    if (t1 > T(0)) {
      T z1 = t1 * D.z + P.z;
      if (zmin <= z1 && z1 <= zmax) return t1;
    } else if (t1 == T(0)) {
      if (zmin <= P.z && P.z <= zmax) return T(0);
    }
  }
  return T(kBig);
}
template <typename T> __device__ RB_HEAVY_FN T asph_dist(const double* p, V3<T> q, V3<T> d) {
  T a = asph_surface<T>(p, 0, q, d), b = asph_surface<T>(p, 1, q, d);
  T c = asph_wall<T>(p, false, q, d), e = asph_wall<T>(p, true, q, d);
  T m = a;
  m = b < m ? b : m;
  m = c < m ? c : m;
  m = e < m ? e : m;
  return m;
}
template <typename T> __device__ RB_HEAVY_FN bool asph_contains(const double* p, V3<T> q) {
  T r = sqrt(q.x * q.x + q.y * q.y);
  if (r > T(p[4]) || r < T(p[5])) return false;
  T f1, f2;
  if (!asph_sag<T>(p, 0, r, f1) || !asph_sag<T>(p, 1, r, f2)) return false;
  return !(q.z < f1 || f2 < q.z);
}
template <typename T> __device__ RB_HEAVY_FN V3<T> asph_normal(const double* p, V3<T> q, V3<T> d) {
  T r = sqrt(q.x * q.x + q.y * q.y);
  T saf0 = T(p[5]) > T(0) ? fabs(r - T(p[5])) : T(kBig);
  T saf1 = fabs(r - T(p[4]));
  T saf2 = T(kBig), saf3 = T(kBig), df1 = T(kBig), df2 = T(kBig), f;
  if (asph_sag<T>(p, 0, r, f) && asph_dsag<T>(p, 0, r, df1)) saf2 = fabs(f - q.z) / sqrt(T(1) + df1 * df1);
  if (asph_sag<T>(p, 1, r, f) && asph_dsag<T>(p, 1, r, df2)) saf3 = fabs(f - q.z) / sqrt(T(1) + df2 * df2);
  int k = 0;
  T m = saf0;
  if (saf1 < m) { k = 1; m = saf1; }
  if (saf2 < m) { k = 2; m = saf2; }
  if (saf3 < m) { k = 3; }
  T n0 = T(0), n2 = T(0);
  if (k <= 1) {
    n0 = T(1);
  } else {
    T g = k == 2 ? df1 : df2;
    if (g == T(0)) {
      n2 = T(1);
    } else {
      T inv = T(1) / sqrt(T(1) + g * g);
      n0 = g * inv;
      n2 = -inv;
    }
  }
  T cp = r > T(0) ? q.x / r : T(1), sp = r > T(0) ? q.y / r : T(0);
  V3<T> n = mk<T>(n0 * cp, n0 * sp, n2);
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// ------------------------------------------------------------ Winston cone (polygonal)
// AGeoWinstonConePoly (src/AGeoWinstonConePoly.cxx) with AGeoWinstonCone2D's
// parabola crossing (src/AGeoWinstonCone2D.cxx:325-428). params: r1 r2 N theta
// dz f cos(theta) sin(theta) tan(pi/N) {cos, sin}(2 pi i / N) * N.
// Trigonometry-free on the device: tan(atan2(pz, px) - theta) is formed from
// the direction components, and the sector tests become projections.
template <typename T> __device__ __forceinline__ bool wins_r(const double* p, T z, T& r) {
  const T dz = T(p[4]), f = T(p[5]), ct = T(p[6]), st = T(p[7]);
  if (fabs(z) > dz + T(1e-10)) return false;
  T t = z + dz;
  T a0 = t * t * st * st - T(4) * f * (t * ct + f);
  T a1 = T(2) * t * st * ct + T(4) * f * st;
  T a2 = ct * ct;
  r = (-a1 + sqrt(a1 * a1 - T(4) * a0 * a2)) / (T(2) * a2) - T(p[1]);
  return true;
}
template <typename T> __device__ __forceinline__ bool wins_drdz(const double* p, T z, T& g) {
  const T dz = T(p[4]), f = T(p[5]), ct = T(p[6]), st = T(p[7]);
  if (fabs(z) > dz + T(1e-10)) return false;
  T t = z + dz;
  T a0 = t * t * st * st - T(4) * f * (t * ct + f);
  T a1 = T(2) * t * st * ct + T(4) * f * st;
  T a2 = ct * ct;
  T da0 = T(2) * t * st * st - T(4) * f * ct, da1 = T(2) * st * ct;
  g = (-da1 + (a1 * da1 - T(2) * da0 * a2) / sqrt(a1 * a1 - T(4) * a0 * a2)) / (T(2) * a2);
  return true;
}
// InsidePolygon: the projection on the side normal of the point's own sector
// (the largest projection) must not exceed the apothem r
template <typename T> __device__ __forceinline__ bool wins_inside_polygon(const double* p, T x, T y, T r) {
  const int n = int(p[2]);
  T m = T(-kBig);
  for (int i = 0; i < n; i++) {
    T v = x * T(p[9 + 2 * i]) + y * T(p[10 + 2 * i]);
    m = v > m ? v : m;
  }
  return !(m > r);
}
// side i's parabola; half_plane: accept |azimuth| <= pi/2 (DistFromInside),
// else |azimuth| <= pi/N (DistFromOutside)
template <typename T>
__device__ T wins_parabola(const double* p, int i, V3<T> P, V3<T> D, bool half_plane) {
  const T r1 = T(p[0]), r2 = T(p[1]), dz = T(p[4]), f = T(p[5]), ct = T(p[6]), st = T(p[7]);
  const T cph = T(p[9 + 2 * i]), sph = T(p[10 + 2 * i]);
  T x = cph * P.x + sph * P.y, y = -sph * P.x + cph * P.y, z = P.z;
  T px = cph * D.x + sph * D.y, py = -sph * D.x + cph * D.y, pz = D.z;
  if (px == T(0) && pz == T(0)) return T(kBig);
  T X = ct * (x + r2) + (z + dz) * st;
  T Z = -st * (x + r2) + (z + dz) * ct + f;
  T num = pz * ct - px * st, den = px * ct + pz * st;  // tanA = num / den
  T Xc[2];
  if (T(2) * dz * fabs(den) < T(kTol) * fabs(num)) {  // direction almost along the parabola axis
    Xc[0] = Xc[1] = X;
  } else {
    T tanA = num / den;
    T tmp = tanA * tanA - (X * tanA - Z) / f;
    if (tmp < T(0)) return T(kBig);
    T sq = sqrt(tmp);
    Xc[0] = T(2) * f * (tanA + sq);
    Xc[1] = T(2) * f * (tanA - sq);
  }
  const T tn = T(p[8]);
  T best = T(kBig);
  for (int k = 0; k < 2; k++) {
    T Zc = Xc[k] * Xc[k] / T(4) / f;
    T xc = ct * Xc[k] - st * (Zc - f) - r2;
    T zc = st * Xc[k] + ct * (Zc - f) - dz;
    T yc;
    if (fabs(px) <= fabs(pz) && fabs(py) <= fabs(pz)) yc = y + (zc - z) * py / pz;
    else if (fabs(py) <= fabs(px) && fabs(pz) <= fabs(px)) yc = y + (xc - x) * py / px;
    else yc = y + (fabs(px) < T(1e-5) ? (zc - z) * py / pz : (xc - x) * py / px);
    T ex = xc - x, ey = yc - y, ez = zc - z;
    if (xc < r2 || r1 < xc || zc < -dz || dz < zc || ex * px + ez * pz < T(0)) continue;
    bool in_sector = half_plane ? xc >= T(0) : (xc > T(0) && fabs(yc) <= xc * tn);
    if (!in_sector) continue;
    T dd = sqrt(ex * ex + ey * ey + ez * ez);
    best = dd < best ? dd : best;
  }
  return best;
}
template <typename T> __device__ RB_HEAVY_FN bool wins_contains(const double* p, V3<T> q) {
  if (fabs(q.z) > T(p[4])) return false;
  T r;
  if (!wins_r<T>(p, q.z, r)) return false;
  return wins_inside_polygon<T>(p, q.x, q.y, r);
}
template <typename T> __device__ RB_HEAVY_FN T wins_dist_in(const double* p, V3<T> q, V3<T> d) {
  const T dz = T(p[4]);
  T best = T(kBig);
  if (d.z < T(0)) best = (-q.z - dz) / d.z;
  else if (d.z > T(0)) best = (dz - q.z) / d.z;
  const int n = int(p[2]);
  for (int i = 0; i < n; i++) {
    T s = wins_parabola<T>(p, i, q, d, true);
    best = s < best ? s : best;
  }
  return best;
}
template <typename T> __device__ RB_HEAVY_FN T wins_dist_out(const double* p, V3<T> q, V3<T> d) {
  const T dz = T(p[4]);
  if (q.z <= -dz) {
    if (d.z <= T(0)) return T(kBig);
    T s = -(dz + q.z) / d.z;
    if (wins_inside_polygon<T>(p, q.x + s * d.x, q.y + s * d.y, T(p[1]))) return s;
  } else if (q.z >= dz) {
    if (d.z >= T(0)) return T(kBig);
    T s = (dz - q.z) / d.z;
    if (wins_inside_polygon<T>(p, q.x + s * d.x, q.y + s * d.y, T(p[0]))) return s;
  }
  T best = T(kBig);
  const int n = int(p[2]);
  for (int i = 0; i < n; i++) {
    T s = wins_parabola<T>(p, i, q, d, false);
    best = s < best ? s : best;
  }
  return best;
}
template <typename T> __device__ RB_HEAVY_FN V3<T> wins_normal(const double* p, V3<T> q, V3<T> d) {
  const int n = int(p[2]);
  T saf0 = fabs(fabs(T(p[4])) - fabs(q.z));
  int k = 0;
  T m = T(-kBig);
  for (int i = 0; i < n; i++) {
    T v = q.x * T(p[9 + 2 * i]) + q.y * T(p[10 + 2 * i]);
    if (v > m) { m = v; k = i; }
  }
  T r, saf1 = T(kBig);
  if (wins_r<T>(p, q.z, r)) saf1 = fabs(r - m);
  V3<T> nn;
  if (saf0 <= saf1) {
    nn = mk<T>(T(0), T(0), T(1));
  } else {
    T g = T(0);
    wins_drdz<T>(p, q.z, g);
    nn = mk<T>(T(p[9 + 2 * k]), T(p[10 + 2 * k]), -g);
    nn = scale(nn, T(1) / sqrt(dot(nn, nn)));
  }
  if (dot(nn, d) < T(0)) nn = scale(nn, T(-1));
  return nn;
}

// ------------------------------------------------------------ Winston cone (2D trough)
// AGeoWinstonCone2D (src/AGeoWinstonCone2D.cxx): the compound-parabolic
// profile |x| <= R(z) of the cone above, extruded over |y| <= dy. params:
// r1 r2 2 theta dz f cos(theta) sin(theta) 0 {1, 0} {-1, 0} dy -- the
// polygonal cone's layout with its two sides at phi = 0 and pi, then dy.
// The side crossings accept the half-plane of their side, as
// DistToParabola(point, dir, phi, pi) does in both distances (:203-322).
template <typename T> __device__ RB_POLY_FN bool wins2d_contains(const double* p, V3<T> q) {
  return fabs(q.y) <= T(p[13]) && wins_contains<T>(p, q);
}
template <typename T> __device__ RB_POLY_FN T wins2d_dist_in(const double* p, V3<T> q, V3<T> d) {
  const T dy = T(p[13]);
  T best = wins_dist_in<T>(p, q, d);
  if (d.y < T(0)) {
    const T s = -(q.y + dy) / d.y;
    best = s < best ? s : best;
  } else if (d.y > T(0)) {
    const T s = (dy - q.y) / d.y;
    best = s < best ? s : best;
  }
  return best;
}
// the end caps (within the apertures and |y| <= dy), the y walls (within
// |z| <= dz and |x| <= R(z)), then the two parabolic sides within |y| <= dy
template <typename T> __device__ RB_POLY_FN T wins2d_dist_out(const double* p, V3<T> q, V3<T> d) {
  const T dz = T(p[4]), dy = T(p[13]);
  if (q.z <= -dz || q.z >= dz) {
    const bool below = q.z <= -dz;
    if (below ? d.z <= T(0) : d.z >= T(0)) return T(kBig);
    const T s = below ? -(dz + q.z) / d.z : (dz - q.z) / d.z;
    if (fabs(q.x + s * d.x) <= T(p[below ? 1 : 0]) && fabs(q.y + s * d.y) <= dy) return s;
  }
  if (q.y <= -dy || q.y >= dy) {
    const bool below = q.y <= -dy;
    if (below ? d.y <= T(0) : d.y >= T(0)) return T(kBig);
    const T s = below ? -(dy + q.y) / d.y : (dy - q.y) / d.y;
    const T zn = q.z + s * d.z;
    T r;
    if (fabs(zn) <= dz && wins_r<T>(p, zn, r) && fabs(q.x + s * d.x) <= r) return s;
  }
  T best = T(kBig);
  for (int i = 0; i < 2; i++) {
    const T s = wins_parabola<T>(p, i, q, d, true);
    if (s < best && fabs(q.y + s * d.y) <= dy) best = s;
  }
  return best;
}
// the nearest of the y walls, the end caps and the parabolic side, in that
// order of preference (ComputeNormal, :120-172)
template <typename T> __device__ RB_POLY_FN V3<T> wins2d_normal(const double* p, V3<T> q, V3<T> d) {
  const T saf0 = fabs(T(p[13]) - fabs(q.y)), saf1 = fabs(T(p[4]) - fabs(q.z));
  T r, saf2 = T(kBig);
  if (wins_r<T>(p, q.z, r)) saf2 = fabs(r - fabs(q.x));
  V3<T> n;
  if (saf0 <= saf1 && saf0 <= saf2) {
    n = mk<T>(T(0), T(1), T(0));
  } else if (saf1 <= saf2) {
    n = mk<T>(T(0), T(0), T(1));
  } else {
    T g = T(0);
    wins_drdz<T>(p, q.z, g);
    n = mk<T>(T(1), T(0), q.x > T(0) ? -g : g);
    n = scale(n, T(1) / sqrt(dot(n, n)));
  }
  if (dot(n, d) < T(0)) n = scale(n, T(-1));
  return n;
}

// Call structure of the shape code. The node-level queries and the
// primitive dispatch are inlined; the composite levels (Csg<T, D>) stay real
// calls, which keeps each boolean level's register allocation to itself.
// Measured on MI355X (85-facet Davies-Cotton, 1e7 rays, FP64): this split
// 15.2 ms; composite level 0 inlined too 20.3 ms; everything inlined 27.5 ms;
// nothing inlined 19.6 ms. A build can override each level.
#ifndef RB_PRIM_FN
#define RB_PRIM_FN __forceinline__
#endif
#ifndef RB_CSG0_FN
#define RB_CSG0_FN __noinline__
#endif
#ifndef RB_CSG_FN
#define RB_CSG_FN __noinline__
#endif
#ifndef RB_NODE_FN
#define RB_NODE_FN __forceinline__
#endif

// ------------------------------------------------------------ dispatch
template <typename T, unsigned SH> __device__ RB_PRIM_FN bool prim_contains(int type, const double* p, V3<T> q) {
  switch (type) {
    case RB_SHAPE_BOX: return box_contains<T>(p, q);
    case RB_SHAPE_TUBE: return tube_contains<T>(p, q);
    case RB_SHAPE_SPHERE: return sphere_contains<T>(p, q);
    case RB_SHAPE_PARABOLOID: return para_contains<T>(p, q);
    case RB_SHAPE_PGON: return pgon_contains<T>(p, q);
    case kShapePgonGeneral: if constexpr ((SH & kShapesExt) != 0u) return poly_contains<T, true>(p, q); break;
    case RB_SHAPE_PCON: if constexpr ((SH & kShapesExt) != 0u) return poly_contains<T, false>(p, q); break;
    case RB_SHAPE_WINSTON2D: if constexpr ((SH & kShapesExt) != 0u) return wins2d_contains<T>(p, q); break;
    case kShapeSphereWedge: if constexpr ((SH & kShapesExt) != 0u) return sphere_contains<T, true>(p, q); break;
    case RB_SHAPE_ASPHERIC: if constexpr ((SH & kShapesSolver) != 0u) return asph_contains<T>(p, q); break;
    case RB_SHAPE_WINSTON: if constexpr ((SH & kShapesSolver) != 0u) return wins_contains<T>(p, q); break;
  }
  return false;
}
template <typename T, unsigned SH> __device__ RB_PRIM_FN T prim_dist_out(int type, const double* p, V3<T> q, V3<T> d) {
  switch (type) {
    case RB_SHAPE_BOX: return box_dist_out<T>(p, q, d);
    case RB_SHAPE_TUBE: return tube_dist_out<T>(p, q, d);
    case RB_SHAPE_SPHERE: return sphere_dist<T, true>(p, q, d);
    case RB_SHAPE_PARABOLOID: return para_dist_out<T>(p, q, d);
    case RB_SHAPE_PGON: return pgon_dist_out<T>(p, q, d);
    case kShapePgonGeneral: if constexpr ((SH & kShapesExt) != 0u) return poly_dist_out<T, true>(p, q, d); break;
    case RB_SHAPE_PCON: if constexpr ((SH & kShapesExt) != 0u) return poly_dist_out<T, false>(p, q, d); break;
    case RB_SHAPE_WINSTON2D: if constexpr ((SH & kShapesExt) != 0u) return wins2d_dist_out<T>(p, q, d); break;
    case kShapeSphereWedge: if constexpr ((SH & kShapesExt) != 0u) return sphere_dist<T, true, true>(p, q, d); break;
    case RB_SHAPE_ASPHERIC: if constexpr ((SH & kShapesSolver) != 0u) return asph_dist<T>(p, q, d); break;
    case RB_SHAPE_WINSTON: if constexpr ((SH & kShapesSolver) != 0u) return wins_dist_out<T>(p, q, d); break;
  }
  return T(kBig);
}
template <typename T, unsigned SH> __device__ RB_PRIM_FN T prim_dist_in(int type, const double* p, V3<T> q, V3<T> d) {
  switch (type) {
    case RB_SHAPE_BOX: return box_dist_in<T>(p, q, d);
    case RB_SHAPE_TUBE: return tube_dist_in<T>(p, q, d);
    case RB_SHAPE_SPHERE: return sphere_dist<T, false>(p, q, d);
    case RB_SHAPE_PARABOLOID: return para_dist_in<T>(p, q, d);
    case RB_SHAPE_PGON: return pgon_dist_in<T>(p, q, d);
    case kShapePgonGeneral: if constexpr ((SH & kShapesExt) != 0u) return poly_dist_in<T, true>(p, q, d); break;
    case RB_SHAPE_PCON: if constexpr ((SH & kShapesExt) != 0u) return poly_dist_in<T, false>(p, q, d); break;
    case RB_SHAPE_WINSTON2D: if constexpr ((SH & kShapesExt) != 0u) return wins2d_dist_in<T>(p, q, d); break;
    case kShapeSphereWedge: if constexpr ((SH & kShapesExt) != 0u) return sphere_dist<T, false, true>(p, q, d); break;
    case RB_SHAPE_ASPHERIC: if constexpr ((SH & kShapesSolver) != 0u) return asph_dist<T>(p, q, d); break;
    case RB_SHAPE_WINSTON: if constexpr ((SH & kShapesSolver) != 0u) return wins_dist_in<T>(p, q, d); break;
  }
  return T(0);
}
template <typename T, unsigned SH> __device__ RB_PRIM_FN V3<T> prim_normal(int type, const double* p, V3<T> q, V3<T> d) {
  switch (type) {
    case RB_SHAPE_BOX: return box_normal<T>(p, q, d);
    case RB_SHAPE_TUBE: return tube_normal<T>(p, q, d);
    case RB_SHAPE_SPHERE: return sphere_normal<T>(p, q, d);
    case RB_SHAPE_PARABOLOID: return para_normal<T>(p, q, d);
    case RB_SHAPE_PGON: return pgon_normal<T>(p, q, d);
    case kShapePgonGeneral: if constexpr ((SH & kShapesExt) != 0u) return poly_normal<T, true>(p, q, d); break;
    case RB_SHAPE_PCON: if constexpr ((SH & kShapesExt) != 0u) return poly_normal<T, false>(p, q, d); break;
    case RB_SHAPE_WINSTON2D: if constexpr ((SH & kShapesExt) != 0u) return wins2d_normal<T>(p, q, d); break;
    case kShapeSphereWedge: if constexpr ((SH & kShapesExt) != 0u) return sphere_normal<T, true>(p, q, d); break;
    case RB_SHAPE_ASPHERIC: if constexpr ((SH & kShapesSolver) != 0u) return asph_normal<T>(p, q, d); break;
    case RB_SHAPE_WINSTON: if constexpr ((SH & kShapesSolver) != 0u) return wins_normal<T>(p, q, d); break;
  }
  return mk<T>(T(0), T(0), T(1));
}

// Composite (TGeoBoolNode) params: op left right lmat[12] rmat[12].
// Depth-templated so that nested expressions unroll without recursion.
// The inside tests of the boundary-to-boundary iteration are taken `push`
// past each boundary point (kBoolPush in FP64; larger in FP32 search).
template <typename T> __device__ __forceinline__ T bool_push() { return T(kBoolPush); }
template <> __device__ __forceinline__ float bool_push<float>() { return 2e-4f; }

// A point in a shape's local box (lo[3], hi[3]; grown on upload), and a
// ray that may enter it: false only when the ray (origin q, direction d,
// t >= 0) certainly misses the box.
template <typename T>
__device__ __forceinline__ bool sbox_has(const double* b, V3<T> q) {
  return q.x >= T(b[0]) && q.x <= T(b[3]) && q.y >= T(b[1]) && q.y <= T(b[4]) && q.z >= T(b[2]) && q.z <= T(b[5]);
}
template <typename T>
__device__ __forceinline__ bool sbox_ray(const double* b, V3<T> q, V3<T> d) {
  const T o[3] = {q.x, q.y, q.z}, v[3] = {d.x, d.y, d.z};
  T t0 = T(0), t1 = T(kBig);
  for (int a = 0; a < 3; a++) {
    const T lo = T(b[a]), hi = T(b[3 + a]);
    if (v[a] == T(0)) {
      if (o[a] < lo || o[a] > hi) return false;
    } else {
      const T inv = T(1) / v[a];
      T ta = (lo - o[a]) * inv, tb = (hi - o[a]) * inv;
      if (ta > tb) { const T x = ta; ta = tb; tb = x; }
      t0 = ta > t0 ? ta : t0;
      t1 = tb < t1 ? tb : t1;
    }
  }
  return t0 <= t1 * T(1.0001) + T(1e-3);
}

template <typename T, int D, unsigned SH, bool BOX = false> struct Csg {
  // child `side` (0 left, 1 right) contains q: outside its box it does not
  __device__ __forceinline__ static bool child_in(const DScene& S, const double* p, int side, V3<T> q) {
    const double* m = p + (side ? 15 : 3);
    const int c = int(p[side ? 2 : 1]);
    const V3<T> lq = m2l<T>(m, m + 9, q);
    return (!BOX || sbox_has<T>(S.sbox + 6 * c, lq)) && Csg<T, D - 1, SH, BOX>::contains(S, c, lq);
  }
  __device__ RB_CSG_FN static bool contains(const DScene& S, int shape, V3<T> q) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    if (sh.type != RB_SHAPE_COMPOSITE) return prim_contains<T, SH>(sh.type, p, q);
    int op = int(p[0]);
    bool a = child_in(S, p, 0, q);
    if (op == RB_OP_UNION) return a || child_in(S, p, 1, q);
    if (op == RB_OP_SUBTRACTION) return a && !child_in(S, p, 1, q);
    return a && child_in(S, p, 1, q);
  }
  // distance to child `side`; an entry (out) along a ray that misses the
  // child's box is kBig without evaluating the child
  __device__ static Hit child(const DScene& S, const double* p, int side, bool out, V3<T> q, V3<T> d) {
    const double* m = p + (side ? 15 : 3);
    int c = int(p[side ? 2 : 1]);
    V3<T> lq = m2l<T>(m, m + 9, q), ld = m2lv<T>(m, d);
    if (BOX && out && !sbox_ray<T>(S.sbox + 6 * c, lq, ld)) return {kBig, 0u, 0u};
    Hit h = out ? Csg<T, D - 1, SH, BOX>::dist_out(S, c, lq, ld) : Csg<T, D - 1, SH, BOX>::dist_in(S, c, lq, ld);
    h.leaf = leaf_push(h.leaf, side);
    return h;
  }
  // Entry distance from outside. `lim`: a caller that rejects any crossing
  // at or beyond lim (the FP64 search: its best + tolerance) gets kBig as
  // soon as the iteration has passed it; the accumulated distance only grows,
  // so this changes no accepted result.
  __device__ RB_CSG_FN static Hit dist_out(const DScene& S, int shape, V3<T> q, V3<T> d, double lim = kBig) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    if (sh.type != RB_SHAPE_COMPOSITE) return {double(prim_dist_out<T, SH>(sh.type, p, q, d)), 0u, 0u};
    const int op = int(p[0]);
    if (op == RB_OP_UNION) {
      Hit a = child(S, p, 0, true, q, d), b = child(S, p, 1, true, q, d);
      return b.s < a.s ? b : a;
    }
    const T push = bool_push<T>();
    T s = T(0);
    Hit last = {0.0, 0u, 0u};
    for (int it = 0; it < 64; it++) {
      V3<T> x = axpy(q, s, d);
      V3<T> xp = axpy(x, push, d);
      Hit h;
      if (op == RB_OP_SUBTRACTION) {
        // inside B: leave B first; A is tested only outside B
        const bool inB = child_in(S, p, 1, xp);
        if (inB) {
          h = child(S, p, 1, false, x, d);
        } else {
          if (child_in(S, p, 0, xp)) return {double(s), last.leaf, last.exiting};
          h = child(S, p, 0, true, x, d);
          if (h.s >= kBig) return {kBig, 0u, 0u};
        }
      } else {
        const bool inA = child_in(S, p, 0, xp);
        const bool inB = child_in(S, p, 1, xp);
        if (inA && inB) return {double(s), last.leaf, last.exiting};
        Hit ha = {0.0, 0u, 0u}, hb = {0.0, 0u, 0u};
        if (!inA) {
          ha = child(S, p, 0, true, x, d);
          if (ha.s >= kBig) return {kBig, 0u, 0u};
        }
        if (!inB) {
          hb = child(S, p, 1, true, x, d);
          if (hb.s >= kBig) return {kBig, 0u, 0u};
        }
        h = (!inA && (inB || ha.s >= hb.s)) ? ha : hb;
      }
      T step = T(h.s);
      last = h;
      if (step < push) step = push;
      s += step;
      if (double(s) >= lim) return {kBig, 0u, 0u};
    }
    return {kBig, 0u, 0u};
  }
  __device__ RB_CSG_FN static Hit dist_in(const DScene& S, int shape, V3<T> q, V3<T> d) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    if (sh.type != RB_SHAPE_COMPOSITE) return {double(prim_dist_in<T, SH>(sh.type, p, q, d)), 0u, 1u};
    const int op = int(p[0]);
    if (op == RB_OP_INTERSECTION) {
      Hit a = child(S, p, 0, false, q, d), b = child(S, p, 1, false, q, d);
      return b.s < a.s ? b : a;
    }
    if (op == RB_OP_SUBTRACTION) {
      Hit a = child(S, p, 0, false, q, d), b = child(S, p, 1, true, q, d);
      return b.s < a.s ? b : a;
    }
    const T push = bool_push<T>();
    T s = T(0);
    Hit last = {0.0, 0u, 1u};
    for (int it = 0; it < 64; it++) {
      V3<T> x = axpy(q, s, d);
      V3<T> xp = axpy(x, push, d);
      bool inA = child_in(S, p, 0, xp);
      bool inB = child_in(S, p, 1, xp);
      if (!inA && !inB) return {double(s), last.leaf, last.exiting};
      Hit ha = {0.0, 0u, 1u}, hb = {0.0, 0u, 1u};
      if (inA) ha = child(S, p, 0, false, x, d);
      if (inB) hb = child(S, p, 1, false, x, d);
      Hit h = (inA && (!inB || ha.s >= hb.s)) ? ha : hb;
      T step = T(h.s) < push ? push : T(h.s);
      last = h;
      s += step;
    }
    return {double(s), last.leaf, last.exiting};
  }
  // normal of the leaf surface `leaf` (path bits consumed from bit 0)
  __device__ RB_CSG_FN static V3<T> normal(const DScene& S, int shape, uint32_t path, int depth, V3<T> q, V3<T> d) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    if (sh.type != RB_SHAPE_COMPOSITE) return prim_normal<T, SH>(sh.type, p, q, d);
    int side = depth > 0 ? int(path & 1u) : 0;
    const double* m = p + (side ? 15 : 3);
    V3<T> n = Csg<T, D - 1, SH, BOX>::normal(S, int(p[side ? 2 : 1]), path >> 1, depth > 0 ? depth - 1 : 0,
                                    m2l<T>(m, m + 9, q), m2lv<T>(m, d));
    return l2mv<T>(m, n);
  }
  // FP64 crossing of the leaf primitive `path` near the estimate s0 along
  // the ray: the primitive's own entry/exit distance from a point backed
  // off by `back` (the mixed form's root polish). Returns kBig on failure.
  __device__ RB_CSG_FN static double polish(const DScene& S, int shape, uint32_t path, int depth, bool exiting, V3<T> q,
                                  V3<T> d, T s0, T back) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    if (sh.type != RB_SHAPE_COMPOSITE) {
      T t0 = s0 - back > T(0) ? s0 - back : T(0);
      V3<T> q0 = axpy(q, t0, d);
      T t = exiting ? prim_dist_in<T, SH>(sh.type, p, q0, d) : prim_dist_out<T, SH>(sh.type, p, q0, d);
      return t >= T(kBig) ? kBig : double(t0 + t);
    }
    int side = depth > 0 ? int(path & 1u) : 0;
    const double* m = p + (side ? 15 : 3);
    return Csg<T, D - 1, SH, BOX>::polish(S, int(p[side ? 2 : 1]), path >> 1, depth > 0 ? depth - 1 : 0, exiting,
                                 m2l<T>(m, m + 9, q), m2lv<T>(m, d), s0, back);
  }
};
// depth exhausted: primitives only (the host rejects deeper expressions)
template <typename T, unsigned SH, bool BOX> struct Csg<T, 0, SH, BOX> {
  __device__ RB_CSG0_FN static bool contains(const DScene& S, int shape, V3<T> q) {
    DShape sh = S.shapes[shape];
    return prim_contains<T, SH>(sh.type, S.params + sh.off, q);
  }
  __device__ RB_CSG0_FN static Hit dist_out(const DScene& S, int shape, V3<T> q, V3<T> d, double = kBig) {
    DShape sh = S.shapes[shape];
    return {double(prim_dist_out<T, SH>(sh.type, S.params + sh.off, q, d)), 0u, 0u};
  }
  __device__ RB_CSG0_FN static Hit dist_in(const DScene& S, int shape, V3<T> q, V3<T> d) {
    DShape sh = S.shapes[shape];
    return {double(prim_dist_in<T, SH>(sh.type, S.params + sh.off, q, d)), 0u, 1u};
  }
  __device__ RB_CSG0_FN static V3<T> normal(const DScene& S, int shape, uint32_t, int, V3<T> q, V3<T> d) {
    DShape sh = S.shapes[shape];
    return prim_normal<T, SH>(sh.type, S.params + sh.off, q, d);
  }
  __device__ RB_CSG0_FN static double polish(const DScene& S, int shape, uint32_t, int, bool exiting, V3<T> q, V3<T> d, T s0,
                                  T back) {
    DShape sh = S.shapes[shape];
    const double* p = S.params + sh.off;
    T t0 = s0 - back > T(0) ? s0 - back : T(0);
    V3<T> q0 = axpy(q, t0, d);
    T t = exiting ? prim_dist_in<T, SH>(sh.type, p, q0, d) : prim_dist_out<T, SH>(sh.type, p, q0, d);
    return t >= T(kBig) ? kBig : double(t0 + t);
  }
};

// Entry points by the shape's own nesting depth, so that a primitive or a
// single boolean level does not pay for the deepest unrolled form. Scenes
// typed DSceneOf<1> compile no deeper form at all.
template <typename T, typename SC>
__device__ __forceinline__ V3<T> shape_normal(const SC& S, int shape, uint32_t leaf, V3<T> q, V3<T> d) {
  const uint32_t path = leaf & 0xffffffu;
  const int depth = int(leaf >> 24);
  switch (S.shapes[shape].depth) {
    case 0: return Csg<T, 0, SC::shape_set, SC::box_on>::normal(S, shape, path, depth, q, d);
    case 1: return Csg<T, 1, SC::shape_set, SC::box_on>::normal(S, shape, path, depth, q, d);
    case 2:
      if constexpr (SC::max_depth >= 3) return Csg<T, 2, SC::shape_set, SC::box_on>::normal(S, shape, path, depth, q, d);
      break;
  }
  return Csg<T, SC::max_depth, SC::shape_set, SC::box_on>::normal(S, shape, path, depth, q, d);
}
template <typename T, typename SC>
__device__ __forceinline__ bool shape_contains(const SC& S, int shape, V3<T> q) {
  switch (S.shapes[shape].depth) {
    case 0: return Csg<T, 0, SC::shape_set, SC::box_on>::contains(S, shape, q);
    case 1: return Csg<T, 1, SC::shape_set, SC::box_on>::contains(S, shape, q);
    case 2:
      if constexpr (SC::max_depth >= 3) return Csg<T, 2, SC::shape_set, SC::box_on>::contains(S, shape, q);
      break;
  }
  return Csg<T, SC::max_depth, SC::shape_set, SC::box_on>::contains(S, shape, q);
}
template <typename T, typename SC>
__device__ __forceinline__ Hit shape_dist(const SC& S, int shape, V3<T> q, V3<T> d, bool out, double lim = kBig) {
  switch (S.shapes[shape].depth) {
    case 0: return out ? Csg<T, 0, SC::shape_set, SC::box_on>::dist_out(S, shape, q, d) : Csg<T, 0, SC::shape_set, SC::box_on>::dist_in(S, shape, q, d);
    case 1: return out ? Csg<T, 1, SC::shape_set, SC::box_on>::dist_out(S, shape, q, d, lim) : Csg<T, 1, SC::shape_set, SC::box_on>::dist_in(S, shape, q, d);
    case 2:
      if constexpr (SC::max_depth >= 3) return out ? Csg<T, 2, SC::shape_set, SC::box_on>::dist_out(S, shape, q, d, lim) : Csg<T, 2, SC::shape_set, SC::box_on>::dist_in(S, shape, q, d);
      break;
  }
  return out ? Csg<T, SC::max_depth, SC::shape_set, SC::box_on>::dist_out(S, shape, q, d, lim) : Csg<T, SC::max_depth, SC::shape_set, SC::box_on>::dist_in(S, shape, q, d);
}
template <typename T, typename SC>
__device__ __forceinline__ double shape_polish(const SC& S, int shape, uint32_t leaf, bool exiting, V3<T> q,
                                               V3<T> d, T s0, T back) {
  const uint32_t path = leaf & 0xffffffu;
  const int depth = int(leaf >> 24);
  switch (S.shapes[shape].depth) {
    case 0: return Csg<T, 0, SC::shape_set, SC::box_on>::polish(S, shape, path, depth, exiting, q, d, s0, back);
    case 1: return Csg<T, 1, SC::shape_set, SC::box_on>::polish(S, shape, path, depth, exiting, q, d, s0, back);
    case 2:
      if constexpr (SC::max_depth >= 3) return Csg<T, 2, SC::shape_set, SC::box_on>::polish(S, shape, path, depth, exiting, q, d, s0, back);
      break;
  }
  return Csg<T, SC::max_depth, SC::shape_set, SC::box_on>::polish(S, shape, path, depth, exiting, q, d, s0, back);
}

// ------------------------------------------------------------ physics
__device__ inline double table_eval(const DScene& S, int table, double x) {
  rb_table_desc t = S.tables[table];
  const double* xy = S.table_xy + 2 * t.offset;
  int n = t.n;
  if (n == 0) return 0.0;
  if (n == 1) return xy[1];
  // binary search on ascending x: first index with x_i >= x
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xy[2 * mid] < x) lo = mid + 1;
    else hi = mid;
  }
  if (lo < n && xy[2 * lo] == x) return xy[2 * lo + 1];
  double xl, yl, xu, yu;
  if (lo == n) {  // above the range: stored (up, low) pair
    const double* e = xy + 2 * n;
    xu = e[0]; yu = e[1]; xl = e[2]; yl = e[3];
  } else if (lo == 0) {  // below the range: stored (low, up) pair
    const double* e = xy + 2 * n + 4;
    xl = e[0]; yl = e[1]; xu = e[2]; yu = e[3];
  } else {
    xl = xy[2 * (lo - 1)]; yl = xy[2 * (lo - 1) + 1]; xu = xy[2 * lo]; yu = xy[2 * lo + 1];
  }
  if (xl == xu) return yl;
  return yu + (x - xu) * (yl - yu) / (xl - xu);
}
// Rarely taken math (absorption depth, focal-surface angle, Schott index) is
// kept out of line: inlined into the tracing loop, the compiler hoists the
// polynomial constants of log/acos/pow out of the loop into registers and
// then spills them for every ray.
#ifndef RB_COLD_FN
#define RB_COLD_FN __noinline__
#endif
__device__ RB_COLD_FN double cold_log(double x) { return log(x); }
__device__ RB_COLD_FN double cold_acos(double x) { return acos(x); }
__device__ RB_COLD_FN double schott_index(const double* a, double lambda) {
  double l = lambda / 1e-4;
  return sqrt(a[0] + a[1] * pow(l, 2.) + a[2] * pow(l, -2.) + a[3] * pow(l, -4.) + a[4] * pow(l, -6.) +
              a[5] * pow(l, -8.));
}
__device__ inline double lens_index(const DScene& S, const rb_volume_desc& v, double lambda) {
  if (v.index_model == RB_INDEX_TABLE) return table_eval(S, v.index_ref, lambda);
  if (v.index_model == RB_INDEX_SELLMEIER) {
    const double* b = S.params + v.index_ref;
    double l = lambda / 1e-4;
    double l2 = l * l;
    return sqrt(1 + b[0] * l2 / (l2 - b[3]) + b[1] * l2 / (l2 - b[4]) + b[2] * l2 / (l2 - b[5]));
  }
  if (v.index_model == RB_INDEX_SCHOTT) return schott_index(S.params + v.index_ref, lambda);
  return 1.0;
}
__device__ inline double lens_k(const DScene& S, const rb_volume_desc& v, double lambda) {
  return v.k_table >= 0 ? table_eval(S, v.k_table, lambda) : 0.0;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
__device__ __forceinline__ void philox(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                       uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t h0 = __umulhi(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
    uint32_t h1 = __umulhi(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
    uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0; c1 = l1; c2 = n2; c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// uniform in (0, 1] from counter (id, draw, stream) under key seed
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t id, uint32_t draw, uint32_t stream) {
  uint32_t c0 = uint32_t(id), c1 = uint32_t(id >> 32), c2 = draw, c3 = stream;
  philox(c0, c1, c2, c3, uint32_t(seed), uint32_t(seed >> 32));
  uint64_t u = ((uint64_t(c1) << 32) | c0) >> 11;
  return double(u + 1) * (1.0 / 9007199254740992.0);
}

}  // namespace rbd
