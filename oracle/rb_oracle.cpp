// rb_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of ROBAST's non-sequential tracer, used as the
// correctness oracle for librbgpu and as the timed CPU baseline. Nothing in
// the product (librbgpu, robast_amd) links, loads or calls this file; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// It restates, from the reference source (read as text):
//   src/AOpticsManager.cxx:335-520  per-ray state machine (trace_ray)
//   src/AOpticsManager.cxx:170-247  DoReflection            (do_reflection)
//   src/AOpticsManager.cxx:52-167   DoFresnel               (do_fresnel)
//   src/AOpticsManager.cxx:523-587  ray-chunk dispatch      (oracle_trace)
//   src/AMirror.cxx:39-60, src/ALens.cxx:36-60, include/ARefractiveIndex.h:36-65,
//   src/ASellmeierFormula.cxx:46-54, src/ASchottFormula.cxx:43-55,
//   src/AFocalSurface.cxx:35-52                              (physics tables)
// and, because ROOT is not in the reference tree, the TGeo navigator and
// primitive-shape behaviour it relies on (TGeoNavigator::FindNextBoundaryAndStep,
// CrossBoundaryAndLocate, Step, FindNode; TGeoBBox/Tube/Sphere/Paraboloid/Pgon;
// TGeoBoolNode) from ROOT's documented semantics. Those ROOT parts are
// "parity unpinned" against ROOT itself; they are pinned against closed forms,
// brute force and invariants in tests/.
//
// Deliberately plain: AoS, recursion, brute-force loops over daughters, no
// BVH, no FMA contraction (build with -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/rbgpu.h"

namespace orc {

constexpr double kBig = 1e30;        // TGeoShape::Big()
constexpr double kTol = 1e-10;       // TGeoShape::Tolerance()
constexpr double kEpsilon = 1e-6;    // src/AOpticsManager.cxx:20
constexpr double kC = 2.99792458e10; // TMath::C() * AOpticsManager::m(), cm/s
// The navigator constants ROOT fixes internally, which only an import of
// ROOT could pin. They are settable (oracle_set_navigator) so that the
// sensitivity of the results to each can be measured; the defaults are the
// rules the device follows.
static double g_step_back = kEpsilon + 1e-6;  // TGeoNavigator::Step() after SetStep(kEpsilon), cm
static double g_push_scale = 100.;             // CrossBoundaryAndLocate push: g_push_scale (1 + |T| + step) kTol
static double g_bool_push = 1e-8;              // inside-test offset of the boolean iteration, cm
constexpr double kPi = 3.14159265358979323846;

// Operation counts for the algorithmic FLOP figure (oracle_trace_counts):
// every transform, primitive distance / inside test, aspheric Newton step,
// Winston side solve, normal, reflection, refraction and bounding-box test
// the trace performs. Off (null) for ordinary traces.
enum OpKind {
  OP_XFORM_P, OP_XFORM_V,
  OP_DIST_BOX, OP_DIST_TUBE, OP_DIST_SPHERE, OP_DIST_PARA, OP_DIST_PGON, OP_DIST_WINS_CAP,
  OP_ASPH_SEED, OP_ASPH_NEWTON, OP_ASPH_WALL, OP_WINS_SIDE,
  OP_IN_BOX, OP_IN_TUBE, OP_IN_SPHERE, OP_IN_PARA, OP_IN_PGON, OP_IN_ASPH, OP_IN_WINS,
  OP_NORMAL, OP_REFLECT, OP_REFRACT, OP_BOX_TEST, OP_STEP,
  OP_COUNT
};
static thread_local uint64_t* g_ops = nullptr;
static inline void count_op(int k) {
  if (g_ops) g_ops[k]++;
}

struct V3 {
  double x, y, z;
};
static inline V3 v3(double x, double y, double z) { return {x, y, z}; }
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 axpy(V3 p, double s, V3 d) { return {p.x + s * d.x, p.y + s * d.y, p.z + s * d.z}; }

// TGeoMatrix conventions: master = R local + T
static inline V3 m2l(const rb_matrix& m, V3 p) {
  count_op(OP_XFORM_P);
  V3 q = sub(p, v3(m.tr[0], m.tr[1], m.tr[2]));
  const double* r = m.rot;
  return {r[0] * q.x + r[3] * q.y + r[6] * q.z, r[1] * q.x + r[4] * q.y + r[7] * q.z,
          r[2] * q.x + r[5] * q.y + r[8] * q.z};
}
static inline V3 m2lv(const rb_matrix& m, V3 q) {
  count_op(OP_XFORM_V);
  const double* r = m.rot;
  return {r[0] * q.x + r[3] * q.y + r[6] * q.z, r[1] * q.x + r[4] * q.y + r[7] * q.z,
          r[2] * q.x + r[5] * q.y + r[8] * q.z};
}
static inline V3 l2mv(const rb_matrix& m, V3 q) {
  const double* r = m.rot;
  return {r[0] * q.x + r[1] * q.y + r[2] * q.z, r[3] * q.x + r[4] * q.y + r[5] * q.z,
          r[6] * q.x + r[7] * q.y + r[8] * q.z};
}

struct Box {
  double lo[3], hi[3];
};

struct Scene {
  const rb_scene_desc* d;
  std::vector<Box> boxes;  // per node, global frame: the bounding-box pre-check
                           // every TGeo shape makes before its exact distance
  const double* P(int shape) const { return d->params + d->shapes[shape].param_offset; }
  int type(int shape) const { return d->shapes[shape].type; }
  const rb_matrix& mat(const double* p) const { return *reinterpret_cast<const rb_matrix*>(p); }
};

// A distance result and the leaf surface it came from. `leaf` encodes the
// path through composite shapes: bit k = right child taken at depth k,
// depth in bits 24..31.
struct Hit {
  double s;
  uint32_t leaf;
};
static inline uint32_t leaf_push(uint32_t leaf, int side) {
  uint32_t depth = leaf >> 24;
  return ((depth + 1) << 24) | (((leaf & 0xffffffu) << 1) | uint32_t(side));
}

// ---------------------------------------------------------------- box
static bool box_contains(const double* p, V3 q) {
  return std::fabs(q.x - p[3]) <= p[0] && std::fabs(q.y - p[4]) <= p[1] &&
         std::fabs(q.z - p[5]) <= p[2];
}
static double box_dist_in(const double* p, V3 q, V3 d) {
  double n[3] = {q.x - p[3], q.y - p[4], q.z - p[5]};
  double dd[3] = {d.x, d.y, d.z};
  double smin = kBig;
  for (int i = 0; i < 3; i++) {
    if (dd[i] != 0) {
      double s = (dd[i] > 0 ? p[i] - n[i] : p[i] + n[i]) / std::fabs(dd[i]);
      if (s < smin) smin = s;
    }
  }
  return smin < 0 ? 0 : smin;
}
static double box_dist_out(const double* p, V3 q, V3 d) {
  double n[3] = {q.x - p[3], q.y - p[4], q.z - p[5]};
  double dd[3] = {d.x, d.y, d.z};
  double saf[3];
  bool in = true;
  for (int i = 0; i < 3; i++) {
    saf[i] = std::fabs(n[i]) - p[i];
    if (saf[i] > 0) in = false;
  }
  if (in) return 0.0;
  for (int i = 0; i < 3; i++) {
    if (saf[i] < 0) continue;
    if (n[i] * dd[i] >= 0) continue;
    double s = saf[i] / std::fabs(dd[i]);
    bool ok = true;
    for (int j = 0; j < 3; j++) {
      if (j == i) continue;
      if (std::fabs(n[j] + s * dd[j]) > p[j]) {
        ok = false;
        break;
      }
    }
    if (ok) return s;
  }
  return kBig;
}
static V3 box_normal(const double* p, V3 q, V3 d) {
  double saf[3] = {std::fabs(std::fabs(q.x - p[3]) - p[0]), std::fabs(std::fabs(q.y - p[4]) - p[1]),
                   std::fabs(std::fabs(q.z - p[5]) - p[2])};
  int i = 0;
  if (saf[1] < saf[i]) i = 1;
  if (saf[2] < saf[i]) i = 2;
  double dd[3] = {d.x, d.y, d.z};
  double n[3] = {0, 0, 0};
  n[i] = dd[i] > 0 ? 1 : -1;
  return v3(n[0], n[1], n[2]);
}

// ---------------------------------------------------------------- tube
// roots of |(q + s d)_xy|^2 = R^2 : s = -b -/+ sqrt(b^2 - c)
static bool cyl_roots(V3 q, V3 d, double R, double& s1, double& s2) {
  double nsq = d.x * d.x + d.y * d.y;
  if (nsq < 1e-20) return false;
  double b = (q.x * d.x + q.y * d.y) / nsq;
  double c = (q.x * q.x + q.y * q.y - R * R) / nsq;
  double disc = b * b - c;
  if (disc <= 0) return false;
  double sq = std::sqrt(disc);
  s1 = -b - sq;
  s2 = -b + sq;
  return true;
}
static bool tube_contains(const double* p, V3 q) {
  if (std::fabs(q.z) > p[2]) return false;
  double r2 = q.x * q.x + q.y * q.y;
  return r2 >= p[0] * p[0] && r2 <= p[1] * p[1];
}
static double tube_dist_in(const double* p, V3 q, V3 d) {
  double rmin = p[0], rmax = p[1], dz = p[2];
  double sz = kBig;
  if (d.z != 0) {
    sz = ((d.z > 0 ? dz : -dz) - q.z) / d.z;
    if (sz <= 0) return 0.0;
  }
  double s1, s2, s = sz;
  if (rmin > 0 && cyl_roots(q, d, rmin, s1, s2) && s1 > 0) s = std::min(s, s1);
  if (cyl_roots(q, d, rmax, s1, s2)) {
    if (s2 > 0) s = std::min(s, s2);
    else s = 0;
  } else if (d.x * d.x + d.y * d.y >= 1e-20) {
    s = 0;  // no crossing of rmax although moving radially: numerically outside
  }
  return s < 0 ? 0 : s;
}
static double tube_dist_out(const double* p, V3 q, V3 d) {
  double rmin = p[0], rmax = p[1], dz = p[2];
  double best = kBig;
  // z planes
  if (std::fabs(q.z) >= dz && q.z * d.z < 0) {
    double s = (std::fabs(q.z) - dz) / std::fabs(d.z);
    double x = q.x + s * d.x, y = q.y + s * d.y, r2 = x * x + y * y;
    if (r2 >= rmin * rmin && r2 <= rmax * rmax) best = s;
  }
  // entering through the outer wall (first root), or out of the bore into
  // the solid (second root of the inner wall), inside the z slab
  double s1, s2;
  if (cyl_roots(q, d, rmax, s1, s2) && s1 >= 0 && s1 < best) {
    if (std::fabs(q.z + s1 * d.z) <= dz) best = s1;
  }
  if (rmin > 0 && cyl_roots(q, d, rmin, s1, s2) && s2 >= 0 && s2 < best) {
    if (std::fabs(q.z + s2 * d.z) <= dz) best = s2;
  }
  return best;
}
static V3 tube_normal(const double* p, V3 q, V3 d) {
  double r = std::sqrt(q.x * q.x + q.y * q.y);
  double saf0 = std::fabs(p[2] - std::fabs(q.z));
  double saf1 = p[0] > 1e-10 ? std::fabs(r - p[0]) : kBig;
  double saf2 = std::fabs(p[1] - r);
  int i = 0;
  double m = saf0;
  if (saf1 < m) { i = 1; m = saf1; }
  if (saf2 < m) { i = 2; }
  if (i == 0) return v3(0, 0, d.z >= 0 ? 1.0 : -1.0);
  if (r <= 0) return v3(1, 0, 0);
  V3 n = v3(q.x / r, q.y / r, 0);
  if (n.x * d.x + n.y * d.y < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- sphere
// rmin rmax theta1 theta2 (deg), full phi (a phi range: sphw_* below).
// Constraint functions (<= 0 inside):
//   g_out = r^2 - rmax^2 ; g_in = rmin^2 - r^2
//   g_t1 = z - c1 r   (theta >= theta1)  when theta1 > 0
//   g_t2 = c2 r - z   (theta <= theta2)  when theta2 < 180
struct SphereP {
  double rmin, rmax, c1, c2;
  bool t1, t2;
};
static SphereP sphere_p(const double* p) {
  SphereP s;
  s.rmin = p[0];
  s.rmax = p[1];
  s.t1 = p[2] > 0;
  s.t2 = p[3] < 180;
  s.c1 = p[4];  // cos(theta1), cos(theta2) as flattened
  s.c2 = p[5];
  return s;
}
static bool sphere_contains(const double* p, V3 q) {
  SphereP s = sphere_p(p);
  double r2 = dot(q, q);
  if (r2 < s.rmin * s.rmin || r2 > s.rmax * s.rmax) return false;
  if (r2 < 1e-20) return true;
  double r = std::sqrt(r2);
  if (s.t1 && q.z > s.c1 * r) return false;
  if (s.t2 && q.z < s.c2 * r) return false;
  return true;
}
// crossings of the cone z = c r (nappe sign(z) == sign(c)); plane z=0 if c==0
static int cone_roots(V3 q, V3 d, double c, double* out) {
  int n = 0;
  if (std::fabs(c) < 1e-15) {
    if (d.z != 0) out[n++] = -q.z / d.z;
    return n;
  }
  double c2 = c * c;
  double A = d.z * d.z - c2 * dot(d, d);
  double B = 2 * (q.z * d.z - c2 * dot(q, d));
  double C = q.z * q.z - c2 * dot(q, q);
  double r[2];
  int m = 0;
  if (std::fabs(A) < 1e-15) {
    if (B != 0) r[m++] = -C / B;
  } else {
    double disc = B * B - 4 * A * C;
    if (disc < 0) return 0;
    double sq = std::sqrt(disc);
    double qq = -0.5 * (B + (B >= 0 ? sq : -sq));
    r[m++] = qq / A;
    if (qq != 0) r[m++] = C / qq;
  }
  for (int i = 0; i < m; i++) {
    double z = q.z + r[i] * d.z;
    if ((c > 0 && z >= 0) || (c < 0 && z <= 0)) out[n++] = r[i];
  }
  if (n == 2 && out[0] > out[1]) std::swap(out[0], out[1]);
  return n;
}
// constraint values and their rate along d at point x
static void sphere_g(const SphereP& s, V3 x, V3 d, double* g, double* dg, bool* act) {
  double r = std::sqrt(dot(x, x));
  double rd = r > 0 ? dot(x, d) / r : 0;
  g[0] = r - s.rmax;           dg[0] = rd;  act[0] = true;
  g[1] = s.rmin - r;           dg[1] = -rd; act[1] = s.rmin > 0;
  g[2] = x.z - s.c1 * r;       dg[2] = d.z - s.c1 * rd; act[2] = s.t1;
  g[3] = s.c2 * r - x.z;       dg[3] = s.c2 * rd - d.z; act[3] = s.t2;
}
static int sphere_candidates(const SphereP& s, V3 q, V3 d, double* t, int* which) {
  int n = 0;
  double b = dot(q, d), c = dot(q, q) - s.rmax * s.rmax, disc = b * b - c;
  if (disc > 0) {
    double sq = std::sqrt(disc);
    t[n] = -b - sq; which[n++] = 0;
    t[n] = -b + sq; which[n++] = 0;
  }
  if (s.rmin > 0) {
    c = dot(q, q) - s.rmin * s.rmin;
    disc = b * b - c;
    if (disc > 0) {
      double sq = std::sqrt(disc);
      t[n] = -b - sq; which[n++] = 1;
      t[n] = -b + sq; which[n++] = 1;
    }
  }
  double r[2];
  if (s.t1) {
    int m = cone_roots(q, d, s.c1, r);
    for (int i = 0; i < m; i++) { t[n] = r[i]; which[n++] = 2; }
  }
  if (s.t2) {
    int m = cone_roots(q, d, s.c2, r);
    for (int i = 0; i < m; i++) { t[n] = r[i]; which[n++] = 3; }
  }
  return n;
}
static double sphere_dist_out(const double* p, V3 q, V3 d) {
  SphereP s = sphere_p(p);
  double t[8];
  int w[8];
  int n = sphere_candidates(s, q, d, t, w);
  double best = kBig;
  for (int i = 0; i < n; i++) {
    if (t[i] < 0 || t[i] >= best) continue;
    V3 x = axpy(q, t[i], d);
    double g[4], dg[4];
    bool act[4];
    sphere_g(s, x, d, g, dg, act);
    if (dg[w[i]] >= 0) continue;  // must be entering through this surface
    bool ok = true;
    for (int k = 0; k < 4; k++) {
      if (k == w[i] || !act[k]) continue;
      if (g[k] > 1e-9 * (1 + s.rmax)) { ok = false; break; }
    }
    if (ok) best = t[i];
  }
  return best;
}
static double sphere_dist_in(const double* p, V3 q, V3 d) {
  SphereP s = sphere_p(p);
  double t[8];
  int w[8];
  int n = sphere_candidates(s, q, d, t, w);
  double best = kBig;
  for (int i = 0; i < n; i++) {
    if (t[i] < 0 || t[i] >= best) continue;
    V3 x = axpy(q, t[i], d);
    double g[4], dg[4];
    bool act[4];
    sphere_g(s, x, d, g, dg, act);
    if (dg[w[i]] <= 0) continue;  // must be leaving through this surface
    best = t[i];
  }
  return best >= kBig ? 0.0 : best;
}
static V3 sphere_normal(const double* p, V3 q, V3 d) {
  SphereP s = sphere_p(p);
  double r = std::sqrt(dot(q, q));
  double dist[4] = {std::fabs(r - s.rmax), s.rmin > 0 ? std::fabs(r - s.rmin) : kBig, kBig, kBig};
  double rho = std::sqrt(q.x * q.x + q.y * q.y);
  double st1 = std::sqrt(std::max(0.0, 1 - s.c1 * s.c1)), st2 = std::sqrt(std::max(0.0, 1 - s.c2 * s.c2));
  // distance to the cone line in the meridional plane: |rho*c - z*sin|
  if (s.t1) dist[2] = std::fabs(rho * s.c1 - q.z * st1);
  if (s.t2) dist[3] = std::fabs(rho * s.c2 - q.z * st2);
  int k = 0;
  for (int i = 1; i < 4; i++)
    if (dist[i] < dist[k]) k = i;
  V3 n;
  if (k <= 1) {
    n = r > 0 ? mul(q, 1 / r) : v3(0, 0, 1);
  } else {
    double c = k == 2 ? s.c1 : s.c2, st = k == 2 ? st1 : st2;
    double cp = rho > 0 ? q.x / rho : 1, sp = rho > 0 ? q.y / rho : 0;
    // gradient of (z*sin - rho*cos) direction
    n = v3(-c * cp, -c * sp, st);
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- paraboloid
// rlo rhi dz a b : z = a r^2 + b, solid = {|z|<=dz, a(z-b) >= a^2 r^2}
static bool para_contains(const double* p, V3 q) {
  if (std::fabs(q.z) > p[2]) return false;
  double a = p[3], b = p[4];
  double aa = a * (q.z - b);
  if (aa < 0) return false;
  return aa >= a * a * (q.x * q.x + q.y * q.y);
}
// roots of a (x^2+y^2) + b - z = 0 along the ray
static int para_roots(const double* p, V3 q, V3 d, double* r) {
  double a = p[3], b = p[4];
  double A = a * (d.x * d.x + d.y * d.y);
  double B = 2 * a * (q.x * d.x + q.y * d.y) - d.z;
  double C = a * (q.x * q.x + q.y * q.y) + b - q.z;
  if (std::fabs(A) < 1e-30) {
    if (B == 0) return 0;
    r[0] = -C / B;
    return 1;
  }
  double disc = B * B - 4 * A * C;
  if (disc < 0) return 0;
  double sq = std::sqrt(disc);
  double qq = -0.5 * (B + (B >= 0 ? sq : -sq));
  int n = 0;
  r[n++] = qq / A;
  if (qq != 0) r[n++] = C / qq;
  if (n == 2 && r[0] > r[1]) std::swap(r[0], r[1]);
  return n;
}
static double para_dist_out(const double* p, V3 q, V3 d) {
  double rlo = p[0], rhi = p[1], dz = p[2], a = p[3];
  double best = kBig;
  if (q.z <= -dz && d.z > 0) {
    double s = -(dz + q.z) / d.z;
    double x = q.x + s * d.x, y = q.y + s * d.y;
    if (x * x + y * y <= rlo * rlo) best = s;
  } else if (q.z >= dz && d.z < 0) {
    double s = (dz - q.z) / d.z;
    double x = q.x + s * d.x, y = q.y + s * d.y;
    if (x * x + y * y <= rhi * rhi) best = s;
  }
  double r[2];
  int n = para_roots(p, q, d, r);
  for (int i = 0; i < n; i++) {
    if (r[i] < 0 || r[i] >= best) continue;
    V3 x = axpy(q, r[i], d);
    if (std::fabs(x.z) > dz) continue;
    // entering: d/ds [a(z-b) - a^2 r^2] > 0
    double rate = a * d.z - 2 * a * a * (x.x * d.x + x.y * d.y);
    if (rate <= 0) continue;
    best = r[i];
    break;
  }
  return best;
}
static double para_dist_in(const double* p, V3 q, V3 d) {
  double dz = p[2], a = p[3];
  double best = kBig;
  if (d.z > 0) best = (dz - q.z) / d.z;
  else if (d.z < 0) best = (-dz - q.z) / d.z;
  if (best < 0) best = 0;
  double r[2];
  int n = para_roots(p, q, d, r);
  for (int i = 0; i < n; i++) {
    if (r[i] < 0 || r[i] >= best) continue;
    V3 x = axpy(q, r[i], d);
    double rate = a * d.z - 2 * a * a * (x.x * d.x + x.y * d.y);
    if (rate >= 0) continue;  // leaving through the parabolic wall
    best = r[i];
    break;
  }
  return best;
}
static V3 para_normal(const double* p, V3 q, V3 d) {
  double dz = p[2], a = p[3];
  double r2 = q.x * q.x + q.y * q.y;
  // distance to the parabolic surface along z, scaled by the slope
  double zs = a * r2 + p[4];
  double grad_len = std::sqrt(1 + 4 * a * a * r2);
  double dpar = std::fabs(q.z - zs) / grad_len;
  double dtop = std::fabs(dz - q.z), dbot = std::fabs(q.z + dz);
  V3 n;
  if (dpar <= dtop && dpar <= dbot) {
    n = v3(2 * a * q.x, 2 * a * q.y, -1);
    n = mul(n, 1 / grad_len);
  } else {
    n = v3(0, 0, 1);
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- pgon
// phi1 dphi nedges nz {z rmin rmax}; rmin == 0 and dphi == 360 supported.
// Each z segment is a convex frustum of a regular prism: half-spaces
//   n_k . (x, y) <= rmax_i + slope_i (z - z_i),  z_i <= z <= z_{i+1}
struct PgonSeg {
  double z0, z1, r0, slope;
};
static int pgon_nedges(const double* p) { return int(p[2]); }
static int pgon_nz(const double* p) { return int(p[3]); }
// outward normal of side k (centre of its phi sector), stored after the sections
static void pgon_face(const double* p, int k, double& cx, double& sy) {
  const double* f = p + 4 + 3 * pgon_nz(p) + 2 * k;
  cx = f[0];
  sy = f[1];
}
static bool pgon_seg(const double* p, int i, PgonSeg& s) {
  const double* z = p + 4;
  s.z0 = z[3 * i];
  s.z1 = z[3 * (i + 1)];
  if (s.z1 - s.z0 <= 0) return false;
  s.r0 = z[3 * i + 2];
  s.slope = (z[3 * (i + 1) + 2] - s.r0) / (s.z1 - s.z0);
  return true;
}
// clip the ray against segment i: returns false if it misses
static bool pgon_clip(const double* p, const PgonSeg& s, V3 q, V3 d, double& t0, double& t1) {
  t0 = -kBig;
  t1 = kBig;
  // z slab
  if (d.z != 0) {
    double a = (s.z0 - q.z) / d.z, b = (s.z1 - q.z) / d.z;
    if (a > b) std::swap(a, b);
    t0 = std::max(t0, a);
    t1 = std::min(t1, b);
  } else if (q.z < s.z0 || q.z > s.z1) {
    return false;
  }
  int ne = pgon_nedges(p);
  for (int k = 0; k < ne; k++) {
    double cx, sy;
    pgon_face(p, k, cx, sy);
    // f(x) = cx x + sy y - slope z - (r0 - slope z0) <= 0
    double f0 = cx * q.x + sy * q.y - s.slope * (q.z - s.z0) - s.r0;
    double fd = cx * d.x + sy * d.y - s.slope * d.z;
    if (fd == 0) {
      if (f0 > 0) return false;
      continue;
    }
    double t = -f0 / fd;
    if (fd > 0) t1 = std::min(t1, t);
    else t0 = std::max(t0, t);
  }
  return t0 <= t1;
}
static bool pgon_contains(const double* p, V3 q) {
  int nz = pgon_nz(p);
  const double* z = p + 4;
  if (q.z < z[0] || q.z > z[3 * (nz - 1)]) return false;
  double div = p[1] / pgon_nedges(p);
  double phi = std::atan2(q.y, q.x) * 180 / kPi;
  while (phi < p[0]) phi += 360.0;
  double ddp = phi - p[0];
  if (ddp > p[1]) return false;
  int ip = std::min(int(ddp / div), pgon_nedges(p) - 1);
  double cx, sy;
  pgon_face(p, ip, cx, sy);
  double r = q.x * cx + q.y * sy;
  for (int i = 0; i + 1 < nz; i++) {
    double za = z[3 * i], zb = z[3 * (i + 1)];
    if (q.z < za || q.z > zb) continue;
    double rmax;
    if (zb - za < 1e-8) {
      rmax = std::max(z[3 * i + 2], z[3 * (i + 1) + 2]);
    } else {
      double f = (q.z - za) / (zb - za);
      rmax = z[3 * i + 2] + f * (z[3 * (i + 1) + 2] - z[3 * i + 2]);
    }
    if (r <= rmax) return true;
  }
  return false;
}
static double pgon_dist_out(const double* p, V3 q, V3 d) {
  int nz = pgon_nz(p);
  double best = kBig;
  for (int i = 0; i + 1 < nz; i++) {
    PgonSeg s;
    if (!pgon_seg(p, i, s)) continue;
    double t0, t1;
    if (!pgon_clip(p, s, q, d, t0, t1)) continue;
    if (t1 <= 0 || t0 >= t1) continue;
    double t = t0 < 0 ? 0 : t0;
    if (t < best) best = t;
  }
  return best;
}
static double pgon_dist_in(const double* p, V3 q, V3 d) {
  int nz = pgon_nz(p);
  double cur = 0;
  for (int guard = 0; guard < 2 * nz + 2; guard++) {
    bool moved = false;
    for (int i = 0; i + 1 < nz; i++) {
      PgonSeg s;
      if (!pgon_seg(p, i, s)) continue;
      double t0, t1;
      if (!pgon_clip(p, s, q, d, t0, t1)) continue;
      if (t0 <= cur + 1e-9 && t1 > cur + 1e-9) {
        cur = t1;
        moved = true;
        break;
      }
    }
    if (!moved) break;
  }
  return cur;
}
static V3 pgon_normal(const double* p, V3 q, V3 d) {
  int nz = pgon_nz(p);
  const double* z = p + 4;
  double best = kBig;
  V3 n = v3(0, 0, 1);
  for (int i = 0; i < nz; i++) {
    double dzp = std::fabs(q.z - z[3 * i]);
    if (dzp < best) {
      best = dzp;
      n = v3(0, 0, 1);
    }
  }
  double div = p[1] / pgon_nedges(p);
  double phi = std::atan2(q.y, q.x) * 180 / kPi;
  while (phi < p[0]) phi += 360.0;
  int ip = std::min(int((phi - p[0]) / div), pgon_nedges(p) - 1);
  for (int i = 0; i + 1 < nz; i++) {
    PgonSeg s;
    if (!pgon_seg(p, i, s)) continue;
    if (q.z < s.z0 - 1e-9 || q.z > s.z1 + 1e-9) continue;
    double cx, sy;
    pgon_face(p, ip, cx, sy);
    double len = std::sqrt(1 + s.slope * s.slope);
    double f = (cx * q.x + sy * q.y - s.slope * (q.z - s.z0) - s.r0) / len;
    if (std::fabs(f) < best) {
      best = std::fabs(f);
      n = v3(cx / len, sy / len, -s.slope / len);
    }
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}


// ---------------------------------------------------------------- TGeoPcon, general TGeoPgon
// TGeoPcon (any rmin, any phi range) and TGeoPgon with rmin > 0 or a phi
// range. Containment follows the shapes' definitions, as ROOT's
// TGeoPcon::Contains / TGeoPgon::Contains state them: the z section, phi by
// atan2 within [phi1, phi1 + dphi], and the radius -- for a Pgon the
// projection on the side of the point's phi sector -- between rmin and rmax
// interpolated linearly in z. Distances are found by brute force over the
// surfaces: every crossing of the ray with a section plane, with the inner
// and outer cone (or side planes) of every segment, or with a phi
// half-plane is a candidate, and the answer is the first candidate at which
// containment changes, tested 1e-9 cm before and after it.
// params: Pcon phi1 dphi nz {z rmin rmax}*nz {cos sin}*2;
//         Pgon phi1 dphi nedges nz {z rmin rmax}*nz {cos sin}*nedges {cos sin}*2
struct GPoly {
  const double* z;
  int nz, ne;
  bool pgon;
  double phi1, dphi;
};
static GPoly gpoly_p(const double* p, bool pgon) {
  GPoly g;
  g.pgon = pgon;
  g.phi1 = p[0];
  g.dphi = p[1];
  g.ne = pgon ? int(p[2]) : 0;
  g.nz = int(p[pgon ? 3 : 2]);
  g.z = p + (pgon ? 4 : 3);
  return g;
}
// full phi and rmin = 0 throughout: the convex-frustum form above
static bool pgon_simple(const double* p) {
  if (p[1] < 360.0) return false;
  for (int k = 0; k < int(p[3]); k++)
    if (p[4 + 3 * k + 1] != 0.0) return false;
  return true;
}
// phi of q in [phi1, phi1 + 360) (degrees)
static double gpoly_phi(const GPoly& g, V3 q) {
  double phi = std::atan2(q.y, q.x) * 180 / kPi;
  while (phi < g.phi1) phi += 360.0;
  while (phi >= g.phi1 + 360.0) phi -= 360.0;
  return phi;
}
static double gpoly_radius(const GPoly& g, V3 q, double phi) {
  if (!g.pgon) return std::sqrt(q.x * q.x + q.y * q.y);
  const double div = std::min(g.dphi, 360.0) / g.ne;
  const int ip = std::min(int((phi - g.phi1) / div), g.ne - 1);
  const double ph = (g.phi1 + div * (ip + 0.5)) * kPi / 180;
  return q.x * std::cos(ph) + q.y * std::sin(ph);
}
static bool gpoly_contains(const GPoly& g, V3 q) {
  if (q.z < g.z[0] || q.z > g.z[3 * (g.nz - 1)]) return false;
  const double phi = gpoly_phi(g, q);
  if (g.dphi < 360.0 && phi - g.phi1 > g.dphi) return false;
  const double r = gpoly_radius(g, q, phi);
  for (int i = 0; i + 1 < g.nz; i++) {
    const double za = g.z[3 * i], zb = g.z[3 * (i + 1)];
    if (q.z < za || q.z > zb) continue;
    double rmin, rmax;
    if (zb - za < 1e-8) {
      rmin = std::min(g.z[3 * i + 1], g.z[3 * (i + 1) + 1]);
      rmax = std::max(g.z[3 * i + 2], g.z[3 * (i + 1) + 2]);
    } else {
      const double f = (q.z - za) / (zb - za);
      rmin = g.z[3 * i + 1] + f * (g.z[3 * (i + 1) + 1] - g.z[3 * i + 1]);
      rmax = g.z[3 * i + 2] + f * (g.z[3 * (i + 1) + 2] - g.z[3 * i + 2]);
    }
    if (r >= rmin && r <= rmax) return true;
  }
  return false;
}
// every parameter t at which the ray meets one of the solid's surfaces
static void gpoly_candidates(const GPoly& g, V3 q, V3 d, std::vector<double>& ts) {
  ts.clear();
  if (d.z != 0)
    for (int k = 0; k < g.nz; k++) ts.push_back((g.z[3 * k] - q.z) / d.z);
  for (int i = 0; i + 1 < g.nz; i++) {
    const double z0 = g.z[3 * i], z1 = g.z[3 * (i + 1)];
    if (!(z1 > z0)) continue;
    for (int col = 1; col <= 2; col++) {
      const double r0 = g.z[3 * i + col], r1 = g.z[3 * (i + 1) + col];
      if (col == 1 && r0 == 0 && r1 == 0) continue;
      const double s = (r1 - r0) / (z1 - z0);
      if (g.pgon) {
        const double div = std::min(g.dphi, 360.0) / g.ne;
        for (int k = 0; k < g.ne; k++) {
          const double ph = (g.phi1 + div * (k + 0.5)) * kPi / 180;
          const double cx = std::cos(ph), sy = std::sin(ph);
          const double f0 = cx * q.x + sy * q.y - s * (q.z - z0) - r0, fd = cx * d.x + sy * d.y - s * d.z;
          if (fd != 0) ts.push_back(-f0 / fd);
        }
      } else {
        // x^2 + y^2 = (r0 + s (z - z0))^2 along the ray: both nappes
        const double R0 = r0 + s * (q.z - z0);
        const double a = d.x * d.x + d.y * d.y - s * s * d.z * d.z;
        const double b = 2 * (q.x * d.x + q.y * d.y - s * d.z * R0);
        const double c = q.x * q.x + q.y * q.y - R0 * R0;
        if (std::fabs(a) < 1e-14) {
          if (b != 0) ts.push_back(-c / b);
        } else {
          const double disc = b * b - 4 * a * c;
          if (disc >= 0) {
            ts.push_back((-b - std::sqrt(disc)) / (2 * a));
            ts.push_back((-b + std::sqrt(disc)) / (2 * a));
          }
        }
      }
    }
  }
  if (g.dphi < 360.0) {
    for (double phd : {g.phi1, g.phi1 + g.dphi}) {
      const double ph = phd * kPi / 180;
      const double nx = -std::sin(ph), ny = std::cos(ph);
      const double fd = nx * d.x + ny * d.y;
      if (fd != 0) ts.push_back(-(nx * q.x + ny * q.y) / fd);
    }
  }
  std::sort(ts.begin(), ts.end());
}
// from outside (entering) or inside (leaving): the first candidate at which
// containment changes that way
static double gpoly_dist(const GPoly& g, V3 q, V3 d, bool leaving) {
  std::vector<double> ts;
  gpoly_candidates(g, q, d, ts);
  const double e = 1e-9;
  for (double t : ts) {
    if (t < -e) continue;
    const bool before = gpoly_contains(g, axpy(q, t - e, d)), after = gpoly_contains(g, axpy(q, t + e, d));
    if (leaving ? (before && !after) : (!before && after)) return std::max(t, 0.0);
  }
  return leaving ? 0.0 : kBig;
}
// the nearest surface: a section plane, the inner or outer surface of the
// point's segment, or a phi half-plane
static V3 gpoly_normal(const GPoly& g, V3 q, V3 d) {
  double best = kBig;
  V3 n = v3(0, 0, 1);
  for (int k = 0; k < g.nz; k++) best = std::min(best, std::fabs(q.z - g.z[3 * k]));
  const double phi = gpoly_phi(g, q);
  double cx, sy;
  if (g.pgon) {
    const double div = std::min(g.dphi, 360.0) / g.ne;
    const int ip = std::min(int((std::min(phi - g.phi1, std::min(g.dphi, 360.0))) / div), g.ne - 1);
    const double ph = (g.phi1 + div * (ip + 0.5)) * kPi / 180;
    cx = std::cos(ph);
    sy = std::sin(ph);
  } else {
    cx = std::cos(phi * kPi / 180);
    sy = std::sin(phi * kPi / 180);
  }
  const double r = q.x * cx + q.y * sy;
  for (int i = 0; i + 1 < g.nz; i++) {
    const double z0 = g.z[3 * i], z1 = g.z[3 * (i + 1)];
    if (!(z1 > z0) || q.z < z0 - 1e-9 || q.z > z1 + 1e-9) continue;
    for (int col = 1; col <= 2; col++) {
      const double r0 = g.z[3 * i + col], r1 = g.z[3 * (i + 1) + col];
      if (col == 1 && r0 == 0 && r1 == 0) continue;
      const double s = (r1 - r0) / (z1 - z0), len = std::sqrt(1 + s * s);
      const double dist = std::fabs(r - r0 - s * (q.z - z0)) / len;
      if (dist < best) {
        best = dist;
        n = v3(cx / len, sy / len, -s / len);
      }
    }
  }
  if (g.dphi < 360.0) {
    for (double phd : {g.phi1, g.phi1 + g.dphi}) {
      const double ph = phd * kPi / 180;
      if (std::cos(ph) * q.x + std::sin(ph) * q.y < 0) continue;
      const double nx = -std::sin(ph), ny = std::cos(ph);
      const double dist = std::fabs(nx * q.x + ny * q.y);
      if (dist < best) {
        best = dist;
        n = v3(nx, ny, 0);
      }
    }
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- aspheric disk
// AGeoAsphericDisk (src/AGeoAsphericDisk.cxx). params: z1 c1 z2 c2 rmax rmin
// kappa1 kappa2 n1 n2 K1[n1] K2[n2]; surface s in {0, 1} = ROOT's 1, 2.
struct Asph {
  double z[2], c[2], kap[2], rmax, rmin;
  int n[2];
  const double* K[2];
};
static Asph asph_p(const double* p) {
  Asph a;
  a.z[0] = p[0]; a.c[0] = p[1]; a.z[1] = p[2]; a.c[1] = p[3];
  a.rmax = p[4]; a.rmin = p[5];
  a.kap[0] = p[6]; a.kap[1] = p[7];
  a.n[0] = int(p[8]); a.n[1] = int(p[9]);
  a.K[0] = p + 10; a.K[1] = p + 10 + a.n[0];
  return a;
}
// CalcF1/CalcF2 (:126-153); false where ROOT throws (1 - p < 0)
static bool asph_sag(const Asph& a, int s, double r, double& f) {
  double p = r * r * a.c[s] * a.c[s] * a.kap[s];
  if (1 - p < 0) return false;
  f = a.z[s] + r * r * a.c[s] / (1 + std::sqrt(1 - p));
  for (int i = 0; i < a.n[s]; i++) f += a.K[s][i] * std::pow(r, 2 * (i + 1));
  return true;
}
// CalcdF1dr/CalcdF2dr (:96-123); false where ROOT throws (1 - p <= 0)
static bool asph_dsag(const Asph& a, int s, double r, double& df) {
  double p = r * r * a.c[s] * a.c[s] * a.kap[s];
  if (1 - p <= 0) return false;
  df = r * a.c[s] / std::sqrt(1 - p);
  for (int i = 0; i < a.n[s]; i++) df += 2 * (i + 1) * a.K[s][i] * std::pow(r, 2 * (i + 1) - 1);
  return true;
}
// DistToAsphere (:415-526): Matsui's spherical seed, then Newton steps
// along the ray until |e| < 1e-10 (at most 100)
static double asph_dist_surface(const Asph& a, int s, V3 P, V3 D) {
  count_op(OP_ASPH_SEED);
  const double c = a.c[s], kap = a.kap[s], d = a.z[s];
  double H2 = P.x * P.x + P.y * P.y;
  double p = -((P.z - d) * D.z + P.x * D.x + P.y * D.y);
  double M = p * D.z + P.z - d;
  double M2 = (P.z - d) * (P.z - d) + H2 - p * p;
  double check = 1 - (M2 * c - 2 * M) * c / D.z / D.z;
  if (check < 0) return kBig;
  double q = p + (M2 * c - 2 * M) / (D.z * (1 + std::sqrt(check)));
  double np[3] = {P.x + q * D.x, P.y + q * D.y, P.z + q * D.z - d};
  for (int i = 0;; i++) {
    if (i > 100) return kBig;
    count_op(OP_ASPH_NEWTON);
    H2 = np[0] * np[0] + np[1] * np[1];
    check = 1 - kap * H2 * c * c;
    if (check < 0) return kBig;
    double l = std::sqrt(check);
    double x = 0;
    if (c != 0) x += (1 - l) / c / kap;
    for (int j = 0; j < a.n[s]; j++) x += a.K[s][j] * std::pow(H2, j + 1);
    double v = 0;
    for (int j = 0; j < a.n[s]; j++) v += 2 * (j + 1) * a.K[s][j] * std::pow(H2, j);
    v = c * kap + l * v;
    double m = -np[0] * v, nn = -np[1] * v;
    double norm = std::sqrt(l * l + m * m + nn * nn);
    l /= norm; m /= norm; nn /= norm;
    check = D.z * l + D.x * m + D.y * nn;
    if (check == 0) return kBig;
    double e = l * (x - np[2]) / check;
    np[0] += e * D.x; np[1] += e * D.y; np[2] += e * D.z;
    if (std::fabs(e) < 1e-10) break;
  }
  np[2] += d;
  check = D.x * (np[0] - P.x) + D.y * (np[1] - P.y) + D.z * (np[2] - P.z);
  if (check < 0) return kBig;
  double rz = std::pow(np[0] * np[0] + np[1] * np[1], 0.5);
  if (rz < a.rmin || rz > a.rmax) return kBig;
  return std::sqrt(std::pow(np[0] - P.x, 2) + std::pow(np[1] - P.y, 2) + std::pow(np[2] - P.z, 2));
}
// DistToInner / DistToOuter (:529-684): the cylinder r = R between the two sags
static double asph_dist_wall(const Asph& a, bool outer, V3 P, V3 D) {
  count_op(OP_ASPH_WALL);
  if (!outer && !(a.rmin > 0)) return kBig;
  const double R = outer ? a.rmax : a.rmin;
  double rsq = P.x * P.x + P.y * P.y, nsq = D.x * D.x + D.y * D.y;
  if (std::sqrt(nsq) < kTol) return kBig;
  double rdotn = P.x * D.x + P.y * D.y;
  double t1i = 1. / nsq, b = t1i * rdotn, cc = t1i * (rsq - R * R);
  double delta = b * b - cc;
  if (delta > 0) delta = std::sqrt(delta);
  else return kBig;
  double t1 = -b + delta, t2 = -b - delta;
  if (t1 < 0 && t2 < 0) return kBig;
  double zmin, zmax;
  if (!asph_sag(a, 0, R, zmin) || !asph_sag(a, 1, R, zmax)) return kBig;
  if (t2 > 0) {
    double z1 = t1 * D.z + P.z, z2 = t2 * D.z + P.z;
    if (z1 < zmin || zmax < z1) t1 = kBig;
    if (z2 < zmin || zmax < z2) t2 = kBig;
    return t1 < t2 ? t1 : t2;
  } else if (t2 == 0) {
    if (t1 > 0) {
      if (zmin <= P.z && P.z <= zmax) return 0;
      double z1 = t1 * D.z + P.z;
      if (zmin <= z1 && z1 <= zmax) return t1;
    } else if (t1 == 0) {
      if (zmin <= P.z && P.z <= zmax) return 0;
    }
  } else {
    if (t1 > 0) {
      double z1 = t1 * D.z + P.z;
      if (zmin <= z1 && z1 <= zmax) return t1;
    } else if (t1 == 0) {
      if (zmin <= P.z && P.z <= zmax) return 0;
    }
  }
  return kBig;
}
// DistFromInside and DistFromOutside are the same minimum (:363-412)
static double asph_dist(const double* p, V3 q, V3 d) {
  Asph a = asph_p(p);
  double v[4] = {asph_dist_surface(a, 0, q, d), asph_dist_surface(a, 1, q, d), asph_dist_wall(a, false, q, d),
                 asph_dist_wall(a, true, q, d)};
  int k = 0;
  for (int i = 1; i < 4; i++)
    if (v[i] < v[k]) k = i;
  return v[k];
}
static bool asph_contains(const double* p, V3 q) {
  Asph a = asph_p(p);
  double r = std::sqrt(q.x * q.x + q.y * q.y);
  if (r > a.rmax || r < a.rmin) return false;
  double f1, f2;
  if (!asph_sag(a, 0, r, f1) || !asph_sag(a, 1, r, f2)) return false;
  return !(q.z < f1 || f2 < q.z);
}
// ComputeNormal (:246-327): closest of the four surfaces, rotated by phi
// ComputeNormal's choice of face (:246-328): the one with the smallest
// safety, 0 inner wall, 1 outer wall, 2 + s surface s; *df receives the
// surfaces' slopes dF/dr (kBig where undefined).
static int asph_face(const Asph& a, V3 q, double* df) {
  double r = std::sqrt(q.x * q.x + q.y * q.y);
  double saf[4];
  saf[0] = a.rmin > 0 ? std::fabs(r - a.rmin) : kBig;
  saf[1] = std::fabs(r - a.rmax);
  df[0] = df[1] = kBig;
  for (int s = 0; s < 2; s++) {
    double f;
    if (!asph_sag(a, s, r, f)) { saf[2 + s] = kBig; continue; }
    if (!asph_dsag(a, s, r, df[s])) { saf[2 + s] = kBig; continue; }
    saf[2 + s] = std::fabs(f - q.z) / std::sqrt(1 + df[s] * df[s]);
  }
  int k = 0;
  for (int i = 1; i < 4; i++)
    if (saf[i] < saf[k]) k = i;
  return k;
}
static V3 asph_normal(const double* p, V3 q, V3 d) {
  Asph a = asph_p(p);
  double phi = std::atan2(q.y, q.x);
  double df[2];
  const int k = asph_face(a, q, df);
  double n0 = 0, n2 = 0;
  if (k <= 1) {
    n0 = 1;
  } else {
    double g = df[k - 2];
    if (g == 0) n2 = 1;
    else { n0 = g / std::sqrt(1 + g * g); n2 = -1 / std::sqrt(1 + g * g); }
  }
  V3 n = v3(n0 * std::cos(phi), n0 * std::sin(phi), n2);
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}


// ---------------------------------------------------------------- Winston cone (polygonal)
// AGeoWinstonConePoly (src/AGeoWinstonConePoly.cxx) over AGeoWinstonCone2D's
// parabola math (src/AGeoWinstonCone2D.cxx). params: r1 r2 polyN theta dz f ...
struct Wins {
  double r1, r2, theta, dz, f;
  int n;
};
static Wins wins_p(const double* p) {
  Wins w;
  w.r1 = p[0]; w.r2 = p[1]; w.n = int(p[2]); w.theta = p[3]; w.dz = p[4]; w.f = p[5];
  return w;
}
// AGeoWinstonCone2D::CalcR (:82-98); false where it throws (|z| > dz + 1e-10)
static bool wins_r(const Wins& w, double z, double& r) {
  if (std::fabs(z) > w.dz + 1e-10) return false;
  double sint = std::sin(w.theta), cost = std::cos(w.theta);
  double t = z + w.dz;
  double a0 = t * t * sint * sint - 4. * w.f * (t * cost + w.f);
  double a1 = 2. * t * sint * cost + 4. * w.f * sint;
  double a2 = cost * cost;
  r = (-a1 + std::sqrt(a1 * a1 - 4. * a0 * a2)) / (2 * a2) - w.r2;
  return true;
}
static bool wins_drdz(const Wins& w, double z, double& g) {
  if (std::fabs(z) > w.dz + 1e-10) return false;
  double sint = std::sin(w.theta), cost = std::cos(w.theta);
  double t = z + w.dz;
  double a0 = t * t * sint * sint - 4. * w.f * (t * cost + w.f);
  double a1 = 2. * t * sint * cost + 4. * w.f * sint;
  double a2 = cost * cost;
  double da0dt = 2 * t * sint * sint - 4 * w.f * cost, da1dt = 2 * sint * cost;
  g = (-da1dt + (a1 * da1dt - 2 * da0dt * a2) / std::sqrt(a1 * a1 - 4 * a0 * a2)) / (2 * a2);
  return true;
}
// AGeoWinstonConePoly::InsidePolygon (:293-309)
static bool wins_inside_polygon(const Wins& w, double x, double y, double r) {
  double th = std::atan2(y, x);
  while (th > kPi / w.n) th -= 2 * kPi / w.n;
  while (th < -kPi / w.n) th += 2 * kPi / w.n;
  return !(std::sqrt(x * x + y * y) * std::cos(th) > r);
}
// AGeoWinstonCone2D::DistToParabola (:325-428): the side at azimuth phi,
// in the tilted parabola frame; `open` is the accepted azimuth width
static double wins_parabola(const Wins& w, V3 P, V3 D, double phi, double open) {
  count_op(OP_WINS_SIDE);
  double x = std::cos(phi) * P.x + std::sin(phi) * P.y;
  double y = -std::sin(phi) * P.x + std::cos(phi) * P.y;
  double z = P.z;
  double px = std::cos(phi) * D.x + std::sin(phi) * D.y;
  double py = -std::sin(phi) * D.x + std::cos(phi) * D.y;
  double pz = D.z;
  if (px == 0 && pz == 0) return kBig;
  double cost = std::cos(w.theta), sint = std::sin(w.theta);
  double X = cost * (x + w.r2) + (z + w.dz) * sint;
  double Z = -sint * (x + w.r2) + (z + w.dz) * cost + w.f;
  double alpha = std::atan2(pz, px);
  double tanA = std::tan(alpha - w.theta);
  double dist[2] = {kBig, kBig};
  double tmp = tanA * tanA - (X * tanA - Z) / w.f;
  if (tmp < 0) return kBig;
  double Xp, Xm;
  if (w.dz * 2 / std::fabs(tanA) < kTol) {
    Xp = X;
    Xm = X;
  } else {
    Xp = 2 * w.f * (tanA + std::sqrt(tmp));
    Xm = 2 * w.f * (tanA - std::sqrt(tmp));
  }
  double Xs[2] = {Xp, Xm};
  for (int k = 0; k < 2; k++) {
    double Xc = Xs[k];
    double Zc = Xc * Xc / 4. / w.f;
    double xc = cost * Xc - sint * (Zc - w.f) - w.r2;
    double zc = sint * Xc + cost * (Zc - w.f) - w.dz;
    double yc;
    if (std::fabs(px) <= std::fabs(pz) && std::fabs(py) <= std::fabs(pz)) {
      yc = y + (zc - z) * py / pz;
    } else if (std::fabs(py) <= std::fabs(px) && std::fabs(pz) <= std::fabs(px)) {
      yc = y + (xc - x) * py / px;
    } else {
      yc = y + (std::fabs(px) < 1e-5 ? (zc - z) * py / pz : (xc - x) * py / px);
    }
    double dx = xc - x, dy = yc - y, dz = zc - z;
    if (xc < w.r2 || w.r1 < xc || zc < -w.dz || w.dz < zc || dx * px + dz * pz < 0) continue;
    if (std::fabs(std::atan2(yc, xc)) <= open / 2.) dist[k] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  return std::min(dist[0], dist[1]);
}
static bool wins_contains(const double* p, V3 q) {
  Wins w = wins_p(p);
  if (std::fabs(q.z) > w.dz) return false;
  double r;
  if (!wins_r(w, q.z, r)) return false;
  return wins_inside_polygon(w, q.x, q.y, r);
}
static double wins_dist_in(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  double best = kBig;
  if (d.z < 0) best = (-q.z - w.dz) / d.z;
  else if (d.z > 0) best = (w.dz - q.z) / d.z;
  for (int i = 0; i < w.n; i++) best = std::min(best, wins_parabola(w, q, d, i * 2 * kPi / w.n, kPi));
  return best;
}
static double wins_dist_out(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  if (q.z <= -w.dz) {
    if (d.z <= 0) return kBig;
    double s = -(w.dz + q.z) / d.z;
    if (wins_inside_polygon(w, q.x + s * d.x, q.y + s * d.y, w.r2)) return s;
  } else if (q.z >= w.dz) {
    if (d.z >= 0) return kBig;
    double s = (w.dz - q.z) / d.z;
    if (wins_inside_polygon(w, q.x + s * d.x, q.y + s * d.y, w.r1)) return s;
  }
  double best = kBig;
  for (int i = 0; i < w.n; i++) best = std::min(best, wins_parabola(w, q, d, i * 2 * kPi / w.n, 2 * kPi / w.n));
  return best;
}
// AGeoWinstonConePoly::ComputeNormal (:67-122)
static V3 wins_normal(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  double saf0 = std::fabs(std::fabs(w.dz) - std::fabs(q.z)), saf1;
  double phi = std::atan2(q.y, q.x);
  while (phi > kPi / w.n) phi -= 2 * kPi / w.n;
  while (phi < -kPi / w.n) phi += 2 * kPi / w.n;
  double r;
  saf1 = wins_r(w, q.z, r) ? std::fabs(r - std::sqrt(q.x * q.x + q.y * q.y) * std::cos(phi)) : kBig;
  V3 n;
  if (saf0 <= saf1) {
    n = v3(0, 0, 1);
  } else {
    phi = std::atan2(q.y, q.x);
    if (phi < -kPi / w.n) phi += 2 * kPi;
    int k = int(std::floor((phi + kPi / w.n) / (2 * kPi / w.n)));
    double g = 0;
    wins_drdz(w, q.z, g);
    n = v3(std::cos(k * 2 * kPi / w.n), std::sin(k * 2 * kPi / w.n), -g);
    double m = std::sqrt(dot(n, n));
    n = mul(n, 1 / m);
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- sphere with a phi range
// TGeoSphere(rmin, rmax, theta1, theta2, phi1, phi2) with dphi < 360: the
// shell segment restated from its definition (r in [rmin, rmax], theta by
// acos(z / r) in [theta1, theta2], phi by atan2 in [phi1, phi1 + dphi]);
// distances by brute force over every crossing of the ray with the two
// spheres, the two theta cones (either nappe) and the two phi planes,
// taking the first at which containment changes, tested 1e-9 cm either side.
// params: as the sphere, then phi1 dphi {cos sin}(phi1) {cos sin}(phi1 + dphi)
static bool sphw_wedge_full(const double* p) { return p[9] >= 360.0; }
static bool sphw_contains(const double* p, V3 q) {
  const double r = std::sqrt(dot(q, q));
  if (r < p[0] || r > p[1]) return false;
  if (r > 0) {
    const double th = std::acos(std::max(-1.0, std::min(1.0, q.z / r))) * 180 / kPi;
    if (th < p[2] || th > p[3]) return false;
  }
  double phi = std::atan2(q.y, q.x) * 180 / kPi;
  while (phi < p[8]) phi += 360.0;
  while (phi >= p[8] + 360.0) phi -= 360.0;
  return phi - p[8] <= p[9];
}
static double sphw_dist(const double* p, V3 q, V3 d, bool leaving) {
  std::vector<double> ts;
  const double b = dot(q, d), qq = dot(q, q);
  for (double R : {p[0], p[1]}) {
    if (!(R > 0)) continue;
    const double disc = b * b - (qq - R * R);
    if (disc >= 0) {
      ts.push_back(-b - std::sqrt(disc));
      ts.push_back(-b + std::sqrt(disc));
    }
  }
  for (int k = 0; k < 2; k++) {
    const double thd = p[2 + k];
    if ((k == 0 && thd <= 0) || (k == 1 && thd >= 180)) continue;
    const double c = std::cos(thd * kPi / 180), c2 = c * c;
    // z^2 = c^2 r^2 along the ray, both nappes
    const double A = d.z * d.z - c2 * dot(d, d), B = 2 * (q.z * d.z - c2 * dot(q, d)), C = q.z * q.z - c2 * qq;
    if (std::fabs(A) < 1e-14) {
      if (B != 0) ts.push_back(-C / B);
    } else {
      const double disc = B * B - 4 * A * C;
      if (disc >= 0) {
        ts.push_back((-B - std::sqrt(disc)) / (2 * A));
        ts.push_back((-B + std::sqrt(disc)) / (2 * A));
      }
    }
  }
  for (double phd : {p[8], p[8] + p[9]}) {
    const double ph = phd * kPi / 180, nx = -std::sin(ph), ny = std::cos(ph);
    const double fd = nx * d.x + ny * d.y;
    if (fd != 0) ts.push_back(-(nx * q.x + ny * q.y) / fd);
  }
  std::sort(ts.begin(), ts.end());
  const double e = 1e-9;
  for (double t : ts) {
    if (t < -e) continue;
    const bool before = sphw_contains(p, axpy(q, t - e, d)), after = sphw_contains(p, axpy(q, t + e, d));
    if (leaving ? (before && !after) : (!before && after)) return std::max(t, 0.0);
  }
  return leaving ? 0.0 : kBig;
}
// the nearest of the spheres, theta cones and phi half-planes
static V3 sphw_normal(const double* p, V3 q, V3 d) {
  const double r = std::sqrt(dot(q, q)), rho = std::sqrt(q.x * q.x + q.y * q.y);
  double best = kBig;
  V3 n = v3(0, 0, 1);
  for (double R : {p[0], p[1]}) {
    if (R > 0 && std::fabs(r - R) < best) {
      best = std::fabs(r - R);
      n = r > 0 ? mul(q, 1 / r) : v3(0, 0, 1);
    }
  }
  const double th = r > 0 ? std::acos(std::max(-1.0, std::min(1.0, q.z / r))) : 0;
  for (int k = 0; k < 2; k++) {
    const double thd = p[2 + k];
    if ((k == 0 && thd <= 0) || (k == 1 && thd >= 180)) continue;
    const double tk = thd * kPi / 180;
    const double dist = r * std::fabs(std::sin(th - tk));
    if (dist < best) {
      best = dist;
      const double cp = rho > 0 ? q.x / rho : 1, sp = rho > 0 ? q.y / rho : 0;
      n = v3(std::cos(tk) * cp, std::cos(tk) * sp, -std::sin(tk));  // the cone's normal, up to sign
    }
  }
  for (double phd : {p[8], p[8] + p[9]}) {
    const double ph = phd * kPi / 180;
    if (std::cos(ph) * q.x + std::sin(ph) * q.y < 0) continue;
    const double dist = std::fabs(-std::sin(ph) * q.x + std::cos(ph) * q.y);
    if (dist < best) {
      best = dist;
      n = v3(-std::sin(ph), std::cos(ph), 0);
    }
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- Winston cone (2D trough)
// AGeoWinstonCone2D (src/AGeoWinstonCone2D.cxx): Contains (:175-190),
// DistFromInside (:203-238), DistFromOutside (:241-322), ComputeNormal
// (:120-172), with DistToParabola at phi = 0 and pi over an open angle of pi.
// params: r1 r2 2 theta dz f cos sin 0 {1 0} {-1 0} dy
static bool wins2d_contains(const double* p, V3 q) {
  Wins w = wins_p(p);
  if (std::fabs(q.y) > p[13] || std::fabs(q.z) > w.dz) return false;
  double r;
  if (!wins_r(w, q.z, r)) return false;
  return std::fabs(q.x) <= r;
}
static double wins2d_dist_in(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  const double dy = p[13];
  double s[4] = {kBig, kBig, kBig, kBig};
  if (d.z < 0) s[0] = -(q.z + w.dz) / d.z;
  else if (d.z > 0) s[0] = (w.dz - q.z) / d.z;
  if (d.y < 0) s[1] = -(q.y + dy) / d.y;
  else if (d.y > 0) s[1] = (dy - q.y) / d.y;
  s[2] = wins_parabola(w, q, d, 0., kPi);
  s[3] = wins_parabola(w, q, d, kPi, kPi);
  return *std::min_element(s, s + 4);
}
static double wins2d_dist_out(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  const double dy = p[13];
  if (q.z <= -w.dz) {
    if (d.z <= 0) return kBig;
    double sn = -(w.dz + q.z) / d.z;
    if (std::fabs(q.x + sn * d.x) <= w.r2 && std::fabs(q.y + sn * d.y) <= dy) return sn;
  } else if (q.z >= w.dz) {
    if (d.z >= 0) return kBig;
    double sn = (w.dz - q.z) / d.z;
    if (std::fabs(q.x + sn * d.x) <= w.r1 && std::fabs(q.y + sn * d.y) <= dy) return sn;
  }
  if (q.y <= -dy) {
    if (d.y <= 0) return kBig;
    double sn = -(dy + q.y) / d.y, zn = q.z + sn * d.z, r;
    if (std::fabs(zn) <= w.dz && wins_r(w, zn, r) && std::fabs(q.x + sn * d.x) <= r) return sn;
  } else if (q.y >= dy) {
    if (d.y >= 0) return kBig;
    double sn = (dy - q.y) / d.y, zn = q.z + sn * d.z, r;
    if (std::fabs(zn) <= w.dz && wins_r(w, zn, r) && std::fabs(q.x + sn * d.x) <= r) return sn;
  }
  double best = kBig;
  for (double phi : {0., kPi}) {
    double sn = wins_parabola(w, q, d, phi, kPi);
    if (std::fabs(q.y + sn * d.y) <= dy) best = std::min(best, sn);
  }
  return best;
}
static V3 wins2d_normal(const double* p, V3 q, V3 d) {
  Wins w = wins_p(p);
  double saf[3] = {std::fabs(std::fabs(p[13]) - std::fabs(q.y)), std::fabs(std::fabs(w.dz) - std::fabs(q.z)), kBig};
  double r;
  if (wins_r(w, q.z, r)) saf[2] = std::fabs(r - std::fabs(q.x));
  const int i = int(std::min_element(saf, saf + 3) - saf);
  V3 n;
  if (i == 0) {
    n = v3(0, 1, 0);
  } else if (i == 1) {
    n = v3(0, 0, 1);
  } else {
    double g = 0;
    wins_drdz(w, q.z, g);
    n = q.x > 0 ? v3(1, 0, -g) : v3(1, 0, g);
    n = mul(n, 1 / std::sqrt(dot(n, n)));
  }
  if (dot(n, d) < 0) n = mul(n, -1);
  return n;
}

// ---------------------------------------------------------------- dispatch
static bool contains(const Scene& sc, int shape, V3 q);
static Hit dist_out(const Scene& sc, int shape, V3 q, V3 d);
static Hit dist_in(const Scene& sc, int shape, V3 q, V3 d);

struct Comp {
  int op, left, right;
  const rb_matrix* lm;
  const rb_matrix* rm;
};
static Comp comp_of(const Scene& sc, int shape) {
  const double* p = sc.P(shape);
  Comp c;
  c.op = int(p[0]);
  c.left = int(p[1]);
  c.right = int(p[2]);
  c.lm = &sc.mat(p + 3);
  c.rm = &sc.mat(p + 15);
  return c;
}

static const int kInOp[] = {-1, OP_IN_BOX, OP_IN_TUBE, OP_IN_SPHERE, OP_IN_PARA, OP_IN_PGON, OP_IN_PGON, -1,
                            OP_IN_ASPH, OP_IN_WINS, OP_IN_WINS};
static const int kDistOp[] = {-1, OP_DIST_BOX, OP_DIST_TUBE, OP_DIST_SPHERE, OP_DIST_PARA, OP_DIST_PGON,
                              OP_DIST_PGON, -1, -1, OP_DIST_WINS_CAP, OP_DIST_WINS_CAP};
static bool contains(const Scene& sc, int shape, V3 q) {
  const double* p = sc.P(shape);
  if (sc.type(shape) >= 0 && sc.type(shape) <= RB_SHAPE_WINSTON2D && kInOp[sc.type(shape)] >= 0)
    count_op(kInOp[sc.type(shape)]);
  switch (sc.type(shape)) {
    case RB_SHAPE_BOX: return box_contains(p, q);
    case RB_SHAPE_TUBE: return tube_contains(p, q);
    case RB_SHAPE_SPHERE: return sphw_wedge_full(p) ? sphere_contains(p, q) : sphw_contains(p, q);
    case RB_SHAPE_PARABOLOID: return para_contains(p, q);
    case RB_SHAPE_PGON: return pgon_simple(p) ? pgon_contains(p, q) : gpoly_contains(gpoly_p(p, true), q);
    case RB_SHAPE_PCON: return gpoly_contains(gpoly_p(p, false), q);
    case RB_SHAPE_WINSTON2D: return wins2d_contains(p, q);
    case RB_SHAPE_ASPHERIC: return asph_contains(p, q);
    case RB_SHAPE_WINSTON: return wins_contains(p, q);
    case RB_SHAPE_COMPOSITE: {
      Comp c = comp_of(sc, shape);
      bool a = contains(sc, c.left, m2l(*c.lm, q));
      if (c.op == RB_OP_UNION) return a || contains(sc, c.right, m2l(*c.rm, q));
      if (c.op == RB_OP_SUBTRACTION) return a && !contains(sc, c.right, m2l(*c.rm, q));
      return a && contains(sc, c.right, m2l(*c.rm, q));
    }
  }
  return false;
}

static Hit child_out(const Scene& sc, int child, const rb_matrix& m, V3 q, V3 d, int side) {
  Hit h = dist_out(sc, child, m2l(m, q), m2lv(m, d));
  h.leaf = leaf_push(h.leaf, side);
  return h;
}
static Hit child_in(const Scene& sc, int child, const rb_matrix& m, V3 q, V3 d, int side) {
  Hit h = dist_in(sc, child, m2l(m, q), m2lv(m, d));
  h.leaf = leaf_push(h.leaf, side);
  return h;
}

// TGeoUnion/TGeoSubtraction/TGeoIntersection::DistFromOutside restated as
// boundary-to-boundary iteration (inside tests taken g_bool_push past each point)
static Hit comp_dist_out(const Scene& sc, int shape, V3 q, V3 d) {
  Comp c = comp_of(sc, shape);
  if (c.op == RB_OP_UNION) {
    Hit a = child_out(sc, c.left, *c.lm, q, d, 0);
    Hit b = child_out(sc, c.right, *c.rm, q, d, 1);
    return b.s < a.s ? b : a;
  }
  double s = 0;
  uint32_t leaf = 0;
  for (int it = 0; it < 64; it++) {
    V3 x = axpy(q, s, d);
    V3 xp = axpy(x, g_bool_push, d);
    bool inA = contains(sc, c.left, m2l(*c.lm, xp));
    bool inB = contains(sc, c.right, m2l(*c.rm, xp));
    double step;
    if (c.op == RB_OP_SUBTRACTION) {
      if (inA && !inB) return {s, leaf};
      if (inB) {
        Hit h = child_in(sc, c.right, *c.rm, x, d, 1);
        step = h.s;
        leaf = h.leaf;
      } else {
        Hit h = child_out(sc, c.left, *c.lm, x, d, 0);
        if (h.s >= kBig) return {kBig, 0};
        step = h.s;
        leaf = h.leaf;
      }
    } else {  // intersection
      if (inA && inB) return {s, leaf};
      Hit ha = {0, 0}, hb = {0, 0};
      if (!inA) {
        ha = child_out(sc, c.left, *c.lm, x, d, 0);
        if (ha.s >= kBig) return {kBig, 0};
      }
      if (!inB) {
        hb = child_out(sc, c.right, *c.rm, x, d, 1);
        if (hb.s >= kBig) return {kBig, 0};
      }
      if (!inA && (inB || ha.s >= hb.s)) {
        step = ha.s;
        leaf = ha.leaf;
      } else {
        step = hb.s;
        leaf = hb.leaf;
      }
    }
    if (step < g_bool_push) step = g_bool_push;
    s += step;
  }
  return {kBig, 0};
}
static Hit comp_dist_in(const Scene& sc, int shape, V3 q, V3 d) {
  Comp c = comp_of(sc, shape);
  if (c.op == RB_OP_INTERSECTION) {
    Hit a = child_in(sc, c.left, *c.lm, q, d, 0);
    Hit b = child_in(sc, c.right, *c.rm, q, d, 1);
    return b.s < a.s ? b : a;
  }
  if (c.op == RB_OP_SUBTRACTION) {
    Hit a = child_in(sc, c.left, *c.lm, q, d, 0);
    Hit b = child_out(sc, c.right, *c.rm, q, d, 1);
    return b.s < a.s ? b : a;
  }
  // union: leave both children
  double s = 0;
  uint32_t leaf = 0;
  for (int it = 0; it < 64; it++) {
    V3 x = axpy(q, s, d);
    V3 xp = axpy(x, g_bool_push, d);
    bool inA = contains(sc, c.left, m2l(*c.lm, xp));
    bool inB = contains(sc, c.right, m2l(*c.rm, xp));
    if (!inA && !inB) return {s, leaf};
    Hit ha = {0, 0}, hb = {0, 0};
    if (inA) ha = child_in(sc, c.left, *c.lm, x, d, 0);
    if (inB) hb = child_in(sc, c.right, *c.rm, x, d, 1);
    Hit h = (inA && (!inB || ha.s >= hb.s)) ? ha : hb;
    double step = h.s < g_bool_push ? g_bool_push : h.s;
    leaf = h.leaf;
    s += step;
  }
  return {s, leaf};
}

static Hit dist_out(const Scene& sc, int shape, V3 q, V3 d) {
  const double* p = sc.P(shape);
  if (sc.type(shape) >= 0 && sc.type(shape) <= RB_SHAPE_WINSTON2D && kDistOp[sc.type(shape)] >= 0)
    count_op(kDistOp[sc.type(shape)]);
  switch (sc.type(shape)) {
    case RB_SHAPE_BOX: return {box_dist_out(p, q, d), 0};
    case RB_SHAPE_TUBE: return {tube_dist_out(p, q, d), 0};
    case RB_SHAPE_SPHERE: return {sphw_wedge_full(p) ? sphere_dist_out(p, q, d) : sphw_dist(p, q, d, false), 0};
    case RB_SHAPE_PARABOLOID: return {para_dist_out(p, q, d), 0};
    case RB_SHAPE_PGON:
      return {pgon_simple(p) ? pgon_dist_out(p, q, d) : gpoly_dist(gpoly_p(p, true), q, d, false), 0};
    case RB_SHAPE_PCON: return {gpoly_dist(gpoly_p(p, false), q, d, false), 0};
    case RB_SHAPE_WINSTON2D: return {wins2d_dist_out(p, q, d), 0};
    case RB_SHAPE_ASPHERIC: return {asph_dist(p, q, d), 0};
    case RB_SHAPE_WINSTON: return {wins_dist_out(p, q, d), 0};
    case RB_SHAPE_COMPOSITE: return comp_dist_out(sc, shape, q, d);
  }
  return {kBig, 0};
}
static Hit dist_in(const Scene& sc, int shape, V3 q, V3 d) {
  const double* p = sc.P(shape);
  if (sc.type(shape) >= 0 && sc.type(shape) <= RB_SHAPE_WINSTON2D && kDistOp[sc.type(shape)] >= 0)
    count_op(kDistOp[sc.type(shape)]);
  switch (sc.type(shape)) {
    case RB_SHAPE_BOX: return {box_dist_in(p, q, d), 0};
    case RB_SHAPE_TUBE: return {tube_dist_in(p, q, d), 0};
    case RB_SHAPE_SPHERE: return {sphw_wedge_full(p) ? sphere_dist_in(p, q, d) : sphw_dist(p, q, d, true), 0};
    case RB_SHAPE_PARABOLOID: return {para_dist_in(p, q, d), 0};
    case RB_SHAPE_PGON:
      return {pgon_simple(p) ? pgon_dist_in(p, q, d) : gpoly_dist(gpoly_p(p, true), q, d, true), 0};
    case RB_SHAPE_PCON: return {gpoly_dist(gpoly_p(p, false), q, d, true), 0};
    case RB_SHAPE_WINSTON2D: return {wins2d_dist_in(p, q, d), 0};
    case RB_SHAPE_ASPHERIC: return {asph_dist(p, q, d), 0};
    case RB_SHAPE_WINSTON: return {wins_dist_in(p, q, d), 0};
    case RB_SHAPE_COMPOSITE: return comp_dist_in(sc, shape, q, d);
  }
  return {0, 0};
}
// normal of the surface `leaf` of `shape` at q, oriented along d
static V3 shape_normal(const Scene& sc, int shape, uint32_t leaf, V3 q, V3 d) {
  count_op(OP_NORMAL);
  int depth = 0;
  int maxdepth = int(leaf >> 24);
  // walk composite path; matrices accumulate as nested master->local maps
  const rb_matrix* stack[32];
  int ns = 0;
  while (sc.type(shape) == RB_SHAPE_COMPOSITE) {
    Comp c = comp_of(sc, shape);
    int side = depth < maxdepth ? int((leaf >> depth) & 1u) : 0;
    const rb_matrix* m = side ? c.rm : c.lm;
    q = m2l(*m, q);
    d = m2lv(*m, d);
    stack[ns++] = m;
    shape = side ? c.right : c.left;
    depth++;
  }
  const double* p = sc.P(shape);
  V3 n;
  switch (sc.type(shape)) {
    case RB_SHAPE_BOX: n = box_normal(p, q, d); break;
    case RB_SHAPE_TUBE: n = tube_normal(p, q, d); break;
    case RB_SHAPE_SPHERE: n = sphw_wedge_full(p) ? sphere_normal(p, q, d) : sphw_normal(p, q, d); break;
    case RB_SHAPE_PARABOLOID: n = para_normal(p, q, d); break;
    case RB_SHAPE_PGON: n = pgon_simple(p) ? pgon_normal(p, q, d) : gpoly_normal(gpoly_p(p, true), q, d); break;
    case RB_SHAPE_PCON: n = gpoly_normal(gpoly_p(p, false), q, d); break;
    case RB_SHAPE_WINSTON2D: n = wins2d_normal(p, q, d); break;
    case RB_SHAPE_ASPHERIC: n = asph_normal(p, q, d); break;
    case RB_SHAPE_WINSTON: n = wins_normal(p, q, d); break;
    default: n = v3(0, 0, 1);
  }
  while (ns > 0) n = l2mv(*stack[--ns], n);
  return n;
}

// ---------------------------------------------------------------- physics
// TGraph::Eval linear interpolation/extrapolation on x-ascending points
static double table_eval(const rb_scene_desc* d, int table, double x) {
  const rb_table_desc& t = d->tables[table];
  const double* xy = d->table_xy + 2 * t.offset;
  int n = t.n;
  if (n == 0) return 0;
  if (n == 1) return xy[1];
  int low = -1, up = -1;
  for (int i = 0; i < n; i++) {
    double xi = xy[2 * i];
    if (xi < x) low = i;
    else if (xi > x) { up = i; break; }
    else return xy[2 * i + 1];
  }
  // outside the range: the insertion-order pairs stored after the points
  const double* e = nullptr;
  if (up == -1) e = xy + 2 * n;            // (up, low) above
  else if (low == -1) e = xy + 2 * n + 4;  // (low, up) below
  double xl, yl, xu, yu;
  if (e) {
    bool above = up == -1;
    xu = e[above ? 0 : 2]; yu = e[above ? 1 : 3];
    xl = e[above ? 2 : 0]; yl = e[above ? 3 : 1];
  } else {
    xl = xy[2 * low]; yl = xy[2 * low + 1]; xu = xy[2 * up]; yu = xy[2 * up + 1];
  }
  if (xl == xu) return yl;
  return yu + (x - xu) * (yl - yu) / (xl - xu);
}
static double mirror_reflectance(const rb_scene_desc* d, const rb_volume_desc& v, double lambda) {
  double r = v.refl_table >= 0 ? table_eval(d, v.refl_table, lambda) : v.refl_const;
  r = r > 1 ? 1 : r;
  r = r < 0 ? 0 : r;
  return r;
}
static double lens_index(const rb_scene_desc* d, const rb_volume_desc& v, double lambda) {
  switch (v.index_model) {
    case RB_INDEX_TABLE: return table_eval(d, v.index_ref, lambda);
    case RB_INDEX_SELLMEIER: {
      const double* b = d->params + v.index_ref;
      double l = lambda / 1e-4;  // AOpticsManager::um()
      double l2 = l * l;
      return std::sqrt(1 + b[0] * l2 / (l2 - b[3]) + b[1] * l2 / (l2 - b[4]) + b[2] * l2 / (l2 - b[5]));
    }
    case RB_INDEX_SCHOTT: {
      const double* a = d->params + v.index_ref;
      double l = lambda / 1e-4;
      return std::sqrt(a[0] + a[1] * std::pow(l, 2.) + a[2] * std::pow(l, -2.) + a[3] * std::pow(l, -4.) +
                       a[4] * std::pow(l, -6.) + a[5] * std::pow(l, -8.));
    }
  }
  return 1.0;
}
static double lens_k(const rb_scene_desc* d, const rb_volume_desc& v, double lambda) {
  return v.k_table >= 0 ? table_eval(d, v.k_table, lambda) : 0.0;
}
static double lens_abs_length(const rb_scene_desc* d, const rb_volume_desc& v, double lambda) {
  if (v.index_model == RB_INDEX_NONE && v.k_table < 0) return INFINITY;
  double k = lens_k(d, v, lambda);
  return k <= 0 ? INFINITY : lambda / (4 * kPi * k);
}

// Philox4x32-10 (Salmon et al., SC'11), key = seed, counter = (id lo, id hi, draw, stream)
static inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
    uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
    uint32_t h0 = uint32_t(p0 >> 32), l0 = uint32_t(p0);
    uint32_t h1 = uint32_t(p1 >> 32), l1 = uint32_t(p1);
    uint32_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = l1; c[2] = n2; c[3] = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
struct Rng {
  uint64_t seed, id;
  uint32_t draw, stream;
  double uniform() {  // (0, 1]
    uint32_t c[4] = {uint32_t(id), uint32_t(id >> 32), draw++, stream};
    philox(c, uint32_t(seed), uint32_t(seed >> 32));
    uint64_t u = ((uint64_t(c[1]) << 32) | c[0]) >> 11;
    return double(u + 1) * (1.0 / 9007199254740992.0);
  }
};

// ---------------------------------------------------------------- navigator
static void build_boxes(Scene& sc) {
  const rb_scene_desc* d = sc.d;
  sc.boxes.resize(d->n_nodes);
  for (int n = 0; n < d->n_nodes; n++) {
    const rb_node_desc& nd = d->nodes[n];
    const rb_shape_desc& sh = d->shapes[d->volumes[nd.volume].shape];
    Box b;
    for (int k = 0; k < 3; k++) { b.lo[k] = 1e300; b.hi[k] = -1e300; }
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
    for (int k = 0; k < 3; k++) {
      double pad = 1e-6 * (1 + std::fabs(b.lo[k]) + std::fabs(b.hi[k]));
      b.lo[k] -= pad;
      b.hi[k] += pad;
    }
    sc.boxes[n] = b;
  }
}
static bool box_has_point(const Box& b, V3 p) {
  return p.x >= b.lo[0] && p.x <= b.hi[0] && p.y >= b.lo[1] && p.y <= b.hi[1] && p.z >= b.lo[2] && p.z <= b.hi[2];
}
// slab test: can the ray meet the box before `limit`?
static bool box_hit(const Box& b, V3 p, V3 d, double limit) {
  count_op(OP_BOX_TEST);
  double t0 = 0, t1 = limit + 1e-6 * (1 + limit);
  double o[3] = {p.x, p.y, p.z}, v[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; k++) {
    if (v[k] == 0) {
      if (o[k] < b.lo[k] || o[k] > b.hi[k]) return false;
      continue;
    }
    double a = (b.lo[k] - o[k]) / v[k], c = (b.hi[k] - o[k]) / v[k];
    if (a > c) std::swap(a, c);
    t0 = std::max(t0, a);
    t1 = std::min(t1, c);
    if (t0 > t1) return false;
  }
  return true;
}

struct Nav {
  const Scene* sc;
  V3 p, d;            // current point and direction (fPoint, fDirection)
  int cur;            // current node, -1 = outside (fIsOutside)
  bool on_boundary;   // fIsOnBoundary
  double step;        // fStep after the last FindNextBoundaryAndStep
  // crossed surface of the last step, for FindNormal
  int hit_node;
  uint32_t hit_leaf;
};

static const rb_node_desc& node(const Scene& sc, int i) { return sc.d->nodes[i]; }
static int node_shape(const Scene& sc, int i) { return sc.d->volumes[node(sc, i).volume].shape; }
static bool node_contains(const Scene& sc, int i, V3 p) {
  return contains(sc, node_shape(sc, i), m2l(node(sc, i).global, p));
}

// TGeoNavigator::SearchNode: from `start` go up until a node contains p
// (skipping the containment test when start == skip), then down into the
// first daughter (in order, excluding `skip`) that contains p.
static int search_node(const Scene& sc, int start, V3 p, int skip, bool forced) {
  int n = start;
  bool inside = forced || (n == skip);
  while (!inside) {
    if (!node_contains(sc, n, p)) {
      if (node(sc, n).parent < 0) return -1;
      skip = n;
      n = node(sc, n).parent;
    } else {
      inside = true;
    }
  }
  for (;;) {
    const rb_node_desc& nd = node(sc, n);
    int found = -1;
    for (int k = 0; k < nd.n_children; k++) {
      int c = sc.d->children[nd.first_child + k];
      if (c == skip) continue;
      if (!box_has_point(sc.boxes[c], p)) continue;
      if (node_contains(sc, c, p)) {
        found = c;
        break;
      }
    }
    if (found < 0) return n;
    n = found;
    skip = -1;
  }
}

// TGeoNavigator::InitTrack / FindNode from the top
static void init_track(Nav& nv, V3 p, V3 d) {
  nv.p = p;
  nv.d = d;
  nv.on_boundary = false;
  const Scene& sc = *nv.sc;
  nv.cur = node_contains(sc, 0, p) ? search_node(sc, 0, p, -1, true) : -1;
}

// TGeoNavigator::CrossBoundaryAndLocate push: 100 (1 + |T| + step) tolerance
static double locate_push(const Scene& sc, int n, double step) {
  const double* t = node(sc, n).global.tr;
  double trmax = 1. + std::fabs(t[0]) + std::fabs(t[1]) + std::fabs(t[2]);
  return g_push_scale * (trmax + step) * kTol;
}

// TGeoNavigator::FindNextBoundaryAndStep(); returns the next node (-1 = null)
static int find_next_boundary_and_step(Nav& nv) {
  count_op(OP_STEP);
  const Scene& sc = *nv.sc;
  double extra = nv.on_boundary ? kTol : 0.0;
  nv.on_boundary = false;
  nv.p = axpy(nv.p, extra, nv.d);
  nv.hit_node = -1;
  nv.hit_leaf = 0;
  if (nv.cur < 0) {
    const rb_node_desc& top = node(sc, 0);
    Hit h = dist_out(sc, node_shape(sc, 0), m2l(top.global, nv.p), m2lv(top.global, nv.d));
    if (h.s < kBig - kTol) {
      nv.step = h.s + extra;
      nv.p = axpy(nv.p, h.s, nv.d);
      nv.on_boundary = true;
      nv.hit_node = 0;
      nv.hit_leaf = h.leaf;
      V3 pp = axpy(nv.p, locate_push(sc, 0, nv.step), nv.d);
      nv.cur = search_node(sc, 0, pp, -1, true);
      return nv.cur;
    }
    nv.step = kBig;
    return -1;
  }
  int cur = nv.cur;
  const rb_node_desc& cn = node(sc, cur);
  Hit hx = dist_in(sc, node_shape(sc, cur), m2l(cn.global, nv.p), m2lv(cn.global, nv.d));
  if (hx.s <= kTol) {
    nv.step = kTol;
    nv.p = axpy(nv.p, nv.step, nv.d);
    nv.on_boundary = true;
    nv.hit_node = cur;
    nv.hit_leaf = hx.leaf;
    V3 pp = axpy(nv.p, locate_push(sc, cur, nv.step), nv.d);
    nv.cur = search_node(sc, cur, pp, cur, false);
    return nv.cur;
  }
  double best = kBig;
  int crossed = -2;
  uint32_t leaf = 0;
  if (hx.s < best - kTol) {
    best = hx.s;
    crossed = -1;
    leaf = hx.leaf;
  }
  for (int k = 0; k < cn.n_children; k++) {
    int c = sc.d->children[cn.first_child + k];
    if (!box_hit(sc.boxes[c], nv.p, nv.d, best)) continue;
    const rb_node_desc& chn = node(sc, c);
    Hit h = dist_out(sc, node_shape(sc, c), m2l(chn.global, nv.p), m2lv(chn.global, nv.d));
    if (h.s < best - kTol) {
      best = h.s;
      crossed = c;
      leaf = h.leaf;
    }
  }
  nv.p = axpy(nv.p, best, nv.d);
  nv.step = best + extra;
  if (crossed == -2) return cur;
  nv.on_boundary = true;
  nv.hit_leaf = leaf;
  if (crossed == -1) {
    nv.hit_node = cur;
    if (cn.parent < 0) {
      nv.cur = -1;
      return -1;
    }
    int mother = cn.parent;
    V3 pp = axpy(nv.p, locate_push(sc, mother, nv.step), nv.d);
    nv.cur = search_node(sc, mother, pp, cur, false);
    return nv.cur;
  }
  nv.hit_node = crossed;
  V3 pp = axpy(nv.p, locate_push(sc, crossed, nv.step), nv.d);
  nv.cur = search_node(sc, crossed, pp, cur, true);
  return nv.cur;
}

// TGeoNavigator::FindNormal: normal of the crossed surface, along d, master frame
static V3 find_normal(const Nav& nv) {
  const Scene& sc = *nv.sc;
  if (nv.hit_node < 0) return v3(0, 0, 1);
  const rb_node_desc& n = node(sc, nv.hit_node);
  V3 lp = m2l(n.global, nv.p), ld = m2lv(n.global, nv.d);
  V3 ln = shape_normal(sc, node_shape(sc, nv.hit_node), nv.hit_leaf, lp, ld);
  return l2mv(n.global, ln);
}

// TGeoNavigator::Step() after SetStep(kEpsilon): moves fStep + 1e-6 and relocates
static void step_back(Nav& nv, V3 back_dir) {
  const Scene& sc = *nv.sc;
  double s = g_step_back;
  nv.p = axpy(nv.p, s, back_dir);
  int start = nv.cur < 0 ? 0 : nv.cur;
  if (nv.cur < 0) {
    nv.cur = node_contains(sc, 0, nv.p) ? search_node(sc, 0, nv.p, -1, true) : -1;
  } else {
    nv.cur = search_node(sc, start, nv.p, -1, false);
  }
  nv.on_boundary = true;
}

// ---------------------------------------------------------------- tracer
struct Ray {
  V3 x;      // last point
  double t;
  V3 d;      // direction (unit)
  double lambda;
  int status, npoints, last_node;
  double* path;     // optional [cap][4]: slot k holds the (k+1)-th point this trace adds
  int32_t* pnode;
  int cap;
  int npoints_in;   // points the ray held when the trace began
  void add_point(V3 p, double tt, int nd) {
    const int k = npoints - npoints_in;
    if (path && k < cap) {
      double* s = path + 4 * k;
      s[0] = p.x; s[1] = p.y; s[2] = p.z; s[3] = tt;
      pnode[k] = nd;
    }
    x = p;
    t = tt;
    last_node = nd;
    npoints++;
  }
  void set_direction(V3 v) {  // ARay::SetDirection normalises (src/ARay.cxx:210-223)
    double mag = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (mag > 0) d = v3(v.x / mag, v.y / mag, v.z / mag);
  }
};

static int kind_of(const Scene& sc, int n) {
  if (n < 0) return RB_KNULL;
  return sc.d->volumes[node(sc, n).volume].kind;
}
static const rb_volume_desc& vol_of(const Scene& sc, int n) { return sc.d->volumes[node(sc, n).volume]; }

struct Opts {
  int limit;
  bool fresnel;
};

// ---- border-surface conditions (include/ABorderSurfaceCondition.h:24-45)
// AOpticalComponent::FindBorderSurfaceCondition (src/AOpticalComponent.cxx:51-65):
// the first condition of the current node's volume whose second component is
// the next node's volume (-1: the null node).
static const rb_border_desc* find_border(const Scene& sc, int cur, int next) {
  if (cur < 0 || sc.d->n_borders <= 0) return nullptr;
  const rb_volume_desc& v = vol_of(sc, cur);
  const int v2 = next < 0 ? -1 : node(sc, next).volume;
  for (int k = 0; k < v.border_count; k++)
    if (sc.d->borders[v.border_first + k].volume2 == v2) return &sc.d->borders[v.border_first + k];
  return nullptr;
}

// Uniforms one stochastic surface sample may consume before the smooth
// normal is kept (the reference's loops are unbounded; the device caps them
// the same way so that every lane ends).
static const uint32_t kMaxSurfaceDraws = 1u << 16;

// TVector3::RotateUz(u) (u a unit vector)
static V3 rotate_uz(V3 v, V3 u) {
  double up = u.x * u.x + u.y * u.y;
  if (up != 0.) {
    up = std::sqrt(up);
    return v3((u.x * u.z * v.x - u.y * v.y + u.x * up * v.z) / up, (u.y * u.z * v.x + u.x * v.y + u.y * up * v.z) / up,
              (u.z * u.z * v.x - v.x + u.z * up * v.z) / up);
  }
  if (u.z < 0.) return v3(-v.x, v.y, -v.z);
  return v;
}

// GetFacetNormal (src/AOpticsManager.cxx:250-301): the smooth normal, or for
// a rough condition the Geant4-style facet normal. TRandom::Gaus is replaced
// by Box-Muller on two Philox uniforms.
static V3 facet_normal(const Nav& nv, const Ray& ray, const rb_border_desc* bc, Rng& rng) {
  V3 normal = find_normal(nv);
  if (!bc || bc->lambertian || bc->sigma == 0.) return normal;
  const double sigma = bc->sigma;
  const double f_max = std::min(1., 4. * sigma);
  const uint32_t stop = rng.draw + kMaxSurfaceDraws;
  while (rng.draw < stop) {
    double alpha = 0.;
    bool ok = false;
    while (rng.draw < stop) {
      const double u1 = rng.uniform(), u2 = rng.uniform();
      alpha = sigma * (std::sqrt(-2. * std::log(u1)) * std::cos(2. * kPi * u2));
      if (!(f_max * rng.uniform() > std::sin(alpha) || alpha >= kPi / 2)) {
        ok = true;
        break;
      }
    }
    if (!ok) break;
    const double phi = 2. * kPi * rng.uniform();
    const double sa = std::sin(alpha), ca = std::cos(alpha), sp = std::sin(phi), cp = std::cos(phi);
    const V3 f = rotate_uz(v3(sa * cp, sa * sp, ca), normal);
    if (dot(ray.d, f) > 0.0) return f;
  }
  return normal;
}

// DoReflection's Lambertian direction (src/AOpticsManager.cxx:209-224)
static V3 lambert_dir(V3 n, Rng& rng) {
  const double y = 0.5 * rng.uniform();
  const double theta = std::asin(std::sqrt(2 * y));
  const double phi = 2. * kPi * rng.uniform();
  const double rad2deg = 180. / kPi, deg2rad = kPi / 180.;
  const double theta_n =
      (n.x == 0. && n.y == 0. && n.z == 0. ? 0. : std::atan2(std::sqrt(n.x * n.x + n.y * n.y), n.z)) * rad2deg;
  const double phi_n = (n.x == 0. && n.y == 0. ? 0. : std::atan2(n.y, n.x)) * rad2deg;
  // TGeoRotation("", phi_n + 90, theta_n + 180, 0): Euler z-x-z, LocalToMaster
  const double sph = std::sin(deg2rad * (phi_n + 90)), cph = std::cos(deg2rad * (phi_n + 90));
  const double sth = std::sin(deg2rad * (theta_n + 180)), cth = std::cos(deg2rad * (theta_n + 180));
  const V3 v = v3(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
  return v3(cph * v.x - cth * sph * v.y + sth * sph * v.z, sph * v.x + cth * cph * v.y - sth * cph * v.z,
            sth * v.y + cth * v.z);
}

// src/AOpticsManager.cxx:170-247
// AMultilayer::CoherentTMMMixed once PreCalculateCoherentTMM has filled its
// TH2Ds (include/AMultilayer.h:110-114, 243-262): TH2::Interpolate(lambda,
// angle), restated from ROOT's documented behaviour -- bilinear between the
// four nearest bin centres, a neighbour bin beyond the axis replaced by the
// edge bin (so the outer half bins are flat along that axis), 0 outside the
// axes. Written out step by step, independently of the device's form.
static double th2_interpolate(const double* vals, const rb_grid2d_desc& g, double x, double y) {
  const double wx = (g.xhi - g.xlo) / g.nx, wy = (g.yhi - g.ylo) / g.ny;
  int bx, by;  // TAxis::FindFixBin
  if (x < g.xlo) bx = 0; else if (!(x < g.xhi)) bx = g.nx + 1; else bx = 1 + int(g.nx * (x - g.xlo) / (g.xhi - g.xlo));
  if (y < g.ylo) by = 0; else if (!(y < g.yhi)) by = g.ny + 1; else by = 1 + int(g.ny * (y - g.ylo) / (g.yhi - g.ylo));
  if (bx > g.nx + 1) bx = g.nx + 1;
  if (by > g.ny + 1) by = g.ny + 1;
  if (bx < 1 || bx > g.nx || by < 1 || by > g.ny) return 0.0;  // "Cannot interpolate outside histogram domain"
  // the quadrant of the bin holding (x, y) picks the lower-left of the four centres
  const bool right = (g.xlo + bx * wx) - x <= wx / 2, upper = (g.ylo + by * wy) - y <= wy / 2;
  const int ax = right ? bx : bx - 1, ay = upper ? by : by - 1;
  const double x1 = g.xlo + (ax - 0.5) * wx, x2 = g.xlo + (ax + 0.5) * wx;
  const double y1 = g.ylo + (ay - 0.5) * wy, y2 = g.ylo + (ay + 0.5) * wy;
  auto clampx = [&](int b) { return std::min(std::max(b, 1), g.nx); };
  auto clampy = [&](int b) { return std::min(std::max(b, 1), g.ny); };
  auto at = [&](int i, int j) { return vals[int64_t(clampy(j) - 1) * g.nx + clampx(i) - 1]; };
  const double q11 = at(ax, ay), q21 = at(ax + 1, ay), q12 = at(ax, ay + 1), q22 = at(ax + 1, ay + 1);
  const double d = (x2 - x1) * (y2 - y1);
  return q11 / d * (x2 - x) * (y2 - y) + q21 / d * (x - x1) * (y2 - y) + q12 / d * (x2 - x) * (y - y1) +
         q22 / d * (x - x1) * (y - y1);
}
static void multilayer_rt(const rb_scene_desc* d, const rb_border_desc* bc, double lambda, double cos1, double& R,
                          double& T) {
  const rb_grid2d_desc& g = d->grids[bc->multilayer];
  const double angle = std::acos(cos1);
  R = th2_interpolate(d->params + g.r_offset, g, lambda, angle);
  T = th2_interpolate(d->params + g.t_offset, g, lambda, angle);
}

static void do_reflection(double n1, Ray& ray, Nav& nv, int current, int next, const V3* normal, Rng& rng) {
  count_op(OP_REFLECT);
  const Scene& sc = *nv.sc;
  double step = nv.step;
  const rb_border_desc* bc = find_border(sc, current, next);
  V3 n = normal ? *normal : facet_normal(nv, ray, bc, rng);
  V3 d1 = ray.d;
  double cos1 = dot(d1, n);
  bool absorbed = false;
  if (kind_of(sc, next) == RB_KMIRROR) {
    double ref, trans;
    const rb_volume_desc& mv = vol_of(sc, next);
    if (bc && bc->multilayer >= 0) {  // src/AOpticsManager.cxx:194-198: the coating, unclamped
      multilayer_rt(sc.d, bc, ray.lambda, cos1, ref, trans);
    } else if (mv.refl_grid >= 0) {  // AMirror::GetReflectance (src/AMirror.cxx:39-60): the TH2 first, clamped
      const rb_grid2d_desc& g = sc.d->grids[mv.refl_grid];
      ref = th2_interpolate(sc.d->params + g.r_offset, g, ray.lambda, std::acos(cos1));
      ref = ref > 1 ? 1 : ref;
      ref = ref < 0 ? 0 : ref;
    } else {
      ref = mirror_reflectance(sc.d, mv, ray.lambda);
    }
    if (ref < rng.uniform()) {
      absorbed = true;
      ray.status = RB_ABSORB;
    }
  }
  V3 d2 = v3(d1.x - 2 * n.x * cos1, d1.y - 2 * n.y * cos1, d1.z - 2 * n.z * cos1);
  if (bc && bc->lambertian) d2 = lambert_dir(n, rng);  // drawn even for an absorbed ray
  if (!absorbed) ray.set_direction(d2);
  double t = ray.t + step / (kC / n1);
  step_back(nv, mul(d1, -1));
  nv.d = ray.d;
  ray.add_point(nv.p, t, next);
}

// src/AOpticsManager.cxx:52-167
static void do_fresnel(double n1, double n2, double k2, Ray& ray, Nav& nv, int current, int next, const Opts& o,
                       Rng& rng) {
  count_op(OP_REFRACT);
  double step = nv.step;
  const rb_border_desc* bc = find_border(*nv.sc, current, next);
  V3 n = facet_normal(nv, ray, bc, rng);
  V3 d1 = ray.d;
  double cos1 = dot(d1, n);
  double sin1 = std::sqrt(1 - cos1 * cos1);
  double sin2 = n1 * sin1 / n2;
  double cos2 = std::sqrt(1 - sin2 * sin2);
  bool absorbed = false;
  if (bc && bc->multilayer >= 0) {  // :79-96: reflect below R, transmit below R + T, else absorb
    double R, T;
    multilayer_rt(nv.sc->d, bc, ray.lambda, cos1, R, T);
    double rnd = rng.uniform();
    if (rnd < R) {
      do_reflection(n1, ray, nv, current, next, &n, rng);
      return;
    } else if (!(rnd < R + T)) {
      absorbed = true;
    }
    goto transmission_process;
  }
  if (sin2 > 1.) {
    do_reflection(n1, ray, nv, current, next, &n, rng);
    return;
  }
  if (o.fresnel) {
    double Rs, Rp;
    auto sq = [](double v) { return v * v; };
    if (k2 <= 0.) {
      double eta1S = n1 * cos1, eta2S = n2 * cos2;
      double eta1P = n1 / cos1, eta2P = n2 / cos2;
      Rs = sq((eta1S - eta2S) / (eta1S + eta2S));
      Rp = sq((eta1P - eta2P) / (eta1P + eta2P));
    } else {
      double eta1S = n1 * cos1, eta1P = n1 / cos1;
      double x1S = eta1S, x1P = eta1P;
      double u = sq(n2) - sq(k2) - sq(n1 * sin1);
      double v = 2 * n2 * k2;
      double tmp = std::sqrt(sq(u) + sq(v));
      double cosxi2 = std::sqrt(1 + u / tmp) / std::sqrt(2.);
      double sinxi2 = std::sqrt(1 - u / tmp) / std::sqrt(2.);
      double x2S = std::sqrt(tmp) * cosxi2, y2S = std::sqrt(tmp) * sinxi2;
      tmp = sq(x2S) + sq(y2S);
      double x2P = (2 * n2 * k2 * y2S + (sq(n2) - sq(k2)) * x2S) / tmp;
      double y2P = (2 * n2 * k2 * x2S - (sq(n2) - sq(k2)) * y2S) / tmp;
      Rs = (sq(x1S - x2S) + sq(y2S)) / (sq(x1S + x2S) + sq(y2S));
      Rp = (sq(x1P - x2P) + sq(y2P)) / (sq(x1P + x2P) + sq(y2P));
    }
    double R = (Rs + Rp) / 2.;
    if (rng.uniform() < R) {
      do_reflection(n1, ray, nv, current, next, &n, rng);
      return;
    }
  }
transmission_process:
  V3 d2;
  if (sin1 != 0) {
    d2 = v3((d1.x - cos1 * n.x) * sin2 / sin1 + n.x * cos2, (d1.y - cos1 * n.y) * sin2 / sin1 + n.y * cos2,
            (d1.z - cos1 * n.z) * sin2 / sin1 + n.z * cos2);
  } else {
    d2 = d1;
  }
  double t = ray.t + step / (kC / n1);
  ray.add_point(nv.p, t, next);
  if (absorbed) {
    ray.status = RB_ABSORB;
  } else {
    ray.set_direction(d2);  // a NaN d2 (transmission beyond the critical angle) is ignored, as by ARay
    nv.d = ray.d;
  }
}

// src/AOpticsManager.cxx:346-519 for one ray
static void trace_ray(const Scene& sc, const Opts& o, Ray& ray, Rng& rng) {
  if (ray.status != RB_RUN) return;
  Nav nv;
  nv.sc = &sc;
  nv.hit_node = -1;
  nv.hit_leaf = 0;
  nv.step = 0;
  init_track(nv, ray.x, ray.d);
  const double lambda = ray.lambda;
  while (ray.status == RB_RUN) {
    V3 x1 = ray.x;
    double t1 = ray.t;
    int current = nv.cur;
    int next = find_next_boundary_and_step(nv);
    double step = nv.step;
    int tc = kind_of(sc, current), tn = kind_of(sc, next);

    if (tc == RB_KLENS) {
      double abs = lens_abs_length(sc.d, vol_of(sc, current), lambda);
      if (abs > 0 && abs != INFINITY) {
        double abs_step = -abs * std::log(rng.uniform());
        if (abs_step < step) {
          double n1 = lens_index(sc.d, vol_of(sc, current), lambda);
          V3 x2 = axpy(x1, abs_step, ray.d);
          ray.add_point(x2, t1 + abs_step / (kC / n1), next);
          ray.status = RB_ABSORB;
          continue;
        }
      }
    }

    bool c_free = tc == RB_KNULL || tc == RB_KOPT || tc == RB_KOTHER;
    if ((c_free || tc == RB_KLENS) && tn == RB_KMIRROR) {
      double n1 = tc == RB_KLENS ? lens_index(sc.d, vol_of(sc, current), lambda) : 1.;
      do_reflection(n1, ray, nv, current, next, nullptr, rng);
    } else if (c_free && tn == RB_KLENS) {
      const rb_volume_desc& v = vol_of(sc, next);
      do_fresnel(1, lens_index(sc.d, v, lambda), lens_k(sc.d, v, lambda), ray, nv, current, next, o, rng);
    } else if ((c_free || tc == RB_KLENS) && (tn == RB_KOBS || tn == RB_KFOCUS)) {
      double n1 = tc == RB_KLENS ? lens_index(sc.d, vol_of(sc, current), lambda) : 1.;
      ray.add_point(nv.p, t1 + step / (kC / n1), next);
    } else if (c_free && (tn == RB_KOTHER || tn == RB_KOPT)) {
      ray.add_point(nv.p, t1 + step / kC, next);
    } else if (tc == RB_KLENS && tn == RB_KLENS) {
      const rb_volume_desc& v = vol_of(sc, next);
      do_fresnel(lens_index(sc.d, vol_of(sc, current), lambda), lens_index(sc.d, v, lambda),
                 lens_k(sc.d, v, lambda), ray, nv, current, next, o, rng);
    } else if (tc == RB_KLENS && (tn == RB_KNULL || tn == RB_KOPT || tn == RB_KOTHER)) {
      do_fresnel(lens_index(sc.d, vol_of(sc, current), lambda), 1, 0, ray, nv, current, next, o, rng);
    }

    if (tn == RB_KNULL) {
      ray.add_point(nv.p, t1 + step / kC, next);
      ray.status = RB_EXIT;
    } else if (tc == RB_KFOCUS || tc == RB_KOBS || tc == RB_KMIRROR || tn == RB_KOBS) {
      ray.status = RB_STOP;
    } else if (tn == RB_KFOCUS) {
      const rb_volume_desc& v = vol_of(sc, next);
      double angle = 0.;
      if (v.qe_angle_table >= 0) {
        V3 n = facet_normal(nv, ray, find_border(sc, current, next), rng);
        angle = std::acos(dot(ray.d, n));
      }
      double qe = v.qe_lambda_table >= 0 ? table_eval(sc.d, v.qe_lambda_table, lambda) : 1.;
      if (v.qe_angle_table >= 0) qe *= table_eval(sc.d, v.qe_angle_table, angle);
      if (qe == 1 || rng.uniform() < qe) ray.status = RB_FOCUS;
      else ray.status = RB_STOP;
    }

    if (ray.status == RB_RUN && ray.npoints >= o.limit) ray.status = RB_SUSPEND;
  }
}

// ---------------------------------------------------------------- dual-number adjoint reference
// Forward-mode restatement of the focal-plane loss gradient, used to check
// the device adjoint. Each focused ray is traced as above with its path
// recorded; the chain is then recomputed with dual numbers whose tangents are
// the parameters of each mirror the ray reflects on -- (tip, tilt, R) of a
// spherical facet, (c, kappa, K_1..K_n) of an AGeoAsphericDisk surface --
// and the focal-surface z. Nothing here is hand-differentiated: the facet
// rotation is rebuilt as M0 Rx(tip) Ry(tilt), the crossings are solved again
// (the sphere's quadratic; Newton steps on the disk surface's sag from the
// recorded crossing, CalcF and CalcdF/dr of src/AGeoAsphericDisk.cxx:96-153),
// and every step follows the forward formulas of reflect() and the
// navigator's step back.
constexpr int kDualPer = 2 + RB_ADJ_MAX_POLY;    // tangents per reflection
constexpr int kDualN = 4 * kDualPer + 1;         // at most 4 reflections + focal z
struct Dual {
  double v;
  double d[kDualN];
};
static inline Dual dc(double v) {
  Dual r;
  r.v = v;
  for (int k = 0; k < kDualN; k++) r.d[k] = 0;
  return r;
}
static inline Dual dvar(double v, int k) {
  Dual r = dc(v);
  r.d[k] = 1;
  return r;
}
static inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
  for (int k = 0; k < kDualN; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
static inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
  for (int k = 0; k < kDualN; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
static inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < kDualN; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
static inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v / b.v;
  for (int k = 0; k < kDualN; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
static inline Dual dsqrt(const Dual& a) {
  Dual r;
  r.v = std::sqrt(a.v);
  for (int k = 0; k < kDualN; k++) r.d[k] = a.d[k] / (2 * r.v);
  return r;
}
static inline Dual dsin(const Dual& a) {
  Dual r;
  r.v = std::sin(a.v);
  const double c = std::cos(a.v);
  for (int k = 0; k < kDualN; k++) r.d[k] = c * a.d[k];
  return r;
}
static inline Dual dcos(const Dual& a) {
  Dual r;
  r.v = std::cos(a.v);
  const double s = -std::sin(a.v);
  for (int k = 0; k < kDualN; k++) r.d[k] = s * a.d[k];
  return r;
}
struct DV3 {
  Dual x, y, z;
};
static inline DV3 dv(V3 a) { return {dc(a.x), dc(a.y), dc(a.z)}; }
static inline DV3 operator+(const DV3& a, const DV3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline DV3 operator-(const DV3& a, const DV3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline DV3 operator*(const DV3& a, const Dual& s) { return {a.x * s, a.y * s, a.z * s}; }
static inline Dual ddot(const DV3& a, const DV3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 val(const DV3& a) { return v3(a.x.v, a.y.v, a.z.v); }
// local = R^T (p - T) and master = R local (+ T), R constant
static inline DV3 dm2lv(const rb_matrix& m, const DV3& q) {
  const double* r = m.rot;
  return {dc(r[0]) * q.x + dc(r[3]) * q.y + dc(r[6]) * q.z, dc(r[1]) * q.x + dc(r[4]) * q.y + dc(r[7]) * q.z,
          dc(r[2]) * q.x + dc(r[5]) * q.y + dc(r[8]) * q.z};
}
static inline DV3 dl2mv(const rb_matrix& m, const DV3& q) {
  const double* r = m.rot;
  return {dc(r[0]) * q.x + dc(r[1]) * q.y + dc(r[2]) * q.z, dc(r[3]) * q.x + dc(r[4]) * q.y + dc(r[5]) * q.z,
          dc(r[6]) * q.x + dc(r[7]) * q.y + dc(r[8]) * q.z};
}
static inline DV3 dtr(const rb_matrix& m) { return dv(v3(m.tr[0], m.tr[1], m.tr[2])); }

// A disk surface's coefficients as duals (c, kappa, K_k), its offset z.
struct DSurf {
  Dual c, kap, K[RB_ADJ_MAX_POLY];
  double z;
  int n;
};
static DSurf dsurf(const Asph& a, int s, int base) {
  DSurf f;
  f.c = base >= 0 ? dvar(a.c[s], base) : dc(a.c[s]);
  f.kap = base >= 0 ? dvar(a.kap[s], base + 1) : dc(a.kap[s]);
  f.n = a.n[s];
  for (int k = 0; k < f.n; k++) f.K[k] = base >= 0 ? dvar(a.K[s][k], base + 2 + k) : dc(a.K[s][k]);
  f.z = a.z[s];
  return f;
}
// G = q_z - CalcF(r) and a = CalcdFdr(r) / r at the local point q
static void dsurf_eval(const DSurf& f, const DV3& q, Dual& G, Dual& a) {
  const Dual w = q.x * q.x + q.y * q.y;
  const Dual sq = dsqrt(dc(1) - w * f.c * f.c * f.kap);
  Dual F = w * f.c / (dc(1) + sq);
  a = f.c / sq;
  Dual wk1 = dc(1);  // w^(k-1)
  for (int k = 1; k <= f.n; k++) {
    F = F + f.K[k - 1] * wk1 * w;
    a = a + dc(2.0 * k) * f.K[k - 1] * wk1;
    wk1 = wk1 * w;
  }
  G = q.z - dc(f.z) - F;
}
// The ray o + t e (local) on the surface: Newton steps in t from t0, the
// recorded crossing, to machine precision; with dual numbers the tangent
// converges with the value.
static Dual dsurf_cross(const DSurf& f, const DV3& o, const DV3& e, double t0) {
  Dual t = dc(t0);
  for (int it = 0; it < 5; it++) {
    const DV3 q = o + e * t;
    Dual G, a;
    dsurf_eval(f, q, G, a);
    const Dual gp = e.z - a * (q.x * e.x + q.y * e.y);
    t = t - G / gp;
  }
  return t;
}

// How the chain's reflection (or the focal hit) is differentiated; as the
// device classifies it.
enum { kDNone = 0, kDFacet = 1, kDAsph = 2, kDPlane = 3 };
static int dual_kind(const rb_scene_desc* d, const rb_adjoint_desc* A, int node, V3 P, int& surf) {
  if (A->node_param) {
    const int32_t* np = A->node_param + 3 * node;
    bool ok = true;
    for (int k = 0; k < 3; k++) ok = ok && np[k] >= 0 && np[k] < A->n_params;
    if (ok) return kDFacet;
  }
  if (!A->surface_param) return kDNone;
  const int32_t first = A->surface_param[2 * node], count = A->surface_param[2 * node + 1];
  const rb_shape_desc& sh = d->shapes[d->volumes[d->nodes[node].volume].shape];
  if (first < 0 || sh.type != RB_SHAPE_ASPHERIC) return kDNone;
  const Asph a = asph_p(d->params + sh.param_offset);
  double df[2];
  const int face = asph_face(a, m2l(d->nodes[node].global, P), df);
  if (face < 2) return kDNone;
  surf = face - 2;
  if (a.n[surf] > RB_ADJ_MAX_POLY || count < 2 + a.n[surf] || int64_t(first) + count > A->n_params) return kDNone;
  return kDAsph;
}

// Loss gradient of one focused ray; returns false when the chain holds a
// reflection that cannot be differentiated (the device adjoint skips those too).
static bool dual_ray(const rb_scene_desc* d, const rb_adjoint_desc* A, V3 origin, V3 d0, const double* path,
                     const int32_t* pnode, int K, V3 hit, double* grad) {
  const double eps = g_step_back;
  if (K > 4) return false;
  int kinds[4] = {kDNone, kDNone, kDNone, kDNone}, surfs[4] = {0, 0, 0, 0};
  // the recorded points are the crossings stepped back by eps along the
  // incoming direction: moved forward again once that direction is known
  DV3 o = dv(origin), din = dv(d0);
  for (int j = 0; j < K; j++) {
    const int node = pnode[j];
    const rb_matrix& g = d->nodes[node].global;
    const V3 rec = axpy(v3(path[4 * j], path[4 * j + 1], path[4 * j + 2]), eps, val(din));
    kinds[j] = dual_kind(d, A, node, rec, surfs[j]);
    if (kinds[j] == kDNone) return false;
    const int base = kDualPer * j;
    DV3 P, n;
    if (kinds[j] == kDFacet) {
      const int32_t* np = A->node_param + 3 * node;
      const double tip0 = A->param_values[np[0]], tilt0 = A->param_values[np[1]];
      // M0 = M Ry(tilt0)^T Rx(tip0)^T: the placement without the tip/tilt
      double rx[9] = {1, 0, 0, 0, std::cos(tip0), -std::sin(tip0), 0, std::sin(tip0), std::cos(tip0)};
      double ry[9] = {std::cos(tilt0), 0, std::sin(tilt0), 0, 1, 0, -std::sin(tilt0), 0, std::cos(tilt0)};
      double e[9], m0[9];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double s = 0;
          for (int k = 0; k < 3; k++) s += rx[c * 3 + k] * ry[k * 3 + r];  // (Rx Ry)^T
          e[r * 3 + c] = s;
        }
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double s = 0;
          for (int k = 0; k < 3; k++) s += g.rot[r * 3 + k] * e[k * 3 + c];
          m0[r * 3 + c] = s;
        }
      const Dual tip = dvar(tip0, base), tilt = dvar(tilt0, base + 1), R = dvar(A->param_values[np[2]], base + 2);
      // local axis of the sphere centre: Rx(tip) Ry(tilt) e_z = (sin tilt, -sin tip cos tilt, cos tip cos tilt)
      const Dual ct = dcos(tilt);
      const DV3 a = {dsin(tilt), dc(0) - dsin(tip) * ct, dcos(tip) * ct};
      DV3 C;
      C.x = dc(g.tr[0]) + (dc(m0[0]) * a.x + dc(m0[1]) * a.y + dc(m0[2]) * a.z) * R;
      C.y = dc(g.tr[1]) + (dc(m0[3]) * a.x + dc(m0[4]) * a.y + dc(m0[5]) * a.z) * R;
      C.z = dc(g.tr[2]) + (dc(m0[6]) * a.x + dc(m0[7]) * a.y + dc(m0[8]) * a.z) * R;
      // |o + t din - C| = R; the root nearest the recorded crossing
      const DV3 oc = o - C;
      const Dual b = ddot(oc, din), c = ddot(oc, oc) - R * R;
      const Dual q = dsqrt(b * b - c);
      const Dual t1 = dc(0) - b - q, t2 = dc(0) - b + q;
      const V3 p1 = axpy(val(o), t1.v, val(din)), p2 = axpy(val(o), t2.v, val(din));
      const V3 e1 = sub(p1, rec), e2 = sub(p2, rec);
      const Dual t = dot(e1, e1) <= dot(e2, e2) ? t1 : t2;
      P = o + din * t;
      n = (P - C) * (dc(1) / R);
    } else {
      const rb_shape_desc& sh = d->shapes[d->volumes[d->nodes[node].volume].shape];
      const DSurf f = dsurf(asph_p(d->params + sh.param_offset), surfs[j], base);
      const DV3 ol = dm2lv(g, o - dtr(g)), el = dm2lv(g, din);
      const Dual t = dsurf_cross(f, ol, el, dot(sub(rec, val(o)), val(din)));
      const DV3 q = ol + el * t;
      Dual G, a;
      dsurf_eval(f, q, G, a);
      const DV3 gl = {dc(0) - a * q.x, dc(0) - a * q.y, dc(1)};
      n = dl2mv(g, gl * (dc(1) / dsqrt(ddot(gl, gl))));
      P = dtr(g) + dl2mv(g, q);
    }
    const Dual k = ddot(din, n);
    DV3 dout = din - n * (dc(2) * k);
    dout = dout * (dc(1) / dsqrt(ddot(dout, dout)));
    o = P - din * dc(eps);
    din = dout;
  }
  // the focal surface: the plane z = hit.z moved by the z parameter, or the
  // focal node's disk surface moved along its z axis
  const Dual dz = A->z_param >= 0 ? dvar(0.0, kDualN - 1) : dc(0.0);
  DV3 h;
  const rb_node_desc& fn = d->nodes[A->focal_node];
  const rb_shape_desc& fsh = d->shapes[d->volumes[fn.volume].shape];
  if (fsh.type == RB_SHAPE_ASPHERIC) {
    const Asph a = asph_p(d->params + fsh.param_offset);
    double df[2];
    const int face = asph_face(a, m2l(fn.global, hit), df);
    DV3 ol = dm2lv(fn.global, o - dtr(fn.global)), el = dm2lv(fn.global, din);
    ol.z = ol.z - dz;
    const double s0 = dot(sub(hit, val(o)), val(din));
    Dual s;
    if (face >= 2) {
      s = dsurf_cross(dsurf(a, face - 2, -1), ol, el, s0);
    } else {  // a wall, r = rmin or rmax: Newton steps on q_x^2 + q_y^2 - r^2
      const double rw = face == 0 ? a.rmin : a.rmax;
      s = dc(s0);
      for (int it = 0; it < 5; it++) {
        const DV3 q = ol + el * s;
        const Dual G = q.x * q.x + q.y * q.y - dc(rw * rw);
        const Dual gp = dc(2) * (q.x * el.x + q.y * el.y);
        s = s - G / gp;
      }
    }
    h = o + din * s;
  } else {
    const Dual s = (dc(hit.z) + dz - o.z) / din.z;
    h = o + din * s;
  }
  // the loss adjoint on (h_x, h_y)
  const double ux = h.x.v - A->centroid_x, uy = h.y.v - A->centroid_y;
  double gx = A->weight * ux, gy = A->weight * uy;
  if (A->mode == RB_ADJ_ENCIRCLED) {
    const double r0 = A->ee_radius, w = A->ee_width;
    const double sg = 1.0 / (1.0 + std::exp(-(r0 * r0 - ux * ux - uy * uy) / (2 * r0 * w)));
    const double f = -A->weight * sg * (1 - sg) / (r0 * w);
    gx = f * ux;
    gy = f * uy;
  }
  auto put = [&](int param, int k) {
    if (param >= 0) grad[param] += gx * h.x.d[k] + gy * h.y.d[k];
  };
  for (int j = 0; j < K; j++) {
    const int base = kDualPer * j;
    if (kinds[j] == kDFacet) {
      const int32_t* np = A->node_param + 3 * pnode[j];
      for (int k = 0; k < 3; k++) put(np[k], base + k);
    } else {
      const rb_shape_desc& sh = d->shapes[d->volumes[d->nodes[pnode[j]].volume].shape];
      const int n = int(d->params[sh.param_offset + 8 + surfs[j]]);
      const int first = A->surface_param[2 * pnode[j]];
      for (int k = 0; k < 2 + n; k++) put(first + k, base + k);
    }
  }
  put(A->z_param, kDualN - 1);
  return true;
}

static void dual_range(const Scene& sc, const Opts& o, const rb_trace_opts* opts, const rb_rays* r,
                       const rb_adjoint_desc* A, int64_t b, int64_t e, double* grad, int64_t* used) {
  constexpr int kCap = 8;
  double path[4 * kCap];
  int32_t pnode[kCap];
  for (int64_t i = b; i < e; i++) {
    if (r->status[i] != RB_RUN) continue;
    Ray ray;
    ray.x = v3(r->x[i], r->y[i], r->z[i]);
    ray.t = r->t[i];
    ray.lambda = r->lambda[i];
    ray.status = RB_RUN;
    ray.npoints = 1;
    ray.npoints_in = 1;
    ray.last_node = -1;
    ray.d = v3(r->dx[i], r->dy[i], r->dz[i]);
    ray.cap = kCap;
    ray.path = path;
    ray.pnode = pnode;
    Rng rng{opts->seed, opts->ray_id_base + uint64_t(i), 0u, 0u};
    trace_ray(sc, o, ray, rng);
    if (ray.status != RB_FOCUS || ray.last_node != A->focal_node || ray.npoints - 1 > kCap) continue;
    // the points between the origin and the focal hit: reflections on
    // mirrors, or boundary crossings that leave the line unchanged
    double rpath[4 * kCap];
    int32_t rnode[kCap];
    int K = 0;
    bool ok = true;
    for (int j = 0; j + 2 < ray.npoints; j++) {
      const int kind = kind_of(sc, pnode[j]);
      if (kind == RB_KMIRROR) {
        if (K == 4) ok = false;
        if (!ok) break;
        for (int c = 0; c < 4; c++) rpath[4 * K + c] = path[4 * j + c];
        rnode[K++] = pnode[j];
      } else if (kind == RB_KLENS) {
        ok = false;  // refraction: outside the model
      }
    }
    if (ok && dual_ray(sc.d, A, v3(r->x[i], r->y[i], r->z[i]), v3(r->dx[i], r->dy[i], r->dz[i]), rpath, rnode, K,
                       ray.x, grad))
      (*used)++;
  }
}

}  // namespace orc

using namespace orc;

static void trace_range(const Scene& sc, const Opts& o, const rb_trace_opts* opts, rb_rays* r, int64_t b,
                        int64_t e) {
  for (int64_t i = b; i < e; i++) {
    Ray ray;
    ray.x = v3(r->x[i], r->y[i], r->z[i]);
    ray.t = r->t[i];
    ray.lambda = r->lambda[i];
    ray.status = r->status[i];
    ray.npoints = r->npoints[i];
    ray.npoints_in = ray.npoints;
    ray.last_node = r->last_node ? r->last_node[i] : -1;
    ray.d = v3(r->dx[i], r->dy[i], r->dz[i]);  // already unit, as ARay stores it
    ray.cap = r->path_capacity;
    ray.path = ray.cap > 0 ? r->path_xyzt + int64_t(4) * ray.cap * i : nullptr;
    ray.pnode = ray.cap > 0 ? r->path_node + int64_t(ray.cap) * i : nullptr;
    Rng rng{opts->seed, opts->ray_id_base + uint64_t(i), 0u, 0u};
    if (ray.status == RB_RUN) trace_ray(sc, o, ray, rng);
    r->x[i] = ray.x.x; r->y[i] = ray.x.y; r->z[i] = ray.x.z; r->t[i] = ray.t;
    r->dx[i] = ray.d.x; r->dy[i] = ray.d.y; r->dz[i] = ray.d.z;
    r->status[i] = ray.status;
    r->npoints[i] = ray.npoints;
    if (r->last_node) r->last_node[i] = ray.last_node;
  }
}

extern "C" {

// Trace n rays with `nthreads` host threads in contiguous chunks of
// (n / nthreads), the last taking the remainder (src/AOpticsManager.cxx:533-541).
// Sets the navigator constants (a value <= 0 keeps the current one) and
// returns the ones in force in out[3] = {step back, push scale, bool push}.
// Not to be called while a trace runs.
void oracle_set_navigator(double step_back, double push_scale, double bool_push, double* out) {
  if (step_back > 0) g_step_back = step_back;
  if (push_scale > 0) g_push_scale = push_scale;
  if (bool_push > 0) g_bool_push = bool_push;
  if (out) {
    out[0] = g_step_back;
    out[1] = g_push_scale;
    out[2] = g_bool_push;
  }
}

int oracle_trace(const rb_scene_desc* desc, const rb_trace_opts* opts, rb_rays* rays, int64_t n, int nthreads) {
  if (!desc || !opts || !rays || n < 0) return -1;
  Scene sc{desc, {}};
  build_boxes(sc);
  Opts o{opts->limit > 0 ? opts->limit : 100, opts->fresnel != 0};
  if (nthreads < 2 || n < 2 * nthreads) {
    trace_range(sc, o, opts, rays, 0, n);
    return 0;
  }
  std::vector<std::thread> th;
  int64_t chunk = n / nthreads;
  for (int k = 0; k < nthreads; k++) {
    int64_t b = chunk * k, e = (k == nthreads - 1) ? n : chunk * (k + 1);
    th.emplace_back(trace_range, std::cref(sc), std::cref(o), opts, rays, b, e);
  }
  for (auto& t : th) t.join();
  return 0;
}

// Dual-number reference of rbgpu_adjoint: adds d(loss)/d(params) of the rays
// (not modified) into grad[n_params]; returns the number of rays that contributed.
int64_t oracle_adjoint_dual(const rb_scene_desc* desc, const rb_trace_opts* opts, const rb_rays* rays, int64_t n,
                            const rb_adjoint_desc* adj, double* grad, int nthreads) {
  if (!desc || !opts || !rays || !adj || !grad || n < 0) return -1;
  Scene sc{desc, {}};
  build_boxes(sc);
  Opts o{opts->limit > 0 ? opts->limit : 100, opts->fresnel != 0};
  nthreads = std::max(1, std::min<int>(nthreads, int(n / 64) + 1));
  std::vector<std::vector<double>> g(nthreads, std::vector<double>(adj->n_params, 0.0));
  std::vector<int64_t> used(nthreads, 0);
  std::vector<std::thread> th;
  int64_t chunk = n / nthreads;
  for (int k = 0; k < nthreads; k++) {
    int64_t b = chunk * k, e = (k == nthreads - 1) ? n : chunk * (k + 1);
    th.emplace_back(dual_range, std::cref(sc), std::cref(o), opts, rays, adj, b, e, g[k].data(), &used[k]);
  }
  for (auto& t : th) t.join();
  int64_t total = 0;
  for (int k = 0; k < nthreads; k++) {
    for (int p = 0; p < adj->n_params; p++) grad[p] += g[k][p];
    total += used[k];
  }
  return total;
}

// Single-threaded trace that also counts the operations it performs
// (OP_COUNT counters added into `counts`); the FLOP figure's input.
int oracle_trace_counts(const rb_scene_desc* desc, const rb_trace_opts* opts, rb_rays* rays, int64_t n,
                        uint64_t* counts, int n_counts) {
  if (!desc || !opts || !rays || !counts || n < 0 || n_counts < OP_COUNT) return -1;
  Scene sc{desc, {}};
  build_boxes(sc);
  Opts o{opts->limit > 0 ? opts->limit : 100, opts->fresnel != 0};
  std::vector<uint64_t> c(OP_COUNT, 0);
  g_ops = c.data();
  trace_range(sc, o, opts, rays, 0, n);
  g_ops = nullptr;
  for (int k = 0; k < OP_COUNT; k++) counts[k] += c[k];
  return OP_COUNT;
}

// Shape-level probes for tests (local frame of `shape`)
int oracle_shape_query(const rb_scene_desc* desc, int shape, const double* p, const double* d, double* out) {
  Scene sc{desc, {}};
  V3 q = v3(p[0], p[1], p[2]), dd = v3(d[0], d[1], d[2]);
  out[0] = contains(sc, shape, q) ? 1.0 : 0.0;
  Hit a = dist_out(sc, shape, q, dd);
  Hit b = dist_in(sc, shape, q, dd);
  out[1] = a.s;
  out[2] = b.s;
  V3 n = shape_normal(sc, shape, a.leaf, axpy(q, a.s < kBig ? a.s : 0, dd), dd);
  out[3] = n.x; out[4] = n.y; out[5] = n.z;
  // the normal where the ray leaves the shape (from an inside point)
  V3 m = shape_normal(sc, shape, b.leaf, axpy(q, b.s < kBig ? b.s : 0, dd), dd);
  out[6] = m.x; out[7] = m.y; out[8] = m.z;
  return 0;
}

// Ray sources (host restatement of rbgpu_generate_rays): GRID follows
// ARayShooter::Rectangle (src/ARayShooter.cxx:401-459) translated to
// (-D sin th, 0, D cos th); DISK draws 4 Philox uniforms per ray (stream 1).
int oracle_generate_rays(const rb_raygen_desc* g, rb_rays* r, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    double ox, oy, oz, dx, dy, dz, lam;
    if (g->kind == RB_RAYGEN_CONE) {
      Rng u{g->seed, uint64_t(i), 0u, 1u};
      double u0 = u.uniform(), u1 = u.uniform(), u2 = u.uniform(), u3 = u.uniform(), u4 = u.uniform();
      double rho = 0.5 * g->size * std::sqrt(u0), al = 2 * kPi * u1;
      double th = g->theta_min + (g->theta_max - g->theta_min) * u2, ph = 2 * kPi * u3;
      lam = g->lambda_min + (g->lambda_max - g->lambda_min) * u4;
      double ri = g->distance * std::tan(th);
      double ix = ri * std::cos(ph), iy = ri * std::sin(ph);
      double vx = ix - rho * std::cos(al), vy = iy - rho * std::sin(al), vz = -g->distance;
      double m = std::sqrt(vx * vx + vy * vy + vz * vz);
      dx = vx / m; dy = vy / m; dz = vz / m;
      double s = g->start / -dz;
      ox = ix - s * dx; oy = iy - s * dy; oz = g->z0 + g->start;
    } else {
      double th, px, py;
      if (g->kind == RB_RAYGEN_GRID) {
        int ns = g->n_side;
        double delta = ns == 1 ? g->size / 2 : g->size / (ns - 1);
        int64_t a = i / ns, b = i % ns;
        px = a * delta - g->size / 2;
        py = b * delta - g->size / 2;
        th = g->theta_min;
        lam = g->lambda_min;
      } else {
        Rng u{g->seed, uint64_t(i), 0u, 1u};
        double u0 = u.uniform(), u1 = u.uniform(), u2 = u.uniform(), u3 = u.uniform();
        th = g->theta_min + (g->theta_max - g->theta_min) * u2;
        lam = g->lambda_min + (g->lambda_max - g->lambda_min) * u3;
        double rr = 0.5 * g->size * std::sqrt(u0), ph = 2 * kPi * u1;
        px = rr * std::cos(ph);
        py = rr * std::sin(ph);
      }
      double st = std::sin(th), ct = std::cos(th), t2 = kPi - th;
      double cx = -g->distance * st, cz = g->distance * ct;
      if (g->kind == RB_RAYGEN_GRID) {
        ox = px + cx; oy = py; oz = cz;
      } else {
        ox = cx + px * ct; oy = py; oz = cz + px * st;
      }
      dx = std::sin(t2); dy = 0; dz = std::cos(t2);
    }
    r->x[i] = ox; r->y[i] = oy; r->z[i] = oz; r->t[i] = 0;
    r->dx[i] = dx; r->dy[i] = dy; r->dz[i] = dz;
    r->lambda[i] = lam;
    r->status[i] = RB_RUN;
    r->npoints[i] = 1;
    if (r->last_node) r->last_node[i] = -1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// ARayShooter (src/ARayShooter.cxx:122-459), restated with its loops: the
// pattern is built in the reference's order and each ray placed with
// TGeoRotation::LocalToMaster (m_i = l0 R[3i] + l1 R[3i+1] + l2 R[3i+2]) and
// TGeoTranslation::LocalToMaster (m_i = l_i + T_i), then ARay's constructor
// (src/ARay.cxx:28-36: AddPoint, SetDirection normalising a non-zero vector,
// :217-223). The random kinds draw Philox uniforms (stream 2, key seed,
// counter id_base + ray) in place of gRandom; the disk and the sphere use
// inverse-CDF draws where the reference rejects (RandomCircle :211-215,
// RandomCone :250-256, TRandom::Sphere): the same uniform measure.
namespace shoot {
static void rot(const double* R, const double* l, double* m) {
  for (int i = 0; i < 3; i++) m[i] = l[0] * R[3 * i] + l[1] * R[3 * i + 1] + l[2] * R[3 * i + 2];
}
static void trn(const double* T, const double* l, double* m) {
  for (int i = 0; i < 3; i++) m[i] = l[i] + T[i];
}
struct Out {
  rb_rays* r;
  int64_t k = 0;
  void add(double lam, const double* x, const double* d) {  // new ARay(0, lam, x, 0, d)
    double mag = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double u[3] = {0, 0, 0};
    if (mag > 0) { u[0] = d[0] / mag; u[1] = d[1] / mag; u[2] = d[2] / mag; }
    r->x[k] = x[0]; r->y[k] = x[1]; r->z[k] = x[2]; r->t[k] = 0;
    r->dx[k] = u[0]; r->dy[k] = u[1]; r->dz[k] = u[2];
    r->lambda[k] = lam;
    r->status[k] = RB_RUN;
    r->npoints[k] = 1;
    if (r->last_node) r->last_node[k] = -1;
    k++;
  }
};
}  // namespace shoot

int64_t oracle_shooter_count(const rb_shooter_desc* g, int64_t n) {
  switch (g->kind) {
    case RB_SHOOT_RECTANGLE:
      return (g->a < 0 || g->b < 0 || g->nx < 1 || g->ny < 1) ? 0 : int64_t(g->nx) * g->ny;
    case RB_SHOOT_CIRCLE: {
      if (g->a < 0 || g->nx < 1 || g->ny < 1) return 0;
      int64_t c = 1;
      for (int64_t i = 0; i < g->nx; i++) c += int64_t(g->ny) * (i + 1);
      return c;
    }
    case RB_SHOOT_RANDOM_CIRCLE: return g->a < 0 ? 0 : std::max<int64_t>(n, 0);
    case RB_SHOOT_RANDOM_RECTANGLE: return (g->a < 0 || g->b < 0 || n < 1) ? 0 : n;
    case RB_SHOOT_RANDOM_CONE:
    case RB_SHOOT_RANDOM_SPHERE:
    case RB_SHOOT_RANDOM_SPHERICAL_CONE: return std::max<int64_t>(n, 0);
  }
  return -1;
}

int oracle_shoot_rays(const rb_shooter_desc* g, rb_rays* r, int64_t n) {
  using namespace shoot;
  if (oracle_shooter_count(g, n) != n) return -1;
  Out out{r};
  const double* R = g->rot;
  const double* T = g->tr;
  double dir[3], new_dir[3];
  rot(R, g->dir, new_dir);  // the v direction, rotated (Rectangle :419-423 and the others alike)
  switch (g->kind) {
    case RB_SHOOT_RECTANGLE: {  // :401-452
      const double dx = g->a, dy = g->b;
      const int nx = g->nx, ny = g->ny;
      double deltax = nx == 1 ? dx / 2 : dx / (nx - 1);
      double deltay = ny == 1 ? dy / 2 : dy / (ny - 1);
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
          double x[3] = {i * deltax - dx / 2, j * deltay - dy / 2, 0}, np[3];
          rot(R, x, np);
          trn(T, np, x);
          out.add(g->lambda, x, new_dir);
        }
      break;
    }
    case RB_SHOOT_CIRCLE: {  // :122-183
      const double rmax = g->a;
      const int nr = g->nx, nphi = g->ny;
      double position[3] = {0, 0, 0}, np[3];
      trn(T, position, np);
      out.add(g->lambda, np, new_dir);
      for (int i = 0; i < nr; i++) {
        double rr = rmax * (i + 1) / nr;
        for (int j = 0; j < nphi * (i + 1); j++) {
          double phi = 2 * kPi / nphi / (i + 1) * j;
          double x[3] = {rr * std::cos(phi), rr * std::sin(phi), 0};
          rot(R, x, np);
          trn(T, np, x);
          out.add(g->lambda, x, new_dir);
        }
      }
      break;
    }
    case RB_SHOOT_RANDOM_CIRCLE:  // :186-237
      for (int64_t i = 0; i < n; i++) {
        Rng u{g->seed, g->id_base + uint64_t(i), 0u, 2u};
        double rho = g->a * std::sqrt(u.uniform()), phi = 2 * kPi * u.uniform();
        double x[3] = {rho * std::cos(phi), rho * std::sin(phi), 0}, np[3];
        rot(R, x, np);
        trn(T, np, x);
        out.add(g->lambda, x, new_dir);
      }
      break;
    case RB_SHOOT_RANDOM_CONE:  // :240-285: from tr toward a disk of radius r at z = d
      for (int64_t i = 0; i < n; i++) {
        Rng u{g->seed, g->id_base + uint64_t(i), 0u, 2u};
        double rho = g->a * std::sqrt(u.uniform()), phi = 2 * kPi * u.uniform();
        double goal[3] = {rho * std::cos(phi), rho * std::sin(phi), g->b}, tmp[3];
        rot(R, goal, tmp);
        trn(T, tmp, goal);
        double start[3] = {0, 0, 0};
        trn(T, start, tmp);
        double d[3] = {goal[0] - tmp[0], goal[1] - tmp[1], goal[2] - tmp[2]};
        out.add(g->lambda, tmp, d);
      }
      break;
    case RB_SHOOT_RANDOM_RECTANGLE:  // :288-335
      for (int64_t i = 0; i < n; i++) {
        Rng u{g->seed, g->id_base + uint64_t(i), 0u, 2u};
        double lo_x = -g->a / 2., lo_y = -g->b / 2.;
        double px = lo_x + (g->a / 2. - lo_x) * u.uniform();  // TRandom::Uniform(x1, x2) = x1 + (x2 - x1) U
        double py = lo_y + (g->b / 2. - lo_y) * u.uniform();
        double x[3] = {px, py, 0}, np[3];
        rot(R, x, np);
        trn(T, np, x);
        out.add(g->lambda, x, new_dir);
      }
      break;
    case RB_SHOOT_RANDOM_SPHERE:  // :338-356: isotropic from tr, no rotation
      for (int64_t i = 0; i < n; i++) {
        Rng u{g->seed, g->id_base + uint64_t(i), 0u, 2u};
        double cz = 2 * u.uniform() - 1, phi = 2 * kPi * u.uniform();
        double s = std::sqrt(std::max(0.0, 1 - cz * cz));
        dir[0] = s * std::cos(phi); dir[1] = s * std::sin(phi); dir[2] = cz;
        double p[3] = {0, 0, 0}, np[3];
        trn(T, p, np);
        out.add(g->lambda, np, dir);
      }
      break;
    case RB_SHOOT_RANDOM_SPHERICAL_CONE: {  // :359-391
      const double c = std::cos(g->a * (kPi / 180.0));
      for (int64_t i = 0; i < n; i++) {
        Rng u{g->seed, g->id_base + uint64_t(i), 0u, 2u};
        double ran = c + (1 - c) * u.uniform();
        double th = std::acos(ran);
        double phi = 0 + (2 * kPi - 0) * u.uniform();
        dir[0] = std::sin(th) * std::cos(phi); dir[1] = std::sin(th) * std::sin(phi); dir[2] = std::cos(th);
        double nd[3], p[3] = {0, 0, 0}, np[3];
        rot(R, dir, nd);
        trn(T, p, np);
        out.add(g->lambda, np, nd);
      }
      break;
    }
    default: return -1;
  }
  return out.k == n ? 0 : -2;
}

// ACorsikaIACTFile::GetRayArray (src/ACorsikaIACTFile.cxx:71-131) over a
// bunch table: the loop over bunches, the telescope/array filter, the shift of
// each bunch from the observation level up to height z along its direction,
// and `for (j = 0; j < photons; j++)` rays per bunch (float compare as there).
// A bunch with lambda == 0 draws its wavelength per ray (Philox stream 3,
// counter id_base + output ray). Returns the ray count; rays are written only
// when `r` is not null and the count fits `capacity`. -1: photons above 2^24.
int64_t oracle_bunches_to_rays(const rb_bunches* b, int64_t nb, const rb_bunch_opts* o, rb_rays* r,
                               int64_t capacity) {
  const double m = 1e2 * 1.0, cm = 1.0, nm = 1e-9 * m, ns = 1e-9 * 1.0;  // AOpticsManager units
  int64_t total = 0;
  for (int64_t i = 0; i < nb; i++) {
    if (o->tel_no != b->tel_no[i] || o->array_no != b->array_no[i]) continue;
    const float photons = b->photons[i];
    if (photons > 16777216.0f) return -1;
    for (int64_t j = 0; float(j) < photons; j++) total++;
  }
  if (!r || total > capacity) return total;
  int64_t k = 0;
  for (int64_t i = 0; i < nb; i++) {
    if (o->tel_no != b->tel_no[i] || o->array_no != b->array_no[i]) continue;
    const float x = b->x[i], y = b->y[i], time = b->time[i], cx = b->cx[i], cy = b->cy[i], cz = b->cz[i];
    const float lambda = b->lambda[i], photons = b->photons[i];
    double airmass = -1. / cz;
    double tel_dist = (o->z - o->tel_z * cm) * airmass;
    double speed = kC / o->refractive_index;
    double px = x * cm - tel_dist * cx;
    double py = y * cm - tel_dist * cy;
    double pt = time * ns - tel_dist / speed;
    for (int64_t j = 0; float(j) < photons; j++, k++) {
      double lam = lambda;
      if (lambda == 0) {
        Rng u{o->seed, o->id_base + uint64_t(k), 0u, 3u};
        lam = 1. / (1. / o->lambda_min - u.uniform() * (1. / o->lambda_min - 1. / o->lambda_max));
      }
      double d[3] = {cx, cy, cz};
      double mag = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double u3[3] = {0, 0, 0};
      if (mag > 0) { u3[0] = d[0] / mag; u3[1] = d[1] / mag; u3[2] = d[2] / mag; }
      r->x[k] = px; r->y[k] = py; r->z[k] = o->z; r->t[k] = pt;
      r->dx[k] = u3[0]; r->dy[k] = u3[1]; r->dz[k] = u3[2];
      r->lambda[k] = lam * nm;
      r->status[k] = RB_RUN;
      r->npoints[k] = 1;
      if (r->last_node) r->last_node[k] = -1;
    }
  }
  return total;
}

// Batch containment probe for brute-force distance checks (local frame of `shape`).
int oracle_contains_batch(const rb_scene_desc* desc, int shape, int64_t n, const double* pts, uint8_t* out) {
  Scene sc{desc, {}};
  for (int64_t i = 0; i < n; i++) out[i] = contains(sc, shape, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2])) ? 1 : 0;
  return 0;
}

double oracle_table_eval(const rb_scene_desc* desc, int table, double x) { return table_eval(desc, table, x); }

double oracle_th2_interpolate(const rb_grid2d_desc* g, const double* vals, double x, double y) {
  return th2_interpolate(vals, *g, x, y);
}

double oracle_lens_index(const rb_scene_desc* desc, int volume, double lambda) {
  return lens_index(desc, desc->volumes[volume], lambda);
}

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key[0], key[1]);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

double oracle_uniform(uint64_t seed, uint64_t id, uint32_t draw, uint32_t stream) {
  Rng r{seed, id, draw, stream};
  return r.uniform();
}

// ---------------------------------------------------------------------------
// Spot analysis after the trace: ROOT's TH2D fill of the focal-plane hits and
// AGeoUtil::ContainmentRadius (src/AGeoUtil.cxx:198-308) over it.
//
// TH2D(nx, xlo, xhi, ny, ylo, yhi) semantics restated from ROOT's documented
// behaviour (ROOT is not in the reference tree; version unpinned):
//   TAxis::FindBin   x < lo -> 0; !(x < hi) -> n + 1; else 1 + int(n (x - lo) / (hi - lo))
//   TAxis::GetBinCenter(b) = lo + (b - 1) w + 0.5 w,  w = (hi - lo) / n
//   TH2::Fill        an entry outside either axis goes to an under/overflow
//                    bin and does not enter the statistics (sum w, x, x^2, y, y^2)
//   TH1::GetMean     sum x / sum w;  GetStdDev = sqrt(max(sum x^2 / sum w - mean^2, 0))
//   TH1::Integral    sum of the in-range bins
// counts[(by - 1) nx + (bx - 1)] holds bin (bx, by); stats = {sum w, sum x,
// sum x^2, sum y, sum y^2}, summed in ray order.
static int find_bin(double x, int n, double lo, double hi) {
  if (x < lo) return 0;
  if (!(x < hi)) return n + 1;
  return 1 + int(n * (x - lo) / (hi - lo));
}
static double bin_center(int b, int n, double lo, double hi) {
  const double w = (hi - lo) / double(n);
  return lo + (b - 1) * w + 0.5 * w;
}

int oracle_hist2d(const double* x, const double* y, const int32_t* status, int64_t n, int32_t want, int nx, double xlo,
                  double xhi, int ny, double ylo, double yhi, uint32_t* counts, double* stats) {
  if (nx <= 0 || ny <= 0 || !(xhi > xlo) || !(yhi > ylo)) return -1;
  std::fill(counts, counts + int64_t(nx) * ny, 0u);
  for (int k = 0; k < 5; k++) stats[k] = 0.0;
  for (int64_t i = 0; i < n; i++) {
    if (status[i] != want) continue;
    const int bx = find_bin(x[i], nx, xlo, xhi), by = find_bin(y[i], ny, ylo, yhi);
    if (bx < 1 || bx > nx || by < 1 || by > ny) continue;
    counts[int64_t(by - 1) * nx + (bx - 1)]++;
    stats[0] += 1.0;
    stats[1] += x[i];
    stats[2] += x[i] * x[i];
    stats[3] += y[i];
    stats[4] += y[i] * y[i];
  }
  return 0;
}

// SumInRadius (src/AGeoUtil.cxx:18-42): contents of the bins whose centres lie
// within r of (x, y).
static double sum_in_radius(const uint32_t* counts, int nx, double xlo, double xhi, int ny, double ylo, double yhi,
                            double x, double y, double r) {
  const double r2 = r * r;
  double total = 0.;
  for (int ix = 1; ix <= nx; ++ix) {
    const double cx = bin_center(ix, nx, xlo, xhi);
    for (int iy = 1; iy <= ny; ++iy) {
      const double cy = bin_center(iy, ny, ylo, yhi);
      const double c = counts[int64_t(iy - 1) * nx + (ix - 1)];
      if (c <= 0) continue;
      const double d2 = (cx - x) * (cx - x) + (cy - y) * (cy - y);
      if (d2 <= r2) total += c;
    }
  }
  return total;
}

// ContainmentRadius (src/AGeoUtil.cxx:198-308): radius and centre of the
// circle holding `fraction` of the histogram. out = {r, x, y}. Each
// hill-climbing run of the centre is capped at kMaxClimb steps, as on the
// device (the reference's loops are uncapped; the cap is never reached on
// real spots, since every step raises the contained count).
static const int kMaxClimb = 1 << 16;
void oracle_containment(const uint32_t* counts, const double* stats, int nx, double xlo, double xhi, int ny,
                        double ylo, double yhi, double fraction, double* out) {
  auto sir = [&](double x, double y, double r) { return sum_in_radius(counts, nx, xlo, xhi, ny, ylo, yhi, x, y, r); };
  int no_shift = 0, no_stable = 0;
  const double sw = stats[0];
  double x = sw == 0 ? 0. : stats[1] / sw;
  double y = sw == 0 ? 0. : stats[3] / sw;
  const double vx = sw == 0 ? 0. : std::max(stats[2] / sw - x * x, 0.0);
  const double vy = sw == 0 ? 0. : std::max(stats[4] / sw - y * y, 0.0);
  double r = std::sqrt(std::sqrt(vx) * std::sqrt(vx) + std::sqrt(vy) * std::sqrt(vy)) * 1.5;
  double dr = 0.1 * r;
  double integral = 0.;
  for (int64_t b = 0; b < int64_t(nx) * ny; b++) integral += counts[b];
  const double sum_goal = integral * fraction;
  for (int i = 0; i < 100 && no_shift < 30; i++) {
    bool stable_r = false, stable_x = true, stable_y = true;
    const double sum0 = sir(x, y, r);
    double next_r = r;
    if (sum0 < sum_goal) {
      const double sum1 = sir(x, y, r + dr);
      if (sum1 == sum0) {
        dr *= 2.;
        continue;
      }
      next_r = r + dr * (sum_goal - sum0) / (sum1 - sum0);
    } else if (sum0 != sum_goal) {
      const double sum1 = sir(x, y, r - dr);
      if (sum1 == sum0) {
        dr *= 2.;
        continue;
      }
      next_r = r - dr * (sum0 - sum_goal) / (sum0 - sum1);
    }
    if (next_r < 0.) next_r = 0.5 * r;
    if (next_r < 0.5 * r) next_r = 0.5 * r;
    if (next_r > 2. * r) next_r = 2. * r;
    stable_r = std::fabs(next_r - r) < 0.0001 * r;
    r = next_r;
    {
      double sum1 = sir(x, y, r);
      dr *= sum0 != sum_goal ? std::fabs((sum1 - sum_goal) / (sum0 - sum_goal)) : 0.5;
      if (dr > 0.5 * r) dr = 0.5 * r;
      if (dr < 0.0005 * r) dr = 0.0005 * r;
      no_shift++;
      for (double dx = 0.25 * r; dx > 0.1 * dr; dx *= 0.25) {
        double sum_x1 = sir(x + dx, y, r);
        double sum_x2 = sir(x - dx, y, r);
        for (int climb = 0; climb < kMaxClimb && sum_x1 > sum1; climb++) {
          no_shift = 0;
          x += dx;
          sum_x2 = sum1;
          sum1 = sum_x1;
          sum_x1 = sir(x + dx, y, r);
          stable_x = false;
        }
        for (int climb = 0; climb < kMaxClimb && sum_x2 > sum1; climb++) {
          no_shift = 0;
          x -= dx;
          sum_x1 = sum1;
          sum1 = sum_x2;
          sum_x2 = sir(x - dx, y, r);
          stable_x = false;
        }
      }
    }
    for (double dy = 0.1 * r; dy > 0.1 * dr; dy *= 0.25) {
      double sum1 = sir(x, y, r);
      double sum_y1 = sir(x, y + dy, r);
      double sum_y2 = sir(x, y - dy, r);
      for (int climb = 0; climb < kMaxClimb && sum_y1 > sum1; climb++) {
        no_shift = 0;
        y += dy;
        sum_y2 = sum1;
        sum1 = sum_y1;
        sum_y1 = sir(x, y + dy, r);
        stable_y = false;
      }
      for (int climb = 0; climb < kMaxClimb && sum_y2 > sum1; climb++) {
        no_shift = 0;
        y -= dy;
        sum_y1 = sum1;
        sum1 = sum_y2;
        sum_y2 = sir(x, y - dy, r);
        stable_y = false;
      }
    }
    if (stable_r && stable_x && stable_y) no_stable++;
    else no_stable = 0;
    if (no_stable >= 4) break;
  }
  out[0] = r;
  out[1] = x;
  out[2] = y;
}

}  // extern "C"
