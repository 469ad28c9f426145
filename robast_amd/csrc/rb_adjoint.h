// rb_adjoint.h -- reverse-mode derivative of a focal-plane loss through a
// recomputed chain of mirror reflections (device).
//
// Forward chain per reflection j (o_j = previous step-back point, or the ray
// origin; din_j its direction):
//   P_j  = o_j + t_j din_j on the mirror surface          (t_j implicit)
//   n_j  = the surface normal at P_j, k_j = din_j . n_j
//   d'_j = din_j - 2 k_j n_j                               (specular reflection)
//   P'_j = P_j - eps din_j                                 (the 2e-6 cm step back)
// Spherical facets: |P_j - C_j| = R_j with facet centre C = T + M (0, 0, R),
// M = M0 Rx(tip) Ry(tilt), so
//   dC/dR = M e_z,  dC/dtip = -R cos(tilt) M e_y,  dC/dtilt = R M e_x.
// AGeoAsphericDisk surfaces, in the node's frame q = M^T (P - T):
//   G(q) = q_z - z - F(w) = 0, w = q_x^2 + q_y^2,
//   F = w c / (1 + sq) + sum_k K_k w^k, sq = sqrt(1 - w c^2 kappa),
//   normal along g = (-a q_x, -a q_y, 1), a = F'(r) / r = c / sq + sum_k 2k K_k w^(k-1),
// so t_j moves with the parameters by dt = (dF/dtheta - g.(do + t de)) / (g.e).
// The chain ends on the focal surface: the plane z = z_cam, or a disk
// surface of the focal node (an implicit crossing like the mirrors').
// The loss per focused ray is a function of its hit (h_x, h_y); its adjoint
// hbar seeds the reverse pass. Spherical-facet steps run in FP32 on
// quantities formed relative to each facet in FP64; aspheric steps and the
// adjoints carried between segments are FP64; gradients are accumulated in
// FP64.
#pragma once

#include "rb_trace.h"

namespace rbd {

constexpr int kMaxAdjParams = 1024;

constexpr int kMaxAsphPoly = RB_ADJ_MAX_POLY;

struct AdjParams {
  const int32_t* node_param;     // [n_nodes][3], or null
  const int32_t* surface_param;  // [n_nodes][2] (first, count), or null
  const double* values;
  int32_t n_params, z_param, focal_node, mode;
  double cx, cy, weight;
  double ee_r0, ee_w;         // encircled: radius and width
  const double* ref_xy;       // moments: per-group reference points, or null
  int64_t group_size;         // moments: rays per group
  int32_t n_groups;           // moments: groups (>= 1)
};
constexpr int kMaxAdjAcc = 6400;  // LDS accumulator entries (FP64)
// accumulator entries per group, and in all
__host__ __device__ __forceinline__ int adj_block(int mode, int n_params) {
  return mode == RB_ADJ_MOMENTS ? 3 * n_params + 4 : mode == RB_ADJ_ENCIRCLED ? n_params + 1 : n_params;
}
__host__ __device__ __forceinline__ int adj_acc_size(int mode, int n_params, int n_groups) {
  return mode == RB_ADJ_MOMENTS ? n_groups * (3 * n_params + 4) : adj_block(mode, n_params);
}

struct V3f {
  float x, y, z;
};
__device__ __forceinline__ V3f f3(V3<double> v) { return {float(v.x), float(v.y), float(v.z)}; }
__device__ __forceinline__ float dotf(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3f addf(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3f mulf(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// Which face of an AGeoAsphericDisk (params p) the local point q lies on,
// by the safety rule of its normal (src/AGeoAsphericDisk.cxx:246-328):
// 0 inner wall, 1 outer wall, 2 surface F1, 3 surface F2.
__device__ inline int asph_face(const double* p, V3<double> q) {
  const double r = sqrt(q.x * q.x + q.y * q.y);
  double m = p[5] > 0.0 ? fabs(r - p[5]) : kBig;
  int k = 0;
  const double s1 = fabs(r - p[4]);
  if (s1 < m) { k = 1; m = s1; }
  for (int s = 0; s < 2; s++) {
    double f, df;
    if (asph_sag<double>(p, s, r, f) && asph_dsag<double>(p, s, r, df)) {
      const double saf = fabs(f - q.z) / sqrt(1.0 + df * df);
      if (saf < m) { k = 2 + s; m = saf; }
    }
  }
  return k;
}

// How a reflection (or the focal hit) is differentiated.
enum { kAdjNone = 0, kAdjFacet = 1, kAdjAsph = 2, kAdjPlane = 3, kAdjWall = 4 };

// Kind of the reflection on `node` at P; for kAdjAsph also the surface
// (0: F1, 1: F2). Parameter indices outside [0, n_params) make it kAdjNone:
// they would read param_values and write the LDS accumulators out of bounds.
__device__ inline int reflection_kind(const DScene& S, const AdjParams& A, int node, V3<double> P, int& surf) {
  if (A.node_param) {
    const int32_t* np = A.node_param + 3 * node;
    bool ok = true;
    for (int k = 0; k < 3; k++) ok = ok && np[k] >= 0 && np[k] < A.n_params;
    if (ok) return kAdjFacet;
  }
  if (!A.surface_param) return kAdjNone;
  const int32_t first = A.surface_param[2 * node], count = A.surface_param[2 * node + 1];
  const DNode& nd = S.nodes[node];
  const DShape sh = S.shapes[nd.shape];
  if (first < 0 || sh.type != RB_SHAPE_ASPHERIC) return kAdjNone;
  const double* p = S.params + sh.off;
  const int face = asph_face(p, m2l<double>(nd.rot, nd.tr, P));
  if (face < 2) return kAdjNone;
  surf = face - 2;
  const int n = int(p[8 + surf]);
  if (n > kMaxAsphPoly || count < 2 + n || int64_t(first) + count > A.n_params) return kAdjNone;
  return kAdjAsph;
}

// Kind of the focal hit: the plane z = z_cam, or, on a disk, a surface
// (surf 0, 1) or a wall (kAdjWall; surf 0 inner, 1 outer).
__device__ inline int focal_kind(const DScene& S, const AdjParams& A, V3<double> hit, int& surf) {
  const DNode& nd = S.nodes[A.focal_node];
  const DShape sh = S.shapes[nd.shape];
  if (sh.type != RB_SHAPE_ASPHERIC) return kAdjPlane;
  const int face = asph_face(S.params + sh.off, m2l<double>(nd.rot, nd.tr, hit));
  surf = face & 1;
  return face < 2 ? kAdjWall : kAdjAsph;
}

// True if every reflection of the chain can be differentiated.
__device__ __forceinline__ bool chain_covered(const DScene& S, const AdjParams& A, const ChainRecord& rec) {
  int surf;
  for (int j = 0; j < rec.n; j++)
    if (reflection_kind(S, A, rec.node[j], rec.P[j], surf) == kAdjNone) return false;
  return true;
}

// A disk surface at w = r^2 with the reference's unnormalised coefficients:
// sag F - z, a = F'(r) / r and their derivatives.
struct AsphAt {
  double sag, a, dadw;             // F - z, a, da/dw
  double sag_c, sag_k, a_c, a_k;   // d/dc, d/dkappa
};
__device__ inline AsphAt asph_at(const double* p, int s, double w) {
  const double c = p[1 + 2 * s], kap = p[6 + s];
  const int n = int(p[8 + s]);
  const double* K = p + 10 + (s ? int(p[8]) : 0);
  const double sq = sqrt(1.0 - w * c * c * kap), u = 1.0 + sq, sq3 = sq * sq * sq;
  AsphAt r;
  r.sag = w * c / u;
  r.a = c / sq;
  r.dadw = c * c * c * kap / (2.0 * sq3);
  r.sag_c = w / u + w * w * c * c * kap / (sq * u * u);
  r.sag_k = w * w * c * c * c / (2.0 * sq * u * u);
  r.a_c = 1.0 / sq3;
  r.a_k = w * c * c * c / (2.0 * sq3);
  double wk1 = 1.0, wk2 = 0.0;  // w^(k-1), w^(k-2)
  for (int k = 1; k <= n; k++) {
    r.sag += K[k - 1] * wk1 * w;
    r.a += 2.0 * k * K[k - 1] * wk1;
    r.dadw += 2.0 * k * (k - 1) * K[k - 1] * wk2;
    wk2 = wk1;
    wk1 *= w;
  }
  return r;
}

// Unit normal direction (unnormalised g = (-a q_x, -a q_y, 1)) of a disk
// surface at the local point q.
__device__ __forceinline__ V3<double> asph_grad(const double* p, int s, V3<double> q, AsphAt& at) {
  at = asph_at(p, s, q.x * q.x + q.y * q.y);
  return mk(-at.a * q.x, -at.a * q.y, 1.0);
}

// Outgoing direction of the reflection on `node` (kind, surf as classified).
__device__ inline V3<double> reflected_dir(const DScene& S, const AdjParams& A, int node, int kind, int surf,
                                           V3<double> P, V3<double> din) {
  const DNode& nd = S.nodes[node];
  V3<double> n;
  if (kind == kAdjFacet) {
    const double R = A.values[A.node_param[3 * node + 2]];
    V3<double> C = mk(nd.tr[0] + nd.rot[2] * R, nd.tr[1] + nd.rot[5] * R, nd.tr[2] + nd.rot[8] * R);
    n = scale(mk(P.x - C.x, P.y - C.y, P.z - C.z), 1.0 / R);
  } else {
    const double* p = S.params + S.shapes[nd.shape].off;
    AsphAt at;
    const V3<double> g = asph_grad(p, surf, m2l<double>(nd.rot, nd.tr, P), at);
    n = l2mv<double>(nd.rot, scale(g, 1.0 / sqrt(dot(g, g))));
  }
  const double k = dot(din, n);
  V3<double> d = mk(din.x - 2 * k * n.x, din.y - 2 * k * n.y, din.z - 2 * k * n.z);
  return scale(d, 1.0 / sqrt(dot(d, d)));
}

// Reverse pass through one reflection on a disk surface (FP64). On entry
// Ppbar and doutbar are the adjoints of the step-back point P' and of the
// outgoing direction; on exit those of the segment's origin o and of its
// incoming direction din. Adds dL/d(c, kappa, K_k) at param first + (0, 1, 1 + k).
__device__ inline void asph_reflection_adjoint(const DScene& S, const DNode& nd, int surf, int first, V3<double> P,
                                               V3<double> din, V3<double> o, double eps, V3<double>& Ppbar,
                                               V3<double>& doutbar, double* acc) {
  const double* p = S.params + S.shapes[nd.shape].off;
  const V3<double> q = m2l<double>(nd.rot, nd.tr, P), e = m2lv<double>(nd.rot, din);
  const double t = dot(mk(P.x - o.x, P.y - o.y, P.z - o.z), din);
  AsphAt at;
  const V3<double> g = asph_grad(p, surf, q, at);
  const double gn = sqrt(dot(g, g));
  const V3<double> n = scale(g, 1.0 / gn);
  const double k = dot(e, n);
  // P' = P - eps din (global)
  V3<double> qbar = m2lv<double>(nd.rot, Ppbar);
  V3<double> dinbar = scale(Ppbar, -eps);
  // e' = e - 2 k n
  const V3<double> dbl = m2lv<double>(nd.rot, doutbar);
  V3<double> ebar = dbl;
  const double kbar = -2.0 * dot(dbl, n);
  V3<double> nbar = scale(dbl, -2.0 * k);
  ebar = axpy(ebar, kbar, n);
  nbar = axpy(nbar, kbar, e);
  // n = g / |g|
  const V3<double> gbar = scale(axpy(nbar, -dot(nbar, n), n), 1.0 / gn);
  // g = (-a q_x, -a q_y, 1), a = a(w; theta), w = q_x^2 + q_y^2
  const double abar = -(gbar.x * q.x + gbar.y * q.y);
  const double wbar = abar * at.dadw;
  qbar.x += -at.a * gbar.x + 2.0 * q.x * wbar;
  qbar.y += -at.a * gbar.y + 2.0 * q.y * wbar;
  // q = ol + t e, t from G(q) = 0
  const double tbar = dot(qbar, e);
  const double ge = dot(g, e);
  V3<double> olbar = axpy(qbar, -tbar / ge, g);
  ebar = axpy(ebar, t, qbar);
  ebar = axpy(ebar, -tbar * t / ge, g);
  const double ts = tbar / ge;  // dL/dF through t
  atomicAdd(&acc[first], ts * at.sag_c + abar * at.a_c);
  atomicAdd(&acc[first + 1], ts * at.sag_k + abar * at.a_k);
  const int n_poly = int(p[8 + surf]);
  const double w = q.x * q.x + q.y * q.y;
  double wk1 = 1.0;
  for (int kk = 1; kk <= n_poly; kk++) {
    atomicAdd(&acc[first + 1 + kk], ts * wk1 * w + abar * 2.0 * kk * wk1);
    wk1 *= w;
  }
  Ppbar = l2mv<double>(nd.rot, olbar);
  doutbar = axpy(dinbar, 1.0, l2mv<double>(nd.rot, ebar));
}

// Accumulates hbar . dh/dparams of one focused ray into `acc` (LDS, FP64):
// the reverse pass seeded with the loss adjoint hbar of the hit's (x, y).
// The chain must be covered (chain_covered).
__device__ void adjoint_ray(const DScene& S, const AdjParams& A, const ChainRecord& rec, V3<double> origin,
                            V3<double> d0, V3<double> hit, V3<double> hbar, double* acc) {
  const int K = rec.n;
  const double eps = kEps + 1e-6;
  // the last segment: from the last step-back point along the last
  // reflection's outgoing direction (recomputed in FP64), or straight from
  // the origin for a ray that reflected nowhere
  int surf = 0;
  V3<double> dlast = d0;
  if (K > 0) {
    const int kind = reflection_kind(S, A, rec.node[K - 1], rec.P[K - 1], surf);
    dlast = reflected_dir(S, A, rec.node[K - 1], kind, surf, rec.P[K - 1], rec.din[K - 1]);
  }
  const V3<double> Pp = K > 0 ? axpy(rec.P[K - 1], -eps, rec.din[K - 1]) : origin;
  // h = P' + s d' on the focal surface G_f(q) = 0, q = M^T (h - T) - dz e_z
  // for a shift dz of the focal node along its z axis; with g_l = dG_f/dq and
  // g_F = M g_l, ds = -(g_F.(dP' + s dd') - g_l.z dz) / (g_F.d'). The plane
  // has g = e_z; a disk surface (-a q_x, -a q_y, 1); a disk wall (q_x, q_y, 0).
  V3<double> gl = mk(0.0, 0.0, 1.0), gF = gl;
  const int fk = focal_kind(S, A, hit, surf);
  if (fk != kAdjPlane) {
    const DNode& fn = S.nodes[A.focal_node];
    const V3<double> q = m2l<double>(fn.rot, fn.tr, hit);
    AsphAt at;
    gl = fk == kAdjAsph ? asph_grad(S.params + S.shapes[fn.shape].off, surf, q, at) : mk(q.x, q.y, 0.0);
    gF = l2mv<double>(fn.rot, gl);
  }
  const double s = dot(mk(hit.x - Pp.x, hit.y - Pp.y, hit.z - Pp.z), dlast);
  const double gd = dot(gF, dlast);
  const double sbar = dot(hbar, dlast);
  V3<double> Ppbar = axpy(hbar, -sbar / gd, gF);
  V3<double> doutbar = axpy(scale(hbar, s), -sbar * s / gd, gF);
  if (A.z_param >= 0) atomicAdd(&acc[A.z_param], sbar * gl.z / gd);

  for (int j = K - 1; j >= 0; j--) {
    const int node = rec.node[j];
    const DNode& nd = S.nodes[node];
    const V3<double> o = j == 0 ? origin : axpy(rec.P[j - 1], -eps, rec.din[j - 1]);
    if (reflection_kind(S, A, node, rec.P[j], surf) == kAdjAsph) {
      asph_reflection_adjoint(S, nd, surf, A.surface_param[2 * node], rec.P[j], rec.din[j], o, eps, Ppbar, doutbar,
                              acc);
      continue;
    }
    const int32_t* np = A.node_param + 3 * node;
    const double R = A.values[np[2]], tilt = A.values[np[1]];
    const V3<double> C = mk(nd.tr[0] + nd.rot[2] * R, nd.tr[1] + nd.rot[5] * R, nd.tr[2] + nd.rot[8] * R);
    const V3f din = f3(rec.din[j]);
    const V3f m = f3(mk(rec.P[j].x - C.x, rec.P[j].y - C.y, rec.P[j].z - C.z));  // P - C, exact in FP64
    const float Rf = float(R);
    const V3f n = mulf(m, 1.0f / Rf);
    const float k = dotf(din, n);
    const float t = float(dot(mk(rec.P[j].x - o.x, rec.P[j].y - o.y, rec.P[j].z - o.z), rec.din[j]));
    const V3f Ppbar_f = f3(Ppbar), doutbar_f = f3(doutbar);
    // step back P' = P - eps din
    V3f Pbar = Ppbar_f;
    V3f dinbar = mulf(Ppbar_f, -float(eps));
    // reflection d' = din - 2 k n
    dinbar = addf(dinbar, doutbar_f);
    const float kbar = -2.0f * dotf(doutbar_f, n);
    V3f nbar = mulf(doutbar_f, -2.0f * k);
    dinbar = addf(dinbar, mulf(n, kbar));
    nbar = addf(nbar, mulf(din, kbar));
    // facet frame: C = T + R ez, w = P - T = m + R ez (the hit relative to the
    // facet vertex, formed in FP64 so that ez + n = w / R keeps its precision)
    const V3f ex = {float(nd.rot[0]), float(nd.rot[3]), float(nd.rot[6])};
    const V3f ey = {float(nd.rot[1]), float(nd.rot[4]), float(nd.rot[7])};
    const V3f ez = {float(nd.rot[2]), float(nd.rot[5]), float(nd.rot[8])};
    const V3f w = f3(mk(rec.P[j].x - nd.tr[0], rec.P[j].y - nd.tr[1], rec.P[j].z - nd.tr[2]));
    // n = (P - C) / R: dn/dP = 1/R, dn/dC = -1/R, dn/dR (C = T + R ez) = -w / R^2
    Pbar = addf(Pbar, mulf(nbar, 1.0f / Rf));
    V3f Cbar = mulf(nbar, -1.0f / Rf);
    float Rbar = -dotf(nbar, w) / (Rf * Rf);
    // P = o + t din with t from |P - C| = R (implicit), m.din = R k:
    // dt = (m.dC - m.do - t m.ddin) / md, and through R: (w.ez) / md
    const float tbar = dotf(Pbar, din);
    V3f obar = Pbar;
    dinbar = addf(dinbar, mulf(Pbar, t));
    const float md = Rf * k;
    obar = addf(obar, mulf(m, -tbar / md));
    dinbar = addf(dinbar, mulf(m, -tbar * t / md));
    Cbar = addf(Cbar, mulf(m, tbar / md));
    Rbar += tbar * dotf(w, ez) / md;
    // tip and tilt move C at fixed R: dC/dtip = -R cos(tilt) ey, dC/dtilt = R ex
    atomicAdd(&acc[np[0]], double(-Rf * cosf(float(tilt)) * dotf(Cbar, ey)));
    atomicAdd(&acc[np[1]], double(Rf * dotf(Cbar, ex)));
    atomicAdd(&acc[np[2]], double(Rbar));
    // into the previous segment: its step-back point and outgoing direction
    Ppbar = mk(double(obar.x), double(obar.y), double(obar.z));
    doutbar = mk(double(dinbar.x), double(dinbar.y), double(dinbar.z));
  }
}

// One focused ray's contribution in the descriptor's mode: the centroid
// form seeds hbar = weight (h - c); the moments form seeds (h - c0), (1, 0)
// and (0, 1) into three consecutive n_params blocks and adds the ray's
// shifted moments (1, h_x - c0_x, h_y - c0_y, |h - c0|^2) after them; the
// encircled form seeds the derivative of weight * sigmoid((r0^2 - |u|^2) /
// (2 r0 w)) and adds the loss itself after the n_params entries.
__device__ void adjoint_accumulate(const DScene& S, const AdjParams& A, const ChainRecord& rec, V3<double> origin,
                                   V3<double> d0, V3<double> hit, int64_t ray, double* acc) {
  if (!chain_covered(S, A, rec)) return;
  double rx = A.cx, ry = A.cy;
  if (A.mode == RB_ADJ_MOMENTS) {
    const int64_t g = ray / A.group_size;
    acc += g * adj_block(A.mode, A.n_params);
    if (A.ref_xy) {
      rx = A.ref_xy[2 * g];
      ry = A.ref_xy[2 * g + 1];
    }
  }
  const double ux = hit.x - rx, uy = hit.y - ry;
  if (A.mode == RB_ADJ_CENTROID) {
    adjoint_ray(S, A, rec, origin, d0, hit, mk(A.weight * ux, A.weight * uy, 0.0), acc);
    return;
  }
  if (A.mode == RB_ADJ_ENCIRCLED) {
    const double u = (A.ee_r0 * A.ee_r0 - ux * ux - uy * uy) / (2.0 * A.ee_r0 * A.ee_w);
    const double sg = 1.0 / (1.0 + exp(-u));
    const double f = -A.weight * sg * (1.0 - sg) / (A.ee_r0 * A.ee_w);
    adjoint_ray(S, A, rec, origin, d0, hit, mk(f * ux, f * uy, 0.0), acc);
    atomicAdd(&acc[A.n_params], A.weight * sg);
    return;
  }
  const int P = A.n_params;
  adjoint_ray(S, A, rec, origin, d0, hit, mk(ux, uy, 0.0), acc);
  adjoint_ray(S, A, rec, origin, d0, hit, mk(1.0, 0.0, 0.0), acc + P);
  adjoint_ray(S, A, rec, origin, d0, hit, mk(0.0, 1.0, 0.0), acc + 2 * P);
  atomicAdd(&acc[3 * P], 1.0);
  atomicAdd(&acc[3 * P + 1], ux);
  atomicAdd(&acc[3 * P + 2], uy);
  atomicAdd(&acc[3 * P + 3], ux * ux + uy * uy);
}

}  // namespace rbd
