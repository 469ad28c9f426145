// rb_adjoint.h -- reverse-mode derivative of a focal-plane loss through a
// recomputed chain of spherical-facet reflections (device).
//
// Forward chain per reflection j (o_j = previous step-back point, or the ray
// origin; din_j its direction):
//   P_j  = o_j + t_j din_j,  |P_j - C_j| = R_j          (sphere crossing)
//   n_j  = (P_j - C_j) / R_j, k_j = din_j . n_j
//   d'_j = din_j - 2 k_j n_j                             (specular reflection)
//   P'_j = P_j - eps din_j                               (the 2e-6 cm step back)
// with facet centre C = T + M (0, 0, R), M = M0 Rx(tip) Ry(tilt), so
//   dC/dR = M e_z,  dC/dtip = -R cos(tilt) M e_y,  dC/dtilt = R M e_x.
// The chain ends on the focal plane z = z_cam: h = P' + s d', s = (z_cam - P'_z)/d'_z.
// The loss per focused ray is w |h - centroid|^2 / 2, so dL/dh = w (h - centroid).
// Per-ray arithmetic is FP32 on quantities formed relative to each facet in
// FP64; gradients are accumulated in FP64.
#pragma once

#include "rb_trace.h"

namespace rbd {

constexpr int kMaxAdjParams = 1024;

struct AdjParams {
  const int32_t* node_param;  // [n_nodes][3]
  const double* values;
  int32_t n_params, z_param, focal_node, mode;
  double cx, cy, weight;
  const double* ref_xy;       // moments: per-group reference points, or null
  int64_t group_size;         // moments: rays per group
  int32_t n_groups;           // moments: groups (>= 1)
};
constexpr int kMaxAdjAcc = 6400;  // LDS accumulator entries (FP64)
// accumulator entries per group, and in all
__host__ __device__ __forceinline__ int adj_block(int mode, int n_params) {
  return mode == 0 ? n_params : 3 * n_params + 4;
}
__host__ __device__ __forceinline__ int adj_acc_size(int mode, int n_params, int n_groups) {
  return mode == 0 ? n_params : n_groups * (3 * n_params + 4);
}

struct V3f {
  float x, y, z;
};
__device__ __forceinline__ V3f f3(V3<double> v) { return {float(v.x), float(v.y), float(v.z)}; }
__device__ __forceinline__ float dotf(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3f addf(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3f mulf(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// True if every reflection of the chain is on a facet with parameters.
// True if every reflection of the chain is on a node whose three
// parameter indices lie in [0, n_params): an index out of that range would
// read param_values and write the LDS accumulators out of bounds.
__device__ __forceinline__ bool chain_covered(const AdjParams& A, const ChainRecord& rec) {
  for (int j = 0; j < rec.n; j++) {
    const int32_t* np = A.node_param + 3 * rec.node[j];
    for (int k = 0; k < 3; k++)
      if (np[k] < 0 || np[k] >= A.n_params) return false;
  }
  return true;
}

// Accumulates hbar . dh/dparams of one focused ray into `acc` (LDS, FP64):
// the reverse pass seeded with the loss adjoint hbar of the hit's (x, y).
__device__ void adjoint_ray(const DScene& S, const AdjParams& A, const ChainRecord& rec, V3<double> origin,
                            V3<double> d0, V3<double> hit, V3f hbar, double* acc) {
  const int K = rec.n;
  // focal plane: reached from the last step-back point along the last
  // reflection's outgoing direction (recomputed in FP64), or straight from
  // the origin for a ray that reflected nowhere
  const double eps = kEps + 1e-6;
  V3<double> dlast = K > 0 ? rec.din[K - 1] : d0;
  if (K > 0) {
    const DNode& nd = S.nodes[rec.node[K - 1]];
    const double R = A.values[A.node_param[3 * rec.node[K - 1] + 2]];
    V3<double> C = mk(nd.tr[0] + nd.rot[2] * R, nd.tr[1] + nd.rot[5] * R, nd.tr[2] + nd.rot[8] * R);
    V3<double> m = mk(rec.P[K - 1].x - C.x, rec.P[K - 1].y - C.y, rec.P[K - 1].z - C.z);
    V3<double> n = scale(m, 1.0 / R);
    double k = dot(dlast, n);
    dlast = mk(dlast.x - 2 * k * n.x, dlast.y - 2 * k * n.y, dlast.z - 2 * k * n.z);
    double mag = sqrt(dot(dlast, dlast));
    dlast = scale(dlast, 1.0 / mag);
  }
  const V3<double> Pp = K > 0 ? axpy(rec.P[K - 1], -eps, rec.din[K - 1]) : origin;
  const float s = float((hit.z - Pp.z) / dlast.z);
  V3f dout = f3(dlast);
  // h = P' + s d', s = (z_cam - P'_z) / d'_z
  float sbar = dotf(hbar, dout);
  V3f Ppbar = hbar;
  V3f doutbar = mulf(hbar, s);
  float zbar = sbar / dout.z;
  Ppbar.z -= sbar / dout.z;
  doutbar.z -= sbar * s / dout.z;
  if (A.z_param >= 0) atomicAdd(&acc[A.z_param], double(zbar));

  for (int j = K - 1; j >= 0; j--) {
    const int node = rec.node[j];
    const DNode& nd = S.nodes[node];
    const int32_t* np = A.node_param + 3 * node;
    const double R = A.values[np[2]], tilt = A.values[np[1]];
    const V3<double> C = mk(nd.tr[0] + nd.rot[2] * R, nd.tr[1] + nd.rot[5] * R, nd.tr[2] + nd.rot[8] * R);
    const V3<double> o = j == 0 ? origin : axpy(rec.P[j - 1], -eps, rec.din[j - 1]);
    const V3f din = f3(rec.din[j]);
    const V3f m = f3(mk(rec.P[j].x - C.x, rec.P[j].y - C.y, rec.P[j].z - C.z));  // P - C, exact in FP64
    const float Rf = float(R);
    const V3f n = mulf(m, 1.0f / Rf);
    const float k = dotf(din, n);
    const float t = float(dot(mk(rec.P[j].x - o.x, rec.P[j].y - o.y, rec.P[j].z - o.z), rec.din[j]));
    // step back P' = P - eps din
    V3f Pbar = Ppbar;
    V3f dinbar = mulf(Ppbar, -float(eps));
    // reflection d' = din - 2 k n
    dinbar = addf(dinbar, doutbar);
    const float kbar = -2.0f * dotf(doutbar, n);
    V3f nbar = mulf(doutbar, -2.0f * k);
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
    Ppbar = obar;
    doutbar = dinbar;
  }
}

// One focused ray's contribution in the descriptor's mode: the centroid
// form seeds hbar = weight (h - c); the moments form seeds (h - c0), (1, 0)
// and (0, 1) into three consecutive n_params blocks and adds the ray's
// shifted moments (1, h_x - c0_x, h_y - c0_y, |h - c0|^2) after them.
__device__ void adjoint_accumulate(const DScene& S, const AdjParams& A, const ChainRecord& rec, V3<double> origin,
                                   V3<double> d0, V3<double> hit, int64_t ray, double* acc) {
  if (!chain_covered(A, rec)) return;
  double rx = A.cx, ry = A.cy;
  if (A.mode != 0) {
    const int64_t g = ray / A.group_size;
    acc += g * adj_block(A.mode, A.n_params);
    if (A.ref_xy) {
      rx = A.ref_xy[2 * g];
      ry = A.ref_xy[2 * g + 1];
    }
  }
  const double ux = hit.x - rx, uy = hit.y - ry;
  if (A.mode == 0) {
    adjoint_ray(S, A, rec, origin, d0, hit, V3f{float(A.weight * ux), float(A.weight * uy), 0.0f}, acc);
    return;
  }
  const int P = A.n_params;
  adjoint_ray(S, A, rec, origin, d0, hit, V3f{float(ux), float(uy), 0.0f}, acc);
  adjoint_ray(S, A, rec, origin, d0, hit, V3f{1.0f, 0.0f, 0.0f}, acc + P);
  adjoint_ray(S, A, rec, origin, d0, hit, V3f{0.0f, 1.0f, 0.0f}, acc + 2 * P);
  atomicAdd(&acc[3 * P], 1.0);
  atomicAdd(&acc[3 * P + 1], ux);
  atomicAdd(&acc[3 * P + 2], uy);
  atomicAdd(&acc[3 * P + 3], ux * ux + uy * uy);
}

}  // namespace rbd
