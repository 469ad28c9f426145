// rb_trace.h -- navigation and the per-ray optics state machine (device).
#pragma once

#include "rb_device.h"

namespace rbd {

#ifdef RB_HOST_EMULATION
inline long rb_emul_fallbacks = 0;
// host views are never staged (stage_n = 0)
inline DBvh bvh_node(const DScene& S, int i) { return S.bvh[i]; }
#else
// LDS copy of the top node's BVH (DScene::stage_lo / stage_n), written by
// stage_bvh() at the entry of every kernel that navigates.
__shared__ __attribute__((aligned(16))) DBvh g_bvh_stage[kBvhStage];  // read as float4 pairs

// BVH entry i, from the LDS copy when staged. The two reads stay in separate
// branches (and types) so that the LDS one is a ds_read, not a flat load
// through a selected generic pointer.
__device__ __forceinline__ DBvh bvh_node(const DScene& S, int i) {
  const unsigned k = unsigned(i - S.stage_lo);
  DBvh b;
  if (k < unsigned(S.stage_n)) {
    const float4* l = reinterpret_cast<const float4*>(&g_bvh_stage[k]);
    const float4 a = l[0], c = l[1];
    b = {{a.x, a.y, a.z}, __float_as_int(a.w), {c.x, c.y, c.z}, __float_as_int(c.w)};
  } else {
    b = S.bvh[i];
  }
  return b;
}
#endif

struct RayIO {
  double *x, *y, *z, *t, *dx, *dy, *dz;
  const double* lambda;
  int32_t *status, *npoints, *last_node;
  double* path_xyzt;
  int32_t* path_node;
  int32_t path_cap;
};

struct TraceParams {
  uint64_t seed, id_base;
  int32_t limit, fresnel;
};

// per-lane ray and navigator state
// Candidate leaves of one BVH traversal, in LDS: a lane's slot k lives at
// [k * kCandStride], so adjacent lanes hit adjacent banks.
#ifndef RB_POSTPONE
#define RB_POSTPONE 1
#endif
#ifndef RB_NCAND
#define RB_NCAND 8
#endif
constexpr int kCand = RB_NCAND;
constexpr int kCandStride = 256;  // the tracing kernels' block size
struct Cand {
  int32_t* node;
  float* tn;
};

struct Lane {
  Cand cand;          // this lane's candidate slots
  uint64_t ray;       // Philox counter id of the ray (ray_id_base + its index)
  V3<double> p, d;    // navigator point (== last recorded point between segments), unit direction
  double t, lambda;
  int32_t cur;        // current node, -1 = outside the top volume
  int32_t status, npoints, last_node;
  uint32_t draw;      // Philox draw counter
  bool on_b;          // TGeoNavigator fIsOnBoundary
  // crossing of the last step
  double step;
  int32_t hit_node;
  uint32_t hit_leaf;
};

// Reflection recorders: trace_lane reports every specular reflection
// (node, boundary point before the step back, incoming direction).
struct NoRecord {
  __device__ __forceinline__ void reflect(int, V3<double>, V3<double>) {}
  __device__ __forceinline__ void refract() {}
};
// The adjoint's recomputed bounce chain: up to kMax reflections (node,
// crossing point before the step back, incoming direction). A longer chain or
// a refraction marks the ray as outside the adjoint's model.
struct ChainRecord {
  static constexpr int kMax = 4;
  int n = 0;
  bool overflow = false;
  __device__ __forceinline__ void refract() { overflow = true; }
  int node[kMax];
  V3<double> P[kMax], din[kMax];
  __device__ __forceinline__ void reflect(int nd, V3<double> p, V3<double> d) {
    if (n < kMax) {
      node[n] = nd;
      P[n] = p;
      din[n] = d;
      n++;
    } else {
      overflow = true;
    }
  }
};

template <typename SC>
__device__ RB_NODE_FN bool node_contains(const SC& S, int n, V3<double> p) {
  const DNode& nd = S.nodes[n];
  return shape_contains<double>(S, nd.shape, m2l<double>(nd.rot, nd.tr, p));
}

__device__ __forceinline__ bool box_has(const DBvh& b, float x, float y, float z) {
  return x >= b.lo[0] && x <= b.hi[0] && y >= b.lo[1] && y <= b.hi[1] && z >= b.lo[2] && z <= b.hi[2];
}

// The neighbour cell of daughter `hint` of node n when the FP32 point (fx,
// fy, fz) lies in its cell box (the cell's record), else a record with
// prim < 0. The hint arrives from the ray's state: bounded before it is read.
template <typename SC>
__device__ __forceinline__ DBvh cell_at(const SC& S, int n, int hint, float fx, float fy, float fz) {
  DBvh none{};
  none.prim = -1;
  if (hint < 0 || hint >= S.n_nodes) return none;
  const DNode& h = S.nodes[hint];
  if (h.parent != n || h.cell < 0) return none;
  const DBvh c = S.cells[h.cell];
  return box_has(c, fx, fy, fz) ? c : none;
}

// First daughter of n (lowest node id, excluding skip) whose shape contains
// p; stackless walk of n's BVH (depth-first order with escape links). When
// p lies in the cell of daughter `hint`, only that cell's neighbours can
// contain it (their boxes are the ones that meet the cell), and they are
// tested instead of the walk: the same candidates, the same answer.
template <typename SC>
__device__ int first_daughter_containing(const SC& S, int n, V3<double> p, int skip, int hint = -1) {
  const DNode& nd = S.nodes[n];
  if (nd.n_children == 0) return -1;
  const float fx = float(p.x), fy = float(p.y), fz = float(p.z);
  int found = 0x7fffffff;
  const DBvh cell = cell_at(S, n, hint, fx, fy, fz);
  if (cell.prim >= 0) {
    for (int k = 0; k < cell.prim; k++) {
      const DBvh e = S.cell_nbrs[cell.escape + k];
      if (e.prim != skip && e.prim < found && box_has(e, fx, fy, fz) && node_contains(S, e.prim, p)) found = e.prim;
    }
    return found == 0x7fffffff ? -1 : found;
  }
  int i = nd.bvh_root;
  const int end = bvh_node(S, i).escape;
  while (i < end) {
    const DBvh b = bvh_node(S, i);
    if (!box_has(b, fx, fy, fz)) {
      i = b.escape;
      continue;
    }
    if (b.prim >= 0 && b.prim != skip && b.prim < found && node_contains(S, b.prim, p)) found = b.prim;
    i++;
  }
  return found == 0x7fffffff ? -1 : found;
}

// TGeoNavigator::SearchNode (up until inside, then down through daughters)
#ifndef RB_SEARCH_FN
#define RB_SEARCH_FN RB_NODE_FN
#endif
// hint: a daughter whose neighbour cell may hold p (the node the ray just
// left or reflected on), used at the level of its parent.
template <typename SC>
__device__ RB_SEARCH_FN int search_node(const SC& S, int start, V3<double> p, int skip, bool forced, int hint = -1) {
  int n = start;
  bool inside = forced || n == skip;
  while (!inside) {
    if (node_contains(S, n, p)) {
      inside = true;
    } else {
      int par = S.nodes[n].parent;
      if (par < 0) return -1;
      skip = n;
      n = par;
    }
  }
  for (int guard = 0; guard < 64; guard++) {
    int c = first_daughter_containing(S, n, p, skip, hint);
    if (c < 0) return n;
    n = c;
    skip = -1;
    hint = -1;
  }
  return n;
}

__device__ __forceinline__ double locate_push(const DScene& S, int n, double step) {
  const double* t = S.nodes[n].tr;
  return 100. * ((1. + fabs(t[0]) + fabs(t[1]) + fabs(t[2])) + step) * kTol;
}

// distance from a ray (origin p, unit d, FP64) to node c's shape, the
// search itself in precision T; the origin is first rebased to `base`
// along the ray (the node's box entry) so FP32 works near the node's frame
// lim: crossings at or beyond it may be reported as kBig (see Csg::dist_out)
template <typename T, typename SC>
__device__ RB_NODE_FN Hit node_dist(const SC& S, int c, V3<double> p, V3<double> d, double base, bool out,
                                    double lim = kBig) {
  const DNode& nd = S.nodes[c];
  V3<double> pb = axpy(p, base, d);
  V3<double> q = m2l<double>(nd.rot, nd.tr, pb), ld = m2lv<double>(nd.rot, d);
  V3<T> qt = mk<T>(T(q.x), T(q.y), T(q.z)), dt = mk<T>(T(ld.x), T(ld.y), T(ld.z));
  Hit h = shape_dist<T>(S, nd.shape, qt, dt, out, lim - base);
  if (h.s < kBig) h.s += base;
  return h;
}

// FP64 root polish of a crossing found in FP32 (leaf surface of node c)
template <typename SC>
__device__ RB_NODE_FN double node_polish(const SC& S, int c, const Hit& h, V3<double> p, V3<double> d) {
  const DNode& nd = S.nodes[c];
  V3<double> q = m2l<double>(nd.rot, nd.tr, p), ld = m2lv<double>(nd.rot, d);
  const double back = 2e-3 + 1e-6 * h.s;
  double s = shape_polish<double>(S, nd.shape, h.leaf, h.exiting != 0, q, ld, h.s, back);
  return (s < kBig && fabs(s - h.s) <= back) ? s : -1.0;
}

// Slab test of a BVH box (FP32, boxes rounded outward on the host): entry
// distance tn, and whether the box is crossed no farther than lim.
__device__ __forceinline__ bool box_pass(const DBvh& b, float ox, float oy, float oz, float ix, float iy, float iz,
                                         float lim, float& tn) {
  float t0x = (b.lo[0] - ox) * ix, t1x = (b.hi[0] - ox) * ix;
  float t0y = (b.lo[1] - oy) * iy, t1y = (b.hi[1] - oy) * iy;
  float t0z = (b.lo[2] - oz) * iz, t1z = (b.hi[2] - oz) * iz;
  tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.0f));
  float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  return tn <= tf * 1.0001f + 1e-3f && tn <= lim;
}

// Gather the leaves (daughter node ids) whose boxes the ray crosses no
// farther than lim, keeping the kCand nearest by box entry, sorted by entry
// distance. Returns the count kept; *cut receives the smallest entry
// distance among the boxes that did not fit (+inf if none was dropped).
__device__ __forceinline__ int gather_leaves(const DScene& S, int root, float ox, float oy, float oz, float ix,
                                             float iy, float iz, float lim, Cand cand, float* cut) {
  int i = root, nc = 0;
  float drop = 3.0e38f;
  const int end = bvh_node(S, i).escape;
  while (i < end) {
    const DBvh b = bvh_node(S, i);
    float tn;
    if (!box_pass(b, ox, oy, oz, ix, iy, iz, fminf(lim, drop), tn)) {
      i = b.escape;
      continue;
    }
    if (b.prim >= 0) {
      // insertion into the sorted slots; a full list drops its farthest
      int k = nc < kCand ? nc : kCand - 1;
      if (nc == kCand) {
        const float last = cand.tn[(kCand - 1) * kCandStride];
        if (tn >= last) {
          drop = fminf(drop, tn);
          i++;
          continue;
        }
        drop = fminf(drop, last);
      } else {
        nc++;
      }
      while (k > 0 && cand.tn[(k - 1) * kCandStride] > tn) {
        cand.tn[k * kCandStride] = cand.tn[(k - 1) * kCandStride];
        cand.node[k * kCandStride] = cand.node[(k - 1) * kCandStride];
        k--;
      }
      cand.tn[k * kCandStride] = tn;
      cand.node[k * kCandStride] = b.prim;
    }
    i++;
  }
  *cut = drop;
  return nc;
}

// Exit distance of the ray (p, d) from a cell box, or -1 when p is not
// inside it (FP64 against the box's FP32 bounds).
__device__ __forceinline__ double cell_exit(const DBvh& c, V3<double> p, V3<double> d) {
  const double pp[3] = {p.x, p.y, p.z}, dd[3] = {d.x, d.y, d.z};
  double t = kBig;
  for (int a = 0; a < 3; a++) {
    const double lo = double(c.lo[a]), hi = double(c.hi[a]);
    if (!(pp[a] >= lo && pp[a] <= hi)) return -1.0;
    if (dd[a] > 0.0) t = fmin(t, (hi - pp[a]) / dd[a]);
    else if (dd[a] < 0.0) t = fmin(t, (lo - pp[a]) / dd[a]);
  }
  return t;
}

// gather_leaves over a neighbour cell's records instead of a BVH walk.
template <typename SC>
__device__ __forceinline__ int gather_cell(const SC& S, const DBvh& cell, float ox, float oy, float oz, float ix,
                                           float iy, float iz, float lim, Cand cand, float* cut) {
  int nc = 0;
  float drop = 3.0e38f;
  for (int j = 0; j < cell.prim; j++) {
    const DBvh b = S.cell_nbrs[cell.escape + j];
    float tn;
    if (!box_pass(b, ox, oy, oz, ix, iy, iz, fminf(lim, drop), tn)) continue;
    int k = nc < kCand ? nc : kCand - 1;
    if (nc == kCand) {
      const float last = cand.tn[(kCand - 1) * kCandStride];
      if (tn >= last) {
        drop = fminf(drop, tn);
        continue;
      }
      drop = fminf(drop, last);
    } else {
      nc++;
    }
    while (k > 0 && cand.tn[(k - 1) * kCandStride] > tn) {
      cand.tn[k * kCandStride] = cand.tn[(k - 1) * kCandStride];
      cand.node[k * kCandStride] = cand.node[(k - 1) * kCandStride];
      k--;
    }
    cand.tn[k * kCandStride] = tn;
    cand.node[k * kCandStride] = b.prim;
  }
  *cut = drop;
  return nc;
}

// FP64 ranking rule: a daughter must beat the best by the tolerance; within
// the tolerance the lower node id wins, as the reference's loop over
// daughters in index order decides.
__device__ __forceinline__ void rank_fp64(int c, const Hit& h, double& best, int& crossed, Hit& hit) {
  if (h.s < best - kTol || (crossed >= 0 && c < crossed && h.s < best + kTol && h.s < kBig)) {
    best = h.s;
    crossed = c;
    hit = h;
  }
}

// Find the nearest boundary of the current node `cur` for the ray at p:
// its own exit (crossed = -1) or a daughter's entry (crossed = node id),
// all shape distances in FP64 (the plain form).
//
// The daughters are tested in two phases: the BVH walk only gathers the
// (up to kCand nearest) leaves whose boxes the ray crosses, then every lane
// tests its k-th candidate together, so the shape code runs with the wave's
// lanes in step instead of one leaf-holding lane at a time. Candidates are
// tested in order of box entry and the test stops at the first entry beyond
// the best crossing (a shape is never met before its box). If a box that
// did not fit could still hold a nearer crossing, the one-pass walk decides.
// The ranking rule makes the result independent of the test order.
//
// `hint` (the daughter the ray last interacted with, or -1) is ranked before
// the walk: a ray that keeps meeting one daughter (a light guide's wall) then
// starts the walk with that crossing as its bound, which prunes the boxes
// beyond it.
template <typename SC>
__device__ void search_boundary_fp64(const SC& S, int cur, V3<double> p, V3<double> d, double& best,
                                     int& crossed, Hit& hit, Cand cand, int hint) {
  const DNode& cn = S.nodes[cur];
  Hit hx = node_dist<double>(S, cur, p, d, 0.0, false);
  best = kBig;
  crossed = -2;
  if (hx.s < best - kTol) {
    best = hx.s;
    crossed = -1;
    hit = hx;
  }
  if (cn.n_children == 0) return;
  // (a ray's last node arrives from the caller: bounded before it is read)
  const bool hinted = hint > 0 && hint < S.n_nodes && S.nodes[hint].parent == cur;
  if (hinted) rank_fp64(hint, node_dist<double>(S, hint, p, d, 0.0, true, best + kTol), best, crossed, hit);
  const float ox = float(p.x), oy = float(p.y), oz = float(p.z);
  const float ix = 1.0f / float(d.x), iy = 1.0f / float(d.y), iz = 1.0f / float(d.z);
#if RB_POSTPONE
  float cut;
  if (hinted) {
    // The hint's neighbour cell: a daughter crossed at a point of the cell
    // box has a box that meets it, so while the best crossing lies inside
    // the box (the ray, convexly, stays inside up to its exit distance) the
    // cell's neighbours are a complete candidate set, and the walk is not
    // needed. Otherwise the walk below ranks every daughter again.
    const DBvh cell = cell_at(S, cur, hint, ox, oy, oz);
    if (cell.prim >= 0) {
      const double texit = cell_exit(cell, p, d);
      if (texit > 0.0) {
        const float lim = float(fmin(best, texit)) * 1.0001f + 1e-2f;
        const int nl = gather_cell(S, cell, ox, oy, oz, ix, iy, iz, lim, cand, &cut);
        for (int k = 0; k < nl; k++) {
          if (cand.tn[k * kCandStride] > float(best) * 1.0001f + 1e-2f) break;
          const int c = cand.node[k * kCandStride];
          if (c == hint) continue;
          rank_fp64(c, node_dist<double>(S, c, p, d, 0.0, true, best + kTol), best, crossed, hit);
        }
        if (best + 1e-6 * (1.0 + best) < texit && !(cut <= float(best) * 1.0001f + 1e-2f)) return;
      }
    }
  }
  const int nc = gather_leaves(S, cn.bvh_root, ox, oy, oz, ix, iy, iz, float(best) * 1.0001f + 1e-2f, cand, &cut);
  for (int k = 0; k < nc; k++) {
    if (cand.tn[k * kCandStride] > float(best) * 1.0001f + 1e-2f) break;
    const int c = cand.node[k * kCandStride];
    if (hinted && c == hint) continue;
    rank_fp64(c, node_dist<double>(S, c, p, d, 0.0, true, best + kTol), best, crossed, hit);
  }
  if (!(cut <= float(best) * 1.0001f + 1e-2f)) return;
  // a dropped box may hold a nearer crossing: rank all daughters afresh
  best = kBig;
  crossed = -2;
  if (hx.s < best - kTol) {
    best = hx.s;
    crossed = -1;
    hit = hx;
  }
#endif
  int i = cn.bvh_root;
  const int end = bvh_node(S, i).escape;
  while (i < end) {
    const DBvh b = bvh_node(S, i);
    float tn;
    if (!box_pass(b, ox, oy, oz, ix, iy, iz, float(best) * 1.0001f + 1e-2f, tn)) {
      i = b.escape;
      continue;
    }
    if (b.prim >= 0) rank_fp64(b.prim, node_dist<double>(S, b.prim, p, d, 0.0, true, best + kTol), best, crossed, hit);
    i++;
  }
}

// Mixed form of the same search. Composite shapes are always measured in
// FP64: their boolean iteration steps past each boundary by an inside-test
// offset, and the FP32 offset (2 um) exceeds features such as a 1 um mirror
// shell, which FP32 would step straight over. Primitives are ranked with FP32 shape
// distances in frames rebased to each daughter's box entry; a candidate
// within kNearF32 of the ray origin (the surface just left, or one the ray
// sits on) is resolved in FP64 on the spot, since FP32 cannot tell which
// side of it the ray is. The winner is then polished in FP64 and checked
// with an FP64 inside test; near-ties and failed checks redo the search
// in FP64. Returns false when the caller must run the FP64 search.
constexpr double kNearF32 = 2e-2;

// One daughter of the mixed search: FP32 distance from its box entry, FP64
// when it is within kNearF32 of the origin; keeps the best and second best.
__device__ __forceinline__ bool fp64_only(const DScene& S, int node) {
  return S.shapes[S.nodes[node].shape].depth > 0;
}
template <typename SC>
__device__ __forceinline__ void rank_mixed(const SC& S, int c, float tn, V3<double> p, V3<double> d,
                                           double& best, double& second, int& crossed, Hit& hit, bool& exact) {
  Hit h;
  bool hexact = fp64_only(S, c);
  if (!hexact) {
    const double base = fmax(0.0, double(tn) * (1.0 - 1e-6) - 1.0);
    h = node_dist<float>(S, c, p, d, base, true);
  }
  if (hexact || h.s < kNearF32) {
    h = node_dist<double>(S, c, p, d, 0.0, true);
    hexact = true;
  }
  if (h.s < best) {
    second = best;
    best = h.s;
    crossed = c;
    hit = h;
    exact = hexact;
  } else if (h.s < second) {
    second = h.s;
  }
}

template <typename SC>
__device__ bool search_boundary_mixed(const SC& S, int cur, V3<double> p, V3<double> d, double& best,
                                      int& crossed, Hit& hit, Cand cand) {
  const DNode& cn = S.nodes[cur];
  best = kBig;
  crossed = -2;
  bool exact = false;
  double second = kBig;
  Hit hx;
  bool hx_exact = fp64_only(S, cur);
  if (!hx_exact) hx = node_dist<float>(S, cur, p, d, 0.0, false);
  if (hx_exact || hx.s < kNearF32) {
    hx = node_dist<double>(S, cur, p, d, 0.0, false);
    hx_exact = true;
  }
  if (hx.s < kBig) {
    best = hx.s;
    crossed = -1;
    hit = hx;
    exact = hx_exact;
  }
  if (cn.n_children > 0) {
    const float ox = float(p.x), oy = float(p.y), oz = float(p.z);
    const float ix = 1.0f / float(d.x), iy = 1.0f / float(d.y), iz = 1.0f / float(d.z);
    bool walk = true;
#if RB_POSTPONE
    float cut;
    const int nc =
        gather_leaves(S, cn.bvh_root, ox, oy, oz, ix, iy, iz, float(fmin(best, second)) * 1.0001f + 1e-2f, cand, &cut);
    for (int k = 0; k < nc; k++) {
      const float tn = cand.tn[k * kCandStride];
      if (tn > float(fmin(best, second)) * 1.0001f + 1e-2f) break;
      rank_mixed(S, cand.node[k * kCandStride], tn, p, d, best, second, crossed, hit, exact);
    }
    walk = cut <= float(fmin(best, second)) * 1.0001f + 1e-2f;
    if (walk) {  // a dropped box may hold a nearer crossing: rank all daughters afresh
      best = kBig;
      second = kBig;
      crossed = -2;
      exact = false;
      if (hx.s < kBig) {
        best = hx.s;
        crossed = -1;
        hit = hx;
        exact = hx_exact;
      }
    }
#endif
    if (walk) {
      int i = cn.bvh_root;
      const int end = bvh_node(S, i).escape;
      while (i < end) {
        const DBvh b = bvh_node(S, i);
        float tn;
        if (!box_pass(b, ox, oy, oz, ix, iy, iz, float(fmin(best, second)) * 1.0001f + 1e-2f, tn)) {
          i = b.escape;
          continue;
        }
        if (b.prim >= 0) rank_mixed(S, b.prim, tn, p, d, best, second, crossed, hit, exact);
        i++;
      }
    }
  }
  if (crossed == -2) return true;
  const double margin = 1e-2 + 1e-6 * best;
  if (second < best + margin) return false;  // near-tie: rank in FP64
  if (exact) return true;
  const int node = crossed == -1 ? cur : crossed;
  double s = node_polish(S, node, hit, p, d);
  if (s < 0.0) return false;
  // the polished crossing must really enter (leave) the node in FP64
  const bool inside = node_contains(S, node, axpy(p, s + kBoolPush, d));
  if (inside != (crossed != -1)) return false;
  best = s;
  hit.s = s;
  return true;
}

// TGeoNavigator::FindNextBoundaryAndStep; returns the next node (-1 = null).
// MIXED: the search runs in FP32 and the winning crossing is polished in FP64.
// HINT: the FP64 search ranks the daughter of the last interaction first,
// and searches and relocates within the neighbour cells of the daughters it
// leaves or reflects on (the results do not change; the work-queue kernels,
// whose rays bounce many times inside light guides, compile it in).
template <bool MIXED, bool HINT = false, typename SC>
__device__ int next_boundary(const SC& S, Lane& L) {
  double extra = L.on_b ? kTol : 0.0;
  L.on_b = false;
  L.p = axpy(L.p, extra, L.d);
  L.hit_node = -1;
  L.hit_leaf = 0;
  if (L.cur < 0) {
    Hit h = node_dist<double>(S, 0, L.p, L.d, 0.0, true);
    if (h.s < kBig - kTol) {
      L.step = h.s + extra;
      L.p = axpy(L.p, h.s, L.d);
      L.on_b = true;
      L.hit_node = 0;
      L.hit_leaf = h.leaf;
      V3<double> pp = axpy(L.p, locate_push(S, 0, L.step), L.d);
      L.cur = search_node(S, 0, pp, -1, true);
      return L.cur;
    }
    L.step = kBig;
    return -1;
  }
  const int cur = L.cur;
  const DNode& cn = S.nodes[cur];
  double best;
  int crossed;
  Hit hit = {kBig, 0u, 0u};
  if (!MIXED || !search_boundary_mixed(S, cur, L.p, L.d, best, crossed, hit, L.cand)) {
#ifdef RB_HOST_EMULATION
    if (MIXED) rb_emul_fallbacks++;
#endif
    search_boundary_fp64(S, cur, L.p, L.d, best, crossed, hit, L.cand, HINT ? L.last_node : -1);
  }
  if (crossed == -1 && best <= kTol) {
    // TGeoNavigator: do not stay on the exit surface; step by the tolerance
    L.step = kTol;
    L.p = axpy(L.p, kTol, L.d);
    L.on_b = true;
    L.hit_node = cur;
    L.hit_leaf = hit.leaf;
    V3<double> pp = axpy(L.p, locate_push(S, cur, L.step), L.d);
    L.cur = search_node(S, cur, pp, cur, false);
    return L.cur;
  }
  L.p = axpy(L.p, best, L.d);
  L.step = best + extra;
  if (crossed == -2) return cur;
  L.on_b = true;
  L.hit_leaf = hit.leaf;
  if (crossed == -1) {
    L.hit_node = cur;
    if (cn.parent < 0) {
      L.cur = -1;
      return -1;
    }
    V3<double> pp = axpy(L.p, locate_push(S, cn.parent, L.step), L.d);
    L.cur = search_node(S, cn.parent, pp, cur, false, HINT ? cur : -1);
    return L.cur;
  }
  L.hit_node = crossed;
  V3<double> pp = axpy(L.p, locate_push(S, crossed, L.step), L.d);
  L.cur = search_node(S, crossed, pp, cur, true);
  return L.cur;
}

template <typename SC>
__device__ V3<double> find_normal(const SC& S, const Lane& L) {
  if (L.hit_node < 0) return mk(0.0, 0.0, 1.0);
  const DNode& nd = S.nodes[L.hit_node];
  V3<double> lp = m2l<double>(nd.rot, nd.tr, L.p), ld = m2lv<double>(nd.rot, L.d);
  V3<double> ln = shape_normal<double>(S, nd.shape, L.hit_leaf, lp, ld);
  return l2mv<double>(nd.rot, ln);
}

__device__ __forceinline__ int kind_of(const DScene& S, int n) { return n < 0 ? RB_KNULL : S.nodes[n].kind; }
__device__ __forceinline__ const rb_volume_desc& vol_of(const DScene& S, int n) {
  return S.volumes[S.nodes[n].volume];
}

__device__ __forceinline__ double draw(const TraceParams& P, uint64_t id, Lane& L) {
  return uniform01(P.seed, id, L.draw++, 0u);
}

// ---- border-surface conditions (ABorderSurfaceCondition)
// AOpticalComponent::FindBorderSurfaceCondition (src/AOpticalComponent.cxx:51-65):
// the first condition of node `cur`'s volume whose second component is node
// `next`'s volume (-1: the null node). Kernels built for scenes without
// conditions compile none of this.
template <typename SC>
__device__ __forceinline__ const rb_border_desc* find_border(const SC& S, int cur, int next) {
  if constexpr (!SC::borders_on) {
    return nullptr;
  } else {
    if (cur < 0) return nullptr;
    const rb_volume_desc& v = vol_of(S, cur);
    const int v2 = next < 0 ? -1 : S.nodes[next].volume;
    for (int k = 0; k < v.border_count; k++)
      if (S.borders[v.border_first + k].volume2 == v2) return &S.borders[v.border_first + k];
    return nullptr;
  }
}

// The draws of one stochastic surface sample are bounded: a rejection loop
// that has not accepted after kMaxSurfaceDraws uniforms keeps the smooth
// normal (the reference's loops are unbounded; the cap only guarantees that
// every lane ends).
constexpr uint32_t kMaxSurfaceDraws = 1u << 16;
struct SurfaceSample {
  V3<double> v;
  uint32_t draw;
};

// TVector3::RotateUz(u) of v (u a unit vector): the frame whose z axis is u.
__device__ inline V3<double> rotate_uz(V3<double> v, V3<double> u) {
  double up = u.x * u.x + u.y * u.y;
  if (up != 0.) {
    up = sqrt(up);
    return mk((u.x * u.z * v.x - u.y * v.y + u.x * up * v.z) / up, (u.y * u.z * v.x + u.x * v.y + u.y * up * v.z) / up,
              (u.z * u.z * v.x - v.x + u.z * up * v.z) / up);
  }
  if (u.z < 0.) return mk(-v.x, v.y, -v.z);
  return v;
}

// GetFacetNormal's roughness branch (src/AOpticsManager.cxx:266-298, after
// Geant4's G4OpBoundaryProcess::GetFacetNormal): alpha ~ Gaus(0, sigma),
// accepted with probability sin(alpha) / min(1, 4 sigma) below pi/2; a
// uniform azimuth; the tilted normal about n, drawn again until it faces the
// ray (d . n' > 0). Gaus is Box-Muller on two Philox uniforms.
__device__ RB_COLD_FN SurfaceSample rough_normal(uint64_t seed, uint64_t id, uint32_t draw, V3<double> n, V3<double> d,
                                                 double sigma) {
  const uint32_t stop = draw + kMaxSurfaceDraws;
  const double f_max = fmin(1., 4. * sigma);
  while (draw < stop) {
    double alpha = 0.;
    bool ok = false;
    while (draw < stop) {
      const double u1 = uniform01(seed, id, draw++, 0u), u2 = uniform01(seed, id, draw++, 0u);
      alpha = sigma * (sqrt(-2. * log(u1)) * cos(2. * kPi * u2));
      if (!(f_max * uniform01(seed, id, draw++, 0u) > sin(alpha) || alpha >= kPi / 2)) {
        ok = true;
        break;
      }
    }
    if (!ok) break;
    const double phi = 2. * kPi * uniform01(seed, id, draw++, 0u);
    const double sa = sin(alpha), ca = cos(alpha), sp = sin(phi), cp = cos(phi);
    const V3<double> f = rotate_uz(mk(sa * cp, sa * sp, ca), n);
    if (dot(d, f) > 0.0) return {f, draw};
  }
  return {n, draw};
}

// DoReflection's Lambertian branch (src/AOpticsManager.cxx:209-224): theta =
// asin(sqrt(2 y)), y ~ U(0, 1/2), phi ~ U(0, 2 pi), turned by
// TGeoRotation(phi_n + 90, theta_n + 180, 0) of the normal's angles (degrees).
__device__ RB_COLD_FN SurfaceSample lambert_dir(uint64_t seed, uint64_t id, uint32_t draw, V3<double> n) {
  const double y = 0.5 * uniform01(seed, id, draw++, 0u);
  const double theta = asin(sqrt(2 * y));
  const double phi = 2. * kPi * uniform01(seed, id, draw++, 0u);
  const double rad2deg = 180. / kPi, deg2rad = kPi / 180.;
  const double theta_n = (n.x == 0. && n.y == 0. && n.z == 0. ? 0. : atan2(sqrt(n.x * n.x + n.y * n.y), n.z)) * rad2deg;
  const double phi_n = (n.x == 0. && n.y == 0. ? 0. : atan2(n.y, n.x)) * rad2deg;
  const double sph = sin(deg2rad * (phi_n + 90)), cph = cos(deg2rad * (phi_n + 90));
  const double sth = sin(deg2rad * (theta_n + 180)), cth = cos(deg2rad * (theta_n + 180));
  const V3<double> v = mk(sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta));
  // TGeoRotation::SetAngles(phi, theta, psi = 0), LocalToMaster
  return {mk(cph * v.x - cth * sph * v.y + sth * sph * v.z, sph * v.x + cth * cph * v.y - sth * cph * v.z,
             sth * v.y + cth * v.z),
          draw};
}

// AOpticsManager::GetFacetNormal (src/AOpticsManager.cxx:250-301): the
// navigator's normal, tilted by the border condition's roughness.
template <typename SC>
__device__ __forceinline__ V3<double> facet_normal(const SC& S, const TraceParams& P, Lane& L,
                                                   const rb_border_desc* bc) {
  V3<double> n = find_normal(S, L);
  if constexpr (SC::borders_on) {
    if (bc && !bc->lambertian && bc->sigma != 0.) {
      const SurfaceSample r = rough_normal(P.seed, L.ray, L.draw, n, L.d, bc->sigma);
      L.draw = r.draw;
      n = r.v;
    }
  }
  return n;
}

// TH2::Interpolate(x, y) over a multilayer table (AMultilayer::CoherentTMMMixed
// after PreCalculateCoherentTMM, include/AMultilayer.h:110-114, 243-262):
// bilinear between the four nearest bin centres, the neighbour bin held
// inside the axis, 0 outside either axis. Restated from ROOT's documented
// behaviour; no FMA contraction, so the oracle's plain evaluation agrees bit
// for bit.
struct Th2Axis {
  int b, i1, i2;
  double c1, c2;
};
__device__ inline Th2Axis th2_axis(double v, int n, double lo, double hi) {
#pragma clang fp contract(off)
  const double w = (hi - lo) / n;
  int b = v < lo ? 0 : !(v < hi) ? n + 1 : 1 + int(n * (v - lo) / (hi - lo));
  b = b > n + 1 ? n + 1 : b;
  const double up = lo + b * w;
  const int b1 = (up - v) > w / 2 ? b - 1 : b;
  Th2Axis a;
  a.b = b;
  a.c1 = lo + (b1 - 0.5) * w;
  a.c2 = lo + (b1 + 0.5) * w;
  a.i1 = b1 < 1 ? 1 : b1 > n ? n : b1;
  a.i2 = b1 + 1 < 1 ? 1 : b1 + 1 > n ? n : b1 + 1;
  return a;
}
__device__ RB_COLD_FN double th2_interpolate(const double* vals, const rb_grid2d_desc& g, double x, double y) {
#pragma clang fp contract(off)
  const Th2Axis ax = th2_axis(x, g.nx, g.xlo, g.xhi), ay = th2_axis(y, g.ny, g.ylo, g.yhi);
  if (ax.b < 1 || ax.b > g.nx || ay.b < 1 || ay.b > g.ny) return 0.0;
  const double q11 = vals[int64_t(ay.i1 - 1) * g.nx + ax.i1 - 1], q21 = vals[int64_t(ay.i1 - 1) * g.nx + ax.i2 - 1];
  const double q12 = vals[int64_t(ay.i2 - 1) * g.nx + ax.i1 - 1], q22 = vals[int64_t(ay.i2 - 1) * g.nx + ax.i2 - 1];
  const double d = (ax.c2 - ax.c1) * (ay.c2 - ay.c1);
  return q11 / d * (ax.c2 - x) * (ay.c2 - y) + q21 / d * (x - ax.c1) * (ay.c2 - y) +
         q12 / d * (ax.c2 - x) * (y - ay.c1) + q22 / d * (x - ax.c1) * (y - ay.c1);
}
// Reflectance and transmittance of border condition bc's multilayer at the
// ray's wavelength and incidence angle acos(cos1).
template <typename SC>
__device__ inline void multilayer_rt(const SC& S, const rb_border_desc* bc, double lambda, double cos1, double& R,
                                     double& T) {
  const rb_grid2d_desc& g = S.grids[bc->multilayer];
  const double angle = cold_acos(cos1);
  R = th2_interpolate(S.params + g.r_offset, g, lambda, angle);
  T = th2_interpolate(S.params + g.t_offset, g, lambda, angle);
}

// The path record's slot k holds the (k+1)-th point this trace adds; the
// ray's point count at entry is still in io.npoints[i] (written back only
// when the trace ends), so it is read there rather than kept in a register.
__device__ __forceinline__ void add_point(const RayIO& io, int64_t i, Lane& L, V3<double> x, double t, int node) {
  if (io.path_cap > 0) {
    const int k = L.npoints - io.npoints[i];
    if (k < io.path_cap) {
      double* s = io.path_xyzt + (int64_t(io.path_cap) * i + k) * 4;
      s[0] = x.x; s[1] = x.y; s[2] = x.z; s[3] = t;
      io.path_node[int64_t(io.path_cap) * i + k] = node;
    }
  }
  L.t = t;
  L.last_node = node;
  L.npoints++;
}

__device__ __forceinline__ void set_direction(Lane& L, V3<double> v) {
  double mag = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (mag > 0) L.d = mk(v.x / mag, v.y / mag, v.z / mag);
}

// AOpticsManager::DoReflection (src/AOpticsManager.cxx:170-247). CELLS: the
// relocation after the step back tests the neighbour cell of the node
// reflected on (the work-queue kernels compile it in).
template <bool CELLS = false, typename Rec, typename SC>
__device__ void do_reflection(const SC& S, const TraceParams& P, const RayIO& io, int64_t i, uint64_t id,
                              Lane& L, double n1, int next, V3<double> n, Rec& rec, const rb_border_desc* bc) {
  double step = L.step;
  V3<double> d1 = L.d;
  double cos1 = dot(d1, n);
  bool absorbed = false;
  if (kind_of(S, next) == RB_KMIRROR) {
    const rb_volume_desc& v = vol_of(S, next);
    double ref = v.refl_table >= 0 ? table_eval(S, v.refl_table, L.lambda) : v.refl_const;
    ref = ref > 1 ? 1 : ref;
    ref = ref < 0 ? 0 : ref;
    if constexpr (SC::borders_on) {
      if (v.refl_grid >= 0) {  // AMirror::GetReflectance's TH2 (src/AMirror.cxx:46-49), clamped
        const rb_grid2d_desc& g = S.grids[v.refl_grid];
        ref = th2_interpolate(S.params + g.r_offset, g, L.lambda, cold_acos(cos1));
        ref = ref > 1 ? 1 : ref;
        ref = ref < 0 ? 0 : ref;
      }
      if (bc && bc->multilayer >= 0) {  // the coating's reflectance, unclamped (:194-198)
        double T;
        multilayer_rt(S, bc, L.lambda, cos1, ref, T);
      }
    }
    if (ref < draw(P, id, L)) {
      absorbed = true;
      L.status = RB_ABSORB;
    }
  }
  V3<double> d2 = mk(d1.x - 2 * n.x * cos1, d1.y - 2 * n.y * cos1, d1.z - 2 * n.z * cos1);
  if constexpr (SC::borders_on) {
    if (bc && bc->lambertian) {  // drawn even for an absorbed ray, as the reference does
      const SurfaceSample l = lambert_dir(P.seed, id, L.draw, n);
      L.draw = l.draw;
      d2 = l.v;
    }
  }
  if (!absorbed) {
    set_direction(L, d2);
    rec.reflect(next, L.p, d1);
  }
  double t = L.t + step / (kC / n1);
  // TGeoNavigator::Step() of kEpsilon backwards moves kEpsilon + 1e-6, then relocates
  L.p = axpy(L.p, -(kEps + 1e-6), d1);
  L.cur = L.cur < 0 ? (node_contains(S, 0, L.p) ? search_node(S, 0, L.p, -1, true) : -1)
                    : search_node(S, L.cur, L.p, -1, false, CELLS ? next : -1);
  L.on_b = true;
  add_point(io, i, L, L.p, t, next);
}

// AOpticsManager::DoFresnel (src/AOpticsManager.cxx:52-167). Refracts the
// ray, or returns true when it is reflected instead (total internal
// reflection, the Fresnel draw, or a multilayer's reflectance): the caller
// then runs DoReflection with the same normal, so each is inlined once. A
// multilayer condition (:79-96) replaces both the TIR test and the Fresnel
// draw: one uniform against R, R + T; above R + T the ray is transmitted
// and absorbed at the surface.
template <typename Rec, typename SC>
__device__ bool do_fresnel(const SC& S, const TraceParams& P, const RayIO& io, int64_t i, uint64_t id, Lane& L,
                           double n1, double n2, double k2, int next, V3<double> n, Rec& rec,
                           const rb_border_desc* bc) {
  double step = L.step;
  V3<double> d1 = L.d;
  double cos1 = dot(d1, n);
  double sin1 = sqrt(1 - cos1 * cos1);
  double sin2 = n1 * sin1 / n2;
  double cos2 = sqrt(1 - sin2 * sin2);
  bool coated = false, absorbed = false;
  if constexpr (SC::borders_on) {
    if (bc && bc->multilayer >= 0) {
      double R, T;
      multilayer_rt(S, bc, L.lambda, cos1, R, T);
      const double u = draw(P, id, L);
      if (u < R) return true;
      coated = true;
      absorbed = !(u < R + T);
    }
  }
  if (!coated && sin2 > 1.) return true;
  if (!coated && P.fresnel) {
    double Rs, Rp;
    if (k2 <= 0.) {
      double eta1S = n1 * cos1, eta2S = n2 * cos2, eta1P = n1 / cos1, eta2P = n2 / cos2;
      double a = (eta1S - eta2S) / (eta1S + eta2S), b = (eta1P - eta2P) / (eta1P + eta2P);
      Rs = a * a;
      Rp = b * b;
    } else {
      double x1S = n1 * cos1, x1P = n1 / cos1;
      double u = n2 * n2 - k2 * k2 - (n1 * sin1) * (n1 * sin1);
      double v = 2 * n2 * k2;
      double tmp = sqrt(u * u + v * v);
      double cosxi2 = sqrt(1 + u / tmp) / sqrt(2.), sinxi2 = sqrt(1 - u / tmp) / sqrt(2.);
      double x2S = sqrt(tmp) * cosxi2, y2S = sqrt(tmp) * sinxi2;
      tmp = x2S * x2S + y2S * y2S;
      double nk = n2 * n2 - k2 * k2;
      double x2P = (2 * n2 * k2 * y2S + nk * x2S) / tmp;
      double y2P = (2 * n2 * k2 * x2S - nk * y2S) / tmp;
      Rs = ((x1S - x2S) * (x1S - x2S) + y2S * y2S) / ((x1S + x2S) * (x1S + x2S) + y2S * y2S);
      Rp = ((x1P - x2P) * (x1P - x2P) + y2P * y2P) / ((x1P + x2P) * (x1P + x2P) + y2P * y2P);
    }
    double R = (Rs + Rp) / 2.;
    if (draw(P, id, L) < R) return true;
  }
  V3<double> d2 = d1;
  if (sin1 != 0) {
    double f = sin2 / sin1;
    d2 = mk((d1.x - cos1 * n.x) * f + n.x * cos2, (d1.y - cos1 * n.y) * f + n.y * cos2,
            (d1.z - cos1 * n.z) * f + n.z * cos2);
  }
  double t = L.t + step / (kC / n1);
  add_point(io, i, L, L.p, t, next);
  if (absorbed) {
    L.status = RB_ABSORB;
  } else {
    set_direction(L, d2);  // beyond the critical angle (a coated surface only) d2 is NaN and is ignored
    rec.refract();
  }
  return false;
}

// AOpticsManager::TraceNonSequential(TObjArray*) for one ray (src/AOpticsManager.cxx:353-518)
// Locate the ray's starting point (TGeoNavigator::InitTrack).
template <typename SC>
__device__ __forceinline__ void begin_lane(const SC& S, Lane& L) {
  L.on_b = false;
  L.cur = node_contains(S, 0, L.p) ? search_node(S, 0, L.p, -1, true) : -1;
}

// One segment of the per-ray state machine: next boundary, then the
// interaction the (current, next) volume kinds select.
template <bool MIXED, bool HINT = false, typename Rec, typename SC>
__device__ void step_lane(const SC& S, const TraceParams& P, const RayIO& io, int64_t i, Lane& L, Rec& rec) {
  const uint64_t id = L.ray;
  const double lambda = L.lambda;
  {
    const V3<double> x1 = L.p;
    const double t1 = L.t;
    const int current = L.cur;
    const int np0 = L.npoints;
    const int next = next_boundary<MIXED, HINT>(S, L);
    const double step = L.step;
    const int tc = kind_of(S, current), tn = kind_of(S, next);

    if (SC::lens_on && tc == RB_KLENS) {
      const rb_volume_desc& v = vol_of(S, current);
      double k = lens_k(S, v, lambda);
      double absl = k <= 0 ? 0.0 : lambda / (4 * kPi * k);
      if (k > 0 && absl > 0 && isfinite(absl)) {
        double abs_step = -absl * cold_log(draw(P, id, L));
        if (abs_step < step) {
          double n1 = lens_index(S, v, lambda);
          L.p = axpy(x1, abs_step, L.d);
          add_point(io, i, L, L.p, t1 + abs_step / (kC / n1), next);
          L.status = RB_ABSORB;
          return;
        }
      }
    }

    // (current, next) transition table; the mirror and the Fresnel cases
    // share one normal and one DoReflection
    const bool c_free = tc == RB_KNULL || tc == RB_KOPT || tc == RB_KOTHER;
    // (scenes traced without the lens physics have no RB_KLENS node)
    const bool c_lens = SC::lens_on && tc == RB_KLENS;
    const bool n_free = tn == RB_KNULL || tn == RB_KOPT || tn == RB_KOTHER;
    const bool mirror = (c_free || c_lens) && tn == RB_KMIRROR;
    const bool fresnel = SC::lens_on && ((c_free && tn == RB_KLENS) || (c_lens && (tn == RB_KLENS || n_free)));
    if (mirror || fresnel) {
      const double n1 = c_lens ? lens_index(S, vol_of(S, current), lambda) : 1.;
      double n2 = 1., k2 = 0.;
      if (SC::lens_on && tn == RB_KLENS) {
        const rb_volume_desc& v = vol_of(S, next);
        n2 = lens_index(S, v, lambda);
        k2 = lens_k(S, v, lambda);
      }
      const rb_border_desc* bc = find_border(S, current, next);
      const V3<double> n = facet_normal(S, P, L, bc);
      if (mirror || do_fresnel(S, P, io, i, id, L, n1, n2, k2, next, n, rec, bc))
        do_reflection<HINT>(S, P, io, i, id, L, n1, next, n, rec, bc);
    } else if ((c_free || c_lens) && (tn == RB_KOBS || tn == RB_KFOCUS)) {
      double n1 = c_lens ? lens_index(S, vol_of(S, current), lambda) : 1.;
      add_point(io, i, L, L.p, t1 + step / (kC / n1), next);
    } else if (c_free && (tn == RB_KOTHER || tn == RB_KOPT)) {
      add_point(io, i, L, L.p, t1 + step / kC, next);
    }

    if (tn == RB_KNULL) {
      add_point(io, i, L, L.p, t1 + step / kC, next);
      L.status = RB_EXIT;
    } else if (tc == RB_KFOCUS || tc == RB_KOBS || tc == RB_KMIRROR || tn == RB_KOBS) {
      // ray->Stop() without AddPoint: the ray's last point stays the one
      // recorded before this step, not the navigator's new position
      L.status = RB_STOP;
      if (L.npoints == np0) L.p = x1;
    } else if (tn == RB_KFOCUS) {
      const rb_volume_desc& v = vol_of(S, next);
      double angle = 0.;
      if (v.qe_angle_table >= 0) angle = cold_acos(dot(L.d, facet_normal(S, P, L, find_border(S, current, next))));
      double qe = v.qe_lambda_table >= 0 ? table_eval(S, v.qe_lambda_table, lambda) : 1.;
      if (v.qe_angle_table >= 0) qe *= table_eval(S, v.qe_angle_table, angle);
      L.status = (qe == 1 || draw(P, id, L) < qe) ? RB_FOCUS : RB_STOP;
    }
    if (L.status == RB_RUN && L.npoints >= P.limit) L.status = RB_SUSPEND;
  }
}

template <bool MIXED, bool HINT = false, typename Rec, typename SC>
__device__ void trace_lane(const SC& S, const TraceParams& P, const RayIO& io, int64_t i, Lane& L, Rec& rec) {
  begin_lane(S, L);
  while (L.status == RB_RUN) step_lane<MIXED, HINT>(S, P, io, i, L, rec);
}

}  // namespace rbd
