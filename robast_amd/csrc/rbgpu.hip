// rbgpu.hip -- librbgpu: C ABI (include/rbgpu.h), scene upload with per-node
// BVH build, ray sources and the trace kernel launch for gfx950.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rbgpu.h"
#include "rb_scene_build.h"
#include "rb_trace.h"
#include "rb_adjoint.h"
#include "rb_analysis.h"

using namespace rbd;

struct rb_ctx {
  int device = 0;
  int cus = 0;                                // compute units (grid sizes of the striding kernels)
  unsigned long long* next_ray = nullptr;     // persistent kernels' ray counter (device)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
  double last_ms = 0, kernel_ms = 0;
  // coherence-order buffers, grown on demand: keys and ray indices (double
  // buffered for the radix sort) and the sort's temporary storage
  void* order_buf = nullptr;
  size_t order_cap = 0;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  void* rec_buf = nullptr;  // the rays' staged copy (RayRec, one line per ray), grown on demand
  size_t rec_cap = 0;
  void* scratch = nullptr;  // this device's copy of the allocation-free entry point's scratch array
  hipError_t scratch_err = hipSuccess;
  int launches = 0;
  std::string err;
};

struct rb_scene {
  rb_ctx* ctx = nullptr;
  void* dmem = nullptr;
  DScene s{};
  int max_depth = 0;  // deepest boolean nesting among the scene's shapes
  bool has_lens = false;  // a placed volume refracts (ALens)
  bool has_cells = false; // some node's daughters have neighbour cells
  bool has_borders = false;  // border-surface conditions are defined
  bool has_ext = false;      // a TGeoPcon, general TGeoPgon or AGeoWinstonCone2D: traced with the deep kernel form
  bool has_solver = false;   // an AGeoAsphericDisk or AGeoWinstonConePoly
  bool has_top_box = false;
  float top_lo[3] = {0, 0, 0}, top_hi[3] = {0, 0, 0};
};

namespace {

// The scene as the kernels compiled for composites at most MAXD deep take it.
template <int MAXD, bool BORDERS = false, bool BOX = false, bool LENS = true, bool SOLVER = true>
DSceneOf<MAXD, BORDERS, (MAXD > 1), BOX, LENS, SOLVER> scene_of(const rb_scene* sc) {
  DSceneOf<MAXD, BORDERS, (MAXD > 1), BOX, LENS, SOLVER> s;
  static_cast<DScene&>(s) = sc->s;
  return s;
}

int fail(rb_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, -2, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

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
// Block size of the tracing kernels: the LDS candidate slots are laid out
// for exactly this many lanes.
constexpr int kTraceBlock = kCandStride;
// Occupancy request for the tracing kernels (waves per SIMD); 0 leaves the
// register budget to the compiler. Two waves (256 registers per lane, the
// rest spilled) measured faster than one wave with 512 on MI355X: the
// second wave hides the latency of the scene and BVH loads.
#ifndef RB_WAVES_PER_EU
#define RB_WAVES_PER_EU 2
#endif
#if RB_WAVES_PER_EU > 0
#define RB_TRACE_ATTR __attribute__((amdgpu_waves_per_eu(RB_WAVES_PER_EU)))
#else
#define RB_TRACE_ATTR
#endif

// A ray's state in one 128-byte line. Traced in coherence order, the rays of
// a wavefront lie far apart in the input, so each of the eleven state arrays
// costs the lane a line of its own; the ordering kernel, which reads the
// rays in input order anyway, writes this copy so that the trace reads one
// line per ray instead, and writes the traced state back into the same line
// (k_unstage then copies the lines to the state arrays, coalesced).
struct alignas(128) RayRec {
  double x, y, z, t, dx, dy, dz, lambda;
  int32_t status, npoints, last_node, pad;
};
static_assert(sizeof(RayRec) == 128, "one line per ray");

__device__ __forceinline__ void load_lane(const RayIO& io, int64_t i, Lane& L) {
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
}
// Whether ray i is running and, if it is, its state into L: from its line of
// the staged copy when there is one, else from the state arrays.
__device__ __forceinline__ bool fetch_ray(const RayIO& io, const RayRec* rr, int64_t i, Lane& L) {
  if (rr) {
    const RayRec& r = rr[i];
    if (r.status != RB_RUN) return false;
    L.p = mk(r.x, r.y, r.z);
    L.d = mk(r.dx, r.dy, r.dz);
    L.t = r.t;
    L.lambda = r.lambda;
    L.status = RB_RUN;
    L.npoints = r.npoints;
    L.last_node = r.last_node;
    L.draw = 0;
    L.step = 0;
    L.hit_node = -1;
    L.hit_leaf = 0;
    return true;
  }
  if (io.status[i] != RB_RUN) return false;
  load_lane(io, i, L);
  return true;
}
// The index handed back through an empty asm: the store addresses are then
// formed again after the trace instead of being kept (and spilled) across it
// from the loads.
#ifndef RB_ADDR_BARRIER
#define RB_ADDR_BARRIER 1
#endif
__device__ __forceinline__ int64_t opaque_index(int64_t i) {
  if (RB_ADDR_BARRIER) asm volatile("" : "+v"(i));
  return i;
}
// The traced state into the ray's staged line (all 128 bytes, in 16-byte
// stores), copied to the state arrays afterwards by k_unstage.
__device__ __forceinline__ void store_staged(RayRec* rr, int64_t i, const Lane& L) {
  double2* o = reinterpret_cast<double2*>(rr + i);
  o[0] = make_double2(L.p.x, L.p.y);
  o[1] = make_double2(L.p.z, L.t);
  o[2] = make_double2(L.d.x, L.d.y);
  o[3] = make_double2(L.d.z, L.lambda);
  int4* q = reinterpret_cast<int4*>(o + 4);
  q[0] = make_int4(L.status, L.npoints, L.last_node, 0);
  q[1] = q[2] = q[3] = make_int4(0, 0, 0, 0);
}
__device__ __forceinline__ void store_lane(const RayIO& io, int64_t i, const Lane& L) {
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

// Copies the top node's BVH into LDS. Called by every navigating kernel at
// entry, by all threads of the block, before any thread returns.
__device__ __forceinline__ void stage_bvh(const DScene& S) {
  for (int k = threadIdx.x; k < S.stage_n; k += blockDim.x) g_bvh_stage[k] = S.bvh[S.stage_lo + k];
  __syncthreads();
}

// Coherence key of each ray: a 30-bit Morton code of the point where it
// enters one of the top volume's daughter boxes (box tests only, over the
// top node's BVH), quantised over the daughters' joint box: the nearest box
// it crosses, or with first_box the first one the walk meets. Rays that are
// not running, or cross no daughter box, get the last cell and sort last.
// The work-queue schedules take first_box: Winston cone array at 1e7 rays
// 83.7 -> 80.7 ms per trace (there every ray's nearest box is the window
// above the cones). The lane kernel keeps the nearest box: Davies-Cotton's
// trace kernel 4.46 against 4.55 ms with first_box (its key kernel is
// 0.05 ms faster), Schwarzschild-Couder unchanged.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
// Where the trace reads the rays' state arrays through the permutation (the
// allocation-free entry point, which has no room for the staged copy), rays
// are reordered only within tiles of 2^kOrderTileBits consecutive rays: the
// tile index forms the key's high bits. A tile's ray state (40 MB at 2^19
// rays) stays in the MALL while its permuted loads and stores run, and a
// tile still holds enough rays for coherent wavefronts. Measured on the
// 1e7-ray Davies-Cotton (FP64, ordering included): input order 14.8 ms;
// global reorder 13.0; tiles of 2^15 / 2^17 / 2^19 / 2^21 rays 14.8 / 13.0 /
// 12.1 / 12.3. Moving the ray state into sorted order and back with
// separate gather/scatter passes instead of permuted loads in the trace
// kernel cost 14.2-14.8 ms: the trace gets faster but the two passes of
// scattered 8-byte accesses cost more than they save.
// With the staged copy (RayRec, written by the key kernel in input order,
// one line read per ray) the order is global (kOrderGlobalBits): 1e7-ray
// Davies-Cotton 7.54 -> 6.82 ms against 2^19-ray tiles without the copy,
// 7.52 with the copy and 2^19-ray tiles, 6.93 with 2^21.
#ifndef RB_ORDER_TILE_BITS
#define RB_ORDER_TILE_BITS 19
#endif
constexpr int kOrderTileBits = RB_ORDER_TILE_BITS;
constexpr int kOrderGlobalBits = 31;  // one tile: batches are below 2^31 rays
// Keys are 32-bit: the tile index in the high bits and as many of the top
// bits of the 30-bit Morton code as fit below it (all 30 for a global order
// or up to 4 tiles; 27 = 9 bits per axis for 2^19-ray tiles at 1e7 rays),
// so that the radix sort moves 8-byte pairs (3 passes of 10 bits for the
// global order, 4 of 8 in the allocation-free entry point) instead of
// 12-byte ones. The walk reads the top node's BVH through the same LDS copy
// as the tracing kernels (stage_bvh / bvh_node): it is a chain of dependent
// loads, and LDS answers them at a fraction of the L2 latency.
__global__ void __launch_bounds__(256) k_order_key(DScene S, RayIO io, int64_t n, float lox, float loy, float loz,
                                                   float sx, float sy, float sz, int tile_bits, int mbits,
                                                   uint32_t* key, uint32_t* idx, RayRec* rr, bool first_box) {
  stage_bvh(S);
  const int root = S.nodes[0].bvh_root;
  const int end = bvh_node(S, root).escape;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t k = 0x3fffffffu;  // not running or no daughter box crossed: the last cell
    const int32_t status = io.status[i];
    const double x = io.x[i], y = io.y[i], z = io.z[i], ddx = io.dx[i], ddy = io.dy[i], ddz = io.dz[i];
    if (rr) {
      // the whole line from registers, in 16-byte stores: no line is left
      // partly written
      const double t = io.t[i], lambda = io.lambda[i];
      const int32_t npoints = io.npoints[i], last_node = io.last_node ? io.last_node[i] : -1;
      double2* o = reinterpret_cast<double2*>(rr + i);
      o[0] = make_double2(x, y);
      o[1] = make_double2(z, t);
      o[2] = make_double2(ddx, ddy);
      o[3] = make_double2(ddz, lambda);
      int4* q = reinterpret_cast<int4*>(o + 4);
      q[0] = make_int4(status, npoints, last_node, 0);
      q[1] = q[2] = q[3] = make_int4(0, 0, 0, 0);
    }
    if (status == RB_RUN) {
      const float ox = float(x), oy = float(y), oz = float(z);
      const float dx = float(ddx), dy = float(ddy), dz = float(ddz);
      const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
      float best = 3.0e38f;
      int b = root;
      while (b < end) {
        const DBvh box = bvh_node(S, b);
        float tn;
        if (!box_pass(box, ox, oy, oz, ix, iy, iz, best, tn)) {
          b = box.escape;
          continue;
        }
        if (box.prim >= 0) {
          best = fminf(best, tn);
          if (first_box) break;
        }
        b++;
      }
      if (best < 3.0e38f) {
        const float px = (ox + best * dx - lox) * sx, py = (oy + best * dy - loy) * sy,
                    pz = (oz + best * dz - loz) * sz;
        const uint32_t qx = uint32_t(fminf(fmaxf(px, 0.0f), 1023.0f)), qy = uint32_t(fminf(fmaxf(py, 0.0f), 1023.0f)),
                       qz = uint32_t(fminf(fmaxf(pz, 0.0f), 1023.0f));
        k = (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
      }
    }
    // tile index above the top mbits (<= 30) Morton bits
    key[i] = (uint32_t(i >> tile_bits) << mbits) | (k >> (30 - mbits));
    idx[i] = uint32_t(i);
  }
}

// The traced rays' staged lines back into the state arrays, in input order
// (coalesced on both sides). Lines of rays that were not running hold their
// input state, which is written back unchanged.
__global__ void __launch_bounds__(256) k_unstage(const RayRec* rr, RayIO io, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2* o = reinterpret_cast<const double2*>(rr + i);
    const double2 a = o[0], b = o[1], c = o[2], d = o[3];
    const int4 q = reinterpret_cast<const int4*>(o + 4)[0];
    io.x[i] = a.x; io.y[i] = a.y; io.z[i] = b.x; io.t[i] = b.y;
    io.dx[i] = c.x; io.dy[i] = c.y; io.dz[i] = d.x;
    io.status[i] = q.x;
    io.npoints[i] = q.y;
    if (io.last_node) io.last_node[i] = q.z;
  }
}

// Persistent form: a grid sized to the machine, each lane taking the next
// ray from a global counter as soon as its current one terminates, so a
// wave's lanes stay busy instead of waiting for the longest path among 64
// fixed rays. The counter is claimed once per wave (one vector atomic for
// all lanes that need a ray) and walks the trace order (`perm`, or the
// input order). A ray is traced from start to end by the lane that claimed
// it, with the same per-ray Philox stream, so results equal k_trace's. Every
// wave leaves when the counter has passed n and none of its lanes holds a ray.
template <bool MIXED, int MAXD, bool BORDERS>
__global__ void __launch_bounds__(kTraceBlock) RB_TRACE_ATTR k_trace_persist(DSceneOf<MAXD, BORDERS, (MAXD > 1), true> S, RayIO io,
                                                                          int64_t n,
                                                                          TraceParams P,
                                                                          unsigned long long* next_ray,
                                                                          const uint32_t* perm, RayRec* rr) {
  __shared__ int32_t cand_node[kCand * kCandStride];
  __shared__ float cand_tn[kCand * kCandStride];
  stage_bvh(S);
  Lane L;
  L.cand = {cand_node + threadIdx.x, cand_tn + threadIdx.x};
  NoRecord rec;
  int64_t i = 0;
  bool active = false, exhausted = false;
  for (;;) {
    const bool need = !active && !exhausted;
    const unsigned long long want = __ballot(need);
    if (want != 0ull) {
      const int leader = __ffsll(want) - 1;
      const unsigned rank = __builtin_amdgcn_mbcnt_hi(unsigned(want >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(want), 0u));
      unsigned long long base = 0;
      if (need && rank == 0) base = atomicAdd(next_ray, (unsigned long long)__popcll(want));
      const unsigned lo = __shfl(unsigned(base), leader), hi = __shfl(unsigned(base >> 32), leader);
      if (need) {
        const int64_t c = int64_t((static_cast<unsigned long long>(hi) << 32) | lo) + rank;
        if (c >= n) {
          exhausted = true;
        } else if (i = perm ? int64_t(perm[c]) : c; fetch_ray(io, rr, i, L)) {
          L.ray = P.id_base + uint64_t(i);
          begin_lane(S, L);
          active = true;
        }
      }
    }
    if (!__any(active)) {
      if (!__any(!exhausted)) break;
      continue;
    }
    if (active) {
      step_lane<MIXED, true>(S, P, io, i, L, rec);
      if (L.status != RB_RUN) {
        if (rr)
          store_staged(rr, i, L);
        else
          store_lane(io, i, L);
        active = false;
      }
    }
  }
}

// Regrouping form: the persistent schedule, with the block's live rays
// regrouped before every step so that each wavefront holds rays in the same
// phase. Each lane's phase key is 0 (in a non-refracting volume, inside the
// neighbour cell of the daughter it last met: a search among that cell's
// neighbours), 1 (in a non-refracting volume otherwise: a search of the
// whole BVH), 2 (inside a lens: its exit and the Fresnel physics) or 3 (no
// ray). A counting sort over the block -- per-wave ballot
// counts, then each lane's destination from the counts of the lower keys and
// waves and its rank among its wave's lanes of its key -- moves every lane's
// ray state through LDS to its new lane. Lanes then step their rays as in
// k_trace_persist; free lanes claim the next rays (one atomic per wave).
// A ray is stepped by whichever lane holds it, always with its own Philox
// stream (keyed by its index) and draw counter, so results equal k_trace's.
// Threads per block of the regrouping form: the rays of one block are
// regrouped together (more rays sort into more uniform wavefronts), and
// every step waits at the block's barriers for its slowest wave. Each 256
// lanes keep their own candidate slots.
#ifndef RB_REGROUP_BLOCK
#define RB_REGROUP_BLOCK 256
#endif
constexpr int kRegroupBlock = RB_REGROUP_BLOCK;
static_assert(kRegroupBlock % 64 == 0 && (kRegroupBlock <= kCandStride || kRegroupBlock % kCandStride == 0),
              "regroup block: whole waves, whole slot groups");
constexpr int kRegroupSlots = (kRegroupBlock + kCandStride - 1) / kCandStride;
struct RegroupLDS {
  double px[kRegroupBlock], py[kRegroupBlock], pz[kRegroupBlock];
  double dx[kRegroupBlock], dy[kRegroupBlock], dz[kRegroupBlock];
  double t[kRegroupBlock], lambda[kRegroupBlock], step[kRegroupBlock];
  uint64_t ray[kRegroupBlock];
  int32_t cur[kRegroupBlock], status[kRegroupBlock], npoints[kRegroupBlock], last_node[kRegroupBlock];
  int32_t hit_node[kRegroupBlock], flags[kRegroupBlock];  // flags: bit 0 active, bit 1 on_b
  uint32_t draw[kRegroupBlock], hit_leaf[kRegroupBlock];
  int32_t count[kRegroupBlock / 64][4];
  int32_t exhausted;
};

template <bool MIXED, int MAXD, bool BORDERS>
__global__ void __launch_bounds__(kRegroupBlock) RB_TRACE_ATTR k_trace_regroup(DSceneOf<MAXD, BORDERS, (MAXD > 1), true> S, RayIO io,
                                                                          int64_t n, TraceParams P,
                                                                          unsigned long long* next_ray,
                                                                          const uint32_t* perm, RayRec* rr) {
  __shared__ int32_t cand_node[kRegroupSlots * kCand * kCandStride];
  __shared__ float cand_tn[kRegroupSlots * kCand * kCandStride];
  __shared__ RegroupLDS G;
  stage_bvh(S);
  if (threadIdx.x == 0) G.exhausted = 0;
  __syncthreads();
  Lane L;
  const int slot = int(threadIdx.x) % kCandStride + int(threadIdx.x) / kCandStride * kCand * kCandStride;
  L.cand = {cand_node + slot, cand_tn + slot};
  NoRecord rec;
  bool active = false;
  const int wave = int(threadIdx.x) >> 6;
  for (;;) {
    // free lanes claim the next rays of the trace order, one atomic per wave
    const bool need = !active && !G.exhausted;
    const unsigned long long want = __ballot(need);
    if (want != 0ull) {
      const int leader = __ffsll(want) - 1;
      const unsigned rank = __builtin_amdgcn_mbcnt_hi(unsigned(want >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(want), 0u));
      unsigned long long base = 0;
      if (need && rank == 0) base = atomicAdd(next_ray, (unsigned long long)__popcll(want));
      const unsigned lo = __shfl(unsigned(base), leader), hi = __shfl(unsigned(base >> 32), leader);
      if (need) {
        const int64_t c = int64_t((static_cast<unsigned long long>(hi) << 32) | lo) + rank;
        int64_t i;
        if (c >= n) {
          G.exhausted = 1;
        } else if (i = perm ? int64_t(perm[c]) : c; fetch_ray(io, rr, i, L)) {
          L.ray = P.id_base + uint64_t(i);
          begin_lane(S, L);
          active = true;
        }
      }
    }
    // the block leaves when none of its lanes holds a ray and the order is used up
    if (!__syncthreads_or(active ? 1 : 0) && G.exhausted) break;
    // counting sort of the block's lanes by phase key
    int key = 3;
    if (active) {
      if (kind_of(S, L.cur) == RB_KLENS) {
        key = 2;
      } else {
        const DBvh cell = cell_at(S, L.cur, L.last_node, float(L.p.x), float(L.p.y), float(L.p.z));
        key = cell.prim >= 0 ? 0 : 1;
      }
    }
    const unsigned long long m0 = __ballot(key == 0), m1 = __ballot(key == 1), m2 = __ballot(key == 2);
    const unsigned long long mine = key == 0 ? m0 : key == 1 ? m1 : key == 2 ? m2 : ~(m0 | m1 | m2);
    if ((threadIdx.x & 63) == 0) {
      G.count[wave][0] = __popcll(m0);
      G.count[wave][1] = __popcll(m1);
      G.count[wave][2] = __popcll(m2);
      G.count[wave][3] = 64 - __popcll(m0) - __popcll(m1) - __popcll(m2);
    }
    __syncthreads();
    int dest = int(__builtin_amdgcn_mbcnt_hi(unsigned(mine >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(mine), 0u)));
    for (int w = 0; w < kRegroupBlock / 64; w++) {
      for (int k = 0; k < key; k++) dest += G.count[w][k];
      if (w < wave) dest += G.count[w][key];
    }
    G.px[dest] = L.p.x; G.py[dest] = L.p.y; G.pz[dest] = L.p.z;
    G.dx[dest] = L.d.x; G.dy[dest] = L.d.y; G.dz[dest] = L.d.z;
    G.t[dest] = L.t; G.lambda[dest] = L.lambda; G.step[dest] = L.step;
    G.ray[dest] = L.ray;
    G.cur[dest] = L.cur; G.status[dest] = L.status; G.npoints[dest] = L.npoints; G.last_node[dest] = L.last_node;
    G.hit_node[dest] = L.hit_node; G.flags[dest] = (active ? 1 : 0) | (L.on_b ? 2 : 0);
    G.draw[dest] = L.draw; G.hit_leaf[dest] = L.hit_leaf;
    __syncthreads();
    const int me = int(threadIdx.x);
    L.p = mk(G.px[me], G.py[me], G.pz[me]);
    L.d = mk(G.dx[me], G.dy[me], G.dz[me]);
    L.t = G.t[me]; L.lambda = G.lambda[me]; L.step = G.step[me];
    L.ray = G.ray[me];
    L.cur = G.cur[me]; L.status = G.status[me]; L.npoints = G.npoints[me]; L.last_node = G.last_node[me];
    L.hit_node = G.hit_node[me];
    active = (G.flags[me] & 1) != 0;
    L.on_b = (G.flags[me] & 2) != 0;
    L.draw = G.draw[me]; L.hit_leaf = G.hit_leaf[me];
    // (the next iteration's first barrier keeps these reads ahead of the next writes)
    if (active) {
      const int64_t i = int64_t(L.ray - P.id_base);
      step_lane<MIXED, true>(S, P, io, i, L, rec);
      if (L.status != RB_RUN) {
        if (rr)
          store_staged(rr, i, L);
        else
          store_lane(io, i, L);
        active = false;
      }
    }
  }
}

// One lane per ray. With `perm`, lane t traces ray perm[t] (the coherence
// order); the ray keeps its own index for its Philox stream and its slots.
template <bool MIXED, int MAXD, bool BORDERS, bool LENS, bool SOLVER>
__global__ void __launch_bounds__(kTraceBlock) RB_TRACE_ATTR k_trace(DSceneOf<MAXD, BORDERS, (MAXD > 1), false, LENS, SOLVER> S, RayIO io, int64_t n,
                                                                  TraceParams P, const uint32_t* perm,
                                                                  RayRec* rr) {
  __shared__ int32_t cand_node[kCand * kCandStride];
  __shared__ float cand_tn[kCand * kCandStride];
  stage_bvh(S);
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = perm ? int64_t(perm[t]) : t;
  Lane L;
  if (!fetch_ray(io, rr, i, L)) return;
  L.cand = {cand_node + threadIdx.x, cand_tn + threadIdx.x};
  L.ray = P.id_base + uint64_t(i);
  NoRecord rec;
  trace_lane<MIXED>(S, P, io, i, L, rec);
  if (rr)
    store_staged(rr, opaque_index(i), L);
  else
    store_lane(io, opaque_index(i), L);
}

// Adjoint of the focal-plane loss: each lane recomputes its ray's bounce
// chain (nothing is stored by the forward pass) and accumulates parameter
// gradients in LDS (FP64), flushed with one global atomic per parameter.
template <bool MIXED, int MAXD, bool LENS = true, bool SOLVER = true>
__global__ void __launch_bounds__(kTraceBlock) RB_TRACE_ATTR k_adjoint(DSceneOf<MAXD, false, (MAXD > 1), false, LENS, SOLVER> S, RayIO io, int64_t n,
                                                                    TraceParams P, AdjParams A, double* grad,
                                                                    const uint32_t* perm, const RayRec* rr) {
  __shared__ double acc[kMaxAdjAcc];
  const int nacc = adj_acc_size(A.mode, A.n_params, A.n_groups);
  __shared__ int32_t cand_node[kCand * kCandStride];
  __shared__ float cand_tn[kCand * kCandStride];
  for (int k = threadIdx.x; k < nacc; k += blockDim.x) acc[k] = 0.0;
  stage_bvh(S);
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = t < n ? (perm ? int64_t(perm[t]) : t) : n;
  Lane L;
  if (i < n && fetch_ray(io, rr, i, L)) {
    L.cand = {cand_node + threadIdx.x, cand_tn + threadIdx.x};
    L.last_node = -1;  // the recomputed chain starts without a hint
    L.ray = P.id_base + uint64_t(i);
    const V3<double> origin = L.p, d0 = L.d;
    ChainRecord rec;
    RayIO ro = io;
    ro.path_cap = 0;
    trace_lane<MIXED>(S, P, ro, i, L, rec);
    if (L.status == RB_FOCUS && !rec.overflow && L.last_node == A.focal_node)
      adjoint_accumulate(S, A, rec, origin, d0, L.p, i, acc);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nacc; k += blockDim.x)
    if (acc[k] != 0.0) atomicAdd(&grad[k], acc[k]);
}

__global__ void __launch_bounds__(256) k_generate(rb_raygen_desc g, RayIO io, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ox, oy, oz, dx, dy, dz, lam;
  if (g.kind == RB_RAYGEN_CONE) {
    double u0 = uniform01(g.seed, uint64_t(i), 0u, 1u), u1 = uniform01(g.seed, uint64_t(i), 1u, 1u);
    double u2 = uniform01(g.seed, uint64_t(i), 2u, 1u), u3 = uniform01(g.seed, uint64_t(i), 3u, 1u);
    double u4 = uniform01(g.seed, uint64_t(i), 4u, 1u);
    double rho = 0.5 * g.size * sqrt(u0), al = 2 * kPi * u1;
    double th = g.theta_min + (g.theta_max - g.theta_min) * u2, ph = 2 * kPi * u3;
    lam = g.lambda_min + (g.lambda_max - g.lambda_min) * u4;
    double ri = g.distance * tan(th);
    double ix = ri * cos(ph), iy = ri * sin(ph);
    double vx = ix - rho * cos(al), vy = iy - rho * sin(al), vz = -g.distance;
    double m = sqrt(vx * vx + vy * vy + vz * vz);
    dx = vx / m; dy = vy / m; dz = vz / m;
    double s = g.start / -dz;
    ox = ix - s * dx; oy = iy - s * dy; oz = g.z0 + g.start;
  } else {
    double th, px, py;
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
      ox = px + cx; oy = py; oz = cz;
    } else {
      ox = cx + px * ct; oy = py; oz = cz + px * st;
    }
    dx = sin(t2); dy = 0; dz = cos(t2);
  }
  io.x[i] = ox;
  io.y[i] = oy;
  io.z[i] = oz;
  io.t[i] = 0;
  io.dx[i] = dx;
  io.dy[i] = dy;
  io.dz[i] = dz;
  const_cast<double*>(io.lambda)[i] = lam;
  io.status[i] = RB_RUN;
  io.npoints[i] = 1;
  if (io.last_node) io.last_node[i] = -1;
}

// ---------------------------------------------------------------- ray sources
// ARayShooter (src/ARayShooter.cxx:122-459) on the device: lane i finds its
// place in the pattern from i alone, places it with the rotation and then the
// translation (TGeoRotation / TGeoTranslation::LocalToMaster), and normalises
// the direction as ARay::SetDirection does (src/ARay.cxx:217-223). FMA
// contraction is off so every position rounds as the written expression.
constexpr int kSourceBlock = 256;

__device__ __forceinline__ void rot3(const double* R, const double* l, double* m) {
#pragma clang fp contract(off)
  for (int r = 0; r < 3; r++) m[r] = l[0] * R[3 * r] + l[1] * R[3 * r + 1] + l[2] * R[3 * r + 2];
}

__device__ __forceinline__ void store_ray(const RayIO& io, int64_t i, const double* p, double t, const double* d,
                                          double lam) {
#pragma clang fp contract(off)
  const double mag = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double u0 = 0, u1 = 0, u2 = 0;  // SetDirection leaves a zero vector unset: TVector3's (0, 0, 0)
  if (mag > 0) {
    u0 = d[0] / mag;
    u1 = d[1] / mag;
    u2 = d[2] / mag;
  }
  io.x[i] = p[0];
  io.y[i] = p[1];
  io.z[i] = p[2];
  io.t[i] = t;
  io.dx[i] = u0;
  io.dy[i] = u1;
  io.dz[i] = u2;
  const_cast<double*>(io.lambda)[i] = lam;
  io.status[i] = RB_RUN;
  io.npoints[i] = 1;
  if (io.last_node) io.last_node[i] = -1;
}

__global__ void __launch_bounds__(kSourceBlock) k_shoot(rb_shooter_desc g, RayIO io, int64_t n) {
#pragma clang fp contract(off)
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t id = g.id_base + uint64_t(i);
    double loc[3] = {0., 0., 0.};        // the pattern point, local frame
    double dl[3] = {g.dir[0], g.dir[1], g.dir[2]};
    bool place = true;                   // false: the ray starts at the translation itself
    bool turn = true;                    // the direction is rotated
    bool cone = false;                   // RandomCone: direction = goal - start
    switch (g.kind) {
      case RB_SHOOT_RECTANGLE: {  // :426-431: i over nx outer, j over ny inner
        const int64_t a = i / g.ny, b = i - a * g.ny;
        const double deltax = g.nx == 1 ? g.a / 2 : g.a / (g.nx - 1);
        const double deltay = g.ny == 1 ? g.b / 2 : g.b / (g.ny - 1);
        loc[0] = double(a) * deltax - g.a / 2;
        loc[1] = double(b) * deltay - g.b / 2;
        break;
      }
      case RB_SHOOT_CIRCLE: {  // :154-177: the centre, then ring k (0-based) of nphi (k + 1) points
        if (i == 0) {
          place = false;
          break;
        }
        const int64_t m = i - 1, nphi = g.ny;
        const int64_t q = m / nphi;  // ring k starts at nphi k (k + 1) / 2: the largest k with k (k + 1) / 2 <= q
        int64_t k = int64_t((sqrt(8.0 * double(q) + 1.0) - 1.0) * 0.5);
        while ((k + 1) * (k + 2) / 2 <= q) k++;
        while (k > 0 && k * (k + 1) / 2 > q) k--;
        const int64_t j = m - nphi * (k * (k + 1) / 2);
        const double r = g.a * double(k + 1) / g.nx;
        const double phi = 2 * kPi / double(nphi) / double(k + 1) * double(j);
        loc[0] = r * cos(phi);
        loc[1] = r * sin(phi);
        break;
      }
      case RB_SHOOT_RANDOM_CIRCLE:
      case RB_SHOOT_RANDOM_CONE: {  // uniform in the disk of radius a (at z = b for the cone)
        const double rho = g.a * sqrt(uniform01(g.seed, id, 0u, 2u));
        const double phi = 2 * kPi * uniform01(g.seed, id, 1u, 2u);
        loc[0] = rho * cos(phi);
        loc[1] = rho * sin(phi);
        if (g.kind == RB_SHOOT_RANDOM_CONE) {
          loc[2] = g.b;
          cone = true;
        }
        break;
      }
      case RB_SHOOT_RANDOM_RECTANGLE: {  // TRandom::Uniform(x1, x2) = x1 + (x2 - x1) U
        const double lo_x = -g.a / 2., lo_y = -g.b / 2.;
        loc[0] = lo_x + (g.a / 2. - lo_x) * uniform01(g.seed, id, 0u, 2u);
        loc[1] = lo_y + (g.b / 2. - lo_y) * uniform01(g.seed, id, 1u, 2u);
        break;
      }
      case RB_SHOOT_RANDOM_SPHERE: {  // isotropic; neither position nor direction is rotated
        const double cz = 2 * uniform01(g.seed, id, 0u, 2u) - 1;
        const double phi = 2 * kPi * uniform01(g.seed, id, 1u, 2u);
        const double s = sqrt(fmax(0.0, 1 - cz * cz));
        dl[0] = s * cos(phi);
        dl[1] = s * sin(phi);
        dl[2] = cz;
        place = false;
        turn = false;
        break;
      }
      default: {  // RB_SHOOT_RANDOM_SPHERICAL_CONE, :365-371
        const double c = cos(g.a * (kPi / 180.0));
        const double th = acos(c + (1 - c) * uniform01(g.seed, id, 0u, 2u));
        const double phi = 0 + (2 * kPi - 0) * uniform01(g.seed, id, 1u, 2u);
        dl[0] = sin(th) * cos(phi);
        dl[1] = sin(th) * sin(phi);
        dl[2] = cos(th);
        place = false;
        break;
      }
    }
    double p[3], d[3];
    const double origin[3] = {0. + g.tr[0], 0. + g.tr[1], 0. + g.tr[2]};
    if (place) {
      rot3(g.rot, loc, p);
      for (int k = 0; k < 3; k++) p[k] = p[k] + g.tr[k];
    } else {
      for (int k = 0; k < 3; k++) p[k] = origin[k];
    }
    if (cone) {  // :268-281: start at the translated origin, aim at the placed goal
      for (int k = 0; k < 3; k++) d[k] = p[k] - origin[k];
      for (int k = 0; k < 3; k++) p[k] = origin[k];
    } else if (turn) {
      rot3(g.rot, dl, d);
    } else {
      for (int k = 0; k < 3; k++) d[k] = dl[k];
    }
    store_ray(io, i, p, 0.0, d, g.lambda);
  }
}

// CORSIKA bunches (src/ACorsikaIACTFile.cxx:71-131): count the rays of each
// bunch, scan, then one lane per output ray finds its bunch by binary search
// over the inclusive scan, so the writes of neighbouring lanes are adjacent.
struct BunchIn {
  const int32_t *tel, *arr;
  const float *x, *y, *time, *cx, *cy, *cz, *lambda, *photons;
};
struct BunchParams {
  double z, tel_z, index, lambda_min, lambda_max;
  uint64_t seed, id_base;
  int32_t tel, arr;
};
constexpr float kMaxBunchPhotons = 16777216.0f;  // 2^24: float(j) < photons counts exactly up to here

__global__ void __launch_bounds__(kSourceBlock) k_bunch_count(BunchIn B, int64_t nb, int32_t tel, int32_t arr,
                                                              int64_t* cnt, int32_t* bad) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nb; i += stride) {
    int64_t c = 0;
    if (B.tel[i] == tel && B.arr[i] == arr) {
      const float p = B.photons[i];
      if (p > kMaxBunchPhotons)
        atomicOr(bad, 1);
      else if (p > 0.f)  // #{j >= 0 : float(j) < p}; NaN and p <= 0 give none
        c = int64_t(ceilf(p));
    }
    cnt[i] = c;
  }
}

// Tile k of the output holds rays [k T, (k + 1) T). tile_b0[k] receives the
// bunch of ray k T, and tile_b0[ntiles] the bunch of the last ray: each bunch
// with rays writes the entries of the tile starts it covers, so every entry is
// written once and no lane searches the whole table.
constexpr int kExpandTile = 2 * kSourceBlock;  // rays per tile
constexpr int kExpandLds = 2048;               // bunch ends a tile stages in LDS (16 KB)

__global__ void __launch_bounds__(kSourceBlock) k_bunch_tiles(const int64_t* cnt, const int64_t* ends, int64_t nb,
                                                              int64_t total, int64_t* tile_b0) {
  const int64_t ntiles = (total + kExpandTile - 1) / kExpandTile;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nb; i += stride) {
    const int64_t c = cnt[i];
    if (c == 0) continue;
    const int64_t e = ends[i], s = e - c;
    for (int64_t k = (s + kExpandTile - 1) / kExpandTile; k * kExpandTile < e; k++) tile_b0[k] = i;
    if (s <= total - 1 && total - 1 < e) tile_b0[ntiles] = i;
  }
}

// One block per tile: the tile's slice of the inclusive scan goes to LDS and
// each lane finds its ray's bunch there, so neighbouring lanes write
// neighbouring rays. A tile spanning more bunches than LDS holds (many
// bunches of other telescopes between the selected ones) searches its slice
// in global memory instead.
__global__ void __launch_bounds__(kSourceBlock) k_bunch_expand(BunchIn B, const int64_t* ends,
                                                               const int64_t* tile_b0, int64_t total, BunchParams Q,
                                                               RayIO io) {
#pragma clang fp contract(off)
  __shared__ int64_t s_ends[kExpandLds];
  const double m = 1e2 * 1.0, cm = 1.0, nm = 1e-9 * m, ns = 1e-9 * 1.0;  // AOpticsManager units
  const int64_t ntiles = (total + kExpandTile - 1) / kExpandTile;
  for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
    const int64_t t0 = k * kExpandTile, t1 = t0 + kExpandTile < total ? t0 + kExpandTile : total;
    const int64_t b0 = tile_b0[k], nbt = tile_b0[k + 1] - b0 + 1;
    const bool staged = nbt <= kExpandLds;
    if (staged)
      for (int64_t j = threadIdx.x; j < nbt; j += blockDim.x) s_ends[j] = ends[b0 + j];
    __syncthreads();
    for (int64_t r = t0 + threadIdx.x; r < t1; r += blockDim.x) {
      int64_t lo = 0, hi = nbt - 1;  // the first bunch of the slice whose inclusive end exceeds r
      while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((staged ? s_ends[mid] : ends[b0 + mid]) > r)
          hi = mid;
        else
          lo = mid + 1;
      }
      const int64_t b = b0 + lo;
      const float cx = B.cx[b], cy = B.cy[b], cz = B.cz[b], lambda = B.lambda[b];
      const double airmass = -1. / cz;
      const double tel_dist = (Q.z - Q.tel_z * cm) * airmass;
      const double speed = kC / Q.index;
      double p[3], d[3] = {cx, cy, cz};
      p[0] = B.x[b] * cm - tel_dist * cx;
      p[1] = B.y[b] * cm - tel_dist * cy;
      p[2] = Q.z;
      const double pt = B.time[b] * ns - tel_dist / speed;
      double lam = lambda;
      if (lambda == 0)
        lam = 1. / (1. / Q.lambda_min -
                    uniform01(Q.seed, Q.id_base + uint64_t(r), 0u, 3u) * (1. / Q.lambda_min - 1. / Q.lambda_max));
      store_ray(io, r, p, pt, d, lam * nm);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- spot analysis
constexpr int kHistBlock = 256;
constexpr int kSearchBlock = 512;  // 8 wavefronts: about one histogram row per thread per sum

// TH2::Fill of the terminal (x, y) of the rays with status H.status: one
// 32-bit atomic per in-range entry. Each block also sums its entries'
// statistics, in a fixed tree order, into partial[block][0..4].
__global__ void __launch_bounds__(kHistBlock) k_hist2d(const double* x, const double* y, const int32_t* status,
                                                       int64_t n, HistParams H, uint32_t* counts, double* partial) {
  double s[5] = {0., 0., 0., 0., 0.};
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (status[i] != H.status) continue;
    const double xi = x[i], yi = y[i];
    const int bx = hist_bin(xi, H.nx, H.xlo, H.xhi), by = hist_bin(yi, H.ny, H.ylo, H.yhi);
    if (bx < 1 || bx > H.nx || by < 1 || by > H.ny) continue;
    atomicAdd(&counts[int64_t(by - 1) * H.nx + (bx - 1)], 1u);
    s[0] += 1.0;
    s[1] += xi;
    s[2] += xi * xi;
    s[3] += yi;
    s[4] += yi * yi;
  }
  __shared__ double red[5][kHistBlock];
  for (int k = 0; k < 5; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = kHistBlock / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) partial[int64_t(blockIdx.x) * 5 + k] = red[k][0];
}

// Compaction of the non-empty bins (after an exclusive scan of the flags):
// entry pos[b] holds bin b; row_start[by - 1] is the first entry of row by.
// Row prefix sums of the histogram: one wavefront per row, an inclusive
// 64-lane scan per chunk of 64 columns carried along the row.
__global__ void __launch_bounds__(kHistBlock) k_hist_rowprefix(const uint32_t* counts, int nx, int ny,
                                                               uint64_t* prefix) {
  const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6, waves = int(blockDim.x) >> 6;
  for (int row = int(blockIdx.x) * waves + wave; row < ny; row += int(gridDim.x) * waves) {
    const uint32_t* c = counts + int64_t(row) * nx;
    uint64_t* p = prefix + int64_t(row) * nx;
    uint64_t carry = 0;
    for (int base = 0; base < nx; base += 64) {
      const int col = base + lane;
      uint64_t v = col < nx ? uint64_t(c[col]) : 0;
      for (int off = 1; off < 64; off <<= 1) {
        const unsigned lo = __shfl_up(unsigned(v), off, 64), hi = __shfl_up(unsigned(v >> 32), off, 64);
        if (lane >= off) v += (uint64_t(hi) << 32) | lo;
      }
      if (col < nx) p[col] = carry + v;
      const unsigned lo = __shfl(unsigned(v), 63, 64), hi = __shfl(unsigned(v >> 32), 63, 64);
      carry += (uint64_t(hi) << 32) | lo;
    }
  }
}

// One block: the statistics (block partials added in block order), the
// integral, then the containment search. out = {r, x, y, sum w, sum x,
// sum x^2, sum y, sum y^2}.
__global__ void __launch_bounds__(kSearchBlock) k_containment(RowPrefix P, const double* partial, int nparts,
                                                              HistParams H, double fraction, double* out) {
  __shared__ double st_sh[5];
  __shared__ uint64_t red[kSearchBlock / 64];
  if (threadIdx.x == 0) {
    double t[5] = {0., 0., 0., 0., 0.};
    for (int b = 0; b < nparts; b++)
      for (int k = 0; k < 5; k++) t[k] += partial[int64_t(b) * 5 + k];
    for (int k = 0; k < 5; k++) st_sh[k] = t[k];
  }
  __syncthreads();
  double st[5];
  for (int k = 0; k < 5; k++) st[k] = st_sh[k];
  uint64_t acc = 0;
  for (int row = threadIdx.x; row < H.ny; row += blockDim.x) acc += P.p[int64_t(row) * H.nx + H.nx - 1];
  const double integral = double(block_sum_u64(acc, red));
  double res[3];
  containment_search(P, H, st, integral, fraction, red, res);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; k++) out[k] = res[k];
    for (int k = 0; k < 5; k++) out[3 + k] = st[k];
  }
}

// ---------------------------------------------------------------- launches
// Clears the persistent schedule's ray counter with a store from a kernel,
// for the entry point whose callers may capture its stream into a graph.
__global__ void k_clear_counter(unsigned long long* c) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *c = 0ull;
}

// The trace kernel for the scene's depth and the lane schedule.
template <bool MIXED, int MAXD, bool BORDERS>
static int launch_trace(rb_ctx* ctx, const rb_scene* scene, int schedule, const RayIO& io, int64_t n,
                        const TraceParams& P, const uint32_t* perm, RayRec* rr, unsigned long long* counter,
                        bool clear_by_kernel) {
  const int block = kTraceBlock;
  const int64_t grid = (n + block - 1) / block;
  if (schedule != RB_SCHEDULE_LANE) {
    const bool regroup = schedule == RB_SCHEDULE_REGROUP;
    const void* fn = regroup ? (const void*)k_trace_regroup<MIXED, MAXD, BORDERS>
                             : (const void*)k_trace_persist<MIXED, MAXD, BORDERS>;
    const int pblock = regroup ? kRegroupBlock : block;
    int per_cu = 0;
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, pblock, 0));
    const int64_t pgrid = std::min<int64_t>((n + pblock - 1) / pblock, int64_t(std::max(per_cu, 1)) * ctx->cus);
    if (clear_by_kernel)
      hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(64), 0, ctx->stream, counter);
    else
      HIPCHK(ctx, hipMemsetAsync(counter, 0, sizeof(unsigned long long), ctx->stream));
    if (regroup)
      hipLaunchKernelGGL((k_trace_regroup<MIXED, MAXD, BORDERS>), dim3(unsigned(pgrid)), dim3(pblock), 0, ctx->stream,
                         scene_of<MAXD, BORDERS, true>(scene), io, n, P, counter, perm, rr);
    else
      hipLaunchKernelGGL((k_trace_persist<MIXED, MAXD, BORDERS>), dim3(unsigned(pgrid)), dim3(block), 0, ctx->stream,
                         scene_of<MAXD, BORDERS, true>(scene), io, n, P, counter, perm, rr);
  } else if (scene->has_lens) {
    hipLaunchKernelGGL((k_trace<MIXED, MAXD, BORDERS, true, true>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<MAXD, BORDERS>(scene), io, n, P, perm, rr);
  } else if (scene->has_solver) {
    // no refracting volume: the form without the lens physics
    hipLaunchKernelGGL((k_trace<MIXED, MAXD, BORDERS, false, true>), dim3(unsigned(grid)), dim3(block), 0,
                       ctx->stream, scene_of<MAXD, BORDERS, false, false>(scene), io, n, P, perm, rr);
  } else {
    // neither a refracting volume nor an aspheric or Winston shape
    hipLaunchKernelGGL((k_trace<MIXED, MAXD, BORDERS, false, false>), dim3(unsigned(grid)), dim3(block), 0,
                       ctx->stream, scene_of<MAXD, BORDERS, false, false, false>(scene), io, n, P, perm, rr);
  }
  HIPCHK(ctx, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- C ABI
// True if the GPU can dereference p: device or managed memory, or host
// memory pinned and mapped by HIP. Pageable host memory would fault in a
// kernel, so callers that promise device pointers are checked with this.
static bool gpu_accessible(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
         (a.type == hipMemoryTypeHost && a.devicePointer != nullptr);
}
static bool rays_accessible(const rb_rays* r, bool with_path) {
  const void* ptrs[13] = {r->x, r->y, r->z, r->t, r->dx, r->dy, r->dz, r->lambda,
                          r->status, r->npoints, r->last_node, r->path_xyzt, r->path_node};
  for (int k = 0; k < (with_path ? 13 : 11); k++)
    if (!gpu_accessible(ptrs[k])) return false;
  return true;
}

// Batches at least this large are traced in coherence order by default.
// Whether the order pays depends on the input: at 2^19 rays, randomly placed
// beams gain (Davies-Cotton 0.65 -> 0.57 ms, Schwarzschild-Couder 1.34 ->
// 0.96), while a grid of parallel rays, already coherent as given, loses
// (paraboloid 0.18 -> 0.35); at 2^17 every case loses. Telling the two apart
// on the device (a count of rays whose key is below their predecessor's,
// read back before the sort) cost 0.16 ms of host round trip per sorted
// batch, more than it saved.
constexpr int64_t kAutoOrderMin = int64_t(1) << 20;
// Scenes traced with a work-queue schedule (a refracting volume) cost tens
// of times more per ray, so the order's fixed cost matters less: the Winston
// cone array (keyed by the first daughter box met) at 2^18 rays 4.40 ms as
// given, 4.28 sorted; 2^19 6.74 and 6.43; 1e6 11.41 and 10.31 (ordering
// included).
constexpr int64_t kAutoOrderMinQueue = int64_t(1) << 18;

// Sort bits of the coherence order of n rays: the tile index in the high
// bits, as many Morton bits below it as fit in 32.
struct OrderBits {
  int tile_bits;  // rays per tile: 2^tile_bits
  int mbits;
  unsigned end_bit;
};
static OrderBits order_bits(int64_t n, int tile_bits) {
  unsigned index_bits = 0;
  while ((int64_t(1) << index_bits) < ((n - 1) >> tile_bits) + 1) index_bits++;
  const int mbits = std::min(30, 32 - int(index_bits));
  return {tile_bits, mbits, unsigned(mbits) + index_bits};
}

// k_order_key over the rays, one per thread. With the staged copy, one ray
// per thread measured 0.58 ms at 1e7 rays against 0.89 for 8 blocks of 256
// per CU striding over them (the BVH staging is then paid per block, from L2).
static int launch_order_key(rb_ctx* ctx, const rb_scene* scene, const RayIO& io, int64_t n, const OrderBits& ob,
                            uint32_t* key, uint32_t* idx, RayRec* rr, bool first_box) {
  float s[3];
  for (int k = 0; k < 3; k++) {
    const float ext = scene->top_hi[k] - scene->top_lo[k];
    s[k] = ext > 0.0f ? 1023.0f / ext : 0.0f;
  }
  const int64_t grid = (n + 255) / 256;
  hipLaunchKernelGGL(k_order_key, dim3(unsigned(grid)), dim3(256), 0, ctx->stream, scene->s, io, n, scene->top_lo[0],
                     scene->top_lo[1], scene->top_lo[2], s[0], s[1], s[2], ob.tile_bits, ob.mbits, key, idx, rr,
                     first_box);
  HIPCHK(ctx, hipGetLastError());
  return 0;
}

// The order's radix sort: 10 bits per onesweep pass, so the global order's
// 30-bit keys take 3 passes instead of 4 (blocks of 1024 threads x 8 keys).
// Measured on the 1e7-ray Davies-Cotton, whole trace: 6.15 ms with rocPRIM's
// default (8 bits), 6.02 with this; 10 bits on 512 x 12, 512 x 24 or
// 1024 x 4 gained nothing, 11 bits lost 0.25 ms.
using OrderSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>, 10,
                                        rocprim::block_radix_rank_algorithm::match>>;

// Coherence order of n rays: keys from k_order_key, then a radix sort of
// (key, index) pairs. *perm receives the sorted indices (device). With rr,
// *rr receives the rays' staged copy (one line per ray, in input order),
// written by the key kernel into a buffer of the context's; where that
// buffer cannot be had (128 B per ray), *rr is null and the order is tiled.
static int order_rays(rb_ctx* ctx, const rb_scene* scene, const RayIO& io, int64_t n, const uint32_t** perm,
                      RayRec** rr = nullptr, bool first_box = false) {
  if (size_t(n) > ctx->order_cap) {
    if (ctx->order_buf) HIPCHK(ctx, hipFree(ctx->order_buf));
    ctx->order_buf = nullptr;
    ctx->order_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->order_buf, 4 * sizeof(uint32_t) * size_t(n)));
    ctx->order_cap = size_t(n);
  }
  uint32_t* key_in = reinterpret_cast<uint32_t*>(ctx->order_buf);
  uint32_t* key_out = key_in + ctx->order_cap;
  uint32_t* idx_in = reinterpret_cast<uint32_t*>(key_out + ctx->order_cap);
  uint32_t* idx_out = idx_in + ctx->order_cap;
  RayRec* rec = nullptr;
  if (rr) {
    if (size_t(n) > ctx->rec_cap) {
      if (ctx->rec_buf) HIPCHK(ctx, hipFree(ctx->rec_buf));
      ctx->rec_buf = nullptr;
      ctx->rec_cap = 0;
      if (hipMalloc(&ctx->rec_buf, sizeof(RayRec) * size_t(n)) == hipSuccess) {
        ctx->rec_cap = size_t(n);
      } else {
        (void)hipGetLastError();  // out of memory: trace without the copy
        ctx->rec_buf = nullptr;
      }
    }
    rec = ctx->rec_cap ? static_cast<RayRec*>(ctx->rec_buf) : nullptr;
    *rr = rec;
  }
  const unsigned count = unsigned(n);
  const OrderBits ob = order_bits(n, rec ? kOrderGlobalBits : kOrderTileBits);
  size_t bytes = 0;
  HIPCHK(ctx, rocprim::radix_sort_pairs<OrderSortConfig>(nullptr, bytes, key_in, key_out, idx_in, idx_out, count, 0u, ob.end_bit,
                                        ctx->stream));
  if (bytes > ctx->sort_tmp_bytes) {
    if (ctx->sort_tmp) HIPCHK(ctx, hipFree(ctx->sort_tmp));
    ctx->sort_tmp = nullptr;
    ctx->sort_tmp_bytes = 0;
    HIPCHK(ctx, hipMalloc(&ctx->sort_tmp, bytes));
    ctx->sort_tmp_bytes = bytes;
  }
  const int rc = launch_order_key(ctx, scene, io, n, ob, key_in, idx_in, rec, first_box);
  if (rc) return rc;
  bytes = ctx->sort_tmp_bytes;
  HIPCHK(ctx, rocprim::radix_sort_pairs<OrderSortConfig>(ctx->sort_tmp, bytes, key_in, key_out, idx_in, idx_out, count, 0u, ob.end_bit,
                                        ctx->stream));
  *perm = idx_out;
  return 0;
}

// ---------------------------------------------------------------- allocation-free trace
// rbgpu_trace_noalloc works in this file-scope device array instead of
// allocating: the persistent schedule's ray counter (first 256 B), the
// coherence order's keys and indices, double buffered (16 B per ray), and
// the radix sort's temporary storage. One copy per GPU, held until the
// process ends; every call on a GPU shares it.
constexpr int64_t kNoAllocMaxRays = RB_NOALLOC_MAX_RAYS;
static_assert(kNoAllocMaxRays == (int64_t(1) << 22), "scratch sized for 2^22 rays");
constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kScratchOrderBytes = 4 * sizeof(uint32_t) * size_t(kNoAllocMaxRays);
constexpr size_t kScratchSortBytes = size_t(1) << 20;
constexpr size_t kScratchBytes = kScratchCounterBytes + kScratchOrderBytes + kScratchSortBytes;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on each GPU, resolved once per device by
// rbgpu_ctx_create with that device current. In host code the array's name
// is not a device address; kernels and copies take this pointer.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static hipError_t resolve_scratch(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// Coherence order inside the scratch array (at most kNoAllocMaxRays rays):
// the double-buffered radix sort, whose temporary storage is only its
// histograms, so the keys and indices need no second copy.
static int order_rays_scratch(rb_ctx* ctx, const rb_scene* scene, const RayIO& io, int64_t n, const uint32_t** perm,
                              bool first_box) {
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch) + kScratchCounterBytes;
  uint32_t* key_a = reinterpret_cast<uint32_t*>(base);
  uint32_t* key_b = key_a + kNoAllocMaxRays;
  uint32_t* idx_a = key_b + kNoAllocMaxRays;
  uint32_t* idx_b = idx_a + kNoAllocMaxRays;
  void* tmp = base + kScratchOrderBytes;
  rocprim::double_buffer<uint32_t> keys(key_a, key_b);
  rocprim::double_buffer<uint32_t> idx(idx_a, idx_b);
  const unsigned count = unsigned(n);
  const OrderBits ob = order_bits(n, kOrderTileBits);
  size_t bytes = 0;
  HIPCHK(ctx, rocprim::radix_sort_pairs(nullptr, bytes, keys, idx, count, 0u, ob.end_bit, ctx->stream));
  if (bytes > kScratchSortBytes)
    return fail(ctx, -2, "rbgpu_trace_noalloc: the radix sort needs " + std::to_string(bytes) +
                             " B of temporary storage, more than the scratch array's " +
                             std::to_string(kScratchSortBytes));
  const int rc = launch_order_key(ctx, scene, io, n, ob, key_a, idx_a, nullptr, first_box);
  if (rc) return rc;
  bytes = kScratchSortBytes;
  HIPCHK(ctx, rocprim::radix_sort_pairs(tmp, bytes, keys, idx, count, 0u, ob.end_bit, ctx->stream));
  *perm = idx.current();
  return 0;
}

// Ordering, schedule and kernel form of one trace, enqueued on the
// context's stream between its timing events. With noalloc the order
// buffers and the ray counter live in the scratch array and nothing is
// allocated; otherwise they are the context's own, grown on demand.
static int trace_launch(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, const RayIO& io, int64_t n,
                        const TraceParams& P, bool noalloc, int* launches) {
  const bool mixed = opts->precision == RB_PRECISION_MIXED;
  const int schedule = opts->schedule != RB_SCHEDULE_AUTO ? opts->schedule
                       : scene->has_lens ? (scene->has_cells ? RB_SCHEDULE_REGROUP : RB_SCHEDULE_PERSISTENT)
                                         : RB_SCHEDULE_LANE;
  const bool persistent = schedule != RB_SCHEDULE_LANE;
  const bool coherent =
      scene->has_top_box && n < (int64_t(1) << 31) &&
      (opts->order == RB_ORDER_COHERENT ||
       (opts->order == RB_ORDER_AUTO && n >= (persistent ? kAutoOrderMinQueue : kAutoOrderMin)));
  const uint32_t* perm = nullptr;
  RayRec* rr = nullptr;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (coherent) {
    // the key: the first daughter box met for the work-queue schedules, the
    // nearest for the lane kernel
    const int rc = noalloc ? order_rays_scratch(ctx, scene, io, n, &perm, persistent)
                           : order_rays(ctx, scene, io, n, &perm, &rr, persistent);
    if (rc) return rc;
  }
  // the key kernel, the radix sort (counted once), the trace and the
  // copy-back of the staged lines
  *launches = (perm ? 2 : 0) + 1 + (rr ? 1 : 0);
  unsigned long long* counter = nullptr;
  if (noalloc) {
    counter = static_cast<unsigned long long*>(ctx->scratch);
  } else {
    if (persistent && !ctx->next_ray) HIPCHK(ctx, hipMalloc(&ctx->next_ray, sizeof(unsigned long long)));
    counter = ctx->next_ray;
  }
  HIPCHK(ctx, hipEventRecord(ctx->evk0, ctx->stream));
  // kernels: composites at most one level deep and no TGeoPcon / general
  // TGeoPgon / AGeoWinstonCone2D (form 0, or 3 with border conditions);
  // deeper composites or those shapes (form 1, or 2 with border conditions).
  // The one-level border form measured 9.45 ms against 11.04 for the deep
  // one on the coated Davies-Cotton (1e7 rays): a 336-byte frame, not 1,688.
  const bool shallow = scene->max_depth <= 1 && !scene->has_ext;
  const int form = scene->has_borders ? (shallow ? 3 : 2) : shallow ? 0 : 1;
  int rc;
  if (form == 3)
    rc = mixed ? launch_trace<true, 1, true>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc)
               : launch_trace<false, 1, true>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc);
  else if (mixed)
    rc = form == 0 ? launch_trace<true, 1, false>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc)
       : form == 1 ? launch_trace<true, kMaxCsgDepth, false>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc)
                   : launch_trace<true, kMaxCsgDepth, true>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc);
  else
    rc = form == 0 ? launch_trace<false, 1, false>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc)
       : form == 1 ? launch_trace<false, kMaxCsgDepth, false>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc)
                   : launch_trace<false, kMaxCsgDepth, true>(ctx, scene, schedule, io, n, P, perm, rr, counter, noalloc);
  if (rc) return rc;
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->evk1, ctx->stream));
  if (rr) {
    const int64_t grid = std::min<int64_t>((n + 255) / 256, int64_t(ctx->cus) * 8);
    hipLaunchKernelGGL(k_unstage, dim3(unsigned(grid)), dim3(256), 0, ctx->stream, rr, io, n);
    HIPCHK(ctx, hipGetLastError());
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  return 0;
}

// Checks shared by the trace entry points; 1 = nothing to trace (n == 0).
static int check_trace_args(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, const rb_rays* rays,
                            int64_t n, const char* who) {
  const std::string w(who);
  if (!scene || !opts || !rays || n < 0) return fail(ctx, -1, w + ": null argument or negative n");
  if (opts->precision != RB_PRECISION_FP64 && opts->precision != RB_PRECISION_MIXED)
    return fail(ctx, -1, w + ": unknown precision mode");
  if (opts->order != RB_ORDER_AUTO && opts->order != RB_ORDER_INPUT && opts->order != RB_ORDER_COHERENT)
    return fail(ctx, -1, w + ": unknown order");
  if (opts->schedule != RB_SCHEDULE_AUTO && opts->schedule != RB_SCHEDULE_LANE &&
      opts->schedule != RB_SCHEDULE_PERSISTENT && opts->schedule != RB_SCHEDULE_REGROUP)
    return fail(ctx, -1, w + ": unknown schedule");
  ctx->launches = 0;
  ctx->last_ms = 0;
  ctx->kernel_ms = 0;
  if (n == 0) return 1;  // an empty batch: nothing to read, arrays may be null
  if (!rays->x || !rays->y || !rays->z || !rays->t || !rays->dx || !rays->dy || !rays->dz || !rays->lambda ||
      !rays->status || !rays->npoints)
    return fail(ctx, -1, w + ": missing ray array");
  if (rays->path_capacity < 0 || (rays->path_capacity > 0 && (!rays->path_xyzt || !rays->path_node)))
    return fail(ctx, -1, w + ": bad path record");
  return 0;
}

static TraceParams trace_params(const rb_trace_opts* opts) {
  TraceParams P;
  P.seed = opts->seed;
  P.id_base = opts->ray_id_base;
  P.limit = opts->limit > 0 ? opts->limit : 100;
  P.fresnel = opts->fresnel ? 1 : 0;
  return P;
}

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
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->evk0) != hipSuccess || hipEventCreate(&c->evk1) != hipSuccess ||
      hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus <= 0) {
    rbgpu_ctx_destroy(c);
    return -2;
  }
  // rbgpu_trace_noalloc's scratch on this device (the device is current);
  // a failure is reported by that entry point and leaves the others usable
  c->scratch_err = resolve_scratch(device, &c->scratch);
  *out = c;
  return 0;
}

void rbgpu_ctx_destroy(rb_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->next_ray) (void)hipFree(c->next_ray);
  if (c->order_buf) (void)hipFree(c->order_buf);
  if (c->sort_tmp) (void)hipFree(c->sort_tmp);
  if (c->rec_buf) (void)hipFree(c->rec_buf);
  if (c->evk0) (void)hipEventDestroy(c->evk0);
  if (c->evk1) (void)hipEventDestroy(c->evk1);
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
  out[9] = sizeof(rb_adjoint_desc);
  out[10] = sizeof(rb_hist2d_desc);
  out[11] = sizeof(rb_border_desc);
  out[12] = sizeof(rb_shooter_desc);
  out[13] = sizeof(rb_bunches);
  out[14] = sizeof(rb_bunch_opts);
  out[15] = sizeof(rb_grid2d_desc);
  return 16;
}

int rbgpu_scene_upload(rb_ctx* ctx, const rb_scene_desc* d, rb_scene** out) {
  if (!ctx || !out) return -1;
  *out = nullptr;
  if (!d) return fail(ctx, -1, "scene: null descriptor");
  std::vector<char> blob;
  SceneLayout lay;
  int rc = pack_scene(d, blob, lay, ctx->err);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
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
  sc->s = scene_view(static_cast<const char*>(sc->dmem), lay);
  sc->s.stage_lo = lay.top_bvh_lo;
  sc->s.stage_n = std::min(lay.top_bvh_n, int32_t(kBvhStage));
  sc->max_depth = lay.max_depth;
  sc->has_lens = lay.has_lens;
  sc->has_cells = lay.has_cells;
  sc->has_borders = lay.has_borders;
  sc->has_ext = lay.has_ext;
  sc->has_solver = lay.has_solver;
  sc->has_top_box = lay.has_top_box;
  for (int k = 0; k < 3; k++) {
    sc->top_lo[k] = lay.top_lo[k];
    sc->top_hi[k] = lay.top_hi[k];
  }
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
  const int chk = check_trace_args(ctx, scene, opts, rays, n, "rbgpu_trace");
  if (chk) return chk > 0 ? 0 : chk;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const TraceParams P = trace_params(opts);

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
    if (!rays_accessible(rays, cap > 0)) return fail(ctx, -1, "rbgpu_trace: device_ptrs set but a ray array is not GPU memory");
    io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
    io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
    io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
    io.path_cap = int32_t(cap);
    io.path_xyzt = rays->path_xyzt;
    io.path_node = rays->path_node;
  }

  int launches = 0;
  const int rc = trace_launch(ctx, scene, opts, io, n, P, false, &launches);
  if (rc) return rc;
  ctx->launches = launches;

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
  float ms = 0, kms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  HIPCHK(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
  ctx->last_ms = ms;
  ctx->kernel_ms = kms;
  return 0;
}

int rbgpu_trace_noalloc(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, rb_rays* rays, int64_t n) {
  if (!ctx) return -1;
  const int chk = check_trace_args(ctx, scene, opts, rays, n, "rbgpu_trace_noalloc");
  if (chk) return chk > 0 ? 0 : chk;
  if (n > kNoAllocMaxRays)
    return fail(ctx, -1, "rbgpu_trace_noalloc: " + std::to_string(n) + " rays exceed the scratch maximum of " +
                             std::to_string(kNoAllocMaxRays) + " (use rbgpu_trace, or split the batch)");
  if (!opts->device_ptrs)
    return fail(ctx, -1, "rbgpu_trace_noalloc: the ray arrays must be device memory (set device_ptrs)");
  if (ctx->scratch_err != hipSuccess || !ctx->scratch)
    return fail(ctx, -2, std::string("rbgpu_trace_noalloc: scratch array address unresolved: ") +
                             hipGetErrorString(ctx->scratch_err));
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!rays_accessible(rays, rays->path_capacity > 0))
    return fail(ctx, -1, "rbgpu_trace_noalloc: a ray array is not GPU memory");
  const TraceParams P = trace_params(opts);
  RayIO io;
  io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
  io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
  io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
  io.path_cap = int32_t(rays->path_capacity);
  io.path_xyzt = rays->path_xyzt;
  io.path_node = rays->path_node;
  int launches = 0;
  const int rc = trace_launch(ctx, scene, opts, io, n, P, true, &launches);
  if (rc) return rc;
  ctx->launches = launches;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0, kms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  HIPCHK(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
  ctx->last_ms = ms;
  ctx->kernel_ms = kms;
  return 0;
}

int rbgpu_generate_rays(rb_ctx* ctx, const rb_raygen_desc* g, rb_rays* rays, int64_t n) {
  if (!ctx) return -1;
  if (!g || !rays || n < 0) return fail(ctx, -1, "rbgpu_generate_rays: null argument");
  if (g->kind != RB_RAYGEN_GRID && g->kind != RB_RAYGEN_DISK && g->kind != RB_RAYGEN_CONE)
    return fail(ctx, -1, "rbgpu_generate_rays: bad kind");
  if (g->kind == RB_RAYGEN_GRID && (g->n_side < 1 || int64_t(g->n_side) * g->n_side < n))
    return fail(ctx, -1, "rbgpu_generate_rays: grid smaller than n");
  if (n == 0) return 0;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!rays->x || !rays->y || !rays->z || !rays->t || !rays->dx || !rays->dy || !rays->dz || !rays->lambda ||
      !rays->status || !rays->npoints)
    return fail(ctx, -1, "rbgpu_generate_rays: missing ray array");
  if (!rays_accessible(rays, false)) return fail(ctx, -1, "rbgpu_generate_rays: a ray array is not GPU memory");
  RayIO io{};
  io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
  io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
  io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
  const int block = kTraceBlock;
  const int64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_generate, dim3(unsigned(grid)), dim3(block), 0, ctx->stream, *g, io, n);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// Patterns above this many rays are refused (int64 index arithmetic in k_shoot
// stays exact far beyond it).
constexpr int64_t kMaxShooterRays = int64_t(1) << 40;

int64_t rbgpu_shooter_count(const rb_shooter_desc* g, int64_t n) {
  if (!g) return -1;
  int64_t c;
  switch (g->kind) {
    case RB_SHOOT_RECTANGLE:  // :407-409
      c = (g->a < 0 || g->b < 0 || g->nx < 1 || g->ny < 1) ? 0 : int64_t(g->nx) * g->ny;
      break;
    case RB_SHOOT_CIRCLE: {  // :128-130; ring i holds nphi (i + 1) points
      if (g->a < 0 || g->nx < 1 || g->ny < 1) return 0;
      const double est = double(g->ny) * double(g->nx) * (double(g->nx) + 1) / 2;
      if (est > double(kMaxShooterRays)) return -1;
      c = 1 + int64_t(g->ny) * (int64_t(g->nx) * (int64_t(g->nx) + 1) / 2);
      break;
    }
    case RB_SHOOT_RANDOM_CIRCLE: c = g->a < 0 ? 0 : std::max<int64_t>(n, 0); break;        // :190-192
    case RB_SHOOT_RANDOM_RECTANGLE: c = (g->a < 0 || g->b < 0 || n < 1) ? 0 : n; break;   // :294-296
    case RB_SHOOT_RANDOM_CONE:
    case RB_SHOOT_RANDOM_SPHERE:
    case RB_SHOOT_RANDOM_SPHERICAL_CONE: c = std::max<int64_t>(n, 0); break;
    default: return -1;
  }
  return c > kMaxShooterRays ? -1 : c;
}

static RayIO source_io(const rb_rays* rays) {
  RayIO io{};
  io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
  io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
  io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
  return io;
}

static bool source_arrays_present(const rb_rays* r) {
  return r->x && r->y && r->z && r->t && r->dx && r->dy && r->dz && r->lambda && r->status && r->npoints;
}

int rbgpu_shoot_rays(rb_ctx* ctx, const rb_shooter_desc* g, rb_rays* rays, int64_t n) {
  if (!ctx) return -1;
  if (!g || !rays || n < 0) return fail(ctx, -1, "rbgpu_shoot_rays: null argument or negative n");
  const int64_t want = rbgpu_shooter_count(g, n);
  if (want < 0) return fail(ctx, -1, "rbgpu_shoot_rays: bad kind or a pattern above 2^40 rays");
  if (want != n)
    return fail(ctx, -1, "rbgpu_shoot_rays: n = " + std::to_string(n) + " but the pattern has " +
                             std::to_string(want) + " rays");
  ctx->last_ms = ctx->kernel_ms = 0;
  ctx->launches = 0;
  if (n == 0) return 0;
  if (!source_arrays_present(rays)) return fail(ctx, -1, "rbgpu_shoot_rays: missing ray array");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!rays_accessible(rays, false)) return fail(ctx, -1, "rbgpu_shoot_rays: a ray array is not GPU memory");
  const RayIO io = source_io(rays);
  const int64_t grid = std::min<int64_t>((n + kSourceBlock - 1) / kSourceBlock, int64_t(ctx->cus) * 64);
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_shoot, dim3(unsigned(grid)), dim3(kSourceBlock), 0, ctx->stream, *g, io, n);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ctx->kernel_ms = ms;
  ctx->launches = 1;
  return 0;
}

int rbgpu_bunches_to_rays(rb_ctx* ctx, const rb_bunches* b, int64_t nb, const rb_bunch_opts* o, rb_rays* rays,
                          int64_t capacity, int64_t* n_rays) {
  if (!ctx) return -1;
  if (!b || !o || !n_rays || nb < 0 || capacity < 0)
    return fail(ctx, -1, "rbgpu_bunches_to_rays: null argument or negative size");
  *n_rays = 0;
  if (!(o->refractive_index > 0) || !(o->lambda_min > 0) || !(o->lambda_max > 0))
    return fail(ctx, -1, "rbgpu_bunches_to_rays: refractive index and wavelength bounds must be positive");
  ctx->last_ms = ctx->kernel_ms = 0;
  ctx->launches = 0;
  if (nb == 0) return 0;
  if (!b->tel_no || !b->array_no || !b->x || !b->y || !b->time || !b->cx || !b->cy || !b->cz || !b->lambda ||
      !b->photons)
    return fail(ctx, -1, "rbgpu_bunches_to_rays: missing bunch array");
  if (rays && !source_arrays_present(rays)) return fail(ctx, -1, "rbgpu_bunches_to_rays: missing ray array");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (rays && !rays_accessible(rays, false))
    return fail(ctx, -1, "rbgpu_bunches_to_rays: a ray array is not GPU memory");
  DevRays dev;
  BunchIn B{b->tel_no, b->array_no, b->x, b->y, b->time, b->cx, b->cy, b->cz, b->lambda, b->photons};
  if (!o->device_ptrs) {
    int32_t *tel, *arr;
    HIPCHK(ctx, dev.alloc(&tel, nb));
    HIPCHK(ctx, dev.alloc(&arr, nb));
    const size_t ib = size_t(nb) * sizeof(int32_t), fb = size_t(nb) * sizeof(float);
    HIPCHK(ctx, hipMemcpyAsync(tel, b->tel_no, ib, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(arr, b->array_no, ib, hipMemcpyHostToDevice, ctx->stream));
    const float* src[8] = {b->x, b->y, b->time, b->cx, b->cy, b->cz, b->lambda, b->photons};
    float* dst[8];
    for (int k = 0; k < 8; k++) {
      HIPCHK(ctx, dev.alloc(&dst[k], nb));
      HIPCHK(ctx, hipMemcpyAsync(dst[k], src[k], fb, hipMemcpyHostToDevice, ctx->stream));
    }
    B = BunchIn{tel, arr, dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6], dst[7]};
  } else {
    const void* ptrs[10] = {B.tel, B.arr, B.x, B.y, B.time, B.cx, B.cy, B.cz, B.lambda, B.photons};
    for (const void* p : ptrs)
      if (!gpu_accessible(p)) return fail(ctx, -1, "rbgpu_bunches_to_rays: device_ptrs set but a bunch array is not GPU memory");
  }
  int64_t *cnt, *ends;
  int32_t* bad;
  HIPCHK(ctx, dev.alloc(&cnt, nb));
  HIPCHK(ctx, dev.alloc(&ends, nb));
  HIPCHK(ctx, dev.alloc(&bad, 1));
  size_t scan_bytes = 0;
  HIPCHK(ctx, rocprim::inclusive_scan(nullptr, scan_bytes, cnt, ends, size_t(nb), rocprim::plus<int64_t>(),
                                      ctx->stream));
  char* scan_tmp;
  HIPCHK(ctx, dev.alloc(&scan_tmp, int64_t(std::max<size_t>(scan_bytes, 1))));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(bad, 0, sizeof(int32_t), ctx->stream));
  const int64_t cgrid = std::min<int64_t>((nb + kSourceBlock - 1) / kSourceBlock, int64_t(ctx->cus) * 16);
  hipLaunchKernelGGL(k_bunch_count, dim3(unsigned(cgrid)), dim3(kSourceBlock), 0, ctx->stream, B, nb, o->tel_no,
                     o->array_no, cnt, bad);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, rocprim::inclusive_scan(scan_tmp, scan_bytes, cnt, ends, size_t(nb), rocprim::plus<int64_t>(),
                                      ctx->stream));
  int64_t total = 0;
  int32_t hbad = 0;
  HIPCHK(ctx, hipMemcpyAsync(&total, ends + (nb - 1), sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (hbad) return fail(ctx, -1, "rbgpu_bunches_to_rays: a selected bunch holds more than 2^24 photons");
  *n_rays = total;
  ctx->launches = 2;
  if (!rays) return 0;
  if (total > capacity)
    return fail(ctx, -1, "rbgpu_bunches_to_rays: " + std::to_string(total) + " rays exceed the capacity " +
                             std::to_string(capacity));
  if (total > 0) {
    BunchParams Q{o->z, o->tel_z, o->refractive_index, o->lambda_min, o->lambda_max, o->seed, o->id_base,
                  o->tel_no, o->array_no};
    const RayIO io = source_io(rays);
    const int64_t ntiles = (total + kExpandTile - 1) / kExpandTile;
    int64_t* tile_b0;
    HIPCHK(ctx, dev.alloc(&tile_b0, ntiles + 1));
    HIPCHK(ctx, hipEventRecord(ctx->evk0, ctx->stream));
    hipLaunchKernelGGL(k_bunch_tiles, dim3(unsigned(cgrid)), dim3(kSourceBlock), 0, ctx->stream, cnt, ends, nb, total,
                       tile_b0);
    HIPCHK(ctx, hipGetLastError());
    const int64_t grid = std::min<int64_t>(ntiles, int64_t(ctx->cus) * 64);
    hipLaunchKernelGGL(k_bunch_expand, dim3(unsigned(grid)), dim3(kSourceBlock), 0, ctx->stream, B, ends, tile_b0,
                       total, Q, io);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->evk1, ctx->stream));
    ctx->launches = 4;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0, kms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (total > 0) HIPCHK(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
  ctx->last_ms = ms;
  ctx->kernel_ms = kms;
  return 0;
}

int rbgpu_adjoint(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, const rb_rays* rays, int64_t n,
                  const rb_adjoint_desc* adj, double* grad) {
  if (!ctx) return -1;
  if (!scene || !opts || !rays || !adj || !grad || n < 0) return fail(ctx, -1, "rbgpu_adjoint: null argument");
  if (!opts->device_ptrs) return fail(ctx, -1, "rbgpu_adjoint: rays must be device pointers");
  if ((!adj->node_param && !adj->surface_param) || !adj->param_values || adj->n_params <= 0 ||
      adj->n_params > kMaxAdjParams)
    return fail(ctx, -1, "rbgpu_adjoint: parameter tables missing or larger than " + std::to_string(kMaxAdjParams));
  if (adj->z_param >= adj->n_params || adj->focal_node >= scene->s.n_nodes)
    return fail(ctx, -1, "rbgpu_adjoint: z parameter or focal node out of range");
  if (adj->mode != RB_ADJ_CENTROID && adj->mode != RB_ADJ_MOMENTS && adj->mode != RB_ADJ_ENCIRCLED)
    return fail(ctx, -1, "rbgpu_adjoint: unknown mode");
  if (adj->mode == RB_ADJ_ENCIRCLED && !(adj->ee_radius > 0 && adj->ee_width > 0))
    return fail(ctx, -1, "rbgpu_adjoint: the encircled-energy radius and width must be positive");
  if (scene->has_borders)
    return fail(ctx, -3, "rbgpu_adjoint: rough or Lambertian border surfaces are not differentiated");
  const int groups = adj->mode == RB_ADJ_MOMENTS && adj->n_groups > 0 ? adj->n_groups : 1;
  const int64_t group_size = adj->mode == RB_ADJ_MOMENTS && adj->n_groups > 0 ? adj->group_size : std::max<int64_t>(n, 1);
  if (adj->mode == RB_ADJ_MOMENTS && adj->n_groups > 0 &&
      (adj->group_size <= 0 || int64_t(adj->n_groups) * adj->group_size < n || !adj->ref_xy))
    return fail(ctx, -1, "rbgpu_adjoint: groups must cover the n rays and carry reference points");
  if (adj_acc_size(adj->mode, adj->n_params, groups) > kMaxAdjAcc)
    return fail(ctx, -1, "rbgpu_adjoint: accumulators exceed " + std::to_string(kMaxAdjAcc) + " entries");
  if (!rays->x || !rays->y || !rays->z || !rays->t || !rays->dx || !rays->dy || !rays->dz || !rays->lambda ||
      !rays->status || !rays->npoints)
    return fail(ctx, -1, "rbgpu_adjoint: missing ray array");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!rays_accessible(rays, false) || (adj->node_param && !gpu_accessible(adj->node_param)) ||
      (adj->surface_param && !gpu_accessible(adj->surface_param)) || !gpu_accessible(adj->param_values) ||
      !gpu_accessible(grad) || (adj->mode == RB_ADJ_MOMENTS && adj->n_groups > 0 && !gpu_accessible(adj->ref_xy)))
    return fail(ctx, -1, "rbgpu_adjoint: rays, parameter table, values and grad must be GPU memory");
  ctx->launches = 0;
  ctx->last_ms = 0;
  ctx->kernel_ms = 0;
  if (n == 0) return 0;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  TraceParams P;
  P.seed = opts->seed;
  P.id_base = opts->ray_id_base;
  P.limit = opts->limit > 0 ? opts->limit : 100;
  P.fresnel = opts->fresnel ? 1 : 0;
  RayIO io{};
  io.x = rays->x; io.y = rays->y; io.z = rays->z; io.t = rays->t;
  io.dx = rays->dx; io.dy = rays->dy; io.dz = rays->dz; io.lambda = rays->lambda;
  io.status = rays->status; io.npoints = rays->npoints; io.last_node = rays->last_node;
  io.path_cap = 0;
  AdjParams A;
  A.node_param = adj->node_param;
  A.surface_param = adj->surface_param;
  A.ee_r0 = adj->ee_radius;
  A.ee_w = adj->ee_width;
  A.values = adj->param_values;
  A.n_params = adj->n_params;
  A.z_param = adj->z_param;
  A.focal_node = adj->focal_node;
  A.mode = adj->mode;
  A.ref_xy = adj->mode == RB_ADJ_MOMENTS && adj->n_groups > 0 ? adj->ref_xy : nullptr;
  A.group_size = group_size;
  A.n_groups = groups;
  A.cx = adj->centroid_x;
  A.cy = adj->centroid_y;
  A.weight = adj->weight;
  const int block = kTraceBlock;
  const int64_t grid = (n + block - 1) / block;
  const bool coherent = scene->has_top_box && n < (int64_t(1) << 31) &&
                        (opts->order == RB_ORDER_COHERENT || (opts->order == RB_ORDER_AUTO && n >= kAutoOrderMin));
  const uint32_t* perm = nullptr;
  RayRec* rr = nullptr;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (coherent) {
    const int rc = order_rays(ctx, scene, io, n, &perm, &rr);
    if (rc) return rc;
  }
  HIPCHK(ctx, hipEventRecord(ctx->evk0, ctx->stream));
  const bool shallow = scene->max_depth <= 1 && !scene->has_ext;
  if (opts->precision == RB_PRECISION_MIXED && shallow)
    hipLaunchKernelGGL((k_adjoint<true, 1>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream, scene_of<1>(scene), io,
                       n, P, A, grad, perm, rr);
  else if (opts->precision == RB_PRECISION_MIXED)
    hipLaunchKernelGGL((k_adjoint<true, kMaxCsgDepth>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<kMaxCsgDepth>(scene), io, n, P, A, grad, perm, rr);
  else if (shallow && !scene->has_lens && !scene->has_solver)  // neither lens physics nor iterative solvers
    hipLaunchKernelGGL((k_adjoint<false, 1, false, false>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<1, false, false, false, false>(scene), io, n, P, A, grad, perm, rr);
  else if (shallow && !scene->has_lens)  // no refracting volume: the form without the lens physics
    hipLaunchKernelGGL((k_adjoint<false, 1, false>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<1, false, false, false>(scene), io, n, P, A, grad, perm, rr);
  else if (shallow)
    hipLaunchKernelGGL((k_adjoint<false, 1>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream, scene_of<1>(scene),
                       io, n, P, A, grad, perm, rr);
  else if (!scene->has_lens)
    hipLaunchKernelGGL((k_adjoint<false, kMaxCsgDepth, false>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<kMaxCsgDepth, false, false, false>(scene), io, n, P, A, grad, perm, rr);
  else
    hipLaunchKernelGGL((k_adjoint<false, kMaxCsgDepth>), dim3(unsigned(grid)), dim3(block), 0, ctx->stream,
                       scene_of<kMaxCsgDepth>(scene), io, n, P, A, grad, perm, rr);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->evk1, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0, kms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  HIPCHK(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
  ctx->last_ms = ms;
  ctx->kernel_ms = kms;
  ctx->launches = (perm ? 2 : 0) + 1;  // key kernel, radix sort, adjoint
  return 0;
}

double rbgpu_last_kernel_ms(const rb_ctx* c) { return c ? c->kernel_ms : 0.0; }

int rbgpu_containment_radius(rb_ctx* ctx, const rb_rays* rays, int64_t n, const rb_hist2d_desc* h,
                             double fraction, double* out, uint32_t* counts, double* stats) {
  if (!ctx) return -1;
  if (!rays || !h || !out || n < 0) return fail(ctx, -1, "rbgpu_containment_radius: null argument or negative n");
  if (h->nx < 1 || h->ny < 1 || int64_t(h->nx) * h->ny > (int64_t(1) << 26))
    return fail(ctx, -1, "rbgpu_containment_radius: bin counts must be >= 1 with nx * ny <= 2^26");
  if (!(h->xhi > h->xlo) || !(h->yhi > h->ylo) || !std::isfinite(h->xhi - h->xlo) || !std::isfinite(h->yhi - h->ylo))
    return fail(ctx, -1, "rbgpu_containment_radius: axis range must be finite with hi > lo");
  if (!(fraction > 0.0 && fraction <= 1.0)) return fail(ctx, -1, "rbgpu_containment_radius: fraction must lie in (0, 1]");
  if (n > 0 && (!rays->x || !rays->y || !rays->status))
    return fail(ctx, -1, "rbgpu_containment_radius: missing x, y or status array");
  ctx->launches = 0;
  ctx->last_ms = 0;
  ctx->kernel_ms = 0;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DevRays dev;
  const double* x = rays->x;
  const double* y = rays->y;
  const int32_t* status = rays->status;
  if (n > 0 && !h->device_ptrs) {
    double *dx, *dy;
    int32_t* ds;
    HIPCHK(ctx, dev.alloc(&dx, n));
    HIPCHK(ctx, dev.alloc(&dy, n));
    HIPCHK(ctx, dev.alloc(&ds, n));
    HIPCHK(ctx, hipMemcpyAsync(dx, x, size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dy, y, size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ds, status, size_t(n) * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    x = dx;
    y = dy;
    status = ds;
  } else if (n > 0 && (!gpu_accessible(x) || !gpu_accessible(y) || !gpu_accessible(status))) {
    return fail(ctx, -1, "rbgpu_containment_radius: device_ptrs set but a ray array is not GPU memory");
  }
  const int nb = h->nx * h->ny;
  const int grid = int(std::max<int64_t>(1, std::min<int64_t>((n + kHistBlock - 1) / kHistBlock, int64_t(ctx->cus) * 4)));
  uint32_t* dcounts;
  double *partial, *dout;
  HIPCHK(ctx, dev.alloc(&dcounts, nb));
  HIPCHK(ctx, dev.alloc(&partial, int64_t(grid) * 5));
  HIPCHK(ctx, dev.alloc(&dout, 8));
  HistParams H{h->nx, h->ny, h->xlo, h->xhi, h->ylo, h->yhi, h->status};
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(dcounts, 0, size_t(nb) * sizeof(uint32_t), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(partial, 0, size_t(grid) * 5 * sizeof(double), ctx->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_hist2d, dim3(unsigned(grid)), dim3(kHistBlock), 0, ctx->stream, x, y, status, n, H, dcounts,
                       partial);
    HIPCHK(ctx, hipGetLastError());
  }
  // per-row prefix sums of the contents
  uint64_t* prefix;
  HIPCHK(ctx, dev.alloc(&prefix, nb));
  const int pgrid = std::max(1, std::min((h->ny + 3) / 4, ctx->cus * 8));
  hipLaunchKernelGGL(k_hist_rowprefix, dim3(unsigned(pgrid)), dim3(kHistBlock), 0, ctx->stream, dcounts, h->nx, h->ny,
                     prefix);
  HIPCHK(ctx, hipGetLastError());
  RowPrefix Pr{prefix};
  HIPCHK(ctx, hipEventRecord(ctx->evk0, ctx->stream));
  hipLaunchKernelGGL(k_containment, dim3(1), dim3(kSearchBlock), 0, ctx->stream, Pr, partial, grid, H, fraction,
                     dout);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->evk1, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  double res[8];
  HIPCHK(ctx, hipMemcpyAsync(res, dout, sizeof(res), hipMemcpyDeviceToHost, ctx->stream));
  if (counts)
    HIPCHK(ctx, hipMemcpyAsync(counts, dcounts, size_t(nb) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0, kms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  HIPCHK(ctx, hipEventElapsedTime(&kms, ctx->evk0, ctx->evk1));
  ctx->last_ms = ms;
  ctx->kernel_ms = kms;
  ctx->launches = n > 0 ? 3 : 2;  // fill, row prefix sums, search
  for (int k = 0; k < 3; k++) out[k] = res[k];
  if (stats)
    for (int k = 0; k < 5; k++) stats[k] = res[3 + k];
  return 0;
}

double rbgpu_last_trace_ms(const rb_ctx* c) { return c ? c->last_ms : 0.0; }
int rbgpu_last_trace_launches(const rb_ctx* c) { return c ? c->launches : 0; }

}  // extern "C"
