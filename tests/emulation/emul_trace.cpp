// TEST INFRASTRUCTURE ONLY: the device tracer's own source
// (robast_amd/csrc/rb_trace.h, rb_device.h, rb_scene_build.h) compiled as
// plain C++ on the host through rb_host_emulation.h, one ray at a time, so
// the device algorithms can be run under AddressSanitizer and
// UndefinedBehaviorSanitizer and compared with the oracle on the CPU.
// Built by tests/test_emulation.py as a program: `emul_trace IN OUT` reads
// the flattened scene, the options and the rays from IN and writes the
// traced rays to OUT (the layout is in tests/test_emulation.py);
// `emul_trace --lane IN OUT` runs the lane kernel's form for scenes with a
// lens too; `emul_trace --oracle IN OUT` traces them with the oracle
// (oracle/rb_oracle.cpp, linked in), so the oracle runs under the same
// sanitizers. Never part of librbgpu.
#define RB_HOST_EMULATION 1
#include <cstdio>
#include <string>
#include <vector>

#include "../../robast_amd/csrc/rb_scene_build.h"
#include "../../robast_amd/csrc/rb_trace.h"

using namespace rbd;

namespace {

// The candidate slots one lane of a 256-lane block owns: slot k at [k * 256].
struct Slots {
  std::vector<int32_t> node = std::vector<int32_t>(size_t(kCand) * kCandStride);
  std::vector<float> tn = std::vector<float>(size_t(kCand) * kCandStride);
};

// HINT: the work-queue kernels' form (the library picks those for scenes
// with a lens): the FP64 search ranks the last daughter first, searches and
// relocates within neighbour cells, and boolean nodes test their children's
// boxes first
template <bool MIXED, int MAXD, bool BORDERS, bool HINT, bool LENS = true, bool SOLVER = true>
void run(const DScene& base, const TraceParams& P, const RayIO& io, int64_t n) {
  DSceneOf<MAXD, BORDERS, (MAXD > 1), HINT, LENS, SOLVER> S;
  static_cast<DScene&>(S) = base;
  Slots slots;
  for (int64_t i = 0; i < n; i++) {
    if (io.status[i] != RB_RUN) continue;
    Lane L;
    L.cand = {slots.node.data(), slots.tn.data()};
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
    L.ray = P.id_base + uint64_t(i);
    NoRecord rec;
    trace_lane<MIXED, HINT>(S, P, io, i, L, rec);
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
}

}  // namespace

extern "C" {

// Traces n host rays in place with the device code; returns 0, or the
// scene packer's error code. *fallbacks counts the mixed form's reruns in
// FP64. Scenes with a lens run the work-queue kernels' form unless
// lane_form.
int emul_trace(const rb_scene_desc* d, const rb_trace_opts* o, rb_rays* r, int64_t n, int64_t* fallbacks,
               bool lane_form) {
  std::vector<char> blob;
  SceneLayout lay;
  std::string err;
  const int rc = pack_scene(d, blob, lay, err);
  if (rc) {
    std::fprintf(stderr, "scene: %s\n", err.c_str());
    return rc;
  }
  const DScene base = scene_view(blob.data(), lay);
  TraceParams P;
  P.seed = o->seed;
  P.id_base = o->ray_id_base;
  P.limit = o->limit > 0 ? o->limit : 100;
  P.fresnel = o->fresnel ? 1 : 0;
  RayIO io;
  io.x = r->x; io.y = r->y; io.z = r->z; io.t = r->t;
  io.dx = r->dx; io.dy = r->dy; io.dz = r->dz; io.lambda = r->lambda;
  io.status = r->status; io.npoints = r->npoints; io.last_node = r->last_node;
  io.path_cap = int32_t(r->path_capacity);
  io.path_xyzt = r->path_xyzt;
  io.path_node = r->path_node;
  rb_emul_fallbacks = 0;
  const bool mixed = o->precision == RB_PRECISION_MIXED;
  const bool shallow = lay.max_depth <= 1 && !lay.has_ext;
  const int form = lay.has_borders ? (shallow ? 3 : 2) : shallow ? 0 : 1;
  const bool hint = lay.has_lens && !lane_form;
  // the lane kernels of scenes without a refracting volume leave the lens
  // physics out (LENS = false), and those of scenes without an aspheric or
  // Winston shape too the iterative solvers (SOLVER = false), as the
  // library's do
  if (!lay.has_lens && !lay.has_solver) {
    if (mixed)
      form == 0 ? run<true, 1, false, false, false, false>(base, P, io, n)
      : form == 1 ? run<true, kMaxCsgDepth, false, false, false, false>(base, P, io, n)
      : form == 2 ? run<true, kMaxCsgDepth, true, false, false, false>(base, P, io, n)
                  : run<true, 1, true, false, false, false>(base, P, io, n);
    else
      form == 0 ? run<false, 1, false, false, false, false>(base, P, io, n)
      : form == 1 ? run<false, kMaxCsgDepth, false, false, false, false>(base, P, io, n)
      : form == 2 ? run<false, kMaxCsgDepth, true, false, false, false>(base, P, io, n)
                  : run<false, 1, true, false, false, false>(base, P, io, n);
  } else if (!lay.has_lens) {
    if (mixed)
      form == 0 ? run<true, 1, false, false, false>(base, P, io, n)
      : form == 1 ? run<true, kMaxCsgDepth, false, false, false>(base, P, io, n)
      : form == 2 ? run<true, kMaxCsgDepth, true, false, false>(base, P, io, n)
                  : run<true, 1, true, false, false>(base, P, io, n);
    else
      form == 0 ? run<false, 1, false, false, false>(base, P, io, n)
      : form == 1 ? run<false, kMaxCsgDepth, false, false, false>(base, P, io, n)
      : form == 2 ? run<false, kMaxCsgDepth, true, false, false>(base, P, io, n)
                  : run<false, 1, true, false, false>(base, P, io, n);
  } else if (form == 3) {
    if (mixed)
      hint ? run<true, 1, true, true>(base, P, io, n) : run<true, 1, true, false>(base, P, io, n);
    else
      hint ? run<false, 1, true, true>(base, P, io, n) : run<false, 1, true, false>(base, P, io, n);
  } else if (mixed)
    form == 0 ? (hint ? run<true, 1, false, true>(base, P, io, n) : run<true, 1, false, false>(base, P, io, n))
    : form == 1 ? (hint ? run<true, kMaxCsgDepth, false, true>(base, P, io, n)
                        : run<true, kMaxCsgDepth, false, false>(base, P, io, n))
                : (hint ? run<true, kMaxCsgDepth, true, true>(base, P, io, n)
                        : run<true, kMaxCsgDepth, true, false>(base, P, io, n));
  else
    form == 0 ? (hint ? run<false, 1, false, true>(base, P, io, n) : run<false, 1, false, false>(base, P, io, n))
    : form == 1 ? (hint ? run<false, kMaxCsgDepth, false, true>(base, P, io, n)
                        : run<false, kMaxCsgDepth, false, false>(base, P, io, n))
                : (hint ? run<false, kMaxCsgDepth, true, true>(base, P, io, n)
                        : run<false, kMaxCsgDepth, true, false>(base, P, io, n));
  if (fallbacks) *fallbacks = rb_emul_fallbacks;
  return 0;
}

// oracle/rb_oracle.cpp, linked into the program
int oracle_trace(const rb_scene_desc* desc, const rb_trace_opts* opts, rb_rays* rays, int64_t n, int nthreads);

}  // extern "C"

namespace {

struct Reader {
  FILE* f;
  bool ok = true;
  template <typename X>
  std::vector<X> array(int64_t& count) {
    int64_t hdr[2] = {0, 0};  // element count, bytes
    ok = ok && std::fread(hdr, sizeof(hdr), 1, f) == 1;
    count = hdr[0];
    std::vector<X> v;
    if (!ok || hdr[0] < 0 || hdr[1] != hdr[0] * int64_t(sizeof(X))) {
      ok = false;
      return v;
    }
    v.resize(size_t(hdr[0]) + 1);  // never empty: data() is then a valid pointer
    if (hdr[1] > 0) ok = ok && std::fread(v.data(), size_t(hdr[1]), 1, f) == 1;
    return v;
  }
};

}  // namespace

int main(int argc, char** argv) {
  const bool use_oracle = argc == 4 && std::string(argv[1]) == "--oracle";
  const bool lane_form = argc == 4 && std::string(argv[1]) == "--lane";
  if (argc != 3 && !use_oracle && !lane_form) {
    std::fprintf(stderr, "usage: emul_trace [--oracle | --lane] IN OUT\n");
    return 2;
  }
  if (use_oracle || lane_form) argv++;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  Reader R{in};
  int64_t c[9], cnt;
  auto nodes = R.array<rb_node_desc>(c[0]);
  auto children = R.array<int32_t>(c[1]);
  auto volumes = R.array<rb_volume_desc>(c[2]);
  auto shapes = R.array<rb_shape_desc>(c[3]);
  auto params = R.array<double>(c[4]);
  auto tables = R.array<rb_table_desc>(c[5]);
  auto table_xy = R.array<double>(c[6]);
  auto borders = R.array<rb_border_desc>(c[7]);
  auto grids = R.array<rb_grid2d_desc>(c[8]);
  auto opts = R.array<rb_trace_opts>(cnt);
  if (!R.ok || cnt != 1) return 3;
  std::vector<double> f[8];
  int64_t n = -1;
  for (int k = 0; k < 8; k++) {
    f[k] = R.array<double>(cnt);
    if (n < 0) n = cnt;
    if (cnt != n) return 3;
  }
  std::vector<int32_t> g[3];
  for (int k = 0; k < 3; k++) {
    g[k] = R.array<int32_t>(cnt);
    if (cnt != n) return 3;
  }
  int64_t npath = 0, nnode = 0;
  auto path_xyzt = R.array<double>(npath);
  auto path_node = R.array<int32_t>(nnode);
  std::fclose(in);
  if (!R.ok || n < 0 || npath != 4 * nnode || (n > 0 && nnode % n)) return 3;
  rb_scene_desc d;
  d.n_nodes = int32_t(c[0]); d.nodes = nodes.data();
  d.n_children = int32_t(c[1]); d.children = children.data();
  d.n_volumes = int32_t(c[2]); d.volumes = volumes.data();
  d.n_shapes = int32_t(c[3]); d.shapes = shapes.data();
  d.n_params = int32_t(c[4]); d.params = params.data();
  d.n_tables = int32_t(c[5]); d.tables = tables.data();
  d.n_table_xy = int32_t(c[6]); d.table_xy = table_xy.data();
  d.n_borders = int32_t(c[7]); d.borders = borders.data();
  d.n_grids = int32_t(c[8]); d.grids = grids.data();
  rb_rays r{};
  r.x = f[0].data(); r.y = f[1].data(); r.z = f[2].data(); r.t = f[3].data();
  r.dx = f[4].data(); r.dy = f[5].data(); r.dz = f[6].data(); r.lambda = f[7].data();
  r.status = g[0].data(); r.npoints = g[1].data(); r.last_node = g[2].data();
  r.path_capacity = n > 0 ? int32_t(nnode / n) : 0;
  r.path_xyzt = r.path_capacity ? path_xyzt.data() : nullptr;
  r.path_node = r.path_capacity ? path_node.data() : nullptr;
  int64_t fallbacks = 0;
  const int rc = use_oracle ? oracle_trace(&d, opts.data(), &r, n, 2)
                            : emul_trace(&d, opts.data(), &r, n, &fallbacks, lane_form);
  if (rc) return 4;
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  bool ok = std::fwrite(&fallbacks, sizeof(fallbacks), 1, out) == 1;
  for (int k = 0; k < 8; k++) ok = ok && (n == 0 || std::fwrite(f[k].data(), sizeof(double) * size_t(n), 1, out) == 1);
  for (int k = 0; k < 3; k++) ok = ok && (n == 0 || std::fwrite(g[k].data(), sizeof(int32_t) * size_t(n), 1, out) == 1);
  if (nnode > 0) {
    ok = ok && std::fwrite(path_xyzt.data(), sizeof(double) * size_t(npath), 1, out) == 1;
    ok = ok && std::fwrite(path_node.data(), sizeof(int32_t) * size_t(nnode), 1, out) == 1;
  }
  std::fclose(out);
  return ok ? 0 : 5;
}
