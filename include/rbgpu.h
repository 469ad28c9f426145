/*
 * rbgpu.h -- C ABI of librbgpu, the MI355X-native non-sequential optical
 * tracer that stands in for ROBAST's AOpticsManager::TraceNonSequential.
 *
 * Everything here is plain C: fixed-size structs, raw pointers and sizes,
 * int return codes (0 = ok, < 0 = error; rbgpu_last_error() has the text).
 * No C++ or torch types cross this boundary, so the ROOT side (a cgo/ctypes/
 * C++ shim) can bind it directly.
 *
 * Units follow ROBAST (include/AOpticsManager.h:59-71): lengths in cm,
 * time in s, wavelength in cm (400 nm = 4e-5).
 */
#ifndef RBGPU_H
#define RBGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ray status: same codes and order as ARay's private enum
 *      (include/ARay.h:26: kRun, kStop, kExit, kFocus, kSuspend, kAbsorb) --- */
enum {
  RB_RUN = 0,
  RB_STOP = 1,
  RB_EXIT = 2,
  RB_FOCUS = 3,
  RB_SUSPEND = 4,
  RB_ABSORB = 5
};

/* ---- optical class of a volume: AOpticsManager's enum
 *      (include/AOpticsManager.h:45-53). A volume's kind is decided by its
 *      exact class, as IsA()== does there (subclasses do not match). ------- */
enum {
  RB_KLENS = 0,
  RB_KOBS = 1,
  RB_KMIRROR = 2,
  RB_KFOCUS = 3,
  RB_KOPT = 4,
  RB_KOTHER = 5,
  RB_KNULL = 6
};

/* ---- shape types and their parameter layout in rb_scene_desc.params ---- */
enum {
  RB_SHAPE_BOX = 1,        /* dx dy dz ox oy oz                     (TGeoBBox)        */
  RB_SHAPE_TUBE = 2,       /* rmin rmax dz                          (TGeoTube)        */
  RB_SHAPE_SPHERE = 3,     /* rmin rmax theta1 theta2 [deg] cos1 cos2 sin1 sin2 phi1 dphi [deg]
                              {cos sin}(phi1) {cos sin}(phi1 + dphi)  (TGeoSphere; dphi >= 360: full) */
  RB_SHAPE_PARABOLOID = 4, /* rlo rhi dz a b                        (TGeoParaboloid)  */
  RB_SHAPE_PGON = 5,       /* phi1 dphi nedges nz {z rmin rmax}*nz {cos sin}*nedges
                              {cos sin}(phi1) {cos sin}(phi1 + dphi)   (TGeoPgon; side k's outward
                              normal at phi1 + dphi/nedges (k + 1/2); z non-decreasing;
                              dphi >= 360 is the full polygon) */
  RB_SHAPE_PCON = 6,       /* phi1 dphi nz {z rmin rmax}*nz {cos sin}(phi1) {cos sin}(phi1 + dphi)
                              (TGeoPcon; z non-decreasing; dphi >= 360 is the full cone) */
  RB_SHAPE_COMPOSITE = 7,  /* op left right lmat[12] rmat[12]       (TGeoBoolNode)    */
  RB_SHAPE_ASPHERIC = 8,   /* z1 c1 z2 c2 rmax rmin kappa1 kappa2 n1 n2 K1[n1] K2[n2]
                              Kn1[n1] Kn2[n2] with Kn_k = K_k rmax^2k (AGeoAsphericDisk) */
  RB_SHAPE_WINSTON = 9,    /* r1 r2 polyN theta dz f cos(theta) sin(theta) tan(pi/N)
                              {cos sin}(2 pi i / N) * N          (AGeoWinstonConePoly) */
  RB_SHAPE_WINSTON2D = 10  /* r1 r2 2 theta dz f cos(theta) sin(theta) 0 1 0 -1 0 dy
                              (AGeoWinstonCone2D: the profile of AGeoWinstonConePoly with two
                              sides, at phi = 0 and pi, extruded over |y| <= dy) */
};

/* boolean operators of RB_SHAPE_COMPOSITE (TGeoUnion/TGeoSubtraction/TGeoIntersection) */
enum { RB_OP_UNION = 0, RB_OP_SUBTRACTION = 1, RB_OP_INTERSECTION = 2 };

/* refractive-index models of a lens volume */
enum {
  RB_INDEX_NONE = 0,      /* n = 1 (ALens without index, src/ALens.cxx:58-60)       */
  RB_INDEX_TABLE = 1,     /* TGraph n(lambda), linear Eval                          */
  RB_INDEX_SELLMEIER = 2, /* B1 B2 B3 C1 C2 C3, lambda in um (src/ASellmeierFormula.cxx:46-54) */
  RB_INDEX_SCHOTT = 3     /* A0..A5, lambda in um (src/ASchottFormula.cxx:43-55)    */
};

/* A 4x3 affine transform in ROOT's convention: master = R * local + T,
 * R row-major (TGeoMatrix::LocalToMaster). */
typedef struct {
  double rot[9];
  double tr[3];
} rb_matrix;

typedef struct {
  int32_t type;          /* RB_SHAPE_* */
  int32_t param_offset;  /* first parameter in rb_scene_desc.params */
  int32_t n_params;
  int32_t pad;
  double bbox_origin[3]; /* local bounding box (TGeoBBox::fOrigin, fDX..fDZ) */
  double bbox_half[3];
} rb_shape_desc;

typedef struct {
  double refl_const;       /* mirror reflectance when refl_table < 0 (default 1) */
  int32_t shape;           /* index into shapes */
  int32_t kind;            /* RB_K* */
  int32_t refl_table;      /* mirror: TGraph R(lambda), -1 = constant            */
  int32_t index_model;     /* lens: RB_INDEX_*                                   */
  int32_t index_ref;       /* lens: table index (TABLE) or param offset (formula) */
  int32_t k_table;         /* lens: TGraph k(lambda), -1 = none (k = 0)          */
  int32_t qe_lambda_table; /* focal surface: TGraph QE(lambda), -1 = 1           */
  int32_t qe_angle_table;  /* focal surface: TGraph QE(angle [rad]), -1 = none   */
  int32_t border_first;    /* this volume's border-surface conditions:           */
  int32_t border_count;    /*   borders[border_first .. + border_count)          */
  int32_t refl_grid;       /* mirror: TH2 R(lambda, angle) in grids (AMirror::SetReflectance(TH2),
                              read with TH2::Interpolate before refl_table), -1 = none */
  int32_t pad;
} rb_volume_desc;

/* ABorderSurfaceCondition(component1, component2) (include/ABorderSurfaceCondition.h:24-45),
 * listed under component1's volume in registration order; the first whose
 * component2 matches the volume entered applies
 * (AOpticalComponent::FindBorderSurfaceCondition, src/AOpticalComponent.cxx:51-65). */
typedef struct {
  int32_t volume2;    /* component2's volume; -1 = the null node (leaving the top volume) */
  int32_t lambertian; /* EnableLambertian(kTRUE): diffuse reflection                 */
  double sigma;       /* SetGaussianRoughness(sigma) [rad]; 0 = smooth                 */
  int32_t multilayer; /* SetMultilayer(): index into rb_scene_desc.grids; -1 = none    */
  int32_t pad;
} rb_border_desc;

/* A multilayer coating (AMultilayer, include/AMultilayer.h) as the device
 * reads it: its unpolarised reflectance and transmittance (CoherentTMMMixed)
 * tabulated at the bin centres of an nx x ny grid over wavelength x [xlo,
 * xhi) (cm) and incidence angle y [ylo, yhi) (rad), as
 * AMultilayer::PreCalculateCoherentTMM fills its TH2Ds
 * (include/AMultilayer.h:243-262), and read back with TH2::Interpolate.
 * params[r_offset + (by - 1) nx + bx - 1] holds R of bin (bx, by), likewise
 * T at t_offset. */
typedef struct {
  int32_t nx, ny;
  double xlo, xhi;
  double ylo, yhi;
  int32_t r_offset, t_offset;
} rb_grid2d_desc;

/* One placed instance of a volume (a path in the TGeo node tree), with its
 * global matrix. Node 0 is the top volume. */
typedef struct {
  rb_matrix global;
  int32_t volume;
  int32_t parent;      /* -1 for the top node */
  int32_t first_child; /* into rb_scene_desc.children */
  int32_t n_children;
} rb_node_desc;

/* A TGraph for linear Eval: table_xy[offset ..] holds the n points sorted by
 * x, then 4 more pairs: the (up, low) points TGraph::Eval extrapolates from
 * above the range and the (low, up) points it uses below it. ROOT picks those
 * by a scan in insertion order, so the flattener records them explicitly. */
typedef struct {
  int32_t offset; /* first (x, y) pair in table_xy */
  int32_t n;      /* number of points (the 4 extrapolation pairs follow) */
} rb_table_desc;

typedef struct {
  int32_t n_nodes;
  const rb_node_desc* nodes;
  int32_t n_children;
  const int32_t* children;
  int32_t n_volumes;
  const rb_volume_desc* volumes;
  int32_t n_shapes;
  const rb_shape_desc* shapes;
  int32_t n_params;
  const double* params;
  int32_t n_tables;
  const rb_table_desc* tables;
  int32_t n_table_xy; /* number of doubles (2 per point) */
  const double* table_xy;
  int32_t n_borders;
  const rb_border_desc* borders;
  int32_t n_grids;
  const rb_grid2d_desc* grids;
} rb_scene_desc;

/* Rays in structure-of-arrays form. On input each ray holds its last point
 * (x, y, z, t), unit direction, wavelength, status and point count, exactly
 * what ARay carries (include/ARay.h:24-35). Rays whose status is not RB_RUN
 * are left untouched (src/AOpticsManager.cxx:349-351). On output the arrays
 * hold the terminal point, final direction, status, point count and the node
 * index of the last AddNode (-1 = null node). */
typedef struct {
  double* x;
  double* y;
  double* z;
  double* t;
  double* dx;
  double* dy;
  double* dz;
  double* lambda;
  int32_t* status;
  int32_t* npoints;
  int32_t* last_node;
  /* optional bounded path record (ARay's TGeoTrack points and node history):
   * path_capacity == 0 disables it; otherwise path_xyzt is [n][capacity][4]
   * and path_node [n][capacity]; slot k holds the (k+1)-th point the trace
   * adds (a ray entering with npoints = m: its point m + k); the points it
   * held before the call are not repeated */
  int32_t path_capacity;
  int32_t pad;
  double* path_xyzt;
  int32_t* path_node;
} rb_rays;

enum { RB_PRECISION_FP64 = 0, RB_PRECISION_MIXED = 1 };

typedef struct {
  int32_t limit;       /* AOpticsManager::fLimit (default 100); SetLimit ignores n <= 0 */
  int32_t fresnel;     /* 1 = Fresnel reflection on (default); 0 = DisableFresnelReflection(kTRUE) */
  int32_t precision;   /* RB_PRECISION_* */
  int32_t device_ptrs; /* 1 = rb_rays arrays are device pointers already in HBM */
  uint64_t seed;       /* Philox4x32-10 key */
  uint64_t ray_id_base;/* global id of ray 0 (Philox counter), for sharded runs */
  int32_t order;       /* RB_ORDER_*: the order rays are traced in (results do not depend on it) */
  int32_t schedule;    /* RB_SCHEDULE_*: how rays are handed to lanes (results do not depend on it) */
} rb_trace_opts;

/* Trace order. Every ray keeps its own Philox stream (keyed by its index),
 * so per-ray results are the same in any order; the order only changes how
 * coherent the rays sharing a wavefront are. COHERENT sorts running rays by
 * a Morton code of the point where each first enters a daughter box of the
 * top volume, and traces them in that order through a permutation; each
 * ray's state is staged in a 128-byte line of a context buffer (128 B per
 * ray, grown on demand), read and written by the trace and copied back to
 * the ray arrays after it. AUTO sorts batches of 2^20 rays or more, or of
 * 2^18 or more when the work-queue schedule traces them. */
enum { RB_ORDER_AUTO = 0, RB_ORDER_INPUT = 1, RB_ORDER_COHERENT = 2 };

/* Lane schedule. LANE: one lane per ray for the whole trace. PERSISTENT: a
 * grid sized to the machine whose lanes take the next ray (in trace order)
 * from a work counter as soon as theirs terminates, so a wavefront does not
 * wait for its longest path. REGROUP: PERSISTENT with the block's live rays
 * sorted by the kind of step they take next before every step, so that each
 * wavefront runs one kind. AUTO picks REGROUP for scenes with refracting
 * volumes (multi-bounce paths of very unequal length) and a volume with at
 * least 256 daughters, PERSISTENT for other scenes with refracting volumes,
 * and LANE otherwise. */
enum { RB_SCHEDULE_AUTO = 0, RB_SCHEDULE_LANE = 1, RB_SCHEDULE_PERSISTENT = 2, RB_SCHEDULE_REGROUP = 3 };

/* ---- synthetic ray sources (device-side ARayShooter equivalents) ---- */
enum {
  RB_RAYGEN_GRID = 0, /* ARayShooter::Square grid (src/ARayShooter.cxx:401-459)        */
  RB_RAYGEN_DISK = 1, /* uniform disk of diameter `size`, normal to the direction,
                         angle ~ U[theta_min, theta_max], lambda ~ U[l_min, l_max]   */
  RB_RAYGEN_CONE = 2  /* converging beam: from a pupil disk of diameter `size` at
                         height `distance` above the image plane z = z0 toward the
                         image of a field angle th ~ U[theta_min, theta_max] at
                         azimuth ~ U[0, 2 pi) (radius distance * tan th); rays start
                         `start` above the image plane                              */
};

typedef struct {
  int32_t kind;
  int32_t n_side;     /* GRID: n rays per side                                  */
  double size;        /* GRID: side length d; DISK: diameter                    */
  double distance;    /* origin of ray plane at distance*(−sin th, 0, cos th)   */
  double theta_min;   /* rad */
  double theta_max;   /* rad */
  double lambda_min;  /* cm  */
  double lambda_max;  /* cm  */
  uint64_t seed;
  double z0;          /* CONE: image plane */
  double start;       /* CONE: start height above the image plane */
} rb_raygen_desc;

/* ---- API ---------------------------------------------------------------- */
typedef struct rb_ctx rb_ctx;
typedef struct rb_scene rb_scene;

/* Create a context on HIP device `device` with its own stream (default
 * priority). Contexts are not to be shared across host threads. */
int rbgpu_ctx_create(int device, rb_ctx** out);
void rbgpu_ctx_destroy(rb_ctx* ctx);
const char* rbgpu_last_error(const rb_ctx* ctx);

/* Upload a flattened scene (the AOpticsManager's closed TGeo tree). The
 * library builds a BVH over each node's daughters on the host and keeps the
 * scene in HBM. Replaces the geometry held by AOpticsManager (a TGeoManager,
 * include/AOpticsManager.h:28) after CloseGeometry(). */
int rbgpu_scene_upload(rb_ctx* ctx, const rb_scene_desc* desc, rb_scene** out);
void rbgpu_scene_free(rb_scene* scene);

/* Trace n rays to termination: the non-sequential state machine of
 * AOpticsManager::TraceNonSequential(TObjArray*) (src/AOpticsManager.cxx:335-520)
 * applied to every running ray, the work ARayArray& dispatch does on host
 * threads (src/AOpticsManager.cxx:523-587). Blocks until done. */
int rbgpu_trace(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts,
                rb_rays* rays, int64_t n);

/* rbgpu_trace without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same kernels, same
 * results ray for ray as rbgpu_trace. The ray arrays must be device memory
 * (opts->device_ptrs set) and n <= RB_NOALLOC_MAX_RAYS, else -1 with a
 * message. The coherence order's keys and indices and the persistent
 * schedule's ray counter live in one file-scope device array of about
 * 65 MiB (RB_NOALLOC_MAX_RAYS x 16 B + 1 MiB), which each GPU the library
 * runs on holds until the process ends. A sorted trace here has no staged
 * ray lines: the kernels read and write the ray arrays through the
 * permutation, sorted within tiles of 2^19 rays. Every call on a GPU shares
 * that array: calls that overlap on one GPU (two contexts, two host threads)
 * race. Calls no hipMalloc, hipFree or allocating library routine.
 * Blocks until done. */
#define RB_NOALLOC_MAX_RAYS 4194304
int rbgpu_trace_noalloc(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts,
                        rb_rays* rays, int64_t n);

/* Fill n rays on the device (pointers in `rays` must be device pointers). */
int rbgpu_generate_rays(rb_ctx* ctx, const rb_raygen_desc* gen, rb_rays* rays,
                        int64_t n);

/* ---- ARayShooter on the device ------------------------------------------
 * The ray sources of src/ARayShooter.cxx, written straight into HBM. Every
 * kind places its local pattern with `rot` (positions, and the direction
 * where the reference rotates it) and then `tr` (positions only), as
 * TGeoRotation/TGeoTranslation::LocalToMaster do there; a null rotation or
 * translation of the reference is the identity / zero here. Directions are
 * normalised as ARay::SetDirection does (src/ARay.cxx:217-223; a zero vector
 * is left (0, 0, 0)). Rays start with t = 0, status RB_RUN, one point.
 * The random kinds draw from Philox streams keyed by (seed, id_base + i):
 * the reference draws from the shared gRandom, so only the distribution can
 * agree, and it does (the disk and sphere are drawn by inverse CDF where the
 * reference rejects: the same uniform measure). */
enum {
  RB_SHOOT_RECTANGLE = 0,             /* Rectangle(lambda, dx=a, dy=b, nx, ny, rot, tr, v) :401-452;
                                         Square(lambda, d, n, ...) :455-459 is a = b = d, nx = ny = n */
  RB_SHOOT_CIRCLE = 1,                /* Circle(lambda, rmax=a, nr=nx, nphi=ny, rot, tr, v) :122-183   */
  RB_SHOOT_RANDOM_CIRCLE = 2,         /* RandomCircle(lambda, rmax=a, n, rot, tr, v) :186-237          */
  RB_SHOOT_RANDOM_CONE = 3,           /* RandomCone(lambda, r=a, d=b, n, rot, tr) :240-285 (dir unused) */
  RB_SHOOT_RANDOM_RECTANGLE = 4,      /* RandomRectangle(lambda, dx=a, dy=b, n, rot, tr, v) :288-335;
                                         RandomSquare(lambda, d, n, ...) :394-398 is a = b = d        */
  RB_SHOOT_RANDOM_SPHERE = 5,         /* RandomSphere(lambda, n, tr) :338-356 (rot and dir unused)     */
  RB_SHOOT_RANDOM_SPHERICAL_CONE = 6  /* RandomSphericalCone(lambda, n, theta=a [deg], rot, tr)
                                         :359-391 (dir unused)                                        */
};

typedef struct {
  int32_t kind;
  int32_t nx;        /* RECTANGLE: nx; CIRCLE: nr                                 */
  int32_t ny;        /* RECTANGLE: ny; CIRCLE: nphi                               */
  int32_t pad;
  double lambda;     /* cm                                                        */
  double a;          /* see the kinds above                                       */
  double b;
  double rot[9];     /* rotation matrix, row-major, local -> master               */
  double tr[3];      /* translation                                               */
  double dir[3];     /* the `v` direction argument ((0, 0, 1) when the reference's
                        v is null)                                                */
  uint64_t seed;     /* Philox key of the random kinds                            */
  uint64_t id_base;  /* stream of ray i is id_base + i                            */
} rb_shooter_desc;

/* Number of rays the reference's function makes for `desc` (n: its ray-count
 * argument, used by the random kinds): nx * ny for RECTANGLE,
 * 1 + nphi * nr (nr + 1) / 2 for CIRCLE, n for the random kinds, and 0 where
 * the reference returns an empty array (negative sizes, counts below 1).
 * Returns -1 for an unknown kind. */
int64_t rbgpu_shooter_count(const rb_shooter_desc* desc, int64_t n);

/* Write the rays of `desc` into `rays` (device pointers); n must equal
 * rbgpu_shooter_count(desc, n). */
int rbgpu_shoot_rays(rb_ctx* ctx, const rb_shooter_desc* desc, rb_rays* rays, int64_t n);

/* ---- CORSIKA photon bunches -> rays ---------------------------------------
 * The conversion of ACorsikaIACTFile::GetRayArray(telNo, arrayNo, z,
 * refractiveIndex) (src/ACorsikaIACTFile.cxx:71-131) over the bunch table its
 * reader fills (one entry per bunch, the TTree branches of :330-342, FP32 as
 * there; cz = -sqrt(1 - cx^2 - cy^2) as the reader computes it, :452). */
typedef struct {
  const int32_t* tel_no;
  const int32_t* array_no;
  const float* x;        /* cm, at the telescope's observation level     */
  const float* y;
  const float* zem;      /* emission height (carried, not used)          */
  const float* time;     /* ns                                           */
  const float* cx;
  const float* cy;
  const float* cz;
  const float* lambda;   /* nm; 0 = not set by CORSIKA                   */
  const float* photons;  /* photons in the bunch (<= 1 << 24)            */
} rb_bunches;

typedef struct {
  int32_t tel_no;            /* telNo                                              */
  int32_t array_no;          /* arrayNo                                            */
  double z;                  /* cm: start height of the photons (argument z)       */
  double tel_z;              /* cm: GetTelescopeZ(telNo)                           */
  double refractive_index;   /* argument refractiveIndex                           */
  double lambda_min;         /* nm: fMinWavelength                                 */
  double lambda_max;         /* nm: upper end of the drawn wavelength; the reference
                                uses fMaxPhotonBunches here (:123-125)             */
  uint64_t seed;             /* Philox key of the wavelength draws                 */
  uint64_t id_base;          /* stream of output ray i is id_base + i              */
  int32_t device_ptrs;       /* the bunch arrays are GPU memory                    */
  int32_t pad;
} rb_bunch_opts;

/* Each bunch of telescope tel_no and array array_no becomes as many rays as
 * the reference's `for (j = 0; j < photons; j++)` makes (ceil(photons), none
 * for photons <= 0), in bunch order: start (x - D cx, y - D cy, z) with
 * D = (z - tel_z) (-1 / cz), time time * 1 ns - D n / c, direction
 * (cx, cy, cz) normalised, wavelength lambda nm, or for lambda == 0
 * 1 / (1 / lambda_min - U (1 / lambda_min - 1 / lambda_max)) nm.
 * *n_rays receives the number of rays. With rays == NULL only the count is
 * computed; otherwise `rays` (device pointers) must hold `capacity` rays and
 * the call fails without writing rays when the count exceeds it. */
int rbgpu_bunches_to_rays(rb_ctx* ctx, const rb_bunches* bunches, int64_t n_bunches, const rb_bunch_opts* opts,
                          rb_rays* rays, int64_t capacity, int64_t* n_rays);

/* ---- adjoint of a focal-plane loss (differentiable alignment) ------------
 * Supported optics: chains of up to 4 reflections, each on
 *  - a spherical facet (node_param): the node shape
 *    `sphere(rmin = R) : T(0, 0, R) * cut` placed with rotation
 *    M = M0 Rx(tip) Ry(tilt) (a Davies-Cotton facet), or
 *  - an AGeoAsphericDisk surface (surface_param): the disk is the node's own
 *    shape; the surface the ray reflects on (F1 or F2, picked as the trace's
 *    normal picks it, src/AGeoAsphericDisk.cxx:246-328) has parameters
 *    (c, kappa, K_1..K_n): its sag F(r) = z + r^2 c / (1 + sqrt(1 - r^2 c^2 kappa))
 *    + sum_k K_k r^2k (:126-153), n <= RB_ADJ_MAX_POLY, with kappa = conic + 1
 *    (SetConicConstants, :1007-1013), so d/dkappa is d/dconic;
 * ending on a focal surface: the plane z = z_cam, or, when the focal node's
 * shape is an AGeoAsphericDisk, the disk face the ray enters (a surface or a
 * wall; z_param then moves the focal node along its own z axis). Rays that
 * reflect elsewhere (a mirror disk's side wall, a node without parameters)
 * or refract give no gradient. The loss seen by each focused ray, on its hit (h_x, h_y) and
 * with u = h - centroid:
 *   CENTROID, MOMENTS: weight * |u|^2 / 2 (weight = 2 / N and centroid = mean
 *     hit give d(spot RMS^2)/dparams);
 *   ENCIRCLED: weight * sigmoid((r0^2 - |u|^2) / (2 r0 w)), a smooth count of
 *     the hits within r0 of the centroid point (width w), which is a fixed
 *     point here (the target); weight = -1 / N makes minimising it maximise
 *     the encircled energy fraction. */
#define RB_ADJ_MAX_POLY 10
typedef struct {
  const int32_t* node_param;   /* device [n_nodes][3]: tip, tilt, R parameter index or -1; an index
                                  outside [0, n_params) counts as -1 (the ray gives no gradient) */
  const double* param_values;  /* device [n_params]: current tip, tilt (rad), R (cm), ... */
  int32_t n_params;
  int32_t z_param;             /* index of the focal-plane z offset parameter, or -1 */
  int32_t focal_node;          /* node id whose entry plane carries z_param, or -1 */
  int32_t mode;                /* RB_ADJ_CENTROID or RB_ADJ_MOMENTS */
  double centroid_x, centroid_y; /* CENTROID: the centroid; MOMENTS with no ref_xy: the reference point */
  double weight;                 /* RB_ADJ_CENTROID only */
  /* RB_ADJ_MOMENTS only: the rays form n_groups consecutive groups of
   * group_size rays (ray i belongs to group i / group_size), each with its
   * own reference point ref_xy[2 g], ref_xy[2 g + 1] (device memory) and its
   * own block of 3 n_params + 4 sums in grad. n_groups = 0 means one group
   * over all n rays, referenced to (centroid_x, centroid_y). */
  const double* ref_xy;
  int64_t group_size;
  int32_t n_groups;
  int32_t pad;
  /* AGeoAsphericDisk mirrors: device [n_nodes][2] = (first, count) or NULL.
   * A node with first >= 0 has the parameters param[first + 0] = c (1/cm),
   * [first + 1] = kappa, [first + 1 + k] = K_k (cm^(1-2k)) of the surface a
   * ray reflects on; count must be >= 2 + n of that surface and
   * first + count <= n_params. The gradient is with respect to those values;
   * the caller keeps them equal to the scene's (param_values is not read for
   * these nodes). */
  const int32_t* surface_param;
  double ee_radius, ee_width;    /* RB_ADJ_ENCIRCLED: r0 > 0 and w > 0 (cm) */
} rb_adjoint_desc;

/* RB_ADJ_CENTROID: grad[p] += weight * sum over focused rays of
 *   (h - centroid) . dh/dp                    (grad has n_params entries).
 * RB_ADJ_MOMENTS: one pass that needs no centroid. With u = h - r, r the
 * group's reference point, each group's block of grad receives
 *   [0, P)    sum u . dh/dp
 *   [P, 2P)   sum dh_x/dp
 *   [2P, 3P)  sum dh_y/dp
 *   3P..3P+3  N, sum u_x, sum u_y, sum |u|^2  (P = n_params),
 * from which the caller forms, with m = (sum u) / N (centroid c = r + m),
 *   spot RMS^2 = sum |u|^2 / N - |m|^2 and its gradient
 *   (2 / N) (sum u.dh/dp - m_x sum dh_x/dp - m_y sum dh_y/dp).
 * A reference point near the centroid keeps the FP32 products small. The
 * accumulators of all groups must fit: groups x (3 P + 4) <= 6400. */
/* RB_ADJ_ENCIRCLED: grad[p] += sum over focused rays of d(loss)/dp, and
 * grad[n_params] += the summed loss (grad has n_params + 1 entries). */
enum { RB_ADJ_CENTROID = 0, RB_ADJ_MOMENTS = 1, RB_ADJ_ENCIRCLED = 2 };

/* Re-trace the n rays (left unchanged) and accumulate into grad (device,
 * FP64; added to, not overwritten) as `adj->mode` selects. Per-ray
 * derivative arithmetic is FP32, accumulation FP64. No counterpart in the
 * reference: Optimize.C fits with Minuit and no gradient
 * (tutorials/Optimize.C:188-205). */
int rbgpu_adjoint(rb_ctx* ctx, const rb_scene* scene, const rb_trace_opts* opts, const rb_rays* rays, int64_t n,
                  const rb_adjoint_desc* adj, double* grad);

/* Milliseconds the last rbgpu_trace (or rbgpu_adjoint,
 * rbgpu_containment_radius, rbgpu_shoot_rays, rbgpu_bunches_to_rays) spent on
 * the device, ordering (histogram fill, bunch counting) included, timed with
 * HIP events on the context's stream, and the number of kernel launches it
 * made. */
double rbgpu_last_trace_ms(const rb_ctx* ctx);
/* Milliseconds of the last call's trace (adjoint, containment search, ray
 * source, bunch expansion) kernels alone, without the ordering (fill,
 * counting) kernels. No counterpart in the reference (measurement). */
double rbgpu_last_kernel_ms(const rb_ctx* ctx);
int rbgpu_last_trace_launches(const rb_ctx* ctx);

/* ---- spot analysis after the trace -------------------------------------- */
/* Binning of the spot histogram, as ROOT's TH2D("", "", nx, xlo, xhi, ny,
 * ylo, yhi) that the reference's callers fill with the focal-plane hits
 * (tutorials/MST.C:272-277, tutorials/unittest_robast.py:590-610). */
typedef struct rb_hist2d_desc {
  int32_t nx, ny;          /* bins per axis; nx * ny <= 1 << 26               */
  double xlo, xhi;         /* x axis range, xhi > xlo                          */
  double ylo, yhi;         /* y axis range, yhi > ylo                          */
  int32_t status;          /* the rays counted (RB_FOCUS: the focal spot)      */
  int32_t device_ptrs;     /* rays->x, y and status are GPU memory             */
} rb_hist2d_desc;

/* AGeoUtil::ContainmentRadius(TH2*, fraction, r, x, y)
 * (src/AGeoUtil.cxx:198-308) of the terminal (x, y) of the n rays with
 * status hist->status. The histogram is filled on the device as TH2::Fill
 * fills it (entries outside either axis go to no bin and stay out of the
 * statistics), and the radius search runs on the device.
 * out[0..2] = r, x, y. If not NULL, counts (host, ny * nx, bin (bx, by) at
 * (by - 1) * nx + bx - 1) and stats (host, {sum w, sum x, sum x^2, sum y,
 * sum y^2}) receive the histogram. 0 < fraction <= 1. */
int rbgpu_containment_radius(rb_ctx* ctx, const rb_rays* rays, int64_t n, const rb_hist2d_desc* hist,
                             double fraction, double* out, uint32_t* counts, double* stats);

/* Writes sizeof() of the ABI structs (rb_matrix, rb_shape_desc,
 * rb_volume_desc, rb_node_desc, rb_table_desc, rb_scene_desc, rb_rays,
 * rb_trace_opts, rb_raygen_desc, rb_adjoint_desc, rb_hist2d_desc,
 * rb_border_desc, rb_shooter_desc, rb_bunches, rb_bunch_opts,
 * rb_grid2d_desc) into out[0..15] for binding checks; returns 16.
 * No counterpart in the reference (binding self-check only). */
int rbgpu_abi_sizes(int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* RBGPU_H */
