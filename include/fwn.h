/*
 * fwn.h -- C ABI of the MI355X fast winding number engine.
 *
 * This is the drop-in boundary under lagrange::winding::FastWindingNumber
 * (modules/winding/include/lagrange/winding/FastWindingNumber.h:24-108).  The reference
 * class owns a float/int copy of the mesh and a HDK_Sample::UT_SolidAngle<float,float>
 * engine (modules/winding/src/FastWindingNumber.cpp:28-85); here the same role is played
 * by an opaque fwn_engine holding a GPU-built tree in HBM.  All functions are plain C:
 * pointers and sizes, status codes, no exceptions cross the boundary.
 *
 * Entry points and the reference interface each replaces:
 *   fwn_create            <- FastWindingNumber(const SurfaceMesh&) / Impl::initialize + UT_SolidAngle::init
 *                            (FastWindingNumber.cpp:31-58, 87-98)
 *   fwn_destroy           <- ~FastWindingNumber (FastWindingNumber.cpp:101)
 *   fwn_solid_angle_point <- FastWindingNumber::solid_angle (FastWindingNumber.cpp:69-76, 110-113)
 *   fwn_is_inside_point   <- FastWindingNumber::is_inside (FastWindingNumber.cpp:60-67, 105-108)
 *   fwn_solid_angle       <- the per-point calls above issued over an array of points, as
 *                            volume::mesh_to_volume does per voxel (modules/volume/src/mesh_to_volume.cpp:167-183)
 *   fwn_sign_voxel_box /
 *   fwn_sign_voxels       <- the interior-test lambda of volume::mesh_to_volume with Sign::WindingNumber
 *                            (modules/volume/src/mesh_to_volume.cpp:175-183), voxel centres made on the GPU
 *   fwn_orientation_criterion <- the per-facet probe pair of modules/winding/examples/fix_orientation.cpp:86-115
 *   fwn_sample_inside     <- the rejection-sampling loop of modules/winding/examples/sample_points_in_mesh.cpp:59-76
 *   fwn_closest_point     <- bvh::TriangleAABBTree::get_closest_point (modules/bvh/src/TriangleAABBTree.cpp:86-127)
 *   fwn_create_multi /
 *   fwn_solid_angle_multi <- no single reference call: many FastWindingNumber objects queried at once
 *   fwn_export_tree, fwn_query_stats, fwn_bruteforce_f64 -- measurement and verification hooks
 *                            (no reference counterpart).
 *
 * Threading: an engine is immutable after creation.  Query calls on one engine serialise on
 * an internal mutex (they share the engine's stream and query scratch); calls on different
 * engines run concurrently.
 */
#ifndef FWN_H
#define FWN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------------- */
enum {
    FWN_OK = 0,
    FWN_INVALID_ARGUMENT = 1,
    FWN_OUT_OF_MEMORY = 2,
    FWN_NO_DEVICE = 3,
    FWN_HIP_ERROR = 4,
    FWN_LIMIT_EXCEEDED = 5
};

/* ---- flags for query calls -------------------------------------------------------------- */
enum {
    FWN_DEVICE_POINTERS = 1 /* Q / outputs are device pointers (already resident in HBM) */
};

typedef struct fwn_engine fwn_engine;

typedef struct fwn_options {
    float accuracy_scale;     /* beta: far field when |q-p|^2 > beta^2 r^2. Reference default 2 */
    float near_fp64_rel;      /* tau = this x bbox diagonal (default 1e-4; 0 disables FP64): a leaf's
                                 triangles use the FP64 solid angle when q is within tau of the
                                 leaf's bounding sphere and, for a single triangle, within tau of
                                 its plane */
    int32_t leaf_max;         /* max triangles per leaf bucket (1..8, default 1: every triangle has
                                 its own far-field expansion, as the reference's items; with the
                                 SAH builder this is the reference layout's tree.  Buckets of 2..8
                                 traverse faster (fwn_create_tuned picks one) but a bucket can take
                                 the far field where single triangles would be summed exactly) */
    int32_t fp64_moment_min;  /* tree nodes with at least this many triangles merge moments in FP64 */
    int32_t device;           /* HIP device ordinal */
    int32_t waves_per_block;  /* query launch shape: 1, 2, 4 or 8 waves per block (default 1) */
    int32_t tree_width;       /* slots per tree node: 2, 4 or 8 (default 4) */
    int32_t launch_mode;      /* FWN_LAUNCH_GRID (default) or FWN_LAUNCH_PERSISTENT */
    int32_t tree_builder;     /* FWN_TREE_SAH (default) or FWN_TREE_LBVH */
    int32_t moment_pass;      /* FWN_MOMENTS_LEVELS (default) or FWN_MOMENTS_CLIMB; both compute the
                                 same merges in the same order per node (bit-identical trees) */
    int32_t reserved[2];
} fwn_options;

/* bottom-up moment pass of the build.  FWN_MOMENTS_LEVELS: one launch per tree depth, deepest
 * first (trees deeper than 512 binary levels fall back to the climb).  FWN_MOMENTS_CLIMB: one
 * thread per leaf climbs to the root; the second child to arrive at a node merges it. */
enum { FWN_MOMENTS_LEVELS = 0, FWN_MOMENTS_CLIMB = 1 };

/* tree builders.  FWN_TREE_LBVH: Morton-ordered binary radix tree (one sort, one pass); a wide
 * node repeatedly opens its largest non-leaf slot until it has tree_width slots.
 * FWN_TREE_SAH: top-down binary splits chosen by the binned (16 bins on the longest centroid
 * axis) box-area cost, every node split down to single triangles; a wide node takes
 * log2(tree_width) whole binary levels (a 4-wide node: children and grandchildren), the
 * layout of the reference's 4-ary tree (modules/winding/src/FastWindingNumber.cpp:57 hands
 * the triangles to the dependency's UT_SolidAngle::init, which builds a 4-ary BVH).  With
 * leaf_max 1 the SAH tree matches the CPU restatement of that build slot for slot. */
enum { FWN_TREE_LBVH = 0, FWN_TREE_SAH = 1 };

/* query launch modes: one workgroup per group of 64-point batches, or a persistent grid whose
 * waves take batches from a work counter, meshes with the deepest trees first */
enum { FWN_LAUNCH_GRID = 0, FWN_LAUNCH_PERSISTENT = 1 };

/* ---- exported tree format (fwn_export_tree) ------------------------------------------------
 * The engine stores, per mesh, a tree of nodes with tree_width (<= FWN_WIDTH) slots each; node
 * n's slot s is at word (n * tree_width + s) * FWN_SLOT_WORDS, and a node's occupied slots come
 * first.  Each slot is 24 32-bit words (96 bytes: one 16-word and one 8-word scalar load).
 * Words 0..22 are float, word 23 is the slot's link:
 *   0-2   p     area-weighted centroid of the slot's triangles
 *   3     thr   beta^2 r2 in float: r2 is the squared distance from p to the farthest corner of
 *               the slot's box (rounded up, so every point of the slot's triangles lies within
 *               sqrt(r2) of p) and beta the engine's accuracy scale; the slot takes the far field
 *               for a query q when |p - q|^2 > thr
 *   4-6   N     sum of area-weighted normals
 *   7-12  A     -3 * (M1xx, M1yy, M1zz, M1xy+M1yx, M1yz+M1zy, M1zx+M1xz), plus tr M1 on xx, yy,
 *               zz, where M1_ji = sum_t A_t (c_t-p)_j n_t,i
 *   13-22 B     7.5 * (S_xxx, S_yyy, S_zzz, 3S_xxy, 3S_xxz, 3S_yyx, 3S_yyz, 3S_zzx, 3S_zzy, 6S_xyz),
 *               plus C_x on xxx, yyx, zzx, C_y on yyy, xxy, zzy and C_z on zzz, xxz, yyz, where
 *               S is the fully symmetrised second moment int (x-p)_j (x-p)_k n_i dA and
 *               C = -4.5 t, t_k = sum_i S_iik
 *   With d = p - q the far field is |d|^-3 (d.N + |d|^-2 (q1 + |d|^-2 q2)), q1 the quadratic form
 *   of A and q2 the cubic of B: the expansion's constant terms tr M1 and C.d are folded into
 *   them through |d|^2 / |d|^2.
 *   23    link  internal slot: (child node << 4) | occupied slots of the child node (bit 31 clear);
 *               leaf slot: FWN_LINK_LEAF | (first << 3) | (count - 1), the slot's triangles being
 *               positions [first, first + count) of the sorted triangle array (count <= 8);
 *               unoccupied slot: FWN_LINK_EMPTY (all other words 0)
 * Triangles are stored in tree order, FWN_TRI_WORDS words each: v0, v1, v2 (9 floats), original
 * facet index (int32), then two floats of the FP64 near-field rule of the leaf whose first
 * triangle this is (0 for the other triangles of a bucket): (sqrt(r2) + tau)^2 and
 * tau^2 |N|^2, with tau the mesh's FP64 distance (fwn_options.near_fp64_rel).
 * Limits: fewer than 2^28 - 1 triangles and 2^27 nodes per engine.
 */
#define FWN_WIDTH 8 /* maximum tree_width */
#define FWN_SLOT_WORDS 24
#define FWN_TRI_WORDS 12
#define FWN_LINK_LEAF 0x80000000u
#define FWN_LINK_EMPTY 0xFFFFFFFFu

enum {
    FWN_W_PX = 0, FWN_W_PY = 1, FWN_W_PZ = 2,
    FWN_W_THR = 3, /* beta^2 r^2 in float: the far-field threshold for the engine's accuracy scale */
    FWN_W_NX = 4, FWN_W_NY = 5, FWN_W_NZ = 6,
    FWN_W_A = 7,  /* 6 words */
    FWN_W_B = 13, /* 10 words */
    FWN_W_LINK = 23
};

/* words of a stored triangle */
enum { FWN_T_FACET = 9, FWN_T_NEAR2 = 10, FWN_T_PLANE = 11 };

typedef struct fwn_tree_info {
    int64_t num_meshes;
    int64_t num_nodes;       /* wide nodes over all meshes */
    int64_t num_triangles;   /* over all meshes */
    int32_t width;           /* slots per node (tree_width) */
    int32_t max_depth;       /* deepest wide-node level (root = 1) */
    double build_ms;         /* device time of the last build */
} fwn_tree_info;

/* per-query work counters written by fwn_query_stats (int32 each) */
enum { FWN_STAT_TESTS = 0, FWN_STAT_FAR = 1, FWN_STAT_TRI32 = 2, FWN_STAT_TRI64 = 3, FWN_STAT_COUNT = 4 };

/* ---- lifetime ------------------------------------------------------------------------- */
int fwn_options_default(fwn_options* opt);

/* One mesh.  V: nV x 3 float (row-major), F: nF x 3 int32, host pointers.  The engine copies
 * what it needs; the caller's arrays may be freed afterwards (FastWindingNumber.cpp:39-57). */
int fwn_create(fwn_engine** out, const float* V, int64_t nV, const int32_t* F, int64_t nF,
               const fwn_options* opt);

/* Many meshes in one engine.  Mesh m owns vertices [v_off[m], v_off[m+1]) and facets
 * [f_off[m], f_off[m+1]) of V / F; facet indices are local to the mesh.  Offsets have
 * n_meshes+1 entries.  All meshes are built together by the same kernels. */
int fwn_create_multi(fwn_engine** out, int32_t n_meshes, const float* V, const int64_t* v_off,
                     const int32_t* F, const int64_t* f_off, const fwn_options* opt);

int fwn_destroy(fwn_engine* e);

/* ---- queries -------------------------------------------------------------------------- */
/* Solid angle (steradians, as FastWindingNumber::solid_angle) of nQ points Q (nQ x 3 float).
 * inside_out (nullable) receives (omega / (4 pi) > 0.5) evaluated in double, byte per point
 * (FastWindingNumber.cpp:66).  flags: FWN_DEVICE_POINTERS.  Results do not depend on how the
 * points are batched: host calls above 2^21 points are answered in pipelined chunks, calls of at
 * most 2,048 points without the Morton sort. */
int fwn_solid_angle(fwn_engine* e, const float* Q, int64_t nQ, float* omega_out, uint8_t* inside_out,
                    int32_t flags);

/* No-allocation variant of fwn_solid_angle (no counterpart in the reference), for callers that
 * cannot allocate device memory during a call: code in HIP graph capture regions, tight serving
 * loops, harnesses that forbid hipMalloc.  Same results as fwn_solid_angle with
 * FWN_DEVICE_POINTERS, bit for bit, for a single-mesh engine and at most FWN_NOALLOC_MAX_POINTS
 * points (more: FWN_LIMIT_EXCEEDED).  Q (nQ x 3 float), omega_out and inside_out (nullable) are
 * device pointers.  The call allocates nothing (no hipMalloc / hipFree, no pinned host memory):
 * its scratch (point keys and order, sort histograms, call metadata and flags, about 17 MB) is one
 * fixed device array of the library, which every GPU the library runs on holds until the process
 * ends.  Calls on one GPU share that array and would race if they overlapped: the library
 * serialises them with a process-wide lock, and each call returns only after its kernels have
 * finished. */
#define FWN_NOALLOC_MAX_POINTS (1 << 20)
int fwn_solid_angle_noalloc(fwn_engine* e, const float* Q, int64_t nQ, float* omega_out, uint8_t* inside_out);

/* Points of every mesh in one launch: points of mesh m are [q_off[m], q_off[m+1]) of Q. */
int fwn_solid_angle_multi(fwn_engine* e, const float* Q, const int64_t* q_off, float* omega_out,
                          uint8_t* inside_out, int32_t flags);

/* Winding-number sign of voxels, batched: the interior test lagrange::volume::mesh_to_volume
 * calls once per voxel with Sign::WindingNumber (modules/volume/src/mesh_to_volume.cpp:172-183).
 * Voxel (i, j, k) has centre float(voxel_size * (i, j, k) + voxel_size / 2), evaluated in double
 * as the reference's transform does (:147-149, :176-181); inside_out[t] = is_inside(centre).
 * voxel_size is absolute (> 0): a relative size is resolved by the caller against its mesh
 * (:132-145).  The centres are generated on the device.  omega_out is optional.
 *
 * fwn_sign_voxel_box: every voxel lo + (x, y, z), 0 <= x < dims[0] (fastest), y < dims[1],
 * z < dims[2]; output t = x + dims[0] (y + dims[1] z).
 * fwn_sign_voxels: n explicit voxels, ijk is n x 3 int32.
 * flags: FWN_DEVICE_POINTERS applies to ijk and the outputs. */
int fwn_sign_voxel_box(fwn_engine* e, const int32_t lo[3], const int32_t dims[3], double voxel_size,
                       uint8_t* inside_out, float* omega_out, int32_t flags);
int fwn_sign_voxels(fwn_engine* e, const int32_t* ijk, int64_t n, double voxel_size, uint8_t* inside_out,
                    float* omega_out, int32_t flags);

/* Orientation criterion of every facet of an input mesh against this engine's (reference) mesh,
 * batched: modules/winding/examples/fix_orientation.cpp:86-115.  Facet f = (a, b, c) of
 * (V: nV x 3 float, F: nF x 3 int32) gets n = normalized((b - a) x (c - a)), g = (a + b + c) / 3,
 * and criterion_out[f] = (omega(g + eps n) - omega(g - eps n)) / (4 pi), the float difference
 * divided in double.  The example flips facets whose criterion exceeds its threshold (0.8).
 * The 2 nF probe points are generated on the device.  flags: FWN_DEVICE_POINTERS (V, F, out). */
int fwn_orientation_criterion(fwn_engine* e, const float* V, int64_t nV, const int32_t* F, int64_t nF,
                              float epsilon, double* criterion_out, int32_t flags);

/* Points uniform in the box [lo, hi] kept when inside the mesh, batched:
 * modules/winding/examples/sample_points_in_mesh.cpp:59-76.  Sample i has coordinates
 * (hi - lo) * u + lo with u from a counter-based generator of (seed, i) (24-bit uniform in
 * [0, 1)); the inside samples are written to points_out (capacity n_samples x 3 floats) in
 * sample order and counted in *n_inside_out.  flags: FWN_DEVICE_POINTERS (points_out). */
int fwn_sample_inside(fwn_engine* e, const float lo[3], const float hi[3], int64_t n_samples, uint64_t seed,
                      float* points_out, int64_t* n_inside_out, int32_t flags);

/* Closest point on the mesh (unsigned distance) of nQ points, over the same tree:
 * lagrange::bvh::TriangleAABBTree::get_closest_point (modules/bvh/src/TriangleAABBTree.cpp:86-127)
 * with the per-triangle squared distance of lagrange::point_triangle_squared_distance (Eberly's
 * region method in double, modules/core/include/lagrange/utils/point_triangle_squared_distance.h).
 * sq_dist_out (double) and facet_out (input facet index) per point; closest_out (nQ x 3 double)
 * is optional.  The reference rounds both to the mesh scalar; here they stay double.  On an
 * empty mesh facet_out is -1 and sq_dist_out +inf.  Equidistant facets may be reported by either
 * index.  flags: FWN_DEVICE_POINTERS. */
int fwn_closest_point(fwn_engine* e, const float* Q, int64_t nQ, double* sq_dist_out, int32_t* facet_out,
                      double* closest_out, int32_t flags);

/* Single point, host memory, as the reference's per-point API.  Concurrent calls on one engine
 * (callers on several threads, as the reference is driven from TBB workers) are answered
 * together: a caller that finds no batch running answers every pending point in one batched
 * query and wakes the others; each gets exactly the value of a batched call. */
int fwn_solid_angle_point(fwn_engine* e, const float q[3], float* omega_out);
int fwn_is_inside_point(fwn_engine* e, const float q[3], uint8_t* inside_out);

/* ---- measurement / verification hooks ---------------------------------------------------- */
int fwn_tree_info_get(const fwn_engine* e, fwn_tree_info* info);

/* Copy the tree to host memory.  slots: num_nodes*width*FWN_SLOT_WORDS words; tris:
 * num_triangles*FWN_TRI_WORDS words; roots: 2 int32 per mesh (root node, occupied slots);
 * tau: one float per mesh (absolute FP64 near-field distance).  Any pointer may be NULL. */
int fwn_export_tree(const fwn_engine* e, uint32_t* slots, uint32_t* tris, int32_t* roots, float* tau);

/* Same traversal as fwn_solid_angle (single mesh: q_off NULL) but also writes FWN_STAT_COUNT
 * int32 work counters per point (stats_out, host or device per flags): slots tested, far-field
 * terms, FP32 triangles, FP64 triangles. */
int fwn_query_stats(fwn_engine* e, const float* Q, const int64_t* q_off, int64_t nQ, float* omega_out,
                    int32_t* stats_out, int32_t flags);

/* FP64 brute force on the GPU: omega_out[i] = sum over all facets of mesh 0 of the exact
 * triangle solid angle (van Oosterom-Strackee, FP64).  Host pointers unless flags says device. */
int fwn_bruteforce_f64(fwn_engine* e, const float* Q, int64_t nQ, double* omega_out, int32_t flags);

/* Launch shape for later query calls on this engine (no reference counterpart: tuning hook).
 * waves_per_block 1, 2, 4 or 8; mode FWN_LAUNCH_GRID or FWN_LAUNCH_PERSISTENT. */
int fwn_set_launch(fwn_engine* e, int32_t waves_per_block, int32_t mode);

/* Autotuner (no reference counterpart): times every launch shape (waves per block x grid or
 * persistent) on the given points (same layout as fwn_solid_angle_multi; q_off NULL for a
 * single-mesh engine), keeps the fastest on the engine, and returns it.  When cache_path is
 * not NULL the choice is read from / appended to that text file, keyed by device name, tree
 * width, leaf size, number of meshes and point count (log2 buckets). */
int fwn_autotune(fwn_engine* e, const float* Q, const int64_t* q_off, int64_t nQ, int32_t flags,
                 const char* cache_path, int32_t* waves_out, int32_t* mode_out, double* best_ms_out);

/* Build-option tuner (no reference counterpart): builds one engine per leaf size in leaf_sizes
 * (other options from opt, NULL = defaults), autotunes each one's launch shape on the given
 * points (as fwn_autotune, q_off NULL with nQ points for a single mesh) and keeps the one whose
 * query call is fastest; the others are destroyed.  Same mesh layout as fwn_create_multi
 * (n_meshes = 1, v_off = {0, nV}, f_off = {0, nF} for one mesh). */
int fwn_create_tuned(fwn_engine** out, int32_t n_meshes, const float* V, const int64_t* v_off, const int32_t* F,
                     const int64_t* f_off, const fwn_options* opt, const float* Q, const int64_t* q_off, int64_t nQ,
                     int32_t flags, const int32_t* leaf_sizes, int32_t n_leaf_sizes, int32_t* leaf_out,
                     double* best_ms_out);

/* Average device duration (ms) of the traversal kernel in the most recent query call and the
 * number of launches it took; measured with HIP events on the engine's stream. */
int fwn_last_timing(const fwn_engine* e, double* traverse_ms, double* total_ms);

/* ---- learned surrogate (no reference counterpart) ------------------------------------------
 * A multi-resolution hash-grid encoding feeding a 4 x 128 ReLU MLP (BF16 MFMA, FP32 accumulate),
 * trained on the GPU for the mesh of a single-mesh engine on winding numbers w = omega / (4 pi)
 * that the exact engine generates.  Queries fall back to the exact engine when the prediction is
 * within `band` of 0.5 or the point lies in a cell of the distance-bound grid (cells touched by a
 * triangle's box grown by dist_rel x bbox diagonal). */
typedef struct fwn_surrogate fwn_surrogate;

typedef struct fwn_surrogate_options {
    int32_t levels;        /* hash-grid levels (16) */
    int32_t log2_table;    /* entries per level = 2^log2_table (19) */
    int32_t base_res;      /* coarsest level resolution (16) */
    int32_t max_res;       /* finest level resolution (2048) */
    int32_t hidden_layers; /* 4 hidden layers of width 128 */
    int32_t batch;         /* training batch (65536) */
    int32_t steps;         /* Adam steps (3000) */
    int32_t grid_res;      /* distance-bound grid cells per axis (1024) */
    float lr;              /* Adam learning rate (5e-3; x0.1 decay over the last third) */
    float band;            /* fallback band around w = 0.5 (0.2) */
    float dist_rel;        /* distance bound as a fraction of the bbox diagonal (2e-3) */
    float near_frac;       /* fraction of training points drawn near the surface (0.5) */
    int64_t n_samples;     /* training points labelled by the exact engine (20,000,000) */
    uint64_t seed;
} fwn_surrogate_options;

typedef struct fwn_surrogate_info {
    int32_t levels;
    int32_t res[16];       /* per-level resolution */
    int32_t dense[16];     /* 1 when the level is stored densely (no hashing) */
    int64_t table_floats;  /* levels x 2^log2_table x 2 */
    int64_t mlp_floats;    /* packed W0 b0 W1 b1 W2 b2 W3 b3 w4 b4 (row-major K x N weights) */
    float delta;           /* absolute distance bound */
    float last_loss;       /* mean squared error of the last training batch */
    double train_ms;       /* sampling + labelling + training, device time */
    double encode_ms;      /* last query / predict call: Morton sort of the points (batches of 65,536 or more) */
    double mlp_ms;         /* last query / predict call: fused encoding + MLP kernel */
    double fallback_ms;    /* last query call: compaction + exact engine on the fallback points */
    double query_ms;       /* last query call: whole call on the device */
} fwn_surrogate_info;

int fwn_surrogate_options_default(fwn_surrogate_options* opt);
/* The surrogate keeps a pointer to `exact`, which must outlive it (destroy the surrogate first). */
int fwn_surrogate_create(fwn_surrogate** out, fwn_engine* exact, const fwn_surrogate_options* opt);
int fwn_surrogate_destroy(fwn_surrogate* s);
int fwn_surrogate_train(fwn_surrogate* s);
int fwn_surrogate_info_get(const fwn_surrogate* s, fwn_surrogate_info* info);

/* Winding numbers w (= omega / (4 pi)) and is_inside of nQ points: surrogate where confident,
 * exact engine for the fallback points (n_fallback_out: how many).  flags: FWN_DEVICE_POINTERS. */
int fwn_surrogate_query(fwn_surrogate* s, const float* Q, int64_t nQ, float* w_out, uint8_t* inside_out,
                        int32_t flags, int64_t* n_fallback_out);

/* Raw network predictions and fallback flags (no exact evaluation). */
int fwn_surrogate_predict(fwn_surrogate* s, const float* Q, int64_t nQ, float* pred_out, uint8_t* fallback_out,
                          int32_t flags);

/* Parameter access (host arrays, FP32 masters): copies out when *_out is not NULL, then in when
 * *_in is not NULL. */
int fwn_surrogate_params(fwn_surrogate* s, float* table_out, float* mlp_out, const float* table_in,
                         const float* mlp_in);

/* One forward + backward pass on n host points with targets y (no update): mean squared error,
 * predictions, and the gradients of the table and the packed MLP (host arrays). */
int fwn_surrogate_gradients(fwn_surrogate* s, const float* Q, const float* y, int64_t n, float* loss_out,
                            float* pred_out, float* g_table_out, float* g_mlp_out);

/* Thread-local message describing the last failure on this thread. */
const char* fwn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FWN_H */
