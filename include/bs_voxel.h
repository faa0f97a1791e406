/*
 * bs_voxel.h -- C ABI of the MI355X implicit-modeling path (mesh -> narrow-band
 * SDF, CSG, offset, dual contouring) of baby_shark's `src/voxel` module.
 *
 * The reference exposes this path only as a Rust API (src/voxel/prelude.rs:1-4);
 * each entry point below names the Rust item it replaces.  Conventions:
 *   - plain pointers and sizes, opaque handles, one owner per handle;
 *   - every call is synchronous on the calling thread and returns bs_status;
 *     BS_ERR_EMPTY is the reference's `None`;
 *   - a bs_ctx owns one HIP stream on one device; calls on one context must not
 *     overlap (the Rust converters are `&mut self`, not re-entrant);
 *   - the last error message of the calling thread is bs_last_error().
 *
 * Volume data layout (device resident, exported canonically):
 *   leaves sorted by origin (x, then y, then z), origin = voxel index & ~7;
 *   mask: 8 uint64 words per leaf, bit (i % 64) of word (i / 64) is voxel
 *         offset i = (x & 7) << 6 | (y & 7) << 3 | (z & 7)  (leaf_node/mod.rs:30-36);
 *   values: 512 float per leaf in offset order.
 */
#ifndef BS_VOXEL_H
#define BS_VOXEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BS_OK = 0,
    BS_ERR_EMPTY = 1,            /* reference returns None (empty mesh)      */
    BS_ERR_INVALID_ARG = 2,      /* bad pointer/size/parameter               */
    BS_ERR_HIP = 3,              /* HIP runtime error                        */
    BS_ERR_UNSUPPORTED_TILE = 4, /* reference panics with todo!() on a tile  */
    BS_ERR_INTERNAL = 5,         /* reference panics (unreachable!/unwrap)   */
    BS_ERR_NO_DEVICE = 6,        /* no HIP device / library built without GPU */
    BS_ERR_IO = 7,               /* a file could not be read or written       */
    BS_ERR_UNSUPPORTED = 8,      /* reserved: a method the library lacks      */
    BS_ERR_CAPACITY = 9          /* input or output beyond a caller-sized or fixed buffer */
} bs_status;

typedef struct bs_ctx bs_ctx;
typedef struct bs_volume bs_volume;

/* Mesh-to-volume parameters (mesh_to_volume.rs:17-50, defaults :223-236). */
typedef struct {
    float voxel_size;     /* with_voxel_size; default 1.0                         */
    int64_t band;         /* with_narrow_band_width; default 0                    */
    float wn_accuracy;    /* accuracy scale of approximate(); reference: 2.0      */
    float sign_threshold; /* wn < threshold => outside; reference: 0.2            */
    float fp64_band;      /* |wn - threshold| < fp64_band is re-evaluated in FP64 */
    int32_t sign_mode;    /* BS_SIGN_PER_VOXEL (the reference) or BS_SIGN_SHARED  */
} bs_m2v_params;

/* Sign modes.  PER_VOXEL evaluates the winding number at every active voxel,
 * as compute_sings does (mesh_to_volume.rs:198-215).  SHARED evaluates it once
 * per group of a leaf's voxels joined by face links with |d1| + |d2| larger
 * than the voxel size (plus a rounding margin): no surface point lies between
 * two such voxels, so for a closed mesh (constant winding number off the
 * surface) the group shares one sign.  Not for open meshes. */
#define BS_SIGN_PER_VOXEL 0
#define BS_SIGN_SHARED 1
/* SURROGATE (no counterpart in the reference; closed meshes only): each leaf's
 * eight 4^3 tiles are gated by the tile classifier set with
 * bs_ctx_set_surrogate.  A tile it culls (predicted smallest |SDF| at or above
 * the threshold) takes one winding-number evaluation at its voxel of largest
 * |d|, r; each of its voxels v with |d_v| > |v - r| (plus a rounding margin)
 * shares r's sign -- no surface point lies in the ball of radius |d_v| about
 * v, which holds r -- and every other voxel, culled tile or not, is
 * evaluated itself.  The exact distances thus judge every culled tile. */
#define BS_SIGN_SURROGATE 2

typedef struct {
    float voxel_size;
    uint64_t n_leaves;
    uint64_t n_active;    /* active voxels in leaves                 */
    uint64_t n_tiles;     /* active internal tiles                   */
} bs_volume_info;

/* Work counters and kernel timings of the last bs_mesh_to_volume* call.  The
 * ms_* phases are HIP-event times on the context's stream (with profiling on):
 * the BVH is built on a second stream beside the distance pass, so ms_bvh is
 * the active-voxel list plus what remains of the BVH build after it. */
typedef struct {
    uint64_t n_tris, n_sub, n_pairs, n_leaves, n_active, n_candidates;
    uint64_t n_bvh_nodes, wn_node_visits, wn_leaf_tris, wn_far_evals, n_fp64_rechecks;
    uint64_t wn_leaf_visits, used_hash, used_lds_stack, bvh_height;
    double ms_total, ms_subdivide, ms_leaves, ms_distance, ms_bvh, ms_winding;
    uint64_t n_sign_evals; /* voxels whose winding number was evaluated          */
    double ms_sign_share;  /* sign groups, representatives and sign copy (SHARED) */
    /* SURROGATE: 4^3 tiles holding an active voxel, tiles culled, culled tiles
     * in which some voxel failed the exact check (misclassified), and the
     * voxels that failed it (evaluated themselves) */
    uint64_t n_sg_tiles, n_sg_culled, n_sg_misclassified, n_sg_fallback;
    /* passes of the winding-number tree build: level passes of the top-down
     * build, or clustering iterations of BS_PATH_PLOC (0 for the Morton split) */
    uint64_t n_bvh_iterations;
} bs_m2v_stats;

/* Work counters of the last bs_volume_offset call. */
typedef struct {
    uint64_t n_kept;      /* leaves left by remove_if                          */
    uint64_t n_pool;      /* leaves after the eight sweeps (empty ones included) */
    uint64_t n_out;       /* leaves of the result                               */
    uint64_t n_planes;    /* leaf planes swept over all sweeps                  */
    uint64_t n_launches;  /* kernel launches of the sweeps                      */
    uint64_t n_capacity;  /* leaf pool capacity of the successful attempt       */
    uint64_t used_hash;   /* leaf map was a hash (box too large for a grid)     */
    uint64_t n_swept;     /* leaf sweeps summed over the eight directions       */
    double ms_device;     /* first to last kernel of the call (stream events)   */
    double ms_wall;       /* the whole call on the host clock                    */
} bs_offset_stats;

/* Counters of the last CSG call (bs_volume_union/subtract/intersect). */
typedef struct {
    uint64_t n_a, n_b;        /* leaves of the two operands                   */
    uint64_t n_out;           /* leaves of the result                         */
    uint64_t n_paired;        /* result leaves combined from both operands    */
    double ms_flood;          /* leaf flood fill of both operands (events)    */
    double ms_assemble;       /* result assembly kernel (events)              */
    double ms_wall;           /* the whole call on the host clock             */
    /* host-clock parts of ms_wall: building the operands' upper levels from
     * their leaves (when they carry none), the upper-level flood fill, CSG
     * and output plan, device allocation and release, and the result's
     * upper levels */
    double ms_topology, ms_plan, ms_memory, ms_compact;
    /* ms_plan split: flags and upper-level flood fill, the tree CSG with the
     * collection of the result's leaves, the output order and plan */
    double ms_plan_fill, ms_plan_tree, ms_plan_order;
} bs_csg_stats;

/* Counters of the last bs_dual_contour* call (see below). */
typedef struct {
    uint64_t n_leaves, n_tris, used_hash;
    double ms_total;  /* device time from the first kernel to the last (stream events) */
} bs_dc_stats;

const char* bs_last_error(void);

/* Context: binds `device`, creates one default-priority HIP stream. */
bs_status bs_ctx_create(int device, bs_ctx** out);
void bs_ctx_destroy(bs_ctx* ctx);
/* The context's hipStream_t (as void*), for callers that time or order work. */
void* bs_ctx_stream(bs_ctx* ctx);
/* Instrumentation of bs_mesh_to_volume*, off by default: BS_PROFILE_EVENTS
 * times each phase with HIP events (adds syncs); BS_PROFILE_COUNTERS counts
 * winding-number node visits (wn_* in the stats) with a slower kernel
 * variant.  Results are identical either way. */
#define BS_PROFILE_EVENTS 1
#define BS_PROFILE_COUNTERS 2
bs_status bs_ctx_set_profiling(bs_ctx* ctx, int flags);
bs_status bs_ctx_last_m2v_stats(const bs_ctx* ctx, bs_m2v_stats* out);
/* Alternate code paths, normally chosen by size (for tests and diagnosis):
 * BS_PATH_FORCE_HASH      leaf allocation through the GPU hash even when a dense
 *                         leaf grid over the band's bounding box (<= 2^28 cells) fits;
 * BS_PATH_FORCE_LDS_STACK winding-number traversal with its LDS stack even for
 *                         trees low enough for the register stack (height < 62);
 * BS_PATH_LBVH            winding-number tree split by Morton order (Karras 2012)
 *                         also for meshes of 2^18 triangles or more, which
 *                         otherwise take the top-down binned build;
 * BS_PATH_PLOC            the clustered build (PLOC, Meister and Bittner 2018)
 *                         for those meshes instead (BS_PATH_LBVH wins if both
 *                         are set).  With any tree the distances are identical;
 *                         the winding numbers differ by the far-field error of
 *                         the trees, so a sign can differ only where the winding
 *                         number is that close to the threshold. */
#define BS_PATH_FORCE_HASH 1
#define BS_PATH_FORCE_LDS_STACK 2
#define BS_PATH_LBVH 4
#define BS_PATH_PLOC 8
bs_status bs_ctx_set_paths(bs_ctx* ctx, int flags);

void bs_m2v_params_default(bs_m2v_params* p);

/* MeshToVolume::convert (mesh_to_volume.rs:52-73).  `tri_xyz` is a host f32
 * soup of n triangles (9 floats each, PolygonSoup order: mesh/polygon_soup/
 * data_structure.rs:78-86).  *out is NULL and BS_ERR_EMPTY is returned where
 * the reference returns None. */
bs_status bs_mesh_to_volume(bs_ctx* ctx, const float* tri_xyz, size_t n, const bs_m2v_params* p, bs_volume** out);
/* Same, with `tri_xyz_dev` already in device memory of the context's device. */
bs_status bs_mesh_to_volume_device(bs_ctx* ctx, const float* tri_xyz_dev, size_t n, const bs_m2v_params* p,
                                   bs_volume** out);

/* Many MeshToVolume::convert calls in one: mesh i is tri_xyz[i] with n_tris[i]
 * triangles (device pointers, or host pointers when host_input != 0) and
 * params[i].  The meshes are spread over n_workers streams of this device
 * (0 = 4), each with its own scratch arena and host thread, largest first, so
 * small meshes' kernels overlap.  out[i] receives the volume or NULL;
 * status[i] (nullable) the per-mesh status (BS_ERR_EMPTY where convert
 * returns None).  Returns BS_OK when every mesh gave BS_OK or BS_ERR_EMPTY,
 * else the first other status.  Results equal per-mesh bs_mesh_to_volume*
 * calls bit for bit.  The reference has no batched entry point (callers
 * loop over convert). */
bs_status bs_batch_mesh_to_volume(bs_ctx* ctx, size_t n_meshes, const float* const* tri_xyz, const size_t* n_tris,
                                  const bs_m2v_params* params, int host_input, int n_workers, bs_volume** out,
                                  bs_status* status);

/* Volume::clone (volume/mod.rs:116-123) and drop.  A dropped volume's leaf
 * arrays go to a per-device pool (at most 4 sets and 8 GiB per device) that
 * later volumes of similar size draw from; bs_volume_free synchronises the
 * volume's device first, as hipFree does. */
bs_status bs_volume_clone(const bs_volume* v, bs_volume** out);
void bs_volume_free(bs_volume* v);

/* Frees every pooled leaf array set on every device.  No counterpart in the
 * reference (Rust drop returns memory at once): for callers that need the
 * device memory back for other work. */
void bs_release_cached_memory(void);

/* Volume::voxel_size and sizes. */
bs_status bs_volume_info_get(const bs_volume* v, bs_volume_info* out);

/* Canonical host export (any pointer may be NULL to skip that array):
 * origins n_leaves*3 int32, masks n_leaves*8 uint64, values n_leaves*512 float;
 * tiles: origin n_tiles*3 int32, edge length n_tiles int32, value n_tiles float. */
bs_status bs_volume_export(const bs_volume* v, int32_t* origins, uint64_t* masks, float* values);
bs_status bs_volume_export_tiles(const bs_volume* v, int32_t* origins, int32_t* sizes, float* values);

/* VolumeBuilder::sphere (builder.rs:21-30) through Volume::from_fn
 * (volume/mod.rs:40-72): f(p) = |p - (cx,cy,cz)| - radius sampled at p = idx*h
 * over floor((c - r - h)/h) .. ceil((c + r + h)/h), kept where |f| <= 2h.
 * The result may have no leaves (e.g. a negative radius). */
bs_status bs_volume_sphere(bs_ctx* ctx, float radius, float cx, float cy, float cz, float voxel_size,
                           bs_volume** out);

/* VolumeBuilder::cuboid (builder.rs:32-51) through Volume::from_fn: inside
 * the box [min, max], minus the distance to the nearest face; outside, the
 * distance to the box; sampled over floor((min - h)/h) .. ceil((max + h)/h),
 * kept where |f| <= 2h. */
bs_status bs_volume_cuboid(bs_ctx* ctx, float min_x, float min_y, float min_z, float max_x, float max_y,
                           float max_z, float voxel_size, bs_volume** out);

/* VolumeBuilder::iwp (builder.rs:53-76) through Volume::from_fn: inside
 * [min, max], cell_size * -(cx + cy + cz - 0.51 (cx cy + cy cz + cz cx) - 1)
 * with c* = cos(p* / cell_size); outside, the distance to the box grown by h;
 * band width 2 (kept where |f| <= 3h).  cos is correctly rounded here; the
 * reference uses the platform libm's cosf, so values may differ by an ulp.
 * cell_size must be finite and > 0 (the reference fills NaN otherwise). */
bs_status bs_volume_iwp(bs_ctx* ctx, float min_x, float min_y, float min_z, float max_x, float max_y, float max_z,
                        float cell_size, float voxel_size, bs_volume** out);

/* A grid built by inserting voxel values one by one (TreeNode::insert on an
 * empty VolumeGrid: root_node/tree_node.rs, internal_node/tree_node.rs:54-79,
 * leaf_node/tree_node.rs:46-50): idx n*3 int32 voxel indices,
 * values n floats, both host.  Each inserted voxel becomes active; other
 * entries of a touched leaf hold 0.0; a repeated index keeps the last value.
 * Indices must lie within +-(2^23 - 16) on every axis. */
bs_status bs_volume_from_voxels(bs_ctx* ctx, const int32_t* idx, const float* values, size_t n, float voxel_size,
                                bs_volume** out);

/* Volume::union / subtract / intersect (volume/mod.rs:74-93).  Both operands
 * are flood filled (FloodFill, root_node/flood_fill.rs:8-42) and the tree CSG
 * runs with `self` as the receiver (root_node/csg.rs:9-58,
 * internal_node/csg.rs:68-162, leaf_node/csg.rs:17-45).  The Rust methods take
 * both volumes by value: on BS_OK `self` holds the result and `other` has been
 * freed (its handle is dangling).  On an error both handles stay owned by the
 * caller and their contents are unspecified.  Like the reference, voxel sizes
 * are not checked; the result keeps self's. */
bs_status bs_volume_union(bs_ctx* ctx, bs_volume* self, bs_volume* other);
bs_status bs_volume_subtract(bs_ctx* ctx, bs_volume* self, bs_volume* other);
bs_status bs_volume_intersect(bs_ctx* ctx, bs_volume* self, bs_volume* other);
bs_status bs_ctx_last_csg_stats(const bs_ctx* ctx, bs_csg_stats* out);

/* FloodFill::flood_fill of the whole grid (root_node/flood_fill.rs:8-42,
 * internal_node/flood_fill.rs:17-75, leaf_node/flood_fill.rs:12-50): inactive
 * leaf entries become +-FLT_MAX with the scan-line sign and stay inactive. */
bs_status bs_volume_flood_fill(bs_ctx* ctx, bs_volume* v);

/* Volume::offset (volume/mod.rs:95-108) in place: entries with |v| > 2h are
 * deactivated (remove_if), the band is extended by fast sweeping
 * (fast_sweep.rs:10-287) up to |D| + 2h on the side of D's sign, empty leaves
 * are removed and D is subtracted from every active value.  A volume whose
 * active tiles are not all beyond 2h gives BS_ERR_INTERNAL (the reference
 * panics at fast_sweep.rs:100); the volume is then unchanged. */
bs_status bs_volume_offset(bs_ctx* ctx, bs_volume* v, float distance);
bs_status bs_ctx_last_offset_stats(const bs_ctx* ctx, bs_offset_stats* out);

/* DualContouringMesher::default().with_voxel_size(voxel_size).mesh(volume)
 * (meshing/dual_contouring.rs:23-83): a triangle soup of *n_tris triangles,
 * 9 floats each, in the sequential order (leaves by origin, voxels x-y-z,
 * edges X-Y-Z; the reference's parallel order is unspecified).  *out_xyz is
 * malloc'd host memory released with bs_free (NULL when there is no
 * triangle).  A volume with active tiles gives BS_ERR_UNSUPPORTED_TILE (the
 * reference stops at todo!()); a crossing voxel with no neighbour along an axis
 * gives BS_ERR_INTERNAL (unreachable!() at :347). */
bs_status bs_dual_contour(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz, size_t* n_tris);
/* Same, with the soup left in device memory (release with bs_device_free). */
bs_status bs_dual_contour_device(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz_dev,
                                 size_t* n_tris);
/* No counterpart in the reference: a variant of bs_dual_contour_device for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It writes the same soup into the
 * caller's device buffer out_xyz_dev of capacity_tris triangles (9 floats
 * each) and allocates no device or host memory: every intermediate buffer
 * lives in a workspace each device holds from the library's load until the
 * process ends (about 50 MB), set up by bs_ctx_create.  Limits: at most
 * BS_DC_INTO_MAX_LEAVES leaves (else BS_ERR_CAPACITY, nothing written).
 * *n_tris receives the soup's triangle count; when it exceeds capacity_tris
 * nothing is written and the call returns BS_ERR_CAPACITY, so the caller can
 * size the buffer and call again.  All calls on one device share the
 * workspace: calls that overlap on one device race. */
#define BS_DC_INTO_MAX_LEAVES 8192
bs_status bs_dual_contour_into(bs_ctx* ctx, const bs_volume* v, float voxel_size, float* out_xyz_dev,
                               size_t capacity_tris, size_t* n_tris);
bs_status bs_ctx_last_dc_stats(const bs_ctx* ctx, bs_dc_stats* out);
/* Vertex placement of the bs_dual_contour* calls on this context.
 * BS_DC_VERTEX_PARTICLE (the default) is the reference's damped particle
 * iteration (find_feature_point, dual_contouring.rs:429-458).  No counterpart
 * in the reference, opt-in: BS_DC_VERTEX_QEF_F64 puts each cell's vertex at
 * the minimiser of the quadric error sum_i (n_i . (x - p_i))^2 over the cell's
 * edge crossings, solved in FP64 (eigen-decomposition of the 3x3 normal
 * matrix, eigenvalues below 0.1 of the largest dropped, so planes, edges and
 * corners keep their sharp features), clamped to the cell.  The triangles
 * (which cells they join, their order) are the same in both modes; only the
 * vertex positions differ. */
#define BS_DC_VERTEX_PARTICLE 0
#define BS_DC_VERTEX_QEF_F64 1
bs_status bs_ctx_set_dc_vertex(bs_ctx* ctx, int mode);
void bs_free(void* p);
void bs_device_free(void* p_dev);

/* MarchingCubesMesher::default().with_voxel_size(voxel_size).mesh(volume)
 * (src/voxel/meshing/marching_cubes.rs:32-63): Marching Cubes 33 over every
 * cube whose eight corners have values, leaves and active tiles in the tree's
 * visit order (a tile meets the cubes of its six faces, :970-994).  *out_xyz
 * receives 9 floats per triangle (malloc'd, release with bs_free).  Each cube
 * emits the tiling the reference's tables (lookup_table.rs) select, in its
 * vertex order; case 6.1.2 uses the centre vertex an earlier cube left, as
 * the reference does. */
bs_status bs_marching_cubes(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz, size_t* n_tris);
/* Same, with the soup left in device memory (release with bs_device_free). */
bs_status bs_marching_cubes_device(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz_dev,
                                   size_t* n_tris);
/* Counters of the last bs_marching_cubes* call (n_leaves, n_tris, used_hash, ms_total). */
bs_status bs_ctx_last_mc_stats(const bs_ctx* ctx, bs_dc_stats* out);

/* VoxelRemesher::remesh (remeshing/voxel.rs:65-81): MeshToVolume with band 0
 * (VoxelRemesher::default, :83-90) and voxel size h, then the mesher of
 * `method`: BS_MESHING_FEATURE_PRESERVING is DualContouringMesher (the soup
 * comes back as bs_dual_contour's, malloc'd, release with bs_free);
 * BS_MESHING_MANIFOLD (the reference's default, :92) is MarchingCubesMesher,
 * as bs_marching_cubes.  tri_xyz is device memory when device_input is
 * non-zero.  An empty mesh gives BS_ERR_EMPTY (the reference's None). */
#define BS_MESHING_FEATURE_PRESERVING 0
#define BS_MESHING_MANIFOLD 1
bs_status bs_voxel_remesh(bs_ctx* ctx, const float* tri_xyz, size_t n, int device_input, float voxel_size,
                          int method, float** out_xyz, size_t* n_tris);

/* Tile classifier of BS_SIGN_SURROGATE (no counterpart in the reference): a
 * 3 -> 64 -> 64 -> 64 -> 1 ReLU network predicting a 4^3 tile's smallest
 * |SDF| in voxels from the tile centre.  BF16 weights and activations, FP32
 * accumulation, FP32 master weights trained with Adam on the device.
 * Parameters (FP32, this order): W1[64][3], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[64], b4; W[out][in].  The network input is
 * (point - shift) * scale. */
typedef struct bs_mlp bs_mlp;
#define BS_MLP_PARAMS 8641
typedef struct {
    uint32_t epochs, batch;
    float lr, lr_final;     /* cosine decay from lr to lr_final over all steps */
    float beta1, beta2, eps;
    uint64_t seed;          /* sample order of each epoch                      */
} bs_mlp_train_params;
bs_status bs_mlp_create(bs_ctx* ctx, uint64_t seed, bs_mlp** out); /* He-uniform weights, zero biases */
void bs_mlp_destroy(bs_mlp* m);
bs_status bs_mlp_set_params(bs_ctx* ctx, bs_mlp* m, const float* params); /* host, resets Adam */
bs_status bs_mlp_get_params(bs_ctx* ctx, const bs_mlp* m, float* params);
bs_status bs_mlp_set_input_map(bs_mlp* m, float shift_x, float shift_y, float shift_z, float scale);
/* out_dev[i] = network(x_dev[3 i .. 3 i + 2]), device arrays */
bs_status bs_mlp_forward_device(bs_ctx* ctx, const bs_mlp* m, const float* x_dev, size_t n, float* out_dev);
/* gradient of mean((network(x) - y)^2) over the n samples (grad: host,
 * BS_MLP_PARAMS floats) and that loss */
bs_status bs_mlp_gradient_device(bs_ctx* ctx, bs_mlp* m, const float* x_dev, const float* y_dev, size_t n,
                                 float* grad, double* loss);
/* mini-batch Adam on the squared error; epoch_loss (host, epochs floats, may
 * be NULL) receives each epoch's mean loss */
bs_status bs_mlp_train_device(bs_ctx* ctx, bs_mlp* m, const float* x_dev, const float* y_dev, size_t n,
                              const bs_mlp_train_params* params, float* epoch_loss);
/* Training samples from a volume: for the eight 4^3 tiles of each leaf
 * (leaf-major, tile = x-half << 2 | y-half << 1 | z-half) the tile centre
 * (x_dev, 3 floats) and the smallest |value| / voxel size over its active
 * voxels (y_dev; -1 for a tile without one).  *n_out = 8 * leaves; capacity
 * below that gives BS_ERR_CAPACITY. */
bs_status bs_volume_tile_samples(bs_ctx* ctx, const bs_volume* v, float* x_dev, float* y_dev, size_t capacity,
                                 size_t* n_out);
/* The classifier and cull threshold (voxels) of BS_SIGN_SURROGATE on this
 * context.  The network is not owned and must outlive its use.  With m NULL
 * a tile is culled when its true smallest |SDF| is at or above the threshold
 * (a perfect classifier, for measurement). */
bs_status bs_ctx_set_surrogate(bs_ctx* ctx, const bs_mlp* m, float threshold_vox);

/* Binary STL (io/stl.rs), host only, no device needed.  bs_stl_read is
 * StlReader (:64-92): *out_xyz receives 9 floats per face (malloc'd, release
 * with bs_free; NULL for zero faces); a missing or short file gives BS_ERR_IO.
 * bs_stl_write is StlWriter (:150-199): unit face normals of (b-a) x (c-a),
 * zeros for degenerate faces. */
bs_status bs_stl_read(const char* path, float** out_xyz, size_t* n_tris);
bs_status bs_stl_write(const char* path, const float* tri_xyz, size_t n_tris);

/* Extension with no reference counterpart: every active tile becomes the
 * leaves it covers, all 512 entries active and equal to the tile value (a
 * 128^3 tile gives 4096 leaves).  The reference's dual contouring stops at
 * todo!() on tiles (dual_contouring.rs:146), and a union leaves active -MAX
 * tiles (internal_node/csg.rs:42-48); densifying first lets bs_dual_contour
 * run on such volumes. */
bs_status bs_volume_densify_tiles(bs_ctx* ctx, bs_volume* v);

/* FloodFill::sign_at (root_node/flood_fill.rs:56-63, internal_node/
 * flood_fill.rs:110-117, leaf_node/flood_fill.rs:72-75) for n voxel indices
 * (idx: n*3 int32, host): out_negative[i] = 1 when the sign is negative.  A
 * missing root entry is positive; a tile slot answers with its value's sign,
 * active or not.  The tile of a slot whose branch a CSG removed is undefined
 * in the reference (it reads pointer bits); here it keeps the slot's earlier
 * tile value. */
bs_status bs_volume_signs_at(bs_ctx* ctx, bs_volume* v, const int32_t* idx, size_t n, uint8_t* out_negative);

/* Device views of the leaf arrays (valid until the volume is freed/consumed). */
bs_status bs_volume_device_leaves(const bs_volume* v, const int32_t** origins, const uint64_t** masks,
                                  const float** values, uint64_t* n_leaves);

#ifdef __cplusplus
}
#endif

#endif /* BS_VOXEL_H */
