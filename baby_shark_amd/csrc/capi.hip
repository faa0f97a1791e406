// C ABI entry points (include/bs_voxel.h): argument checks, error mapping and
// the canonical host export of volumes.
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <thread>
#include <string>

#include "common.hpp"
#include "stl.hpp"

namespace bs {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

void mesh_to_volume_device(bs_ctx* ctx, const float* d_tris, uint64_t n, const bs_m2v_params& prm, bs_volume** out);
void init_context_streams(bs_ctx* ctx);
void volume_sphere(bs_ctx* ctx, float radius, float cx, float cy, float cz, float h, bs_volume** out);
void volume_cuboid(bs_ctx* ctx, const float mn[3], const float mx[3], float h, bs_volume** out);
void volume_iwp(bs_ctx* ctx, const float mn[3], const float mx[3], float cell_size, float h, bs_volume** out);
void volume_csg(bs_ctx* ctx, bs_volume* a, bs_volume* b, CsgOp op);
void volume_flood_fill(bs_ctx* ctx, bs_volume* v);
void volume_densify_tiles(bs_ctx* ctx, bs_volume* v);
void volume_offset(bs_ctx* ctx, bs_volume* vol, float distance, bs_offset_stats* stats);
void dual_contour_device(bs_ctx* ctx, bs_volume* vol, float voxel_size, float** out_dev, uint64_t* n_tris,
                         bs_dc_stats* stats);
void dual_contour_into(bs_ctx* ctx, const bs_volume* vol, float voxel_size, float* out_dev, uint64_t capacity_tris,
                       uint64_t* n_tris);
void dc_noalloc_init(int device);
void marching_cubes_device(bs_ctx* ctx, bs_volume* vol, float voxel_size, float** out_dev, uint64_t* n_tris,
                           bs_dc_stats* stats);
void volume_from_voxels(bs_ctx* ctx, const int32_t* idx, const float* values, uint64_t n, float h, bs_volume** out);
void volume_signs_at(bs_ctx* ctx, bs_volume* v, const int32_t* idx, uint64_t n, uint8_t* out_neg);

template <class F>
static bs_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return BS_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host out of memory";
        return BS_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return BS_ERR_INTERNAL;
    }
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        BS_HIP(hipGetDevice(&prev));
        if (prev != dev) BS_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

static void free_volume_arrays(bs_volume* v) { free_leaves(v); }

}  // namespace bs

using namespace bs;

extern "C" {

const char* bs_last_error(void) { return g_last_error.c_str(); }

bs_status bs_ctx_create(int device, bs_ctx** out) {
    if (!out) return BS_ERR_INVALID_ARG;
    *out = nullptr;
    return guarded([&] {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) throw Error(BS_ERR_NO_DEVICE, "no HIP device available");
        BS_REQUIRE(device >= 0 && device < count, BS_ERR_INVALID_ARG, "device index out of range");
        DeviceGuard g(device);
        dc_noalloc_init(device);  // this device's address of bs_dual_contour_into's workspace
        auto* c = new bs_ctx();
        c->device = device;
        hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (se != hipSuccess) {
            delete c;
            throw Error(BS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(se));
        }
        *out = c;
    });
}

// the second stream, its events and the pinned buffer of a context
static void release_side(bs_ctx* c) {
    if (c->stream2) {
        (void)hipStreamSynchronize(c->stream2);
        (void)hipStreamDestroy(c->stream2);
        c->stream2 = nullptr;
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_leaves) (void)hipEventDestroy(c->ev_leaves);
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    c->ev_fork = c->ev_join = c->ev_leaves = nullptr;
    c->host_pinned = nullptr;
}

void bs_ctx_destroy(bs_ctx* ctx) {
    if (!ctx) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_side(ctx);
    ctx->scratch.release();
    for (auto& w : ctx->workers) {
        (void)hipStreamSynchronize(w->stream);
        release_side(w.get());
        w->scratch.release();
        (void)hipStreamDestroy(w->stream);
    }
    ctx->workers.clear();
    (void)hipStreamDestroy(ctx->stream);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete ctx;
}

void* bs_ctx_stream(bs_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

bs_status bs_ctx_set_profiling(bs_ctx* ctx, int enabled) {  // flags
    if (!ctx) return BS_ERR_INVALID_ARG;
    if (enabled & ~(BS_PROFILE_EVENTS | BS_PROFILE_COUNTERS)) return BS_ERR_INVALID_ARG;
    ctx->profiling = (enabled & BS_PROFILE_EVENTS) != 0;
    ctx->counting = (enabled & BS_PROFILE_COUNTERS) != 0;
    return BS_OK;
}

bs_status bs_ctx_set_paths(bs_ctx* ctx, int flags) {
    if (!ctx || (flags & ~(BS_PATH_FORCE_HASH | BS_PATH_FORCE_LDS_STACK | BS_PATH_LBVH | BS_PATH_PLOC))) {
        g_last_error = "unknown BS_PATH_* flag";
        return BS_ERR_INVALID_ARG;
    }
    ctx->paths = flags;
    return BS_OK;
}

bs_status bs_ctx_set_dc_vertex(bs_ctx* ctx, int mode) {
    if (!ctx || (mode != BS_DC_VERTEX_PARTICLE && mode != BS_DC_VERTEX_QEF_F64)) {
        g_last_error = "unknown BS_DC_VERTEX_* mode";
        return BS_ERR_INVALID_ARG;
    }
    ctx->dc_vertex = mode;
    return BS_OK;
}

bs_status bs_ctx_last_m2v_stats(const bs_ctx* ctx, bs_m2v_stats* out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = ctx->last_m2v;
    return BS_OK;
}

void bs_m2v_params_default(bs_m2v_params* p) {
    if (!p) return;
    p->voxel_size = 1.0f;     // mesh_to_volume.rs:226
    p->band = 0;              // mesh_to_volume.rs:229
    p->wn_accuracy = 2.0f;    // mesh_to_volume.rs:264
    p->sign_threshold = 0.2f; // mesh_to_volume.rs:266
    p->fp64_band = 0.1f;
    p->sign_mode = BS_SIGN_PER_VOXEL;
}

bs_status bs_mesh_to_volume_device(bs_ctx* ctx, const float* tri_xyz_dev, size_t n, const bs_m2v_params* p,
                                   bs_volume** out) {
    if (!ctx || !out || !p || (n > 0 && !tri_xyz_dev)) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        mesh_to_volume_device(ctx, tri_xyz_dev, n, *p, out);
    });
}

bs_status bs_mesh_to_volume(bs_ctx* ctx, const float* tri_xyz, size_t n, const bs_m2v_params* p, bs_volume** out) {
    if (!ctx || !out || !p || (n > 0 && !tri_xyz)) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        if (n == 0) throw Error(BS_ERR_EMPTY, "empty mesh: the reference returns None");
        float* d = ctx->scratch.as<float>(S_INPUT, 9 * (uint64_t)n);
        BS_HIP(hipMemcpyAsync(d, tri_xyz, sizeof(float) * 9 * n, hipMemcpyHostToDevice, ctx->stream));
        mesh_to_volume_device(ctx, d, n, *p, out);
    });
}

bs_status bs_batch_mesh_to_volume(bs_ctx* ctx, size_t n_meshes, const float* const* tri_xyz, const size_t* n_tris,
                                  const bs_m2v_params* params, int host_input, int n_workers, bs_volume** out,
                                  bs_status* status) {
    if (!ctx || (n_meshes > 0 && (!tri_xyz || !n_tris || !params || !out))) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    for (size_t i = 0; i < n_meshes; ++i) {
        out[i] = nullptr;
        if (status) status[i] = BS_OK;
    }
    if (n_workers <= 0) n_workers = 4;
    n_workers = (int)std::min<size_t>((size_t)std::min(n_workers, 16), std::max<size_t>(n_meshes, 1));
    std::vector<bs_status> st(n_meshes, BS_OK);
    std::vector<std::string> msg(n_meshes);
    const bs_status setup = guarded([&] {
        DeviceGuard g(ctx->device);
        while ((int)ctx->workers.size() < n_workers) {
            auto w = std::make_unique<bs_ctx>();
            w->device = ctx->device;
            BS_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
            bs_ctx* wp = w.get();
            ctx->workers.push_back(std::move(w));
            // every stream, event and pinned buffer a worker uses exists before
            // the worker threads start: none is created while others launch
            init_context_streams(wp);
        }
    });
    if (setup != BS_OK) return setup;
    // largest meshes first: the arenas reach their size early and the tail is short
    std::vector<size_t> order(n_meshes);
    for (size_t i = 0; i < n_meshes; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return n_tris[a] > n_tris[b]; });
    std::atomic<size_t> next{0};
    auto work = [&](int wi) {
        bs_ctx* w = ctx->workers[wi].get();
        w->paths = ctx->paths;
        w->profiling = false;
        w->counting = false;
        for (size_t k; (k = next.fetch_add(1)) < n_meshes;) {
            const size_t i = order[k];
            st[i] = guarded([&] {
                DeviceGuard g(w->device);
                if (n_tris[i] == 0) throw Error(BS_ERR_EMPTY, "empty mesh: the reference returns None");
                BS_REQUIRE(tri_xyz[i] != nullptr, BS_ERR_INVALID_ARG, "null mesh pointer");
                const float* d = tri_xyz[i];
                if (host_input) {
                    float* buf = w->scratch.as<float>(S_INPUT, 9 * (uint64_t)n_tris[i]);
                    BS_HIP(hipMemcpyAsync(buf, tri_xyz[i], sizeof(float) * 9 * n_tris[i], hipMemcpyHostToDevice,
                                          w->stream));
                    d = buf;
                }
                mesh_to_volume_device(w, d, n_tris[i], params[i], &out[i]);
            });
            if (st[i] != BS_OK) msg[i] = g_last_error;
        }
    };
    std::vector<std::thread> threads;
    for (int wi = 1; wi < n_workers; ++wi) threads.emplace_back(work, wi);
    work(0);
    for (auto& t : threads) t.join();
    bs_status first = BS_OK;
    for (size_t i = 0; i < n_meshes; ++i) {
        if (status) status[i] = st[i];
        if (st[i] != BS_OK && st[i] != BS_ERR_EMPTY && first == BS_OK) {
            first = st[i];
            g_last_error = "mesh " + std::to_string(i) + ": " + msg[i];
        }
    }
    if (first == BS_OK) g_last_error.clear();
    return first;
}

bs_status bs_volume_clone(const bs_volume* v, bs_volume** out) {
    if (!v || !out) return BS_ERR_INVALID_ARG;
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(v->device);
        auto* c = new bs_volume();
        c->device = v->device;
        c->voxel_size = v->voxel_size;
        c->n_leaves = v->n_leaves;
        c->tiles = v->tiles;
        if (v->topo) c->topo = v->topo->clone();
        try {
            if (v->n_leaves) {
                alloc_leaves(c, v->n_leaves);
                BS_HIP(hipMemcpy(c->origins, v->origins, sizeof(int32_t) * 4 * v->n_leaves, hipMemcpyDeviceToDevice));
                BS_HIP(hipMemcpy(c->masks, v->masks, sizeof(uint64_t) * 8 * v->n_leaves, hipMemcpyDeviceToDevice));
                BS_HIP(hipMemcpy(c->values, v->values, sizeof(float) * 512 * v->n_leaves, hipMemcpyDeviceToDevice));
            }
        } catch (...) {
            free_volume_arrays(c);
            delete c;
            throw;
        }
        *out = c;
    });
}

void bs_release_cached_memory(void) { release_cached_leaves(); }

void bs_volume_free(bs_volume* v) {
    if (!v) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(v->device);
    free_volume_arrays(v);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete v;
}

bs_status bs_volume_info_get(const bs_volume* v, bs_volume_info* out) {
    if (!v || !out) return BS_ERR_INVALID_ARG;
    return guarded([&] {
        DeviceGuard g(v->device);
        out->voxel_size = v->voxel_size;
        out->n_leaves = v->n_leaves;
        out->n_tiles = v->tiles.size();
        uint64_t act = 0;
        if (v->n_leaves) {
            std::vector<uint64_t> m(8 * v->n_leaves);
            BS_HIP(hipMemcpy(m.data(), v->masks, sizeof(uint64_t) * m.size(), hipMemcpyDeviceToHost));
            for (uint64_t w : m) act += (uint64_t)__builtin_popcountll(w);
        }
        out->n_active = act;
    });
}

bs_status bs_volume_export(const bs_volume* v, int32_t* origins, uint64_t* masks, float* values) {
    if (!v) return BS_ERR_INVALID_ARG;
    return guarded([&] {
        DeviceGuard g(v->device);
        const uint64_t n = v->n_leaves;
        if (!n) return;
        if (origins) {
            std::vector<int32_t> o(4 * n);
            BS_HIP(hipMemcpy(o.data(), v->origins, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) origins[3 * i + k] = o[4 * i + k];
        }
        if (masks) BS_HIP(hipMemcpy(masks, v->masks, sizeof(uint64_t) * 8 * n, hipMemcpyDeviceToHost));
        if (values) BS_HIP(hipMemcpy(values, v->values, sizeof(float) * 512 * n, hipMemcpyDeviceToHost));
    });
}

bs_status bs_volume_export_tiles(const bs_volume* v, int32_t* origins, int32_t* sizes, float* values) {
    if (!v) return BS_ERR_INVALID_ARG;
    for (size_t i = 0; i < v->tiles.size(); ++i) {
        const bs::TileRec& t = v->tiles[i];
        if (origins) {
            origins[3 * i] = t.x;
            origins[3 * i + 1] = t.y;
            origins[3 * i + 2] = t.z;
        }
        if (sizes) sizes[i] = t.size;
        if (values) values[i] = t.value;
    }
    return BS_OK;
}

bs_status bs_volume_sphere(bs_ctx* ctx, float radius, float cx, float cy, float cz, float voxel_size,
                           bs_volume** out) {
    if (!ctx || !out) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        volume_sphere(ctx, radius, cx, cy, cz, voxel_size, out);
    });
}

bs_status bs_volume_cuboid(bs_ctx* ctx, float min_x, float min_y, float min_z, float max_x, float max_y,
                           float max_z, float voxel_size, bs_volume** out) {
    if (!ctx || !out) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        const float mn[3] = {min_x, min_y, min_z}, mx[3] = {max_x, max_y, max_z};
        volume_cuboid(ctx, mn, mx, voxel_size, out);
    });
}

bs_status bs_volume_iwp(bs_ctx* ctx, float min_x, float min_y, float min_z, float max_x, float max_y, float max_z,
                        float cell_size, float voxel_size, bs_volume** out) {
    if (!ctx || !out) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        const float mn[3] = {min_x, min_y, min_z}, mx[3] = {max_x, max_y, max_z};
        volume_iwp(ctx, mn, mx, cell_size, voxel_size, out);
    });
}

bs_status bs_volume_from_voxels(bs_ctx* ctx, const int32_t* idx, const float* values, size_t n, float voxel_size,
                                bs_volume** out) {
    if (!ctx || !out || (n > 0 && (!idx || !values))) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        volume_from_voxels(ctx, idx, values, n, voxel_size, out);
    });
}

static bs_status csg_entry(bs_ctx* ctx, bs_volume* self, bs_volume* other, CsgOp op) {
    if (!ctx || !self || !other) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    const bs_status st = guarded([&] {
        DeviceGuard g(ctx->device);
        volume_csg(ctx, self, other, op);
    });
    if (st == BS_OK) bs_volume_free(other);
    return st;
}

bs_status bs_volume_union(bs_ctx* ctx, bs_volume* self, bs_volume* other) {
    return csg_entry(ctx, self, other, CSG_UNION);
}
bs_status bs_volume_subtract(bs_ctx* ctx, bs_volume* self, bs_volume* other) {
    return csg_entry(ctx, self, other, CSG_SUBTRACT);
}
bs_status bs_volume_intersect(bs_ctx* ctx, bs_volume* self, bs_volume* other) {
    return csg_entry(ctx, self, other, CSG_INTERSECT);
}

bs_status bs_ctx_last_csg_stats(const bs_ctx* ctx, bs_csg_stats* out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = ctx->last_csg;
    return BS_OK;
}

bs_status bs_volume_flood_fill(bs_ctx* ctx, bs_volume* v) {
    if (!ctx || !v) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    return guarded([&] {
        DeviceGuard g(ctx->device);
        BS_REQUIRE(v->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
        volume_flood_fill(ctx, v);
    });
}

bs_status bs_volume_offset(bs_ctx* ctx, bs_volume* v, float distance) {
    if (!ctx || !v) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    return guarded([&] {
        DeviceGuard g(ctx->device);
        volume_offset(ctx, v, distance, &ctx->last_offset);
    });
}

bs_status bs_ctx_last_offset_stats(const bs_ctx* ctx, bs_offset_stats* out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = ctx->last_offset;
    return BS_OK;
}

bs_status bs_dual_contour_device(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz_dev,
                                 size_t* n_tris) {
    if (!ctx || !v || !out_xyz_dev || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz_dev = nullptr;
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        uint64_t n = 0;
        dual_contour_device(ctx, const_cast<bs_volume*>(v), voxel_size, out_xyz_dev, &n, &ctx->last_dc);
        *n_tris = (size_t)n;
    });
}

bs_status bs_dual_contour_into(bs_ctx* ctx, const bs_volume* v, float voxel_size, float* out_xyz_dev,
                               size_t capacity_tris, size_t* n_tris) {
    if (!ctx || !v || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        uint64_t n = 0;
        try {
            dual_contour_into(ctx, v, voxel_size, out_xyz_dev, capacity_tris, &n);
        } catch (...) {
            *n_tris = (size_t)n;
            throw;
        }
        *n_tris = (size_t)n;
    });
}

bs_status bs_dual_contour(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz, size_t* n_tris) {
    if (!ctx || !v || !out_xyz || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz = nullptr;
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        float* d = nullptr;
        uint64_t n = 0;
        dual_contour_device(ctx, const_cast<bs_volume*>(v), voxel_size, &d, &n, &ctx->last_dc);
        if (!n) return;
        float* h = static_cast<float*>(std::malloc(sizeof(float) * 9 * n));
        if (!h) {
            (void)hipFree(d);
            throw Error(BS_ERR_INTERNAL, "host out of memory");
        }
        hipError_t e = hipMemcpy(h, d, sizeof(float) * 9 * n, hipMemcpyDeviceToHost);
        (void)hipFree(d);
        if (e != hipSuccess) {
            std::free(h);
            BS_HIP(e);
        }
        *out_xyz = h;
        *n_tris = (size_t)n;
    });
}

// a device soup of n triangles to malloc'd host memory; the device copy is freed
static float* soup_to_host(float* d, uint64_t n) {
    float* h = static_cast<float*>(std::malloc(sizeof(float) * 9 * n));
    if (!h) {
        (void)hipFree(d);
        throw Error(BS_ERR_INTERNAL, "host out of memory");
    }
    hipError_t e = hipMemcpy(h, d, sizeof(float) * 9 * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        std::free(h);
        BS_HIP(e);
    }
    return h;
}

bs_status bs_marching_cubes_device(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz_dev,
                                   size_t* n_tris) {
    if (!ctx || !v || !out_xyz_dev || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz_dev = nullptr;
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        uint64_t n = 0;
        marching_cubes_device(ctx, const_cast<bs_volume*>(v), voxel_size, out_xyz_dev, &n, &ctx->last_mc);
        *n_tris = (size_t)n;
    });
}

bs_status bs_marching_cubes(bs_ctx* ctx, const bs_volume* v, float voxel_size, float** out_xyz, size_t* n_tris) {
    if (!ctx || !v || !out_xyz || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz = nullptr;
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        float* d = nullptr;
        uint64_t n = 0;
        marching_cubes_device(ctx, const_cast<bs_volume*>(v), voxel_size, &d, &n, &ctx->last_mc);
        if (!n) return;
        *out_xyz = soup_to_host(d, n);
        *n_tris = (size_t)n;
    });
}

bs_status bs_ctx_last_mc_stats(const bs_ctx* ctx, bs_dc_stats* out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = ctx->last_mc;
    return BS_OK;
}

bs_status bs_voxel_remesh(bs_ctx* ctx, const float* tri_xyz, size_t n, int device_input, float voxel_size,
                          int method, float** out_xyz, size_t* n_tris) {
    if (!ctx || !out_xyz || !n_tris || (n > 0 && !tri_xyz)) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz = nullptr;
    *n_tris = 0;
    return guarded([&] {
        DeviceGuard g(ctx->device);
        BS_REQUIRE(method == BS_MESHING_FEATURE_PRESERVING || method == BS_MESHING_MANIFOLD, BS_ERR_INVALID_ARG,
                   "unknown meshing method");
        if (n == 0) throw Error(BS_ERR_EMPTY, "empty mesh: the reference returns None");
        // VoxelRemesher::default(): MeshToVolume with band 0 (remeshing/voxel.rs:83-90)
        bs_m2v_params p;
        bs_m2v_params_default(&p);
        p.voxel_size = voxel_size;
        p.band = 0;
        const float* d = tri_xyz;
        if (!device_input) {
            float* dd = ctx->scratch.as<float>(S_INPUT, 9 * (uint64_t)n);
            BS_HIP(hipMemcpyAsync(dd, tri_xyz, sizeof(float) * 9 * n, hipMemcpyHostToDevice, ctx->stream));
            d = dd;
        }
        bs_volume* vol = nullptr;
        mesh_to_volume_device(ctx, d, n, p, &vol);
        std::unique_ptr<bs_volume, void (*)(bs_volume*)> hold(vol, [](bs_volume* v) {
            free_volume_arrays(v);
            delete v;
        });
        float* dt = nullptr;
        uint64_t nt = 0;
        if (method == BS_MESHING_MANIFOLD)
            marching_cubes_device(ctx, vol, voxel_size, &dt, &nt, &ctx->last_mc);
        else
            dual_contour_device(ctx, vol, voxel_size, &dt, &nt, &ctx->last_dc);
        if (!nt) return;
        *out_xyz = soup_to_host(dt, nt);
        *n_tris = (size_t)nt;
    });
}

bs_status bs_ctx_last_dc_stats(const bs_ctx* ctx, bs_dc_stats* out) {
    if (!ctx || !out) return BS_ERR_INVALID_ARG;
    *out = ctx->last_dc;
    return BS_OK;
}

bs_status bs_stl_read(const char* path, float** out_xyz, size_t* n_tris) {
    if (!path || !out_xyz || !n_tris) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    *out_xyz = nullptr;
    *n_tris = 0;
    return guarded([&] {
        std::vector<float> xyz;
        const std::string err = stl_read(path, xyz);
        if (!err.empty()) throw Error(BS_ERR_IO, err);
        if (xyz.empty()) return;
        float* p = static_cast<float*>(std::malloc(sizeof(float) * xyz.size()));
        if (!p) throw std::bad_alloc();
        std::memcpy(p, xyz.data(), sizeof(float) * xyz.size());
        *out_xyz = p;
        *n_tris = xyz.size() / 9;
    });
}

bs_status bs_stl_write(const char* path, const float* tri_xyz, size_t n_tris) {
    if (!path || (n_tris > 0 && !tri_xyz)) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    return guarded([&] {
        const std::string err = stl_write(path, tri_xyz, n_tris);
        if (!err.empty()) throw Error(BS_ERR_IO, err);
    });
}

void bs_free(void* p) { std::free(p); }

void bs_device_free(void* p) {
    if (p) (void)hipFree(p);
}

bs_status bs_volume_densify_tiles(bs_ctx* ctx, bs_volume* v) {
    if (!ctx || !v) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    return guarded([&] {
        DeviceGuard g(ctx->device);
        BS_REQUIRE(v->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
        volume_densify_tiles(ctx, v);
    });
}

bs_status bs_volume_signs_at(bs_ctx* ctx, bs_volume* v, const int32_t* idx, size_t n, uint8_t* out_negative) {
    if (!ctx || !v || (n > 0 && (!idx || !out_negative))) {
        g_last_error = "null argument";
        return BS_ERR_INVALID_ARG;
    }
    return guarded([&] {
        DeviceGuard g(ctx->device);
        BS_REQUIRE(v->device == ctx->device, BS_ERR_INVALID_ARG, "volume lives on another device");
        volume_signs_at(ctx, v, idx, n, out_negative);
    });
}

bs_status bs_volume_device_leaves(const bs_volume* v, const int32_t** origins, const uint64_t** masks,
                                  const float** values, uint64_t* n_leaves) {
    if (!v) return BS_ERR_INVALID_ARG;
    if (origins) *origins = v->origins;
    if (masks) *masks = v->masks;
    if (values) *values = v->values;
    if (n_leaves) *n_leaves = v->n_leaves;
    return BS_OK;
}

}  // extern "C"
