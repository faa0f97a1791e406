// Shared host/device pieces of the bs_voxel library: error handling, the
// per-context scratch arena, the device volume record and small device helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <memory>
#include <vector>

#include "../../include/bs_voxel.h"
#include "topology.hpp"

namespace bs {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    bs_status status;
    Error(bs_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);

#define BS_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::bs::Error(BS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                              __FILE__ + ":" + std::to_string(__LINE__));              \
    } while (0)

#define BS_REQUIRE(cond, status, msg)                          \
    do {                                                       \
        if (!(cond)) throw ::bs::Error((status), (msg));       \
    } while (0)

// ---------------------------------------------------------------- scratch arena
// Named device buffers owned by a context that only ever grow, so a steady
// stream of same-sized calls allocates nothing after the first.
enum ScratchSlot : int {
    S_ROWS, S_SUBS, S_ROW_OFF, S_SUB_OFF, S_SUBTRIS, S_RECS, S_LEAFCNT, S_LEAFOFF, S_HASH_KEYS, S_HASH_VALS,
    S_KEYS, S_KEYS_SORTED, S_CAND_LEAF, S_CAND_CNT, S_CAND_OFF, S_CAND_LIST, S_CURSOR, S_ACT_CNT, S_ACT_OFF,
    S_ACT_LIST, S_COUNTERS, S_CUB, S_MORTON, S_MORTON_SORTED, S_IDX, S_IDX_SORTED, S_BVH_TRIS, S_BVH_NODES,
    S_BVH_LINKS, S_BVH_PARENT, S_BVH_FLAGS, S_BVH_INIT, S_BVH_BOX, S_RECHECK, S_ARGMIN, S_MISC0, S_MISC1,
    S_MISC2, S_MISC3, S_INPUT, S_BVH_KIDS, S_FLAGS_A, S_FLAGS_B, S_CSG_PLAN, S_DC_MAP, S_DC_KEYS, S_DC_FAIL, S_DC_CELLS, S_SIGN_ROOTS, S_REP_CNT, S_REP_OFF, S_REPS,
    S_BVH_HEIGHT, S_BVH_HEIGHT_S, S_BVH_INODE, S_BVH_INODE_S, S_BVH_MISC, S_BVH_STARTS, S_CUB_B,
    S_DC_CELLMASK, S_DC_TRICNT, S_DC_TRIOFF, S_MC_TAB, S_MC_UNITS, S_MC_TILES, S_MC_T8, S_MC_T8K, S_MC_T128,
    S_MC_T128K, S_MC_TORG, S_MC_LANES, S_LEAF_BOX, S_SG_PRED, S_MLP_PART, S_MLP_GRAD, S_MLP_PERM, S_MC_FLAGS, S_MC_CV, S_MC_V12,
    S_PLOC_BOX, S_PLOC_INFO, S_PLOC_C0, S_PLOC_C1, S_PLOC_NN, S_PLOC_KEEP, S_PLOC_POS, S_PLOC_RING,
    S_TD_PRIMS, S_TD_FR, S_TD_BINS, S_TD_TASKS, S_TD_TCNT, S_TD_SPLIT, S_TD_SMALL, S_TD_HIST,
    S_LEAF_MKEY, S_LEAF_MKEY_S, S_LEAF_IDS, S_LEAF_PERM, S_COUNT
};

struct Scratch {
    void* ptr[S_COUNT] = {};
    size_t cap[S_COUNT] = {};
    void* get(int slot, size_t bytes) {
        if (bytes == 0) bytes = 256;
        if (cap[slot] < bytes) {
            if (ptr[slot]) BS_HIP(hipFree(ptr[slot]));
            ptr[slot] = nullptr;
            size_t b = bytes + bytes / 4 + 256;
            BS_HIP(hipMalloc(&ptr[slot], b));
            cap[slot] = b;
        }
        return ptr[slot];
    }
    template <class T> T* as(int slot, size_t count) { return static_cast<T*>(get(slot, count * sizeof(T))); }
    void release() {
        for (int i = 0; i < S_COUNT; ++i)
            if (ptr[i]) {
                (void)hipFree(ptr[i]);
                ptr[i] = nullptr;
                cap[i] = 0;
            }
    }
};

}  // namespace bs

// ---------------------------------------------------------------- handles
struct bs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool profiling = false;  // HIP events around the mesh-to-volume phases
    bool counting = false;   // winding-traversal visit counters (a slower kernel variant)
    int paths = 0;  // BS_PATH_* flags
    int dc_vertex = BS_DC_VERTEX_PARTICLE;  // dual-contouring vertex placement
    bs::Scratch scratch;
    bs_m2v_stats last_m2v = {};
    bs_offset_stats last_offset = {};
    bs_dc_stats last_dc = {};
    bs_dc_stats last_mc = {};
    bs_csg_stats last_csg = {};
    std::vector<std::unique_ptr<bs_ctx>> workers;  // sub-contexts of batched calls (own stream, scratch)
    // second stream for work that overlaps the main one (the BVH build), with
    // its events and a pinned readback buffer; created on first use
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_leaves = nullptr;
    uint32_t* host_pinned = nullptr;  // 512 words
    // tile classifier of BS_SIGN_SURROGATE (not owned), and its cull threshold
    // in voxels; without a network the gate uses each tile's true smallest |d|
    const struct bs_mlp* surrogate = nullptr;
    float surrogate_threshold = 0.0f;
};

struct bs_volume {
    int device = 0;
    float voxel_size = 1.0f;
    uint64_t n_leaves = 0;
    int32_t* origins = nullptr;  // device, n_leaves*4 (x, y, z, 0)
    uint64_t* masks = nullptr;   // device, n_leaves*8
    float* values = nullptr;     // device, n_leaves*512
    std::vector<bs::TileRec> tiles;           // active internal tiles (host, few)
    std::unique_ptr<bs::Topology> topo;       // upper levels; built from the leaves when absent
};

namespace bs {

// Leaf arrays of a volume (origins, masks, values) for n leaves on the current
// device, drawn from and returned to a per-device pool of released sets.  A
// release waits for `pending` (a stream that may still run kernels on the
// arrays, after a call failed part-way) and for nothing else.
void alloc_leaves(bs_volume* v, uint64_t n);
void free_leaves(bs_volume* v, hipStream_t pending = nullptr);
void release_cached_leaves();

// ---------------------------------------------------------------- device helpers
constexpr int LEAF_DIM = 8;
constexpr int LEAF_VOX = 512;
constexpr int KEY_BIAS = 1 << 20;  // leaf coordinates are biased into 21 bits

// Leaf key: biased leaf coordinates (origin >> 3) packed x|y|z into 63 bits so
// that unsigned key order is lexicographic (x, y, z) order of origins.
__host__ __device__ inline uint64_t leaf_key(int lx, int ly, int lz) {
    return ((uint64_t)(uint32_t)(lx + KEY_BIAS) << 42) | ((uint64_t)(uint32_t)(ly + KEY_BIAS) << 21) |
           (uint64_t)(uint32_t)(lz + KEY_BIAS);
}
__host__ __device__ inline void leaf_key_decode(uint64_t k, int& lx, int& ly, int& lz) {
    lx = (int)((k >> 42) & 0x1FFFFF) - KEY_BIAS;
    ly = (int)((k >> 21) & 0x1FFFFF) - KEY_BIAS;
    lz = (int)(k & 0x1FFFFF) - KEY_BIAS;
}
constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ inline uint32_t hash_u64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k;
}

// Open addressing, linear probing; capacity is a power of two.  Probing is
// bounded by the capacity: a full table reports failure instead of spinning.
__device__ inline bool hash_insert(uint64_t* keys, uint32_t cap_mask, uint64_t key) {
    uint32_t h = hash_u64(key) & cap_mask;
    for (uint32_t probe = 0; probe <= cap_mask; ++probe) {
        unsigned long long prev = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)EMPTY_KEY,
                                            (unsigned long long)key);
        if (prev == EMPTY_KEY || prev == key) return true;
        h = (h + 1) & cap_mask;
    }
    return false;
}
constexpr uint32_t NOT_FOUND = 0xFFFFFFFFu;
__device__ inline uint32_t hash_find(const uint64_t* keys, uint32_t cap_mask, uint64_t key) {
    uint32_t h = hash_u64(key) & cap_mask;
    for (uint32_t probe = 0; probe <= cap_mask; ++probe) {
        uint64_t k = keys[h];
        if (k == key) return h;
        if (k == EMPTY_KEY) return NOT_FOUND;
        h = (h + 1) & cap_mask;
    }
    return NOT_FOUND;
}

// Correctly rounded f32 square root and division.  HIP's __fsqrt_rn is the
// native (approximate) v_sqrt_f32 unless OCML's rounded operations are
// enabled; __builtin_sqrtf and `/` are correctly rounded under
// -fhip-fp32-correctly-rounded-divide-sqrt, which the build passes.
__device__ inline float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ inline float div_rn(float a, float b) { return a / b; }

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Mask word x of a leaf (bits y*8+z) from the 8-bit z-rows held by lanes
// (x, 0..7); valid in every lane of the group.
__device__ inline uint64_t gather_word(uint32_t row, int y) {
    uint64_t w = (uint64_t)row << (8 * y);
    w |= shfl_xor_u64(w, 1);
    w |= shfl_xor_u64(w, 2);
    w |= shfl_xor_u64(w, 4);
    return w;
}

// Two HIP events around a stretch of work on a stream.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        BS_HIP(hipEventCreate(&a));
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            throw Error(BS_ERR_HIP, "hipEventCreate failed");
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { BS_HIP(hipEventRecord(a, s)); }
    void stop(hipStream_t s) { BS_HIP(hipEventRecord(b, s)); }
    double ms() {
        BS_HIP(hipEventSynchronize(b));
        float f = 0.0f;
        BS_HIP(hipEventElapsedTime(&f, a, b));
        return f;
    }
};

inline unsigned div_up(uint64_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

}  // namespace bs
