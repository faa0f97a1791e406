// Device lookup of leaves by leaf coordinate (origin >> 3), shared by the
// kernels that walk across leaf borders.
#pragma once
#include <algorithm>
#include <climits>

#include "common.hpp"

namespace bs {

// Leaf lookup by leaf coordinate (origin >> 3): a dense cell grid over the
// reachable box, or an open-addressing hash when that box is too large.  A
// cell holds the pool slot, -1 when empty, -2 while a leaf is being created
// (read as empty: a new leaf has no active voxel yet).
struct LeafMap {
    int dense;
    int32_t* cells;
    int32_t x0, y0, z0;
    uint32_t nx, ny, nz;
    uint64_t* keys;  // hash only
    uint32_t cap_mask;

    __device__ int32_t find(int32_t lx, int32_t ly, int32_t lz) const {
        if (dense) {
            const uint32_t ix = (uint32_t)(lx - x0), iy = (uint32_t)(ly - y0), iz = (uint32_t)(lz - z0);
            if (ix >= nx || iy >= ny || iz >= nz) return -1;
            const int32_t s = cells[((uint64_t)ix * ny + iy) * nz + iz];
            return s >= 0 ? s : -1;
        }
        const uint64_t key = leaf_key(lx, ly, lz);
        uint32_t h = hash_u64(key) & cap_mask;
        for (uint32_t probe = 0; probe <= cap_mask; ++probe) {
            const uint64_t k = keys[h];
            if (k == key) {
                const int32_t s = cells[h];
                return s >= 0 ? s : -1;
            }
            if (k == EMPTY_KEY) return -1;
            h = (h + 1) & cap_mask;
        }
        return -1;
    }

    // Claims the cell of a leaf that does not exist yet: the cell address, or
    // nullptr when the leaf exists (or is being created) already.  *fail is
    // set when the coordinate is outside the map.
    __device__ int32_t* claim(int32_t lx, int32_t ly, int32_t lz, uint32_t* fail) const {
        if (dense) {
            const uint32_t ix = (uint32_t)(lx - x0), iy = (uint32_t)(ly - y0), iz = (uint32_t)(lz - z0);
            if (ix >= nx || iy >= ny || iz >= nz) {
                atomicOr(fail, 1u);
                return nullptr;
            }
            int32_t* c = &cells[((uint64_t)ix * ny + iy) * nz + iz];
            return atomicCAS(c, -1, -2) == -1 ? c : nullptr;
        }
        const uint64_t key = leaf_key(lx, ly, lz);
        uint32_t h = hash_u64(key) & cap_mask;
        for (uint32_t probe = 0; probe <= cap_mask; ++probe) {
            const unsigned long long prev =
                atomicCAS((unsigned long long*)&keys[h], (unsigned long long)EMPTY_KEY, (unsigned long long)key);
            if (prev == EMPTY_KEY) return &cells[h];
            if (prev == key) return nullptr;
            h = (h + 1) & cap_mask;
        }
        atomicOr(fail, 2u);
        return nullptr;
    }

    // Registers a leaf before any sweep (single writer per key).
    __device__ bool put(int32_t lx, int32_t ly, int32_t lz, int32_t slot) const {
        uint32_t fail = 0;
        int32_t* c = claim(lx, ly, lz, &fail);
        if (!c) return false;
        *c = slot;
        return true;
    }
};

// Registers leaves [0, n) of an origin array (x, y, z, 0 per leaf) in a
// map; *fail is set when a key is taken twice or does not fit.
static __global__ void k_leafmap_insert(uint64_t n, const int4* __restrict__ origins, LeafMap map, uint32_t* fail) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 o = origins[i];
    if (!map.put(o.x >> 3, o.y >> 3, o.z >> 3, (int32_t)i)) atomicOr(fail, 1u);
}

// Leaf-coordinate box (origin >> 3) of leaves [0, n): box[0..2] the minimum,
// box[3..5] the maximum; box must hold k_leaf_box_init's start values.
static __global__ void k_leaf_box_init(int32_t* box) {
    if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? INT32_MAX : INT32_MIN;
}
static __global__ __launch_bounds__(256) void k_leaf_box(uint64_t n, const int4* __restrict__ origins, int32_t* box) {
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int4 o = origins[i];
        const int32_t c[3] = {o.x >> 3, o.y >> 3, o.z >> 3};
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], c[k]);
            hi[k] = max(hi[k], c[k]);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = min(lo[k], __shfl_xor(lo[k], off, 64));
            hi[k] = max(hi[k], __shfl_xor(hi[k], off, 64));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&box[k], lo[k]);
            atomicMax(&box[3 + k], hi[k]);
        }
}

// The leaf-coordinate box of a volume's n > 0 leaves, computed on the device
// (six integers come back instead of every origin).
inline void leaf_box(bs_ctx* ctx, const int32_t* origins, uint64_t n, int32_t lo[3], int32_t hi[3]) {
    hipStream_t st = ctx->stream;
    int32_t* d = ctx->scratch.as<int32_t>(S_LEAF_BOX, 6);
    k_leaf_box_init<<<1, 64, 0, st>>>(d);
    BS_HIP(hipGetLastError());
    k_leaf_box<<<(unsigned)std::min<uint64_t>(div_up(n, 256), 1024), 256, 0, st>>>(
        n, reinterpret_cast<const int4*>(origins), d);
    BS_HIP(hipGetLastError());
    int32_t h[6];
    BS_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
    BS_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 3; ++k) {
        lo[k] = h[k];
        hi[k] = h[3 + k];
    }
}

}  // namespace bs
