// The learned tile classifier of the surrogate sign mode: a 3 -> 64 -> 64 ->
// 64 -> 1 ReLU network (BF16 weights, FP32 accumulation) that predicts a
// tile's smallest |SDF| (in voxels) from the tile centre, with its inference
// and training kernels (surrogate.hip).
#pragma once
#include "common.hpp"

struct bs_mlp {
    int device = 0;
    float* params = nullptr;  // device, BS_MLP_PARAMS FP32 master weights
    float* adam_m = nullptr;  // device, first moments
    float* adam_v = nullptr;  // device, second moments
    void* packed = nullptr;   // device, BF16 operand fragments + biases (MlpPacked)
    uint64_t adam_t = 0;      // Adam steps taken
    float in_shift[3] = {0.0f, 0.0f, 0.0f};
    float in_scale = 1.0f;    // network input = (point - in_shift) * in_scale
};

namespace bs {

// Shared-sign roots of the surrogate mode for every leaf of `vol` (values and
// masks written by the distance pass): tiles are the eight 4^3 blocks of a
// leaf.  A culled tile -- predicted smallest |SDF| >= threshold voxels, or,
// with no network, its true smallest |SDF| >= threshold -- takes its active
// voxel of largest |d| as representative r, and each of its voxels v with
// |d_v| > |v - r| (1 + 2^-10) + extra shares r's sign (no surface point lies
// in the ball around v that holds r); every other voxel is its own root.
// roots / rep_cnt as the distance pass's shared-sign groups.  counters
// receives 4 values: tiles with an active voxel, culled tiles, culled tiles
// with a voxel that failed the check, voxels that failed it.
void surrogate_groups(bs_ctx* ctx, hipStream_t st, const bs_volume* vol, float h, float extra, uint16_t* roots,
                      uint32_t* rep_cnt, unsigned long long* counters);

}  // namespace bs
