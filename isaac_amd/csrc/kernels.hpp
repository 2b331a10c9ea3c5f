// Launch entry points of the gfx950 kernels (host side of the HIP library).
#pragma once

#include "common.hpp"

namespace isaac_hip
{
    hipError_t launch_classify(const ClassifyParams& P, hipStream_t s);
    hipError_t launch_persistent_copy(
        const float* src,
        const int64_t stride[3],
        int dim,
        const ChainDesc& chain,
        int gX,
        int gY,
        int gZ,
        int guard,
        float* dst,
        hipStream_t s);
    // the fused march (clear, particles, iso + volume), the SSAO pass and shading
    // + volume blend; the host records its per-kernel events between them
    hipError_t launch_march(const FrameParams& F, const ParticleParams& PP, bool counting, hipStream_t s);
    hipError_t launch_ssao(const FrameParams& F, hipStream_t s);
    hipError_t launch_shade(const FrameParams& F, hipStream_t s);

    // ---- precision form: field bricks and the post-classification march
    // src: local voxel (0,0,0) of an AoS float source (dim components, element strides);
    // scale: nbricks*dim floats; out: nbricks*dim*648 pairs (float2 or 2 x 16 bit)
    hipError_t launch_bricks(
        const float* src,
        const int64_t stride[3],
        int dim,
        int X,
        int Y,
        int Z,
        int dtype,
        float* scale,
        void* out,
        hipStream_t s);
    // batched form: the persistent march over all frames (grid = persistent workgroups),
    // then SSAO and shading over a (maxTiles, nframes) grid
    hipError_t launch_batch_march(const BatchParams& B, int dtype, bool linear, size_t lds, int grid, hipStream_t s);
    hipError_t launch_batch_ssao(const FrameParams* frames, int nframes, int maxTiles, hipStream_t s);
    hipError_t launch_batch_shade(const FrameParams* frames, int nframes, int maxTiles, hipStream_t s);
    hipError_t launch_brick_march(const FrameParams& F, const PostParams& Q, const ParticleParams& PP, bool counting, hipStream_t s);

    // ---- particle cell lists
    struct BinParams
    {
        const float* pos; // n x 3
        const float* attr; // n x dim
        const float* radius; // n or nullptr
        int32_t dim;
        int64_t n;
        int32_t grid[3];
        float cw[3]; // cell width = local / grid (float)
        ChainDesc chain;
        const float4* lut; // T entries, device
        int32_t T;
        int32_t filter; // drop particles whose TF alpha < 0.5
    };

    struct BinBuffers
    {
        int32_t* cellOf; // n
        uint32_t* rank; // n: arrival rank of the particle inside its cell
        float4* rel; // n: rel.xyz, lut index (or -1 if invisible) in .w
        unsigned long long* count64; // ncells: all particles (low 32 bits), visible ones (high 32 bits)
        int32_t* count; // ncells (all particles)
        int32_t* offset; // ncells + 1
        int32_t* voffset; // ncells + 1
        int32_t* perm; // n
        float4* recs; // visible records in cell order
        float* recRadius; // per record radius or nullptr
        unsigned long long* scanTmp; // scan workspace (bin_scan_workspace_elems entries)
    };

    size_t bin_scan_workspace_elems(int64_t ncells);
    hipError_t launch_bin(const BinParams& P, const BinBuffers& B, hipStream_t s);
    hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
} // namespace isaac_hip
