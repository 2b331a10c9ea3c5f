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
    hipError_t launch_tex_bricks(const uint2* tex, uint2* out, int X, int Y, int Z, hipStream_t s);
    // shade: the occlusion kernel also shades its pixels (only where ssao_shades(F))
    hipError_t launch_ssao(const FrameParams& F, hipStream_t s, bool shade = false);
    bool ssao_shades(const FrameParams& F);
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
    hipError_t launch_batch_march(const BatchParams& B, int dtype, bool linear, int sig, size_t lds, int grid, hipStream_t s);
    // component counts of the enabled sources as base-5 digits, slot 0 lowest (-1: no source)
    int post_signature(const PostParams& Q);
    bool natural_chains(const PostParams& Q);
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
        unsigned long long* count64; // 2 * ncells uint32: all particles per cell, then a flag per cell (1: holds visible particles)
        int64_t ncells;
        int32_t* count; // ncells (all particles)
        int32_t* offset; // ncells + 1
        int2* vrange; // ncells: [start, end) of the cell's visible records in recs (start = offset)
        int32_t* perm; // n
        float4* recs; // n: each cell's visible records from its offset on, ascending particle index
        float* recRadius; // per record radius or nullptr
        unsigned long long* scanTmp; // scan workspace (bin_scan_workspace_elems entries)
        unsigned long long* visMask; // ceil(n / 64): visible-records build, bit i % 64 of word i / 64: particle i is visible
    };

    // ---- field sources: advection noise (updateNoiseTexture, lib/isaac.hpp:499-562) and
    // the per-frame advection texture (GenerateAdvectionTextureKernel,
    // lib/isaac/isaac_common_kernel.hpp:500-550)
    // noise (X*Y*Z floats, x fastest) = the seeded, blurred, scaled noise; tmp: scratch of the same size
    hipError_t launch_noise(float* noise, float* tmp, int X, int Y, int Z, const float scale[3], uint32_t seedPoints, hipStream_t s);

    struct AdvectParams
    {
        const float* field; // local voxel (0,0,0) of the float3 field source
        int64_t stride[3]; // element strides (elements of 3 floats)
        const float* noise; // X*Y*Z
        const uint8_t* history; // last frame's advection texture, X*Y*Z
        uint8_t* out; // this frame's, X*Y*Z
        int32_t X, Y, Z;
        float stepSize; // advectionStepFactor * ISAAC_MAX_ADVECTION_STEP_SIZE
        float scale[3];
        float historyWeight;
        int32_t seeding; // 1: the noise seeds this frame (timeStep % period < seedingTime)
    };
    hipError_t launch_advect(const AdvectParams& P, hipStream_t s);

    // ---- min/max request (MinMaxKernel / MinMaxParticleKernel,
    // lib/isaac/isaac_min_max_kernel.hpp:22-97).  Every value carries the position
    // of its element in the reference's visiting order (column x + y*X, then z, then
    // the particle's place in its cell), and ties keep the earlier element, so the
    // parallel reduction returns the bits the sequential loops return (+0 vs -0).
    // NaN never replaces a value; the start values are FLT_MAX / -FLT_MAX.
    struct MinMaxPartial
    {
        float vmin, vmax;
        uint64_t kmin, kmax; // 0 = the start value
    };

    constexpr int kMinMaxGroups = 1024; // partial results per source (workgroups)

    struct MinMaxVolumeParams
    {
        const float* data; // local voxel (0,0,0)
        int64_t stride[3]; // element strides (elements of dim floats)
        int32_t dim;
        int32_t applyChain; // 0: data already holds the chain values (texture of a non-persistent source)
        int32_t X, Y, Z;
        uint32_t divX_mul, divX_shift, divXY_mul, divXY_shift;
        int64_t nvox; // < 2^32
        ChainDesc chain;
    };

    struct MinMaxParticleParams
    {
        const float* pos; // n x 3
        const float* attr; // n x dim
        int32_t dim;
        int64_t n;
        int32_t grid[3];
        float cw[3]; // cell width = local / grid (float), as in the cell lists
        ChainDesc chain;
    };

    // the workgroup's (256 threads) combined partial, valid in thread 0
    __device__ inline MinMaxPartial mm_block_reduce(MinMaxPartial a)
    {
        for(int off = 32; off > 0; off >>= 1)
        {
            const float v0 = __shfl_xor(a.vmin, off), v1 = __shfl_xor(a.vmax, off);
            const uint64_t k0 = __shfl_xor(a.kmin, off), k1 = __shfl_xor(a.kmax, off);
            mm_take_min(a.vmin, a.kmin, v0, k0);
            mm_take_max(a.vmax, a.kmax, v1, k1);
        }
        __shared__ MinMaxPartial waves[4];
        if((threadIdx.x & 63) == 0)
            waves[threadIdx.x >> 6] = a;
        __syncthreads();
        if(threadIdx.x == 0)
            for(int w = 1; w < 4; w++)
            {
                mm_take_min(a.vmin, a.kmin, waves[w].vmin, waves[w].kmin);
                mm_take_max(a.vmax, a.kmax, waves[w].vmax, waves[w].kmax);
            }
        return a;
    }

    // partials: kMinMaxGroups entries of workspace; out: (min, max) of the source
    hipError_t launch_min_max_volume(const MinMaxVolumeParams& P, MinMaxPartial* partials, float2* out, hipStream_t s);
    hipError_t launch_min_max_particles(const MinMaxParticleParams& P, MinMaxPartial* partials, float2* out, hipStream_t s);
    // reduces `count` partials into out (one workgroup)
    hipError_t launch_min_max_final(const MinMaxPartial* partials, int count, float2* out, hipStream_t s);

    size_t bin_scan_workspace_elems(int64_t ncells);
    hipError_t launch_bin(const BinParams& P, const BinBuffers& B, hipStream_t s);
    // only the render records of the visible particles (count, offset, vrange, recs per
    // cell over visible particles); count, offset and perm then do not hold the full list
    hipError_t launch_bin_visible(const BinParams& P, const BinBuffers& B, hipStream_t s);
    hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
} // namespace isaac_hip
