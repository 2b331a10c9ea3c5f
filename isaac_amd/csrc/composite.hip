// Ordered sort-last compositing of the images of several sub-volumes rendered on
// one GPU: the per-GPU part of the reference's IceT step (lib/isaac.hpp:800-831,
// 2015-2056), without the exchange between ranks.
//
//   * composite order: each sub-volume's distance to the camera, as the reference
//     computes it for icetCompositeOrder (lib/isaac.hpp:2021-2048), sorted
//     ascending, equal distances in sub-volume order (std::multimap);
//   * blending: IceT's ICET_COMPOSITE_MODE_BLEND on ICET_IMAGE_COLOR_RGBA_UBYTE
//     (the mode and format set at lib/isaac.hpp:821-823), premultiplied "over" on
//     bytes: dest = (IceTUByte)(dest * (255 - src.a) / 255 + src), integer
//     division, per channel, src the nearer image;
//   * association: a pairwise tree over the composite order (images 2i and 2i+1
//     first, then the pairs, ...), which is binary swap's
//     (ICET_SINGLE_IMAGE_STRATEGY_BSWAP, lib/isaac.hpp:811) for power-of-two
//     counts.  Byte rounding makes "over" not associative, so the tree is part of
//     the result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "isaac_hip.h"

namespace isaac_hip
{
    constexpr int kMaxCompositeImages = 64;
    constexpr int kCompositeLevels = 7; // 2^6 = 64 images -> levels 0..6

    struct CompositeParams
    {
        const uint32_t* img[kMaxCompositeImages]; // RGBA8 images in composite order (nearest first)
        int32_t n;
        int64_t pixels;
        uint32_t* out;
    };

    // IceT's premultiplied "over" on RGBA bytes: front over back
    __device__ inline uint32_t over_ubyte(uint32_t front, uint32_t back)
    {
        const uint32_t f = 255u - (front >> 24);
        uint32_t r = 0;
#pragma unroll
        for(int c = 0; c < 4; c++)
        {
            const uint32_t d = (back >> (8 * c)) & 0xffu, s = (front >> (8 * c)) & 0xffu;
            r |= ((d * f / 255u + s) & 0xffu) << (8 * c);
        }
        return r;
    }

    // 4 pixels per thread; a binary counter over the images gives the pairwise tree:
    // after image k, ctz(k + 1) pairs close (stack level j holds a finished block of 2^j
    // images); what is left at the end closes from the back
    __global__ __launch_bounds__(256) void composite_kernel(const CompositeParams P)
    {
        const int64_t quads = (P.pixels + 3) / 4;
        for(int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < quads; q += int64_t(gridDim.x) * blockDim.x)
        {
            const int64_t p0 = q * 4;
            const bool full = p0 + 4 <= P.pixels;
            uint32_t st[kCompositeLevels][4];
#pragma unroll
            for(int j = 0; j < kCompositeLevels; j++)
#pragma unroll
                for(int e = 0; e < 4; e++)
                    st[j][e] = 0u;
            for(int k = 0; k < P.n; k++)
            {
                uint32_t t[4];
                if(full)
                {
                    const uint4 v = *reinterpret_cast<const uint4*>(P.img[k] + p0);
                    t[0] = v.x;
                    t[1] = v.y;
                    t[2] = v.z;
                    t[3] = v.w;
                }
                else
                {
#pragma unroll
                    for(int e = 0; e < 4; e++)
                        t[e] = p0 + e < P.pixels ? P.img[k][p0 + e] : 0u;
                }
                const int m = __builtin_ctz(uint32_t(k + 1));
#pragma unroll
                for(int j = 0; j < kCompositeLevels; j++)
                    if(j < m)
#pragma unroll
                        for(int e = 0; e < 4; e++)
                            t[e] = over_ubyte(st[j][e], t[e]);
#pragma unroll
                for(int j = 0; j < kCompositeLevels; j++)
                    if(j == m)
#pragma unroll
                        for(int e = 0; e < 4; e++)
                            st[j][e] = t[e];
            }
            // blocks left for the set bits of n: the lowest is the back-most
            uint32_t r[4] = {0u, 0u, 0u, 0u};
            bool have = false;
#pragma unroll
            for(int j = 0; j < kCompositeLevels; j++)
            {
                if(!((uint32_t(P.n) >> j) & 1u))
                    continue;
#pragma unroll
                for(int e = 0; e < 4; e++)
                    r[e] = have ? over_ubyte(st[j][e], r[e]) : st[j][e];
                have = true;
            }
            if(full)
                *reinterpret_cast<uint4*>(P.out + p0) = make_uint4(r[0], r[1], r[2], r[3]);
            else
#pragma unroll
                for(int e = 0; e < 4; e++)
                    if(p0 + e < P.pixels)
                        P.out[p0 + e] = r[e];
        }
    }
} // namespace isaac_hip

using namespace isaac_hip;

extern "C"
{
    int isaac_hip_composite_order(
        int32_t n,
        const int32_t* position,
        const int32_t* local_size,
        const int32_t global_size[3],
        const float scale[3],
        const double modelview[16],
        int32_t* order)
    {
        if(n < 1 || !position || !local_size || !global_size || !scale || !modelview || !order)
            return ISAAC_HIP_ERR_ARG;
        // the constructor's float conversions (lib/isaac.hpp:659-661,796-797)
        uint64_t gs[3];
        for(int i = 0; i < 3; i++)
        {
            if(global_size[i] < 1)
                return ISAAC_HIP_ERR_ARG;
            gs[i] = uint64_t(float(global_size[i]) * scale[i]);
        }
        const uint64_t maxScaled = std::max(gs[0], std::max(gs[1], gs[2]));
        if(maxScaled < 2)
            return ISAAC_HIP_ERR_ARG;
        std::vector<float> dist(static_cast<size_t>(n));
        for(int r = 0; r < n; r++)
        {
            double pt[4];
            for(int i = 0; i < 3; i++)
            {
                const int32_t ps = int32_t(float(position[3 * r + i]) * scale[i]);
                const uint64_t ls = uint64_t(float(local_size[3 * r + i]) * scale[i]);
                // (positionScaled + (localSizeScaled - globalSizeScaled) / 2) / (maxSizeScaled / 2), the
                // divisor an integer division (lib/isaac.hpp:2021-2025)
                pt[i] = (double(ps) + (double(ls) - double(gs[i])) / 2.0) / double(maxScaled / 2);
            }
            pt[3] = 1.0;
            // glm: modelview * point (column-major), then length
            double v[4];
            for(int row = 0; row < 4; row++)
                v[row] = (modelview[row] * pt[0] + modelview[4 + row] * pt[1])
                    + (modelview[8 + row] * pt[2] + modelview[12 + row] * pt[3]);
            dist[size_t(r)] = float(std::sqrt((v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3])));
        }
        std::vector<int32_t> idx(static_cast<size_t>(n));
        for(int r = 0; r < n; r++)
            idx[size_t(r)] = r;
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return dist[size_t(a)] < dist[size_t(b)]; });
        std::copy(idx.begin(), idx.end(), order);
        return 0;
    }

    int isaac_hip_composite(
        int32_t n,
        const uint8_t* const* images,
        const int32_t* order,
        int64_t pixels,
        uint8_t* out,
        void* stream)
    {
        if(n < 1 || n > kMaxCompositeImages || !images || !order || !out || pixels < 0)
            return ISAAC_HIP_ERR_ARG;
        CompositeParams P{};
        std::vector<bool> seen(static_cast<size_t>(n), false);
        for(int k = 0; k < n; k++)
        {
            const int32_t o = order[k];
            if(o < 0 || o >= n || seen[size_t(o)] || !images[o])
                return ISAAC_HIP_ERR_ARG;
            seen[size_t(o)] = true;
            // 16-byte loads of 4 pixels: images and output start 16-byte aligned
            if(reinterpret_cast<uintptr_t>(images[o]) & 15u)
                return ISAAC_HIP_ERR_ARG;
            P.img[k] = reinterpret_cast<const uint32_t*>(images[o]);
        }
        if(reinterpret_cast<uintptr_t>(out) & 15u)
            return ISAAC_HIP_ERR_ARG;
        P.n = n;
        P.pixels = pixels;
        P.out = reinterpret_cast<uint32_t*>(out);
        if(pixels == 0)
            return 0;
        int64_t blocks = ((pixels + 3) / 4 + 255) / 256;
        if(blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(composite_kernel, dim3(unsigned(blocks)), dim3(256), 0, (hipStream_t) stream, P);
        return hipGetLastError() == hipSuccess ? 0 : ISAAC_HIP_ERR_DEVICE;
    }
}
