// gfx950 kernels of the per-rank render path.
//
//   classify_kernel        MergeToCombinedTextureKernel  lib/isaac/isaac_common_kernel.hpp:553-615,690-758
//   march_kernel           ClearBufferKernel + ParticleRenderKernel + CombinedIsoRenderKernel +
//                          the march of CombinedVolumeRenderKernel, fused per pixel
//                          lib/isaac/isaac_common_kernel.hpp:195-217, lib/isaac/isaac_particle_kernel.hpp:117-260,
//                          lib/isaac/isaac_iso_kernel.hpp:25-128, lib/isaac/isaac_volume_kernel.hpp:23-82
//   ssao_tiled_kernel      SSAOKernel (live 7x7 part)     lib/isaac/isaac_ssao_kernel.hpp:161-182
//                          (with the shading of its pixels fused in render modes 0, 1, 3)
//   shade_kernel           ShadingKernel + the volume pass's final blend
//                          lib/isaac/isaac_common_kernel.hpp:220-320, lib/isaac/isaac_volume_kernel.hpp:83-86
//
// The fusions keep the reference's quantisation points: every stage that the
// reference stores as RGBA8 in the G-buffer is rounded to RGBA8 in registers
// at the same point, so the fused kernels produce the stage sequence's bytes.
//
// Launch shape: 256-thread workgroups = four wave64s, each wave an 8x8 pixel
// tile (16x16 pixels per workgroup), so the 64 rays of a wave sample
// neighbouring voxels; the march deals its tiles to the 8 XCDs (each with its
// own L2) in super-tiles or bands (march_tile), and can run one-wave workgroups.
#include "kernels.hpp"

namespace isaac_hip
{
#define DITHER_TABLE(type, name, n) __constant__ type name[n]
#include "dither_tables.inc"
#undef DITHER_TABLE

    // ------------------------------------------------------------------ functor chain
    // pow stays out of line: inlined, its ~70 instructions per component and
    // functor slot made up most of the code of the marches that evaluate the
    // chain per sample (rarely used, and the call is only taken for it).
    __device__ __noinline__ float chain_pow(float a, float b)
    {
        return powf(a, b);
    }

    // f2(f1(f0(v,p0),p1),p2).x with the feature dimension propagating
    // (lib/isaac/isaac_functor_chain.hpp:96-108, lib/isaac/isaac_functors.hpp:42-183).
    __device__ inline float apply_chain(const ChainDesc& c, float v0, float v1, float v2, float v3, int n)
    {
        float v[4] = {v0, v1, v2, v3};
#pragma unroll
        for(int k = 0; k < kMaxFunctors; k++)
        {
            const int op = c.op[k];
            if(op == OP_ADD)
            {
#pragma unroll
                for(int i = 0; i < 4; i++)
                    if(i < n)
                        v[i] = v[i] + c.p[k][i];
            }
            else if(op == OP_MUL)
            {
#pragma unroll
                for(int i = 0; i < 4; i++)
                    if(i < n)
                        v[i] = v[i] * c.p[k][i];
            }
            else if(op == OP_LENGTH)
            {
                float d;
                if(n == 1)
                    d = v[0] * v[0];
                else if(n == 2)
                    d = v[0] * v[0] + v[1] * v[1];
                else if(n == 3)
                    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                else
                    d = (v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3]);
                v[0] = sqrtf(d);
                n = 1;
            }
            else if(op == OP_POW)
            {
#pragma unroll
                for(int i = 0; i < 4; i++)
                    if(i < n)
                        v[i] = chain_pow(v[i], c.p[k][i]);
            }
            else if(op == OP_SUM)
            {
                float s = 0.0f;
#pragma unroll
                for(int i = 0; i < 4; i++)
                    if(i < n)
                        s = s + v[i];
                v[0] = s;
                n = 1;
            }
        }
        return v[0];
    }

    // x / d for a divisor fixed per launch, with r = RN(1/d) computed once on the
    // host: x*r plus one fma residual correction (Markstein's correction), three
    // instructions instead of the correctly rounded division sequence.  Equal to the
    // IEEE quotient for finite operands away from overflow and underflow
    // (tests/cpp/divq_check.c: 0 of 2e8 seeded quotients differ, a quarter of them
    // next to rounding boundaries; for d = 255 every x in [0, 256],
    // tests/cpp/div255_check.c).
    __device__ inline float div_by(float x, float d, float r)
    {
        const float q0 = x * r;
        const float e = __builtin_fmaf(-q0, d, x);
        return __builtin_fmaf(e, r, q0);
    }

    __device__ inline int tf_index(float v, int T)
    {
        return tf_index_of(v * float(T), T);
    }

    // ------------------------------------------------------------------ classification
    __device__ inline uint32_t dither_b4(int mode, int x, int y, int z, float r, float g, float b, float a)
    {
        if(mode == 0)
        {
            return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
        }
        // threshold of the mode's mask (applyDither3D, lib/isaac/isaac_dither_kernel.hpp:20-323):
        // 1 ordered 4^3, 2 blue noise 4^3, 3 blue noise 8^3, any other mode blue noise 16^3
        float d;
        if(mode == 1)
            d = float(kDitherOrdered4[(x & 3) + (y & 3) * 4 + (z & 3) * 16]) / 64.0f;
        else if(mode == 2)
            d = float(kDitherBlue4[(x & 3) + (y & 3) * 4 + (z & 3) * 16]) / 64.0f;
        else if(mode == 3)
            d = float(kDitherBlue8[(x & 7) + (y & 7) * 8 + (z & 7) * 64]) / 512.0f;
        else
            d = float(kDitherBlue16[(x & 15) + (y & 15) * 16 + (z & 15) * 256]) / 4096.0f;
        auto q = [d](float v) -> uint32_t
        {
            float s = v * 255.0f;
            int i = (int) s;
            i += ((s - floorf(s)) > d) ? 1 : 0;
            return (uint32_t) iclamp(i, 0, 255);
        };
        return q(r) | (q(g) << 8) | (q(b) << 16) | (q(a) << 24);
    }

    // One workgroup stages up to nLut transfer functions (T float4 each) in LDS
    // and then walks voxels with a grid-stride loop (x fastest, coalesced
    // source reads and texel writes).
    __global__ __launch_bounds__(256) void classify_kernel(const ClassifyParams P)
    {
        extern __shared__ float4 lds_lut[];
        const int T = P.T;
        for(int i = threadIdx.x; i < P.nLut * T; i += blockDim.x)
            lds_lut[i] = P.luts[i];
        __syncthreads();
        const int64_t stride = int64_t(gridDim.x) * blockDim.x;
        for(int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < P.nvox; idx += stride)
        {
            const uint32_t u = (uint32_t) idx;
            const uint32_t zq = fastdiv(u, P.divXY_mul, P.divXY_shift);
            const uint32_t rem = u - zq * uint32_t(P.X * P.Y);
            const uint32_t yq = fastdiv(rem, P.divX_mul, P.divX_shift);
            const int x = int(rem - yq * uint32_t(P.X)), y = int(yq), z = int(zq);
            float vr = 0.f, vg = 0.f, vb = 0.f, va = 0.f;
            float ir = 0.f, ig = 0.f, ib = 0.f, ia = 0.f;
            for(int s = 0; s < P.nsrc; s++)
            {
                const ClassifySource& S = P.src[s];
                if(!S.active)
                    continue;
                const float* e = S.data + (int64_t(x) * S.stride[0] + int64_t(y) * S.stride[1] + int64_t(z) * S.stride[2]) * S.dim;
                float c0 = e[0];
                float c1 = S.dim > 1 ? e[1] : 0.f;
                float c2 = S.dim > 2 ? e[2] : 0.f;
                float c3 = S.dim > 3 ? e[3] : 0.f;
                const float tv = apply_chain(S.chain, c0, c1, c2, c3, S.dim);
                const int li = tf_index(tv, T);
                const float4 col = (S.lutSlot < P.nLut) ? lds_lut[S.lutSlot * T + li] : P.luts[S.lutSlot * T + li];
                // field sources: alpha blended with the advection texture
                // (MergeAdvectionToCombinedTextureIterator, lib/isaac/isaac_common_kernel.hpp:655-656)
                float ca = col.w;
                if(S.adv)
                    ca = ca * div_by(float(S.adv[idx]), 255.0f, 1.0f / 255.0f);
                if(S.weight > 0.0f)
                {
                    float a = ca * S.weight;
                    vr = vr + col.x * a;
                    vg = vg + col.y * a;
                    vb = vb + col.z * a;
                    va = va + a;
                }
                if(S.iso > 0.0f)
                {
                    float a = div_by(ca, S.iso, S.invIso) * 0.5f;
                    if(ia < a)
                    {
                        ir = col.x;
                        ig = col.y;
                        ib = col.z;
                        ia = a;
                    }
                }
            }
            if(P.totalWeight != 0.0f)
            {
                vr = div_by(vr, P.totalWeight, P.invTotalWeight);
                vg = div_by(vg, P.totalWeight, P.invTotalWeight);
                vb = div_by(vb, P.totalWeight, P.invTotalWeight);
                va = div_by(va, P.totalWeight, P.invTotalWeight);
            }
            else
            {
                vr = vg = vb = va = 0.0f;
            }
            const uint32_t vol = dither_b4(P.ditherMode, x, y, z, vr, vg, vb, va);
            const uint32_t q0 = unorm8(ir), q1 = unorm8(ig), q2 = unorm8(ib), q3 = unorm8(ia);
            // one 8-byte texel per voxel: combined volume RGBA8, combined iso RGBA8
            const uint32_t iso = q0 | (q1 << 8) | (q2 << 16) | (q3 << 24);
            P.tex[idx] = make_uint2(vol, iso);
            if(P.texH)
                P.texH[idx] = half_texel(vol, iso);
        }
    }

    // non-persistent source -> functor-applied float texture incl. guard
    // (UpdatePersistendTextureKernel, lib/isaac/isaac_common_kernel.hpp:432-451)
    __global__ __launch_bounds__(256) void persistent_copy_kernel(
        const float* src,
        int64_t sx,
        int64_t sy,
        int64_t sz,
        int dim,
        ChainDesc chain,
        int gX,
        int gY,
        int gZ,
        int guard,
        float* dst)
    {
        const int64_t n = int64_t(gX) * gY * gZ;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        {
            int x = int(i % gX), y = int((i / gX) % gY), z = int(i / (int64_t(gX) * gY));
            int cx = x - guard, cy = y - guard, cz = z - guard;
            const float* e = src + (int64_t(cx) * sx + int64_t(cy) * sy + int64_t(cz) * sz) * dim;
            float c0 = e[0], c1 = dim > 1 ? e[1] : 0.f, c2 = dim > 2 ? e[2] : 0.f, c3 = dim > 3 ? e[3] : 0.f;
            dst[i] = apply_chain(chain, c0, c1, c2, c3, dim);
        }
    }

    // ------------------------------------------------------------------ rays
    struct Ray
    {
        F3 dir, start;
        float startDepth, endDepth;
        bool isClipped;
        F3 clipN;
    };

    // pixelToRay (lib/isaac/isaac_common_kernel.hpp:51-92); glm mat4*vec4 = (m0*x+m1*y)+(m2*z+m3*w)
    __device__ inline Ray pixel_to_ray(const FrameParams& F, int px, int py)
    {
        const float vx = float(px) / F.fbW * 2.0f - 1.0f;
        const float vy = float(py) / F.fbH * 2.0f - 1.0f;
        float s[4], e[4];
#pragma unroll
        for(int r = 0; r < 4; r++)
        {
            const float a0 = F.inv[0 * 4 + r] * vx;
            const float a1 = F.inv[1 * 4 + r] * vy;
            const float s0 = a0 + a1;
            const float bs = F.inv[2 * 4 + r] * -1.0f;
            const float be = F.inv[2 * 4 + r] * 1.0f;
            const float a3 = F.inv[3 * 4 + r] * 1.0f;
            s[r] = s0 + (bs + a3);
            e[r] = s0 + (be + a3);
        }
        Ray ray;
        F3 st = f3(s[0] / s[3], s[1] / s[3], s[2] / s[3]);
        F3 en = f3(e[0] / e[3], e[1] / e[3], e[2] / e[3]);
        st = muls(st, F.maxSize);
        en = muls(en, F.maxSize);
        ray.dir = normalize3(sub(en, st));
        ray.start = st;
        ray.isClipped = false;
        ray.clipN = f3(0.f, 0.f, 0.f);
        ray.startDepth = 0.f;
        ray.endDepth = 0.f;
        return ray;
    }

    // clipRay (lib/isaac/isaac_common_kernel.hpp:94-193)
    __device__ inline bool clip_ray(const FrameParams& F, Ray& ray)
    {
        ray.start = add(ray.start, f3(F.off[0], F.off[1], F.off[2]));
        float bmin[3], bmax[3];
        const float st[3] = {ray.start.x, ray.start.y, ray.start.z};
        const float dr[3] = {ray.dir.x, ray.dir.y, ray.dir.z};
#pragma unroll
        for(int i = 0; i < 3; i++)
        {
            bmin[i] = (F.bbMin[i] - st[i]) / dr[i];
            bmax[i] = (F.bbMax[i] - st[i]) / dr[i];
            if(bmax[i] < bmin[i])
            {
                float t = bmax[i];
                bmax[i] = bmin[i];
                bmin[i] = t;
            }
        }
        ray.startDepth = gmax(bmin[0], gmax(bmin[1], bmin[2]));
        ray.endDepth = gmin(bmax[0], gmin(bmax[1], bmax[2]));
        float cn[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for(int i = 0; i < 3; i++)
        {
            const float sgn = gsign(dr[i]);
            const bool lowOuter = F.outerLow[i] && (sgn + 1.0f) != 0.0f;
            const bool highOuter = F.outerHigh[i] && (sgn - 1.0f) != 0.0f;
            if(bmin[i] == ray.startDepth && (lowOuter || highOuter))
            {
                ray.isClipped = true;
                cn[i] = sgn;
            }
        }
        ray.clipN = f3(cn[0], cn[1], cn[2]);
        for(int i = 0; i < F.nclip; i++)
        {
            const F3 n = f3(F.cnorm[i][0], F.cnorm[i][1], F.cnorm[i][2]);
            const F3 p = f3(F.cpos[i][0], F.cpos[i][1], F.cpos[i][2]);
            const float d = dot3(ray.dir, n);
            const float depth = (dot3(p, n) - dot3(ray.start, n)) / d;
            if(d > 0.0f)
            {
                if(ray.endDepth < depth)
                    return false;
                if(ray.startDepth <= depth)
                {
                    ray.clipN = n;
                    ray.isClipped = true;
                    ray.startDepth = depth;
                }
            }
            else
            {
                if(ray.startDepth > depth)
                    return false;
                if(ray.endDepth > depth)
                    ray.endDepth = depth;
            }
        }
        ray.startDepth = gmax(ray.startDepth, kZNear);
        if(ray.startDepth > ray.endDepth || isinf(ray.startDepth) || isinf(ray.endDepth))
            return false;
        return true;
    }

    // ------------------------------------------------------------------ texture sampling
    // x / 255 in binary32 as x*r plus one fma residual correction, r = RN(1/255).
    // For every float x in [0, 256] this equals the IEEE quotient bit for bit
    // (tests/test_div255.py checks all 1,132,462,081 inputs); trilinear results
    // of RGBA8 texels always lie in [0, 255].  Three instructions instead of the
    // correctly rounded division sequence.
    __device__ inline float div255(float x)
    {
        const float r = 1.0f / 255.0f;
        const float q0 = x * r;
        const float e = __builtin_fmaf(-q0, 255.0f, x);
        return __builtin_fmaf(e, r, q0);
    }

    // The 8 trilinear taps of one sample position (or the single NEAREST tap),
    // each an 8-byte texel: .x combined volume RGBA8, .y combined iso RGBA8.
    // Tap k holds voxel (x0 + (k>>2), y0 + ((k>>1)&1), z0 + (k&1)), clamped.
    struct Taps
    {
        uint2 t[8];
        float fx, fy, fz;
    };

    // two adjacent texels (8-byte aligned; one 16-byte load)
    struct TexPair
    {
        uint2 a, b;
    };

    // index of texel (x, y, z) in the brick layout (FrameParams::texB)
    __device__ inline uint32_t brick_texel(const FrameParams& F, uint32_t x, uint32_t y, uint32_t z)
    {
        const uint32_t b = (z >> 3) * F.nbxy + (y >> 3) * F.nbx + (x >> 3);
        return b * uint32_t(kTexBrickTexels) + (z & 7u) * uint32_t(kTexBrickSlice) + (y & 7u) * uint32_t(kTexBrickRow)
            + (x & 7u);
    }

    // Sampler<LINEAR|NEAREST, CLAMP> addressing (lib/isaac/isaac_texture.hpp:182-259).  Every
    // coordinate is clamped into the texture, so a fetch at any position is in bounds.
    // BRICK reads the brick layout (texels x and x + 1 of a row stay adjacent there:
    // x & 7 = 7 is followed by the apron texel).
    template<bool LINEAR, bool BRICK>
    __device__ inline void fetch_taps(const FrameParams& F, float px, float py, float pz, Taps& T)
    {
        const uint2* __restrict__ tex = BRICK ? F.texB : F.tex;
        const int X = F.X, Y = F.Y, Z = F.Z;
        auto index = [&](uint32_t x, uint32_t y, uint32_t z)
        { return BRICK ? brick_texel(F, x, y, z) : x + uint32_t(X) * (y + uint32_t(Y) * z); };
        if(!LINEAR)
        {
            const uint32_t x = iclamp(int(px), 0, X - 1), y = iclamp(int(py), 0, Y - 1), z = iclamp(int(pz), 0, Z - 1);
            T.t[0] = tex[index(x, y, z)];
            T.fx = T.fy = T.fz = 0.0f;
            return;
        }
        const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
        const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
        T.fx = cx - flx;
        T.fy = cy - fly;
        T.fz = cz - flz;
        const int bx = int(flx), by = int(fly), bz = int(flz);
        const uint32_t x0 = iclamp(bx, 0, X - 1), x1 = iclamp(bx + 1, 0, X - 1);
        const uint32_t y0 = iclamp(by, 0, Y - 1), y1 = iclamp(by + 1, 0, Y - 1);
        const uint32_t z0 = iclamp(bz, 0, Z - 1), z1 = iclamp(bz + 1, 0, Z - 1);
        // texels x0 and x0+1 are adjacent: one 16-byte load per (y, z) row; at the
        // x edge (x1 == x0) the second texel (the next row's first, the padding
        // texel after the last, or an apron texel) is replaced by the first
        const TexPair p00 = *reinterpret_cast<const TexPair*>(tex + index(x0, y0, z0));
        const TexPair p01 = *reinterpret_cast<const TexPair*>(tex + index(x0, y0, z1));
        const TexPair p10 = *reinterpret_cast<const TexPair*>(tex + index(x0, y1, z0));
        const TexPair p11 = *reinterpret_cast<const TexPair*>(tex + index(x0, y1, z1));
        const bool dup = x1 == x0;
        T.t[0] = p00.a;
        T.t[1] = p01.a;
        T.t[2] = p10.a;
        T.t[3] = p11.a;
        T.t[4] = dup ? p00.a : p00.b;
        T.t[5] = dup ? p01.a : p01.b;
        T.t[6] = dup ? p10.a : p10.b;
        T.t[7] = dup ? p11.a : p11.b;
    }

    // one normalised channel: trilinear of the bytes (z innermost, x outermost,
    // lib/isaac/isaac_helper.hpp:124-144), then / 255
    template<bool LINEAR, bool ISO>
    __device__ inline float channel(const Taps& T, int sh)
    {
        auto d = [&](int k) { return float(((ISO ? T.t[k].y : T.t[k].x) >> sh) & 0xffu); };
        if(!LINEAR)
            return div255(d(0));
        const float oz = 1.0f - T.fz, oy = 1.0f - T.fy, ox = 1.0f - T.fx;
        const float a0 = d(0) * oz + d(1) * T.fz, a1 = d(2) * oz + d(3) * T.fz;
        const float A = a0 * oy + a1 * T.fy;
        const float b0 = d(4) * oz + d(5) * T.fz, b1 = d(6) * oz + d(7) * T.fz;
        const float B = b0 * oy + b1 * T.fy;
        return div255(A * ox + B * T.fx);
    }

    // bitwise OR of the taps' volume (.x) or iso (.y) texels
    template<bool LINEAR, bool ISO>
    __device__ inline uint32_t taps_or(const Taps& T)
    {
        uint32_t o = ISO ? T.t[0].y : T.t[0].x;
        if(LINEAR)
        {
#pragma unroll
            for(int k = 1; k < 8; k++)
                o |= ISO ? T.t[k].y : T.t[k].x;
        }
        return o;
    }

    template<bool LINEAR, bool BRICK>
    __device__ inline float iso_alpha_at(const FrameParams& F, float px, float py, float pz)
    {
        Taps T;
        fetch_taps<LINEAR, BRICK>(F, px, py, pz, T);
        return channel<LINEAR, true>(T, 24);
    }

    // ------------------------------------------------------------------ march sampling
    // The march's per-step samples weigh the eight texels once per position: each
    // channel is sum_k byte_k * w_k (one multiply and seven fmas) with the corner
    // weights w_k = wx * wy * wz, instead of seven lerps per channel, and at the x
    // edge (both x taps on texel x0, CLAMP) the first texel takes weight 1 instead of
    // the texels being swapped.  In exact arithmetic this is the reference's trilinear
    // (lib/isaac/isaac_helper.hpp:124-144); in binary32 the result differs from its
    // nested lerps by rounding only.  Texel indices are 32-bit; rows y+1 and z+1
    // are the row of y and z plus a stride or plus 0 at a clamped edge.
    struct Corners
    {
        uint2 t[8]; // k = 4 x + 2 y + z (x outermost, as the reference's lerps nest)
        float w[8];
    };

    // the corners of the sample at c: for LINEAR the position minus 0.5 (the texel
    // centres' offset), for NEAREST the position itself; BRICK reads the brick layout
    // OFF32: the texture spans < 4 GiB, so the loads take a 32-bit byte offset from the
    // texture's (uniform) base address instead of a 64-bit address per lane.  HALF reads
    // the binary16 copy of the linear layout (same texel indices) instead.
    template<bool LINEAR, bool BRICK, bool OFF32 = false, bool HALF = false>
    __device__ inline void fetch_corners(const FrameParams& F, uint32_t XY, float cx, float cy, float cz, Corners& C)
    {
        static_assert(!(HALF && BRICK), "the binary16 copy has the linear layout");
        const uint2* __restrict__ tex = HALF ? F.texH : (BRICK ? F.texB : F.tex);
        const int X = F.X, Y = F.Y, Z = F.Z;
        if(!LINEAR)
        {
            const uint32_t x = iclamp(int(cx), 0, X - 1), y = iclamp(int(cy), 0, Y - 1), z = iclamp(int(cz), 0, Z - 1);
            C.t[0] = tex[BRICK ? brick_texel(F, x, y, z) : x + uint32_t(X) * y + XY * z];
            return;
        }
        const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
        const int bx = int(flx), by = int(fly), bz = int(flz);
        // taps x0 and x0 + 1 come from one 16-byte load; where both are x0 (bx < 0 or
        // bx >= X - 1) the second gets weight 0
        const bool xin = uint32_t(bx) < uint32_t(X - 1);
        const float fx = xin ? cx - flx : 0.0f, fy = cy - fly, fz = cz - flz;
        const uint32_t x0 = iclamp(bx, 0, X - 1), y0 = iclamp(by, 0, Y - 1), z0 = iclamp(bz, 0, Z - 1);
        // rows y + 1 and z + 1: the next row / slice, or the same one at a clamped edge
        // (in the brick layout the next row of y & 7 = 7 is the apron row)
        const uint32_t dy = uint32_t(by) < uint32_t(Y - 1) ? (BRICK ? uint32_t(kTexBrickRow) : uint32_t(X)) : 0u;
        const uint32_t dz = uint32_t(bz) < uint32_t(Z - 1) ? (BRICK ? uint32_t(kTexBrickSlice) : XY) : 0u;
        const uint32_t i00 = BRICK ? brick_texel(F, x0, y0, z0) : x0 + __umul24(y0, uint32_t(X)) + z0 * XY;
        uint4 p00, p01, p10, p11;
        if constexpr(OFF32)
        {
            const char* base = reinterpret_cast<const char*>(tex);
            const uint32_t o00 = i00 << 3, o01 = o00 + (dz << 3), o10 = o00 + (dy << 3), o11 = o10 + (dz << 3);
            p00 = *reinterpret_cast<const uint4*>(base + o00);
            p01 = *reinterpret_cast<const uint4*>(base + o01);
            p10 = *reinterpret_cast<const uint4*>(base + o10);
            p11 = *reinterpret_cast<const uint4*>(base + o11);
        }
        else
        {
            const uint32_t i01 = i00 + dz, i10 = i00 + dy, i11 = i10 + dz;
            p00 = *reinterpret_cast<const uint4*>(tex + i00);
            p01 = *reinterpret_cast<const uint4*>(tex + i01);
            p10 = *reinterpret_cast<const uint4*>(tex + i10);
            p11 = *reinterpret_cast<const uint4*>(tex + i11);
        }
        C.t[0] = make_uint2(p00.x, p00.y);
        C.t[1] = make_uint2(p01.x, p01.y);
        C.t[2] = make_uint2(p10.x, p10.y);
        C.t[3] = make_uint2(p11.x, p11.y);
        C.t[4] = make_uint2(p00.z, p00.w);
        C.t[5] = make_uint2(p01.z, p01.w);
        C.t[6] = make_uint2(p10.z, p10.w);
        C.t[7] = make_uint2(p11.z, p11.w);
        const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
        const float w00 = oy * oz, w01 = oy * fz, w10 = fy * oz, w11 = fy * fz;
        C.w[0] = ox * w00;
        C.w[1] = ox * w01;
        C.w[2] = ox * w10;
        C.w[3] = ox * w11;
        C.w[4] = fx * w00;
        C.w[5] = fx * w01;
        C.w[6] = fx * w10;
        C.w[7] = fx * w11;
    }

    // one channel in byte units (0..255): the volume (.x) or iso (.y) texel's byte at `sh`
    template<bool LINEAR, bool ISO>
    __device__ inline float corner_sum(const Corners& C, int sh)
    {
        auto d = [&](int k) { return float(((ISO ? C.t[k].y : C.t[k].x) >> sh) & 0xffu); };
        if(!LINEAR)
            return d(0);
        float v = d(0) * C.w[0];
#pragma unroll
        for(int k = 1; k < 8; k++)
            v = __builtin_fmaf(d(k), C.w[k], v);
        return v;
    }

    // channel ch (0..3: r, g, b, a) of the volume RGBA from corners of the binary16 copy,
    // in byte units: the same sum as corner_sum<true, false>(C, 8 * ch) (each binary16 value
    // is its byte exactly; the first product goes through an fma with +0, which returns the
    // product itself as both factors are >= +0); the alpha is read through its magnitude
    template<int CH>
    __device__ inline float half_sum(const Corners& C)
    {
        auto d = [&](int k)
        {
            const uint32_t u = CH < 2 ? C.t[k].x : C.t[k].y;
            const float v = float(__ushort_as_half((unsigned short) ((CH & 1) ? (u >> 16) : (u & 0xffffu))));
            return CH == 3 ? fabsf(v) : v;
        };
        float v = __builtin_fmaf(d(0), C.w[0], 0.0f);
#pragma unroll
        for(int k = 1; k < 8; k++)
            v = __builtin_fmaf(d(k), C.w[k], v);
        return v;
    }

    template<bool LINEAR, bool ISO>
    __device__ inline uint32_t corners_or(const Corners& C)
    {
        uint32_t o = ISO ? C.t[0].y : C.t[0].x;
        if(LINEAR)
        {
#pragma unroll
            for(int k = 1; k < 8; k++)
                o |= ISO ? C.t[k].y : C.t[k].x;
        }
        return o;
    }

    struct F4
    {
        float x, y, z, w;
    };

    // ------------------------------------------------------------------ pixel tiling
    // workgroup -> 16x16 tile, XCD-banded; returns false for threads outside the viewport
    // pixel of this thread in 16x16 tile `b` of the viewport (row-major tiles)
    // pixel of this lane in 16x16 tile b, whose 8x8 quarter `wave` (0..3) its wave renders
    __device__ inline bool tile_pixel_at(const FrameParams& F, uint32_t b, uint32_t wave, int& px, int& py)
    {
        const uint32_t tx = b % uint32_t(F.tilesX), ty = b / uint32_t(F.tilesX);
        const uint32_t lane = threadIdx.x & 63u;
        px = F.vx0 + int(tx * 16u + (wave & 1u) * 8u + (lane & 7u));
        py = F.vy0 + int(ty * 16u + (wave >> 1) * 8u + (lane >> 3));
        return b < uint32_t(F.tilesX) * uint32_t(F.tilesY) && px < F.vx1 && py < F.vy1;
    }

    __device__ inline bool tile_pixel_at(const FrameParams& F, uint32_t b, int& px, int& py)
    {
        return tile_pixel_at(F, b, threadIdx.x >> 6, px, py);
    }

    // The march's workgroups: one 16x16 tile of 4 waves each, or (F.waveGroups == 1)
    // one wave each: workgroup b is then quarter (b / 8) % 4 of the tile of logical
    // workgroup (b / 32) * 8 + b % 8, and runs on the XCD b % 8 of that logical
    // workgroup, so the dealing of tiles to the XCDs is unchanged while the
    // dispatcher fills any SIMD with a free wave slot.
    __device__ inline uint32_t march_block(const FrameParams& F)
    {
        return F.waveGroups == 1 ? ((blockIdx.x >> 5) << 3) | (blockIdx.x & 7u) : blockIdx.x;
    }
    __device__ inline uint32_t march_wave(const FrameParams& F)
    {
        return F.waveGroups == 1 ? (blockIdx.x >> 3) & 3u : threadIdx.x >> 6;
    }

    __device__ inline uint32_t xcd_tile(const FrameParams& F, uint32_t b)
    {
        const uint32_t nb = uint32_t(F.tilesX) * uint32_t(F.tilesY);
        const uint32_t full = nb & ~7u;
        if(b < full)
            b = (b & 7u) * (full >> 3) + (b >> 3);
        return b;
    }

    // Super-tiles of SxS tiles (S = F.tileOrder: 2, 4, 8 or 16; 4 is 64x64 pixels)
    // dealt round-robin to the XCDs: workgroup b runs on XCD b % 8 and takes tile
    // (b / 8) % S^2 of that XCD's super-tile number (b / 8) / S^2.  Every XCD gets
    // super-tiles from all over the frame (balanced when the work is concentrated,
    // e.g. a volume in the middle of the frame) while the rays of one super-tile
    // share its L2.  Grid: march_grid workgroups; padding workgroups map to no tile.
    __device__ inline uint32_t supertile_tile(const FrameParams& F, uint32_t b)
    {
        const uint32_t S = uint32_t(F.tileOrder), sh = uint32_t(__ffs(int(S)) - 1);
        const uint32_t tX = uint32_t(F.tilesX), tY = uint32_t(F.tilesY);
        const uint32_t sx = (tX + S - 1u) >> sh, sy = (tY + S - 1u) >> sh;
        const uint32_t i = b >> 3;
        const uint32_t s = (b & 7u) + 8u * (i >> (2u * sh)), w = i & (S * S - 1u);
        if(s >= sx * sy)
            return 0xffffffffu;
        const uint32_t tx = ((s % sx) << sh) + (w & (S - 1u)), ty = ((s / sx) << sh) + (w >> sh);
        if(tx >= tX || ty >= tY)
            return 0xffffffffu;
        return ty * tX + tx;
    }

    __device__ inline uint32_t march_tile(const FrameParams& F)
    {
        const uint32_t b = march_block(F);
        return F.tileOrder != 0 ? supertile_tile(F, b) : xcd_tile(F, b);
    }

    __device__ inline bool tile_pixel(const FrameParams& F, int& px, int& py)
    {
        return tile_pixel_at(F, xcd_tile(F, blockIdx.x), px, py);
    }

    __device__ inline void block_count_add(unsigned long long* dst, unsigned long long v)
    {
        __shared__ unsigned long long red[4];
        // wave reduction through DPP-free shuffles
        for(int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
        if((threadIdx.x & 63) == 0)
            red[threadIdx.x >> 6] = v;
        __syncthreads();
        if(threadIdx.x == 0)
        {
            unsigned long long s = 0;
            for(uint32_t w = 0; w < (blockDim.x >> 6); w++)
                s += red[w];
            if(s)
                atomicAdd(dst, s);
        }
        __syncthreads();
    }

    // ------------------------------------------------------------------ particles
    // ParticleRenderKernel (lib/isaac/isaac_particle_kernel.hpp:117-260): 3D-DDA over
    // the particle cell grid from the clipped ray start; the first cell holding a
    // visible sphere hit ends the walk (hit point of the nearest sphere in it).
    template<bool COUNT>
    __device__ inline void particle_dda(
        const ParticleParams& PP,
        const Ray& ray,
        uint32_t& color,
        float& depth,
        F3& normal,
        unsigned long long& nCells,
        unsigned long long& nTests)
    {
        const F3 ps = f3(PP.pscale[0], PP.pscale[1], PP.pscale[2]);
        const int ds[3] = {int(gsign(ray.dir.x)), int(gsign(ray.dir.y)), int(gsign(ray.dir.z))};
        const F3 cur = add(ray.start, muls(ray.dir, ray.startDepth));
        const F3 cdiv = divv(cur, ps);
        uint32_t cell[3];
        cell[0] = uint32_t(iclamp(int(cdiv.x), 0, PP.grid[0] - 1));
        cell[1] = uint32_t(iclamp(int(cdiv.y), 0, PP.grid[1] - 1));
        cell[2] = uint32_t(iclamp(int(cdiv.z), 0, PP.grid[2] - 1));
        const float rayLength = ray.endDepth - ray.startDepth;
        float tested = 0.0f;
        const float curA[3] = {cur.x, cur.y, cur.z};
        const float dA[3] = {ray.dir.x, ray.dir.y, ray.dir.z};
        const float psA[3] = {ps.x, ps.y, ps.z};
        float t[3], dt[3];
#pragma unroll
        for(int i = 0; i < 3; i++)
        {
            t[i] = ((float(cell[i]) + float(ds[i] > 0 ? ds[i] : 0)) * psA[i] - curA[i]) / dA[i];
            dt[i] = psA[i] / dA[i] * float(ds[i]);
            if(dA[i] == 0.0f)
                t[i] = FLT_MAX;
        }
        bool hit = false;
        float pdepth = FLT_MAX;
        float4 pcol = make_float4(0.f, 0.f, 0.f, 0.f);
        F3 pn = f3(0.f, 0.f, 0.f);
        const uint32_t gx = uint32_t(PP.grid[0]), gy = uint32_t(PP.grid[1]), gz = uint32_t(PP.grid[2]);
        auto inside = [&](const uint32_t (&c)[3]) { return c[0] < gx && c[1] < gy && c[2] < gz; };
        auto cell_id = [&](const uint32_t (&c)[3]) { return int(c[0] + gx * (c[1] + gy * c[2])); };
        // record ranges of the current cell, requested one cell ahead: the walk's next
        // cell depends on the ray only, so the range of cell i + 1 is in flight while
        // the spheres of cell i are tested
        int2 kr[4] = {};
        if(inside(cell))
        {
            const int cid = cell_id(cell);
#pragma unroll
            for(int s = 0; s < 4; s++)
                if(s < PP.nsrc)
                    kr[s] = PP.src[s].vrange[cid];
        }
        while(inside(cell) && !hit && tested <= rayLength)
        {
            if(COUNT)
                nCells++;
            // the walk's next cell (ties between axes go towards z)
            uint32_t next[3] = {cell[0], cell[1], cell[2]};
            float nextTested;
            const bool ax0 = t[0] < t[1] && t[0] < t[2];
            const bool ax1 = !ax0 && t[1] < t[0] && t[1] < t[2];
            if(ax0)
            {
                next[0] += uint32_t(ds[0]);
                nextTested = t[0];
            }
            else if(ax1)
            {
                next[1] += uint32_t(ds[1]);
                nextTested = t[1];
            }
            else
            {
                next[2] += uint32_t(ds[2]);
                nextTested = t[2];
            }
            int2 krn[4] = {};
            if(inside(next))
            {
                const int nid = cell_id(next);
#pragma unroll
                for(int s = 0; s < 4; s++)
                    if(s < PP.nsrc)
                        krn[s] = PP.src[s].vrange[nid];
            }
            const F3 cf = f3(float(cell[0]), float(cell[1]), float(cell[2]));
#pragma unroll
            for(int s = 0; s < 4; s++)
            {
                if(s >= PP.nsrc)
                    break;
                const ParticleSourceDev& S = PP.src[s];
                for(int k = kr[s].x; k < kr[s].y; k++)
                {
                    if(COUNT)
                        nTests++;
                    const float4 rec = S.recs[k];
                    const F3 ppos = mul(add(f3(rec.x, rec.y, rec.z), cf), ps);
                    const F3 L = sub(ppos, cur);
                    const float radius = (S.radius ? S.radius[k] : S.constRadius) * S.weight;
                    const float r2 = radius * radius;
                    const float tca = dot3(L, ray.dir);
                    const float d2 = dot3(L, L) - tca * tca;
                    if(d2 <= r2)
                    {
                        const float thc = sqrtf(r2 - d2);
                        const float t0 = tca - thc;
                        const float t1 = tca + thc;
                        if(t1 >= 0.0f && t0 < pdepth)
                        {
                            const float4 val = S.lut[__float_as_int(rec.w)];
                            if(val.w >= 0.5f)
                            {
                                pcol = val;
                                pdepth = t0;
                                hit = true;
                                pn = sub(add(cur, muls(ray.dir, t0)), ppos);
                                if(t0 < 0.0f && ray.isClipped)
                                    pn = f3(-ray.clipN.x, -ray.clipN.y, -ray.clipN.z);
                            }
                        }
                    }
                }
            }
            cell[0] = next[0];
            cell[1] = next[1];
            cell[2] = next[2];
            tested = nextTested;
            if(ax0)
                t[0] = t[0] + dt[0];
            else if(ax1)
                t[1] = t[1] + dt[1];
            else
                t[2] = t[2] + dt[2];
#pragma unroll
            for(int s = 0; s < 4; s++)
                kr[s] = krn[s];
        }
        if(hit)
        {
            normal = normalize3(pn);
            color = to_byte4(pcol.x, pcol.y, pcol.z, 1.0f);
            depth = pdepth + ray.startDepth;
        }
    }

    // ------------------------------------------------------------------ fused march
    // One pass per pixel: clear, particle DDA, then ONE traversal of the ray that
    // serves both the iso march (CombinedIsoRenderKernel) and the volume march
    // (CombinedVolumeRenderKernel).  Both reference kernels sample the same
    // positions startU + stepVec*i from the same startSteps; the volume stops at
    // floor(min(clipEnd, depth)/step) where depth is the particle/iso depth, so
    // the iso decision at step i is made before the volume consumes step i.  The
    // premultiplied volume colour goes to volColor and is blended over the shaded
    // solid colour by shade_kernel after SSAO, which keeps the reference's pass
    // order (particles, iso, SSAO, shading, volume) for every pixel.
    // The taps of step i+1 are fetched before step i is composited (clamped
    // addresses, so a fetch past the last step stays in bounds).
    template<bool LINEAR, bool COUNT, bool PREFETCH, bool BRICK, bool OFF32 = false, bool HALF = false>
    __device__ __attribute__((always_inline)) inline void march_body(const FrameParams& F, const ParticleParams& PP)
    {
        int px, py;
        const bool valid = tile_pixel_at(F, march_tile(F), march_wave(F), px, py);
        unsigned long long nCells = 0, nTests = 0, nIso = 0, nIsoHits = 0, nIsoRays = 0, nPartRays = 0;
        unsigned long long nVol = 0, nVolRays = 0, nSteps = 0;
        if(valid)
        {
            const size_t p = size_t(py) * size_t(F.W) + size_t(px);
            uint32_t color = F.bgClear;
            F3 normal = f3(0.f, 0.f, 0.f);
            float depth = FLT_MAX;
            float4 vcol = make_float4(0.f, 0.f, 0.f, -1.f);
            Ray ray = pixel_to_ray(F, px, py);
            const bool inside = clip_ray(F, ray);
            // ---- particles
            if(inside && PP.nsrc > 0)
            {
                nPartRays++;
                particle_dda<COUNT>(PP, ray, color, depth, normal, nCells, nTests);
            }
            // ---- iso + volume, one traversal
            const int mode = F.renderMode;
            // ShadingKernel zeroes the depth in these modes, so the volume pass never blends there
            const bool depthZeroed = mode == 2 || (mode >= 4 && mode <= 8);
            if(inside)
            {
                const float endAfterParticles = gmin(ray.endDepth, depth);
                bool isoOn = F.isoEnabled && endAfterParticles > ray.startDepth;
                bool volBlend = F.volumeEnabled && !depthZeroed && endAfterParticles > ray.startDepth;
                if(isoOn || volBlend)
                {
                    const F3 sc = f3(F.scale[0], F.scale[1], F.scale[2]);
                    const float stepU = F.step * (len3(ray.dir) / len3(divv(ray.dir, sc)));
                    const float factor = stepU / F.minSize * 2.0f * F.totalWeightAll;
                    const int startSteps = int(ceilf(ray.startDepth / stepU));
                    const int endIso = int(floorf(endAfterParticles / stepU));
                    int endVol = endIso;
                    const F3 stepVec = divv(muls(ray.dir, stepU), sc);
                    const F3 startU = divv(ray.start, sc);
                    if(isoOn)
                        nIsoRays++;
                    float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
                    // The traversal state is kept in plain numbers rather than flags, so
                    // the loop carries no lane masks: the iso test runs while i <= isoEnd and
                    // the volume while i <= volEnd (either set below the first step to switch
                    // it off); hitStep is the step of the iso hit (-1: none); oldValue is the
                    // previous step's iso sum, -1 where it was not formed; blendW is the
                    // volume colour's w marker (0: the volume pass blends, -1: it does not).
                    const int off = -0x7fffffff;
                    int isoEnd = isoOn ? endIso : off;
                    int volEnd = volBlend ? endVol : off;
                    int hitStep = -1;
                    float oldValue = 0.0f;
                    float blendW = volBlend ? 0.0f : -1.0f;
                    int i = startSteps;
                    int last = isoOn ? endIso : endVol;
                    if(i <= last)
                    {
                        // volume samples in byte units times factor / 255
                        const float vscale = factor * (1.0f / 255.0f);
                        const uint32_t XY = uint32_t(F.X) * uint32_t(F.Y);
                        // sample i at startC + stepVec * i (one fma per axis), shifted by the
                        // texel centres' -0.5 for the trilinear sampler
                        const float sh = LINEAR ? 0.5f : 0.0f;
                        const F3 startC = f3(startU.x - sh, startU.y - sh, startU.z - sh);
                        // the step's corners (from the binary16 copy for HALF) and, for the iso
                        // sums of HALF, the combined texture's corners at the same position
                        auto at = [&](int k, Corners& C)
                        {
                            const float fk = float(k);
                            fetch_corners<LINEAR, BRICK, OFF32, HALF>(F, XY, __builtin_fmaf(stepVec.x, fk, startC.x),
                                                                      __builtin_fmaf(stepVec.y, fk, startC.y),
                                                                      __builtin_fmaf(stepVec.z, fk, startC.z), C);
                        };
                        auto atIso = [&](int k, Corners& C)
                        {
                            const float fk = float(k);
                            fetch_corners<LINEAR, BRICK, OFF32, false>(F, XY, __builtin_fmaf(stepVec.x, fk, startC.x),
                                                                       __builtin_fmaf(stepVec.y, fk, startC.y),
                                                                       __builtin_fmaf(stepVec.z, fk, startC.z), C);
                        };
                        // one step: the iso test and the volume sample of the corners c of step i
                        auto process = [&](const Corners& cur) __attribute__((always_inline))
                        {
                            if(COUNT)
                                nSteps++;
                            if(i <= isoEnd)
                            {
                                // the iso alpha in byte units; alpha >= 0.5 is a sum >= 127.5, which
                                // needs an alpha byte >= 128 (the corner weights sum to 1 up to a few
                                // ulps, so eight bytes <= 127 stay below 127.1).  Without such a byte
                                // the sum is not formed: it is exactly 0 when all alpha bytes are 0,
                                // else it is formed again from the step's corners if the next step hits
                                // (HALF: the flag is the sign of the binary16 alpha, and the sum
                                // comes from the combined texture's corners of this step)
                                const uint32_t alphas = corners_or<LINEAR, true>(cur);
                                const bool high = (alphas >> 31) != 0u;
                                float sum = 0.0f;
                                if(HALF)
                                {
                                    if(high)
                                    {
                                        Corners ic;
                                        atIso(i, ic);
                                        sum = corner_sum<LINEAR, true>(ic, 24);
                                    }
                                }
                                else if(high)
                                    sum = corner_sum<LINEAR, true>(cur, 24);
                                if(COUNT)
                                    nIso++;
                                float tmpSum = oldValue;
                                oldValue = (high || (!HALF && (alphas >> 24) == 0u)) ? sum : -1.0f;
                                if(sum >= 127.5f)
                                {
                                    if(tmpSum < 0.0f)
                                    {
                                        Corners pc;
                                        atIso(i - 1, pc);
                                        tmpSum = corner_sum<LINEAR, true>(pc, 24);
                                    }
                                    const bool first = ray.isClipped && i == startSteps;
                                    const float tt = float(i) * stepU;
                                    const float value = div255(sum), tmp = div255(tmpSum);
                                    float d;
                                    if(first)
                                        d = ray.startDepth;
                                    else
                                        d = tt + stepU * (0.5f - tmp) / (value - tmp);
                                    // colour and gradient at the hit are sampled after the loop
                                    hitStep = i;
                                    depth = d;
                                    if(COUNT)
                                        nIsoHits++;
                                    isoEnd = off;
                                    // the volume pass ends at min(clip end, iso depth)
                                    const float ve = gmin(ray.endDepth, d);
                                    if(!(ve > ray.startDepth))
                                    {
                                        volEnd = off;
                                        blendW = -1.0f;
                                    }
                                    else if(volEnd != off)
                                        volEnd = int(floorf(ve / stepU));
                                }
                            }
                            // (eight zero volume texels give the sample (0,0,0,0), which leaves
                            // the accumulated colour unchanged, so no test for them: the branch
                            // cost more than the arithmetic it skipped)
                            if(i <= volEnd)
                            {
                                if(COUNT)
                                    nVol++;
                                const float vr = HALF ? half_sum<0>(cur) : corner_sum<LINEAR, false>(cur, 0);
                                const float vg = HALF ? half_sum<1>(cur) : corner_sum<LINEAR, false>(cur, 8);
                                const float vb = HALF ? half_sum<2>(cur) : corner_sum<LINEAR, false>(cur, 16);
                                const float va = HALF ? half_sum<3>(cur) : corner_sum<LINEAR, false>(cur, 24);
                                const float w = (1.0f - ca) * vscale;
                                cr = cr + w * vr;
                                cg = cg + w * vg;
                                cb = cb + w * vb;
                                ca = ca + w * va;
                                if(ca > 0.99f)
                                    volEnd = off;
                            }
                        };
                        Corners cur;
                        at(i, cur);
                        if constexpr(PREFETCH)
                        {
                            // two buffers in turn: the next step's corners are requested before
                            // this step's are composited, with no copy between the buffers (a hit
                            // can extend the volume pass by one step past the range requested)
                            Corners nxt;
                            for(;;)
                            {
                                bool ahead = i + 1 <= last;
                                if(ahead)
                                    at(i + 1, nxt);
                                process(cur);
                                last = isoEnd > volEnd ? isoEnd : volEnd;
                                if(i + 1 > last)
                                    break;
                                i++;
                                if(!ahead)
                                    at(i, nxt);
                                ahead = i + 1 <= last;
                                if(ahead)
                                    at(i + 1, cur);
                                process(nxt);
                                last = isoEnd > volEnd ? isoEnd : volEnd;
                                if(i + 1 > last)
                                    break;
                                i++;
                                if(!ahead)
                                    at(i, cur);
                            }
                        }
                        else
                        {
                            for(;;)
                            {
                                process(cur);
                                last = isoEnd > volEnd ? isoEnd : volEnd;
                                if(i + 1 > last)
                                    break;
                                i++;
                                at(i, cur);
                            }
                        }
                    }
                    const bool isoHit = hitStep >= 0;
                    const bool isoFirst = isoHit && ray.isClipped && hitStep == startSteps;
                    if(isoHit)
                    {
                        // hit colour and central-difference gradient of the iso alpha
                        // (lib/isaac/isaac_iso_kernel.hpp:97-119)
                        const F3 pu = divv(add(ray.start, muls(ray.dir, depth)), sc);
                        Taps h;
                        fetch_taps<LINEAR, BRICK>(F, pu.x, pu.y, pu.z, h);
                        const float hr = channel<LINEAR, true>(h, 0), hg = channel<LINEAR, true>(h, 8);
                        const float hb = channel<LINEAR, true>(h, 16);
                        F3 g;
                        g.x = iso_alpha_at<LINEAR, BRICK>(F, pu.x + 1.0f, pu.y, pu.z) - iso_alpha_at<LINEAR, BRICK>(F, pu.x - 1.0f, pu.y, pu.z);
                        g.y = iso_alpha_at<LINEAR, BRICK>(F, pu.x, pu.y + 1.0f, pu.z) - iso_alpha_at<LINEAR, BRICK>(F, pu.x, pu.y - 1.0f, pu.z);
                        g.z = iso_alpha_at<LINEAR, BRICK>(F, pu.x, pu.y, pu.z + 1.0f) - iso_alpha_at<LINEAR, BRICK>(F, pu.x, pu.y, pu.z - 1.0f);
                        float gl = len3(g);
                        if(isoFirst || gl == 0.0f)
                        {
                            g = ray.clipN;
                            gl = 1.0f;
                        }
                        normal = f3(-g.x / gl, -g.y / gl, -g.z / gl);
                        color = to_byte4(hr, hg, hb, 1.0f);
                    }
                    if(blendW >= 0.0f)
                    {
                        nVolRays++;
                        vcol = make_float4(cr, cg, cb, ca);
                    }
                }
            }
            F.color[p] = color;
            F.depth[p] = depth;
            F.normal[3 * p + 0] = normal.x;
            F.normal[3 * p + 1] = normal.y;
            F.normal[3 * p + 2] = normal.z;
            F.volColor[p] = vcol;
            if(!F.aoEnabled)
                F.ao[p] = 0.0f;
        }
        if(COUNT)
        {
            block_count_add(F.counters + 0, nVol);
            block_count_add(F.counters + 1, nIso);
            block_count_add(F.counters + 2, nIsoHits);
            block_count_add(F.counters + 3, nCells);
            block_count_add(F.counters + 4, nTests);
            block_count_add(F.counters + 5, nVolRays);
            block_count_add(F.counters + 6, nIsoRays);
            block_count_add(F.counters + 7, nPartRays);
            block_count_add(F.counters + 8, nSteps);
            // the wave runs the traversal loop as long as its longest lane
            int wm = int(nSteps);
            for(int off = 32; off > 0; off >>= 1)
                wm = max(wm, __shfl_xor(wm, off, 64));
            block_count_add(F.counters + 9, (threadIdx.x & 63u) == 0 ? (unsigned long long) wm : 0ull);
        }
    }

    // The march at the compiler's register allocation, and at occupancy targets of
    // 6 and 8 waves per SIMD (at most 80 / 64 VGPRs); the launcher picks one by the
    // tuning key ISAAC_HIP_TUNE_MARCH_WAVES.
    template<bool LINEAR, bool COUNT, bool PREFETCH, bool BRICK, bool OFF32 = false, bool HALF = false>
    __global__ __launch_bounds__(256) void march_kernel(const FrameParams F, const ParticleParams PP)
    {
        march_body<LINEAR, COUNT, PREFETCH, BRICK, OFF32, HALF>(F, PP);
    }
    template<bool LINEAR, bool BRICK, bool OFF32, bool HALF = false>
    __global__ __launch_bounds__(256, 6) void march_kernel_w6(const FrameParams F, const ParticleParams PP)
    {
        march_body<LINEAR, false, false, BRICK, OFF32, HALF>(F, PP);
    }
    template<bool LINEAR, bool BRICK, bool OFF32, bool HALF = false>
    __global__ __launch_bounds__(256, 8) void march_kernel_w8(const FrameParams F, const ParticleParams PP)
    {
        march_body<LINEAR, false, false, BRICK, OFF32, HALF>(F, PP);
    }

    // ------------------------------------------------------------------ SSAO
    __device__ __attribute__((always_inline)) inline void ssao_body(const FrameParams& F, uint32_t tile)
    {
        int px, py;
        if(!tile_pixel_at(F, tile, px, py))
            return;
        const float ref = F.depth[size_t(py) * F.W + px];
        float occ = 0.0f;
#pragma unroll
        for(int i = -3; i <= 3; ++i)
        {
            const int sx = iclamp(px + i * 10, 0, F.W - 1);
#pragma unroll
            for(int j = -3; j <= 3; ++j)
            {
                const int sy = iclamp(py + j * 10, 0, F.H - 1);
                if(F.depth[size_t(sy) * F.W + sx] < ref)
                    occ = occ + 1.0f;
            }
        }
        F.ao[size_t(py) * F.W + px] = gclamp(occ / 42.0f, 0.0f, 1.0f);
    }

    // ssao_body's count with the depth taps from LDS: a workgroup takes a 64x32 block
    // of the viewport and stages the 124x92 depth region around it (the 30-pixel reach
    // of the 7x7 taps at a stride of 10), each region texel read at its framebuffer-
    // clamped position -- so region[px + 10 i - x0 + 30] is the clamped tap itself and
    // the 49 taps are LDS reads at fixed offsets.  The count is exact (at most 49
    // increments of 1), so the occlusion equals ssao_body's.
    constexpr int kSsaoBW = 64, kSsaoBH = 32, kSsaoReach = 30;
    constexpr int kSsaoRW = kSsaoBW + 2 * kSsaoReach, kSsaoRH = kSsaoBH + 2 * kSsaoReach; // 124 x 92
    // SHADE: the block also shades its pixels with the occlusion term it just formed (render
    // modes that leave the depth buffer alone, so no block changes a depth another one reads)
    // the G-buffer values the shading of one pixel reads (besides occlusion and depth)
    struct ShadeIn
    {
        uint32_t cb;
        F3 normal;
        float4 v;
    };

    __device__ __attribute__((always_inline)) inline ShadeIn shade_load(const FrameParams& F, size_t p);
    __device__ __attribute__((always_inline)) inline void shade_finish(const FrameParams& F, int px, int py, const ShadeIn& in,
                                                                      float ao, float depth);

    template<bool SHADE>
    __global__ __launch_bounds__(256) void ssao_tiled_kernel(const FrameParams F)
    {
        __shared__ float region[kSsaoRW * kSsaoRH];
        const int nbX = (F.vx1 - F.vx0 + kSsaoBW - 1) / kSsaoBW;
        const int x0 = F.vx0 + int(blockIdx.x % uint32_t(nbX)) * kSsaoBW;
        const int y0 = F.vy0 + int(blockIdx.x / uint32_t(nbX)) * kSsaoBH;
        // staging: thread t loads column t % 124 of rows t / 124 + 2 k (248 threads)
        const int col = int(threadIdx.x) % kSsaoRW, r0 = int(threadIdx.x) / kSsaoRW;
        if(r0 < 2)
        {
            const int gx = iclamp(x0 - kSsaoReach + col, 0, F.W - 1);
            const float* src = F.depth + gx;
#pragma unroll 4
            for(int r = r0; r < kSsaoRH; r += 2)
            {
                const int gy = iclamp(y0 - kSsaoReach + r, 0, F.H - 1);
                region[r * kSsaoRW + col] = src[size_t(gy) * F.W];
            }
        }
        __syncthreads();
#pragma unroll
        for(int k = 0; k < kSsaoBW * kSsaoBH / 256; k++)
        {
            const int q = int(threadIdx.x) + 256 * k;
            const int px = x0 + (q & (kSsaoBW - 1)), py = y0 + q / kSsaoBW;
            if(px >= F.vx1 || py >= F.vy1)
                continue;
            // (SHADE: this pixel's G-buffer values are requested before its taps are counted)
            ShadeIn in;
            if(SHADE)
                in = shade_load(F, size_t(py) * size_t(F.W) + size_t(px));
            // the tap (-3, -3) of this pixel; every tap is a fixed offset from it
            const float* t0 = region + (py - y0) * kSsaoRW + (px - x0);
            const float ref = t0[kSsaoReach * kSsaoRW + kSsaoReach];
            float occ = 0.0f;
#pragma unroll
            for(int i = 0; i < 7; ++i)
#pragma unroll
                for(int j = 0; j < 7; ++j)
                    if(t0[(j * 10) * kSsaoRW + i * 10] < ref)
                        occ = occ + 1.0f;
            const float ao = gclamp(occ / 42.0f, 0.0f, 1.0f);
            F.ao[size_t(py) * F.W + px] = ao;
            if(SHADE)
                shade_finish(F, px, py, in, ao, ref);
        }
    }

    // ------------------------------------------------------------------ shading + blend
    // ShadingKernel (lib/isaac/isaac_common_kernel.hpp:220-320), then the volume
    // pass's final blend color + (1 - color.a) * solid
    // (lib/isaac/isaac_volume_kernel.hpp:83-86) for pixels whose volume pass ran.
    __device__ __attribute__((always_inline)) inline ShadeIn shade_load(const FrameParams& F, size_t p)
    {
        ShadeIn in;
        in.cb = F.color[p];
        in.normal = f3(F.normal[3 * p + 0], F.normal[3 * p + 1], F.normal[3 * p + 2]);
        in.v = F.volColor[p];
        return in;
    }

    // one pixel, given its G-buffer values, occlusion term and depth
    __device__ __attribute__((always_inline)) inline void shade_finish(const FrameParams& F, int px, int py, const ShadeIn& in,
                                                                      float ao, float depth)
    {
        const size_t p = size_t(py) * size_t(F.W) + size_t(px);
        const uint32_t cb = in.cb;
        F4 color = F4{div255(float(cb & 0xffu)), div255(float((cb >> 8) & 0xffu)), div255(float((cb >> 16) & 0xffu)),
                      div255(float(cb >> 24))};
        const F3 normal = in.normal;
        const float aoStrength = 1.0f - ao;
        const int mode = F.renderMode;
        uint32_t solid = cb;
        bool depthChanged = false;
        if(mode == 7)
            color = F4{F.rankHue[0], F.rankHue[1], F.rankHue[2], color.w};
        if(mode < 3 || mode == 7)
        {
            const Ray ray = pixel_to_ray(F, px, py);
            const F3 lightDir = f3(-ray.dir.x, -ray.dir.y, -ray.dir.z);
            float lightFactor = fabsf(dot3(normal, lightDir));
            const F3 half = normalize3(add(lightDir, lightDir));
            float spec = dot3(normal, half);
            const float s2 = spec * spec;
            spec = s2 * s2;
            spec = spec * 0.5f;
            if(mode == 1)
                spec = 0.0f;
            const float w = F.aoWeight;
            const float aoF = (1.0f - w) + w * aoStrength;
            lightFactor = lightFactor + aoF;
            lightFactor = lightFactor * 0.5f;
            solid = to_byte4(gmin(color.x * lightFactor + spec, 1.0f), gmin(color.y * lightFactor + spec, 1.0f),
                             gmin(color.z * lightFactor + spec, 1.0f), color.w);
            if(mode == 2 || mode == 7)
            {
                depth = 0.0f;
                depthChanged = true;
            }
        }
        else if(mode == 3)
            solid = to_byte4(F.bg[0], F.bg[1], F.bg[2], color.w);
        else if(mode == 4)
        {
            solid = to_byte4(normal.x * 0.5f + 0.5f, normal.y * 0.5f + 0.5f, normal.z * 0.5f + 0.5f, color.w);
            depth = 0.0f;
            depthChanged = true;
        }
        else if(mode == 5)
        {
            const float d = depth / F.maxGlobalSizeScaled;
            solid = to_byte4(d, d, d, color.w);
            depth = 0.0f;
            depthChanged = true;
        }
        else if(mode == 6)
        {
            const float w = F.aoWeight;
            const float aoF = (1.0f - w) + w * aoStrength;
            solid = to_byte4(aoF, aoF, aoF, color.w);
            depth = 0.0f;
            depthChanged = true;
        }
        else if(mode == 8)
        {
            solid = to_byte4(F.rankHue[0], F.rankHue[1], F.rankHue[2], 1.0f);
            depth = 0.0f;
            depthChanged = true;
        }
        uint32_t out = solid;
        const float4 v = in.v;
        const float sr = div255(float(solid & 0xffu)), sg = div255(float((solid >> 8) & 0xffu));
        const float sb = div255(float((solid >> 16) & 0xffu)), sa = div255(float(solid >> 24));
        float4 fo = make_float4(sr, sg, sb, sa);
        if(v.w >= 0.0f)
        {
            const float om = 1.0f - v.w;
            fo = make_float4(v.x + om * sr, v.y + om * sg, v.z + om * sb, v.w + om * sa);
            out = to_byte4(fo.x, fo.y, fo.z, fo.w);
        }
        if(F.outFloat)
            F.outFloat[p] = make_float4(gclamp(fo.x, 0.f, 1.f), gclamp(fo.y, 0.f, 1.f), gclamp(fo.z, 0.f, 1.f), gclamp(fo.w, 0.f, 1.f));
        F.color[p] = out;
        if(depthChanged)
            F.depth[p] = depth;
    }

    __device__ __attribute__((always_inline)) inline void shade_body(const FrameParams& F, uint32_t tile)
    {
        int px, py;
        if(!tile_pixel_at(F, tile, px, py))
            return;
        const size_t p = size_t(py) * size_t(F.W) + size_t(px);
        shade_finish(F, px, py, shade_load(F, p), F.aoEnabled ? F.ao[p] : 0.0f, F.depth[p]);
    }

    __global__ __launch_bounds__(256) void shade_kernel(const FrameParams F)
    {
        shade_body(F, xcd_tile(F, blockIdx.x));
    }

    // ------------------------------------------------------------------ precision form
    // Field bricks (FP32 / FP16 / BF16 + FP32 scale) rendered with
    // post-classification: interpolate the raw field, then functor chain and
    // transfer function per sample (the semantics of the reference's legacy
    // renderer, lib/isaac/isaac_volume_kernel.hpp:101-168), blended across
    // sources as the classification pass does (lib/isaac/isaac_common_kernel.hpp:574-612)
    // but without the RGBA8 quantisation.  The ray setup (pixelToRay + clipRay)
    // runs in binary64; interpolation and compositing accumulate in binary32.

    struct RayD
    {
        double dir[3], start[3];
        double startDepth, endDepth;
        bool isClipped;
        double clipN[3];
    };

    __device__ inline bool ray_setup_d(const FrameParams& F, const PostParams& Q, int px, int py, RayD& r)
    {
        const double vx = double(px) / double(F.fbW) * 2.0 - 1.0;
        const double vy = double(py) / double(F.fbH) * 2.0 - 1.0;
        double s[4], e[4];
        for(int k = 0; k < 4; k++)
        {
            const double s0 = Q.invd[0 * 4 + k] * vx + Q.invd[1 * 4 + k] * vy;
            s[k] = s0 + (Q.invd[3 * 4 + k] - Q.invd[2 * 4 + k]);
            e[k] = s0 + (Q.invd[3 * 4 + k] + Q.invd[2 * 4 + k]);
        }
        double d[3], len2 = 0.0;
        for(int k = 0; k < 3; k++)
        {
            r.start[k] = s[k] / s[3] * Q.maxSized;
            d[k] = e[k] / e[3] * Q.maxSized - r.start[k];
            len2 += d[k] * d[k];
        }
        const double il = 1.0 / sqrt(len2);
        double bmin[3], bmax[3];
        for(int k = 0; k < 3; k++)
        {
            r.dir[k] = d[k] * il;
            r.start[k] += double(F.off[k]);
            bmin[k] = (double(F.bbMin[k]) - r.start[k]) / r.dir[k];
            bmax[k] = (double(F.bbMax[k]) - r.start[k]) / r.dir[k];
            if(bmax[k] < bmin[k])
            {
                const double t = bmax[k];
                bmax[k] = bmin[k];
                bmin[k] = t;
            }
            r.clipN[k] = 0.0;
        }
        r.isClipped = false;
        r.startDepth = fmax(bmin[0], fmax(bmin[1], bmin[2]));
        r.endDepth = fmin(bmax[0], fmin(bmax[1], bmax[2]));
        for(int k = 0; k < 3; k++)
        {
            const double sgn = (r.dir[k] > 0.0) ? 1.0 : ((r.dir[k] < 0.0) ? -1.0 : 0.0);
            if(bmin[k] == r.startDepth && ((F.outerLow[k] && sgn + 1.0 != 0.0) || (F.outerHigh[k] && sgn - 1.0 != 0.0)))
            {
                r.isClipped = true;
                r.clipN[k] = sgn;
            }
        }
        for(int i = 0; i < F.nclip; i++)
        {
            double dn = 0.0, pn = 0.0, sn = 0.0;
            for(int k = 0; k < 3; k++)
            {
                dn += r.dir[k] * double(F.cnorm[i][k]);
                pn += Q.cposd[i][k] * double(F.cnorm[i][k]);
                sn += r.start[k] * double(F.cnorm[i][k]);
            }
            const double t = (pn - sn) / dn;
            if(dn > 0.0)
            {
                if(r.endDepth < t)
                    return false;
                if(r.startDepth <= t)
                {
                    for(int k = 0; k < 3; k++)
                        r.clipN[k] = double(F.cnorm[i][k]);
                    r.isClipped = true;
                    r.startDepth = t;
                }
            }
            else
            {
                if(r.startDepth > t)
                    return false;
                if(r.endDepth > t)
                    r.endDepth = t;
            }
        }
        r.startDepth = fmax(r.startDepth, double(kZNear));
        return !(r.startDepth > r.endDepth || isinf(r.startDepth) || isinf(r.endDepth));
    }

    __device__ inline float half_lo(uint32_t v)
    {
        return __half2float(__ushort_as_half((unsigned short) (v & 0xffffu)));
    }
    __device__ inline float half_hi(uint32_t v)
    {
        return __half2float(__ushort_as_half((unsigned short) (v >> 16)));
    }

    // Tap geometry of one sample position, shared by every source (all sources
    // of a rank have the local box's brick grid).  Sampler LINEAR|NEAREST with
    // CLAMP addressing (lib/isaac/isaac_texture.hpp:182-259): the taps of a sample
    // all come from the brick holding its clamped (x0, y0, z0) tap; where x1 = x0
    // (clamped at the edge) the second element of each pair gets weight 0.
    struct BrickTaps
    {
        uint32_t brick; // brick index
        uint32_t i00, i10, i01, i11; // pair index in a component block of rows (y0,z0) (y1,z0) (y0,z1) (y1,z1)
        // trilinear weights of the corners (x, y, z) = 000 001 010 011 100 101 110 111,
        // shared by every component of every source at this position; where both x
        // taps are texel x0 (CLAMP at the x edge) the x1 corners weigh 0
        float w[8];
    };

    template<bool LINEAR>
    __device__ inline BrickTaps brick_taps(const PostParams& Q, int X, int Y, int Z, float px, float py, float pz)
    {
        BrickTaps t;
        if(!LINEAR)
        {
            const uint32_t x = iclamp(int(px), 0, X - 1), y = iclamp(int(py), 0, Y - 1), z = iclamp(int(pz), 0, Z - 1);
            t.brick = ((z >> 3) * uint32_t(Q.nby) + (y >> 3)) * uint32_t(Q.nbx) + (x >> 3);
            t.i00 = ((z & 7) * 9 + (y & 7)) * 8 + (x & 7);
            t.i10 = t.i01 = t.i11 = t.i00;
            return t;
        }
        const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
        const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
        const int bx = int(flx), by = int(fly), bz = int(flz);
        const uint32_t x0 = iclamp(bx, 0, X - 1), x1 = iclamp(bx + 1, 0, X - 1);
        const uint32_t y0 = iclamp(by, 0, Y - 1), y1 = iclamp(by + 1, 0, Y - 1);
        const uint32_t z0 = iclamp(bz, 0, Z - 1), z1 = iclamp(bz + 1, 0, Z - 1);
        const uint32_t Bx = x0 >> 3, By = y0 >> 3, Bz = z0 >> 3;
        const uint32_t lx = x0 - (Bx << 3);
        const uint32_t ly0 = y0 - (By << 3), ly1 = y1 - (By << 3);
        const uint32_t lz0 = z0 - (Bz << 3), lz1 = z1 - (Bz << 3);
        t.brick = (Bz * uint32_t(Q.nby) + By) * uint32_t(Q.nbx) + Bx;
        t.i00 = lx + (lz0 * 9 + ly0) * 8;
        t.i10 = lx + (lz0 * 9 + ly1) * 8;
        t.i01 = lx + (lz1 * 9 + ly0) * 8;
        t.i11 = lx + (lz1 * 9 + ly1) * 8;
        const float fx = x1 == x0 ? 0.0f : cx - flx, fy = cy - fly, fz = cz - flz;
        const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
        const float w00 = oy * oz, w01 = oy * fz, w10 = fy * oz, w11 = fy * fz;
        t.w[0] = ox * w00;
        t.w[1] = ox * w01;
        t.w[2] = ox * w10;
        t.w[3] = ox * w11;
        t.w[4] = fx * w00;
        t.w[5] = fx * w01;
        t.w[6] = fx * w10;
        t.w[7] = fx * w11;
        return t;
    }

    // the DIM pairs of one pair row (all components of a source side by side):
    // lo[c] = v_c[x], hi[c] = v_c[x+1], still divided by the brick scale.  One
    // struct load per row, which the compiler emits as one wide global load.
    template<int N>
    struct Words
    {
        uint32_t w[N];
    };

    // 32-bit words of one pair row: 2 per component in FP32, 1 (two 16-bit halves) otherwise
    template<int DT, int DIM>
    constexpr int row_words()
    {
        return DT == kDtF32 ? 2 * DIM : DIM;
    }

    // The raw words of one source's sample: its 1 (NEAREST) or 4 (LINEAR) pair
    // rows and its brick scales, sized for 4 components.  Fetching is split
    // from interpolation so that the rows of every source of a sample are
    // requested before the first of them is used.
    template<int DT>
    struct SlotRaw
    {
        uint32_t r[4][row_words<DT, 4>()];
        uint32_t sc[4];
    };

    template<int DT, bool LINEAR, int DIM>
    __device__ inline void slot_fetch_dim(const BrickSource& S, const BrickTaps& t, SlotRaw<DT>& R)
    {
        constexpr int W = row_words<DT, DIM>();
        const Words<W>* rows = reinterpret_cast<const Words<W>*>(S.data) + t.brick * uint32_t(kBrickPairs);
        const uint32_t idx[4] = {t.i00, t.i10, t.i01, t.i11};
#pragma unroll
        for(int k = 0; k < (LINEAR ? 4 : 1); k++)
        {
            const Words<W> a = rows[idx[k]];
#pragma unroll
            for(int w = 0; w < W; w++)
                R.r[k][w] = a.w[w];
        }
        const Words<DIM> sc = reinterpret_cast<const Words<DIM>*>(S.scale)[t.brick];
#pragma unroll
        for(int c = 0; c < DIM; c++)
            R.sc[c] = sc.w[c];
    }

    template<int DT, bool LINEAR>
    __device__ __attribute__((always_inline)) inline void slot_fetch(const BrickSource& S, const BrickTaps& t, SlotRaw<DT>& R)
    {
        switch(S.dim)
        {
        case 1: slot_fetch_dim<DT, LINEAR, 1>(S, t, R); break;
        case 2: slot_fetch_dim<DT, LINEAR, 2>(S, t, R); break;
        case 3: slot_fetch_dim<DT, LINEAR, 3>(S, t, R); break;
        default: slot_fetch_dim<DT, LINEAR, 4>(S, t, R); break;
        }
    }

    // lo[c] = v_c[x], hi[c] = v_c[x+1] of one fetched row (still divided by the brick scale)
    template<int DT, int DIM>
    __device__ inline void decode_row(const uint32_t* w, float lo[DIM], float hi[DIM])
    {
#pragma unroll
        for(int c = 0; c < DIM; c++)
        {
            if constexpr(DT == kDtF32)
            {
                lo[c] = __uint_as_float(w[2 * c]);
                hi[c] = __uint_as_float(w[2 * c + 1]);
            }
            else if constexpr(DT == kDtF16)
            {
                lo[c] = half_lo(w[c]);
                hi[c] = half_hi(w[c]);
            }
            else
            {
                lo[c] = __uint_as_float(w[c] << 16);
                hi[c] = __uint_as_float(w[c] & 0xffff0000u);
            }
        }
    }

    // raw sample of every component of one source with DIM components (v[c], c < DIM)
    template<int DT, bool LINEAR, int DIM>
    __device__ inline void slot_interp(const SlotRaw<DT>& R, const BrickTaps& t, float v[4])
    {
        if(!LINEAR)
        {
            float a[DIM], b[DIM];
            decode_row<DT, DIM>(R.r[0], a, b);
#pragma unroll
            for(int c = 0; c < DIM; c++)
                v[c] = a[c] * __uint_as_float(R.sc[c]);
            return;
        }
        float d000[DIM], d100[DIM], d010[DIM], d110[DIM], d001[DIM], d101[DIM], d011[DIM], d111[DIM];
        decode_row<DT, DIM>(R.r[0], d000, d100);
        decode_row<DT, DIM>(R.r[1], d010, d110);
        decode_row<DT, DIM>(R.r[2], d001, d101);
        decode_row<DT, DIM>(R.r[3], d011, d111);
        // sum of the corners times their weights (one multiply and seven fmas per
        // component), then the brick scale; equal to the nested lerps in exact arithmetic
#pragma unroll
        for(int c = 0; c < DIM; c++)
        {
            float a = d000[c] * t.w[0];
            a = __builtin_fmaf(d001[c], t.w[1], a);
            a = __builtin_fmaf(d010[c], t.w[2], a);
            a = __builtin_fmaf(d011[c], t.w[3], a);
            a = __builtin_fmaf(d100[c], t.w[4], a);
            a = __builtin_fmaf(d101[c], t.w[5], a);
            a = __builtin_fmaf(d110[c], t.w[6], a);
            a = __builtin_fmaf(d111[c], t.w[7], a);
            v[c] = a * __uint_as_float(R.sc[c]);
        }
    }

    // the functor-chain value of one source's sample, dispatched on its component
    // count (fetch and interpolation in the same case: one source at a time)
    template<int DT, bool LINEAR, int DIM>
    __device__ inline void slot_sample(const BrickSource& S, const BrickTaps& t, float v[4])
    {
        SlotRaw<DT> R;
        slot_fetch_dim<DT, LINEAR, DIM>(S, t, R);
        slot_interp<DT, LINEAR, DIM>(R, t, v);
    }

    template<int DT, bool LINEAR>
    __device__ __attribute__((always_inline)) inline float brick_value(const BrickSource& S, const BrickTaps& t)
    {
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        switch(S.dim)
        {
        case 1: slot_sample<DT, LINEAR, 1>(S, t, v); break;
        case 2: slot_sample<DT, LINEAR, 2>(S, t, v); break;
        case 3: slot_sample<DT, LINEAR, 3>(S, t, v); break;
        default: slot_sample<DT, LINEAR, 4>(S, t, v); break;
        }
        // one chain per source slot, with the component count as a value
        return apply_chain(S.chain, v[0], v[1], v[2], v[3], S.dim);
    }

    // Source signature: the component counts of the enabled sources in slot
    // order, as base-5 digits (slot 0 lowest), 0 = resolved at run time.  A
    // march compiled for one signature fetches the rows of all its sources
    // before interpolating any of them, so their latencies overlap.
    constexpr bool sig_natural(int sig)
    {
        return sig >= kSigNatural;
    }

    constexpr int sig_count(int sig)
    {
        int n = 0;
        for(sig %= kSigNatural; sig > 0; sig /= 5)
            n++;
        return n;
    }

    constexpr int sig_dim(int sig, int slot)
    {
        sig %= kSigNatural;
        for(int i = 0; i < slot; i++)
            sig /= 5;
        return sig % 5;
    }

    // the natural chain's value of a D-component sample: the value itself, or the length
    // with apply_chain's association ((v0 v0 + v1 v1) + v2 v2)
    template<int D>
    __device__ inline float natural_value(const float (&v)[4])
    {
        if constexpr(D == 1)
            return v[0];
        else if constexpr(D == 2)
            return sqrtf(v[0] * v[0] + v[1] * v[1]);
        else if constexpr(D == 3)
            return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        else
            return sqrtf((v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3]));
    }

    template<int DT, bool LINEAR, int SIG, int S>
    __device__ __attribute__((always_inline)) inline void sig_fetch(const PostParams& Q, const BrickTaps& t, SlotRaw<DT>* R)
    {
        if constexpr(S < sig_count(SIG))
        {
            slot_fetch_dim<DT, LINEAR, sig_dim(SIG, S)>(Q.src[S], t, R[S]);
            sig_fetch<DT, LINEAR, SIG, S + 1>(Q, t, R);
        }
    }

    struct PostSample
    {
        float r, g, b, a; // premultiplied, normalised volume colour
        float ir, ig, ib, ia; // iso colour and alpha (alpha / threshold * 0.5, maximum over sources)
    };

    // one source's classified sample into the running sums (weights in slot order)
    __device__ __attribute__((always_inline)) inline void post_accumulate(PostSample& o, const BrickSource& S, const float4 col)
    {
        if(S.weight > 0.0f)
        {
            const float a = col.w * S.weight;
            o.r = o.r + col.x * a;
            o.g = o.g + col.y * a;
            o.b = o.b + col.z * a;
            o.a = o.a + a;
        }
        if(S.iso > 0.0f)
        {
            const float a = col.w * S.isoScale;
            if(o.ia < a)
            {
                o.ir = col.x;
                o.ig = col.y;
                o.ib = col.z;
                o.ia = a;
            }
        }
    }

    template<int DT, bool LINEAR, int SIG, int S>
    __device__ __attribute__((always_inline)) inline void sig_classify(
        const PostParams& Q, const float4* lut, const BrickTaps& t, const SlotRaw<DT>* R, PostSample& o)
    {
        if constexpr(S < sig_count(SIG))
        {
            constexpr int D = sig_dim(SIG, S);
            float v[4] = {0.f, 0.f, 0.f, 0.f};
            slot_interp<DT, LINEAR, D>(R[S], t, v);
            const float tv = sig_natural(SIG) ? natural_value<D>(v) : apply_chain(Q.src[S].chain, v[0], v[1], v[2], v[3], D);
            post_accumulate(o, Q.src[S], lut[S * Q.T + tf_index(tv, Q.T)]);
            sig_classify<DT, LINEAR, SIG, S + 1>(Q, lut, t, R, o);
        }
    }

    template<int DT, bool LINEAR, int SIG>
    __device__ __attribute__((always_inline)) inline PostSample post_classify(const PostParams& Q, const float4* lut, int X, int Y, int Z, float px, float py, float pz)
    {
        const BrickTaps t = brick_taps<LINEAR>(Q, X, Y, Z, px, py, pz);
        PostSample o{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if constexpr(SIG != 0)
        {
            SlotRaw<DT> R[sig_count(SIG)];
            sig_fetch<DT, LINEAR, SIG, 0>(Q, t, R);
            sig_classify<DT, LINEAR, SIG, 0>(Q, lut, t, R, o);
        }
        else
        {
            // unrolled over the source slots so every parameter-block index is constant
#pragma unroll
            for(int s = 0; s < kMaxBrickSources; s++)
            {
                if(s >= Q.nsrc)
                    break;
                const BrickSource& S = Q.src[s];
                const float tv = brick_value<DT, LINEAR>(S, t);
                post_accumulate(o, S, lut[s * Q.T + tf_index(tv, Q.T)]);
            }
        }
        // invTotalWeight = 1 / (sum of volume-source weights), 0 when that sum is 0
        o.r = o.r * Q.invTotalWeight;
        o.g = o.g * Q.invTotalWeight;
        o.b = o.b * Q.invTotalWeight;
        o.a = o.a * Q.invTotalWeight;
        return o;
    }

    // Post-classification march over field bricks.  Same pass structure as
    // march_kernel (clear, particle DDA, one traversal serving iso and volume,
    // G-buffer + premultiplied volume colour out); sample positions
    // startU + stepVec*i with startSteps / end steps taken from the binary64 ray.
    // the enabled sources' transfer functions, staged in LDS in source order
    __device__ inline void stage_luts(float4* lds_lut, const PostParams& Q)
    {
#pragma unroll
        for(int j = 0; j < kMaxBrickSources; j++)
        {
            if(j >= Q.nsrc)
                break;
            const float4* src = Q.luts + size_t(Q.src[j].lutSlot) * Q.T;
            for(int i = threadIdx.x; i < Q.T; i += blockDim.x)
                lds_lut[j * Q.T + i] = src[i];
        }
    }

    template<int DT, bool LINEAR, bool COUNT, int SIG = 0>
    __device__ __attribute__((always_inline)) inline void brick_march_body(
        const FrameParams& F,
        const PostParams& Q,
        const ParticleParams& PP,
        const float4* lut,
        uint32_t tile)
    {
        int px, py;
        const bool valid = tile_pixel_at(F, tile, px, py);
        unsigned long long nVol = 0, nIso = 0, nIsoHits = 0, nVolRays = 0, nIsoRays = 0;
        unsigned long long nCells = 0, nTests = 0, nPartRays = 0, nSteps = 0;
        if(valid)
        {
            const size_t p = size_t(py) * size_t(F.W) + size_t(px);
            uint32_t color = F.bgClear;
            F3 normal = f3(0.f, 0.f, 0.f);
            float depth = FLT_MAX;
            float4 vcol = make_float4(0.f, 0.f, 0.f, -1.f);
            RayD rd;
            const bool inside = ray_setup_d(F, Q, px, py, rd);
            Ray ray;
            ray.dir = f3(float(rd.dir[0]), float(rd.dir[1]), float(rd.dir[2]));
            ray.start = f3(float(rd.start[0]), float(rd.start[1]), float(rd.start[2]));
            ray.startDepth = float(rd.startDepth);
            ray.endDepth = float(rd.endDepth);
            ray.isClipped = rd.isClipped;
            ray.clipN = f3(float(rd.clipN[0]), float(rd.clipN[1]), float(rd.clipN[2]));
            if(inside && PP.nsrc > 0)
            {
                nPartRays++;
                particle_dda<COUNT>(PP, ray, color, depth, normal, nCells, nTests);
            }
            const int mode = F.renderMode;
            const bool depthZeroed = mode == 2 || (mode >= 4 && mode <= 8);
            if(inside)
            {
                const double endD = fmin(rd.endDepth, double(depth));
                bool isoOn = F.isoEnabled && endD > rd.startDepth;
                bool volBlend = F.volumeEnabled && !depthZeroed && endD > rd.startDepth;
                if(isoOn || volBlend)
                {
                    const F3 sc = f3(F.scale[0], F.scale[1], F.scale[2]);
                    const double ex = rd.dir[0] / double(sc.x), ey = rd.dir[1] / double(sc.y), ez = rd.dir[2] / double(sc.z);
                    const double stepUd = double(F.step)
                        * (sqrt(rd.dir[0] * rd.dir[0] + rd.dir[1] * rd.dir[1] + rd.dir[2] * rd.dir[2]) / sqrt(ex * ex + ey * ey + ez * ez));
                    const float stepU = float(stepUd);
                    const float factor = stepU / F.minSize * 2.0f * F.totalWeightAll;
                    const int startSteps = int(ceil(rd.startDepth / stepUd));
                    const int endIso = int(floor(endD / stepUd));
                    int endVol = endIso;
                    const F3 stepVec = divv(muls(ray.dir, stepU), sc);
                    const F3 startU = divv(ray.start, sc);
                    if(isoOn)
                        nIsoRays++;
                    float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
                    float oldValue = 0.0f;
                    // traversal state as step bounds, not flags (no lane masks carried by the
                    // loop): the iso test runs while i <= isoEnd, the volume while i <= volEnd,
                    // either set below every step to switch it off; hitStep is the hit's step
                    const int off = -0x7fffffff;
                    int isoEnd = isoOn ? endIso : off;
                    int volEnd = volBlend ? endVol : off;
                    int hitStep = -1;
                    for(int i = startSteps;; i++)
                    {
                        if(i > (isoEnd > volEnd ? isoEnd : volEnd))
                            break;
                        if(COUNT)
                            nSteps++;
                        const float fi = float(i);
                        const PostSample sm = post_classify<DT, LINEAR, SIG>(
                            Q, lut, F.X, F.Y, F.Z, startU.x + stepVec.x * fi, startU.y + stepVec.y * fi, startU.z + stepVec.z * fi);
                        if(i <= isoEnd)
                        {
                            const float value = gmin(sm.ia, 1.0f);
                            if(COUNT)
                                nIso++;
                            const float tmp = oldValue;
                            oldValue = value;
                            if(value >= 0.5f)
                            {
                                const bool first = ray.isClipped && i == startSteps;
                                depth = first ? ray.startDepth : fi * stepU + stepU * (0.5f - tmp) / (value - tmp);
                                hitStep = i;
                                isoEnd = off;
                                if(COUNT)
                                    nIsoHits++;
                                // the volume pass ends at min(clip end, iso depth)
                                const double ve = fmin(rd.endDepth, double(depth));
                                if(!(ve > rd.startDepth))
                                    volEnd = off;
                                else if(volEnd != off)
                                    volEnd = int(floor(ve / stepUd));
                            }
                        }
                        if(i <= volEnd)
                        {
                            if(COUNT)
                                nVol++;
                            const float w = 1.0f - ca;
                            cr = cr + w * (sm.r * factor);
                            cg = cg + w * (sm.g * factor);
                            cb = cb + w * (sm.b * factor);
                            ca = ca + w * (sm.a * factor);
                            if(ca > 0.99f)
                                volEnd = off;
                        }
                    }
                    const bool isoHit = hitStep >= 0;
                    const bool isoFirst = isoHit && ray.isClipped && hitStep == startSteps;
                    // a hit at or before the ray's start leaves no volume segment to blend
                    if(isoHit && !(fmin(rd.endDepth, double(depth)) > rd.startDepth))
                        volBlend = false;
                    if(isoHit)
                    {
                        const F3 pu = divv(add(ray.start, muls(ray.dir, depth)), sc);
                        // hit colour, then the six central-difference samples of the iso value
                        // (one loop, so the per-sample classification is inlined once)
                        float hr = 0.f, hg = 0.f, hb = 0.f;
                        F3 g = f3(0.f, 0.f, 0.f);
#pragma unroll 1
                        for(int d = -1; d < 6; d++)
                        {
                            const float sx = d == 0 ? 1.0f : (d == 1 ? -1.0f : 0.0f);
                            const float sy = d == 2 ? 1.0f : (d == 3 ? -1.0f : 0.0f);
                            const float sz = d == 4 ? 1.0f : (d == 5 ? -1.0f : 0.0f);
                            const PostSample h = post_classify<DT, LINEAR, SIG>(Q, lut, F.X, F.Y, F.Z, pu.x + sx, pu.y + sy, pu.z + sz);
                            const float v = gmin(h.ia, 1.0f);
                            if(d < 0)
                            {
                                hr = h.ir;
                                hg = h.ig;
                                hb = h.ib;
                            }
                            else if(d == 0)
                                g.x = v;
                            else if(d == 1)
                                g.x = g.x - v;
                            else if(d == 2)
                                g.y = v;
                            else if(d == 3)
                                g.y = g.y - v;
                            else if(d == 4)
                                g.z = v;
                            else
                                g.z = g.z - v;
                        }
                        float gl = len3(g);
                        if(isoFirst || gl == 0.0f)
                        {
                            g = ray.clipN;
                            gl = 1.0f;
                        }
                        normal = f3(-g.x / gl, -g.y / gl, -g.z / gl);
                        color = to_byte4(hr, hg, hb, 1.0f);
                    }
                    if(volBlend)
                    {
                        nVolRays++;
                        vcol = make_float4(cr, cg, cb, ca);
                    }
                }
            }
            F.color[p] = color;
            F.depth[p] = depth;
            F.normal[3 * p + 0] = normal.x;
            F.normal[3 * p + 1] = normal.y;
            F.normal[3 * p + 2] = normal.z;
            F.volColor[p] = vcol;
            if(!F.aoEnabled)
                F.ao[p] = 0.0f;
        }
        if(COUNT)
        {
            block_count_add(F.counters + 0, nVol);
            block_count_add(F.counters + 1, nIso);
            block_count_add(F.counters + 2, nIsoHits);
            block_count_add(F.counters + 3, nCells);
            block_count_add(F.counters + 4, nTests);
            block_count_add(F.counters + 5, nVolRays);
            block_count_add(F.counters + 6, nIsoRays);
            block_count_add(F.counters + 7, nPartRays);
            block_count_add(F.counters + 8, nSteps);
            // the wave runs the traversal loop as long as its longest lane
            int wm = int(nSteps);
            for(int off = 32; off > 0; off >>= 1)
                wm = max(wm, __shfl_xor(wm, off, 64));
            block_count_add(F.counters + 9, (threadIdx.x & 63u) == 0 ? (unsigned long long) wm : 0ull);
        }
    }

    template<int DT, bool LINEAR, bool COUNT>
    __global__ __launch_bounds__(256) void brick_march_kernel(const FrameParams F, const PostParams Q, const ParticleParams PP)
    {
        extern __shared__ float4 lds_lut[];
        stage_luts(lds_lut, Q);
        __syncthreads();
        brick_march_body<DT, LINEAR, COUNT>(F, Q, PP, lds_lut, march_tile(F));
    }

    // trilinear march compiled for one source signature (see sig_count)
    template<int DT, int SIG>
    __global__ __launch_bounds__(256) void brick_march_sig_kernel(const FrameParams F, const PostParams Q, const ParticleParams PP)
    {
        extern __shared__ float4 lds_lut[];
        stage_luts(lds_lut, Q);
        __syncthreads();
        brick_march_body<DT, true, false, SIG>(F, Q, PP, lds_lut, march_tile(F));
    }

    // ------------------------------------------------------------------ batched form
    // One persistent launch over many frames (sub-volume x view): workgroups pop
    // chunks of 16x16 tiles from a global counter until the work is exhausted
    // (every workgroup reaches the exit when the counter passes the last chunk).
    // Frames share the transfer functions (staged once per workgroup).
    template<int DT, bool LINEAR, int SIG = 0>
    __global__ __launch_bounds__(256) void batch_march_kernel(const BatchParams B)
    {
        extern __shared__ float4 lds_lut[];
        __shared__ uint32_t s_work;
        stage_luts(lds_lut, B.posts[0]);
        for(;;)
        {
            __syncthreads();
            if(threadIdx.x == 0)
                s_work = atomicAdd(B.counter, 1u);
            __syncthreads();
            // uniform across the workgroup: keep the work index in a scalar register so
            // the frame's parameter blocks are read with scalar loads
            const uint32_t t0 = uint32_t(__builtin_amdgcn_readfirstlane(int(s_work))) * uint32_t(B.chunk);
            if(t0 >= B.ntiles)
                break;
            const uint32_t t1 = min(t0 + uint32_t(B.chunk), B.ntiles);
            // frame of tile t0: largest i with tileStart[i] <= t0
            int lo = 0, hi = B.nframes - 1;
            while(lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if(B.tileStart[mid] <= t0)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            int item = __builtin_amdgcn_readfirstlane(lo);
            for(uint32_t t = t0; t < t1; t++)
            {
                while(t >= B.tileStart[item + 1])
                    item++;
                item = __builtin_amdgcn_readfirstlane(item);
                brick_march_body<DT, LINEAR, false, SIG>(B.frames[item], B.posts[item], B.parts[item], lds_lut, t - B.tileStart[item]);
            }
        }
    }

    // per-frame passes of the batch: grid (tiles of the largest frame, frames)
    __global__ __launch_bounds__(256) void batch_ssao_kernel(const FrameParams* frames)
    {
        const FrameParams& F = frames[blockIdx.y];
        if(F.aoEnabled)
            ssao_body(F, blockIdx.x);
    }

    __global__ __launch_bounds__(256) void batch_shade_kernel(const FrameParams* frames)
    {
        shade_body(frames[blockIdx.y], blockIdx.x);
    }


    // per brick and component: scale = max |v| over the 9^3 region (1 if all zero)
    __device__ inline uint32_t to_bf16_rne(float f)
    {
        const uint32_t u = __float_as_uint(f);
        if((u & 0x7f800000u) == 0x7f800000u)
            return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u); // inf / nan
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    }

    // One workgroup per brick: the brick's 9^3 region (x, y, z clamped to the local
    // box) of all DIM components is read once into LDS; wave c < DIM reduces the
    // max |v| of component c (the brick scale: 1 when the region is all zero or the
    // form is FP32); then every thread writes pairs (v[x], v[x+1]) of the region,
    // divided by the scale and rounded for 16-bit forms.  Output element
    // (brick*648 + pair)*DIM + c, pair = (lz*9 + ly)*8 + lx.
    template<int DT, int DIM>
    __global__ __launch_bounds__(256) void brick_build_kernel(
        const float* src,
        int64_t sx,
        int64_t sy,
        int64_t sz,
        int X,
        int Y,
        int Z,
        int nbx,
        int nby,
        float* scale,
        void* out)
    {
        __shared__ float reg[729 * DIM];
        __shared__ float sc[DIM], rc[DIM];
        __shared__ int64_t rowOff[81]; // element offset of region row (ly, lz), x = 0
        const int64_t brick = blockIdx.x;
        const int Bx = int(brick % nbx), By = int((brick / nbx) % nby), Bz = int(brick / (int64_t(nbx) * nby));
        if(threadIdx.x < 81)
        {
            const int lz = int(threadIdx.x) / 9, ly = int(threadIdx.x) - lz * 9;
            const int y = min(By * 8 + ly, Y - 1), z = min(Bz * 8 + lz, Z - 1);
            rowOff[threadIdx.x] = (int64_t(y) * sy + int64_t(z) * sz) * DIM;
        }
        __syncthreads();
        const int x0 = Bx * 8, xLast = X - 1 - x0; // last local x inside the box
        const int64_t sxD = sx * DIM;
        for(int i = threadIdx.x; i < 729 * DIM; i += 256)
        {
            const int vox = i / DIM, c = i - vox * DIM;
            const int row = vox / 9, lx = vox - row * 9;
            const int x = x0 + min(lx, xLast);
            reg[i] = src[rowOff[row] + int64_t(x) * sxD + c];
        }
        __syncthreads();
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if(wave < DIM)
        {
            float m = 0.0f;
            if(DT != kDtF32)
                for(int vox = lane; vox < 729; vox += 64)
                    m = fmaxf(m, fabsf(reg[vox * DIM + wave]));
            for(int off = 32; off > 0; off >>= 1)
                m = fmaxf(m, __shfl_xor(m, off));
            const float s = m > 0.0f ? m : 1.0f;
            if(lane == 0)
            {
                sc[wave] = s;
                // RN(1/s) for the quotients, 0 where s is outside [2^-60, 2^60] (there
                // the quotients take the IEEE division)
                rc[wave] = (s >= 0x1p-60f && s <= 0x1p+60f) ? 1.0f / s : 0.0f;
                scale[brick * DIM + wave] = s;
            }
        }
        __syncthreads();
        for(int e = threadIdx.x; e < kBrickPairs * DIM; e += 256)
        {
            const int pair = e / DIM, c = e - pair * DIM;
            const int vox0 = (pair >> 3) * 9 + (pair & 7);
            const float v0 = reg[vox0 * DIM + c], v1 = reg[(vox0 + 1) * DIM + c];
            const int64_t i = brick * (kBrickPairs * DIM) + e;
            if(DT == kDtF32)
                reinterpret_cast<float2*>(out)[i] = make_float2(v0, v1);
            else
            {
                // v / s as x*r plus one fma residual step (div_by), with v's sign (the
                // residual step turns -0 into +0): the IEEE quotient for |v| <= s with s in
                // [2^-60, 2^60]; for BF16, whose range keeps tiny quotients, only where
                // |v| >= 2^-60 (the residual stays normal)
                const float s = sc[c], r = rc[c];
                auto quot = [&](float v) {
                    const bool fast = r != 0.0f && (DT == kDtF16 || fabsf(v) >= 0x1p-60f);
                    return fast ? copysignf(div_by(v, s, r), v) : v / s;
                };
                const float q0 = quot(v0), q1 = quot(v1);
                uint32_t w;
                if(DT == kDtF16)
                    w = uint32_t(__half_as_ushort(__float2half_rn(q0))) | (uint32_t(__half_as_ushort(__float2half_rn(q1))) << 16);
                else
                    w = to_bf16_rne(q0) | (to_bf16_rne(q1) << 16);
                reinterpret_cast<uint32_t*>(out)[i] = w;
            }
        }
    }

    template<int DT>
    hipError_t launch_brick_build(
        const float* src, const int64_t stride[3], int dim, int X, int Y, int Z, int nbx, int nby, int64_t nbricks, float* scale,
        void* out, hipStream_t s)
    {
        const dim3 grid = dim3(static_cast<unsigned>(nbricks)), block = dim3(256);
        switch(dim)
        {
        case 1: hipLaunchKernelGGL((brick_build_kernel<DT, 1>), grid, block, 0, s, src, stride[0], stride[1], stride[2], X, Y, Z, nbx, nby, scale, out); break;
        case 2: hipLaunchKernelGGL((brick_build_kernel<DT, 2>), grid, block, 0, s, src, stride[0], stride[1], stride[2], X, Y, Z, nbx, nby, scale, out); break;
        case 3: hipLaunchKernelGGL((brick_build_kernel<DT, 3>), grid, block, 0, s, src, stride[0], stride[1], stride[2], X, Y, Z, nbx, nby, scale, out); break;
        case 4: hipLaunchKernelGGL((brick_build_kernel<DT, 4>), grid, block, 0, s, src, stride[0], stride[1], stride[2], X, Y, Z, nbx, nby, scale, out); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }

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
        hipStream_t s)
    {
        const int nbx = (X + 7) / 8, nby = (Y + 7) / 8, nbz = (Z + 7) / 8;
        const int64_t nbricks = int64_t(nbx) * nby * nbz;
        if(nbricks == 0)
            return hipSuccess;
        if(nbricks > int64_t(0x7fffffff))
            return hipErrorInvalidValue;
        if(dtype == kDtF32)
            return launch_brick_build<kDtF32>(src, stride, dim, X, Y, Z, nbx, nby, nbricks, scale, out, s);
        if(dtype == kDtF16)
            return launch_brick_build<kDtF16>(src, stride, dim, X, Y, Z, nbx, nby, nbricks, scale, out, s);
        return launch_brick_build<kDtBf16>(src, stride, dim, X, Y, Z, nbx, nby, nbricks, scale, out, s);
    }

    static_assert(sizeof(FrameParams) + sizeof(PostParams) + sizeof(ParticleParams) <= 4096, "kernel argument block");

    // workgroups of a march launch: one per tile, or the padded super-tile grid
    unsigned march_grid(const FrameParams& F)
    {
        const unsigned nb = unsigned(F.tilesX) * unsigned(F.tilesY);
        const unsigned S = unsigned(F.tileOrder);
        if(S == 0 || nb == 0)
            return nb;
        const unsigned ns = ((unsigned(F.tilesX) + S - 1u) / S) * ((unsigned(F.tilesY) + S - 1u) / S);
        return (ns + 7u) / 8u * 8u * S * S;
    }

    // a source's chain is its natural one: no operations for a scalar source, one
    // `length` (then nothing) for a vector source
    bool natural_chain(const ChainDesc& c, int dim)
    {
        for(int k = 0; k < kMaxFunctors; k++)
        {
            const int want = (k == 0 && dim > 1) ? OP_LENGTH : OP_IDEM;
            if(c.op[k] != want)
                return false;
        }
        return true;
    }

    bool natural_chains(const PostParams& Q)
    {
        for(int i = 0; i < Q.nsrc; i++)
            if(!natural_chain(Q.src[i].chain, Q.src[i].dim))
                return false;
        return true;
    }

    int post_signature(const PostParams& Q)
    {
        int sig = 0;
        for(int i = Q.nsrc - 1; i >= 0; i--)
            sig = sig * 5 + Q.src[i].dim;
        return Q.nsrc > 0 ? sig : -1;
    }

    hipError_t launch_brick_march(const FrameParams& F, const PostParams& Q, const ParticleParams& PP, bool counting, hipStream_t s)
    {
        const unsigned nb = march_grid(F);
        if(nb == 0)
            return hipSuccess;
        const dim3 grid(nb), block(256);
        const size_t lds = size_t(Q.nsrc) * Q.T * sizeof(float4);
        if(F.linear && !counting)
        {
            // signatures with a compiled march: density (1), vector (3), and pairs of them
            const int sig = post_signature(Q);
            const int v = Q.dtype * 8 + (sig == 1 ? 0 : sig == 3 ? 1 : sig == 6 ? 2 : sig == 8 ? 3 : sig == 16 ? 4 : 7);
            const bool natural = natural_chains(Q);
            switch(v)
            {
#define ISAAC_SIG_CASE(dt, k, sg)                                                                                       \
    case dt * 8 + k:                                                                                                    \
        if(natural)                                                                                                     \
            hipLaunchKernelGGL((brick_march_sig_kernel<dt, sg + kSigNatural>), grid, block, lds, s, F, Q, PP);          \
        else                                                                                                            \
            hipLaunchKernelGGL((brick_march_sig_kernel<dt, sg>), grid, block, lds, s, F, Q, PP);                        \
        return hipGetLastError();
                ISAAC_SIG_CASE(0, 0, 1)
                ISAAC_SIG_CASE(0, 1, 3)
                ISAAC_SIG_CASE(0, 2, 6)
                ISAAC_SIG_CASE(0, 3, 8)
                ISAAC_SIG_CASE(0, 4, 16)
                ISAAC_SIG_CASE(1, 0, 1)
                ISAAC_SIG_CASE(1, 1, 3)
                ISAAC_SIG_CASE(1, 2, 6)
                ISAAC_SIG_CASE(1, 3, 8)
                ISAAC_SIG_CASE(1, 4, 16)
                ISAAC_SIG_CASE(2, 0, 1)
                ISAAC_SIG_CASE(2, 1, 3)
                ISAAC_SIG_CASE(2, 2, 6)
                ISAAC_SIG_CASE(2, 3, 8)
                ISAAC_SIG_CASE(2, 4, 16)
#undef ISAAC_SIG_CASE
            default: break;
            }
        }
        const int v = (Q.dtype * 4) | (F.linear ? 2 : 0) | (counting ? 1 : 0);
        switch(v)
        {
        case 0: hipLaunchKernelGGL((brick_march_kernel<0, false, false>), grid, block, lds, s, F, Q, PP); break;
        case 1: hipLaunchKernelGGL((brick_march_kernel<0, false, true>), grid, block, lds, s, F, Q, PP); break;
        case 2: hipLaunchKernelGGL((brick_march_kernel<0, true, false>), grid, block, lds, s, F, Q, PP); break;
        case 3: hipLaunchKernelGGL((brick_march_kernel<0, true, true>), grid, block, lds, s, F, Q, PP); break;
        case 4: hipLaunchKernelGGL((brick_march_kernel<1, false, false>), grid, block, lds, s, F, Q, PP); break;
        case 5: hipLaunchKernelGGL((brick_march_kernel<1, false, true>), grid, block, lds, s, F, Q, PP); break;
        case 6: hipLaunchKernelGGL((brick_march_kernel<1, true, false>), grid, block, lds, s, F, Q, PP); break;
        case 7: hipLaunchKernelGGL((brick_march_kernel<1, true, true>), grid, block, lds, s, F, Q, PP); break;
        case 8: hipLaunchKernelGGL((brick_march_kernel<2, false, false>), grid, block, lds, s, F, Q, PP); break;
        case 9: hipLaunchKernelGGL((brick_march_kernel<2, false, true>), grid, block, lds, s, F, Q, PP); break;
        case 10: hipLaunchKernelGGL((brick_march_kernel<2, true, false>), grid, block, lds, s, F, Q, PP); break;
        default: hipLaunchKernelGGL((brick_march_kernel<2, true, true>), grid, block, lds, s, F, Q, PP); break;
        }
        return hipGetLastError();
    }

    hipError_t launch_batch_march(const BatchParams& B, int dtype, bool linear, int sig, size_t lds, int grid, hipStream_t s)
    {
        if(B.nframes <= 0 || grid <= 0)
            return hipSuccess;
        const bool natural = sig_natural(sig);
        sig %= kSigNatural;
        if(linear && (sig == 1 || sig == 3 || sig == 16))
        {
            const int k = dtype * 3 + (sig == 1 ? 0 : sig == 3 ? 1 : 2);
            switch(k)
            {
#define ISAAC_SIG_CASE(dt, j, sg)                                                                                       \
    case dt * 3 + j:                                                                                                    \
        if(natural)                                                                                                     \
            hipLaunchKernelGGL((batch_march_kernel<dt, true, sg + kSigNatural>), dim3(grid), dim3(256), lds, s, B);     \
        else                                                                                                            \
            hipLaunchKernelGGL((batch_march_kernel<dt, true, sg>), dim3(grid), dim3(256), lds, s, B);                   \
        return hipGetLastError();
                ISAAC_SIG_CASE(0, 0, 1)
                ISAAC_SIG_CASE(0, 1, 3)
                ISAAC_SIG_CASE(0, 2, 16)
                ISAAC_SIG_CASE(1, 0, 1)
                ISAAC_SIG_CASE(1, 1, 3)
                ISAAC_SIG_CASE(1, 2, 16)
                ISAAC_SIG_CASE(2, 0, 1)
                ISAAC_SIG_CASE(2, 1, 3)
                ISAAC_SIG_CASE(2, 2, 16)
#undef ISAAC_SIG_CASE
            default: break;
            }
        }
        const int v = dtype * 2 + (linear ? 1 : 0);
        switch(v)
        {
        case 0: hipLaunchKernelGGL((batch_march_kernel<0, false>), dim3(grid), dim3(256), lds, s, B); break;
        case 1: hipLaunchKernelGGL((batch_march_kernel<0, true>), dim3(grid), dim3(256), lds, s, B); break;
        case 2: hipLaunchKernelGGL((batch_march_kernel<1, false>), dim3(grid), dim3(256), lds, s, B); break;
        case 3: hipLaunchKernelGGL((batch_march_kernel<1, true>), dim3(grid), dim3(256), lds, s, B); break;
        case 4: hipLaunchKernelGGL((batch_march_kernel<2, false>), dim3(grid), dim3(256), lds, s, B); break;
        default: hipLaunchKernelGGL((batch_march_kernel<2, true>), dim3(grid), dim3(256), lds, s, B); break;
        }
        return hipGetLastError();
    }

    hipError_t launch_batch_ssao(const FrameParams* frames, int nframes, int maxTiles, hipStream_t s)
    {
        if(nframes <= 0 || maxTiles <= 0)
            return hipSuccess;
        hipLaunchKernelGGL(batch_ssao_kernel, dim3(unsigned(maxTiles), unsigned(nframes)), dim3(256), 0, s, frames);
        return hipGetLastError();
    }

    hipError_t launch_batch_shade(const FrameParams* frames, int nframes, int maxTiles, hipStream_t s)
    {
        if(nframes <= 0 || maxTiles <= 0)
            return hipSuccess;
        hipLaunchKernelGGL(batch_shade_kernel, dim3(unsigned(maxTiles), unsigned(nframes)), dim3(256), 0, s, frames);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ launchers
    // the 4 voxels of quad q (x, y, z of its first voxel): every source's components,
    // f[s][k * D + c] for voxel k, component c (16-byte loads)
    template<int SIG>
    __device__ __attribute__((always_inline)) inline void cls_fetch(
        const ClassifyParams& P, int64_t q, float (&f)[sig_count(SIG)][16], int& x, int& y, int& z)
    {
        constexpr int N = sig_count(SIG);
        const uint32_t u = uint32_t(q) << 2;
        const uint32_t zq = fastdiv(u, P.divXY_mul, P.divXY_shift);
        const uint32_t rem = u - zq * uint32_t(P.X * P.Y);
        const uint32_t yq = fastdiv(rem, P.divX_mul, P.divX_shift);
        x = int(rem - yq * uint32_t(P.X));
        y = int(yq);
        z = int(zq);
#pragma unroll
        for(int s = 0; s < N; s++)
        {
            const int D = sig_dim(SIG, s);
            const ClassifySource& S = P.src[s];
            const float4* e = reinterpret_cast<const float4*>(
                S.data + (int64_t(x) + int64_t(y) * S.stride[1] + int64_t(z) * S.stride[2]) * D);
#pragma unroll
            for(int j = 0; j < D; j++)
            {
                const float4 w = e[j];
                f[s][4 * j + 0] = w.x;
                f[s][4 * j + 1] = w.y;
                f[s][4 * j + 2] = w.z;
                f[s][4 * j + 3] = w.w;
            }
        }
    }

    // classification of one fetched quad, its 4 texels stored with two 16-byte stores
    template<int SIG>
    __device__ __attribute__((always_inline)) inline void cls_emit(
        const ClassifyParams& P, const float4* lds_lut, const float (&f)[sig_count(SIG)][16], int x, int y, int z, int64_t q)
    {
        constexpr int N = sig_count(SIG);
        const int T = P.T;
        const uint32_t u = uint32_t(q) << 2;
        uint32_t texel[8];
#pragma unroll
        for(int k = 0; k < 4; k++)
        {
            float vr = 0.f, vg = 0.f, vb = 0.f, va = 0.f;
            float ir = 0.f, ig = 0.f, ib = 0.f, ia = 0.f;
#pragma unroll
            for(int s = 0; s < N; s++)
            {
                const int D = sig_dim(SIG, s);
                const ClassifySource& S = P.src[s];
                const float c0 = f[s][k * D];
                const float c1 = D > 1 ? f[s][k * D + 1] : 0.f;
                const float c2 = D > 2 ? f[s][k * D + 2] : 0.f;
                // natural chains (identity, or one `length`) evaluated inline, in the
                // interpreter's order, instead of interpreting the chain per voxel
                const float tv = sig_natural(SIG) ? (D == 1 ? c0 : sqrtf(c0 * c0 + c1 * c1 + c2 * c2))
                                                  : apply_chain(S.chain, c0, c1, c2, 0.f, D);
                const float4 col = lds_lut[S.lutSlot * T + tf_index(tv, T)];
                const float ca = col.w;
                if(S.weight > 0.0f)
                {
                    float a = ca * S.weight;
                    vr = vr + col.x * a;
                    vg = vg + col.y * a;
                    vb = vb + col.z * a;
                    va = va + a;
                }
                if(S.iso > 0.0f)
                {
                    float a = div_by(ca, S.iso, S.invIso) * 0.5f;
                    if(ia < a)
                    {
                        ir = col.x;
                        ig = col.y;
                        ib = col.z;
                        ia = a;
                    }
                }
            }
            if(P.totalWeight != 0.0f)
            {
                vr = div_by(vr, P.totalWeight, P.invTotalWeight);
                vg = div_by(vg, P.totalWeight, P.invTotalWeight);
                vb = div_by(vb, P.totalWeight, P.invTotalWeight);
                va = div_by(va, P.totalWeight, P.invTotalWeight);
            }
            else
            {
                vr = vg = vb = va = 0.0f;
            }
            texel[2 * k] = dither_b4(P.ditherMode, x + k, y, z, vr, vg, vb, va);
            const uint32_t q0 = unorm8(ir), q1 = unorm8(ig), q2 = unorm8(ib), q3 = unorm8(ia);
            texel[2 * k + 1] = q0 | (q1 << 8) | (q2 << 16) | (q3 << 24);
        }
        uint4* o = reinterpret_cast<uint4*>(P.tex + u);
        if(P.texH)
        {
            // the march steps over the binary16 copy and reads the combined texture only at
            // steps that can hit the iso-surface: its stores are non-temporal, so they do not
            // displace the copy from the caches
            typedef uint32_t v4u __attribute__((ext_vector_type(4)));
            __builtin_nontemporal_store(v4u{texel[0], texel[1], texel[2], texel[3]}, reinterpret_cast<v4u*>(o));
            __builtin_nontemporal_store(v4u{texel[4], texel[5], texel[6], texel[7]}, reinterpret_cast<v4u*>(o) + 1);
        }
        else
        {
            o[0] = make_uint4(texel[0], texel[1], texel[2], texel[3]);
            o[1] = make_uint4(texel[4], texel[5], texel[6], texel[7]);
        }
        if(P.texH)
        {
            uint2 h[4];
#pragma unroll
            for(int k = 0; k < 4; k++)
                h[k] = half_texel(texel[2 * k], texel[2 * k + 1]);
            uint4* oh = reinterpret_cast<uint4*>(P.texH + u);
            oh[0] = make_uint4(h[0].x, h[0].y, h[1].x, h[1].y);
            oh[1] = make_uint4(h[2].x, h[2].y, h[3].x, h[3].y);
        }
    }

    // Classification compiled for the signature of its active sources (component
    // counts 1 or 3, at most two sources, all transfer functions in LDS, x-contiguous
    // sources with rows of a multiple of 4 voxels, no advection texture): each
    // thread classifies 4 consecutive voxels of a row per iteration (every source's
    // values in 16-byte loads; requesting the next quad's values before classifying the
    // current one held 32 more VGPRs, 90 against 71, and was 11 % slower than the
    // occupancy that frees) and stores the 4 texels with two 16-byte stores.  Same
    // arithmetic, in the same order, as classify_kernel.  Workgroups of 512 threads:
    // eight waves share one LDS copy of the transfer functions, which halves the
    // staging reads against 256-thread workgroups and lets 8 waves per SIMD be
    // resident with two 1024-entry functions (32 KB per workgroup).
    template<int SIG>
    __global__ __launch_bounds__(512) void classify_sig_kernel(const ClassifyParams P)
    {
        extern __shared__ float4 lds_lut[];
        for(int i = threadIdx.x; i < P.nLut * P.T; i += blockDim.x)
            lds_lut[i] = P.luts[i];
        __syncthreads();
        constexpr int N = sig_count(SIG);
        const int64_t nq = P.nvox >> 2;
        const int64_t stride = int64_t(gridDim.x) * blockDim.x;
        int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        if(q >= nq)
            return;
        for(; q < nq; q += stride)
        {
            float f[N][16];
            int x, y, z;
            cls_fetch<SIG>(P, q, f, x, y, z);
            cls_emit<SIG>(P, lds_lut, f, x, y, z, q);
        }
    }


    // signature of the active sources for classify_sig_kernel, 0 when the generic kernel applies;
    // Pc receives the active sources in slot order
    int classify_signature(const ClassifyParams& P, ClassifyParams& Pc)
    {
        if(P.X % 4 != 0 || (reinterpret_cast<uintptr_t>(P.tex) & 15u) != 0)
            return 0;
        Pc = P;
        int n = 0, sig = 0, w = 1;
        for(int s = 0; s < P.nsrc; s++)
        {
            const ClassifySource& S = P.src[s];
            if(!S.active)
                continue;
            if(n == 2 || (S.dim != 1 && S.dim != 3) || S.adv || S.lutSlot >= P.nLut || S.stride[0] != 1 || S.stride[1] % 4 != 0
               || S.stride[2] % 4 != 0 || (reinterpret_cast<uintptr_t>(S.data) & 15u) != 0)
                return 0;
            Pc.src[n++] = S;
            sig += S.dim * w;
            w *= 5;
        }
        Pc.nsrc = n;
        return n > 0 ? sig : 0;
    }

    hipError_t launch_classify(const ClassifyParams& P, hipStream_t s)
    {
        if(P.nvox <= 0)
            return hipSuccess;
        const size_t lds = size_t(P.nLut) * P.T * sizeof(float4);
        ClassifyParams Pc;
        const int sig = classify_signature(P, Pc);
        if(sig)
        {
            int64_t blocks = (P.nvox / 4 + 511) / 512;
            if(blocks > 2048)
                blocks = 2048;
            const dim3 grid = dim3(unsigned(blocks)), block(512);
            bool natural = true;
            for(int k = 0; k < Pc.nsrc; k++)
                natural = natural && natural_chain(Pc.src[k].chain, Pc.src[k].dim);
#define ISAAC_CLS_SIG(sg)                                                                                            \
    case sg:                                                                                                         \
        if(natural)                                                                                                  \
            hipLaunchKernelGGL((classify_sig_kernel<sg + kSigNatural>), grid, block, lds, s, Pc);                     \
        else                                                                                                         \
            hipLaunchKernelGGL((classify_sig_kernel<sg>), grid, block, lds, s, Pc);                                   \
        return hipGetLastError();
            switch(sig)
            {
                ISAAC_CLS_SIG(1)
                ISAAC_CLS_SIG(3)
                ISAAC_CLS_SIG(6)
                ISAAC_CLS_SIG(8)
                ISAAC_CLS_SIG(16)
                ISAAC_CLS_SIG(18)
            default: break;
            }
#undef ISAAC_CLS_SIG
        }
        int64_t blocks = (P.nvox + 255) / 256;
        if(blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(classify_kernel, dim3(unsigned(blocks)), dim3(256), lds, s, P);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ field-source advection
    // halton (lib/isaac/isaac_helper.hpp:57-69)
    __device__ inline float halton(uint32_t index, uint32_t base)
    {
        float result = 0.0f, f = 1.0f;
        while(index > 0)
        {
            f = f / float(base);
            result = result + f * float(index % base);
            index = index / base;
        }
        return result;
    }

    // HaltonSeedingKernel (lib/isaac/isaac_common_kernel.hpp:334-348): point i = 1..count
    // at (halton(i,3), halton(i,5), halton(i,7)) * size, truncated, set to 1 (one
    // thread per point; points that land on one voxel all store 1)
    __global__ __launch_bounds__(256) void halton_seed_kernel(float* noise, int X, int Y, int Z, uint32_t count)
    {
        for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x + 1; i <= count; i += gridDim.x * blockDim.x)
        {
            const int x = int(float(X) * halton(i, 3)), y = int(float(Y) * halton(i, 5)), z = int(float(Z) * halton(i, 7));
            noise[(size_t(z) * Y + y) * X + x] = 1.0f;
        }
    }

    // Sampler<LINEAR, REPEAT> on a guard-less float texture (lib/isaac/isaac_texture.hpp:182-259)
    __device__ inline int wrap(int c, int n)
    {
        return (n + (c % n)) % n;
    }

    __device__ inline float sample_repeat(const float* t, int X, int Y, int Z, float px, float py, float pz)
    {
        const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
        const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
        const float fx = cx - flx, fy = cy - fly, fz = cz - flz;
        const int x0 = int(flx), y0 = int(fly), z0 = int(flz);
        auto v = [&](int dx, int dy, int dz) {
            return t[(size_t(wrap(z0 + dz, Z)) * Y + wrap(y0 + dy, Y)) * X + wrap(x0 + dx, X)];
        };
        const float oz = 1.0f - fz, oy = 1.0f - fy, ox = 1.0f - fx;
        const float a0 = v(0, 0, 0) * oz + v(0, 0, 1) * fz, a1 = v(0, 1, 0) * oz + v(0, 1, 1) * fz;
        const float A = a0 * oy + a1 * fy;
        const float b0 = v(1, 0, 0) * oz + v(1, 0, 1) * fz, b1 = v(1, 1, 0) * oz + v(1, 1, 1) * fz;
        const float B = b0 * oy + b1 * fy;
        return A * ox + B * fx;
    }

    // GaussBlur7Kernel (lib/isaac/isaac_common_kernel.hpp:381-411) along one axis
    // sample position coord + mask * k / scale, as the reference evaluates it
    __global__ __launch_bounds__(256) void gauss7_kernel(const float* src, float* dst, int X, int Y, int Z, float3 mask, float3 scale)
    {
        const float w[4] = {64.0f, 40.5f, 10.0f, 1.0f};
        const uint32_t n = uint32_t(X) * uint32_t(Y) * uint32_t(Z);
        for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        {
            const int x = int(i % uint32_t(X)), y = int((i / uint32_t(X)) % uint32_t(Y)), z = int(i / (uint32_t(X) * uint32_t(Y)));
            float r = 0.0f;
            for(int k = -3; k < 4; k++)
            {
                const float px = float(x) + mask.x * float(k) / scale.x, py = float(y) + mask.y * float(k) / scale.y,
                            pz = float(z) + mask.z * float(k) / scale.z;
                r = r + sample_repeat(src, X, Y, Z, px, py, pz) * w[k < 0 ? -k : k];
            }
            dst[i] = r / 167.0f;
        }
    }

    // MultiplyClampKernel (lib/isaac/isaac_common_kernel.hpp:413-429)
    __global__ __launch_bounds__(256) void multiply_clamp_kernel(const float* src, float* dst, int64_t n, float value, float lo, float hi)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
            dst[i] = fminf(fmaxf(src[i] * value, lo), hi);
    }

    hipError_t launch_noise(float* noise, float* tmp, int X, int Y, int Z, const float scale[3], uint32_t seedPoints, hipStream_t s)
    {
        const int64_t n = int64_t(X) * Y * Z;
        if(n <= 0)
            return hipSuccess;
        if(n >= (int64_t(1) << 32))
            return hipErrorInvalidValue;
        const unsigned grid = unsigned(std::min<int64_t>((n + 255) / 256, 8192));
        const float3 sc = make_float3(scale[0], scale[1], scale[2]);
        hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(noise), 0u, size_t(n), s);
        if(e != hipSuccess)
            return e;
        if(seedPoints > 0)
            hipLaunchKernelGGL(halton_seed_kernel, dim3(std::min<uint32_t>((seedPoints + 255) / 256, 4096)), dim3(256), 0, s,
                               noise, X, Y, Z, seedPoints);
        // three separable passes noise -> tmp (x) -> noise (y) -> tmp (z), then noise = clamp(tmp * 30, 0, 1)
        hipLaunchKernelGGL(gauss7_kernel, dim3(grid), dim3(256), 0, s, (const float*) noise, tmp, X, Y, Z, make_float3(1.0f, 0.0f, 0.0f), sc);
        hipLaunchKernelGGL(gauss7_kernel, dim3(grid), dim3(256), 0, s, (const float*) tmp, noise, X, Y, Z, make_float3(0.0f, 1.0f, 0.0f), sc);
        hipLaunchKernelGGL(gauss7_kernel, dim3(grid), dim3(256), 0, s, (const float*) noise, tmp, X, Y, Z, make_float3(0.0f, 0.0f, 1.0f), sc);
        hipLaunchKernelGGL(multiply_clamp_kernel, dim3(grid), dim3(256), 0, s, (const float*) tmp, noise, n, 30.0f, 0.0f, 1.0f);
        return hipGetLastError();
    }

    // Sampler<LINEAR, VALUE> on a byte texture whose guard (and beyond) reads 0
    __device__ inline float sample_bytes_zero(const uint8_t* t, int X, int Y, int Z, float px, float py, float pz)
    {
        const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
        const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
        const float fx = cx - flx, fy = cy - fly, fz = cz - flz;
        const int x0 = int(flx), y0 = int(fly), z0 = int(flz);
        auto v = [&](int dx, int dy, int dz) {
            const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
            if(x < 0 || y < 0 || z < 0 || x >= X || y >= Y || z >= Z)
                return 0.0f;
            return float(t[(size_t(z) * Y + y) * X + x]);
        };
        const float oz = 1.0f - fz, oy = 1.0f - fy, ox = 1.0f - fx;
        const float a0 = v(0, 0, 0) * oz + v(0, 0, 1) * fz, a1 = v(0, 1, 0) * oz + v(0, 1, 1) * fz;
        const float A = a0 * oy + a1 * fy;
        const float b0 = v(1, 0, 0) * oz + v(1, 0, 1) * fz, b1 = v(1, 1, 0) * oz + v(1, 1, 1) * fz;
        const float B = b0 * oy + b1 * fy;
        return A * ox + B * fx;
    }

    // GenerateAdvectionTextureKernel (lib/isaac/isaac_common_kernel.hpp:500-550): the
    // unit field direction moves the history back by stepSize / scale; the noise seeds
    // where the seeding window is open; blended value dithered with the 16^3 mask
    __global__ __launch_bounds__(256) void advect_kernel(const AdvectParams P)
    {
        const uint32_t n = uint32_t(P.X) * uint32_t(P.Y) * uint32_t(P.Z);
        for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        {
            const int x = int(i % uint32_t(P.X)), y = int((i / uint32_t(P.X)) % uint32_t(P.Y)), z = int(i / (uint32_t(P.X) * uint32_t(P.Y)));
            const float* e = P.field + (int64_t(x) * P.stride[0] + int64_t(y) * P.stride[1] + int64_t(z) * P.stride[2]) * 3;
            float vx = e[0], vy = e[1], vz = e[2];
            float len = sqrtf((vx * vx + vy * vy) + vz * vz);
            len = (len < FLT_MIN) ? FLT_MIN : len;
            vx = vx / len;
            vy = vy / len;
            vz = vz / len;
            const float ox = vx * P.stepSize / P.scale[0], oy = vy * P.stepSize / P.scale[1], oz = vz * P.stepSize / P.scale[2];
            const float px = (float(x) + 0.5f) - ox, py = (float(y) + 0.5f) - oy, pz = (float(z) + 0.5f) - oz;
            const float history = sample_bytes_zero(P.history, P.X, P.Y, P.Z, px, py, pz) / 255.0f;
            const float noise = P.seeding ? P.noise[i] : 0.0f;
            const float v = noise + history * P.historyWeight * (1.0f - noise);
            P.out[i] = uint8_t(dither_b4(5, x, y, z, v, 0.f, 0.f, 0.f) & 0xffu);
        }
    }

    hipError_t launch_advect(const AdvectParams& P, hipStream_t s)
    {
        const int64_t n = int64_t(P.X) * P.Y * P.Z;
        if(n <= 0)
            return hipSuccess;
        if(n >= (int64_t(1) << 32))
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(advect_kernel, dim3(unsigned(std::min<int64_t>((n + 255) / 256, 8192))), dim3(256), 0, s, P);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ min/max request
    // MinMaxKernel (lib/isaac/isaac_min_max_kernel.hpp:22-57) as a grid-stride
    // reduction over the voxels; the visiting-order key of voxel (x, y, z) is
    // (x + y*X) * Z + z + 1 (columns in the host loop's order, z inside a column).
    __global__ __launch_bounds__(256) void min_max_volume_kernel(const MinMaxVolumeParams P, MinMaxPartial* partials)
    {
        MinMaxPartial a{FLT_MAX, -FLT_MAX, 0, 0};
        const int64_t stride = int64_t(gridDim.x) * blockDim.x;
        for(int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < P.nvox; idx += stride)
        {
            const uint32_t u = (uint32_t) idx;
            const uint32_t zq = fastdiv(u, P.divXY_mul, P.divXY_shift);
            const uint32_t rem = u - zq * uint32_t(P.X * P.Y);
            const uint32_t yq = fastdiv(rem, P.divX_mul, P.divX_shift);
            const int x = int(rem - yq * uint32_t(P.X)), y = int(yq), z = int(zq);
            const float* e = P.data + (int64_t(x) * P.stride[0] + int64_t(y) * P.stride[1] + int64_t(z) * P.stride[2]) * P.dim;
            float v = e[0];
            if(P.applyChain)
                v = apply_chain(P.chain, e[0], P.dim > 1 ? e[1] : 0.f, P.dim > 2 ? e[2] : 0.f, P.dim > 3 ? e[3] : 0.f, P.dim);
            const uint64_t key = (uint64_t(x) + uint64_t(y) * uint64_t(P.X)) * uint64_t(P.Z) + uint64_t(z) + 1;
            mm_take_min(a.vmin, a.kmin, v, key);
            mm_take_max(a.vmax, a.kmax, v, key);
        }
        a = mm_block_reduce(a);
        if(threadIdx.x == 0)
            partials[blockIdx.x] = a;
    }

    __global__ __launch_bounds__(256) void min_max_final_kernel(const MinMaxPartial* partials, int count, float2* out)
    {
        MinMaxPartial a{FLT_MAX, -FLT_MAX, 0, 0};
        for(int i = threadIdx.x; i < count; i += blockDim.x)
        {
            mm_take_min(a.vmin, a.kmin, partials[i].vmin, partials[i].kmin);
            mm_take_max(a.vmax, a.kmax, partials[i].vmax, partials[i].kmax);
        }
        a = mm_block_reduce(a);
        if(threadIdx.x == 0)
            *out = make_float2(a.vmin, a.vmax);
    }

    hipError_t launch_min_max_final(const MinMaxPartial* partials, int count, float2* out, hipStream_t s)
    {
        hipLaunchKernelGGL(min_max_final_kernel, dim3(1), dim3(256), 0, s, partials, count, out);
        return hipGetLastError();
    }

    hipError_t launch_min_max_volume(const MinMaxVolumeParams& P, MinMaxPartial* partials, float2* out, hipStream_t s)
    {
        if(P.nvox <= 0 || P.nvox >= (int64_t(1) << 32))
            return hipErrorInvalidValue;
        const int groups = int(std::min<int64_t>((P.nvox + 255) / 256, kMinMaxGroups));
        hipLaunchKernelGGL(min_max_volume_kernel, dim3(groups), dim3(256), 0, s, P, partials);
        hipError_t e = hipGetLastError();
        return e != hipSuccess ? e : launch_min_max_final(partials, groups, out, s);
    }

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
        hipStream_t s)
    {
        int64_t n = int64_t(gX) * gY * gZ;
        int64_t blocks = (n + 255) / 256;
        if(blocks > 4096)
            blocks = 4096;
        if(blocks < 1)
            return hipSuccess;
        hipLaunchKernelGGL(
            persistent_copy_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, src, stride[0], stride[1], stride[2], dim,
            chain, gX, gY, gZ, guard, dst);
        return hipGetLastError();
    }

    // the trilinear march over the binary16 copy (texture layout 16), at every schedule
    template<bool OFF32>
    void launch_march_half(const FrameParams& F, const ParticleParams& PP, bool counting, bool prefetch, dim3 grid, dim3 block,
                           hipStream_t s)
    {
        if(counting)
            hipLaunchKernelGGL((march_kernel<true, true, false, false, OFF32, true>), grid, block, 0, s, F, PP);
        else if(prefetch)
            hipLaunchKernelGGL((march_kernel<true, false, true, false, OFF32, true>), grid, block, 0, s, F, PP);
        else if(F.waves == 6)
            hipLaunchKernelGGL((march_kernel_w6<true, false, OFF32, true>), grid, block, 0, s, F, PP);
        else if(F.waves == 8)
            hipLaunchKernelGGL((march_kernel_w8<true, false, OFF32, true>), grid, block, 0, s, F, PP);
        else
            hipLaunchKernelGGL((march_kernel<true, false, false, false, OFF32, true>), grid, block, 0, s, F, PP);
    }

    template<bool BRICK>
    hipError_t launch_march_layout(const FrameParams& F, const ParticleParams& PP, bool counting, hipStream_t s)
    {
        const unsigned nb = march_grid(F);
        if(nb == 0)
            return hipSuccess;
        // one-wave workgroups: 4 per logical workgroup, the logical grid rounded to 8
        const dim3 grid(F.waveGroups == 1 ? (nb + 7u) / 8u * 32u : nb), block(F.waveGroups == 1 ? 64 : 256);
        // next-step prefetch is compiled for the linear layout only
        const bool prefetch = F.prefetch && !BRICK;
        if constexpr(!BRICK)
        {
            if(F.texH && F.linear)
            {
                if(F.texOff32)
                    launch_march_half<true>(F, PP, counting, prefetch, grid, block, s);
                else
                    launch_march_half<false>(F, PP, counting, prefetch, grid, block, s);
                return hipGetLastError();
            }
        }
        if(!counting && !prefetch && (F.waves == 6 || F.waves == 8))
        {
            if(F.waves == 6)
            {
                if(F.linear && F.texOff32)
                    hipLaunchKernelGGL((march_kernel_w6<true, BRICK, true>), grid, block, 0, s, F, PP);
                else if(F.linear)
                    hipLaunchKernelGGL((march_kernel_w6<true, BRICK, false>), grid, block, 0, s, F, PP);
                else
                    hipLaunchKernelGGL((march_kernel_w6<false, BRICK, false>), grid, block, 0, s, F, PP);
            }
            else
            {
                if(F.linear && F.texOff32)
                    hipLaunchKernelGGL((march_kernel_w8<true, BRICK, true>), grid, block, 0, s, F, PP);
                else if(F.linear)
                    hipLaunchKernelGGL((march_kernel_w8<true, BRICK, false>), grid, block, 0, s, F, PP);
                else
                    hipLaunchKernelGGL((march_kernel_w8<false, BRICK, false>), grid, block, 0, s, F, PP);
            }
            return hipGetLastError();
        }
        const int v = (F.linear ? 4 : 0) | (counting ? 2 : 0) | (prefetch ? 1 : 0);
        switch(v)
        {
        case 0: hipLaunchKernelGGL((march_kernel<false, false, false, BRICK>), grid, block, 0, s, F, PP); break;
        case 2: hipLaunchKernelGGL((march_kernel<false, true, false, BRICK>), grid, block, 0, s, F, PP); break;
        case 4:
            if(F.texOff32)
                hipLaunchKernelGGL((march_kernel<true, false, false, BRICK, true>), grid, block, 0, s, F, PP);
            else
                hipLaunchKernelGGL((march_kernel<true, false, false, BRICK, false>), grid, block, 0, s, F, PP);
            break;
        case 6: hipLaunchKernelGGL((march_kernel<true, true, false, BRICK>), grid, block, 0, s, F, PP); break;
        default:
            if constexpr(!BRICK)
            {
                if(v == 1)
                    hipLaunchKernelGGL((march_kernel<false, false, true, false>), grid, block, 0, s, F, PP);
                else if(v == 3)
                    hipLaunchKernelGGL((march_kernel<false, true, true, false>), grid, block, 0, s, F, PP);
                else if(v == 5 && F.texOff32)
                    hipLaunchKernelGGL((march_kernel<true, false, true, false, true>), grid, block, 0, s, F, PP);
                else if(v == 5)
                    hipLaunchKernelGGL((march_kernel<true, false, true, false>), grid, block, 0, s, F, PP);
                else
                    hipLaunchKernelGGL((march_kernel<true, true, true, false>), grid, block, 0, s, F, PP);
            }
            break;
        }
        return hipGetLastError();
    }

    hipError_t launch_march(const FrameParams& F, const ParticleParams& PP, bool counting, hipStream_t s)
    {
        return F.texB ? launch_march_layout<true>(F, PP, counting, s) : launch_march_layout<false>(F, PP, counting, s);
    }

    // Linear texture -> brick layout (FrameParams::texB): one workgroup per brick, the
    // 729 texels of brick b (apron and parts beyond the volume take the clamped texel)
    // written in order.
    __global__ __launch_bounds__(256) void tex_brick_kernel(
        const uint2* __restrict__ tex,
        uint2* __restrict__ out,
        int X,
        int Y,
        int Z,
        uint32_t nbx,
        uint32_t nby)
    {
        const uint32_t b = blockIdx.x;
        const uint32_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
        for(uint32_t t = threadIdx.x; t < uint32_t(kTexBrickTexels); t += 256u)
        {
            const uint32_t lz = t / uint32_t(kTexBrickSlice), ly = (t / uint32_t(kTexBrickRow)) % uint32_t(kTexBrickRow);
            const uint32_t lx = t % uint32_t(kTexBrickRow);
            const uint32_t x = min(bx * uint32_t(kTexBrick) + lx, uint32_t(X - 1));
            const uint32_t y = min(by * uint32_t(kTexBrick) + ly, uint32_t(Y - 1));
            const uint32_t z = min(bz * uint32_t(kTexBrick) + lz, uint32_t(Z - 1));
            out[size_t(b) * kTexBrickTexels + t] = tex[x + size_t(X) * (y + size_t(Y) * z)];
        }
    }

    hipError_t launch_tex_bricks(const uint2* tex, uint2* out, int X, int Y, int Z, hipStream_t s)
    {
        const uint32_t nbx = uint32_t(X + kTexBrick - 1) / kTexBrick, nby = uint32_t(Y + kTexBrick - 1) / kTexBrick;
        const uint32_t nbz = uint32_t(Z + kTexBrick - 1) / kTexBrick;
        if(nbx * nby * nbz == 0)
            return hipSuccess;
        hipLaunchKernelGGL(tex_brick_kernel, dim3(nbx * nby * nbz), dim3(256), 0, s, tex, out, X, Y, Z, nbx, nby);
        return hipGetLastError();
    }

    bool ssao_shades(const FrameParams& F)
    {
        // render modes 0, 1 and 3 leave the depth buffer as the march wrote it
        return F.aoEnabled && (F.renderMode == 0 || F.renderMode == 1 || F.renderMode == 3);
    }

    hipError_t launch_ssao(const FrameParams& F, hipStream_t s, bool shade)
    {
        const unsigned bw = unsigned(F.vx1 - F.vx0 + kSsaoBW - 1) / kSsaoBW;
        const unsigned bh = unsigned(F.vy1 - F.vy0 + kSsaoBH - 1) / kSsaoBH;
        if(bw * bh == 0)
            return hipSuccess;
        if(shade)
            hipLaunchKernelGGL(ssao_tiled_kernel<true>, dim3(bw * bh), dim3(256), 0, s, F);
        else
            hipLaunchKernelGGL(ssao_tiled_kernel<false>, dim3(bw * bh), dim3(256), 0, s, F);
        return hipGetLastError();
    }

    hipError_t launch_shade(const FrameParams& F, hipStream_t s)
    {
        const unsigned nb = unsigned(F.tilesX) * unsigned(F.tilesY);
        if(nb == 0)
            return hipSuccess;
        hipLaunchKernelGGL(shade_kernel, dim3(nb), dim3(256), 0, s, F);
        return hipGetLastError();
    }
} // namespace isaac_hip
