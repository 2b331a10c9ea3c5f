// gfx950 kernels of the per-rank render path.
//
//   classify_kernel        MergeToCombinedTextureKernel  lib/isaac/isaac_common_kernel.hpp:553-615,690-758
//   solid_kernel           ClearBufferKernel + ParticleRenderKernel + CombinedIsoRenderKernel, fused per pixel
//                          lib/isaac/isaac_common_kernel.hpp:195-217, lib/isaac/isaac_particle_kernel.hpp:117-260,
//                          lib/isaac/isaac_iso_kernel.hpp:25-128
//   ssao_kernel            SSAOKernel (live 7x7 part)     lib/isaac/isaac_ssao_kernel.hpp:161-182
//   shade_volume_kernel    ShadingKernel + CombinedVolumeRenderKernel, fused per pixel
//                          lib/isaac/isaac_common_kernel.hpp:220-320, lib/isaac/isaac_volume_kernel.hpp:23-88
//
// The fusions keep the reference's quantisation points: every stage that the
// reference stores as RGBA8 in the G-buffer is rounded to RGBA8 in registers
// at the same point, so the fused kernels produce the stage sequence's bytes.
//
// Launch shape: 256-thread workgroups = four wave64s, each wave an 8x8 pixel
// tile (16x16 pixels per workgroup), so the 64 rays of a wave sample
// neighbouring voxels.  Workgroups are remapped so that each of the 8 XCDs
// (each with its own L2) renders a contiguous band of tiles.
#include "kernels.hpp"

namespace isaac_hip
{
    __constant__ uint8_t kDitherMask4[64]
        = {17, 46, 55, 1,  57, 7,  38, 31, 9,  52, 2,  54, 36, 26, 19, 44, 53, 20, 47, 22, 29, 60,
           62, 12, 45, 14, 26, 33, 8,  39, 11, 63, 13, 56, 6,  50, 40, 32, 23, 48, 21, 42, 51, 5,
           61, 3,  34, 27, 41, 10, 30, 37, 4,  35, 15, 59, 49, 24, 43, 18, 25, 64, 58, 16};

    // ------------------------------------------------------------------ functor chain
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
                        v[i] = powf(v[i], c.p[k][i]);
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

    __device__ inline int tf_index(float v, int T)
    {
        float lv = roundf(v * float(T));
        if(lv != lv)
            return 0;
        return (int) gclamp(lv, 0.0f, float(T - 1));
    }

    // ------------------------------------------------------------------ classification
    __device__ inline uint32_t dither_b4(int mode, int x, int y, int z, float r, float g, float b, float a)
    {
        if(mode == 0)
        {
            uint32_t q0 = (uint32_t) roundf(gclamp(r, 0.0f, 1.0f) * 255.0f);
            uint32_t q1 = (uint32_t) roundf(gclamp(g, 0.0f, 1.0f) * 255.0f);
            uint32_t q2 = (uint32_t) roundf(gclamp(b, 0.0f, 1.0f) * 255.0f);
            uint32_t q3 = (uint32_t) roundf(gclamp(a, 0.0f, 1.0f) * 255.0f);
            return q0 | (q1 << 8) | (q2 << 16) | (q3 << 24);
        }
        const float d = float(kDitherMask4[(x & 3) + (y & 3) * 4 + (z & 3) * 16]) / 64.0f;
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
                if(S.weight > 0.0f)
                {
                    float a = col.w * S.weight;
                    vr = vr + col.x * a;
                    vg = vg + col.y * a;
                    vb = vb + col.z * a;
                    va = va + a;
                }
                if(S.iso > 0.0f)
                {
                    float a = col.w / S.iso * 0.5f;
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
                vr = vr / P.totalWeight;
                vg = vg / P.totalWeight;
                vb = vb / P.totalWeight;
                va = va / P.totalWeight;
            }
            else
            {
                vr = vg = vb = va = 0.0f;
            }
            P.volTex[idx] = dither_b4(P.ditherMode, x, y, z, vr, vg, vb, va);
            uint32_t q0 = (uint32_t) roundf(gclamp(ir, 0.0f, 1.0f) * 255.0f);
            uint32_t q1 = (uint32_t) roundf(gclamp(ig, 0.0f, 1.0f) * 255.0f);
            uint32_t q2 = (uint32_t) roundf(gclamp(ib, 0.0f, 1.0f) * 255.0f);
            uint32_t q3 = (uint32_t) roundf(gclamp(ia, 0.0f, 1.0f) * 255.0f);
            P.isoTex[idx] = q0 | (q1 << 8) | (q2 << 16) | (q3 << 24);
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
    struct F4
    {
        float x, y, z, w;
    };

    __device__ inline F4 unpack4(uint32_t v)
    {
        return F4{float(v & 0xffu), float((v >> 8) & 0xffu), float((v >> 16) & 0xffu), float(v >> 24)};
    }

    // Sampler<LINEAR|NEAREST, CLAMP, normalized> on RGBA8 (lib/isaac/isaac_texture.hpp:135-306)
    template<bool LINEAR>
    __device__ inline F4 sample4(const uint32_t* __restrict__ tex, int X, int Y, int Z, float px, float py, float pz)
    {
        F4 r;
        if(!LINEAR)
        {
            const int x = iclamp(int(px), 0, X - 1), y = iclamp(int(py), 0, Y - 1), z = iclamp(int(pz), 0, Z - 1);
            r = unpack4(tex[uint32_t(x) + uint32_t(X) * (uint32_t(y) + uint32_t(Y) * uint32_t(z))]);
        }
        else
        {
            const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
            const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
            const float fx = cx - flx, fy = cy - fly, fz = cz - flz;
            const int bx = int(flx), by = int(fly), bz = int(flz);
            const uint32_t x0 = iclamp(bx, 0, X - 1), x1 = iclamp(bx + 1, 0, X - 1);
            const uint32_t y0 = iclamp(by, 0, Y - 1), y1 = iclamp(by + 1, 0, Y - 1);
            const uint32_t z0 = iclamp(bz, 0, Z - 1), z1 = iclamp(bz + 1, 0, Z - 1);
            const uint32_t r00 = uint32_t(X) * (y0 + uint32_t(Y) * z0), r01 = uint32_t(X) * (y0 + uint32_t(Y) * z1);
            const uint32_t r10 = uint32_t(X) * (y1 + uint32_t(Y) * z0), r11 = uint32_t(X) * (y1 + uint32_t(Y) * z1);
            const uint32_t t000 = tex[x0 + r00], t001 = tex[x0 + r01], t010 = tex[x0 + r10], t011 = tex[x0 + r11];
            const uint32_t t100 = tex[x1 + r00], t101 = tex[x1 + r01], t110 = tex[x1 + r10], t111 = tex[x1 + r11];
            const float oz = 1.0f - fz, oy = 1.0f - fy, ox = 1.0f - fx;
            auto lin = [](float o, float t, float a, float b) { return a * o + b * t; };
            auto tri = [&](int sh) -> float
            {
                const float d000 = float((t000 >> sh) & 0xffu), d001 = float((t001 >> sh) & 0xffu);
                const float d010 = float((t010 >> sh) & 0xffu), d011 = float((t011 >> sh) & 0xffu);
                const float d100 = float((t100 >> sh) & 0xffu), d101 = float((t101 >> sh) & 0xffu);
                const float d110 = float((t110 >> sh) & 0xffu), d111 = float((t111 >> sh) & 0xffu);
                const float a0 = lin(oz, fz, d000, d001), a1 = lin(oz, fz, d010, d011);
                const float A = lin(oy, fy, a0, a1);
                const float b0 = lin(oz, fz, d100, d101), b1 = lin(oz, fz, d110, d111);
                const float B = lin(oy, fy, b0, b1);
                return lin(ox, fx, A, B);
            };
            r = F4{tri(0), tri(8), tri(16), tri(24)};
        }
        return F4{r.x / 255.0f, r.y / 255.0f, r.z / 255.0f, r.w / 255.0f};
    }

    // alpha channel only (the iso march); identical arithmetic on channel 3
    template<bool LINEAR>
    __device__ inline float sample_alpha(const uint32_t* __restrict__ tex, int X, int Y, int Z, float px, float py, float pz)
    {
        float r;
        if(!LINEAR)
        {
            const int x = iclamp(int(px), 0, X - 1), y = iclamp(int(py), 0, Y - 1), z = iclamp(int(pz), 0, Z - 1);
            r = float(tex[uint32_t(x) + uint32_t(X) * (uint32_t(y) + uint32_t(Y) * uint32_t(z))] >> 24);
        }
        else
        {
            const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
            const float flx = floorf(cx), fly = floorf(cy), flz = floorf(cz);
            const float fx = cx - flx, fy = cy - fly, fz = cz - flz;
            const int bx = int(flx), by = int(fly), bz = int(flz);
            const uint32_t x0 = iclamp(bx, 0, X - 1), x1 = iclamp(bx + 1, 0, X - 1);
            const uint32_t y0 = iclamp(by, 0, Y - 1), y1 = iclamp(by + 1, 0, Y - 1);
            const uint32_t z0 = iclamp(bz, 0, Z - 1), z1 = iclamp(bz + 1, 0, Z - 1);
            const uint32_t r00 = uint32_t(X) * (y0 + uint32_t(Y) * z0), r01 = uint32_t(X) * (y0 + uint32_t(Y) * z1);
            const uint32_t r10 = uint32_t(X) * (y1 + uint32_t(Y) * z0), r11 = uint32_t(X) * (y1 + uint32_t(Y) * z1);
            const float d000 = float(tex[x0 + r00] >> 24), d001 = float(tex[x0 + r01] >> 24);
            const float d010 = float(tex[x0 + r10] >> 24), d011 = float(tex[x0 + r11] >> 24);
            const float d100 = float(tex[x1 + r00] >> 24), d101 = float(tex[x1 + r01] >> 24);
            const float d110 = float(tex[x1 + r10] >> 24), d111 = float(tex[x1 + r11] >> 24);
            const float oz = 1.0f - fz, oy = 1.0f - fy, ox = 1.0f - fx;
            const float a0 = d000 * oz + d001 * fz, a1 = d010 * oz + d011 * fz;
            const float A = a0 * oy + a1 * fy;
            const float b0 = d100 * oz + d101 * fz, b1 = d110 * oz + d111 * fz;
            const float B = b0 * oy + b1 * fy;
            r = A * ox + B * fx;
        }
        return r / 255.0f;
    }

    // ------------------------------------------------------------------ pixel tiling
    // workgroup -> 16x16 tile, XCD-banded; returns false for threads outside the viewport
    __device__ inline bool tile_pixel(const FrameParams& F, int& px, int& py)
    {
        const uint32_t nb = uint32_t(F.tilesX) * uint32_t(F.tilesY);
        uint32_t b = blockIdx.x;
        const uint32_t full = nb & ~7u;
        if(b < full)
            b = (b & 7u) * (full >> 3) + (b >> 3);
        const uint32_t tx = b % uint32_t(F.tilesX), ty = b / uint32_t(F.tilesX);
        const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
        px = F.vx0 + int(tx * 16u + (wave & 1u) * 8u + (lane & 7u));
        py = F.vy0 + int(ty * 16u + (wave >> 1) * 8u + (lane >> 3));
        return px < F.vx1 && py < F.vy1;
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
            unsigned long long s = red[0] + red[1] + red[2] + red[3];
            if(s)
                atomicAdd(dst, s);
        }
        __syncthreads();
    }

    // ------------------------------------------------------------------ solid pass
    template<bool LINEAR, bool COUNT>
    __global__ __launch_bounds__(256) void solid_kernel(const FrameParams F, const ParticleParams PP)
    {
        int px, py;
        const bool valid = tile_pixel(F, px, py);
        unsigned long long nCells = 0, nTests = 0, nIso = 0, nIsoHits = 0, nIsoRays = 0, nPartRays = 0;
        if(valid)
        {
            const size_t p = size_t(py) * size_t(F.W) + size_t(px);
            uint32_t color = F.bgClear;
            F3 normal = f3(0.f, 0.f, 0.f);
            float depth = FLT_MAX;
            Ray ray = pixel_to_ray(F, px, py);
            const bool inside = clip_ray(F, ray);
            // ---- particles (ParticleRenderKernel): 3D-DDA over the particle cell grid
            if(inside && PP.nsrc > 0)
            {
                nPartRays++;
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
                while(cell[0] < uint32_t(PP.grid[0]) && cell[1] < uint32_t(PP.grid[1]) && cell[2] < uint32_t(PP.grid[2])
                      && !hit && tested <= rayLength)
                {
                    if(COUNT)
                        nCells++;
                    const int cid = int(cell[0] + uint32_t(PP.grid[0]) * (cell[1] + uint32_t(PP.grid[1]) * cell[2]));
                    const F3 cf = f3(float(cell[0]), float(cell[1]), float(cell[2]));
                    for(int s = 0; s < PP.nsrc; s++)
                    {
                        const ParticleSourceDev& S = PP.src[s];
                        const int k0 = S.voff[cid], k1 = S.voff[cid + 1];
                        for(int k = k0; k < k1; k++)
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
                    if(t[0] < t[1] && t[0] < t[2])
                    {
                        cell[0] += uint32_t(ds[0]);
                        tested = t[0];
                        t[0] = t[0] + dt[0];
                    }
                    else if(t[1] < t[0] && t[1] < t[2])
                    {
                        cell[1] += uint32_t(ds[1]);
                        tested = t[1];
                        t[1] = t[1] + dt[1];
                    }
                    else
                    {
                        cell[2] += uint32_t(ds[2]);
                        tested = t[2];
                        t[2] = t[2] + dt[2];
                    }
                }
                if(hit)
                {
                    normal = normalize3(pn);
                    color = to_byte4(pcol.x, pcol.y, pcol.z, 1.0f);
                    depth = pdepth + ray.startDepth;
                }
            }
            // ---- iso (CombinedIsoRenderKernel)
            if(inside && F.isoEnabled)
            {
                const float endDepth = gmin(ray.endDepth, depth);
                if(endDepth > ray.startDepth)
                {
                    nIsoRays++;
                    const F3 sc = f3(F.scale[0], F.scale[1], F.scale[2]);
                    const float stepU = F.step * (len3(ray.dir) / len3(divv(ray.dir, sc)));
                    const int startSteps = int(ceilf(ray.startDepth / stepU));
                    const int endSteps = int(floorf(endDepth / stepU));
                    const F3 stepVec = divv(muls(ray.dir, stepU), sc);
                    const F3 startU = divv(ray.start, sc);
                    float oldValue = 0.0f;
                    for(int i = startSteps; i <= endSteps; i++)
                    {
                        const float fi = float(i);
                        const float qx = startU.x + stepVec.x * fi;
                        const float qy = startU.y + stepVec.y * fi;
                        const float qz = startU.z + stepVec.z * fi;
                        const float value = sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, qx, qy, qz);
                        if(COUNT)
                            nIso++;
                        const float tmp = oldValue;
                        oldValue = value;
                        if(value < 0.5f)
                            continue;
                        const bool first = ray.isClipped && i == startSteps;
                        const float tt = fi * stepU;
                        float d;
                        if(first)
                            d = ray.startDepth;
                        else
                            d = tt + stepU * (0.5f - tmp) / (value - tmp);
                        const F3 pu = divv(add(ray.start, muls(ray.dir, d)), sc);
                        F4 hc = sample4<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x, pu.y, pu.z);
                        F3 g;
                        g.x = sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x + 1.0f, pu.y, pu.z)
                            - sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x - 1.0f, pu.y, pu.z);
                        g.y = sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x, pu.y + 1.0f, pu.z)
                            - sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x, pu.y - 1.0f, pu.z);
                        g.z = sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x, pu.y, pu.z + 1.0f)
                            - sample_alpha<LINEAR>(F.isoTex, F.X, F.Y, F.Z, pu.x, pu.y, pu.z - 1.0f);
                        float gl = len3(g);
                        if(first || gl == 0.0f)
                        {
                            g = ray.clipN;
                            gl = 1.0f;
                        }
                        normal = f3(-g.x / gl, -g.y / gl, -g.z / gl);
                        color = to_byte4(hc.x, hc.y, hc.z, 1.0f);
                        depth = d;
                        if(COUNT)
                            nIsoHits++;
                        break;
                    }
                }
            }
            F.color[p] = color;
            F.depth[p] = depth;
            F.normal[3 * p + 0] = normal.x;
            F.normal[3 * p + 1] = normal.y;
            F.normal[3 * p + 2] = normal.z;
            if(!F.aoEnabled)
                F.ao[p] = 0.0f;
        }
        if(COUNT)
        {
            block_count_add(F.counters + 1, nIso);
            block_count_add(F.counters + 2, nIsoHits);
            block_count_add(F.counters + 3, nCells);
            block_count_add(F.counters + 4, nTests);
            block_count_add(F.counters + 6, nIsoRays);
            block_count_add(F.counters + 7, nPartRays);
        }
    }

    // ------------------------------------------------------------------ SSAO
    __global__ __launch_bounds__(256) void ssao_kernel(const FrameParams F)
    {
        int px, py;
        if(!tile_pixel(F, px, py))
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

    // ------------------------------------------------------------------ shading + volume
    template<bool LINEAR, bool COUNT>
    __global__ __launch_bounds__(256) void shade_volume_kernel(const FrameParams F)
    {
        int px, py;
        const bool valid = tile_pixel(F, px, py);
        unsigned long long nSamples = 0, nRays = 0;
        if(valid)
        {
            const size_t p = size_t(py) * size_t(F.W) + size_t(px);
            const uint32_t cb = F.color[p];
            F4 color = F4{float(cb & 0xffu) / 255.0f, float((cb >> 8) & 0xffu) / 255.0f,
                          float((cb >> 16) & 0xffu) / 255.0f, float(cb >> 24) / 255.0f};
            const F3 normal = f3(F.normal[3 * p + 0], F.normal[3 * p + 1], F.normal[3 * p + 2]);
            float depth = F.depth[p];
            const float ao = F.aoEnabled ? F.ao[p] : 0.0f;
            const float aoStrength = 1.0f - ao;
            const int mode = F.renderMode;
            Ray ray = pixel_to_ray(F, px, py);
            uint32_t solid = cb;
            bool depthChanged = false;
            if(mode == 7)
                color = F4{F.rankHue[0], F.rankHue[1], F.rankHue[2], color.w};
            if(mode < 3 || mode == 7)
            {
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
            if(F.volumeEnabled && clip_ray(F, ray))
            {
                const float endDepth = gmin(ray.endDepth, depth);
                if(endDepth > ray.startDepth)
                {
                    nRays++;
                    const F3 sc = f3(F.scale[0], F.scale[1], F.scale[2]);
                    const float stepU = F.step * (len3(ray.dir) / len3(divv(ray.dir, sc)));
                    const float factor = stepU / F.minSize * 2.0f * F.totalWeightAll;
                    const int startSteps = int(ceilf(ray.startDepth / stepU));
                    const int endSteps = int(floorf(endDepth / stepU));
                    const F3 stepVec = divv(muls(ray.dir, stepU), sc);
                    const F3 startU = divv(ray.start, sc);
                    float cr = 0.f, cg = 0.f, cbl = 0.f, ca = 0.f;
                    for(int i = startSteps; i <= endSteps; i++)
                    {
                        const float fi = float(i);
                        const float qx = startU.x + stepVec.x * fi;
                        const float qy = startU.y + stepVec.y * fi;
                        const float qz = startU.z + stepVec.z * fi;
                        F4 v = sample4<LINEAR>(F.volTex, F.X, F.Y, F.Z, qx, qy, qz);
                        if(COUNT)
                            nSamples++;
                        v.x = v.x * factor;
                        v.y = v.y * factor;
                        v.z = v.z * factor;
                        v.w = v.w * factor;
                        const float w = 1.0f - ca;
                        cr = cr + w * v.x;
                        cg = cg + w * v.y;
                        cbl = cbl + w * v.z;
                        ca = ca + w * v.w;
                        if(ca > 0.99f)
                            break;
                    }
                    const float sr = float(solid & 0xffu) / 255.0f, sg = float((solid >> 8) & 0xffu) / 255.0f;
                    const float sb = float((solid >> 16) & 0xffu) / 255.0f, sa = float(solid >> 24) / 255.0f;
                    const float om = 1.0f - ca;
                    out = to_byte4(cr + om * sr, cg + om * sg, cbl + om * sb, ca + om * sa);
                }
            }
            F.color[p] = out;
            if(depthChanged)
                F.depth[p] = depth;
        }
        if(COUNT)
        {
            block_count_add(F.counters + 0, nSamples);
            block_count_add(F.counters + 5, nRays);
        }
    }

    // ------------------------------------------------------------------ launchers
    hipError_t launch_classify(const ClassifyParams& P, hipStream_t s)
    {
        if(P.nvox <= 0)
            return hipSuccess;
        const size_t lds = size_t(P.nLut) * P.T * sizeof(float4);
        int64_t blocks = (P.nvox + 255) / 256;
        if(blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(classify_kernel, dim3(unsigned(blocks)), dim3(256), lds, s, P);
        return hipGetLastError();
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

    hipError_t launch_frame(const FrameParams& F, const ParticleParams& PP, bool counting, hipStream_t s)
    {
        const unsigned nb = unsigned(F.tilesX) * unsigned(F.tilesY);
        if(nb == 0)
            return hipSuccess;
        const dim3 grid(nb), block(256);
        if(F.linear)
        {
            if(counting)
                hipLaunchKernelGGL((solid_kernel<true, true>), grid, block, 0, s, F, PP);
            else
                hipLaunchKernelGGL((solid_kernel<true, false>), grid, block, 0, s, F, PP);
        }
        else
        {
            if(counting)
                hipLaunchKernelGGL((solid_kernel<false, true>), grid, block, 0, s, F, PP);
            else
                hipLaunchKernelGGL((solid_kernel<false, false>), grid, block, 0, s, F, PP);
        }
        if(F.aoEnabled)
            hipLaunchKernelGGL(ssao_kernel, grid, block, 0, s, F);
        if(F.linear)
        {
            if(counting)
                hipLaunchKernelGGL((shade_volume_kernel<true, true>), grid, block, 0, s, F);
            else
                hipLaunchKernelGGL((shade_volume_kernel<true, false>), grid, block, 0, s, F);
        }
        else
        {
            if(counting)
                hipLaunchKernelGGL((shade_volume_kernel<false, true>), grid, block, 0, s, F);
            else
                hipLaunchKernelGGL((shade_volume_kernel<false, false>), grid, block, 0, s, F);
        }
        return hipGetLastError();
    }
} // namespace isaac_hip
