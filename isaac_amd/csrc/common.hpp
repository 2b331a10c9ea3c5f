// Shared host/device definitions of the gfx950 render path.
//
// All arithmetic helpers spell out the association order of the reference's
// glm expressions; the library is compiled with -ffp-contract=off so each
// + - * / below is exactly one IEEE-754 binary32 operation, as on the
// reference's CPU path.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace isaac_hip
{
    constexpr int kMaxSources = 16;
    constexpr int kMaxAdvectionStep = 3; // ISAAC_MAX_ADVECTION_STEP_SIZE (lib/isaac/isaac_defines.hpp:21)
    constexpr int kMaxFunctors = 3;
    constexpr int kMaxClipping = 16;
    constexpr float kZNear = 0.002f; // ISAAC_Z_NEAR lib/isaac/isaac_defines.hpp:53
    // added to a source signature of a compiled march: every source's functor chain is
    // its natural one (none for a scalar source, one `length` for a vector source),
    // evaluated inline in the interpreter's operation order instead of interpreted
    constexpr int kSigNatural = 1000;
    constexpr int kTexBrick = 8; // side of a texture brick (texture layout 8)
    constexpr int kTexHalf = 16; // texture layout 16: the march steps over a binary16 copy of the volume RGBA
    constexpr int kTexBrickRow = kTexBrick + 1, kTexBrickSlice = kTexBrickRow * kTexBrickRow;
    constexpr int kTexBrickTexels = kTexBrickSlice * kTexBrickRow; // 729

    enum Op : int32_t
    {
        OP_IDEM = 0,
        OP_ADD = 1,
        OP_MUL = 2,
        OP_LENGTH = 3,
        OP_POW = 4,
        OP_SUM = 5
    };

    // ------------------------------------------------------------------ kernel parameter blocks
    struct ChainDesc
    {
        int32_t op[kMaxFunctors];
        float p[kMaxFunctors][4];
    };

    struct ClassifySource
    {
        const float* data; // local voxel (0,0,0)
        int64_t stride[3]; // element strides
        int32_t dim;
        int32_t active; // weight + iso > 0
        float weight;
        float iso;
        float invIso; // RN(1 / iso) (0 when iso is 0)
        ChainDesc chain;
        int32_t lutSlot; // index of the LUT in the staged table
        const uint8_t* adv; // field source: advection texture (alpha *= adv / 255), else nullptr
    };

    struct ClassifyParams
    {
        int32_t X, Y, Z;
        int32_t nsrc;
        float totalWeight;
        float invTotalWeight; // RN(1 / totalWeight) (0 when totalWeight is 0)
        int32_t ditherMode;
        int32_t T;
        int32_t nLut; // LUTs staged in LDS (0 = read from global)
        uint32_t divX_mul, divX_shift, divXY_mul, divXY_shift;
        const float4* luts; // nsrc * T
        uint2* tex; // per voxel: combined volume RGBA8 (.x), combined iso RGBA8 (.y)
        uint2* texH; // or nullptr: per voxel the volume RGBA8 as four binary16 values (see half_texel)
        int64_t nvox;
        ClassifySource src[kMaxSources];
    };

    // Per-frame ray / G-buffer state (drawCallBack, lib/isaac.hpp:2140-2283).
    struct FrameParams
    {
        float inv[16]; // inverse MVP, column major
        int32_t W, H;
        int32_t vx0, vy0, vx1, vy1; // viewport clipped to the framebuffer
        int32_t tilesX, tilesY; // 16x16 pixel tiles over the viewport
        float fbW, fbH;
        float maxSize; // maxGlobalSizeScaled * 0.5
        float off[3]; // position_offset (clipRay)
        float bbMin[3], bbMax[3]; // local box in scaled voxels
        int32_t outer[3]; // face i is an outer face for every sign (see clipRay)
        int32_t outerLow[3], outerHigh[3];
        int32_t nclip;
        float cpos[kMaxClipping][3]; // plane positions, scaled and offset
        float cnorm[kMaxClipping][3];
        float scale[3];
        float step;
        int32_t linear;
        int32_t prefetch; // fetch the next step's taps before compositing the current one
        int32_t waves; // march occupancy target: 0 compiler default, 6 or 8 waves per SIMD
        int32_t tileOrder; // march tile order: 0 one band of tile rows per XCD, S: SxS-tile super-tiles dealt to the XCDs
        int32_t waveGroups; // combined-form march workgroups: 1 one wave each, else 4 waves (one 16x16 tile)
        // textures
        int32_t X, Y, Z;
        const uint2* tex; // interleaved combined textures, x fastest
        // or nullptr: the volume RGBA of the same texels as binary16 values, the alpha's sign
        // set where the iso alpha byte is >= 128 (half_texel); the trilinear march steps over it
        const uint2* texH;
        // the same texels in 8^3 bricks with a one-texel apron on the + side (9^3 texels per
        // brick, x fastest inside, bricks x fastest; apron and partial bricks hold the
        // clamped texel), or nullptr: every trilinear sample reads one brick
        const uint2* texB;
        uint32_t nbx, nbxy; // bricks per row and per slice of the brick layout
        int32_t texOff32; // the layout the march reads spans < 4 GiB: its loads take 32-bit byte offsets
        int32_t isoEnabled;
        int32_t volumeEnabled;
        float minSize;
        float totalWeightAll;
        // shading
        int32_t renderMode;
        int32_t aoEnabled;
        float aoWeight;
        float bg[4]; // bgColor (channel reversed)
        uint32_t bgClear; // transformColor(bg with w = 0)
        float maxGlobalSizeScaled;
        float rankHue[4]; // getHSVA(halton(rank,3)*2pi,1,1,1) for render modes 7/8
        // G-buffer
        uint32_t* color;
        float* depth;
        float* normal; // 3 floats per pixel
        float* ao;
        float4* volColor; // premultiplied volume colour per pixel; w < 0: the volume pass did not blend
        float4* outFloat; // optional: final colour before RGBA8 quantisation (precision studies)
        unsigned long long* counters;
    };

    // ------------------------------------------------------------------ precision form (field bricks)
    // A source re-laid into 8^3 bricks with a one-voxel apron on the + side, stored
    // as x-pairs (v[x], v[x+1]) for x in 0..7, y, z in 0..8 (648 pair rows per brick,
    // the source's components side by side in each row), so the 8 trilinear taps of
    // any sample come from one brick with 4 aligned pair loads per component.  Pairs
    // are float2 (F32) or two 16-bit values (F16, BF16) holding v / scale, with one
    // FP32 scale per brick and component.
    constexpr int kBrick = 8;
    constexpr int kBrickPairs = 8 * 9 * 9;
    constexpr int kMaxBrickSources = 4;
    constexpr int kDtF32 = 0, kDtF16 = 1, kDtBf16 = 2; // values of isaac_hip_dtype

    struct BrickSource
    {
        const void* data; // [brick][648 pair rows][component]
        const float* scale; // [brick][component]
        int32_t dim;
        int32_t active; // weight + iso > 0
        float weight;
        float iso;
        float isoScale; // 0.5 / iso: iso value = alpha * isoScale
        ChainDesc chain;
        int32_t lutSlot; // source index of its transfer function in luts
    };

    struct PostParams
    {
        int32_t nsrc;
        int32_t dtype; // ISAAC_HIP_F32 / F16 / BF16
        int32_t nbx, nby, nbz;
        float invTotalWeight; // 1 / sum of volume-source weights (0 if that sum is 0)
        int32_t T;
        const float4* luts; // all sources' transfer functions, T entries each
        double invd[16]; // inverse MVP in binary64 (ray setup in FP64)
        double maxSized; // maxGlobalSizeScaled * 0.5
        double cposd[kMaxClipping][3]; // clipping plane positions, scaled and offset, binary64
        BrickSource src[kMaxBrickSources];
    };

    // A batch of frames (sub-volume x view) rendered by one persistent launch; the
    // per-frame parameter blocks live in device memory.
    struct FrameParams;
    struct ParticleParams;
    struct BatchParams
    {
        const struct FrameParams* frames;
        const PostParams* posts;
        const struct ParticleParams* parts;
        const uint32_t* tileStart; // nframes + 1 prefix sums of 16x16 tiles per frame
        int32_t nframes;
        uint32_t ntiles;
        int32_t chunk; // tiles per pop of the work counter
        uint32_t* counter; // zeroed before the launch
    };

    struct ParticleSourceDev
    {
        const int2* vrange; // ncells: [start, end) of the cell's visible records in recs
        const float4* recs; // rel.xyz (cell units), lut index bits in .w
        const float* radius; // per rec or nullptr
        float constRadius;
        float weight;
        const float4* lut; // T entries
    };

    struct ParticleParams
    {
        int32_t nsrc; // active (weight != 0) particle sources
        int32_t grid[3];
        float pscale[3];
        ParticleSourceDev src[4];
    };

    // ------------------------------------------------------------------ device math
    struct F3
    {
        float x, y, z;
    };

    __host__ __device__ inline F3 f3(float x, float y, float z)
    {
        return F3{x, y, z};
    }
    __host__ __device__ inline F3 add(F3 a, F3 b)
    {
        return F3{a.x + b.x, a.y + b.y, a.z + b.z};
    }
    __host__ __device__ inline F3 sub(F3 a, F3 b)
    {
        return F3{a.x - b.x, a.y - b.y, a.z - b.z};
    }
    __host__ __device__ inline F3 mul(F3 a, F3 b)
    {
        return F3{a.x * b.x, a.y * b.y, a.z * b.z};
    }
    __host__ __device__ inline F3 divv(F3 a, F3 b)
    {
        return F3{a.x / b.x, a.y / b.y, a.z / b.z};
    }
    __host__ __device__ inline F3 muls(F3 a, float s)
    {
        return F3{a.x * s, a.y * s, a.z * s};
    }
    __host__ __device__ inline F3 divs(F3 a, float s)
    {
        return F3{a.x / s, a.y / s, a.z / s};
    }
    // glm dot for vec3: (x*x' + y*y') + z*z'
    __host__ __device__ inline float dot3(F3 a, F3 b)
    {
        float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
        return tx + ty + tz;
    }
    __host__ __device__ inline float len3(F3 a)
    {
        return sqrtf(dot3(a, a));
    }
    __host__ __device__ inline F3 normalize3(F3 a)
    {
        return muls(a, 1.0f / sqrtf(dot3(a, a)));
    }
    __host__ __device__ inline float gmin(float a, float b)
    {
        return (b < a) ? b : a;
    }
    __host__ __device__ inline float gmax(float a, float b)
    {
        return (a < b) ? b : a;
    }
    __host__ __device__ inline float gclamp(float x, float lo, float hi)
    {
        return gmin(gmax(x, lo), hi);
    }
    __host__ __device__ inline float gsign(float x)
    {
        return float((0.0f < x) ? 1 : 0) - float((x < 0.0f) ? 1 : 0);
    }
    // v clamped to [lo, hi] (lo <= hi): one max and one min
    __host__ __device__ inline int iclamp(int v, int lo, int hi)
    {
        const int a = v < lo ? lo : v;
        return a > hi ? hi : a;
    }

    // (uint32_t) roundf(gclamp(v, 0, 1) * 255): the clamp is one median-of-three, and
    // round half away from zero of x = v * 255 is trunc(x + h) with h = 0x1.fffffep-2, the
    // largest binary32 below 0.5: for x < 0.5 the sum stays below 1 (with h = 0.5 it rounds
    // to 1 just below x = 0.5), and for x >= 0.5 it crosses the next integer exactly when
    // x's fraction is >= 0.5 (every binary32 v in [0, 1] checked, tests/cpp/round_check.c).
    // A NaN v gives 0 here (the median of NaN, 0, 1 is the minimum of the other two) as the
    // conversion of NaN gives 0 in the direct form.
    __device__ inline uint32_t unorm8(float v)
    {
        const float x = __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f) * 255.0f;
        return (uint32_t) (x + 0x1.fffffep-2f);
    }

    // The march texel of texture layout 16: the volume RGBA8 bytes as binary16 values
    // (every integer 0..255 is exact there), so a trilinear sum takes each one straight into
    // an fma with mixed operand formats, and the sign of the alpha set where the iso alpha
    // byte is >= 128 (the only steps where the iso alpha can reach 0.5; the march reads
    // the alpha through its magnitude).  .x = (r, g), .y = (b, a) with the first in the low half.
    __device__ inline uint2 half_texel(uint32_t vol, uint32_t iso)
    {
        auto h = [](uint32_t b) { return uint32_t(__half_as_ushort(__float2half_rn(float(b)))); };
        const uint32_t flag = (iso >> 31) << 15;
        return make_uint2(h(vol & 0xffu) | (h((vol >> 8) & 0xffu) << 16),
                          h((vol >> 16) & 0xffu) | ((h(vol >> 24) | flag) << 16));
    }

    // transfer-function index of x = value * T: roundf(x), NaN -> 0, clamped to
    // [0, T - 1] (lib/isaac/isaac_common_kernel.hpp:590-592), as trunc(med3(x + h, 0, T - 1))
    // with h = 0x1.fffffep-2 as in unorm8 (every binary32 x in [-2, T + 2] and the
    // infinities checked for T = 1024, tests/cpp/round_check.c); a NaN x gives NaN + h, whose
    // median with 0 and T - 1 is 0
    __device__ inline int tf_index_of(float x, int T)
    {
        return (int) __builtin_amdgcn_fmed3f(x + 0x1.fffffep-2f, 0.0f, float(T - 1));
    }

    // transformColor float4 -> byte4 (truncating, lib/isaac/isaac_helper.hpp:71-74)
    __host__ __device__ inline uint32_t to_byte4(float r, float g, float b, float a)
    {
        uint32_t q0 = (uint32_t) (gclamp(r, 0.0f, 1.0f) * 255.0f);
        uint32_t q1 = (uint32_t) (gclamp(g, 0.0f, 1.0f) * 255.0f);
        uint32_t q2 = (uint32_t) (gclamp(b, 0.0f, 1.0f) * 255.0f);
        uint32_t q3 = (uint32_t) (gclamp(a, 0.0f, 1.0f) * 255.0f);
        return q0 | (q1 << 8) | (q2 << 16) | (q3 << 24);
    }

    // fast unsigned division by an invariant divisor d >= 1 (round-up multiplier method)
    __host__ inline void make_fastdiv(uint32_t d, uint32_t& mul, uint32_t& shift)
    {
        uint32_t s = 0;
        while((1ull << s) < d)
            s++;
        uint64_t m = ((1ull << (32 + s)) + d - 1) / d;
        mul = (uint32_t) m; // m < 2^33; low 32 bits, the implicit 2^32 is added back in fastdiv
        shift = s;
    }
    __device__ inline uint32_t fastdiv(uint32_t n, uint32_t mul, uint32_t shift)
    {
        uint64_t hi = __umulhi(n, mul);
        return (uint32_t) ((hi + n) >> shift);
    }

    // min/max with the visiting-order key of each value: glm's min(x, y) = (y < x) ? y : x
    // keeps x on ties, so among equal values the earliest element wins; a NaN y fails
    // both comparisons and is never taken
    __device__ inline void mm_take_min(float& v, uint64_t& k, float v2, uint64_t k2)
    {
        if(v2 < v || (v2 == v && k2 < k))
        {
            v = v2;
            k = k2;
        }
    }

    __device__ inline void mm_take_max(float& v, uint64_t& k, float v2, uint64_t k2)
    {
        if(v < v2 || (v2 == v && k2 < k))
        {
            v = v2;
            k = k2;
        }
    }
} // namespace isaac_hip
