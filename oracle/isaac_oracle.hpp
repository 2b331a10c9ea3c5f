// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A from-scratch C++17 restatement of ISAAC's per-rank render path (the
// "optimized" two-stage renderer plus particles, SSAO and shading), written
// from reading the reference sources as text.  It is the correctness checker
// for the HIP implementation: only tests, the smoke check and the benchmark's
// CPU-baseline leg may load it.  The product never links or calls it.
//
// Templated on the arithmetic type R:
//   R = float  : reference-faithful FP32 path (the parity oracle).  Built with
//                -ffp-contract=off so every + - * / is one IEEE-754 binary32
//                operation in the same order as the reference expression.
//   R = double : the FP64 oracle (same algorithm, same quantisation points,
//                double arithmetic) used for PSNR statements.
//
// Citations are reference file:line (paths relative to the reference root).
// Third-party arithmetic (glm 0.9.9.9-dev @ 6ad79aae, not vendored in the
// reference) is restated from glm's published algorithms: perspectiveRH_NO,
// compute_inverse<4,4>, mat4*vec4, mat4*mat4, dot/length/normalize.
//
// Pinning (each checked by tests/test_oracle_*.py):
//   transfer LUT / functor chain / dither ........ closed-form KATs
//   trilinear sampler ........................... exact on affine fields, CLAMP borders
//   pixelToRay / clipRay ........................ textbook slab + plane test, FP64 vs FP32
//   volume compositing .......................... constant-volume closed form, early-exit count
//   iso ......................................... analytic plane/sphere fields
//   particles ................................... brute-force nearest sphere on tiny inputs
//   binning ..................................... brute-force stable sort (bit-exact)
//   SSAO ........................................ brute force on tiny depth images
//   shading ..................................... per-pixel closed form
//   camera inverse .............................. M * inv(M) = I, FP64 vs FP32
// glm's exact FP32 rounding for inverse/normalize is "parity unpinned" (no
// glm build is available); the restatement follows glm's published op order.
//
// ORACLE_MUTANT (default 0) plants one known mistake, for the tests that show the
// closed-form checks reject it; the shipped liboracle.so is always built with 0:
//   1 classification divides by the weights of ALL sources (particles included)
//   2 the volume march multiplies by the weights of the volume sources only
//   3 iso ties go to the later source (<= instead of <)
//   4 iso texel alpha is not rescaled by the threshold
//   5 render mode 7 keeps the texel colour instead of the rank hue
//   6 render mode 2 keeps the depth (the volume is drawn over the solid)
#pragma once
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace isaac_oracle
{
    constexpr int MAX_SOURCES = 16; // ISAAC_MAX_SOURCES lib/isaac/isaac_defines.hpp:36-38
    constexpr int MAX_FUNCTORS = 3; // ISAAC_MAX_FUNCTORS lib/isaac/isaac_defines.hpp:27-29
    constexpr int MAX_CLIPPING = 16; // ISAAC_MAX_CLIPPING lib/isaac/isaac_defines.hpp:44-46
    constexpr float Z_NEAR = 0.002f; // ISAAC_Z_NEAR lib/isaac/isaac_defines.hpp:53
    constexpr float Z_FAR = 100.0f; // ISAAC_Z_FAR lib/isaac/isaac_defines.hpp:54

    // ------------------------------------------------------------------ vectors
    template<class R>
    struct V2
    {
        R x, y;
    };
    template<class R>
    struct V3
    {
        R x, y, z;
        R& operator[](int i)
        {
            return i == 0 ? x : (i == 1 ? y : z);
        }
        R operator[](int i) const
        {
            return i == 0 ? x : (i == 1 ? y : z);
        }
    };
    template<class R>
    struct V4
    {
        R x, y, z, w;
        R& operator[](int i)
        {
            return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w));
        }
        R operator[](int i) const
        {
            return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w));
        }
    };
    struct B4
    {
        uint8_t x, y, z, w;
    };

    template<class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
    template<class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    template<class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
    template<class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
    template<class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
    template<class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
    template<class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
    template<class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
    template<class R> inline V4<R> operator+(V4<R> a, V4<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
    template<class R> inline V4<R> operator-(V4<R> a, V4<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
    template<class R> inline V4<R> operator*(V4<R> a, V4<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
    template<class R> inline V4<R> operator*(V4<R> a, R s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
    template<class R> inline V4<R> operator*(R s, V4<R> a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
    template<class R> inline V4<R> operator/(V4<R> a, R s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }

    // glm compute_dot: vec3 = (x+y)+z, vec4 = (x+y)+(z+w)
    template<class R> inline R dot(V3<R> a, V3<R> b)
    {
        V3<R> t = a * b;
        return t.x + t.y + t.z;
    }
    template<class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
    // glm normalize: v * inversesqrt(dot(v,v)), inversesqrt(x) = 1/sqrt(x)
    template<class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }
    // glm::min / max / clamp scalar semantics
    template<class R> inline R gmin(R a, R b) { return (b < a) ? b : a; }
    template<class R> inline R gmax(R a, R b) { return (a < b) ? b : a; }
    template<class R> inline R gclamp(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
    template<class R> inline R gsign(R x) { return R((R(0) < x) ? 1 : 0) - R((x < R(0)) ? 1 : 0); }
    template<class R> inline R gfract(R x) { return x - std::floor(x); }

    // ------------------------------------------------------------------ matrices (column major m[col][row], glm)
    template<class R>
    struct M4
    {
        R m[4][4];
    };

    template<class R> inline M4<R> mat_identity()
    {
        M4<R> r{};
        for(int i = 0; i < 4; i++)
            r.m[i][i] = R(1);
        return r;
    }

    // glm mat4*mat4: Result[c] = A0*B[c][0] + A1*B[c][1] + A2*B[c][2] + A3*B[c][3] (left to right)
    template<class R> inline M4<R> mat_mul(const M4<R>& a, const M4<R>& b)
    {
        M4<R> r;
        for(int c = 0; c < 4; c++)
            for(int row = 0; row < 4; row++)
            {
                R s = a.m[0][row] * b.m[c][0];
                s = s + a.m[1][row] * b.m[c][1];
                s = s + a.m[2][row] * b.m[c][2];
                s = s + a.m[3][row] * b.m[c][3];
                r.m[c][row] = s;
            }
        return r;
    }

    // glm mat4*vec4: (m0*x + m1*y) + (m2*z + m3*w)
    template<class R> inline V4<R> mat_vec(const M4<R>& m, V4<R> v)
    {
        V4<R> r;
        for(int row = 0; row < 4; row++)
        {
            R a0 = m.m[0][row] * v.x;
            R a1 = m.m[1][row] * v.y;
            R s0 = a0 + a1;
            R a2 = m.m[2][row] * v.z;
            R a3 = m.m[3][row] * v.w;
            R s1 = a2 + a3;
            r[row] = s0 + s1;
        }
        return r;
    }

    // glm compute_inverse<4,4> (cofactor/adjugate form, glm 0.9.9 op order)
    template<class R> inline M4<R> mat_inverse(const M4<R>& M)
    {
        auto m = [&](int c, int r) { return M.m[c][r]; };
        R c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
        R c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
        R c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
        R c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
        R c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
        R c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
        R c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
        R c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
        R c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
        R c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
        R c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
        R c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
        R c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
        R c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
        R c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
        R c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
        R c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
        R c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
        V4<R> f0{c00, c00, c02, c03}, f1{c04, c04, c06, c07}, f2{c08, c08, c10, c11};
        V4<R> f3{c12, c12, c14, c15}, f4{c16, c16, c18, c19}, f5{c20, c20, c22, c23};
        V4<R> v0{m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
        V4<R> v1{m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
        V4<R> v2{m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
        V4<R> v3{m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
        V4<R> i0 = v1 * f0 - v2 * f1 + v3 * f2;
        V4<R> i1 = v0 * f0 - v2 * f3 + v3 * f4;
        V4<R> i2 = v0 * f1 - v1 * f3 + v3 * f5;
        V4<R> i3 = v0 * f2 - v1 * f4 + v2 * f5;
        V4<R> sa{R(1), R(-1), R(1), R(-1)}, sb{R(-1), R(1), R(-1), R(1)};
        V4<R> col[4] = {i0 * sa, i1 * sb, i2 * sa, i3 * sb};
        V4<R> row0{col[0].x, col[1].x, col[2].x, col[3].x};
        V4<R> mc0{M.m[0][0], M.m[0][1], M.m[0][2], M.m[0][3]};
        V4<R> d0 = mc0 * row0;
        R d1 = (d0.x + d0.y) + (d0.z + d0.w);
        R inv = R(1) / d1;
        M4<R> r;
        for(int c = 0; c < 4; c++)
        {
            V4<R> t = col[c] * inv;
            r.m[c][0] = t.x;
            r.m[c][1] = t.y;
            r.m[c][2] = t.z;
            r.m[c][3] = t.w;
        }
        return r;
    }

    // glm::perspective (RH, depth -1..1) in double.  The default controller passes
    // fovy = 45.0, which glm 0.9.9 interprets as radians
    // (lib/isaac/isaac_controllers.hpp:33-37).
    inline void perspective(double fovy, double aspect, double zNear, double zFar, double out[16])
    {
        double tanHalf = std::tan(fovy / 2.0);
        for(int i = 0; i < 16; i++)
            out[i] = 0.0;
        out[0 * 4 + 0] = 1.0 / (aspect * tanHalf);
        out[1 * 4 + 1] = 1.0 / tanHalf;
        out[2 * 4 + 2] = -(zFar + zNear) / (zFar - zNear);
        out[2 * 4 + 3] = -1.0;
        out[3 * 4 + 2] = -(2.0 * zFar * zNear) / (zFar - zNear);
    }

    // --------------------------------------------------------------- scene/state
    // Sizes as stored in SimulationSizeStruct (lib/isaac/isaac_types.hpp:119-131),
    // filled at lib/isaac.hpp:2181-2202 from the ctor conversions lib/isaac.hpp:659-661,796-797.
    struct Sizes
    {
        int64_t globalSize[3];
        int64_t maxGlobalSize;
        int32_t position[3];
        int64_t localSize[3];
        int64_t localParticleSize[3];
        int64_t globalSizeScaled[3];
        int64_t maxGlobalSizeScaled;
        int32_t positionScaled[3];
        int64_t localSizeScaled[3];
    };

    inline Sizes make_sizes(const int64_t g[3], const int64_t l[3], const int64_t p[3], const int32_t pos[3], const float scale[3])
    {
        Sizes s{};
        for(int i = 0; i < 3; i++)
        {
            s.globalSize[i] = g[i];
            s.localSize[i] = l[i];
            s.localParticleSize[i] = p[i];
            s.position[i] = pos[i];
            s.globalSizeScaled[i] = (int64_t) (float(g[i]) * scale[i]);
            s.localSizeScaled[i] = (int64_t) (float(l[i]) * scale[i]);
            s.positionScaled[i] = (int32_t) (float(pos[i]) * scale[i]);
        }
        s.maxGlobalSize = std::max(g[0], std::max(g[1], g[2]));
        s.maxGlobalSizeScaled = std::max(s.globalSizeScaled[0], std::max(s.globalSizeScaled[1], s.globalSizeScaled[2]));
        return s;
    }

    struct ClipPlane
    {
        float position[3];
        float normal[3]; // already scaled by `scale` and normalised (editClipping lib/isaac.hpp:958-986)
    };

    struct VolumeSource
    {
        const float* data = nullptr; // points at local voxel (0,0,0)
        int dim = 1; // featureDim 1..4 (floats per element, interleaved)
        int64_t stride[3] = {0, 0, 0}; // element strides x,y,z
    };

    struct ParticleSource
    {
        const float* pos = nullptr; // count x 3, local unscaled voxel coordinates
        const float* attr = nullptr; // count x dim
        int dim = 1;
        const float* radius = nullptr; // per particle, or nullptr -> const_radius
        float const_radius = 0.5f;
        int64_t count = 0;
    };

    struct Chain
    {
        int bytecode[MAX_FUNCTORS] = {0, 0, 0};
        float params[MAX_FUNCTORS][4] = {};
    };

    struct State
    {
        Sizes sizes{};
        float scale[3] = {1, 1, 1};
        int fbW = 0, fbH = 0;
        int viewport[4] = {0, 0, 0, 0};
        int T = 1024;
        std::vector<VolumeSource> vol;
        std::vector<ParticleSource> part;
        // per source, volume sources first, then particle sources
        std::vector<const float*> lut; // T float4 each
        std::vector<Chain> chain;
        std::vector<float> weight;
        std::vector<float> iso;
        std::vector<const uint8_t*> adv; // per volume source: advection texture of a field source, else null
        float step = 0.5f;
        bool interpolation = false;
        int ditherMode = 1;
        bool aoEnabled = false;
        float aoWeight = 0.5f;
        int renderMode = 0;
        int rank = 0;
        std::vector<ClipPlane> clip;
        double proj[16];
        double mv[16];
        float bg[4] = {0, 0, 0, 0};
    };

    struct Counters
    {
        uint64_t volSamples = 0; // volume march iterations (1 trilinear/nearest sample each)
        uint64_t isoSamples = 0; // iso march iterations
        uint64_t isoHits = 0; // iso hits (7 extra samples each)
        uint64_t cellsVisited = 0; // particle DDA cells
        uint64_t sphereTests = 0; // particle ray-sphere tests
        uint64_t volRays = 0; // rays that entered the volume march loop
        uint64_t isoRays = 0;
        uint64_t partRays = 0;
    };

    // --------------------------------------------------------------- functors
    // pool order idem, add, mul, length, pow, sum (lib/isaac/isaac_functors.hpp:185-208)
    enum Op
    {
        OP_IDEM = 0,
        OP_ADD = 1,
        OP_MUL = 2,
        OP_LENGTH = 3,
        OP_POW = 4,
        OP_SUM = 5
    };

    // applyFunctorChain: f2(f1(f0(v,p0),p1),p2).x with the feature dimension
    // propagating through the chain (lib/isaac/isaac_functor_chain.hpp:96-108,
    // lib/isaac/isaac_functors.hpp:42-183).
    template<class R>
    inline R apply_chain_vals(const Chain& c, const R* v_in, int dim)
    {
        R v[4];
        for(int i = 0; i < dim; i++)
            v[i] = v_in[i];
        int n = dim;
        for(int k = 0; k < MAX_FUNCTORS; k++)
        {
            const float* p = c.params[k];
            switch(c.bytecode[k])
            {
            case OP_IDEM:
                break;
            case OP_ADD:
                for(int i = 0; i < n; i++)
                    v[i] = v[i] + R(p[i]);
                break;
            case OP_MUL:
                for(int i = 0; i < n; i++)
                    v[i] = v[i] * R(p[i]);
                break;
            case OP_LENGTH:
            {
                R d;
                if(n == 1)
                    d = v[0] * v[0];
                else if(n == 2)
                    d = v[0] * v[0] + v[1] * v[1];
                else if(n == 3)
                    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                else
                    d = (v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3]);
                v[0] = std::sqrt(d);
                n = 1;
                break;
            }
            case OP_POW:
                for(int i = 0; i < n; i++)
                    v[i] = std::pow(v[i], R(p[i]));
                break;
            case OP_SUM:
            {
                R s = R(0);
                for(int i = 0; i < n; i++)
                    s = s + v[i];
                v[0] = s;
                n = 1;
                break;
            }
            default:
                break;
            }
        }
        return v[0];
    }

    template<class R>
    inline R apply_chain(const Chain& c, const float* v_in, int dim)
    {
        R v[4];
        for(int i = 0; i < dim; i++)
            v[i] = R(v_in[i]);
        return apply_chain_vals<R>(c, v, dim);
    }

    // TF lookup index: clamp(round(v*T), 0, T-1) (lib/isaac/isaac_common_kernel.hpp:590-592)
    template<class R>
    inline int64_t tf_index(R v, int T)
    {
        R lv = std::round(v * R(T));
        if(lv != lv)
            return 0; // NaN: index 0, what a GPU float->int conversion yields
        return (int64_t) gclamp(lv, R(0), R(T - 1));
    }

    template<class R>
    inline V4<R> lut_at(const float* lut, int64_t i)
    {
        return {R(lut[4 * i + 0]), R(lut[4 * i + 1]), R(lut[4 * i + 2]), R(lut[4 * i + 3])};
    }

    // --------------------------------------------------------------- colour conversions
    // transformColor float4 -> byte4 truncates (lib/isaac/isaac_helper.hpp:71-74)
    template<class R>
    inline B4 to_byte(V4<R> c)
    {
        auto q = [](R v) { return (uint8_t) (gclamp(v, R(0), R(1)) * R(255)); };
        return {q(c.x), q(c.y), q(c.z), q(c.w)};
    }
    template<class R>
    inline V4<R> from_byte(B4 b)
    {
        return {R(b.x) / R(255), R(b.y) / R(255), R(b.z) / R(255), R(b.w) / R(255)};
    }

    // dither masks (reference data, lib/isaac/isaac_dither_kernel.hpp:22-323, LGPL-3.0)
#define DITHER_TABLE(type, name, n) static const type name[n]
#include "dither_tables.inc"
#undef DITHER_TABLE

    // applyDither3D (lib/isaac/isaac_dither_kernel.hpp:7-328): mode 0 rounds; modes 1
    // (ordered 4^3), 2 (blue noise 4^3), 3 (blue noise 8^3) and every other mode
    // (blue noise 16^3) add one where the fraction exceeds the mask threshold
    template<class R>
    inline B4 dither(int mode, int x, int y, int z, V4<R> c)
    {
        if(mode == 0)
        {
            auto q = [](R v) { return (uint8_t) std::round(gclamp(v, R(0), R(1)) * R(255)); };
            return {q(c.x), q(c.y), q(c.z), q(c.w)};
        }
        R d;
        if(mode == 1)
            d = R(kDitherOrdered4[x % 4 + (y % 4) * 4 + (z % 4) * 16]) / R(64);
        else if(mode == 2)
            d = R(kDitherBlue4[x % 4 + (y % 4) * 4 + (z % 4) * 16]) / R(64);
        else if(mode == 3)
            d = R(kDitherBlue8[x % 8 + (y % 8) * 8 + (z % 8) * 64]) / R(512);
        else
            d = R(kDitherBlue16[x % 16 + (y % 16) * 16 + (z % 16) * 256]) / R(4096);
        auto q = [d](R v)
        {
            R s = v * R(255);
            int i = (int) s;
            i += (gfract(s) > d) ? 1 : 0;
            return (uint8_t) std::min(std::max(i, 0), 255);
        };
        return {q(c.x), q(c.y), q(c.z), q(c.w)};
    }

    // --------------------------------------------------------------- textures
    struct Tex3B4
    {
        std::vector<B4> data;
        int64_t X = 0, Y = 0, Z = 0;
        B4 at(int64_t x, int64_t y, int64_t z) const
        {
            return data[x + y * X + z * X * Y];
        }
    };

    // Sampler<filter, CLAMP, normalized>::sample on a byte4 texture
    // (lib/isaac/isaac_texture.hpp:135-306; trilinear lib/isaac/isaac_helper.hpp:124-144)
    template<class R>
    inline V4<R> sample_b4(const Tex3B4& t, V3<R> c, bool linear)
    {
        auto cl = [](int64_t v, int64_t n) { return std::min(std::max(v, int64_t(0)), n - 1); };
        auto fetch = [&](int64_t x, int64_t y, int64_t z)
        {
            B4 b = t.at(cl(x, t.X), cl(y, t.Y), cl(z, t.Z));
            return V4<R>{R(b.x), R(b.y), R(b.z), R(b.w)};
        };
        V4<R> r;
        if(!linear)
        {
            r = fetch((int64_t) (int) c.x, (int64_t) (int) c.y, (int64_t) (int) c.z);
        }
        else
        {
            V3<R> cc{c.x - R(0.5), c.y - R(0.5), c.z - R(0.5)};
            R fx = std::floor(cc.x), fy = std::floor(cc.y), fz = std::floor(cc.z);
            int64_t bx = (int) fx, by = (int) fy, bz = (int) fz;
            V3<R> f{cc.x - fx, cc.y - fy, cc.z - fz};
            V4<R> d[2][2][2];
            for(int z = 0; z < 2; z++)
                for(int y = 0; y < 2; y++)
                    for(int x = 0; x < 2; x++)
                        d[x][y][z] = fetch(bx + x, by + y, bz + z);
            auto lin = [](R t_, V4<R> a, V4<R> b) { return a * (R(1) - t_) + b * t_; };
            V4<R> a0 = lin(f.z, d[0][0][0], d[0][0][1]);
            V4<R> a1 = lin(f.z, d[0][1][0], d[0][1][1]);
            V4<R> A = lin(f.y, a0, a1);
            V4<R> b0 = lin(f.z, d[1][0][0], d[1][0][1]);
            V4<R> b1 = lin(f.z, d[1][1][0], d[1][1][1]);
            V4<R> B = lin(f.y, b0, b1);
            r = lin(f.x, A, B);
        }
        return r / R(255);
    }

    // --------------------------------------------------------------- rays
    template<class R>
    struct Ray
    {
        V3<R> dir, start, end;
        R startDepth, endDepth;
        bool isClipped;
        V3<R> clippingNormal;
    };

    // pixelToRay (lib/isaac/isaac_common_kernel.hpp:51-92)
    template<class R>
    inline Ray<R> pixel_to_ray(const M4<R>& invMVP, const Sizes& sz, R px, R py, R W, R H)
    {
        R vx = px / W * R(2) - R(1);
        R vy = py / H * R(2) - R(1);
        V4<R> s = mat_vec(invMVP, V4<R>{vx, vy, R(-1), R(1)});
        V4<R> e = mat_vec(invMVP, V4<R>{vx, vy, R(1), R(1)});
        Ray<R> ray;
        ray.start = V3<R>{s.x / s.w, s.y / s.w, s.z / s.w};
        ray.end = V3<R>{e.x / e.w, e.y / e.w, e.z / e.w};
        R maxSize = R(sz.maxGlobalSizeScaled) * R(0.5);
        ray.start = ray.start * maxSize;
        ray.end = ray.end * maxSize;
        ray.dir = normalize(ray.end - ray.start);
        ray.isClipped = false;
        ray.clippingNormal = V3<R>{0, 0, 0};
        ray.startDepth = 0;
        ray.endDepth = 0;
        return ray;
    }

    // clipRay (lib/isaac/isaac_common_kernel.hpp:94-193)
    template<class R>
    inline bool clip_ray(Ray<R>& ray, const Sizes& sz, const std::vector<ClipPlane>& planes)
    {
        const int n = (int) planes.size();
        V3<R> cpos[MAX_CLIPPING], cnorm[MAX_CLIPPING];
        for(int i = 0; i < n; i++)
        {
            V3<R> gs{R(sz.globalSizeScaled[0]), R(sz.globalSizeScaled[1]), R(sz.globalSizeScaled[2])};
            V3<R> p{R(planes[i].position[0]), R(planes[i].position[1]), R(planes[i].position[2])};
            cpos[i] = p * gs * R(0.5);
            cnorm[i] = V3<R>{R(planes[i].normal[0]), R(planes[i].normal[1]), R(planes[i].normal[2])};
        }
        V3<R> off;
        for(int i = 0; i < 3; i++)
            off[i] = R(int32_t(sz.globalSizeScaled[i]) / 2 - sz.positionScaled[i]);
        ray.start = ray.start + off;
        ray.end = ray.end + off;
        for(int i = 0; i < n; i++)
            cpos[i] = cpos[i] + off;

        V3<R> bmin, bmax;
        for(int i = 0; i < 3; i++)
        {
            int32_t lo = std::max(int32_t(0), -sz.positionScaled[i]);
            int32_t hi = std::min(int32_t(sz.localSizeScaled[i]), int32_t(sz.globalSizeScaled[i]) - sz.positionScaled[i]);
            bmin[i] = (R(lo) - ray.start[i]) / ray.dir[i];
            bmax[i] = (R(hi) - ray.start[i]) / ray.dir[i];
        }
        for(int i = 0; i < 3; i++)
            if(bmax[i] < bmin[i])
                std::swap(bmax[i], bmin[i]);
        ray.startDepth = gmax(bmin.x, gmax(bmin.y, bmin.z));
        ray.endDepth = gmin(bmax.x, gmin(bmax.y, bmax.z));

        for(int i = 0; i < 3; i++)
        {
            float sgn = float(gsign(ray.dir[i]));
            bool lowOuter = sz.position[i] == 0 && (sgn + 1.0f) != 0.0f;
            bool highOuter = (sz.position[i] + sz.localSize[i] == sz.globalSize[i]) && (sgn - 1.0f) != 0.0f;
            if(bmin[i] == ray.startDepth && (lowOuter || highOuter))
            {
                ray.isClipped = true;
                ray.clippingNormal[i] = R(sgn);
            }
        }

        for(int i = 0; i < n; i++)
        {
            R d = dot(ray.dir, cnorm[i]);
            R depth = (dot(cpos[i], cnorm[i]) - dot(ray.start, cnorm[i])) / d;
            if(d > 0)
            {
                if(ray.endDepth < depth)
                    return false;
                if(ray.startDepth <= depth)
                {
                    ray.clippingNormal = cnorm[i];
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
        ray.startDepth = gmax(ray.startDepth, R(Z_NEAR));
        if(ray.startDepth > ray.endDepth || std::isinf(ray.startDepth) || std::isinf(ray.endDepth))
            return false;
        return true;
    }

    // --------------------------------------------------------------- classification
    // MergeToCombinedTextureKernel (lib/isaac/isaac_common_kernel.hpp:553-615,690-758);
    // totalWeight over volume(+field) sources only (lib/isaac.hpp:1918-1922).
    template<class R>
    inline void classify(const State& st, Tex3B4& volTex, Tex3B4& isoTex)
    {
        const int64_t X = st.sizes.localSize[0], Y = st.sizes.localSize[1], Z = st.sizes.localSize[2];
        volTex.X = isoTex.X = X;
        volTex.Y = isoTex.Y = Y;
        volTex.Z = isoTex.Z = Z;
        volTex.data.assign(X * Y * Z, B4{0, 0, 0, 0});
        isoTex.data.assign(X * Y * Z, B4{0, 0, 0, 0});
        const int nv = (int) st.vol.size();
        R totalWeight = R(0);
        {
            float tw = 0.0f;
            for(int s = 0; s < nv; s++)
                tw += st.weight[s];
#if ORACLE_MUTANT == 1
            for(size_t s = size_t(nv); s < st.weight.size(); s++)
                tw += st.weight[s];
#endif
            totalWeight = R(tw);
        }
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for(int64_t z = 0; z < Z; z++)
            for(int64_t y = 0; y < Y; y++)
                for(int64_t x = 0; x < X; x++)
                {
                    V4<R> vc{0, 0, 0, 0}, ic{0, 0, 0, 0};
                    for(int s = 0; s < nv; s++)
                    {
                        const R w = R(st.weight[s]);
                        const R thr = R(st.iso[s]);
                        if(!(w + thr > R(0)))
                            continue;
                        const VolumeSource& src = st.vol[s];
                        const float* e = src.data + (x * src.stride[0] + y * src.stride[1] + z * src.stride[2]) * src.dim;
                        R tv = apply_chain<R>(st.chain[s], e, src.dim);
                        V4<R> color = lut_at<R>(st.lut[s], tf_index(tv, st.T));
                        // field source: alpha * advection / 255
                        // (MergeAdvectionToCombinedTextureIterator, lib/isaac/isaac_common_kernel.hpp:655-656)
                        if(size_t(s) < st.adv.size() && st.adv[size_t(s)])
                            color.w = color.w * (R(st.adv[size_t(s)][x + y * X + z * X * Y]) / R(255));
                        if(w > R(0))
                        {
                            V4<R> c = color;
                            c.w = c.w * w;
                            c.x = c.x * c.w;
                            c.y = c.y * c.w;
                            c.z = c.z * c.w;
                            vc = vc + c;
                        }
                        if(thr > R(0))
                        {
                            V4<R> c = color;
#if ORACLE_MUTANT != 4
                            c.w = c.w / thr * R(0.5);
#endif
#if ORACLE_MUTANT == 3
                            if(ic.w <= c.w)
#else
                            if(ic.w < c.w)
#endif
                                ic = c;
                        }
                    }
                    // all-zero weights would divide by zero in the reference; defined here as an empty voxel
                    if(totalWeight != R(0))
                        vc = vc / totalWeight;
                    else
                        vc = V4<R>{0, 0, 0, 0};
                    const int64_t idx = x + y * X + z * X * Y;
                    volTex.data[idx] = dither<R>(st.ditherMode, (int) x, (int) y, (int) z, vc);
                    auto rq = [](R v) { return (uint8_t) std::round(gclamp(v, R(0), R(1)) * R(255)); };
                    isoTex.data[idx] = B4{rq(ic.x), rq(ic.y), rq(ic.z), rq(ic.w)};
                }
    }

    // --------------------------------------------------------------- particle binning
    // The reference leaves binning to the application (getIterator(cell),
    // example/example.cpp:148-228).  The oracle's cell list is a stable
    // counting sort of particle indices by cell id: cell = clamp(floor(p / cw)),
    // cw = localSize / particleGrid (float), rel = p / cw - cell (cell units).
    struct CellList
    {
        std::vector<int32_t> cellOf; // per particle
        std::vector<int32_t> count; // per cell
        std::vector<int64_t> offset; // ncells + 1
        std::vector<int64_t> perm; // sorted particle indices
        std::vector<float> rel; // per sorted slot, 3 floats
    };

    inline int32_t cell_coord(float p, float cw, int64_t n)
    {
        float q = p / cw;
        int32_t c = (int32_t) std::floor(q);
        if(c < 0)
            c = 0;
        if(c > n - 1)
            c = (int32_t) (n - 1);
        return c;
    }

    inline CellList bin_particles(const ParticleSource& ps, const Sizes& sz)
    {
        CellList cl;
        const int64_t gx = sz.localParticleSize[0], gy = sz.localParticleSize[1], gz = sz.localParticleSize[2];
        const int64_t ncell = gx * gy * gz;
        const float cw[3] = {float(sz.localSize[0]) / float(gx), float(sz.localSize[1]) / float(gy), float(sz.localSize[2]) / float(gz)};
        cl.cellOf.resize(ps.count);
        cl.count.assign(ncell, 0);
        for(int64_t i = 0; i < ps.count; i++)
        {
            int32_t cx = cell_coord(ps.pos[3 * i + 0], cw[0], gx);
            int32_t cy = cell_coord(ps.pos[3 * i + 1], cw[1], gy);
            int32_t cz = cell_coord(ps.pos[3 * i + 2], cw[2], gz);
            int32_t c = (int32_t) (cx + cy * gx + cz * gx * gy);
            cl.cellOf[i] = c;
            cl.count[c]++;
        }
        cl.offset.assign(ncell + 1, 0);
        for(int64_t c = 0; c < ncell; c++)
            cl.offset[c + 1] = cl.offset[c] + cl.count[c];
        std::vector<int64_t> cursor(cl.offset.begin(), cl.offset.end() - 1);
        cl.perm.resize(ps.count);
        cl.rel.resize(3 * ps.count);
        for(int64_t i = 0; i < ps.count; i++)
        {
            int32_t c = cl.cellOf[i];
            int64_t slot = cursor[c]++;
            cl.perm[slot] = i;
            int64_t cx = c % gx, cy = (c / gx) % gy, cz = c / (gx * gy);
            cl.rel[3 * slot + 0] = ps.pos[3 * i + 0] / cw[0] - float(cx);
            cl.rel[3 * slot + 1] = ps.pos[3 * i + 1] / cw[1] - float(cy);
            cl.rel[3 * slot + 2] = ps.pos[3 * i + 2] / cw[2] - float(cz);
        }
        return cl;
    }

    // --------------------------------------------------------------- field-source advection
    // glm 0.9.9 scalar min/max: min(x, y) = (y < x) ? y : x, max(x, y) = (x < y) ? y : x
    // (a NaN y never replaces x; on ties the running value stays)
    inline float glm_min(float x, float y)
    {
        return (y < x) ? y : x;
    }

    inline float glm_max(float x, float y)
    {
        return (x < y) ? y : x;
    }

    // halton (lib/isaac/isaac_helper.hpp:57-69)
    inline float halton(uint32_t index, uint32_t base)
    {
        float result = 0.0f;
        float f = 1.0f;
        while(index > 0)
        {
            f /= float(base);
            result += f * float(index % base);
            index /= base;
        }
        return result;
    }

    // linear / bilinear / trilinear of the 2x2x2 taps v[x][y][z] (lib/isaac/isaac_helper.hpp:124-144)
    inline float lin(float t, float a, float b)
    {
        return a * (1.0f - t) + b * t;
    }

    inline float trilin(float fx, float fy, float fz, const float v[2][2][2])
    {
        const float a = lin(fy, lin(fz, v[0][0][0], v[0][0][1]), lin(fz, v[0][1][0], v[0][1][1]));
        const float b = lin(fy, lin(fz, v[1][0][0], v[1][0][1]), lin(fz, v[1][1][0], v[1][1][1]));
        return lin(fx, a, b);
    }

    // Sampler<LINEAR, border>::sample on a guard-less texture (lib/isaac/isaac_texture.hpp:182-259):
    // coord - 0.5, taps at floor + {0,1}^3; REPEAT wraps with the reference's modulo,
    // VALUE reads 0 outside
    template<class T, bool REPEAT>
    inline float sample_linear(const T* t, int64_t X, int64_t Y, int64_t Z, float px, float py, float pz)
    {
        const float cx = px - 0.5f, cy = py - 0.5f, cz = pz - 0.5f;
        const int64_t x0 = (int64_t) std::floor(cx), y0 = (int64_t) std::floor(cy), z0 = (int64_t) std::floor(cz);
        float v[2][2][2];
        for(int dz = 0; dz < 2; dz++)
            for(int dy = 0; dy < 2; dy++)
                for(int dx = 0; dx < 2; dx++)
                {
                    int64_t x = x0 + dx, y = y0 + dy, z = z0 + dz;
                    if(REPEAT)
                    {
                        x = (X + (x % X)) % X;
                        y = (Y + (y % Y)) % Y;
                        z = (Z + (z % Z)) % Z;
                        v[dx][dy][dz] = float(t[x + y * X + z * X * Y]);
                    }
                    else
                    {
                        const bool in = x >= 0 && y >= 0 && z >= 0 && x < X && y < Y && z < Z;
                        v[dx][dy][dz] = in ? float(t[x + y * X + z * X * Y]) : 0.0f;
                    }
                }
        return trilin(cx - std::floor(cx), cy - std::floor(cy), cz - std::floor(cz), v);
    }

    // updateNoiseTexture (lib/isaac.hpp:499-562): clear, Halton points set to 1
    // (HaltonSeedingKernel, lib/isaac/isaac_common_kernel.hpp:334-348), GaussBlur7 along
    // x, y, z (lib/isaac/isaac_common_kernel.hpp:381-411), then clamp(v * 30, 0, 1)
    inline std::vector<float> make_noise(int64_t X, int64_t Y, int64_t Z, const float scale[3], uint32_t seedPoints)
    {
        const size_t n = size_t(X * Y * Z);
        std::vector<float> a(n, 0.0f), b(n, 0.0f);
        for(uint32_t i = 1; i <= seedPoints; ++i)
        {
            const int64_t x = (int64_t) (float(X) * halton(i, 3));
            const int64_t y = (int64_t) (float(Y) * halton(i, 5));
            const int64_t z = (int64_t) (float(Z) * halton(i, 7));
            a[size_t(x + y * X + z * X * Y)] = 1.0f;
        }
        const float g7[4] = {64.0f, 40.5f, 10.0f, 1.0f};
        auto blur = [&](const std::vector<float>& src, std::vector<float>& dst, float mx, float my, float mz)
        {
            for(int64_t z = 0; z < Z; z++)
                for(int64_t y = 0; y < Y; y++)
                    for(int64_t x = 0; x < X; x++)
                    {
                        float r = 0.0f;
                        for(int k = -3; k < 4; k++)
                        {
                            const float px = float(x) + mx * float(k) / scale[0];
                            const float py = float(y) + my * float(k) / scale[1];
                            const float pz = float(z) + mz * float(k) / scale[2];
                            r += sample_linear<float, true>(src.data(), X, Y, Z, px, py, pz) * g7[std::abs(k)];
                        }
                        dst[size_t(x + y * X + z * X * Y)] = r / 167.0f;
                    }
        };
        blur(a, b, 1, 0, 0);
        blur(b, a, 0, 1, 0);
        blur(a, b, 0, 0, 1);
        for(size_t i = 0; i < n; i++)
            a[i] = std::min(std::max(b[i] * 30.0f, 0.0f), 1.0f);
        return a;
    }

    // GenerateAdvectionTextureKernel (lib/isaac/isaac_common_kernel.hpp:500-550) over the
    // local voxels; the advection textures' guard is never written on one rank, so the
    // history reads 0 outside the local box
    inline std::vector<uint8_t> advect(
        const VolumeSource& field,
        const float* noise,
        const uint8_t* history,
        int64_t X,
        int64_t Y,
        int64_t Z,
        const float scale[3],
        float stepSize,
        float historyWeight,
        bool seeding)
    {
        std::vector<uint8_t> out(size_t(X * Y * Z));
        for(int64_t z = 0; z < Z; z++)
            for(int64_t y = 0; y < Y; y++)
                for(int64_t x = 0; x < X; x++)
                {
                    const float* e = field.data + (x * field.stride[0] + y * field.stride[1] + z * field.stride[2]) * 3;
                    float v[3] = {e[0], e[1], e[2]};
                    float len = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
                    len = glm_max(len, std::numeric_limits<float>::min());
                    for(int k = 0; k < 3; k++)
                        v[k] /= len;
                    float off[3];
                    for(int k = 0; k < 3; k++)
                        off[k] = v[k] * stepSize / scale[k];
                    const float px = (float(x) + 0.5f) - off[0], py = (float(y) + 0.5f) - off[1], pz = (float(z) + 0.5f) - off[2];
                    const float hist = sample_linear<uint8_t, false>(history, X, Y, Z, px, py, pz) / 255.0f;
                    const size_t i = size_t(x + y * X + z * X * Y);
                    const float nz = seeding ? noise[i] : 0.0f;
                    const V4<float> c{nz + hist * historyWeight * (1.0f - nz), 0.0f, 0.0f, 0.0f};
                    out[i] = dither<float>(5, (int) x, (int) y, (int) z, c).x;
                }
        return out;
    }

    // --------------------------------------------------------------- min/max request
    // MinMaxKernel (lib/isaac/isaac_min_max_kernel.hpp:22-57): one thread per (x, y)
    // column walks z over the functor-chain values; the host then reduces the columns
    // in order x + y*X with strict comparisons (CalcMinMaxIterator,
    // lib/isaac.hpp:372-432).  Non-persistent sources read their texture, which holds
    // the same chain values.
    inline void min_max_volume(const VolumeSource& vs, const Chain& chain, const Sizes& sz, float& outMin, float& outMax)
    {
        const int64_t X = sz.localSize[0], Y = sz.localSize[1], Z = sz.localSize[2];
        std::vector<float> colMin(size_t(X * Y)), colMax(size_t(X * Y));
        for(int64_t y = 0; y < Y; y++)
            for(int64_t x = 0; x < X; x++)
            {
                float mn = std::numeric_limits<float>::max();
                float mx = -std::numeric_limits<float>::max();
                for(int64_t z = 0; z < Z; z++)
                {
                    const float* e = vs.data + (x * vs.stride[0] + y * vs.stride[1] + z * vs.stride[2]) * vs.dim;
                    const float v = apply_chain<float>(chain, e, vs.dim);
                    mn = glm_min(mn, v);
                    mx = glm_max(mx, v);
                }
                colMin[size_t(x + y * X)] = mn;
                colMax[size_t(x + y * X)] = mx;
            }
        outMin = std::numeric_limits<float>::max();
        outMax = -std::numeric_limits<float>::max();
        for(size_t i = 0; i < colMin.size(); i++)
        {
            if(colMin[i] < outMin)
                outMin = colMin[i];
            if(colMax[i] > outMax)
                outMax = colMax[i];
        }
    }

    // MinMaxParticleKernel (lib/isaac/isaac_min_max_kernel.hpp:60-97): per (x, y)
    // column of the particle grid, z ascending, the particles of each cell in list
    // order (here: ascending particle index); host reduction lib/isaac.hpp:434-497.
    inline void min_max_particles(const ParticleSource& ps, const Chain& chain, const Sizes& sz, float& outMin, float& outMax)
    {
        const CellList cl = bin_particles(ps, sz);
        const int64_t gx = sz.localParticleSize[0], gy = sz.localParticleSize[1], gz = sz.localParticleSize[2];
        std::vector<float> colMin(size_t(gx * gy)), colMax(size_t(gx * gy));
        for(int64_t y = 0; y < gy; y++)
            for(int64_t x = 0; x < gx; x++)
            {
                float mn = std::numeric_limits<float>::max();
                float mx = -std::numeric_limits<float>::max();
                for(int64_t z = 0; z < gz; z++)
                {
                    const int64_t c = x + y * gx + z * gx * gy;
                    for(int64_t k = cl.offset[size_t(c)]; k < cl.offset[size_t(c) + 1]; k++)
                    {
                        const int64_t i = cl.perm[size_t(k)];
                        const float v = apply_chain<float>(chain, ps.attr + i * ps.dim, ps.dim);
                        mn = glm_min(mn, v);
                        mx = glm_max(mx, v);
                    }
                }
                colMin[size_t(x + y * gx)] = mn;
                colMax[size_t(x + y * gx)] = mx;
            }
        outMin = std::numeric_limits<float>::max();
        outMax = -std::numeric_limits<float>::max();
        for(size_t i = 0; i < colMin.size(); i++)
        {
            if(colMin[i] < outMin)
                outMin = colMin[i];
            if(colMax[i] > outMax)
                outMax = colMax[i];
        }
    }

    // --------------------------------------------------------------- G-buffer
    template<class R>
    struct GBuffer
    {
        int W = 0, H = 0;
        std::vector<B4> color;
        std::vector<R> depth;
        std::vector<V3<R>> normal;
        std::vector<R> ao;
    };

    template<class R>
    struct Frame
    {
        M4<R> invMVP;
        V4<R> bgColor; // channel reversed (lib/isaac.hpp:2245-2249)
    };

    template<class R>
    inline Frame<R> make_frame(const State& st)
    {
        Frame<R> f;
        // drawCallBack: matrices converted to isaac_float, inverse in float (lib/isaac.hpp:2174-2178)
        M4<R> P, MV;
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
            {
                P.m[c][r] = R(st.proj[c * 4 + r]);
                MV.m[c][r] = R(st.mv[c * 4 + r]);
            }
        f.invMVP = mat_inverse(mat_mul(P, MV));
        f.bgColor = V4<R>{R(st.bg[3]), R(st.bg[2]), R(st.bg[1]), R(st.bg[0])};
        return f;
    }

    template<class R, class F>
    inline void for_viewport(const State& st, F&& fn)
    {
        const int x0 = st.viewport[0], y0 = st.viewport[1];
        const int x1 = std::min(st.fbW, st.viewport[0] + st.viewport[2]);
        const int y1 = std::min(st.fbH, st.viewport[1] + st.viewport[3]);
#pragma omp parallel for schedule(dynamic, 1)
        for(int y = y0; y < y1; y++)
            for(int x = x0; x < x1; x++)
                fn(x, y);
    }

    // ClearBufferKernel (lib/isaac/isaac_common_kernel.hpp:195-217)
    template<class R>
    inline void pass_clear(const State& st, const Frame<R>& fr, GBuffer<R>& g)
    {
        V4<R> bg = fr.bgColor;
        bg.w = R(0);
        B4 bgb = to_byte(bg);
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                g.color[p] = bgb;
                g.normal[p] = V3<R>{0, 0, 0};
                g.depth[p] = std::numeric_limits<R>::max();
                g.ao[p] = R(0);
            });
    }

    // The walk of one ray (ParticleRenderKernel, lib/isaac/isaac_particle_kernel.hpp:117-260):
    // 3D-DDA over the particle cell grid from the clipped ray start; the walk
    // ends in the first cell that yields a visible hit.  On a hit returns true with
    // the RGBA8 colour, the normal and the depth (from the ray origin).
    template<class R>
    inline bool particle_ray(
        const State& st,
        const std::vector<CellList>& lists,
        const Ray<R>& ray,
        B4& colorOut,
        V3<R>& normalOut,
        R& depthOut,
        uint64_t& myCells,
        uint64_t& myTests)
    {
        const int nv = (int) st.vol.size();
        const int np = (int) st.part.size();
        const Sizes& sz = st.sizes;
        // scale = localSizeScaled / localParticleSize (lib/isaac.hpp:2312)
        V3<R> pscale;
        for(int i = 0; i < 3; i++)
            pscale[i] = R(float(sz.localSizeScaled[i]) / float(sz.localParticleSize[i]));
        const int64_t gx = sz.localParticleSize[0], gy = sz.localParticleSize[1], gz = sz.localParticleSize[2];
        V4<R> pColor{0, 0, 0, 0};
        R depth = std::numeric_limits<R>::max();
        bool hit = false;
        int32_t ds[3] = {(int32_t) gsign(ray.dir.x), (int32_t) gsign(ray.dir.y), (int32_t) gsign(ray.dir.z)};
        V3<R> cur = ray.start + ray.dir * ray.startDepth;
        uint32_t cell[3];
        V3<R> cdiv = cur / pscale;
        for(int i = 0; i < 3; i++)
        {
            int32_t c = (int32_t) cdiv[i];
            int32_t hi = (int32_t) (sz.localParticleSize[i] - 1);
            c = std::min(std::max(c, int32_t(0)), hi);
            cell[i] = (uint32_t) c;
        }
        R rayLength = ray.endDepth - ray.startDepth;
        R tested = R(0);
        V3<R> t, dt;
        for(int i = 0; i < 3; i++)
        {
            t[i] = ((R(cell[i]) + R(std::max(ds[i], 0))) * pscale[i] - cur[i]) / ray.dir[i];
            dt[i] = pscale[i] / ray.dir[i] * R(ds[i]);
        }
        for(int i = 0; i < 3; i++)
            if(ray.dir[i] == R(0))
                t[i] = std::numeric_limits<R>::max();
        V3<R> normal{0, 0, 0};
        while(cell[0] < (uint64_t) gx && cell[1] < (uint64_t) gy && cell[2] < (uint64_t) gz && !hit && tested <= rayLength)
        {
            myCells++;
            const int64_t cid = cell[0] + cell[1] * gx + cell[2] * gx * gy;
            for(int s = 0; s < np; s++)
            {
                const float w = st.weight[nv + s];
                if(w == 0.0f)
                    continue;
                const ParticleSource& ps = st.part[s];
                const CellList& cl = lists[s];
                for(int64_t k = cl.offset[cid]; k < cl.offset[cid + 1]; k++)
                {
                    myTests++;
                    const int64_t pi = cl.perm[k];
                    V3<R> rel{R(cl.rel[3 * k + 0]), R(cl.rel[3 * k + 1]), R(cl.rel[3 * k + 2])};
                    V3<R> ppos = (rel + V3<R>{R(cell[0]), R(cell[1]), R(cell[2])}) * pscale;
                    V3<R> L = ppos - cur;
                    R radius = R(ps.radius ? ps.radius[pi] : ps.const_radius) * R(w);
                    R r2 = radius * radius;
                    R tca = dot(L, ray.dir);
                    R d2 = dot(L, L) - tca * tca;
                    if(d2 <= r2)
                    {
                        R thc = std::sqrt(r2 - d2);
                        R t0 = tca - thc;
                        R t1 = tca + thc;
                        if(t1 >= R(0) && t0 < depth)
                        {
                            R v = apply_chain<R>(st.chain[nv + s], ps.attr + pi * ps.dim, ps.dim);
                            V4<R> val = lut_at<R>(st.lut[nv + s], tf_index(v, st.T));
                            if(val.w >= R(0.5))
                            {
                                pColor = val;
                                depth = t0;
                                hit = true;
                                normal = cur + t0 * ray.dir - ppos;
                                if(t0 < R(0) && ray.isClipped)
                                    normal = -ray.clippingNormal;
                            }
                        }
                    }
                }
            }
            if(t.x < t.y && t.x < t.z)
            {
                cell[0] += (uint32_t) ds[0];
                tested = t.x;
                t.x = t.x + dt.x;
            }
            else if(t.y < t.x && t.y < t.z)
            {
                cell[1] += (uint32_t) ds[1];
                tested = t.y;
                t.y = t.y + dt.y;
            }
            else
            {
                cell[2] += (uint32_t) ds[2];
                tested = t.z;
                t.z = t.z + dt.z;
            }
        }
        if(!hit)
            return false;
        pColor.w = R(1);
        colorOut = to_byte(pColor);
        normalOut = normalize(normal);
        depthOut = depth + ray.startDepth;
        return true;
    }

    inline bool any_particles(const State& st)
    {
        const int nv = (int) st.vol.size();
        bool any = false;
        for(size_t s = 0; s < st.part.size(); s++)
            any |= st.weight[nv + s] != 0.0f;
        return any;
    }

    // ParticleRenderKernel + MergeParticleSourceIterator (lib/isaac/isaac_particle_kernel.hpp:32-260)
    template<class R>
    inline void pass_particles(
        const State& st,
        const Frame<R>& fr,
        const std::vector<CellList>& lists,
        GBuffer<R>& g,
        Counters& cnt)
    {
        if(!any_particles(st))
            return;
        const Sizes& sz = st.sizes;
        uint64_t cells = 0, tests = 0, rays = 0;
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                Ray<R> ray = pixel_to_ray(fr.invMVP, sz, R(x), R(y), R(g.W), R(g.H));
                if(!clip_ray(ray, sz, st.clip))
                    return;
                ray.endDepth = gmin(ray.endDepth, g.depth[p]);
                uint64_t myCells = 0, myTests = 0;
                B4 col;
                V3<R> n;
                R d;
                const bool hit = particle_ray(st, lists, ray, col, n, d, myCells, myTests);
#pragma omp atomic
                cells += myCells;
#pragma omp atomic
                tests += myTests;
#pragma omp atomic
                rays += 1;
                if(!hit)
                    return;
                g.color[p] = col;
                g.normal[p] = n;
                g.depth[p] = d;
            });
        cnt.cellsVisited += cells;
        cnt.sphereTests += tests;
        cnt.partRays += rays;
    }

    inline void scaled(const float s[3], float out[3])
    {
        out[0] = s[0];
        out[1] = s[1];
        out[2] = s[2];
    }

    // CombinedIsoRenderKernel (lib/isaac/isaac_iso_kernel.hpp:25-128)
    template<class R>
    inline void pass_iso(const State& st, const Frame<R>& fr, const Tex3B4& isoTex, GBuffer<R>& g, Counters& cnt)
    {
        const Sizes& sz = st.sizes;
        const V3<R> scale{R(st.scale[0]), R(st.scale[1]), R(st.scale[2])};
        uint64_t samples = 0, hits = 0, rays = 0;
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                Ray<R> ray = pixel_to_ray(fr.invMVP, sz, R(x), R(y), R(g.W), R(g.H));
                if(!clip_ray(ray, sz, st.clip))
                    return;
                ray.endDepth = gmin(ray.endDepth, g.depth[p]);
                if(ray.endDepth <= ray.startDepth)
                    return;
                R stepU = R(st.step) * (length(ray.dir) / length(ray.dir / scale));
                int32_t startSteps = (int32_t) std::ceil(ray.startDepth / stepU);
                int32_t endSteps = (int32_t) std::floor(ray.endDepth / stepU);
                V3<R> stepVec = stepU * ray.dir / scale;
                V3<R> startU = ray.start / scale;
                bool hit = false;
                R depth = ray.endDepth;
                V4<R> hitColor{0, 0, 0, 0};
                V3<R> hitNormal{0, 0, 0};
                R oldValue = R(0);
                uint64_t my = 0;
                for(int32_t i = startSteps; i <= endSteps && !hit; i++)
                {
                    V3<R> pos = startU + stepVec * R(i);
                    bool first = ray.isClipped && i == startSteps;
                    R tt = R(i) * stepU;
                    R value = sample_b4<R>(isoTex, pos, st.interpolation).w;
                    my++;
                    R tmp = oldValue;
                    oldValue = value;
                    if(value < R(0.5))
                        continue;
                    if(first)
                        depth = ray.startDepth;
                    else
                        depth = tt + stepU * (R(0.5) - tmp) / (value - tmp);
                    hit = true;
                    V3<R> np = ray.start + ray.dir * depth;
                    V3<R> pu = np / scale;
                    hitColor = sample_b4<R>(isoTex, pu, st.interpolation);
                    hitColor.w = R(1);
                    V3<R> grad;
                    grad.x = sample_b4<R>(isoTex, pu + V3<R>{1, 0, 0}, st.interpolation).w
                        - sample_b4<R>(isoTex, pu - V3<R>{1, 0, 0}, st.interpolation).w;
                    grad.y = sample_b4<R>(isoTex, pu + V3<R>{0, 1, 0}, st.interpolation).w
                        - sample_b4<R>(isoTex, pu - V3<R>{0, 1, 0}, st.interpolation).w;
                    grad.z = sample_b4<R>(isoTex, pu + V3<R>{0, 0, 1}, st.interpolation).w
                        - sample_b4<R>(isoTex, pu - V3<R>{0, 0, 1}, st.interpolation).w;
                    R gl = length(grad);
                    if(first || gl == R(0))
                    {
                        grad = ray.clippingNormal;
                        gl = R(1);
                    }
                    hitNormal = (-grad) / gl;
                }
#pragma omp atomic
                samples += my;
#pragma omp atomic
                rays += 1;
                if(hit)
                {
#pragma omp atomic
                    hits += 1;
                    g.color[p] = to_byte(hitColor);
                    g.normal[p] = hitNormal;
                    g.depth[p] = depth;
                }
            });
        cnt.isoSamples += samples;
        cnt.isoHits += hits;
        cnt.isoRays += rays;
    }

    // SSAOKernel live part (lib/isaac/isaac_ssao_kernel.hpp:161-182): 7x7 taps at
    // 10-pixel stride, CLAMP border, count of closer taps / 42.
    template<class R>
    inline void pass_ssao(const State& st, GBuffer<R>& g)
    {
        std::vector<R> depth = g.depth;
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                R ref = depth[p];
                R occ = R(0);
                for(int i = -3; i <= 3; ++i)
                    for(int j = -3; j <= 3; ++j)
                    {
                        int sx = std::min(std::max(x + i * 10, 0), g.W - 1);
                        int sy = std::min(std::max(y + j * 10, 0), g.H - 1);
                        if(depth[size_t(sy) * g.W + sx] < ref)
                            occ = occ + R(1);
                    }
                g.ao[p] = gclamp(occ / R(42), R(0), R(1));
            });
    }

    // getHSVA (lib/isaac/isaac_helper.hpp:221-264); h converted to float by the caller
    template<class R>
    inline V4<R> hsva(R h, R s, R v, R a)
    {
        const double PI3 = M_PI / 3.0;
        int hi = (int) std::floor(double(h) / PI3);
        R f = R(double(h) / PI3 - double(R(hi)));
        R pp = v * (R(1) - s);
        R q = v * (R(1) - s * f);
        R t = v * (R(1) - s * (R(1) - f));
        V4<R> r{0, 0, 0, a};
        switch(hi)
        {
        case 0:
        case 6:
            r.x = v, r.y = t, r.z = pp;
            break;
        case 1:
            r.x = q, r.y = v, r.z = pp;
            break;
        case 2:
            r.x = pp, r.y = v, r.z = t;
            break;
        case 3:
            r.x = pp, r.y = q, r.z = v;
            break;
        case 4:
            r.x = t, r.y = pp, r.z = v;
            break;
        case 5:
            r.x = v, r.y = pp, r.z = q;
            break;
        }
        return r;
    }

    // ShadingKernel (lib/isaac/isaac_common_kernel.hpp:220-320)
    template<class R>
    inline void pass_shading(const State& st, const Frame<R>& fr, GBuffer<R>& g)
    {
        const Sizes& sz = st.sizes;
        const int mode = st.renderMode;
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                V4<R> color = from_byte<R>(g.color[p]);
                V3<R> normal = g.normal[p];
                R aoStrength = R(1) - g.ao[p];
#if ORACLE_MUTANT != 5
                if(mode == 7)
                    color = hsva<R>(R(halton(st.rank, 3) * float(2 * M_PI)), 1, 1, color.w);
#endif
                if(mode < 3 || mode == 7)
                {
                    Ray<R> ray = pixel_to_ray(fr.invMVP, sz, R(x), R(y), R(g.W), R(g.H));
                    V3<R> lightDir = -ray.dir;
                    R lightFactor = std::abs(dot(normal, lightDir));
                    V3<R> half = normalize(-ray.dir + lightDir);
                    R spec = dot(normal, half);
                    spec = R(std::pow(double(spec), 4.0));
                    spec = spec * R(0.5);
                    if(mode == 1)
                        spec = R(0);
                    R w = R(st.aoWeight);
                    R aoF = (R(1) - w) + w * aoStrength;
                    lightFactor = lightFactor + aoF;
                    lightFactor = lightFactor * R(0.5);
                    V4<R> sc = color * lightFactor;
                    V3<R> shaded{gmin(sc.x + spec, R(1)), gmin(sc.y + spec, R(1)), gmin(sc.z + spec, R(1))};
                    g.color[p] = to_byte(V4<R>{shaded.x, shaded.y, shaded.z, color.w});
#if ORACLE_MUTANT == 6
                    if(mode == 7)
#else
                    if(mode == 2 || mode == 7)
#endif
                        g.depth[p] = R(0);
                }
                else if(mode == 3)
                {
                    V4<R> bg = fr.bgColor;
                    bg.w = color.w;
                    g.color[p] = to_byte(bg);
                }
                else if(mode == 4)
                {
                    V3<R> n = normal * R(0.5) + V3<R>{R(0.5), R(0.5), R(0.5)};
                    g.color[p] = to_byte(V4<R>{n.x, n.y, n.z, color.w});
                    g.depth[p] = R(0);
                }
                else if(mode == 5)
                {
                    R d = g.depth[p] / R(sz.maxGlobalSizeScaled);
                    g.color[p] = to_byte(V4<R>{d, d, d, color.w});
                    g.depth[p] = R(0);
                }
                else if(mode == 6)
                {
                    R w = R(st.aoWeight);
                    R aoF = (R(1) - w) + w * aoStrength;
                    g.color[p] = to_byte(V4<R>{aoF, aoF, aoF, color.w});
                    g.depth[p] = R(0);
                }
                else if(mode == 8)
                {
                    g.color[p] = to_byte(hsva<R>(R(halton(st.rank, 3) * float(2 * M_PI)), 1, 1, 1));
                    g.depth[p] = R(0);
                }
            });
    }

    // CombinedVolumeRenderKernel (lib/isaac/isaac_volume_kernel.hpp:23-88); totalWeight
    // over ALL sources incl. particles (lib/isaac.hpp:2449-2453).
    template<class R>
    inline void pass_volume(const State& st, const Frame<R>& fr, const Tex3B4& volTex, GBuffer<R>& g, Counters& cnt)
    {
        const Sizes& sz = st.sizes;
        float twf = 0.0f;
#if ORACLE_MUTANT == 2
        for(size_t s = 0; s < st.vol.size(); s++)
            twf += st.weight[s];
#else
        for(float w : st.weight)
            twf += w;
#endif
        if(!(twf > 0.0f))
            return;
        const R totalWeight = R(twf);
        const V3<R> scale{R(st.scale[0]), R(st.scale[1]), R(st.scale[2])};
        uint64_t samples = 0, rays = 0;
        for_viewport<R>(
            st,
            [&](int x, int y)
            {
                size_t p = size_t(y) * g.W + x;
                Ray<R> ray = pixel_to_ray(fr.invMVP, sz, R(x), R(y), R(g.W), R(g.H));
                if(!clip_ray(ray, sz, st.clip))
                    return;
                ray.endDepth = gmin(ray.endDepth, g.depth[p]);
                if(ray.endDepth <= ray.startDepth)
                    return;
                R minSize = R(std::min(int(sz.globalSize[0]), std::min(int(sz.globalSize[1]), int(sz.globalSize[2]))));
                R stepU = R(st.step) * (length(ray.dir) / length(ray.dir / scale));
                R factor = stepU / minSize * R(2) * totalWeight;
                int32_t startSteps = (int32_t) std::ceil(ray.startDepth / stepU);
                int32_t endSteps = (int32_t) std::floor(ray.endDepth / stepU);
                V3<R> stepVec = stepU * ray.dir / scale;
                V3<R> startU = ray.start / scale;
                V4<R> color{0, 0, 0, 0};
                uint64_t my = 0;
                for(int32_t i = startSteps; i <= endSteps; i++)
                {
                    V3<R> pos = startU + stepVec * R(i);
                    V4<R> value = sample_b4<R>(volTex, pos, st.interpolation);
                    my++;
                    value = value * factor;
                    R weight = R(1) - color.w;
                    color = color + weight * value;
                    if(color.w > R(0.99))
                        break;
                }
                V4<R> solid = from_byte<R>(g.color[p]);
                color = color + (R(1) - color.w) * solid;
                g.color[p] = to_byte(color);
#pragma omp atomic
                samples += my;
#pragma omp atomic
                rays += 1;
            });
        cnt.volSamples += samples;
        cnt.volRays += rays;
    }

    // Whole drawCallBack + classification (lib/isaac.hpp:1916-1947, 2140-2543).
    template<class R>
    struct FrameOut
    {
        GBuffer<R> g;
        std::vector<B4> pre; // colour after particles + iso, before shading
        std::vector<R> preDepth; // depth before shading
        std::vector<B4> solid; // colour after shading, before the volume pass
        Counters counters;
    };

    template<class R>
    inline void render(
        const State& st,
        const Tex3B4& volTex,
        const Tex3B4& isoTex,
        const std::vector<CellList>& lists,
        FrameOut<R>& out)
    {
        GBuffer<R>& g = out.g;
        g.W = st.fbW;
        g.H = st.fbH;
        const size_t N = size_t(st.fbW) * st.fbH;
        if(g.color.size() != N)
        {
            g.color.assign(N, B4{0, 0, 0, 0});
            g.depth.assign(N, R(0));
            g.normal.assign(N, V3<R>{0, 0, 0});
            g.ao.assign(N, R(0));
        }
        Frame<R> fr = make_frame<R>(st);
        pass_clear(st, fr, g);
        pass_particles(st, fr, lists, g, out.counters);
        bool anyIso = false;
        for(size_t s = 0; s < st.vol.size(); s++)
            anyIso |= st.iso[s] > 0.0f;
        if(anyIso)
            pass_iso(st, fr, isoTex, g, out.counters);
        if(st.aoEnabled && st.aoWeight > 0.0f)
            pass_ssao(st, g);
        out.pre = g.color;
        out.preDepth = g.depth;
        pass_shading(st, fr, g);
        out.solid = g.color;
        pass_volume(st, fr, volTex, g, out.counters);
    }
} // namespace isaac_oracle
