// CPU ORACLE — TEST INFRASTRUCTURE ONLY.  Checker for the library's brick
// (post-classification) render form; never part of the product path.
//
// The brick form has no counterpart among the reference's default kernels.  Its
// semantics are those of the reference's legacy renderer
// (lib/isaac/isaac_volume_kernel.hpp:101-168, lib/isaac/isaac_iso_kernel.hpp:131-300:
// interpolate the raw source, then functor chain and transfer function per
// sample), with the sources blended per sample as the optimized classification
// blends them per voxel (lib/isaac/isaac_common_kernel.hpp:574-612) but without
// the RGBA8 rounding.  Pass order, ray setup, step bounds, compositing and
// shading are those of the optimized renderer restated in isaac_oracle.hpp.
//
// Two precisions:
//   RS: ray setup (camera inverse, pixelToRay, clipRay, step bounds) — double
//   R : sampling, classification and compositing — float (the library's FP32
//       arithmetic) or double (the FP64 reference image)
// Field representation (dtype):
//   RAW  (-1) the float source itself (the FP64 reference)
//   F32  ( 0) bricks holding the float source, scale 1
//   F16  ( 1) bricks holding RN_f16(v / s), s = max |v| over the brick's 9^3 region
//   BF16 ( 2) bricks holding RN_bf16(v / s)
// A sample interpolates the stored values of its 8 taps, all taken from the brick
// of the clamped (x0, y0, z0) tap (so the apron voxels carry that brick's scale),
// and multiplies by that brick's scale.
#pragma once

#include "isaac_oracle.hpp"

#include <cstring>
#include <stdexcept>

namespace isaac_oracle
{
    enum BrickDtype
    {
        BRICK_RAW = -1,
        BRICK_F32 = 0,
        BRICK_F16 = 1,
        BRICK_BF16 = 2
    };

    // IEEE binary16 round-to-nearest-even from binary32, written from the format
    // definition (bias 15, 10 fraction bits, subnormals, overflow to infinity).
    inline uint16_t f32_to_f16_rne(float f)
    {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        const uint32_t sign = (u >> 16) & 0x8000u;
        const uint32_t a = u & 0x7fffffffu;
        if(a >= 0x7f800000u)
            return uint16_t(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
        if(a >= 0x477ff000u) // >= 65520: rounds to infinity
            return uint16_t(sign | 0x7c00u);
        if(a < 0x38800000u) // below 2^-14: subnormal result, exact scaling by 2^24 then RNE
        {
            float v;
            std::memcpy(&v, &a, 4);
            const uint32_t m = (uint32_t) std::nearbyint(v * 16777216.0f);
            return uint16_t(sign | m);
        }
        const uint32_t e = (a >> 23) - 127u + 15u;
        const uint32_t mant = a & 0x7fffffu;
        uint32_t h = (e << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if(rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
            h++;
        return uint16_t(sign | h);
    }

    inline float f16_to_f32(uint16_t h)
    {
        const uint32_t sign = uint32_t(h & 0x8000u) << 16;
        const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
        float v;
        if(e == 0)
            v = std::ldexp(float(m), -24);
        else if(e == 31)
            v = m ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
        else
            v = std::ldexp(float(m | 0x400u), int(e) - 25);
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u |= sign;
        std::memcpy(&v, &u, 4);
        return v;
    }

    // bfloat16 = the upper 16 bits of binary32, rounded to nearest even
    inline uint16_t f32_to_bf16_rne(float f)
    {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        if((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu))
            return uint16_t((u >> 16) | 0x40u);
        const uint32_t lsb = (u >> 16) & 1u;
        return uint16_t((u + 0x7fffu + lsb) >> 16);
    }

    inline float bf16_to_f32(uint16_t b)
    {
        const uint32_t u = uint32_t(b) << 16;
        float v;
        std::memcpy(&v, &u, 4);
        return v;
    }

    inline float quantize(float v, float s, int dtype)
    {
        if(dtype == BRICK_F16)
            return f16_to_f32(f32_to_f16_rne(v / s));
        if(dtype == BRICK_BF16)
            return bf16_to_f32(f32_to_bf16_rne(v / s));
        return v;
    }

    // One field as the brick form sees it: a source (raw or functor-applied) of
    // `dim` components plus per-brick scales.
    struct BrickField
    {
        const float* data = nullptr; // voxel (0,0,0), AoS dim floats per voxel
        int64_t stride[3] = {0, 0, 0};
        int dim = 1;
        int dtype = BRICK_RAW;
        int64_t X = 0, Y = 0, Z = 0, nbx = 0, nby = 0, nbz = 0;
        std::vector<float> scale; // [brick][component]

        float at(int64_t x, int64_t y, int64_t z, int c) const
        {
            return data[(x * stride[0] + y * stride[1] + z * stride[2]) * dim + c];
        }
    };

    inline BrickField make_brick_field(const float* data, const int64_t stride[3], int dim, const Sizes& sz, int dtype)
    {
        BrickField f;
        f.data = data;
        for(int k = 0; k < 3; k++)
            f.stride[k] = stride[k];
        f.dim = dim;
        f.dtype = dtype;
        f.X = sz.localSize[0];
        f.Y = sz.localSize[1];
        f.Z = sz.localSize[2];
        f.nbx = (f.X + 7) / 8;
        f.nby = (f.Y + 7) / 8;
        f.nbz = (f.Z + 7) / 8;
        const int64_t nb = f.nbx * f.nby * f.nbz;
        f.scale.assign(size_t(nb * dim), 1.0f);
        if(dtype == BRICK_F16 || dtype == BRICK_BF16)
        {
#pragma omp parallel for schedule(static)
            for(int64_t b = 0; b < nb; b++)
            {
                const int64_t bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
                for(int c = 0; c < dim; c++)
                {
                    float m = 0.0f;
                    for(int lz = 0; lz < 9; lz++)
                        for(int ly = 0; ly < 9; ly++)
                            for(int lx = 0; lx < 9; lx++)
                            {
                                const int64_t x = std::min(bx * 8 + lx, f.X - 1), y = std::min(by * 8 + ly, f.Y - 1);
                                const int64_t z = std::min(bz * 8 + lz, f.Z - 1);
                                m = std::max(m, std::fabs(f.at(x, y, z, c)));
                            }
                    f.scale[size_t(b * dim + c)] = m > 0.0f ? m : 1.0f;
                }
            }
        }
        return f;
    }

    // Sampler<LINEAR|NEAREST, CLAMP> of all components of a brick field at a
    // position in voxel units (lib/isaac/isaac_texture.hpp:182-259 addressing;
    // trilinear as z innermost, x outermost, lib/isaac/isaac_helper.hpp:124-144).
    template<class R>
    inline void sample_field(const BrickField& f, V3<R> pos, bool linear, R out[4])
    {
        auto cl = [](int64_t v, int64_t n) { return std::min(std::max(v, int64_t(0)), n - 1); };
        if(!linear)
        {
            const int64_t x = cl((int64_t) (int32_t) pos.x, f.X), y = cl((int64_t) (int32_t) pos.y, f.Y);
            const int64_t z = cl((int64_t) (int32_t) pos.z, f.Z);
            const int64_t b = ((z / 8) * f.nby + y / 8) * f.nbx + x / 8;
            for(int c = 0; c < f.dim; c++)
            {
                if(f.dtype == BRICK_RAW)
                    out[c] = R(f.at(x, y, z, c));
                else
                {
                    const float s = f.scale[size_t(b * f.dim + c)];
                    out[c] = R(quantize(f.at(x, y, z, c), s, f.dtype)) * R(s);
                }
            }
            return;
        }
        const V3<R> cpos{pos.x - R(0.5), pos.y - R(0.5), pos.z - R(0.5)};
        const R flx = std::floor(cpos.x), fly = std::floor(cpos.y), flz = std::floor(cpos.z);
        const R fx = cpos.x - flx, fy = cpos.y - fly, fz = cpos.z - flz;
        const int64_t bx = (int64_t) (int32_t) flx, by = (int64_t) (int32_t) fly, bz = (int64_t) (int32_t) flz;
        const int64_t x0 = cl(bx, f.X), x1 = cl(bx + 1, f.X), y0 = cl(by, f.Y), y1 = cl(by + 1, f.Y);
        const int64_t z0 = cl(bz, f.Z), z1 = cl(bz + 1, f.Z);
        const int64_t b = ((z0 / 8) * f.nby + y0 / 8) * f.nbx + x0 / 8;
        const R ox = R(1) - fx, oy = R(1) - fy, oz = R(1) - fz;
        for(int c = 0; c < f.dim; c++)
        {
            const float s = f.dtype == BRICK_RAW ? 1.0f : f.scale[size_t(b * f.dim + c)];
            auto d = [&](int64_t x, int64_t y, int64_t z)
            {
                const float v = f.at(x, y, z, c);
                return R(f.dtype == BRICK_RAW ? v : quantize(v, s, f.dtype));
            };
            const R a0 = d(x0, y0, z0) * oz + d(x0, y0, z1) * fz;
            const R a1 = d(x0, y1, z0) * oz + d(x0, y1, z1) * fz;
            const R A = a0 * oy + a1 * fy;
            const R b0 = d(x1, y0, z0) * oz + d(x1, y0, z1) * fz;
            const R b1 = d(x1, y1, z0) * oz + d(x1, y1, z1) * fz;
            const R B = b0 * oy + b1 * fy;
            const R v = A * ox + B * fx;
            out[c] = f.dtype == BRICK_RAW ? v : v * R(s);
        }
    }

    // Per-sample classification: volume colour premultiplied and normalised by the
    // volume-source weight sum; iso colour and value max_s(alpha_s / thr_s * 0.5).
    template<class R>
    struct PostSample
    {
        V4<R> vol;
        V4<R> iso; // rgb colour, w = iso value (not yet clamped)
    };

    struct PostScene
    {
        const State* st = nullptr;
        std::vector<BrickField> fields; // one per volume source (empty data = disabled)
        std::vector<bool> persistent;
    };

    template<class R>
    inline PostSample<R> post_classify(const PostScene& ps, V3<R> pos)
    {
        const State& st = *ps.st;
        const int nv = (int) st.vol.size();
        PostSample<R> o{{0, 0, 0, 0}, {0, 0, 0, 0}};
        float twf = 0.0f;
        for(int s = 0; s < nv; s++)
            twf += st.weight[s];
        for(int s = 0; s < nv; s++)
        {
            if(!(st.weight[s] + st.iso[s] > 0.0f))
                continue;
            const BrickField& f = ps.fields[s];
            R v[4] = {0, 0, 0, 0};
            sample_field<R>(f, pos, st.interpolation, v);
            // persistent: chain on the interpolated raw value; non-persistent: the
            // field already holds the functor-applied scalar
            const R tv = ps.persistent[s] ? apply_chain_vals<R>(st.chain[s], v, f.dim) : v[0];
            const V4<R> col = lut_at<R>(st.lut[s], tf_index(tv, st.T));
            if(st.weight[s] > 0.0f)
            {
                const R a = col.w * R(st.weight[s]);
                o.vol.x = o.vol.x + col.x * a;
                o.vol.y = o.vol.y + col.y * a;
                o.vol.z = o.vol.z + col.z * a;
                o.vol.w = o.vol.w + a;
            }
            if(st.iso[s] > 0.0f)
            {
                // iso value = alpha * (0.5 / threshold), the scale taken once per source
                const R a = col.w * (R(0.5) / R(st.iso[s]));
                if(o.iso.w < a)
                    o.iso = V4<R>{col.x, col.y, col.z, a};
            }
        }
        // normalised by the volume-source weight sum through its reciprocal
        const R inv = twf != 0.0f ? R(1) / R(twf) : R(0);
        o.vol = o.vol * inv;
        return o;
    }

    template<class R>
    inline R post_iso_value(const PostScene& ps, V3<R> pos)
    {
        return gmin(post_classify<R>(ps, pos).iso.w, R(1));
    }

    // a ray in march precision R from the binary64 setup
    template<class R>
    inline Ray<R> to_march_ray(const Ray<double>& d)
    {
        Ray<R> r;
        r.dir = V3<R>{R(d.dir.x), R(d.dir.y), R(d.dir.z)};
        r.start = V3<R>{R(d.start.x), R(d.start.y), R(d.start.z)};
        r.end = V3<R>{R(d.end.x), R(d.end.y), R(d.end.z)};
        r.startDepth = R(d.startDepth);
        r.endDepth = R(d.endDepth);
        r.isClipped = d.isClipped;
        r.clippingNormal = V3<R>{R(d.clippingNormal.x), R(d.clippingNormal.y), R(d.clippingNormal.z)};
        return r;
    }

    struct StepBounds
    {
        double stepUd;
        int32_t startSteps;
    };

    inline StepBounds step_bounds(const State& st, const Ray<double>& ray)
    {
        const V3<double> scale{double(st.scale[0]), double(st.scale[1]), double(st.scale[2])};
        StepBounds b;
        b.stepUd = double(st.step) * (length(ray.dir) / length(ray.dir / scale));
        b.startSteps = (int32_t) std::ceil(ray.startDepth / b.stepUd);
        return b;
    }

    template<class R>
    struct PostOut
    {
        GBuffer<R> g;
        std::vector<V4<R>> finalColor; // before RGBA8 quantisation, clamped to [0, 1]
        std::vector<B4> pre;
        Counters counters;
        int stride = 1;
    };

    // The whole frame in the brick form.  `stride` > 1 renders only pixels with
    // x % stride == 0 and y % stride == 0 (sampled error studies).  With SSAO on, the
    // stride must divide the SSAO tap spacing (10 pixels) and the viewport must be the
    // whole framebuffer; then pixels with (W-1-x) % stride == 0 or (H-1-y) % stride == 0
    // are rendered too, so every depth tap of a rendered pixel (x + 10i clamped to
    // [0, W-1]) is itself a rendered pixel and its occlusion is exact.
    template<class R>
    inline void render_post(const PostScene& ps, const std::vector<CellList>& lists, int stride, PostOut<R>& out)
    {
        const State& st = *ps.st;
        const Sizes& sz = st.sizes;
        GBuffer<R>& g = out.g;
        g.W = st.fbW;
        g.H = st.fbH;
        out.stride = stride;
        const size_t N = size_t(st.fbW) * st.fbH;
        g.color.assign(N, B4{0, 0, 0, 0});
        g.depth.assign(N, R(0));
        g.normal.assign(N, V3<R>{0, 0, 0});
        g.ao.assign(N, R(0));
        out.finalColor.assign(N, V4<R>{0, 0, 0, 0});
        // binary64 camera inverse and rays
        Frame<double> frd = make_frame<double>(st);
        Frame<R> fr = make_frame<R>(st);
        const V3<R> scale{R(st.scale[0]), R(st.scale[1]), R(st.scale[2])};
        const bool ssao = st.aoEnabled && st.aoWeight > 0.0f;
        if(ssao && stride != 1
           && (10 % stride != 0 || st.viewport[0] != 0 || st.viewport[1] != 0 || st.viewport[2] < st.fbW
               || st.viewport[3] < st.fbH))
            throw std::runtime_error("sampled rendering with SSAO needs a stride dividing 10 and the whole framebuffer");
        auto kept = [&](int v, int n) { return v % stride == 0 || (ssao && (n - 1 - v) % stride == 0); };
        auto each = [&](auto&& fn)
        {
            const int x0 = st.viewport[0], y0 = st.viewport[1];
            const int x1 = std::min(st.fbW, st.viewport[0] + st.viewport[2]);
            const int y1 = std::min(st.fbH, st.viewport[1] + st.viewport[3]);
#pragma omp parallel for schedule(dynamic, 1)
            for(int y = y0; y < y1; y++)
            {
                if(!kept(y, st.fbH))
                    continue;
                for(int x = x0; x < x1; x++)
                    if(kept(x, st.fbW))
                        fn(x, y);
            }
        };
        auto ray_of = [&](int x, int y, Ray<double>& rd)
        {
            rd = pixel_to_ray(frd.invMVP, sz, double(x), double(y), double(g.W), double(g.H));
            return clip_ray(rd, sz, st.clip);
        };
        // clear
        V4<R> bg = fr.bgColor;
        bg.w = R(0);
        const B4 bgb = to_byte(bg);
        each(
            [&](int x, int y)
            {
                const size_t p = size_t(y) * g.W + x;
                g.color[p] = bgb;
                g.depth[p] = std::numeric_limits<R>::max();
            });
        // particles
        if(any_particles(st))
        {
            uint64_t cells = 0, tests = 0, rays = 0;
            each(
                [&](int x, int y)
                {
                    Ray<double> rd;
                    if(!ray_of(x, y, rd))
                        return;
                    const size_t p = size_t(y) * g.W + x;
                    Ray<R> ray = to_march_ray<R>(rd);
                    uint64_t c = 0, t = 0;
                    B4 col;
                    V3<R> n;
                    R d;
                    if(particle_ray(st, lists, ray, col, n, d, c, t))
                    {
                        g.color[p] = col;
                        g.normal[p] = n;
                        g.depth[p] = d;
                    }
#pragma omp atomic
                    cells += c;
#pragma omp atomic
                    tests += t;
#pragma omp atomic
                    rays += 1;
                });
            out.counters.cellsVisited += cells;
            out.counters.sphereTests += tests;
            out.counters.partRays += rays;
        }
        // iso
        bool anyIso = false;
        for(size_t s = 0; s < st.vol.size(); s++)
            anyIso |= st.iso[s] > 0.0f;
        if(anyIso)
        {
            uint64_t samples = 0, hits = 0, rays = 0;
            each(
                [&](int x, int y)
                {
                    Ray<double> rd;
                    if(!ray_of(x, y, rd))
                        return;
                    const size_t p = size_t(y) * g.W + x;
                    const double endD = std::min(rd.endDepth, double(g.depth[p]));
                    if(!(endD > rd.startDepth))
                        return;
                    const Ray<R> ray = to_march_ray<R>(rd);
                    const StepBounds sb = step_bounds(st, rd);
                    const R stepU = R(sb.stepUd);
                    const int32_t endSteps = (int32_t) std::floor(endD / sb.stepUd);
                    const V3<R> stepVec = stepU * ray.dir / scale;
                    const V3<R> startU = ray.start / scale;
                    R oldValue = R(0);
                    uint64_t my = 0;
                    bool hit = false;
                    for(int32_t i = sb.startSteps; i <= endSteps && !hit; i++)
                    {
                        const V3<R> pos = startU + stepVec * R(i);
                        const R value = post_iso_value<R>(ps, pos);
                        my++;
                        const R tmp = oldValue;
                        oldValue = value;
                        if(value < R(0.5))
                            continue;
                        const bool first = ray.isClipped && i == sb.startSteps;
                        const R depth = first ? ray.startDepth : R(i) * stepU + stepU * (R(0.5) - tmp) / (value - tmp);
                        hit = true;
                        const V3<R> pu = (ray.start + ray.dir * depth) / scale;
                        V4<R> hc = post_classify<R>(ps, pu).iso;
                        hc.w = R(1);
                        V3<R> grad;
                        grad.x = post_iso_value<R>(ps, pu + V3<R>{1, 0, 0}) - post_iso_value<R>(ps, pu - V3<R>{1, 0, 0});
                        grad.y = post_iso_value<R>(ps, pu + V3<R>{0, 1, 0}) - post_iso_value<R>(ps, pu - V3<R>{0, 1, 0});
                        grad.z = post_iso_value<R>(ps, pu + V3<R>{0, 0, 1}) - post_iso_value<R>(ps, pu - V3<R>{0, 0, 1});
                        R gl = length(grad);
                        if(first || gl == R(0))
                        {
                            grad = ray.clippingNormal;
                            gl = R(1);
                        }
                        g.color[p] = to_byte(hc);
                        g.normal[p] = (-grad) / gl;
                        g.depth[p] = depth;
                    }
#pragma omp atomic
                    samples += my;
#pragma omp atomic
                    rays += 1;
                    if(hit)
                    {
#pragma omp atomic
                        hits += 1;
                    }
                });
            out.counters.isoSamples += samples;
            out.counters.isoHits += hits;
            out.counters.isoRays += rays;
        }
        if(ssao)
            pass_ssao(st, g);
        out.pre = g.color;
        pass_shading(st, fr, g);
        // volume
        float twf = 0.0f;
        for(float w : st.weight)
            twf += w;
        const R totalWeight = R(twf);
        const R minSize = R(std::min(int(sz.globalSize[0]), std::min(int(sz.globalSize[1]), int(sz.globalSize[2]))));
        uint64_t vsamples = 0, vrays = 0;
        each(
            [&](int x, int y)
            {
                const size_t p = size_t(y) * g.W + x;
                const V4<R> solid = from_byte<R>(g.color[p]);
                V4<R> fin = solid;
                Ray<double> rd;
                if(twf > 0.0f && ray_of(x, y, rd))
                {
                    const double endD = std::min(rd.endDepth, double(g.depth[p]));
                    if(endD > rd.startDepth)
                    {
                        const Ray<R> ray = to_march_ray<R>(rd);
                        const StepBounds sb = step_bounds(st, rd);
                        const R stepU = R(sb.stepUd);
                        const R factor = stepU / minSize * R(2) * totalWeight;
                        const int32_t endSteps = (int32_t) std::floor(endD / sb.stepUd);
                        const V3<R> stepVec = stepU * ray.dir / scale;
                        const V3<R> startU = ray.start / scale;
                        V4<R> color{0, 0, 0, 0};
                        uint64_t my = 0;
                        for(int32_t i = sb.startSteps; i <= endSteps; i++)
                        {
                            const V3<R> pos = startU + stepVec * R(i);
                            const V4<R> value = post_classify<R>(ps, pos).vol * factor;
                            my++;
                            const R w = R(1) - color.w;
                            color = color + w * value;
                            if(color.w > R(0.99))
                                break;
                        }
                        fin = color + (R(1) - color.w) * solid;
                        g.color[p] = to_byte(fin);
#pragma omp atomic
                        vsamples += my;
#pragma omp atomic
                        vrays += 1;
                    }
                }
                out.finalColor[p] = V4<R>{gclamp(fin.x, R(0), R(1)), gclamp(fin.y, R(0), R(1)), gclamp(fin.z, R(0), R(1)),
                                          gclamp(fin.w, R(0), R(1))};
            });
        out.counters.volSamples += vsamples;
        out.counters.volRays += vrays;
    }
} // namespace isaac_oracle
