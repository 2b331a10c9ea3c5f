// CPU ORACLE C entry points — TEST INFRASTRUCTURE ONLY (loaded by tests/, the
// smoke check and the benchmark's CPU-baseline leg through ctypes).
#include "post_oracle.hpp"

#include <chrono>
#include <omp.h>

using namespace isaac_oracle;


// Flat scene description shared with oracle/oracle.py (ctypes mirror).
struct OracleScene
{
    int64_t global_size[3];
    int64_t local_size[3];
    int64_t particle_grid[3];
    int32_t position[3];
    float scale[3];
    int32_t fb_w, fb_h;
    int32_t viewport[4];
    int32_t transfer_size;
    int32_t n_volume;
    int32_t n_particle;
    const float* vol_data[MAX_SOURCES];
    int32_t vol_dim[MAX_SOURCES];
    int64_t vol_stride[MAX_SOURCES][3];
    const float* part_pos[MAX_SOURCES];
    const float* part_attr[MAX_SOURCES];
    int32_t part_dim[MAX_SOURCES];
    const float* part_radius[MAX_SOURCES];
    float part_const_radius[MAX_SOURCES];
    int64_t part_count[MAX_SOURCES];
    const float* lut[2 * MAX_SOURCES];
    int32_t bytecode[2 * MAX_SOURCES][MAX_FUNCTORS];
    float params[2 * MAX_SOURCES][MAX_FUNCTORS][4];
    float weight[2 * MAX_SOURCES];
    float iso[2 * MAX_SOURCES];
    float step;
    int32_t interpolation;
    int32_t dither_mode;
    int32_t ao_enabled;
    float ao_weight;
    int32_t render_mode;
    int32_t n_clip;
    float clip_pos[MAX_CLIPPING][3];
    float clip_normal[MAX_CLIPPING][3];
    double proj[16];
    double mv[16];
    float bg[4];
    int32_t threads; // 0 = OpenMP default
    int32_t vol_nonpersistent[MAX_SOURCES]; // brick form: 1 = sample the functor-applied field
    const uint8_t* vol_adv[MAX_SOURCES]; // field sources: advection texture (X*Y*Z bytes), else null
};

struct OraclePostOutputs
{
    uint8_t* rgba; // W*H*4
    float* rgba_f; // W*H*4 final colour before quantisation, clamped
    uint8_t* pre; // W*H*4 after particles + iso, before shading
    float* depth; // W*H
    float* normal; // W*H*3
    float* ao; // W*H
    uint64_t counters[8];
    double seconds;
};

struct OracleOutputs
{
    uint8_t* rgba; // W*H*4 final framebuffer
    uint8_t* solid; // W*H*4 after shading, before volume
    uint8_t* pre; // W*H*4 after particles + iso, before shading
    float* pre_depth; // W*H depth before shading
    float* depth; // W*H (as float even for the FP64 oracle)
    float* normal; // W*H*3
    float* ao; // W*H
    uint8_t* vol_tex; // local voxels * 4
    uint8_t* iso_tex; // local voxels * 4
    uint64_t counters[8]; // volSamples isoSamples isoHits cellsVisited sphereTests volRays isoRays partRays
    double seconds_classify;
    double seconds_render;
    double seconds_bin;
};


namespace
{
    State make_state(const OracleScene* sc)
    {
        State st;
        st.scale[0] = sc->scale[0];
        st.scale[1] = sc->scale[1];
        st.scale[2] = sc->scale[2];
        st.sizes = make_sizes(sc->global_size, sc->local_size, sc->particle_grid, sc->position, sc->scale);
        st.fbW = sc->fb_w;
        st.fbH = sc->fb_h;
        for(int i = 0; i < 4; i++)
            st.viewport[i] = sc->viewport[i];
        st.T = sc->transfer_size;
        const int nv = sc->n_volume, np = sc->n_particle;
        st.vol.resize(nv);
        for(int s = 0; s < nv; s++)
        {
            st.vol[s].data = sc->vol_data[s];
            st.vol[s].dim = sc->vol_dim[s];
            for(int k = 0; k < 3; k++)
                st.vol[s].stride[k] = sc->vol_stride[s][k];
        }
        st.adv.assign(size_t(nv), nullptr);
        for(int s = 0; s < nv; s++)
            st.adv[size_t(s)] = sc->vol_adv[s];
        st.part.resize(np);
        for(int s = 0; s < np; s++)
        {
            st.part[s].pos = sc->part_pos[s];
            st.part[s].attr = sc->part_attr[s];
            st.part[s].dim = sc->part_dim[s];
            st.part[s].radius = sc->part_radius[s];
            st.part[s].const_radius = sc->part_const_radius[s];
            st.part[s].count = sc->part_count[s];
        }
        st.lut.resize(nv + np);
        st.chain.resize(nv + np);
        st.weight.resize(nv + np);
        st.iso.resize(nv + np);
        for(int s = 0; s < nv + np; s++)
        {
            st.lut[s] = sc->lut[s];
            for(int k = 0; k < MAX_FUNCTORS; k++)
            {
                st.chain[s].bytecode[k] = sc->bytecode[s][k];
                for(int c = 0; c < 4; c++)
                    st.chain[s].params[k][c] = sc->params[s][k][c];
            }
            st.weight[s] = sc->weight[s];
            st.iso[s] = sc->iso[s];
        }
        st.step = sc->step;
        st.interpolation = sc->interpolation != 0;
        st.ditherMode = sc->dither_mode;
        st.aoEnabled = sc->ao_enabled != 0;
        st.aoWeight = sc->ao_weight;
        st.renderMode = sc->render_mode;
        st.clip.resize(sc->n_clip);
        for(int i = 0; i < sc->n_clip; i++)
            for(int k = 0; k < 3; k++)
            {
                st.clip[i].position[k] = sc->clip_pos[i][k];
                st.clip[i].normal[k] = sc->clip_normal[i][k];
            }
        for(int i = 0; i < 16; i++)
        {
            st.proj[i] = sc->proj[i];
            st.mv[i] = sc->mv[i];
        }
        for(int i = 0; i < 4; i++)
            st.bg[i] = sc->bg[i];
        return st;
    }

    template<class R>
    int render_impl(const OracleScene* sc, OracleOutputs* out)
    {
        if(sc->threads > 0)
            omp_set_num_threads(sc->threads);
        State st = make_state(sc);
        using clk = std::chrono::steady_clock;
        auto t0 = clk::now();
        Tex3B4 volTex, isoTex;
        classify<R>(st, volTex, isoTex);
        auto t1 = clk::now();
        std::vector<CellList> lists;
        for(auto& ps : st.part)
            lists.push_back(bin_particles(ps, st.sizes));
        auto t2 = clk::now();
        FrameOut<R> fo;
        render<R>(st, volTex, isoTex, lists, fo);
        auto t3 = clk::now();
        out->seconds_classify = std::chrono::duration<double>(t1 - t0).count();
        out->seconds_bin = std::chrono::duration<double>(t2 - t1).count();
        out->seconds_render = std::chrono::duration<double>(t3 - t2).count();
        const size_t N = size_t(st.fbW) * st.fbH;
        if(out->rgba)
            std::memcpy(out->rgba, fo.g.color.data(), N * 4);
        if(out->solid)
            std::memcpy(out->solid, fo.solid.data(), N * 4);
        if(out->pre)
            std::memcpy(out->pre, fo.pre.data(), N * 4);
        for(size_t i = 0; i < N; i++)
        {
            if(out->depth)
                out->depth[i] = float(fo.g.depth[i]);
            if(out->pre_depth)
                out->pre_depth[i] = float(fo.preDepth[i]);
            if(out->ao)
                out->ao[i] = float(fo.g.ao[i]);
            if(out->normal)
            {
                out->normal[3 * i + 0] = float(fo.g.normal[i].x);
                out->normal[3 * i + 1] = float(fo.g.normal[i].y);
                out->normal[3 * i + 2] = float(fo.g.normal[i].z);
            }
        }
        if(out->vol_tex)
            std::memcpy(out->vol_tex, volTex.data.data(), volTex.data.size() * 4);
        if(out->iso_tex)
            std::memcpy(out->iso_tex, isoTex.data.data(), isoTex.data.size() * 4);
        const Counters& c = fo.counters;
        uint64_t cv[8]
            = {c.volSamples, c.isoSamples, c.isoHits, c.cellsVisited, c.sphereTests, c.volRays, c.isoRays, c.partRays};
        for(int i = 0; i < 8; i++)
            out->counters[i] = cv[i];
        return 0;
    }

    template<class R>
    int render_post_impl(const OracleScene* sc, int dtype, int stride, OraclePostOutputs* out)
    {
        if(sc->threads > 0)
            omp_set_num_threads(sc->threads);
        if(stride < 1)
            return -1;
        State st = make_state(sc);
        using clk = std::chrono::steady_clock;
        auto t0 = clk::now();
        PostScene ps;
        ps.st = &st;
        const int nv = (int) st.vol.size();
        std::vector<std::vector<float>> applied(nv);
        for(int s = 0; s < nv; s++)
        {
            const VolumeSource& v = st.vol[s];
            const bool np = sc->vol_nonpersistent[s] != 0;
            ps.persistent.push_back(!np);
            if(!(st.weight[s] + st.iso[s] > 0.0f))
            {
                ps.fields.emplace_back();
                continue;
            }
            if(!np)
            {
                ps.fields.push_back(make_brick_field(v.data, v.stride, v.dim, st.sizes, dtype));
                continue;
            }
            // non-persistent: the functor-applied float field (UpdatePersistendTextureKernel,
            // lib/isaac/isaac_common_kernel.hpp:432-451)
            const int64_t X = st.sizes.localSize[0], Y = st.sizes.localSize[1], Z = st.sizes.localSize[2];
            std::vector<float>& a = applied[s];
            a.resize(size_t(X * Y * Z));
#pragma omp parallel for schedule(static)
            for(int64_t z = 0; z < Z; z++)
                for(int64_t y = 0; y < Y; y++)
                    for(int64_t x = 0; x < X; x++)
                    {
                        const float* e = v.data + (x * v.stride[0] + y * v.stride[1] + z * v.stride[2]) * v.dim;
                        a[size_t((z * Y + y) * X + x)] = apply_chain<float>(st.chain[s], e, v.dim);
                    }
            const int64_t st1[3] = {1, X, X * Y};
            ps.fields.push_back(make_brick_field(a.data(), st1, 1, st.sizes, dtype));
        }
        std::vector<CellList> lists;
        for(auto& p : st.part)
            lists.push_back(bin_particles(p, st.sizes));
        PostOut<R> po;
        try
        {
            render_post<R>(ps, lists, stride, po);
        }
        catch(const std::exception&)
        {
            return -2;
        }
        out->seconds = std::chrono::duration<double>(clk::now() - t0).count();
        const size_t N = size_t(st.fbW) * st.fbH;
        if(out->rgba)
            std::memcpy(out->rgba, po.g.color.data(), N * 4);
        if(out->pre)
            std::memcpy(out->pre, po.pre.data(), N * 4);
        for(size_t i = 0; i < N; i++)
        {
            if(out->rgba_f)
            {
                out->rgba_f[4 * i + 0] = float(po.finalColor[i].x);
                out->rgba_f[4 * i + 1] = float(po.finalColor[i].y);
                out->rgba_f[4 * i + 2] = float(po.finalColor[i].z);
                out->rgba_f[4 * i + 3] = float(po.finalColor[i].w);
            }
            if(out->depth)
                out->depth[i] = float(po.g.depth[i]);
            if(out->ao)
                out->ao[i] = float(po.g.ao[i]);
            if(out->normal)
            {
                out->normal[3 * i + 0] = float(po.g.normal[i].x);
                out->normal[3 * i + 1] = float(po.g.normal[i].y);
                out->normal[3 * i + 2] = float(po.g.normal[i].z);
            }
        }
        const Counters& c = po.counters;
        uint64_t cv[8]
            = {c.volSamples, c.isoSamples, c.isoHits, c.cellsVisited, c.sphereTests, c.volRays, c.isoRays, c.partRays};
        for(int i = 0; i < 8; i++)
            out->counters[i] = cv[i];
        return 0;
    }
} // namespace

extern "C"
{
    int oracle_render_f32(const OracleScene* sc, OracleOutputs* out)
    {
        return render_impl<float>(sc, out);
    }

    int oracle_render_f64(const OracleScene* sc, OracleOutputs* out)
    {
        return render_impl<double>(sc, out);
    }

    // Stable counting-sort cell list for one particle source.
    int oracle_bin_particles(
        const float* pos,
        int64_t count,
        const int64_t local_size[3],
        const int64_t grid[3],
        int32_t* cell_of,
        int32_t* counts,
        int64_t* offsets,
        int64_t* perm,
        float* rel)
    {
        int64_t g[3] = {local_size[0], local_size[1], local_size[2]};
        int32_t p0[3] = {0, 0, 0};
        float s1[3] = {1, 1, 1};
        Sizes sz = make_sizes(g, local_size, grid, p0, s1);
        ParticleSource ps;
        ps.pos = pos;
        ps.count = count;
        CellList cl = bin_particles(ps, sz);
        const int64_t ncell = grid[0] * grid[1] * grid[2];
        std::memcpy(cell_of, cl.cellOf.data(), sizeof(int32_t) * count);
        std::memcpy(counts, cl.count.data(), sizeof(int32_t) * ncell);
        std::memcpy(offsets, cl.offset.data(), sizeof(int64_t) * (ncell + 1));
        std::memcpy(perm, cl.perm.data(), sizeof(int64_t) * count);
        std::memcpy(rel, cl.rel.data(), sizeof(float) * 3 * count);
        return 0;
    }

    // min/max request: per source, volume sources first (min_max_volume /
    // min_max_particles); mins/maxs hold n_volume + n_particle floats
    int oracle_min_max(const OracleScene* sc, float* mins, float* maxs)
    {
        State st = make_state(sc);
        const int nv = sc->n_volume;
        for(int s = 0; s < nv; s++)
            min_max_volume(st.vol[s], st.chain[s], st.sizes, mins[s], maxs[s]);
        for(int s = 0; s < sc->n_particle; s++)
            min_max_particles(st.part[s], st.chain[nv + s], st.sizes, mins[nv + s], maxs[nv + s]);
        return 0;
    }

    // field-source advection: the noise texture and one advection step
    int oracle_noise(const int64_t local[3], const float scale[3], uint32_t seed_points, float* out)
    {
        const std::vector<float> n = make_noise(local[0], local[1], local[2], scale, seed_points);
        std::memcpy(out, n.data(), n.size() * sizeof(float));
        return 0;
    }

    int oracle_advect(
        const float* field,
        const int64_t local[3],
        const float scale[3],
        float step_size,
        float history_weight,
        int seeding,
        const float* noise,
        const uint8_t* history,
        uint8_t* out)
    {
        VolumeSource f;
        f.data = field;
        f.dim = 3;
        f.stride[0] = 1;
        f.stride[1] = local[0];
        f.stride[2] = local[0] * local[1];
        const std::vector<uint8_t> a
            = advect(f, noise, history, local[0], local[1], local[2], scale, step_size, history_weight, seeding != 0);
        std::memcpy(out, a.data(), a.size());
        return 0;
    }

    // Stage-level entry points used by known-answer tests.
    void oracle_inverse_f32(const float* m, float* out)
    {
        M4<float> a;
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
                a.m[c][r] = m[c * 4 + r];
        M4<float> b = mat_inverse(a);
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
                out[c * 4 + r] = b.m[c][r];
    }

    void oracle_inverse_f64(const double* m, double* out)
    {
        M4<double> a;
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
                a.m[c][r] = m[c * 4 + r];
        M4<double> b = mat_inverse(a);
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
                out[c * 4 + r] = b.m[c][r];
    }

    void oracle_perspective(double fovy, double aspect, double zn, double zf, double* out)
    {
        perspective(fovy, aspect, zn, zf, out);
    }

    float oracle_apply_chain(const int32_t bytecode[3], const float params[12], const float* v, int dim)
    {
        Chain c;
        for(int k = 0; k < 3; k++)
        {
            c.bytecode[k] = bytecode[k];
            for(int i = 0; i < 4; i++)
                c.params[k][i] = params[k * 4 + i];
        }
        return apply_chain<float>(c, v, dim);
    }

    void oracle_dither(int mode, int x, int y, int z, const float c[4], uint8_t out[4])
    {
        B4 b = dither<float>(mode, x, y, z, V4<float>{c[0], c[1], c[2], c[3]});
        out[0] = b.x;
        out[1] = b.y;
        out[2] = b.z;
        out[3] = b.w;
    }

    // Trilinear/nearest sample of a byte4 texture (normalised to [0,1]).
    void oracle_sample_b4(const uint8_t* tex, int64_t X, int64_t Y, int64_t Z, const float p[3], int linear, float out[4])
    {
        Tex3B4 t;
        t.X = X;
        t.Y = Y;
        t.Z = Z;
        t.data.resize(X * Y * Z);
        std::memcpy(t.data.data(), tex, X * Y * Z * 4);
        V4<float> r = sample_b4<float>(t, V3<float>{p[0], p[1], p[2]}, linear != 0);
        out[0] = r.x;
        out[1] = r.y;
        out[2] = r.z;
        out[3] = r.w;
    }

    // pixelToRay + clipRay for one pixel; returns 1 if the ray hits the local box.
    int oracle_ray(const OracleScene* sc, int precision, float px, float py, double out[14])
    {
        State st = make_state(sc);
        if(precision == 64)
        {
            Frame<double> fr = make_frame<double>(st);
            Ray<double> r = pixel_to_ray(fr.invMVP, st.sizes, double(px), double(py), double(st.fbW), double(st.fbH));
            int hit = clip_ray(r, st.sizes, st.clip) ? 1 : 0;
            double o[14] = {r.dir.x, r.dir.y, r.dir.z, r.start.x, r.start.y, r.start.z, r.startDepth, r.endDepth,
                            double(r.isClipped), r.clippingNormal.x, r.clippingNormal.y, r.clippingNormal.z, 0, 0};
            std::memcpy(out, o, sizeof(o));
            return hit;
        }
        Frame<float> fr = make_frame<float>(st);
        Ray<float> r = pixel_to_ray(fr.invMVP, st.sizes, px, py, float(st.fbW), float(st.fbH));
        int hit = clip_ray(r, st.sizes, st.clip) ? 1 : 0;
        double o[14] = {r.dir.x, r.dir.y, r.dir.z, r.start.x, r.start.y, r.start.z, r.startDepth, r.endDepth,
                        double(r.isClipped), r.clippingNormal.x, r.clippingNormal.y, r.clippingNormal.z, 0, 0};
        std::memcpy(out, o, sizeof(o));
        return hit;
    }

    // Brick form (post-classification) frame.  precision 32: ray setup in binary64,
    // sampling and compositing in binary32 (the library's arithmetic); 64: all binary64.
    // dtype -1 samples the float sources directly, 0/1/2 the F32/F16/BF16 bricks.
    int oracle_render_post(const OracleScene* sc, int precision, int dtype, int stride, OraclePostOutputs* out)
    {
        if(dtype < BRICK_RAW || dtype > BRICK_BF16)
            return -1;
        return precision == 64 ? render_post_impl<double>(sc, dtype, stride, out) : render_post_impl<float>(sc, dtype, stride, out);
    }

    // per-brick, per-component scales of one source (nbricks*dim floats)
    int oracle_brick_scales(const float* data, const int64_t stride[3], int dim, const int64_t local[3], int dtype, float* out)
    {
        Sizes sz{};
        for(int k = 0; k < 3; k++)
            sz.localSize[k] = local[k];
        BrickField f = make_brick_field(data, stride, dim, sz, dtype);
        std::memcpy(out, f.scale.data(), f.scale.size() * sizeof(float));
        return 0;
    }

    // Brick-form samples of one source at n positions (voxel units, x y z), all components
    int oracle_sample_field(
        const float* data,
        const int64_t stride[3],
        int dim,
        const int64_t local[3],
        int dtype,
        int precision,
        int linear,
        int64_t n,
        const double* pos,
        double* out)
    {
        Sizes sz{};
        for(int k = 0; k < 3; k++)
            sz.localSize[k] = local[k];
        BrickField f = make_brick_field(data, stride, dim, sz, dtype);
        for(int64_t i = 0; i < n; i++)
        {
            if(precision == 64)
            {
                double v[4] = {0, 0, 0, 0};
                sample_field<double>(f, V3<double>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, linear != 0, v);
                for(int c = 0; c < 4; c++)
                    out[4 * i + c] = v[c];
            }
            else
            {
                float v[4] = {0, 0, 0, 0};
                sample_field<float>(f, V3<float>{float(pos[3 * i]), float(pos[3 * i + 1]), float(pos[3 * i + 2])}, linear != 0, v);
                for(int c = 0; c < 4; c++)
                    out[4 * i + c] = v[c];
            }
        }
        return 0;
    }

    void oracle_f32_to_f16(const float* in, uint16_t* out, int64_t n)
    {
        for(int64_t i = 0; i < n; i++)
            out[i] = f32_to_f16_rne(in[i]);
    }

    void oracle_f16_to_f32(const uint16_t* in, float* out, int64_t n)
    {
        for(int64_t i = 0; i < n; i++)
            out[i] = f16_to_f32(in[i]);
    }

    void oracle_f32_to_bf16(const float* in, uint16_t* out, int64_t n)
    {
        for(int64_t i = 0; i < n; i++)
            out[i] = f32_to_bf16_rne(in[i]);
    }

    int oracle_max_threads()
    {
        return omp_get_max_threads();
    }
}
