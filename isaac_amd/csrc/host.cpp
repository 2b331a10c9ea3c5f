// Host side of the gfx950 render library: the C-ABI declared in
// include/isaac_hip.h.  It owns the device buffers of one rank's render state
// (the reference's IsaacVisualization members, lib/isaac.hpp:565-945) and
// turns the steering setters into kernel parameter blocks.
#include "context.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

using namespace isaac_hip;

namespace
{
    const char* kFunctorNames[6] = {"idem", "add", "mul", "length", "pow", "sum"}; // lib/isaac/isaac_functors.hpp:185-208

    // ------------------------------------------------------------------ host math (glm op order)
    template<class R>
    struct M4T
    {
        R m[4][4];
    };
    using M4 = M4T<float>;

    template<class R>
    M4T<R> mat_mul(const M4T<R>& a, const M4T<R>& b)
    {
        M4T<R> r;
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

    // glm compute_inverse<4,4> (lib/isaac.hpp:2177 calls glm::inverse); binary32 for the
    // combined form, binary64 for the ray setup of the brick form
    template<class R>
    M4T<R> mat_inverse(const M4T<R>& M)
    {
        auto m = [&](int c, int r) { return M.m[c][r]; };
        const R c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
        const R c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
        const R c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
        const R c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
        const R c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
        const R c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
        const R c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
        const R c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
        const R c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
        const R c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
        const R c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
        const R c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
        const R c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
        const R c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
        const R c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
        const R c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
        const R c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
        const R c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
        const R f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
        const R f3_[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
        const R v0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
        const R v1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
        const R v2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
        const R v3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
        R col[4][4];
        const R sa[4] = {R(1), R(-1), R(1), R(-1)}, sb[4] = {R(-1), R(1), R(-1), R(1)};
        for(int k = 0; k < 4; k++)
        {
            const R i0 = v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k];
            const R i1 = v0[k] * f0[k] - v2[k] * f3_[k] + v3[k] * f4[k];
            const R i2 = v0[k] * f1[k] - v1[k] * f3_[k] + v3[k] * f5[k];
            const R i3 = v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k];
            col[0][k] = i0 * sa[k];
            col[1][k] = i1 * sb[k];
            col[2][k] = i2 * sa[k];
            col[3][k] = i3 * sb[k];
        }
        const R d0[4] = {M.m[0][0] * col[0][0], M.m[0][1] * col[1][0], M.m[0][2] * col[2][0], M.m[0][3] * col[3][0]};
        const R d1 = (d0[0] + d0[1]) + (d0[2] + d0[3]);
        const R inv = R(1) / d1;
        M4T<R> r;
        for(int c = 0; c < 4; c++)
            for(int k = 0; k < 4; k++)
                r.m[c][k] = col[c][k] * inv;
        return r;
    }

    float halton(uint32_t index, uint32_t base)
    {
        float result = 0, f = 1;
        while(index > 0)
        {
            f /= float(base);
            result += f * float(index % base);
            index /= base;
        }
        return result;
    }

    // getHSVA (lib/isaac/isaac_helper.hpp:221-264)
    void hsva(float h, float s, float v, float a, float out[4])
    {
        const double pi3 = M_PI / 3.0;
        const int hi = int(std::floor(double(h) / pi3));
        const float f = float(double(h) / pi3 - double(float(hi)));
        const float p = v * (1.0f - s);
        const float q = v * (1.0f - s * f);
        const float t = v * (1.0f - s * (1.0f - f));
        float r = 0, g = 0, b = 0;
        switch(hi)
        {
        case 0:
        case 6:
            r = v, g = t, b = p;
            break;
        case 1:
            r = q, g = v, b = p;
            break;
        case 2:
            r = p, g = v, b = t;
            break;
        case 3:
            r = p, g = q, b = v;
            break;
        case 4:
            r = t, g = p, b = v;
            break;
        case 5:
            r = v, g = p, b = q;
            break;
        }
        out[0] = r;
        out[1] = g;
        out[2] = b;
        out[3] = a;
    }

    // updateFunctions' per-source parser (lib/isaac.hpp:1086-1186)
    int32_t parse_chain(const std::string& src, int32_t bytecode[3], float params[12])
    {
        int32_t err = 0;
        float P[3][4] = {};
        int32_t bc[3] = {0, 0, 0};
        std::string source = src;
        int elem = 0;
        bool again = true;
        while(again)
        {
            size_t pos = source.find('|');
            std::string token;
            if(pos == std::string::npos)
            {
                again = false;
                token = source;
            }
            else
            {
                token = source.substr(0, pos);
                source.erase(0, pos + 1);
            }
            if(elem >= 3)
            {
                err = 1;
                break;
            }
            token.erase(std::remove_if(token.begin(), token.end(), [](unsigned char ch) { return std::isspace(ch); }), token.end());
            const size_t tb = token.find('(');
            if(tb == std::string::npos)
            {
                for(int k = 0; k < 4; k++)
                    P[elem][k] = 0.0f;
            }
            else
            {
                const size_t te = token.find(')');
                if(te == std::string::npos)
                {
                    err = -1;
                    break;
                }
                if(te - tb == 1)
                {
                    for(int k = 0; k < 4; k++)
                        P[elem][k] = 0.0f;
                }
                else
                {
                    std::string parameters = token.substr(tb + 1, te - tb - 1);
                    float* arr = P[elem];
                    int pe = 0;
                    bool pagain = true;
                    while(pagain)
                    {
                        size_t pp = parameters.find(',');
                        std::string par;
                        if(pp == std::string::npos)
                        {
                            pagain = false;
                            par = parameters;
                        }
                        else
                        {
                            par = parameters.substr(0, pp);
                            parameters.erase(0, pp + 1);
                        }
                        if(pe >= 4)
                        {
                            err = 2;
                            break;
                        }
                        try
                        {
                            arr[pe] = std::stof(par);
                        }
                        catch(const std::exception&)
                        {
                            err = -2;
                            pe++;
                            break;
                        }
                        pe++;
                    }
                    for(; pe < 4; pe++)
                        arr[pe] = arr[pe - 1];
                }
            }
            if(tb != std::string::npos)
                token = token.substr(0, tb);
            bool found = false;
            for(int k = 0; k < 6; k++)
                if(token == kFunctorNames[k])
                {
                    bc[elem] = k;
                    found = true;
                    break;
                }
            if(!found)
            {
                err = -2;
                break;
            }
            elem++;
        }
        for(; elem < 3; elem++)
        {
            bc[elem] = 0;
            for(int k = 0; k < 4; k++)
                P[elem][k] = 0.0f;
        }
        for(int k = 0; k < 3; k++)
        {
            bytecode[k] = bc[k];
            for(int c = 0; c < 4; c++)
                params[k * 4 + c] = P[k][c];
        }
        return err;
    }

    // updateTransfer (lib/isaac.hpp:1244-1266); std::map keeps the first of duplicate keys
    void build_lut(int T, int n, const uint32_t* pos, const float* rgba, float* lut)
    {
        std::map<uint32_t, const float*> d;
        for(int i = 0; i < n; i++)
            d.insert(std::make_pair(pos[i], rgba + 4 * i));
        std::fill(lut, lut + size_t(T) * 4, 0.0f);
        if(d.empty())
            return;
        auto before = d.begin();
        auto next = before;
        for(next++; next != d.end(); next++)
        {
            const uint32_t width = next->first - before->first;
            for(size_t j = 0; j < size_t(width) && size_t(j + before->first) < size_t(T); j++)
                for(int c = 0; c < 4; c++)
                    lut[(before->first + j) * 4 + c]
                        = (before->second[c] * float(width - j) + next->second[c] * float(j)) / float(width);
            before = next;
        }
    }

} // namespace


namespace
{
    void derive_sizes(isaac_hip_ctx* c)
    {
        // the ctor's float conversions (lib/isaac.hpp:659-661,796-797)
        for(int i = 0; i < 3; i++)
        {
            c->gsScaled[i] = int64_t(float(c->global[i]) * c->scale[i]);
            c->lsScaled[i] = int64_t(float(c->local[i]) * c->scale[i]);
            c->posScaled[i] = int32_t(float(c->position[i]) * c->scale[i]);
        }
        c->maxGsScaled = std::max(c->gsScaled[0], std::max(c->gsScaled[1], c->gsScaled[2]));
    }

    // timing marker k on the context's stream: every marker (timing events 1), the two
    // around the march (2, markers 4 and 5) or none (0)
    hipError_t mark(isaac_hip_ctx* c, int k)
    {
        const bool on = c->timingEvents == 1 || (c->timingEvents == 2 && (k == 4 || k == 5));
        return on ? hipEventRecord(c->ev[k], c->stream) : hipSuccess;
    }

    // assigns v to dst and bumps the context's generation when the value changes
    template<class T>
    void set_tracked(isaac_hip_ctx* c, T& dst, const T& v)
    {
        if(std::memcmp(&dst, &v, sizeof(T)) != 0)
            c->generation++;
        dst = v;
    }

    int alloc_textures(isaac_hip_ctx* c)
    {
        const size_t nvox = size_t(c->local[0]) * size_t(c->local[1]) * size_t(c->local[2]);
        if(c->tex)
            (void) hipFree(c->tex);
        c->tex = nullptr;
        if(c->texH)
            (void) hipFree(c->texH);
        c->texH = nullptr;
        // one texel of padding: the march reads texels x and x+1 with one 16-byte load,
        // also at the last texel (whose x+1 it then ignores)
        c->generation++;
        HIPCHK(c, hipMalloc(&c->tex, (nvox + 1) * 8));
        HIPCHK(c, hipMemsetAsync(c->tex, 0, (nvox + 1) * 8, c->stream));
        c->texVox = nvox;
        c->classified = false;
        return 0;
    }

    int ensure_np_texture(isaac_hip_ctx* c, VolSlot& v)
    {
        const int64_t g = v.desc.guard_size;
        const size_t n = size_t(c->local[0] + 2 * g) * size_t(c->local[1] + 2 * g) * size_t(c->local[2] + 2 * g);
        if(v.npTex && v.npBytes == n * 4)
            return 0;
        if(v.npTex)
            (void) hipFree(v.npTex);
        v.npTex = nullptr;
        c->generation++;
        HIPCHK(c, hipMalloc(&v.npTex, n * 4));
        v.npBytes = n * 4;
        return 0;
    }

    // the advection textures of a field source (zero-filled: the history of the
    // first frame and the region outside the local box read as 0)
    int ensure_advection(isaac_hip_ctx* c, VolSlot& v)
    {
        const size_t n = size_t(c->local[0]) * size_t(c->local[1]) * size_t(c->local[2]);
        if(v.advVox != n)
            for(uint8_t*& a : v.adv)
                if(a)
                {
                    (void) hipFree(a);
                    a = nullptr;
                }
        v.advVox = n;
        for(uint8_t*& a : v.adv)
        {
            if(a)
                continue;
            c->generation++;
            HIPCHK(c, hipMalloc(&a, (n + 3) / 4 * 4));
            HIPCHK(c, launch_fill_u32(reinterpret_cast<uint32_t*>(a), 0u, (n + 3) / 4, c->stream));
        }
        return 0;
    }

    // the noise of the advection seeds (updateNoiseTexture, lib/isaac.hpp:499-562),
    // rebuilt when the seed count or the local size changes
    int ensure_noise(isaac_hip_ctx* c)
    {
        const size_t n = size_t(c->local[0]) * size_t(c->local[1]) * size_t(c->local[2]);
        if(c->noise && c->noiseVox != n)
        {
            (void) hipFree(c->noise);
            (void) hipFree(c->noiseTmp);
            c->noise = c->noiseTmp = nullptr;
        }
        if(!c->noise)
        {
            c->generation++;
            HIPCHK(c, hipMalloc(&c->noise, n * 4));
            HIPCHK(c, hipMalloc(&c->noiseTmp, n * 4));
            c->noiseVox = n;
            c->noiseDirty = true;
        }
        if(c->noiseDirty)
        {
            HIPCHK(c, launch_noise(c->noise, c->noiseTmp, int(c->local[0]), int(c->local[1]), int(c->local[2]), c->scale,
                                   uint32_t(c->seedPoints), c->stream));
            c->noiseDirty = false;
        }
        return 0;
    }

    int ensure_bricks(isaac_hip_ctx* c, VolSlot& v, int dim)
    {
        const int64_t nbx = (c->local[0] + 7) / 8, nby = (c->local[1] + 7) / 8, nbz = (c->local[2] + 7) / 8;
        const int64_t nb = nbx * nby * nbz;
        const size_t pairBytes = c->brickDtype == ISAAC_HIP_F32 ? 8 : 4;
        const size_t bytes = size_t(nb) * size_t(dim) * size_t(kBrickPairs) * pairBytes;
        const size_t ns = size_t(nb) * size_t(dim);
        if(uint64_t(ns) * uint64_t(kBrickPairs) >= (uint64_t(1) << 32))
            return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "brick form: more than 2^32 pairs per source (32-bit pair index)");
        if(v.brick && v.brickBytes == bytes && v.scaleCount == ns)
            return 0;
        if(v.brick)
            (void) hipFree(v.brick);
        if(v.brickScale)
            (void) hipFree(v.brickScale);
        v.brick = nullptr;
        v.brickScale = nullptr;
        v.brickBytes = v.scaleCount = 0;
        c->generation++;
        HIPCHK(c, hipMalloc(&v.brick, std::max<size_t>(bytes, 8)));
        HIPCHK(c, hipMalloc(&v.brickScale, std::max<size_t>(ns, 1) * 4));
        v.brickBytes = bytes;
        v.scaleCount = ns;
        v.nbricks = nb;
        v.brickDim = dim;
        return 0;
    }

    void free_part(PartSlot& p)
    {
        for(void*& b : p.blocks)
        {
            if(b)
                (void) hipFree(b);
            b = nullptr;
        }
        p.B = BinBuffers{};
        p.capN = p.capCells = 0;
    }

    int ensure_part_buffers(isaac_hip_ctx* c, PartSlot& p)
    {
        const int64_t n = std::max<int64_t>(p.desc.count, 1);
        const int64_t nc = c->pgrid[0] * c->pgrid[1] * c->pgrid[2];
        const bool needRad = p.desc.radius != nullptr;
        if(p.capN >= n && p.capCells == nc && (p.B.recRadius != nullptr) == needRad)
            return 0;
        free_part(p);
        c->generation++;
        const size_t sizes[12] = {size_t(n) * 4, size_t(n) * 4, size_t(n) * 16, size_t(nc) * 8, size_t(nc) * 4,
                                  size_t(nc + 1) * 4, size_t(nc) * 8, size_t(n) * 4, size_t(n) * 16,
                                  needRad ? size_t(n) * 4 : 0, bin_scan_workspace_elems(nc) * 8, size_t((n + 63) / 64) * 8};
        for(int i = 0; i < 12; i++)
            if(sizes[i])
                HIPCHK(c, hipMalloc(&p.blocks[i], sizes[i]));
        p.B.cellOf = (int32_t*) p.blocks[0];
        p.B.rank = (uint32_t*) p.blocks[1];
        p.B.rel = (float4*) p.blocks[2];
        p.B.count64 = (unsigned long long*) p.blocks[3];
        p.B.ncells = nc;
        p.B.count = (int32_t*) p.blocks[4];
        p.B.offset = (int32_t*) p.blocks[5];
        p.B.vrange = (int2*) p.blocks[6];
        p.B.perm = (int32_t*) p.blocks[7];
        p.B.recs = (float4*) p.blocks[8];
        p.B.recRadius = (float*) p.blocks[9];
        p.B.scanTmp = (unsigned long long*) p.blocks[10];
        p.B.visMask = (unsigned long long*) p.blocks[11];
        p.capN = n;
        p.capCells = nc;
        return 0;
    }

    ChainDesc to_chain(const int32_t bc[3], const float params[12])
    {
        ChainDesc d{};
        for(int k = 0; k < 3; k++)
        {
            d.op[k] = bc[k];
            for(int i = 0; i < 4; i++)
                d.p[k][i] = params[k * 4 + i];
        }
        return d;
    }

    int check_idx(isaac_hip_ctx* c, int idx)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(idx < 0 || idx >= c->nsrc)
            return fail(c, ISAAC_HIP_ERR_ARG, "source index out of range");
        return 0;
    }

    float elapsed_us(hipEvent_t a, hipEvent_t b)
    {
        float ms = 0.f;
        if(hipEventElapsedTime(&ms, a, b) != hipSuccess)
            return 0.f;
        return ms * 1000.0f;
    }
} // namespace

namespace isaac_hip
{
    // Kernel parameter blocks of one frame of context c (drawCallBack's setup,
    // lib/isaac.hpp:2140-2283) rendered into the G-buffer targets t.
    int build_frame(
        isaac_hip_ctx* c,
        const double proj[16],
        const double mv[16],
        const float bg[4],
        const int32_t viewport[4],
        const FrameTargets& t,
        FrameParams& F,
        PostParams& Q,
        ParticleParams& PP)
    {
        const int W = t.W, H = t.H;
        F = FrameParams{};
        Q = PostParams{};
        PP = ParticleParams{};
        // inverse MVP in binary32 (lib/isaac.hpp:2174-2178)
        M4 P, MV;
        for(int col = 0; col < 4; col++)
            for(int r = 0; r < 4; r++)
            {
                P.m[col][r] = float(proj[col * 4 + r]);
                MV.m[col][r] = float(mv[col * 4 + r]);
            }
        M4 inv = mat_inverse(mat_mul(P, MV));
        for(int col = 0; col < 4; col++)
            for(int r = 0; r < 4; r++)
                F.inv[col * 4 + r] = inv.m[col][r];
        F.W = W;
        F.H = H;
        int vp[4] = {0, 0, W, H};
        if(viewport)
            for(int i = 0; i < 4; i++)
                vp[i] = viewport[i];
        F.vx0 = std::max(0, vp[0]);
        F.vy0 = std::max(0, vp[1]);
        F.vx1 = std::min(W, vp[0] + vp[2]);
        F.vy1 = std::min(H, vp[1] + vp[3]);
        if(F.vx1 < F.vx0)
            F.vx1 = F.vx0;
        if(F.vy1 < F.vy0)
            F.vy1 = F.vy0;
        F.tilesX = (F.vx1 - F.vx0 + 15) / 16;
        F.tilesY = (F.vy1 - F.vy0 + 15) / 16;
        F.fbW = float(W);
        F.fbH = float(H);
        F.maxSize = float(c->maxGsScaled) * 0.5f;
        for(int i = 0; i < 3; i++)
        {
            F.off[i] = float(int32_t(c->gsScaled[i]) / 2 - c->posScaled[i]);
            F.bbMin[i] = float(std::max(int32_t(0), -c->posScaled[i]));
            F.bbMax[i] = float(std::min(int32_t(c->lsScaled[i]), int32_t(c->gsScaled[i]) - c->posScaled[i]));
            F.outerLow[i] = c->position[i] == 0;
            F.outerHigh[i] = (int64_t(c->position[i]) + c->local[i] == c->global[i]);
            F.scale[i] = c->scale[i];
        }
        F.nclip = c->nclip;
        for(int i = 0; i < c->nclip; i++)
            for(int k = 0; k < 3; k++)
            {
                F.cpos[i][k] = c->clipPos[i][k] * float(c->gsScaled[k]) * 0.5f + F.off[k];
                F.cnorm[i][k] = c->clipNormal[i][k];
            }
        F.step = c->step;
        F.linear = c->interpolation ? 1 : 0;
        F.prefetch = c->tunePrefetch;
        F.waves = c->tuneWaves;
        F.tileOrder = c->tuneTiles;
        F.waveGroups = (c->form == ISAAC_HIP_FORM_COMBINED && c->tuneMarchThreads == 64) ? 1 : 4;
        F.X = int32_t(c->local[0]);
        F.Y = int32_t(c->local[1]);
        F.Z = int32_t(c->local[2]);
        F.tex = c->tex;
        if(c->builtTexLayout == kTexHalf && c->texH)
            F.texH = c->texH;
        if(c->builtTexLayout == kTexBrick && c->texBricks)
        {
            F.texB = c->texBricks;
            F.nbx = uint32_t((c->local[0] + kTexBrick - 1) / kTexBrick);
            F.nbxy = F.nbx * uint32_t((c->local[1] + kTexBrick - 1) / kTexBrick);
        }
        {
            // bytes of the layout the march reads, its padding texel included
            const uint64_t texels = F.texB ? uint64_t(c->texBrickCount) * kTexBrickTexels
                                           : uint64_t(c->local[0]) * c->local[1] * c->local[2];
            F.texOff32 = (c->tuneMarchOffsets == 32 && (texels + 1) * sizeof(uint2) <= 0xffffffffull) ? 1 : 0;
        }
        const int nvf = c->nv + c->nf;
        F.isoEnabled = 0;
        for(int s = 0; s < nvf; s++)
            if(c->iso[s] > 0.0f)
                F.isoEnabled = 1;
        float twAll = 0.0f;
        for(int s = 0; s < c->nsrc; s++)
            twAll += c->weight[s];
        F.totalWeightAll = twAll;
        F.volumeEnabled = twAll > 0.0f;
        F.minSize = float(std::min(int(c->global[0]), std::min(int(c->global[1]), int(c->global[2]))));
        F.renderMode = c->renderMode;
        F.aoEnabled = (c->aoEnabled && c->aoWeight > 0.0f) ? 1 : 0;
        F.aoWeight = c->aoWeight;
        float bgv[4] = {0, 0, 0, 0};
        if(bg)
            for(int i = 0; i < 4; i++)
                bgv[i] = bg[i];
        // bgColor = {bg[3], bg[2], bg[1], bg[0]} (lib/isaac.hpp:2245-2249)
        F.bg[0] = bgv[3];
        F.bg[1] = bgv[2];
        F.bg[2] = bgv[1];
        F.bg[3] = bgv[0];
        F.bgClear = to_byte4(F.bg[0], F.bg[1], F.bg[2], 0.0f);
        F.maxGlobalSizeScaled = float(c->maxGsScaled);
        hsva(halton(0, 3) * float(2 * M_PI), 1.0f, 1.0f, 1.0f, F.rankHue);
        F.color = t.color;
        F.depth = t.depth;
        F.normal = t.normal;
        F.ao = t.ao;
        F.volColor = t.volColor;
        F.counters = t.counters;
        // particles
        PP.nsrc = 0;
        for(int i = 0; i < 3; i++)
        {
            PP.grid[i] = int32_t(c->pgrid[i]);
            PP.pscale[i] = float(c->lsScaled[i]) / float(c->pgrid[i]);
        }
        const int T = c->cfg.transfer_size;
        for(int s = 0; s < c->np; s++)
        {
            const PartSlot& p = c->parts[s];
            const int gi = nvf + s;
            if(c->weight[gi] == 0.0f || !p.set)
                continue;
            if(p.capCells != c->pgrid[0] * c->pgrid[1] * c->pgrid[2])
                return fail(c, ISAAC_HIP_ERR_STATE, "particle cell lists are stale; call isaac_hip_update_buffers");
            if(PP.nsrc >= 4)
                return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "at most 4 active particle sources per frame");
            ParticleSourceDev& S = PP.src[PP.nsrc++];
            S.vrange = p.B.vrange;
            S.recs = p.B.recs;
            S.radius = p.B.recRadius;
            S.constRadius = p.desc.constant_radius;
            S.weight = c->weight[gi];
            S.lut = c->lutDev + size_t(gi) * T;
        }
        F.outFloat = t.outFloat;
        if(c->form == ISAAC_HIP_FORM_BRICKS)
        {
            // binary64 camera inverse for the ray setup
            M4T<double> Pd, MVd;
            for(int col = 0; col < 4; col++)
                for(int r = 0; r < 4; r++)
                {
                    Pd.m[col][r] = proj[col * 4 + r];
                    MVd.m[col][r] = mv[col * 4 + r];
                }
            const M4T<double> invd = mat_inverse(mat_mul(Pd, MVd));
            for(int col = 0; col < 4; col++)
                for(int r = 0; r < 4; r++)
                    Q.invd[col * 4 + r] = invd.m[col][r];
            Q.maxSized = double(c->maxGsScaled) * 0.5;
            for(int i = 0; i < c->nclip; i++)
                for(int k = 0; k < 3; k++)
                    Q.cposd[i][k] = double(c->clipPos[i][k]) * double(c->gsScaled[k]) * 0.5 + double(F.off[k]);
            Q.dtype = c->brickDtype;
            Q.nbx = int32_t((c->local[0] + 7) / 8);
            Q.nby = int32_t((c->local[1] + 7) / 8);
            Q.nbz = int32_t((c->local[2] + 7) / 8);
            Q.T = T;
            Q.luts = c->lutDev;
            float tw = 0.0f;
            for(int s = 0; s < nvf; s++)
                tw += c->weight[s];
            Q.invTotalWeight = tw != 0.0f ? 1.0f / tw : 0.0f;
            for(int s = 0; s < nvf; s++)
            {
                const VolSlot& v = c->vols[s];
                if(!v.set || c->weight[s] + c->iso[s] <= 0.0f)
                    continue;
                if(!v.bricked)
                    return fail(c, ISAAC_HIP_ERR_STATE, "source " + std::to_string(s) + " enabled after update_buffers; call it again");
                BrickSource& B = Q.src[Q.nsrc++];
                B.data = v.brick;
                B.scale = v.brickScale;
                B.dim = v.brickDim;
                B.active = 1;
                B.weight = c->weight[s];
                B.iso = c->iso[s];
                B.isoScale = B.iso > 0.0f ? 0.5f / B.iso : 0.0f;
                B.chain = v.desc.persistent ? c->chain[s] : ChainDesc{};
                B.lutSlot = s;
            }
            // the enabled sources' transfer functions are staged in LDS (64 KiB per workgroup)
            if(size_t(Q.nsrc) * size_t(T) * sizeof(float4) > 64 * 1024)
                return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "brick form: enabled sources x transfer size above 4096 entries");
        }
        return 0;
    }
    // the visible particles' records of every enabled particle source (launch_bin_visible)
    int bin_particles(isaac_hip_ctx* c)
    {
        const int T = c->cfg.transfer_size;
        const int nvf = c->nv + c->nf;
        // particle cell lists
        for(int s = 0; s < c->np; s++)
        {
            PartSlot& p = c->parts[s];
            const int gi = nvf + s;
            if(!p.set || c->weight[gi] == 0.0f)
                continue;
            int rc = ensure_part_buffers(c, p);
            if(rc)
                return rc;
            BinParams P{};
            P.pos = p.desc.position;
            P.attr = p.desc.attribute;
            P.radius = p.desc.radius;
            P.dim = p.desc.feature_dim;
            P.n = p.desc.count;
            for(int i = 0; i < 3; i++)
            {
                P.grid[i] = int32_t(c->pgrid[i]);
                P.cw[i] = float(c->local[i]) / float(c->pgrid[i]);
            }
            P.chain = c->chain[gi];
            P.lut = c->lutDev + size_t(gi) * T;
            P.T = T;
            P.filter = 1;
            HIPCHK(c, launch_bin_visible(P, p.B, c->stream));
            p.last = P;
            p.lastValid = true;
            p.fullValid = false;
        }
        return 0;
    }

} // namespace isaac_hip

extern "C"
{
    const char* isaac_hip_version(void)
    {
        return "isaac-hip 0.1.0 (gfx950)";
    }

    void isaac_hip_default_projection(int32_t w, int32_t h, double out[16])
    {
        // glm::perspective(45.0 [radians in glm 0.9.9], w/h, ISAAC_Z_NEAR, ISAAC_Z_FAR), RH, depth -1..1
        const double fovy = 45.0, aspect = double(w) / double(h), zn = double(kZNear), zf = double(100.0f);
        const double t = std::tan(fovy / 2.0);
        for(int i = 0; i < 16; i++)
            out[i] = 0.0;
        out[0] = 1.0 / (aspect * t);
        out[5] = 1.0 / t;
        out[10] = -(zf + zn) / (zf - zn);
        out[11] = -1.0;
        out[14] = -(2.0 * zf * zn) / (zf - zn);
    }

    void isaac_hip_modelview(const double look_at[3], const double rotation[9], double distance, double out[16])
    {
        // distanceM * rotationM * translationM in double, glm column-major
        double T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {look_at[0], look_at[1], look_at[2], 1}};
        double R[4][4] = {};
        for(int c = 0; c < 3; c++)
            for(int r = 0; r < 3; r++)
                R[c][r] = rotation[c * 3 + r];
        R[3][3] = 1.0;
        double D[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, distance, 1}};
        auto mm = [](const double A[4][4], const double B[4][4], double C[4][4])
        {
            for(int c = 0; c < 4; c++)
                for(int r = 0; r < 4; r++)
                {
                    double s = A[0][r] * B[c][0];
                    s = s + A[1][r] * B[c][1];
                    s = s + A[2][r] * B[c][2];
                    s = s + A[3][r] * B[c][3];
                    C[c][r] = s;
                }
        };
        double DR[4][4], M[4][4];
        mm(D, R, DR);
        mm(DR, T, M);
        for(int c = 0; c < 4; c++)
            for(int r = 0; r < 4; r++)
                out[c * 4 + r] = M[c][r];
    }

    int32_t isaac_hip_parse_functor_chain(const char* chain, int32_t bytecode[3], float params[12])
    {
        return parse_chain(chain ? std::string(chain) : std::string(), bytecode, params);
    }

    int isaac_hip_build_transfer_lut(int32_t T, int32_t n, const uint32_t* positions, const float* rgba, float* lut)
    {
        if(T <= 0 || n < 0 || !lut || (n > 0 && (!positions || !rgba)))
            return ISAAC_HIP_ERR_ARG;
        build_lut(T, n, positions, rgba, lut);
        return 0;
    }

    int isaac_hip_default_transfer(
        int32_t i,
        int32_t n_sources,
        int32_t is_particle,
        int32_t T,
        uint32_t positions[2],
        float rgba[8])
    {
        if(n_sources <= 0 || i < 0 || i >= n_sources)
            return ISAAC_HIP_ERR_ARG;
        // hue = 2*pi*i/S computed as float(2i) * M_PI / float(S) (lib/isaac.hpp:778)
        const float h = float(double(float(2 * i)) * M_PI / double(float(n_sources)));
        hsva(h, 1.0f, 1.0f, is_particle ? 1.0f : 0.0f, rgba);
        hsva(h, 1.0f, 1.0f, 1.0f, rgba + 4);
        positions[0] = 0;
        positions[1] = uint32_t(T);
        return 0;
    }

    int isaac_hip_create(const isaac_hip_config* cfg, int device, isaac_hip_ctx** out)
    {
        if(!cfg || !out)
            return ISAAC_HIP_ERR_ARG;
        *out = nullptr;
        if(cfg->n_volume_sources < 0 || cfg->n_particle_sources < 0 || cfg->n_field_sources < 0)
            return ISAAC_HIP_ERR_ARG;
        if(cfg->n_volume_sources + cfg->n_field_sources > ISAAC_HIP_MAX_SOURCES
           || cfg->n_particle_sources > ISAAC_HIP_MAX_SOURCES)
            return ISAAC_HIP_ERR_ARG;
        if(cfg->framebuffer_size[0] <= 0 || cfg->framebuffer_size[1] <= 0 || cfg->transfer_size <= 0)
            return ISAAC_HIP_ERR_ARG;
        for(int i = 0; i < 3; i++)
            if(cfg->global_size[i] <= 0 || cfg->local_size[i] <= 0 || cfg->local_particle_size[i] <= 0
               || cfg->local_size[i] > (1 << 20) || cfg->scale[i] <= 0.0f)
                return ISAAC_HIP_ERR_ARG;
        auto* c = new isaac_hip_ctx();
        c->cfg = *cfg;
        c->device = device;
        hipError_t e = hipSetDevice(device);
        if(e != hipSuccess)
        {
            delete c;
            return ISAAC_HIP_ERR_DEVICE;
        }
        if(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        {
            delete c;
            return ISAAC_HIP_ERR_DEVICE;
        }
        c->ownStream = true;
        for(auto& ev : c->ev)
            (void) hipEventCreate(&ev);
        c->nv = cfg->n_volume_sources;
        c->nf = cfg->n_field_sources;
        c->np = cfg->n_particle_sources;
        c->nsrc = c->nv + c->nf + c->np;
        for(int i = 0; i < 3; i++)
        {
            c->global[i] = cfg->global_size[i];
            c->local[i] = cfg->local_size[i];
            c->pgrid[i] = cfg->local_particle_size[i];
            c->position[i] = cfg->position[i];
            c->scale[i] = cfg->scale[i];
        }
        derive_sizes(c);
        c->vols.resize(c->nv + c->nf);
        c->parts.resize(c->np);
        c->chain.resize(c->nsrc);
        c->chainErr.assign(c->nsrc, 0);
        c->lutHost.assign(c->nsrc, std::vector<float>(size_t(cfg->transfer_size) * 4, 0.0f));
        c->weight.assign(c->nsrc, 2.0f); // ISAAC_DEFAULT_WEIGHT
        c->iso.assign(c->nsrc, 0.0f);
        const int T = cfg->transfer_size;
        for(int i = 0; i < c->nsrc; i++)
        {
            // default functions and transfer functions (lib/isaac.hpp:720-731,759-794)
            const bool isPart = i >= c->nv + c->nf;
            const bool isField = i >= c->nv && !isPart;
            int32_t bc[3];
            float pr[12];
            parse_chain(isField ? "length" : "idem", bc, pr);
            c->chain[i] = to_chain(bc, pr);
            uint32_t pos[2];
            float rgba[8];
            isaac_hip_default_transfer(i, c->nsrc, isPart ? 1 : 0, T, pos, rgba);
            build_lut(T, 2, pos, rgba, c->lutHost[i].data());
        }
        int rc = 0;
        const size_t npx = size_t(cfg->framebuffer_size[0]) * size_t(cfg->framebuffer_size[1]);
        if((e = hipMalloc(&c->lutDev, size_t(std::max(c->nsrc, 1)) * T * sizeof(float4))) != hipSuccess
           || (e = hipMalloc(&c->color, npx * 4)) != hipSuccess || (e = hipMalloc(&c->depth, npx * 4)) != hipSuccess
           || (e = hipMalloc(&c->normal, npx * 12)) != hipSuccess || (e = hipMalloc(&c->ao, npx * 4)) != hipSuccess
           || (e = hipMalloc(&c->volColor, npx * 16)) != hipSuccess
           || (e = hipMalloc(&c->counters, 64 * sizeof(unsigned long long))) != hipSuccess)
        {
            isaac_hip_destroy(c);
            return ISAAC_HIP_ERR_ALLOC;
        }
        (void) hipMemsetAsync(c->color, 0, npx * 4, c->stream);
        (void) hipMemsetAsync(c->depth, 0, npx * 4, c->stream);
        (void) hipMemsetAsync(c->normal, 0, npx * 12, c->stream);
        (void) hipMemsetAsync(c->ao, 0, npx * 4, c->stream);
        (void) hipMemsetAsync(c->counters, 0, 64 * sizeof(unsigned long long), c->stream);
        rc = alloc_textures(c);
        if(rc != 0)
        {
            isaac_hip_destroy(c);
            return rc;
        }
        (void) hipStreamSynchronize(c->stream);
        *out = c;
        return 0;
    }

    void isaac_hip_destroy(isaac_hip_ctx* c)
    {
        if(!c)
            return;
        (void) hipSetDevice(c->device);
        if(c->stream)
            (void) hipStreamSynchronize(c->stream);
        for(auto& v : c->vols)
        {
            if(v.npTex)
                (void) hipFree(v.npTex);
            if(v.brick)
                (void) hipFree(v.brick);
            if(v.brickScale)
                (void) hipFree(v.brickScale);
            for(uint8_t* a : v.adv)
                if(a)
                    (void) hipFree(a);
        }
        for(auto& p : c->parts)
            free_part(p);
        aonet_destroy(c);
        void* bufs[] = {c->lutDev,   c->tex,        c->texBricks, c->texH,     c->color,    c->depth,
                        c->normal,   c->ao,         c->volColor,  c->counters, c->outFloat,
                        c->mmPartials, c->mmResult, c->noise,     c->noiseTmp};
        for(void* b : bufs)
            if(b)
                (void) hipFree(b);
        for(auto& ev : c->ev)
            if(ev)
                (void) hipEventDestroy(ev);
        if(c->videoWs)
            (void) hipFree(c->videoWs);
        if(c->ownStream && c->stream)
            (void) hipStreamDestroy(c->stream);
        delete c;
    }

    const char* isaac_hip_last_error(const isaac_hip_ctx* c)
    {
        return c ? c->err.c_str() : "null context";
    }

    int isaac_hip_set_stream(isaac_hip_ctx* c, void* s)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        (void) hipStreamSynchronize(c->stream);
        c->generation++;
        if(c->ownStream && c->stream)
            (void) hipStreamDestroy(c->stream);
        if(s)
        {
            c->stream = (hipStream_t) s;
            c->ownStream = false;
        }
        else
        {
            if(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
                return fail(c, ISAAC_HIP_ERR_DEVICE, "stream creation failed");
            c->ownStream = true;
        }
        return 0;
    }

    int isaac_hip_set_volume_source(isaac_hip_ctx* c, int idx, const isaac_hip_volume_desc* d)
    {
        if(!c || !d)
            return ISAAC_HIP_ERR_ARG;
        if(idx < 0 || idx >= c->nv + c->nf)
            return fail(c, ISAAC_HIP_ERR_ARG, "volume source index out of range");
        if(d->dtype != ISAAC_HIP_F32)
            return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "the plain render path reads float32 sources");
        if(d->feature_dim < 1 || d->feature_dim > 4 || !d->data || d->guard_size < 0)
            return fail(c, ISAAC_HIP_ERR_ARG, "bad volume source descriptor");
        if(!c->vols[idx].set || std::memcmp(&c->vols[idx].desc, d, sizeof(*d)) != 0)
            c->generation++;
        c->vols[idx].desc = *d;
        c->vols[idx].set = true;
        return 0;
    }

    int isaac_hip_set_particle_source(isaac_hip_ctx* c, int idx, const isaac_hip_particle_desc* d)
    {
        if(!c || !d)
            return ISAAC_HIP_ERR_ARG;
        if(idx < 0 || idx >= c->np)
            return fail(c, ISAAC_HIP_ERR_ARG, "particle source index out of range");
        if(d->feature_dim < 1 || d->feature_dim > 4 || d->count < 0 || d->count > INT32_MAX
           || (d->count > 0 && (!d->position || !d->attribute)))
            return fail(c, ISAAC_HIP_ERR_ARG, "bad particle source descriptor");
        if(!c->parts[idx].set || std::memcmp(&c->parts[idx].desc, d, sizeof(*d)) != 0)
            c->generation++;
        c->parts[idx].desc = *d;
        c->parts[idx].set = true;
        return 0;
    }

    int isaac_hip_set_function(isaac_hip_ctx* c, int idx, const char* chain, int32_t* error_code)
    {
        int rc = check_idx(c, idx);
        if(rc)
            return rc;
        int32_t bc[3];
        float pr[12];
        const int32_t e = parse_chain(chain ? std::string(chain) : std::string(), bc, pr);
        set_tracked(c, c->chain[idx], to_chain(bc, pr));
        set_tracked(c, c->chainErr[idx], e);
        if(error_code)
            *error_code = e;
        return 0;
    }

    int isaac_hip_set_transfer_points(isaac_hip_ctx* c, int idx, int n, const uint32_t* positions, const float* rgba)
    {
        int rc = check_idx(c, idx);
        if(rc)
            return rc;
        if(n < 0 || (n > 0 && (!positions || !rgba)))
            return fail(c, ISAAC_HIP_ERR_ARG, "bad transfer points");
        std::vector<float> lut(c->lutHost[idx].size());
        build_lut(c->cfg.transfer_size, n, positions, rgba, lut.data());
        if(lut != c->lutHost[idx])
        {
            c->lutHost[idx] = lut;
            c->lutDirty = true;
            c->generation++;
        }
        return 0;
    }

    int isaac_hip_set_transfer_lut(isaac_hip_ctx* c, int idx, const float* lut)
    {
        int rc = check_idx(c, idx);
        if(rc)
            return rc;
        if(!lut)
            return fail(c, ISAAC_HIP_ERR_ARG, "null lut");
        if(std::memcmp(c->lutHost[idx].data(), lut, size_t(c->cfg.transfer_size) * 16) != 0)
        {
            std::memcpy(c->lutHost[idx].data(), lut, size_t(c->cfg.transfer_size) * 16);
            c->lutDirty = true;
            c->generation++;
        }
        return 0;
    }

    int isaac_hip_set_weight(isaac_hip_ctx* c, int idx, float w)
    {
        int rc = check_idx(c, idx);
        if(rc)
            return rc;
        set_tracked(c, c->weight[idx], w);
        return 0;
    }

    int isaac_hip_set_iso_threshold(isaac_hip_ctx* c, int idx, float t)
    {
        int rc = check_idx(c, idx);
        if(rc)
            return rc;
        set_tracked(c, c->iso[idx], t);
        return 0;
    }

    int isaac_hip_set_step(isaac_hip_ctx* c, float step)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        set_tracked(c, c->step, step < 0.01f ? 0.01f : step); // lib/isaac.hpp:1629-1633
        return 0;
    }

    int isaac_hip_set_interpolation(isaac_hip_ctx* c, int on)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        set_tracked(c, c->interpolation, on != 0);
        return 0;
    }

    int isaac_hip_set_advection(isaac_hip_ctx* c, float step_factor, float history_weight, int32_t seeding_period, int32_t seeding_duration)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(seeding_period < 1 || !(step_factor == step_factor) || !(history_weight == history_weight))
            return fail(c, ISAAC_HIP_ERR_ARG, "advection: seeding period >= 1 and finite factors expected");
        set_tracked(c, c->advStepFactor, step_factor);
        set_tracked(c, c->advHistoryWeight, history_weight);
        set_tracked(c, c->advSeedPeriod, seeding_period);
        set_tracked(c, c->advSeedTime, seeding_duration);
        return 0;
    }

    int isaac_hip_set_seed_points(isaac_hip_ctx* c, int32_t seed_points)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        const int32_t n = seed_points < 0 ? 0 : seed_points;
        if(n != c->seedPoints)
            c->noiseDirty = true;
        set_tracked(c, c->seedPoints, n);
        return 0;
    }

    int isaac_hip_get_advection(const isaac_hip_ctx* c, int field_idx, const void** texture, const void** noise, int32_t* time_step)
    {
        if(!c || field_idx < 0 || field_idx >= c->nf)
            return ISAAC_HIP_ERR_ARG;
        const VolSlot& v = c->vols[size_t(c->nv + field_idx)];
        if(texture)
            *texture = v.adv[v.advCur];
        if(noise)
            *noise = c->noise;
        if(time_step)
            *time_step = c->timeStep;
        return 0;
    }

    int isaac_hip_set_dither_mode(isaac_hip_ctx* c, int mode)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        // every value is valid: as in applyDither3D, modes other than 0-3 use the 16^3 mask
        set_tracked(c, c->ditherMode, mode);
        return 0;
    }

    int isaac_hip_set_ao(isaac_hip_ctx* c, int enabled, float weight)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        set_tracked(c, c->aoEnabled, enabled != 0);
        set_tracked(c, c->aoWeight, weight);
        return 0;
    }

    int isaac_hip_set_render_mode(isaac_hip_ctx* c, int mode)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        set_tracked(c, c->renderMode, mode);
        return 0;
    }

    int isaac_hip_edit_clipping(isaac_hip_ctx* c, uint32_t nr, float px, float py, float pz, float nx, float ny, float nz)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(nr >= uint32_t(kMaxClipping))
            return 0;
        // editClipping (lib/isaac.hpp:958-986)
        const float sx = nx * c->scale[0], sy = ny * c->scale[1], sz = nz * c->scale[2];
        const float l = sqrtf(sx * sx + sy * sy + sz * sz);
        if(l == 0.0f)
            return 0;
        set_tracked(c, c->clipPos[nr][0], px);
        set_tracked(c, c->clipPos[nr][1], py);
        set_tracked(c, c->clipPos[nr][2], pz);
        set_tracked(c, c->clipNormal[nr][0], sx / l);
        set_tracked(c, c->clipNormal[nr][1], sy / l);
        set_tracked(c, c->clipNormal[nr][2], sz / l);
        c->clipSaved[nr][0] = nx;
        c->clipSaved[nr][1] = ny;
        c->clipSaved[nr][2] = nz;
        return 1;
    }

    int isaac_hip_add_clipping(isaac_hip_ctx* c, float px, float py, float pz, float nx, float ny, float nz)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(isaac_hip_edit_clipping(c, uint32_t(c->nclip), px, py, pz, nx, ny, nz) == 1)
        {
            c->nclip++;
            c->generation++;
            return 1;
        }
        return 0;
    }

    int isaac_hip_remove_clipping(isaac_hip_ctx* c, uint32_t nr)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(nr >= uint32_t(c->nclip))
            return 0;
        c->generation++;
        c->nclip--;
        for(int i = int(nr); i < c->nclip; i++)
            for(int k = 0; k < 3; k++)
            {
                c->clipPos[i][k] = c->clipPos[i + 1][k];
                c->clipNormal[i][k] = c->clipNormal[i + 1][k];
                c->clipSaved[i][k] = c->clipSaved[i + 1][k];
            }
        return 0;
    }

    int isaac_hip_clipping_count(const isaac_hip_ctx* c)
    {
        return c ? c->nclip : ISAAC_HIP_ERR_ARG;
    }

    int isaac_hip_get_clipping(const isaac_hip_ctx* c, uint32_t nr, float position[3], float normal[3])
    {
        if(!c || !position || !normal || nr >= uint32_t(c->nclip))
            return ISAAC_HIP_ERR_ARG;
        for(int k = 0; k < 3; k++)
        {
            position[k] = c->clipPos[nr][k];
            normal[k] = c->clipSaved[nr][k];
        }
        return 0;
    }

    int isaac_hip_update_position(isaac_hip_ctx* c, const int32_t position[3])
    {
        if(!c || !position)
            return ISAAC_HIP_ERR_ARG;
        for(int i = 0; i < 3; i++)
            set_tracked(c, c->position[i], position[i]);
        derive_sizes(c);
        return 0;
    }

    int isaac_hip_update_local_size(isaac_hip_ctx* c, const int64_t ls[3])
    {
        if(!c || !ls)
            return ISAAC_HIP_ERR_ARG;
        for(int i = 0; i < 3; i++)
            if(ls[i] <= 0 || ls[i] > (1 << 20))
                return fail(c, ISAAC_HIP_ERR_ARG, "bad local size");
        for(int i = 0; i < 3; i++)
            set_tracked(c, c->local[i], ls[i]);
        derive_sizes(c);
        return alloc_textures(c);
    }

    int isaac_hip_update_local_particle_size(isaac_hip_ctx* c, const int64_t ps[3])
    {
        if(!c || !ps)
            return ISAAC_HIP_ERR_ARG;
        for(int i = 0; i < 3; i++)
            if(ps[i] <= 0)
                return fail(c, ISAAC_HIP_ERR_ARG, "bad particle grid");
        for(int i = 0; i < 3; i++)
            set_tracked(c, c->pgrid[i], ps[i]);
        return 0;
    }

    int isaac_hip_global_size_scaled(const isaac_hip_ctx* c, int32_t out[3])
    {
        if(!c || !out)
            return ISAAC_HIP_ERR_ARG;
        for(int i = 0; i < 3; i++)
            out[i] = int32_t(c->gsScaled[i]);
        return 0;
    }

    int isaac_hip_update_buffers(isaac_hip_ctx* c)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        HIPCHK(c, hipSetDevice(c->device));
        const int T = c->cfg.transfer_size;
        const int nvf = c->nv + c->nf;
        for(int s = 0; s < nvf; s++)
            if(!c->vols[s].set && (c->weight[s] + c->iso[s] > 0.0f))
                return fail(c, ISAAC_HIP_ERR_STATE, "volume source " + std::to_string(s) + " has no data");
        if(c->lutDirty)
        {
            std::vector<float> all(size_t(c->nsrc) * T * 4);
            for(int s = 0; s < c->nsrc; s++)
                std::memcpy(all.data() + size_t(s) * T * 4, c->lutHost[s].data(), size_t(T) * 16);
            HIPCHK(c, hipMemcpyAsync(c->lutDev, all.data(), all.size() * 4, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            c->lutDirty = false;
        }
        HIPCHK(c, mark(c, 0));
        // non-persistent sources (UpdatePersistentTextureIterator, lib/isaac.hpp:249-290)
        for(int s = 0; s < nvf; s++)
        {
            VolSlot& v = c->vols[s];
            const bool enabled = c->weight[s] != 0.0f || c->iso[s] != 0.0f;
            if(!v.set || v.desc.persistent || !enabled)
                continue;
            int rc = ensure_np_texture(c, v);
            if(rc)
                return rc;
            const int g = v.desc.guard_size;
            HIPCHK(
                c,
                launch_persistent_copy(
                    (const float*) v.desc.data, v.desc.stride, v.desc.feature_dim, c->chain[s], int(c->local[0] + 2 * g),
                    int(c->local[1] + 2 * g), int(c->local[2] + 2 * g), g, v.npTex, c->stream));
        }
        // field sources (UpdateAdvectionTextureIterator, lib/isaac.hpp:292-370): every
        // buffer update swaps the advection textures and advances the time step
        // (lib/isaac.hpp:1821-1826), then each enabled field source advects its
        // history along its unit field direction and reseeds from the noise
        if(c->nf > 0)
        {
            int rc = ensure_noise(c);
            if(rc)
                return rc;
            for(int s = c->nv; s < nvf; s++)
                c->vols[s].advCur ^= 1;
            c->timeStep++;
            const bool seeding = (c->timeStep % c->advSeedPeriod) < c->advSeedTime;
            for(int s = c->nv; s < nvf; s++)
            {
                VolSlot& v = c->vols[s];
                const bool enabled = c->weight[s] != 0.0f || c->iso[s] != 0.0f;
                if(!v.set || !enabled)
                    continue;
                if(v.desc.feature_dim != 3)
                    return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "field sources are float3 vector fields");
                rc = ensure_advection(c, v);
                if(rc)
                    return rc;
                AdvectParams A{};
                A.field = (const float*) v.desc.data;
                for(int k = 0; k < 3; k++)
                {
                    A.stride[k] = v.desc.stride[k];
                    A.scale[k] = c->scale[k];
                }
                A.noise = c->noise;
                A.history = v.adv[v.advCur ^ 1];
                A.out = v.adv[v.advCur];
                A.X = int32_t(c->local[0]);
                A.Y = int32_t(c->local[1]);
                A.Z = int32_t(c->local[2]);
                A.stepSize = c->advStepFactor * float(kMaxAdvectionStep);
                A.historyWeight = c->advHistoryWeight;
                A.seeding = seeding ? 1 : 0;
                HIPCHK(c, launch_advect(A, c->stream));
            }
        }
        HIPCHK(c, mark(c, 1));
        if(c->form == ISAAC_HIP_FORM_BRICKS)
        {
            // brick form: re-lay every enabled volume/field source (the raw
            // source when persistent, its functor-applied texture otherwise)
            int active = 0;
            for(int s = 0; s < nvf; s++)
            {
                VolSlot& v = c->vols[s];
                v.bricked = false;
                if(!v.set || c->weight[s] + c->iso[s] <= 0.0f)
                    continue;
                if(s >= c->nv)
                    return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "brick form: field sources (advection) render in the combined form");
                if(++active > kMaxBrickSources)
                    return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "brick form: at most 4 enabled volume/field sources");
                const float* src;
                int64_t st[3];
                int dim;
                if(v.desc.persistent)
                {
                    src = (const float*) v.desc.data;
                    dim = v.desc.feature_dim;
                    for(int k = 0; k < 3; k++)
                        st[k] = v.desc.stride[k];
                }
                else
                {
                    const int64_t g = v.desc.guard_size;
                    const int64_t gX = c->local[0] + 2 * g, gY = c->local[1] + 2 * g;
                    src = v.npTex + (g + g * gX + g * gX * gY);
                    dim = 1;
                    st[0] = 1;
                    st[1] = gX;
                    st[2] = gX * gY;
                }
                int rc = ensure_bricks(c, v, dim);
                if(rc)
                    return rc;
                HIPCHK(c, launch_bricks(src, st, dim, int(c->local[0]), int(c->local[1]), int(c->local[2]), c->brickDtype,
                                        v.brickScale, v.brick, c->stream));
                v.bricked = true;
            }
            // the particle records after the GPU-heavy brick build, so the host enqueues
            // their short launches while the build runs
            HIPCHK(c, mark(c, 2));
            if(int rc = bin_particles(c))
                return rc;
            HIPCHK(c, mark(c, 3));
            c->evRecorded[0] = c->timingEvents == 1;
            c->classified = true;
            c->builtForm = ISAAC_HIP_FORM_BRICKS;
            return 0;
        }
        // classification (lib/isaac.hpp:1916-1947)
        ClassifyParams P{};
        P.X = int32_t(c->local[0]);
        P.Y = int32_t(c->local[1]);
        P.Z = int32_t(c->local[2]);
        P.nvox = int64_t(P.X) * P.Y * P.Z;
        if(P.nvox >= (int64_t(1) << 32))
            return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "local volume above 2^32 voxels");
        P.nsrc = nvf;
        float tw = 0.0f;
        for(int s = 0; s < nvf; s++)
            tw += c->weight[s];
        P.totalWeight = tw;
        P.invTotalWeight = tw != 0.0f ? 1.0f / tw : 0.0f;
        P.ditherMode = c->ditherMode;
        P.T = T;
        make_fastdiv(uint32_t(P.X), P.divX_mul, P.divX_shift);
        make_fastdiv(uint32_t(P.X * P.Y), P.divXY_mul, P.divXY_shift);
        P.luts = c->lutDev;
        P.tex = c->tex;
        if(c->tuneTexLayout == kTexHalf)
        {
            if(!c->texH)
            {
                // one padding texel, as for the combined texture
                HIPCHK(c, hipMalloc(&c->texH, (c->texVox + 1) * sizeof(uint2)));
                HIPCHK(c, hipMemsetAsync(c->texH + c->texVox, 0, sizeof(uint2), c->stream));
            }
            P.texH = c->texH;
        }
        const int maxLds = int((64 * 1024) / (size_t(T) * sizeof(float4)));
        P.nLut = std::min(nvf, std::max(0, maxLds));
        for(int s = 0; s < nvf; s++)
        {
            ClassifySource& S = P.src[s];
            const VolSlot& v = c->vols[s];
            S.weight = c->weight[s];
            S.iso = c->iso[s];
            S.invIso = S.iso > 0.0f ? 1.0f / S.iso : 0.0f;
            S.active = (S.weight + S.iso > 0.0f) ? 1 : 0;
            S.lutSlot = s;
            if(!v.set)
            {
                S.active = 0;
                S.data = nullptr;
                S.dim = 1;
                continue;
            }
            if(v.desc.persistent)
            {
                S.data = (const float*) v.desc.data;
                S.dim = v.desc.feature_dim;
                for(int k = 0; k < 3; k++)
                    S.stride[k] = v.desc.stride[k];
                S.chain = c->chain[s];
            }
            else
            {
                const int64_t g = v.desc.guard_size;
                const int64_t gX = c->local[0] + 2 * g, gY = c->local[1] + 2 * g;
                S.data = v.npTex ? v.npTex + (g + g * gX + g * gX * gY) : nullptr;
                S.dim = 1;
                S.stride[0] = 1;
                S.stride[1] = gX;
                S.stride[2] = gX * gY;
                S.chain = ChainDesc{};
                if(!S.data)
                    S.active = 0;
            }
            S.adv = (s >= c->nv && S.active) ? v.adv[v.advCur] : nullptr;
        }
        HIPCHK(c, launch_classify(P, c->stream));
        if(c->tuneTexLayout == kTexBrick)
        {
            const size_t nb = size_t((c->local[0] + kTexBrick - 1) / kTexBrick) * size_t((c->local[1] + kTexBrick - 1) / kTexBrick)
                * size_t((c->local[2] + kTexBrick - 1) / kTexBrick);
            if(nb != c->texBrickCount)
            {
                if(c->texBricks)
                    HIPCHK(c, hipFree(c->texBricks));
                c->texBricks = nullptr;
                c->texBrickCount = 0;
                // one padding texel: the last brick's last row is read 16 bytes at a time too
                HIPCHK(c, hipMalloc(&c->texBricks, (nb * kTexBrickTexels + 1) * sizeof(uint2)));
                HIPCHK(c, hipMemsetAsync(c->texBricks + nb * kTexBrickTexels, 0, sizeof(uint2), c->stream));
                c->texBrickCount = nb;
            }
            HIPCHK(c, launch_tex_bricks(c->tex, c->texBricks, int(c->local[0]), int(c->local[1]), int(c->local[2]), c->stream));
        }
        // the particle records after the classification, so the host enqueues their short
        // launches while the classification runs
        HIPCHK(c, mark(c, 2));
        if(int rc = bin_particles(c))
            return rc;
        HIPCHK(c, mark(c, 3));
        c->evRecorded[0] = c->timingEvents == 1;
        c->classified = true;
        c->builtForm = ISAAC_HIP_FORM_COMBINED;
        c->builtTexLayout = c->tuneTexLayout;
        return 0;
    }

    int isaac_hip_render(
        isaac_hip_ctx* c,
        const double proj[16],
        const double mv[16],
        const float bg[4],
        const int32_t viewport[4],
        uint8_t* out_rgba,
        int out_is_device)
    {
        if(!c || !proj || !mv)
            return ISAAC_HIP_ERR_ARG;
        HIPCHK(c, hipSetDevice(c->device));
        if(!c->classified)
            return fail(c, ISAAC_HIP_ERR_STATE, "isaac_hip_update_buffers must run before the first render");
        if(c->builtForm != c->form)
            return fail(c, ISAAC_HIP_ERR_STATE, "render form changed; call isaac_hip_update_buffers");
        if(c->form == ISAAC_HIP_FORM_COMBINED && c->builtTexLayout != c->tuneTexLayout)
            return fail(c, ISAAC_HIP_ERR_STATE, "texture layout changed; call isaac_hip_update_buffers");
        const int W = c->cfg.framebuffer_size[0], H = c->cfg.framebuffer_size[1];
        FrameTargets tg{W, H, c->color, c->depth, c->normal, c->ao, c->volColor, c->floatOut ? c->outFloat : nullptr, c->counters};
        FrameParams F;
        PostParams Q;
        ParticleParams PP;
        {
            int rc = build_frame(c, proj, mv, bg, viewport, tg, F, Q, PP);
            if(rc)
                return rc;
        }
        if(c->counting)
            HIPCHK(c, launch_fill_u32(reinterpret_cast<uint32_t*>(c->counters), 0u, 2 * ISAAC_HIP_CNT_COUNT, c->stream));
        HIPCHK(c, mark(c, 4));
        if(c->form == ISAAC_HIP_FORM_BRICKS)
            HIPCHK(c, launch_brick_march(F, Q, PP, c->counting, c->stream));
        else
            HIPCHK(c, launch_march(F, PP, c->counting, c->stream));
        HIPCHK(c, mark(c, 5));
        bool fused = false;
        if(F.aoEnabled)
        {
            if(c->aoMode == ISAAC_HIP_AO_LEARNED)
            {
                int rc = aonet_occlusion(c);
                if(rc)
                    return rc;
            }
            else
            {
                // the occlusion kernel shades its own pixels where the mode allows
                fused = c->tuneFusedShading && ssao_shades(F);
                HIPCHK(c, launch_ssao(F, c->stream, fused));
            }
        }
        HIPCHK(c, mark(c, 6));
        if(!fused)
            HIPCHK(c, launch_shade(F, c->stream));
        HIPCHK(c, mark(c, 7));
        if(out_rgba)
        {
            const size_t bytes = size_t(W) * size_t(H) * 4;
            HIPCHK(c, hipMemcpyAsync(out_rgba, c->color, bytes, out_is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, mark(c, 8));
        c->evRecorded[1] = c->timingEvents != 0;
        c->evMarchOnly = c->timingEvents == 2;
        c->rendered = true;
        if(out_rgba && !out_is_device)
            HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }

    int isaac_hip_video_frame_message(isaac_hip_ctx* c, int32_t quality, char* msg, int64_t capacity, int64_t* length)
    {
        if(!c || !msg || !length || quality < 0)
            return ISAAC_HIP_ERR_ARG;
        if(!c->rendered)
            return fail(c, ISAAC_HIP_ERR_STATE, "no frame rendered yet");
        HIPCHK(c, hipSetDevice(c->device));
        const int W = c->cfg.framebuffer_size[0], H = c->cfg.framebuffer_size[1];
        const uint8_t* rgba = static_cast<const uint8_t*>(static_cast<void*>(c->color));
        const int64_t need = quality > 0 ? isaac_hip_jpeg_workspace_bytes(W, H) : (8 * int64_t(W) * H * 4 + 5) / 6 + 16;
        if(need < 0)
            return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "video frames are 1..65535 pixels a side");
        if(c->videoWsBytes < need)
        {
            if(c->videoWs)
                (void) hipFree(c->videoWs);
            c->videoWs = nullptr;
            c->videoWsBytes = 0;
            HIPCHK(c, hipMalloc(&c->videoWs, size_t(need)));
            c->videoWsBytes = need;
        }
        const int rc = quality > 0
            ? isaac_hip_jpeg_frame_message(rgba, W, H, quality, c->videoWs, c->videoWsBytes, msg, capacity, length, c->stream)
            : isaac_hip_frame_message(rgba, W, H, static_cast<char*>(c->videoWs), msg, capacity, length, c->stream);
        return rc ? fail(c, rc, rc == ISAAC_HIP_ERR_ARG ? "message buffer too small" : "video frame encoding failed") : 0;
    }

    int64_t isaac_hip_video_frame_capacity(const isaac_hip_ctx* c)
    {
        if(!c)
            return -1;
        const int64_t n = int64_t(c->cfg.framebuffer_size[0]) * c->cfg.framebuffer_size[1] * 4;
        // raw RGBA base64, or a JPEG (at most twice the raw size with its markers) in base64
        return (8 * (2 * n + 4096) + 5) / 6 + 64;
    }

    int isaac_hip_synchronize(isaac_hip_ctx* c)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }

    int isaac_hip_get_timings(const isaac_hip_ctx* c, isaac_hip_timings* out)
    {
        if(!c || !out)
            return ISAAC_HIP_ERR_ARG;
        (void) hipStreamSynchronize(c->stream);
        *out = isaac_hip_timings{};
        if(c->evRecorded[0])
        {
            // update order: source copies (ev 0-1), classification or brick build (1-2),
            // particle records (2-3)
            out->source_copy_us = elapsed_us(c->ev[0], c->ev[1]);
            out->optimization_buffer_us = elapsed_us(c->ev[1], c->ev[2]);
            out->binning_us = elapsed_us(c->ev[2], c->ev[3]);
            out->buffer_us = out->source_copy_us + out->binning_us;
        }
        if(c->evRecorded[1])
            out->march_us = elapsed_us(c->ev[4], c->ev[5]);
        if(c->evRecorded[1] && !c->evMarchOnly)
        {
            out->ssao_us = elapsed_us(c->ev[5], c->ev[6]);
            out->shade_us = elapsed_us(c->ev[6], c->ev[7]);
            out->kernel_us = elapsed_us(c->ev[4], c->ev[7]);
            out->copy_us = elapsed_us(c->ev[7], c->ev[8]);
        }
        return 0;
    }

    int isaac_hip_copy_to_host(isaac_hip_ctx* c, void* dst, const void* src, size_t bytes)
    {
        if(!c || (!dst && bytes) || (!src && bytes))
            return ISAAC_HIP_ERR_ARG;
        HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }

    int isaac_hip_set_tuning(isaac_hip_ctx* c, int32_t key, int32_t value)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        switch(key)
        {
        case ISAAC_HIP_TUNE_MARCH_PREFETCH:
            c->tunePrefetch = value != 0;
            return 0;
        case ISAAC_HIP_TUNE_MARCH_WAVES:
            if(value != 0 && value != 6 && value != 8)
                return fail(c, ISAAC_HIP_ERR_ARG, "march waves: 0, 6 or 8");
            c->tuneWaves = value;
            return 0;
        case ISAAC_HIP_TUNE_MARCH_TILES:
            if(value != 0 && value != 2 && value != 4 && value != 8 && value != 16)
                return fail(c, ISAAC_HIP_ERR_ARG, "march tile order: 0 (XCD bands) or a super-tile side of 2, 4, 8 or 16 tiles");
            c->tuneTiles = value;
            return 0;
        case ISAAC_HIP_TUNE_MARCH_THREADS:
            if(value != 64 && value != 256)
                return fail(c, ISAAC_HIP_ERR_ARG, "march workgroup: 64 or 256 threads");
            c->tuneMarchThreads = value;
            return 0;
        case ISAAC_HIP_TUNE_FUSED_SHADING:
            c->tuneFusedShading = value != 0;
            return 0;
        case ISAAC_HIP_TUNE_TIMING_EVENTS:
            if(value < 0 || value > 2)
                return fail(c, ISAAC_HIP_ERR_ARG, "timing events: 0 (none), 1 (all) or 2 (around the march)");
            c->timingEvents = value;
            return 0;
        case ISAAC_HIP_TUNE_MARCH_OFFSETS:
            if(value != 32 && value != 64)
                return fail(c, ISAAC_HIP_ERR_ARG, "march load addressing: 32 (byte offsets) or 64 (addresses)");
            c->tuneMarchOffsets = value;
            return 0;
        case ISAAC_HIP_TUNE_TEX_LAYOUT:
            if(value != 0 && value != kTexBrick && value != kTexHalf)
                return fail(c, ISAAC_HIP_ERR_ARG, "texture layout: 0 (linear), 8 (8^3 bricks) or 16 (binary16 copy)");
            c->tuneTexLayout = value;
            return 0;
        default:
            return fail(c, ISAAC_HIP_ERR_ARG, "unknown tuning key");
        }
    }

    int isaac_hip_set_render_form(isaac_hip_ctx* c, int32_t form, int32_t brick_dtype)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(form != ISAAC_HIP_FORM_COMBINED && form != ISAAC_HIP_FORM_BRICKS)
            return fail(c, ISAAC_HIP_ERR_ARG, "unknown render form");
        if(brick_dtype != ISAAC_HIP_F32 && brick_dtype != ISAAC_HIP_F16 && brick_dtype != ISAAC_HIP_BF16)
            return fail(c, ISAAC_HIP_ERR_ARG, "brick dtype must be ISAAC_HIP_F32, ISAAC_HIP_F16 or ISAAC_HIP_BF16");
        set_tracked(c, c->form, int(form));
        set_tracked(c, c->brickDtype, int(brick_dtype));
        return 0;
    }

    int isaac_hip_get_bricks(
        const isaac_hip_ctx* c,
        int source,
        const void** pairs,
        const void** scales,
        int64_t* nbricks,
        int32_t* pair_bytes)
    {
        if(!c || source < 0 || source >= c->nv + c->nf)
            return ISAAC_HIP_ERR_ARG;
        const VolSlot& v = c->vols[source];
        if(pairs)
            *pairs = v.bricked ? v.brick : nullptr;
        if(scales)
            *scales = v.bricked ? v.brickScale : nullptr;
        if(nbricks)
            *nbricks = v.bricked ? v.nbricks : 0;
        if(pair_bytes)
            *pair_bytes = c->brickDtype == ISAAC_HIP_F32 ? 8 : 4;
        return 0;
    }

    int isaac_hip_set_float_output(isaac_hip_ctx* c, int enabled)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(enabled && !c->outFloat)
        {
            const size_t npx = size_t(c->cfg.framebuffer_size[0]) * size_t(c->cfg.framebuffer_size[1]);
            HIPCHK(c, hipMalloc(&c->outFloat, npx * sizeof(float4)));
            HIPCHK(c, launch_fill_u32(reinterpret_cast<uint32_t*>(c->outFloat), 0u, npx * 4, c->stream));
        }
        c->floatOut = enabled != 0;
        return 0;
    }

    int isaac_hip_get_float_image(const isaac_hip_ctx* c, const void** rgba_f32)
    {
        if(!c || !rgba_f32)
            return ISAAC_HIP_ERR_ARG;
        *rgba_f32 = c->floatOut ? c->outFloat : nullptr;
        return 0;
    }

    int isaac_hip_set_counting(isaac_hip_ctx* c, int enabled)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        c->counting = enabled != 0;
        return 0;
    }

    int isaac_hip_get_counters(isaac_hip_ctx* c, uint64_t out[ISAAC_HIP_CNT_COUNT])
    {
        if(!c || !out)
            return ISAAC_HIP_ERR_ARG;
        unsigned long long h[16];
        HIPCHK(c, hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for(int i = 0; i < ISAAC_HIP_CNT_COUNT; i++)
            out[i] = h[i];
        return 0;
    }

    int isaac_hip_get_gbuffer(const isaac_hip_ctx* c, const void** color, const void** depth, const void** normal, const void** ao)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(color)
            *color = c->color;
        if(depth)
            *depth = c->depth;
        if(normal)
            *normal = c->normal;
        if(ao)
            *ao = c->ao;
        return 0;
    }

    int isaac_hip_get_combined_textures(const isaac_hip_ctx* c, const void** texels, int32_t* texel_bytes)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(texels)
            *texels = c->tex;
        if(texel_bytes)
            *texel_bytes = 8;
        return 0;
    }

    int isaac_hip_min_max(isaac_hip_ctx* c, float* min_out, float* max_out, int32_t count)
    {
        if(!c || !min_out || !max_out)
            return ISAAC_HIP_ERR_ARG;
        const int nvf = c->nv + c->nf, total = c->nsrc;
        if(count < total)
            return fail(c, ISAAC_HIP_ERR_ARG, "min_max: count below the number of sources");
        HIPCHK(c, hipSetDevice(c->device));
        if(!c->mmPartials)
            HIPCHK(c, hipMalloc(&c->mmPartials, sizeof(MinMaxPartial) * kMinMaxGroups));
        if(!c->mmResult)
            HIPCHK(c, hipMalloc(&c->mmResult, sizeof(float2) * std::max(total, 1)));
        std::vector<char> computed(static_cast<size_t>(total), 0);
        for(int s = 0; s < nvf; s++)
        {
            const VolSlot& v = c->vols[s];
            MinMaxVolumeParams P{};
            P.X = int32_t(c->local[0]);
            P.Y = int32_t(c->local[1]);
            P.Z = int32_t(c->local[2]);
            P.nvox = int64_t(P.X) * P.Y * P.Z;
            if(!v.set || P.nvox == 0)
                continue;
            if(P.nvox >= (int64_t(1) << 32))
                return fail(c, ISAAC_HIP_ERR_UNSUPPORTED, "min_max: local volume above 2^32 voxels");
            make_fastdiv(uint32_t(P.X), P.divX_mul, P.divX_shift);
            make_fastdiv(uint32_t(P.X * P.Y), P.divXY_mul, P.divXY_shift);
            if(v.desc.persistent)
            {
                P.data = (const float*) v.desc.data;
                P.dim = v.desc.feature_dim;
                for(int k = 0; k < 3; k++)
                    P.stride[k] = v.desc.stride[k];
                P.applyChain = 1;
                P.chain = c->chain[s];
            }
            else
            {
                // the texture the last buffer update filled with the chain values
                if(!v.npTex)
                    continue;
                const int64_t g = v.desc.guard_size;
                const int64_t gX = c->local[0] + 2 * g, gY = c->local[1] + 2 * g;
                P.data = v.npTex + (g + g * gX + g * gX * gY);
                P.dim = 1;
                P.stride[0] = 1;
                P.stride[1] = gX;
                P.stride[2] = gX * gY;
                P.applyChain = 0;
            }
            HIPCHK(c, launch_min_max_volume(P, c->mmPartials, c->mmResult + s, c->stream));
            computed[size_t(s)] = 1;
        }
        for(int s = 0; s < c->np; s++)
        {
            const PartSlot& p = c->parts[s];
            if(!p.set)
                continue;
            MinMaxParticleParams P{};
            P.pos = p.desc.position;
            P.attr = p.desc.attribute;
            P.dim = p.desc.feature_dim;
            P.n = p.desc.count;
            for(int i = 0; i < 3; i++)
            {
                P.grid[i] = int32_t(c->pgrid[i]);
                P.cw[i] = float(c->local[i]) / float(c->pgrid[i]);
            }
            P.chain = c->chain[nvf + s];
            HIPCHK(c, launch_min_max_particles(P, c->mmPartials, c->mmResult + nvf + s, c->stream));
            computed[size_t(nvf + s)] = 1;
        }
        std::vector<float2> res(static_cast<size_t>(total));
        if(total)
            HIPCHK(c, hipMemcpyAsync(res.data(), c->mmResult, sizeof(float2) * total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for(int s = 0; s < total; s++)
        {
            min_out[s] = computed[size_t(s)] ? res[size_t(s)].x : FLT_MAX;
            max_out[s] = computed[size_t(s)] ? res[size_t(s)].y : -FLT_MAX;
        }
        return 0;
    }

    int isaac_hip_get_cell_list(isaac_hip_ctx* c, int idx, const void** counts, const void** offsets, const void** perm)
    {
        if(!c || idx < 0 || idx >= c->np)
            return ISAAC_HIP_ERR_ARG;
        PartSlot& p = c->parts[idx];
        if(!p.lastValid)
            return fail(c, ISAAC_HIP_ERR_STATE, "no cell list: the particle source was not binned by update_buffers");
        if(!p.fullValid)
        {
            // the full list of all particles from the last update's inputs; its render
            // records equal the visible build's
            HIPCHK(c, hipSetDevice(c->device));
            HIPCHK(c, launch_bin(p.last, p.B, c->stream));
            p.fullValid = true;
        }
        if(counts)
            *counts = p.B.count;
        if(offsets)
            *offsets = p.B.offset;
        if(perm)
            *perm = p.B.perm;
        return 0;
    }
}
