// Host side of the learned occlusion pass: the U-Net's parameters, workspaces,
// training loop (forward, backward, Adam) and the inference call the render
// path makes in place of the SSAO kernel.  The network has no counterpart in
// the reference; it approximates the reference's SSAO term
// (lib/isaac/isaac_ssao_kernel.hpp:161-182) from the G-buffer.
#include "aonet.hpp"
#include "context.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace isaac_hip
{
    namespace
    {
        struct LayerDef
        {
            int cin, cout;
        };
        // e0a e0b e1a e1b e2a e2b e3a e3b d2 d1 d0 (channels of the flat parameter layout)
        constexpr int kNumLayers = 11;
        constexpr LayerDef kLayers[kNumLayers] = {{8, 16},  {16, 16}, {16, 32}, {32, 32},      {32, 64},     {64, 64},
                                                  {64, 64}, {64, 64}, {64 + 64, 64}, {64 + 32, 32}, {32 + 16, 16}};
        // level (resolution divisor 2^level) of each layer
        constexpr int kLevel[kNumLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0};
        constexpr int kFeatures = 8;
        constexpr int kHeadIn = 16;
        constexpr int kMaxBatch = 64;

        // forward weight row length: 9*cin rounded up to the forward MFMA's K of 32
        int kpad(int cin)
        {
            return (9 * cin + 31) / 32 * 32;
        }
    } // namespace

    // One set of activation (and optionally gradient) buffers for n images of h x w.
    struct AoWorkspace
    {
        int n = 0, h = 0, w = 0;
        bool grads = false;
        std::vector<void*> allocs;
        _Float16* x0 = nullptr;
        _Float16* act[kNumLayers] = {};
        _Float16 *p0 = nullptr, *p1 = nullptr, *p2 = nullptr;
        float* target = nullptr;
        float* out = nullptr;
        float* loss = nullptr;
        int32_t* origins = nullptr;
        float* g[kNumLayers] = {};
        float *gp0 = nullptr, *gp1 = nullptr, *gp2 = nullptr, *gc2 = nullptr, *gc1 = nullptr, *gc0 = nullptr;
        _Float16* dz = nullptr;

        void release()
        {
            for(void* p : allocs)
                (void) hipFree(p);
            allocs.clear();
            *this = AoWorkspace{};
        }

        template<class T>
        hipError_t get(T*& p, size_t count)
        {
            void* q = nullptr;
            hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
            if(e != hipSuccess)
                return e;
            allocs.push_back(q);
            p = static_cast<T*>(q);
            return hipSuccess;
        }

        hipError_t ensure(int n_, int h_, int w_, bool grads_)
        {
            if(n_ <= n && h_ == h && w_ == w && (grads || !grads_))
                return hipSuccess;
            release();
            const size_t full = size_t(n_) * h_ * w_, half = full / 4, quarter = full / 16, eighth = full / 64;
            const size_t lvl[4] = {full, half, quarter, eighth};
            hipError_t e = hipSuccess;
            auto chk = [&](hipError_t r) {
                if(e == hipSuccess)
                    e = r;
            };
            chk(get(x0, full * kFeatures));
            for(int i = 0; i < kNumLayers; i++)
                chk(get(act[i], lvl[kLevel[i]] * kLayers[i].cout));
            chk(get(p0, half * 16));
            chk(get(p1, quarter * 32));
            chk(get(p2, eighth * 64));
            chk(get(target, full));
            chk(get(out, full));
            chk(get(loss, 1));
            chk(get(origins, size_t(2) * n_));
            if(grads_)
            {
                for(int i = 0; i < kNumLayers; i++)
                    chk(get(g[i], lvl[kLevel[i]] * kLayers[i].cout));
                chk(get(gp0, half * 16));
                chk(get(gp1, quarter * 32));
                chk(get(gp2, eighth * 64));
                chk(get(gc2, quarter * 128));
                chk(get(gc1, half * 96));
                chk(get(gc0, full * 48));
                chk(get(dz, full * 16));
            }
            if(e != hipSuccess)
            {
                release();
                return e;
            }
            n = n_;
            h = h_;
            w = w_;
            grads = grads_;
            return hipSuccess;
        }
    };

    struct AoNet
    {
        int64_t nparams = 0;
        int64_t woff[kNumLayers] = {}, boff[kNumLayers] = {}, headOff = 0;
        int kp[kNumLayers] = {};
        float *master = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
        _Float16* wf[kNumLayers] = {};
        _Float16* wb[kNumLayers] = {};
        int64_t step = 0;
        bool ready = false;
        size_t maxLds = 64 * 1024; // LDS bytes per workgroup on this device
        AoWorkspace infer, train;
        int batchCount = 0, batchSide = 0;

        ~AoNet()
        {
            infer.release();
            train.release();
            for(void* p : {(void*) master, (void*) grad, (void*) m, (void*) v})
                if(p)
                    (void) hipFree(p);
            for(int i = 0; i < kNumLayers; i++)
            {
                if(wf[i])
                    (void) hipFree(wf[i]);
                if(wb[i])
                    (void) hipFree(wb[i]);
            }
        }
    };

    int64_t aonet_param_count()
    {
        int64_t n = 0;
        for(int i = 0; i < kNumLayers; i++)
            n += int64_t(kLayers[i].cout) * 9 * kLayers[i].cin + kLayers[i].cout;
        return n + kHeadIn + 1;
    }

    namespace
    {
        int create_net(isaac_hip_ctx* c)
        {
            if(c->aonet)
                return 0;
            AoNet* net = new AoNet();
            int64_t off = 0;
            for(int i = 0; i < kNumLayers; i++)
            {
                net->woff[i] = off;
                off += int64_t(kLayers[i].cout) * 9 * kLayers[i].cin;
                net->boff[i] = off;
                off += kLayers[i].cout;
                net->kp[i] = kpad(kLayers[i].cin);
            }
            net->headOff = off;
            net->nparams = off + kHeadIn + 1;
            int lds = 0;
            // the device's LDS per workgroup (160 KiB on gfx950), at most 160 KiB
            if(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && lds > 0)
                net->maxLds = std::min<size_t>(size_t(lds), 160 * 1024);
            hipError_t e = hipSuccess;
            auto chk = [&](hipError_t r) {
                if(e == hipSuccess)
                    e = r;
            };
            const size_t pb = size_t(net->nparams) * sizeof(float);
            chk(hipMalloc(&net->master, pb));
            chk(hipMalloc(&net->grad, pb));
            chk(hipMalloc(&net->m, pb));
            chk(hipMalloc(&net->v, pb));
            for(int i = 0; i < kNumLayers && e == hipSuccess; i++)
            {
                chk(hipMalloc(&net->wf[i], size_t(kLayers[i].cout) * net->kp[i] * sizeof(_Float16)));
                chk(hipMalloc(&net->wb[i], size_t(kLayers[i].cin) * 9 * kLayers[i].cout * sizeof(_Float16)));
            }
            if(e != hipSuccess)
            {
                delete net;
                return hip_fail(c, e, "occlusion network allocation");
            }
            c->aonet = net;
            return 0;
        }

        hipError_t pack(AoNet* net, hipStream_t s)
        {
            for(int i = 0; i < kNumLayers; i++)
            {
                hipError_t e = launch_pack_weights(net->master + net->woff[i], net->wf[i], net->wb[i], kLayers[i].cin,
                                                   kLayers[i].cout, net->kp[i], s);
                if(e != hipSuccess)
                    return e;
            }
            return hipSuccess;
        }

        // input of a decoder layer: upsampled `up` (ca channels) followed by the skip tensor
        hipError_t conv_fwd(
            AoNet* net,
            int i,
            const _Float16* in,
            _Float16* out,
            int n,
            int h,
            int w,
            hipStream_t s,
            const _Float16* up = nullptr,
            int ca = 0,
            _Float16* pool = nullptr)
        {
            ConvArgs a{};
            a.pool = pool;
            a.x = in;
            a.up = up;
            a.ca = ca;
            a.w = net->wf[i];
            a.bias = net->master + net->boff[i];
            a.y = out;
            a.n = n;
            a.h = h;
            a.w_ = w;
            a.cin = kLayers[i].cin;
            a.kp = net->kp[i];
            return launch_conv_fwd(a, kLayers[i].cout, net->maxLds, s);
        }

        // U-Net forward over ws.x0 (n images of h x w, multiples of 8); the pools are
        // fused into the layers before them.  With `ao`, the head is fused into the
        // last layer (clamped occlusion of the imgW x imgH corner of each image into
        // ao) and that layer's activation is not stored.
        hipError_t forward(
            AoNet* net,
            AoWorkspace& ws,
            int n,
            int h,
            int w,
            hipStream_t s,
            float* ao = nullptr,
            int imgW = 0,
            int imgH = 0)
        {
            const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
            hipError_t e;
#define AOCHK(x)                                                                                                      \
    if((e = (x)) != hipSuccess)                                                                                       \
    return e
            AOCHK(conv_fwd(net, 0, ws.x0, ws.act[0], n, h, w, s));
            AOCHK(conv_fwd(net, 1, ws.act[0], ws.act[1], n, h, w, s, nullptr, 0, ws.p0));
            AOCHK(conv_fwd(net, 2, ws.p0, ws.act[2], n, h1, w1, s));
            AOCHK(conv_fwd(net, 3, ws.act[2], ws.act[3], n, h1, w1, s, nullptr, 0, ws.p1));
            AOCHK(conv_fwd(net, 4, ws.p1, ws.act[4], n, h2, w2, s));
            AOCHK(conv_fwd(net, 5, ws.act[4], ws.act[5], n, h2, w2, s, nullptr, 0, ws.p2));
            AOCHK(conv_fwd(net, 6, ws.p2, ws.act[6], n, h3, w3, s));
            AOCHK(conv_fwd(net, 7, ws.act[6], ws.act[7], n, h3, w3, s));
            AOCHK(conv_fwd(net, 8, ws.act[5], ws.act[8], n, h2, w2, s, ws.act[7], 64));
            AOCHK(conv_fwd(net, 9, ws.act[3], ws.act[9], n, h1, w1, s, ws.act[8], 64));
            if(!ao)
                return conv_fwd(net, 10, ws.act[1], ws.act[10], n, h, w, s, ws.act[9], 32);
            ConvArgs a{};
            a.x = ws.act[1];
            a.up = ws.act[9];
            a.ca = 32;
            a.w = net->wf[10];
            a.bias = net->master + net->boff[10];
            a.y = nullptr;
            a.n = n;
            a.h = h;
            a.w_ = w;
            a.cin = kLayers[10].cin;
            a.kp = net->kp[10];
            a.head = ao;
            a.headW = net->master + net->headOff;
            a.imgW = imgW;
            a.imgH = imgH;
            return launch_conv_fwd(a, kLayers[10].cout, net->maxLds, s);
        }

        // backward of layer i given ws.g[i] (gradient wrt its ReLU output): weight and
        // bias gradients, and (when gin != nullptr) += gradient wrt its input
        hipError_t layer_bwd(
            AoNet* net,
            AoWorkspace& ws,
            int i,
            const _Float16* in,
            float* gin,
            int n,
            int h,
            int w,
            hipStream_t s,
            const _Float16* up = nullptr,
            int ca = 0)
        {
            const int cin = kLayers[i].cin, cout = kLayers[i].cout;
            const int64_t npx = int64_t(n) * h * w;
            hipError_t e;
            AOCHK(launch_relu_bwd(ws.g[i], ws.act[i], ws.dz, npx * cout, s));
            AOCHK(launch_conv3x3_dw(in, up, ca, ws.dz, net->grad + net->woff[i], n, h, w, cin, cout, s));
            AOCHK(launch_bias_grad(ws.dz, net->grad + net->boff[i], npx, cout, s));
            if(gin)
            {
                ConvArgs a{};
                a.x = ws.dz;
                a.up = nullptr;
                a.ca = 0;
                a.w = net->wb[i];
                a.bias = nullptr;
                a.y = gin;
                a.n = n;
                a.h = h;
                a.w_ = w;
                a.cin = cout;
                a.kp = 9 * cout;
                AOCHK(launch_conv3x3_bwd(a, cin, net->maxLds, s));
            }
            return hipSuccess;
        }

        hipError_t zero(float* p, size_t count, hipStream_t s)
        {
            return launch_fill_u32(reinterpret_cast<uint32_t*>(p), 0u, count, s);
        }

        // loss and parameter gradients (into net->grad, zeroed first) of ws over n images
        hipError_t backward(AoNet* net, AoWorkspace& ws, int n, int h, int w, hipStream_t s)
        {
            const size_t full = size_t(n) * h * w, half = full / 4, quarter = full / 16, eighth = full / 64;
            const size_t lvl[4] = {full, half, quarter, eighth};
            const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
            hipError_t e;
            AOCHK(zero(net->grad, size_t(net->nparams), s));
            AOCHK(zero(ws.loss, 1, s));
            for(int i = 0; i < kNumLayers; i++)
                AOCHK(zero(ws.g[i], lvl[kLevel[i]] * kLayers[i].cout, s));
            AOCHK(zero(ws.gp0, half * 16, s));
            AOCHK(zero(ws.gp1, quarter * 32, s));
            AOCHK(zero(ws.gp2, eighth * 64, s));
            AOCHK(zero(ws.gc2, quarter * 128, s));
            AOCHK(zero(ws.gc1, half * 96, s));
            AOCHK(zero(ws.gc0, full * 48, s));
            AOCHK(launch_head_train(ws.act[10], net->master + net->headOff, ws.target, ws.loss, net->grad + net->headOff,
                                    ws.g[10], ws.out, int64_t(full), s));
            AOCHK(layer_bwd(net, ws, 10, ws.act[1], ws.gc0, n, h, w, s, ws.act[9], 32));
            AOCHK(launch_upcat_bwd(ws.gc0, ws.g[9], ws.g[1], n, h, w, 32, 16, s));
            AOCHK(layer_bwd(net, ws, 9, ws.act[3], ws.gc1, n, h1, w1, s, ws.act[8], 64));
            AOCHK(launch_upcat_bwd(ws.gc1, ws.g[8], ws.g[3], n, h1, w1, 64, 32, s));
            AOCHK(layer_bwd(net, ws, 8, ws.act[5], ws.gc2, n, h2, w2, s, ws.act[7], 64));
            AOCHK(launch_upcat_bwd(ws.gc2, ws.g[7], ws.g[5], n, h2, w2, 64, 64, s));
            AOCHK(layer_bwd(net, ws, 7, ws.act[6], ws.g[6], n, h3, w3, s));
            AOCHK(layer_bwd(net, ws, 6, ws.p2, ws.gp2, n, h3, w3, s));
            AOCHK(launch_pool2_bwd(ws.gp2, ws.g[5], n, h2, w2, 64, s));
            AOCHK(layer_bwd(net, ws, 5, ws.act[4], ws.g[4], n, h2, w2, s));
            AOCHK(layer_bwd(net, ws, 4, ws.p1, ws.gp1, n, h2, w2, s));
            AOCHK(launch_pool2_bwd(ws.gp1, ws.g[3], n, h1, w1, 32, s));
            AOCHK(layer_bwd(net, ws, 3, ws.act[2], ws.g[2], n, h1, w1, s));
            AOCHK(layer_bwd(net, ws, 2, ws.p0, ws.gp0, n, h1, w1, s));
            AOCHK(launch_pool2_bwd(ws.gp0, ws.g[1], n, h, w, 16, s));
            AOCHK(layer_bwd(net, ws, 1, ws.act[0], ws.g[0], n, h, w, s));
            AOCHK(layer_bwd(net, ws, 0, ws.x0, nullptr, n, h, w, s));
            return hipSuccess;
#undef AOCHK
        }

        int upload_features(isaac_hip_ctx* c, AoWorkspace& ws, int n, int h, int w, const float* features, const float* target)
        {
            const size_t full = size_t(n) * h * w;
            std::vector<_Float16> x(full * kFeatures);
            for(size_t i = 0; i < x.size(); i++)
                x[i] = _Float16(features[i]);
            HIPCHK(c, hipMemcpyAsync(ws.x0, x.data(), x.size() * sizeof(_Float16), hipMemcpyHostToDevice, c->stream));
            if(target)
                HIPCHK(c, hipMemcpyAsync(ws.target, target, full * sizeof(float), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            return 0;
        }

        int check_dims(isaac_hip_ctx* c, int n, int h, int w)
        {
            if(n <= 0 || h <= 0 || w <= 0 || h % 8 || w % 8)
                return fail(c, ISAAC_HIP_ERR_ARG, "occlusion network: n > 0 and h, w multiples of 8");
            if(!c->aonet || !c->aonet->ready)
                return fail(c, ISAAC_HIP_ERR_STATE, "occlusion network has no parameters (isaac_hip_aonet_set_params)");
            return 0;
        }
    } // namespace

    // the render path's occlusion pass: G-buffer depth/normal -> U-Net -> c->ao
    int aonet_occlusion(isaac_hip_ctx* c)
    {
        AoNet* net = c->aonet;
        if(!net || !net->ready)
            return fail(c, ISAAC_HIP_ERR_STATE, "learned occlusion selected without network parameters");
        const int W = c->cfg.framebuffer_size[0], H = c->cfg.framebuffer_size[1];
        const int Wp = (W + 7) / 8 * 8, Hp = (H + 7) / 8 * 8;
        HIPCHK(c, net->infer.ensure(1, Hp, Wp, false));
        FeatureArgs f{};
        f.depth = c->depth;
        f.normal = c->normal;
        f.imgW = W;
        f.imgH = H;
        f.invScale = 1.0f / float(c->maxGsScaled);
        f.x = net->infer.x0;
        f.n = 1;
        f.h = Hp;
        f.w = Wp;
        HIPCHK(c, launch_features(f, c->stream));
        HIPCHK(c, forward(net, net->infer, 1, Hp, Wp, c->stream, c->ao, W, H));
        return 0;
    }

    AoWorkspace* aonet_workspace_create()
    {
        return new AoWorkspace();
    }

    void aonet_workspace_destroy(AoWorkspace* ws)
    {
        if(!ws)
            return;
        ws->release();
        delete ws;
    }

    int aonet_reserve_frames(isaac_hip_ctx* c, AoWorkspace* ws, int chunk, int W, int H, bool* reallocated)
    {
        AoNet* net = c->aonet;
        if(!net || !net->ready)
            return fail(c, ISAAC_HIP_ERR_STATE, "learned occlusion selected without network parameters");
        if(!ws)
            return fail(c, ISAAC_HIP_ERR_ARG, "no occlusion workspace");
        const int Wp = (W + 7) / 8 * 8, Hp = (H + 7) / 8 * 8;
        const _Float16* before = ws->x0;
        HIPCHK(c, ws->ensure(chunk, Hp, Wp, false));
        if(reallocated)
            *reallocated = ws->x0 != before;
        return 0;
    }

    // learned occlusion of nframes frame-major G-buffer images (W x H each), `chunk` frames
    // per network evaluation; allocates nothing once aonet_reserve_frames has run
    int aonet_occlusion_frames(
        isaac_hip_ctx* c,
        AoWorkspace* ws,
        const float* depth,
        const float* normal,
        float* ao,
        int nframes,
        int W,
        int H,
        const float* invScaleDev,
        int chunk,
        hipStream_t s)
    {
        AoNet* net = c->aonet;
        if(!net || !net->ready)
            return fail(c, ISAAC_HIP_ERR_STATE, "learned occlusion selected without network parameters");
        const int Wp = (W + 7) / 8 * 8, Hp = (H + 7) / 8 * 8;
        if(!ws || ws->n < chunk || ws->h != Hp || ws->w != Wp)
            return fail(c, ISAAC_HIP_ERR_STATE, "batch occlusion workspace not reserved");
        const size_t img = size_t(W) * H;
        for(int f0 = 0; f0 < nframes; f0 += chunk)
        {
            const int n = std::min(chunk, nframes - f0);
            FeatureArgs fa{};
            fa.depth = depth + img * f0;
            fa.normal = normal + img * 3 * f0;
            fa.imgW = W;
            fa.imgH = H;
            fa.invScale = 1.0f;
            fa.invScaleItem = invScaleDev + f0;
            fa.imgStride = int64_t(img);
            fa.x = ws->x0;
            fa.n = n;
            fa.h = Hp;
            fa.w = Wp;
            HIPCHK(c, launch_features(fa, s));
            HIPCHK(c, forward(net, *ws, n, Hp, Wp, s, ao + img * f0, W, H));
        }
        return 0;
    }

    void aonet_destroy(isaac_hip_ctx* c)
    {
        delete c->aonet;
        c->aonet = nullptr;
    }
} // namespace isaac_hip

using namespace isaac_hip;

extern "C"
{
    int64_t isaac_hip_aonet_param_count(void)
    {
        return aonet_param_count();
    }

    int isaac_hip_set_ao_mode(isaac_hip_ctx* c, int32_t mode)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(mode != ISAAC_HIP_AO_PLAIN && mode != ISAAC_HIP_AO_LEARNED)
            return fail(c, ISAAC_HIP_ERR_ARG, "unknown occlusion mode");
        if(mode == ISAAC_HIP_AO_LEARNED && !(c->aonet && c->aonet->ready))
            return fail(c, ISAAC_HIP_ERR_STATE, "learned occlusion needs network parameters (isaac_hip_aonet_set_params)");
        c->generation++;
        c->aoMode = mode;
        return 0;
    }

    int isaac_hip_aonet_set_params(isaac_hip_ctx* c, const float* params, int64_t n)
    {
        if(!c || !params || n != aonet_param_count())
            return ISAAC_HIP_ERR_ARG;
        HIPCHK(c, hipSetDevice(c->device));
        int rc = create_net(c);
        if(rc)
            return rc;
        AoNet* net = c->aonet;
        const size_t pb = size_t(n) * sizeof(float);
        HIPCHK(c, hipMemcpyAsync(net->master, params, pb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, zero(net->m, size_t(n), c->stream));
        HIPCHK(c, zero(net->v, size_t(n), c->stream));
        HIPCHK(c, pack(net, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        net->step = 0;
        net->ready = true;
        return 0;
    }

    int isaac_hip_aonet_get_params(isaac_hip_ctx* c, float* params, int64_t n)
    {
        if(!c || !params || n != aonet_param_count())
            return ISAAC_HIP_ERR_ARG;
        if(!c->aonet || !c->aonet->ready)
            return fail(c, ISAAC_HIP_ERR_STATE, "occlusion network has no parameters");
        HIPCHK(c, hipMemcpyAsync(params, c->aonet->master, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }

    int isaac_hip_aonet_clear_batch(isaac_hip_ctx* c)
    {
        if(!c)
            return ISAAC_HIP_ERR_ARG;
        if(c->aonet)
            c->aonet->batchCount = 0;
        return 0;
    }

    int isaac_hip_aonet_add_crops(isaac_hip_ctx* c, int32_t n, const int32_t* xy, int32_t side)
    {
        if(!c || n <= 0 || !xy || side <= 0 || side % 8)
            return ISAAC_HIP_ERR_ARG;
        if(!c->aonet || !c->aonet->ready)
            return fail(c, ISAAC_HIP_ERR_STATE, "occlusion network has no parameters");
        AoNet* net = c->aonet;
        if(net->batchCount > 0 && side != net->batchSide)
            return fail(c, ISAAC_HIP_ERR_ARG, "all crops of a batch have one size");
        if(net->batchCount + n > kMaxBatch)
            return fail(c, ISAAC_HIP_ERR_ARG, "training batch holds at most 64 crops");
        const int W = c->cfg.framebuffer_size[0], H = c->cfg.framebuffer_size[1];
        for(int i = 0; i < n; i++)
            if(xy[2 * i] < 0 || xy[2 * i + 1] < 0 || xy[2 * i] + side > W || xy[2 * i + 1] + side > H)
                return fail(c, ISAAC_HIP_ERR_ARG, "crop outside the framebuffer");
        HIPCHK(c, net->train.ensure(kMaxBatch, side, side, true));
        net->batchSide = side;
        const size_t per = size_t(side) * side;
        HIPCHK(c, hipMemcpyAsync(net->train.origins, xy, size_t(2) * n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        FeatureArgs f{};
        f.depth = c->depth;
        f.normal = c->normal;
        f.ao = c->ao;
        f.imgW = W;
        f.imgH = H;
        f.invScale = 1.0f / float(c->maxGsScaled);
        f.origin = net->train.origins;
        f.x = net->train.x0 + per * kFeatures * net->batchCount;
        f.target = net->train.target + per * net->batchCount;
        f.n = n;
        f.h = side;
        f.w = side;
        HIPCHK(c, launch_features(f, c->stream));
        // the origins buffer is reused by the next call
        HIPCHK(c, hipStreamSynchronize(c->stream));
        net->batchCount += n;
        return 0;
    }

    int isaac_hip_aonet_train_step(isaac_hip_ctx* c, float lr, float* loss)
    {
        if(!c || !(lr > 0.0f))
            return ISAAC_HIP_ERR_ARG;
        if(!c->aonet || !c->aonet->ready || c->aonet->batchCount == 0)
            return fail(c, ISAAC_HIP_ERR_STATE, "empty training batch");
        AoNet* net = c->aonet;
        const int n = net->batchCount, s = net->batchSide;
        HIPCHK(c, forward(net, net->train, n, s, s, c->stream));
        HIPCHK(c, backward(net, net->train, n, s, s, c->stream));
        net->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, float(net->step)), c2 = 1.0f - std::pow(b2, float(net->step));
        const float gscale = 1.0f / float(double(n) * s * s);
        HIPCHK(c, launch_adam(net->master, net->grad, net->m, net->v, net->nparams, lr, b1, b2, 1e-8f, c1, c2, gscale, c->stream));
        HIPCHK(c, pack(net, c->stream));
        float l = 0.0f;
        HIPCHK(c, hipMemcpyAsync(&l, net->train.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if(loss)
            *loss = l * gscale;
        return 0;
    }

    int isaac_hip_aonet_loss_grad(
        isaac_hip_ctx* c,
        int32_t n,
        int32_t h,
        int32_t w,
        const float* features,
        const float* target,
        float* loss,
        float* grad)
    {
        if(!c || !features || !target)
            return ISAAC_HIP_ERR_ARG;
        int rc = check_dims(c, n, h, w);
        if(rc)
            return rc;
        AoNet* net = c->aonet;
        HIPCHK(c, net->train.ensure(std::max(n, kMaxBatch), h, w, true));
        net->batchCount = 0; // the batch buffers are overwritten
        rc = upload_features(c, net->train, n, h, w, features, target);
        if(rc)
            return rc;
        HIPCHK(c, forward(net, net->train, n, h, w, c->stream));
        HIPCHK(c, backward(net, net->train, n, h, w, c->stream));
        float l = 0.0f;
        HIPCHK(c, hipMemcpyAsync(&l, net->train.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if(grad)
            HIPCHK(c, hipMemcpyAsync(grad, net->grad, size_t(net->nparams) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if(loss)
            *loss = l;
        return 0;
    }

    int isaac_hip_aonet_forward(isaac_hip_ctx* c, int32_t n, int32_t h, int32_t w, const float* features, float* out)
    {
        if(!c || !features || !out)
            return ISAAC_HIP_ERR_ARG;
        int rc = check_dims(c, n, h, w);
        if(rc)
            return rc;
        AoNet* net = c->aonet;
        HIPCHK(c, net->train.ensure(std::max(n, kMaxBatch), h, w, true));
        net->batchCount = 0;
        const size_t full = size_t(n) * h * w;
        std::vector<float> zeros(full, 0.0f);
        rc = upload_features(c, net->train, n, h, w, features, zeros.data());
        if(rc)
            return rc;
        HIPCHK(c, forward(net, net->train, n, h, w, c->stream));
        HIPCHK(c, backward(net, net->train, n, h, w, c->stream)); // the head kernel writes the raw output
        HIPCHK(c, hipMemcpyAsync(out, net->train.out, full * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }
}
