// Batched form: many sub-volumes (one render context each, brick form) x many
// camera views rendered by one persistent march launch, one occlusion pass over
// all frames and one shading launch, optionally captured in one hipGraph and
// replayed.  No counterpart in the reference, which renders one frame of one
// sub-volume per drawCallBack (lib/isaac.hpp:2140-2543) and keeps one live
// instance per template instantiation (`static myself`, lib/isaac.hpp:1307,3029).
#include "context.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using namespace isaac_hip;

struct isaac_hip_batch
{
    int device = 0;
    // the batch's own stream, ordered after each context's stream at every render
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<isaac_hip_ctx*> ctx;
    std::vector<hipEvent_t> ctxEv; // one per context: its stream's position at the render
    std::vector<uint64_t> gen; // each context's generation when the views were set
    isaac_hip::AoWorkspace* aoWs = nullptr; // learned occlusion activations of this batch
    int W = 0, H = 0, nviews = 0, nframes = 0;
    int tilesPerFrame = 0;
    int chunk = 4, groupsPerCU = 4, aoChunk = 16;
    int numCU = 256;
    // host copies of the per-frame parameter blocks
    std::vector<FrameParams> hF;
    std::vector<PostParams> hQ;
    std::vector<ParticleParams> hP;
    std::vector<float> hInvScale;
    // device
    FrameParams* dF = nullptr;
    PostParams* dQ = nullptr;
    ParticleParams* dP = nullptr;
    uint32_t* dTileStart = nullptr;
    uint32_t* dCounter = nullptr;
    float* dInvScale = nullptr;
    uint32_t* color = nullptr;
    float* depth = nullptr;
    float* normal = nullptr;
    float* ao = nullptr;
    float4* volColor = nullptr;
    size_t capFrames = 0;
    bool learned = false;
    bool aoOn = false;
    int dtype = 0;
    bool linear = true;
    int sig = 0; // source signature of every frame (0: mixed)
    size_t lutLds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graphStale = true;
    hipEvent_t ev[2] = {};
    bool timed = false;
};

namespace
{
    int bfail(isaac_hip_batch* b, int code, const std::string& msg)
    {
        if(b)
            b->err = msg;
        return code;
    }

#define BCHK(b, call)                                                                                                 \
    do                                                                                                                \
    {                                                                                                                 \
        hipError_t e_ = (call);                                                                                       \
        if(e_ != hipSuccess)                                                                                          \
            return bfail(b, ISAAC_HIP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while(0)

    void free_frames(isaac_hip_batch* b)
    {
        void* bufs[] = {b->dF, b->dQ, b->dP, b->dTileStart, b->dInvScale, b->color, b->depth, b->normal, b->ao, b->volColor};
        for(void* p : bufs)
            if(p)
                (void) hipFree(p);
        b->dF = nullptr;
        b->dQ = nullptr;
        b->dP = nullptr;
        b->dTileStart = nullptr;
        b->dInvScale = nullptr;
        b->color = nullptr;
        b->depth = nullptr;
        b->normal = nullptr;
        b->ao = nullptr;
        b->volColor = nullptr;
        b->capFrames = 0;
    }

    void drop_graph(isaac_hip_batch* b)
    {
        if(b->exec)
            (void) hipGraphExecDestroy(b->exec);
        if(b->graph)
            (void) hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->graphStale = true;
    }

    int ensure_frames(isaac_hip_batch* b, int nframes)
    {
        if(size_t(nframes) <= b->capFrames)
            return 0;
        drop_graph(b);
        free_frames(b);
        const size_t px = size_t(b->W) * b->H * nframes;
        BCHK(b, hipMalloc(&b->dF, sizeof(FrameParams) * nframes));
        BCHK(b, hipMalloc(&b->dQ, sizeof(PostParams) * nframes));
        BCHK(b, hipMalloc(&b->dP, sizeof(ParticleParams) * nframes));
        BCHK(b, hipMalloc(&b->dTileStart, sizeof(uint32_t) * (nframes + 1)));
        BCHK(b, hipMalloc(&b->dInvScale, sizeof(float) * nframes));
        BCHK(b, hipMalloc(&b->color, px * 4));
        BCHK(b, hipMalloc(&b->depth, px * 4));
        BCHK(b, hipMalloc(&b->normal, px * 12));
        BCHK(b, hipMalloc(&b->ao, px * 4));
        BCHK(b, hipMalloc(&b->volColor, px * 16));
        BCHK(b, launch_fill_u32(reinterpret_cast<uint32_t*>(b->ao), 0u, px, b->stream));
        b->capFrames = size_t(nframes);
        return 0;
    }

    // the sequence of one batched frame set: persistent march, occlusion, shading
    int enqueue(isaac_hip_batch* b)
    {
        BCHK(b, launch_fill_u32(b->dCounter, 0u, 1, b->stream));
        BatchParams B{};
        B.frames = b->dF;
        B.posts = b->dQ;
        B.parts = b->dP;
        B.tileStart = b->dTileStart;
        B.nframes = b->nframes;
        B.ntiles = uint32_t(b->tilesPerFrame) * uint32_t(b->nframes);
        B.chunk = b->chunk;
        B.counter = b->dCounter;
        const int grid = std::max(1, b->numCU * b->groupsPerCU);
        BCHK(b, launch_batch_march(B, b->dtype, b->linear, b->sig, b->lutLds, grid, b->stream));
        if(b->aoOn)
        {
            if(b->learned)
            {
                int rc = aonet_occlusion_frames(b->ctx[0], b->aoWs, b->depth, b->normal, b->ao, b->nframes, b->W, b->H,
                                                b->dInvScale, b->aoChunk, b->stream);
                if(rc)
                    return bfail(b, rc, b->ctx[0]->err);
            }
            else
                BCHK(b, launch_batch_ssao(b->dF, b->nframes, b->tilesPerFrame, b->stream));
        }
        BCHK(b, launch_batch_shade(b->dF, b->nframes, b->tilesPerFrame, b->stream));
        return 0;
    }
} // namespace

extern "C"
{
    int isaac_hip_batch_create(isaac_hip_ctx* const* ctxs, int32_t n, int32_t width, int32_t height, isaac_hip_batch** out)
    {
        if(!ctxs || n <= 0 || width <= 0 || height <= 0 || !out)
            return ISAAC_HIP_ERR_ARG;
        *out = nullptr;
        for(int i = 0; i < n; i++)
            if(!ctxs[i])
                return ISAAC_HIP_ERR_ARG;
        isaac_hip_batch* b = new isaac_hip_batch();
        b->ctx.assign(ctxs, ctxs + n);
        b->device = ctxs[0]->device;
        b->W = width;
        b->H = height;
        b->tilesPerFrame = ((width + 15) / 16) * ((height + 15) / 16);
        b->ctxEv.assign(size_t(n), nullptr);
        b->gen.assign(size_t(n), 0);
        bool ok = hipSetDevice(b->device) == hipSuccess && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess
                  && hipMalloc(&b->dCounter, 64) == hipSuccess && hipEventCreate(&b->ev[0]) == hipSuccess
                  && hipEventCreate(&b->ev[1]) == hipSuccess;
        for(int i = 0; ok && i < n; i++)
            ok = hipEventCreateWithFlags(&b->ctxEv[i], hipEventDisableTiming) == hipSuccess;
        if(!ok)
        {
            isaac_hip_batch_destroy(b);
            return ISAAC_HIP_ERR_DEVICE;
        }
        int cus = 0;
        if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && cus > 0)
            b->numCU = cus;
        *out = b;
        return 0;
    }

    void isaac_hip_batch_destroy(isaac_hip_batch* b)
    {
        if(!b)
            return;
        (void) hipSetDevice(b->device);
        if(b->stream)
            (void) hipStreamSynchronize(b->stream);
        drop_graph(b);
        free_frames(b);
        aonet_workspace_destroy(b->aoWs);
        if(b->dCounter)
            (void) hipFree(b->dCounter);
        for(hipEvent_t e : b->ev)
            if(e)
                (void) hipEventDestroy(e);
        for(hipEvent_t e : b->ctxEv)
            if(e)
                (void) hipEventDestroy(e);
        if(b->stream)
            (void) hipStreamDestroy(b->stream);
        delete b;
    }

    const char* isaac_hip_batch_last_error(const isaac_hip_batch* b)
    {
        return b ? b->err.c_str() : "null batch";
    }

    int isaac_hip_batch_set_tuning(isaac_hip_batch* b, int32_t key, int32_t value)
    {
        if(!b)
            return ISAAC_HIP_ERR_ARG;
        switch(key)
        {
        case ISAAC_HIP_BATCH_TUNE_CHUNK:
            if(value < 1 || value > 64)
                return bfail(b, ISAAC_HIP_ERR_ARG, "chunk: 1..64 tiles");
            b->chunk = value;
            break;
        case ISAAC_HIP_BATCH_TUNE_GROUPS_PER_CU:
            if(value < 1 || value > 32)
                return bfail(b, ISAAC_HIP_ERR_ARG, "groups per CU: 1..32");
            b->groupsPerCU = value;
            break;
        case ISAAC_HIP_BATCH_TUNE_AO_CHUNK:
            if(value < 1 || value > 256)
                return bfail(b, ISAAC_HIP_ERR_ARG, "occlusion chunk: 1..256 frames");
            b->aoChunk = value;
            break;
        default:
            return bfail(b, ISAAC_HIP_ERR_ARG, "unknown batch tuning key");
        }
        b->graphStale = true;
        return 0;
    }

    int isaac_hip_batch_set_views(isaac_hip_batch* b, int32_t nviews, const double* proj, const double* mv, const float bg[4])
    {
        if(!b || nviews <= 0 || !proj || !mv)
            return ISAAC_HIP_ERR_ARG;
        const int n = int(b->ctx.size());
        const int nframes = n * nviews;
        isaac_hip_ctx* c0 = b->ctx[0];
        // every context: brick form, prepared, same device, same steering (shared transfer functions)
        for(int i = 0; i < n; i++)
        {
            isaac_hip_ctx* c = b->ctx[i];
            if(c->device != b->device)
                return bfail(b, ISAAC_HIP_ERR_ARG, "contexts on different devices");
            if(c->form != ISAAC_HIP_FORM_BRICKS || c->builtForm != ISAAC_HIP_FORM_BRICKS || !c->classified)
                return bfail(b, ISAAC_HIP_ERR_STATE, "context " + std::to_string(i) + ": brick form with update_buffers done");
            if(c->brickDtype != c0->brickDtype || c->interpolation != c0->interpolation || c->nsrc != c0->nsrc
               || c->cfg.transfer_size != c0->cfg.transfer_size || c->aoEnabled != c0->aoEnabled)
                return bfail(b, ISAAC_HIP_ERR_ARG, "context " + std::to_string(i) + ": render state differs from context 0");
            for(int s = 0; s < c->nsrc; s++)
                if(c->weight[s] != c0->weight[s] || c->iso[s] != c0->iso[s] || c->lutHost[s] != c0->lutHost[s]
                   || std::memcmp(&c->chain[s], &c0->chain[s], sizeof(ChainDesc)) != 0)
                    return bfail(b, ISAAC_HIP_ERR_ARG, "context " + std::to_string(i) + ": steering differs from context 0");
        }
        HIPCHK(c0, hipSetDevice(b->device));
        // the contexts' bricks are built on their own streams
        for(int i = 0; i < n; i++)
            BCHK(b, hipStreamSynchronize(b->ctx[i]->stream));
        int rc = ensure_frames(b, nframes);
        if(rc)
            return rc;
        b->hF.resize(nframes);
        b->hQ.resize(nframes);
        b->hP.resize(nframes);
        b->hInvScale.resize(nframes);
        const size_t img = size_t(b->W) * b->H;
        for(int f = 0; f < nframes; f++)
        {
            isaac_hip_ctx* c = b->ctx[f / nviews];
            FrameTargets t{b->W,
                           b->H,
                           b->color + img * f,
                           b->depth + img * f,
                           b->normal + img * 3 * f,
                           b->ao + img * f,
                           b->volColor + img * f,
                           nullptr,
                           c->counters};
            const int32_t vp[4] = {0, 0, b->W, b->H};
            rc = build_frame(c, proj + 16 * size_t(f), mv + 16 * size_t(f), bg, vp, t, b->hF[f], b->hQ[f], b->hP[f]);
            if(rc)
                return bfail(b, rc, c->err);
            b->hF[f].waves = 0;
            b->hF[f].prefetch = 0;
            b->hF[f].waveGroups = 4;
            b->hInvScale[f] = 1.0f / float(c->maxGsScaled);
            // the persistent march stages frame 0's transfer functions for every frame:
            // every frame must bind the same sources in the same slots
            const PostParams& q0 = b->hQ[0];
            const PostParams& q = b->hQ[f];
            bool same = q.nsrc == q0.nsrc && q.dtype == q0.dtype && q.T == q0.T;
            for(int s = 0; same && s < q.nsrc; s++)
                same = q.src[s].active == q0.src[s].active && q.src[s].dim == q0.src[s].dim
                       && q.src[s].lutSlot == q0.src[s].lutSlot;
            if(!same)
                return bfail(b, ISAAC_HIP_ERR_ARG,
                             "context " + std::to_string(f / nviews) + ": bound sources differ from context 0's");
        }
        std::vector<uint32_t> ts(nframes + 1);
        for(int f = 0; f <= nframes; f++)
            ts[f] = uint32_t(f) * uint32_t(b->tilesPerFrame);
        BCHK(b, hipMemcpyAsync(b->dF, b->hF.data(), sizeof(FrameParams) * nframes, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->dQ, b->hQ.data(), sizeof(PostParams) * nframes, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->dP, b->hP.data(), sizeof(ParticleParams) * nframes, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->dTileStart, ts.data(), sizeof(uint32_t) * (nframes + 1), hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->dInvScale, b->hInvScale.data(), sizeof(float) * nframes, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
        if(nframes != b->nframes || nviews != b->nviews)
            b->graphStale = true;
        b->nframes = nframes;
        b->nviews = nviews;
        // the source signature selects a march compiled for it when every frame shares it
        int sig = post_signature(b->hQ[0]);
        bool natural = true;
        for(int f = 0; f < nframes; f++)
        {
            if(post_signature(b->hQ[f]) != sig)
                sig = 0;
            natural = natural && natural_chains(b->hQ[f]);
        }
        if(sig > 0 && natural)
            sig += kSigNatural;
        const size_t lutLds = size_t(b->hQ[0].nsrc) * size_t(b->hQ[0].T) * sizeof(float4);
        // launch arguments the captured graph holds
        if(c0->brickDtype != b->dtype || c0->interpolation != b->linear || sig != b->sig || lutLds != b->lutLds)
            b->graphStale = true;
        b->dtype = c0->brickDtype;
        b->linear = c0->interpolation;
        b->sig = sig;
        b->lutLds = lutLds;
        const bool aoOn = b->hF[0].aoEnabled != 0;
        const bool learned = aoOn && c0->aoMode == ISAAC_HIP_AO_LEARNED;
        if(aoOn != b->aoOn || learned != b->learned)
            b->graphStale = true;
        b->aoOn = aoOn;
        b->learned = learned;
        if(learned)
        {
            if(!b->aoWs)
                b->aoWs = aonet_workspace_create();
            bool moved = false;
            rc = aonet_reserve_frames(c0, b->aoWs, b->aoChunk, b->W, b->H, &moved);
            if(rc)
                return bfail(b, rc, c0->err);
            if(moved)
                drop_graph(b);
        }
        for(int i = 0; i < n; i++)
            b->gen[i] = b->ctx[i]->generation;
        return 0;
    }

    int isaac_hip_batch_render(isaac_hip_batch* b, int32_t mode)
    {
        if(!b || b->nframes == 0)
            return b ? bfail(b, ISAAC_HIP_ERR_STATE, "no views set (isaac_hip_batch_set_views)") : ISAAC_HIP_ERR_ARG;
        BCHK(b, hipSetDevice(b->device));
        // the frames' parameter blocks point into the contexts' buffers as they were
        // when the views were set: refuse if any was reallocated or re-steered since
        for(size_t i = 0; i < b->ctx.size(); i++)
            if(b->ctx[i]->generation != b->gen[i])
                return bfail(b, ISAAC_HIP_ERR_STATE,
                             "context " + std::to_string(i)
                                 + " changed since isaac_hip_batch_set_views (buffers reallocated, stream or steering changed)");
        if(b->learned)
        {
            bool moved = false;
            int rc = aonet_reserve_frames(b->ctx[0], b->aoWs, b->aoChunk, b->W, b->H, &moved);
            if(rc)
                return bfail(b, rc, b->ctx[0]->err);
            if(moved)
                drop_graph(b);
        }
        // after everything already queued on the contexts' streams (brick and cell-list
        // builds of update_buffers, network training on context 0)
        for(size_t i = 0; i < b->ctx.size(); i++)
        {
            BCHK(b, hipEventRecord(b->ctxEv[i], b->ctx[i]->stream));
            BCHK(b, hipStreamWaitEvent(b->stream, b->ctxEv[i], 0));
        }
        BCHK(b, hipEventRecord(b->ev[0], b->stream));
        if(mode == ISAAC_HIP_BATCH_PERSISTENT)
        {
            int rc = enqueue(b);
            if(rc)
                return rc;
        }
        else if(mode == ISAAC_HIP_BATCH_GRAPH)
        {
            if(b->graphStale || !b->exec)
            {
                drop_graph(b);
                BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = enqueue(b);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->stream, &g);
                if(rc)
                {
                    if(g)
                        (void) hipGraphDestroy(g);
                    return rc;
                }
                if(e != hipSuccess)
                    return bfail(b, ISAAC_HIP_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
                b->graph = g;
                BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->graphStale = false;
            }
            BCHK(b, hipGraphLaunch(b->exec, b->stream));
        }
        else if(mode == ISAAC_HIP_BATCH_PER_FRAME)
        {
            // the same frames with one launch sequence per frame (the unbatched way)
            for(int f = 0; f < b->nframes; f++)
            {
                const FrameParams& F = b->hF[f];
                BCHK(b, launch_brick_march(F, b->hQ[f], b->hP[f], false, b->stream));
                if(b->aoOn)
                {
                    if(b->learned)
                    {
                        const size_t img = size_t(b->W) * b->H;
                        int rc = aonet_occlusion_frames(b->ctx[0], b->aoWs, b->depth + img * f, b->normal + img * 3 * f,
                                                        b->ao + img * f, 1, b->W, b->H, b->dInvScale + f, 1, b->stream);
                        if(rc)
                            return bfail(b, rc, b->ctx[0]->err);
                    }
                    else
                        BCHK(b, launch_ssao(F, b->stream));
                }
                BCHK(b, launch_shade(F, b->stream));
            }
        }
        else
            return bfail(b, ISAAC_HIP_ERR_ARG, "unknown batch mode");
        BCHK(b, hipEventRecord(b->ev[1], b->stream));
        // and the contexts' later work (a rebuild of their bricks or cell lists) after the
        // frames that read their buffers
        for(size_t i = 0; i < b->ctx.size(); i++)
            BCHK(b, hipStreamWaitEvent(b->ctx[i]->stream, b->ev[1], 0));
        b->timed = true;
        return 0;
    }

    int isaac_hip_batch_synchronize(isaac_hip_batch* b)
    {
        if(!b)
            return ISAAC_HIP_ERR_ARG;
        BCHK(b, hipStreamSynchronize(b->stream));
        return 0;
    }

    int isaac_hip_batch_last_time_us(isaac_hip_batch* b, double* us)
    {
        if(!b || !us)
            return ISAAC_HIP_ERR_ARG;
        if(!b->timed)
            return bfail(b, ISAAC_HIP_ERR_STATE, "nothing rendered yet");
        BCHK(b, hipEventSynchronize(b->ev[1]));
        float ms = 0.0f;
        BCHK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        *us = double(ms) * 1e3;
        return 0;
    }

    int isaac_hip_batch_get_images(const isaac_hip_batch* b, const void** rgba, int32_t* nframes)
    {
        if(!b)
            return ISAAC_HIP_ERR_ARG;
        if(rgba)
            *rgba = b->color;
        if(nframes)
            *nframes = b->nframes;
        return 0;
    }
}
