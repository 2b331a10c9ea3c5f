// Internal definition of the render context shared by the host translation
// units of the library (host.cpp: rendering and steering; aonet_host.cpp: the
// learned occlusion pass).  Not part of the C-ABI.
#pragma once

#include "isaac_hip.h"
#include "kernels.hpp"

#include <string>
#include <vector>

namespace isaac_hip
{
    struct AoNet; // learned occlusion network (aonet_host.cpp)
    struct AoWorkspace; // activations of one set of network evaluations (aonet_host.cpp)

    struct VolSlot
    {
        isaac_hip_volume_desc desc{};
        bool set = false;
        float* npTex = nullptr; // non-persistent functor-applied texture (with guard)
        size_t npBytes = 0;
        void* brick = nullptr; // brick form: pairs
        float* brickScale = nullptr;
        size_t brickBytes = 0, scaleCount = 0;
        int64_t nbricks = 0;
        int brickDim = 0;
        bool bricked = false; // bricks hold the source as of the last update_buffers
        uint8_t* adv[2] = {nullptr, nullptr}; // field source: advection textures (current, history)
        int advCur = 0;
        size_t advVox = 0;
    };

    struct PartSlot
    {
        isaac_hip_particle_desc desc{};
        bool set = false;
        int64_t capN = 0, capCells = 0;
        BinBuffers B{};
        void* blocks[14] = {};
        // the frame builds only the visible particles' records; the full cell list of
        // all particles is built on request (isaac_hip_get_cell_list) with the inputs
        // of the last update_buffers
        BinParams last{};
        bool lastValid = false, fullValid = false;
    };
} // namespace isaac_hip

struct isaac_hip_ctx
{
    using VolSlot = isaac_hip::VolSlot;
    using PartSlot = isaac_hip::PartSlot;
    isaac_hip_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    // video frame encoding (isaac_hip_video_frame_message): device workspace, lazily sized
    void* videoWs = nullptr;
    int64_t videoWsBytes = 0;
    bool rendered = false; // a frame is in the colour buffer
    // bumped whenever a device buffer that a frame's parameter blocks point to is
    // (re)allocated or freed, the stream changes, or steering state changes: a batch
    // (batch_host.cpp) that snapshotted this context's frames refuses to render
    // until its views are set again
    uint64_t generation = 0;
    std::string err;
    int nv = 0, nf = 0, np = 0, nsrc = 0;
    std::vector<VolSlot> vols;
    std::vector<PartSlot> parts;
    std::vector<isaac_hip::ChainDesc> chain;
    std::vector<int32_t> chainErr;
    std::vector<std::vector<float>> lutHost;
    float4* lutDev = nullptr;
    bool lutDirty = true;
    std::vector<float> weight, iso;
    float step = 0.5f;
    bool interpolation = false;
    int ditherMode = 1;
    bool aoEnabled = false;
    float aoWeight = 0.5f;
    int renderMode = 0;
    int nclip = 0;
    float clipPos[isaac_hip::kMaxClipping][3];
    float clipNormal[isaac_hip::kMaxClipping][3];
    float clipSaved[isaac_hip::kMaxClipping][3];
    // sizes
    int64_t global[3], local[3], pgrid[3];
    int32_t position[3];
    float scale[3];
    int64_t gsScaled[3], lsScaled[3], maxGsScaled;
    int32_t posScaled[3];
    // device buffers
    uint2* tex = nullptr; // per voxel: combined volume RGBA8, combined iso RGBA8
    size_t texVox = 0;
    uint2* texBricks = nullptr; // the same texels in 8^3 bricks with a +1 apron (texture layout 8)
    uint2* texH = nullptr; // the volume RGBA as binary16 per texel (texture layout 16), texVox + 1 texels
    size_t texBrickCount = 0;
    uint32_t* color = nullptr;
    float* depth = nullptr;
    float* normal = nullptr;
    float* ao = nullptr;
    float4* volColor = nullptr;
    unsigned long long* counters = nullptr;
    bool counting = false;
    bool classified = false;
    int tunePrefetch = 0;
    int tuneWaves = 0;
    int tuneTiles = 4; // super-tiles of 4x4 tiles (measured fastest at 1080p)
    int tuneTexLayout = 0; // march texture: 0 linear (x fastest), 8 bricks of 8^3 texels, 16 linear binary16 copy
    int tuneMarchThreads = 256; // combined-form march workgroup: 256 (4 waves, one tile) or 64 (one wave)
    int tuneMarchOffsets = 32; // combined-form march loads: 32-bit byte offsets where the texture allows, or 64-bit addresses
    int builtTexLayout = -1; // layout whose texels the last combined-form update_buffers wrote
    int form = ISAAC_HIP_FORM_COMBINED;
    int brickDtype = ISAAC_HIP_F32;
    int builtForm = -1; // form prepared by the last update_buffers
    bool floatOut = false;
    float4* outFloat = nullptr;
    int aoMode = ISAAC_HIP_AO_PLAIN;
    isaac_hip::AoNet* aonet = nullptr;
    // field sources: advection state (lib/isaac.hpp:606-644 defaults)
    float* noise = nullptr;
    float* noiseTmp = nullptr;
    size_t noiseVox = 0;
    bool noiseDirty = true;
    int32_t seedPoints = 1000;
    float advStepFactor = 1.0f, advHistoryWeight = 0.95f;
    int32_t advSeedPeriod = 1, advSeedTime = 1;
    int32_t timeStep = 0;
    // min/max request: per-workgroup partials and one (min, max) per source
    isaac_hip::MinMaxPartial* mmPartials = nullptr;
    float2* mmResult = nullptr;
    // events: 0 update start, 1 after source copies, 2 after classification (or brick
    // build), 3 after the particle records; 4 draw start, 5 after the march, 6 after the
    // occlusion pass, 7 after shading, 8 after the framebuffer copy
    hipEvent_t ev[12] = {};
    bool evRecorded[2] = {false, false};
    bool tuneFusedShading = true; // the occlusion kernel shades its pixels where the render mode allows
    int timingEvents = 1; // timing markers between the frame's kernels: 1 all, 2 around the march only, 0 none
    bool evMarchOnly = false; // the last render recorded the march's markers only
};

namespace isaac_hip
{
    inline int fail(isaac_hip_ctx* c, int code, const std::string& msg)
    {
        if(c)
            c->err = msg;
        return code;
    }

    // G-buffer targets of one frame
    struct FrameTargets
    {
        int W, H;
        uint32_t* color;
        float* depth;
        float* normal;
        float* ao;
        float4* volColor;
        float4* outFloat;
        unsigned long long* counters;
    };
    // kernel parameter blocks of one frame (host.cpp)
    int build_frame(
        isaac_hip_ctx* c,
        const double proj[16],
        const double mv[16],
        const float bg[4],
        const int32_t viewport[4],
        const FrameTargets& t,
        FrameParams& F,
        PostParams& Q,
        ParticleParams& PP);

    // learned occlusion pass (aonet_host.cpp)
    int aonet_occlusion(isaac_hip_ctx* c);
    // a batch's own activation workspace, evaluated with context c's network
    AoWorkspace* aonet_workspace_create();
    void aonet_workspace_destroy(AoWorkspace* ws);
    // sizes ws for `chunk` frames of W x H; *reallocated tells whether its buffers moved
    int aonet_reserve_frames(isaac_hip_ctx* c, AoWorkspace* ws, int chunk, int W, int H, bool* reallocated);
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
        hipStream_t s);
    void aonet_destroy(isaac_hip_ctx* c);
    int64_t aonet_param_count();

    inline int hip_fail(isaac_hip_ctx* c, hipError_t e, const char* what)
    {
        return fail(c, ISAAC_HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }

#define HIPCHK(ctx, call)                                                                                             \
    do                                                                                                                \
    {                                                                                                                 \
        hipError_t e_ = (call);                                                                                       \
        if(e_ != hipSuccess)                                                                                          \
            return hip_fail(ctx, e_, #call);                                                                          \
    } while(0)
} // namespace isaac_hip
