/* isaac_hip.h — C-ABI of the MI355X (gfx950) in-situ ray-casting renderer.
 *
 * One context = one rank's IsaacVisualization render state on one GPU: the
 * per-voxel classification pass, the particle/iso/SSAO/shading/volume passes
 * of the IceT draw callback, and the steering setters that feed them.  There
 * is no IceT, MPI, JSON, JPEG or TCP behind this interface: the caller (the
 * C++ shim include/isaac_mi355x.hpp, Python through ctypes, or a cgo/JNI
 * binding) owns those.
 *
 * Conventions
 *   - every function returns 0 on success and a negative ISAAC_HIP_ERR_* on
 *     failure; isaac_hip_last_error(ctx) then holds a message.  Setters that
 *     the reference exposes as bool return 1/0 in that position instead.
 *   - device pointers are plain `const void*` / `const float*`; streams are
 *     passed as `void*` (a hipStream_t).  No torch or HIP types appear here.
 *   - calls on one context are not thread safe; several contexts may coexist
 *     (the reference's `static myself` limit, lib/isaac.hpp:1307,3029, does
 *     not apply).
 *   - source indices: volume sources 0..nv-1, then field sources, then
 *     particle sources (the reference's combined source list order,
 *     lib/isaac.hpp:720-731).
 */
#ifndef ISAAC_HIP_H
#define ISAAC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C"
{
#endif

#define ISAAC_HIP_MAX_SOURCES 16 /* ISAAC_MAX_SOURCES   lib/isaac/isaac_defines.hpp:36-38 */
#define ISAAC_HIP_MAX_FUNCTORS 3 /* ISAAC_MAX_FUNCTORS  lib/isaac/isaac_defines.hpp:27-29 */
#define ISAAC_HIP_MAX_CLIPPING 16 /* ISAAC_MAX_CLIPPING  lib/isaac/isaac_defines.hpp:44-46 */

    enum isaac_hip_status
    {
        ISAAC_HIP_OK = 0,
        ISAAC_HIP_ERR_ARG = -1,
        ISAAC_HIP_ERR_DEVICE = -2,
        ISAAC_HIP_ERR_STATE = -3,
        ISAAC_HIP_ERR_UNSUPPORTED = -4,
        ISAAC_HIP_ERR_ALLOC = -5
    };

    enum isaac_hip_dtype
    {
        ISAAC_HIP_F32 = 0,
        ISAAC_HIP_F16 = 1,
        ISAAC_HIP_BF16 = 2
    };

    /* counters filled when counting is enabled (isaac_hip_set_counting) */
    enum isaac_hip_counter
    {
        ISAAC_HIP_CNT_VOL_SAMPLES = 0,
        ISAAC_HIP_CNT_ISO_SAMPLES = 1,
        ISAAC_HIP_CNT_ISO_HITS = 2,
        ISAAC_HIP_CNT_CELLS_VISITED = 3,
        ISAAC_HIP_CNT_SPHERE_TESTS = 4,
        ISAAC_HIP_CNT_VOL_RAYS = 5,
        ISAAC_HIP_CNT_ISO_RAYS = 6,
        ISAAC_HIP_CNT_PART_RAYS = 7,
        ISAAC_HIP_CNT_MARCH_STEPS = 8, /* fused iso+volume traversal steps (one gather of the taps each) */
        ISAAC_HIP_CNT_MARCH_WAVE_STEPS = 9, /* sum over waves of their longest lane's traversal steps:
                                               MARCH_STEPS / (64 * this) is the march's lane utilisation */
        ISAAC_HIP_CNT_COUNT = 10
    };

    typedef struct isaac_hip_ctx isaac_hip_ctx;

    /* Replaces the IsaacVisualization constructor's size arguments
     * (lib/isaac.hpp:565-583) and the source-list template parameters
     * (lib/isaac.hpp:58-69). */
    typedef struct isaac_hip_config
    {
        int32_t framebuffer_size[2]; /* framebufferSize (W, H) */
        int64_t global_size[3]; /* globalSize */
        int64_t local_size[3]; /* localSize */
        int64_t local_particle_size[3]; /* localParticleSize (particle cell grid) */
        int32_t position[3]; /* position of the local box in the global box */
        float scale[3]; /* scale */
        int32_t transfer_size; /* T_transferSize */
        int32_t n_volume_sources; /* size of T_VolumeSourceList */
        int32_t n_field_sources; /* size of T_FieldSourceList: float3 vector fields whose alpha follows the
                                    advection texture (combined form) */
        int32_t n_particle_sources; /* size of T_ParticleList */
    } isaac_hip_config;

    /* A volume source as the precompiled library sees it: the reference's
     * `operator[](isaac_int3)` (example/example.cpp:39-143) reduced to a strided
     * device array.  Element (x,y,z) lives at data + (x*sx + y*sy + z*sz)*feature_dim
     * components; coordinates -guard_size .. local+guard_size-1 must be readable. */
    typedef struct isaac_hip_volume_desc
    {
        const void* data; /* device pointer at local voxel (0,0,0) */
        int32_t dtype; /* ISAAC_HIP_F32 */
        int32_t feature_dim; /* 1..4 (featureDim) */
        int64_t stride[3]; /* element strides x, y, z */
        int32_t guard_size; /* guardSize */
        int32_t persistent; /* persistent: read at every classification; else copied
                               functor-applied into a library texture (UpdatePersistendTextureKernel) */
    } isaac_hip_volume_desc;

    /* A particle source: raw particles in local voxel coordinates; the library
     * builds the per-cell lists that the reference's getIterator(cell) yields
     * (example/example.cpp:148-228). */
    typedef struct isaac_hip_particle_desc
    {
        const float* position; /* device, count x 3, local voxel coordinates */
        const float* attribute; /* device, count x feature_dim (getAttribute) */
        int32_t feature_dim; /* 1..4 */
        const float* radius; /* device, count radii, or NULL */
        float constant_radius; /* getRadius() when radius == NULL */
        int64_t count;
    } isaac_hip_particle_desc;

    /* Per-phase wall times in microseconds (the reference's public timing
     * members, lib/isaac.hpp:2129-2137) */
    typedef struct isaac_hip_timings
    {
        double kernel_us; /* drawCallBack kernels */
        double copy_us; /* framebuffer copy */
        double buffer_us; /* source update + particle binning */
        double optimization_buffer_us; /* classification */
        /* per kernel (GPU time between events on the context stream) */
        double source_copy_us; /* non-persistent source textures */
        double binning_us; /* particle cell lists */
        double march_us; /* clear + particles + iso + volume march (one traversal) */
        double ssao_us; /* SSAO */
        double shade_us; /* shading + volume blend */
    } isaac_hip_timings;

    /* ---------------------------------------------------------------- lifetime */
    int isaac_hip_create(const isaac_hip_config* cfg, int device, isaac_hip_ctx** out);
    void isaac_hip_destroy(isaac_hip_ctx* ctx);
    const char* isaac_hip_last_error(const isaac_hip_ctx* ctx);
    int isaac_hip_set_stream(isaac_hip_ctx* ctx, void* hip_stream);

    /* ---------------------------------------------------------------- sources */
    int isaac_hip_set_volume_source(isaac_hip_ctx* ctx, int idx, const isaac_hip_volume_desc* desc);
    int isaac_hip_set_particle_source(isaac_hip_ctx* ctx, int idx, const isaac_hip_particle_desc* desc);

    /* ---------------------------------------------------------------- steering
     * (the JSON keys of doVisualization, lib/isaac.hpp:1564-1795, as setters) */
    /* updateFunctions for one source (lib/isaac.hpp:1079-1241); *error_code gets
     * functions[i].errorCode: 0 ok, 1 too many functors, 2 too many params,
     * -1 missing ')', -2 bad number / unknown functor */
    int isaac_hip_set_function(isaac_hip_ctx* ctx, int idx, const char* chain, int32_t* error_code);
    /* updateTransfer from control points (lib/isaac.hpp:1244-1266) */
    int isaac_hip_set_transfer_points(isaac_hip_ctx* ctx, int idx, int n, const uint32_t* positions, const float* rgba);
    /* a ready LUT of transfer_size RGBA floats (host memory) */
    int isaac_hip_set_transfer_lut(isaac_hip_ctx* ctx, int idx, const float* lut_rgba);
    int isaac_hip_set_weight(isaac_hip_ctx* ctx, int idx, float weight);
    int isaac_hip_set_iso_threshold(isaac_hip_ctx* ctx, int idx, float threshold);
    int isaac_hip_set_step(isaac_hip_ctx* ctx, float step); /* clamped to >= 0.01 */
    int isaac_hip_set_interpolation(isaac_hip_ctx* ctx, int trilinear);
    /* Field-source advection (the JSON keys "advection step", "advection weight",
     * "advection seeding period", "advection seeding duration", lib/isaac.hpp:1643-1661;
     * defaults 1, 0.95, 1, 1).  Each isaac_hip_update_buffers swaps every field source's
     * advection texture with its history, advances the time step and, per enabled field
     * source, writes GenerateAdvectionTextureKernel (lib/isaac/isaac_common_kernel.hpp:500-550):
     * history sampled at voxel centre - unit field direction * step_factor * 3 / scale,
     * noise where time_step % seeding_period < seeding_duration, dithered to a byte.
     * The classification multiplies the field source's TF alpha by texture / 255
     * (lib/isaac/isaac_common_kernel.hpp:655-656).  Field sources with weight or iso
     * are rejected by the brick form. */
    int isaac_hip_set_advection(isaac_hip_ctx* ctx, float step_factor, float history_weight, int32_t seeding_period,
                                int32_t seeding_duration);
    /* "seed points" (lib/isaac.hpp:1636-1642): Halton points of the noise texture
     * (updateNoiseTexture, lib/isaac.hpp:499-562), clamped to >= 0, default 1000;
     * the noise is rebuilt at the next buffer update when the count changes */
    int isaac_hip_set_seed_points(isaac_hip_ctx* ctx, int32_t seed_points);
    /* applyDither3D (lib/isaac/isaac_dither_kernel.hpp:7-328) of the classification:
     * 0 round, 1 ordered 4^3, 2 blue noise 4^3, 3 blue noise 8^3, others blue noise 16^3 */
    int isaac_hip_set_dither_mode(isaac_hip_ctx* ctx, int mode);
    int isaac_hip_set_ao(isaac_hip_ctx* ctx, int enabled, float weight);
    int isaac_hip_set_render_mode(isaac_hip_ctx* ctx, int mode); /* 0..8 */
    /* clipping planes (lib/isaac.hpp:958-1017): add/edit return 1 on success, 0 as the reference's false */
    int isaac_hip_add_clipping(isaac_hip_ctx* ctx, float px, float py, float pz, float nx, float ny, float nz);
    int isaac_hip_edit_clipping(
        isaac_hip_ctx* ctx,
        uint32_t nr,
        float px,
        float py,
        float pz,
        float nx,
        float ny,
        float nz);
    int isaac_hip_remove_clipping(isaac_hip_ctx* ctx, uint32_t nr);
    int isaac_hip_clipping_count(const isaac_hip_ctx* ctx);
    /* plane nr as set: its position and its normal before scaling and normalisation
     * (clipping.elem[nr].position and clippingSavedNormals[nr], the values the
     * periodic message reports, lib/isaac.hpp:2799-2816) */
    int isaac_hip_get_clipping(const isaac_hip_ctx* ctx, uint32_t nr, float position[3], float normal[3]);
    /* updatePosition / updateLocalSize / updateLocalParticleSize (lib/isaac.hpp:1056-1076) */
    int isaac_hip_update_position(isaac_hip_ctx* ctx, const int32_t position[3]);
    int isaac_hip_update_local_size(isaac_hip_ctx* ctx, const int64_t local_size[3]);
    int isaac_hip_update_local_particle_size(isaac_hip_ctx* ctx, const int64_t local_particle_size[3]);
    /* globalSizeScaled (lib/isaac.hpp:2181-2202), the size the register message reports */
    int isaac_hip_global_size_scaled(const isaac_hip_ctx* ctx, int32_t out[3]);

    /* ---------------------------------------------------------------- per frame
     * doVisualization's buffer section (lib/isaac.hpp:1805-1948): non-persistent
     * source copies, particle cell lists, classification into the combined
     * RGBA8 volume/iso textures.  Runs on the context stream. */
    int isaac_hip_update_buffers(isaac_hip_ctx* ctx);
    /* drawCallBack (lib/isaac.hpp:2140-2543): proj/mv column-major doubles as
     * IceT passes them, bg as IceT passes it, viewport = readbackViewport
     * (x, y, w, h).  The framebuffer (RGBA8, premultiplied, row 0 = bottom) is
     * copied to out_rgba (W*H*4 bytes, host or device per out_is_device) unless
     * out_rgba is NULL.  A host copy waits for completion; otherwise the call
     * returns after enqueueing on the context stream. */
    int isaac_hip_render(
        isaac_hip_ctx* ctx,
        const double proj[16],
        const double modelview[16],
        const float bg[4],
        const int32_t viewport[4],
        uint8_t* out_rgba,
        int out_is_device);
    int isaac_hip_synchronize(isaac_hip_ctx* ctx);
    /* The min/max request (sendMinMax, lib/isaac.hpp:1951-2006, one rank): per
     * source, in the combined order (volume, field, particle), the min and max of
     * its functor-chain values over the local volume (MinMaxKernel,
     * lib/isaac/isaac_min_max_kernel.hpp:22-57; CalcMinMaxIterator,
     * lib/isaac.hpp:372-432) or over its particles (MinMaxParticleKernel,
     * lib/isaac/isaac_min_max_kernel.hpp:60-97; lib/isaac.hpp:434-497).  Persistent
     * volume sources and particle sources are read through their descriptors (the
     * pointers must be valid), non-persistent sources from the texture the last
     * isaac_hip_update_buffers filled.  Comparisons follow glm's min/max: NaN is
     * skipped, the start values are FLT_MAX / -FLT_MAX (so +inf never becomes the
     * min nor -inf the max), and ties keep the first element in the reference's
     * visiting order.  A source without values (unset, no particles, a
     * non-persistent source never updated) reports FLT_MAX / -FLT_MAX.
     * min_out / max_out: host, count >= number of sources.  Blocks until done. */
    int isaac_hip_min_max(isaac_hip_ctx* ctx, float* min_out, float* max_out, int32_t count);

    /* ---------------------------------------------------------------- tuning
     * Launch-schedule knobs with no counterpart in the reference (results do not
     * change, only speed); used by the autotuner. */
    enum isaac_hip_tuning_key
    {
        ISAAC_HIP_TUNE_MARCH_PREFETCH = 0, /* 1: fetch step i+1 before compositing step i (default 0) */
        ISAAC_HIP_TUNE_MARCH_WAVES = 1, /* march occupancy target: 0 (compiler default), 6 or 8 waves per SIMD */
        ISAAC_HIP_TUNE_MARCH_TILES = 2, /* march tile order: 0 one band of tile rows per XCD; 2, 4 (default), 8 or
                                          16: super-tiles of that many 16x16-pixel tiles a side, dealt
                                          round-robin to the XCDs */
        ISAAC_HIP_TUNE_TEX_LAYOUT = 3, /* texture the combined-form march reads: 0 (default) the linear texture
                                          (x fastest); 8: a copy in 8^3 bricks with a one-texel apron, built by
                                          update_buffers after classification; 16: the classification also
                                          writes the volume RGBA of every texel as four binary16 values, which
                                          the trilinear march steps over (the iso sums of steps that can reach
                                          the threshold, hit colours and the nearest sampler read the linear
                                          texture).  The linear texture stays what isaac_hip_get_texture
                                          returns; a change needs update_buffers before the next render;
                                          next-step prefetch applies to layouts 0 and 16 */
        ISAAC_HIP_TUNE_MARCH_THREADS = 4, /* combined-form march workgroup: 256 (default; 4 waves render one 16x16
                                            tile) or 64 (one wave per workgroup, one 8x8 quarter of a tile) */
        ISAAC_HIP_TUNE_MARCH_OFFSETS = 5, /* combined-form trilinear march loads: 32 (default) a 32-bit byte offset
                                            from the texture's base where the texture spans < 4 GiB (larger ones
                                            always take 64-bit addresses); 64: a 64-bit address per lane */
        ISAAC_HIP_TUNE_TIMING_EVENTS = 6, /* 1 (default): update_buffers and render record timing events between
                                            their kernels (isaac_hip_get_timings); 2: only the two around the
                                            march (get_timings reports march_us, zeros elsewhere); 0: none (each
                                            event costs a few microseconds on the stream; all timings zero) */
        ISAAC_HIP_TUNE_FUSED_SHADING = 7 /* 1 (default): with the plain occlusion pass in render modes 0, 1 and 3
                                            (those that keep the depth buffer), the occlusion kernel shades its
                                            own pixels (get_timings: ssao_us holds both, shade_us is 0); 0: a
                                            separate shading kernel */
    };
    int isaac_hip_set_tuning(isaac_hip_ctx* ctx, int32_t key, int32_t value);

    /* ---------------------------------------------------------------- render forms
     * No counterpart in the reference (its build picks the optimized renderer;
     * the legacy post-classification renderer, lib/isaac/isaac_volume_kernel.hpp:91-632,
     * is compiled out by default).
     *   COMBINED  the reference's optimized renderer: classification into the
     *             combined RGBA8 volume / iso textures, then the march (default).
     *   BRICKS    each volume/field source re-laid into 8^3 bricks (one-voxel apron)
     *             in brick_dtype (ISAAC_HIP_F32, F16 or BF16; 16-bit forms store
     *             v / s with one FP32 scale s = max |v| per brick and component);
     *             the march interpolates the raw field and applies the functor
     *             chain and transfer function per sample (post-classification),
     *             accumulating in FP32, with the camera inverse and ray/box/plane
     *             setup in FP64.  At most 4 volume/field sources may be enabled.
     * Takes effect at the next isaac_hip_update_buffers. */
    enum isaac_hip_form
    {
        ISAAC_HIP_FORM_COMBINED = 0,
        ISAAC_HIP_FORM_BRICKS = 1
    };
    int isaac_hip_set_render_form(isaac_hip_ctx* ctx, int32_t form, int32_t brick_dtype);
    /* device pointers of volume/field source i's bricks: *pairs = nbricks * 648 * feature_dim
     * pairs, ordered [brick][pair row][component] (brick index = (bz*nby + by)*nbx + bx,
     * x fastest; pair row (lz*9 + ly)*8 + lx holds (v[x], v[x+1]) at voxel
     * (8bx+lx, 8by+ly, 8bz+lz), coordinates clamped into the local box), each pair
     * *pair_bytes (8 or 4) bytes; *scales = nbricks * feature_dim floats.  Valid after
     * update_buffers in BRICKS form. */
    int isaac_hip_get_bricks(
        const isaac_hip_ctx* ctx,
        int source,
        const void** pairs,
        const void** scales,
        int64_t* nbricks,
        int32_t* pair_bytes);
    /* When enabled, render also writes the final colour before RGBA8 quantisation
     * (float RGBA per pixel, clamped to [0,1]) to a device buffer whose pointer
     * isaac_hip_get_float_image returns (precision studies against an FP64 render). */
    int isaac_hip_set_float_output(isaac_hip_ctx* ctx, int enabled);
    int isaac_hip_get_float_image(const isaac_hip_ctx* ctx, const void** rgba_f32);

    /* ---------------------------------------------------------------- batched form
     * No counterpart in the reference (one frame of one sub-volume per drawCallBack,
     * lib/isaac.hpp:2140-2543).  A batch renders frames = contexts x views, all at
     * width x height: one persistent march launch over every 16x16 tile of every
     * frame (workgroups pop chunks of tiles from a global counter), one occlusion
     * pass over all frames (plain SSAO, or the learned pass of context 0, evaluated
     * over several frames at a time) and one shading launch.  Each context is one
     * sub-volume in the brick form with update_buffers done, and all contexts share
     * the steering state of context 0 (transfer functions, chains, weights, iso,
     * interpolation, brick dtype, occlusion) and bind the same sources.  Contexts must
     * outlive the batch.  The batch renders on a stream of its own; each render waits
     * for the work already queued on every context's stream (update_buffers, network
     * training), so bricks rebuilt in place are read after their rebuild, and the work
     * queued on those streams afterwards waits for the render.  set_views
     * snapshots the contexts' buffers and steering: after any context reallocates a
     * buffer (update_buffers with more particles or another local size, brick dtype
     * change), changes stream or steering, render returns ISAAC_HIP_ERR_STATE until
     * set_views runs again.  The learned occlusion pass uses context 0's network with
     * activations owned by the batch, so batches of different sizes can share
     * contexts. */
    typedef struct isaac_hip_batch isaac_hip_batch;
    enum isaac_hip_batch_mode
    {
        ISAAC_HIP_BATCH_PERSISTENT = 0, /* the launch sequence, enqueued directly */
        ISAAC_HIP_BATCH_GRAPH = 1, /* the same sequence captured once in a hipGraph, then replayed */
        ISAAC_HIP_BATCH_PER_FRAME = 2 /* one march / occlusion / shading sequence per frame (unbatched reference) */
    };
    enum isaac_hip_batch_tuning_key
    {
        ISAAC_HIP_BATCH_TUNE_CHUNK = 0, /* tiles per pop of the work counter (default 4) */
        ISAAC_HIP_BATCH_TUNE_GROUPS_PER_CU = 1, /* persistent workgroups per compute unit (default 4) */
        ISAAC_HIP_BATCH_TUNE_AO_CHUNK = 2 /* frames per evaluation of the learned occlusion network (default 16) */
    };
    int isaac_hip_batch_create(isaac_hip_ctx* const* ctxs, int32_t n_ctx, int32_t width, int32_t height, isaac_hip_batch** out);
    void isaac_hip_batch_destroy(isaac_hip_batch* batch);
    const char* isaac_hip_batch_last_error(const isaac_hip_batch* batch);
    int isaac_hip_batch_set_tuning(isaac_hip_batch* batch, int32_t key, int32_t value);
    /* frame f = (context f / n_views, view f % n_views); proj and mv hold 16 column-major
     * doubles per frame (n_ctx * n_views each); bg as for isaac_hip_render */
    int isaac_hip_batch_set_views(isaac_hip_batch* batch, int32_t n_views, const double* proj, const double* mv, const float bg[4]);
    /* enqueue all frames on the stream (mode: ISAAC_HIP_BATCH_*); returns after enqueueing */
    int isaac_hip_batch_render(isaac_hip_batch* batch, int32_t mode);
    int isaac_hip_batch_synchronize(isaac_hip_batch* batch);
    /* GPU time of the last isaac_hip_batch_render (waits for it) */
    int isaac_hip_batch_last_time_us(isaac_hip_batch* batch, double* us);
    /* device pointer of the frames' RGBA8 images, frame-major (nframes * width * height * 4 bytes) */
    int isaac_hip_batch_get_images(const isaac_hip_batch* batch, const void** rgba, int32_t* nframes);

    /* ---------------------------------------------------------------- sort-last compositing
     * The per-GPU part of the reference's IceT step (lib/isaac.hpp:800-831, 2015-2056) for
     * several sub-volumes rendered on one GPU; no exchange between processes.
     *
     * isaac_hip_composite_order replaces the distance sort feeding icetCompositeOrder
     * (lib/isaac.hpp:2021-2048): for sub-volume r (position / local_size: n x 3 int32,
     * global_size, scale and the modelview as every context of the set has them) the distance
     * float(|modelview * p|) of p = (positionScaled + (localSizeScaled - globalSizeScaled) / 2)
     * / (maxSizeScaled / 2) (integer division), w = 1; order[0..n) = sub-volumes by ascending
     * distance, equal distances in index order.  Host only.
     *
     * isaac_hip_composite replaces icetDrawFrame's compositing (ICET_COMPOSITE_MODE_BLEND,
     * RGBA_UBYTE, ordered; lib/isaac.hpp:811-823): out = images[order[0]] over images[order[1]]
     * over ..., IceT's byte "over" dest = (dest * (255 - src.a)) / 255 + src per channel
     * (integer division, result taken mod 256), associated as a pairwise tree over the order
     * (binary swap's association for power-of-two counts).  images: host array of n device
     * pointers to RGBA8 images of `pixels` pixels, 16-byte aligned; out: device, 16-byte
     * aligned; 1 <= n <= 64; enqueued on `stream` (0: the null stream). */
    int isaac_hip_composite_order(int32_t n, const int32_t* position, const int32_t* local_size, const int32_t global_size[3],
                                  const float scale[3], const double modelview[16], int32_t* order);
    int isaac_hip_composite(int32_t n, const uint8_t* const* images, const int32_t* order, int64_t pixels, uint8_t* out,
                            void* stream);

    /* ---------------------------------------------------------------- frame message
     * The video frame serverSendFrame sends without libjpeg (ISAAC_JPEG off,
     * lib/isaac/isaac_communicator.hpp:217-287; the JPEG path is not built).
     *
     * isaac_hip_base64: n bytes (device, 4-byte aligned) -> ceil(8n/6) characters (device,
     * 16-byte aligned): boost base64_from_binary<transform_width<., 6, 8>> as the reference
     * uses it (isaac_communicator.hpp:251-263): standard alphabet, most significant bits
     * first, the last partial 6-bit group zero-filled, no '=' padding.  Enqueued on stream.
     *
     * isaac_hip_frame_message: msg (host) = {"payload": "data:image/raw-rgba;base64,<base64 of
     * the width*height RGBA8 device image>"} with a terminating 0; *length = its length
     * without the 0 (also set when capacity is too small, which returns ISAAC_HIP_ERR_ARG);
     * scratch: device buffer of ceil(8n/6) characters, n = 4*width*height.  Blocks until
     * the message is on the host. */
    int isaac_hip_base64(const uint8_t* in, int64_t n, char* out, void* stream);
    int isaac_hip_frame_message(const uint8_t* rgba, int32_t width, int32_t height, char* scratch, char* msg, int64_t capacity,
                                int64_t* length, void* stream);

    /* The JPEG path of serverSendFrame (the reference's default): libjpeg-turbo with
     * JCS_EXT_RGBX input, jpeg_set_defaults and jpeg_set_quality(quality, FALSE)
     * (isaac_communicator.hpp:221-250), reproduced byte for byte on the GPU: baseline
     * 4:2:0 JPEG of the width x height RGBX frame (device, 4-byte aligned; the 4th byte
     * is ignored), quality 1..100 (out-of-range values clamp, as jpeg_set_quality does).
     * workspace: device, 256-byte aligned, isaac_hip_jpeg_workspace_bytes(width, height)
     * bytes (-1 for sizes outside 1..65535); the file goes to jpeg (host, capacity bytes),
     * *length its size (also set when capacity is too small, which returns
     * ISAAC_HIP_ERR_ARG).  Blocks until the file is on the host.
     * isaac_hip_jpeg_frame_message: {"payload": "data:image/jpeg;base64,<file>"} with a
     * terminating 0 into msg (host), base64 as for isaac_hip_base64. */
    /* The video frame of the context's last render (its framebuffer, as isaac_hip_render left
     * it), the message serverSendFrame sends after each doVisualization
     * (lib/isaac.hpp:2845-2861): quality 1..100 the JPEG path, 0 the raw-RGBA path.  The
     * device workspace is the context's own (allocated on first use, sized for the frame).
     * isaac_hip_video_frame_capacity: a msg size that always suffices.  Blocks until done. */
    int isaac_hip_video_frame_message(isaac_hip_ctx* ctx, int32_t quality, char* msg, int64_t capacity, int64_t* length);
    int64_t isaac_hip_video_frame_capacity(const isaac_hip_ctx* ctx);
    int64_t isaac_hip_jpeg_workspace_bytes(int32_t width, int32_t height);
    int isaac_hip_jpeg_encode(const uint8_t* rgbx, int32_t width, int32_t height, int32_t quality, void* workspace,
                              int64_t workspace_bytes, uint8_t* jpeg, int64_t capacity, int64_t* length, void* stream);
    int isaac_hip_jpeg_frame_message(const uint8_t* rgbx, int32_t width, int32_t height, int32_t quality, void* workspace,
                                     int64_t workspace_bytes, char* msg, int64_t capacity, int64_t* length, void* stream);

    /* No counterpart in the reference.  For callers that cannot allocate device memory
     * (serving loops and harnesses that forbid hipMalloc): isaac_hip_jpeg_encode without
     * the workspace argument.  The same kernels work in the library's own device scratch
     * of ISAAC_HIP_JPEG_NOALLOC_SCRATCH_BYTES (44 MiB), which every GPU the library runs
     * on holds until the process ends; frames whose isaac_hip_jpeg_workspace_bytes exceeds
     * it (larger than 1920x1080 or 2048x1024) return ISAAC_HIP_ERR_ARG.  The stream must
     * belong to the current device (0: its null stream).  Calls no hipMalloc or hipFree.
     * Every call on one GPU uses that GPU's one scratch: calls that overlap on one GPU
     * (from other threads or on other streams) race, so the caller serialises them.
     * Blocks until the file is on the host. */
#define ISAAC_HIP_JPEG_NOALLOC_SCRATCH_BYTES (44ll << 20)
    int isaac_hip_jpeg_encode_noalloc(const uint8_t* rgbx, int32_t width, int32_t height, int32_t quality, uint8_t* jpeg,
                                      int64_t capacity, int64_t* length, void* stream);

    /* ---------------------------------------------------------------- learned occlusion
     * No counterpart in the reference.  A U-Net over the G-buffer (input: depth /
     * maxGlobalSizeScaled, background flag, normal; 3x3 convolutions with 16/32/64/64
     * channels over four levels, FP16 weights and activations, FP32 accumulation on the
     * matrix cores; 1x1 linear head) that stands in for the SSAO pass
     * (SSAOKernel, lib/isaac/isaac_ssao_kernel.hpp:161-182) when the occlusion mode is
     * LEARNED.  It is trained on the GPU from crops of frames rendered with the plain
     * pass: the loss is sum((net - ao_plain)^2), the optimiser Adam.
     * Parameters: one flat float vector; per convolution layer (e0a e0b e1a e1b e2a e2b
     * e3a e3b d2 d1 d0) the weights [cout][9][cin] (tap = ky*3 + kx) then the bias
     * [cout]; last the head's 16 weights and its bias. */
    enum isaac_hip_ao_mode
    {
        ISAAC_HIP_AO_PLAIN = 0, /* the reference's 7x7 depth-comparison pass (default) */
        ISAAC_HIP_AO_LEARNED = 1
    };
    int isaac_hip_set_ao_mode(isaac_hip_ctx* ctx, int32_t mode);
    int64_t isaac_hip_aonet_param_count(void);
    int isaac_hip_aonet_set_params(isaac_hip_ctx* ctx, const float* params, int64_t n); /* host; resets Adam */
    int isaac_hip_aonet_get_params(isaac_hip_ctx* ctx, float* params, int64_t n);
    /* training batch: square crops (side a multiple of 8, at most 64 crops) of the
     * current G-buffer, whose ao must hold the plain pass's result (render with
     * ambient occlusion on and mode PLAIN first); xy = n (x0, y0) pairs (host) */
    int isaac_hip_aonet_clear_batch(isaac_hip_ctx* ctx);
    int isaac_hip_aonet_add_crops(isaac_hip_ctx* ctx, int32_t n, const int32_t* xy, int32_t side);
    /* one Adam step over the batch; *loss = mean squared error before the step */
    int isaac_hip_aonet_train_step(isaac_hip_ctx* ctx, float lr, float* loss);
    /* checks: loss and gradient (host arrays, n*h*w*8 features, n*h*w targets, h and
     * w multiples of 8) and the raw network output; both overwrite the batch */
    int isaac_hip_aonet_loss_grad(
        isaac_hip_ctx* ctx,
        int32_t n,
        int32_t h,
        int32_t w,
        const float* features,
        const float* target,
        float* loss,
        float* grad);
    int isaac_hip_aonet_forward(isaac_hip_ctx* ctx, int32_t n, int32_t h, int32_t w, const float* features, float* out);

    /* ---------------------------------------------------------------- introspection */
    int isaac_hip_get_timings(const isaac_hip_ctx* ctx, isaac_hip_timings* out);
    int isaac_hip_set_counting(isaac_hip_ctx* ctx, int enabled);
    int isaac_hip_get_counters(isaac_hip_ctx* ctx, uint64_t out[ISAAC_HIP_CNT_COUNT]);
    /* device pointers of the G-buffer (colour RGBA8, depth f32, normal 3xf32, ao f32), W*H each */
    int isaac_hip_get_gbuffer(
        const isaac_hip_ctx* ctx,
        const void** color,
        const void** depth,
        const void** normal,
        const void** ao);
    /* device pointer of the combined textures (lib/isaac.hpp:646-647): one texel per
     * local voxel, x fastest; each texel is *texel_bytes (8) bytes: the combined volume
     * RGBA8 followed by the combined iso RGBA8 */
    int isaac_hip_get_combined_textures(const isaac_hip_ctx* ctx, const void** texels, int32_t* texel_bytes);
    /* device pointers of field source field_idx's current advection texture (local
     * voxels, x fastest, one byte each) and of the noise texture (floats), and the
     * time step of the last buffer update */
    int isaac_hip_get_advection(
        const isaac_hip_ctx* ctx,
        int field_idx,
        const void** texture,
        const void** noise,
        int32_t* time_step);
    /* device pointers of a particle source's exact cell list: per-cell count
     * (int32) and exclusive offset (int32, ncells+1), and the stable
     * permutation of particle indices (int32).  update_buffers builds only what
     * the frame reads (each cell's records of the particles whose transfer-function
     * alpha is >= 0.5, in ascending index); the first call after it builds the full
     * list on the context stream from the particle buffers of that update, which
     * must still hold the same data.  ISAAC_HIP_ERR_STATE before any update. */
    int isaac_hip_get_cell_list(
        isaac_hip_ctx* ctx,
        int particle_idx,
        const void** counts,
        const void** offsets,
        const void** perm);

    /* ---------------------------------------------------------------- host helpers (no GPU needed) */
    /* DefaultController projection: glm::perspective(45.0, w/h, 0.002, 100)
     * (lib/isaac/isaac_controllers.hpp:33-37), column-major */
    void isaac_hip_default_projection(int32_t w, int32_t h, double out[16]);
    /* updateModelview (lib/isaac.hpp:2878-2889); rotation column-major 3x3 */
    void isaac_hip_modelview(const double look_at[3], const double rotation[9], double distance, double out[16]);
    /* the parser of updateFunctions for one chain; returns errorCode */
    int32_t isaac_hip_parse_functor_chain(const char* chain, int32_t bytecode[3], float params[12]);
    /* updateTransfer's LUT build; returns 0 */
    int isaac_hip_build_transfer_lut(int32_t T, int32_t n, const uint32_t* positions, const float* rgba, float* lut);
    /* default TF control points of source i of n (lib/isaac.hpp:772-792): 2 points */
    int isaac_hip_default_transfer(
        int32_t i,
        int32_t n_sources,
        int32_t is_particle,
        int32_t T,
        uint32_t positions[2],
        float rgba[8]);
    /* ---------------------------------------------------------------- server messages
     * The JSON messages the master rank sends to the ISAAC server, as the reference
     * builds them with jansson and writes them with json_dumps(root, 0): the register
     * message of the constructor (lib/isaac.hpp:834-892, sent by init, 1276-1292), the
     * periodic state message of each visualization step (lib/isaac.hpp:2546-2838) and
     * the exit message of the destructor (lib/isaac.hpp:2110-2118); and the bytes
     * serverSend puts on the socket (lib/isaac/isaac_communicator.hpp:125-190).  Host
     * only: no GPU work, no context.  jansson's writer as restated here: objects keep
     * insertion order, ", " between items and ": " after keys, no indentation;
     * integers as %lld; reals as %.17g with ".0" appended when the text has neither
     * '.' nor 'e', and a '+' and leading zeros removed from the exponent; strings with
     * \" \\ \b \f \n \r \t and \u00XX (uppercase hex) for other control characters,
     * '/' and non-ASCII bytes unescaped.  Every message function writes the text and a
     * terminating 0 into msg (host, capacity bytes) and sets *length to the text's
     * length (also when capacity is too small, which returns ISAAC_HIP_ERR_ARG). */
    typedef struct isaac_hip_steering_state
    {
        int32_t n_sources; /* combined order: volume, field, particle sources */
        int32_t n_volume_field; /* volume + field sources (those with an iso threshold) */
        const char* const* source_names; /* getName() of each source */
        const int32_t* feature_dims;
        const char* const* functions; /* functor chain text of each source */
        const int32_t* function_errors; /* its errorCode */
        int32_t transfer_size; /* T */
        const float* transfer_lut; /* n_sources * T RGBA floats */
        const int32_t* transfer_point_counts; /* control points per source */
        const uint32_t* transfer_point_values; /* all sources' points, concatenated, ascending per source */
        const float* transfer_point_rgba; /* 4 per point */
        const float* weights; /* n_sources */
        const float* iso_thresholds; /* n_volume_field */
        const float* min_max; /* (min, max) per source */
        float step;
        int32_t interpolation;
        int32_t render_optimization;
        int32_t seed_points;
        float advection_step;
        float advection_weight;
        int32_t advection_seeding_period;
        int32_t advection_seeding_duration;
        int32_t advection_on_pause;
        int32_t advection_border;
        float background[3];
        int32_t n_clipping;
        const float* clipping_position; /* 3 per plane */
        const float* clipping_normal; /* 3 per plane: the normals as set (clippingSavedNormals) */
        int32_t ao_enabled;
        float ao_weight;
        int32_t render_mode;
        int32_t dither_mode;
        double projection[16]; /* column-major */
        double look_at[3];
        double rotation[9]; /* column-major */
        double distance;
        const char* metadata; /* JSON text of the "metadata" object; NULL: {} */
    } isaac_hip_steering_state;

    /* sections of the periodic message: the reference's send* flags */
    enum isaac_hip_send_flag
    {
        ISAAC_HIP_SEND_PROJECTION = 1u << 0,
        ISAAC_HIP_SEND_LOOK_AT = 1u << 1,
        ISAAC_HIP_SEND_ROTATION = 1u << 2,
        ISAAC_HIP_SEND_DISTANCE = 1u << 3,
        ISAAC_HIP_SEND_TRANSFER = 1u << 4,
        ISAAC_HIP_SEND_FUNCTIONS = 1u << 5,
        ISAAC_HIP_SEND_WEIGHT = 1u << 6,
        ISAAC_HIP_SEND_INTERPOLATION = 1u << 7,
        ISAAC_HIP_SEND_RENDER_OPTIMIZATION = 1u << 8,
        ISAAC_HIP_SEND_STEP = 1u << 9,
        ISAAC_HIP_SEND_SEED_POINTS = 1u << 10,
        ISAAC_HIP_SEND_ADVECTION_STEP = 1u << 11,
        ISAAC_HIP_SEND_ADVECTION_WEIGHT = 1u << 12,
        ISAAC_HIP_SEND_ADVECTION_SEEDING_PERIOD = 1u << 13,
        ISAAC_HIP_SEND_ADVECTION_SEEDING_DURATION = 1u << 14,
        ISAAC_HIP_SEND_ADVECTION_ON_PAUSE = 1u << 15,
        ISAAC_HIP_SEND_ADVECTION_BORDER = 1u << 16,
        ISAAC_HIP_SEND_ISO_THRESHOLD = 1u << 17,
        ISAAC_HIP_SEND_MIN_MAX = 1u << 18,
        ISAAC_HIP_SEND_BACKGROUND = 1u << 19,
        ISAAC_HIP_SEND_CLIPPING = 1u << 20,
        ISAAC_HIP_SEND_AO = 1u << 21,
        ISAAC_HIP_SEND_RENDER_MODE = 1u << 22,
        ISAAC_HIP_SEND_DITHER_MODE = 1u << 23,
        ISAAC_HIP_SEND_INIT = 1u << 24 /* sendInitJson set by a request ("init" / "retry") */
    };

    /* The master rank's message state across steps: the init root (jsonInitRoot), which
     * starts as the register message's root and takes the camera, interpolation, step,
     * advection, background, occlusion and mode values of every periodic message that
     * sends them; a periodic message that sends any of these (or SEND_INIT) carries it
     * as "init". */
    typedef struct isaac_hip_session isaac_hip_session;
    int isaac_hip_session_create(isaac_hip_session** out);
    void isaac_hip_session_destroy(isaac_hip_session* session);
    /* the register message; it becomes the session's init root.  framebuffer: the
     * compositing buffer size; global_size_scaled: globalSizeScaled */
    int isaac_hip_register_message(
        isaac_hip_session* session,
        const isaac_hip_steering_state* state,
        const char* name,
        int32_t nodes,
        const int32_t framebuffer[2],
        const int32_t global_size_scaled[3],
        char* msg,
        int64_t capacity,
        int64_t* length);
    /* the periodic message with meta number meta_nr and the sections in send (flags);
     * ISAAC_HIP_ERR_STATE before the register message */
    int isaac_hip_period_message(
        isaac_hip_session* session,
        const isaac_hip_steering_state* state,
        int64_t meta_nr,
        uint32_t send,
        char* msg,
        int64_t capacity,
        int64_t* length);
    int isaac_hip_exit_message(char* msg, int64_t capacity, int64_t* length);
    /* serverSend's bytes for one call: with uid >= 0 (starting) "{\"uid\": <uid>"; then,
     * when content is not NULL, content with its first byte replaced by ',' and without
     * its last byte; then "} " when finishing.  A periodic message goes out starting and
     * not finishing, the video frame message after it not starting and finishing (or,
     * without a frame, content NULL and finishing), so the server reads one object
     * (lib/isaac.hpp:2836-2861). */
    int isaac_hip_server_send_bytes(int32_t uid, const char* content, int32_t finishing, char* out, int64_t capacity,
                                    int64_t* length);

    /* library version string */
    const char* isaac_hip_version(void);
    /* blocking copy of `bytes` from device memory to host memory on the context
     * stream (reads back G-buffer / texture / cell-list pointers) */
    int isaac_hip_copy_to_host(isaac_hip_ctx* ctx, void* dst, const void* src_device, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* ISAAC_HIP_H */
