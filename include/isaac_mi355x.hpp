// isaac_mi355x.hpp — header-only C++17 drop-in for one rank's
// IsaacVisualization (lib/isaac.hpp:58-3036), forwarding to the gfx950 C-ABI
// in isaac_hip.h.  Link with -lisaac_hip.
//
// What stays the same as the reference interface:
//   - the constructor arguments (name, master, server, sizes, position,
//     source lists, scale; lib/isaac.hpp:565-583); the alpaka host/acc/stream
//     triple becomes a device index and an optional hipStream_t (void*)
//   - source concepts: featureDim, guardSize, persistent, getName(),
//     update(bool enabled, void* pointer) (example/example.cpp:39-228)
//   - steering state and setters: functions[i].source + updateFunctions(),
//     transfer points + updateTransfer(), sourceWeight, isoThreshold, step,
//     interpolation, ditherMode, renderMode, ambientOcclusion,
//     add/edit/removeClipping, updatePosition/LocalSize/LocalParticleSize
//   - doVisualization(metaTargets, pointer, redraw) and the public timing
//     members kernelTime, copyTime, bufferTime, optimizationBufferTime
// What differs, because a precompiled GPU library cannot call a user's
// device-side operator[]:
//   - a volume source also provides `isaac::VolumeView view() const` (device
//     pointer at local voxel (0,0,0) plus element strides);
//   - a particle source provides `isaac::ParticleView view() const` (raw
//     particles; the library builds the per-cell lists that getIterator(cell)
//     returned in the reference);
//   - source lists are std::tuple instead of boost::fusion::list;
//   - there is no server, JSON, MPI or IceT: doVisualization renders the
//     local sub-volume and keeps the RGBA8 frame in `image()`.  It still
//     returns json_t* (always null: no client metadata arrives without a
//     server) and getJsonMetaRoot() returns a null json_t*, so the
//     reference's `json_decref(vis->doVisualization(META_MASTER))` and
//     `json_object_set_new(vis->getJsonMetaRoot(), ...)` compile and run
//     unchanged: jansson accepts a null json_t* in both; without jansson,
//     include isaac_json_stub.h for no-op stand-ins.
#pragma once

#include "isaac_hip.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

// jansson's root type (struct json_t, typedef json_t); declared here so the
// signatures match the reference whether or not jansson.h is included
struct json_t;
typedef struct json_t json_t;

namespace isaac
{
    using isaac_float = float;
    using isaac_double = double;
    using isaac_int = int32_t;
    using isaac_uint = uint32_t;

    struct isaac_size2
    {
        int64_t x, y;
    };
    struct isaac_size3
    {
        int64_t x, y, z;
    };
    struct isaac_int3
    {
        int32_t x, y, z;
    };
    struct isaac_float3
    {
        float x, y, z;
    };
    struct isaac_float4
    {
        float x, y, z, w;
    };

    enum IsaacVisualizationMetaEnum
    {
        META_MERGE = 0,
        META_MASTER = 1
    };

    struct VolumeView
    {
        const float* data; // device pointer at local voxel (0,0,0)
        int64_t stride[3]; // element strides x, y, z
    };

    struct ParticleView
    {
        const float* position; // device, count x 3, local voxel coordinates
        const float* attribute; // device, count x featureDim
        const float* radius; // device, count, or nullptr
        float constantRadius; // getRadius() when radius == nullptr
        int64_t count;
    };

    struct FunctionsStruct
    {
        std::string source;
        isaac_int errorCode = 0;
    };

    struct AOParams
    {
        isaac_float weight = 0.5f;
        bool isEnabled = false;
    };

    // MinMax / MinMaxArray (lib/isaac/isaac_types.hpp:223-234)
    struct MinMax
    {
        isaac_float min;
        isaac_float max;
    };

    template<int T_n>
    struct MinMaxArray
    {
        isaac_float min[T_n > 0 ? T_n : 1];
        isaac_float max[T_n > 0 ? T_n : 1];
    };

    // DefaultController (lib/isaac/isaac_controllers.hpp:23-43)
    struct DefaultController
    {
        static const int passCount = 1;
        static void projection(int32_t w, int32_t h, double out[16])
        {
            isaac_hip_default_projection(w, h, out);
        }
    };

    class IsaacError : public std::runtime_error
    {
    public:
        using std::runtime_error::runtime_error;
    };

    template<typename T_VolumeSourceList, typename T_FieldSourceList, typename T_ParticleList, int T_transferSize, typename T_Controller = DefaultController>
    class IsaacVisualization
    {
        static constexpr int vSourceListSize = int(std::tuple_size<T_VolumeSourceList>::value);
        static constexpr int fSourceListSize = int(std::tuple_size<T_FieldSourceList>::value);
        static constexpr int pSourceListSize = int(std::tuple_size<T_ParticleList>::value);

    public:
        static constexpr int volumeFieldSourceListSize = vSourceListSize + fSourceListSize;
        static constexpr int combinedSourceListSize = volumeFieldSourceListSize + pSourceListSize;

        IsaacVisualization(
            int device,
            void* stream,
            const std::string name,
            const isaac_int master,
            const std::string serverUrl,
            const isaac_uint serverPort,
            const isaac_size2 framebufferSize,
            const isaac_size3 globalSize,
            const isaac_size3 localSize,
            const isaac_size3 localParticleSize,
            const isaac_int3 position,
            T_VolumeSourceList& volumeSources,
            T_FieldSourceList& fieldSources,
            T_ParticleList& particleSources,
            isaac_float3 scale)
            : name(name)
            , master(master)
            , framebufferSize(framebufferSize)
            , volumeSources(volumeSources)
            , fieldSources(fieldSources)
            , particleSources(particleSources)
        {
            (void) serverUrl;
            (void) serverPort;
            isaac_hip_config cfg{};
            cfg.framebuffer_size[0] = int32_t(framebufferSize.x);
            cfg.framebuffer_size[1] = int32_t(framebufferSize.y);
            const int64_t g[3] = {globalSize.x, globalSize.y, globalSize.z};
            const int64_t l[3] = {localSize.x, localSize.y, localSize.z};
            const int64_t p[3] = {localParticleSize.x, localParticleSize.y, localParticleSize.z};
            for(int i = 0; i < 3; i++)
            {
                cfg.global_size[i] = g[i];
                cfg.local_size[i] = l[i];
                cfg.local_particle_size[i] = p[i];
            }
            cfg.position[0] = position.x;
            cfg.position[1] = position.y;
            cfg.position[2] = position.z;
            cfg.scale[0] = scale.x;
            cfg.scale[1] = scale.y;
            cfg.scale[2] = scale.z;
            cfg.transfer_size = T_transferSize;
            cfg.n_volume_sources = vSourceListSize;
            cfg.n_field_sources = fSourceListSize;
            cfg.n_particle_sources = pSourceListSize;
            int rc = isaac_hip_create(&cfg, device, &ctx);
            if(rc != 0)
                throw IsaacError("isaac_hip_create failed (" + std::to_string(rc) + ")");
            if(stream)
                check(isaac_hip_set_stream(ctx, stream), "set_stream");
            // defaults (lib/isaac.hpp:608-609,696-699,720-794,2987-2988)
            for(int i = 0; i < combinedSourceListSize; i++)
            {
                functions[i].source = (i >= vSourceListSize && i < volumeFieldSourceListSize) ? "length" : "idem";
                sourceWeight[i] = 2.0f;
                isoThreshold[i] = 0.0f;
                uint32_t pos[2];
                float rgba[8];
                isaac_hip_default_transfer(i, combinedSourceListSize, i >= volumeFieldSourceListSize, T_transferSize, pos, rgba);
                transferPoints[i][pos[0]] = isaac_float4{rgba[0], rgba[1], rgba[2], rgba[3]};
                transferPoints[i][pos[1]] = isaac_float4{rgba[4], rgba[5], rgba[6], rgba[7]};
            }
            T_Controller::projection(int32_t(framebufferSize.x), int32_t(framebufferSize.y), projection);
            rotation = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            lookAt = {0, 0, 0};
            distance = -4.5;
            updateModelview();
            updateFunctions();
            updateTransfer();
            frame.resize(size_t(framebufferSize.x) * size_t(framebufferSize.y) * 4);
        }

        ~IsaacVisualization()
        {
            isaac_hip_session_destroy(session);
            isaac_hip_destroy(ctx);
        }

        IsaacVisualization(const IsaacVisualization&) = delete;
        IsaacVisualization& operator=(const IsaacVisualization&) = delete;

        // no server: always succeeds (lib/isaac.hpp:1276-1295 returns -1 only on connection failure)
        int init(int communicatorBehaviour = 0)
        {
            (void) communicatorBehaviour;
            return 0;
        }

        void updateFunctions()
        {
            for(int i = 0; i < combinedSourceListSize; i++)
            {
                int32_t err = 0;
                check(isaac_hip_set_function(ctx, i, functions[i].source.c_str(), &err), "set_function");
                functions[i].errorCode = err;
            }
        }

        void updateTransfer()
        {
            for(int i = 0; i < combinedSourceListSize; i++)
            {
                std::vector<uint32_t> pos;
                std::vector<float> rgba;
                for(const auto& kv : transferPoints[i])
                {
                    pos.push_back(kv.first);
                    rgba.insert(rgba.end(), {kv.second.x, kv.second.y, kv.second.z, kv.second.w});
                }
                check(isaac_hip_set_transfer_points(ctx, i, int(pos.size()), pos.data(), rgba.data()), "set_transfer");
            }
        }

        bool editClipping(isaac_uint nr, isaac_float px, isaac_float py, isaac_float pz, isaac_float nx, isaac_float ny, isaac_float nz)
        {
            return isaac_hip_edit_clipping(ctx, nr, px, py, pz, nx, ny, nz) == 1;
        }

        void addClipping(isaac_float px, isaac_float py, isaac_float pz, isaac_float nx, isaac_float ny, isaac_float nz)
        {
            isaac_hip_add_clipping(ctx, px, py, pz, nx, ny, nz);
        }

        void removeClipping(isaac_uint nr)
        {
            isaac_hip_remove_clipping(ctx, nr);
        }

        void updatePosition(const isaac_int3 position)
        {
            const int32_t p[3] = {position.x, position.y, position.z};
            check(isaac_hip_update_position(ctx, p), "update_position");
        }

        void updateLocalSize(const isaac_size3 localSize)
        {
            const int64_t l[3] = {localSize.x, localSize.y, localSize.z};
            check(isaac_hip_update_local_size(ctx, l), "update_local_size");
        }

        void updateLocalParticleSize(const isaac_size3 localParticleSize)
        {
            const int64_t l[3] = {localParticleSize.x, localParticleSize.y, localParticleSize.z};
            check(isaac_hip_update_local_particle_size(ctx, l), "update_local_particle_size");
        }

        // camera steering (the JSON keys of lib/isaac.hpp:1440-1513)
        void setRotation(const std::array<double, 9>& columnMajor)
        {
            rotation = columnMajor;
            updateModelview();
        }
        void setLookAt(double x, double y, double z)
        {
            lookAt = {x, y, z};
            updateModelview();
        }
        void setDistance(double d)
        {
            distance = d;
            updateModelview();
        }

        // getJsonMetaRoot (lib/isaac.hpp:1269-1273): the metadata root the reference
        // sends to the server; there is no server here, so it is null (jansson's
        // json_object_set_new on it returns -1 and releases the value)
        json_t* getJsonMetaRoot()
        {
            return nullptr;
        }

        // doVisualization (lib/isaac.hpp:1298-2104) for this rank: buffers when
        // redrawing, then the draw callback; the frame lands in image().  Returns the
        // client metadata the reference forwards (lib/isaac/mainpage.dox:113-117):
        // always null here, and json_decref(nullptr) is a no-op
        json_t* doVisualization(IsaacVisualizationMetaEnum metaTargets = META_MASTER, void* pointer = nullptr, bool redraw = true)
        {
            (void) metaTargets;
            redraw = redraw || advectionOnPause; // lib/isaac.hpp:1303
            pushState();
            if(redraw)
            {
                bindSources(pointer);
                check(isaac_hip_update_buffers(ctx), "update_buffers");
            }
            // the min/max request (lib/isaac.hpp:1951-2006, one rank)
            if(sendMinMax)
            {
                check(
                    isaac_hip_min_max(ctx, minMaxArray.min, minMaxArray.max, combinedSourceListSize > 0 ? combinedSourceListSize : 1),
                    "min_max");
                sendMinMax = false;
            }
            const int32_t vp[4] = {0, 0, int32_t(framebufferSize.x), int32_t(framebufferSize.y)};
            check(isaac_hip_render(ctx, projection, modelview, backgroundColor, vp, frame.data(), 0), "render");
            isaac_hip_timings t{};
            isaac_hip_get_timings(ctx, &t);
            kernelTime = uint64_t(t.kernel_us);
            copyTime = uint64_t(t.copy_us);
            bufferTime = uint64_t(t.buffer_us);
            optimizationBufferTime = uint64_t(t.optimization_buffer_us);
            return nullptr;
        }

        // RGBA8, premultiplied, row 0 = bottom (the IceT image of lib/isaac.hpp:2533-2539)
        const std::vector<uint8_t>& image() const
        {
            return frame;
        }

        // the message serverSendFrame sends for the last frame (lib/isaac.hpp:2845-2861):
        // {"payload": "data:image/jpeg;base64,..."} (quality 1..100, the communicator's default
        // 90) or the raw-RGBA payload (quality 0); JPEG and base64 computed on the GPU
        std::string videoFrameMessage(int quality = 90)
        {
            const int64_t cap = isaac_hip_video_frame_capacity(ctx);
            if(cap <= 0)
                throw std::runtime_error("isaac: video frame capacity");
            message.resize(size_t(cap));
            int64_t n = 0;
            check(isaac_hip_video_frame_message(ctx, quality, message.data(), cap, &n), "video_frame_message");
            return std::string(message.data(), size_t(n));
        }

        // The master rank's messages to the ISAAC server, as text (jansson's
        // json_dumps(root, 0)): the register message init() sends
        // (lib/isaac.hpp:834-892, 1276-1292), the periodic message of one step with
        // the sections in `send` (ISAAC_HIP_SEND_* flags, the reference's send*
        // members; lib/isaac.hpp:2546-2838), numbered by metaNr, which it advances,
        // and the exit message of the destructor (lib/isaac.hpp:2110-2118).
        // `metadata` is the JSON text of the metadata object (getJsonMetaRoot).
        std::string registerMessage(int nodes = 1)
        {
            if(!session)
                check(isaac_hip_session_create(&session) == 0 ? 0 : -1, "session_create");
            MessageState st(*this);
            const int32_t fb[2] = {int32_t(framebufferSize.x), int32_t(framebufferSize.y)};
            int32_t gs[3] = {0, 0, 0};
            check(isaac_hip_global_size_scaled(ctx, gs), "global_size_scaled");
            return messageText([&](char* m, int64_t cap, int64_t* n)
                               { return isaac_hip_register_message(session, &st.s, name.c_str(), nodes, fb, gs, m, cap, n); });
        }

        std::string periodMessage(uint32_t send)
        {
            if(!session)
                throw IsaacError("periodMessage: registerMessage first (init)");
            MessageState st(*this);
            const int64_t nr = metaNr++;
            return messageText([&](char* m, int64_t cap, int64_t* n)
                               { return isaac_hip_period_message(session, &st.s, nr, send, m, cap, n); });
        }

        static std::string exitMessage()
        {
            return messageText([](char* m, int64_t cap, int64_t* n) { return isaac_hip_exit_message(m, cap, n); });
        }

        std::string metadata = "{}";
        int64_t metaNr = 0;

        isaac_hip_ctx* handle() const
        {
            return ctx;
        }

        // steering state, reference names
        FunctionsStruct functions[combinedSourceListSize > 0 ? combinedSourceListSize : 1];
        std::map<isaac_uint, isaac_float4> transferPoints[combinedSourceListSize > 0 ? combinedSourceListSize : 1];
        isaac_float sourceWeight[combinedSourceListSize > 0 ? combinedSourceListSize : 1];
        isaac_float isoThreshold[combinedSourceListSize > 0 ? combinedSourceListSize : 1];
        isaac_float step = 0.5f;
        // field-source advection steering (lib/isaac.hpp:610,639-643)
        isaac_int seedPoints = 1000;
        isaac_float advectionStepFactor = 1.0f;
        isaac_float advectionHistoryWeight = 0.95f;
        isaac_int advectionSeedingPeriod = 1;
        isaac_int advectionSeedingTime = 1;
        bool advectionOnPause = false;
        // set to request min/max of every source with the next doVisualization; the
        // result lands in minMaxArray (combined source order)
        bool sendMinMax = false;
        MinMaxArray<combinedSourceListSize> minMaxArray;
        bool interpolation = false;
        isaac_int ditherMode = 1;
        isaac_int renderMode = 0;
        AOParams ambientOcclusion;
        float backgroundColor[4] = {0, 0, 0, 1};
        double projection[16];
        double modelview[16];
        std::array<double, 3> lookAt;
        std::array<double, 9> rotation;
        double distance;
        // per-phase microseconds (lib/isaac.hpp:2129-2137)
        uint64_t kernelTime = 0, copyTime = 0, bufferTime = 0, optimizationBufferTime = 0;

    private:
        void check(int rc, const char* what)
        {
            if(rc < 0)
                throw IsaacError(std::string(what) + ": " + isaac_hip_last_error(ctx));
        }

        template<typename F>
        static std::string messageText(F&& write)
        {
            std::vector<char> buf(1 << 16);
            int64_t n = 0;
            int rc = write(buf.data(), int64_t(buf.size()), &n);
            if(rc == ISAAC_HIP_ERR_ARG && n + 1 > int64_t(buf.size()))
            {
                buf.resize(size_t(n) + 1);
                rc = write(buf.data(), int64_t(buf.size()), &n);
            }
            if(rc != 0)
                throw IsaacError("server message: error " + std::to_string(rc));
            return std::string(buf.data(), size_t(n));
        }

        // the visualization state the messages report, in the C-ABI's form
        struct MessageState
        {
            isaac_hip_steering_state s{};
            std::vector<std::string> names, funcs;
            std::vector<const char*> namePtr, funcPtr;
            std::vector<int32_t> dims, errs, counts;
            std::vector<uint32_t> values;
            std::vector<float> lut, rgba, minMax, clipPos, clipNormal;

            template<typename T_Tuple, std::size_t... I>
            void collect(const T_Tuple& t, std::index_sequence<I...>)
            {
                (names.push_back(std::tuple_element_t<I, T_Tuple>::getName()), ...);
                (dims.push_back(int32_t(std::tuple_element_t<I, T_Tuple>::featureDim)), ...);
                (void) t;
            }

            explicit MessageState(const IsaacVisualization& v)
            {
                collect(v.volumeSources, std::make_index_sequence<vSourceListSize>{});
                collect(v.fieldSources, std::make_index_sequence<fSourceListSize>{});
                collect(v.particleSources, std::make_index_sequence<pSourceListSize>{});
                const int n = combinedSourceListSize;
                lut.resize(size_t(std::max(n, 1)) * T_transferSize * 4);
                for(int i = 0; i < n; i++)
                {
                    funcs.push_back(v.functions[i].source);
                    errs.push_back(v.functions[i].errorCode);
                    std::vector<uint32_t> pos;
                    std::vector<float> col;
                    for(const auto& kv : v.transferPoints[i])
                    {
                        pos.push_back(kv.first);
                        col.insert(col.end(), {kv.second.x, kv.second.y, kv.second.z, kv.second.w});
                    }
                    isaac_hip_build_transfer_lut(T_transferSize, int32_t(pos.size()), pos.data(), col.data(),
                                                 lut.data() + size_t(i) * T_transferSize * 4);
                    counts.push_back(int32_t(pos.size()));
                    values.insert(values.end(), pos.begin(), pos.end());
                    rgba.insert(rgba.end(), col.begin(), col.end());
                    minMax.push_back(v.minMaxArray.min[i]);
                    minMax.push_back(v.minMaxArray.max[i]);
                }
                for(const std::string& x : names)
                    namePtr.push_back(x.c_str());
                for(const std::string& x : funcs)
                    funcPtr.push_back(x.c_str());
                const int nclip = isaac_hip_clipping_count(v.ctx);
                for(int i = 0; i < nclip; i++)
                {
                    float p[3], q[3];
                    isaac_hip_get_clipping(v.ctx, uint32_t(i), p, q);
                    clipPos.insert(clipPos.end(), p, p + 3);
                    clipNormal.insert(clipNormal.end(), q, q + 3);
                }
                s.n_sources = n;
                s.n_volume_field = volumeFieldSourceListSize;
                s.source_names = namePtr.data();
                s.feature_dims = dims.data();
                s.functions = funcPtr.data();
                s.function_errors = errs.data();
                s.transfer_size = T_transferSize;
                s.transfer_lut = lut.data();
                s.transfer_point_counts = counts.data();
                s.transfer_point_values = values.data();
                s.transfer_point_rgba = rgba.data();
                s.weights = v.sourceWeight;
                s.iso_thresholds = v.isoThreshold;
                s.min_max = minMax.data();
                s.step = v.step;
                s.interpolation = v.interpolation ? 1 : 0;
                s.render_optimization = 1;
                s.seed_points = v.seedPoints;
                s.advection_step = v.advectionStepFactor;
                s.advection_weight = v.advectionHistoryWeight;
                s.advection_seeding_period = v.advectionSeedingPeriod;
                s.advection_seeding_duration = v.advectionSeedingTime;
                s.advection_on_pause = v.advectionOnPause ? 1 : 0;
                s.advection_border = 0;
                for(int k = 0; k < 3; k++)
                    s.background[k] = v.backgroundColor[k];
                s.n_clipping = nclip;
                s.clipping_position = clipPos.data();
                s.clipping_normal = clipNormal.data();
                s.ao_enabled = v.ambientOcclusion.isEnabled ? 1 : 0;
                s.ao_weight = v.ambientOcclusion.weight;
                s.render_mode = v.renderMode;
                s.dither_mode = v.ditherMode;
                for(int k = 0; k < 16; k++)
                    s.projection[k] = v.projection[k];
                for(int k = 0; k < 3; k++)
                    s.look_at[k] = v.lookAt[size_t(k)];
                for(int k = 0; k < 9; k++)
                    s.rotation[k] = v.rotation[size_t(k)];
                s.distance = v.distance;
                s.metadata = v.metadata.c_str();
            }
        };

        isaac_hip_session* session = nullptr;

        void updateModelview()
        {
            isaac_hip_modelview(lookAt.data(), rotation.data(), distance, modelview);
        }

        void pushState()
        {
            for(int i = 0; i < combinedSourceListSize; i++)
            {
                check(isaac_hip_set_weight(ctx, i, sourceWeight[i]), "set_weight");
                check(isaac_hip_set_iso_threshold(ctx, i, isoThreshold[i]), "set_iso_threshold");
            }
            check(isaac_hip_set_step(ctx, step), "set_step");
            check(isaac_hip_set_interpolation(ctx, interpolation ? 1 : 0), "set_interpolation");
            check(isaac_hip_set_dither_mode(ctx, ditherMode), "set_dither_mode");
            check(isaac_hip_set_ao(ctx, ambientOcclusion.isEnabled ? 1 : 0, ambientOcclusion.weight), "set_ao");
            check(isaac_hip_set_render_mode(ctx, renderMode), "set_render_mode");
            check(isaac_hip_set_advection(ctx, advectionStepFactor, advectionHistoryWeight, advectionSeedingPeriod,
                                          advectionSeedingTime),
                  "set_advection");
            check(isaac_hip_set_seed_points(ctx, seedPoints), "set_seed_points");
        }

        template<typename T_Source>
        void bindVolume(int idx, T_Source& s, void* pointer)
        {
            const bool enabled = sourceWeight[idx] != 0.0f || isoThreshold[idx] != 0.0f;
            s.update(enabled, pointer);
            const VolumeView v = s.view();
            isaac_hip_volume_desc d{};
            d.data = v.data;
            d.dtype = ISAAC_HIP_F32;
            d.feature_dim = int32_t(T_Source::featureDim);
            for(int k = 0; k < 3; k++)
                d.stride[k] = v.stride[k];
            d.guard_size = int32_t(T_Source::guardSize);
            d.persistent = T_Source::persistent ? 1 : 0;
            check(isaac_hip_set_volume_source(ctx, idx, &d), "set_volume_source");
        }

        template<typename T_Source>
        void bindParticles(int idx, T_Source& s, void* pointer)
        {
            s.update(sourceWeight[volumeFieldSourceListSize + idx] != 0.0f, pointer);
            const ParticleView v = s.view();
            isaac_hip_particle_desc d{};
            d.position = v.position;
            d.attribute = v.attribute;
            d.feature_dim = int32_t(T_Source::featureDim);
            d.radius = v.radius;
            d.constant_radius = v.constantRadius;
            d.count = v.count;
            check(isaac_hip_set_particle_source(ctx, idx, &d), "set_particle_source");
        }

        template<std::size_t... I>
        void bindVolumes(void* pointer, std::index_sequence<I...>)
        {
            (bindVolume(int(I), std::get<I>(volumeSources), pointer), ...);
        }
        template<std::size_t... I>
        void bindFields(void* pointer, std::index_sequence<I...>)
        {
            (bindVolume(vSourceListSize + int(I), std::get<I>(fieldSources), pointer), ...);
        }
        template<std::size_t... I>
        void bindParticleList(void* pointer, std::index_sequence<I...>)
        {
            (bindParticles(int(I), std::get<I>(particleSources), pointer), ...);
        }

        void bindSources(void* pointer)
        {
            bindVolumes(pointer, std::make_index_sequence<vSourceListSize>{});
            bindFields(pointer, std::make_index_sequence<fSourceListSize>{});
            bindParticleList(pointer, std::make_index_sequence<pSourceListSize>{});
        }

        std::string name;
        isaac_int master;
        isaac_size2 framebufferSize;
        T_VolumeSourceList& volumeSources;
        T_FieldSourceList& fieldSources;
        T_ParticleList& particleSources;
        isaac_hip_ctx* ctx = nullptr;
        std::vector<uint8_t> frame;
        std::vector<char> message;
    };
} // namespace isaac
