// The JSON messages of the master rank to the ISAAC server (register, periodic state,
// exit) and serverSend's framing, written as the reference's jansson calls produce
// them (lib/isaac.hpp:834-892, 2110-2118, 2546-2838; json_dumps(root, 0);
// lib/isaac/isaac_communicator.hpp:125-190).  Host code only.
//
// A message is built as an ordered list of (key, JSON text) pairs; each value is
// rendered once, in jansson's format (see include/isaac_hip.h).
#include "isaac_hip.h"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace
{
    // ------------------------------------------------------------------ jansson's writer
    // "" for NaN and infinities: json_real refuses them, and jansson's append / set of
    // the resulting NULL adds nothing (jobj and jarr skip empty values)
    std::string jreal(double v)
    {
        if(!std::isfinite(v))
            return std::string();
        char b[64];
        const int n = std::snprintf(b, sizeof(b), "%.17g", v);
        std::string s(b, size_t(n));
        const size_t e = s.find('e');
        if(e == std::string::npos)
        {
            if(s.find('.') == std::string::npos)
                s += ".0";
            return s;
        }
        // the exponent without '+' and without leading zeros: "1e+22" -> "1e22", "1e-05" -> "1e-5"
        size_t d = e + 1;
        const bool neg = s[d] == '-';
        if(s[d] == '+' || s[d] == '-')
            d++;
        while(d + 1 < s.size() && s[d] == '0')
            d++;
        return s.substr(0, e + 1) + (neg ? "-" : "") + s.substr(d);
    }

    std::string jint(long long v)
    {
        char b[32];
        int n = std::snprintf(b, sizeof(b), "%lld", v);
        return std::string(b, size_t(n));
    }

    std::string jbool(bool v)
    {
        return v ? "true" : "false";
    }

    std::string jstr(const char* s)
    {
        std::string o = "\"";
        for(const unsigned char* p = reinterpret_cast<const unsigned char*>(s ? s : ""); *p; p++)
        {
            const unsigned char c = *p;
            switch(c)
            {
            case '\\': o += "\\\\"; break;
            case '"': o += "\\\""; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if(c < 0x20)
                {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04X", unsigned(c));
                    o += b;
                }
                else
                    o += char(c);
            }
        }
        return o + "\"";
    }

    using Items = std::vector<std::pair<std::string, std::string>>;

    std::string jobj(const Items& items)
    {
        std::string o = "{";
        bool first = true;
        for(const auto& kv : items)
        {
            if(kv.second.empty())
                continue;
            if(!first)
                o += ", ";
            first = false;
            o += jstr(kv.first.c_str()) + ": " + kv.second;
        }
        return o + "}";
    }

    std::string jarr(const std::vector<std::string>& v)
    {
        std::string o = "[";
        bool first = true;
        for(const std::string& x : v)
        {
            if(x.empty())
                continue;
            if(!first)
                o += ", ";
            first = false;
            o += x;
        }
        return o + "]";
    }

    template<typename T>
    std::string jreals(const T* p, int n)
    {
        std::vector<std::string> v;
        for(int i = 0; i < n; i++)
            v.push_back(jreal(double(p[i])));
        return jarr(v);
    }

    // json_object_set on an ordered object: replace in place, else append (a refused
    // value leaves the object as it was)
    void set(Items& items, const std::string& key, const std::string& value)
    {
        if(value.empty())
            return;
        for(auto& kv : items)
            if(kv.first == key)
            {
                kv.second = value;
                return;
            }
        items.emplace_back(key, value);
    }

    int emit(const std::string& text, char* msg, int64_t capacity, int64_t* length)
    {
        if(!length)
            return ISAAC_HIP_ERR_ARG;
        *length = int64_t(text.size());
        if(!msg || capacity < int64_t(text.size()) + 1)
            return ISAAC_HIP_ERR_ARG;
        std::memcpy(msg, text.c_str(), text.size() + 1);
        return 0;
    }

    // the functor pool in bytecode order (lib/isaac/isaac_functors.hpp:42-219)
    struct FunctorInfo
    {
        const char* name;
        const char* description;
        bool usesParameter;
    };
    const FunctorInfo kFunctors[] = {
        {"idem", "Does nothing. Keeps the feature dimension.", false},
        {"add", "Summarizes the input with a constant parameter. Keeps the feature dimension.", true},
        {"mul", "Multiplies the input with a constant parameter. Keeps the feature dimension.", true},
        {"length", "Calculates the length of an input. Reduces the feature dimension to 1.", false},
        {"pow", "Calculates the power of the input with a constant exponent. Keeps the feature dimension.", true},
        {"sum", "Calculates the sum of all components. Reduces the feature dimension to 1.", false},
    };

    // ISAAC_PROTOCOL_VERSION_MAJOR / _MINOR (lib/isaac/isaac_version.hpp:7-8)
    constexpr int kProtocolMajor = 1, kProtocolMinor = 0;

    bool valid_state(const isaac_hip_steering_state* s)
    {
        if(!s || s->n_sources < 0 || s->n_sources > ISAAC_HIP_MAX_SOURCES || s->n_volume_field < 0
           || s->n_volume_field > s->n_sources || s->n_clipping < 0 || s->n_clipping > ISAAC_HIP_MAX_CLIPPING)
            return false;
        if(s->n_sources > 0 && (!s->source_names || !s->feature_dims || !s->weights))
            return false;
        return true;
    }

    std::string metadata(const isaac_hip_steering_state* s)
    {
        return s->metadata && *s->metadata ? std::string(s->metadata) : std::string("{}");
    }
} // namespace

struct isaac_hip_session
{
    Items init; // jsonInitRoot
    bool registered = false;
};

extern "C"
{
    int isaac_hip_session_create(isaac_hip_session** out)
    {
        if(!out)
            return ISAAC_HIP_ERR_ARG;
        *out = new(std::nothrow) isaac_hip_session();
        return *out ? 0 : ISAAC_HIP_ERR_ALLOC;
    }

    void isaac_hip_session_destroy(isaac_hip_session* session)
    {
        delete session;
    }

    int isaac_hip_register_message(
        isaac_hip_session* session,
        const isaac_hip_steering_state* s,
        const char* name,
        int32_t nodes,
        const int32_t framebuffer[2],
        const int32_t global_size_scaled[3],
        char* msg,
        int64_t capacity,
        int64_t* length)
    {
        if(!session || !valid_state(s) || !framebuffer || !global_size_scaled)
            return ISAAC_HIP_ERR_ARG;
        Items r;
        // recreateJSON: "metadata" first (lib/isaac.hpp:2870-2875)
        r.emplace_back("metadata", metadata(s));
        r.emplace_back("type", jstr("register"));
        r.emplace_back("name", jstr(name));
        r.emplace_back("nodes", jint(nodes));
        r.emplace_back("framebuffer width", jint(framebuffer[0]));
        r.emplace_back("framebuffer height", jint(framebuffer[1]));
        r.emplace_back("max functors", jint(ISAAC_HIP_MAX_FUNCTORS));
        std::vector<std::string> functors;
        for(const FunctorInfo& f : kFunctors)
            functors.push_back(jobj({{"name", jstr(f.name)}, {"description", jstr(f.description)},
                                     {"uses parameter", jbool(f.usesParameter)}}));
        r.emplace_back("functors", jarr(functors));
        r.emplace_back("projection count", jint(1));
        r.emplace_back("projection", jreals(s->projection, 16));
        r.emplace_back("position", jreals(s->look_at, 3));
        r.emplace_back("rotation", jreals(s->rotation, 9));
        r.emplace_back("distance", jreal(s->distance));
        std::vector<std::string> sources;
        for(int i = 0; i < s->n_sources; i++)
            sources.push_back(jobj({{"name", jstr(s->source_names[i])}, {"feature dimension", jint(s->feature_dims[i])}}));
        r.emplace_back("sources", jarr(sources));
        r.emplace_back("interpolation", jbool(s->interpolation != 0));
        r.emplace_back("step", jreal(s->step));
        r.emplace_back("dimension", jint(3));
        r.emplace_back("width", jint(global_size_scaled[0]));
        r.emplace_back("height", jint(global_size_scaled[1]));
        r.emplace_back("depth", jint(global_size_scaled[2]));
        r.emplace_back("protocol", jarr({jint(kProtocolMajor), jint(kProtocolMinor)}));
        r.emplace_back("ao isEnabled", jbool(s->ao_enabled != 0));
        r.emplace_back("ao weight", jreal(s->ao_weight));
        const std::string text = jobj(r);
        const int rc = emit(text, msg, capacity, length);
        if(rc == 0)
        {
            // init(): jsonInitRoot = jsonRoot (lib/isaac.hpp:1288-1291)
            session->init = r;
            session->registered = true;
        }
        return rc;
    }

    int isaac_hip_period_message(
        isaac_hip_session* session,
        const isaac_hip_steering_state* s,
        int64_t meta_nr,
        uint32_t send,
        char* msg,
        int64_t capacity,
        int64_t* length)
    {
        if(!session || !valid_state(s))
            return ISAAC_HIP_ERR_ARG;
        if(!session->registered)
            return ISAAC_HIP_ERR_STATE; // init() sends the register message first
        Items r;
        Items init = session->init;
        bool sendInit = (send & ISAAC_HIP_SEND_INIT) != 0;
        auto both = [&](const char* key, const std::string& v)
        {
            r.emplace_back(key, v);
            set(init, key, v);
            sendInit = true;
        };
        r.emplace_back("metadata", metadata(s));
        r.emplace_back("type", jstr("period"));
        r.emplace_back("meta nr", jint(meta_nr));
        if(send & ISAAC_HIP_SEND_PROJECTION)
            both("projection", jreals(s->projection, 16));
        if(send & ISAAC_HIP_SEND_LOOK_AT)
            both("position", jreals(s->look_at, 3));
        if(send & ISAAC_HIP_SEND_ROTATION)
            both("rotation", jreals(s->rotation, 9));
        if(send & ISAAC_HIP_SEND_DISTANCE)
            both("distance", jreal(s->distance));
        if(send & ISAAC_HIP_SEND_TRANSFER)
        {
            if(s->n_sources > 0 && (!s->transfer_lut || !s->transfer_point_counts || s->transfer_size < 0))
                return ISAAC_HIP_ERR_ARG;
            std::vector<std::string> arr, pts;
            int64_t off = 0;
            for(int i = 0; i < s->n_sources; i++)
            {
                std::vector<std::string> lut;
                lut.reserve(size_t(s->transfer_size));
                for(int j = 0; j < s->transfer_size; j++)
                {
                    const float* c = s->transfer_lut + (size_t(i) * size_t(s->transfer_size) + size_t(j)) * 4;
                    // isaac_uint(x * isaac_float(255)) per channel
                    std::vector<std::string> col;
                    for(int k = 0; k < 4; k++)
                        col.push_back(jint((long long) (uint32_t) (c[k] * 255.0f)));
                    lut.push_back(jarr(col));
                }
                arr.push_back(jarr(lut));
                std::vector<std::string> p;
                const int n = s->transfer_point_counts[i];
                if(n < 0 || (n > 0 && (!s->transfer_point_values || !s->transfer_point_rgba)))
                    return ISAAC_HIP_ERR_ARG;
                for(int k = 0; k < n; k++, off++)
                {
                    const float* c = s->transfer_point_rgba + off * 4;
                    p.push_back(jobj({{"value", jint(s->transfer_point_values[off])}, {"r", jreal(c[0])}, {"g", jreal(c[1])},
                                      {"b", jreal(c[2])}, {"a", jreal(c[3])}}));
                }
                pts.push_back(jarr(p));
            }
            r.emplace_back("transfer array", jarr(arr));
            r.emplace_back("transfer points", jarr(pts));
        }
        if(send & ISAAC_HIP_SEND_FUNCTIONS)
        {
            if(s->n_sources > 0 && (!s->functions || !s->function_errors))
                return ISAAC_HIP_ERR_ARG;
            std::vector<std::string> f;
            for(int i = 0; i < s->n_sources; i++)
                f.push_back(jobj({{"source", jstr(s->functions[i])}, {"error", jint(s->function_errors[i])}}));
            r.emplace_back("functions", jarr(f));
        }
        if(send & ISAAC_HIP_SEND_WEIGHT)
            r.emplace_back("weight", jreals(s->weights, s->n_sources));
        if(send & ISAAC_HIP_SEND_INTERPOLATION)
            both("interpolation", jbool(s->interpolation != 0));
        if(send & ISAAC_HIP_SEND_RENDER_OPTIMIZATION)
            both("render optimization", jbool(s->render_optimization != 0));
        if(send & ISAAC_HIP_SEND_STEP)
            both("step", jreal(s->step));
        if(send & ISAAC_HIP_SEND_SEED_POINTS)
            both("seed points", jint(s->seed_points));
        if(send & ISAAC_HIP_SEND_ADVECTION_STEP)
            both("advection step", jreal(s->advection_step));
        if(send & ISAAC_HIP_SEND_ADVECTION_WEIGHT)
            both("advection weight", jreal(s->advection_weight));
        if(send & ISAAC_HIP_SEND_ADVECTION_SEEDING_PERIOD)
            both("advection seeding period", jint(s->advection_seeding_period));
        if(send & ISAAC_HIP_SEND_ADVECTION_SEEDING_DURATION)
            both("advection seeding duration", jint(s->advection_seeding_duration));
        if(send & ISAAC_HIP_SEND_ADVECTION_ON_PAUSE)
            both("advection on pause", jbool(s->advection_on_pause != 0));
        if(send & ISAAC_HIP_SEND_ADVECTION_BORDER)
            both("advection border", jbool(s->advection_border != 0));
        if(send & ISAAC_HIP_SEND_ISO_THRESHOLD)
        {
            if(s->n_volume_field > 0 && !s->iso_thresholds)
                return ISAAC_HIP_ERR_ARG;
            r.emplace_back("iso threshold", jreals(s->iso_thresholds, s->n_volume_field));
        }
        if(send & ISAAC_HIP_SEND_MIN_MAX)
        {
            if(s->n_sources > 0 && !s->min_max)
                return ISAAC_HIP_ERR_ARG;
            std::vector<std::string> mm;
            for(int i = 0; i < s->n_sources; i++)
                mm.push_back(jobj({{"min", jreal(s->min_max[2 * i])}, {"max", jreal(s->min_max[2 * i + 1])}}));
            r.emplace_back("minmax", jarr(mm));
        }
        if(send & ISAAC_HIP_SEND_BACKGROUND)
            both("background color", jreals(s->background, 3));
        if(send & ISAAC_HIP_SEND_CLIPPING)
        {
            if(s->n_clipping > 0 && (!s->clipping_position || !s->clipping_normal))
                return ISAAC_HIP_ERR_ARG;
            std::vector<std::string> c;
            for(int i = 0; i < s->n_clipping; i++)
                c.push_back(jobj({{"position", jreals(s->clipping_position + 3 * i, 3)},
                                  {"normal", jreals(s->clipping_normal + 3 * i, 3)}}));
            r.emplace_back("clipping", jarr(c));
        }
        if(send & ISAAC_HIP_SEND_AO)
        {
            r.emplace_back("ao isEnabled", jbool(s->ao_enabled != 0));
            r.emplace_back("ao weight", jreal(s->ao_weight));
            set(init, "ao isEnabled", jbool(s->ao_enabled != 0));
            set(init, "ao weight", jreal(s->ao_weight));
            sendInit = true;
        }
        if(send & ISAAC_HIP_SEND_RENDER_MODE)
            both("render mode", jint(s->render_mode));
        if(send & ISAAC_HIP_SEND_DITHER_MODE)
            both("dither mode", jint(s->dither_mode));
        // DefaultController::sendFeedback adds nothing (lib/isaac/isaac_controllers.hpp:40-42)
        if(sendInit)
            r.emplace_back("init", jobj(init));
        const int rc = emit(jobj(r), msg, capacity, length);
        if(rc == 0)
            session->init = std::move(init);
        return rc;
    }

    int isaac_hip_exit_message(char* msg, int64_t capacity, int64_t* length)
    {
        // a fresh root: {"type": "exit"} (lib/isaac.hpp:2112-2115)
        return emit(jobj({{"type", jstr("exit")}}), msg, capacity, length);
    }

    int isaac_hip_server_send_bytes(int32_t uid, const char* content, int32_t finishing, char* out, int64_t capacity,
                                    int64_t* length)
    {
        std::string o;
        if(uid >= 0)
            o = "{\"uid\": " + std::to_string(uid);
        if(content)
        {
            // the first l = strlen - 1 bytes go out, the first replaced by ','
            const size_t n = std::strlen(content);
            if(n < 1)
                return ISAAC_HIP_ERR_ARG;
            if(n >= 2)
            {
                o += ',';
                o.append(content + 1, n - 2);
            }
        }
        if(finishing)
            o += "} ";
        return emit(o, out, capacity, length);
    }
}
