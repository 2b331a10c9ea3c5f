// isaac_json_stub.h — no-op stand-ins for the jansson calls an ISAAC
// application makes around the render call, for builds without jansson.
//
// The reference's documented pattern (lib/isaac/mainpage.dox:105-117,
// example/example.cpp:429-437,567) is
//
//   json_object_set_new(vis->getJsonMetaRoot(), "kernel_time", json_integer(t));
//   json_decref(vis->doVisualization(META_MASTER));
//
// With the drop-in (isaac_mi355x.hpp) both roots are null json_t*; link
// jansson and these calls behave as jansson defines them on null, or include
// this header instead of jansson.h and they do nothing.  Do not include both.
#pragma once

#include <cstdint>

struct json_t;
typedef struct json_t json_t;

static inline void json_decref(json_t* json)
{
    (void) json;
}
static inline int json_object_set_new(json_t* object, const char* key, json_t* value)
{
    (void) object;
    (void) key;
    (void) value;
    return -1;
}
static inline json_t* json_string(const char* value)
{
    (void) value;
    return nullptr;
}
static inline json_t* json_integer(long long value)
{
    (void) value;
    return nullptr;
}
static inline json_t* json_real(double value)
{
    (void) value;
    return nullptr;
}
