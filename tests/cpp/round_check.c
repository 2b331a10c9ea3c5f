/* Exhaustive checks of the rounding shortcuts used by the gfx950 classification
 * and transfer-function lookups (isaac_amd/csrc/common.hpp unorm8, tf_index_of):
 *  (1) byte quantisation: for every binary32 v in [0, 1],
 *        x = v * 255;  (uint32_t)(x + h)  ==  (uint32_t) roundf(x),
 *      h = 0x1.fffffep-2 (the largest binary32 below 0.5); and, as a check that the
 *      loop sees the inputs where it matters, x + 0.5 in place of x + h does differ
 *  (2) transfer-function index: for every binary32 x in [-2, T + 2] (T = 1024)
 *      and +-0, +-inf, +-FLT_MAX,
 *        (int) med3(x + h, 0, T - 1)   (med3 as v_med3_f32: NaN -> the minimum of the others)
 *      equals the reference form lv = roundf(x); NaN -> 0; (int) clamp(lv, 0, T - 1),
 *      and, for NaN x, both are 0 (checked on one quiet and one signalling pattern).
 * Prints the mismatch counts; exit status 0 when both are 0. */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float f_of(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static float gmin(float a, float b) { return (b < a) ? b : a; }
static float gmax(float a, float b) { return (a < b) ? b : a; }
static float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

static int tf_ref(float x, int T)
{
    const float lv = roundf(x);
    if(lv != lv)
        return 0;
    return (int) gclamp(lv, 0.0f, (float) (T - 1));
}

/* v_med3_f32: the median, or, when an input is NaN, the minimum of the others */
static float hw_med3(float a, float b, float c)
{
    if(a != a)
        return fminf(b, c);
    return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
}
static int tf_new(float x, int T)
{
    return (int) hw_med3(x + 0x1.fffffep-2f, 0.0f, (float) (T - 1));
}

static uint64_t check_tf(float x, int T, uint64_t* n)
{
    (*n)++;
    const int a = tf_ref(x, T), b = tf_new(x, T);
    if(a != b)
    {
        printf("tf_index T=%d x=%a ref=%d new=%d\n", T, x, a, b);
        return 1;
    }
    return 0;
}

int main(void)
{
    uint64_t bad8 = 0, n8 = 0, half8 = 0;
    const uint32_t one = 0x3f800000u;
    for(uint32_t u = 0; u <= one; u++)
    {
        const float x = f_of(u) * 255.0f;
        const uint32_t a = (uint32_t) roundf(x);
        const uint32_t b = (uint32_t) (x + 0x1.fffffep-2f);
        if((uint32_t) (x + 0.5f) != a)
            half8++;
        n8++;
        if(a != b)
        {
            if(bad8 < 5)
                printf("unorm8 v=%a ref=%u new=%u\n", f_of(u), a, b);
            bad8++;
        }
    }
    uint64_t badt = 0, nt = 0;
    const int T = 1024;
    const float hi = (float) T + 2.0f;
    uint32_t uhi;
    memcpy(&uhi, &hi, 4);
    for(uint32_t u = 0; u <= uhi; u++)
        badt += check_tf(f_of(u), T, &nt);
    const float lo = -2.0f;
    uint32_t ulo;
    memcpy(&ulo, &lo, 4);
    for(uint32_t u = 0x80000000u; u <= ulo; u++)
        badt += check_tf(f_of(u), T, &nt);
    const float special[] = {INFINITY, -INFINITY, FLT_MAX, -FLT_MAX, 0.0f, -0.0f};
    for(int i = 0; i < 6; i++)
        badt += check_tf(special[i], T, &nt);
    if(tf_new(f_of(0x7fc00000u), T) != 0 || tf_new(f_of(0x7fa00000u), T) != 0)
        badt++;
    printf("unorm8 with x + 0.5 in place of x + h: %llu inputs differ from roundf\n", (unsigned long long) half8);
    printf("unorm8: %llu inputs, mismatches %llu; tf_index: %llu inputs, mismatches %llu\n", (unsigned long long) n8,
           (unsigned long long) bad8, (unsigned long long) nt, (unsigned long long) badt);
    return (bad8 == 0 && badt == 0 && half8 > 0) ? 0 : 1;
}
