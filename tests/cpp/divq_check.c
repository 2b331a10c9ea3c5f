/* Check of the general division used by the gfx950 classification and brick-build
 * kernels: for binary32 x and d > 0 with r = RN(1/d),
 *     q0 = RN(x * r); e = fma(-q0, d, x); q = copysign(fma(e, r, q0), x)
 * equals RN(x / d) bit for bit (for x >= 0 the copysign changes nothing).  Divisors:
 * 64 with extreme mantissas (all ones, lowest bit, half) over 16 binades, the rest
 * log-uniform in [1e-3, 1e3]; per divisor, seeded x uniform in [0, d], uniform in
 * [0, 1], and x = RN(q * d) for q within 2 ulps of a random float in [0, 1] (inputs
 * whose quotient sits next to a rounding boundary), every other one negated, and
 * x = +0 and -0.  Prints the mismatch count; exit status 0 when it is 0. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static uint32_t u_of(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static uint64_t next(uint64_t* s)
{
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

static double unit(uint64_t* s)
{
    return (double) (next(s) >> 11) / 9007199254740992.0;
}

int main(int argc, char** argv)
{
    const int nd = argc > 1 ? atoi(argv[1]) : 2000;
    const int nx = argc > 2 ? atoi(argv[2]) : 100000;
    uint64_t bad = 0, n = 0;
    for(int j = 0; j < nd; j++)
    {
        uint64_t s = 0x2545F4914F6CDD1Dull * (uint64_t) (j + 1);
        float d;
        if(j < 64)
        {
            const uint32_t e = 120u + (uint32_t) (j % 16);
            const uint32_t m = (j & 16) ? 0x7fffffu : ((j & 32) ? 0x000001u : 0x400000u);
            d = f_of((e << 23) | m);
        }
        else
            d = (float) (1e-3 * pow(1e6, unit(&s)));
        const float r = 1.0f / d;
        for(int i = 0; i < nx; i++)
        {
            float x;
            if(i < nx / 4)
                x = (float) (unit(&s) * d);
            else if(i < nx / 2)
                x = (float) unit(&s);
            else
            {
                float q = (float) unit(&s);
                q = f_of(u_of(q) + (uint32_t) (next(&s) % 5) - 2u);
                x = q * d;
            }
            if(i & 1)
                x = -x;
            if(i < 2)
                x = i ? -0.0f : 0.0f;
            const float q0 = x * r;
            const float e = fmaf(-q0, d, x);
            const float q = copysignf(fmaf(e, r, q0), x);
            const float ref = x / d;
            if(u_of(q) != u_of(ref))
            {
                if(bad < 5)
                    printf("x=%a d=%a ref=%a got=%a\n", x, d, ref, q);
                bad++;
            }
            n++;
        }
    }
    printf("divisors %d, inputs %llu: mismatches %llu\n", nd, (unsigned long long) n, (unsigned long long) bad);
    return bad == 0 ? 0 : 1;
}
