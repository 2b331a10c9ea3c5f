/* Exhaustive check of the division used by the gfx950 march kernels:
 * for every binary32 x in [0, 256], RN(x / 255) equals
 *     q0 = RN(x * r); e = fma(-q0, 255, x); q = fma(e, r, q0),   r = RN(1/255).
 * Exit status 0 when all 1,132,462,081 inputs agree. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void)
{
    const float d = 255.0f;
    const float r = 1.0f / 255.0f;
    const float top = 256.0f;
    uint32_t hi;
    memcpy(&hi, &top, 4);
    uint64_t bad = 0, n = 0;
    for(uint32_t u = 0; u <= hi; u++)
    {
        float x;
        memcpy(&x, &u, 4);
        const float q0 = x * r;
        const float e = fmaf(-q0, d, x);
        const float q = fmaf(e, r, q0);
        const float ref = x / d;
        if(memcmp(&q, &ref, 4) != 0)
        {
            if(bad < 5)
                printf("x=%a ref=%a got=%a\n", x, ref, q);
            bad++;
        }
        n++;
    }
    printf("checked %llu inputs, mismatches %llu\n", (unsigned long long) n, (unsigned long long) bad);
    return bad != 0;
}
