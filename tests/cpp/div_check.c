/* Correct rounding of the one-correction quotient used by the anisotropic-kernel walks (div_h in
 * fltk_sph_ice_amd/csrc/aniso.hip): q = fma(fma(-q0, h, r), y, q0) with y = RN(1 / h) and q0 = RN(r y), compared
 * bit for bit with the IEEE quotient r / h.  Usage: div_check <mode> h...  (mode 1: the one-correction quotient,
 * mode 0: q0 alone, the uncorrected product).  For each h every float r in [h / 4, 2h] is checked, and every
 * 4099th float in [2^-75, h / 4).  Prints the mismatch count per h; exits 1 when any h has one. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float fu(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t uf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static long check(float h, int mode, uint32_t lo, uint32_t hi, uint32_t stride) {
    volatile float hv = h;
    const float y = 1.0f / hv;
    long bad = 0;
    for (uint64_t u = lo; u <= hi; u += stride) {
        const float r = fu((uint32_t)u);
        volatile float ref = r / hv;
        const float q0 = r * y;
        const float q = mode ? fmaf(fmaf(-q0, h, r), y, q0) : q0;
        if (uf(q) != uf(ref)) ++bad;
    }
    return bad;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const int mode = atoi(argv[1]);
    int any = 0;
    for (int a = 2; a < argc; ++a) {
        const float h = strtof(argv[a], NULL);
        const uint32_t q4 = uf(0.25f * h), top = uf(2.0f * h), tiny = uf(ldexpf(1.0f, -75));
        const long bad = check(h, mode, q4, top, 1) + check(h, mode, tiny, q4 - 1, 4099);
        printf("%.9g %ld\n", h, bad);
        any |= bad != 0;
    }
    return any;
}
