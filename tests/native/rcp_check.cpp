// Compares the RCPSS emulation used by the kernels (x86_fp.h) with this
// CPU's own RCPSS instruction, using the table of the model named by the
// first argument ("intel" or "amd").  Exit status 0 when every tested input
// agrees bit for bit; prints the first few disagreements otherwise.
//
// Inputs: every mantissa for a set of exponents near the ends and the middle
// of the range, both signs; 4096 pseudo-random mantissas for every exponent;
// zeros, sub-normals, infinities and NaNs.
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "x86_fp.h"

static uint32_t live(uint32_t b) {
    float x;
    memcpy(&x, &b, 4);
    const float r = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
    uint32_t o;
    memcpy(&o, &r, 4);
    return o;
}

static int model = ifx::kRcpIntel;

static uint32_t emul(uint32_t b) {
    float x;
    memcpy(&x, &b, 4);
    return ifx::f2u(ifx::rcp_ps(x, model));
}

static long bad = 0, checked = 0;
static void check(uint32_t b) {
    const uint32_t a = live(b), e = emul(b);
    checked++;
    const bool nan = ((a >> 23) & 0xff) == 0xff && (a & 0x7fffff) && ((e >> 23) & 0xff) == 0xff && (e & 0x7fffff);
    if (a != e && !nan) {
        if (bad < 8) printf("input %08x: cpu %08x emulation %08x\n", b, a, e);
        bad++;
    }
}

int main(int argc, char** argv) {
    if (argc > 1 && strcmp(argv[1], "amd") == 0) model = ifx::kRcpAmd;
    printf("model %s\n", model == ifx::kRcpAmd ? "amd" : "intel");
    const uint32_t exps[] = {1, 2, 3, 64, 100, 125, 126, 127, 128, 129, 200, 250, 251, 252, 253, 254};
    for (uint32_t s = 0; s < 2; s++)
        for (uint32_t e : exps)
            for (uint32_t m = 0; m < (1u << 23); m++) check((s << 31) | (e << 23) | m);
    uint32_t x = 0x12345678u;
    for (uint32_t s = 0; s < 2; s++)
        for (uint32_t e = 0; e < 256; e++)
            for (int i = 0; i < 4096; i++) {
                x ^= x << 13;
                x ^= x >> 17;
                x ^= x << 5;
                check((s << 31) | (e << 23) | (x & 0x7fffffu));
            }
    const uint32_t specials[] = {0x00000000u, 0x80000000u, 0x00000001u, 0x807fffffu,
                                 0x7f800000u, 0xff800000u, 0x7fc00000u};
    for (uint32_t b : specials) check(b);
    printf("checked %ld inputs, %ld disagreements\n", checked, bad);
    return bad ? 1 : 0;
}
