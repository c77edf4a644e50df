// TEST INFRASTRUCTURE: exposes the renderer's libm-matching sinf, cosf,
// asinf and acosf (psychopath_amd/csrc/libm_f32.h) built for the host, and
// counts the inputs on which they differ from the host C library's.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../psychopath_amd/csrc/libm_f32.h"

extern "C" {

float mc_asinf(float x) { return psy::psy_asinf(x); }
float mc_acosf(float x) { return psy::psy_acosf(x); }
float mc_sinf(float x) { return psy::psy_sinf(x); }
float mc_cosf(float x) { return psy::psy_cosf(x); }

// As mc_mismatches, for sinf and cosf.
uint64_t mc_mismatches_sincos(uint32_t lo, uint32_t hi, uint32_t step, uint32_t* out) {
    uint64_t bad = 0;
    out[0] = out[1] = 0xffffffffu;
    for (uint64_t u = lo; u <= hi; u += step) {
        for (uint32_t sgn = 0; sgn < 2; sgn++) {
            const uint32_t bits = (uint32_t)u | (sgn ? 0x80000000u : 0u);
            float x;
            memcpy(&x, &bits, 4);
            const float r[4] = {sinf(x), psy::psy_sinf(x), cosf(x), psy::psy_cosf(x)};
            uint32_t b[4];
            memcpy(b, r, 16);
            if (b[0] != b[1] && !(r[0] != r[0] && r[1] != r[1])) {
                if (out[0] == 0xffffffffu) out[0] = bits;
                bad++;
            }
            if (b[2] != b[3] && !(r[2] != r[2] && r[3] != r[3])) {
                if (out[1] == 0xffffffffu) out[1] = bits;
                bad++;
            }
        }
    }
    return bad;
}

// Inputs with bit patterns lo, lo+step, ... <= hi, each with both signs
// (any NaN matches any NaN).
// Returns the mismatch count; out[0], out[1] receive the first mismatching
// input's bits for asinf and acosf (0xffffffff if none).
uint64_t mc_mismatches(uint32_t lo, uint32_t hi, uint32_t step, uint32_t* out) {
    uint64_t bad = 0;
    out[0] = out[1] = 0xffffffffu;
    for (uint64_t u = lo; u <= hi; u += step) {
        for (uint32_t sgn = 0; sgn < 2; sgn++) {
            const uint32_t bits = (uint32_t)u | (sgn ? 0x80000000u : 0u);
            float x;
            memcpy(&x, &bits, 4);
            const float a = asinf(x), b = psy::psy_asinf(x), c = acosf(x), d = psy::psy_acosf(x);
            uint32_t ua, ub, uc, ud;
            memcpy(&ua, &a, 4);
            memcpy(&ub, &b, 4);
            memcpy(&uc, &c, 4);
            memcpy(&ud, &d, 4);
            if (ua != ub && !(a != a && b != b)) {
                if (out[0] == 0xffffffffu) out[0] = bits;
                bad++;
            }
            if (uc != ud && !(c != c && d != d)) {
                if (out[1] == 0xffffffffu) out[1] = bits;
                bad++;
            }
        }
    }
    return bad;
}
}
