// Device helpers shared by the encoder kernels (kernels.h) and the parameter-file
// kernels (params_io.h, built in a translation unit of its own): bit reversal, the
// layout of the domain-constant buffer, and [s] G1 from the generator's fixed-base table.
#pragma once
#include "curve.h"

namespace zgr {

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t x, int bits) {
    return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// consts buffer (8 u32 per entry, Fr mont)
enum {
    K_OMEGA = 0,
    K_OMEGA_INV = 1,
    K_NINV = 2,
    K_CF = 3,  // K_CF + c, c < 3
    K_COUNT = 8
};


// [s] G1 for a canonical scalar s, via the fixed-base table
__device__ __forceinline__ g1_xyzz fixed_base_mul(const uint32_t* table, const fr& s) {
    g1_xyzz acc = xyzz_inf();
    for (int w = 0; w < 32; w++) {
        uint32_t d = (s.v[w >> 2] >> (8 * (w & 3))) & 0xffu;
        g1_aff P = aff_load(table + 16 * (size_t)(w * 256 + (d ? d : 1)));
        g1_xyzz nxt = xyzz_madd(acc, P);
        acc = xyzz_select(d != 0, nxt, acc);
    }
    return acc;
}

}  // namespace zgr
