// split_bf16.hpp -- FP32 values as BF16 pairs for the split-BF16x3 products
// (device code only).
#pragma once

#include <hip/hip_runtime.h>

namespace cir {

typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// Two FP32 values as BF16 hi = bf16(x) (round to nearest even) and lo =
// bf16(x - hi), each pair packed in 32 bits (a in the low half).  hi is
// widened back to FP32 from the packed bits (a BF16 value is the top half
// of its FP32 value), so each pair costs two conversions, two bit operations
// and two subtractions.
__device__ __forceinline__ void split_bf16x2(float a, float b, unsigned& hi, unsigned& lo) {
    const bf16x2 h = __builtin_convertvector((float2v){a, b}, bf16x2);
    const unsigned hu = __builtin_bit_cast(unsigned, h);
    const float ha = __uint_as_float(hu << 16), hb = __uint_as_float(hu & 0xffff0000u);
    const bf16x2 l = __builtin_convertvector((float2v){a - ha, b - hb}, bf16x2);
    hi = hu;
    lo = __builtin_bit_cast(unsigned, l);
}

// Four FP32 values split as above, packed 4 to a 64-bit register pair: the
// operand of the 16x16x16 BF16 MFMA, or 4 consecutive k of an 8-k operand.
__device__ __forceinline__ void split_bf16(float a, float b, float c, float d, shortx4& hi, shortx4& lo) {
    unsigned h0, l0, h1, l1;
    split_bf16x2(a, b, h0, l0);
    split_bf16x2(c, d, h1, l1);
    hi = __builtin_bit_cast(shortx4, (unsigned long long)h0 | ((unsigned long long)h1 << 32));
    lo = __builtin_bit_cast(shortx4, (unsigned long long)l0 | ((unsigned long long)l1 << 32));
}

}  // namespace cir
