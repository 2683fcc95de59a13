// pr_codec.hpp -- PageRank contribution formats (stored value <-> float).
//
// FP32 (PRv, accumulated in FP64 as PRv.C:127 does), bfloat16 with
// round-to-nearest-even, bfloat16 by truncation (scustomfp<true,8,7>),
// IEEE half with a per-iteration power-of-two scale, customfp<6,10>
// (customfp.h:95-117 encode, :141-169 decode) and the signed
// scustomfp<true,6,9,true,0> of the delta PageRank (APR_vary.C:743-750).
#pragma once
#include "internal.hpp"

namespace gpr {

template <int FMT> struct Codec;
template <> struct Codec<GPR_CONTRIB_F32> {
    using store_t = float;
    using acc_t = double;
    __device__ static float dec(float v) { return v; }
    __device__ static float enc(float v) { return v; }
};
// bfloat16, round to nearest even
template <> struct Codec<GPR_CONTRIB_BF16> {
    using store_t = uint16_t;
    using acc_t = float;
    __device__ static float dec(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
    __device__ static uint16_t enc(float f) {
        uint32_t u = __float_as_uint(f);
        // a NaN stays a (quiet) NaN instead of rounding into the infinity
        if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
};
// scustomfp<true,8,7,false,0>: the upper half of the float, truncated
// (customfp.h:95-117 with S=true, E=8, M=7, B=0)
template <> struct Codec<GPR_CONTRIB_BF16T> {
    using store_t = uint16_t;
    using acc_t = float;
    __device__ static float dec(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
    __device__ static uint16_t enc(float f) { return (uint16_t)(__float_as_uint(f) >> 16); }
};
// IEEE half (v_cvt_f16_f32, round to nearest even), values pre-scaled by
// 2^k so that the largest contribution sits near the top of the range
template <> struct Codec<GPR_CONTRIB_F16> {
    using store_t = uint16_t;
    using acc_t = float;
    __device__ static float dec(uint16_t h) {
        _Float16 x;
        __builtin_memcpy(&x, &h, 2);
        return (float)x;
    }
    __device__ static uint16_t enc(float f) {
        _Float16 x = (_Float16)f;
        uint16_t h;
        __builtin_memcpy(&h, &x, 2);
        return h;
    }
};
// customfp<6,10>: no sign, the low 6 exponent bits and top 10 mantissa bits,
// truncated; decode restores exponent bits 01 above them, so the format
// covers [2^-63, 2) (customfp.h:95-117 encode, :141-169 decode)
template <> struct Codec<GPR_CONTRIB_CFP16> {
    using store_t = uint16_t;
    using acc_t = float;
    __device__ static float dec(uint16_t h) { return __uint_as_float(((uint32_t)h | 0x10000u) << 13); }
    __device__ static uint16_t enc(float f) { return (uint16_t)(__float_as_uint(f) >> 13); }
};
// scustomfp<true,6,9,true,0>: sign, the low 6 exponent bits and the top 9
// mantissa bits, truncated (customfp.h:95-117; the vector narrowing of
// target/conversion.h:467-497 gives the same bits).  Decode puts exponent
// bits 01 above the six, so magnitudes cover [2^-63, 2), and code 0 decodes
// to 0 as the vector widening does for a format that may hold zero
// (target/conversion.h:150-189).
template <> struct Codec<GPR_CONTRIB_SCFP16> {
    using store_t = uint16_t;
    using acc_t = float;
    __device__ static float dec(uint16_t h) {
        const uint32_t t = ((uint32_t)(h & 0x8000u) << 16) | 0x20000000u | ((uint32_t)(h & 0x7fffu) << 14);
        return h ? __uint_as_float(t) : 0.0f;
    }
    __device__ static uint16_t enc(float f) {
        const uint32_t u = __float_as_uint(f);
        return (uint16_t)(((u >> 14) & 0x7fffu) | ((u >> 16) & 0x8000u));
    }
};

} // namespace gpr
