// Counter-based random stream of the MPCD step (host and device).
//
// Every draw is a pure function of (seed, system, purpose, step, cell, index),
// so a draw does not depend on which lane, block or launch shape produces it,
// and a recomputation (adjoint replay, checkpoint resume) reproduces it.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11).  Key = (seed low word, seed high word),
// counter = (index, cell, step, purpose<<24 | system).
//
// Float transforms use only IEEE-exact operations (+ - * / sqrt floor and bit
// moves), so the host and the GPU produce identical bits for identical inputs
// when compiled without floating-point contraction.  They replace the
// reference's stateful Xoshiro128+ transforms (common/random.hpp:126-164,
// 237-253) with the same formulas where the reference has one:
//   uniform      u32 * 2^-32                      (random.hpp:126-129)
//   Box-Muller   sqrt(-2 log(u+FLT_MIN)) (cos, sin)(pi t)   (random.hpp:141-164)
//   unit vector  Marsaglia 1972 rejection         (random.hpp:237-253)
#pragma once

#include <stdint.h>

#ifndef __HIP_DEVICE_COMPILE__
#include <cstring>
#endif

#define MPCD_HD __host__ __device__ __forceinline__

#pragma clang fp contract(off)

namespace mpcd_hip {

enum Purpose : uint32_t {
    P_SHIFT = 1,
    P_GHOST_COUNT = 2,
    P_GHOST_POS = 3,
    P_GHOST_VEL = 4,
    P_AXIS = 5,
    P_EXT = 6,
    P_THERM = 7,
    P_INIT_POS = 8,
    P_INIT_VEL = 9,
};

struct Words {
    uint32_t w[4];
};

MPCD_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
    return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}

// full 64-bit product a * b.  On the GPU one v_mad_u64_u32 gives both halves
// (the compiler otherwise issues v_mul_lo_u32 and v_mul_hi_u32, two
// quarter-rate instructions); on the host the plain product.
MPCD_HD uint64_t mul_wide(uint32_t a, uint32_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    uint64_t r;
    uint64_t carry;   // the instruction's carry-out (always clear: the addend is 0)
    asm volatile("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
    (void)carry;
    return r;
#else
    return static_cast<uint64_t>(a) * b;
#endif
}

MPCD_HD Words philox(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t const p0 = mul_wide(0xD2511F53u, c0), p1 = mul_wide(0xCD9E8D57u, c2);
        uint32_t const hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        uint32_t const hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        uint32_t const n0 = hi1 ^ c1 ^ k0;
        uint32_t const n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {{c0, c1, c2, c3}};
}

struct Stream {
    uint32_t k0, k1, sys;
    MPCD_HD Words at(uint32_t purpose, uint32_t index, uint32_t cell, uint32_t step) const {
        return philox(k0, k1, index, cell, step, (purpose << 24) | (sys & 0xffffffu));
    }
};

MPCD_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
MPCD_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// uniform on [0,1] (the endpoint 1 occurs, as in the reference transform)
MPCD_HD float uniform(uint32_t u) { return static_cast<float>(u) * (1.0f / 4294967296.0f); }

// natural log of a positive normal float: x = m 2^e with m in (sqrt(1/2), sqrt(2)],
// log m = 2 atanh((m-1)/(m+1)) summed to the f^9 term; ln 2 split hi/lo.
MPCD_HD float log_pos(float x) {
    uint32_t const b = f2u(x);
    int e = static_cast<int>((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float const f = (m - 1.0f) / (m + 1.0f);
    float const s = f * f;
    float const p = f * (2.0f + s * (0.666666667f + s * (0.4f + s * (0.285714286f + s * 0.222222222f))));
    float const fe = static_cast<float>(e);
    return fe * 0.693145751953125f + (fe * 1.42860682e-06f + p);
}

// sin(pi t) and cos(pi t) for t in [0, 2]: reduce to |r| <= 1/4 by quarter
// turns, then Taylor series of sin and cos on |pi r| <= pi/4.
MPCD_HD void sincospi(float t, float* s_out, float* c_out) {
    float const q = floorf(2.0f * t + 0.5f);
    float const r = t - 0.5f * q;
    float const x = r * 3.14159265f;
    float const x2 = x * x;
    float const s = x + x * x2 * (-0.166666667f + x2 * (0.00833333333f + x2 * (-0.000198412698f + x2 * 2.75573192e-06f)));
    float const c = 1.0f + x2 * (-0.5f + x2 * (0.0416666667f + x2 * (-0.00138888889f + x2 * (2.48015873e-05f + x2 * -2.75573192e-07f))));
    int const qi = static_cast<int>(q) & 3;
    float so, co;
    if (qi == 0) { so = s; co = c; }
    else if (qi == 1) { so = c; co = -s; }
    else if (qi == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *s_out = so;
    *c_out = co;
}

// exp(x): x = k ln2 + r, |r| <= ln2/2, degree-7 Taylor polynomial times 2^k.
MPCD_HD float exp_f(float x) {
    if (x != x) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 89.0f) return __builtin_huge_valf();
    float const k = floorf(x * 1.44269504f + 0.5f);
    float const r = (x - k * 0.693145751953125f) - k * 1.42860682e-06f;
    float const p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f + r * (0.00138888889f + r * 0.000198412698f))))));
    int const ki = static_cast<int>(k);
    if (ki >= -126) {
        if (ki > 127) return __builtin_huge_valf();
        return p * u2f(static_cast<uint32_t>(ki + 127) << 23);
    }
    return (p * u2f(static_cast<uint32_t>(ki + 127 + 64) << 23)) * u2f(static_cast<uint32_t>(127 - 64) << 23);
}

// Box-Muller pair from two words: (cos part, sin part)
MPCD_HD void box_muller(uint32_t a, uint32_t b, float* g0, float* g1) {
    float const u1 = log_pos(uniform(a) + 1.17549435e-38f);
    float const t = static_cast<float>(b) * (2.0f / 4294967296.0f);
    float s, c;
    sincospi(t, &s, &c);
    float const rad = sqrtf(-2.0f * u1);
    *g0 = c * rad;
    *g1 = s * rad;
}

// grid shift of a step: three uniforms minus 1/2 (cpu_backend.cpp:24-26)
MPCD_HD void grid_shift(Stream const& rs, uint32_t step, float* out) {
    Words const r = rs.at(P_SHIFT, 0, 0, step);
    out[0] = uniform(r.w[0]) - 0.5f;
    out[1] = uniform(r.w[1]) - 0.5f;
    out[2] = uniform(r.w[2]) - 0.5f;
}

}  // namespace mpcd_hip
