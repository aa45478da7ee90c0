// ll_common.hpp -- device arithmetic shared by every Lucas-Lehmer kernel (gfx950).
//
// Two element types run through the same transform code:
//   * C64: an FP64 complex number (the real FFT of N weighted digits is run as a complex
//     FFT of N/2 packed pairs, see ll_batch.hip);
//   * GL:  an element of the Goldilocks field GF(q), q = 2^64 - 2^32 + 1 (the exact
//     integer transform of cuda/ll_ntt.cu:38-61), built here from 32-bit mul_lo / mul_hi
//     pieces with the reductions 2^64 == 2^32 - 1 and 2^96 == -1 (mod q).
//
// Roots of unity: both types use w16 = the primitive 16th root of unity that the DFT-16
// butterflies hard-wire.  For C64 that is exp(-2 pi i / 16) (forward).  For GL it is
// 2^12 (2 has multiplicative order 192 in GF(q), so 2^12 has order 16), which turns every
// twiddle inside a radix-16 butterfly into a shift.  The host picks the N-th root of
// unity for the outer twiddles so that its (N/16)-th power is exactly 2^12.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef int64_t i64;
typedef uint32_t u32;
typedef int32_t i32;

#define LL_DEV __device__ __forceinline__
#define LL_HD __host__ __device__ __forceinline__

// Digit layout of a length-N IBDWT of 2^p - 1, for any N: digit d has b_d = c_{d+1} - c_d
// bits, c_d = ceil(d p / N), and weight exponent e_d = c_d N - d p.  Both follow from the
// residue r_d = d p mod N alone: with P = p mod N and F = floor(p / N),
//   b_d = F + [r_d + P >= N] + [r_{d+1} > 0] - [r_d > 0],   e_d = (N - r_d) mod N,
//   r_{d+1} = (r_d + P) mod N,   r_{d+D} = (r_d + (D p mod N)) mod N,
// so walking digits needs no division.
struct DigitWalk {
    u32 n, P, F;
    LL_HD u32 add(u32 r, u32 delta) const {          // (r + delta) mod N, r, delta < N
        const u32 v = r + delta;
        return v >= n ? v - n : v;
    }
    LL_HD u32 next(u32 r) const { return add(r, P); }
    LL_HD int bits(u32 r) const {
        const u32 rn = next(r);
        return (int)(F + (r + P >= n ? 1u : 0u) + (rn > 0 ? 1u : 0u)) - (r > 0 ? 1 : 0);
    }
    LL_HD u32 wexp(u32 r) const { return r ? n - r : 0u; }
};

// The same layout walked by weight exponent, for p coprime to N (a prime p > 7): digit d has
// e_d = (-d p) mod N and b_d = F + [e_d < P] bits (the closed form of the ceiling definition
// above), and e_{d+1} = (e_d - P) mod N, so each digit costs one compare and one modular
// subtraction.
struct ExpWalk {
    u32 n, P, F;
    // (e - delta) mod N for e, delta < N: with e < delta the difference wraps to at least
    // 2^32 - N and adding N gives the residue; otherwise the difference is below N
    LL_HD u32 sub(u32 e, u32 delta) const {
        const u32 d = e - delta, w = d + n;
        return w < d ? w : d;
    }
    LL_HD u32 next(u32 e) const { return sub(e, P); }
    LL_HD int bits(u32 e) const { return (int)(F + (e < P ? 1u : 0u)); }
    LL_HD u32 of_residue(u32 r) const { return r ? n - r : 0u; }   // e_d from r = d p mod N
};

// Returns v unchanged, but opaque to the optimiser.  Index math derived from it is then
// recomputed where it is used instead of being hoisted out of the squaring loop and kept
// live in registers for its whole length (twiddle and weight addresses are loop-invariant,
// and holding them all would cost more VGPRs than the residue itself).
LL_DEV int ll_opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
LL_DEV u32 ll_opaque(u32 v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Global-memory pointers with their address space spelled out, so loads through them are
// global_load (32-bit offsets from a uniform base) rather than flat_load of 64-bit addresses.
#define LL_GLOBAL __attribute__((address_space(1)))
template <class T>
using gptr = const T LL_GLOBAL *;
template <class T>
LL_DEV gptr<T> as_global(const void *p) {
    return (gptr<T>)p;
}

// Load element `idx` of a read-only table.  The address is the base plus a 32-bit byte
// offset, which keeps a uniform base in SGPRs (global_load vaddr, saddr) with no 64-bit
// address arithmetic per lane; tables are far below 4 GB.
template <class T>
LL_DEV T ld_table(gptr<T> base, u32 idx) {
    return *(gptr<T>)((const char LL_GLOBAL *)base + idx * (u32)sizeof(T));
}

// ------------------------------------------------------------------------------------------
// FP64 complex
// ------------------------------------------------------------------------------------------
struct C64 {
    double re, im;
};
LL_HD C64 cmk(double r, double i) { C64 c; c.re = r; c.im = i; return c; }
LL_DEV C64 cadd(C64 a, C64 b) { return cmk(a.re + b.re, a.im + b.im); }
LL_DEV C64 csub(C64 a, C64 b) { return cmk(a.re - b.re, a.im - b.im); }
LL_DEV C64 cmul(C64 a, C64 b) {
    return cmk(__fma_rn(a.re, b.re, -a.im * b.im), __fma_rn(a.re, b.im, a.im * b.re));
}
// a * conj(b)
LL_DEV C64 cmulc(C64 a, C64 b) {
    return cmk(__fma_rn(a.re, b.re, a.im * b.im), __fma_rn(a.im, b.re, -a.re * b.im));
}
LL_DEV C64 csqr(C64 a) { return cmk(__fma_rn(a.re, a.re, -a.im * a.im), 2.0 * a.re * a.im); }
LL_DEV C64 cconj(C64 a) { return cmk(a.re, -a.im); }
LL_DEV C64 ld_table(gptr<C64> base, u32 idx) {
    gptr<double> d = (gptr<double>)((const char LL_GLOBAL *)base + idx * 16u);
    return cmk(d[0], d[1]);
}

// ------------------------------------------------------------------------------------------
// Goldilocks field, q = 2^64 - 2^32 + 1.  Values are kept fully reduced in [0, q).
// ------------------------------------------------------------------------------------------
#define GL_Q 0xFFFFFFFF00000001ull
#define GL_EPS 0xFFFFFFFFull   // 2^64 mod q

LL_DEV u64 gl_add(u64 a, u64 b) {
    u64 s = a + b;
    u64 t = s + GL_EPS;                 // s + 2^64 - q
    // overflow (s < a): true sum = s + 2^64 == s + EPS (mod q), and s + EPS < q here
    // no overflow: subtract q when s >= q, i.e. add EPS with wrap
    bool ovf = s < a;
    bool ge = s >= GL_Q;
    return (ovf || ge) ? t : s;
}
LL_DEV u64 gl_sub(u64 a, u64 b) {
    u64 d = a - b;
    return (a < b) ? d - GL_EPS : d;    // borrow: d + q = d - EPS (mod 2^64)
}
LL_DEV u64 gl_neg(u64 a) { return a ? GL_Q - a : 0; }

// Reduce a 128-bit value (x3:x2:x1:x0 in 32-bit limbs, lo = x1:x0) modulo q:
//   x = lo + x2 * 2^64 + x3 * 2^96 == lo + x2 * (2^32 - 1) - x3.
LL_DEV u64 gl_reduce128(u64 lo, u64 hi) {
    u32 x2 = (u32)hi, x3 = (u32)(hi >> 32);
    u64 t0 = lo - (u64)x3;
    if (lo < (u64)x3) t0 -= GL_EPS;                 // borrow: add q
    u64 t1 = ((u64)x2 << 32) - (u64)x2;             // x2 * (2^32 - 1) < q
    u64 r = t0 + t1;
    if (r < t0) r += GL_EPS;                        // carry out of 2^64
    if (r >= GL_Q) r -= GL_Q;
    return r;
}

// 64x64 -> 128 product from four 32x32 -> 64 products (v_mad_u64_u32 / mul_lo / mul_hi).
LL_DEV void mul64x64(u64 a, u64 b, u64 &lo, u64 &hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 p01 = (u64)a0 * b1;
    u64 p10 = (u64)a1 * b0;
    u64 p11 = (u64)a1 * b1;
    u64 mid = (p00 >> 32) + (u32)p01 + (u32)p10;          // < 3 * 2^32
    lo = (mid << 32) | (u32)p00;
    hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

LL_DEV u64 gl_mul(u64 a, u64 b) {
    u64 lo, hi;
    mul64x64(a, b, lo, hi);
    return gl_reduce128(lo, hi);
}
LL_DEV u64 gl_sqr(u64 a) { return gl_mul(a, a); }

// a * 2^k mod q for a compile-time 0 <= k < 192 (2^96 == -1, so k >= 96 negates).
template <int K>
LL_DEV u64 gl_mul_pow2(u64 a) {
    constexpr int k0 = ((K % 192) + 192) % 192;
    constexpr bool neg = k0 >= 96;
    constexpr int k = neg ? k0 - 96 : k0;
    u64 r;
    if constexpr (k == 0) {
        r = a;
    } else if constexpr (k < 64) {
        u64 lo = a << k, hi = a >> (64 - k);
        r = gl_reduce128(lo, hi);
    } else {
        // a * 2^k = (a * 2^(k-32)) * 2^32, both shifts < 64
        u64 t = gl_reduce128(a << (k - 32), a >> (96 - k));
        r = gl_reduce128(t << 32, t >> 32);
    }
    if constexpr (neg) r = gl_neg(r);
    return r;
}

// Signed digit <-> field element.
LL_DEV u64 gl_from_i64(i64 x) { return x < 0 ? GL_Q - (u64)(-x) : (u64)x; }
LL_DEV i64 gl_to_i64(u64 v) { return (v > (GL_Q >> 1)) ? -(i64)(GL_Q - v) : (i64)v; }

// ------------------------------------------------------------------------------------------
// Element traits: add/sub, multiplication by a table twiddle, and by w16^e (compile time).
// INV selects the inverse transform (conjugate / inverse roots).
// ------------------------------------------------------------------------------------------
struct FC64 {
    typedef C64 T;
    typedef C64 W;                     // twiddle type
    static LL_DEV T add(T a, T b) { return cadd(a, b); }
    static LL_DEV T sub(T a, T b) { return csub(a, b); }
    template <bool INV>
    static LL_DEV T twiddle(T a, W w) { return INV ? cmulc(a, w) : cmul(a, w); }
    // multiply by w16^E, w16 = exp(-2 pi i/16) (forward) or its conjugate (inverse)
    template <int E, bool INV>
    static LL_DEV T mul_w16(T a) {
        constexpr int e0 = ((E % 16) + 16) % 16;
        constexpr int e = INV ? (16 - e0) % 16 : e0;
        constexpr double C1 = 0.92387953251128673848259;   // cos(pi/8)
        constexpr double S1 = 0.38268343236508978177923;   // sin(pi/8)
        constexpr double H = 0.70710678118654752440084;    // sqrt(1/2)
        // exp(-2 pi i e/16) = cos(2 pi e/16) - i sin(2 pi e/16)
        if constexpr (e == 0) return a;
        else if constexpr (e == 4) return cmk(a.im, -a.re);           // * -i
        else if constexpr (e == 8) return cmk(-a.re, -a.im);          // * -1
        else if constexpr (e == 12) return cmk(-a.im, a.re);          // * +i
        else if constexpr (e == 2) return cmk(H * (a.re + a.im), H * (a.im - a.re));
        else if constexpr (e == 6) return cmk(H * (a.im - a.re), -H * (a.re + a.im));
        else if constexpr (e == 10) return cmk(-H * (a.re + a.im), H * (a.re - a.im));
        else if constexpr (e == 14) return cmk(H * (a.re - a.im), H * (a.re + a.im));
        else {
            constexpr double cs[16] = {1, C1, H, S1, 0, -S1, -H, -C1, -1, -C1, -H, -S1, 0, S1, H, C1};
            constexpr double sn[16] = {0, S1, H, C1, 1, C1, H, S1, 0, -S1, -H, -C1, -1, -C1, -H, -S1};
            // (a.re + i a.im) (c - i s)
            constexpr double c = cs[e], s = sn[e];
            return cmk(__fma_rn(a.re, c, a.im * s), __fma_rn(a.im, c, -a.re * s));
        }
    }
};

struct FGL {
    typedef u64 T;
    typedef u64 W;
    static LL_DEV T add(T a, T b) { return gl_add(a, b); }
    static LL_DEV T sub(T a, T b) { return gl_sub(a, b); }
    template <bool INV>
    static LL_DEV T twiddle(T a, W w) { return gl_mul(a, w); }   // host stores inverse twiddles
    // w16 = 2^12 forward, 2^-12 = 2^180 inverse
    template <int E, bool INV>
    static LL_DEV T mul_w16(T a) {
        constexpr int e0 = ((E % 16) + 16) % 16;
        constexpr int k = INV ? (192 - 12 * e0) % 192 : (12 * e0) % 192;
        return gl_mul_pow2<k>(a);
    }
};

// ---- exact re-layout of a residue from N digits to N' digits (checkpoint rollback) -------------
// `dig` holds the N canonical digits of X in [0, 2^p): 0 <= dig[d] < 2^(c_{d+1} - c_d), with
// c_d = ceil(d p / N) the position of digit d's lowest bit.  Digit k of the N'-digit form takes
// bits [c'_k, c'_{k+1}) of X and is balanced by moving a carry of one into digit k + 1 when it is
// at least half its range; the top digit's carry goes to digit 0 (2^p == 1).  Every digit of the
// result is therefore in [-2^(b-1), 2^(b-1)], computed on its own (no carry chain).  The same
// code runs on the device (the large engine's lg_relayout, the batch engine's bt_resume) and on
// the host (llh_relayout_host, for tests).
LL_HD u64 cpos(u64 d, u64 p, u64 n) { return (d * p + n - 1) / n; }   // ceil(d p / n), d p < 2^53

// bits [lo, hi) of X (hi - lo < 64)
LL_HD u64 bits_at(const i64 *dig, u64 p, u64 n, u64 lo, u64 hi) {
    u64 v = 0;
    u64 d = (lo * n) / p;               // c_d <= lo < c_{d+1}
    u64 cd = cpos(d, p, n);
    for (u64 a = lo; a < hi; ++d) {
        const u64 cn = cpos(d + 1, p, n);
        const u64 b = hi < cn ? hi : cn;
        const u64 piece = ((u64)dig[d] >> (a - cd)) & ((1ull << (b - a)) - 1);
        v |= piece << (a - lo);
        a = b;
        cd = cn;
    }
    return v;
}

LL_HD i64 relayout_digit(const i64 *dig, u64 p, u64 n, u64 n2, u64 k) {
    const u64 km = k ? k - 1 : n2 - 1;
    const u64 lo = cpos(k, p, n2), hi = cpos(k + 1, p, n2);
    const u64 lo2 = cpos(km, p, n2), hi2 = cpos(km + 1, p, n2);
    const int w = (int)(hi - lo), w2 = (int)(hi2 - lo2);
    const u64 v = bits_at(dig, p, n, lo, hi), v2 = bits_at(dig, p, n, lo2, hi2);
    const i64 c = (w > 0 && v >= (1ull << (w - 1))) ? 1 : 0;
    const i64 c2 = (w2 > 0 && v2 >= (1ull << (w2 - 1))) ? 1 : 0;
    return (i64)v - c * ((i64)1 << w) + c2;
}
