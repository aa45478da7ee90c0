// ll_tables.cpp -- host-side size rules and transform tables.
//
//  * fft_length_conservative: lucaslehmer.cpp:244-254 (smallest {1,3,5} * 16 * 2^k length
//    with p/N <= max bits per word).
//  * ntt_length_ref / ntt_max_bits: cuda/ll_ntt.cu:172-180 (N * 2^(2B-1) < 2^62).
//  * Goldilocks roots: cuda/ll_ntt.cu:182-195 (generator 7, N-th root of 2 through the
//    192N-th root of unity); the root of unity is additionally aligned so that
//    w_N^(N/16) == 2^12, the constant the radix-16 butterflies hard-wire.
//  * FP64 twiddles and weights are evaluated in long double with an exact octant
//    reduction of j/n and rounded once to double.
#include "ll_internal.hpp"
#include <cmath>
#include <mutex>

namespace llh {

static const u64 Q = 0xFFFFFFFF00000001ull;
typedef unsigned __int128 u128;

u64 gl_mul_h(u64 a, u64 b) { return (u64)(((u128)a * b) % Q); }
u64 gl_pow_h(u64 a, u64 e) {
    u64 r = 1;
    a %= Q;
    while (e) {
        if (e & 1) r = gl_mul_h(r, a);
        a = gl_mul_h(a, a);
        e >>= 1;
    }
    return r;
}
u64 gl_inv_h(u64 a) { return gl_pow_h(a, Q - 2); }

u64 gl_root2_h(u64 n) {
    if ((Q - 1) % (192 * n)) return 0;
    u64 om = gl_pow_h(7, (Q - 1) / (192 * n));
    u64 r192 = gl_pow_h(om, n), pw = 1, s = 0;
    for (; s < 192; ++s) {
        if (pw == 2) break;
        pw = gl_mul_h(pw, r192);
    }
    if (s == 192) return 0;
    return gl_pow_h(om, s);
}

u64 gl_root_unity_aligned(u64 n) {
    if (n <= 16) return gl_pow_h(2, 192 / n);          // 2^(192/n) has order n
    u64 w = gl_pow_h(7, (Q - 1) / n);                    // some primitive n-th root
    u64 t = gl_pow_h(w, n / 16);                         // a primitive 16th root = 2^(12u), u odd
    for (u64 u = 1; u < 16; u += 2) {
        if (gl_pow_h(2, 12 * u) == t) {
            // w' = w^(u^-1 mod n)  =>  w'^(n/16) = 2^12
            u64 inv = 1;
            for (u64 c = 1; c < n; c += 2)
                if ((u128)c * u % n == 1) { inv = c; break; }
            return gl_pow_h(w, inv);
        }
    }
    return 0;
}

u64 fft_length_conservative(u64 p, double max_bpw) {
    u64 best = 0;
    const u64 ms[3] = {1, 3, 5};
    for (u64 m : ms)
        for (u64 N = m * 16; N < ((u64)1 << 40); N <<= 1)
            if ((double)p / (double)N <= max_bpw) {
                if (!best || N < best) best = N;
                break;
            }
    return best;
}

int ntt_max_bits(u64 n) {
    int l = 0;
    while (((u64)1 << l) < n) ++l;
    return (63 - l) / 2;
}

u64 ntt_length_ref(u64 p) {
    for (u64 N = 64; N <= 8192; N <<= 1)
        if ((p + N - 1) / N <= (u64)ntt_max_bits(N)) return N;
    return 0;
}

}  // namespace llh
