// ll_dft.hpp -- radix-2/4/8/16 butterflies and one register-blocked Stockham pass.
//
// A length-M transform is run by G threads that each hold R = M/G elements in registers.
// Pass i has radix r_i (r_i divides R); a thread runs R/r_i butterflies per pass.  The
// Stockham ("autosort") formulation keeps the output in natural order:
//
//   pass (radix r, L = product of earlier radices), butterfly index i in [0, M/r):
//     k = i mod L;  x_s = A[i + s M/r] * w_{rL}^{s k};  y = DFT_r(x);
//     B[(i - k) r + k + q L] = y_q
//
// With r_first = r_last = R, thread t's inputs to the first pass are elements
// t + G s (s < R) and its outputs of the last pass are elements t + G q: both are the
// same "column" layout, so the digits stay in registers between iterations and only the
// middle hand-offs go through LDS.  LDS indices are padded by one element per 2^PS so the
// stride-2^PS stores of a radix-2^PS pass do not pile onto one bank.
#pragma once
#include <type_traits>
#include "ll_common.hpp"

// One pad element per 2^PS: the stride-2^PS stores of a radix-2^PS first pass land on distinct
// banks for every 8 consecutive lanes (16-byte stores are served 8 lanes at a time, 128 bytes).
template <int PS>
LL_HD constexpr int lds_pad_s(int i) { return i + (i >> PS); }

// Twiddles of a transform with radices R1 .. R5 (0 = no such pass) in per-pass order: pass p >= 2
// (L = R1 ... R_{p-1}) keeps w_M^{s k M / (R_p L)}, s in [1, R_p), k in [0, L), at
// O_p + (s - 1) L + k, so the lanes of a pass (consecutive k) read consecutive entries.
template <int R1, int R2, int R3, int R4, int R5>
struct PassTable {
    static constexpr int L2 = R1;
    static constexpr int L3 = L2 * (R2 ? R2 : 1);
    static constexpr int L4 = L3 * (R3 ? R3 : 1);
    static constexpr int L5 = L4 * (R4 ? R4 : 1);
    static constexpr int O2 = 0;
    static constexpr int O3 = O2 + (R2 ? (R2 - 1) * L2 : 0);
    static constexpr int O4 = O3 + (R3 ? (R3 - 1) * L3 : 0);
    static constexpr int O5 = O4 + (R4 ? (R4 - 1) * L4 : 0);
    static constexpr int SIZE = O5 + (R5 ? (R5 - 1) * L5 : 0);
    static constexpr int M = L5 * (R5 ? R5 : 1);
    // full-table exponent of per-pass entry e
    LL_HD static constexpr u32 exponent(int e) {
        return e < O3 ? one(e - O2, R2, L2) : e < O4 ? one(e - O3, R3, L3) : e < O5 ? one(e - O4, R4, L4)
                                                                             : one(e - O5, R5, L5);
    }
    LL_HD static constexpr u32 one(int d, int r, int L) {
        return (u32)((d / L + 1) * (d % L) * (M / (r * L)));
    }
};

// Copy the per-pass twiddle table of PassTable<R1..R5> out of the full table src (w_M^j, j < M)
// into LDS, all NT threads of the workgroup.
template <class W, int NT, int R1, int R2, int R3, int R4, int R5>
LL_DEV void stage_pass_table(W *dst, const void *src, int tid) {
    typedef PassTable<R1, R2, R3, R4, R5> PT;
    const W LL_GLOBAL *g = (const W LL_GLOBAL *)src;
    constexpr int NI = (PT::SIZE + NT - 1) / NT;
    // All of a thread's loads are issued before its first store, so their round trips overlap
    // (a guarded load-then-store per entry waits out each load in turn).  Lanes past the end
    // load the last entry and store nothing.
    W v[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = tid + NT * i;
        v[i] = ld_table(g, PT::exponent(PT::SIZE % NT == 0 || e < PT::SIZE ? e : PT::SIZE - 1));
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int e = tid + NT * i;
        if (PT::SIZE % NT == 0 || e < PT::SIZE) dst[e] = v[i];
    }
}

// Twiddle table access: a global table (gptr) or a copy staged in LDS (generic pointer to
// __shared__ memory, which the compiler lowers to ds_read once the kernel is inlined).
template <class W, class TP>
LL_DEV W tw_at(TP tw, u32 idx) {
    if constexpr (std::is_same<TP, gptr<W>>::value) return ld_table(tw, idx);
    else return tw[idx];
}

template <class F, bool INV>
LL_DEV void dft2(typename F::T &a0, typename F::T &a1) {
    typename F::T t = F::sub(a0, a1);
    a0 = F::add(a0, a1);
    a1 = t;
}

template <class F, bool INV>
LL_DEV void dft4(typename F::T &a0, typename F::T &a1, typename F::T &a2, typename F::T &a3) {
    typedef typename F::T T;
    T t0 = F::add(a0, a2), t1 = F::sub(a0, a2);
    T t2 = F::add(a1, a3);
    T t3 = F::template mul_w16<4, INV>(F::sub(a1, a3));
    a0 = F::add(t0, t2);
    a2 = F::sub(t0, t2);
    a1 = F::add(t1, t3);
    a3 = F::sub(t1, t3);
}

// Odd-length DFT of FP64 complex data (r = 3, 5, 7), symmetric form: with
// a_s = x_s + x_{r-s}, b_s = x_s - x_{r-s} (s = 1 .. (r-1)/2) and theta = 2 pi s q / r,
//   y_q = x_0 + sum a_s cos(theta) - i sum b_s sin(theta),   y_{r-q} = the same with +i
// (the inverse transform swaps the two).
template <int r>
struct OddRoots {
    // cos / sin of 2 pi j / r, j < r (plain constants, long double literals rounded once)
    static constexpr double c(int j) {
        return r == 3 ? (j % 3 == 0 ? 1.0 : -0.5)
             : r == 5 ? (j % 5 == 0 ? 1.0 : (j % 5 == 1 || j % 5 == 4) ? 0.309016994374947424102293417183 : -0.809016994374947424102293417183)
             : (j % 7 == 0 ? 1.0 : (j % 7 == 1 || j % 7 == 6) ? 0.623489801858733530525004884004
                               : (j % 7 == 2 || j % 7 == 5) ? -0.222520933956314404288902564497
                                                            : -0.900968867902419126236102319507);
    }
    static constexpr double s(int j) {
        return r == 3 ? (j % 3 == 0 ? 0.0 : (j % 3 == 1 ? 0.866025403784438646763723170753 : -0.866025403784438646763723170753))
             : r == 5 ? (j % 5 == 0 ? 0.0 : j % 5 == 1 ? 0.951056516295153572116439333379 : j % 5 == 2 ? 0.587785252292473129168705954639
                        : j % 5 == 3 ? -0.587785252292473129168705954639 : -0.951056516295153572116439333379)
             : (j % 7 == 0 ? 0.0 : j % 7 == 1 ? 0.781831482468029808708444526674 : j % 7 == 2 ? 0.974927912181823607018131682994
                : j % 7 == 3 ? 0.433883739117558120475768332849 : j % 7 == 4 ? -0.433883739117558120475768332849
                : j % 7 == 5 ? -0.974927912181823607018131682994 : -0.781831482468029808708444526674);
    }
};

template <int r, bool INV>
LL_DEV void dft_odd(C64 *x) {
    constexpr int h = (r - 1) / 2;
    C64 a[h], b[h];
#pragma unroll
    for (int s = 1; s <= h; ++s) {
        a[s - 1] = cadd(x[s], x[r - s]);
        b[s - 1] = csub(x[s], x[r - s]);
    }
    C64 y[r];
    C64 y0 = x[0];
#pragma unroll
    for (int s = 0; s < h; ++s) y0 = cadd(y0, a[s]);
    y[0] = y0;
#pragma unroll
    for (int q = 1; q <= h; ++q) {
        C64 A = x[0], B = cmk(0.0, 0.0);
#pragma unroll
        for (int s = 1; s <= h; ++s) {
            constexpr int dummy = 0;
            (void)dummy;
            const double cs = OddRoots<r>::c(s * q), sn = OddRoots<r>::s(s * q);
            A = cmk(__fma_rn(a[s - 1].re, cs, A.re), __fma_rn(a[s - 1].im, cs, A.im));
            B = cmk(__fma_rn(b[s - 1].re, sn, B.re), __fma_rn(b[s - 1].im, sn, B.im));
        }
        // -i B = (B.im, -B.re)
        if constexpr (!INV) {
            y[q] = cmk(A.re + B.im, A.im - B.re);
            y[r - q] = cmk(A.re - B.im, A.im + B.re);
        } else {
            y[q] = cmk(A.re - B.im, A.im + B.re);
            y[r - q] = cmk(A.re + B.im, A.im - B.re);
        }
    }
#pragma unroll
    for (int q = 0; q < r; ++q) x[q] = y[q];
}

// y[q] = sum_s x[s] w_r^{s q} in place, natural order in and out.
template <class F, int r, bool INV>
LL_DEV void dft_r(typename F::T *x) {
    typedef typename F::T T;
    if constexpr (r == 3 || r == 5 || r == 7) {
        static_assert(std::is_same<T, C64>::value, "odd radices are FP64 only");
        dft_odd<r, INV>(x);
    } else if constexpr (r == 1) {
        return;
    } else if constexpr (r == 2) {
        dft2<F, INV>(x[0], x[1]);
    } else if constexpr (r == 4) {
        dft4<F, INV>(x[0], x[1], x[2], x[3]);
    } else if constexpr (r == 8) {
        // s = b + 2a, q = c + 4d:  Z_b = DFT4(x[b + 2a]),  Z_b[c] *= w8^{bc},  y = DFT2 over b
        T z0[4] = {x[0], x[2], x[4], x[6]};
        T z1[4] = {x[1], x[3], x[5], x[7]};
        dft4<F, INV>(z0[0], z0[1], z0[2], z0[3]);
        dft4<F, INV>(z1[0], z1[1], z1[2], z1[3]);
        z1[1] = F::template mul_w16<2, INV>(z1[1]);
        z1[2] = F::template mul_w16<4, INV>(z1[2]);
        z1[3] = F::template mul_w16<6, INV>(z1[3]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = F::add(z0[c], z1[c]);
            x[c + 4] = F::sub(z0[c], z1[c]);
        }
    } else if constexpr (r == 16) {
        // s = b + 4a, q = c + 4d:  Z_b = DFT4(x[b + 4a]),  Z_b[c] *= w16^{bc},  y = DFT4 over b
        T z[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            z[b][0] = x[b]; z[b][1] = x[b + 4]; z[b][2] = x[b + 8]; z[b][3] = x[b + 12];
            dft4<F, INV>(z[b][0], z[b][1], z[b][2], z[b][3]);
        }
        z[1][1] = F::template mul_w16<1, INV>(z[1][1]);
        z[1][2] = F::template mul_w16<2, INV>(z[1][2]);
        z[1][3] = F::template mul_w16<3, INV>(z[1][3]);
        z[2][1] = F::template mul_w16<2, INV>(z[2][1]);
        z[2][2] = F::template mul_w16<4, INV>(z[2][2]);
        z[2][3] = F::template mul_w16<6, INV>(z[2][3]);
        z[3][1] = F::template mul_w16<3, INV>(z[3][1]);
        z[3][2] = F::template mul_w16<6, INV>(z[3][2]);
        z[3][3] = F::template mul_w16<9, INV>(z[3][3]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            dft4<F, INV>(z[0][c], z[1][c], z[2][c], z[3][c]);
            x[c] = z[0][c]; x[c + 4] = z[1][c]; x[c + 8] = z[2][c]; x[c + 12] = z[3][c];
        }
    } else {
        static_assert(r == 16, "unsupported radix");
    }
}

// One Stockham pass over the LDS buffer `lds` (padded, M elements) for a group of G
// threads; t is the thread's index inside its group.  `tw` holds w_M^j, j < M (for the
// inverse transform the C64 path conjugates, the GL path is handed the inverse table).
// FIRST: inputs come from registers;  LAST: outputs stay in registers.
// OFF >= 0: `tw` is a per-pass table (PassTable) and this pass's entries start at OFF.
// PS: LDS pad shift (lds_pad_s), P = 2^PS.
//
// Addressing: G is a multiple of L (L <= M / R = G) so k = (t + G b) mod L = t mod L, and the
// padded LDS index of every access is a per-thread base plus a compile-time constant (an
// immediate offset): for C a multiple of P, pad(t + C) = pad(t) + C (P + 1) / P.
// PS pads the loads, PSW the stores (the next pass loads with PSW as its PS).
template <class F, int M, int G, int R, int r, int L, bool FIRST, bool LAST, bool INV, int OFF = -1, int PS = 4,
          int PSW = PS, class TP>
LL_DEV void stockham_pass(typename F::T (&x)[R], typename F::T *lds, int t, TP tw) {
    static_assert(R % r == 0, "radix must divide the per-thread count");
    static_assert(!FIRST || r == R, "first pass must use the full per-thread radix");
    static_assert(!LAST || (r == R && L * r == M), "last pass must use the full radix");
    static_assert(G % L == 0, "group size must be a multiple of the pass span");
    constexpr int NB = R / r;
    constexpr int STRIDE = M / r;
    constexpr int P = 1 << PS;
    t = ll_opaque(t);
    // The twiddles depend on k = t mod L only (G is a multiple of L), so they are the same for
    // every butterfly of the thread.  They are loaded before the barrier: the table is
    // read-only, and the loads stay in flight while the workgroup meets and reads the LDS.
    typedef typename F::W W;
    W wv[L > 1 ? r - 1 : 1];
    if constexpr (L > 1) {
        const int k = (int)((u32)t % (u32)L);     // t >= 0: unsigned remainder
        constexpr int TSTEP = M / (r * L);
#pragma unroll
        for (int s = 1; s < r; ++s) {
            const u32 e = OFF >= 0 ? (u32)(OFF + (s - 1) * L + k) : (u32)(s * k * TSTEP);
            wv[s - 1] = tw_at<W>(tw, e);
        }
    }
    if constexpr (!FIRST) {
        __syncthreads();
        if constexpr (G % P == 0 && STRIDE % P == 0) {
            const int rb = lds_pad_s<PS>(t);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int s = 0; s < r; ++s) x[b * r + s] = lds[rb + (G * b + s * STRIDE) / P * (P + 1)];
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int s = 0; s < r; ++s) x[b * r + s] = lds[lds_pad_s<PS>(t + G * b + s * STRIDE)];
        }
    }
    if constexpr (L > 1) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int s = 1; s < r; ++s) x[b * r + s] = F::template twiddle<INV>(x[b * r + s], wv[s - 1]);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) dft_r<F, r, INV>(&x[b * r]);
    if constexpr (!LAST) {
        constexpr int PW = 1 << PSW;
        __syncthreads();
        // output q of butterfly i = t + G b goes to (i - k) r + k + q L
        if constexpr (L == 1 && r == PW) {
            // i r + q, q < PW: padded i (r + 1) + q
            const int wb = t * (r + 1);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < r; ++q) lds[wb + G * b * (r + 1) + q] = x[b * r + q];
        } else if constexpr (L % PW == 0) {
            // (t - k) r, G b r and q L are multiples of PW; k < L adds k + k / PW
            const int k = (int)((u32)t % (u32)L);
            const int wb = (int)((u32)(t - k) * (u32)r / (u32)PW * (u32)(PW + 1)) + k + (k >> PSW);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < r; ++q) lds[wb + (G * b * r + q * L) / PW * (PW + 1)] = x[b * r + q];
        } else if constexpr (PW % L == 0 && (r * L) % PW == 0) {
            // (t - k) r and G b r are multiples of r L, so of PW; k + q L stays inside one
            // PW-aligned block (L divides PW), which adds q L / PW pad elements
            const int k = (int)((u32)t % (u32)L);
            const int wb = (int)((u32)(t - k) * (u32)r / (u32)PW * (u32)(PW + 1)) + k;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < r; ++q) lds[wb + G * b * r / PW * (PW + 1) + q * L + q * L / PW] = x[b * r + q];
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int i = t + G * b;
                const int k = i % L;
#pragma unroll
                for (int q = 0; q < r; ++q) lds[lds_pad_s<PSW>((i - k) * r + k + q * L)] = x[b * r + q];
            }
        }
    }
}

// Whole transform: two, three or four passes (R3 == 0: two; R4 == 0: three).  PP: `tw` is the
// per-pass table of PassTable<R1, R2, R3, R4, 0>; PS: LDS pad shift of the first hand-off, whose
// stores are strided by R1.  PSM: pad shift of the later hand-offs, whose stores are runs of L
// consecutive elements; it is used where both of their stores keep constant offsets, else PS.
template <class F, int M, int G, int R, int R1, int R2, int R3, int R4, bool INV, bool PP = false, int PS = 4,
          int PSM = PS, class TP>
LL_DEV void transform(typename F::T (&x)[R], typename F::T *lds, int t, TP tw) {
    typedef PassTable<R1, R2, R3, R4, 0> PT;
    constexpr int O2 = PP ? PT::O2 : -1, O3 = PP ? PT::O3 : -1, O4 = PP ? PT::O4 : -1;
    constexpr int PM = 1 << PSM;
    constexpr int QS = ((R1 * R2) % PM == 0 && PM % R1 == 0 && G % PM == 0) ? PSM : PS;
    if constexpr (R4 != 0) {
        static_assert(R1 * R2 * R3 * R4 == M, "radices must multiply to M");
        stockham_pass<F, M, G, R, R1, 1, true, false, INV, -1, PS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R2, R1, false, false, INV, O2, PS, QS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R3, R1 * R2, false, false, INV, O3, QS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R4, R1 * R2 * R3, false, true, INV, O4, QS>(x, lds, t, tw);
    } else if constexpr (R3 == 0) {
        static_assert(R1 * R2 == M, "radices must multiply to M");
        stockham_pass<F, M, G, R, R1, 1, true, false, INV, -1, PS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R2, R1, false, true, INV, O2, PS>(x, lds, t, tw);
    } else {
        static_assert(R1 * R2 * R3 == M, "radices must multiply to M");
        stockham_pass<F, M, G, R, R1, 1, true, false, INV, -1, PS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R2, R1, false, false, INV, O2, PS, QS>(x, lds, t, tw);
        stockham_pass<F, M, G, R, R3, R1 * R2, false, true, INV, O3, QS>(x, lds, t, tw);
    }
}


// ---- all-LDS mixed-radix Stockham transform (any radix list, any thread count) -----------
// Used where the register-blocked form does not apply (lengths with an odd factor).  The data
// sits in LDS in natural order (padded); every pass reads its butterflies' inputs, barriers,
// and writes the outputs in place.  A thread runs ceil((H/r)/G) butterflies per pass.
template <class F, int H, int G, int r, int L, bool INV, int OFF = -1, int PS = 4, class TP>
LL_DEV void mixed_pass(typename F::T *lds, int t, TP tw) {
    typedef typename F::T T;
    constexpr int NBF = H / r;
    constexpr int NB = (NBF + G - 1) / G;
    constexpr int TSTEP = H / (r * L);
    constexpr int P = 1 << PS;
    static_assert(H % (r * L) == 0, "radix schedule does not divide the length");
    // Fast addressing as in stockham_pass: with G a multiple of L, k = (t + G b) mod L = t mod L;
    // with G, NBF and L multiples of P every padded index is a per-thread base plus a constant.
    constexpr bool FAST = G % L == 0 && G % P == 0 && NBF % P == 0 && (L == 1 ? r == P : L % P == 0);
    T y[NB][r];
    t = ll_opaque(t);
    const int kt = (int)((u32)t % (u32)L);
    // FAST: every butterfly of the thread has k = kt, so its twiddles are loaded once, before
    // the barrier (read-only table; the loads stay in flight across it)
    typedef typename F::W W;
    constexpr bool PRE = FAST && L > 1;
    W wv[PRE ? r - 1 : 1];
    if constexpr (PRE) {
#pragma unroll
        for (int s = 1; s < r; ++s) {
            const u32 e = OFF >= 0 ? (u32)(OFF + (s - 1) * L + kt) : (u32)(s * kt * TSTEP);
            wv[s - 1] = tw_at<W>(tw, e);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int i = t + G * b;
        if (NBF % G == 0 || i < NBF) {
            const int k = FAST ? kt : i % L;
            if constexpr (FAST) {
                const int rb = lds_pad_s<PS>(t);
#pragma unroll
                for (int s = 0; s < r; ++s) y[b][s] = lds[rb + (G * b + s * NBF) / P * (P + 1)];
            } else {
#pragma unroll
                for (int s = 0; s < r; ++s) y[b][s] = lds[lds_pad_s<PS>(i + s * NBF)];
            }
            if constexpr (PRE) {
#pragma unroll
                for (int s = 1; s < r; ++s) y[b][s] = F::template twiddle<INV>(y[b][s], wv[s - 1]);
            } else if constexpr (L > 1) {
#pragma unroll
                for (int s = 1; s < r; ++s) {
                    const u32 e = OFF >= 0 ? (u32)(OFF + (s - 1) * L + k) : (u32)(s * k * TSTEP);
                    y[b][s] = F::template twiddle<INV>(y[b][s], tw_at<W>(tw, e));
                }
            }
            dft_r<F, r, INV>(y[b]);
        }
    }
    __syncthreads();
    // output q of butterfly i goes to (i - k) r + k + q L
    const int wb = L == 1 ? t * (r + 1) : (int)((u32)(t - kt) * (u32)r / (u32)P * (u32)(P + 1)) + kt + (kt >> PS);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int i = t + G * b;
        if (NBF % G == 0 || i < NBF) {
            if constexpr (FAST) {
#pragma unroll
                for (int q = 0; q < r; ++q) {
                    const int off = L == 1 ? G * b * (r + 1) + q : (G * b * r + q * L) / P * (P + 1);
                    lds[wb + off] = y[b][q];
                }
            } else {
                const int k = i % L;
#pragma unroll
                for (int q = 0; q < r; ++q) lds[lds_pad_s<PS>((i - k) * r + k + q * L)] = y[b][q];
            }
        }
    }
}

// x (column layout: thread t holds elements t + G s, s < R) -> transform -> x (same layout)
template <class F, int H, int G, int R, bool INV, int R1, int R2, int R3, int R4, bool PP = false, int PS = 4,
          class TP>
LL_DEV void mixed_transform(typename F::T (&x)[R], typename F::T *lds, int t, TP tw) {
    static_assert(G * R == H, "column layout must cover the length");
    static_assert(R1 * (R2 ? R2 : 1) * (R3 ? R3 : 1) * (R4 ? R4 : 1) == H, "radices must multiply to H");
    typedef PassTable<R1, R2, R3, R4, 0> PT;
    const int tt = ll_opaque(t);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < R; ++s) lds[lds_pad_s<PS>(tt + G * s)] = x[s];
    mixed_pass<F, H, G, R1, 1, INV, -1, PS>(lds, t, tw);
    if constexpr (R2) mixed_pass<F, H, G, R2, R1, INV, PP ? PT::O2 : -1, PS>(lds, t, tw);
    if constexpr (R3) mixed_pass<F, H, G, R3, R1 * R2, INV, PP ? PT::O3 : -1, PS>(lds, t, tw);
    if constexpr (R4) mixed_pass<F, H, G, R4, R1 * R2 * R3, INV, PP ? PT::O4 : -1, PS>(lds, t, tw);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < R; ++s) x[s] = lds[lds_pad_s<PS>(tt + G * s)];
}


// ---- register first/last passes around odd middle passes -------------------------------
// For a length W = 8 * m * 8 (m = product of M1 M2 M3, any of 2..16 including 3, 5, 7):
// the first radix-8 pass reads the column layout from registers, the middle passes run
// through LDS (mixed_pass), the last radix-8 pass leaves the column layout in registers.
template <class F, int W, int G, int R, bool INV, int M1, int M2, int M3, bool PP = false, int PS = 4, class TP>
LL_DEV void hybrid_transform(typename F::T (&x)[R], typename F::T *lds, int t, TP tw) {
    static_assert(R == 8 && G * R == W, "hybrid transform runs 8 elements per thread");
    static_assert(8 * M1 * (M2 ? M2 : 1) * (M3 ? M3 : 1) * 8 == W, "radices must multiply to W");
    // radix list 8, M1, [M2], [M3], 8 packed without gaps for the per-pass table
    constexpr int P3 = M2 ? M2 : 8, P4 = M2 ? (M3 ? M3 : 8) : 0, P5 = (M2 && M3) ? 8 : 0;
    typedef PassTable<8, M1, P3, P4, P5> PT;
    constexpr int OL = !PP ? -1 : (M2 == 0 ? PT::O3 : M3 == 0 ? PT::O4 : PT::O5);
    stockham_pass<F, W, G, R, 8, 1, true, false, INV, -1, PS>(x, lds, t, tw);
    mixed_pass<F, W, G, M1, 8, INV, PP ? PT::O2 : -1, PS>(lds, t, tw);
    if constexpr (M2 != 0) mixed_pass<F, W, G, M2, 8 * M1, INV, PP ? PT::O3 : -1, PS>(lds, t, tw);
    if constexpr (M3 != 0) mixed_pass<F, W, G, M3, 8 * M1 * M2, INV, PP ? PT::O4 : -1, PS>(lds, t, tw);
    stockham_pass<F, W, G, R, 8, W / 8, false, true, INV, OL, PS>(x, lds, t, tw);
}
