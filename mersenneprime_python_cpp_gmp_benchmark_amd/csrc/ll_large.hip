// ll_large.hip -- Lucas-Lehmer for one large exponent: the residue lives in HBM and each
// squaring is three kernel launches of a four-step transform.
//
// Counterpart of the FP64 loop lucaslehmer.cpp:281-311 (and of the NTT loop of
// cuda/ll_ntt.cu:91-163, which the reference only runs up to N = 8192) for transform lengths
// 2^15 .. 2^22 elements.  Layout: the M transform elements form an H x W matrix, element
// j = w + W h (row h, column w; rows contiguous).  With k = k_h + H k_w:
//
//   Z[k] = sum_w w_W^{w k_w} w_M^{w k_h} sum_h z[w + W h] w_H^{h k_h}
//
//  K1 cols_fwd  (blocks of CW columns): carry-in of the previous squaring, weights,
//               H-point FFT per column, twiddle w_M^{w k_h}; in place.
//  K2 rows      (rows k_h and H - k_h together): W-point FFT per row, square (FP64: the
//               packed real spectrum pairs row k_h with row H - k_h), inverse W-point FFT.
//  K3 cols_inv  (blocks of CW columns): twiddle w_M^{-w k_h}, inverse H-point FFT, unweight,
//               round, per-iteration max round-off, -2, and an exact balanced carry along each
//               row segment of CW elements; the segment's carry-out goes to a carry buffer
//               and K1 of the next squaring adds it to the next segment (the last segment of
//               row h feeds row h + 1, and the top one wraps to digit 0: 2^p == 1).
//
// Digits: FP64 packs digits 2j, 2j+1 into element j (N = 2M); GL holds digit j in element j
// (N = M).  Weights and the w_M twiddles come from two-level tables (product of a low and
// a high table entry), so every table stays cache resident at every length.
#include "ll_common.hpp"
#include "ll_dft.hpp"
#include "ll_internal.hpp"
#include <type_traits>

namespace {

constexpr int kCW = 4;       // columns per column-kernel workgroup
constexpr int kR = 8;        // elements per thread in every large-engine FFT

// radix schedule of a 2^K-point transform with 8 elements per thread (first = last = 8)
template <int K>
struct Sched {
    static_assert(K >= 6 && K <= 11, "transform length out of range");
    static constexpr int R1 = 8;
    static constexpr int R2 = K == 6 ? 8 : K == 7 ? 2 : K == 8 ? 4 : K == 9 ? 8 : K == 10 ? 2 : 4;
    static constexpr int R3 = K == 6 ? 0 : 8;
    static constexpr int R4 = K <= 9 ? 0 : 8;
};

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
constexpr bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Column (H-point) transform schedule.  Powers of two use the register-blocked Stockham
// passes; lengths with an odd factor (3, 5, 7, 15) use the all-LDS mixed-radix passes.
template <int H>
struct ColSched {
    static constexpr bool MIXED = !is_pow2(H);
    static constexpr int G = H / kR;
    static constexpr int R1 = MIXED ? (H % 15 == 0 ? 3 : H % 7 == 0 ? 7 : H % 5 == 0 ? 5 : 3) : Sched<ilog2c(H)>::R1;
    static constexpr int REST = H / R1;                 // after the first odd pass (MIXED)
    static constexpr int R2 = MIXED ? (H % 15 == 0 ? 5 : (REST % 16 == 0 && REST > 128 ? 16 : (REST % 8 == 0 ? 8 : 4)))
                                    : Sched<ilog2c(H)>::R2;
    static constexpr int REST2 = REST / R2;
    static constexpr int R3 = MIXED ? (REST2 >= 16 && REST2 % 16 == 0 && REST2 > 64 ? 16 : REST2 >= 8 ? 8 : REST2)
                                    : Sched<ilog2c(H)>::R3;
    static constexpr int REST3 = R3 ? REST2 / R3 : 1;
    static constexpr int R4 = MIXED ? (REST3 > 1 ? REST3 : 0) : Sched<ilog2c(H)>::R4;
};

// Row (W-point) transform schedule: powers of two use the register Stockham passes; W = 64 m
// with m in {3, 5, 6, 7, 10, 12, 14, 15, 20, 24, 28, 30} uses hybrid_transform with the middle
// factor m split into radices (odd ones first).
template <int W>
struct RowSched {
    static constexpr bool HYBRID = !is_pow2(W);
    static constexpr int MID = HYBRID ? W / 64 : 1;
    static constexpr int odd = MID % 15 == 0 ? 15 : MID % 7 == 0 ? 7 : MID % 5 == 0 ? 5 : MID % 3 == 0 ? 3 : 1;
    static constexpr int pw = MID / odd;                       // power-of-two part
    static constexpr int M1 = odd == 15 ? 3 : odd;
    static constexpr int M2 = odd == 15 ? 5 : (pw > 1 ? pw : 0);
    static constexpr int M3 = odd == 15 ? (pw > 1 ? pw : 0) : 0;
};

template <bool IS_GL>
struct ElemOf {
    typedef typename std::conditional<IS_GL, FGL, FC64>::type F;
    typedef typename F::T T;
    typedef typename std::conditional<IS_GL, u64, double>::type WT;   // weight table type
    typedef typename std::conditional<IS_GL, i64, double>::type CT;   // carry type
};

}  // namespace

namespace llh {

// Kernel arguments (by value).  Table pointers are device addresses.
struct LargeArgs {
    void *elem;              // [M] elements
    void *carry;             // [nblk][H] row-segment carry-outs (double or i64)
    unsigned long long *err; // per-iteration max round-off slots (double bits)
    const void *twW, *twWi;  // w_W^j, j < W   (GL: forward, inverse)
    const void *twH, *twHi;  // w_H^j, j < H
    const void *tmL, *tmH;   // w_M^e = tmL[e & (S-1)] * tmH[e >> s]
    const void *tmLi, *tmHi; // GL inverse of the above
    const void *wfL, *wfH;   // forward weight of exponent e: wfL[e & (Sw-1)] * wfH[e >> sw]
    const void *wiL, *wiH;   // inverse weight (with the transform's 1/M or 1/N)
    u32 p, log_n, log_s, log_sw;     // log_n: log2 N when N is a power of two, else 0
    u32 err_slot;
    u64 n;                            // digit count N
    double inv_n;                     // 1.0 / N
    u32 p_mod_n, p_div_n;             // p mod N, floor(p / N)
};

}  // namespace llh

namespace {

using llh::LargeArgs;

template <class T>
LL_DEV T tw2(const void *lo, const void *hi, u32 e, u32 log_s);

template <>
LL_DEV C64 tw2<C64>(const void *lo, const void *hi, u32 e, u32 log_s) {
    return cmul(ld_table(as_global<C64>(lo), e & ((1u << log_s) - 1)), ld_table(as_global<C64>(hi), e >> log_s));
}
template <>
LL_DEV u64 tw2<u64>(const void *lo, const void *hi, u32 e, u32 log_s) {
    return gl_mul(ld_table(as_global<u64>(lo), e & ((1u << log_s) - 1)), ld_table(as_global<u64>(hi), e >> log_s));
}
LL_DEV double wt2d(const void *lo, const void *hi, u32 e, u32 log_s) {
    return ld_table(as_global<double>(lo), e & ((1u << log_s) - 1)) * ld_table(as_global<double>(hi), e >> log_s);
}

// Digit layout for any N: c_d = ceil(d p / N), b_d = c_{d+1} - c_d, weight exponent
// e_d = c_d N - d p in [0, N).  x = d p < 2^53, so one double-precision quotient estimate
// is within one of floor(x / N) and a single correction makes it exact.
struct NDiv {
    u64 n;
    double inv;
    u32 log_n;   // nonzero when N is a power of two
};
LL_DEV void divmod_n(u64 x, const NDiv &D, u64 &q, u64 &r) {
    if (D.log_n) {
        q = x >> D.log_n;
        r = x & (D.n - 1);
        return;
    }
    q = (u64)((double)x * D.inv);
    i64 rr = (i64)(x - q * D.n);
    if (rr < 0) { --q; rr += (i64)D.n; }
    else if (rr >= (i64)D.n) { ++q; rr -= (i64)D.n; }
    r = (u64)rr;
}
LL_DEV u64 cbit(u64 d, u32 p, const NDiv &D) {
    u64 q, r;
    divmod_n(d * p, D, q, r);
    return q + (r != 0);
}
LL_DEV int bits_of(u64 d, u32 p, const NDiv &D) { return (int)(cbit(d + 1, p, D) - cbit(d, p, D)); }
LL_DEV u32 wexp(u64 d, u32 p, const NDiv &D) {
    u64 q, r;
    divmod_n(d * p, D, q, r);
    return r ? (u32)(D.n - r) : 0u;
}
LL_HD NDiv ndiv_of(const LargeArgs &A) { return NDiv{A.n, A.inv_n, A.log_n}; }

// Walks digit positions by the residue r_d = d p mod N alone: with P = p mod N, F = p div N,
//   b_d = F + [r_d + P >= N] + [r_{d+1} > 0] - [r_d > 0],   e_d = (N - r_d) mod N,
//   r_{d+1} = (r_d + P) mod N,   r_{d+D} = (r_d + (D p mod N)) mod N.
struct DigitWalk {
    u32 n, P, F;
    LL_DEV u32 add(u32 r, u32 delta) const {          // (r + delta) mod N, r, delta < N
        const u32 v = r + delta;
        return v >= n ? v - n : v;
    }
    LL_DEV u32 next(u32 r) const { return add(r, P); }
    LL_DEV int bits(u32 r) const {
        const u32 rn = next(r);
        return (int)(F + (r + P >= n ? 1u : 0u) + (rn > 0 ? 1u : 0u)) - (r > 0 ? 1 : 0);
    }
    LL_DEV u32 wexp(u32 r) const { return r ? n - r : 0u; }
};
LL_DEV DigitWalk walk_of(const LargeArgs &A) { return DigitWalk{(u32)A.n, A.p_mod_n, A.p_div_n}; }
// d p mod N for one position (one exact division; used once per thread or per row)
LL_DEV u32 rmod(u64 d, u32 p, const NDiv &D) {
    u64 q, r;
    divmod_n(d * p, D, q, r);
    return (u32)r;
}

// balanced split of v at b bits: v = x + c 2^b, x in [-2^(b-1), 2^(b-1))
LL_DEV double split_d(double v, int b, double &c) {
    c = floor(__fma_rn(v, ldexp(1.0, -b), 0.5));
    return __fma_rn(-c, ldexp(1.0, b), v);
}
LL_DEV i64 split_i(i64 v, int b, i64 &c) {
    c = (v + ((i64)1 << (b - 1))) >> b;
    return v - (c << b);
}

// ---- K1 -----------------------------------------------------------------------------------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(kCW * (H / kR)) lg_cols_fwd(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    typedef typename E::CT CT;
    typedef ColSched<H> S;
    constexpr int GH = H / kR, NT = kCW * GH;
    constexpr int REG = H + H / 16 + (IS_GL ? 8 : 4);
    constexpr int NBLK = W / kCW;
    __shared__ __attribute__((aligned(16))) T lds_all[kCW * REG];
    const int tid = threadIdx.x, c = tid % kCW, t = tid / kCW;
    const int blk = blockIdx.x, w = blk * kCW + c;
    T *lds = lds_all + c * REG;
    T *elem = reinterpret_cast<T *>(A.elem);
    const CT *carry = reinterpret_cast<const CT *>(A.carry);
    const u32 p = A.p;
    const NDiv log_n = ndiv_of(A);

    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[w + W * (t + GH * s)];
    // ---- carry-in: transpose to rows, add the previous segment's carry-out, re-split
#pragma unroll
    for (int s = 0; s < kR; ++s) lds_all[(t + GH * s) * kCW + c] = x[s];
    __syncthreads();
    const DigitWalk DW = walk_of(A);
    for (int h = tid; h < H; h += NT) {
        CT cin;
        if (blk > 0) cin = carry[(blk - 1) * H + h];
        else cin = h > 0 ? carry[(NBLK - 1) * H + h - 1] : carry[(NBLK - 1) * H + H - 1];
        T *row = lds_all + h * kCW;
        const u64 j0 = (u64)blk * kCW + (u64)W * h;
        u32 r = rmod(IS_GL ? j0 : 2 * j0, p, log_n);
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            if constexpr (IS_GL) {
                const int b = DW.bits(r);
                r = DW.next(r);
                i64 v = (i64)row[cc] + cin;
                if (cc == kCW - 1) { row[cc] = (u64)v; break; }
                i64 cc2;
                v = split_i(v, b, cc2);
                row[cc] = (u64)v;
                cin = cc2;
            } else {
                const int b0 = DW.bits(r);
                r = DW.next(r);
                const int b1 = DW.bits(r);
                r = DW.next(r);
                double c1, c2;
                const double x0 = split_d(row[cc].re + cin, b0, c1);
                double x1 = row[cc].im + c1;
                if (cc == kCW - 1) { row[cc] = cmk(x0, x1); break; }
                x1 = split_d(x1, b1, c2);
                row[cc] = cmk(x0, x1);
                cin = c2;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = lds_all[(t + GH * s) * kCW + c];
    // ---- weights (digit residues walked down the column: step = W GH elements)
    {
        const int tt = ll_opaque(t);
        const u64 j0 = (u64)w + (u64)W * tt;
        u32 r = rmod(IS_GL ? j0 : 2 * j0, p, log_n);
        const u32 step = rmod((IS_GL ? 1ull : 2ull) * W * GH, p, log_n);
#pragma unroll
        for (int s = 0; s < kR; ++s) {
            if constexpr (IS_GL) {
                x[s] = gl_mul(gl_from_i64((i64)x[s]), tw2<u64>(A.wfL, A.wfH, DW.wexp(r), A.log_sw));
            } else {
                const u32 r1 = DW.next(r);
                x[s] = cmk(x[s].re * wt2d(A.wfL, A.wfH, DW.wexp(r), A.log_sw),
                           x[s].im * wt2d(A.wfL, A.wfH, DW.wexp(r1), A.log_sw));
            }
            r = DW.add(r, step);
        }
    }
    if constexpr (S::MIXED)
        mixed_transform<F, H, GH, kR, false, S::R1, S::R2, S::R3, S::R4>(x, lds, t, as_global<T>(A.twH));
    else
        transform<F, H, GH, kR, S::R1, S::R2, S::R3, S::R4, false>(x, lds, t, as_global<T>(A.twH));
    // ---- twiddle w_M^{w k_h}, k_h = t + GH q, and store
    {
        const int tt = ll_opaque(t);
#pragma unroll
        for (int q = 0; q < kR; ++q) {
            const u32 kh = tt + GH * q;
            const u32 e = (u32)w * kh;
            if constexpr (IS_GL) x[q] = gl_mul(x[q], tw2<u64>(A.tmL, A.tmH, e, A.log_s));
            else x[q] = cmul(x[q], tw2<C64>(A.tmL, A.tmH, e, A.log_s));
            elem[w + W * kh] = x[q];
        }
    }
}

template <class F, int W, int GW, bool INV>
LL_DEV void row_transform(typename F::T (&x)[kR], typename F::T *lds, int t, gptr<typename F::W> tw) {
    typedef RowSched<W> RS;
    if constexpr (RS::HYBRID) {
        hybrid_transform<F, W, GW, kR, INV, RS::M1, RS::M2, RS::M3>(x, lds, t, tw);
    } else {
        typedef Sched<ilog2c(W)> S;
        transform<F, W, GW, kR, S::R1, S::R2, S::R3, S::R4, INV>(x, lds, t, tw);
    }
}

// ---- K2 -----------------------------------------------------------------------------------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(2 * (W / kR)) lg_rows(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    typedef RowSched<W> RS;
    constexpr int GW = W / kR;
    constexpr int REG = W + W / 16 + 4;
    __shared__ __attribute__((aligned(16))) T lds_all[2 * REG];
    const int tid = threadIdx.x, g = tid / GW, t = tid % GW;
    const int i = blockIdx.x;
    const int row = g == 0 ? i : (i == 0 ? H / 2 : H - i);
    T *lds = lds_all + g * REG;
    T *elem = reinterpret_cast<T *>(A.elem) + (size_t)W * row;

    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[t + GW * s];
    row_transform<F, W, GW, false>(x, lds, t, as_global<T>(A.twW));
    if constexpr (IS_GL) {
#pragma unroll
        for (int q = 0; q < kR; ++q) x[q] = gl_sqr(x[q]);
        row_transform<F, W, GW, true>(x, lds, t, as_global<T>(A.twWi));
    } else {
        const int tt = ll_opaque(t);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kR; ++q) lds[lds_pad(tt + GW * q)] = x[q];
        __syncthreads();
        // partner of (row, k_w): row 0 -> (0, -k_w); row H/2 -> (H/2, W-1-k_w);
        // otherwise -> (H - row, W-1-k_w), held by the other group of this workgroup
        const T *plds = (row == 0 || row == H / 2) ? lds : lds_all + (1 - g) * REG;
#pragma unroll
        for (int q = 0; q < kR; ++q) {
            const int kw = tt + GW * q;
            const int kwp = row == 0 ? (kw == 0 ? 0 : W - kw) : (W - 1 - kw);
            const C64 zm = plds[lds_pad(kwp)];
            const C64 D = csub(x[q], cconj(zm));
            const u32 k = (u32)row + (u32)H * (u32)kw;
            const C64 wk = tw2<C64>(A.tmL, A.tmH, k, A.log_s);
            const C64 AD = cmul(cmk(1.0 + wk.re, wk.im), csqr(D));
            const C64 Z2 = csqr(x[q]);
            x[q] = cmk(__fma_rn(-0.25, AD.re, Z2.re), __fma_rn(-0.25, AD.im, Z2.im));
        }
        row_transform<F, W, GW, true>(x, lds, t, as_global<T>(A.twW));
    }
#pragma unroll
    for (int s = 0; s < kR; ++s) elem[t + GW * s] = x[s];
}

// ---- K3 -----------------------------------------------------------------------------------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(kCW * (H / kR)) lg_cols_inv(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    typedef typename E::CT CT;
    typedef ColSched<H> S;
    constexpr int GH = H / kR, NT = kCW * GH;
    constexpr int REG = H + H / 16 + (IS_GL ? 8 : 4);
    __shared__ __attribute__((aligned(16))) T lds_all[kCW * REG];
    __shared__ unsigned long long s_err;
    const int tid = threadIdx.x, c = tid % kCW, t = tid / kCW;
    const int blk = blockIdx.x, w = blk * kCW + c;
    T *lds = lds_all + c * REG;
    T *elem = reinterpret_cast<T *>(A.elem);
    CT *carry = reinterpret_cast<CT *>(A.carry);
    const u32 p = A.p;
    const NDiv log_n = ndiv_of(A);
    if (tid == 0) s_err = 0;

    T x[kR];
    {
        const int tt = ll_opaque(t);
#pragma unroll
        for (int s = 0; s < kR; ++s) {
            const u32 kh = tt + GH * s;
            const u32 e = (u32)w * kh;
            x[s] = elem[w + W * kh];
            if constexpr (IS_GL) x[s] = gl_mul(x[s], tw2<u64>(A.tmLi, A.tmHi, e, A.log_s));
            else x[s] = cmulc(x[s], tw2<C64>(A.tmL, A.tmH, e, A.log_s));
        }
    }
    if constexpr (S::MIXED)
        mixed_transform<F, H, GH, kR, true, S::R1, S::R2, S::R3, S::R4>(x, lds, t, as_global<T>(A.twH));
    else
        transform<F, H, GH, kR, S::R1, S::R2, S::R3, S::R4, true>(x, lds, t, as_global<T>(IS_GL ? A.twHi : A.twH));
    // ---- unweight and round (rows h = t + GH q)
    double eloc = 0.0;
    const DigitWalk DW = walk_of(A);
    {
        const int tt = ll_opaque(t);
        const u64 j0 = (u64)w + (u64)W * tt;
        u32 r = rmod(IS_GL ? j0 : 2 * j0, p, log_n);
        const u32 step = rmod((IS_GL ? 1ull : 2ull) * W * GH, p, log_n);
#pragma unroll
        for (int q = 0; q < kR; ++q) {
            const u64 j = j0 + (u64)W * GH * q;
            const u32 rq = r;
            r = DW.add(r, step);
            if constexpr (IS_GL) {
                i64 v = gl_to_i64(gl_mul(x[q], tw2<u64>(A.wiL, A.wiH, DW.wexp(rq), A.log_sw)));
                if (j == 0) v -= 2;
                x[q] = (u64)v;
            } else {
                const double v0 = x[q].re * wt2d(A.wiL, A.wiH, DW.wexp(rq), A.log_sw);
                const double v1 = x[q].im * wt2d(A.wiL, A.wiH, DW.wexp(DW.next(rq)), A.log_sw);
                double r0 = __builtin_rint(v0);
                const double r1 = __builtin_rint(v1);
                eloc = fmax(eloc, fmax(fabs(v0 - r0), fabs(v1 - r1)));
                if (j == 0) r0 -= 2.0;
                x[q] = cmk(r0, r1);
            }
        }
    }
    // ---- balanced carry along each row segment of kCW elements
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kR; ++q) lds_all[(t + GH * q) * kCW + c] = x[q];
    __syncthreads();
    for (int h = tid; h < H; h += NT) {
        T *row = lds_all + h * kCW;
        const u64 j0 = (u64)blk * kCW + (u64)W * h;
        u32 r = rmod(IS_GL ? j0 : 2 * j0, p, log_n);
        CT cy = 0;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            if constexpr (IS_GL) {
                i64 c2;
                const i64 v = split_i((i64)row[cc] + cy, DW.bits(r), c2);
                r = DW.next(r);
                row[cc] = (u64)v;
                cy = c2;
            } else {
                double c1, c2;
                const int b0 = DW.bits(r);
                r = DW.next(r);
                const int b1 = DW.bits(r);
                r = DW.next(r);
                const double x0 = split_d(row[cc].re + cy, b0, c1);
                const double x1 = split_d(row[cc].im + c1, b1, c2);
                row[cc] = cmk(x0, x1);
                cy = c2;
            }
        }
        carry[blk * H + h] = cy;
        // the segment is contiguous in memory
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) elem[blk * kCW + cc + (size_t)W * h] = row[cc];
    }
    if constexpr (!IS_GL) {
        atomicMax(&s_err, (unsigned long long)__double_as_longlong(eloc));
        __syncthreads();
        if (tid == 0) atomicMax(A.err + A.err_slot, s_err);
    }
}

// ---- carry-in only (after the last squaring, before normalisation) ----------------------------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(256) lg_apply_carry(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::T T;
    typedef typename E::CT CT;
    constexpr int NBLK = W / kCW;
    const int blk = blockIdx.x;
    T *elem = reinterpret_cast<T *>(A.elem);
    const CT *carry = reinterpret_cast<const CT *>(A.carry);
    const u32 p = A.p;
    const NDiv log_n = ndiv_of(A);
    for (int h = threadIdx.x; h < H; h += 256) {
        CT cin;
        if (blk > 0) cin = carry[(blk - 1) * H + h];
        else cin = h > 0 ? carry[(NBLK - 1) * H + h - 1] : carry[(NBLK - 1) * H + H - 1];
        T *row = elem + blk * kCW + (size_t)W * h;
        const u64 j0 = (u64)blk * kCW + (u64)W * h;
        for (int cc = 0; cc < kCW; ++cc) {
            if constexpr (IS_GL) {
                i64 v = (i64)row[cc] + cin;
                if (cc == kCW - 1) { row[cc] = (u64)v; break; }
                i64 c2;
                v = split_i(v, bits_of(j0 + cc, p, log_n), c2);
                row[cc] = (u64)v;
                cin = c2;
            } else {
                const u64 d0 = 2 * (j0 + cc);
                double c1, c2;
                const double x0 = split_d(row[cc].re + cin, bits_of(d0, p, log_n), c1);
                double x1 = row[cc].im + c1;
                if (cc == kCW - 1) { row[cc] = cmk(x0, x1); break; }
                x1 = split_d(x1, bits_of(d0 + 1, p, log_n), c2);
                row[cc] = cmk(x0, x1);
                cin = c2;
            }
        }
    }
}

}  // namespace

// ---- exact normalisation of N digits (one workgroup) --------------------------------------------
// Digits (near balanced) are read as i64 from the element array; each thread normalises a
// contiguous segment to non-negative digits; carries then ripple between segments in rounds
// (usually one digit deep) until none is left, the top segment's carry wrapping to digit 0.
// Output: res64 of the canonical residue and whether it is zero (the all-ones pattern
// M_p counts as zero).  Restates the GMP rebuild of lucaslehmer.cpp:312-323 on the device.
namespace {
constexpr int kNormThreads = 1024;

struct NormOut {
    unsigned long long res64;
    int zero;
    int status;
};

template <bool IS_GL>
__global__ void __launch_bounds__(kNormThreads) lg_normalize(const void *elem_v, i64 *dig, u32 p, NDiv log_n,
                                                           NormOut *out) {
    const u64 N = log_n.n;
    const u64 K = (N + kNormThreads - 1) / kNormThreads;
    const int t = threadIdx.x;
    const u64 d0 = (u64)t * K, d1 = d0 + K < N ? d0 + K : N;
    __shared__ i64 s_cout[kNormThreads];
    __shared__ int s_any, s_ones, s_zero;
    // gather digits to the i64 scratch array
    for (u64 d = d0; d < d1; ++d) {
        if constexpr (IS_GL) dig[d] = (i64)reinterpret_cast<const u64 *>(elem_v)[d];
        else dig[d] = (i64)reinterpret_cast<const double *>(elem_v)[d];
    }
    // local normalisation to non-negative digits
    i64 cy = 0;
    for (u64 d = d0; d < d1; ++d) {
        const int b = bits_of(d, p, log_n);
        const i64 v = dig[d] + cy;
        dig[d] = v & (((i64)1 << b) - 1);
        cy = v >> b;
    }
    s_cout[t] = cy;
    __syncthreads();
    for (int round = 0; round < 4 * kNormThreads + 8; ++round) {
        __syncthreads();              // every thread has read the previous round's flag
        if (t == 0) s_any = 0;
        __syncthreads();
        i64 cin = s_cout[(t + kNormThreads - 1) % kNormThreads];
        __syncthreads();
        cy = 0;
        if (cin != 0 && d0 < d1) {
            for (u64 d = d0; d < d1 && cin != 0; ++d) {
                const int b = bits_of(d, p, log_n);
                const i64 v = dig[d] + cin;
                dig[d] = v & (((i64)1 << b) - 1);
                cin = v >> b;
            }
            cy = cin;
        }
        s_cout[t] = cy;
        if (cy != 0) s_any = 1;
        __syncthreads();
        if (!s_any) break;
    }
    if (t == 0) { s_ones = 1; s_zero = 1; }
    __syncthreads();
    int ones = 1, zero = 1;
    for (u64 d = d0; d < d1; ++d) {
        const u64 v = (u64)dig[d];
        if (v) zero = 0;
        if (v != ((1ull << bits_of(d, p, log_n)) - 1)) ones = 0;
    }
    if (!ones) s_ones = 0;
    if (!zero) s_zero = 0;
    __syncthreads();
    if (t == 0) {
        u64 r = 0;
        for (u64 d = 0; d < N; ++d) {
            const u64 cd = cbit(d, p, log_n);
            if (cd >= 64) break;
            r |= (u64)dig[d] << cd;
        }
        const bool z = s_zero || s_ones;
        out->res64 = z ? 0 : r;
        out->zero = z;
        out->status = s_any ? 1 : 0;
    }
}

}  // namespace

namespace llh {

struct LargeLaunch {
    void (*fwd)(LargeArgs);
    void (*rows)(LargeArgs);
    void (*inv)(LargeArgs);
    void (*apply)(LargeArgs);
    int w_len, h_len;   // W and H (either may carry the odd factor)
};

template <bool IS_GL, int W, int H>
static LargeLaunch make_launch() {
    return LargeLaunch{lg_cols_fwd<IS_GL, W, H>, lg_rows<IS_GL, W, H>, lg_cols_inv<IS_GL, W, H>,
                       lg_apply_carry<IS_GL, W, H>, W, H};
}

// element count M -> (W, H) split and kernels.  FP64: M = m 2^k with m in {1, 3, 5, 7, 15},
// the odd factor in the row length W; GL: powers of two (the field has no 3rd or 7th roots of
// unity, and no N-th root of 2 for N = 3 * 2^k).
bool large_launch_for(int transform, u64 M, LargeLaunch &L) {
    const bool gl = transform == kNTT;
#define LL_P2(MV, W, H)                                                   \
    case MV:                                                              \
        L = gl ? make_launch<true, W, H>() : make_launch<false, W, H>();  \
        return true;
#define LL_ODD(MV, W, H)                                                  \
    case MV:                                                              \
        if (gl) return false;                                             \
        L = make_launch<false, W, H>();                                   \
        return true;
    switch (M) {
        LL_P2(1ull << 15, 256, 128)
        LL_P2(1ull << 16, 256, 256)
        LL_P2(1ull << 17, 512, 256)
        LL_P2(1ull << 18, 512, 512)
        LL_P2(1ull << 19, 1024, 512)
        LL_P2(1ull << 20, 1024, 1024)
        LL_P2(1ull << 21, 2048, 1024)
        LL_ODD(3ull << 14, 384, 128)
        LL_ODD(3ull << 15, 384, 256)
        LL_ODD(3ull << 16, 768, 256)
        LL_ODD(3ull << 17, 768, 512)
        LL_ODD(3ull << 18, 768, 1024)
        LL_ODD(3ull << 19, 1536, 1024)
        LL_ODD(5ull << 13, 320, 128)
        LL_ODD(5ull << 14, 640, 128)
        LL_ODD(5ull << 15, 640, 256)
        LL_ODD(5ull << 16, 640, 512)
        LL_ODD(5ull << 17, 640, 1024)
        LL_ODD(5ull << 18, 1280, 1024)
        LL_ODD(7ull << 13, 448, 128)
        LL_ODD(7ull << 14, 448, 256)
        LL_ODD(7ull << 15, 896, 256)
        LL_ODD(7ull << 16, 896, 512)
        LL_ODD(7ull << 17, 896, 1024)
        LL_ODD(7ull << 18, 1792, 1024)
        LL_ODD(15ull << 13, 960, 128)
        LL_ODD(15ull << 14, 960, 256)
        LL_ODD(15ull << 15, 960, 512)
        LL_ODD(15ull << 16, 960, 1024)
        LL_ODD(15ull << 17, 1920, 1024)
        default: return false;
    }
#undef LL_P2
#undef LL_ODD
}

hipError_t large_step(const LargeLaunch &L, const LargeArgs &A, hipStream_t st) {
    const int W = L.w_len, H = L.h_len;
    const dim3 gcol(W / kCW), bcol(kCW * (H / kR));
    const dim3 grow(H / 2), brow(2 * (W / kR));
    hipLaunchKernelGGL(L.fwd, gcol, bcol, 0, st, A);
    hipLaunchKernelGGL(L.rows, grow, brow, 0, st, A);
    hipLaunchKernelGGL(L.inv, gcol, bcol, 0, st, A);
    return hipGetLastError();
}

hipError_t large_apply_carry(const LargeLaunch &L, const LargeArgs &A, hipStream_t st) {
    hipLaunchKernelGGL(L.apply, dim3(L.w_len / kCW), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t large_normalize(int transform, const void *elem, i64 *dig, const LargeArgs &A, void *out,
                           hipStream_t st) {
    const NDiv D = ndiv_of(A);
    if (transform == kNTT)
        hipLaunchKernelGGL(lg_normalize<true>, dim3(1), dim3(kNormThreads), 0, st, elem, dig, A.p, D,
                           (NormOut *)out);
    else
        hipLaunchKernelGGL(lg_normalize<false>, dim3(1), dim3(kNormThreads), 0, st, elem, dig, A.p, D,
                           (NormOut *)out);
    return hipGetLastError();
}

}  // namespace llh
