// ll_large_impl.hpp -- kernels of the large engine: Lucas-Lehmer for one large exponent, the
// residue in HBM and each squaring two kernel launches of a four-step transform.  Included by
// ll_large.hip (normalisation, re-layout, dispatch) and by the ll_large_g*.hip units, each of
// which instantiates a group of the schedules listed in ll_large_scheds.inc.
//
// Counterpart of the FP64 loop lucaslehmer.cpp:281-311 (and of the NTT loop of
// cuda/ll_ntt.cu:91-163, which the reference only runs up to N = 8192) for transform lengths
// 2^15 .. 2^21 elements.  The M transform elements form an H x W matrix, element j = w + W h
// (row h, column w), stored pair-interleaved in HBM (elem_index): a 128-byte line holds the same
// LC columns of rows h and H - h (LC = 4 FP64 complex or 8 GL words; row 0 pairs with row H/2).
// The row pass then reads its row pair as whole lines, and a column-pass block of LC columns
// reads whole lines too; in plain row-major order two blocks on different XCDs each read and
// wrote half of every line, which took twice the time for the same bytes.  With k = k_h + H k_w:
//
//   Z[k] = sum_w w_W^{w k_w} w_M^{w k_h} sum_h z[w + W h] w_H^{h k_h}
//
//  rows       (rows k_h and H - k_h together): W-point FFT per row, square (FP64: the packed
//             real spectrum pairs row k_h with row H - k_h), inverse W-point FFT.
//  cols_fused (blocks of CW columns): twiddle w_M^{-w k_h}, inverse H-point FFT, unweight,
//             round, per-iteration max round-off, -2, an exact balanced carry along each row
//             segment of CW elements; the segment's carry-out is handed to the next block
//             inside the launch (the last segment of row h feeds row h + 1, and the top one
//             wraps to digit 0: 2^p == 1); then the carry-in with a re-split, weights, the
//             forward H-point FFT and the twiddle w_M^{w k_h} of the next squaring.
//  cols_fwd / cols_inv: the forward and the inverse half alone, which open and close a block
//             of squarings (the carries between blocks wait in a buffer).
//
// Digits: FP64 packs digits 2j, 2j+1 into element j (N = 2M); GL holds digit j in element j
// (N = M).  Weights and the w_M twiddles come from two-level tables (product of a low and
// a high table entry), so every table stays cache resident at every length; the FFT
// twiddles come from global tables in per-pass order (a pass's lanes read consecutive
// entries), loaded before each pass's barrier.  Staging them in LDS instead cost every
// workgroup its own copy and 8-16 KB of LDS, and one squaring took 30.3 us at N = 2^20 against
// 27.8 us read from L2.
#include "ll_common.hpp"
#include "ll_dft.hpp"
#include "ll_internal.hpp"
#pragma once
#include <type_traits>

namespace {

constexpr int kR = 8;        // elements per thread in every large-engine FFT

// radix schedule of a 2^K-point transform with 8 elements per thread (first = last = 8)
template <int K>
struct Sched {
    static_assert(K >= 6 && K <= 12, "transform length out of range");
    static constexpr int R1 = 8;
    static constexpr int R2 = K == 6 ? 8 : K == 7 ? 2 : K == 8 ? 4 : K == 9 ? 8 : K == 10 ? 2 : K == 11 ? 4 : 8;
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

// Radix lists (R1 .. R5, 0 = no pass) of the two transforms, for the per-pass twiddle tables.
template <int H>
struct ColRadix {
    typedef ColSched<H> S;
    static constexpr int R1 = S::R1, R2 = S::R2, R3 = S::R3, R4 = S::R4, R5 = 0;
};
template <int W>
struct RowRadix {
    typedef RowSched<W> RS;
    typedef Sched<RS::HYBRID ? 10 : ilog2c(W)> S;
    static constexpr int R1 = RS::HYBRID ? 8 : S::R1;
    static constexpr int R2 = RS::HYBRID ? RS::M1 : S::R2;
    static constexpr int R3 = RS::HYBRID ? (RS::M2 ? RS::M2 : 8) : S::R3;
    static constexpr int R4 = RS::HYBRID ? (RS::M2 ? (RS::M3 ? RS::M3 : 8) : 0) : S::R4;
    static constexpr int R5 = RS::HYBRID ? (RS::M2 && RS::M3 ? 8 : 0) : 0;
};

// LDS pad shift of the large-engine transforms (one pad element per 8: radix-8 passes)
constexpr int kPS = 3;

template <bool IS_GL>
struct ElemOf {
    typedef typename std::conditional<IS_GL, FGL, FC64>::type F;
    typedef typename F::T T;
    typedef typename std::conditional<IS_GL, u64, double>::type WT;   // weight table type
    typedef typename std::conditional<IS_GL, i64, double>::type CT;   // carry type
    static constexpr u32 LC = IS_GL ? 8 : 4;                           // columns per half line
};

// Storage index of element (h, w) of the H x W matrix (H even; W a multiple of LC): the line of
// row pair P = (h < H/2 ? h : H - h) mod H/2 holds LC columns of row P (slot 0) and of row
// H - P (slot 1; for P = 0, row H/2).
LL_HD u64 elem_index(u32 h, u32 w, u32 W, u32 H, u32 LC) {
    const u32 hh = H / 2;
    const u32 P = h < hh ? h : (h == hh ? 0u : H - h);
    return (u64)P * 2 * W + (u64)(w / LC) * (2 * LC) + (h < hh ? 0 : LC) + w % LC;
}

}  // namespace


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

// A two-level table, entry e = lo[e mod 2^ls] * hi[e >> ls], read from global memory or from a
// copy staged in LDS.
LL_DEV C64 tmul(C64 a, C64 b) { return cmul(a, b); }
LL_DEV u64 tmul(u64 a, u64 b) { return gl_mul(a, b); }
LL_DEV double tmul(double a, double b) { return a * b; }
template <class T, bool LDS>
struct Tab2 {
    const T *lo, *hi;
    u32 ls;
    LL_DEV T operator()(u32 e) const {
        const u32 a = e & ((1u << ls) - 1), b = e >> ls;
        if constexpr (LDS) return tmul(lo[a], hi[b]);
        else return tmul(ld_table(as_global<T>(lo), a), ld_table(as_global<T>(hi), b));
    }
};
template <class T>
LL_HD Tab2<T, false> gtab(const void *lo, const void *hi, u32 ls) {
    return Tab2<T, false>{(const T *)lo, (const T *)hi, ls};
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
LL_HD NDiv ndiv_of(const LargeArgs &A) { return NDiv{A.n, A.inv_n, A.log_n}; }

LL_DEV ExpWalk ewalk_of(const LargeArgs &A) { return ExpWalk{(u32)A.n, A.p_mod_n, A.p_div_n}; }
// d p mod N for one position (one exact division; used once per thread or per row)
LL_DEV u32 rmod(u64 d, u32 p, const NDiv &D) {
    u64 q, r;
    divmod_n(d * p, D, q, r);
    return (u32)r;
}

// balanced split of v at b bits: v = x + c 2^b, |x| <= 2^(b-1); both steps exact for integer v
// below 2^53 (a power-of-two scaling, then the exact difference)
LL_DEV double split_d(double v, int b, double &c) {
    c = __builtin_rint(ldexp(v, -b));
    return v - ldexp(c, b);
}
LL_DEV i64 split_i(i64 v, int b, i64 &c) {
    c = (v + ((i64)1 << (b - 1))) >> b;
    return v - (c << b);
}

// ---- cross-workgroup hand-off ---------------------------------------------------------------
// Values another workgroup reads inside the same launch are written and read with agent-scope
// atomics, which go to the coherence point shared by all XCDs (the per-XCD L2s are not coherent
// with each other inside a launch).
LL_DEV void st_coh(u64 *p, u64 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
LL_DEV u64 ld_coh_u(const u64 *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
LL_DEV void st_coh(double *p, double v) { st_coh((u64 *)p, (u64)__double_as_longlong(v)); }
LL_DEV void st_coh(i64 *p, i64 v) { st_coh((u64 *)p, (u64)v); }
LL_DEV void ld_coh(const double *p, double &v) { v = __longlong_as_double((long long)ld_coh_u((const u64 *)p)); }
LL_DEV void ld_coh(const i64 *p, i64 &v) { v = (i64)ld_coh_u((const u64 *)p); }

// Element stores that the next launch reads: non-temporal, so they stream to memory during the
// launch instead of sitting dirty in this XCD's L2 until the write-back at the launch's end.
LL_DEV void st_stream(C64 *p, C64 v) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 t;
    t.x = v.re;
    t.y = v.im;
    __builtin_nontemporal_store(t, reinterpret_cast<d2 *>(p));
}
LL_DEV void st_stream(u64 *p, u64 v) { __builtin_nontemporal_store(v, p); }

// wait for this thread's vector-memory operations only (vmcnt = 0; expcnt and lgkmcnt at
// their "no wait" maxima, gfx9 encoding)
LL_DEV void wait_vmem() { __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8)); }

// Tagged carry words of the fused column pass: bits 63..48 hold the launch's tag (LargeArgs::tag),
// bits 47..0 the signed carry.  Every launch rewrites every word, so a word holds either this
// launch's value or the one before's; the host gives consecutive launches different nonzero tags
// (the words start cleared), so a reader polls its own words until the tag matches, with no flag
// and no barrier between publishing and reading.
constexpr u64 kTagShift = 48;
constexpr i64 kCarryLim = (i64)1 << 47;
LL_DEV u64 tag_word(u32 seq, i64 c) {
    return ((u64)(seq & 0xFFFFu) << kTagShift) | ((u64)c & ((1ull << kTagShift) - 1));
}
LL_DEV i64 untag(u64 w) { return (i64)(w << 16) >> 16; }
LL_DEV i64 carry_i(double c) { return (i64)c; }      // integral, |c| < 2^53
LL_DEV i64 carry_i(i64 c) { return c; }
LL_DEV void carry_from(i64 v, double &c) { c = (double)v; }
LL_DEV void carry_from(i64 v, i64 &c) { c = v; }

// Polls until the word at p carries tag seq, for at most about two seconds of wall clock; past
// that it raises *fault and stops waiting, so the launch always drains (and once *fault is up,
// later waits give up within a few polls).
constexpr u64 kSpinTicks = 200000000ull;   // 2 s at the usual 100 MHz wall clock
LL_DEV bool tag_ok(u64 w, u32 seq) { return (u32)(w >> kTagShift) == (seq & 0xFFFFu); }
LL_DEV u64 wait_tagged(const u64 *p, u32 seq, unsigned *fault) {
    u64 w = ld_coh_u(p);
    if (tag_ok(w, seq)) return w;
    const u64 t0 = wall_clock64();
    for (u32 k = 1;; ++k) {
        __builtin_amdgcn_s_sleep(1);
        w = ld_coh_u(p);
        if (tag_ok(w, seq)) return w;
        if (wall_clock64() - t0 > kSpinTicks ||
            ((k & 63) == 0 && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            atomicOr(fault, 1u);
            return w & ~((1ull << kTagShift) - 1);    // carry 0: the run is reported as failed
        }
    }
}

// Exponents of the per-pass twiddle tables of the H-point and W-point transforms (host side:
// the host builds the tables in this order, so kernel and table cannot disagree).
template <int H>
u32 col_pass_exponent(int e) {
    typedef ColRadix<H> RR;
    return PassTable<RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>::exponent(e);
}
template <int W>
u32 row_pass_exponent(int e) {
    typedef RowRadix<W> RR;
    return PassTable<RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>::exponent(e);
}
template <int H>
constexpr int col_pass_size() {
    typedef ColRadix<H> RR;
    return PassTable<RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>::SIZE;
}
template <int W>
constexpr int row_pass_size() {
    typedef RowRadix<W> RR;
    return PassTable<RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>::SIZE;
}

// ---- column kernels ---------------------------------------------------------------------------
// A workgroup owns a block of CW adjacent columns (W / CW blocks).  Thread (c, t), with
// c = tid % CW, holds column w = blk CW + c at rows t + GH q: the column layout of the H-point
// transform.  Between the inverse and the forward transform the digits sit in LDS in the
// row-segment layout lds_all[h CW + c] (a row segment is CW elements, 2 CW digits for FP64,
// contiguous digits), and thread tid walks the segments of rows h = tid + NT i.
// row segments per thread (CW > kR: the first H threads walk one segment each, the rest idle)
LL_HD constexpr int nrow(int cw) { return cw <= kR ? kR / cw : 1; }
LL_HD constexpr int nwt(bool gl) { return gl ? kR : 2 * kR; }  // digit weights per thread

template <bool IS_GL, int W, int H, int CW>
struct Col {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    typedef typename E::CT CT;
    typedef typename E::WT WT;
    typedef ColSched<H> S;
    static constexpr int GH = H / kR, NT = CW * GH, NBLK = W / CW;
    static constexpr int REG = H + H / 8 + (IS_GL ? 8 : 4);
    // Row-segment layout in LDS: segment h at h RSS, RSS = CW + 1 elements (odd), so the lanes of
    // a wave, which walk consecutive segments, start on distinct banks (at a stride of CW the
    // 16-byte reads of 16 lanes fell on the same four banks)
    static constexpr int RSS = CW + 1;
    static constexpr int LDS = CW * REG > H * RSS ? CW * REG : H * RSS;   // elements of lds_all
    static constexpr int NROW = nrow(CW);
    static constexpr int NWT = nwt(IS_GL);
    // every thread walks NROW segments (CW <= kR), or the first H threads one each (CW > kR)
    static constexpr bool ALLSEG = CW <= kR;
    static_assert(ALLSEG ? H == NT * NROW : NT == H * (CW / kR), "row segments must split evenly over the workgroup");
    LL_DEV static bool owns(int h) { return ALLSEG || h < H; }
};

// w_M^{w k_h} at the thread's positions k_h = t + GH q (table: forward or GL inverse); the
// exponent w k_h < W H = M needs no reduction
template <bool IS_GL, int W, int H, int CW, class TAB>
LL_DEV void col_twiddles(typename Col<IS_GL, W, H, CW>::T (&tm)[kR], const TAB &tab, int w, int t) {
    typedef Col<IS_GL, W, H, CW> C;
#pragma unroll
    for (int q = 0; q < kR; ++q) tm[q] = tab((u32)w * (u32)(t + C::GH * q));
}

// Digit weights (table: forward or inverse) at the thread's elements j = w + W (t + GH q);
// FP64 elements hold digits 2j and 2j + 1.
template <bool IS_GL, int W, int H, int CW, class TAB>
LL_DEV void col_weights(typename Col<IS_GL, W, H, CW>::WT (&wt)[nwt(IS_GL)], const TAB &tab, const LargeArgs &A,
                        int w, int t) {
    typedef Col<IS_GL, W, H, CW> C;
    const ExpWalk EW = ewalk_of(A);
    const NDiv D = ndiv_of(A);
    const u64 j0 = (u64)w + (u64)W * t;
    u32 e = EW.of_residue(rmod(IS_GL ? j0 : 2 * j0, A.p, D));
    const u32 step = rmod((IS_GL ? 1ull : 2ull) * W * C::GH, A.p, D);
#pragma unroll
    for (int q = 0; q < kR; ++q) {
        if constexpr (IS_GL) {
            wt[q] = tab(e);
        } else {
            wt[2 * q] = tab(e);
            wt[2 * q + 1] = tab(EW.next(e));
        }
        e = EW.sub(e, step);
    }
}

// x: column layout, already multiplied by the inverse w_M twiddles.  Inverse H-point transform,
// unweight, round (max round-off into eloc), -2 at digit 0, then the row-segment layout in LDS
// with an exact balanced carry along each segment; cy[i] is the carry out of row tid + NT i.
template <bool IS_GL, int W, int H, int CW, class TP>
LL_DEV void col_inverse(typename Col<IS_GL, W, H, CW>::T (&x)[kR], typename Col<IS_GL, W, H, CW>::T *lds_all, TP tw,
                        const typename Col<IS_GL, W, H, CW>::WT (&wi)[nwt(IS_GL)], const LargeArgs &A,
                        int blk, int c, int t, int tid,
                        typename Col<IS_GL, W, H, CW>::CT (&cy)[nrow(CW)], double &eloc) {
    typedef Col<IS_GL, W, H, CW> C;
    typedef typename C::F F;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::S S;
    T *lds = lds_all + c * C::REG;
    if constexpr (S::MIXED)
        mixed_transform<F, H, C::GH, kR, true, S::R1, S::R2, S::R3, S::R4, true, kPS>(x, lds, t, tw);
    else
        transform<F, H, C::GH, kR, S::R1, S::R2, S::R3, S::R4, true, true, kPS>(x, lds, t, tw);
    const bool has_d0 = blk == 0 && c == 0 && t == 0;
#pragma unroll
    for (int q = 0; q < kR; ++q) {
        if constexpr (IS_GL) {
            i64 v = gl_to_i64(gl_mul(x[q], wi[q]));
            if (q == 0 && has_d0) v -= 2;
            x[q] = (u64)v;
        } else {
            const double v0 = x[q].re * wi[2 * q];
            const double v1 = x[q].im * wi[2 * q + 1];
            double r0 = __builtin_rint(v0);
            const double r1 = __builtin_rint(v1);
            eloc = fmax(eloc, fmax(fabs(v0 - r0), fabs(v1 - r1)));
            if (q == 0 && has_d0) r0 -= 2.0;
            x[q] = cmk(r0, r1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kR; ++q) lds_all[(t + C::GH * q) * C::RSS + c] = x[q];
    __syncthreads();
    const ExpWalk EW = ewalk_of(A);
    const NDiv D = ndiv_of(A);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        cy[i] = 0;
        if (!C::owns(h)) continue;
        T *row = lds_all + h * C::RSS;
        const u64 j0 = (u64)blk * CW + (u64)W * h;
        u32 e = EW.of_residue(rmod(IS_GL ? j0 : 2 * j0, A.p, D));
        CT cv = 0;
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
            if constexpr (IS_GL) {
                i64 c2;
                const i64 v = split_i((i64)row[cc] + cv, EW.bits(e), c2);
                e = EW.next(e);
                row[cc] = (u64)v;
                cv = c2;
            } else {
                double c1, c2;
                const int b0 = EW.bits(e);
                e = EW.next(e);
                const int b1 = EW.bits(e);
                e = EW.next(e);
                const double x0 = split_d(row[cc].re + cv, b0, c1);
                const double x1 = split_d(row[cc].im + c1, b1, c2);
                row[cc] = cmk(x0, x1);
                cv = c2;
            }
        }
        cy[i] = cv;
    }
}

// lds_all: row-segment layout; cin[i]: carry into the segment of row tid + NT i.  Adds the
// carry with a balanced re-split along the segment (the last element keeps what is left), returns
// to the column layout, weights, forward H-point transform, w_M twiddles, stores.
template <bool IS_GL, int W, int H, int CW, class TP>
LL_DEV void col_forward(typename Col<IS_GL, W, H, CW>::T *lds_all, TP tw,
                        const typename Col<IS_GL, W, H, CW>::WT (&wf)[nwt(IS_GL)],
                        const typename Col<IS_GL, W, H, CW>::T (&tm)[kR], const LargeArgs &A, int blk, int c, int t,
                        int tid, const typename Col<IS_GL, W, H, CW>::CT (&cin)[nrow(CW)]) {
    typedef Col<IS_GL, W, H, CW> C;
    typedef typename C::F F;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::S S;
    const ExpWalk EW = ewalk_of(A);
    const NDiv D = ndiv_of(A);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        if (!C::owns(h)) continue;
        T *row = lds_all + h * C::RSS;
        const u64 j0 = (u64)blk * CW + (u64)W * h;
        u32 e = EW.of_residue(rmod(IS_GL ? j0 : 2 * j0, A.p, D));
        CT cv = cin[i];
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
            if constexpr (IS_GL) {
                const int b = EW.bits(e);
                e = EW.next(e);
                i64 v = (i64)row[cc] + cv;
                if (cc == CW - 1) { row[cc] = (u64)v; break; }
                i64 c2;
                v = split_i(v, b, c2);
                row[cc] = (u64)v;
                cv = c2;
            } else {
                const int b0 = EW.bits(e);
                e = EW.next(e);
                const int b1 = EW.bits(e);
                e = EW.next(e);
                double c1, c2;
                const double x0 = split_d(row[cc].re + cv, b0, c1);
                double x1 = row[cc].im + c1;
                if (cc == CW - 1) { row[cc] = cmk(x0, x1); break; }
                x1 = split_d(x1, b1, c2);
                row[cc] = cmk(x0, x1);
                cv = c2;
            }
        }
    }
    __syncthreads();
    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) {
        const T v = lds_all[(t + C::GH * s) * C::RSS + c];
        if constexpr (IS_GL) x[s] = gl_mul(gl_from_i64((i64)v), wf[s]);
        else x[s] = cmk(v.re * wf[2 * s], v.im * wf[2 * s + 1]);
    }
    T *lds = lds_all + c * C::REG;
    if constexpr (S::MIXED)
        mixed_transform<F, H, C::GH, kR, false, S::R1, S::R2, S::R3, S::R4, true, kPS>(x, lds, t, tw);
    else
        transform<F, H, C::GH, kR, S::R1, S::R2, S::R3, S::R4, false, true, kPS>(x, lds, t, tw);
    T *elem = reinterpret_cast<T *>(A.elem);
    const int w = blk * CW + c;
#pragma unroll
    for (int q = 0; q < kR; ++q) {
        if constexpr (IS_GL) x[q] = gl_mul(x[q], tm[q]);
        else x[q] = cmul(x[q], tm[q]);
        st_stream(elem + elem_index(t + C::GH * q, w, W, H, C::E::LC), x[q]);
    }
}

// load the column block with the inverse w_M twiddles applied (FP64: conjugate of the forward)
template <bool IS_GL, int W, int H, int CW>
LL_DEV void col_load_inv(typename Col<IS_GL, W, H, CW>::T (&x)[kR], const typename Col<IS_GL, W, H, CW>::T (&tmi)[kR],
                         const LargeArgs &A, int w, int t) {
    typedef Col<IS_GL, W, H, CW> C;
    const typename C::T *elem = reinterpret_cast<const typename C::T *>(A.elem);
#pragma unroll
    for (int s = 0; s < kR; ++s) {
        x[s] = elem[elem_index(t + C::GH * s, w, W, H, C::E::LC)];
        if constexpr (IS_GL) x[s] = gl_mul(x[s], tmi[s]);
        else x[s] = cmulc(x[s], tmi[s]);
    }
}

// Lanes per segment of the round-off reduction of an NT-thread workgroup: the largest power of
// two up to 64 that divides NT, so every segment lies wholly inside the workgroup and no lane
// reads a lane that is not there.
LL_HD constexpr int seg_lanes(int nt) { return nt % 64 == 0 ? 64 : nt % 32 == 0 ? 32 : nt % 16 == 0 ? 16 : nt % 8 == 0 ? 8 : 1; }
// maximum of v over each segment of SEG lanes (every lane gets its segment's)
template <int SEG>
LL_DEV double seg_max(double v) {
#pragma unroll
    for (int off = SEG / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, SEG));
    return v;
}

template <int NT>
LL_DEV void err_publish(unsigned long long *s_err, double eloc, const LargeArgs &A, int tid) {
    constexpr int SEG = seg_lanes(NT);
    const double em = seg_max<SEG>(eloc);
    if (tid % SEG == 0) atomicMax(s_err, (unsigned long long)__double_as_longlong(em));
    __syncthreads();
    if (tid == 0) atomicMax(A.err + A.err_slot, *s_err);
}

// ---- K1: forward column pass (first squaring of a block: carry-in from the carry buffer) --------
template <bool IS_GL, int W, int H, int CW>
__global__ void __launch_bounds__(CW * (H / kR)) lg_cols_fwd(LargeArgs A) {
    typedef Col<IS_GL, W, H, CW> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    __shared__ __attribute__((aligned(16))) T lds_all[C::LDS];
    const int tid = threadIdx.x, c = tid % CW, t = tid / CW;
    const int blk = blockIdx.x, w = blk * CW + c;
    const T *elem = reinterpret_cast<const T *>(A.elem);
    const CT *carry = reinterpret_cast<const CT *>(A.carry);
    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[elem_index(t + C::GH * s, w, W, H, C::E::LC)];
    CT cin[C::NROW];
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        cin[i] = !C::owns(h) ? CT(0)
                 : blk > 0   ? carry[(blk - 1) * H + h]
                             : carry[(C::NBLK - 1) * H + (h > 0 ? h - 1 : H - 1)];
    }
    WT wf[C::NWT];
    col_weights<IS_GL, W, H, CW>(wf, gtab<WT>(A.wfL, A.wfH, A.log_sw), A, w, t);
    T tm[kR];
    col_twiddles<IS_GL, W, H, CW>(tm, gtab<T>(A.tmL, A.tmH, A.log_s), w, t);
#pragma unroll
    for (int s = 0; s < kR; ++s) lds_all[(t + C::GH * s) * C::RSS + c] = x[s];
    __syncthreads();
    col_forward<IS_GL, W, H, CW>(lds_all, as_global<T>(A.twH), wf, tm, A, blk, c, t, tid, cin);
}

// ---- K3: inverse column pass (last squaring of a block: carry-outs to the carry buffer) ---------
template <bool IS_GL, int W, int H, int CW>
__global__ void __launch_bounds__(CW * (H / kR)) lg_cols_inv(LargeArgs A) {
    typedef Col<IS_GL, W, H, CW> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    __shared__ __attribute__((aligned(16))) T lds_all[C::LDS];
    __shared__ unsigned long long s_err;
    const int tid = threadIdx.x, c = tid % CW, t = tid / CW;
    const int blk = blockIdx.x, w = blk * CW + c;
    if (tid == 0) s_err = 0;
    T tmi[kR];
    col_twiddles<IS_GL, W, H, CW>(tmi, gtab<T>(IS_GL ? A.tmLi : A.tmL, IS_GL ? A.tmHi : A.tmH, A.log_s), w, t);
    T x[kR];
    col_load_inv<IS_GL, W, H, CW>(x, tmi, A, w, t);
    WT wi[C::NWT];
    col_weights<IS_GL, W, H, CW>(wi, gtab<WT>(A.wiL, A.wiH, A.log_sw), A, w, t);
    CT cy[C::NROW];
    double eloc = 0.0;
    col_inverse<IS_GL, W, H, CW>(x, lds_all, as_global<T>(IS_GL ? A.twHi : A.twH), wi, A, blk, c, t, tid, cy, eloc);
    T *elem = reinterpret_cast<T *>(A.elem);
    CT *carry = reinterpret_cast<CT *>(A.carry);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        if (!C::owns(h)) continue;
        carry[blk * H + h] = cy[i];
        const T *row = lds_all + h * C::RSS;
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) elem[elem_index(h, blk * CW + cc, W, H, C::E::LC)] = row[cc];
    }
    if constexpr (!IS_GL) err_publish<C::NT>(&s_err, eloc, A, tid);
}

// ---- K3+K1 fused: inverse column pass of one squaring and forward column pass of the next --------
// Workgroup b runs the inverse pass of block b, publishes its row-segment carry-outs as tagged
// words, waits for block b - 1's (block 0: block NBLK - 1's, one row down, the top one wrapping
// to digit 0), adds them and runs the forward pass.  Progress: every workgroup publishes before
// it waits, so a waiting workgroup only needs its predecessor to be resident or dispatched
// later.  With the plain order, b waits on b - 1, which in-order dispatch has already placed,
// and block 0 on the last block, which needs one free CU once the others finish.  With the
// XCD-aware order (below) the predecessor of the workgroup at blockIdx x < 8 is dispatched near
// the end; the launch still drains as long as 9 workgroups fit on the GPU at once, because
// every workgroup that waits has published and the predecessors it waits on only wait on
// published blocks.  The launch's tag comes with its arguments, so nothing is waited on before
// the loads go out (a launch counter read with an atomic cost a device-scope round trip first).
// Column block of workgroup b in the fused pass.  Workgroups are dealt to the 8 XCDs in turn
// (b mod 8); mapping runs of NBLK / 8 consecutive blocks to one XCD keeps each block's
// predecessor, whose carries and flag it waits for, on the same XCD for all but 8 hand-offs,
// and neighbouring blocks, whose 64-byte row segments share HBM lines, on one L2.  Below 128
// blocks (a quarter of the CUs or fewer in use) the plain order measured faster, and so it did
// with several jobs' kernels running at once (LargeArgs::xcd_blocks is then 0).
template <int NBLK>
LL_DEV int xcd_block(int b) {
    if constexpr (NBLK % 8 == 0 && NBLK >= 128) return (b % 8) * (NBLK / 8) + b / 8;
    else return b;
}

template <bool IS_GL, int W, int H, int CW>
__global__ void __launch_bounds__(CW * (H / kR)) lg_cols_fused(LargeArgs A) {
    typedef Col<IS_GL, W, H, CW> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    constexpr int NBLK = C::NBLK;
    __shared__ __attribute__((aligned(16))) T lds_all[C::LDS];
    __shared__ unsigned long long s_err;
    const int tid = threadIdx.x, c = tid % CW, t = tid / CW;
    const int blk = A.xcd_blocks ? xcd_block<NBLK>(blockIdx.x) : (int)blockIdx.x, w = blk * CW + c;
    if (tid == 0) s_err = 0;
    // Every global load of the thread goes out before the first wait: the element, twiddle and
    // weight loads; each FFT pass then loads its table twiddles before its barrier.
    T tm[kR];
    col_twiddles<IS_GL, W, H, CW>(tm, gtab<T>(A.tmL, A.tmH, A.log_s), w, t);
    T x[kR];
    if constexpr (IS_GL) {
        T tmi[kR];
        col_twiddles<IS_GL, W, H, CW>(tmi, gtab<T>(A.tmLi, A.tmHi, A.log_s), w, t);
        col_load_inv<IS_GL, W, H, CW>(x, tmi, A, w, t);
    } else {
        col_load_inv<IS_GL, W, H, CW>(x, tm, A, w, t);
    }
    WT wi[C::NWT], wf[C::NWT];
    col_weights<IS_GL, W, H, CW>(wi, gtab<WT>(A.wiL, A.wiH, A.log_sw), A, w, t);
    col_weights<IS_GL, W, H, CW>(wf, gtab<WT>(A.wfL, A.wfH, A.log_sw), A, w, t);
    CT cy[C::NROW];
    double eloc = 0.0;
    col_inverse<IS_GL, W, H, CW>(x, lds_all, as_global<T>(IS_GL ? A.twHi : A.twH), wi, A, blk, c, t, tid, cy, eloc);
    const u32 seq = A.tag;
    u64 *tc = reinterpret_cast<u64 *>(A.tcarry);
    bool wide = false;
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        if (!C::owns(tid + C::NT * i)) continue;
        const i64 ci = carry_i(cy[i]);
        wide |= ci >= kCarryLim || ci < -kCarryLim;
        st_coh(tc + blk * H + tid + C::NT * i, tag_word(seq, ci));
    }
    if (wide) atomicOr(A.fault, 2u);            // a carry the tagged word cannot hold: fail the run
    // the workgroup's round-off maximum: reduced over segments of up to a wave first, then one
    // LDS atomic per segment (one per thread queued the whole workgroup's atomics on one LDS
    // word, in front of the forward pass's LDS traffic)
    if constexpr (!IS_GL) {
        constexpr int SEG = seg_lanes(C::NT);
        const double em = seg_max<SEG>(eloc);
        if (tid % SEG == 0) atomicMax(&s_err, (unsigned long long)__double_as_longlong(em));
    }
    // the first poll of every word is issued before any is waited on: one round trip, not NROW
    CT cin[C::NROW];
    u64 cwd[C::NROW];
    int src[C::NROW];
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        src[i] = blk > 0 ? (blk - 1) * H + h : (NBLK - 1) * H + (h > 0 ? h - 1 : H - 1);
        cwd[i] = C::owns(h) ? ld_coh_u(tc + src[i]) : tag_word(seq, 0);
    }
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        if (!tag_ok(cwd[i], seq)) cwd[i] = wait_tagged(tc + src[i], seq, A.fault);
        carry_from(untag(cwd[i]), cin[i]);
    }
    col_forward<IS_GL, W, H, CW>(lds_all, as_global<T>(A.twH), wf, tm, A, blk, c, t, tid, cin);
    // the workgroup's round-off maximum (s_err, complete after the barriers of col_forward) goes
    // out last, off the path of the carry hand-off, with no barrier of its own
    if constexpr (!IS_GL)
        if (tid == 0) atomicMax(A.err + A.err_slot, s_err);
}

template <class F, int W, int GW, bool INV, class TP>
LL_DEV void row_transform(typename F::T (&x)[kR], typename F::T *lds, int t, TP tw) {
    typedef RowSched<W> RS;
    if constexpr (RS::HYBRID) {
        hybrid_transform<F, W, GW, kR, INV, RS::M1, RS::M2, RS::M3, true, kPS>(x, lds, t, tw);
    } else {
        typedef Sched<ilog2c(W)> S;
        transform<F, W, GW, kR, S::R1, S::R2, S::R3, S::R4, INV, true, kPS>(x, lds, t, tw);
    }
}

// ---- K2 -----------------------------------------------------------------------------------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(2 * (W / kR)) lg_rows(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    constexpr int GW = W / kR, NT = 2 * GW;
    constexpr int REG = W + W / 8 + 4;
    __shared__ __attribute__((aligned(16))) T lds_all[2 * REG];
    const int tid = threadIdx.x, g = tid / GW, t = tid % GW;
    const int i = blockIdx.x;
    const int row = g == 0 ? i : (i == 0 ? H / 2 : H - i);
    T *lds = lds_all + g * REG;
    // rows i and H - i (i = 0: rows 0 and H/2) are slots 0 and 1 of row pair i's lines
    T *elem = reinterpret_cast<T *>(A.elem) + (size_t)i * 2 * W + g * E::LC;
    static_assert(GW % E::LC == 0, "a thread's elements must start a half line alike");
    const int tl = (t / (int)E::LC) * 2 * (int)E::LC + t % (int)E::LC;   // element t of the row

    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[tl + 2 * GW * s];
    C64 wk[IS_GL ? 1 : kR];
    if constexpr (!IS_GL) {
#pragma unroll
        for (int q = 0; q < kR; ++q) wk[q] = tw2<C64>(A.tmL, A.tmH, (u32)row + (u32)H * (u32)(t + GW * q), A.log_s);
    }
    row_transform<F, W, GW, false>(x, lds, t, as_global<T>(A.twW));
    if constexpr (IS_GL) {
#pragma unroll
        for (int q = 0; q < kR; ++q) x[q] = gl_sqr(x[q]);
        row_transform<F, W, GW, true>(x, lds, t, as_global<T>(A.twWi));
    } else {
        const int tt = ll_opaque(t);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kR; ++q) lds[tt + GW * q] = x[q];     // unpadded: contiguous stores, mirrored loads
        __syncthreads();
        // partner of (row, k_w): row 0 -> (0, -k_w); row H/2 -> (H/2, W-1-k_w);
        // otherwise -> (H - row, W-1-k_w), held by the other group of this workgroup
        const T *plds = (row == 0 || row == H / 2) ? lds : lds_all + (1 - g) * REG;
        const int kwb = (row == 0 ? W : W - 1) - tt;
#pragma unroll
        for (int q = 0; q < kR; ++q) {
            // k_w = tt + GW q is 0 only for q = 0, tt = 0 (row 0 pairs it with itself)
            const int kwp = q == 0 && row == 0 && tt == 0 ? 0 : kwb - GW * q;
            const C64 zm = plds[kwp];
            const C64 D = csub(x[q], cconj(zm));
            const C64 AD = cmul(cmk(1.0 + wk[q].re, wk[q].im), csqr(D));
            const C64 Z2 = csqr(x[q]);
            x[q] = cmk(__fma_rn(-0.25, AD.re, Z2.re), __fma_rn(-0.25, AD.im, Z2.im));
        }
        row_transform<F, W, GW, true>(x, lds, t, as_global<T>(A.twW));
    }
#pragma unroll
    for (int s = 0; s < kR; ++s) st_stream(elem + tl + 2 * GW * s, x[s]);
}

// ---- carry-in only (after the last squaring, before normalisation) ----------------------------
template <bool IS_GL, int W, int H, int CW>
__global__ void __launch_bounds__(256) lg_apply_carry(LargeArgs A) {
    typedef ElemOf<IS_GL> E;
    typedef typename E::T T;
    typedef typename E::CT CT;
    constexpr int NBLK = W / CW;
    const int blk = blockIdx.x;
    T *elem = reinterpret_cast<T *>(A.elem);
    const CT *carry = reinterpret_cast<const CT *>(A.carry);
    const u32 p = A.p;
    const NDiv log_n = ndiv_of(A);
    for (int h = threadIdx.x; h < H; h += 256) {
        CT cin;
        if (blk > 0) cin = carry[(blk - 1) * H + h];
        else cin = h > 0 ? carry[(NBLK - 1) * H + h - 1] : carry[(NBLK - 1) * H + H - 1];
        const u64 j0 = (u64)blk * CW + (u64)W * h;
        for (int cc = 0; cc < CW; ++cc) {
            T *row = elem + elem_index(h, blk * CW + cc, W, H, E::LC) - cc;   // row[cc]: element (h, blk CW + cc)
            if constexpr (IS_GL) {
                i64 v = (i64)row[cc] + cin;
                if (cc == CW - 1) { row[cc] = (u64)v; break; }
                i64 c2;
                v = split_i(v, bits_of(j0 + cc, p, log_n), c2);
                row[cc] = (u64)v;
                cin = c2;
            } else {
                const u64 d0 = 2 * (j0 + cc);
                double c1, c2;
                const double x0 = split_d(row[cc].re + cin, bits_of(d0, p, log_n), c1);
                double x1 = row[cc].im + c1;
                if (cc == CW - 1) { row[cc] = cmk(x0, x1); break; }
                x1 = split_d(x1, bits_of(d0 + 1, p, log_n), c2);
                row[cc] = cmk(x0, x1);
                cin = c2;
            }
        }
    }
}

}  // namespace

namespace {
// The five kernels of one schedule (the row pass depends on the split only) and the order of its
// twiddle tables.
template <bool IS_GL, int W, int H, int CW>
llh::LargeLaunch make_large_launch() {
    return llh::LargeLaunch{lg_cols_fwd<IS_GL, W, H, CW>, lg_rows<IS_GL, W, H>, lg_cols_inv<IS_GL, W, H, CW>,
                            lg_cols_fused<IS_GL, W, H, CW>, lg_apply_carry<IS_GL, W, H, CW>, W, H, CW,
                            row_pass_exponent<W>, col_pass_exponent<H>, row_pass_size<W>(), col_pass_size<H>()};
}
}  // namespace
