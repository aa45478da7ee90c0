// ll_large.hip -- Lucas-Lehmer for one large exponent: the residue lives in HBM and each
// squaring is two kernel launches of a four-step transform.
//
// Counterpart of the FP64 loop lucaslehmer.cpp:281-311 (and of the NTT loop of
// cuda/ll_ntt.cu:91-163, which the reference only runs up to N = 8192) for transform lengths
// 2^15 .. 2^21 elements.  Layout: the M transform elements form an H x W matrix, element
// j = w + W h (row h, column w; rows contiguous).  With k = k_h + H k_w:
//
//   Z[k] = sum_w w_W^{w k_w} w_M^{w k_h} sum_h z[w + W h] w_H^{h k_h}
//
//  rows       (rows k_h and H - k_h together): W-point FFT per row, square (FP64: the packed
//             real spectrum pairs row k_h with row H - k_h), inverse W-point FFT.
//  cols_fused (blocks of kCW columns): twiddle w_M^{-w k_h}, inverse H-point FFT, unweight,
//             round, per-iteration max round-off, -2, an exact balanced carry along each row
//             segment of kCW elements; the segment's carry-out is handed to the next block
//             inside the launch (the last segment of row h feeds row h + 1, and the top one
//             wraps to digit 0: 2^p == 1); then the carry-in with a re-split, weights, the
//             forward H-point FFT and the twiddle w_M^{w k_h} of the next squaring.
//  cols_fwd / cols_inv: the forward and the inverse half alone, which open and close a block
//             of squarings (the carries between blocks wait in a buffer).
//
// Digits: FP64 packs digits 2j, 2j+1 into element j (N = 2M); GL holds digit j in element j
// (N = M).  Weights and the w_M twiddles come from two-level tables (product of a low and
// a high table entry), so every table stays cache resident at every length; the FFT
// twiddles are staged in LDS in per-pass order.
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
};

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

// Tagged carry words of the fused column pass: bits 63..48 hold the launch's sequence number
// (mod 2^16), bits 47..0 the signed carry.  Every launch rewrites every word, so a word holds
// either this launch's value or the one before's, whose tag differs; a reader polls its own
// words until the tag matches, with no flag and no barrier between publishing and reading.
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

// Per-pass twiddle tables of the H-point and W-point transforms, copied into LDS out of the full
// tables (all NT threads of the workgroup).
template <class T, int H, int NT>
LL_DEV void stage_col_table(T *dst, const void *src, int tid) {
    typedef ColRadix<H> RR;
    stage_pass_table<T, NT, RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>(dst, src, tid);
}
template <class T, int W, int NT>
LL_DEV void stage_row_table(T *dst, const void *src, int tid) {
    typedef RowRadix<W> RR;
    stage_pass_table<T, NT, RR::R1, RR::R2, RR::R3, RR::R4, RR::R5>(dst, src, tid);
}

// ---- column kernels ---------------------------------------------------------------------------
// A workgroup owns a block of kCW adjacent columns (W / kCW blocks).  Thread (c, t), with
// c = tid % kCW, holds column w = blk kCW + c at rows t + GH q: the column layout of the H-point
// transform.  Between the inverse and the forward transform the digits sit in LDS in the
// row-segment layout lds_all[h kCW + c] (a row segment is kCW elements, 2 kCW digits for FP64,
// contiguous digits), and thread tid walks the segments of rows h = tid + NT i.
constexpr int kNRow = kR / kCW;                              // row segments per thread
LL_HD constexpr int nwt(bool gl) { return gl ? kR : 2 * kR; }  // digit weights per thread

template <bool IS_GL, int W, int H>
struct Col {
    typedef ElemOf<IS_GL> E;
    typedef typename E::F F;
    typedef typename E::T T;
    typedef typename E::CT CT;
    typedef typename E::WT WT;
    typedef ColSched<H> S;
    static constexpr int GH = H / kR, NT = kCW * GH, NBLK = W / kCW;
    static constexpr int REG = H + H / 8 + (IS_GL ? 8 : 4);
    static constexpr int NROW = kNRow;
    static constexpr int NWT = nwt(IS_GL);
    static_assert(H == NT * NROW, "row segments must split evenly over the workgroup");
};

// w_M^{w k_h} at the thread's positions k_h = t + GH q (table: forward or GL inverse); the
// exponent w k_h < W H = M needs no reduction
template <bool IS_GL, int W, int H, class TAB>
LL_DEV void col_twiddles(typename Col<IS_GL, W, H>::T (&tm)[kR], const TAB &tab, int w, int t) {
    typedef Col<IS_GL, W, H> C;
#pragma unroll
    for (int q = 0; q < kR; ++q) tm[q] = tab((u32)w * (u32)(t + C::GH * q));
}

// Digit weights (table: forward or inverse) at the thread's elements j = w + W (t + GH q);
// FP64 elements hold digits 2j and 2j + 1.
template <bool IS_GL, int W, int H, class TAB>
LL_DEV void col_weights(typename Col<IS_GL, W, H>::WT (&wt)[nwt(IS_GL)], const TAB &tab, const LargeArgs &A,
                        int w, int t) {
    typedef Col<IS_GL, W, H> C;
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
template <bool IS_GL, int W, int H, class TP>
LL_DEV void col_inverse(typename Col<IS_GL, W, H>::T (&x)[kR], typename Col<IS_GL, W, H>::T *lds_all, TP tw,
                        const typename Col<IS_GL, W, H>::WT (&wi)[nwt(IS_GL)], const LargeArgs &A,
                        int blk, int c, int t, int tid,
                        typename Col<IS_GL, W, H>::CT (&cy)[kNRow], double &eloc) {
    typedef Col<IS_GL, W, H> C;
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
    for (int q = 0; q < kR; ++q) lds_all[(t + C::GH * q) * kCW + c] = x[q];
    __syncthreads();
    const ExpWalk EW = ewalk_of(A);
    const NDiv D = ndiv_of(A);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        T *row = lds_all + h * kCW;
        const u64 j0 = (u64)blk * kCW + (u64)W * h;
        u32 e = EW.of_residue(rmod(IS_GL ? j0 : 2 * j0, A.p, D));
        CT cv = 0;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
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
template <bool IS_GL, int W, int H, class TP>
LL_DEV void col_forward(typename Col<IS_GL, W, H>::T *lds_all, TP tw,
                        const typename Col<IS_GL, W, H>::WT (&wf)[nwt(IS_GL)],
                        const typename Col<IS_GL, W, H>::T (&tm)[kR], const LargeArgs &A, int blk, int c, int t,
                        int tid, const typename Col<IS_GL, W, H>::CT (&cin)[kNRow]) {
    typedef Col<IS_GL, W, H> C;
    typedef typename C::F F;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::S S;
    const ExpWalk EW = ewalk_of(A);
    const NDiv D = ndiv_of(A);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        T *row = lds_all + h * kCW;
        const u64 j0 = (u64)blk * kCW + (u64)W * h;
        u32 e = EW.of_residue(rmod(IS_GL ? j0 : 2 * j0, A.p, D));
        CT cv = cin[i];
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) {
            if constexpr (IS_GL) {
                const int b = EW.bits(e);
                e = EW.next(e);
                i64 v = (i64)row[cc] + cv;
                if (cc == kCW - 1) { row[cc] = (u64)v; break; }
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
                if (cc == kCW - 1) { row[cc] = cmk(x0, x1); break; }
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
        const T v = lds_all[(t + C::GH * s) * kCW + c];
        if constexpr (IS_GL) x[s] = gl_mul(gl_from_i64((i64)v), wf[s]);
        else x[s] = cmk(v.re * wf[2 * s], v.im * wf[2 * s + 1]);
    }
    T *lds = lds_all + c * C::REG;
    if constexpr (S::MIXED)
        mixed_transform<F, H, C::GH, kR, false, S::R1, S::R2, S::R3, S::R4, true, kPS>(x, lds, t, tw);
    else
        transform<F, H, C::GH, kR, S::R1, S::R2, S::R3, S::R4, false, true, kPS>(x, lds, t, tw);
    T *elem = reinterpret_cast<T *>(A.elem);
    const int w = blk * kCW + c;
#pragma unroll
    for (int q = 0; q < kR; ++q) {
        if constexpr (IS_GL) x[q] = gl_mul(x[q], tm[q]);
        else x[q] = cmul(x[q], tm[q]);
        st_stream(elem + w + W * (t + C::GH * q), x[q]);
    }
}

// load the column block with the inverse w_M twiddles applied (FP64: conjugate of the forward)
template <bool IS_GL, int W, int H>
LL_DEV void col_load_inv(typename Col<IS_GL, W, H>::T (&x)[kR], const typename Col<IS_GL, W, H>::T (&tmi)[kR],
                         const LargeArgs &A, int w, int t) {
    typedef Col<IS_GL, W, H> C;
    const typename C::T *elem = reinterpret_cast<const typename C::T *>(A.elem);
#pragma unroll
    for (int s = 0; s < kR; ++s) {
        x[s] = elem[w + W * (t + C::GH * s)];
        if constexpr (IS_GL) x[s] = gl_mul(x[s], tmi[s]);
        else x[s] = cmulc(x[s], tmi[s]);
    }
}

LL_DEV void err_publish(unsigned long long *s_err, double eloc, const LargeArgs &A, int tid) {
    atomicMax(s_err, (unsigned long long)__double_as_longlong(eloc));
    __syncthreads();
    if (tid == 0) atomicMax(A.err + A.err_slot, *s_err);
}

// ---- K1: forward column pass (first squaring of a block: carry-in from the carry buffer) --------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(kCW * (H / kR)) lg_cols_fwd(LargeArgs A) {
    typedef Col<IS_GL, W, H> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    __shared__ __attribute__((aligned(16))) T lds_all[kCW * C::REG];
    __shared__ __attribute__((aligned(16))) T s_tw[H];
    const int tid = threadIdx.x, c = tid % kCW, t = tid / kCW;
    const int blk = blockIdx.x, w = blk * kCW + c;
    const T *elem = reinterpret_cast<const T *>(A.elem);
    const CT *carry = reinterpret_cast<const CT *>(A.carry);
    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[w + W * (t + C::GH * s)];
    CT cin[C::NROW];
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        cin[i] = blk > 0 ? carry[(blk - 1) * H + h] : carry[(C::NBLK - 1) * H + (h > 0 ? h - 1 : H - 1)];
    }
    WT wf[C::NWT];
    col_weights<IS_GL, W, H>(wf, gtab<WT>(A.wfL, A.wfH, A.log_sw), A, w, t);
    T tm[kR];
    col_twiddles<IS_GL, W, H>(tm, gtab<T>(A.tmL, A.tmH, A.log_s), w, t);
    stage_col_table<T, H, C::NT>(s_tw, A.twH, tid);
#pragma unroll
    for (int s = 0; s < kR; ++s) lds_all[(t + C::GH * s) * kCW + c] = x[s];
    __syncthreads();
    col_forward<IS_GL, W, H>(lds_all, (const T *)s_tw, wf, tm, A, blk, c, t, tid, cin);
}

// ---- K3: inverse column pass (last squaring of a block: carry-outs to the carry buffer) ---------
template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(kCW * (H / kR)) lg_cols_inv(LargeArgs A) {
    typedef Col<IS_GL, W, H> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    __shared__ __attribute__((aligned(16))) T lds_all[kCW * C::REG];
    __shared__ __attribute__((aligned(16))) T s_tw[H];
    __shared__ unsigned long long s_err;
    const int tid = threadIdx.x, c = tid % kCW, t = tid / kCW;
    const int blk = blockIdx.x, w = blk * kCW + c;
    if (tid == 0) s_err = 0;
    T tmi[kR];
    col_twiddles<IS_GL, W, H>(tmi, gtab<T>(IS_GL ? A.tmLi : A.tmL, IS_GL ? A.tmHi : A.tmH, A.log_s), w, t);
    T x[kR];
    col_load_inv<IS_GL, W, H>(x, tmi, A, w, t);
    WT wi[C::NWT];
    col_weights<IS_GL, W, H>(wi, gtab<WT>(A.wiL, A.wiH, A.log_sw), A, w, t);
    stage_col_table<T, H, C::NT>(s_tw, IS_GL ? A.twHi : A.twH, tid);
    CT cy[C::NROW];
    double eloc = 0.0;
    col_inverse<IS_GL, W, H>(x, lds_all, (const T *)s_tw, wi, A, blk, c, t, tid, cy, eloc);
    T *elem = reinterpret_cast<T *>(A.elem);
    CT *carry = reinterpret_cast<CT *>(A.carry);
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        carry[blk * H + h] = cy[i];
        const T *row = lds_all + h * kCW;
#pragma unroll
        for (int cc = 0; cc < kCW; ++cc) elem[blk * kCW + cc + (size_t)W * h] = row[cc];
    }
    if constexpr (!IS_GL) err_publish(&s_err, eloc, A, tid);
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
// published blocks.  The tag is the launch's sequence number: every workgroup adds one to a
// counter that only grows, so ticket / NBLK + 1 is the same for all workgroups of one launch
// and one more than the launch before's.
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

template <bool IS_GL, int W, int H>
__global__ void __launch_bounds__(kCW * (H / kR)) lg_cols_fused(LargeArgs A) {
    typedef Col<IS_GL, W, H> C;
    typedef typename C::T T;
    typedef typename C::CT CT;
    typedef typename C::WT WT;
    constexpr int NBLK = C::NBLK;
    __shared__ __attribute__((aligned(16))) T lds_all[kCW * C::REG];
    __shared__ __attribute__((aligned(16))) T s_tw[H];
    __shared__ __attribute__((aligned(16))) T s_twi[IS_GL ? H : 1];
    __shared__ unsigned long long s_err;
    __shared__ u32 s_seq;
    const int tid = threadIdx.x, c = tid % kCW, t = tid / kCW;
    const int blk = A.xcd_blocks ? xcd_block<NBLK>(blockIdx.x) : (int)blockIdx.x, w = blk * kCW + c;
    if (tid == 0) {
        s_seq = (u32)(atomicAdd(A.ticket, 1ull) / (unsigned long long)NBLK) + 1u;
        s_err = 0;
    }
    stage_col_table<T, H, C::NT>(s_tw, A.twH, tid);
    if constexpr (IS_GL) stage_col_table<T, H, C::NT>(s_twi, A.twHi, tid);
    T tm[kR];
    col_twiddles<IS_GL, W, H>(tm, gtab<T>(A.tmL, A.tmH, A.log_s), w, t);
    T x[kR];
    if constexpr (IS_GL) {
        T tmi[kR];
        col_twiddles<IS_GL, W, H>(tmi, gtab<T>(A.tmLi, A.tmHi, A.log_s), w, t);
        col_load_inv<IS_GL, W, H>(x, tmi, A, w, t);
    } else {
        col_load_inv<IS_GL, W, H>(x, tm, A, w, t);
    }
    WT wi[C::NWT], wf[C::NWT];
    col_weights<IS_GL, W, H>(wi, gtab<WT>(A.wiL, A.wiH, A.log_sw), A, w, t);
    col_weights<IS_GL, W, H>(wf, gtab<WT>(A.wfL, A.wfH, A.log_sw), A, w, t);
    CT cy[C::NROW];
    double eloc = 0.0;
    col_inverse<IS_GL, W, H>(x, lds_all, (const T *)(IS_GL ? s_twi : s_tw), wi, A, blk, c, t, tid, cy, eloc);
    const u32 seq = s_seq;   // written before the barriers inside col_inverse
    u64 *tc = reinterpret_cast<u64 *>(A.tcarry);
    bool wide = false;
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const i64 ci = carry_i(cy[i]);
        wide |= ci >= kCarryLim || ci < -kCarryLim;
        st_coh(tc + blk * H + tid + C::NT * i, tag_word(seq, ci));
    }
    if (wide) atomicOr(A.fault, 2u);            // a carry the tagged word cannot hold: fail the run
    if constexpr (!IS_GL) atomicMax(&s_err, (unsigned long long)__double_as_longlong(eloc));
    // the first poll of every word is issued before any is waited on: one round trip, not NROW
    CT cin[C::NROW];
    u64 cw[C::NROW];
    int src[C::NROW];
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        const int h = tid + C::NT * i;
        src[i] = blk > 0 ? (blk - 1) * H + h : (NBLK - 1) * H + (h > 0 ? h - 1 : H - 1);
        cw[i] = ld_coh_u(tc + src[i]);
    }
#pragma unroll
    for (int i = 0; i < C::NROW; ++i) {
        if (!tag_ok(cw[i], seq)) cw[i] = wait_tagged(tc + src[i], seq, A.fault);
        carry_from(untag(cw[i]), cin[i]);
    }
    col_forward<IS_GL, W, H>(lds_all, (const T *)s_tw, wf, tm, A, blk, c, t, tid, cin);
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
    __shared__ __attribute__((aligned(16))) T s_tw[W];
    __shared__ __attribute__((aligned(16))) T s_twi[IS_GL ? W : 1];
    const int tid = threadIdx.x, g = tid / GW, t = tid % GW;
    const int i = blockIdx.x;
    const int row = g == 0 ? i : (i == 0 ? H / 2 : H - i);
    T *lds = lds_all + g * REG;
    T *elem = reinterpret_cast<T *>(A.elem) + (size_t)W * row;

    T x[kR];
#pragma unroll
    for (int s = 0; s < kR; ++s) x[s] = elem[t + GW * s];
    C64 wk[IS_GL ? 1 : kR];
    if constexpr (!IS_GL) {
#pragma unroll
        for (int q = 0; q < kR; ++q) wk[q] = tw2<C64>(A.tmL, A.tmH, (u32)row + (u32)H * (u32)(t + GW * q), A.log_s);
    }
    stage_row_table<T, W, NT>(s_tw, A.twW, tid);
    if constexpr (IS_GL) stage_row_table<T, W, NT>(s_twi, A.twWi, tid);
    row_transform<F, W, GW, false>(x, lds, t, (const T *)s_tw);
    if constexpr (IS_GL) {
#pragma unroll
        for (int q = 0; q < kR; ++q) x[q] = gl_sqr(x[q]);
        row_transform<F, W, GW, true>(x, lds, t, (const T *)s_twi);
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
        row_transform<F, W, GW, true>(x, lds, t, (const T *)s_tw);
    }
#pragma unroll
    for (int s = 0; s < kR; ++s) st_stream(elem + t + GW * s, x[s]);
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

namespace {

// One thread per output digit; the output is the element array of the N'-digit engine (FP64:
// N' doubles, element j = digits 2j, 2j + 1; GL: N' words holding the signed digit).
template <bool IS_GL>
__global__ void __launch_bounds__(256) lg_relayout(const i64 *dig, u32 p, u64 n, u64 n2, void *elem2) {
    const u64 k = (u64)blockIdx.x * 256ull + threadIdx.x;
    if (k >= n2) return;
    const i64 v = relayout_digit(dig, p, n, n2, k);
    if constexpr (IS_GL) reinterpret_cast<u64 *>(elem2)[k] = (u64)v;
    else reinterpret_cast<double *>(elem2)[k] = (double)v;
}

}  // namespace

namespace llh {

hipError_t large_relayout(int transform, const i64 *dig, u32 p, u64 n, u64 n2, void *elem2, hipStream_t st) {
    const dim3 grid((unsigned)((n2 + 255) / 256));
    if (transform == kNTT) hipLaunchKernelGGL(lg_relayout<true>, grid, dim3(256), 0, st, dig, p, n, n2, elem2);
    else hipLaunchKernelGGL(lg_relayout<false>, grid, dim3(256), 0, st, dig, p, n, n2, elem2);
    return hipGetLastError();
}

void relayout_host(const i64 *dig, u64 p, u64 n, u64 n2, i64 *out) {
    for (u64 k = 0; k < n2; ++k) out[k] = relayout_digit(dig, p, n, n2, k);
}

template <bool IS_GL, int W, int H>
static LargeLaunch make_launch() {
    return LargeLaunch{lg_cols_fwd<IS_GL, W, H>, lg_rows<IS_GL, W, H>, lg_cols_inv<IS_GL, W, H>,
                       lg_cols_fused<IS_GL, W, H>, lg_apply_carry<IS_GL, W, H>, W, H};
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

int large_blocks(const LargeLaunch &L) { return L.w_len / kCW; }

// k squarings: forward column pass, then per squaring the row pass and the fused column pass
// (the last squaring ends with the inverse column pass, leaving its carries in the carry buffer).
// Squaring i reports its round-off in err slot i.
hipError_t large_squarings(const LargeLaunch &L, LargeArgs A, int k, hipStream_t st) {
    const int W = L.w_len, H = L.h_len;
    const dim3 gcol(W / kCW), bcol(kCW * (H / kR));
    const dim3 grow(H / 2), brow(2 * (W / kR));
    hipLaunchKernelGGL(L.fwd, gcol, bcol, 0, st, A);
    for (int i = 0; i < k; ++i) {
        A.err_slot = (u32)i;
        hipLaunchKernelGGL(L.rows, grow, brow, 0, st, A);
        if (i + 1 < k) hipLaunchKernelGGL(L.fused, gcol, bcol, 0, st, A);
        else hipLaunchKernelGGL(L.inv, gcol, bcol, 0, st, A);
    }
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
