// ll_batch.hip -- the batched, LDS-resident Lucas-Lehmer engine: every exponent of a list
// in a few persistent launches, whatever its transform length.
//
// Counterpart of cuda/ll_ntt.cu:74-167 (ll_kernel, one block per exponent and one launch
// per length group) and of the FP64 loop lucaslehmer.cpp:281-311, re-designed for gfx950:
//
//  * Size classes: a class has M transform elements (FP64: N = 2M digits, GL: N = M) and is
//    run by groups of G = M/R threads (R elements per thread, R = 8; R = 4 for M < 64).  M is a
//    power of two (16 .. 4096) or 8 m 8 with m = 3, 5, 7 times a power of two (FP64 only).
//  * Families: classes share a workgroup size that their G divide, one persistent kernel per
//    family: powers of two up to M = 2048 and the one-word lane 256 (one M = 2048 exponent per
//    workgroup, so its barriers wait on its own waves), 3 * 2^k 384, 5 * 2^k 320, 7 * 2^k 448,
//    and M = 4096 512.  A workgroup holds WG/G exponents of a class at once, all stepping in
//    lock-step.
//  * Workgroups pull exponents from one queue per class (an atomic counter over items the
//    host sorted longest-first), classes in descending length order, then the one-word lane
//    (p < 64).  The families' launches run on separate streams, so one family's tail is
//    back-filled by the next.
//  * The residue lives in registers in a "column" layout (thread t holds elements t + G s);
//    only the transform's middle hand-offs, the real-FFT pairing and the carry hand-off go
//    through LDS.  Nothing touches HBM between the first and the last squaring except the
//    read-only twiddle tables (L2-resident, per-pass order so every wave load is contiguous);
//    the FP64 digit weights come from two-level tables staged in LDS.
//  * FP64 (FC64): the N weighted real digits are packed as N/2 complex numbers
//    z_j = y_{2j} + i y_{2j+1}; after the forward complex FFT the spectrum of the real
//    signal is squared in packed form:  Z'_k = Z_k^2 - (1 + w_M^k) (Z_k - conj Z_{M-k})^2 / 4,
//    whose inverse complex FFT is the packed cyclic self-convolution.
//  * Goldilocks (FGL): an N-point NTT, pointwise square, inverse NTT; exact.
//  * Balanced carry (lucaslehmer.cpp:297-310): computed in parallel.  Each digit rounds and
//    splits locally, carries move one element to the right through LDS (the top element's
//    carry wraps to digit 0 because 2^p == 1), and a second local split leaves every digit
//    within a few units of balanced.  The representation stays exact; the final residue is
//    normalised exactly once, on the device, at the end of the test.
#include "ll_common.hpp"
#include "ll_dft.hpp"
#include "ll_internal.hpp"
#include <type_traits>
#include <vector>

namespace {

constexpr u32 kNone = 0xFFFFFFFFu;
constexpr double kErrLimit = 0.4;   // round-off above this aborts the exponent (rollback longer)

constexpr int kMaxGroups = 64;     // WG / min G (256 / 4; the other families need fewer)

struct Ctl {
    u32 item[kMaxGroups];
    u32 bad[kMaxGroups];
    u32 rb[kMaxGroups];
    unsigned long long err[kMaxGroups];
    u32 ev, any;
};
constexpr u32 kCheck = 64;          // squarings between round-off checks

// Size classes.  Powers of two: radices R1 .. R4 (first = last = R; 0 = no pass).  Others:
// the hybrid schedule 8, H1, [H2], 8 (H1 odd; register radix-8 first and last passes, the
// middle ones through LDS).  WG: the family's workgroup size.
struct Cls {
    int M, G, R, WG;
    int R1, R2, R3, R4;
    int H1, H2;
};
constexpr Cls kCls[] = {
    {0, 0, 0, 256, 0, 0, 0, 0, 0, 0},          // 0: one-word lane (p < 64)
    {16, 4, 4, 256, 4, 4, 0, 0, 0, 0},         // 1
    {32, 8, 4, 256, 4, 2, 4, 0, 0, 0},         // 2
    {64, 8, 8, 256, 8, 8, 0, 0, 0, 0},         // 3
    {128, 16, 8, 256, 8, 2, 8, 0, 0, 0},       // 4
    {256, 32, 8, 256, 8, 4, 8, 0, 0, 0},       // 5
    {512, 64, 8, 256, 8, 8, 8, 0, 0, 0},       // 6
    {1024, 128, 8, 256, 8, 2, 8, 8, 0, 0},     // 7
    {2048, 256, 8, 256, 8, 4, 8, 8, 0, 0},     // 8
    {4096, 512, 8, 512, 8, 8, 8, 8, 0, 0},     // 9
    {192, 24, 8, 384, 0, 0, 0, 0, 3, 0},       // 10
    {384, 48, 8, 384, 0, 0, 0, 0, 3, 2},       // 11
    {768, 96, 8, 384, 0, 0, 0, 0, 3, 4},       // 12
    {1536, 192, 8, 384, 0, 0, 0, 0, 3, 8},     // 13
    {3072, 384, 8, 384, 0, 0, 0, 0, 3, 16},    // 14
    {320, 40, 8, 320, 0, 0, 0, 0, 5, 0},       // 15
    {640, 80, 8, 320, 0, 0, 0, 0, 5, 2},       // 16
    {1280, 160, 8, 320, 0, 0, 0, 0, 5, 4},     // 17
    {2560, 320, 8, 320, 0, 0, 0, 0, 5, 8},     // 18
    {448, 56, 8, 448, 0, 0, 0, 0, 7, 0},       // 19
    {896, 112, 8, 448, 0, 0, 0, 0, 7, 2},      // 20
    {1792, 224, 8, 448, 0, 0, 0, 0, 7, 4},     // 21
    {3584, 448, 8, 448, 0, 0, 0, 0, 7, 8},     // 22
    // Schedule variants: the same length with the middle radices in another order (another
    // pass span L for each, so other twiddle and LDS patterns).  The host runs a length at the
    // variant the autotuner measured fastest (kClsBase maps a variant to its length's class).
    {1024, 128, 8, 256, 8, 8, 2, 8, 0, 0},     // 23: class 7 as 8 8 2 8
    {2048, 256, 8, 256, 8, 8, 4, 8, 0, 0},     // 24: class 8 as 8 8 4 8
    {1280, 160, 8, 320, 0, 0, 0, 0, 4, 5},     // 25: class 17 as 8 4 5 8
    {2560, 320, 8, 320, 0, 0, 0, 0, 8, 5},     // 26: class 18 as 8 8 5 8
};
constexpr int kNumCls = (int)(sizeof(kCls) / sizeof(kCls[0]));
constexpr int kClsBase[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                            7, 8, 17, 18};
static_assert(sizeof(kClsBase) / sizeof(kClsBase[0]) == kNumCls, "every class has a base class");
// Kernel families: 0 the powers of two up to M = 2048 and the one-word lane (one exponent of
// M = 2048 per 4-wave workgroup, so a barrier waits on that exponent's waves only), 1 3 * 2^k,
// 2 5 * 2^k, 3 7 * 2^k, 4 the power of two M = 4096 (G = 512).
constexpr int kNumFam = 5;
constexpr int kFamilyWG[kNumFam] = {256, 384, 320, 448, 512};

// radix list of class C for the per-pass twiddle table (PassTable<R1 .. R5>)
template <int C>
struct ClsRadix {
    static constexpr Cls K = kCls[C];
    static constexpr bool P2 = K.H1 == 0;
    static constexpr int R1 = P2 ? K.R1 : 8;
    static constexpr int R2 = P2 ? K.R2 : K.H1;
    static constexpr int R3 = P2 ? K.R3 : (K.H2 ? K.H2 : 8);
    static constexpr int R4 = P2 ? K.R4 : (K.H2 ? 8 : 0);
    typedef PassTable<R1, R2, R3, R4, 0> PT;
    static_assert(PT::M == K.M, "schedule must cover the class length");
};

// LDS of a family's kernel: every group's residue buffer (M + M/R padded elements of 16 bytes
// at most, WG/G groups) and the control block.  Referenced directly (not through pointer
// arguments) so every access is a ds_* instruction.
template <int WG>
__shared__ __attribute__((aligned(16))) char g_smem[WG * 8 * 16 * 9 / 8];
__shared__ Ctl g_ctl;
// FP64 two-level weight tables of the class being run (ClassDesc::wtab, slots of kWtabSlot),
// in slots of kSlot<WG> entries: the 256-thread family (N <= 4096) needs 64 per slot, which
// keeps its LDS at 40 KB, four workgroups per CU
template <int WG>
constexpr int kSlot = WG == 256 ? 64 : llh::kWtabSlot;
template <int WG>
__shared__ double g_wtab[4 * kSlot<WG>];

// Exact canonical residue of N near-balanced digits (natural order, i64, in LDS):
// full carry propagation with wrap (2^p == 1), then Res64 = low 64 bits of the value in
// [0, 2^p - 1), with the all-ones pattern (== M_p == 0) folded to zero.
// Restates the GMP rebuild of lucaslehmer.cpp:312-323 / cuda/ll_ntt.cu:238-251 without GMP.
__device__ int normalize_res64(i64 *d, u32 p, u32 N, u64 &res, int &zero) {
    const DigitWalk DW{N, p % N, p / N};
    i64 carry = 0;
    for (int round = 0; round < 4; ++round) {
        u32 r = 0;
        for (u32 j = 0; j < N; ++j) {
            if (round > 0 && carry == 0) break;
            const int b = DW.bits(r);
            r = DW.next(r);
            i64 v = d[j] + carry;
            d[j] = v & (((i64)1 << b) - 1);
            carry = v >> b;
        }
        if (carry == 0) break;
    }
    if (carry != 0) return -1;
    u64 res_ = 0, cj = 0;
    int z = 1, ones = 1;
    u32 r = 0;
    for (u32 j = 0; j < N; ++j) {
        const int b = DW.bits(r);
        r = DW.next(r);
        u64 dj = (u64)d[j];
        if (dj) z = 0;
        if (dj != (((u64)1 << b) - 1)) ones = 0;
        if (cj < 64) res_ |= dj << cj;
        cj += (u64)b;
    }
    if (ones) { res_ = 0; z = 1; }
    res = res_;
    zero = z;
    return 0;
}

LL_HD constexpr int ilog2_ceil(u32 v) { return v <= 1 ? 0 : 1 + ilog2_ceil((v + 1) / 2); }

// A wave-uniform pointer, moved into SGPRs (readfirstlane) so loads through it take a scalar
// base and a 32-bit vector offset.  run_class is a real call and receives the class descriptor
// through memory, which would otherwise leave its table pointers in VGPRs.
template <class T>
LL_DEV gptr<T> uniform_table(const void *p) {
    const u64 v = (u64)p;
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)v), hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
    return (gptr<T>)(((u64)hi << 32) | lo);
}

// Bits and weight exponents (as 4 e, the walk's scaling) of the two digits 2j, 2j + 1 of FP64
// element j.
struct Pair {
    int b0, b1;
    u32 e0, e1;
};

// TWO: the FP64 carry takes two hand-off rounds (ClassDesc::carry2, set by the host for the
// exponents with fewer than log2 N + 3 bits per digit); otherwise one.  CK: checkpoints
// (ClassDesc::ck_slots).  A separate instantiation: with the checkpoint code merely present the
// loop ran 2.3% slower (C3: 5.70 s against 5.56 s, same box), though its instructions are the same.
template <bool IS_GL, int WG, int C, bool TWO, bool CK>
__device__ __attribute__((noinline)) void run_class(const llh::ClassDesc &cd, const llh::BatchItem *__restrict__ items,
                          llh::BatchResult *__restrict__ out, u32 *counter) {
    Ctl &ctl = g_ctl;
    constexpr Cls K = kCls[C];
    constexpr int M = K.M, G = K.G, R = K.R;
    constexpr bool P2 = K.H1 == 0;
    static_assert(!IS_GL || P2, "the Goldilocks classes are powers of two");
    static_assert(K.WG == WG, "class run by another family's kernel");
    typedef typename std::conditional<IS_GL, FGL, FC64>::type F;
    typedef typename F::T T;
    typedef typename F::W W;
    typedef typename std::conditional<IS_GL, u64, double>::type WT;
    constexpr int NG = WG / G;
    constexpr int PS = R == 8 ? 3 : 2;            // LDS pad shift: one pad element per R
    // FP64 hand-offs after the first use one pad element per 32: their 16-byte stores are runs of
    // consecutive elements, and the LDS serves contiguous 16-byte loads of a wave without bank
    // conflicts, where one pad per 8 (needed by the first pass's stride-8 stores) costs two passes
    constexpr int PSM = R == 8 ? 5 : PS;
    static_assert(R == 8 || R == 4, "per-thread counts 4 and 8");
    static_assert(WG % G == 0, "groups must tile the workgroup");
    constexpr int REG = M + (M >> PS);
    static_assert(NG <= kMaxGroups, "too many groups per workgroup");
    static_assert(NG * REG * (int)sizeof(T) <= (int)sizeof(g_smem<WG>), "LDS buffer too small");
    constexpr u32 N = IS_GL ? (u32)M : 2u * (u32)M;
    static_assert(G * R == M, "a group's threads hold the whole residue");

    const int tid = threadIdx.x, g = tid / G, t = tid % G;
    T *lds = reinterpret_cast<T *>(g_smem<WG>) + g * REG;
    // digit count and weight-table split are fixed by the class (the host builds the two-level
    // weights with the same S = 2^ceil(ceil(log2 N) / 2))
    constexpr int log_n = ilog2_ceil(N);          // exact for the powers of two, the only users
    constexpr u32 nmask = N - 1;
    constexpr int log_s = (ilog2_ceil(N) + 1) / 2;
    constexpr u32 smask = (1u << log_s) - 1;
    static_assert(IS_GL || ((1 << log_s) <= kSlot<WG> && ((N + smask) >> log_s) <= (u32)kSlot<WG>),
                  "weight table slots too small for the class");
    const gptr<W> tw = uniform_table<W>(cd.tw);
    const gptr<W> twp = uniform_table<W>(cd.twp);
    const gptr<W> twpi = IS_GL ? uniform_table<W>(cd.twpi) : twp;
    const gptr<WT> wf = uniform_table<WT>(cd.wf);
    const gptr<WT> wi = uniform_table<WT>(cd.wi);
    if constexpr (!IS_GL) {
        // The low slots (0, 2) are stored as their S/2 even entries, then the S/2 odd ones.
        // A digit pair's exponents are e0 = (-2 j p) mod N (even, as N is even) and
        // e1 = e0 - p mod N (odd), and across the lanes of a wave e0 steps by -2 P mod N.  In
        // natural order the low indices e0 mod S then fall on every other bank pair (two-way
        // conflicts on each read); in the halves the index (e0 mod S) / 2 steps by the odd P
        // and every lane of a pass has a bank pair of its own.
        const gptr<double> wt = uniform_table<double>(cd.wtab);
#pragma unroll
        for (int i = tid; i < 4 * kSlot<WG>; i += WG) {
            const u32 slot = (u32)i / kSlot<WG>, pos = (u32)i % kSlot<WG>;
            const u32 k = slot % 2 != 0 || pos > smask ? pos : pos <= (smask >> 1) ? 2 * pos : 2 * pos - smask;
            g_wtab<WG>[i] = ld_table(wt, slot * llh::kWtabSlot + k);
        }
    }
    // Weight of the digit whose walk value is E = 4 e (e its weight exponent, of parity ODD;
    // forward: slots 0, 1; inverse: slots 2, 3).  With the low slot stored as even entries,
    // then odd ones, entry e mod S sits at byte 4 (e mod S) for even e and at byte
    // 4 (S - 1) + 4 (e mod S) for odd e: both are E masked to 4 (S - 1), plus a constant.
    auto wlds = [&](u32 E, int inv, int odd) {
        const char *lo = reinterpret_cast<const char *>(&g_wtab<WG>[(2 * inv) * kSlot<WG>]);
        const char *hi = reinterpret_cast<const char *>(&g_wtab<WG>[(2 * inv + 1) * kSlot<WG>]);
        const u32 lob = (u32)odd * 4u * smask + (E & (4u * smask));
        const u32 hib = (E >> (log_s + 2)) << 3;
        return *reinterpret_cast<const double *>(lo + lob) * *reinterpret_cast<const double *>(hi + hib);
    };
    // Digits 2j, 2j + 1 of element j = t + G s, s = 0 .. R-1, in order.  With P = p mod N and
    // F = p div N (p > 64 is prime, so coprime to N, and 0 < P < N), the digit of weight
    // exponent e ((-d p) mod N for digit index d) has F + [e < P] bits, and the next digit's
    // exponent is (e - P) mod N; element j + G's low digit has (e - 2 G p) mod N.  This equals
    // the residue walk of DigitWalk digit for digit.  The walk carries E = 4 e (and 4 P,
    // 4 step, mod 4 N), which is the low table's byte offset up to a mask.
    struct Layout {
        u32 P4, F, step4, E;
    };
    constexpr u32 N4 = 4u * N;
    auto sub_mod = [](u32 a, u32 b) -> u32 {      // (a - b) mod 4N for a, b < 4N
        if constexpr (P2) {
            return (a - b) & (N4 - 1);
        } else {
            // for a < b the difference wraps to at least 2^32 - 4N and d + 4N is the residue;
            // otherwise d < 4N <= d + 4N: the smaller of the two is (a - b) mod 4N
            const u32 d = a - b;
            return min(d, d + N4);
        }
    };
    auto layout = [&](u32 pp, int tt) {
        Layout L;
        const u32 P = pp % N;
        L.F = pp / N;
        L.P4 = 4u * P;
        L.step4 = 8u * (u32)G * (P % (u32)R);      // 4 ((2 G P) mod N), as N = 2 G R
        L.E = sub_mod(0u, 4u * ((2u * (u32)tt * P) % N));
        return L;
    };
    auto pair_at = [&](Layout &L) {
        Pair q;
        q.e0 = L.E;
        q.e1 = sub_mod(q.e0, L.P4);
        q.b0 = (int)(L.F + (q.e0 < L.P4 ? 1u : 0u));
        q.b1 = (int)(L.F + (q.e1 < L.P4 ? 1u : 0u));
        L.E = sub_mod(q.e0, L.step4);
        return q;
    };

    // v = c 2^b + d with the balanced digit |d| <= 2^(b-1): both steps exact for integer v
    // below 2^53 (the scaling is a power of two, d is the exact difference)
    auto split = [](double v, int b, double &c) {
        c = __builtin_rint(ldexp(v, -b));
        return v - ldexp(c, b);
    };

    u32 item = kNone, p = N + 1;
    u64 target = 0, it = 0, t0 = 0;
    double emax = 0.0;
    T x[R];
    // This group's checkpoint slot (null: checkpoints off), recomputed where it is used so that
    // nothing of it stays live across the squaring loop.
    constexpr bool ck_on = CK;
    auto slot_of = [&]() { return cd.ck_slots + (size_t)blockIdx.x * (16 * WG) + (size_t)g * N; };
    // Slots, rollback and resume entries hold the elements' raw words in natural order (FP64:
    // element j = digits 2j, 2j + 1 as integral doubles; GL: the signed digit's word), so saving
    // and restoring a residue is a plain copy with no conversion in this kernel.
    auto write_slot = [&]() {
        T *slot = reinterpret_cast<T *>(slot_of());
#pragma unroll
        for (int s = 0; s < R; ++s) slot[t + G * s] = x[s];
    };

    auto init_item = [&]() {
        u32 init = 0;
        if (item != kNone) {
            const llh::BatchItem bi = items[item];
            p = bi.p;
            target = bi.iters;
            it = bi.it0;
            init = bi.init;
        } else {
            p = N + 1;          // a valid layout for an idle group; its digits stay zero
            target = 0;
            it = 0;
        }
        emax = 0.0;
        if (init) {             // resume: the near-balanced digits of a checkpoint, re-split at N
            const T *src = reinterpret_cast<const T *>(cd.resume + (size_t)(init - 1) * llh::kResumeStride);
#pragma unroll
            for (int s = 0; s < R; ++s) x[s] = src[t + G * s];
        } else {
#pragma unroll
            for (int s = 0; s < R; ++s) {
                if constexpr (IS_GL) x[s] = 0; else x[s] = cmk(0.0, 0.0);
            }
            if (item != kNone && t == 0) {
                if constexpr (IS_GL) x[0] = 4; else x[0] = cmk(4.0, 0.0);
            }
        }
        if (ck_on && item != kNone) write_slot();    // the start state is the first checkpoint
        t0 = wall_clock64();
    };

    // Squarings to the next workgroup event (a group reaches its target, or the periodic
    // round-off check every kCheck squarings) and whether any group has an exponent.  Both are
    // uniform over the workgroup, so the squaring loop itself needs no reduction; a group whose
    // round-off went above the limit is caught at the next event (its maximum is sticky).
    u32 ev = 1, ev0 = 1;    // squarings to the next event; ev0: the interval's full length
    bool any = false;
    auto schedule = [&]() {
        __syncthreads();
        if (tid == 0) {
            ctl.ev = kCheck;
            ctl.any = 0;
        }
        __syncthreads();
        if (t == 0 && item != kNone) {
            atomicMin(&ctl.ev, target > it + kCheck ? kCheck : (target > it ? (u32)(target - it) : 1u));
            ctl.any = 1;
        }
        __syncthreads();
        ev = ctl.ev;
        ev0 = ev;
        any = ctl.any != 0;
    };

    // first fetch for every group
    if (t == 0) {
        u32 k = atomicAdd(counter, 1u);
        ctl.item[g] = (k < cd.count) ? cd.first + k : kNone;
    }
    __syncthreads();
    item = ctl.item[g];
    init_item();
    schedule();

    while (any) {
        const bool has = item != kNone;
        double eit = 0.0;

        if constexpr (!IS_GL) {
            // ---- weight: z_j = x_{2j} a_{2j} + i x_{2j+1} a_{2j+1}
            {
                const int tw0 = ll_opaque(t);
                Layout L = layout(ll_opaque(p), tw0);
                double wa[R], wb[R];
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const Pair q = pair_at(L);
                    wa[s] = wlds(q.e0, 0, 0);
                    wb[s] = wlds(q.e1, 0, 1);
                }
#pragma unroll
                for (int s = 0; s < R; ++s) x[s] = cmk(x[s].re * wa[s], x[s].im * wb[s]);
            }
            if constexpr (P2)
                transform<F, M, G, R, K.R1, K.R2, K.R3, K.R4, false, true, PS, PSM>(x, lds, t, twp);
            else
                hybrid_transform<F, M, G, R, false, K.H1, K.H2, 0, true, PS>(x, lds, t, twp);
            // ---- square the real spectrum in packed form (pairs k, M-k).  The exchange is
            // stored unpadded: the stores and the mirrored loads are contiguous across the
            // lanes, which the LDS serves without bank conflicts (a pad element per 16 made
            // every 16-byte load of it take two passes).
            const int tq = ll_opaque(t);
            C64 wq[R];      // the pairing's table loads stay in flight across the exchange
#pragma unroll
            for (int q = 0; q < R; ++q) wq[q] = ld_table(tw, (u32)(tq + G * q));
            __syncthreads();
#pragma unroll
            for (int q = 0; q < R; ++q) lds[tq + G * q] = x[q];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int k = tq + G * q;
                const int km = q == 0 && tq == 0 ? 0 : M - k;     // k = 0 only for q = 0, t = 0
                const C64 zm = lds[km];
                const C64 D = csub(x[q], cconj(zm));
                const C64 w = wq[q];
                const C64 A = cmk(1.0 + w.re, w.im);
                const C64 AD = cmul(A, csqr(D));
                const C64 Z2 = csqr(x[q]);
                x[q] = cmk(__fma_rn(-0.25, AD.re, Z2.re), __fma_rn(-0.25, AD.im, Z2.im));
            }
            if constexpr (P2)
                transform<F, M, G, R, K.R1, K.R2, K.R3, K.R4, true, true, PS, PSM>(x, lds, t, twp);
            else
                hybrid_transform<F, M, G, R, true, K.H1, K.H2, 0, true, PS>(x, lds, t, twp);
            // ---- unweight, round, error, -2, carry within the pair
            double cout[R];
            {
                const int tu = ll_opaque(t);
                Layout L = layout(ll_opaque(p), tu);
                double wa[R], wb[R];
                int ba[R], bb[R];
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const Pair q = pair_at(L);
                    wa[s] = wlds(q.e0, 1, 0);
                    wb[s] = wlds(q.e1, 1, 1);
                    ba[s] = q.b0;
                    bb[s] = q.b1;
                }
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const u32 j = tu + G * s;
                    const double v0 = x[s].re * wa[s], v1 = x[s].im * wb[s];
                    double r0 = __builtin_rint(v0), r1 = __builtin_rint(v1);
                    eit = fmax(eit, fmax(fabs(v0 - r0), fabs(v1 - r1)));
                    if (s == 0 && tu == 0 && has) r0 -= 2.0;     // j = 0 only for s = 0, t = 0
                    double cA, cB;
                    const double x0 = split(r0, ba[s], cA);
                    r1 += cA;
                    const double x1 = split(r1, bb[s], cB);
                    x[s] = cmk(x0, x1);
                    cout[s] = cB;
                }
            }
            // ---- carry hand-offs: element j takes the carry out of element j-1 (wrap at 0).
            // Round 1 moves the large carries and re-splits the pair's low digit; what it leaves
            // for the high digit is at most about N/4 (typically sqrt(N)/4) against a balanced
            // digit of 2^(b-1).  With floor(p/N) >= log2 N + 3 that slack is below 1/16 of a
            // digit and stays bounded over the squarings (d' = d0 (1 + d)^2 with d0 <= 1/16), so
            // the high digit keeps it.  With fewer bits per digit (TWO) round 1 also splits the
            // high digit and round 2 moves that carry into the next pair's low digit, which keeps
            // the digits balanced down to a few bits per digit.
            double *ldsd = reinterpret_cast<double *>(lds);
            const int tc = ll_opaque(t);
            __syncthreads();
#pragma unroll
            for (int s = 0; s < R; ++s) ldsd[tc + G * s] = cout[s];
            __syncthreads();
            {
                Layout L = layout(ll_opaque(p), tc);
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const u32 j = tc + G * s;
                    const double cin = ldsd[s == 0 && tc == 0 ? M - 1 : j - 1];
                    const Pair q = pair_at(L);
                    double cC, cD;
                    const double x0 = split(x[s].re + cin, q.b0, cC);
                    if constexpr (TWO) {
                        const double x1 = split(x[s].im + cC, q.b1, cD);
                        x[s] = cmk(x0, x1);
                        cout[s] = cD;
                    } else {
                        x[s] = cmk(x0, x[s].im + cC);
                    }
                }
            }
            if constexpr (TWO) {
                __syncthreads();
#pragma unroll
                for (int s = 0; s < R; ++s) ldsd[tc + G * s] = cout[s];
                __syncthreads();
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const u32 j = tc + G * s;
                    x[s].re += ldsd[s == 0 && tc == 0 ? M - 1 : j - 1];
                }
            }
        } else {
            // ---- Goldilocks: weight, NTT, square, inverse NTT
            const int tw0 = ll_opaque(t);
            const u32 pw = ll_opaque(p);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tw0 + G * s;
                const u32 e = (0u - j * pw) & nmask;
                x[s] = gl_mul(gl_from_i64((i64)x[s]), ld_table(wf, e));
            }
            transform<F, M, G, R, K.R1, K.R2, K.R3, K.R4, false, true, PS>(x, lds, t, twp);
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = gl_sqr(x[q]);
            transform<F, M, G, R, K.R1, K.R2, K.R3, K.R4, true, true, PS>(x, lds, t, twpi);
            i64 cout[R];
            const int tu = ll_opaque(t);
            const u32 pu = ll_opaque(p);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tu + G * s;
                const u32 e = (0u - j * pu) & nmask;
                const int b = (int)((pu >> log_n) + (e < (pu & nmask) ? 1u : 0u));   // as pair_at
                i64 v = gl_to_i64(gl_mul(x[s], ld_table(wi, e)));
                if (s == 0 && tu == 0 && has) v -= 2;
                const i64 c = (v + ((i64)1 << (b - 1))) >> b;
                x[s] = (u64)(v - (c << b));
                cout[s] = c;
            }
            i64 *ldsi = reinterpret_cast<i64 *>(lds);
#pragma unroll
            for (int round = 0; round < 2; ++round) {
                const int tc = ll_opaque(t);
                const u32 pc = ll_opaque(p);
                __syncthreads();
#pragma unroll
                for (int s = 0; s < R; ++s) ldsi[tc + G * s] = cout[s];
                __syncthreads();
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const u32 j = tc + G * s;
                    const i64 cin = ldsi[(j - 1) & nmask];
                    i64 v = (i64)x[s] + cin;
                    if (round == 0) {
                        const u32 e = (0u - j * pc) & nmask;
                        const int b = (int)((pc >> log_n) + (e < (pc & nmask) ? 1u : 0u));
                        const i64 c = (v + ((i64)1 << (b - 1))) >> b;
                        v -= c << b;
                        cout[s] = c;
                    }
                    x[s] = (u64)v;
                }
            }
        }

        if (has) {
            ++it;
            emax = fmax(emax, eit);
        }
        if (--ev == 0) {
            const bool done = has && it >= target;
            const bool bad = has && emax > kErrLimit;
            // ---- event: flags, error reduction, exact normalisation, refill
            if (t == 0) { ctl.bad[g] = 0; ctl.err[g] = 0; }
            __syncthreads();
            if (bad) ctl.bad[g] = 1;
            __syncthreads();
            const bool gbad = has && ctl.bad[g] != 0;
            const bool gfin = done || gbad;
            // Checkpoints.  Events are uniform over the workgroup, so a group's last good
            // checkpoint (the previous event, or its start state if it began then) is ev0
            // squarings back.
            if constexpr (ck_on) {
                if (gbad && t == 0) ctl.rb[g] = atomicAdd(cd.rb_count, 1u);
                __syncthreads();
                if (gbad) {             // the last good checkpoint goes to a rollback entry
                    const u32 k = ctl.rb[g];
                    if (k < cd.rb_cap) {
                        const T *slot = reinterpret_cast<const T *>(slot_of());
                        T *dst = reinterpret_cast<T *>(cd.rb_area + (size_t)k * llh::kResumeStride);
#pragma unroll
                        for (int s = 0; s < R; ++s) dst[t + G * s] = slot[t + G * s];
                    }
                } else if (has && !done) {      // within the limit so far: a new checkpoint
                    write_slot();
                }
            }
            i64 *dig = reinterpret_cast<i64 *>(lds);
            if (gfin) {
                atomicMax(&ctl.err[g], (unsigned long long)__double_as_longlong(emax));
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const u32 j = t + G * s;
                    if constexpr (IS_GL) {
                        dig[j] = (i64)x[s];
                    } else {
                        dig[2 * j] = (i64)x[s].re;
                        dig[2 * j + 1] = (i64)x[s].im;
                    }
                }
            }
            __syncthreads();
            if (gfin && t == 0) {
                llh::BatchResult r;
                r.maxerr = __longlong_as_double((long long)ctl.err[g]);
                r.iters_done = it;
                r.t0 = t0;
                r.res64 = 0;
                r.verdict = 0;
                r.ck_iter = it - ev0;
                r.ck = 0;
                r.pad = 0;
                if (gbad) {
                    r.status = llh::kStRoundoff;
                    if (ck_on && ctl.rb[g] < cd.rb_cap) r.ck = ctl.rb[g] + 1;
                } else {
                    u64 res = 0;
                    int zero = 0;
                    r.status = normalize_res64(dig, p, N, res, zero) == 0 ? llh::kStOk : llh::kStNotRun;
                    r.res64 = res;
                    r.verdict = zero;
                }
                r.t1 = wall_clock64();
                out[items[item].slot] = r;
                u32 k = atomicAdd(counter, 1u);
                ctl.item[g] = (k < cd.count) ? cd.first + k : kNone;
            }
            __syncthreads();
            if (gfin) {
                item = ctl.item[g];
                init_item();
            }
            schedule();
        }
    }
}

// p < 64: the residue is one machine word; (s^2 - 2) mod M_p with a 128-bit square.
__device__ void run_tiny(const llh::ClassDesc &cd, const llh::BatchItem *__restrict__ items,
                         llh::BatchResult *__restrict__ out, u32 *counter) {
    for (;;) {
        const u32 k = atomicAdd(counter, 1u);
        if (k >= cd.count) break;
        const llh::BatchItem bi = items[cd.first + k];
        const u64 t0 = wall_clock64();
        const u32 p = bi.p;
        const u64 Mp = (((u64)1) << p) - 1;
        u64 s = 4;
        for (u64 i = 0; i < bi.iters; ++i) {
            u64 lo, hi;
            mul64x64(s, s, lo, hi);
            // x = s^2 + M_p - 2 (never negative), then fold with 2^p == 1
            unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
            v += (unsigned __int128)(Mp - 2);
            while (v >> p) v = (v & Mp) + (v >> p);
            u64 r = (u64)v;
            if (r >= Mp) r -= Mp;
            s = r;
        }
        llh::BatchResult r;
        r.verdict = s == 0;
        r.status = llh::kStOk;
        r.res64 = s;
        r.maxerr = 0.0;
        r.iters_done = bi.iters;
        r.t0 = t0;
        r.t1 = wall_clock64();
        r.ck_iter = 0;
        r.ck = 0;
        r.pad = 0;
        out[bi.slot] = r;
    }
}

// class C's branch of a family kernel (nothing for classes of other families or, for GL,
// lengths with an odd factor)
template <bool IS_GL, int WG, int C>
__device__ __forceinline__ void run_if(const llh::ClassDesc &cd, const llh::BatchItem *__restrict__ items,
                                       llh::BatchResult *__restrict__ out, u32 *counter) {
    if constexpr (C > 0 && kCls[C].WG == WG && (!IS_GL || kCls[C].H1 == 0)) {
        if constexpr (IS_GL) {
            run_class<true, WG, C, true, false>(cd, items, out, counter);
        } else if (cd.ck_slots) {
            if (cd.carry2) run_class<false, WG, C, true, true>(cd, items, out, counter);
            else run_class<false, WG, C, false, true>(cd, items, out, counter);
        } else {
            if (cd.carry2) run_class<false, WG, C, true, false>(cd, items, out, counter);
            else run_class<false, WG, C, false, false>(cd, items, out, counter);
        }
    }
}

template <bool IS_GL, int WG>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(WG == 384 ? 3 : 4, WG == 384 ? 3 : 4)))
ll_batch_kernel(const llh::ClassDesc *__restrict__ classes, int ncls,
                const llh::BatchItem *__restrict__ items, llh::BatchResult *__restrict__ out,
                u32 *counters) {
    for (int c = 0; c < ncls; ++c) {
        const llh::ClassDesc cd = classes[c];
        switch (cd.cls) {
#define LL_CLASS(C) \
    case C: run_if<IS_GL, WG, C>(cd, items, out, counters + c); break;
            LL_CLASS(1) LL_CLASS(2) LL_CLASS(3) LL_CLASS(4) LL_CLASS(5) LL_CLASS(6) LL_CLASS(7)
            LL_CLASS(8) LL_CLASS(9) LL_CLASS(10) LL_CLASS(11) LL_CLASS(12) LL_CLASS(13) LL_CLASS(14)
            LL_CLASS(15) LL_CLASS(16) LL_CLASS(17) LL_CLASS(18) LL_CLASS(19) LL_CLASS(20) LL_CLASS(21)
            LL_CLASS(22) LL_CLASS(23) LL_CLASS(24) LL_CLASS(25) LL_CLASS(26)
#undef LL_CLASS
            case 0:
                if constexpr (WG == 256) run_tiny(cd, items, out, counters + c);
                break;
            default: break;
        }
        __syncthreads();
    }
}
static_assert(kNumCls == 27, "the class switch lists every class");

template <int C>
static void pass_exponents(std::vector<u32> &ex) {
    typedef typename ClsRadix<C>::PT PT;
    ex.resize(PT::SIZE);
    for (int e = 0; e < PT::SIZE; ++e) ex[e] = PT::exponent(e);
}

}  // namespace

namespace llh {

int batch_num_classes() { return kNumCls; }

// class of transform length M (0 = none); the length's base class, not a schedule variant
int batch_class_of_m(u64 M) {
    for (int c = 1; c < kNumCls; ++c)
        if ((u64)kCls[c].M == M && kClsBase[c] == c) return c;
    return 0;
}
int batch_class_base(int c) { return c >= 0 && c < kNumCls ? kClsBase[c] : 0; }
// radix list of class c: R1..R4 (power of two) or 8, H1, [H2], 8; unused entries 0
void batch_class_radices(int c, int r[4]) {
    const Cls &k = kCls[c > 0 && c < kNumCls ? c : 0];
    if (k.H1 == 0) {
        r[0] = k.R1; r[1] = k.R2; r[2] = k.R3; r[3] = k.R4;
    } else {
        r[0] = 8; r[1] = k.H1; r[2] = k.H2 ? k.H2 : 8; r[3] = k.H2 ? 8 : 0;
    }
}
u64 batch_class_m(int c) { return c > 0 && c < kNumCls ? (u64)kCls[c].M : 0; }
bool batch_class_pow2(int c) { return c > 0 && c < kNumCls && kCls[c].H1 == 0; }
int batch_class_family(int c) {
    const int wg = c >= 0 && c < kNumCls ? kCls[c].WG : kFamilyWG[0];
    for (int f = 0; f < kNumFam; ++f)
        if (kFamilyWG[f] == wg) return f;
    return 0;
}
int batch_class_factor(int c) {
    const int f = batch_class_family(c);
    return f == 4 ? 0 : f;
}
int batch_num_families() { return kNumFam; }
int batch_class_groups(int c) { return c > 0 && c < kNumCls ? kCls[c].WG / kCls[c].G : 1; }

// Full-table exponent of every entry of class c's per-pass twiddle table.
bool batch_pass_exponents(int c, std::vector<u32> &ex) {
    switch (c) {
#define LL_PX(C) \
    case C: pass_exponents<C>(ex); return true;
        LL_PX(1) LL_PX(2) LL_PX(3) LL_PX(4) LL_PX(5) LL_PX(6) LL_PX(7) LL_PX(8) LL_PX(9) LL_PX(10) LL_PX(11)
        LL_PX(12) LL_PX(13) LL_PX(14) LL_PX(15) LL_PX(16) LL_PX(17) LL_PX(18) LL_PX(19) LL_PX(20) LL_PX(21)
        LL_PX(22) LL_PX(23) LL_PX(24) LL_PX(25) LL_PX(26)
#undef LL_PX
        default: return false;
    }
}

int batch_family_wg(int family) { return family >= 0 && family < kNumFam ? kFamilyWG[family] : 0; }

// Launch family `family`'s persistent batch kernel on `stream`; grid = resident workgroups.
hipError_t launch_batch(int transform, int family, const ClassDesc *d_classes, int ncls, const BatchItem *d_items,
                        BatchResult *d_out, u32 *d_counters, int grid, hipStream_t stream) {
    const bool gl = transform == kNTT;
    if (family < 0 || family >= kNumFam || (gl && family != 0 && family != 4)) return hipErrorInvalidValue;
    const int wg = kFamilyWG[family];
#define LL_LAUNCH(GL, W)                                                                                         \
    hipLaunchKernelGGL((ll_batch_kernel<GL, W>), dim3(grid), dim3(wg), 0, stream, d_classes, ncls, d_items, d_out, \
                       d_counters)
    switch (family) {
        case 1: LL_LAUNCH(false, 384); break;
        case 2: LL_LAUNCH(false, 320); break;
        case 3: LL_LAUNCH(false, 448); break;
        case 4:
            if (gl) LL_LAUNCH(true, 512);
            else LL_LAUNCH(false, 512);
            break;
        default:
            if (gl) LL_LAUNCH(true, 256);
            else LL_LAUNCH(false, 256);
            break;
    }
#undef LL_LAUNCH
    return hipGetLastError();
}

// One workgroup per resume request: normalise the checkpoint's near-balanced digits at the old
// length to canonical digits (thread 0, the exact carry of normalize_res64), then re-split them
// at the new length into the request's resume entry (relayout_digit, every thread).  ok[e] = 0
// when the carries did not settle; the item then starts again from s_0.
// Entries hold raw element words (FP64: integral doubles); the digits are made i64 in place
// first, and the resume entry is written back in the engine's word format.
template <bool IS_GL>
__global__ void __launch_bounds__(256) bt_resume(const llh::ResumeReq *reqs, i64 *rb_area, i64 *resume, int *ok) {
    const llh::ResumeReq q = reqs[blockIdx.x];
    i64 *d = rb_area + (size_t)q.rb * llh::kResumeStride;
    if constexpr (!IS_GL) {
        const double *dd = reinterpret_cast<const double *>(d);
        for (u32 k = threadIdx.x; k < q.n; k += 256) d[k] = (i64)dd[k];
    }
    __syncthreads();
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        u64 res;
        int zero;
        s_ok = normalize_res64(d, q.p, q.n, res, zero) == 0;
        ok[blockIdx.x] = s_ok;
    }
    __syncthreads();
    if (!s_ok) return;
    i64 *dst = resume + (size_t)blockIdx.x * llh::kResumeStride;
    for (u32 k = threadIdx.x; k < q.n2; k += 256) {
        const i64 v = relayout_digit(d, q.p, q.n, q.n2, k);
        if constexpr (IS_GL) dst[k] = v;
        else reinterpret_cast<double *>(dst)[k] = (double)v;
    }
}

hipError_t batch_resume(int transform, const ResumeReq *d_reqs, int n, i64 *rb_area, i64 *resume, int *d_ok,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (transform == kNTT) hipLaunchKernelGGL(bt_resume<true>, dim3(n), dim3(256), 0, st, d_reqs, rb_area, resume, d_ok);
    else hipLaunchKernelGGL(bt_resume<false>, dim3(n), dim3(256), 0, st, d_reqs, rb_area, resume, d_ok);
    return hipGetLastError();
}

int batch_blocks_per_cu(int transform, int family) {
    const bool gl = transform == kNTT;
    if (family < 0 || family >= kNumFam || (gl && family != 0 && family != 4)) return 1;
    const int wg = kFamilyWG[family];
    int nb = 0;
    hipError_t e;
#define LL_OCC(GL, W) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ll_batch_kernel<GL, W>, wg, 0)
    switch (family) {
        case 1: LL_OCC(false, 384); break;
        case 2: LL_OCC(false, 320); break;
        case 3: LL_OCC(false, 448); break;
        case 4:
            if (gl) LL_OCC(true, 512);
            else LL_OCC(false, 512);
            break;
        default:
            if (gl) LL_OCC(true, 256);
            else LL_OCC(false, 256);
            break;
    }
#undef LL_OCC
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

}  // namespace llh
