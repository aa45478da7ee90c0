// ll_large.hip -- the large engine's host side: schedule dispatch, the squaring loop's launch
// sequence, and the kernels that do not depend on the schedule (exact normalisation of the
// digits, re-layout of a checkpoint at another length).  The squaring kernels are templates in
// ll_large_impl.hpp, instantiated per schedule by the ll_large_g*.hip units.
#include "ll_large_impl.hpp"
#include <vector>


// ---- exact normalisation of N digits ----------------------------------------------------------
// The digits (near balanced, after the pending carries are applied) become the canonical digits
// of the residue, each in [0, 2^b_d), in dig; then Res64 and whether the residue is zero (the
// all-ones pattern M_p counts as zero).  Restates the GMP rebuild of lucaslehmer.cpp:312-323 on
// the device, in four launches:
//   local  (one thread per segment of kNormK digits) canonical digits of each segment alone and
//          the segment's carry out;
//   carry  (one workgroup) the segment carries, exactly: thread t adds each carry at the first
//          digit of the next of its segments and lets it run on while digits overflow; what
//          leaves its range goes to thread t + 1 in rounds until none is left (the top digit's
//          carry wraps to digit 0: 2^p == 1 mod M_p);
//   check  (one thread per segment) the all-zero and all-ones flags;
//   final  (one thread) Res64 from the lowest digits.
// The segment carries sit in the output buffer, kNormSegOffset bytes after the NormOut.
namespace {
constexpr int kNormK = 32;
constexpr int kNormThreads = 1024;
constexpr size_t kNormSegOffset = 256;

struct NormOut {
    unsigned long long res64;
    int zero;       // the residue is zero (all digits zero, or all ones)
    int status;     // 1: the carry rounds did not settle
    int ones;       // the check pass's all-ones flag
};

LL_DEV i64 digit_mask(int b) { return ((i64)1 << b) - 1; }

// adds c at digit a and carries on through digits [a, b) while it overflows; returns what leaves
LL_DEV i64 add_run(i64 *dig, u64 a, u64 b, i64 c, u32 p, const NDiv &D) {
    for (u64 d = a; d < b && c != 0; ++d) {
        const int bits = bits_of(d, p, D);
        const i64 v = dig[d] + c;
        dig[d] = v & digit_mask(bits);
        c = v >> bits;
    }
    return c;
}

// Storage index of digit d's word: FP64 digit d is part d % 2 of element d / 2 (a word of two
// doubles), GL digit d is element d; the elements are stored by elem_index of the W x H layout.
template <bool IS_GL>
LL_DEV u64 digit_word(u64 d, u32 W, u32 H) {
    const u64 j = IS_GL ? d : d / 2;
    const u64 e = elem_index((u32)(j / W), (u32)(j % W), W, H, ElemOf<IS_GL>::LC);
    return IS_GL ? e : 2 * e + (d & 1);
}

template <bool IS_GL>
__global__ void __launch_bounds__(256) lg_norm_local(const void *elem_v, i64 *dig, i64 *segc, u32 p, NDiv D,
                                                     u32 W, u32 H) {
    const u64 s = (u64)blockIdx.x * 256u + threadIdx.x;
    if (s >= D.n / kNormK) return;
    const u64 d0 = s * kNormK;
    i64 v[kNormK];
#pragma unroll
    for (int k = 0; k < kNormK; ++k) {
        if constexpr (IS_GL) v[k] = (i64)reinterpret_cast<const u64 *>(elem_v)[digit_word<true>(d0 + k, W, H)];
        else v[k] = (i64)reinterpret_cast<const double *>(elem_v)[digit_word<false>(d0 + k, W, H)];
    }
    i64 c = 0;
    u64 lo = cbit(d0, p, D);
#pragma unroll
    for (int k = 0; k < kNormK; ++k) {
        const u64 hi = cbit(d0 + k + 1, p, D);
        const int bits = (int)(hi - lo);
        lo = hi;
        const i64 x = v[k] + c;
        dig[d0 + k] = x & digit_mask(bits);
        c = x >> bits;
    }
    segc[s] = c;
}

__global__ void __launch_bounds__(kNormThreads) lg_norm_carry(i64 *dig, const i64 *segc, u32 p, NDiv D,
                                                             NormOut *out) {
    const u64 nseg = D.n / kNormK;
    const u64 S = (nseg + kNormThreads - 1) / kNormThreads;
    const int t = threadIdx.x;
    const u64 s0 = (u64)t * S < nseg ? (u64)t * S : nseg, s1 = s0 + S < nseg ? s0 + S : nseg;
    const u64 d0 = s0 * kNormK, d1 = s1 * kNormK;
    __shared__ i64 s_cout[kNormThreads];
    __shared__ int s_any;
    i64 r = 0;                       // carry into segment s from the one below it
    for (u64 s = s0; s < s1; ++s) {
        const i64 left = r != 0 ? add_run(dig, s * kNormK, (s + 1) * kNormK, r, p, D) : 0;
        r = segc[s] + left;
    }
    s_cout[t] = r;
    for (int round = 0; round < 4 * kNormThreads + 8; ++round) {
        __syncthreads();             // every thread has read the previous round's flag
        if (t == 0) s_any = 0;
        __syncthreads();
        const i64 cin = s_cout[(t + kNormThreads - 1) % kNormThreads];
        __syncthreads();
        const i64 cy = cin == 0 ? 0 : d0 < d1 ? add_run(dig, d0, d1, cin, p, D) : cin;
        s_cout[t] = cy;
        if (cy != 0) s_any = 1;
        __syncthreads();
        if (!s_any) break;
    }
    if (t == 0) {
        out->status = s_any ? 1 : 0;
        out->zero = 1;
        out->ones = 1;
    }
}

// The flags are combined over the workgroup first, so each workgroup sends at most two atomics
// (one per thread kept 32 K of them queued on the same two words at N = 2^20).
__global__ void __launch_bounds__(256) lg_norm_check(const i64 *dig, u32 p, NDiv D, NormOut *out) {
    const u64 s = (u64)blockIdx.x * 256u + threadIdx.x;
    bool zero = true, ones = true;
    if (s < D.n / kNormK) {
        const u64 d0 = s * kNormK;
        u64 lo = cbit(d0, p, D);
        for (int k = 0; k < kNormK; ++k) {
            const u64 hi = cbit(d0 + k + 1, p, D);
            const i64 v = dig[d0 + k];
            zero = zero && v == 0;
            ones = ones && v == digit_mask((int)(hi - lo));
            lo = hi;
        }
    }
    const int all_zero = __syncthreads_and(zero ? 1 : 0);
    const int all_ones = __syncthreads_and(ones ? 1 : 0);
    if (threadIdx.x == 0) {
        if (!all_zero) atomicAnd(&out->zero, 0);
        if (!all_ones) atomicAnd(&out->ones, 0);
    }
}

__global__ void lg_norm_final(const i64 *dig, u32 p, NDiv D, NormOut *out) {
    u64 r = 0;
    for (u64 d = 0; d < D.n; ++d) {
        const u64 cd = cbit(d, p, D);
        if (cd >= 64) break;
        r |= (u64)dig[d] << cd;
    }
    const bool z = out->zero || out->ones;
    out->res64 = z ? 0 : r;
    out->zero = z ? 1 : 0;
}

}  // namespace

namespace {

// One thread per output digit; the output is the element array of the N'-digit engine (FP64:
// N' doubles, element j = digits 2j, 2j + 1; GL: N' words holding the signed digit).
template <bool IS_GL>
__global__ void __launch_bounds__(256) lg_relayout(const i64 *dig, u32 p, u64 n, u64 n2, void *elem2, u32 W2,
                                                   u32 H2) {
    const u64 k = (u64)blockIdx.x * 256ull + threadIdx.x;
    if (k >= n2) return;
    const i64 v = relayout_digit(dig, p, n, n2, k);
    if constexpr (IS_GL) reinterpret_cast<u64 *>(elem2)[digit_word<true>(k, W2, H2)] = (u64)v;
    else reinterpret_cast<double *>(elem2)[digit_word<false>(k, W2, H2)] = (double)v;
}

}  // namespace

namespace llh {

// elem2: the element array of the N'-digit job, whose schedule has a W2 x H2 element matrix
hipError_t large_relayout(int transform, const i64 *dig, u32 p, u64 n, u64 n2, void *elem2, u32 w2, u32 h2,
                          hipStream_t st) {
    const dim3 grid((unsigned)((n2 + 255) / 256));
    if (transform == kNTT)
        hipLaunchKernelGGL(lg_relayout<true>, grid, dim3(256), 0, st, dig, p, n, n2, elem2, w2, h2);
    else hipLaunchKernelGGL(lg_relayout<false>, grid, dim3(256), 0, st, dig, p, n, n2, elem2, w2, h2);
    return hipGetLastError();
}

void relayout_host(const i64 *dig, u64 p, u64 n, u64 n2, i64 *out) {
    for (u64 k = 0; k < n2; ++k) out[k] = relayout_digit(dig, p, n, n2, k);
}


// ---- schedules ------------------------------------------------------------------------------
bool large_launch_group0(int transform, u64 M, int W, int H, int CW, LargeLaunch &L);
bool large_launch_group1(int transform, u64 M, int W, int H, int CW, LargeLaunch &L);
bool large_launch_group2(int transform, u64 M, int W, int H, int CW, LargeLaunch &L);
bool large_launch_group3(int transform, u64 M, int W, int H, int CW, LargeLaunch &L);

namespace {
struct SchedRow {
    int group, transforms;
    u64 m;
    int w, h, cw;
};
const SchedRow kSchedRows[] = {
#define LLS(G, TR, MV, WV, HV, CWV) {G, TR, MV, WV, HV, CWV},
#include "ll_large_scheds.inc"
#undef LLS
};
bool row_runs(const SchedRow &r, int transform) { return (r.transforms & (transform == kNTT ? 2 : 1)) != 0; }
}  // namespace

// The schedules of M elements for a transform, in list order (the first is the default).
std::vector<LargeSched> large_schedules(int transform, u64 M) {
    std::vector<LargeSched> v;
    for (const SchedRow &r : kSchedRows)
        if (r.m == M && row_runs(r, transform)) v.push_back(LargeSched{r.w, r.h, r.cw});
    return v;
}

// Kernels of schedule `sched` (index into large_schedules) for M elements.  FP64: M = m 2^k with
// m in {1, 3, 5, 7, 15}, the odd factor in the row length W; GL: powers of two (the field has no
// 3rd or 7th roots of unity, and no N-th root of 2 for N = 3 * 2^k).
bool large_launch_for(int transform, u64 M, int sched, LargeLaunch &L) {
    int k = 0;
    for (const SchedRow &r : kSchedRows) {
        if (r.m != M || !row_runs(r, transform)) continue;
        if (k++ != sched) continue;
        switch (r.group) {
            case 0: return large_launch_group0(transform, M, r.w, r.h, r.cw, L);
            case 1: return large_launch_group1(transform, M, r.w, r.h, r.cw, L);
            case 2: return large_launch_group2(transform, M, r.w, r.h, r.cw, L);
            case 3: return large_launch_group3(transform, M, r.w, r.h, r.cw, L);
            default: return false;
        }
    }
    return false;
}

int large_blocks(const LargeLaunch &L) { return L.w_len / L.cw; }

// k squarings: forward column pass, then per squaring the row pass and the fused column pass
// (the last squaring ends with the inverse column pass, leaving its carries in the carry buffer).
// Squaring i reports its round-off in err slot i; its fused pass tags its carry words A.tag + i.
hipError_t large_squarings(const LargeLaunch &L, LargeArgs A, int k, hipStream_t st) {
    const int W = L.w_len, H = L.h_len;
    const dim3 gcol(W / L.cw), bcol(L.cw * (H / kR));
    const dim3 grow(H / 2), brow(2 * (W / kR));
    const u32 tag0 = A.tag;
    hipLaunchKernelGGL(L.fwd, gcol, bcol, 0, st, A);
    for (int i = 0; i < k; ++i) {
        A.err_slot = (u32)i;
        A.tag = tag0 + (u32)i;
        hipLaunchKernelGGL(L.rows, grow, brow, 0, st, A);
        if (i + 1 < k) hipLaunchKernelGGL(L.fused, gcol, bcol, 0, st, A);
        else hipLaunchKernelGGL(L.inv, gcol, bcol, 0, st, A);
    }
    return hipGetLastError();
}

hipError_t large_apply_carry(const LargeLaunch &L, const LargeArgs &A, hipStream_t st) {
    hipLaunchKernelGGL(L.apply, dim3(L.w_len / L.cw), dim3(256), 0, st, A);
    return hipGetLastError();
}

size_t large_normalize_bytes(u64 n) { return kNormSegOffset + (size_t)(n / kNormK) * sizeof(i64); }

// out: large_normalize_bytes(N) bytes, a NormOut followed by the segment carries.  N is a
// multiple of kNormK (every large length is).
hipError_t large_normalize(int transform, const void *elem, i64 *dig, const LargeArgs &A, void *out,
                           hipStream_t st) {
    const NDiv D = ndiv_of(A);
    NormOut *no = (NormOut *)out;
    i64 *segc = (i64 *)((char *)out + kNormSegOffset);
    const dim3 g((unsigned)((A.n / kNormK + 255) / 256));
    if (transform == kNTT)
        hipLaunchKernelGGL(lg_norm_local<true>, g, dim3(256), 0, st, elem, dig, segc, A.p, D, A.w_len, A.h_len);
    else hipLaunchKernelGGL(lg_norm_local<false>, g, dim3(256), 0, st, elem, dig, segc, A.p, D, A.w_len, A.h_len);
    hipLaunchKernelGGL(lg_norm_carry, dim3(1), dim3(kNormThreads), 0, st, dig, (const i64 *)segc, A.p, D, no);
    hipLaunchKernelGGL(lg_norm_check, g, dim3(256), 0, st, (const i64 *)dig, A.p, D, no);
    hipLaunchKernelGGL(lg_norm_final, dim3(1), dim3(1), 0, st, (const i64 *)dig, A.p, D, no);
    return hipGetLastError();
}

}  // namespace llh
