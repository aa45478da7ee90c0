// ll_batch.hip -- the batched, LDS-resident Lucas-Lehmer engine: every exponent of a list
// in ONE persistent launch, whatever its transform length.
//
// Counterpart of cuda/ll_ntt.cu:74-167 (ll_kernel, one block per exponent and one launch
// per length group) and of the FP64 loop lucaslehmer.cpp:281-311, re-designed for gfx950:
//
//  * A workgroup is 256 threads.  A size class with M transform elements is run by groups
//    of G = M/R threads (R elements per thread, R = 16 for M >= 256), so one workgroup holds
//    256/G exponents of that class at once, all stepping in lock-step.
//  * Workgroups pull exponents from one queue per class (an atomic counter over items the
//    host sorted longest-first), classes in descending length order, then the one-word lane
//    (p < 64) -- one launch for the whole list, no per-length launches.
//  * The residue lives in registers in a "column" layout (thread t holds elements t + G s);
//    only the transform's middle hand-offs, the real-FFT pairing and the carry hand-off go
//    through LDS.  Nothing touches HBM between the first and the last squaring except the
//    read-only twiddle and weight tables (L2-resident).
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

constexpr int kWG = 512;
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr double kErrLimit = 0.4;   // round-off above this aborts the exponent (rollback at 2N)

constexpr int kMaxGroups = 128;

struct Ctl {
    u32 item[kMaxGroups];
    u32 bad[kMaxGroups];
    unsigned long long err[kMaxGroups];
};

// LDS of the batch kernels: every group's residue buffer (padded) and the control block.
// Referenced directly (not through pointer arguments) so every access is a ds_* instruction.
// (a group of G threads holds M = G R elements, padded one per R: lds_pad_s<log2 R>)
constexpr int kSmemBytes = kWG * 8 * 16 * 9 / 8;
__shared__ __attribute__((aligned(16))) char g_smem[kSmemBytes];
__shared__ Ctl g_ctl;
// FP64 two-level weight tables of the class being run (ClassDesc::wtab)
__shared__ double g_wtab[4 * llh::kWtabSlot];

// Radix schedule of each size class (index log2 M): M, G = M / R threads per exponent, R
// elements per thread, radices R1 .. R4 (first = last = R; 0 = no pass).
struct Sch {
    int M, G, R, R1, R2, R3, R4;
};
constexpr Sch kSch[13] = {
    {0, 0, 0, 0, 0, 0, 0},          {0, 0, 0, 0, 0, 0, 0},       {0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0},          {16, 4, 4, 4, 4, 0, 0},      {32, 8, 4, 4, 2, 4, 0},
    {64, 8, 8, 8, 8, 0, 0},         {128, 16, 8, 8, 2, 8, 0},    {256, 32, 8, 8, 4, 8, 0},
    {512, 64, 8, 8, 8, 8, 0},       {1024, 128, 8, 8, 2, 8, 8},  {2048, 256, 8, 8, 4, 8, 8},
    {4096, 512, 8, 8, 8, 8, 8},
};

// Exact canonical residue of N near-balanced digits (natural order, i64, in LDS):
// full carry propagation with wrap (2^p == 1), then Res64 = low 64 bits of the value in
// [0, 2^p - 1), with the all-ones pattern (== M_p == 0) folded to zero.
// Restates the GMP rebuild of lucaslehmer.cpp:312-323 / cuda/ll_ntt.cu:238-251 without GMP.
__device__ int normalize_res64(i64 *d, u32 p, int log_n, u64 &res, int &zero) {
    const u32 N = 1u << log_n;
    i64 carry = 0;
    for (int round = 0; round < 4; ++round) {
        u64 cj = 0;
        for (u32 j = 0; j < N; ++j) {
            if (round > 0 && carry == 0) break;
            u64 cn = ((u64)(j + 1) * p + N - 1) >> log_n;
            int b = (int)(cn - cj);
            cj = cn;
            i64 v = d[j] + carry;
            d[j] = v & (((i64)1 << b) - 1);
            carry = v >> b;
        }
        if (carry == 0) break;
    }
    if (carry != 0) return -1;
    u64 r = 0, cj = 0;
    int z = 1, ones = 1;
    for (u32 j = 0; j < N; ++j) {
        u64 cn = ((u64)(j + 1) * p + N - 1) >> log_n;
        int b = (int)(cn - cj);
        u64 dj = (u64)d[j];
        if (dj) z = 0;
        if (dj != (((u64)1 << b) - 1)) ones = 0;
        if (cj < 64) r |= dj << cj;
        cj = cn;
    }
    if (ones) { r = 0; z = 1; }
    res = r;
    zero = z;
    return 0;
}

template <bool IS_GL, int M, int G, int R, int R1, int R2, int R3, int R4>
__device__ __attribute__((noinline)) void run_class(const llh::ClassDesc &cd, const llh::BatchItem *__restrict__ items,
                          llh::BatchResult *__restrict__ out, u32 *counter) {
    Ctl &ctl = g_ctl;
    typedef typename std::conditional<IS_GL, FGL, FC64>::type F;
    typedef typename F::T T;
    typedef typename F::W W;
    typedef typename std::conditional<IS_GL, u64, double>::type WT;
    constexpr int NG = kWG / G;
    constexpr int PS = R == 8 ? 3 : 2;            // LDS pad shift: one pad element per R
    static_assert(R == 8 || R == 4, "per-thread counts 4 and 8");
    constexpr int REG = M + (M >> PS);
    static_assert(kWG / G * REG * (int)sizeof(T) <= kSmemBytes, "LDS buffer too small");
    static_assert(NG <= kMaxGroups, "too many groups per workgroup");

    const int tid = threadIdx.x, g = tid / G, t = tid % G;
    T *lds = reinterpret_cast<T *>(g_smem) + g * REG;
    const int log_n = cd.log_n;
    const u32 N = 1u << log_n, nmask = N - 1;
    const gptr<W> tw = as_global<W>(cd.tw);
    const gptr<W> twp = as_global<W>(cd.twp);
    const gptr<W> twpi = IS_GL ? as_global<W>(cd.twpi) : twp;
    const gptr<WT> wf = as_global<WT>(cd.wf);
    const gptr<WT> wi = as_global<WT>(cd.wi);
    const u32 smask = (1u << cd.log_s) - 1;
    const int log_s = cd.log_s;
    if constexpr (!IS_GL) {
        static_assert(4 * llh::kWtabSlot <= kWG, "one weight entry per thread");
        if (tid < 4 * llh::kWtabSlot) g_wtab[tid] = ld_table(as_global<double>(cd.wtab), (u32)tid);
    }
    // weight of exponent e (forward: slots 0, 1; inverse: slots 2, 3)
    auto wlds = [&](u32 e, int inv) {
        return g_wtab[(2 * inv) * llh::kWtabSlot + (e & smask)] * g_wtab[(2 * inv + 1) * llh::kWtabSlot + (e >> log_s)];
    };

    u32 item = kNone, p = N + 1;
    u64 target = 0, it = 0, t0 = 0;
    double emax = 0.0;
    T x[R];

    auto init_item = [&]() {
        if (item != kNone) {
            const llh::BatchItem bi = items[item];
            p = bi.p;
            target = bi.iters;
        } else {
            p = N + 1;          // a valid layout for an idle group; its digits stay zero
            target = 0;
        }
        it = 0;
        emax = 0.0;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if constexpr (IS_GL) x[s] = 0; else x[s] = cmk(0.0, 0.0);
        }
        if (item != kNone && t == 0) {
            if constexpr (IS_GL) x[0] = 4; else x[0] = cmk(4.0, 0.0);
        }
        t0 = wall_clock64();
    };

    // first fetch for every group
    if (t == 0) {
        u32 k = atomicAdd(counter, 1u);
        ctl.item[g] = (k < cd.count) ? cd.first + k : kNone;
    }
    __syncthreads();
    item = ctl.item[g];
    init_item();

    for (;;) {
        const bool has = item != kNone;
        if (!__syncthreads_or(has)) break;
        double eit = 0.0;

        if constexpr (!IS_GL) {
            // ---- weight: z_j = x_{2j} a_{2j} + i x_{2j+1} a_{2j+1}
            const int tw0 = ll_opaque(t);
            const u32 pw = ll_opaque(p);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tw0 + G * s;
                const u32 dp0 = 2u * j * pw;
                const u32 e0 = (0u - dp0) & nmask, e1 = (0u - dp0 - pw) & nmask;
                x[s] = cmk(x[s].re * wlds(e0, 0), x[s].im * wlds(e1, 0));
            }
            transform<F, M, G, R, R1, R2, R3, R4, false, true, PS>(x, lds, t, twp);
            // ---- square the real spectrum in packed form (pairs k, M-k)
            const int tq = ll_opaque(t);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lds_pad(tq + G * q)] = x[q];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int k = tq + G * q;
                const int km = (M - k) & (M - 1);
                const C64 zm = lds[lds_pad(km)];
                const C64 D = csub(x[q], cconj(zm));
                const C64 w = ld_table(tw, (u32)k);
                const C64 A = cmk(1.0 + w.re, w.im);
                const C64 AD = cmul(A, csqr(D));
                const C64 Z2 = csqr(x[q]);
                x[q] = cmk(__fma_rn(-0.25, AD.re, Z2.re), __fma_rn(-0.25, AD.im, Z2.im));
            }
            transform<F, M, G, R, R1, R2, R3, R4, true, true, PS>(x, lds, t, twp);
            // ---- unweight, round, error, -2, carry within the pair
            double cout[R];
            const int tu = ll_opaque(t);
            const u32 pu = ll_opaque(p);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tu + G * s;
                const u32 dp0 = 2u * j * pu;
                const u32 c0 = (dp0 + N - 1) >> log_n, c1 = (dp0 + pu + N - 1) >> log_n,
                          c2 = (dp0 + 2u * pu + N - 1) >> log_n;
                const int b0 = (int)(c1 - c0), b1 = (int)(c2 - c1);
                const u32 e0 = (0u - dp0) & nmask, e1 = (0u - dp0 - pu) & nmask;
                const double v0 = x[s].re * wlds(e0, 1), v1 = x[s].im * wlds(e1, 1);
                double r0 = __builtin_rint(v0), r1 = __builtin_rint(v1);
                eit = fmax(eit, fmax(fabs(v0 - r0), fabs(v1 - r1)));
                if (j == 0 && has) r0 -= 2.0;
                const double cA = floor(__fma_rn(r0, ldexp(1.0, -b0), 0.5));
                const double x0 = __fma_rn(-cA, ldexp(1.0, b0), r0);
                r1 += cA;
                const double cB = floor(__fma_rn(r1, ldexp(1.0, -b1), 0.5));
                const double x1 = __fma_rn(-cB, ldexp(1.0, b1), r1);
                x[s] = cmk(x0, x1);
                cout[s] = cB;
            }
            // ---- carry hand-offs: element j takes the carry out of element j-1 (wrap at 0).
            // Round 1 moves the large carries and re-splits both digits of the pair; round 2
            // moves the remaining small carry into the next pair's low digit.  Two rounds keep
            // the digits near balanced even at a few bits per digit (one round lets the slack,
            // ~sqrt(N)/4, feed back into the next squaring and grow).
            double *ldsd = reinterpret_cast<double *>(lds);
            const int tc = ll_opaque(t);
            const u32 pc = ll_opaque(p);
            __syncthreads();
#pragma unroll
            for (int s = 0; s < R; ++s) ldsd[tc + G * s] = cout[s];
            __syncthreads();
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tc + G * s;
                const double cin = ldsd[(j - 1) & (M - 1)];
                const u32 dp0 = 2u * j * pc;
                const u32 c0 = (dp0 + N - 1) >> log_n, c1 = (dp0 + pc + N - 1) >> log_n,
                          c2 = (dp0 + 2u * pc + N - 1) >> log_n;
                const int b0 = (int)(c1 - c0), b1 = (int)(c2 - c1);
                double x0 = x[s].re + cin;
                const double cC = floor(__fma_rn(x0, ldexp(1.0, -b0), 0.5));
                x0 = __fma_rn(-cC, ldexp(1.0, b0), x0);
                double x1 = x[s].im + cC;
                const double cD = floor(__fma_rn(x1, ldexp(1.0, -b1), 0.5));
                x1 = __fma_rn(-cD, ldexp(1.0, b1), x1);
                x[s] = cmk(x0, x1);
                cout[s] = cD;
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < R; ++s) ldsd[tc + G * s] = cout[s];
            __syncthreads();
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tc + G * s;
                x[s].re += ldsd[(j - 1) & (M - 1)];
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
            transform<F, M, G, R, R1, R2, R3, R4, false, true, PS>(x, lds, t, twp);
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = gl_sqr(x[q]);
            transform<F, M, G, R, R1, R2, R3, R4, true, true, PS>(x, lds, t, twpi);
            i64 cout[R];
            const int tu = ll_opaque(t);
            const u32 pu = ll_opaque(p);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const u32 j = tu + G * s;
                const u32 dp = j * pu;
                const u32 e = (0u - dp) & nmask;
                const int b = (int)(((dp + pu + N - 1) >> log_n) - ((dp + N - 1) >> log_n));
                i64 v = gl_to_i64(gl_mul(x[s], ld_table(wi, e)));
                if (j == 0 && has) v -= 2;
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
                        const u32 dp = j * pc;
                        const int b = (int)(((dp + pc + N - 1) >> log_n) - ((dp + N - 1) >> log_n));
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
        const bool done = has && it >= target;
        const bool bad = has && eit > kErrLimit;
        if (__syncthreads_or(done || bad)) {
            // ---- finish phase: flags, error reduction, exact normalisation, refill
            if (t == 0) { ctl.bad[g] = 0; ctl.err[g] = 0; }
            __syncthreads();
            if (bad) ctl.bad[g] = 1;
            __syncthreads();
            const bool gbad = has && ctl.bad[g] != 0;
            const bool gfin = done || gbad;
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
                if (gbad) {
                    r.status = llh::kStRoundoff;
                } else {
                    u64 res = 0;
                    int zero = 0;
                    r.status = normalize_res64(dig, p, log_n, res, zero) == 0 ? llh::kStOk : llh::kStNotRun;
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
        out[bi.slot] = r;
    }
}

template <bool IS_GL>
__global__ void __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(4, 4)))
ll_batch_kernel(const llh::ClassDesc *__restrict__ classes, int ncls,
                const llh::BatchItem *__restrict__ items, llh::BatchResult *__restrict__ out,
                u32 *counters) {
    for (int c = 0; c < ncls; ++c) {
        const llh::ClassDesc cd = classes[c];
        switch (cd.log_m) {
#define LL_CLASS(L)                                                                                       \
    case L:                                                                                               \
        run_class<IS_GL, kSch[L].M, kSch[L].G, kSch[L].R, kSch[L].R1, kSch[L].R2, kSch[L].R3, kSch[L].R4>( \
            cd, items, out, counters + c);                                                                \
        break;
            LL_CLASS(12) LL_CLASS(11) LL_CLASS(10) LL_CLASS(9) LL_CLASS(8) LL_CLASS(7) LL_CLASS(6)
            LL_CLASS(5) LL_CLASS(4)
#undef LL_CLASS
            case 0: run_tiny(cd, items, out, counters + c); break;
            default: break;
        }
        __syncthreads();
    }
}

}  // namespace

namespace llh {

// Launch the persistent batch kernel on `stream`; grid = resident workgroups.
hipError_t launch_batch(int transform, const ClassDesc *d_classes, int ncls, const BatchItem *d_items,
                        BatchResult *d_out, u32 *d_counters, int grid, hipStream_t stream) {
    if (transform == kNTT)
        hipLaunchKernelGGL(ll_batch_kernel<true>, dim3(grid), dim3(kWG), 0, stream, d_classes, ncls,
                           d_items, d_out, d_counters);
    else
        hipLaunchKernelGGL(ll_batch_kernel<false>, dim3(grid), dim3(kWG), 0, stream, d_classes, ncls,
                           d_items, d_out, d_counters);
    return hipGetLastError();
}

template <int L>
static void pass_exponents(std::vector<u32> &ex) {
    typedef PassTable<kSch[L].R1, kSch[L].R2, kSch[L].R3, kSch[L].R4, 0> PT;
    static_assert(PT::M == kSch[L].M, "schedule must cover the class length");
    ex.resize(PT::SIZE);
    for (int e = 0; e < PT::SIZE; ++e) ex[e] = PT::exponent(e);
}

// Full-table exponent of every entry of the class's per-pass twiddle table.
bool batch_pass_exponents(int log_m, std::vector<u32> &ex) {
    switch (log_m) {
        case 4: pass_exponents<4>(ex); return true;
        case 5: pass_exponents<5>(ex); return true;
        case 6: pass_exponents<6>(ex); return true;
        case 7: pass_exponents<7>(ex); return true;
        case 8: pass_exponents<8>(ex); return true;
        case 9: pass_exponents<9>(ex); return true;
        case 10: pass_exponents<10>(ex); return true;
        case 11: pass_exponents<11>(ex); return true;
        case 12: pass_exponents<12>(ex); return true;
        default: return false;
    }
}

int batch_blocks_per_cu(int transform) {
    int nb = 0;
    hipError_t e = transform == kNTT
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ll_batch_kernel<true>, kWG, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ll_batch_kernel<false>, kWG, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

}  // namespace llh
