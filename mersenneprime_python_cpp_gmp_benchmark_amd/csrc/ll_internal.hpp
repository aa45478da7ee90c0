// ll_internal.hpp -- host/device structures of the batched engine and the host helpers
// (digit layout, size rules, tables) shared by every translation unit of libllhip.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string>

namespace llh {

typedef uint64_t u64;
typedef int64_t i64;
typedef uint32_t u32;
typedef int32_t i32;

// transform selector (C ABI values)
enum Transform : int { kFP64 = 0, kNTT = 1 };

// per-exponent status
enum ItemStatus : int {
    kStOk = 0,          // ran min(p-2, iters) squarings, verdict/res64 valid
    kStRoundoff = 1,    // FP64 round-off above the threshold at some iteration (redo at 2N)
    kStNotRun = 2,      // not run (no supported length)
    kStTrivial = 3,     // p < 3 or even: answered by the argument rules, nothing ran
};

// One work item of the persistent launch.  log_n == 0 marks the one-word lane (p < 64).
struct BatchItem {
    u32 p;
    u32 log_n;
    u64 iters;          // squarings to run (already min(p-2, cap))
    u32 slot;           // index into the result array
    u32 pad;
};

struct BatchResult {
    i32 verdict;        // 1 when s == 0 after the run
    i32 status;         // ItemStatus
    u64 res64;          // low 64 bits of the canonical residue
    double maxerr;      // FP64: worst |v - rint(v)| over the run (0 for the exact transforms)
    u64 iters_done;
    u64 t0, t1;         // device wall clock at start / end of the item
};

// One size class of a launch: items [first, first+count) share log_n and tables.
struct ClassDesc {
    i32 log_m;          // log2 of the transform length in elements (C64: N/2, GL: N)
    i32 log_n;          // log2 of the digit count N
    u32 first;
    u32 count;
    const void *tw;     // w_M^j, j < M  (C64: {re,im} doubles; GL: u64)
    const void *twi;    // GL: inverse roots; C64: unused (conjugated in registers)
    const void *wf;     // forward weights indexed by e = (-j p) mod N: 2^(e/N) / w2^e
    const void *wi;     // inverse weights incl. the transform's 1/M or 1/N scale
};

// ---- host helpers (ll_tables.cpp) ---------------------------------------------------------
u64 fft_length_conservative(u64 p, double max_bpw);   // lucaslehmer.cpp:244-254
u64 ntt_length_ref(u64 p);                            // cuda/ll_ntt.cu:176-180 (extended)
int ntt_max_bits(u64 n);                              // cuda/ll_ntt.cu:172-175

// Goldilocks host arithmetic
u64 gl_mul_h(u64 a, u64 b);
u64 gl_pow_h(u64 a, u64 e);
u64 gl_inv_h(u64 a);
u64 gl_root2_h(u64 n);                 // n-th root of 2 (0 if none)
u64 gl_root_unity_aligned(u64 n);      // primitive n-th root w with w^(n/16) == 2^12

void set_error(const std::string &msg);

}  // namespace llh
