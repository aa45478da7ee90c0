// ll_internal.hpp -- host/device structures of the batched engine and the host helpers
// (digit layout, size rules, tables) shared by every translation unit of libllhip.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

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
    u64 iters;          // squarings to run in total (already min(p-2, cap))
    u32 slot;           // index into the result array
    u32 it0;            // squarings already done: the item resumes from a checkpoint
    u32 init;           // resume entry + 1 (ClassDesc::resume, kResumeStride digits each); 0: s_0
    u32 pad;
};

struct BatchResult {
    i32 verdict;        // 1 when s == 0 after the run
    i32 status;         // ItemStatus
    u64 res64;          // low 64 bits of the canonical residue
    double maxerr;      // FP64: worst |v - rint(v)| over the run (0 for the exact transforms)
    u64 iters_done;     // squarings done when the item stopped
    u64 t0, t1;         // device wall clock at start / end of the item
    u64 ck_iter;        // kStRoundoff: squarings done at the last good checkpoint
    u32 ck;             // kStRoundoff: its rollback entry + 1 (ClassDesc::rb_area); 0: none
    u32 pad;
};

// Digits per entry of the batch engine's rollback and resume areas (its longest length).
constexpr u32 kResumeStride = 8192;

// A batch exponent that continues at length n2 from rollback entry rb (its checkpoint at n).
struct ResumeReq {
    u32 p, n, n2, rb;
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
    const void *twp;    // tw in per-pass order of the class's radix schedule (PassTable)
    const void *twpi;   // GL: twi in per-pass order
    const void *wtab;   // FP64: two-level weights, 4 x kWtabSlot doubles: forward low | forward
                        // high | inverse low | inverse high, w(e) = low[e & (S-1)] * high[e >> log_s]
    i32 log_s;          // S = 2^log_s
    i32 cls;            // size class (ll_batch.hip); 0 = the one-word lane
    i32 carry2;         // FP64: two carry hand-off rounds (fewer than log2 N + 3 bits per digit)
    // Checkpoints (FP64; all null: off).  At every round-off event that finds a group's run
    // within the limit, the group writes its elements' raw words (natural order) to its slot,
    // workgroup b at ck_slots + b * 16 * WG words, group g at + g * N; when an event finds the
    // limit passed, the group copies the slot to entry atomicAdd(rb_count) < rb_cap of rb_area.
    i64 *ck_slots;
    i64 *rb_area;
    u32 *rb_count;
    u32 rb_cap;
    u32 pad2;
    const i64 *resume;  // resume entries of the items that continue from a checkpoint
};
constexpr int kWtabSlot = 128;

// ---- batched engine size classes and launches (ll_batch.hip) ------------------------------
int batch_num_classes();
int batch_class_of_m(u64 M);            // class of transform length M (0 = none; the base class)
int batch_class_base(int c);            // the base class of a schedule variant (itself otherwise)
void batch_class_radices(int c, int r[4]);   // radix list (unused entries 0)
u64 batch_class_m(int c);               // transform length of class c
bool batch_class_pow2(int c);
int batch_num_families();
int batch_class_family(int c);          // kernel family: 0 powers of two to M = 2048, 1 3 * 2^k,
                                        // 2 5 * 2^k, 3 7 * 2^k, 4 the power of two M = 4096
int batch_class_factor(int c);          // length family (bit of llh_set_batch_families): 0 powers
                                        // of two, 1 3 * 2^k, 2 5 * 2^k, 3 7 * 2^k
int batch_class_groups(int c);          // exponents per workgroup
int batch_family_wg(int family);
bool batch_pass_exponents(int c, std::vector<u32> &ex);   // per-pass twiddle order of class c

// ---- large engine (ll_large.hip) ----------------------------------------------------------
// Kernel arguments (by value).  Table pointers are device addresses.
struct LargeArgs {
    void *elem;              // [M] elements
    void *carry;             // [nblk][H] row-segment carry-outs (double or i64)
    unsigned long long *err; // per-iteration max round-off slots (double bits)
    const void *twW, *twWi;  // twiddles of the W-point row transform in per-pass order (GL: forward, inverse)
    const void *twH, *twHi;  // the same for the H-point column transform
    const void *tmL, *tmH;   // w_M^e = tmL[e & (S-1)] * tmH[e >> s]
    const void *tmLi, *tmHi; // GL inverse of the above
    const void *wfL, *wfH;   // forward weight of exponent e: wfL[e & (Sw-1)] * wfH[e >> sw]
    const void *wiL, *wiH;   // inverse weight (with the transform's 1/M or 1/N)
    u32 p, log_n, log_s, log_sw;     // log_n: log2 N when N is a power of two, else 0
    u32 err_slot;
    u64 n;                            // digit count N
    double inv_n;                     // 1.0 / N
    u32 p_mod_n, p_div_n;             // p mod N, floor(p / N)
    u32 tag;                          // fused column pass: this launch's carry-word tag (16 bits, nonzero)
    void *tcarry;                     // fused column pass: [nblk][H] tagged carry words
    unsigned *fault;                  // raised when a carry hand-off fails (timeout, carry range)
    u32 xcd_blocks;                   // fused column pass: XCD-aware block order (job runs alone)
    u32 w_len, h_len;                 // W and H of the schedule (the element layout's shape)
};

struct LargeLaunch {
    void (*fwd)(LargeArgs);
    void (*rows)(LargeArgs);
    void (*inv)(LargeArgs);
    void (*fused)(LargeArgs);
    void (*apply)(LargeArgs);
    int w_len, h_len;   // W and H (W carries the odd factor)
    int cw;             // columns per column-pass workgroup
    // Per-pass twiddle tables of the row and column transforms: entry e (e < *_tab) holds
    // w_W^{row_exp(e)} (w_H^{col_exp(e)}), the order in which the kernels' passes read them.
    u32 (*row_exp)(int e);
    u32 (*col_exp)(int e);
    int row_tab, col_tab;
};

// One schedule of the large engine: the H x W split of the M elements and the column-pass
// block width (ll_large_scheds.inc).
struct LargeSched {
    int w, h, cw;
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
