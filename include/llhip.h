/*
 * llhip.h -- C ABI of the MI355X Lucas-Lehmer engine (libllhip.so).
 *
 * Plain scalars, caller-owned host arrays, int status codes; no exception crosses the
 * boundary.  All device memory is owned by the library.  Calls are thread-safe: GPU work
 * is serialised on one HIP stream per device (the reference's CUDA module serialises on the
 * default stream, cuda/ll_ntt.cu:229-231).  The pybind11 module `lucaslehmer` releases the
 * GIL around every call, as the reference does (lucaslehmer.cpp:38, :280).
 *
 * Transforms: LLH_FP64 = FP64 Crandall-Fagin IBDWT (real FFT, per-iteration round-off
 * check, rollback to 2N above 0.4); LLH_NTT = exact Goldilocks-field IBDWT.
 * Res64 is always the low 64 bits of the canonical residue in [0, 2^p - 1).
 */
#ifndef LLHIP_H
#define LLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LLH_OK 0
#define LLH_ERR_ARG -1      /* invalid argument (bad transform, NULL output, ...) */
#define LLH_ERR_HIP -2      /* HIP runtime error; llh_last_error() has the text */
#define LLH_ERR_UNSUPPORTED -3

#define LLH_FP64 0
#define LLH_NTT 1

/* per-exponent status values written by llh_batch */
#define LLH_ST_OK 0         /* ran min(p-2, max_iters) squarings; verdict and res64 valid */
#define LLH_ST_ROUNDOFF 1   /* FP64 round-off stayed above 0.4 up to the longest length */
#define LLH_ST_NOT_RUN 2    /* no supported transform length (cf. seconds < 0 in cuda/ll_ntt.cu:256-263) */
#define LLH_ST_TRIVIAL 3    /* p < 3 or even: answered by lucaslehmer.cpp:30-32's rules */

/* Select the device (default 0) and create the library's stream.  No reference counterpart:
 * the reference's CUDA module uses the current device implicitly. */
int llh_init(int device);

/* Text of the last error on this thread ("" if none). */
const char *llh_last_error(void);

/* Device name, compute units, global memory bytes, gfx arch name.
 * Replaces device_info() cuda/ll_ntt.cu:269-272. */
int llh_device_info(char *name, size_t name_len, int *compute_units, uint64_t *global_mem_bytes,
                    char *arch, size_t arch_len);

/* Conservative FP64 length rule: smallest {1,3,5}*16*2^k with p/N <= max_bpw
 * (max_bpw <= 0: the current bits-per-word setting).  Replaces fft_length lucaslehmer.cpp:244-254. */
uint64_t llh_fft_length(uint64_t p, double max_bpw);

/* Length the batched engine would use for p with `transform` (0 = not in the batched engine).
 * For LLH_NTT this extends ntt_length cuda/ll_ntt.cu:176-180 down to N = 16. */
uint64_t llh_batch_length(uint64_t p, int transform);

/* FP64 length of the `lucaslehmer` module's FFT functions: the conservative rule's length
 * (llh_fft_length, lucaslehmer.cpp:244-254) when an engine runs it, else the shortest length of
 * the engines above it (max_bpw <= 0: the bits-per-word knob).  0 = none. */
uint64_t llh_fp64_length(uint64_t p, double max_bpw);

/* Reference NTT length rule (N in 64..8192, 0 = too large).  cuda/ll_ntt.cu:176-180. */
uint64_t llh_ntt_length(uint64_t p);

/* LLH_NTT: the reference's bound, 8192 * max_bits_per_word(8192) = 204,800, the largest p its
 * length rule covers (max_exponent cuda/ll_ntt.cu:279).  Every p above it runs too, on the large
 * engine.  LLH_FP64 (no reference counterpart): the batched FP64 engine's bound, as
 * llh_batch_max_exponent. */
uint64_t llh_max_exponent(int transform);

/* Largest exponent `transform` handles in the batched (LDS-resident) engine: NTT 4096 * 25 =
 * 102,400; FP64 the batch bits-per-digit cap times 8192.  Larger p run on the large engine.
 * No reference counterpart (the reference's single kernel has one bound, cuda/ll_ntt.cu:279). */
uint64_t llh_batch_max_exponent(int transform);

/* Bits-per-word cap of the FP64 length rule (default 18).
 * Replaces set_fft_bits_per_word lucaslehmer.cpp:368-369 (atomic here, not a plain static). */
int llh_set_fft_bits_per_word(double bits);
double llh_get_fft_bits_per_word(void);

/* Bits-per-digit cap of the batched FP64 engine's lengths (default 21: its worst round-off over
 * every prime <= 100001 is then 0.21, half the 0.4 rollback threshold).  No reference
 * counterpart; the reference uses one cap for every length. */
int llh_set_batch_bits_per_word(double bits);
double llh_get_batch_bits_per_word(void);

/* Length families the batched FP64 engine may pick: bit 0 powers of two (always on), bit 1
 * 3 * 2^k, bit 2 5 * 2^k, bit 3 7 * 2^k (default 0xF).  No reference counterpart: the reference's
 * GPU batch runs powers of two only (cuda/ll_ntt.cu:176-180). */
int llh_set_batch_families(int mask);
int llh_get_batch_families(void);

/* Measured cost (device microseconds per exponent-squaring at full occupancy) of batched
 * length n; the engine runs, of the lengths within the bits-per-digit cap (up to twice the
 * shortest), the cheapest.  Defaults: the autotuner's table for MI355X; us = 0 forgets a cost.
 * No reference counterpart (the reference sizes each exponent by the cap alone,
 * lucaslehmer.cpp:244-254). */
int llh_set_batch_cost(uint64_t n, int transform, double us);
double llh_get_batch_cost(uint64_t n, int transform);

/* llh_batch with exponent i run at length n_force[i] (n_force NULL or an entry 0: the engine's
 * choice) and flags bit 0 = no rollback.  No reference counterpart: the entry point of the
 * autotuner and of callers that pick lengths themselves (the learned length of nselect). */
int llh_batch_n(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, const uint64_t *n_force, int flags,
                int32_t *verdict, uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used,
                int32_t *status);

/* Many exponents in one persistent launch.  Replaces lucas_lehmer_batch cuda/ll_ntt.cu:259-267
 * (and the per-thread engine calls of PerfNumMultiCLL.py:189-193).  max_iters = 0 runs the full
 * p-2 squarings.  Outputs are indexed like ps[] (the reference returns them grouped by length);
 * any output pointer except verdict/res64/status may be NULL.  seconds = device time of that
 * exponent; n_used = transform length (digits) actually used (after rollbacks). */
int llh_batch(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, int32_t *verdict,
              uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used, int32_t *status);

/* llh_batch for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and outputs as llh_batch; its device memory (the launch's items, results, class
 * entries and queue counters, about 1.2 MB) is one file-scope array per GPU, and the batch tables
 * are built by llh_init, so call llh_init first: the call then allocates nothing.  Limits: at most
 * 16384 exponents (LLH_ERR_ARG beyond), and every p within the batched engine (FP64: p <= the
 * batch cap * 8192; NTT: p <= 102,400; LLH_ERR_UNSUPPORTED otherwise, before any work).  Calls
 * that overlapped on one GPU would race on the array; the library's lock serialises a process's
 * calls, and each process has its own array. */
int llh_batch_noalloc(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, int32_t *verdict,
                      uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used, int32_t *status);

/* Device time of the kernel launches made by the last llh_batch / llh_lucas_lehmer call
 * (HIP events on the library's stream) and how many launches it took (1 + rollback launches).
 * No reference counterpart: measurement hook for the roofline report. */
int llh_last_launch_stats(double *kernel_ms, uint32_t *launches);

/* Squarings the engines ran in the last llh_batch / llh_lucas_lehmer call (those of runs that
 * rolled back included), how many rollbacks it took, how many of them resumed from a checkpoint,
 * and the iteration the last of those resumed from.  No reference counterpart: the reference
 * aborts at round-off 0.35 and redoes the whole test with GMP (lucaslehmer.cpp:296, :331-335). */
int llh_last_run_stats(uint64_t *squarings, uint32_t *rollbacks, uint32_t *resumed, uint64_t *resumed_at);

/* Large-engine FP64 checkpoints (default 1; 0 = off, a rollback restarts from s_0): the state
 * is copied on the device every `blocks` blocks of 200 squarings.  When the round-off exceeds
 * 0.4, the last checkpoint reached within 0.4 is normalised on the device, re-split exactly at
 * the next length, and the test resumes from its iteration.  No reference counterpart
 * (lucaslehmer.cpp:296 aborts the whole FP64 test). */
int llh_set_checkpoint_blocks(int blocks);
int llh_get_checkpoint_blocks(void);

/* Batch-engine FP64 checkpoints (default off): each exponent's digits are written at every
 * round-off check (64 squarings), and an exponent that rolls back resumes at the next length
 * from its last check within 0.4, re-split on the device.  Off, it restarts from s_0.  The
 * checkpointing kernels run about 2% slower, and at the default bits-per-digit cap no batched
 * exponent rolls back.  llh_batch_noalloc runs without them.  No reference counterpart. */
int llh_set_batch_checkpoints(int on);
int llh_get_batch_checkpoints(void);

/* Host run of the per-digit arithmetic of the device re-layout (the checkpoint rollback's
 * N -> N' step): `from` holds the n_from canonical digits of X in [0, 2^p) (digit d has the
 * c_{d+1} - c_d bits of X from c_d = ceil(d p / n_from)); `to` receives n_to near-balanced digits
 * (|digit| <= 2^(b-1)) of X mod 2^p - 1.  No device call.  No reference counterpart: it exists so
 * the re-layout can be checked on the CPU against the oracle (orc_limbs_to_digits). */
int llh_relayout_host(uint64_t p, uint64_t n_from, const int64_t *from, uint64_t n_to, int64_t *to);

/* One exponent at an explicit transform length n_force (one of llh_supported_lengths; 0 = the
 * engine's length rule).  flags bit 0: no rollback (an FP64 round-off above 0.4 ends the run with
 * LLH_ST_ROUNDOFF instead of restarting at 2N).  seconds = device time.  No reference
 * counterpart: the reference fixes N by its size rules (lucaslehmer.cpp:247, cuda/ll_ntt.cu:176);
 * this is the hook for the learned length choice and for round-off measurements. */
int llh_lucas_lehmer_n(uint64_t p, uint64_t max_iters, int transform, uint64_t n_force, int flags,
                       int32_t *verdict, uint64_t *res64, double *maxerr, uint32_t *n_used, int32_t *status,
                       double *seconds);

/* Transform lengths (digit counts N) the engines run for `transform`, ascending: the batched
 * engine's powers of two and the large engine's m * 2^k (FP64: m in {1, 3, 5, 7, 15}).
 * Writes up to cap values, returns the count.  Counterpart of the reference's length ladder
 * {1, 3, 5} * 2^k (lucaslehmer.cpp:250-252). */
size_t llh_supported_lengths(int transform, uint64_t *out, size_t cap);

/* Schedules of the batched engine at length n (digit count): the radix list of each, 4 entries
 * per schedule (R1..R4 of a power-of-two length; 8, the middle radices, 8 of an odd-factor one;
 * unused entries 0), index 0 the length's base schedule.  Writes up to cap schedules, returns the
 * count (0: not a batched length) or a negative status.  llh_set_batch_schedule runs length n at
 * schedule `sched` from now on (-1: the measured choice compiled into the library);
 * llh_get_batch_schedule returns the index in force.  No reference counterpart: the reference
 * fixes its radix-2 schedule and 256-thread blocks (cuda/ll_ntt.cu:99-127, :224). */
int llh_batch_schedules(int transform, uint64_t n, int32_t *radices, int cap);
int llh_set_batch_schedule(int transform, uint64_t n, int sched);
int llh_get_batch_schedule(int transform, uint64_t n);

/* Schedules of the large engine at length n (digit count): the H x W split of its transform
 * elements and the column-pass block width CW.  Writes up to cap entries into w, h, cw (any may be
 * null) and returns the count (0: not a large-engine length), or a negative status.  No reference
 * counterpart: the reference fixes its one schedule (cuda/ll_ntt.cu:224, 256-thread blocks). */
int llh_large_schedules(int transform, uint64_t n, int32_t *w, int32_t *h, int32_t *cw, int cap);

/* Runs length n at schedule `sched` (an index of llh_large_schedules) from now on; -1 returns it
 * to the measured choice compiled into the library (the autotuner's table; the list's first
 * schedule where the table has no entry).  llh_get_large_schedule returns the index in force.
 * No reference counterpart (see llh_large_schedules). */
int llh_set_large_schedule(int transform, uint64_t n, int sched);
int llh_get_large_schedule(int transform, uint64_t n);

/* Per-iteration maximum round-off of the last large-engine FP64 run (one value per squaring);
 * writes up to cap values and returns how many there are.  The reference only keeps the
 * maximum over the run (lucaslehmer.cpp:289-295). */
size_t llh_last_roundoff_trace(double *out, size_t cap);

/* Device-to-device copy of `bytes` bytes, `repeats` times; *ms = time per copy.  No reference
 * counterpart: the known-size transfer the HBM-traffic counters are calibrated against. */
int llh_device_copy(uint64_t bytes, int repeats, double *ms);

/* ---- round-off model: MLP 3 -> 64 -> 64 -> 1 (FP32, tanh), trained on the GPU ----------------
 * No reference counterpart (the reference fixes N with a bits-per-word cap, lucaslehmer.cpp:246).
 * Inputs are rows of 3 floats (the caller's normalised log2 p, log2 N, bits per digit), targets
 * the normalised log10 of the worst round-off.  Parameters: W1[64][3], b1[64], W2[64][64],
 * b2[64], w3[64], b3 (llh_mlp_param_count floats). */
int llh_mlp_param_count(void);
/* gradient of the mean squared error over n samples (and the loss) at `params` */
int llh_mlp_grad(const float *params, const float *X, const float *y, size_t n, float *grad, float *loss);
/* full-batch Adam for `epochs` steps from `params` (updated in place); loss_hist[epochs] optional */
int llh_mlp_train(float *params, const float *X, const float *y, size_t n, int epochs, float lr, float *loss_hist);
/* predictions for n inputs */
int llh_mlp_predict(const float *params, const float *X, size_t n, float *y_out);

/* One exponent.  transform = LLH_NTT replaces lucas_lehmer_n lucaslehmer.cpp:29-56 (the exact
 * engine; Res64 of gmp_res64_after :364-365); LLH_FP64 replaces lucas_lehmer_fft_raw
 * lucaslehmer.cpp:256-329 (maxerr of fft_info :337-342 / fft_res64_after :361-363).
 * status as in llh_batch. */
int llh_lucas_lehmer(uint64_t p, uint64_t max_iters, int transform, int32_t *verdict,
                     uint64_t *res64, double *maxerr, uint32_t *n_used, int32_t *status);

#ifdef __cplusplus
}
#endif

#endif /* LLHIP_H */
