/*
 * ll_oracle.c -- CPU oracle for the Lucas-Lehmer squaring loop mod 2^p - 1.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the timed
 * CPU baseline).  The product path (the HIP library and the lucaslehmer module) never
 * links, loads or calls it.
 *
 * What it restates (plain C + GMP 6.x, no FFTW: FFTW is absent from this image):
 *   orc_ll_gmp            lucaslehmer.cpp:29-56  (GMP square, -2, shift-and-add mod 2^p-1),
 *                         Res64 of lucaslehmer.cpp:19-24, made canonical in [0, M_p).
 *   orc_primes            lucaslehmer.cpp:71-80  (sieve of Eratosthenes on [lo, hi]).
 *   orc_fft_length        lucaslehmer.cpp:244-254 (smallest {1,3,5}*16*2^k with p/N <= bpw).
 *   orc_ntt_length        cuda/ll_ntt.cu:172-180  (N*2^(2B-1) < 2^62 exactness rule).
 *   orc_gl_*              cuda/ll_ntt.cu:38-44, 182-195 (Goldilocks field, tables, root of 2).
 *   orc_ll_ntt            cuda/ll_ntt.cu:74-167, 198-251 (exact NTT IBDWT, host restatement).
 *   orc_ll_fp64           lucaslehmer.cpp:256-329 (FP64 IBDWT; own mixed-radix complex FFT
 *                         in place of FFTW's r2c/c2r; same weights, rounding, error, carry).
 *   orc_digits_res64      lucaslehmer.cpp:312-323 / cuda/ll_ntt.cu:238-251 (exact GMP rebuild).
 *   orc_ll_gmp_many       PerfNumMultiCLL.py:189-193 model: one exponent per thread.
 *   orc_ll_gmp_split      lucaslehmer.cpp:83-179 (split-Karatsuba squaring on 3 or 9 threads).
 *   orc_bench_square      lucaslehmer.cpp:181-205 (seconds per squaring, plain or split).
 *
 * Pinning (what checks this oracle, other than retyping it): the reference's known-answer
 * tests (tests/test_lucaslehmer.py), known_mp_list.txt verdicts, the README.md:200-206
 * interim residue (p = 3,021,377, 20,000 squarings -> ce811e3772129824), Python big-int
 * brute force on small p, and cross-agreement of the three independent engines here
 * (GMP, NTT, FP64) on every tested exponent.
 */
#define _GNU_SOURCE
#include <gmp.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <time.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <complex.h>

typedef uint64_t u64;
typedef int64_t i64;
typedef unsigned __int128 u128;

/* ----------------------------------------------------------------------------------------
 * Canonical Res64 of an mpz in any sign: low 64 bits of (s mod M_p) in [0, M_p).
 * lucaslehmer.cpp:24 takes mpz_get_ui (the low limb of |s|); the two agree whenever s is
 * already canonical, which is every case the reference reaches except the degenerate
 * negative intermediate (s_k in {0,1} before squaring), which this flags.
 * -------------------------------------------------------------------------------------- */
static u64 low64_of(const mpz_t s) {
    /* s >= 0 here */
    u64 lo = 0;
    size_t n = mpz_size(s);
    if (n >= 1) lo = (u64)mpz_getlimbn(s, 0);
    return lo;
}

const char *orc_gmp_version(void) { return gmp_version; }

/* lucaslehmer.cpp:29-56.  Returns 1 prime / 0 composite.  res64 out = canonical Res64.
 * *degenerate set to 1 if the reference's truncating reduction would have left a negative
 * intermediate (then mpz_get_ui would report |s|; the oracle reports the canonical value). */
int orc_ll_gmp(u64 p, u64 max_iters, u64 *res64, int *degenerate) {
    if (degenerate) *degenerate = 0;
    if (res64) *res64 = 0;
    if (p < 2) return 0;
    if (p == 2) return 1;
    if (p % 2 == 0) return 0;
    mpz_t M, s, hi;
    mpz_init(M); mpz_init(s); mpz_init(hi);
    mpz_set_ui(M, 1); mpz_mul_2exp(M, M, p); mpz_sub_ui(M, M, 1);
    mpz_set_ui(s, 4);
    for (u64 i = 0; i + 2 < p && (!max_iters || i < max_iters); ++i) {
        mpz_mul(s, s, s);
        mpz_sub_ui(s, s, 2);
        if (mpz_sgn(s) < 0) {               /* s was 0 or 1: reference leaves it negative */
            if (degenerate) *degenerate = 1;
            mpz_add(s, s, M);
            continue;
        }
        mpz_tdiv_q_2exp(hi, s, p);
        mpz_tdiv_r_2exp(s, s, p);
        mpz_add(s, s, hi);
        if (mpz_cmp(s, M) >= 0) mpz_sub(s, s, M);
    }
    u64 r = low64_of(s);
    int prime = (mpz_cmp_ui(s, 0) == 0);
    if (res64) *res64 = r;
    mpz_clear(M); mpz_clear(s); mpz_clear(hi);
    return prime;
}

/* One step (x*x - 2) mod M_p on little-endian u64 limb arrays of nlimbs = ceil(p/64).
 * Used by single-squaring unit tests of the device transforms. */
void orc_square_step(u64 p, const u64 *in, u64 *out) {
    size_t nl = (size_t)((p + 63) / 64);
    mpz_t M, s, hi;
    mpz_init(M); mpz_init(s); mpz_init(hi);
    mpz_set_ui(M, 1); mpz_mul_2exp(M, M, p); mpz_sub_ui(M, M, 1);
    mpz_import(s, nl, -1, 8, 0, 0, in);
    mpz_mod(s, s, M);
    mpz_mul(s, s, s);
    mpz_sub_ui(s, s, 2);
    mpz_mod(s, s, M);
    memset(out, 0, nl * 8);
    size_t cnt = 0;
    mpz_export(out, &cnt, -1, 8, 0, 0, s);
    mpz_clear(M); mpz_clear(s); mpz_clear(hi);
}

/* s_k as little-endian u64 limbs (ceil(p/64) of them), canonical in [0, M_p). */
void orc_ll_state(u64 p, u64 iters, u64 *out) {
    size_t nl = (size_t)((p + 63) / 64);
    mpz_t M, s;
    mpz_init(M); mpz_init(s);
    mpz_set_ui(M, 1); mpz_mul_2exp(M, M, p); mpz_sub_ui(M, M, 1);
    mpz_set_ui(s, 4);
    for (u64 i = 0; i < iters; ++i) {
        mpz_mul(s, s, s);
        mpz_sub_ui(s, s, 2);
        mpz_mod(s, s, M);
    }
    memset(out, 0, nl * 8);
    size_t cnt = 0;
    mpz_export(out, &cnt, -1, 8, 0, 0, s);
    mpz_clear(M); mpz_clear(s);
}

/* lucaslehmer.cpp:71-80.  Writes up to cap primes in [max(lo,2), hi]; returns the count
 * (the full count even when it exceeds cap). */
size_t orc_primes(u64 lo, u64 hi, u64 *out, size_t cap) {
    if (hi < 2) return 0;
    unsigned char *comp = calloc(hi + 1, 1);
    for (u64 i = 2; i * i <= hi; ++i)
        if (!comp[i]) for (u64 j = i * i; j <= hi; j += i) comp[j] = 1;
    size_t n = 0;
    for (u64 i = (lo < 2 ? 2 : lo); i <= hi; ++i)
        if (!comp[i]) { if (n < cap && out) out[n] = i; ++n; }
    free(comp);
    return n;
}

/* lucaslehmer.cpp:244-254: smallest N in {1,3,5}*16*2^k with p/N <= max_bpw. */
u64 orc_fft_length(u64 p, double max_bpw) {
    u64 best = 0;
    const u64 ms[3] = {1, 3, 5};
    for (int i = 0; i < 3; ++i)
        for (u64 N = ms[i] * 16; N < ((u64)1 << 40); N <<= 1)
            if ((double)p / (double)N <= max_bpw) { if (!best || N < best) best = N; break; }
    return best;
}

/* cuda/ll_ntt.cu:172-180. */
int orc_ntt_max_bits(u64 N) {
    int l = 0; while (((u64)1 << l) < N) ++l;
    return (63 - l) / 2;
}
u64 orc_ntt_length(u64 p) {
    for (u64 N = 64; N <= 8192; N <<= 1)
        if ((p + N - 1) / N <= (u64)orc_ntt_max_bits(N)) return N;
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * Goldilocks field GF(q), q = 2^64 - 2^32 + 1 (cuda/ll_ntt.cu:38-44).  Plain 128-bit
 * remainder; this is the textbook definition the device's 32-bit-limb reduction must match.
 * ------------------------------------------------------------------------------------- */
#define GL_Q 0xFFFFFFFF00000001ull
u64 orc_gl_mul(u64 a, u64 b) { return (u64)(((u128)a * b) % GL_Q); }
u64 orc_gl_add(u64 a, u64 b) { u128 s = (u128)a + b; return (u64)(s % GL_Q); }
u64 orc_gl_sub(u64 a, u64 b) { return a >= b ? (a - b) % GL_Q : (u64)(((u128)a + GL_Q - b) % GL_Q); }
u64 orc_gl_pow(u64 a, u64 e) { u64 r = 1; a %= GL_Q; while (e) { if (e & 1) r = orc_gl_mul(r, a); a = orc_gl_mul(a, a); e >>= 1; } return r; }
u64 orc_gl_inv(u64 a) { return orc_gl_pow(a, GL_Q - 2); }
/* Primitive n-th root of unity from generator 7 (cuda/ll_ntt.cu:184-185). */
u64 orc_gl_root_unity(u64 n) { return orc_gl_pow(7, (GL_Q - 1) / n); }
/* N-th root of 2 (cuda/ll_ntt.cu:188-192): om = 7^((q-1)/(192N)), 2 = (om^N)^s, w = om^s.
 * Returns 0 if none exists (q-1 not divisible by 192N). */
u64 orc_gl_root2(u64 N) {
    if ((GL_Q - 1) % (192 * N)) return 0;
    u64 om = orc_gl_pow(7, (GL_Q - 1) / (192 * N));
    u64 r192 = orc_gl_pow(om, N), pw = 1, s = 0;
    for (; s < 192; ++s) { if (pw == 2) break; pw = orc_gl_mul(pw, r192); }
    if (s == 192) return 0;
    return orc_gl_pow(om, s);
}

/* Digit layout shared by both transforms: c_j = ceil(j p / N), b_j = c_{j+1} - c_j. */
static inline u64 ceil_jp(u64 j, u64 p, u64 N) { return (u64)(((u128)j * p + N - 1) / N); }
void orc_digit_layout(u64 p, u64 N, int *bits, u64 *wexp) {
    for (u64 j = 0; j < N; ++j) {
        u64 lo = ceil_jp(j, p, N), hi = ceil_jp(j + 1, p, N);
        if (bits) bits[j] = (int)(hi - lo);
        if (wexp) wexp[j] = lo * N - j * p;      /* weight exponent in [0, N): a_j = 2^(wexp/N) */
    }
}

/* Exact rebuild lucaslehmer.cpp:312-323: s = sum x_j 2^c_j mod M_p; canonical Res64. */
int orc_digits_res64(u64 p, u64 N, const i64 *x, u64 *res64) {
    mpz_t M, s, t;
    mpz_init(M); mpz_init(s); mpz_init(t);
    mpz_set_ui(M, 1); mpz_mul_2exp(M, M, p); mpz_sub_ui(M, M, 1);
    for (u64 j = 0; j < N; ++j) {
        if (!x[j]) continue;
        mpz_set_si(t, (long)x[j]);
        mpz_mul_2exp(t, t, ceil_jp(j, p, N));
        mpz_add(s, s, t);
    }
    mpz_mod(s, s, M);
    if (res64) *res64 = low64_of(s);
    int z = mpz_cmp_ui(s, 0) == 0;
    mpz_clear(M); mpz_clear(s); mpz_clear(t);
    return z;
}

/* Canonical residue s (limbs) -> balanced digits x_j at length N (exact re-layout). */
void orc_limbs_to_digits(u64 p, u64 N, const u64 *limbs, i64 *x) {
    size_t nl = (size_t)((p + 63) / 64);
    mpz_t s, t; mpz_init(s); mpz_init(t);
    mpz_import(s, nl, -1, 8, 0, 0, limbs);
    i64 carry = 0;
    for (u64 j = 0; j < N; ++j) {
        u64 lo = ceil_jp(j, p, N), hi = ceil_jp(j + 1, p, N);
        int b = (int)(hi - lo);
        mpz_tdiv_q_2exp(t, s, lo);
        mpz_tdiv_r_2exp(t, t, b);
        i64 v = (i64)mpz_get_ui(t) + carry;
        carry = (v + ((i64)1 << (b - 1))) >> b;
        x[j] = v - (carry << b);
    }
    x[0] += carry;                                 /* 2^p == 1 */
    mpz_clear(s); mpz_clear(t);
}

/* Balanced carry with the wrap (lucaslehmer.cpp:297-310), shared by both CPU transforms. */
static void carry_balanced(u64 N, const int *bits, i64 *x, i64 carry) {
    for (u64 j = 0; j < N; ++j) {
        i64 v = x[j] + carry; int b = bits[j];
        carry = (v + ((i64)1 << (b - 1))) >> b;
        x[j] = v - (carry << b);
    }
    while (carry) {
        for (u64 j = 0; j < N && carry; ++j) {
            i64 v = x[j] + carry; int b = bits[j];
            carry = (v + ((i64)1 << (b - 1))) >> b;
            x[j] = v - (carry << b);
        }
    }
}

/* ---------------------------------------------------------------------------------------
 * Exact Goldilocks NTT IBDWT (cuda/ll_ntt.cu:74-167 semantics, on the host):
 * weight -> DIF NTT -> square -> DIT inverse NTT -> unweight (with 1/N) -> centre-lift ->
 * -2 and balanced carry with wrap.  Any power-of-two N whose bits/word keep the
 * convolution below q/2 (the caller checks; orc_ntt_safe tells).  Returns verdict of the
 * canonical residue after min(p-2, iters) squarings; digits out optional.
 * ------------------------------------------------------------------------------------- */
int orc_ntt_safe(u64 p, u64 N) {
    /* max digit magnitude 2^(B-1), B = ceil(p/N); convolution |.| <= N 2^(2B-2) < q/2 */
    u64 B = (p + N - 1) / N;
    int l = 0; while (((u64)1 << l) < N) ++l;
    return (2 * B - 2 + (u64)l) < 63;
}

int orc_ll_ntt(u64 p, u64 N, u64 iters, u64 *res64, i64 *digits_out) {
    if (res64) *res64 = 0;
    if (p < 2) return 0;
    if (p == 2) return 1;
    if (p % 2 == 0) return 0;
    if (N & (N - 1)) return -2;
    u64 root2 = orc_gl_root2(N);
    if (!root2) return -2;
    int *bits = malloc(N * sizeof(int)); u64 *we = malloc(N * 8);
    orc_digit_layout(p, N, bits, we);
    for (u64 j = 0; j < N; ++j) if (bits[j] < 1) { free(bits); free(we); return -2; }
    u64 *w = malloc(N * 8), *wi = malloc(N * 8), *tw = malloc((N / 2 + 1) * 8), *a = malloc(N * 8);
    i64 *x = calloc(N, 8);
    u64 invN = orc_gl_inv(N % GL_Q);
    for (u64 j = 0; j < N; ++j) {
        w[j] = orc_gl_pow(root2, we[j]);
        wi[j] = orc_gl_mul(orc_gl_inv(w[j]), invN);
    }
    u64 om = orc_gl_root_unity(N), t = 1;
    for (u64 k = 0; k <= N / 2; ++k) { tw[k] = t; t = orc_gl_mul(t, om); }
    x[0] = 4;
    u64 half = N / 2;
    for (u64 it = 0; it + 2 < p && (!iters || it < iters); ++it) {
        for (u64 j = 0; j < N; ++j) {
            u64 f = x[j] < 0 ? GL_Q - (u64)(-x[j]) : (u64)x[j];
            a[j] = orc_gl_mul(f, w[j]);
        }
        for (u64 len = half; len >= 1; len >>= 1) {          /* DIF */
            u64 step = half / len;
            for (u64 i = 0; i < half; ++i) {
                u64 grp = i / len, jj = i - grp * len;
                u64 ia = grp * 2 * len + jj, ib = ia + len;
                u64 u = a[ia], v = a[ib];
                a[ia] = orc_gl_add(u, v);
                a[ib] = orc_gl_mul(orc_gl_sub(u, v), tw[jj * step]);
            }
        }
        for (u64 j = 0; j < N; ++j) a[j] = orc_gl_mul(a[j], a[j]);
        for (u64 len = 1; len < N; len <<= 1) {                /* DIT with w^-k = -w^(N/2-k) */
            u64 step = half / len;
            for (u64 i = 0; i < half; ++i) {
                u64 grp = i / len, jj = i - grp * len;
                u64 ia = grp * 2 * len + jj, ib = ia + len;
                u64 k = jj * step;
                u64 wk = k ? (GL_Q - tw[half - k]) : 1;
                u64 u = a[ia], v = orc_gl_mul(a[ib], wk);
                a[ia] = orc_gl_add(u, v);
                a[ib] = orc_gl_sub(u, v);
            }
        }
        for (u64 j = 0; j < N; ++j) {
            u64 v = orc_gl_mul(a[j], wi[j]);
            x[j] = (v > (GL_Q >> 1)) ? -(i64)(GL_Q - v) : (i64)v;
        }
        carry_balanced(N, bits, x, -2);
    }
    int z = orc_digits_res64(p, N, x, res64);
    if (digits_out) memcpy(digits_out, x, N * 8);
    free(bits); free(we); free(w); free(wi); free(tw); free(a); free(x);
    return z;
}

/* ---------------------------------------------------------------------------------------
 * FP64 IBDWT (lucaslehmer.cpp:256-329).  FFTW's r2c/c2r are replaced by a mixed-radix
 * complex FFT of the real weighted signal (N = m * 2^k, m odd, naive DFT_m stage), which
 * yields the same spectrum.  Twiddles and weights come from long-double libm calls,
 * rounded to double.  err = max |v - nearbyint(v)| per iteration; returns -1 when an
 * iteration's error exceeds abort_err (> 0), as the reference does at 0.35.
 * ------------------------------------------------------------------------------------- */
typedef double complex cplx;

typedef struct { u64 N, m, K; cplx *tw; cplx *twm; cplx *scratch; } FftPlan;

static cplx cexpi_ld(long double num, long double den, int sign) {
    /* exp(sign * 2*pi*i * num/den) in long double, rounded to double */
    long double a = 2.0L * 3.14159265358979323846264338327950288L * num / den;
    return (double)cosl(a) + I * (double)(sign * sinl(a));
}

static void plan_init(FftPlan *P, u64 N) {
    P->N = N; P->m = N; P->K = 1;
    while (!(P->m & 1)) { P->m >>= 1; P->K <<= 1; }
    P->tw = malloc(N * sizeof(cplx));
    for (u64 j = 0; j < N; ++j) P->tw[j] = cexpi_ld((long double)j, (long double)N, -1);
    P->twm = malloc(P->m * P->m * sizeof(cplx));
    for (u64 a = 0; a < P->m; ++a)
        for (u64 b = 0; b < P->m; ++b) P->twm[a * P->m + b] = cexpi_ld((long double)((a * b) % P->m), (long double)P->m, -1);
    P->scratch = malloc(N * sizeof(cplx));
}
static void plan_free(FftPlan *P) { free(P->tw); free(P->twm); free(P->scratch); }

/* in-place radix-2 DIT FFT of length K on a strided view (stride st), twiddles from the
 * length-N table (W_K^e = W_N^(e*N/K)).  sign < 0 forward, > 0 inverse (conjugate). */
static void fft2_strided(cplx *v, u64 K, u64 st, const FftPlan *P, int inverse) {
    /* bit reversal */
    for (u64 i = 1, j = 0; i < K; ++i) {
        u64 bit = K >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = v[i * st]; v[i * st] = v[j * st]; v[j * st] = t; }
    }
    for (u64 len = 2; len <= K; len <<= 1) {
        u64 tstep = P->N / len;
        for (u64 i = 0; i < K; i += len)
            for (u64 k = 0; k < len / 2; ++k) {
                cplx w = P->tw[k * tstep];
                if (inverse) w = conj(w);
                cplx u = v[(i + k) * st], t = v[(i + k + len / 2) * st] * w;
                v[(i + k) * st] = u + t;
                v[(i + k + len / 2) * st] = u - t;
            }
    }
}

/* X[k] = sum_n x[n] W_N^{nk} (forward) or conj twiddles (inverse, unscaled). */
static void fft_mixed(cplx *x, const FftPlan *P, int inverse) {
    u64 N = P->N, m = P->m, K = P->K;
    /* n = m*n2 + n1 ; k = k1 + K*k2 */
    for (u64 n1 = 0; n1 < m; ++n1) fft2_strided(x + n1, K, m, P, inverse);   /* A[n1][k1] at x[m*k1+n1] */
    for (u64 k1 = 0; k1 < K; ++k1)
        for (u64 n1 = 0; n1 < m; ++n1) {
            cplx w = P->tw[(n1 * k1) % N];
            if (inverse) w = conj(w);
            x[m * k1 + n1] *= w;
        }
    cplx *out = P->scratch;
    for (u64 k1 = 0; k1 < K; ++k1)
        for (u64 k2 = 0; k2 < m; ++k2) {
            cplx acc = 0;
            for (u64 n1 = 0; n1 < m; ++n1) {
                cplx w = P->twm[n1 * m + k2];
                if (inverse) w = conj(w);
                acc += x[m * k1 + n1] * w;
            }
            out[k1 + K * k2] = acc;
        }
    memcpy(x, out, N * sizeof(cplx));
}

int orc_ll_fp64(u64 p, u64 N, u64 iters, double abort_err, u64 *res64, double *maxerr_out, i64 *digits_out) {
    if (res64) *res64 = 0;
    if (maxerr_out) *maxerr_out = 0;
    if (p < 2) return 0;
    if (p == 2) return 1;
    if (p % 2 == 0) return 0;
    int *bits = malloc(N * sizeof(int)); u64 *we = malloc(N * 8);
    orc_digit_layout(p, N, bits, we);
    for (u64 j = 0; j < N; ++j) if (bits[j] < 1) { free(bits); free(we); return -2; }
    double *w = malloc(N * 8), *winv = malloc(N * 8);
    for (u64 j = 0; j < N; ++j) {
        long double e = exp2l((long double)we[j] / (long double)N);
        w[j] = (double)e;
        winv[j] = (double)(1.0L / (e * (long double)N));
    }
    FftPlan P; plan_init(&P, N);
    cplx *y = malloc(N * sizeof(cplx));
    i64 *x = calloc(N, 8);
    x[0] = 4;
    double maxerr = 0; int verdict = 0;
    for (u64 it = 0; it + 2 < p && (!iters || it < iters); ++it) {
        for (u64 j = 0; j < N; ++j) y[j] = (double)x[j] * w[j];
        fft_mixed(y, &P, 0);
        for (u64 k = 0; k < N; ++k) y[k] = y[k] * y[k];
        fft_mixed(y, &P, 1);
        double err = 0;
        for (u64 j = 0; j < N; ++j) {
            double v = creal(y[j]) * winv[j], r = nearbyint(v);
            double e = fabs(v - r); if (e > err) err = e;
            x[j] = (i64)r;
        }
        if (err > maxerr) maxerr = err;
        if (abort_err > 0 && err > abort_err) { verdict = -1; goto done; }
        carry_balanced(N, bits, x, -2);
    }
    verdict = orc_digits_res64(p, N, x, res64);
    if (digits_out) memcpy(digits_out, x, N * 8);
done:
    if (maxerr_out) *maxerr_out = maxerr;
    plan_free(&P); free(bits); free(we); free(w); free(winv); free(y); free(x);
    return verdict;
}

/* ---------------------------------------------------------------------------------------
 * Many exponents, one per thread (PerfNumMultiCLL.py:189-193).  engine 0 = GMP, 1 = FP64
 * at the conservative length (lucaslehmer.cpp:331-335 GMP redo on -1), 2 = NTT at
 * orc_ntt_length (0 -> GMP).
 * ------------------------------------------------------------------------------------- */
typedef struct {
    const u64 *ps; size_t n; u64 iters; int engine; double bpw;
    int *verdict; u64 *res; volatile size_t next; pthread_mutex_t mx;
} ManyJob;

static void *many_worker(void *arg) {
    ManyJob *J = arg;
    for (;;) {
        pthread_mutex_lock(&J->mx);
        size_t i = J->next++;
        pthread_mutex_unlock(&J->mx);
        if (i >= J->n) break;
        u64 p = J->ps[i], r = 0; int v;
        if (J->engine == 1 && p >= 512) {
            v = orc_ll_fp64(p, orc_fft_length(p, J->bpw), J->iters, 0.35, &r, NULL, NULL);
            if (v < 0) v = orc_ll_gmp(p, J->iters, &r, NULL);
        } else if (J->engine == 2 && p >= 64 && orc_ntt_length(p)) {
            v = orc_ll_ntt(p, orc_ntt_length(p), J->iters, &r, NULL);
        } else {
            v = orc_ll_gmp(p, J->iters, &r, NULL);
        }
        J->verdict[i] = v; J->res[i] = r;
    }
    return NULL;
}

int orc_ll_many(const u64 *ps, size_t n, u64 iters, int engine, double bpw, int threads,
                int *verdict, u64 *res) {
    ManyJob J = {ps, n, iters, engine, bpw > 0 ? bpw : 18.0, verdict, res, 0, PTHREAD_MUTEX_INITIALIZER};
    if (threads < 1) threads = 1;
    pthread_t *th = malloc(sizeof(pthread_t) * threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, many_worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    return 0;
}

/* ----------------------------------------------------------------------------------------
 * Split-Karatsuba squaring and bench_square (lucaslehmer.cpp:83-205): one test on several
 * threads.  With s = a 2^k + b (k = half the bit length),
 *   s^2 = a^2 2^2k + ((a+b)^2 - a^2 - b^2) 2^k + b^2,
 * three independent half-size squarings: depth 1 runs them on three threads (two workers and
 * the caller), depth 2 splits each again on that lane's own two workers (nine).  Below 4096
 * bits a squaring is not split (lucaslehmer.cpp:125).  Workers spin on a generation counter,
 * as the reference's Pool (lucaslehmer.cpp:99-124), so a split costs no thread creation.
 * -------------------------------------------------------------------------------------- */
typedef struct SplitPool SplitPool;
typedef struct {
    mpz_ptr r;
    mpz_srcptr x;
    SplitPool *sub;          /* depth 2: this lane's own pool; NULL = plain square */
} SqJob;
typedef struct {
    mpz_t a, b, t1, t2, t3;
} SqScratch;
struct SplitPool {
    pthread_t th[2];
    SqJob job[2];
    SqScratch scr[2];        /* scratch of the lanes the workers run (depth 2) */
    unsigned long gen;       /* accessed with __atomic builtins */
    int pending;
    int stop;
};

static void square_split(mpz_ptr r, mpz_srcptr x, SqScratch *s, SplitPool *pool, SplitPool *lane_sub[3],
                         SqScratch lane_scr[3]);

static void sq_run(mpz_ptr r, mpz_srcptr x, SplitPool *sub, SqScratch *scr) {
    if (sub) square_split(r, x, scr, sub, NULL, NULL);
    else mpz_mul(r, x, x);
}

static void *split_worker(void *arg) {
    void **a = (void **)arg;
    SplitPool *P = (SplitPool *)a[0];
    const int i = (int)(intptr_t)a[1];
    free(arg);
    unsigned long seen = 0;
    for (;;) {
        unsigned spins = 0;
        while (__atomic_load_n(&P->gen, __ATOMIC_ACQUIRE) == seen && !__atomic_load_n(&P->stop, __ATOMIC_RELAXED)) {
            if (++spins < 4096) __builtin_ia32_pause();
            else { sched_yield(); spins = 0; }
        }
        if (__atomic_load_n(&P->stop, __ATOMIC_RELAXED)) return NULL;
        seen = __atomic_load_n(&P->gen, __ATOMIC_ACQUIRE);
        sq_run(P->job[i].r, P->job[i].x, P->job[i].sub, &P->scr[i]);
        __atomic_fetch_sub(&P->pending, 1, __ATOMIC_ACQ_REL);
    }
}

static SplitPool *pool_new(void) {
    SplitPool *P = (SplitPool *)calloc(1, sizeof(SplitPool));
    for (int i = 0; i < 2; ++i) {
        mpz_init(P->scr[i].a); mpz_init(P->scr[i].b); mpz_init(P->scr[i].t1); mpz_init(P->scr[i].t2);
        mpz_init(P->scr[i].t3);
        void **arg = (void **)malloc(2 * sizeof(void *));
        arg[0] = P;
        arg[1] = (void *)(intptr_t)i;
        pthread_create(&P->th[i], NULL, split_worker, arg);
    }
    return P;
}

static void pool_free(SplitPool *P) {
    if (!P) return;
    __atomic_store_n(&P->stop, 1, __ATOMIC_RELEASE);
    for (int i = 0; i < 2; ++i) pthread_join(P->th[i], NULL);
    for (int i = 0; i < 2; ++i) {
        mpz_clear(P->scr[i].a); mpz_clear(P->scr[i].b); mpz_clear(P->scr[i].t1); mpz_clear(P->scr[i].t2);
        mpz_clear(P->scr[i].t3);
    }
    free(P);
}

static void pool_wait(SplitPool *P) {
    unsigned spins = 0;
    while (__atomic_load_n(&P->pending, __ATOMIC_ACQUIRE) != 0) {
        if (++spins < 4096) __builtin_ia32_pause();
        else { sched_yield(); spins = 0; }
    }
}

/* r = x^2; pool != NULL splits on it (its two workers take a^2 and b^2, the caller (a+b)^2).
 * lane_sub / lane_scr (depth 2): the three lanes' own pools and scratch. */
static void square_split(mpz_ptr r, mpz_srcptr x, SqScratch *s, SplitPool *pool, SplitPool *lane_sub[3],
                         SqScratch lane_scr[3]) {
    const size_t L = mpz_sizeinbase(x, 2), k = L / 2;
    if (!pool || L < 4096) { mpz_mul(r, x, x); return; }
    mpz_tdiv_q_2exp(s->a, x, k);
    mpz_tdiv_r_2exp(s->b, x, k);
    mpz_add(s->t3, s->a, s->b);
    pool->job[0] = (SqJob){s->t1, s->a, lane_sub ? lane_sub[1] : NULL};
    pool->job[1] = (SqJob){s->t2, s->b, lane_sub ? lane_sub[2] : NULL};
    __atomic_store_n(&pool->pending, 2, __ATOMIC_RELEASE);
    __atomic_fetch_add(&pool->gen, 1, __ATOMIC_ACQ_REL);
    if (lane_sub) square_split(s->t3, s->t3, &lane_scr[0], lane_sub[0], NULL, NULL);
    else mpz_mul(s->t3, s->t3, s->t3);
    pool_wait(pool);
    mpz_sub(s->t3, s->t3, s->t1);
    mpz_sub(s->t3, s->t3, s->t2);
    mpz_mul_2exp(r, s->t1, 2 * k);
    mpz_mul_2exp(s->t3, s->t3, k);
    mpz_add(r, r, s->t3);
    mpz_add(r, r, s->t2);
}

typedef struct {
    SplitPool *top, *lane[3];
    SqScratch top_scr, lane_scr[3];
} SplitCtx;

static void split_ctx_init(SplitCtx *c, int depth) {
    memset(c, 0, sizeof(*c));
    SqScratch *all[4] = {&c->top_scr, &c->lane_scr[0], &c->lane_scr[1], &c->lane_scr[2]};
    for (int i = 0; i < 4; ++i) {
        mpz_init(all[i]->a); mpz_init(all[i]->b); mpz_init(all[i]->t1); mpz_init(all[i]->t2); mpz_init(all[i]->t3);
    }
    if (depth >= 1) c->top = pool_new();
    if (depth >= 2)
        for (int i = 0; i < 3; ++i) c->lane[i] = pool_new();
}

static void split_ctx_free(SplitCtx *c) {
    pool_free(c->top);
    for (int i = 0; i < 3; ++i) pool_free(c->lane[i]);
    SqScratch *all[4] = {&c->top_scr, &c->lane_scr[0], &c->lane_scr[1], &c->lane_scr[2]};
    for (int i = 0; i < 4; ++i) {
        mpz_clear(all[i]->a); mpz_clear(all[i]->b); mpz_clear(all[i]->t1); mpz_clear(all[i]->t2); mpz_clear(all[i]->t3);
    }
}

static void split_square(SplitCtx *c, mpz_ptr r, mpz_srcptr x) {
    square_split(r, x, &c->top_scr, c->top, c->lane[0] ? c->lane : NULL, c->lane_scr);
}

/* lucaslehmer.cpp:147-179: the LL loop with split squarings (depth 0 = orc_ll_gmp's loop). */
int orc_ll_gmp_split(u64 p, u64 max_iters, int depth, u64 *res64) {
    if (p < 2) { *res64 = 0; return 0; }
    if (p == 2) { *res64 = 0; return 1; }
    if (p % 2 == 0) { *res64 = 0; return 0; }
    SplitCtx c;
    split_ctx_init(&c, depth);
    mpz_t M, s, sq, hi;
    mpz_init(M); mpz_init(s); mpz_init(sq); mpz_init(hi);
    mpz_set_ui(M, 1); mpz_mul_2exp(M, M, p); mpz_sub_ui(M, M, 1);
    mpz_set_ui(s, 4);
    for (u64 i = 0; i + 2 < p && (!max_iters || i < max_iters); ++i) {
        split_square(&c, sq, s);
        mpz_sub_ui(sq, sq, 2);
        if (mpz_sgn(sq) < 0) mpz_add(sq, sq, M);
        mpz_tdiv_q_2exp(hi, sq, p);
        mpz_tdiv_r_2exp(s, sq, p);
        mpz_add(s, s, hi);
        if (mpz_cmp(s, M) >= 0) mpz_sub(s, s, M);
    }
    *res64 = low64_of(s);
    const int prime = mpz_cmp_ui(s, 0) == 0;
    mpz_clear(M); mpz_clear(s); mpz_clear(sq); mpz_clear(hi);
    split_ctx_free(&c);
    return prime;
}

/* lucaslehmer.cpp:181-205: seconds per squaring of a random `bits`-bit number (depth 0 plain,
 * 1 and 2 split), the clock around the squarings only. */
double orc_bench_square(u64 bits, int iters, int depth) {
    SplitCtx c;
    split_ctx_init(&c, depth);
    mpz_t x, r;
    mpz_init(x); mpz_init(r);
    gmp_randstate_t st;
    gmp_randinit_default(st);
    mpz_urandomb(x, st, bits);
    gmp_randclear(st);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < iters; ++i) split_square(&c, r, x);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    mpz_clear(x); mpz_clear(r);
    split_ctx_free(&c);
    return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) / (iters > 0 ? iters : 1);
}
