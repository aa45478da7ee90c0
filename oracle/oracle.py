"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module, as the checker or as the timed CPU baseline.  The product
(the HIP library and the ``lucaslehmer`` module) never imports it.

Every function restates a reference function; the C source cites file:line for each.
Small-p pure-Python restatements (``py_ll``) sit here too, for brute-force pinning.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile (gcc + /opt/conda GMP)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P64 = ctypes.POINTER(_u64)
        PI64 = ctypes.POINTER(_i64)
        L.orc_gmp_version.restype = ctypes.c_char_p
        L.orc_ll_gmp.argtypes = [_u64, _u64, P64, ctypes.POINTER(ctypes.c_int)]
        L.orc_ll_gmp.restype = ctypes.c_int
        L.orc_square_step.argtypes = [_u64, P64, P64]
        L.orc_ll_state.argtypes = [_u64, _u64, P64]
        L.orc_primes.argtypes = [_u64, _u64, P64, ctypes.c_size_t]
        L.orc_primes.restype = ctypes.c_size_t
        L.orc_fft_length.argtypes = [_u64, ctypes.c_double]
        L.orc_fft_length.restype = _u64
        L.orc_ntt_length.argtypes = [_u64]
        L.orc_ntt_length.restype = _u64
        L.orc_ntt_max_bits.argtypes = [_u64]
        L.orc_ntt_max_bits.restype = ctypes.c_int
        for f in ("orc_gl_mul", "orc_gl_add", "orc_gl_sub", "orc_gl_pow"):
            getattr(L, f).argtypes = [_u64, _u64]
            getattr(L, f).restype = _u64
        L.orc_gl_inv.argtypes = [_u64]
        L.orc_gl_inv.restype = _u64
        L.orc_gl_root_unity.argtypes = [_u64]
        L.orc_gl_root_unity.restype = _u64
        L.orc_gl_root2.argtypes = [_u64]
        L.orc_gl_root2.restype = _u64
        L.orc_digit_layout.argtypes = [_u64, _u64, ctypes.POINTER(ctypes.c_int), P64]
        L.orc_digits_res64.argtypes = [_u64, _u64, PI64, P64]
        L.orc_digits_res64.restype = ctypes.c_int
        L.orc_limbs_to_digits.argtypes = [_u64, _u64, P64, PI64]
        L.orc_ntt_safe.argtypes = [_u64, _u64]
        L.orc_ntt_safe.restype = ctypes.c_int
        L.orc_ll_ntt.argtypes = [_u64, _u64, _u64, P64, PI64]
        L.orc_ll_ntt.restype = ctypes.c_int
        L.orc_ll_fp64.argtypes = [_u64, _u64, _u64, ctypes.c_double, P64,
                                  ctypes.POINTER(ctypes.c_double), PI64]
        L.orc_ll_fp64.restype = ctypes.c_int
        L.orc_ll_many.argtypes = [P64, ctypes.c_size_t, _u64, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int), P64]
        L.orc_ll_many.restype = ctypes.c_int
        L.orc_ll_gmp_split.argtypes = [_u64, _u64, ctypes.c_int, P64]
        L.orc_ll_gmp_split.restype = ctypes.c_int
        L.orc_bench_square.argtypes = [_u64, ctypes.c_int, ctypes.c_int]
        L.orc_bench_square.restype = ctypes.c_double
        _lib = L
    return _lib


def gmp_version() -> str:
    return lib().orc_gmp_version().decode()


def ll_gmp(p: int, iters: int = 0):
    """(verdict, res64, degenerate) after min(p-2, iters) squarings (iters=0: full test)."""
    r = _u64(0)
    d = ctypes.c_int(0)
    v = lib().orc_ll_gmp(p, iters, ctypes.byref(r), ctypes.byref(d))
    return bool(v), r.value, bool(d.value)


def ll_state(p: int, iters: int) -> int:
    """s_iters as a Python int (canonical, in [0, 2^p-1))."""
    nl = (p + 63) // 64
    buf = (_u64 * nl)()
    lib().orc_ll_state(p, iters, buf)
    return sum(int(buf[i]) << (64 * i) for i in range(nl))


def square_step(p: int, x: int) -> int:
    """(x*x - 2) mod (2^p - 1) through GMP."""
    nl = (p + 63) // 64
    inb = (_u64 * nl)(*[(x >> (64 * i)) & (2**64 - 1) for i in range(nl)])
    outb = (_u64 * nl)()
    lib().orc_square_step(p, inb, outb)
    return sum(int(outb[i]) << (64 * i) for i in range(nl))


def primes(lo: int, hi: int) -> list[int]:
    n = lib().orc_primes(lo, hi, None, 0)
    buf = (_u64 * max(n, 1))()
    lib().orc_primes(lo, hi, buf, n)
    return [int(buf[i]) for i in range(n)]


def fft_length(p: int, bpw: float = 18.0) -> int:
    return int(lib().orc_fft_length(p, bpw))


def ntt_length(p: int) -> int:
    return int(lib().orc_ntt_length(p))


def ntt_safe(p: int, n: int) -> bool:
    return bool(lib().orc_ntt_safe(p, n))


def gl_mul(a, b):
    return int(lib().orc_gl_mul(a, b))


def gl_root2(n):
    return int(lib().orc_gl_root2(n))


def digit_layout(p: int, n: int):
    bits = (ctypes.c_int * n)()
    we = (_u64 * n)()
    lib().orc_digit_layout(p, n, bits, we)
    return list(bits), [int(v) for v in we]


def digits_res64(p: int, digits) -> tuple[bool, int]:
    n = len(digits)
    arr = (_i64 * n)(*digits)
    r = _u64(0)
    z = lib().orc_digits_res64(p, n, arr, ctypes.byref(r))
    return bool(z), r.value


def limbs_to_digits(p: int, n: int, value: int) -> list[int]:
    nl = (p + 63) // 64
    inb = (_u64 * nl)(*[(value >> (64 * i)) & (2**64 - 1) for i in range(nl)])
    out = (_i64 * n)()
    lib().orc_limbs_to_digits(p, n, inb, out)
    return list(out)


def ll_ntt(p: int, n: int, iters: int = 0, want_digits: bool = False):
    """(verdict, res64[, digits]) of the host NTT IBDWT restatement at length n."""
    r = _u64(0)
    dig = (_i64 * n)() if want_digits else None
    v = lib().orc_ll_ntt(p, n, iters, ctypes.byref(r), dig)
    if want_digits:
        return v, r.value, list(dig)
    return v, r.value


def ll_fp64(p: int, n: int, iters: int = 0, abort_err: float = 0.0, want_digits: bool = False):
    """(verdict or -1, res64, maxerr[, digits]) of the host FP64 IBDWT restatement."""
    r = _u64(0)
    e = ctypes.c_double(0)
    dig = (_i64 * n)() if want_digits else None
    v = lib().orc_ll_fp64(p, n, iters, abort_err, ctypes.byref(r), ctypes.byref(e), dig)
    if want_digits:
        return v, r.value, e.value, list(dig)
    return v, r.value, e.value


def ll_many(ps, iters: int = 0, engine: int = 0, threads: int = 1, bpw: float = 18.0):
    """[(verdict, res64)] for many exponents, one exponent per thread (engine 0 GMP,
    1 FP64 at the conservative length, 2 NTT)."""
    n = len(ps)
    arr = (_u64 * n)(*ps)
    ver = (ctypes.c_int * n)()
    res = (_u64 * n)()
    lib().orc_ll_many(arr, n, iters, engine, bpw, threads, ver, res)
    return [(bool(ver[i]), int(res[i])) for i in range(n)]


def ll_gmp_split(p: int, iters: int = 0, depth: int = 1) -> tuple[bool, int]:
    """(verdict, res64) with split-Karatsuba squarings on 3 (depth 1) or 9 (depth 2) threads
    (lucaslehmer.cpp:83-179)."""
    r = _u64(0)
    v = lib().orc_ll_gmp_split(p, iters, depth, ctypes.byref(r))
    return bool(v), int(r.value)


def bench_square(bits: int, iters: int = 20, depth: int = 0) -> float:
    """Seconds per GMP squaring of a random `bits`-bit number, plain or split
    (lucaslehmer.cpp:181-205)."""
    return float(lib().orc_bench_square(bits, iters, depth))


def py_ll(p: int, iters: int = 0) -> tuple[bool, int]:
    """Pure-Python Lucas-Lehmer with exact big ints (small p only): (s == 0, s mod 2^64)."""
    if p < 2:
        return False, 0
    if p == 2:
        return True, 0
    if p % 2 == 0:
        return False, 0
    M = (1 << p) - 1
    s = 4
    k = p - 2 if not iters else min(p - 2, iters)
    for _ in range(k):
        s = (s * s - 2) % M
    return s == 0, s & (2**64 - 1)
