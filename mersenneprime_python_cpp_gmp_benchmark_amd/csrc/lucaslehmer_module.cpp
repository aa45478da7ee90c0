// lucaslehmer_module.cpp -- drop-in for the reference's pybind11 module `lucaslehmer`
// (lucaslehmer.cpp:344-387): same names, arguments, return types and error behaviour,
// every Lucas-Lehmer loop running on the GPU through the C ABI of include/llhip.h.
//
// Engine mapping:
//   * the reference's exact GMP engine (lucas_lehmer, lucas_lehmer_res, gmp_res64_after,
//     and the split-GMP variants, which give the same results) -> the exact Goldilocks NTT
//     engine on the device (FP64 IBDWT only where the NTT has no length for p; its
//     round-off is checked every iteration and rolled back to 2N above 0.4);
//   * the FP64 IBDWT engine (lucas_lehmer_fft*, fft_res64_after, fft_info) -> the FP64 IBDWT
//     on the device; an exponent whose round-off stays too large is redone on the exact
//     engine, as lucaslehmer.cpp:331-335 redoes it with GMP.
// The GIL is released around every device call (lucaslehmer.cpp:38, :280) and the module is
// flagged free-threading safe (:344).  Res64 of the last call is thread-local (:23).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "llhip.h"

namespace py = pybind11;

static thread_local unsigned long long g_last_res64 = 0;

namespace {

struct Run {
    int32_t verdict = 0, status = 0;
    uint64_t res64 = 0;
    double maxerr = 0, seconds = 0;
    uint32_t n = 0;
};

void check(int rc) {
    if (rc != LLH_OK) throw std::runtime_error(std::string("llhip: ") + llh_last_error());
}

// One exponent.  The FP64 engine runs at the reference's conservative length for the module's
// bits-per-word knob (set_fft_bits_per_word, default 18, lucaslehmer.cpp:244-254), or, where no
// engine has that length, the shortest engine length above it; it still rolls back to a longer
// length if the round-off check fires.
Run run_one(unsigned long p, unsigned long iters, int transform) {
    Run r;
    int rc;
    {
        py::gil_scoped_release release;
        const uint64_t n = transform == LLH_FP64 ? llh_fp64_length(p, 0.0) : 0;
        rc = llh_lucas_lehmer_n(p, iters, transform, n, 0, &r.verdict, &r.res64, &r.maxerr, &r.n, &r.status,
                                &r.seconds);
    }
    check(rc);
    return r;
}

// exact result: NTT engine, else FP64 engine (checked), else error
Run run_exact(unsigned long p, unsigned long iters) {
    Run r = run_one(p, iters, LLH_NTT);
    if (r.status == LLH_ST_NOT_RUN) r = run_one(p, iters, LLH_FP64);
    if (r.status == LLH_ST_NOT_RUN || r.status == LLH_ST_ROUNDOFF)
        throw std::runtime_error("lucaslehmer: exponent " + std::to_string(p) + " is beyond the device engines");
    return r;
}

// FP64 engine with the exact redo of lucaslehmer.cpp:331-335
Run run_fft(unsigned long p, unsigned long iters) {
    Run r = run_one(p, iters, LLH_FP64);
    if (r.status == LLH_ST_ROUNDOFF || r.status == LLH_ST_NOT_RUN) {
        Run e = run_exact(p, iters);
        e.maxerr = r.maxerr;
        return e;
    }
    return r;
}

bool trivial(unsigned long p, bool &v) {          // lucaslehmer.cpp:30-32
    if (p < 2) { v = false; return true; }
    if (p == 2) { v = true; return true; }
    if (p % 2 == 0) { v = false; return true; }
    return false;
}

// The argument rules set Res64 to 0 (the reference leaves the thread's previous value there).
bool lucas_lehmer(unsigned long p) {
    bool v;
    if (trivial(p, v)) { g_last_res64 = 0; return v; }
    Run r = run_exact(p, 0);
    g_last_res64 = r.res64;
    return r.verdict != 0;
}

bool lucas_lehmer_fft(unsigned long p) {
    bool v;
    if (trivial(p, v)) { g_last_res64 = 0; return v; }
    Run r = run_fft(p, 0);
    g_last_res64 = r.res64;
    return r.verdict != 0;
}

std::vector<unsigned long> primes_in_range(unsigned long lo, unsigned long hi) {   // lucaslehmer.cpp:71-80
    std::vector<unsigned long> out;
    if (hi < 2) return out;
    std::vector<bool> comp(hi + 1, false);
    for (unsigned long i = 2; i * i <= hi; ++i)
        if (!comp[i])
            for (unsigned long j = i * i; j <= hi; j += i) comp[j] = true;
    for (unsigned long i = (lo < 2 ? 2 : lo); i <= hi; ++i)
        if (!comp[i]) out.push_back(i);
    return out;
}

// decimal digits of 2^p - 1 (display only): base-1e9 doubling
std::string mersenne_str(unsigned long p) {
    std::vector<uint32_t> d{1};
    const uint32_t B = 1000000000u;
    for (unsigned long i = 0; i < p; ++i) {
        uint32_t c = 0;
        for (auto &x : d) {
            uint64_t v = (uint64_t)x * 2 + c;
            x = (uint32_t)(v % B);
            c = (uint32_t)(v / B);
        }
        if (c) d.push_back(c);
    }
    // 2^p - 1: 2^p is never a multiple of 10, so the low limb is >= 1
    d[0] -= 1;
    std::string s = std::to_string(d.back());
    char buf[16];
    for (size_t k = d.size() - 1; k-- > 0;) {
        std::snprintf(buf, sizeof buf, "%09u", d[k]);
        s += buf;
    }
    return s;
}

size_t mersenne_digits(unsigned long p) {
    if (p == 0) return 1;                        // "0"
    // 2^p - 1 has as many digits as 2^p, which is never a power of ten
    return (size_t)std::floor((long double)p * 0.301029995663981195213738894724493027L) + 1;
}

double bench_square(unsigned long bits, int iters, int depth) {
    (void)depth;                                 // the split depths are CPU GMP threading
    unsigned long p = bits | 1ul;
    if (p < 65) p = 65;
    Run r = run_one(p, (unsigned long)(iters > 0 ? iters : 1), LLH_FP64);
    return r.seconds / (double)(iters > 0 ? iters : 1);
}

}  // namespace

PYBIND11_MODULE(lucaslehmer, m, py::mod_gil_not_used()) {
    m.doc() = "Lucas-Lehmer Mersenne prime test on an MI355X (gfx950): exact Goldilocks NTT and FP64 IBDWT "
              "engines on the GPU (releases the GIL; free-threading safe)";
    m.def("lucas_lehmer", &lucas_lehmer, py::arg("p"), "True iff 2^p - 1 is prime (Lucas-Lehmer). p should be prime.");
    m.def("lucas_lehmer_split", [](unsigned long p, int) { return lucas_lehmer(p); }, py::arg("p"),
          py::arg("depth") = 1, "Same result as lucas_lehmer (the reference's thread split is a CPU GMP detail).");
    m.def("lucas_lehmer_fft", &lucas_lehmer_fft, py::arg("p"),
          "Lucas-Lehmer on the FP64 IBDWT engine (exact: redone on the exact engine on rounding trouble).");
    m.def("lucas_lehmer_res",
          [](unsigned long p) {
              bool r = lucas_lehmer(p);
              return std::make_pair(r, g_last_res64);
          },
          py::arg("p"), "(is_prime, res64) on the exact engine: res64 = low 64 bits of the final residue.");
    m.def("lucas_lehmer_split_res",
          [](unsigned long p, int) {
              bool r = lucas_lehmer(p);
              return std::make_pair(r, g_last_res64);
          },
          py::arg("p"), py::arg("depth") = 1, "(is_prime, res64), same as lucas_lehmer_res.");
    m.def("lucas_lehmer_fft_res",
          [](unsigned long p) {
              bool r = lucas_lehmer_fft(p);
              return std::make_pair(r, g_last_res64);
          },
          py::arg("p"), "(is_prime, res64) on the FP64 engine (exact redo on rounding trouble).");
    m.def("fft_res64_after",
          [](unsigned long p, unsigned long iters) {
              bool v;
              if (trivial(p, v)) return std::make_pair(g_last_res64 = 0, 0.0);
              Run r = run_fft(p, iters ? iters : 0);
              g_last_res64 = r.res64;
              return std::make_pair(g_last_res64, r.maxerr);
          },
          py::arg("p"), py::arg("iters"),
          "(res64, worst rounding error) after only `iters` squarings of the LL sequence for p.");
    m.def("gmp_res64_after",
          [](unsigned long p, unsigned long iters) {
              bool v;
              if (trivial(p, v)) return g_last_res64 = 0;
              Run r = run_exact(p, iters);
              g_last_res64 = r.res64;
              return g_last_res64;
          },
          py::arg("p"), py::arg("iters"), "res64 after only `iters` squarings on the exact engine (the arbiter).");
    m.def("fft_info",
          [](unsigned long p) {
              bool v;
              if (trivial(p, v)) {
                  uint64_t n = llh_fft_length(p, 0);
                  return std::make_tuple((size_t)n, n ? (double)p / (double)n : 0.0, 0.0);
              }
              Run r = run_one(p, 0, LLH_FP64);
              size_t n = r.n;
              return std::make_tuple(n, n ? (double)p / (double)n : 0.0, r.maxerr);
          },
          py::arg("p"), "(FFT length, bits per word, worst rounding error) of an FP64-engine test of p.");
    m.attr("fft_backend") = std::string("hip-gfx950-fp64");
    m.def("set_fft_bits_per_word", [](double b) { check(llh_set_fft_bits_per_word(b)); }, py::arg("bits"),
          "Max bits per FFT word (default 18). Higher = shorter FFT = faster, more rounding error.");
    m.def("bench_square", &bench_square, py::arg("bits"), py::arg("iters") = 20, py::arg("depth") = 0,
          "seconds per squaring (one FP64 IBDWT iteration on the GPU) of a `bits`-bit residue");
    m.def("primes", &primes_in_range, py::arg("lo"), py::arg("hi"), "List of primes in [lo, hi] (sieve of Eratosthenes).");
    m.def("mersenne_digits", &mersenne_digits, py::arg("p"), "Number of decimal digits of 2^p - 1.");
    m.def("mersenne_str", &mersenne_str, py::arg("p"), "Decimal string of 2^p - 1.");
    m.attr("gmp_version") = std::string("none (exact engine: Goldilocks NTT on gfx950)");
    struct LucasLehmer {};
    py::class_<LucasLehmer>(m, "LucasLehmer")
        .def(py::init<>())
        .def("sa_lucaslehmer", [](LucasLehmer &, unsigned long p, py::object) { return lucas_lehmer(p); },
             py::arg("N_"), py::arg("S_") = py::none())
        .def("sa_getListOfPrimes", [](LucasLehmer &, py::list primes, unsigned long lo, unsigned long hi) {
            for (auto v : primes_in_range(lo, hi)) primes.append(v);
            return true;
        });
}
