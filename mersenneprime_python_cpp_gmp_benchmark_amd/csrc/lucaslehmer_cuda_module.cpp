// lucaslehmer_cuda_module.cpp -- the GPU batch module under the reference's import name,
// `lucaslehmer_cuda` (cuda/ll_ntt.cu:274-281), backed by libllhip's persistent batch engine.
//
//   lucas_lehmer_batch(exponents)  cuda/ll_ntt.cu:256-267, :276-277: [(p, is_prime, seconds, res64)].
//       seconds < 0 marks an exponent handed back to the caller (the reference's "redo on the
//       CPU"); here only an exponent no engine length covers, or an FP64 run whose round-off
//       stays above 0.4 up to the longest length, is handed back.  The reference rejects p < 64
//       and p > 204,800; this engine runs them (p < 64 in a one-word lane, larger p on the
//       large engine).  Rows come back in input order (the reference groups them by length).
//       Two optional arguments extend the call: iters (0 = the full p - 2 squarings) and
//       transform ("ntt", the reference's exact transform, or "fp64").
//   ntt_length(p)                 cuda/ll_ntt.cu:176-180, the reference rule itself.
//   max_exponent()                cuda/ll_ntt.cu:279: 8192 * max_bits_per_word(8192) = 204,800.
//   device_info()                 cuda/ll_ntt.cu:269-272: (name, compute units, global bytes).
//   batch_length(p, transform)    no reference counterpart: the length this engine runs p at.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "llhip.h"

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc != LLH_OK) throw std::runtime_error(std::string("llhip: ") + llh_last_error());
}

int parse_transform(const std::string &t) {
    if (t == "ntt") return LLH_NTT;
    if (t == "fp64") return LLH_FP64;
    throw std::invalid_argument("transform must be 'ntt' or 'fp64'");
}

std::vector<std::tuple<uint64_t, bool, double, uint64_t>> batch(std::vector<uint64_t> ps, uint64_t iters,
                                                                 const std::string &transform) {
    const int tr = parse_transform(transform);
    const size_t n = ps.size();
    std::vector<int32_t> v(n), st(n);
    std::vector<uint64_t> r(n);
    std::vector<double> s(n);
    int rc;
    {
        py::gil_scoped_release release;
        rc = llh_batch(ps.data(), n, iters, tr, v.data(), r.data(), s.data(), nullptr, nullptr, st.data());
    }
    check(rc);
    std::vector<std::tuple<uint64_t, bool, double, uint64_t>> out(n);
    for (size_t i = 0; i < n; ++i) {
        const bool ok = st[i] == LLH_ST_OK || st[i] == LLH_ST_TRIVIAL;
        out[i] = std::make_tuple(ps[i], ok && v[i] != 0, ok ? s[i] : -1.0, ok ? r[i] : 0ull);
    }
    return out;
}

}  // namespace

PYBIND11_MODULE(lucaslehmer_cuda, m, py::mod_gil_not_used()) {
    m.doc() = "Lucas-Lehmer batches on an MI355X: every exponent of a list in one persistent launch "
              "(exact Goldilocks NTT IBDWT by default, FP64 IBDWT on request)";
    m.def("lucas_lehmer_batch", &batch, py::arg("exponents"), py::arg("iters") = 0, py::arg("transform") = "ntt",
          "[(p, is_prime, seconds, res64)] in input order; seconds < 0 marks an exponent handed back.");
    m.def("ntt_length", [](uint64_t p) { return llh_ntt_length(p); }, py::arg("p"),
          "Goldilocks transform length of the reference rule for p (0 = above 204,800).");
    m.def("max_exponent", []() { return llh_max_exponent(LLH_NTT); },
          "Largest p the reference rule's lengths cover (204,800); larger p also run, on the large engine.");
    m.def("batch_length", [](uint64_t p, const std::string &t) { return llh_batch_length(p, parse_transform(t)); },
          py::arg("p"), py::arg("transform") = "ntt",
          "Length this engine's batch runs p at (0 = the large engine takes it).");
    m.def("device_info",
          []() {
              char name[256], arch[64];
              int cus = 0;
              uint64_t mem = 0;
              check(llh_device_info(name, sizeof name, &cus, &mem, arch, sizeof arch));
              return std::make_tuple(std::string(name), cus, (size_t)mem);
          },
          "(name, compute units, global memory bytes)");
}
