// lucaslehmer_hip_module.cpp -- counterpart of the reference's `lucaslehmer_cuda` module
// (cuda/ll_ntt.cu:274-281): lucas_lehmer_batch / ntt_length / max_exponent / device_info,
// backed by the one-launch persistent batch engine of libllhip.
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

// [(p, is_prime, seconds, res64)] in input order; seconds < 0 marks an exponent the batched
// engine cannot take (redo elsewhere), as cuda/ll_ntt.cu:256-263.
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

PYBIND11_MODULE(lucaslehmer_hip, m, py::mod_gil_not_used()) {
    m.doc() = "Lucas-Lehmer batches on an MI355X: all exponents in one persistent launch "
              "(exact Goldilocks NTT by default, FP64 IBDWT on request)";
    m.def("lucas_lehmer_batch", &batch, py::arg("exponents"), py::arg("iters") = 0, py::arg("transform") = "ntt",
          "[(p, is_prime, seconds, res64)] for a list of exponents; seconds < 0 marks one the engine cannot take.");
    m.def("ntt_length", [](uint64_t p) { return llh_batch_length(p, LLH_NTT); }, py::arg("p"),
          "Transform length the batched NTT engine would use for p (0 = too large).");
    m.def("fft_length", [](uint64_t p) { return llh_batch_length(p, LLH_FP64); }, py::arg("p"),
          "Transform length the batched FP64 engine would use for p (0 = too large).");
    m.def("max_exponent", []() { return llh_max_exponent(LLH_NTT); }, "Largest p the batched NTT engine handles.");
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
