// ll_capi.hip -- the C ABI (include/llhip.h): device state, table cache, batch orchestration.
//
// Host-side counterpart of cuda/ll_ntt.cu:198-267 (run_batch / lucas_lehmer_batch): group
// exponents by transform length, build the length's tables once (kept for the process, as
// the reference keeps FFTW plans, lucaslehmer.cpp:228-242), one persistent launch for the
// whole list, results scattered back in input order.  FP64 exponents whose round-off passes
// 0.4 are re-queued at twice the length (rollback) in a follow-up launch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/llhip.h"
#include "ll_internal.hpp"

namespace llh {

int mlp_n_params();
struct MlpBuffers {
    float *P, *X, *T, *Y, *H1, *H2, *DZ1, *DZ2, *DY, *G, *Mm, *Vv, *loss;
};
hipError_t mlp_alloc(MlpBuffers &B, size_t n);
void mlp_free(MlpBuffers &B);
hipError_t mlp_forward(const MlpBuffers &B, int n, bool with_backward, hipStream_t st);
hipError_t mlp_gradient(const MlpBuffers &B, int n, hipStream_t st);
hipError_t mlp_step(const MlpBuffers &B, float lr, int t, hipStream_t st);

hipError_t launch_batch(int transform, int family, const ClassDesc *d_classes, int ncls, const BatchItem *d_items,
                        BatchResult *d_out, u32 *d_counters, int grid, hipStream_t stream);
int batch_blocks_per_cu(int transform, int family);
hipError_t batch_resume(int transform, const ResumeReq *d_reqs, int n, i64 *rb_area, i64 *resume, int *d_ok,
                        hipStream_t st);

bool large_launch_for(int transform, u64 M, int sched, LargeLaunch &L);
std::vector<LargeSched> large_schedules(int transform, u64 M);
int large_blocks(const LargeLaunch &L);
hipError_t large_squarings(const LargeLaunch &L, LargeArgs A, int k, hipStream_t st);
hipError_t large_apply_carry(const LargeLaunch &L, const LargeArgs &A, hipStream_t st);
hipError_t large_normalize(int transform, const void *elem, i64 *dig, const LargeArgs &A, void *out, hipStream_t st);
size_t large_normalize_bytes(u64 n);
hipError_t large_relayout(int transform, const i64 *dig, u32 p, u64 n, u64 n2, void *elem2, u32 w2, u32 h2,
                          hipStream_t st);
void relayout_host(const i64 *dig, u64 p, u64 n, u64 n2, i64 *out);

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

struct TableSet {
    void *tw = nullptr, *twi = nullptr, *wf = nullptr, *wi = nullptr;
    void *twp = nullptr, *twpi = nullptr, *wtab = nullptr;
    int log_s = 0;
};

struct DeviceState {
    int id = -1;
    hipStream_t stream = nullptr;
    hipStream_t fstream[4] = {nullptr, nullptr, nullptr, nullptr};   // batch families (0: stream)
    int cus = 0;
    double wclk_hz = 1e8;
    int bpc[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};              // workgroups per CU, per kernel family
    std::map<std::pair<int, int>, TableSet> tables;                   // (transform, class)
    void *d_items = nullptr, *d_out = nullptr, *d_cls = nullptr, *d_cnt = nullptr;
    size_t cap_items = 0, cap_cls = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evf[5] = {};              // batch launch timing / fork-join
    unsigned char *scratch = nullptr;                                  // device address of `scratch`
    std::string scratch_err;                                           // why it is null
    // batch checkpoints: per-group slots (one region per family), rollback entries, resume
    // entries, the rollback counter and the resume requests
    void *ck = nullptr, *rb = nullptr, *res = nullptr, *rbc = nullptr, *req = nullptr;
    size_t cap_ck = 0, cap_rb = 0, cap_res = 0, cap_rbc = 0, cap_req = 0;
};

struct LargeTables {
    void *twW = nullptr, *twWi = nullptr, *twH = nullptr, *twHi = nullptr;
    void *tmL = nullptr, *tmH = nullptr, *tmLi = nullptr, *tmHi = nullptr;
    void *wfL = nullptr, *wfH = nullptr, *wiL = nullptr, *wiH = nullptr;
    u32 log_s = 0, log_sw = 0;
};

static std::recursive_mutex g_mx;
static DeviceState g_dev;
static std::map<std::tuple<int, u64, int>, LargeTables> g_large_tables;   // (transform, N, W)
// Schedule of each large length (index into large_schedules); a length without an entry runs the
// measured choice compiled in from ll_large_sched_defaults.inc, else the list's first schedule.
static std::map<std::pair<int, u64>, int> g_large_sched;
static std::vector<double> g_err_trace;     // per-iteration max round-off of the last large run
static double g_last_kernel_ms = 0.0;     // kernel time of the last llh_batch call (all its launches)
static u32 g_last_launches = 0;
static u64 g_last_squarings = 0;          // squarings both engines ran in the last call (kept or not)
static u32 g_last_rollbacks = 0;          // rollbacks in the last call
static u32 g_last_resumed = 0;            // of those, how many resumed from a checkpoint
static u64 g_last_resumed_at = 0;         // iteration the last of them resumed from
static std::atomic<double> g_bpw{18.0};        // reference knob: the conservative length rule
// Bits-per-digit cap of the batched FP64 engine (lengths up to 8192).  The worst round-off over
// every prime <= 100001 roughly doubles per half bit: 0.05 at 19.75-20, 0.10 at 20.25-20.5,
// 0.21 at 20.75-21, 0.40 at 21.25-21.5 bits/digit (then rollbacks start).  21 keeps it at half
// the 0.4 rollback threshold, so short lengths take more bits than the reference's 18.
static std::atomic<double> g_batch_bpw{21.0};
// Length families the batched engine may pick (bit f: family f; powers of two always).
static std::atomic<int> g_batch_families{0xF};
// Checkpoints.  The large engine copies its state every k graph blocks (0 = off: a rollback
// restarts from s_0).  The batch engine's checkpoints (written at every round-off check) are
// off by default: its checkpointing kernels run C3 2.3% slower, and at the default cap no batch
// exponent rolls back.
static std::atomic<int> g_ckpt_blocks{1};
static std::atomic<int> g_batch_ckpt{0};

// Measured cost of each batched length: device microseconds per exponent-squaring with the
// GPU full of exponents of that length (the autotuner's table; ll_batch_costs.inc holds the
// values measured on an MI355X).  Among the lengths within the bits-per-digit cap the engine
// runs the cheapest, not merely the shortest.
struct BatchCost {
    u64 n;
    int transform;
    double us;
};
static const BatchCost kBatchCostDefaults[] = {
#include "ll_batch_costs.inc"
    {0, 0, 0.0}};
static std::mutex g_cost_mx;
static std::map<std::pair<int, u64>, double> &batch_costs() {
    static std::map<std::pair<int, u64>, double> m = [] {
        std::map<std::pair<int, u64>, double> d;
        for (const BatchCost &c : kBatchCostDefaults)
            if (c.n) d[{c.transform, c.n}] = c.us;
        return d;
    }();
    return m;
}
static double batch_cost(int transform, u64 N) {
    std::lock_guard<std::mutex> lk(g_cost_mx);
    auto &m = batch_costs();
    auto it = m.find({transform, N});
    return it == m.end() ? 0.0 : it->second;
}

// batched engine limits
static const int kFp64MaxLog = 13;                    // N <= 8192 (M <= 4096)
static const int kNttMaxLog = 12;                     // N = 16 .. 4096

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string("HIP: ") + hipGetErrorString(e_) + " at " #x);         \
            return LLH_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// Device memory of llh_batch_noalloc: the launch's items, results, class entries and queue
// counters, in one file-scope array (each GPU the library runs on holds its own copy until the
// process ends).  Host code reaches it only through the address hipGetSymbolAddress returns at
// set-up (DeviceState::scratch), never through the name.
constexpr size_t kScratchItems = 16384;          // exponents per call
constexpr size_t kScratchClasses = 128;          // class entries per launch
constexpr size_t kScratchItemBytes = kScratchItems * sizeof(BatchItem);
constexpr size_t kScratchOutBytes = kScratchItems * sizeof(BatchResult);
constexpr size_t kScratchClsBytes = kScratchClasses * sizeof(ClassDesc);
constexpr size_t kScratchBytes = kScratchItemBytes + kScratchOutBytes + kScratchClsBytes + kScratchClasses * 4;
static_assert(kScratchItemBytes % 256 == 0 && kScratchOutBytes % 256 == 0 && kScratchClsBytes % 256 == 0,
              "scratch sections stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The per-class queue counters are cleared by this kernel rather than by hipMemsetAsync, so a
// caller may capture the launch sequence in a graph.
__global__ void fill_u32(u32 *p, u32 v, u32 n) {
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

static int build_batch_tables();

static int init_locked(int device) {
    if (g_dev.id == device) return LLH_OK;
    if (g_dev.id >= 0) {
        set_error("llh_init: already initialised on another device");
        return LLH_ERR_ARG;
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("llh_init: no such device");
        return LLH_ERR_ARG;
    }
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    g_dev.fstream[0] = g_dev.stream;
    for (int f = 1; f < 4; ++f) HIPCHK(hipStreamCreateWithFlags(&g_dev.fstream[f], hipStreamNonBlocking));
    hipDeviceProp_t pr;
    HIPCHK(hipGetDeviceProperties(&pr, device));
    g_dev.cus = pr.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        g_dev.wclk_hz = (double)khz * 1e3;
    for (int f = 0; f < batch_num_families() && f < 5; ++f) {
        g_dev.bpc[0][f] = batch_blocks_per_cu(kFP64, f);
        g_dev.bpc[1][f] = batch_blocks_per_cu(kNTT, f);
    }
    HIPCHK(hipEventCreate(&g_dev.ev0));
    HIPCHK(hipEventCreate(&g_dev.ev1));
    for (int f = 0; f < 5; ++f) HIPCHK(hipEventCreateWithFlags(&g_dev.evf[f], hipEventDisableTiming));
    // the array's address on this device; a failure disables llh_batch_noalloc only
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se == hipSuccess) {
        g_dev.scratch = (unsigned char *)sp;
    } else {
        g_dev.scratch = nullptr;
        g_dev.scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
        (void)hipGetLastError();
    }
    g_dev.id = device;
    // every batch class's tables, so that llh_batch_noalloc never has to allocate them
    return build_batch_tables();
}

static int ensure_init() {
    if (g_dev.id >= 0) {
        HIPCHK(hipSetDevice(g_dev.id));
        return LLH_OK;
    }
    return init_locked(0);
}

// cos(2 pi j / n), sin(2 pi j / n) with an exact quadrant/octant reduction, long double.
static void sincos_frac(u64 j, u64 n, long double &c, long double &s) {
    j %= n;
    const long double HALFPI = 1.57079632679489661923132169163975144L;
    u64 q = (4 * j) / n;               // quadrant
    u64 rem = 4 * j - q * n;           // angle = (q + rem/n) * pi/2
    long double cs, sn;
    if (2 * rem <= n) {
        long double phi = HALFPI * (long double)rem / (long double)n;
        cs = cosl(phi);
        sn = sinl(phi);
    } else {
        long double phi = HALFPI * (long double)(n - rem) / (long double)n;
        cs = sinl(phi);
        sn = cosl(phi);
    }
    switch (q) {
        case 0: c = cs; s = sn; break;
        case 1: c = -sn; s = cs; break;
        case 2: c = -cs; s = -sn; break;
        default: c = sn; s = -cs; break;
    }
}

// Caller holds g_mx.
static int ceil_log2(u64 v);

// Tables of batch class c: FP64 (N = 2M): the full w_M table (real-spectrum pairing), the
// per-pass twiddles and the two-level weights; GL (N = M): full and per-pass roots and the
// root-of-two weights.
static int get_tables(int transform, int c, TableSet &out) {
    auto key = std::make_pair(transform, c);
    auto it = g_dev.tables.find(key);
    if (it != g_dev.tables.end()) {
        out = it->second;
        return LLH_OK;
    }
    const u64 M = batch_class_m(c);
    std::vector<u32> ex;
    if (!M || !batch_pass_exponents(c, ex)) {
        set_error("no batch schedule for this length");
        return LLH_ERR_UNSUPPORTED;
    }
    TableSet ts;
    if (transform == kFP64) {
        const u64 N = 2 * M;
        std::vector<double> tw(2 * M);
        for (u64 j = 0; j < M; ++j) {
            long double cs, sn;
            sincos_frac(j, M, cs, sn);
            tw[2 * j] = (double)cs;
            tw[2 * j + 1] = (double)(-sn);          // w_M^j = exp(-2 pi i j / M)
        }
        std::vector<double> twp(2 * ex.size() + 2);
        for (size_t e = 0; e < ex.size(); ++e) {
            twp[2 * e] = tw[2 * ex[e]];
            twp[2 * e + 1] = tw[2 * ex[e] + 1];
        }
        // two-level weights: e = lo + S hi, w(e) = 2^(lo/N) 2^(S hi/N) (inverse with 1/M)
        const int log_s = (ceil_log2(N) + 1) / 2;
        const u64 S = (u64)1 << log_s, HS = (N + S - 1) >> log_s;
        if (S > (u64)kWtabSlot || HS > (u64)kWtabSlot) {
            set_error("weight table slot too small");
            return LLH_ERR_UNSUPPORTED;
        }
        std::vector<double> wtab(4 * kWtabSlot, 0.0);
        for (u64 i = 0; i < S; ++i) {
            wtab[i] = (double)exp2l((long double)i / (long double)N);
            wtab[2 * kWtabSlot + i] = (double)exp2l(-(long double)i / (long double)N);
        }
        for (u64 i = 0; i < HS; ++i) {
            wtab[kWtabSlot + i] = (double)exp2l((long double)(i * S) / (long double)N);
            wtab[3 * kWtabSlot + i] = (double)(exp2l(-(long double)(i * S) / (long double)N) / (long double)M);
        }
        ts.log_s = log_s;
        HIPCHK(hipMalloc(&ts.tw, tw.size() * 8));
        HIPCHK(hipMalloc(&ts.twp, twp.size() * 8));
        HIPCHK(hipMalloc(&ts.wtab, wtab.size() * 8));
        HIPCHK(hipMemcpy(ts.tw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.twp, twp.data(), twp.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.wtab, wtab.data(), wtab.size() * 8, hipMemcpyHostToDevice));
    } else {
        const u64 N = M;
        const u64 w = gl_root_unity_aligned(N), wv = gl_inv_h(w);
        const u64 r2 = gl_root2_h(N), r2i = gl_inv_h(r2), ninv = gl_inv_h(N);
        if (!batch_class_pow2(c) || !w || !r2) {
            set_error("no Goldilocks roots for this length");
            return LLH_ERR_UNSUPPORTED;
        }
        std::vector<u64> tw(N), twi(N), wf(N), wi(N);
        u64 a = 1, b = 1, cc = 1, d = ninv;
        for (u64 j = 0; j < N; ++j) {
            tw[j] = a; twi[j] = b; wf[j] = cc; wi[j] = d;
            a = gl_mul_h(a, w); b = gl_mul_h(b, wv); cc = gl_mul_h(cc, r2); d = gl_mul_h(d, r2i);
        }
        std::vector<u64> twp(ex.size() + 1), twpi(ex.size() + 1);
        for (size_t e = 0; e < ex.size(); ++e) {
            twp[e] = tw[ex[e]];
            twpi[e] = twi[ex[e]];
        }
        HIPCHK(hipMalloc(&ts.tw, N * 8));
        HIPCHK(hipMalloc(&ts.twi, N * 8));
        HIPCHK(hipMalloc(&ts.wf, N * 8));
        HIPCHK(hipMalloc(&ts.wi, N * 8));
        HIPCHK(hipMalloc(&ts.twp, twp.size() * 8));
        HIPCHK(hipMalloc(&ts.twpi, twpi.size() * 8));
        HIPCHK(hipMemcpy(ts.tw, tw.data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.twi, twi.data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.wf, wf.data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.wi, wi.data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.twp, twp.data(), twp.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ts.twpi, twpi.data(), twpi.size() * 8, hipMemcpyHostToDevice));
    }
    g_dev.tables[key] = ts;
    out = ts;
    return LLH_OK;
}

// Batched-engine lengths (N, base class), ascending: FP64 N = 2M over every length, GL N = M over
// the power-of-two lengths.  Schedule variants of a length are not listed (batch_class_of picks).
static const std::vector<std::pair<u64, int>> &batch_ladder(int transform) {
    static const std::vector<std::pair<u64, int>> lad[2] = {
        [] {
            std::vector<std::pair<u64, int>> v;
            for (int c = 1; c < batch_num_classes(); ++c)
                if (batch_class_base(c) == c) v.push_back({2 * batch_class_m(c), c});
            std::sort(v.begin(), v.end());
            return v;
        }(),
        [] {
            std::vector<std::pair<u64, int>> v;
            for (int c = 1; c < batch_num_classes(); ++c)
                if (batch_class_pow2(c) && batch_class_base(c) == c) v.push_back({batch_class_m(c), c});
            std::sort(v.begin(), v.end());
            return v;
        }()};
    return lad[transform == kNTT ? 1 : 0];
}

// The schedules of base class b (b itself first, then its variants in class order); the GL
// engine runs the power-of-two ones only.
static std::vector<int> batch_schedules_of(int transform, int b) {
    std::vector<int> v;
    if (b <= 0) return v;
    v.push_back(b);
    for (int c = 1; c < batch_num_classes(); ++c)
        if (c != b && batch_class_base(c) == b && (transform == kFP64 || batch_class_pow2(c))) v.push_back(c);
    return v;
}

// Schedule of each batched length (index into batch_schedules_of); a length without an entry runs
// the measured choice compiled in from ll_batch_sched_defaults.inc, else its base class.
static std::mutex g_bsched_mx;
static std::map<std::pair<int, u64>, int> g_batch_sched;
struct BatchSchedDefault {
    int transform;
    u64 n;
    int index;
};
static const BatchSchedDefault kBatchSchedDefaults[] = {
#define LLB(TR, NV, IDX) {TR, NV, IDX},
#include "ll_batch_sched_defaults.inc"
#undef LLB
    {-1, 0, 0}};

static int batch_sched_index(int transform, u64 N) {
    {
        std::lock_guard<std::mutex> lk(g_bsched_mx);
        auto it = g_batch_sched.find(std::make_pair(transform, N));
        if (it != g_batch_sched.end()) return it->second;
    }
    for (const BatchSchedDefault &d : kBatchSchedDefaults)
        if (d.transform == transform && d.n == N) return d.index;
    return 0;
}

// Tables of every batch class of both transforms (the NTT runs the power-of-two classes only).
static int build_batch_tables() {
    for (int c = 1; c < batch_num_classes(); ++c) {
        TableSet ts;
        int rc = get_tables(kFP64, c, ts);
        if (!rc && batch_class_pow2(c)) rc = get_tables(kNTT, c, ts);
        if (rc) return rc;
    }
    return LLH_OK;
}

// the kernels index digits with 32-bit products j * p, j < N
static bool batch_fits(u64 p, u64 N) { return (N + 2) * p < ((u64)1 << 32); }

// Batch class that runs length N: its schedule in force (0 = not a batched length).
static int batch_class_of(int transform, u64 N) {
    for (const auto &e : batch_ladder(transform))
        if (e.first == N) {
            const std::vector<int> v = batch_schedules_of(transform, e.second);
            const int i = batch_sched_index(transform, N);
            return i >= 0 && i < (int)v.size() ? v[i] : e.second;
        }
    return 0;
}

// The batched engine's length for p (0 = not in the batched engine).  Candidates: FP64 the
// lengths with p / N <= the batch bits-per-digit cap, GL those within the exactness bound, in
// the enabled families, up to twice the shortest.  Of these the one with the lowest measured
// cost runs (a length without a measured cost never displaces a shorter one).
static u64 batch_length(u64 p, int transform) {
    if (p < 64) return 0;
    const double bpw = g_batch_bpw.load();
    const int fam = g_batch_families.load() | 1;
    u64 best = 0, shortest = 0;
    double best_cost = 0.0;
    for (const auto &e : batch_ladder(transform)) {
        const u64 N = e.first;
        if (!((fam >> batch_class_factor(e.second)) & 1)) continue;
        const bool ok = transform == kFP64 ? (double)p / (double)N <= bpw
                                           : (p + N - 1) / N <= (u64)ntt_max_bits(N);
        if (!ok) continue;
        if (!shortest) {
            if (!batch_fits(p, N)) return 0;
            shortest = N;
        }
        if (N > 2 * shortest || !batch_fits(p, N)) break;
        const double c = batch_cost(transform, N);
        if (!best) {
            best = N;
            best_cost = c;
        } else if (c > 0.0 && best_cost > 0.0 && c < best_cost) {
            best = N;
            best_cost = c;
        }
    }
    return best;
}

// next batched length above N for p (0 = none): the rollback target
static u64 next_batch_length(u64 p, u64 N, int transform) {
    const int fam = g_batch_families.load() | 1;
    for (const auto &e : batch_ladder(transform))
        if (e.first > N && ((fam >> batch_class_factor(e.second)) & 1)) return batch_fits(p, e.first) ? e.first : 0;
    return 0;
}

static int grow(void **ptr, size_t &cap, size_t need) {
    if (need <= cap) return LLH_OK;
    if (*ptr) HIPCHK(hipFree(*ptr));
    *ptr = nullptr;
    size_t c = std::max(need, cap * 2);
    HIPCHK(hipMalloc(ptr, c));
    cap = c;
    return LLH_OK;
}

struct Pending {
    u64 p;
    u64 iters;
    size_t idx;
    int cls;            // batched engine: size class (0 = the one-word lane, p < 64)
    u64 n = 0;          // the length (batched or large engine)
    u64 it0 = 0;        // squarings done at the checkpoint it resumes from
    u32 init = 0;       // resume entry + 1 (0: from s_0)
    i64 rb = -1;        // rollback entry holding that checkpoint (before the re-layout)
    u64 n_old = 0;      // the checkpoint's length
};

constexpr u32 kRollbackCap = 256;      // checkpointed rollbacks per launch (more restart from s_0)

// The persistent launches over `work` (one per class family, on separate streams so each
// family's tail is back-filled by the others); fills res[i] for work[i].
static int run_launch(int transform, std::vector<Pending> &work, std::vector<BatchResult> &res,
                      bool noalloc = false) {
    const size_t n = work.size();
    res.assign(n, BatchResult());
    if (!n) return LLH_OK;
    // family launch order: odd-factor families (the longest exponents) first, then the power of
    // two M = 4096, the shorter powers of two and the one-word lane last
    auto fam_rank = [](int c) {
        const int f = batch_class_family(c);
        return f == 0 ? 4 : f == 4 ? 3 : f - 1;
    };
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    // per family: longest first, by length descending, p descending; the one-word lane last
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Pending &x = work[a], &y = work[b];
        const int fx = fam_rank(x.cls), fy = fam_rank(y.cls);
        if (fx != fy) return fx < fy;
        const u64 nx = x.cls ? x.n : 0, ny = y.cls ? y.n : 0;
        if (nx != ny) return nx > ny;
        return x.p > y.p;
    });
    std::vector<BatchItem> items(n);
    std::vector<ClassDesc> cls;
    std::vector<int> fam_of;            // family of each class entry
    for (size_t k = 0; k < n; ++k) {
        const Pending &w = work[order[k]];
        items[k].p = (u32)w.p;
        items[k].log_n = 0;
        items[k].iters = w.iters;
        items[k].slot = (u32)order[k];
        items[k].it0 = (u32)w.it0;
        items[k].init = w.init;
        items[k].pad = 0;
        // FP64 carry rounds (ll_batch.hip): one from log2 N + 3 bits per digit, else two.  Within
        // a class the items run p-descending, so each class splits into at most two entries.
        const int two = transform == kNTT || !w.cls || w.p / w.n < (u64)(ceil_log2(w.n) + 3) ? 1 : 0;
        if (cls.empty() || cls.back().cls != w.cls || cls.back().carry2 != two) {
            ClassDesc cd;
            std::memset(&cd, 0, sizeof(cd));
            cd.cls = w.cls;
            cd.carry2 = two;
            const u64 M = batch_class_m(w.cls);
            cd.log_m = M ? ceil_log2(M) : 0;
            cd.log_n = w.cls ? ceil_log2(w.n) : 0;     // exact for the power-of-two classes
            cd.first = (u32)k;
            cd.count = 0;
            if (w.cls) {
                TableSet ts;
                int rc = get_tables(transform, w.cls, ts);
                if (rc) return rc;
                cd.tw = ts.tw; cd.twi = ts.twi; cd.wf = ts.wf; cd.wi = ts.wi;
                cd.twp = ts.twp; cd.twpi = ts.twpi; cd.wtab = ts.wtab; cd.log_s = ts.log_s;
            }
            cd.ck_slots = nullptr;
            cd.rb_area = nullptr;
            cd.rb_count = nullptr;
            cd.rb_cap = 0;
            cd.resume = (const i64 *)g_dev.res;
            cls.push_back(cd);
            fam_of.push_back(batch_class_family(w.cls));
        }
        cls.back().count++;
    }
    int rc;
    BatchItem *d_items;
    BatchResult *d_out;
    ClassDesc *d_cls;
    u32 *d_cnt;
    if (noalloc) {                      // the file-scope array, through its device address
        if (!g_dev.scratch) {
            set_error("llh_batch_noalloc: " + g_dev.scratch_err);
            return LLH_ERR_HIP;
        }
        if (n > kScratchItems || cls.size() > kScratchClasses) {
            set_error("llh_batch_noalloc: more than 16384 exponents or 128 class entries in one launch");
            return LLH_ERR_ARG;
        }
        d_items = (BatchItem *)g_dev.scratch;
        d_out = (BatchResult *)(g_dev.scratch + kScratchItemBytes);
        d_cls = (ClassDesc *)(g_dev.scratch + kScratchItemBytes + kScratchOutBytes);
        d_cnt = (u32 *)(g_dev.scratch + kScratchItemBytes + kScratchOutBytes + kScratchClsBytes);
    } else {
        if ((rc = grow(&g_dev.d_items, g_dev.cap_items, n * sizeof(BatchItem) + n * sizeof(BatchResult) + 64)))
            return rc;
        if ((rc = grow(&g_dev.d_cls, g_dev.cap_cls, cls.size() * (sizeof(ClassDesc) + 16) + 64))) return rc;
        d_items = (BatchItem *)g_dev.d_items;
        d_out = (BatchResult *)((char *)g_dev.d_items + ((n * sizeof(BatchItem) + 63) / 64) * 64);
        d_cls = (ClassDesc *)g_dev.d_cls;
        d_cnt = (u32 *)((char *)g_dev.d_cls + ((cls.size() * sizeof(ClassDesc) + 63) / 64) * 64);
    }
    hipStream_t st = g_dev.stream;
    const int tr = transform == kNTT ? 1 : 0;
    // FP64 checkpoints: one slot region per kernel family (its largest grid x 16 WG digits)
    if (transform == kFP64 && !noalloc && g_batch_ckpt.load()) {
        size_t off[5], tot = 0;
        for (int f = 0; f < 5; ++f) {
            off[f] = tot;
            tot += (size_t)g_dev.cus * (size_t)std::max(g_dev.bpc[tr][f], 1) * 16 * (size_t)batch_family_wg(f) * 8;
        }
        if ((rc = grow(&g_dev.ck, g_dev.cap_ck, tot)) ||
            (rc = grow(&g_dev.rb, g_dev.cap_rb, (size_t)kRollbackCap * kResumeStride * 8)) ||
            (rc = grow(&g_dev.rbc, g_dev.cap_rbc, 256)))
            return rc;
        for (size_t c = 0; c < cls.size(); ++c) {
            if (!cls[c].cls) continue;
            cls[c].ck_slots = (i64 *)((char *)g_dev.ck + off[fam_of[c]]);
            cls[c].rb_area = (i64 *)g_dev.rb;
            cls[c].rb_count = (u32 *)g_dev.rbc;
            cls[c].rb_cap = kRollbackCap;
        }
        hipLaunchKernelGGL(fill_u32, dim3(1), dim3(128), 0, st, (u32 *)g_dev.rbc, 0u, 1u);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(d_items, items.data(), n * sizeof(BatchItem), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_cls, cls.data(), cls.size() * sizeof(ClassDesc), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(128), 0, st, d_cnt, 0u, (u32)cls.size());
    HIPCHK(hipGetLastError());
    // kernel family f runs on stream f (the M = 4096 family shares the 3 * 2^k family's stream:
    // four streams fill the four hardware queues, and the cost table never picks 3 * 2^k there)
    constexpr int kFam = 5;
    hipEvent_t ev0 = g_dev.ev0, ev1 = g_dev.ev1, *evf = g_dev.evf;
    HIPCHK(hipEventRecord(ev0, st));
    bool joined[kFam] = {false, false, false, false, false};
    for (size_t i = 0; i < cls.size();) {
        // the run of class entries of one family
        size_t j = i;
        size_t waves = 0;
        while (j < cls.size() && fam_of[j] == fam_of[i]) {
            const int g = batch_class_groups(cls[j].cls);
            waves += cls[j].cls ? (cls[j].count + g - 1) / g : 1;
            ++j;
        }
        const int f = fam_of[i];
        hipStream_t fs = g_dev.fstream[f == 4 ? 1 : f];
        if (fs != st) HIPCHK(hipStreamWaitEvent(fs, ev0, 0));
        const int grid = (int)std::min<size_t>(std::max<size_t>(waves, 1), (size_t)g_dev.cus * g_dev.bpc[tr][f]);
        HIPCHK(launch_batch(transform, f, d_cls + i, (int)(j - i), d_items, d_out, d_cnt + i, grid, fs));
        if (fs != st) {
            HIPCHK(hipEventRecord(evf[f], fs));
            joined[f] = true;
        }
        i = j;
    }
    for (int f = 0; f < kFam; ++f)
        if (joined[f]) HIPCHK(hipStreamWaitEvent(st, evf[f], 0));
    HIPCHK(hipEventRecord(ev1, st));
    HIPCHK(hipMemcpyAsync(res.data(), d_out, n * sizeof(BatchResult), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    g_last_kernel_ms += ms;
    g_last_launches += 1;
    return LLH_OK;
}

// ---- large engine ------------------------------------------------------------------------
static const int kGraphBlock = 200;       // squarings per captured graph / round-off check

template <class V>
static int upload(void **dst, const std::vector<V> &v) {
    HIPCHK(hipMalloc(dst, v.size() * sizeof(V)));
    HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return LLH_OK;
}

static int ceil_log2(u64 v) {
    int l = 0;
    while (((u64)1 << l) < v) ++l;
    return l;
}

// Measured choice of schedule per large length (autotune.py --schedules writes the file).
struct LargeSchedDefault {
    int transform;
    u64 n;
    int w, h, cw;
};
static const LargeSchedDefault kLargeSchedDefaults[] = {
#define LLD(TR, NV, WV, HV, CWV) {TR, NV, WV, HV, CWV},
#include "ll_large_sched_defaults.inc"
#undef LLD
    {-1, 0, 0, 0, 0}};

// Schedule index the large engine runs length N at (caller holds g_mx).
static int large_sched_of(int transform, u64 N) {
    auto it = g_large_sched.find(std::make_pair(transform, N));
    if (it != g_large_sched.end()) return it->second;
    const u64 M = transform == kFP64 ? N / 2 : N;
    for (const LargeSchedDefault &d : kLargeSchedDefaults) {
        if (d.transform != transform || d.n != N) continue;
        const std::vector<LargeSched> v = large_schedules(transform, M);
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].w == d.w && v[i].h == d.h && v[i].cw == d.cw) return (int)i;
    }
    return 0;
}

// A transform's twiddles w^j (j < n, `words` values each) rearranged into the per-pass order the
// kernels read: entry e holds w^{exp(e)}, e < size.
template <class V>
static std::vector<V> pass_order(const std::vector<V> &nat, int words, u32 (*exp)(int), int size) {
    std::vector<V> v((size_t)size * words);
    for (int e = 0; e < size; ++e)
        for (int k = 0; k < words; ++k) v[(size_t)e * words + k] = nat[(size_t)exp(e) * words + k];
    return v;
}

static int get_large_tables(int transform, u64 N, int sched, LargeTables &out) {
    const u64 M0 = transform == kFP64 ? N / 2 : N;
    const std::vector<LargeSched> sv = large_schedules(transform, M0);
    if (sched < 0 || sched >= (int)sv.size()) {
        set_error("no such large-engine schedule for this length");
        return LLH_ERR_UNSUPPORTED;
    }
    auto key = std::make_tuple(transform, N, sv[sched].w);
    auto it = g_large_tables.find(key);
    if (it != g_large_tables.end()) {
        out = it->second;
        return LLH_OK;
    }
    const u64 M = M0;
    LargeLaunch L;
    if (!large_launch_for(transform, M, sched, L)) {
        set_error("no large-engine kernels for this length");
        return LLH_ERR_UNSUPPORTED;
    }
    const u64 W = (u64)L.w_len, H = (u64)L.h_len;
    LargeTables T;
    T.log_s = (u32)((ceil_log2(M) + 1) / 2);
    T.log_sw = (u32)((ceil_log2(N) + 1) / 2);
    const u64 S = (u64)1 << T.log_s, Sw = (u64)1 << T.log_sw;
    const u64 nS = (M + S - 1) / S, nSw = (N + Sw - 1) / Sw;
    int rc;
    if (transform == kFP64) {
        auto roots = [](u64 n, u64 count, u64 stride) {
            std::vector<double> v(2 * count);
            for (u64 j = 0; j < count; ++j) {
                long double c, s;
                sincos_frac(j * stride, n, c, s);
                v[2 * j] = (double)c;
                v[2 * j + 1] = (double)(-s);
            }
            return v;
        };
        if ((rc = upload(&T.twW, pass_order(roots(W, W, 1), 2, L.row_exp, L.row_tab)))) return rc;
        if ((rc = upload(&T.twH, pass_order(roots(H, H, 1), 2, L.col_exp, L.col_tab)))) return rc;
        if ((rc = upload(&T.tmL, roots(M, S, 1)))) return rc;
        if ((rc = upload(&T.tmH, roots(M, nS, S)))) return rc;
        std::vector<double> wfL(Sw), wfH(nSw), wiL(Sw), wiH(nSw);
        for (u64 e = 0; e < Sw; ++e) {
            long double v = exp2l((long double)e / (long double)N);
            wfL[e] = (double)v;
            wiL[e] = (double)(1.0L / v);
        }
        for (u64 i = 0; i < nSw; ++i) {
            long double v = exp2l((long double)(i * Sw) / (long double)N);
            wfH[i] = (double)v;
            wiH[i] = (double)(1.0L / (v * (long double)M));
        }
        if ((rc = upload(&T.wfL, wfL)) || (rc = upload(&T.wfH, wfH)) || (rc = upload(&T.wiL, wiL)) ||
            (rc = upload(&T.wiH, wiH)))
            return rc;
    } else {
        const u64 w = gl_root_unity_aligned(M), wv = gl_inv_h(w);
        const u64 r2 = gl_root2_h(N), r2i = gl_inv_h(r2), ninv = gl_inv_h(N);
        if (!w || !r2) {
            set_error("no Goldilocks roots for this length");
            return LLH_ERR_UNSUPPORTED;
        }
        auto powers = [](u64 base, u64 count, u64 scale) {
            std::vector<u64> v(count);
            u64 a = scale;
            for (u64 j = 0; j < count; ++j) { v[j] = a; a = gl_mul_h(a, base); }
            return v;
        };
        const u64 wW = gl_pow_h(w, H), wH = gl_pow_h(w, W);   // w_W = w_M^H, w_H = w_M^W
        if ((rc = upload(&T.twW, pass_order(powers(wW, W, 1), 1, L.row_exp, L.row_tab))) ||
            (rc = upload(&T.twWi, pass_order(powers(gl_inv_h(wW), W, 1), 1, L.row_exp, L.row_tab))) ||
            (rc = upload(&T.twH, pass_order(powers(wH, H, 1), 1, L.col_exp, L.col_tab))) ||
            (rc = upload(&T.twHi, pass_order(powers(gl_inv_h(wH), H, 1), 1, L.col_exp, L.col_tab))) ||
            (rc = upload(&T.tmL, powers(w, S, 1))) || (rc = upload(&T.tmH, powers(gl_pow_h(w, S), nS, 1))) ||
            (rc = upload(&T.tmLi, powers(wv, S, 1))) || (rc = upload(&T.tmHi, powers(gl_pow_h(wv, S), nS, 1))) ||
            (rc = upload(&T.wfL, powers(r2, Sw, 1))) || (rc = upload(&T.wfH, powers(gl_pow_h(r2, Sw), nSw, 1))) ||
            (rc = upload(&T.wiL, powers(r2i, Sw, 1))) ||
            (rc = upload(&T.wiH, powers(gl_pow_h(r2i, Sw), nSw, ninv))))
            return rc;
    }
    g_large_tables[key] = T;
    out = T;
    return LLH_OK;
}



// Buffers of finished large-engine jobs, kept for the next job of the same size: a hipFree
// between jobs would wait for the work the other streams still have in flight.
struct BufCache {
    std::multimap<size_t, void *> dev, host;
    void *get(size_t bytes, bool pinned) {
        auto &m = pinned ? host : dev;
        auto it = m.find(bytes);
        if (it == m.end()) return nullptr;
        void *p = it->second;
        m.erase(it);
        return p;
    }
    void put(void *p, size_t bytes, bool pinned) {
        if (p) (pinned ? host : dev).insert({bytes, p});
    }
};
static BufCache g_bufs;
static int cached_malloc(void **p, size_t bytes) {
    if ((*p = g_bufs.get(bytes, false))) return LLH_OK;
    HIPCHK(hipMalloc(p, bytes));
    return LLH_OK;
}
static int cached_host_malloc(void **p, size_t bytes) {
    if ((*p = g_bufs.get(bytes, true))) return LLH_OK;
    HIPCHK(hipHostMalloc(p, bytes, hipHostMallocDefault));
    return LLH_OK;
}

// Result of the large engine's device normalisation (the NormOut at the start of its buffer).
struct NormOutH {
    unsigned long long res64;
    int zero;
    int status;
    int ones;
};

// One exponent on the large engine at length N, as a resumable job.  Squarings are queued in
// blocks of kGraphBlock (a captured hipGraph each), up to two blocks in flight, so the device
// never waits for the host's round-off check of the block before.  After each block the state
// (the element array and the pending row-segment carries) is copied into one of three
// checkpoint slots; a block whose round-off stays within 0.4 makes its checkpoint the last good
// one.  Above 0.4 the job stops with kStRoundoff, and rollback_state() turns the last good
// checkpoint into canonical digits (carries applied, normalised on the device) from which the
// next length resumes at that checkpoint's iteration.  Several jobs run at once on the
// library's streams: one exponent of this size leaves most of the GPU idle.
struct Resume {
    void *dig = nullptr;     // canonical digits of the checkpoint at length n (device, n i64)
    size_t bytes = 0;
    u64 n = 0, done = 0;     // its length and the squarings it has done
};


struct LargeJob {
    struct Block {
        u64 k;          // squarings in the block
        int e;          // round-off slot set (0 or 1)
        int ck;         // checkpoint slot written after it (-1: none)
        u64 done_after;
    };
    u64 p = 0, iters = 0, N = 0, done = 0, queued = 0;
    u64 nblocks = 0;                  // blocks queued so far
    int transform = kFP64;
    hipStream_t st = nullptr;
    LargeLaunch L;
    LargeArgs A;
    void *elem = nullptr, *carry = nullptr, *dig = nullptr, *nout = nullptr, *sync = nullptr;
    void *errs[2] = {nullptr, nullptr};      // per-squaring round-off slots of a block in flight
    double *eh[2] = {nullptr, nullptr};      // pinned: those slots + the fault word
    hipEvent_t evb[2] = {nullptr, nullptr};  // recorded after each in-flight block's read-back
    void *ck_elem[3] = {}, *ck_carry[3] = {};
    u64 ck_done[3] = {0, 0, 0};
    int ck_next = 0, ck_good = -1;
    NormOutH *noh = nullptr;                 // pinned
    size_t errs_bytes = 0, eh_bytes = 0;
    size_t bytes_elem = 0, bytes_carry = 0, bytes_dig = 0, bytes_sync = 0, bytes_nout = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::deque<Block> fl;                    // blocks in flight, oldest first
    std::vector<double> trace;
    BatchResult r = BatchResult();
    bool fault = false;
    u64 squarings = 0;                       // squarings queued on the device (kept or not)

    // 0 = started; kStNotRun in r.status = no kernels for this length.  alone: no other large
    // job runs beside this one (the fused column pass then orders its blocks by XCD; with
    // concurrent jobs the workgroups are not dealt to XCDs in block order).  from: resume from
    // these canonical digits (a checkpoint at another length) instead of s_0 = 4.
    int begin(u64 p_, u64 iters_, int transform_, u64 N_, hipStream_t st_, bool alone, const Resume *from) {
        p = p_; iters = iters_; transform = transform_; N = N_; st = st_;
        done = queued = nblocks = 0;
        ck_next = 0; ck_good = -1;
        fl.clear();
        fault = false;
        squarings = 0;
        r = BatchResult();
        r.status = kStOk;
        const u64 M = transform == kFP64 ? N / 2 : N;
        const int sched = N ? large_sched_of(transform, N) : 0;
        if (!N || !large_launch_for(transform, M, sched, L) || (N + 2) * (double)p >= 9.0e15) {
            r.status = kStNotRun;
            return LLH_OK;
        }
        LargeTables T;
        int rc = get_large_tables(transform, N, sched, T);
        if (rc) return rc;
        const u64 H = (u64)L.h_len;
        const int log_n = (N & (N - 1)) ? 0 : ceil_log2(N);
        const size_t esz = transform == kFP64 ? 16 : 8;
        const u64 nblk = (u64)large_blocks(L);
        errs_bytes = kGraphBlock * 8;
        eh_bytes = (kGraphBlock + 1) * 8;
        // sync: 8 unused bytes, the hand-off fault word, the fused pass's tagged carry words
        bytes_sync = 16 + nblk * H * 8;
        bytes_elem = M * esz;
        bytes_carry = nblk * H * 8;
        bytes_dig = N * 8;
        bytes_nout = large_normalize_bytes(N);
        if ((rc = cached_malloc(&elem, bytes_elem)) || (rc = cached_malloc(&carry, bytes_carry)) ||
            (rc = cached_malloc(&dig, bytes_dig)) || (rc = cached_malloc(&nout, bytes_nout)) ||
            (rc = cached_malloc(&sync, bytes_sync)) || (rc = cached_host_malloc((void **)&noh, sizeof(NormOutH))))
            return rc;
        for (int e = 0; e < 2; ++e)
            if ((rc = cached_malloc(&errs[e], errs_bytes)) || (rc = cached_host_malloc((void **)&eh[e], eh_bytes)))
                return rc;
        if (transform == kFP64 && g_ckpt_blocks.load() > 0)
            for (int c = 0; c < 3; ++c)
                if ((rc = cached_malloc(&ck_elem[c], bytes_elem)) || (rc = cached_malloc(&ck_carry[c], bytes_carry)))
                    return rc;
        HIPCHK(hipMemsetAsync(carry, 0, bytes_carry, st));
        HIPCHK(hipMemsetAsync(sync, 0, bytes_sync, st));
        if (from) {
            HIPCHK(large_relayout(transform, (const i64 *)from->dig, (u32)p, from->n, N, elem, (u32)L.w_len,
                                  (u32)L.h_len, st));
            done = queued = from->done;
        } else {
            HIPCHK(hipMemsetAsync(elem, 0, M * esz, st));
            if (transform == kFP64) {
                const double four = 4.0;
                HIPCHK(hipMemcpyAsync(elem, &four, 8, hipMemcpyHostToDevice, st));
            } else {
                const u64 four = 4;
                HIPCHK(hipMemcpyAsync(elem, &four, 8, hipMemcpyHostToDevice, st));
            }
        }
        A.elem = elem; A.carry = carry; A.err = (unsigned long long *)errs[0];
        A.twW = T.twW; A.twWi = T.twWi; A.twH = T.twH; A.twHi = T.twHi;
        A.tmL = T.tmL; A.tmH = T.tmH; A.tmLi = T.tmLi; A.tmHi = T.tmHi;
        A.wfL = T.wfL; A.wfH = T.wfH; A.wiL = T.wiL; A.wiH = T.wiH;
        A.p = (u32)p; A.log_n = (u32)log_n; A.log_s = T.log_s; A.log_sw = T.log_sw; A.err_slot = 0;
        A.n = N; A.inv_n = 1.0 / (double)N;
        A.p_mod_n = (u32)(p % N); A.p_div_n = (u32)(p / N);
        A.tag = 1;
        A.fault = (unsigned *)((char *)sync + 8);
        A.tcarry = (char *)sync + 16;
        A.xcd_blocks = alone ? 1u : 0u;
        A.w_len = (u32)L.w_len;
        A.h_len = (u32)L.h_len;
        trace.clear();
        trace.reserve(iters - done);
        for (int e = 0; e < 2; ++e) HIPCHK(hipEventCreateWithFlags(&evb[e], hipEventDisableTiming));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        HIPCHK(hipEventRecord(ev0, st));
        if (from) HIPCHK(hipStreamSynchronize(st));   // the caller may reuse from->dig afterwards
        return LLH_OK;
    }

    bool running() const { return r.status == kStOk && done < iters && !fault; }

    // queues the next block: its squarings, the read-back of their round-off, a checkpoint copy
    int launch_block() {
        const u64 k = std::min<u64>(kGraphBlock, iters - queued);
        const int e = (int)(nblocks % 2);
        LargeArgs Ab = A;
        Ab.err = (unsigned long long *)errs[e];
        // carry-word tags 1 + 256 e + i: consecutive fused launches differ, within a block by i and
        // across blocks by the parity e (only the last block is short, so no two short blocks meet),
        // and no tag is 0, the cleared words' tag
        Ab.tag = 1u + 256u * (u32)e;
        HIPCHK(hipMemsetAsync(errs[e], 0, errs_bytes, st));
        if (k == (u64)kGraphBlock) {
            if (!gexec[e]) {
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipError_t er = large_squarings(L, Ab, (int)k, st);
                if (er != hipSuccess) {
                    hipGraph_t g2;
                    (void)hipStreamEndCapture(st, &g2);
                    set_error(std::string("HIP: ") + hipGetErrorString(er) + " while capturing");
                    return LLH_ERR_HIP;
                }
                HIPCHK(hipStreamEndCapture(st, &graph[e]));
                HIPCHK(hipGraphInstantiate(&gexec[e], graph[e], nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(gexec[e], st));
        } else {
            HIPCHK(large_squarings(L, Ab, (int)k, st));
        }
        HIPCHK(hipMemcpyAsync(eh[e], errs[e], errs_bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&eh[e][kGraphBlock], A.fault, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(evb[e], st));
        queued += k;
        squarings += k;
        ++nblocks;
        int ck = -1;
        const int every = g_ckpt_blocks.load();
        if (ck_elem[0] && every > 0 && (nblocks % (u64)every == 0) && queued < iters) {
            ck = ck_next;
            ck_next = (ck_next + 1) % 3;
            HIPCHK(hipMemcpyAsync(ck_elem[ck], elem, bytes_elem, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(ck_carry[ck], carry, bytes_carry, hipMemcpyDeviceToDevice, st));
            ck_done[ck] = queued;
        }
        fl.push_back(Block{k, e, ck, queued});
        return LLH_OK;
    }

    // queues blocks until two are in flight (or every squaring is queued)
    int top_up() {
        while (fl.size() < 2 && queued < iters) {
            int rc = launch_block();
            if (rc) return rc;
        }
        return LLH_OK;
    }

    // waits for the oldest block in flight; round-off bookkeeping and the rollback / fault decisions
    int wait_oldest() {
        if (fl.empty()) return LLH_OK;
        const Block b = fl.front();
        fl.pop_front();
        HIPCHK(hipEventSynchronize(evb[b.e]));
        unsigned fw;
        std::memcpy(&fw, &eh[b.e][kGraphBlock], sizeof fw);
        if (fw) {
            fault = true;
            r.status = kStNotRun;
            return LLH_OK;
        }
        double bmax = 0;
        for (u64 i = 0; i < b.k; ++i) {
            trace.push_back(eh[b.e][i]);
            bmax = std::max(bmax, eh[b.e][i]);
        }
        r.maxerr = std::max(r.maxerr, bmax);
        if (transform == kFP64 && bmax > 0.4) {
            r.status = kStRoundoff;
            return LLH_OK;
        }
        done = b.done_after;
        if (b.ck >= 0) ck_good = b.ck;
        return LLH_OK;
    }

    // After kStRoundoff: the canonical digits of the last good checkpoint, normalised on the
    // device (false: no checkpoint, or its normalisation did not settle; restart from s_0).  On
    // success the job's digit buffer moves to `out`.
    int rollback_state(Resume &out, bool &ok) {
        ok = false;
        HIPCHK(hipStreamSynchronize(st));            // the blocks still in flight end
        if (ck_good < 0 || fault) return LLH_OK;
        LargeArgs Ac = A;
        Ac.elem = ck_elem[ck_good];
        Ac.carry = ck_carry[ck_good];
        noh->status = 1;
        HIPCHK(large_apply_carry(L, Ac, st));
        HIPCHK(large_normalize(transform, Ac.elem, (i64 *)dig, Ac, nout, st));
        HIPCHK(hipMemcpyAsync(noh, nout, sizeof(NormOutH), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (noh->status) return LLH_OK;
        out.dig = dig;
        out.bytes = bytes_dig;
        out.n = N;
        out.done = ck_done[ck_good];
        dig = nullptr;
        ok = true;
        return LLH_OK;
    }

    int finish() {
        if (!ev0) {                       // never started (no kernels for this length)
            release();
            return LLH_OK;
        }
        noh->res64 = 0; noh->zero = 0; noh->status = 0;
        if (r.status == kStOk) {
            HIPCHK(large_apply_carry(L, A, st));
            HIPCHK(large_normalize(transform, elem, (i64 *)dig, A, nout, st));
            HIPCHK(hipMemcpyAsync(noh, nout, sizeof(NormOutH), hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipStreamSynchronize(st));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        r.iters_done = done;
        r.t0 = 0;
        r.t1 = (u64)((double)ms * 1e-3 * g_dev.wclk_hz);
        if (r.status == kStOk) {
            if (noh->status) r.status = kStNotRun;
            r.res64 = noh->res64;
            r.verdict = noh->zero;
        }
        release();
        if (fault) {
            set_error("large engine: a carry hand-off between workgroups failed (a wait timed out, or a carry "
                      "outgrew its 48-bit word)");
            return LLH_ERR_HIP;
        }
        return LLH_OK;
    }

    void release() {
        for (int e = 0; e < 2; ++e) {
            if (gexec[e]) (void)hipGraphExecDestroy(gexec[e]);
            if (graph[e]) (void)hipGraphDestroy(graph[e]);
            if (evb[e]) (void)hipEventDestroy(evb[e]);
            g_bufs.put(errs[e], errs_bytes, false);
            g_bufs.put(eh[e], eh_bytes, true);
            gexec[e] = nullptr; graph[e] = nullptr; evb[e] = nullptr; errs[e] = nullptr; eh[e] = nullptr;
        }
        for (int c = 0; c < 3; ++c) {
            g_bufs.put(ck_elem[c], bytes_elem, false);
            g_bufs.put(ck_carry[c], bytes_carry, false);
            ck_elem[c] = ck_carry[c] = nullptr;
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        g_bufs.put(elem, bytes_elem, false);
        g_bufs.put(carry, bytes_carry, false);
        g_bufs.put(dig, bytes_dig, false);
        g_bufs.put(nout, bytes_nout, false);
        g_bufs.put(sync, bytes_sync, false);
        g_bufs.put(noh, sizeof(NormOutH), true);
        ev0 = ev1 = nullptr;
        elem = carry = dig = nout = sync = nullptr;
        noh = nullptr;
        fl.clear();
    }
    ~LargeJob() { release(); }
};

// Lengths the large engine runs, ascending.  FP64: {1, 3, 5, 7, 15} * 2^k for N in
// [2^15, 2^22]; GL: powers of two 2^15 .. 2^21.
static std::vector<u64> large_lengths(int transform) {
    std::vector<u64> v;
    LargeLaunch L;
    for (u64 m : {1ull, 3ull, 5ull, 7ull, 15ull})
        for (int k = 10; k <= 22; ++k) {
            const u64 N = m << k;
            const u64 M = transform == kFP64 ? N / 2 : N;
            if (large_launch_for(transform, M, 0, L)) v.push_back(N);
        }
    std::sort(v.begin(), v.end());
    return v;
}

// Default large-engine length for p.  FP64: the reference's conservative rule
// (lucaslehmer.cpp:244-254: smallest {1,3,5} * 2^k with p/N <= bits-per-word cap), raised to the
// smallest large-engine length at or above it (the large engine starts at N = 2^16, while the
// rule gives 12288 .. 49152 for 172,032 < p <= 884,736, beyond the batched lengths' cap);
// GL: the smallest power of two within the exactness bound (cuda/ll_ntt.cu:172-175).
static u64 large_length(u64 p, int transform) {
    if (transform == kFP64) {
        const u64 N = fft_length_conservative(p, g_bpw.load());
        if (!N) return 0;
        for (u64 n : large_lengths(transform))
            if (n >= N) return n;
        return 0;
    }
    for (u64 N : large_lengths(transform))
        if ((p + N - 1) / N <= (u64)ntt_max_bits(N)) return N;
    return 0;
}

// the next supported length above N (0 = none): the rollback target
static u64 next_large_length(u64 N, int transform) {
    for (u64 n : large_lengths(transform))
        if (n > N) return n;
    return 0;
}

static const int kLargeJobs = 4;

static int run_large_list(std::vector<Pending> &large, int transform, bool no_rollback, int32_t *verdict,
                          uint64_t *res64, int32_t *status, uint32_t *n_used, std::vector<double> &secs,
                          std::vector<double> &errs) {
    std::sort(large.begin(), large.end(), [](const Pending &a, const Pending &b) {
        return (double)a.iters * (double)a.n > (double)b.iters * (double)b.n;
    });
    size_t next = 0;
    struct Slot {
        LargeJob job;
        size_t item = 0;
        bool busy = false;
    };
    std::vector<Slot> slots(std::min<size_t>(kLargeJobs, large.size()));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // Every way out (an error included) first waits for the slots' streams, so no buffer goes
    // back to the cache, and no graph is destroyed, while kernels queued on it may still run;
    // then the events go.  (Declared after `slots`, so it runs before the slots are destroyed.)
    struct Drain {
        size_t n;
        hipEvent_t &a, &b;
        ~Drain() {
            for (size_t i = 0; i < n; ++i) (void)hipStreamSynchronize(g_dev.fstream[i]);
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } drain{slots.size(), e0, e1};
    int rc;
    // the last exponent to finish leaves its round-off trace (a single exponent: its own)
    auto record = [&](Slot &s) {
        const Pending &w = large[s.item];
        const BatchResult &r = s.job.r;
        secs[w.idx] += (double)(r.t1 - r.t0) / g_dev.wclk_hz;
        errs[w.idx] = std::max(errs[w.idx], r.maxerr);
        verdict[w.idx] = r.verdict;
        res64[w.idx] = r.res64;
        status[w.idx] = r.status;
        if (n_used) n_used[w.idx] = (uint32_t)s.job.N;
        g_err_trace = s.job.trace;
    };
    // start job for item i in slot s at length N (queues its first blocks), from s_0 or a checkpoint
    auto start = [&](Slot &s, size_t i, u64 N, const Resume *from) -> int {
        s.item = i;
        s.busy = true;
        int r2 = s.job.begin(large[i].p, large[i].iters, transform, N, g_dev.fstream[&s - slots.data()],
                             slots.size() == 1, from);
        if (r2) return r2;
        if (s.job.running()) return s.job.top_up();
        return LLH_OK;
    };
    auto fill = [&](Slot &s) -> int {
        while (next < large.size()) {
            const size_t i = next++;
            if (large[i].n == 0) {
                status[large[i].idx] = kStNotRun;
                continue;
            }
            return start(s, i, large[i].n, nullptr);
        }
        s.busy = false;
        return LLH_OK;
    };
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, g_dev.stream));
    for (Slot &s : slots) {
        HIPCHK(hipStreamWaitEvent(g_dev.fstream[&s - slots.data()], e0, 0));
        if ((rc = fill(s))) return rc;
    }
    for (bool any = true; any;) {
        any = false;
        for (Slot &s : slots) {
            if (!s.busy) continue;
            any = true;
            LargeJob &j = s.job;
            if (j.running()) {
                if ((rc = j.wait_oldest())) return rc;
                if (j.running()) {
                    if ((rc = j.top_up())) return rc;
                    continue;
                }
            }
            // the job has stopped: finished, rolled back or not run
            g_last_squarings += j.squarings;
            j.squarings = 0;
            if (j.r.status == kStRoundoff && transform == kFP64 && !no_rollback) {
                const u64 nx = next_large_length(j.N, transform);   // rollback: resume longer
                if (nx) {
                    Resume from;
                    bool ok = false;
                    if ((rc = j.rollback_state(from, ok))) return rc;
                    if ((rc = j.finish())) return rc;
                    const Pending &w = large[s.item];
                    secs[w.idx] += (double)(j.r.t1 - j.r.t0) / g_dev.wclk_hz;
                    errs[w.idx] = std::max(errs[w.idx], j.r.maxerr);
                    g_last_rollbacks += 1;
                    if (ok) {
                        g_last_resumed += 1;
                        g_last_resumed_at = from.done;
                    }
                    rc = start(s, s.item, nx, ok ? &from : nullptr);
                    g_bufs.put(from.dig, from.bytes, false);
                    if (rc) return rc;
                    continue;
                }
            }
            if ((rc = j.finish())) return rc;
            record(s);
            if ((rc = fill(s))) return rc;
        }
    }
    // device time of the whole list: from e0 to the last stream's end
    for (size_t i = 1; i < slots.size(); ++i) {
        HIPCHK(hipEventRecord(e1, g_dev.fstream[i]));
        HIPCHK(hipStreamWaitEvent(g_dev.stream, e1, 0));
    }
    HIPCHK(hipEventRecord(e1, g_dev.stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (!slots.empty()) {
        g_last_kernel_ms += ms;
        g_last_launches += 1;
    }
    return LLH_OK;
}

// Re-splits the checkpoints of the rolled-back exponents at their next lengths (device kernel
// bt_resume) into resume entries; the next launch starts each from its entry and its iteration.
// A checkpoint whose normalisation does not settle starts again from s_0.
static int resume_batch(int transform, std::vector<Pending> &again) {
    std::vector<ResumeReq> reqs;
    std::vector<size_t> which;
    for (size_t k = 0; k < again.size(); ++k)
        if (again[k].rb >= 0) {
            reqs.push_back(ResumeReq{(u32)again[k].p, (u32)again[k].n_old, (u32)again[k].n, (u32)again[k].rb});
            which.push_back(k);
        }
    if (reqs.empty()) return LLH_OK;
    int rc;
    const size_t nr = reqs.size();
    if ((rc = grow(&g_dev.res, g_dev.cap_res, (size_t)kRollbackCap * kResumeStride * 8)) ||
        (rc = grow(&g_dev.req, g_dev.cap_req, nr * (sizeof(ResumeReq) + 4) + 64)))
        return rc;
    hipStream_t st = g_dev.stream;
    ResumeReq *d_req = (ResumeReq *)g_dev.req;
    int *d_ok = (int *)((char *)g_dev.req + ((nr * sizeof(ResumeReq) + 63) / 64) * 64);
    std::vector<int> ok(nr, 0);
    HIPCHK(hipMemcpyAsync(d_req, reqs.data(), nr * sizeof(ResumeReq), hipMemcpyHostToDevice, st));
    HIPCHK(batch_resume(transform, d_req, (int)nr, (i64 *)g_dev.rb, (i64 *)g_dev.res, d_ok, st));
    HIPCHK(hipMemcpyAsync(ok.data(), d_ok, nr * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (size_t e = 0; e < nr; ++e) {
        Pending &w = again[which[e]];
        if (ok[e]) {
            w.init = (u32)e + 1;
            g_last_resumed += 1;
            g_last_resumed_at = w.it0;
        } else {
            w.it0 = 0;
        }
        w.rb = -1;
    }
    return LLH_OK;
}

static int batch_impl(const u64 *ps, size_t n, u64 max_iters, int transform, int32_t *verdict,
                      uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used, int32_t *status,
                      u64 force_n = 0, bool no_rollback = false, const u64 *force_each = nullptr,
                      bool noalloc = false) {
    if (transform != kFP64 && transform != kNTT) {
        set_error("transform must be LLH_FP64 or LLH_NTT");
        return LLH_ERR_ARG;
    }
    if (n && (!ps || !verdict || !res64 || !status)) {
        set_error("NULL argument");
        return LLH_ERR_ARG;
    }
    if (noalloc && n > kScratchItems) {
        set_error("llh_batch_noalloc: at most 16384 exponents per call");
        return LLH_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    int rc = ensure_init();
    if (rc) return rc;
    g_last_kernel_ms = 0.0;
    g_last_launches = 0;
    g_last_squarings = 0;
    g_last_rollbacks = 0;
    g_last_resumed = 0;
    g_last_resumed_at = 0;
    std::vector<Pending> work, large;
    std::vector<double> secs(n, 0.0), errs(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const u64 p = ps[i];
        verdict[i] = 0;
        res64[i] = 0;
        status[i] = kStTrivial;
        if (n_used) n_used[i] = 0;
        if (p < 2) continue;                                   // lucaslehmer.cpp:30
        if (p == 2) { verdict[i] = 1; continue; }              // :31 (M_2 = 3 is prime)
        if (p % 2 == 0) continue;                              // :32
        if (p >= ((u64)1 << 31)) { status[i] = kStNotRun; continue; }
        const u64 it = max_iters ? std::min<u64>(max_iters, p - 2) : p - 2;
        if (p < 64) {                                          // the one-word lane
            work.push_back(Pending{p, it, i, 0, 1});
            continue;
        }
        const u64 fn = force_each && force_each[i] ? force_each[i] : force_n;
        if (fn) {
            if (fn < 16 || fn > ((u64)1 << 24)) {
                set_error("forced length must be in [16, 2^24]");
                return LLH_ERR_ARG;
            }
            const int c = batch_class_of(transform, fn);
            if (c && batch_fits(p, fn))
                work.push_back(Pending{p, it, i, c, fn});
            else
                large.push_back(Pending{p, it, i, -1, fn});
            continue;
        }
        const u64 bn = batch_length(p, transform);
        if (bn == 0) {
            large.push_back(Pending{p, it, i, -1, large_length(p, transform)});
            continue;
        }
        work.push_back(Pending{p, it, i, batch_class_of(transform, bn), bn});
    }
    if (noalloc && !large.empty()) {
        set_error("llh_batch_noalloc: p = " + std::to_string(large[0].p) +
                  " is beyond the batched engine (its device memory is allocated per call)");
        return LLH_ERR_UNSUPPORTED;
    }
    // exponents beyond the batched engine: the large engine, up to kLargeJobs at once (one per
    // stream), longest first; a job that rolls back restarts at the next length in its slot
    if ((rc = run_large_list(large, transform, no_rollback, verdict, res64, status, n_used, secs, errs))) return rc;
    while (!work.empty()) {
        std::vector<BatchResult> res;
        rc = run_launch(transform, work, res, noalloc);
        if (rc) return rc;
        std::vector<Pending> again;
        for (size_t k = 0; k < work.size(); ++k) {
            const Pending &w = work[k];
            const BatchResult &r = res[k];
            secs[w.idx] += (double)(r.t1 - r.t0) / g_dev.wclk_hz;
            errs[w.idx] = std::max(errs[w.idx], r.maxerr);
            if (r.iters_done > w.it0) g_last_squarings += r.iters_done - w.it0;
            if (r.status == kStRoundoff && w.cls > 0 && !no_rollback) {
                const u64 nx = next_batch_length(w.p, w.n, transform);   // rollback: continue longer
                if (nx) {
                    Pending np{w.p, w.iters, w.idx, batch_class_of(transform, nx), nx};
                    if (r.ck && nx <= kResumeStride) {        // from the last good checkpoint
                        np.rb = (i64)r.ck - 1;
                        np.it0 = r.ck_iter;
                        np.n_old = w.n;
                    }
                    again.push_back(np);
                    g_last_rollbacks += 1;
                    continue;
                }
            }
            verdict[w.idx] = r.verdict;
            res64[w.idx] = r.res64;
            status[w.idx] = r.status;
            if (n_used) n_used[w.idx] = (uint32_t)w.n;
        }
        if ((rc = resume_batch(transform, again))) return rc;
        work.swap(again);
    }
    for (size_t i = 0; i < n; ++i) {
        if (seconds) seconds[i] = secs[i];
        if (maxerr) maxerr[i] = errs[i];
    }
    return LLH_OK;
}

}  // namespace llh

using namespace llh;

extern "C" {

int llh_init(int device) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    return init_locked(device);
}

const char *llh_last_error(void) { return g_err.c_str(); }

int llh_device_info(char *name, size_t name_len, int *compute_units, uint64_t *global_mem_bytes, char *arch,
                    size_t arch_len) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    int rc = ensure_init();
    if (rc) return rc;
    hipDeviceProp_t pr;
    HIPCHK(hipGetDeviceProperties(&pr, g_dev.id));
    if (name && name_len) {
        std::strncpy(name, pr.name, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (arch && arch_len) {
        std::strncpy(arch, pr.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    if (compute_units) *compute_units = pr.multiProcessorCount;
    if (global_mem_bytes) *global_mem_bytes = (uint64_t)pr.totalGlobalMem;
    return LLH_OK;
}

uint64_t llh_fft_length(uint64_t p, double max_bpw) {
    return fft_length_conservative(p, max_bpw > 0 ? max_bpw : g_bpw.load());
}

uint64_t llh_batch_length(uint64_t p, int transform) {
    if (p < 64) return p >= 3 ? 1 : 0;
    return batch_length(p, transform);
}

uint64_t llh_ntt_length(uint64_t p) { return ntt_length_ref(p); }

uint64_t llh_fp64_length(uint64_t p, double max_bpw) {
    if (p < 64) return 0;
    if (!(max_bpw > 0)) max_bpw = g_bpw.load();
    // the reference's length (lucaslehmer.cpp:244-254) when an engine runs it, else the shortest
    // engine length above it (every such length keeps p / N <= max_bpw)
    const u64 c = fft_length_conservative(p, max_bpw);
    if (!c) return 0;
    for (const auto &e : batch_ladder(kFP64))
        if (e.first >= c && batch_fits(p, e.first)) return e.first;
    for (u64 N : large_lengths(kFP64))
        if (N >= c) return N;
    return 0;
}

uint64_t llh_max_exponent(int transform) {
    // NTT: the reference's bound, 8192 * max_bits_per_word(8192) (cuda/ll_ntt.cu:279)
    if (transform == kNTT) return (uint64_t)8192 * (uint64_t)ntt_max_bits(8192);
    return (uint64_t)std::floor(g_batch_bpw.load() * (double)((uint64_t)1 << kFp64MaxLog));
}

uint64_t llh_batch_max_exponent(int transform) {
    if (transform == kNTT) return ((uint64_t)1 << kNttMaxLog) * (uint64_t)ntt_max_bits((uint64_t)1 << kNttMaxLog);
    return (uint64_t)std::floor(g_batch_bpw.load() * (double)((uint64_t)1 << kFp64MaxLog));
}

int llh_set_fft_bits_per_word(double bits) {
    if (!(bits > 0) || bits > 64) {
        set_error("bits per word must be in (0, 64]");
        return LLH_ERR_ARG;
    }
    g_bpw.store(bits);
    return LLH_OK;
}

double llh_get_fft_bits_per_word(void) { return g_bpw.load(); }

int llh_set_batch_bits_per_word(double bits) {
    if (!(bits > 0) || bits > 64) {
        set_error("bits per word must be in (0, 64]");
        return LLH_ERR_ARG;
    }
    g_batch_bpw.store(bits);
    return LLH_OK;
}

double llh_get_batch_bits_per_word(void) { return g_batch_bpw.load(); }

int llh_set_batch_families(int mask) {
    if (mask & ~0xF) {
        set_error("family mask has bits 0..3 only");
        return LLH_ERR_ARG;
    }
    g_batch_families.store(mask | 1);
    return LLH_OK;
}

int llh_get_batch_families(void) { return g_batch_families.load(); }

int llh_set_batch_cost(uint64_t n, int transform, double us) {
    if ((transform != kFP64 && transform != kNTT) || !batch_class_of(transform, n) || !(us >= 0)) {
        set_error("llh_set_batch_cost: not a batched length, or a negative cost");
        return LLH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_cost_mx);
    if (us == 0.0) batch_costs().erase({transform, n});
    else batch_costs()[{transform, n}] = us;
    return LLH_OK;
}

double llh_get_batch_cost(uint64_t n, int transform) { return batch_cost(transform, n); }

int llh_batch_n(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, const uint64_t *n_force, int flags,
                int32_t *verdict, uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used,
                int32_t *status) {
    try {
        return batch_impl(ps, n, max_iters, transform, verdict, res64, seconds, maxerr, n_used, status, 0,
                          (flags & 1) != 0, n_force);
    } catch (const std::exception &e) {
        set_error(e.what());
        return LLH_ERR_HIP;
    }
}

int llh_batch_noalloc(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, int32_t *verdict,
                      uint64_t *res64, double *seconds, double *maxerr, uint32_t *n_used, int32_t *status) {
    try {
        return batch_impl(ps, n, max_iters, transform, verdict, res64, seconds, maxerr, n_used, status, 0, false,
                          nullptr, true);
    } catch (const std::exception &e) {
        set_error(e.what());
        return LLH_ERR_HIP;
    }
}

int llh_batch(const uint64_t *ps, size_t n, uint64_t max_iters, int transform, int32_t *verdict, uint64_t *res64,
              double *seconds, double *maxerr, uint32_t *n_used, int32_t *status) {
    try {
        return batch_impl(ps, n, max_iters, transform, verdict, res64, seconds, maxerr, n_used, status);
    } catch (const std::exception &e) {
        set_error(e.what());
        return LLH_ERR_HIP;
    }
}

int llh_lucas_lehmer_n(uint64_t p, uint64_t max_iters, int transform, uint64_t n_force, int flags,
                       int32_t *verdict, uint64_t *res64, double *maxerr, uint32_t *n_used, int32_t *status,
                       double *seconds) {
    if (n_force && (n_force < 16 || n_force > ((uint64_t)1 << 24))) {
        set_error("n_force must be in [16, 2^24]");
        return LLH_ERR_ARG;
    }
    double sec = 0, err = 0;
    uint32_t nu = 0;
    int32_t v = 0, st = 0;
    uint64_t r = 0;
    int rc;
    try {
        rc = batch_impl(&p, 1, max_iters, transform, &v, &r, &sec, &err, &nu, &st, n_force, (flags & 1) != 0);
    } catch (const std::exception &e) {
        set_error(e.what());
        return LLH_ERR_HIP;
    }
    if (rc) return rc;
    if (verdict) *verdict = v;
    if (res64) *res64 = r;
    if (maxerr) *maxerr = err;
    if (n_used) *n_used = nu;
    if (status) *status = st;
    if (seconds) *seconds = sec;
    return LLH_OK;
}

size_t llh_supported_lengths(int transform, uint64_t *out, size_t cap) {
    std::vector<u64> v;
    for (const auto &e : batch_ladder(transform)) v.push_back(e.first);
    for (u64 n : large_lengths(transform))
        if (std::find(v.begin(), v.end(), n) == v.end()) v.push_back(n);
    std::sort(v.begin(), v.end());
    if (out)
        for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
    return v.size();
}

int llh_batch_schedules(int transform, uint64_t n, int32_t *radices, int cap) {
    if (transform != kFP64 && transform != kNTT) {
        set_error("llh_batch_schedules: unknown transform");
        return LLH_ERR_ARG;
    }
    int b = 0;
    for (const auto &e : batch_ladder(transform))
        if (e.first == n) b = e.second;
    const std::vector<int> v = batch_schedules_of(transform, b);
    for (int i = 0; i < (int)v.size() && i < cap; ++i)
        if (radices) batch_class_radices(v[i], radices + 4 * i);
    return (int)v.size();
}

int llh_set_batch_schedule(int transform, uint64_t n, int sched) {
    const int count = llh_batch_schedules(transform, n, nullptr, 0);
    if (count < 0) return count;
    if (count == 0 || sched < -1 || sched >= count) {
        set_error("llh_set_batch_schedule: no such schedule for this length");
        return LLH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_bsched_mx);
    if (sched < 0) g_batch_sched.erase(std::make_pair(transform, (u64)n));
    else g_batch_sched[std::make_pair(transform, (u64)n)] = sched;
    return LLH_OK;
}

int llh_get_batch_schedule(int transform, uint64_t n) {
    const int count = llh_batch_schedules(transform, n, nullptr, 0);
    if (count <= 0) {
        if (count == 0) set_error("llh_get_batch_schedule: not a batched length");
        return count < 0 ? count : LLH_ERR_ARG;
    }
    const int i = batch_sched_index(transform, n);
    return i >= 0 && i < count ? i : 0;
}

int llh_large_schedules(int transform, uint64_t n, int32_t *w, int32_t *h, int32_t *cw, int cap) {
    if (transform != kFP64 && transform != kNTT) {
        set_error("llh_large_schedules: unknown transform");
        return LLH_ERR_ARG;
    }
    const u64 M = transform == kFP64 ? n / 2 : n;
    const std::vector<LargeSched> v = large_schedules(transform, M);
    for (int i = 0; i < (int)v.size() && i < cap; ++i) {
        if (w) w[i] = v[i].w;
        if (h) h[i] = v[i].h;
        if (cw) cw[i] = v[i].cw;
    }
    return (int)v.size();
}

int llh_set_large_schedule(int transform, uint64_t n, int sched) {
    if (transform != kFP64 && transform != kNTT) {
        set_error("llh_set_large_schedule: unknown transform");
        return LLH_ERR_ARG;
    }
    const u64 M = transform == kFP64 ? n / 2 : n;
    const int count = (int)large_schedules(transform, M).size();
    if (sched < -1 || sched >= count) {
        set_error("llh_set_large_schedule: no such schedule for this length");
        return LLH_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    if (sched < 0) g_large_sched.erase(std::make_pair(transform, (u64)n));
    else g_large_sched[std::make_pair(transform, (u64)n)] = sched;
    return LLH_OK;
}

int llh_get_large_schedule(int transform, uint64_t n) {
    const u64 M = transform == kFP64 ? n / 2 : n;
    if ((transform != kFP64 && transform != kNTT) || large_schedules(transform, M).empty()) {
        set_error("llh_get_large_schedule: not a large-engine length");
        return LLH_ERR_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    return large_sched_of(transform, n);
}

size_t llh_last_roundoff_trace(double *out, size_t cap) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    const size_t n = g_err_trace.size();
    if (out)
        for (size_t i = 0; i < n && i < cap; ++i) out[i] = g_err_trace[i];
    return n;
}

int llh_device_copy(uint64_t bytes, int repeats, double *ms) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    int rc = ensure_init();
    if (rc) return rc;
    if (bytes == 0 || repeats < 1) {
        set_error("device_copy: bytes and repeats must be positive");
        return LLH_ERR_ARG;
    }
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    HIPCHK(hipMalloc(&b, bytes));
    hipStream_t st = g_dev.stream;
    HIPCHK(hipMemsetAsync(a, 1, bytes, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    for (int i = 0; i < repeats; ++i) HIPCHK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipStreamSynchronize(st));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return LLH_OK;
}

int llh_mlp_param_count(void) { return mlp_n_params(); }

// shared set-up of the MLP entry points: buffers sized for n samples, parameters and data
static int mlp_setup(MlpBuffers &B, const float *params, const float *X, const float *y, size_t n) {
    std::memset(&B, 0, sizeof(B));
    int rc = ensure_init();
    if (rc) return rc;
    if (!params || !X || n == 0 || n > (1u << 24)) {
        set_error("mlp: bad arguments");
        return LLH_ERR_ARG;
    }
    HIPCHK(mlp_alloc(B, n));
    hipStream_t st = g_dev.stream;
    HIPCHK(hipMemcpyAsync(B.P, params, mlp_n_params() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.X, X, 3 * n * 4, hipMemcpyHostToDevice, st));
    if (y) HIPCHK(hipMemcpyAsync(B.T, y, n * 4, hipMemcpyHostToDevice, st));
    return LLH_OK;
}

int llh_mlp_grad(const float *params, const float *X, const float *y, size_t n, float *grad, float *loss) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    MlpBuffers B;
    int rc = mlp_setup(B, params, X, y, n);
    if (rc) { mlp_free(B); return rc; }
    if (!y || !grad) { mlp_free(B); set_error("mlp: NULL output"); return LLH_ERR_ARG; }
    hipStream_t st = g_dev.stream;
    float l = 0;
    HIPCHK(mlp_forward(B, (int)n, true, st));
    HIPCHK(mlp_gradient(B, (int)n, st));
    HIPCHK(hipMemcpyAsync(grad, B.G, mlp_n_params() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&l, B.loss, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (loss) *loss = l;
    mlp_free(B);
    return LLH_OK;
}

int llh_mlp_train(float *params, const float *X, const float *y, size_t n, int epochs, float lr, float *loss_hist) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    MlpBuffers B;
    int rc = mlp_setup(B, params, X, y, n);
    if (rc) { mlp_free(B); return rc; }
    if (!y || epochs < 1) { mlp_free(B); set_error("mlp: bad arguments"); return LLH_ERR_ARG; }
    hipStream_t st = g_dev.stream;
    HIPCHK(hipMemsetAsync(B.Mm, 0, mlp_n_params() * 4, st));
    HIPCHK(hipMemsetAsync(B.Vv, 0, mlp_n_params() * 4, st));
    std::vector<float> hist(loss_hist ? epochs : 0);
    for (int t = 1; t <= epochs; ++t) {
        HIPCHK(mlp_forward(B, (int)n, true, st));
        HIPCHK(mlp_gradient(B, (int)n, st));
        if (loss_hist) HIPCHK(hipMemcpyAsync(&hist[t - 1], B.loss, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(mlp_step(B, lr, t, st));
    }
    HIPCHK(hipMemcpyAsync(params, B.P, mlp_n_params() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (loss_hist) std::memcpy(loss_hist, hist.data(), epochs * 4);
    mlp_free(B);
    return LLH_OK;
}

int llh_mlp_predict(const float *params, const float *X, size_t n, float *y_out) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    MlpBuffers B;
    int rc = mlp_setup(B, params, X, nullptr, n);
    if (rc) { mlp_free(B); return rc; }
    if (!y_out) { mlp_free(B); set_error("mlp: NULL output"); return LLH_ERR_ARG; }
    hipStream_t st = g_dev.stream;
    HIPCHK(mlp_forward(B, (int)n, false, st));
    HIPCHK(hipMemcpyAsync(y_out, B.Y, n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    mlp_free(B);
    return LLH_OK;
}

int llh_last_run_stats(uint64_t *squarings, uint32_t *rollbacks, uint32_t *resumed, uint64_t *resumed_at) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    if (squarings) *squarings = g_last_squarings;
    if (rollbacks) *rollbacks = g_last_rollbacks;
    if (resumed) *resumed = g_last_resumed;
    if (resumed_at) *resumed_at = g_last_resumed_at;
    return LLH_OK;
}

int llh_set_checkpoint_blocks(int blocks) {
    if (blocks < 0) {
        set_error("checkpoint interval must be >= 0 blocks");
        return LLH_ERR_ARG;
    }
    g_ckpt_blocks.store(blocks);
    return LLH_OK;
}

int llh_get_checkpoint_blocks(void) { return g_ckpt_blocks.load(); }

int llh_set_batch_checkpoints(int on) {
    g_batch_ckpt.store(on ? 1 : 0);
    return LLH_OK;
}

int llh_get_batch_checkpoints(void) { return g_batch_ckpt.load(); }

int llh_relayout_host(uint64_t p, uint64_t n_from, const int64_t *from, uint64_t n_to, int64_t *to) {
    if (!from || !to || p < 3 || n_from < 1 || n_to < 1 || n_from > p || n_to > p || n_from > ((u64)1 << 24) ||
        n_to > ((u64)1 << 24) || p >= ((u64)1 << 31)) {
        set_error("llh_relayout_host: bad arguments");
        return LLH_ERR_ARG;
    }
    relayout_host(from, p, n_from, n_to, to);
    return LLH_OK;
}

int llh_last_launch_stats(double *kernel_ms, uint32_t *launches) {
    std::lock_guard<std::recursive_mutex> lk(g_mx);
    if (kernel_ms) *kernel_ms = g_last_kernel_ms;
    if (launches) *launches = g_last_launches;
    return LLH_OK;
}

int llh_lucas_lehmer(uint64_t p, uint64_t max_iters, int transform, int32_t *verdict, uint64_t *res64,
                     double *maxerr, uint32_t *n_used, int32_t *status) {
    double sec = 0, err = 0;
    uint32_t nu = 0;
    int32_t v = 0, st = 0;
    uint64_t r = 0;
    int rc = llh_batch(&p, 1, max_iters, transform, &v, &r, &sec, &err, &nu, &st);
    if (rc) return rc;
    if (verdict) *verdict = v;
    if (res64) *res64 = r;
    if (maxerr) *maxerr = err;
    if (n_used) *n_used = nu;
    if (status) *status = st;
    return LLH_OK;
}

}  // extern "C"
