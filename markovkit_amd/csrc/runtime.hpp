// Host runtime shared by the library's translation units: status and error mapping, device
// buffers, the library stream, the per-kernel-class profiler, operator upload in the coded
// sliced or plain layout, and the handle types behind the C ABI.
#pragma once

#include "../../include/markovkit_amd.h"
#include "kernels.hpp"
#include "setup.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <tuple>
#include <string>
#include <vector>

namespace mk::rt {

using mk::i64;
using mk::u32;
using mk::u64;
namespace gpu = mk::gpu;

struct device_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline std::string &last_error() {
    thread_local std::string e;
    return e;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        const hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                                     \
            throw ::mk::rt::device_error(std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while (0)

inline void check_launch(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw device_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
mk_status guarded(F &&f) {
    try {
        f();
        return MK_OK;
    } catch (const std::invalid_argument &e) {
        last_error() = e.what();
        return MK_ERR_INVALID_ARGUMENT;
    } catch (const std::overflow_error &e) {
        last_error() = e.what();
        return MK_ERR_OVERFLOW;
    } catch (const device_error &e) {
        last_error() = e.what();
        return MK_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        last_error() = "out of host memory";
        return MK_ERR_RUNTIME;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return MK_ERR_RUNTIME;
    }
}

inline void require(bool ok, const char *what) {
    if (!ok) throw std::invalid_argument(what);
}

// ------------------------------------------------------------------ device memory
struct dbuf {
    void *p = nullptr;
    std::size_t bytes = 0;
    dbuf() = default;
    explicit dbuf(std::size_t n) : bytes(n) {
        if (n) HIPCHK(hipMalloc(&p, n));
    }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    dbuf &operator=(dbuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

template <typename T>
T *dp(const dbuf &b) {
    return static_cast<T *>(b.p);
}

template <typename T>
dbuf upload(const std::vector<T> &v) {
    dbuf b(v.size() * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return b;
}

// ------------------------------------------------------------------ library state
// the library stream of the calling host thread (each thread gets its own, default priority,
// so solves issued from several threads overlap on the device); streams live until the process
// ends
inline hipStream_t stream() {
    thread_local hipStream_t s = [] {
        hipStream_t t = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        // tuning override of the operator kernels' XCD interleave chunk (0 = identity mapping)
        if (const char *env = std::getenv("MK_XCD_CHUNK")) gpu::set_xcd_chunk(std::atoi(env));
        return t;
    }();
    return s;
}

enum kclass {
    K_SPMV,
    K_FIRST,
    K_SWEEP,
    K_RESIDUAL,
    K_RESTRICT,
    K_PROLONG,
    K_TERMINAL,
    K_DOT,
    K_UPDATE,
    K_OTHER,
    K_COARSE, // every multigrid launch below the finest level (the other multigrid classes are
              // the finest level's)
    K_COUNT
};
const char *const kclass_names[K_COUNT] = {"shifted_spmv", "smooth_first", "smooth_sweep",
                                           "residual",     "restrict",     "prolong",
                                           "terminal",     "gmres_dot",    "gmres_update",
                                           "gmres_other",  "coarse_levels"};

struct profiler {
    std::uint32_t mask = 0;
    struct rec {
        int cls;
        hipEvent_t a, b;
        double bytes;
        i64 count = 1; // launches the interval covers
    };
    std::vector<rec> pending;
    std::vector<hipEvent_t> pool;
    i64 launches[K_COUNT] = {};
    double ms[K_COUNT] = {};
    double bytes[K_COUNT] = {};

    hipEvent_t event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    // a span: one event pair around a sequence of launches of class span_cls, whose own
    // launches record no events (their bytes add up into the span)
    int span_cls = -1;
    i64 span_launches = 0;
    double span_bytes = 0.0;

    template <typename F>
    void run(int cls, double nbytes, F &&launch) {
        if (!(mask & (1u << cls))) {
            if (cls == span_cls) {
                span_bytes += nbytes;
                ++span_launches;
            }
            launch();
            check_launch(kclass_names[cls]);
            return;
        }
        rec r{cls, event(), event(), nbytes, 1};
        HIPCHK(hipEventRecord(r.a, stream()));
        launch();
        check_launch(kclass_names[cls]);
        HIPCHK(hipEventRecord(r.b, stream()));
        pending.push_back(r);
        if (pending.size() > 4096) resolve();
    }
    // time body's launches of class cls as one interval: for runs of small latency-bound
    // kernels, where an event pair per launch would cost as much as the launch
    template <typename F>
    void span(int cls, F &&body) {
        if (!(mask & (1u << cls)) || span_cls >= 0) {
            body();
            return;
        }
        rec r{cls, event(), event(), 0.0, 0};
        HIPCHK(hipEventRecord(r.a, stream()));
        mask &= ~(1u << cls);
        span_cls = cls;
        span_bytes = 0.0;
        span_launches = 0;
        try {
            body();
        } catch (...) {
            mask |= 1u << cls;
            span_cls = -1;
            throw;
        }
        mask |= 1u << cls;
        span_cls = -1;
        HIPCHK(hipEventRecord(r.b, stream()));
        r.bytes = span_bytes;
        r.count = span_launches;
        pending.push_back(r);
        if (pending.size() > 4096) resolve();
    }
    void resolve() {
        for (auto &r : pending) {
            HIPCHK(hipEventSynchronize(r.b));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
            launches[r.cls] += r.count;
            ms[r.cls] += t;
            bytes[r.cls] += r.bytes;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void reset() {
        resolve();
        std::fill(std::begin(launches), std::end(launches), 0);
        std::fill(std::begin(ms), std::end(ms), 0.0);
        std::fill(std::begin(bytes), std::end(bytes), 0.0);
    }
};

inline profiler &prof() {
    static profiler p;
    return p;
}

inline void drain() {
    HIPCHK(hipStreamSynchronize(stream()));
    prof().resolve();
}

// ------------------------------------------------------------------ gather CSR on device
template <typename VT>
struct dev_csr {
    i64 n = 0;
    u64 nnz = 0;
    bool coded = false;
    u64 padded = 0; // stored entries of the coded layout, slice padding included
    dbuf blocks, ptr, col, val, exit, soff, ent, rlen, rcode, vdict, edict;
    gpu::csr_view<VT> view;
};

// every array the selected kernel reads must exist before anything is launched on it
template <typename VT>
void check_view(const gpu::csr_view<VT> &v, u64 nnz) {
    bool ok = v.n == 0 || v.n_blocks > 0;
    if (v.coded) {
        ok = ok && v.n_slices == (v.n + gpu::kSlice - 1) / gpu::kSlice &&
             v.n_blocks * (gpu::kThreads / gpu::kSlice) >= v.n_slices &&
             (v.n == 0 || (v.soff && v.rlen && v.vdict && (v.exit || (v.rcode && v.edict)))) &&
             (nnz == 0 || v.ent) && v.n_vdict > 0 && v.n_vdict <= gpu::kDictMax &&
             (v.exit || (v.n_edict > 0 && v.n_edict <= gpu::kDictMax));
    } else {
        ok = ok && (v.n == 0 || (v.blocks && v.ptr && v.exit)) && (nnz == 0 || (v.col && v.val));
    }
    if (!ok) throw std::logic_error("operator upload left a required device array missing");
}

// upload an operator (gather CSR + exit rates) in the coded sliced layout when it fits, else
// in the plain CSR-stream layout; with exit_array the coded layout also takes operators whose
// exit rates overflow the dictionary (kept per row; only the k_csr_sliced passes, row_exit and
// gather_row read that form)
template <typename VT>
dev_csr<VT> upload_csr(const mk::generator_csr &g, bool exit_array = false) {
    dev_csr<VT> d;
    d.n = g.n;
    d.nnz = g.ptr.empty() ? 0 : g.ptr.back();
    if (d.nnz >= (u64{1} << 32)) throw std::overflow_error("operator has more than 2^32 entries");
    d.view.n = g.n;
    u64 max_len = 0;
    for (i64 i = 0; i < g.n; ++i) max_len = std::max<u64>(max_len, g.ptr[i + 1] - g.ptr[i]);

    static_assert(mk::kSliceRows == gpu::kSlice && mk::kCodedColumnBits == static_cast<int>(gpu::kColBits),
                  "host and device agree on the coded layout");
    mk::sliced_operator<VT> so;
    if (mk::encode_sliced<VT>(g, so, exit_array)) {
        const i64 n_slices = static_cast<i64>(so.soff.size()) - 1;
        d.coded = true;
        d.padded = so.padded;
        d.soff = upload(so.soff);
        d.ent = upload(so.ent);
        d.rlen = upload(so.rlen);
        d.rcode = upload(so.rcode);
        d.vdict = upload(so.vdict);
        d.edict = upload(so.edict);
        d.view.n_slices = n_slices;
        d.view.n_blocks = (n_slices + (gpu::kThreads / gpu::kSlice) - 1) / (gpu::kThreads / gpu::kSlice);
        d.view.soff = dp<u32>(d.soff);
        d.view.ent = dp<u32>(d.ent);
        d.view.rlen = dp<std::uint8_t>(d.rlen);
        d.view.rcode = dp<std::uint8_t>(d.rcode);
        d.view.vdict = dp<VT>(d.vdict);
        d.view.edict = dp<VT>(d.edict);
        d.view.n_vdict = static_cast<int>(so.vdict.size());
        d.view.n_edict = static_cast<int>(so.edict.size());
        d.view.max_width = static_cast<int>(max_len);
        if (!so.exit.empty()) {
            d.exit = upload(so.exit);
            d.view.exit = dp<VT>(d.exit);
        }
        d.view.coded = 1;
        check_view(d.view, d.nnz);
        return d;
    }
    // plain CSR-stream: row blocks of up to kRowsPerBlock rows whose entries fit the LDS tile;
    // a row longer than the tile gets a block of its own
    std::vector<u32> blocks{0};
    i64 r = 0;
    while (r < g.n) {
        const i64 r0 = r;
        const u64 p0 = g.ptr[r0];
        ++r;
        while (r < g.n && r - r0 < gpu::kRowsPerBlock &&
               g.ptr[r + 1] - p0 <= static_cast<u64>(gpu::kTile))
            ++r;
        blocks.push_back(static_cast<u32>(r));
    }
    d.blocks = upload(blocks);
    d.view.n_blocks = static_cast<i64>(blocks.size()) - 1;
    d.view.blocks = dp<u32>(d.blocks);
    std::vector<u32> ptr(g.n + 1);
    for (i64 i = 0; i <= g.n; ++i) ptr[i] = static_cast<u32>(g.ptr[i]);
    d.ptr = upload(ptr);
    d.col = upload(g.col);
    std::vector<VT> v(g.val.begin(), g.val.end()), e(g.exit.begin(), g.exit.end());
    d.val = upload(v);
    d.exit = upload(e);
    d.view.ptr = dp<u32>(d.ptr);
    d.view.col = dp<u32>(d.col);
    d.view.val = dp<VT>(d.val);
    d.view.exit = dp<VT>(d.exit);
    check_view(d.view, d.nnz);
    return d;
}

// algorithmic bytes of one pass over an operator in its implemented layout: the encoded
// entries and per-row data once, x once (xs bytes per row), and `vectors` other per-row
// vectors of ts bytes (rhs, output, ...)
template <typename VT>
double op_bytes(const dev_csr<VT> &A, std::size_t xs, std::size_t ts, int vectors) {
    const double n = double(A.n), nnz = double(A.nnz);
    double b = A.coded ? double(A.padded) * 4.0 + n * 2.0 + double(A.view.n_slices + 1) * 4.0
                       : nnz * double(sizeof(VT) + 4) + (n + 1) * 4.0 + n * sizeof(VT);
    if (A.coded && A.view.exit) b += n * (double(sizeof(VT)) - 1.0); // per-row exit, no code
    return b + n * double(xs) + n * double(ts) * vectors;
}

// unique id of a handle (never reused, unlike its address)
inline u64 next_uid() {
    static std::atomic<u64> counter{1};
    return counter.fetch_add(1);
}

// Captured launch sequences, keyed by everything their kernel arguments depend on. A key is
// captured on its second use (one-off configurations stay eager) and replayed afterwards.
// Off by default (with the fused multigrid tail, replay measured no faster than eager launches);
// MK_GRAPHS=1 or mk_graphs_enable(1) turns it on; never while profiling (events cannot be
// recorded inside a capture).
struct graph_key {
    u64 a = 0, b = 0;
    i64 c = 0, d = 0, e = 0, f = 0;
    u64 g = 0; // bit pattern of a scalar argument (a shift)
    bool operator<(const graph_key &o) const {
        return std::tie(a, b, c, d, e, f, g) < std::tie(o.a, o.b, o.c, o.d, o.e, o.f, o.g);
    }
};

inline std::atomic<int> &graphs_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_GRAPHS");
        return (env && env[0] == '1') ? 1 : 0;
    }()};
    return on;
}

// graphs are only ever used for operators of at most kGraphRows rows, where launch latency is a
// large share of an iteration (before the fused tail: -20% per FGMRES iteration at 969 states,
// -14% at 58,905, +8% at 814,385)
constexpr i64 kGraphRows = 262144;

// the fused coarse-level tail of the V-cycle (default on; MK_MG_TAIL=0 turns it off at start)
// Device-side generator assembly of RDME boxes (default on; MK_DEVICE_SETUP=0 assembles on the
// host). Both give the same operator, exit and leak arrays bit for bit.
inline std::atomic<int> &device_setup_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_DEVICE_SETUP");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// levels of at most this many rows run in the fused tail (MK_TAIL_ROWS overrides, for tuning)
// the levels whose coarse correction takes coarse_cycles visits: the top cycle_levels of them
// (all when cycle_levels is 0); the ones below visit their coarse level once
inline i64 cycle_levels(const mk::multigrid_options &o) {
    return o.cycle_levels > 0 ? static_cast<i64>(o.cycle_levels) : i64{1} << 30;
}

inline i64 tail_rows() {
    static const i64 rows = [] {
        const char *env = std::getenv("MK_TAIL_ROWS");
        const long v = env ? std::strtol(env, nullptr, 10) : 0;
        return v > 0 ? static_cast<i64>(v) : gpu::kTailRows;
    }();
    return rows;
}

// LDS-resident fused tail (default on where the levels fit; MK_TAIL_LDS=0 turns it off)
inline std::atomic<int> &tail_lds_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_TAIL_LDS");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// the next red-black level's first half-sweep fused into the kernel that makes its right-hand
// side (default on; MK_FUSED_FIRST=0 turns it off); the same values either way
inline std::atomic<int> &fused_first_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_FUSED_FIRST");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// FGMRES in one workgroup for problems whose whole hierarchy fits the LDS tail (default on;
// MK_TINY_GMRES=0 turns it off): the same iteration, not the same rounding, as the host loop
inline std::atomic<int> &tiny_gmres_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_TINY_GMRES");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// a direct coarse level's dense solve fused with the prolongation above it (default on;
// MK_COARSE_FUSE=0 turns it off); the same values either way
inline std::atomic<int> &coarse_fuse_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_COARSE_FUSE");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

inline std::atomic<int> &tail_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_MG_TAIL");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// the row limit of graph replay (MK_GRAPH_ROWS overrides, for tuning)
inline i64 graph_rows() {
    static const i64 rows = [] {
        const char *env = std::getenv("MK_GRAPH_ROWS");
        const long long v = env ? std::strtoll(env, nullptr, 10) : 0;
        return v > 0 ? static_cast<i64>(v) : kGraphRows;
    }();
    return rows;
}

inline bool graphs_enabled(i64 rows = 0) {
    return graphs_switch().load() != 0 && prof().mask == 0 && rows <= graph_rows();
}

struct graph_cache {
    std::map<graph_key, hipGraphExec_t> exec;
    std::map<graph_key, int> uses;
    ~graph_cache() { clear(); }
    void clear() {
        for (auto &kv : exec) (void)hipGraphExecDestroy(kv.second);
        exec.clear();
        uses.clear();
    }
    // run body on s: replay its graph, capture it, or launch it eagerly
    // force: capture whatever the process-wide switch says (still never while profiling)
    template <typename F>
    void run(const graph_key &k, hipStream_t s, F &&body, i64 rows, bool force = false) {
        if (force ? prof().mask != 0 : !graphs_enabled(rows)) {
            body();
            return;
        }
        auto it = exec.find(k);
        if (it == exec.end()) {
            if (++uses[k] < 2) {
                body();
                return;
            }
            if (exec.size() >= 4096) clear();
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(s, &g));
            hipGraphExec_t e = nullptr;
            const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(err);
            it = exec.emplace(k, e).first;
        }
        HIPCHK(hipGraphLaunch(it->second, s));
    }
};

inline u64 bits(double x) {
    u64 u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}

// Givens-rotated least-squares problem of one restart cycle (column-major upper Hessenberg H of
// size (m + 1) x m, stored row-major with leading dimension m)
struct hessenberg_ls {
    i64 m;
    std::vector<double> H, cs, sn, g, y;
    explicit hessenberg_ls(i64 m_) : m(m_), H((m_ + 1) * m_), cs(m_), sn(m_), g(m_ + 1), y(m_) {}
    void start(double beta) {
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
    }
    // column j = (col[0 .. j], hn); returns |g[j + 1]|, the residual norm of the cycle so far
    double add_column(i64 j, const double *col, double hn) {
        for (i64 k = 0; k <= j; ++k) H[k * m + j] = col[k];
        H[(j + 1) * m + j] = hn;
        for (i64 k = 0; k < j; ++k) {
            const double a = H[k * m + j], c = H[(k + 1) * m + j];
            H[k * m + j] = cs[k] * a + sn[k] * c;
            H[(k + 1) * m + j] = -sn[k] * a + cs[k] * c;
        }
        const double a = H[j * m + j], c = H[(j + 1) * m + j];
        const double d = std::hypot(a, c);
        cs[j] = d > 0.0 ? a / d : 1.0;
        sn[j] = d > 0.0 ? c / d : 0.0;
        H[j * m + j] = d;
        H[(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        return std::abs(g[j + 1]);
    }
    // y = R^{-1} g over the first j columns
    void solve(i64 j) {
        for (i64 i = j; i-- > 0;) {
            double sum = g[i];
            for (i64 k = i + 1; k < j; ++k) sum -= H[i * m + k] * y[k];
            y[i] = sum / H[i * m + i];
        }
    }
};

} // namespace mk::rt

// ------------------------------------------------------------------ handle types
struct mk_box_s {
    mk::population_space space;
};

struct mk_domain_s {
    mk::voxel_domain dom;
};

struct mk_generator_s {
    mk::u64 uid = mk::rt::next_uid();
    mk::i64 n = 0;
    mk::i64 n_jumps = 0;
    std::vector<double> exit, leak;
    double norm1 = 0.0; // ||Q||_1
    mk::rt::dev_csr<double> A;
    mk::rt::dbuf d_x, d_y; // staging for host-buffer calls
};

namespace mk::rt {

// The generator of an RDME box assembled on the device (assemble.hip): the same gather-form
// operator, exit and leak as mk::assemble_generator, encoded in place. Returns false, with G
// untouched, for boxes outside its limits (more than kMaxMolecules molecules per state, more
// than kMaxSpecies species, 65,536 voxels or more); the caller then assembles on the host.
bool assemble_generator_device(const mk::population_space &space, const mk::voxel_domain &domain,
                               const mk::reaction_tables &reactions, bool absorbing,
                               mk_generator_s &G);

// The same rows as mk::assemble_generator, assembled on the device and copied back to a host
// CSR (the multigrid levels' operators). False for spaces outside the kernel's limits.
bool assemble_csr_device(const mk::population_space &space, const mk::voxel_domain &domain,
                         const mk::reaction_tables &reactions, bool absorbing, mk::generator_csr &g);

// G <- the generator g: device upload in the coded or plain layout, exit and leak rates, ||Q||_1
void finish_generator(mk_generator_s &G, mk::generator_csr &&g);

// A right preconditioner of the flexible GMRES: z = M (scale v), z in the precision's type
// (FP32 or FP64), scale read from device memory (1 when null). apply launches on the library
// stream and must not synchronise (GMRES may capture it in a graph).
struct precond {
    mk::u64 uid = 0;
    mk::i64 rows = 0;
    double shift = 0.0;
    int precision = MK_FP64;
    std::function<void(const double *v, void *z, const double *scale)> apply;
    mk_sector_mg_s *mg = nullptr; // a sector hierarchy behind apply (null: another preconditioner)
};

// FGMRES on (shift I - Q) x = b with b, x in device memory (M null: unpreconditioned)
mk_solver_result gmres_device(mk_generator_s &A, double shift, const precond *M, const double *b,
                              double *x, const mk_gmres_options &o);

// level_hooks::transfer on the device (sector levels of at least 4096 states)
bool transfer_device(const mk::population_space &fine, const mk::population_space &coarse,
                     const mk::voxel_domain &dom, const std::vector<i64> &parent,
                     const std::vector<double> &log_fact, const std::vector<double> &log_children,
                     std::vector<i64> &coarse_of, std::vector<double> &log_weight,
                     std::vector<std::uint8_t> *colour);

template <typename T>
struct level_dev {
    i64 n = 0;
    dev_csr<T> A;
    // transfer to the next level
    dbuf coarse_of, weight;
    i64 n_fine_linked = 0; // rows that have a coarse parent
    dev_csr<T> R;          // restriction: next level's rows x this level's rows, rates 1
    i64 n_next = 0;
    // terminal blocks on this level
    i64 n_term = 0, term_max = 0;
    dbuf tstart, tsize, toff, tinv;
    // a coarse level whose rows all lie in terminal blocks (and no coarser level): its cycle
    // from zero is x = T^-1 rhs, the value the reference's smooth / solve / smooth sequence
    // reaches up to rounding, since the block solve is exact
    bool direct = false;
    // host copies: block starts, sizes, offsets and the dense blocks of (diag(exit) - C)
    // without the shift, laid out like tinv (re-inverted per shift by batched solves)
    std::vector<u32> h_tstart, h_tsize;
    std::vector<u64> h_toff;
    std::vector<double> h_tq;
    // aggregate-block smoother (n_agg > 0): every sector's blocks, aggregate by aggregate, with
    // member and affected-row ranges per aggregate on the host; rows of sectors that are
    // terminal on this level keep the Jacobi smoother (jrows)
    i64 n_agg = 0;
    std::vector<u64> h_agg_members, h_agg_rows;
    dbuf bs_members, bs_member_block, bs_block_first, bs_inv_off, bs_inv, bs_rows, escr;
    i64 n_jrows = 0;
    dbuf jrows;
    // red-black Gauss-Seidel: the level split by colour into row-mapped operators, and the
    // colour of every row
    bool rb = false;
    dev_csr<T> Ac[2];
    dbuf rowmap[2], colour;
    // the level numbers each sector's states colour 0 first; on the top level, perm maps a row
    // to its index in the caller's (box) order, pos a box index to its row, and exit_box holds,
    // per box index, the exit rate of its row where that row has colour 0 (-1 for colour 1)
    dbuf perm, pos, exit_box;
    // ILU(0): reciprocal pivots (for the hierarchy's shift) and the level schedules of the
    // forward and backward solves: rows grouped by dependency depth (group k is
    // rows[ptr[k] .. ptr[k + 1])), launched as single wide groups or runs of small groups
    struct ilu_launch {
        int g0 = 0, g1 = 0;
        i64 wide_rows = 0; // > 0: one group of this many rows over a grid
    };
    bool ilu = false;
    dbuf pinv, fw_rows, fw_ptr, bw_rows, bw_ptr;
    std::vector<ilu_launch> fw_launch, bw_launch;
    // work vectors
    dbuf rhs, x, xt, res, crhs, cx;
    int pre = 2, post = 2;
};

} // namespace mk::rt

struct mk_sector_mg_s {
    mk::u64 uid = mk::rt::next_uid();
    mk::multigrid_options opt;
    double shift = 0.0;
    int precision = MK_FP64;
    mk::i64 rows = 0, n_sectors = 0, max_levels = 0;
    std::vector<mk::rt::level_dev<double>> L64;
    std::vector<mk::rt::level_dev<float>> L32;
    mk::rt::dbuf d_b, d_x64, d_out; // staging for host-buffer calls
    // level descriptors of the fused tail (levels tail_start .. end run in one workgroup), and
    // its LDS-resident form when the levels fit (tail_lds_bytes > 0): tail_image, the arena's
    // constant part (its first tail_image_bytes: the descriptors with LDS offsets, then every
    // constant array at its LDS offset, copied there by tail_copies) assembled in HBM, so that
    // the kernel stages it with one contiguous read
    mk::rt::dbuf tail_desc;
    mk::i64 tail_start = -1;
    bool tail_w = false; // the level above the tail takes coarse_cycles visits: the tail runs them
    mk::rt::dbuf tail_image;
    std::vector<mk::gpu::lds_copy> tail_copies;
    mk::u32 tail_image_bytes = 0;
    mk::u32 tail_lds_bytes = 0;
    // a hierarchy over one population_sector (population_multigrid) keeps the host form of its
    // levels: rows, entries and the transfers, in rank order and FP64
    std::shared_ptr<mk::sector_hierarchy> host_levels;
    // rows of every sector on every level ([level][sector]; 0 where a sector has no such level):
    // a leading block of whole sectors is a prefix of every level
    std::vector<std::vector<mk::i64>> sector_rows;
};

