// Whole-box multigrid on the device: the hierarchy of the reference's paper driver
// rdme/cpp/rdme_full_multigrid.cpp:298-880, with every level holding the complete
// reaction-diffusion operator of a single species on a Cartesian grid (births and deaths of the
// Schlogl family, diffusion, absorbing population cap).
//
// Levels. Level l has cells / 2^l cells per axis, diffusivity D 2^l and cell volume 2^(dl) / V;
// its operator is the RDME generator of that coarse grid with the rates re-evaluated at the
// coarse volume (assembled by the library's generator assembly, on the device for levels of
// 4096 states or more). The driver's exact bottom solve is the Thomas algorithm on the one-voxel
// level; a level with at most terminal_rows states may instead be solved exactly with a dense
// inverse, which ends the recursion there (the W-cycle's many visits of tiny levels cost launch
// latency on the GPU, not work).
//
// Smoothers. Lexicographic Gauss-Seidel and the population-block ILU of the driver are
// sequential; the GPU runs Gauss-Seidel in four-colour order instead. The colour of a state is
// 2 (total mod 2) + (sum over molecules of the coordinate sum of their voxel, mod 2): a diffusion
// jump changes the second bit and a birth or death the first, so a colour's rows only read the
// other colours and one launch updates them all. Each level's rows are stored colour by colour
// (ascending state index within a colour), so a colour pass streams contiguous rows; the finest
// level's vectors are permuted on entry and exit. Weighted Jacobi (w = 2/3) is the driver's own.
//
// Cycle (rdme_full_multigrid.cpp:819-878): pre-smoothing (colours 0..3), residual, restriction
// by summing children, the coarse correction (V: one coarse cycle, W: two, the exact bottom
// solve once; or the K-cycle's at most two GCR steps), prolongation with the multinomial
// weights, post-smoothing (colours 3..0). The standalone iteration of the driver's run()
// repeats cycles until |b - A x|_1 / shift < tolerance; without the K-cycle one cycle is
// captured in a HIP graph and replayed.
#include "runtime.hpp"

#include <cmath>
#include <numeric>

using namespace mk::rt;

namespace mk::gpu {
namespace {

constexpr int kBoxThreads = 256;
constexpr int kDenseChunk = 32; // columns of the dense inverse per workgroup

// the one-voxel level: forward elimination with the stored multipliers, back substitution. The
// workgroup stages the coefficients and b in LDS (n <= kThomasLds), then one lane runs the two
// recurrences from LDS; longer chains run from global memory.
constexpr int kThomasLds = 1536;
__global__ void __launch_bounds__(256) k_box_thomas(i64 n, const double *lower, const double *upper,
                                                    const double *pivot, const double *b, double *x) {
    __shared__ double lo[kThomasLds], up[kThomasLds], pv[kThomasLds], xs[kThomasLds];
    if (n <= kThomasLds) {
        for (i64 i = threadIdx.x; i < n; i += blockDim.x) {
            lo[i] = lower[i];
            up[i] = upper[i];
            pv[i] = pivot[i];
            xs[i] = b[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (i64 i = 1; i < n; ++i) xs[i] = xs[i] - lo[i] * xs[i - 1];
            xs[n - 1] /= pv[n - 1];
            for (i64 i = n - 1; i > 0; --i) xs[i - 1] = (xs[i - 1] - up[i - 1] * xs[i]) / pv[i - 1];
        }
        __syncthreads();
        for (i64 i = threadIdx.x; i < n; i += blockDim.x) x[i] = xs[i];
        return;
    }
    if (threadIdx.x != 0) return;
    x[0] = b[0];
    for (i64 i = 1; i < n; ++i) x[i] = b[i] - lower[i] * x[i - 1];
    x[n - 1] /= pivot[n - 1];
    for (i64 i = n - 1; i > 0; --i) x[i - 1] = (x[i - 1] - upper[i - 1] * x[i]) / pivot[i - 1];
}

// restriction with one wavefront per coarse state, for levels whose coarse states have many
// children: lane l sums children l, l + 64, ... in order, then the lanes' sums are added in a
// fixed tree (deterministic)
__global__ void __launch_bounds__(kBoxThreads)
    k_box_restrict_wave(i64 nc, const u32 *cptr, const u32 *cidx, const double *fine, double *coarse) {
    const i64 c = blockIdx.x * static_cast<i64>(kBoxThreads / 64) + (threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63u;
    if (c >= nc) return;
    double s = 0.0;
    for (u32 q = cptr[c] + lane; q < cptr[c + 1]; q += 64) s += fine[cidx[q]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) coarse[c] = s;
}

// ---- small levels: one workgroup runs a whole visit's smoothing and transfer, with a barrier
// between the phases (about 20 launches per visit become 2)
constexpr int kSmallThreads = 1024;

struct small_level {
    csr_view<double> A; // the level operator, rows in the level's order
    i64 n = 0;
    u32 cb[5] = {};     // colour c holds rows cb[c] .. cb[c + 1]
    double shift = 0.0;
    int sweeps = 0, jacobi = 0;
    const double *b = nullptr;
    double *x = nullptr, *r = nullptr, *tmp = nullptr;
    i64 nc = 0;         // the next level: rows, children lists, parents and weights
    int wide = 0;       // restriction by wavefronts (as k_box_restrict_wave)
    const u32 *cptr = nullptr, *cidx = nullptr, *parent = nullptr;
    const double *weight = nullptr;
    double *coarse_b = nullptr, *coarse_x = nullptr;
};

// sum over row i's entries of rate * x[column] in entry order, and row i's exit rate
__device__ __forceinline__ double row_sum(const csr_view<double> &A, u32 i, const double *x, double &ex) {
    double sum = 0.0;
    if (A.coded) {
        const u32 len = A.rlen[i];
        const u32 base = A.soff[i >> 6] + (i & 63u);
        ex = A.exit ? A.exit[i] : A.edict[A.rcode[i]];
        for (u32 k = 0; k < len; ++k) {
            const u32 e = A.ent[base + k * kSlice];
            sum += A.vdict[e >> kColBits] * x[e & kColMask];
        }
    } else {
        ex = A.exit[i];
        for (u32 q = A.ptr[i]; q < A.ptr[i + 1]; ++q) sum += A.val[q] * x[A.col[q]];
    }
    return sum;
}

// the level's smoother (all sweeps), with the formulas of the per-level kernels
__device__ void small_smooth(const small_level &L, bool reverse) {
    const u32 t = threadIdx.x, step = blockDim.x;
    for (int sweep = 0; sweep < L.sweeps; ++sweep) {
        if (L.jacobi) {
            const double omega = 2.0 / 3.0;
            for (u32 i = t; i < L.n; i += step) {
                double ex = 0.0;
                const double sum = row_sum(L.A, i, L.x, ex);
                const double xr = L.x[i], d = L.shift + ex;
                const double res = L.b[i] - (d * xr - sum);
                L.tmp[i] = xr + omega * res / d;
            }
            __syncthreads();
            for (u32 i = t; i < L.n; i += step) L.x[i] = L.tmp[i];
            __syncthreads();
        } else {
            for (int k = 0; k < 4; ++k) {
                const int c = reverse ? 3 - k : k;
                for (u32 i = L.cb[c] + t; i < L.cb[c + 1]; i += step) {
                    double ex = 0.0;
                    const double sum = row_sum(L.A, i, L.x, ex);
                    L.x[i] = (L.b[i] + sum) / (L.shift + ex);
                }
                __syncthreads();
            }
        }
    }
}

// pre-smoothing, residual, restriction to the next level and its zero iterate
__global__ void __launch_bounds__(kSmallThreads) k_box_small_down(small_level L) {
    small_smooth(L, false);
    const u32 t = threadIdx.x, step = blockDim.x;
    for (u32 i = t; i < L.n; i += step) {
        double ex = 0.0;
        const double sum = row_sum(L.A, i, L.x, ex);
        L.r[i] = L.b[i] - ((L.shift + ex) * L.x[i] - sum);
    }
    __syncthreads();
    if (L.wide) {
        // one wavefront per coarse state, the summation order of k_box_restrict_wave
        const u32 lane = t & 63u, wave = t >> 6, waves = step >> 6;
        for (u32 c = wave; c < L.nc; c += waves) {
            double sum = 0.0;
            for (u32 q = L.cptr[c] + lane; q < L.cptr[c + 1]; q += 64) sum += L.r[L.cidx[q]];
            for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
            if (lane == 0) {
                L.coarse_b[c] = sum;
                L.coarse_x[c] = 0.0;
            }
        }
        return;
    }
    for (u32 c = t; c < L.nc; c += step) {
        double sum = 0.0;
        for (u32 q = L.cptr[c]; q < L.cptr[c + 1]; ++q) sum += L.r[L.cidx[q]];
        L.coarse_b[c] = sum;
        L.coarse_x[c] = 0.0;
    }
}

// prolongation of the next level's correction and post-smoothing
__global__ void __launch_bounds__(kSmallThreads) k_box_small_up(small_level L) {
    const u32 t = threadIdx.x, step = blockDim.x;
    for (u32 i = t; i < L.n; i += step) L.x[i] += L.weight[i] * L.coarse_x[L.parent[i]];
    __syncthreads();
    small_smooth(L, true);
}

// partial[c][i] = sum over the columns j of chunk c of inv[j][i] b[j] (inverse column-major:
// the lanes of a wavefront read consecutive rows of one column); one wavefront per 64 rows and
// column chunk, so a 561-state level already spreads over 162 workgroups
__global__ void __launch_bounds__(64)
    k_box_dense(i64 n, const double *inv, const double *b, double *partial) {
    const i64 i = blockIdx.x * static_cast<i64>(64) + threadIdx.x;
    const i64 j0 = static_cast<i64>(blockIdx.y) * kDenseChunk;
    const i64 j1 = j0 + kDenseChunk < n ? j0 + kDenseChunk : n;
    if (i >= n) return;
    double s = 0.0;
    for (i64 j = j0; j < j1; ++j) s += inv[j * n + i] * b[j];
    partial[static_cast<i64>(blockIdx.y) * n + i] = s;
}

__global__ void __launch_bounds__(kBoxThreads)
    k_box_dense_sum(i64 n, int chunks, const double *partial, double *x) {
    const i64 i = blockIdx.x * static_cast<i64>(kBoxThreads) + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += partial[static_cast<i64>(c) * n + i];
    x[i] = s;
}

// partial[block] = sum of |r| over the block's grid-stride share
__global__ void __launch_bounds__(kBoxThreads) k_box_abs_partial(i64 n, const double *r, double *partial) {
    __shared__ double red[kBoxThreads / 64];
    double s = 0.0;
    for (i64 i = blockIdx.x * static_cast<i64>(kBoxThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kBoxThreads)
        s += fabs(r[i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBoxThreads / 64; ++w) t += red[w];
        partial[blockIdx.x] = t;
    }
}

// b = s p (backward Euler) or b = 2 s p - A p (Crank-Nicolson, ap = A p), rounded as written
__global__ void __launch_bounds__(kBoxThreads)
    k_box_rhs(i64 n, double s, const double *p, const double *ap, double *b) {
#pragma clang fp contract(off)
    const double s2 = 2.0 * s;
    for (i64 i = blockIdx.x * static_cast<i64>(kBoxThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kBoxThreads)
        b[i] = ap ? s2 * p[i] - ap[i] : s * p[i];
}

// u = r * pinv (the diagonal solve that starts a backward Gauss-Seidel sweep)
__global__ void __launch_bounds__(kBoxThreads)
    k_box_scale(i64 n, const double *r, const double *pinv, double *u) {
    for (i64 i = blockIdx.x * static_cast<i64>(kBoxThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kBoxThreads)
        u[i] = r[i] * pinv[i];
}

__global__ void __launch_bounds__(kBoxThreads) k_box_copy(i64 n, const double *src, double *dst) {
    for (i64 i = blockIdx.x * static_cast<i64>(kBoxThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kBoxThreads)
        dst[i] = src[i];
}

int grid_of(i64 n) {
    const i64 g = (n + kBoxThreads - 1) / kBoxThreads;
    return static_cast<int>(std::max<i64>(1, std::min<i64>(g, 65535)));
}

} // namespace

void box_thomas(i64 n, const double *lower, const double *upper, const double *pivot,
                const double *b, double *x, hipStream_t s) {
    hipLaunchKernelGGL(k_box_thomas, dim3(1), dim3(256), 0, s, n, lower, upper, pivot, b, x);
}

void box_restrict_wave(i64 nc, const u32 *cptr, const u32 *cidx, const double *fine, double *coarse,
                       hipStream_t s) {
    if (nc == 0) return;
    const i64 blocks = (nc + kBoxThreads / 64 - 1) / (kBoxThreads / 64);
    hipLaunchKernelGGL(k_box_restrict_wave, dim3(static_cast<unsigned>(blocks)), dim3(kBoxThreads), 0, s,
                       nc, cptr, cidx, fine, coarse);
}

void box_dense(i64 n, const double *inv, const double *b, double *partial, double *x, hipStream_t s) {
    const int chunks = static_cast<int>((n + kDenseChunk - 1) / kDenseChunk);
    const int tiles = static_cast<int>((n + 63) / 64);
    const int rows = static_cast<int>((n + kBoxThreads - 1) / kBoxThreads);
    hipLaunchKernelGGL(k_box_dense, dim3(tiles, chunks), dim3(64), 0, s, n, inv, b, partial);
    hipLaunchKernelGGL(k_box_dense_sum, dim3(rows), dim3(kBoxThreads), 0, s, n, chunks, partial, x);
}

// doubles of the partial-sum buffer box_dense needs for an n-state level
i64 box_dense_partials(i64 n) { return (n + kDenseChunk - 1) / kDenseChunk * n; }

void box_abs_partial(i64 n, const double *r, double *partial, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_box_abs_partial, dim3(nb), dim3(kBoxThreads), 0, s, n, r, partial);
}

void box_rhs(i64 n, double shift, const double *p, const double *ap, double *b, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_box_rhs, dim3(grid_of(n)), dim3(kBoxThreads), 0, s, n, shift, p, ap, b);
}

void box_small_down(const small_level &L, hipStream_t s) {
    hipLaunchKernelGGL(k_box_small_down, dim3(1), dim3(kSmallThreads), 0, s, L);
}

void box_small_up(const small_level &L, hipStream_t s) {
    hipLaunchKernelGGL(k_box_small_up, dim3(1), dim3(kSmallThreads), 0, s, L);
}

void box_scale(i64 n, const double *r, const double *pinv, double *u, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_box_scale, dim3(grid_of(n)), dim3(kBoxThreads), 0, s, n, r, pinv, u);
}

void box_copy(i64 n, const double *src, double *dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_box_copy, dim3(grid_of(n)), dim3(kBoxThreads), 0, s, n, src, dst);
}

} // namespace mk::gpu

namespace mk::rt {

// one level of the device hierarchy
struct box_level {
    i64 n = 0;
    i64 voxels = 0;
    int terminal = 0;  // 0 smoothed, 1 Thomas, 2 dense inverse
    u32 cb[5] = {};    // colour c holds rows cb[c] .. cb[c + 1]
    dev_csr<double> A; // the level operator in colour order (residual, Jacobi)
    dev_csr<double> Ac[4];
    dbuf rowmap;       // 0 .. n - 1; colour c's operator maps its local rows from rowmap + cb[c]
    // transfer to the next level, in the two levels' row orders
    dbuf parent, weight, cptr, cidx;
    dbuf b, x, r, tmp; // b, x of coarse levels; r the residual; tmp the Jacobi target
    dbuf lower, upper, pivot, inv, partial;
    dbuf kv1, kw1, kv2, kw2, kr; // K-cycle vectors (the next level's size)
    u64 nnz = 0;
    bool wide_restrict = false;  // coarse states average more than 16 children
    // the driver's sequential smoothers (rows in the box order): level-scheduled triangular
    // solves. Lexicographic Gauss-Seidel: one schedule over the whole level on A, pinv = 1 /
    // diagonal. Population-block ILU: the in-block operator Ain, pinv = 1 / ILU pivot, one
    // schedule per population block, and one row-range operator per block for its residual.
    struct tri_launch {
        int g0 = 0, g1 = 0;
        i64 wide_rows = 0; // > 0: one group of this many rows over a grid
    };
    dev_csr<double> Ain;
    std::vector<dev_csr<double>> Ablock;
    std::vector<std::vector<tri_launch>> fw, bw; // per block (one entry for Gauss-Seidel)
    dbuf pinv, u, fw_rows, fw_ptr, bw_rows, bw_ptr;
    bool small = false; // smoothing and transfers of a visit run in one workgroup each way
};

} // namespace mk::rt

struct mk_box_mg_s {
    mk::u64 uid = mk::rt::next_uid();
    mk_box_mg_params p{};
    mk::i64 rows = 0;
    std::vector<mk::rt::box_level> L;
    mk::rt::dbuf perm0;            // level-0 row r holds the state perm0[r] of the box order
    mk::rt::dbuf b0, x0;           // the finest level's vectors in its row order
    mk::rt::dbuf red, scal;        // reduction partials, device scalars
    mk::rt::dbuf step_b, step_x;   // time-step right-hand side and iterate (box order)
    double *host = nullptr;        // pinned copy of one scalar
    std::unique_ptr<mk_generator_s> gen; // the finest operator in the box order
    mk::rt::graph_cache graphs;
    ~mk_box_mg_s() {
        graphs.clear();
        if (host) (void)hipHostFree(host);
    }
};

namespace mk::rt {
namespace {

// levels of at most this many states run a visit's smoothing and transfers in two
// single-workgroup kernels (four-colour Gauss-Seidel and Jacobi; MK_BOX_SMALL=0 turns it off at
// start, mk_box_mg_small_enable at run time)
constexpr i64 kSmallRows = 2048;
i64 small_rows() {
    static const i64 rows = [] {
        const char *env = std::getenv("MK_BOX_SMALL_ROWS"); // tuning override
        const long v = env ? std::strtol(env, nullptr, 10) : 0;
        return v > 0 ? static_cast<i64>(v) : kSmallRows;
    }();
    return rows;
}
std::atomic<int> &box_small_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_BOX_SMALL");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

// MK_BOX_GRAPHS=0 launches the standalone iteration's cycles eagerly
std::atomic<int> &box_graph_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_BOX_GRAPHS");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

bool sequential(const mk_box_mg_params &p) {
    return p.smoother == MK_BOX_GS_LEX || p.smoother == MK_BOX_BLOCK_ILU;
}

struct host_level {
    int cells = 0;
    double diffusion = 0.0;
    mk::population_space space;
    mk::voxel_domain dom;
    mk::generator_csr g;            // box order
    std::vector<std::uint8_t> colour;
    std::vector<u32> ord, pos;      // row r = state ord[r]; pos[state] = r
    std::vector<i64> parent;        // box order of the next level
    std::vector<double> weight;
    int terminal = 0;
};

// the driver's Schlogl rate tables (rdme_full_multigrid.cpp:400-402): births +1, deaths -1,
// rates[(r V + v) T + m] for m molecules in voxel v
mk::reaction_tables schlogl_tables(const mk_box_mg_params &p, i64 V) {
    mk::reaction_tables t;
    t.n = 2;
    t.changes = {1, -1};
    t.stride = {1};
    t.T = p.cap + 1;
    t.rates.assign(static_cast<std::size_t>(2 * V * t.T), 0.0);
    const double vol = 1.0 / static_cast<double>(V);
    for (i64 v = 0; v < V; ++v)
        for (i64 m = 0; m <= p.cap; ++m) {
            const double x = static_cast<double>(m);
            t.rates[static_cast<std::size_t>(v * t.T + m)] = p.c3 * vol + p.c1 * x * (x - 1) / (2 * vol);
            t.rates[static_cast<std::size_t>((V + v) * t.T + m)] =
                p.c4 * x + p.c2 * x * (x - 1) * (x - 2) / (6 * vol * vol);
        }
    return t;
}

i64 voxels_of(const mk_box_mg_params &p, int cells) {
    i64 v = 1;
    for (int d = 0; d < p.dimension; ++d) v *= cells;
    return v;
}

// sum over axes of the voxel's coordinate (axis 0 fastest)
std::vector<int> coordinate_sums(int dim, int cells, i64 V) {
    std::vector<int> out(V, 0);
    for (i64 v = 0; v < V; ++v) {
        i64 rest = v;
        for (int d = 0; d < dim; ++d) {
            out[v] += static_cast<int>(rest % cells);
            rest /= cells;
        }
    }
    return out;
}

void build_level_operator(const mk_box_mg_params &p, host_level &h) {
    const i64 V = voxels_of(p, h.cells);
    h.space = mk::population_space::box(V, {p.cap});
    std::vector<i64> dims(p.dimension, h.cells);
    std::vector<double> spacing(p.dimension, 1.0 / h.cells);
    h.dom = mk::voxel_domain::grid(dims, spacing, nullptr, {h.diffusion});
    const auto tables = schlogl_tables(p, V);
    if (!(device_setup_switch().load() != 0 && h.space.size() >= 4096 &&
          assemble_csr_device(h.space, h.dom, tables, true, h.g)))
        h.g = mk::assemble_generator(h.space, h.dom, tables, true);
}

// colours of every state and the colour-major order; throws when a stored coupling joins two
// states of one colour (a reaction that does not change the population by one)
void colour_order(const mk_box_mg_params &p, host_level &h, bool check) {
    const i64 n = h.g.n, V = h.space.voxels();
    const auto csum = coordinate_sums(p.dimension, h.cells, V);
    h.colour.assign(n, 0);
    const i64 chunk = 1 << 14;
    mk::parallel_tasks((n + chunk - 1) / chunk, [&](i64 t) {
        std::vector<int> x(V);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i) {
            h.space.state(i, x.data());
            int total = 0, parity = 0;
            for (i64 v = 0; v < V; ++v) {
                total += x[v];
                parity += x[v] * csum[v];
            }
            h.colour[i] = static_cast<std::uint8_t>(((total & 1) << 1) | (parity & 1));
        }
    });
    if (check) {
        std::atomic<int> bad{0};
        mk::parallel_tasks((n + chunk - 1) / chunk, [&](i64 t) {
            for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i)
                for (u64 q = h.g.ptr[i]; q < h.g.ptr[i + 1]; ++q)
                    if (h.colour[h.g.col[q]] == h.colour[i]) bad.store(1);
        });
        if (bad.load())
            throw std::invalid_argument(
                "box multigrid: a coupling joins two states of one colour; the four-colour "
                "Gauss-Seidel needs reactions that change the population by one");
    }
    h.ord.resize(n);
    h.pos.resize(n);
    u32 at = 0;
    for (int c = 0; c < 4; ++c)
        for (i64 i = 0; i < n; ++i)
            if (h.colour[i] == c) h.ord[at++] = static_cast<u32>(i);
    for (i64 r = 0; r < n; ++r) h.pos[h.ord[r]] = static_cast<u32>(r);
}

// the multinomial transfer of fine level f to level c (rdme_full_multigrid.cpp:555-572):
// the coarse state adds the counts of each 2^d block of voxels; weight = exp(log of the
// multinomial probability of the fine arrangement), normalised over the coarse state's children
// in ascending fine order
void build_transfer(const mk_box_mg_params &p, host_level &f, const host_level &c) {
    const i64 V = f.space.voxels(), Vc = c.space.voxels(), n = f.g.n;
    std::vector<i64> vparent(V);
    for (i64 v = 0; v < V; ++v) {
        i64 rest = v, stride = 1, cv = 0;
        for (int d = 0; d < p.dimension; ++d) {
            cv += ((rest % f.cells) / 2) * stride;
            rest /= f.cells;
            stride *= c.cells;
        }
        vparent[v] = cv;
    }
    std::vector<double> lf(p.cap + 1);
    for (int k = 0; k <= p.cap; ++k) lf[k] = std::lgamma(k + 1.0);
    const double log_children = p.dimension * std::log(2.0);
    f.parent.assign(n, 0);
    f.weight.assign(n, 0.0);
    const i64 chunk = 1 << 14;
    mk::parallel_tasks((n + chunk - 1) / chunk, [&](i64 t) {
        std::vector<int> x(V), cn(Vc);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i) {
            f.space.state(i, x.data());
            std::fill(cn.begin(), cn.end(), 0);
            int total = 0;
            double lw = 0.0;
            for (i64 v = 0; v < V; ++v) total += x[v];
            lw = -total * log_children;
            for (i64 v = 0; v < V; ++v) {
                cn[vparent[v]] += x[v];
                lw -= lf[x[v]];
            }
            for (int k : cn) lw += lf[k];
            const i64 j = c.space.find(cn.data());
            if (j < 0) throw std::logic_error("box multigrid: coarse state outside the coarse box");
            f.parent[i] = j;
            f.weight[i] = std::exp(lw);
        }
    });
    std::vector<double> sums(c.g.n, 0.0);
    for (i64 i = 0; i < n; ++i) sums[f.parent[i]] += f.weight[i];
    for (i64 i = 0; i < n; ++i) f.weight[i] /= sums[f.parent[i]];
}

// the operator with rows and columns in the level's row order (entries ascending by column)
mk::generator_csr permuted(const host_level &h) {
    const i64 n = h.g.n;
    mk::generator_csr out;
    out.n = n;
    out.n_jumps = h.g.n_jumps;
    out.ptr.assign(n + 1, 0);
    for (i64 r = 0; r < n; ++r) out.ptr[r + 1] = out.ptr[r] + (h.g.ptr[h.ord[r] + 1] - h.g.ptr[h.ord[r]]);
    out.col.resize(out.ptr[n]);
    out.val.resize(out.ptr[n]);
    out.exit.resize(n);
    out.leak.resize(n);
    const i64 chunk = 1 << 14;
    mk::parallel_tasks((n + chunk - 1) / chunk, [&](i64 t) {
        std::vector<std::pair<u32, double>> e;
        for (i64 r = t * chunk; r < std::min(n, (t + 1) * chunk); ++r) {
            const i64 i = h.ord[r];
            e.clear();
            for (u64 q = h.g.ptr[i]; q < h.g.ptr[i + 1]; ++q) e.push_back({h.pos[h.g.col[q]], h.g.val[q]});
            std::sort(e.begin(), e.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            u64 at = out.ptr[r];
            for (auto &[c, v] : e) {
                out.col[at] = c;
                out.val[at++] = v;
            }
            out.exit[r] = h.g.exit[i];
            out.leak[r] = h.g.leak[i];
        }
    });
    return out;
}

// rows [r0, r1) of g (columns unchanged)
mk::generator_csr row_range(const mk::generator_csr &g, i64 r0, i64 r1) {
    mk::generator_csr out;
    out.n = r1 - r0;
    out.ptr.assign(out.n + 1, 0);
    for (i64 r = r0; r < r1; ++r) out.ptr[r - r0 + 1] = g.ptr[r + 1] - g.ptr[r0];
    out.col.assign(g.col.begin() + g.ptr[r0], g.col.begin() + g.ptr[r1]);
    out.val.assign(g.val.begin() + g.ptr[r0], g.val.begin() + g.ptr[r1]);
    out.exit.assign(g.exit.begin() + r0, g.exit.begin() + r1);
    out.leak.assign(g.leak.begin() + r0, g.leak.begin() + r1);
    return out;
}


// Level schedule of triangular solves: rows grouped by (block, dependency depth); per block the
// launches cover its groups: groups of at least kWideGroup rows get a grid each, runs of smaller
// groups run in one workgroup with a barrier between groups
constexpr i64 kWideGroup = 2048;

void tri_schedule(const std::vector<i64> &depth, const std::vector<i64> &block_begin, dbuf &rows_d,
                  dbuf &ptr_d, std::vector<std::vector<box_level::tri_launch>> &launches) {
    const i64 nb = static_cast<i64>(block_begin.size()) - 1;
    std::vector<i64> gbase(nb + 1, 0);
    for (i64 k = 0; k < nb; ++k) {
        i64 deepest = -1;
        for (i64 i = block_begin[k]; i < block_begin[k + 1]; ++i) deepest = std::max(deepest, depth[i]);
        gbase[k + 1] = gbase[k] + deepest + 1;
    }
    const i64 groups = gbase[nb];
    if (groups >= (i64{1} << 31)) throw std::overflow_error("box multigrid: schedule too deep");
    std::vector<u32> ptr(groups + 1, 0), rows(depth.size());
    std::vector<i64> group_of(depth.size());
    for (i64 k = 0; k < nb; ++k)
        for (i64 i = block_begin[k]; i < block_begin[k + 1]; ++i) group_of[i] = gbase[k] + depth[i];
    for (i64 g : group_of) ++ptr[g + 1];
    for (i64 g = 0; g < groups; ++g) ptr[g + 1] += ptr[g];
    std::vector<u32> fill(ptr.begin(), ptr.end() - 1);
    for (std::size_t i = 0; i < group_of.size(); ++i) rows[fill[group_of[i]]++] = static_cast<u32>(i);
    rows_d = upload(rows);
    ptr_d = upload(ptr);
    launches.assign(nb, {});
    for (i64 k = 0; k < nb; ++k) {
        for (i64 g = gbase[k]; g < gbase[k + 1];) {
            const i64 size = ptr[g + 1] - ptr[g];
            if (size >= kWideGroup) {
                launches[k].push_back({static_cast<int>(g), static_cast<int>(g + 1), size});
                ++g;
                continue;
            }
            i64 e = g;
            while (e < gbase[k + 1] && ptr[e + 1] - ptr[e] < kWideGroup) ++e;
            launches[k].push_back({static_cast<int>(g), static_cast<int>(e), 0});
            g = e;
        }
    }
}

// the two sequential smoothers of the driver on level h (box order), device data into d
void build_sequential(const mk_box_mg_params &p, const host_level &h, box_level &d) {
    const i64 n = h.g.n;
    const auto &g = h.g;
    std::vector<double> pinv(n);
    std::vector<i64> fdepth(n, 0), bdepth(n, 0);
    std::vector<i64> block_begin;
    if (p.smoother == MK_BOX_GS_LEX) {
        // (D + L) and (D + U) of the whole level: every lower (upper) column is a dependency
        for (i64 i = 0; i < n; ++i) {
            pinv[i] = 1.0 / (p.shift + g.exit[i]);
            for (u64 q = g.ptr[i]; q < g.ptr[i + 1] && g.col[q] < i; ++q)
                fdepth[i] = std::max(fdepth[i], fdepth[g.col[q]] + 1);
        }
        for (i64 i = n; i-- > 0;)
            for (u64 q = g.ptr[i + 1]; q-- > g.ptr[i] && g.col[q] > i;)
                bdepth[i] = std::max(bdepth[i], bdepth[g.col[q]] + 1);
        block_begin = {0, n};
        tri_schedule(fdepth, block_begin, d.fw_rows, d.fw_ptr, d.fw);
        tri_schedule(bdepth, block_begin, d.bw_rows, d.bw_ptr, d.bw);
        d.pinv = upload(pinv);
        d.u = dbuf(n * sizeof(double));
        return;
    }
    // population blocks: rows of total k are [offset(k), offset(k + 1)) in the box order
    block_begin.assign(p.cap + 2, 0);
    for (int k = 0; k <= p.cap; ++k) block_begin[k] = h.space.begin(k);
    block_begin[p.cap + 1] = n;
    std::vector<int> block_of(n);
    for (int k = 0; k <= p.cap; ++k)
        for (i64 i = block_begin[k]; i < block_begin[k + 1]; ++i) block_of[i] = k;
    // the in-block operator: the diffusion entries (same population); exit rates stay the
    // level's, so that its diagonal is the level operator's
    mk::generator_csr in;
    in.n = n;
    in.ptr.assign(n + 1, 0);
    for (i64 i = 0; i < n; ++i) {
        u64 c = 0;
        for (u64 q = g.ptr[i]; q < g.ptr[i + 1]; ++q) c += block_of[g.col[q]] == block_of[i];
        in.ptr[i + 1] = in.ptr[i] + c;
    }
    in.col.reserve(in.ptr[n]);
    in.val.reserve(in.ptr[n]);
    for (i64 i = 0; i < n; ++i)
        for (u64 q = g.ptr[i]; q < g.ptr[i + 1]; ++q)
            if (block_of[g.col[q]] == block_of[i]) {
                in.col.push_back(g.col[q]);
                in.val.push_back(g.val[q]);
            }
    in.exit = g.exit;
    in.leak = g.leak;
    // structured ILU(0) pivots (rdme_full_multigrid.cpp:210-233): the block graph is bipartite,
    // so only the diagonal changes: p_i = a_ii - sum_{j < i in the block} (a_ij / p_j) a_ji
    std::vector<double> pivot(n);
    for (i64 i = 0; i < n; ++i) {
        double dd = p.shift + in.exit[i];
        for (u64 q = in.ptr[i]; q < in.ptr[i + 1] && in.col[q] < i; ++q) {
            const u32 j = in.col[q];
            const auto b = in.col.begin() + in.ptr[j], e = in.col.begin() + in.ptr[j + 1];
            const auto it = std::lower_bound(b, e, static_cast<u32>(i));
            if (it == e || *it != static_cast<u32>(i))
                throw std::logic_error("box multigrid: a diffusion jump has no reverse jump");
            // a_ij = -val(q), a_ji = -val(j -> i entry): (a_ij / p_j) a_ji = val val' / p_j
            dd -= in.val[q] / pivot[j] * in.val[in.ptr[j] + (it - b)];
            fdepth[i] = std::max(fdepth[i], fdepth[j] + 1);
        }
        if (!(dd > 0.0) || !std::isfinite(dd)) throw std::runtime_error("box multigrid: ILU pivot");
        pivot[i] = dd;
        pinv[i] = 1.0 / dd;
    }
    for (i64 i = n; i-- > 0;)
        for (u64 q = in.ptr[i + 1]; q-- > in.ptr[i] && in.col[q] > i;)
            bdepth[i] = std::max(bdepth[i], bdepth[in.col[q]] + 1);
    tri_schedule(fdepth, block_begin, d.fw_rows, d.fw_ptr, d.fw);
    tri_schedule(bdepth, block_begin, d.bw_rows, d.bw_ptr, d.bw);
    d.pinv = upload(pinv);
    d.u = dbuf(n * sizeof(double));
    d.Ain = upload_csr<double>(in, true);
    std::vector<u32> rm(n);
    std::iota(rm.begin(), rm.end(), 0u);
    d.rowmap = upload(rm);
    d.Ablock.resize(p.cap + 1);
    for (int k = 0; k <= p.cap; ++k) {
        if (block_begin[k + 1] == block_begin[k]) continue;
        d.Ablock[k] = upload_csr<double>(row_range(g, block_begin[k], block_begin[k + 1]), true);
        d.Ablock[k].view.rowmap = dp<u32>(d.rowmap) + block_begin[k];
    }
}

void build(mk_box_mg_s &M) {
    const auto &p = M.p;
    std::vector<host_level> H;
    // levels down to the terminal one
    int cells = p.cells;
    double diffusion = p.diffusion;
    while (true) {
        host_level h;
        h.cells = cells;
        h.diffusion = diffusion;
        build_level_operator(p, h);
        const bool dense = p.terminal_rows > 0 && h.g.n <= p.terminal_rows;
        h.terminal = dense ? 2 : (cells == 1 ? 1 : 0);
        if (h.terminal || sequential(p)) {
            // terminal levels and the driver's sequential smoothers keep the box order
            if (!h.terminal && cells % 2) throw std::invalid_argument("box multigrid: cells must be a power of two");
            h.ord.resize(h.g.n);
            std::iota(h.ord.begin(), h.ord.end(), 0u);
            h.pos = h.ord;
        } else {
            if (cells % 2) throw std::invalid_argument("box multigrid: cells must be a power of two");
            colour_order(p, h, p.smoother == MK_BOX_GS);
        }
        H.push_back(std::move(h));
        if (H.back().terminal) break;
        cells /= 2;
        diffusion *= 2;
    }
    for (std::size_t l = 0; l + 1 < H.size(); ++l) build_transfer(p, H[l], H[l + 1]);

    const hipStream_t s = stream();
    M.rows = H[0].g.n;
    M.L.resize(H.size());
    for (std::size_t l = 0; l < H.size(); ++l) {
        auto &h = H[l];
        auto &d = M.L[l];
        const i64 n = h.g.n;
        d.n = n;
        d.voxels = h.space.voxels();
        d.terminal = h.terminal;
        d.nnz = h.g.ptr.back();
        if (l > 0) {
            d.b = dbuf(n * sizeof(double));
            d.x = dbuf(n * sizeof(double));
        }
        if (h.terminal == 1) {
            // Thomas coefficients of the tridiagonal one-voxel chain
            // (rdme_full_multigrid.cpp:519-535) from A = (shift + exit) - rates
            std::vector<double> lo(n, 0.0), up(n, 0.0), pv(n, 0.0);
            for (i64 i = 0; i < n; ++i) {
                pv[i] = p.shift + h.g.exit[i];
                for (u64 q = h.g.ptr[i]; q < h.g.ptr[i + 1]; ++q) {
                    if (static_cast<i64>(h.g.col[q]) + 1 == i) lo[i] = -h.g.val[q];
                    else if (static_cast<i64>(h.g.col[q]) == i + 1) up[i] = -h.g.val[q];
                    else throw std::logic_error("box multigrid: one-voxel level not tridiagonal");
                }
                if (i > 0) {
                    lo[i] /= pv[i - 1];
                    pv[i] -= lo[i] * up[i - 1];
                }
            }
            d.lower = upload(lo);
            d.upper = upload(up);
            d.pivot = upload(pv);
            continue;
        }
        if (h.terminal == 2) {
            std::vector<double> a(static_cast<std::size_t>(n) * n, 0.0), inv;
            for (i64 i = 0; i < n; ++i) {
                a[static_cast<std::size_t>(i) * n + i] = p.shift + h.g.exit[i];
                for (u64 q = h.g.ptr[i]; q < h.g.ptr[i + 1]; ++q)
                    a[static_cast<std::size_t>(i) * n + h.g.col[q]] -= h.g.val[q];
            }
            if (!mk::invert_dense(std::move(a), n, inv))
                throw std::runtime_error("box multigrid: singular terminal level");
            std::vector<double> colmajor(inv.size());
            for (i64 i = 0; i < n; ++i)
                for (i64 j = 0; j < n; ++j)
                    colmajor[static_cast<std::size_t>(j) * n + i] = inv[static_cast<std::size_t>(i) * n + j];
            d.inv = upload(colmajor);
            d.partial = dbuf(static_cast<std::size_t>(gpu::box_dense_partials(n)) * sizeof(double));
            continue;
        }
        const auto gp = sequential(p) ? h.g : permuted(h);
        if (sequential(p)) build_sequential(p, h, d);
        for (int c = 0; c < 4 && !sequential(p); ++c) {
            d.cb[c + 1] = d.cb[c];
            while (d.cb[c + 1] < n && h.colour[h.ord[d.cb[c + 1]]] == c) ++d.cb[c + 1];
        }
        if (!sequential(p)) {
            if (d.cb[4] != n) throw std::logic_error("box multigrid: colour order");
            std::vector<u32> rm(n);
            std::iota(rm.begin(), rm.end(), 0u);
            d.rowmap = upload(rm);
        }
        for (int c = 0; c < 4 && !sequential(p); ++c) {
            if (d.cb[c + 1] == d.cb[c]) continue;
            d.Ac[c] = upload_csr<double>(row_range(gp, d.cb[c], d.cb[c + 1]), true);
            d.Ac[c].view.rowmap = dp<u32>(d.rowmap) + d.cb[c];
        }
        d.A = upload_csr<double>(gp, true);
        d.small = n <= small_rows() && (p.smoother == MK_BOX_GS || p.smoother == MK_BOX_JACOBI);
        d.r = dbuf(n * sizeof(double));
        if (p.smoother == MK_BOX_JACOBI) d.tmp = dbuf(n * sizeof(double));
        // transfer in row orders: parent'[pos_f[i]] = pos_c[parent[i]]
        const auto &hc = H[l + 1];
        const i64 nc = hc.g.n;
        std::vector<u32> par(n);
        std::vector<double> w(n);
        for (i64 i = 0; i < n; ++i) {
            par[h.pos[i]] = hc.pos[h.parent[i]];
            w[h.pos[i]] = h.weight[i];
        }
        std::vector<u32> cptr(nc + 1, 0), cidx(n);
        for (i64 r = 0; r < n; ++r) ++cptr[par[r] + 1];
        for (i64 c = 0; c < nc; ++c) cptr[c + 1] += cptr[c];
        {
            std::vector<u32> fill(cptr.begin(), cptr.end() - 1);
            for (i64 r = 0; r < n; ++r) cidx[fill[par[r]]++] = static_cast<u32>(r);
        }
        d.wide_restrict = n > 16 * nc;
        d.parent = upload(par);
        d.weight = upload(w);
        d.cptr = upload(cptr);
        d.cidx = upload(cidx);
        if (p.cycle_type == MK_CYCLE_K && !hc.terminal) {
            for (dbuf *v : {&d.kv1, &d.kw1, &d.kv2, &d.kw2, &d.kr}) *v = dbuf(nc * sizeof(double));
        }
    }
    M.perm0 = upload(H[0].ord);
    M.b0 = dbuf(M.rows * sizeof(double));
    M.x0 = dbuf(M.rows * sizeof(double));
    M.red = dbuf(static_cast<std::size_t>(gpu::reduction_blocks(M.rows)) * sizeof(double));
    M.scal = dbuf(8 * sizeof(double));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&M.host), 8 * sizeof(double)));
    M.gen = std::make_unique<mk_generator_s>();
    finish_generator(*M.gen, std::move(H[0].g));
    (void)s;
}

// ------------------------------------------------------------------ cycle
int level_class(std::size_t l, int fine) { return l == 0 ? fine : K_COARSE; }

double dot(mk_box_mg_s &M, const double *a, const double *b, i64 n) {
    const hipStream_t s = stream();
    const int nb = gpu::reduction_blocks(n);
    prof().run(K_DOT, 16.0 * n, [&] {
        gpu::multi_dot(a, n, 1, b, n, dp<double>(M.red), nb, s);
        gpu::reduce_partials(dp<double>(M.red), nb, 1, dp<double>(M.scal), s);
    });
    HIPCHK(hipMemcpyAsync(M.host, M.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    return M.host[0];
}

void smooth(mk_box_mg_s &M, std::size_t l, const double *b, double *x, bool reverse) {
    auto &lv = M.L[l];
    auto &P = prof();
    const hipStream_t s = stream();
    const double shift = M.p.shift;
    auto forward = [&](const dev_csr<double> &A, const box_level::tri_launch &g, const double *r) {
        gpu::ilu_forward<double>(A.view, dp<u32>(lv.fw_rows), dp<u32>(lv.fw_ptr), g.g0, g.g1, g.wide_rows,
                                 dp<double>(lv.pinv), r, dp<double>(lv.u), s);
    };
    auto backward = [&](const dev_csr<double> &A, const box_level::tri_launch &g) {
        gpu::ilu_backward<double>(A.view, dp<u32>(lv.bw_rows), dp<u32>(lv.bw_ptr), g.g0, g.g1, g.wide_rows,
                                  dp<double>(lv.pinv), dp<double>(lv.u), x, s);
    };
    for (int sweep = 0; sweep < M.p.sweeps; ++sweep) {
        if (M.p.smoother == MK_BOX_GS_LEX) {
            // x += (D + L)^-1 (b - A x) forward, x += (D + U)^-1 (b - A x) backward
            double *r = dp<double>(lv.r), *u = dp<double>(lv.u);
            P.run(level_class(l, K_SWEEP), op_bytes(lv.A, 8, 8, 2), [&] {
                gpu::level_residual<double>(lv.A.view, shift, b, x, r, s);
                if (!reverse) {
                    for (const auto &g : lv.fw[0]) forward(lv.A, g, r);
                    gpu::axpy<double>(lv.n, 1.0, u, x, s);
                } else {
                    gpu::box_scale(lv.n, r, dp<double>(lv.pinv), u, s);
                    for (const auto &g : lv.bw[0]) backward(lv.A, g);
                }
            });
        } else if (M.p.smoother == MK_BOX_BLOCK_ILU) {
            // population blocks in order (reversed after the coarse correction): the block's
            // residual with the current x, then its ILU(0) solve added to x
            double *r = dp<double>(lv.r);
            P.run(level_class(l, K_SWEEP), op_bytes(lv.A, 8, 8, 2) + op_bytes(lv.Ain, 8, 8, 2), [&] {
                const int nb = static_cast<int>(lv.Ablock.size());
                for (int v = 0; v < nb; ++v) {
                    const int k = reverse ? nb - 1 - v : v;
                    if (lv.Ablock[k].n == 0) continue;
                    gpu::level_residual<double>(lv.Ablock[k].view, shift, b, x, r, s);
                    for (const auto &g : lv.fw[k]) forward(lv.Ain, g, r);
                    for (const auto &g : lv.bw[k]) backward(lv.Ain, g);
                }
            });
        } else if (M.p.smoother == MK_BOX_GS) {
            for (int k = 0; k < 4; ++k) {
                const int c = reverse ? 3 - k : k;
                if (lv.cb[c + 1] == lv.cb[c]) continue;
                P.run(level_class(l, K_SWEEP), op_bytes(lv.Ac[c], 8, 8, 1) + (lv.cb[c + 1] - lv.cb[c]) * 4.0,
                      [&] { gpu::gs_half<double>(lv.Ac[c].view, shift, b, x, nullptr, s); });
            }
        } else {
            double *t = dp<double>(lv.tmp);
            P.run(level_class(l, K_SWEEP), op_bytes(lv.A, 8, 8, 2),
                  [&] { gpu::jacobi_sweep<double>(lv.A.view, shift, 2.0 / 3.0, b, x, t, s); });
            P.run(level_class(l, K_OTHER), lv.n * 16.0, [&] { gpu::box_copy(lv.n, t, x, s); });
        }
    }
}

void cycle(mk_box_mg_s &M, std::size_t l, const double *b, double *x);

// K-cycle coarse solve of level l + 1 (rdme_full_multigrid.cpp:856-878)
void krylov_coarse(mk_box_mg_s &M, std::size_t l) {
    auto &lv = M.L[l];
    auto &c = M.L[l + 1];
    auto &P = prof();
    const hipStream_t s = stream();
    const i64 n = c.n;
    const double shift = M.p.shift;
    const double *g = dp<double>(c.b);
    double *xc = dp<double>(c.x);
    double *v1 = dp<double>(lv.kv1), *w1 = dp<double>(lv.kw1), *v2 = dp<double>(lv.kv2),
           *w2 = dp<double>(lv.kw2), *r = dp<double>(lv.kr);
    const double gg = dot(M, g, g, n);
    if (!(gg > 0.0)) return;
    auto apply_coarse = [&](const double *in, double *v, double *w) {
        P.run(K_COARSE, n * 8.0, [&] { gpu::fill<double>(n, 0.0, v, s); });
        cycle(M, l + 1, in, v);
        P.run(K_COARSE, op_bytes(c.A, 8, 8, 1),
              [&] { gpu::shifted_apply<double>(c.A.view, shift, v, nullptr, w, nullptr, s); });
    };
    apply_coarse(g, v1, w1);
    const double ww1 = dot(M, w1, w1, n);
    if (!(ww1 > 0.0)) return;
    const double a1 = dot(M, w1, g, n) / ww1;
    P.run(K_COARSE, n * 48.0, [&] {
        gpu::axpy<double>(n, a1, v1, xc, s);
        gpu::box_copy(n, g, r, s);
        gpu::axpy<double>(n, -a1, w1, r, s);
    });
    if (dot(M, r, r, n) <= 0.0625 * gg) return;
    apply_coarse(r, v2, w2);
    const double beta = dot(M, w2, w1, n) / ww1;
    P.run(K_COARSE, n * 48.0, [&] {
        gpu::axpy<double>(n, -beta, v1, v2, s);
        gpu::axpy<double>(n, -beta, w1, w2, s);
    });
    const double ww2 = dot(M, w2, w2, n);
    if (!(ww2 > 0.0)) return;
    const double a2 = dot(M, w2, r, n) / ww2;
    P.run(K_COARSE, n * 24.0, [&] { gpu::axpy<double>(n, a2, v2, xc, s); });
}

// the descriptor of a small level's fused kernels
gpu::small_level small_desc(mk_box_mg_s &M, std::size_t l, const double *b, double *x) {
    auto &lv = M.L[l];
    auto &c = M.L[l + 1];
    gpu::small_level S;
    S.A = lv.A.view;
    S.n = lv.n;
    for (int k = 0; k < 5; ++k) S.cb[k] = lv.cb[k];
    S.shift = M.p.shift;
    S.sweeps = M.p.sweeps;
    S.jacobi = M.p.smoother == MK_BOX_JACOBI ? 1 : 0;
    S.b = b;
    S.x = x;
    S.r = dp<double>(lv.r);
    S.tmp = dp<double>(lv.tmp);
    S.nc = c.n;
    S.wide = lv.wide_restrict ? 1 : 0;
    S.cptr = dp<u32>(lv.cptr);
    S.cidx = dp<u32>(lv.cidx);
    S.parent = dp<u32>(lv.parent);
    S.weight = dp<double>(lv.weight);
    S.coarse_b = dp<double>(c.b);
    S.coarse_x = dp<double>(c.x);
    return S;
}

void coarse_correction(mk_box_mg_s &M, std::size_t l);

// one cycle of A_l x = b from the guess x (level l's row order)
void cycle(mk_box_mg_s &M, std::size_t l, const double *b, double *x) {
    auto &lv = M.L[l];
    auto &P = prof();
    const hipStream_t s = stream();
    if (lv.small && !lv.terminal && box_small_switch().load() != 0) {
        const auto S = small_desc(M, l, b, x);
        const double bytes = 2.0 * M.p.sweeps * op_bytes(lv.A, 8, 8, 1);
        P.run(level_class(l, K_SWEEP), bytes + op_bytes(lv.A, 8, 8, 2),
              [&] { gpu::box_small_down(S, s); });
        coarse_correction(M, l);
        P.run(level_class(l, K_SWEEP), bytes + lv.n * 28.0, [&] { gpu::box_small_up(S, s); });
        return;
    }
    if (lv.terminal == 1) {
        P.run(K_TERMINAL, lv.n * 40.0, [&] {
            gpu::box_thomas(lv.n, dp<double>(lv.lower), dp<double>(lv.upper), dp<double>(lv.pivot), b, x, s);
        });
        return;
    }
    if (lv.terminal == 2) {
        P.run(K_TERMINAL, double(lv.n) * lv.n * 8.0, [&] {
            gpu::box_dense(lv.n, dp<double>(lv.inv), b, dp<double>(lv.partial), x, s);
        });
        return;
    }
    const double shift = M.p.shift;
    smooth(M, l, b, x, false);
    P.run(level_class(l, K_RESIDUAL), op_bytes(lv.A, 8, 8, 2),
          [&] { gpu::level_residual<double>(lv.A.view, shift, b, x, dp<double>(lv.r), s); });
    auto &c = M.L[l + 1];
    P.run(level_class(l, K_RESTRICT), lv.n * 12.0 + c.n * 12.0, [&] {
        if (lv.wide_restrict) {
            gpu::box_restrict_wave(c.n, dp<u32>(lv.cptr), dp<u32>(lv.cidx), dp<double>(lv.r),
                                   dp<double>(c.b), s);
        } else {
            gpu::restrict_children<double>(c.n, dp<u32>(lv.cptr), dp<u32>(lv.cidx), dp<double>(lv.r),
                                           dp<double>(c.b), s);
        }
        gpu::fill<double>(c.n, 0.0, dp<double>(c.x), s);
    });
    coarse_correction(M, l);
    P.run(level_class(l, K_PROLONG), lv.n * 28.0, [&] {
        gpu::prolong_add<double>(lv.n, dp<u32>(lv.parent), dp<double>(lv.weight), dp<double>(c.x), x, s);
    });
    smooth(M, l, b, x, true);
}

// the correction of level l from level l + 1 (whose b is set and x zero): the exact bottom
// solve once, the K-cycle's GCR steps, or `cycles` coarse cycles
void coarse_correction(mk_box_mg_s &M, std::size_t l) {
    auto &c = M.L[l + 1];
    if (M.p.cycle_type == MK_CYCLE_K && !c.terminal) {
        krylov_coarse(M, l);
    } else {
        const int visits = c.terminal ? 1 : M.p.cycles;
        for (int v = 0; v < visits; ++v) cycle(M, l + 1, dp<double>(c.b), dp<double>(c.x));
    }
}

// |b - A x|_1 / shift on the finest level (row order; the finest level is not terminal)
double error_bound(mk_box_mg_s &M) {
    auto &lv = M.L[0];
    const hipStream_t s = stream();
    const int nb = gpu::reduction_blocks(M.rows);
    prof().run(K_RESIDUAL, op_bytes(lv.A, 8, 8, 2), [&] {
        gpu::level_residual<double>(lv.A.view, M.p.shift, dp<double>(M.b0), dp<double>(M.x0),
                                    dp<double>(lv.r), s);
        gpu::box_abs_partial(M.rows, dp<double>(lv.r), dp<double>(M.red), nb, s);
        gpu::reduce_partials(dp<double>(M.red), nb, 1, dp<double>(M.scal), s);
    });
    HIPCHK(hipMemcpyAsync(M.host, M.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    return M.host[0] / M.p.shift;
}

void run_cycle(mk_box_mg_s &M) {
    const hipStream_t s = stream();
    auto body = [&] { cycle(M, 0, dp<double>(M.b0), dp<double>(M.x0)); };
    if (M.p.cycle_type != MK_CYCLE_K && box_graph_switch().load() != 0 && prof().mask == 0) {
        // the small-level switch selects the kernels captured: it is part of the key
        graph_key key{M.uid, 1, box_small_switch().load(), 0, 0, 0, 0};
        // graph_cache replays only when graphs are enabled globally; capture here directly
        auto it = M.graphs.exec.find(key);
        if (it == M.graphs.exec.end()) {
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
            it = M.graphs.exec.emplace(key, e).first;
        }
        HIPCHK(hipGraphLaunch(it->second, s));
        return;
    }
    body();
}

void enter(mk_box_mg_s &M, const double *b, const double *x, const double *scale = nullptr) {
    const hipStream_t s = stream();
    prof().run(K_OTHER, M.rows * 20.0, [&] {
        gpu::permute_in<double, double>(M.rows, dp<u32>(M.perm0), b, dp<double>(M.b0), s, scale);
        if (x) {
            gpu::permute_in<double, double>(M.rows, dp<u32>(M.perm0), x, dp<double>(M.x0), s);
        } else {
            gpu::fill<double>(M.rows, 0.0, dp<double>(M.x0), s);
        }
    });
}

void leave(mk_box_mg_s &M, double *x) {
    prof().run(K_OTHER, M.rows * 20.0,
               [&] { gpu::permute_out<double>(M.rows, dp<u32>(M.perm0), dp<double>(M.x0), x, stream()); });
}

void check_params(const mk_box_mg_params &p) {
    require(p.dimension >= 1 && p.dimension <= 3, "box multigrid: dimension must be 1, 2 or 3");
    require(p.cells >= 1 && (p.cells & (p.cells - 1)) == 0, "box multigrid: cells must be a power of two");
    require(p.cap >= 0, "box multigrid: negative population cap");
    require(p.shift > 0.0 && std::isfinite(p.shift), "box multigrid: shift must be positive");
    require(p.diffusion >= 0.0 && p.c1 >= 0.0 && p.c2 >= 0.0 && p.c3 >= 0.0 && p.c4 >= 0.0,
            "box multigrid: negative rate constant");
    require(p.smoother == MK_BOX_GS || p.smoother == MK_BOX_JACOBI || p.smoother == MK_BOX_GS_LEX ||
                p.smoother == MK_BOX_BLOCK_ILU,
            "box multigrid: unknown smoother");
    require(p.sweeps >= 1, "box multigrid: at least one sweep");
    require(p.cycles == 1 || p.cycles == 2, "box multigrid: cycles must be 1 (V) or 2 (W)");
    require(p.cycle_type == MK_CYCLE_STANDARD || p.cycle_type == MK_CYCLE_K,
            "box multigrid: unknown cycle type");
    require(p.terminal_rows >= 0 && p.terminal_rows <= 8192,
            "box multigrid: terminal_rows must lie in [0, 8192]");
    i64 V = 1;
    for (int d = 0; d < p.dimension; ++d) {
        require(V <= (i64{1} << 20) / p.cells, "box multigrid: too many voxels");
        V *= p.cells;
    }
}

// the standalone iteration from the guess x (vectors in the box order)
mk_box_mg_result solve(mk_box_mg_s &M, const double *b, double *x, double tolerance,
                       int32_t max_iterations, double *history, int64_t history_cap) {
    require(max_iterations >= 0, "box multigrid: negative iteration limit");
    require(!M.L[0].terminal, "box multigrid: the finest level is solved exactly; use apply");
    enter(M, b, x);
    mk_box_mg_result r{0, 0, 0.0, 0};
    double err = error_bound(M);
    int64_t h = 0;
    while (r.iterations < max_iterations) {
        if (history && h < history_cap) history[h] = err;
        ++h;
        if (err < tolerance || !std::isfinite(err) || err > 1e8) break;
        run_cycle(M);
        ++r.iterations;
        err = error_bound(M);
    }
    leave(M, x);
    drain();
    r.error = err;
    r.converged = err < tolerance ? 1 : 0;
    r.history = h;
    return r;
}

} // namespace
} // namespace mk::rt

// ====================================================================== C ABI
extern "C" {

mk_status mk_box_mg_create(const mk_box_mg_params *params, mk_box_mg *out) {
    return guarded([&] {
        require(params && out, "box multigrid: bad arguments");
        check_params(*params);
        auto M = std::make_unique<mk_box_mg_s>();
        M->p = *params;
        build(*M);
        *out = M.release();
    });
}

mk_status mk_box_mg_destroy(mk_box_mg mg) {
    delete mg;
    return MK_OK;
}

mk_status mk_box_mg_info(mk_box_mg mg, int64_t *rows, int32_t *levels) {
    return guarded([&] {
        require(mg, "box multigrid: null handle");
        if (rows) *rows = mg->rows;
        if (levels) *levels = static_cast<int32_t>(mg->L.size());
    });
}

mk_status mk_box_mg_level(mk_box_mg mg, int32_t level, int64_t *rows, int64_t *voxels,
                          int64_t *entries, int32_t *terminal) {
    return guarded([&] {
        require(mg && level >= 0 && level < static_cast<int32_t>(mg->L.size()),
                "box multigrid: no such level");
        const auto &lv = mg->L[level];
        if (rows) *rows = lv.n;
        if (voxels) *voxels = lv.voxels;
        if (entries) *entries = static_cast<int64_t>(lv.nnz);
        if (terminal) *terminal = lv.terminal;
    });
}

mk_status mk_box_mg_generator(mk_box_mg mg, mk_generator *gen) {
    return guarded([&] {
        require(mg && gen, "box multigrid: bad arguments");
        *gen = mg->gen.get();
    });
}

mk_status mk_box_mg_cycle_device(mk_box_mg mg, const double *d_b, double *d_x) {
    return guarded([&] {
        require(mg && d_b && d_x, "box multigrid: bad arguments");
        enter(*mg, d_b, d_x);
        cycle(*mg, 0, dp<double>(mg->b0), dp<double>(mg->x0));
        leave(*mg, d_x);
        drain();
    });
}

mk_status mk_box_mg_apply_device(mk_box_mg mg, const double *d_b, double *d_z) {
    return guarded([&] {
        require(mg && d_b && d_z, "box multigrid: bad arguments");
        enter(*mg, d_b, nullptr);
        cycle(*mg, 0, dp<double>(mg->b0), dp<double>(mg->x0));
        leave(*mg, d_z);
        drain();
    });
}

mk_status mk_box_mg_solve_device(mk_box_mg mg, const double *d_b, double *d_x, double tolerance,
                                 int32_t max_iterations, double *history, int64_t history_cap,
                                 mk_box_mg_result *result) {
    return guarded([&] {
        require(mg && d_b && d_x && result, "box multigrid: bad arguments");
        *result = solve(*mg, d_b, d_x, tolerance, max_iterations, history, history_cap);
    });
}

mk_status mk_box_mg_step_device(mk_box_mg mg, double *d_p, int32_t crank_nicolson,
                                double tolerance, int32_t max_iterations, mk_box_mg_result *result) {
    return guarded([&] {
        require(mg && d_p && result, "box multigrid: bad arguments");
        auto &M = *mg;
        const i64 n = M.rows;
        if (!M.step_b.p) {
            M.step_b = dbuf(n * sizeof(double));
            M.step_x = dbuf(n * sizeof(double));
        }
        const hipStream_t s = stream();
        double *b = dp<double>(M.step_b), *x = dp<double>(M.step_x);
        prof().run(K_OTHER, op_bytes(M.gen->A, 8, 8, 2), [&] {
            if (crank_nicolson) {
                gpu::shifted_apply<double>(M.gen->A.view, M.p.shift, d_p, nullptr, x, nullptr, s);
                gpu::box_rhs(n, M.p.shift, d_p, x, b, s);
            } else {
                gpu::box_rhs(n, M.p.shift, d_p, nullptr, b, s);
            }
            gpu::box_copy(n, d_p, x, s);
        });
        *result = solve(M, b, x, tolerance, max_iterations, nullptr, 0);
        if (result->converged) {
            prof().run(K_OTHER, n * 16.0, [&] { gpu::box_copy(n, x, d_p, s); });
            drain();
        }
    });
}

mk_status mk_box_mg_solve(mk_box_mg mg, const double *b, double *x, double tolerance,
                          int32_t max_iterations, double *history, int64_t history_cap,
                          mk_box_mg_result *result) {
    mk_status st = MK_OK;
    const mk_status outer = guarded([&] {
        require(mg && b && x, "box multigrid: bad arguments");
        const std::size_t bytes = mg->rows * sizeof(double);
        dbuf db(bytes), dx(bytes);
        HIPCHK(hipMemcpy(db.p, b, bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
        st = mk_box_mg_solve_device(mg, dp<double>(db), dp<double>(dx), tolerance, max_iterations,
                                    history, history_cap, result);
        if (st == MK_OK) HIPCHK(hipMemcpy(x, dx.p, bytes, hipMemcpyDeviceToHost));
    });
    return outer != MK_OK ? outer : st;
}

mk_status mk_gmres_box_device(mk_generator gen, double shift, mk_box_mg mg, const double *d_b,
                              double *d_x, const mk_gmres_options *options,
                              mk_solver_result *result) {
    return guarded([&] {
        require(gen && mg && d_b && d_x && options && result, "gmres: bad arguments");
        require(mg->p.cycle_type == MK_CYCLE_STANDARD,
                "gmres: the K-cycle is nonlinear and runs only as a standalone iteration");
        mk_box_mg_s *m = mg;
        precond M{m->uid, m->rows, m->p.shift, MK_FP64, [m](const double *v, void *z, const double *scale) {
                      enter(*m, v, nullptr, scale);
                      cycle(*m, 0, dp<double>(m->b0), dp<double>(m->x0));
                      leave(*m, static_cast<double *>(z));
                  }};
        *result = gmres_device(*gen, shift, &M, d_b, d_x, *options);
    });
}

mk_status mk_gmres_box(mk_generator gen, double shift, mk_box_mg mg, const double *b, double *x,
                       const mk_gmres_options *options, mk_solver_result *result) {
    mk_status st = MK_OK;
    const mk_status outer = guarded([&] {
        require(gen && mg && b && x, "gmres: bad arguments");
        const std::size_t bytes = gen->n * sizeof(double);
        dbuf db(bytes), dx(bytes);
        HIPCHK(hipMemcpy(db.p, b, bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
        st = mk_gmres_box_device(gen, shift, mg, dp<double>(db), dp<double>(dx), options, result);
        if (st == MK_OK) HIPCHK(hipMemcpy(x, dx.p, bytes, hipMemcpyDeviceToHost));
    });
    return outer != MK_OK ? outer : st;
}

mk_status mk_box_mg_small_enable(int32_t on) {
    box_small_switch().store(on ? 1 : 0);
    return MK_OK;
}

mk_status mk_box_mg_graphs_enable(int32_t on) {
    box_graph_switch().store(on ? 1 : 0);
    return MK_OK;
}

} // extern "C"
