// Batched resolvent-solve kernels (gfx950, wave64): lane = system, wavefront = one state row of
// 64 systems. See batch_kernels.hpp for the layout.
#include "batch_kernels.hpp"

#include <algorithm>

namespace mk::gpu::batch {

namespace {

constexpr int kWaves = kThreads / kLanes;
constexpr int kEntryUnroll = 8;
constexpr int kRowsPerPart = 256;
constexpr int kMaxParts = 512;
// parts_wide: the row partitions of the reductions of a small operand (a workgroup per partition
// and lane group, so that one launch covers the CUs)
constexpr int kWideParts = 1024;

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

enum class bmode { shifted, shifted_residual, sweep, residual, gather_sum, gs };

// limit (when not null): lane j works on the rows r < limit[j] of the operator's level (0: an
// inactive system). Rows are the level's rows: a row-mapped operator (A.rowmap) maps its row q
// to level row rowmap[q], which indexes the vectors and the limit.
template <typename VT, typename XT, typename T>
struct bop_args {
    csr_view<VT> A;
    const double *shifts;
    T omega;
    const XT *x;
    const T *rhs;
    T *out;
    double *partial;
    i64 Bp;
    const int *limit;
    T *zero = nullptr;      // gs: rows also cleared in this vector (the residual after the last
                            // pre-smoothing half-sweep)
    T *nat = nullptr;       // gs: rows also stored to nat[perm[row]] (the caller's order)
    const u32 *perm = nullptr;
    const double *scale = nullptr; // shifted: the product of lane j times scale[j]
};

// sum over the row's entries of rate * x[column][j], ascending entry order; row r is
// wave-uniform, so the entry, length and dictionary loads are scalar
template <typename VT, typename XT, typename T>
__device__ __forceinline__ T row_gather(const csr_view<VT> &A, i64 r, const XT *x, i64 Bp, i64 j,
                                        bool on, T &exit_r) {
    T sum = 0;
    if (A.coded) {
        const i64 slice = r >> 6;
        const u32 len = A.rlen[r];
        const u32 base = A.soff[slice] + static_cast<u32>(r & 63);
        exit_r = static_cast<T>(A.exit ? A.exit[r] : A.edict[A.rcode[r]]);
        for (u32 k0 = 0; k0 < len; k0 += kEntryUnroll) {
            u32 e[kEntryUnroll];
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u)
                e[u] = k0 + u < len ? A.ent[base + (k0 + u) * kSlice] : 0u;
            XT xv[kEntryUnroll];
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u)
                xv[u] = (on && k0 + u < len) ? x[static_cast<i64>(e[u] & kColMask) * Bp + j] : XT(0);
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u)
                if (k0 + u < len) sum += static_cast<T>(A.vdict[e[u] >> kColBits]) * static_cast<T>(xv[u]);
        }
    } else {
        const u32 p0 = A.ptr[r], p1 = A.ptr[r + 1];
        exit_r = A.exit ? static_cast<T>(A.exit[r]) : T(0);
        for (u32 p = p0; p < p1; p += kEntryUnroll) {
            u32 c[kEntryUnroll];
            T v[kEntryUnroll];
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u) {
                c[u] = p + u < p1 ? A.col[p + u] : 0u;
                v[u] = p + u < p1 ? static_cast<T>(A.val[p + u]) : T(0);
            }
            XT xv[kEntryUnroll];
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u)
                xv[u] = (on && p + u < p1) ? x[static_cast<i64>(c[u]) * Bp + j] : XT(0);
#pragma unroll
            for (int u = 0; u < kEntryUnroll; ++u)
                if (p + u < p1) sum += v[u] * static_cast<T>(xv[u]);
        }
    }
    return sum;
}

// grid (ceil(n / 4), Bp / 64): one state row per wavefront, so the dependent entry -> gather
// chains of many rows are in flight at once
template <bmode M, typename VT, typename XT, typename T>
__global__ void __launch_bounds__(kThreads) k_bop(bop_args<VT, XT, T> a) {
    __shared__ double red[kWaves][kLanes];
    const int lane = threadIdx.x & 63, wave = wave_id();
    const i64 j = static_cast<i64>(blockIdx.y) * kLanes + lane;
    const i64 r = static_cast<i64>(blockIdx.x) * kWaves + wave;
    double sq = 0.0;
    if (r < a.A.n) {
        const i64 row = a.A.rowmap ? static_cast<i64>(a.A.rowmap[r]) : r;
        const bool on = !a.limit || row < a.limit[j];
        T ex = 0;
        const T sum = row_gather<VT, XT, T>(a.A, r, a.x, a.Bp, j, on, ex);
        if (on) {
            const i64 o = row * a.Bp + j;
            if constexpr (M == bmode::gather_sum) {
                a.out[o] = sum;
            } else if constexpr (M == bmode::gs) {
                // x_row = (rhs + sum over the other colour) / (s_j + exit), in place
                const T v = (a.rhs[o] + sum) / (static_cast<T>(a.shifts[j]) + ex);
                a.out[o] = v;
                if (a.zero) a.zero[o] = T(0);
                if (a.nat) a.nat[static_cast<i64>(a.perm[row]) * a.Bp + j] = v;
            } else {
                const T xr = static_cast<T>(a.x[o]);
                const T d = static_cast<T>(a.shifts[j]) + ex;
                const T ax = d * xr - sum;
                if constexpr (M == bmode::shifted) {
                    a.out[o] = a.scale ? ax * static_cast<T>(a.scale[j]) : ax;
                } else if constexpr (M == bmode::shifted_residual) {
                    const T res = a.rhs[o] - ax;
                    a.out[o] = res;
                    sq = static_cast<double>(res) * static_cast<double>(res);
                } else if constexpr (M == bmode::sweep) {
                    const T res = a.rhs[o] - ax;
                    a.out[o] = xr + a.omega * res / d;
                } else {
                    a.out[o] = a.rhs[o] - ax;
                }
            }
        }
    }
    if constexpr (M == bmode::shifted_residual) {
        red[wave][lane] = sq;
        __syncthreads();
        if (wave == 0 && a.partial)
            a.partial[static_cast<i64>(blockIdx.x) * a.Bp + j] =
                ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    }
}

i64 row_blocks(i64 n) { return std::max<i64>(1, (n + kWaves - 1) / kWaves); }

template <bmode M, typename VT, typename XT, typename T>
void launch(const bop_args<VT, XT, T> &a, hipStream_t s) {
    if (a.A.n == 0) return;
    hipLaunchKernelGGL((k_bop<M, VT, XT, T>), dim3(static_cast<u32>(row_blocks(a.A.n)), a.Bp / kLanes),
                       dim3(kThreads), 0, s, a);
}

int op_parts(i64 n) {
    return static_cast<int>(std::max<i64>(1, std::min<i64>((n + kRowsPerPart - 1) / kRowsPerPart, kMaxParts)));
}

template <typename T>
__device__ __forceinline__ T row_exit(const csr_view<T> &A, i64 r) {
    return A.exit ? A.exit[r] : A.edict[A.rcode[r]];
}

int grid_for(i64 n) {
    const i64 g = (n + kThreads - 1) / kThreads;
    return static_cast<int>(std::max<i64>(1, std::min<i64>(g, 16384)));
}

// elementwise kernels: grid (ceil(n / 4), Bp / 64), wavefront = row, lane = system
__device__ __forceinline__ void row_lane(i64 &r, i64 &j) {
    r = static_cast<i64>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    j = static_cast<i64>(blockIdx.y) * kLanes + (threadIdx.x & 63);
}

template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_jacobi_first(csr_view<T> A, const double *shifts, T omega, const IN *in,
                   const double *scale, T *rhs_out, T *x, i64 Bp, const int *limit) {
    i64 r, j;
    row_lane(r, j);
    if (r >= A.n || (limit && r >= limit[j])) return;
    const i64 o = r * Bp + j;
    const IN sc = scale ? static_cast<IN>(scale[j]) : IN(1);
    const T b = static_cast<T>(in[o] * sc);
    if (rhs_out) rhs_out[o] = b;
    x[o] = omega * b / (static_cast<T>(shifts[j]) + row_exit(A, r));
}

// the first red-black half-sweep from x = 0: rhs[i] = (T)(scale_j in[perm[i]]) (perm null:
// identity), stored to rhs_out if given; x[i] = rhs[i] / (s_j + exit_i) on colour 0 (colour-1
// entries are set by the next half-sweep)
template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_rb_first(csr_view<T> A, const u8 *colour, const double *shifts, const IN *in,
               const double *scale, T *rhs_out, T *x, const u32 *perm, i64 Bp, const int *limit) {
    i64 r, j;
    row_lane(r, j);
    if (r >= A.n || (limit && r >= limit[j])) return;
    const i64 o = r * Bp + j;
    const i64 src = perm ? static_cast<i64>(perm[r]) * Bp + j : o;
    const IN sc = scale ? static_cast<IN>(scale[j]) : IN(1);
    const T b = static_cast<T>(in[src] * sc);
    if (rhs_out) rhs_out[o] = b;
    if (colour[r] == 0) x[o] = b / (static_cast<T>(shifts[j]) + row_exit(A, r));
}

// x[i] += weight[i] coarse[coarse_of[i]] for the n rows i = rows[q] (all rows when rows is null)
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_prolong_add(i64 n, const u32 *coarse_of, const T *weight, const T *coarse, T *x, i64 Bp,
                  const int *limit, const u32 *rows) {
    i64 q, j;
    row_lane(q, j);
    if (q >= n) return;
    const i64 i = rows ? static_cast<i64>(rows[q]) : q;
    if (limit && i >= limit[j]) return;
    const u32 c = coarse_of[i];
    if (c != 0xFFFFFFFFu) x[i * Bp + j] += weight[i] * coarse[static_cast<i64>(c) * Bp + j];
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_terminal(i64 nb, const u32 *start, const u32 *size, const u64 *offset, const T *inv,
               u64 per_shift, const int *shift_index, i64 max_size, const T *r, T *x, i64 Bp,
               const int *limit) {
    const i64 total = nb * max_size * Bp;
    for (i64 g = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; g < total;
         g += static_cast<i64>(gridDim.x) * kThreads) {
        const i64 j = g % Bp, bi = g / Bp;
        const i64 b = bi / max_size, i = bi % max_size;
        const u32 m = size[b];
        const i64 s0 = start[b];
        // a block lies in one sector: it is inside a lane's rows or past them
        if (i >= m || (limit && s0 >= limit[j])) continue;
        const T *row = inv + static_cast<u64>(shift_index[j]) * per_shift + offset[b] +
                       static_cast<u64>(i) * m;
        T acc = 0;
        for (u32 c = 0; c < m; ++c) acc += row[c] * r[(s0 + c) * Bp + j];
        x[(s0 + i) * Bp + j] += acc;
    }
}

// ---- Krylov
// grid (np, chunks of 8 basis vectors, Bp / 64)
__global__ void __launch_bounds__(kThreads)
    k_multi_dot(const double *V, int count, const double *w, i64 n, i64 Bp, double *partial,
                int np, const int *limit) {
    __shared__ double red[kWaves][kDotChunk][kLanes];
    const int lane = threadIdx.x & 63, wave = wave_id();
    const i64 j = static_cast<i64>(blockIdx.z) * kLanes + lane;
    const i64 rows = limit ? min(n, static_cast<i64>(limit[j])) : n;
    const int k0 = blockIdx.y * kDotChunk;
    const int kc = min(kDotChunk, count - k0);
    const i64 ld = n * Bp;
    double acc[kDotChunk];
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) acc[k] = 0.0;
    {
        // rows r, r + stride, ... in order; four rows' loads in flight at a time (the sums
        // keep the row order)
        const i64 stride = static_cast<i64>(np) * kWaves;
        i64 r = static_cast<i64>(blockIdx.x) * kWaves + wave;
        for (; r + 3 * stride < rows; r += 4 * stride) {
            double wr[4], v[4][kDotChunk];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const i64 o = (r + u * stride) * Bp + j;
                wr[u] = w[o];
#pragma unroll
                for (int k = 0; k < kDotChunk; ++k) v[u][k] = k < kc ? V[(k0 + k) * ld + o] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < kDotChunk; ++k)
                    if (k < kc) acc[k] += v[u][k] * wr[u];
        }
        for (; r < rows; r += stride) {
            const i64 o = r * Bp + j;
            const double wr = w[o];
#pragma unroll
            for (int k = 0; k < kDotChunk; ++k)
                if (k < kc) acc[k] += V[(k0 + k) * ld + o] * wr;
        }
    }
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) red[wave][k][lane] = acc[k];
    __syncthreads();
    if (wave == 0) {
        const i64 m = static_cast<i64>(count) * Bp;
        for (int k = 0; k < kc; ++k)
            partial[static_cast<i64>(blockIdx.x) * m + (k0 + k) * Bp + j] =
                ((red[0][k][lane] + red[1][k][lane]) + red[2][k][lane]) + red[3][k][lane];
    }
}

// first stage of a long reduction: tmp[g * m + o] = sum of partitions g * kGroup ..
// (g + 1) * kGroup - 1 for output o; grid (ceil(np / kGroup), ceil(m / 64)), waves split the
// group's partitions, lanes own outputs
constexpr int kGroup = 256;

__global__ void __launch_bounds__(kThreads)
    k_reduce_groups(const double *partial, int np, i64 m, double *tmp) {
    __shared__ double red[kWaves][kLanes];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const i64 o = static_cast<i64>(blockIdx.y) * kLanes + lane;
    const int p0 = blockIdx.x * kGroup, p1 = min(np, p0 + kGroup);
    double acc = 0.0;
    if (o < m) {
        // this wave's partitions in order, four loads in flight at a time
        int p = p0 + wave;
        for (; p + 3 * kWaves < p1; p += 4 * kWaves) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = partial[static_cast<i64>(p + u * kWaves) * m + o];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u];
        }
        for (; p < p1; p += kWaves) acc += partial[static_cast<i64>(p) * m + o];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && o < m)
        tmp[static_cast<i64>(blockIdx.x) * m + o] =
            ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ void __launch_bounds__(kThreads)
    k_reduce(const double *partial, int np, i64 m, double *out) {
    for (i64 o = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; o < m;
         o += static_cast<i64>(gridDim.x) * kThreads) {
        // partitions in order, eight loads in flight at a time
        double s = 0.0;
        int p = 0;
        for (; p + 8 <= np; p += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partial[static_cast<i64>(p + u) * m + o];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; p < np; ++p) s += partial[static_cast<i64>(p) * m + o];
        out[o] = s;
    }
}

// delayed CGS2 step, grid (np, Bp / 64): partition p takes rows p, p + np, ...; wave wv keeps
// the basis slots k = 4 kk + wv (clamped to count - 1, zero coefficients past count)
template <int PW>
__global__ void __launch_bounds__(kThreads)
    k_dcgs_update(double *V, int count, const double *sc, const double *tg, const double *beta,
                  const double *w, i64 n, i64 Bp, double *partial, int np, double *eta_out,
                  const int *limit) {
    // rows per pair of barriers: four while few basis values sit in registers, else one (more
    // rows' values in registers cost occupancy: measured slower from PW = 3 on)
    constexpr int kRows = PW <= 2 ? 4 : 1;
    __shared__ double part[kRows][2][kWaves][kLanes];
    __shared__ double uw[kRows][2][kLanes];
    __shared__ double red[2][kLanes];
    constexpr int kSlots = PW > 0 ? PW : 1;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const i64 j = static_cast<i64>(blockIdx.y) * kLanes + lane;
    const i64 rows = limit ? min(n, static_cast<i64>(limit[j])) : n;
    const bool on = rows > 0;
    const i64 ld = n * Bp;
    double cs[kSlots], ct[kSlots], acc[kSlots];
    const double *src[kSlots];
#pragma unroll
    for (int kk = 0; kk < PW; ++kk) {
        const int k = kk * kWaves + wave;
        cs[kk] = (on && k < count) ? sc[static_cast<i64>(k) * Bp + j] : 0.0;
        ct[kk] = (on && k < count) ? tg[static_cast<i64>(k) * Bp + j] : 0.0;
        src[kk] = V + static_cast<i64>(min(k, count - 1)) * ld;
        acc[kk] = 0.0;
    }
    double eta = 0.0, ib = 0.0;
    if (on) {
        double st = 0.0;
        for (int k = 0; k < count; ++k) st += sc[static_cast<i64>(k) * Bp + j] * tg[static_cast<i64>(k) * Bp + j];
        const double b = beta[j];
        eta = (tg[static_cast<i64>(count) * Bp + j] - st) / b;
        ib = 1.0 / b;
        if (blockIdx.x == 0 && wave == 0) eta_out[j] = eta;
    }
    double *const q_out = V + static_cast<i64>(count) * ld;
    double *const u_out = V + static_cast<i64>(count + 1) * ld;
    double own = 0.0; // wave 2: q . u', wave 3: u' . u'
    // every wave runs every row (the loop holds barriers); lanes past their rows compute zeros.
    // RB rows per pair of barriers (rows r0, r0 + np, ...; the sums keep the row order)
    constexpr int RB = kRows;
    for (i64 r0 = blockIdx.x; r0 < n; r0 += RB * static_cast<i64>(np)) {
        double v[RB][kSlots];
        bool live[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const i64 r = r0 + u * static_cast<i64>(np);
            const i64 o = r * Bp + j;
            live[u] = r < rows;
#pragma unroll
            for (int kk = 0; kk < PW; ++kk) v[u][kk] = live[u] ? src[kk][o] : 0.0;
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int kk = 0; kk < PW; ++kk) {
                a += cs[kk] * v[u][kk];
                b += ct[kk] * v[u][kk];
            }
            part[u][0][wave][lane] = a;
            part[u][1][wave][lane] = b;
            if (wave < 2) uw[u][wave][lane] = live[u] ? (wave == 0 ? q_out[o] : w[o]) : 0.0;
        }
        __syncthreads();
        double q[RB], up[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const double S = ((part[u][0][0][lane] + part[u][0][1][lane]) + part[u][0][2][lane]) + part[u][0][3][lane];
            const double T = ((part[u][1][0][lane] + part[u][1][1][lane]) + part[u][1][2][lane]) + part[u][1][3][lane];
            q[u] = live[u] ? (uw[u][0][lane] - S) * ib : 0.0;
            up[u] = live[u] ? (uw[u][1][lane] - T) - eta * q[u] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const i64 o = (r0 + u * static_cast<i64>(np)) * Bp + j;
            if (live[u]) {
                if (wave == 0) q_out[o] = q[u];
                if (wave == 1) u_out[o] = up[u];
            }
            if (wave == 2) own += q[u] * up[u];
            if (wave == 3) own += up[u] * up[u];
#pragma unroll
            for (int kk = 0; kk < PW; ++kk) acc[kk] += v[u][kk] * up[u];
        }
    }
    const i64 m = static_cast<i64>(count + 2) * Bp;
    double *out = partial + static_cast<i64>(blockIdx.x) * m;
#pragma unroll
    for (int kk = 0; kk < PW; ++kk) {
        const int k = kk * kWaves + wave;
        if (k < count) out[static_cast<i64>(k) * Bp + j] = acc[kk];
    }
    if (wave >= 2) red[wave - 2][lane] = own;
    __syncthreads();
    if (wave == 0) {
        out[static_cast<i64>(count) * Bp + j] = red[0][lane];
        out[static_cast<i64>(count + 1) * Bp + j] = red[1][lane];
    }
}

__global__ void __launch_bounds__(kThreads)
    k_basis_update(double *w, const double *V, int count, const double *h, i64 n, i64 Bp,
                   const int *limit) {
    i64 r, j;
    row_lane(r, j);
    if (r >= n || (limit && r >= limit[j])) return;
    const i64 o = r * Bp + j, ld = n * Bp;
    double wi = w[o];
    for (int k = 0; k < count; ++k) wi -= h[static_cast<i64>(k) * Bp + j] * V[k * ld + o];
    w[o] = wi;
}

template <typename ZT>
__global__ void __launch_bounds__(kThreads)
    k_combine(double *x, const ZT *Z, const int *cnt, const double *y, i64 n, i64 Bp) {
    i64 r, j;
    row_lane(r, j);
    if (r >= n) return;
    const int c = cnt[j];
    if (c <= 0) return;
    const i64 o = r * Bp + j, ld = n * Bp;
    double s = x[o];
    for (int k = 0; k < c; ++k) s += y[static_cast<i64>(k) * Bp + j] * static_cast<double>(Z[k * ld + o]);
    x[o] = s;
}

// 64 x 64 tiles through LDS; grid (ceil(n / 64), Bp / 64)
constexpr int kT = 64;

__global__ void __launch_bounds__(kThreads)
    k_pack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp) {
    __shared__ double t[kT][kT + 1];
    const i64 i0 = static_cast<i64>(blockIdx.x) * kT, j0 = static_cast<i64>(blockIdx.y) * kT;
    for (int e = threadIdx.x; e < kT * kT; e += kThreads) {
        const int jj = e / kT, ii = e % kT; // read along states (contiguous in src)
        const i64 i = i0 + ii, j = j0 + jj;
        t[jj][ii] = (i < n && j < nsys) ? src[j * n + i] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kT * kT; e += kThreads) {
        const int ii = e / kT, jj = e % kT; // write along systems (contiguous in dst)
        const i64 i = i0 + ii, j = j0 + jj;
        if (i < n) dst[i * Bp + j] = t[jj][ii];
    }
}

__global__ void __launch_bounds__(kThreads)
    k_unpack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp) {
    __shared__ double t[kT][kT + 1];
    const i64 i0 = static_cast<i64>(blockIdx.x) * kT, j0 = static_cast<i64>(blockIdx.y) * kT;
    for (int e = threadIdx.x; e < kT * kT; e += kThreads) {
        const int ii = e / kT, jj = e % kT; // read along systems
        const i64 i = i0 + ii, j = j0 + jj;
        t[ii][jj] = i < n ? src[i * Bp + j] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kT * kT; e += kThreads) {
        const int jj = e / kT, ii = e % kT; // write along states
        const i64 i = i0 + ii, j = j0 + jj;
        if (i < n && j < nsys) dst[j * n + i] = t[ii][jj];
    }
}

// ---- block Krylov (groups of kBlock consecutive lanes share one basis) ----
// Projections of a block basis on a block: H[i][j][a] = sum_r V_i[r][4 g + a] W[r][j] for every
// lane j of group g = j / 4, i < count, on FP64 matrix cores. A wave takes 16 lanes (4 groups)
// and 4 rows per v_mfma_f64_16x16x4f64 step, the rows being the reduction dimension:
//   A[m][k] = V_i[r + k][16 s + m] (lane l: m = l % 16, k = l / 16),
//   B[k][n] = W[r + k][16 s + n]   (the same element position),
//   D[m][n] = sum_k A[m][k] B[k][n] (lane l, register q: m = 4 q + l / 16, n = l % 16),
// so lane l's register q = (l % 16) / 4 holds the group-diagonal entry m = 4 q + l / 16 of column n:
// H[i][16 s + n][a = l / 16]. The off-diagonal 4 x 4 blocks (lanes of different groups) are
// computed and dropped. grid (np partitions of the rows, basis chunks of kDotChunk, Bp / 64).
typedef double mk_double4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double pick4(const mk_double4 &v, int q) {
    return q == 0 ? v[0] : q == 1 ? v[1] : q == 2 ? v[2] : v[3];
}

__global__ void __launch_bounds__(kThreads)
    k_block_dot(const double *V, int count, const double *W, i64 n, i64 Bp, double *partial, int np,
                const int *limit) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    const i64 col = static_cast<i64>(blockIdx.z) * kLanes + wave * 16 + (lane & 15);
    const int kr = lane >> 4;
    const i64 lim = limit ? min(n, static_cast<i64>(limit[col])) : n;
    const int k0 = blockIdx.y * kDotChunk;
    const int kc = min(kDotChunk, count - k0);
    const i64 ld = n * Bp;
    const i64 per = ((n + np - 1) / np + 3) / 4 * 4;
    const i64 r0 = static_cast<i64>(blockIdx.x) * per, r1 = min(n, r0 + per);
    mk_double4 acc[kDotChunk];
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) acc[k] = mk_double4{0.0, 0.0, 0.0, 0.0};
    for (i64 r = r0; r < r1; r += 4) {
        const i64 row = r + kr;
        const bool ok = row < r1 && row < lim;
        const i64 o = row * Bp + col;
        const double w = ok ? W[o] : 0.0;
        double v[kDotChunk];
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k) v[k] = (ok && k < kc) ? V[(k0 + k) * ld + o] : 0.0;
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k)
            acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[k], w, acc[k], 0, 0, 0);
    }
    const int q = (lane & 15) >> 2;
    const i64 m = static_cast<i64>(count) * Bp * kBlock;
    double *out = partial + static_cast<i64>(blockIdx.x) * m;
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k)
        if (k < kc) out[((k0 + k) * Bp + col) * kBlock + kr] = pick4(acc[k], q);
}

// out[r][j] = sign sum_{i < count} sum_a V_i[r][4 g + a] C[i][j][a] (+ out[r][j] unless assign), for
// the rows r < limit[j] of every lane j (group g = j / 4): the block updates W -= V H, the scaling
// V_{k+1} = W R^-1 and the solution update X += Z Y
template <typename IN>
__global__ void __launch_bounds__(kThreads)
    k_block_apply(double *out, const IN *V, int count, const double *C, i64 n, i64 Bp, double sign,
                  int assign, const int *limit, const int *counts) {
    i64 r, j;
    row_lane(r, j);
    if (r >= n || (limit && r >= limit[j])) return;
    const i64 ld = n * Bp, base = r * Bp + (j & ~static_cast<i64>(kBlock - 1));
    const int cnt = counts ? counts[j] : count;
    double s = 0.0;
    for (int i = 0; i < cnt; ++i) {
        const IN *v = V + i * ld + base;
        const double *c = C + (static_cast<i64>(i) * Bp + j) * kBlock;
        s += ((static_cast<double>(v[0]) * c[0] + static_cast<double>(v[1]) * c[1]) +
              static_cast<double>(v[2]) * c[2]) + static_cast<double>(v[3]) * c[3];
    }
    const i64 o = r * Bp + j;
    out[o] = assign ? sign * s : out[o] + sign * s;
}

// y[r][j] = x[r][j] / sqrt(norm2[j]) for rows r < limit[j]; 0 where norm2[j] is not positive
__global__ void __launch_bounds__(kThreads)
    k_scale_lanes(const double *x, const double *norm2, double *y, i64 n, i64 Bp, const int *limit) {
    i64 r, j;
    row_lane(r, j);
    if (r >= n || (limit && r >= limit[j])) return;
    const double q = norm2[j];
    const i64 o = r * Bp + j;
    y[o] = q > 0.0 ? x[o] / std::sqrt(q) : 0.0;
}

// dst[r][j] = r < rows ? src[r] : 0 for one lane j
__global__ void __launch_bounds__(kThreads)
    k_put_lane(const double *src, i64 rows, double *dst, i64 n, i64 Bp, i64 j) {
    for (i64 r = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; r < n;
         r += static_cast<i64>(gridDim.x) * kThreads)
        dst[r * Bp + j] = r < rows ? src[r] : 0.0;
}

__global__ void __launch_bounds__(kThreads)
    k_get_lane(const double *src, i64 rows, i64 Bp, i64 j, double *dst) {
    for (i64 r = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; r < rows;
         r += static_cast<i64>(gridDim.x) * kThreads)
        dst[r] = src[r * Bp + j];
}

// X_k[r][j] = 0 for r >= rows[j], k < count (vectors ld apart), lanes with sel[j] != 0 (all when
// sel is null); grid (ceil(n / 4), Bp / 64)
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_clear_tails(T *X, int count, i64 ld, i64 n, i64 Bp, const int *rows, const int *sel) {
    i64 r, j;
    row_lane(r, j);
    if (r >= n || r < rows[j] || (sel && !sel[j])) return;
    for (int k = 0; k < count; ++k) X[k * ld + r * Bp + j] = T(0);
}

// the delayed-CGS2 candidate's test per lane: s = sn[0 .. count], |u|^2 = sn[count + 1]; the
// Pythagorean norm |u|^2 - |s|^2 is kept unless |s|^2 > |u|^2 / 4 (flag: the candidate gets an
// explicit second projection); lflag[j] = the lane's limit where flagged, else 0
__global__ void __launch_bounds__(kThreads)
    k_dcgs_flags(const double *sn, int count, i64 Bp, const int *limit, int *lflag, double *flag) {
    const i64 j = static_cast<i64>(blockIdx.x) * kThreads + threadIdx.x;
    if (j >= Bp) return;
    const int lim = limit ? limit[j] : 1;
    bool f = false;
    if (lim > 0) {
        double ss = 0.0;
        for (int k = 0; k <= count; ++k) ss += sn[static_cast<i64>(k) * Bp + j] * sn[static_cast<i64>(k) * Bp + j];
        f = !(ss <= 0.25 * sn[static_cast<i64>(count + 1) * Bp + j]);
    }
    lflag[j] = f ? lim : 0;
    flag[j] = f ? 1.0 : 0.0;
}

// the next normalisation per lane from the first (sn1) or, where flagged (flag not null), the
// second (sn2) projections: beta = sqrt(max(|u|^2 - |s|^2, 0)); s is copied to scur for the next
// update; beta_out = beta, dbeta = beta (infinity when zero: the update then finalises a zero
// vector), scale = 1 / beta (0 when zero), beta2 = beta^2; flag_out (when given) records the
// test of k_dcgs_flags on sn1
__global__ void __launch_bounds__(kThreads)
    k_dcgs_norms(const double *sn1, const double *sn2, const double *flag, int count, i64 Bp,
                 const int *limit, double *scur, double *beta_out, double *dbeta, double *scale,
                 double *beta2, double *flag_out) {
    const i64 j = static_cast<i64>(blockIdx.x) * kThreads + threadIdx.x;
    if (j >= Bp) return;
    const bool on = !limit || limit[j] > 0;
    const double *sn = flag && flag[j] != 0.0 ? sn2 : sn1;
    double ss = 0.0;
    for (int k = 0; k <= count; ++k) {
        const double v = on ? sn[static_cast<i64>(k) * Bp + j] : 0.0;
        scur[static_cast<i64>(k) * Bp + j] = v;
        ss += v * v;
    }
    const double u2 = on ? sn[static_cast<i64>(count + 1) * Bp + j] : 0.0;
    if (flag_out) flag_out[j] = on && !(ss <= 0.25 * u2) ? 1.0 : 0.0;
    const double b = std::sqrt(std::max(u2 - ss, 0.0));
    beta_out[j] = b;
    dbeta[j] = b > 0.0 ? b : __builtin_inf();
    scale[j] = b > 0.0 ? 1.0 / b : 0.0;
    beta2[j] = b * b;
}

} // namespace

void dcgs_flags(const double *sn, int count, i64 Bp, const int *limit, int *lflag, double *flag,
                hipStream_t s) {
    hipLaunchKernelGGL(k_dcgs_flags, dim3(static_cast<u32>((Bp + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, s, sn, count, Bp, limit, lflag, flag);
}

void dcgs_norms(const double *sn1, const double *sn2, const double *flag, int count, i64 Bp,
                const int *limit, double *scur, double *beta_out, double *dbeta, double *scale,
                double *beta2, hipStream_t s, double *flag_out) {
    hipLaunchKernelGGL(k_dcgs_norms, dim3(static_cast<u32>((Bp + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, s, sn1, sn2, flag, count, Bp, limit, scur, beta_out, dbeta, scale, beta2,
                       flag_out);
}

void put_lane(const double *src, i64 rows, double *dst, i64 n, i64 Bp, i64 j, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_put_lane, dim3(grid_for(n)), dim3(kThreads), 0, s, src, rows, dst, n, Bp, j);
}

void scale_lanes(const double *x, const double *norm2, double *y, i64 n, i64 Bp, const int *limit,
                 hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_scale_lanes, dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes), dim3(kThreads), 0,
                       s, x, norm2, y, n, Bp, limit);
}

void get_lane(const double *src, i64 rows, i64 Bp, i64 j, double *dst, hipStream_t s) {
    if (rows == 0) return;
    hipLaunchKernelGGL(k_get_lane, dim3(grid_for(rows)), dim3(kThreads), 0, s, src, rows, Bp, j, dst);
}

template <typename T>
void clear_tails(T *X, int count, i64 ld, i64 n, i64 Bp, const int *rows, const int *sel,
                 hipStream_t s) {
    if (n == 0 || count <= 0) return;
    hipLaunchKernelGGL((k_clear_tails<T>), dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes),
                       dim3(kThreads), 0, s, X, count, ld, n, Bp, rows, sel);
}
template void clear_tails<float>(float *, int, i64, i64, i64, const int *, const int *, hipStream_t);
template void clear_tails<double>(double *, int, i64, i64, i64, const int *, const int *, hipStream_t);

void block_dot(const double *V, int count, const double *W, i64 n, i64 Bp, double *partial,
               int np, const int *limit, hipStream_t s) {
    if (count <= 0 || n == 0) return;
    hipLaunchKernelGGL(k_block_dot, dim3(np, (count + kDotChunk - 1) / kDotChunk, Bp / kLanes),
                       dim3(kThreads), 0, s, V, count, W, n, Bp, partial, np, limit);
}

template <typename IN>
void block_apply(double *out, const IN *V, int count, const double *C, i64 n, i64 Bp, double sign,
                 bool assign, const int *limit, hipStream_t s, const int *counts) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_block_apply<IN>), dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes),
                       dim3(kThreads), 0, s, out, V, count, C, n, Bp, sign, assign ? 1 : 0, limit,
                       counts);
}
template void block_apply<double>(double *, const double *, int, const double *, i64, i64, double,
                                  bool, const int *, hipStream_t, const int *);
template void block_apply<float>(double *, const float *, int, const double *, i64, i64, double,
                                 bool, const int *, hipStream_t, const int *);

int parts(i64 n) { return op_parts(n); }

int parts_wide(i64 n) {
    const i64 cap = std::max<i64>(1, (n + kWaves * 4 - 1) / (kWaves * 4)); // at least 4 rows per wave
    return static_cast<int>(std::max<i64>(op_parts(n), std::min<i64>(kWideParts, cap)));
}

i64 residual_parts(i64 n) { return row_blocks(n); }

template <typename XT>
void shifted_apply(const csr_view<double> &A, const double *shifts, const XT *x, const double *b,
                   double *y, double *partial, i64 Bp, const int *limit, hipStream_t s,
                   const double *scale) {
    bop_args<double, XT, double> a{A, shifts, 0.0, x, b, y, partial, Bp, limit};
    a.scale = b ? nullptr : scale;
    if (b) {
        launch<bmode::shifted_residual>(a, s);
    } else {
        launch<bmode::shifted>(a, s);
    }
}

template <typename T, typename IN>
void jacobi_first(const csr_view<T> &A, const double *shifts, double omega, const IN *in,
                  const double *scale, T *rhs_out, T *x, i64 Bp, const int *limit,
                  hipStream_t s) {
    if (A.n == 0) return;
    hipLaunchKernelGGL((k_jacobi_first<T, IN>), dim3(static_cast<u32>(row_blocks(A.n)), Bp / kLanes),
                       dim3(kThreads), 0, s, A,
                       shifts, static_cast<T>(omega), in, scale, rhs_out, x, Bp, limit);
}

template <typename T>
void jacobi_sweep(const csr_view<T> &A, const double *shifts, double omega, const T *rhs,
                  const T *x, T *x_out, i64 Bp, const int *limit, hipStream_t s) {
    launch<bmode::sweep>(bop_args<T, T, T>{A, shifts, static_cast<T>(omega), x, rhs, x_out, nullptr,
                                           Bp, limit}, s);
}

template <typename T>
void level_residual(const csr_view<T> &A, const double *shifts, const T *rhs, const T *x, T *r,
                    i64 Bp, const int *limit, hipStream_t s) {
    launch<bmode::residual>(bop_args<T, T, T>{A, shifts, T(0), x, rhs, r, nullptr, Bp, limit}, s);
}

template <typename T>
void gather_sum(const csr_view<T> &R, const T *x, T *y, i64 Bp, const int *limit, hipStream_t s) {
    launch<bmode::gather_sum>(bop_args<T, T, T>{R, nullptr, T(0), x, nullptr, y, nullptr, Bp, limit}, s);
}

template <typename T>
void prolong_add(i64 n, const u32 *coarse_of, const T *weight, const T *coarse, T *x, i64 Bp,
                 const int *limit, hipStream_t s, const u32 *rows) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_prolong_add<T>), dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes),
                       dim3(kThreads), 0, s, n,
                       coarse_of, weight, coarse, x, Bp, limit, rows);
}

template <typename T, typename IN>
void rb_first(const csr_view<T> &A, const u8 *colour, const double *shifts, const IN *in,
              const double *scale, T *rhs_out, T *x, const u32 *perm, i64 Bp, const int *limit,
              hipStream_t s) {
    if (A.n == 0) return;
    hipLaunchKernelGGL((k_rb_first<T, IN>), dim3(static_cast<u32>(row_blocks(A.n)), Bp / kLanes),
                       dim3(kThreads), 0, s, A, colour, shifts, in, scale, rhs_out, x, perm, Bp, limit);
}

template <typename T>
void gs_half(const csr_view<T> &Ac, const double *shifts, const T *rhs, T *x, T *res_zero, i64 Bp,
             const int *limit, hipStream_t s, T *nat, const u32 *perm) {
    bop_args<T, T, T> a{Ac, shifts, T(0), x, rhs, x, nullptr, Bp, limit};
    a.zero = res_zero;
    a.nat = nat;
    a.perm = perm;
    launch<bmode::gs>(a, s);
}

template <typename T>
void terminal_solve(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                    const T *inverse, u64 per_shift, const int *shift_index, i64 max_size,
                    const T *r, T *x, i64 Bp, const int *limit, hipStream_t s) {
    if (n_blocks == 0) return;
    hipLaunchKernelGGL((k_terminal<T>), dim3(grid_for(n_blocks * max_size * Bp)), dim3(kThreads), 0,
                       s, n_blocks, start, size, offset, inverse, per_shift, shift_index, max_size,
                       r, x, Bp, limit);
}

void multi_dot(const double *V, int count, const double *w, i64 n, i64 Bp, double *partial,
               int np, const int *limit, hipStream_t s) {
    if (count <= 0 || n == 0) return;
    hipLaunchKernelGGL(k_multi_dot, dim3(np, (count + kDotChunk - 1) / kDotChunk, Bp / kLanes),
                       dim3(kThreads), 0, s, V, count, w, n, Bp, partial, np, limit);
}

void reduce_partials(const double *partial, int np, i64 m, double *out, hipStream_t s,
                     double *tmp) {
    if (m <= 0) return;
    if (tmp && np > kGroup) {
        const int ng = (np + kGroup - 1) / kGroup;
        hipLaunchKernelGGL(k_reduce_groups, dim3(ng, static_cast<u32>((m + kLanes - 1) / kLanes)),
                           dim3(kThreads), 0, s, partial, np, m, tmp);
        hipLaunchKernelGGL(k_reduce, dim3(grid_for(m)), dim3(kThreads), 0, s, tmp, ng, m, out);
        return;
    }
    hipLaunchKernelGGL(k_reduce, dim3(grid_for(m)), dim3(kThreads), 0, s, partial, np, m, out);
}

i64 reduce_tmp_size(int np, i64 m) { return np > kGroup ? (np + kGroup - 1) / kGroup * m : 0; }

void dcgs_update(double *V, int count, const double *sc, const double *tg, const double *beta,
                 const double *w, i64 n, i64 Bp, double *partial, int np, double *eta,
                 const int *limit, hipStream_t s) {
    if (count < 0 || count >= kFusedBasis || n == 0) return;
    const dim3 grid(np, Bp / kLanes);
    switch ((count + kWaves - 1) / kWaves) {
#define MK_BDCGS(P)                                                                              \
    case P:                                                                                      \
        hipLaunchKernelGGL(k_dcgs_update<P>, grid, dim3(kThreads), 0, s, V, count, sc, tg, beta, \
                           w, n, Bp, partial, np, eta, limit);                                  \
        break;
        MK_BDCGS(0) MK_BDCGS(1) MK_BDCGS(2) MK_BDCGS(3) MK_BDCGS(4) MK_BDCGS(5) MK_BDCGS(6)
        MK_BDCGS(7) MK_BDCGS(8) MK_BDCGS(9) MK_BDCGS(10) MK_BDCGS(11) MK_BDCGS(12)
        MK_BDCGS(13) MK_BDCGS(14) MK_BDCGS(15) MK_BDCGS(16)
#undef MK_BDCGS
    }
}

void basis_update(double *w, const double *V, int count, const double *h, i64 n, i64 Bp,
                  const int *limit, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_basis_update, dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes),
                       dim3(kThreads), 0, s, w, V, count,
                       h, n, Bp, limit);
}

template <typename ZT>
void combine(double *x, const ZT *Z, const int *cnt, const double *y, i64 n, i64 Bp,
             hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_combine<ZT>), dim3(static_cast<u32>(row_blocks(n)), Bp / kLanes),
                       dim3(kThreads), 0, s, x, Z, cnt, y,
                       n, Bp);
}

void pack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_pack, dim3((n + kT - 1) / kT, Bp / kT), dim3(kThreads), 0, s, src, n,
                       nsys, dst, Bp);
}

void unpack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_unpack, dim3((n + kT - 1) / kT, Bp / kT), dim3(kThreads), 0, s, src, n,
                       nsys, dst, Bp);
}

template void shifted_apply<double>(const csr_view<double> &, const double *, const double *,
                                    const double *, double *, double *, i64, const int *,
                                    hipStream_t, const double *);
template void shifted_apply<float>(const csr_view<double> &, const double *, const float *,
                                   const double *, double *, double *, i64, const int *,
                                   hipStream_t, const double *);
#define MK_BLEVEL(T)                                                                              \
    template void jacobi_first<T, double>(const csr_view<T> &, const double *, double,            \
                                          const double *, const double *, T *, T *, i64,         \
                                          const int *, hipStream_t);                              \
    template void jacobi_sweep<T>(const csr_view<T> &, const double *, double, const T *,         \
                                  const T *, T *, i64, const int *, hipStream_t);                 \
    template void level_residual<T>(const csr_view<T> &, const double *, const T *, const T *,    \
                                    T *, i64, const int *, hipStream_t);                          \
    template void gather_sum<T>(const csr_view<T> &, const T *, T *, i64, const int *,            \
                                hipStream_t);                                                     \
    template void prolong_add<T>(i64, const u32 *, const T *, const T *, T *, i64, const int *,   \
                                 hipStream_t, const u32 *);                                       \
    template void rb_first<T, double>(const csr_view<T> &, const u8 *, const double *,            \
                                      const double *, const double *, T *, T *, const u32 *, i64, \
                                      const int *, hipStream_t);                                  \
    template void gs_half<T>(const csr_view<T> &, const double *, const T *, T *, T *, i64,       \
                             const int *, hipStream_t, T *, const u32 *);                         \
    template void terminal_solve<T>(i64, const u32 *, const u32 *, const u64 *, const T *, u64,   \
                                    const int *, i64, const T *, T *, i64, const int *,          \
                                    hipStream_t);                                                 \
    template void combine<T>(double *, const T *, const int *, const double *, i64, i64,         \
                             hipStream_t);
MK_BLEVEL(float)
MK_BLEVEL(double)
#undef MK_BLEVEL
template void rb_first<float, float>(const csr_view<float> &, const u8 *, const double *,
                                     const float *, const double *, float *, float *, const u32 *,
                                     i64, const int *, hipStream_t);
template void jacobi_first<float, float>(const csr_view<float> &, const double *, double,
                                         const float *, const double *, float *, float *, i64,
                                         const int *, hipStream_t);

} // namespace mk::gpu::batch
