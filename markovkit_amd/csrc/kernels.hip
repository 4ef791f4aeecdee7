// Device kernels of the resolvent solve for gfx950 (CDNA4, wave64). See kernels.hpp for the
// operator layouts. All reductions run in a fixed order (per-wave shuffles, then waves in
// index order, then workgroup partials in index order), so results do not depend on timing.
#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>
#include <type_traits>

namespace mk::gpu {

namespace {

constexpr int kWaves = kThreads / 64;
constexpr int kUnroll = kTile / kThreads; // entries per thread per row block

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum over the workgroup; the result is valid in thread 0
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) r += lds[w];
    }
    __syncthreads();
    return r;
}

int g_xcd_chunk = 16;

// logical row block of workgroup bid: chunk > 0 interleaves chunks of `chunk` blocks over the
// XCDs (the launch grid is padded to a multiple of 8 * chunk; blocks past nb exit)
__device__ __forceinline__ u32 xcd_block(u32 bid, u32 chunk) {
    if (chunk == 0) return bid;
    const u32 xcd = bid & 7u, slot = bid >> 3;
    return ((slot / chunk) * 8u + xcd) * chunk + slot % chunk;
}

u32 padded_grid(i64 nb, u32 chunk) {
    if (chunk == 0) return static_cast<u32>(nb);
    const i64 unit = 8 * static_cast<i64>(chunk);
    return static_cast<u32>((nb + unit - 1) / unit * unit);
}

enum class csr_mode { shifted, shifted_residual, sweep, residual, gather_sum, gs };

template <typename VT, typename XT, typename T>
struct csr_args {
    csr_view<VT> A;
    double shift;
    T omega;
    const XT *x;
    const T *rhs;    // b (shifted_residual) or rhs (sweep, residual)
    T *out;
    double *partial; // shifted_residual: sum of out^2 per row block
    T *aux;          // gs: residual entries to zero (optional)
    T *nat;          // gs: also store the new value at nat[perm[row]] (optional)
    const u32 *perm;
    u32 chunk;       // XCD interleave chunk (set by the launcher)
    // residual and gather_sum: the next level's first red-black half-sweep fused in (first_x
    // not null): where first_colour[r] == 0, first_x[r] = out[r] / (shift + exit of first_A's
    // row r), the formula of k_rb_first
    const std::uint8_t *first_colour = nullptr;
    csr_view<T> first_A{};
    T *first_x = nullptr;
};

// vector entry of the operator's local row r
template <typename VT>
__device__ __forceinline__ u32 vec_row(const csr_view<VT> &A, u32 r) {
    return A.rowmap ? A.rowmap[r] : r;
}

// the first red-black half-sweep from x = 0 of the level a.first_A, on row r with right-hand
// side b
template <typename VT, typename XT, typename T>
__device__ __forceinline__ void first_half(const csr_args<VT, XT, T> &a, u32 r, T b) {
    if (a.first_x && a.first_colour[r] == 0) {
        const T ex = a.first_A.exit ? a.first_A.exit[r] : a.first_A.edict[a.first_A.rcode[r]];
        a.first_x[r] = b / (static_cast<T>(a.shift) + ex);
    }
}

// the operands of vector entry r that the row's result needs besides its entries: the own x
// value (every mode but gs and gather_sum), the right-hand side (the modes that subtract from
// one) and the natural-order index of a gs row that is also stored there. The sliced kernel
// loads them with the row lengths, so their latency overlaps the entry loads and the gathers.
template <typename T>
struct row_operands {
    T x = 0;
    T rhs = 0;
    u32 perm = 0;
};

template <csr_mode M, typename VT, typename XT, typename T>
__device__ __forceinline__ row_operands<T> load_row_operands(const csr_args<VT, XT, T> &a, u32 r) {
    row_operands<T> o;
    if constexpr (M != csr_mode::gs && M != csr_mode::gather_sum) o.x = static_cast<T>(a.x[r]);
    if constexpr (M == csr_mode::gs || M == csr_mode::shifted_residual || M == csr_mode::sweep ||
                  M == csr_mode::residual) {
        o.rhs = a.rhs[r];
    }
    if constexpr (M == csr_mode::gs) {
        if (a.nat) o.perm = a.perm[r];
    }
    return o;
}

template <csr_mode M, typename VT, typename XT, typename T>
__device__ __forceinline__ double finish(const csr_args<VT, XT, T> &a, u32 r, T sum, T exit_r,
                                         const row_operands<T> &o) {
    if constexpr (M == csr_mode::gather_sum) {
        a.out[r] = sum;
        first_half(a, r, sum);
        return 0.0;
    }
    if constexpr (M == csr_mode::gs) {
        // in place: the row's entries all have the other colour
        const T v = (o.rhs + sum) / (static_cast<T>(a.shift) + exit_r);
        a.out[r] = v;
        if (a.aux) a.aux[r] = T(0);
        if (a.nat) a.nat[o.perm] = v;
        return 0.0;
    }
    const T xr = o.x;
    const T d = static_cast<T>(a.shift) + exit_r;
    const T ax = d * xr - sum;
    if constexpr (M == csr_mode::shifted) {
        a.out[r] = ax;
        return 0.0;
    } else if constexpr (M == csr_mode::shifted_residual) {
        const T res = o.rhs - ax;
        a.out[r] = res;
        return static_cast<double>(res) * static_cast<double>(res);
    } else if constexpr (M == csr_mode::sweep) {
        const T res = o.rhs - ax;
        a.out[r] = xr + a.omega * res / d;
        return 0.0;
    } else {
        const T res = o.rhs - ax;
        a.out[r] = res;
        first_half(a, r, res);
        return 0.0;
    }
}

// coded sliced operator: one wavefront per 64-row slice, four slices per workgroup. A chunk is U
// entries per lane: 16 in general, 12 or 8 when every row of the operator fits one such chunk
// (a chunk loads all its slots, so a narrower one issues fewer redundant loads).

// EX: exit rates per row (a.A.exit) instead of the exit dictionary.
// A slice's bounds are wave-uniform (scalar loads), and its entries are stored for all 64 lanes
// up to the slice's width (its longest row), so the entry loads depend on the slice bounds only:
// the first chunk is in flight together with the rows' lengths and codes and the dictionary
// staging, and only the x gathers wait for the row lengths. Each row still sums its entries in
// storage order.
template <csr_mode M, typename VT, typename XT, typename T, bool EX, int U, bool MULTI>
__global__ void __launch_bounds__(kThreads) k_csr_sliced(csr_args<VT, XT, T> a) {
    __shared__ T vdict[kDictMax];
    __shared__ T edict[kDictMax];
    __shared__ double red64[kWaves];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const u32 b = xcd_block(blockIdx.x, a.chunk);
    if (b >= a.A.n_blocks) return;
    static_assert(kDictMax <= kThreads, "one dictionary entry per thread");
    // dictionary values first: the stores into LDS then wait for these loads only
    const T vd = tid < a.A.n_vdict ? static_cast<T>(a.A.vdict[tid]) : T(0);
    const T ed = tid < a.A.n_edict ? static_cast<T>(a.A.edict[tid]) : T(0);
    const i64 slice = static_cast<i64>(b) * kWaves + wave;
    u32 base = 0, width = 0;
    if (slice < a.A.n_slices) {
        base = a.A.soff[slice];
        width = (a.A.soff[slice + 1] - base) / kSlice;
    }
    const u32 *ent = a.A.ent + base + lane;
    // a chunk loads all its slots, clamped to the slice's width: branch-free, so the loads of a
    // chunk are in flight together (the compiler waits at every join of per-slot branches)
    u32 e[U];
    auto load_entries = [&](u32 k0) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            e[u] = __builtin_nontemporal_load(ent + min(k0 + u, width - 1) * kSlice);
    };
    if (width > 0) load_entries(0);
    const i64 r = slice * kSlice + lane;
    const bool has_row = slice < a.A.n_slices && r < a.A.n;
    u32 len = 0, ec = 0, vr = 0;
    T ex = 0;
    row_operands<T> own;
    if (has_row) {
        len = a.A.rlen[r];
        if constexpr (EX) {
            ex = static_cast<T>(a.A.exit[r]);
        } else {
            ec = a.A.rcode[r];
        }
        vr = vec_row(a.A, static_cast<u32>(r));
        own = load_row_operands<M>(a, vr);
    }
    // gathers: guarded per lane in the working precision; for x in another precision the
    // compiler would convert each gathered value inside its guard and wait for it there, so
    // there every slot of the chunk loads (a slot past the row's end reads x[0], in bounds
    // whenever the slice has an entry) and only the sum leaves it out
    XT xv[U];
    auto gather = [&](u32 k0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (std::is_same_v<XT, T>) {
                xv[u] = k0 + u < len ? a.x[e[u] & kColMask] : XT(0);
            } else {
                xv[u] = a.x[k0 + u < len ? e[u] & kColMask : 0u];
            }
        }
    };
    // MULTI (for rows longer than one chunk): a chunk's rate codes are kept four to a register,
    // so once its gathers are issued the entry registers take the next chunk's loads, which are
    // then in flight while this chunk's gathers return. Without MULTI the chunks run one after
    // another (fewer registers).
    constexpr int kCodeWords = MULTI ? (U + 3) / 4 : 1;
    u32 codes[kCodeWords];
    auto keep_codes = [&]() {
#pragma unroll
        for (int q = 0; q < kCodeWords; ++q) codes[q] = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) codes[u / 4] |= (e[u] >> kColBits) << (8 * (u % 4));
    };
    auto code = [&](int u) -> u32 {
        if constexpr (MULTI) return (codes[u / 4] >> (8 * (u % 4))) & 0xFFu;
        return e[u] >> kColBits;
    };
    T sum = 0;
    auto accumulate = [&](u32 k0) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < len) sum += vdict[code(u)] * static_cast<T>(xv[u]);
    };
    if (width > 0) {
        gather(0);
        if constexpr (MULTI) {
            keep_codes();
            if (static_cast<u32>(U) < width) load_entries(U);
        }
    }
    // the dictionaries reach LDS after the first gathers are issued (their stores wait for every
    // load before them)
    if (tid < a.A.n_vdict) vdict[tid] = vd;
    if (tid < a.A.n_edict) edict[tid] = ed;
    __syncthreads();
    if (width > 0) accumulate(0);
    for (u32 k0 = U; k0 < width; k0 += U) {
        if constexpr (MULTI) {
            gather(k0);
            keep_codes();
            if (k0 + U < width) load_entries(k0 + U);
        } else {
            load_entries(k0);
            gather(k0);
        }
        accumulate(k0);
    }
    double sq = 0.0;
    if (has_row) {
        if constexpr (EX) {
            sq = finish<M>(a, vr, sum, ex, own);
        } else {
            sq = finish<M>(a, vr, sum, edict[ec], own);
        }
    }
    if constexpr (M == csr_mode::shifted_residual) {
        const double t = block_sum(sq, red64);
        if (tid == 0 && a.partial) a.partial[b] = t;
    }
}

// plain operator: one workgroup per row block; a block holding one row longer than the tile
// reduces that row across the workgroup
template <csr_mode M, typename VT, typename XT, typename T>
__global__ void __launch_bounds__(kThreads) k_csr_plain(csr_args<VT, XT, T> a) {
    __shared__ T prod[kTile];
    __shared__ T red[kWaves];
    __shared__ double red64[kWaves];
    const u32 b = xcd_block(blockIdx.x, a.chunk);
    if (b >= a.A.n_blocks) return;
    const u32 r0 = a.A.blocks[b], r1 = a.A.blocks[b + 1];
    const u32 p0 = a.A.ptr[r0], p1 = a.A.ptr[r1];
    const int tid = threadIdx.x;
    double sq = 0.0;
    if (p1 - p0 <= static_cast<u32>(kTile)) {
        u32 c[kUnroll];
        VT v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u32 p = p0 + tid + u * kThreads;
            c[u] = p < p1 ? __builtin_nontemporal_load(a.A.col + p) : 0u;
            v[u] = p < p1 ? __builtin_nontemporal_load(a.A.val + p) : VT(0);
        }
        XT xv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u32 p = p0 + tid + u * kThreads;
            xv[u] = p < p1 ? a.x[c[u]] : XT(0);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const u32 q = tid + u * kThreads;
            if (p0 + q < p1) prod[q] = static_cast<T>(v[u]) * static_cast<T>(xv[u]);
        }
        __syncthreads();
        const u32 r = r0 + tid;
        if (r < r1) {
            T sum = 0;
            const u32 q1 = a.A.ptr[r + 1];
            for (u32 q = a.A.ptr[r]; q < q1; ++q) sum += prod[q - p0];
            const u32 vr = vec_row(a.A, r);
            sq = finish<M>(a, vr, sum, static_cast<T>(a.A.exit[r]), load_row_operands<M>(a, vr));
        }
    } else {
        T sum = 0;
        for (u32 p = p0 + tid; p < p1; p += kThreads)
            sum += static_cast<T>(a.A.val[p]) * static_cast<T>(a.x[a.A.col[p]]);
        sum = block_sum(sum, red);
        if (tid == 0) {
            const u32 vr = vec_row(a.A, r0);
            sq = finish<M>(a, vr, sum, static_cast<T>(a.A.exit[r0]), load_row_operands<M>(a, vr));
        }
    }
    if constexpr (M == csr_mode::shifted_residual) {
        const double t = block_sum(sq, red64);
        if (tid == 0 && a.partial) a.partial[b] = t;
    }
}

template <csr_mode M, typename VT, typename XT, typename T>
void launch_csr(csr_args<VT, XT, T> a, hipStream_t s) {
    if (a.A.n_blocks == 0) return;
    a.chunk = static_cast<u32>(g_xcd_chunk);
    const dim3 grid(padded_grid(a.A.n_blocks, a.chunk));
    if (a.A.coded) {
        const int w = a.A.max_width;
        const bool ex = a.A.exit != nullptr;
        const dim3 block(kThreads);
        if (w > 0 && w <= 8) {
            if (ex) hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, true, 8, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, false, 8, false>), grid, block, 0, s, a);
        } else if (w > 0 && w <= 12) {
            if (ex) hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, true, 12, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, false, 12, false>), grid, block, 0, s, a);
        } else if (w > 0 && w <= 16) {
            if (ex) hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, true, 16, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, false, 16, false>), grid, block, 0, s, a);
        } else if constexpr (M == csr_mode::gather_sum) {
            // (the restriction's wide rows gain nothing from the pipelined chunks)
            if (ex) hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, true, 16, false>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, false, 16, false>), grid, block, 0, s, a);
        } else {
            if (ex) hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, true, 16, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_csr_sliced<M, VT, XT, T, false, 16, true>), grid, block, 0, s, a);
        }
    } else {
        hipLaunchKernelGGL((k_csr_plain<M, VT, XT, T>), grid, dim3(kThreads), 0, s, a);
    }
}

template <typename T>
__device__ __forceinline__ T row_exit(const csr_view<T> &A, i64 r) {
    return A.exit ? A.exit[r] : A.edict[A.rcode[r]];
}

template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_jacobi_first(csr_view<T> A, T shift, T omega, const IN *rhs_in, const double *scale_ptr,
                   T *rhs_out, T *x) {
    const IN scale = scale_ptr ? static_cast<IN>(*scale_ptr) : IN(1);
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < A.n;
         i += static_cast<i64>(gridDim.x) * kThreads) {
        const T b = static_cast<T>(rhs_in[i] * scale);
        if (rhs_out) rhs_out[i] = b;
        x[i] = omega * b / (shift + row_exit(A, i));
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_restrict(i64 nc, const u32 *cptr, const u32 *cidx, const T *fine, T *coarse) {
    for (i64 c = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; c < nc;
         c += static_cast<i64>(gridDim.x) * kThreads) {
        T s = 0;
        const u32 q1 = cptr[c + 1];
        for (u32 q = cptr[c]; q < q1; ++q) s += fine[cidx[q]];
        coarse[c] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_prolong_add(i64 n, const u32 *rows, const u32 *coarse_of, const T *weight, const T *coarse,
                  const T *coarse2, T *x) {
    for (i64 q = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; q < n;
         q += static_cast<i64>(gridDim.x) * kThreads) {
        const i64 i = rows ? rows[q] : q;
        const u32 c = coarse_of[i];
        if (c == 0xFFFFFFFFu) continue;
        T cv = coarse[c];
        if (coarse2) cv += coarse2[c]; // the coarse iterate's last correction, added first
        x[i] += weight[i] * cv;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_terminal(i64 nb, const u32 *start, const u32 *size, const u64 *offset, const T *inv,
               i64 max_size, const T *r, T *x, int assign) {
    const i64 g = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x;
    if (g >= nb * max_size) return;
    const i64 b = g / max_size, i = g % max_size;
    const u32 m = size[b];
    if (i >= m) return;
    const T *row = inv + offset[b] + static_cast<u64>(i) * m;
    const T *rr = r + start[b];
    T s = 0;
#pragma unroll 8
    for (u32 j = 0; j < m; ++j) s += row[j] * rr[j];
    if (assign) {
        x[start[b] + i] = s;
    } else {
        x[start[b] + i] += s;
    }
}

// the same product for blocks of more than kWideTerminal rows: one wave per row, lanes over
// columns (coalesced reads of the row), lane partial sums over j = lane, lane + 64, ... and a
// fixed shuffle tree; grid nb * max_size waves
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_terminal_wide(i64 nb, const u32 *start, const u32 *size, const u64 *offset, const T *inv,
                    i64 max_size, const T *r, T *x, int assign) {
    const i64 g = (blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (g >= nb * max_size) return;
    const i64 b = g / max_size, i = g % max_size;
    const u32 m = size[b];
    if (i >= m) return;
    const T *row = inv + offset[b] + static_cast<u64>(i) * m;
    const T *rr = r + start[b];
    T s = 0;
#pragma unroll 4
    for (u32 j = lane; j < m; j += 64) s += row[j] * rr[j];
    s = wave_sum(s);
    if (lane == 0) {
        if (assign) {
            x[start[b] + i] = s;
        } else {
            x[start[b] + i] += s;
        }
    }
}

// the dense solve of one coarse row (a wave, lanes over the row of the inverse, as
// k_terminal_wide) followed by the prolongation of that value to its colour-0 children
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_terminal_prolong(i64 nb, const u32 *start, const u32 *size, const u64 *offset, const T *inv,
                       i64 max_size, const T *r, T *xc, csr_view<T> R, const u8 *fine_colour,
                       const T *weight, T *x_fine) {
    const i64 g = (blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (g >= nb * max_size) return;
    const i64 b = g / max_size, i = g % max_size;
    const u32 m = size[b];
    if (i >= m) return;
    const T *row = inv + offset[b] + static_cast<u64>(i) * m;
    const T *rr = r + start[b];
    T s = 0;
#pragma unroll 4
    for (u32 j = lane; j < m; j += 64) s += row[j] * rr[j];
    s = wave_sum(s);
    const u32 c = start[b] + static_cast<u32>(i);
    if (lane == 0) xc[c] = s;
    // the children of c, one per lane
    if (R.coded) {
        const u32 len = R.rlen[c];
        const u32 base = R.soff[c >> 6] + (c & 63u);
        for (u32 k = lane; k < len; k += 64) {
            const u32 f = R.ent[base + k * kSlice] & kColMask;
            if (fine_colour[f] == 0) x_fine[f] += weight[f] * s;
        }
    } else {
        const u32 p0 = R.ptr[c], p1 = R.ptr[c + 1];
        for (u32 p = p0 + lane; p < p1; p += 64) {
            const u32 f = R.col[p];
            if (fine_colour[f] == 0) x_fine[f] += weight[f] * s;
        }
    }
}

// ---- aggregate-block smoother (thread per listed row)
// sum over row i's entries of rate * x[column], ascending entry order, and row i's exit rate
template <typename T>
__device__ __forceinline__ T gather_row(const csr_view<T> &A, u32 i, const T *x, T &exit_i) {
    T sum = 0;
    if (A.coded) {
        const u32 len = A.rlen[i];
        const u32 base = A.soff[i >> 6] + (i & 63u);
        exit_i = A.exit ? A.exit[i] : A.edict[A.rcode[i]];
        // unrolled so that several entries' loads are in flight; the sum keeps entry order
#pragma unroll 4
        for (u32 k = 0; k < len; ++k) {
            const u32 e = A.ent[base + k * kSlice];
            sum += A.vdict[e >> kColBits] * x[e & kColMask];
        }
    } else {
        exit_i = A.exit[i];
        for (u32 p = A.ptr[i]; p < A.ptr[i + 1]; ++p) sum += A.val[p] * x[A.col[p]];
    }
    return sum;
}

// e = inverse_b r[members of b] for every member of the aggregate's blocks; x += e, escr = e
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_block_solve(const u32 *members, const u32 *member_block, const u32 *block_first,
                  const u64 *inv_off, const T *inv, u64 m0, u64 m1, const T *r, T *x, T *escr) {
    const u64 g = m0 + blockIdx.x * static_cast<u64>(kThreads) + threadIdx.x;
    if (g >= m1) return;
    const u32 b = member_block[g];
    const u32 f = block_first[b], m = block_first[b + 1] - f;
    const T *row = inv + inv_off[b] + static_cast<u64>(g - f) * m;
    T e = 0;
    for (u32 c = 0; c < m; ++c) e += row[c] * r[members[f + c]];
    const u32 i = members[g];
    x[i] += e;
    escr[i] = e;
}

// r -= A e on the rows the aggregate's correction reaches (e = escr, zero off the members)
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_block_residual(csr_view<T> A, T shift, const u32 *rows, u64 q0, u64 q1, const T *escr, T *r) {
    const u64 q = q0 + blockIdx.x * static_cast<u64>(kThreads) + threadIdx.x;
    if (q >= q1) return;
    const u32 i = rows[q];
    T ex = 0;
    const T sum = gather_row(A, i, escr, ex);
    r[i] -= (shift + ex) * escr[i] - sum;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_block_zero(const u32 *members, u64 m0, u64 m1, T *escr) {
    const u64 g = m0 + blockIdx.x * static_cast<u64>(kThreads) + threadIdx.x;
    if (g < m1) escr[members[g]] = T(0);
}

// Jacobi sweep on listed rows: out[i] = x[i] + w (rhs[i] - (A x)[i]) / (shift + exit_i)
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_sweep_rows(csr_view<T> A, T shift, T omega, const T *rhs, const T *x, T *out,
                 const u32 *rows, u64 nrows) {
    const u64 q = blockIdx.x * static_cast<u64>(kThreads) + threadIdx.x;
    if (q >= nrows) return;
    const u32 i = rows[q];
    T ex = 0;
    const T sum = gather_row(A, i, x, ex);
    const T d = shift + ex;
    out[i] = x[i] + omega * (rhs[i] - (d * x[i] - sum)) / d;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_copy_rows(const u32 *rows, u64 nrows, const T *src, T *dst) {
    const u64 q = blockIdx.x * static_cast<u64>(kThreads) + threadIdx.x;
    if (q < nrows) dst[rows[q]] = src[rows[q]];
}

// ---- red-black Gauss-Seidel: the first half-sweep from x = 0
template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_rb_first(csr_view<T> A, const u8 *colour, T shift, const IN *rhs_in, const double *scale_ptr,
               const u32 *perm, T *rhs_out, T *x) {
    const IN scale = scale_ptr ? static_cast<IN>(*scale_ptr) : IN(1);
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < A.n;
         i += static_cast<i64>(gridDim.x) * kThreads) {
        const T b = static_cast<T>(rhs_in[perm ? perm[i] : i] * scale);
        if (rhs_out) rhs_out[i] = b;
        if (colour[i] == 0) x[i] = b / (shift + row_exit(A, i));
    }
}

// the same from the caller's order: box index j goes to row pos[j], whose exit rate (colour 0)
// or -1 (colour 1) is exit_box[j]
template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_rb_first_scatter(i64 n, const T *exit_box, T shift, const IN *rhs_in, const double *scale_ptr,
                       const u32 *pos, T *rhs_out, T *x) {
    const IN scale = scale_ptr ? static_cast<IN>(*scale_ptr) : IN(1);
    for (i64 j = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; j < n;
         j += static_cast<i64>(gridDim.x) * kThreads) {
        const u32 i = pos[j];
        const T ex = exit_box[j];
        const T b = static_cast<T>(rhs_in[j] * scale);
        rhs_out[i] = b;
        if (ex >= T(0)) x[i] = b / (shift + ex);
    }
}

// out[i] = (T)(in[perm[i]] * scale) and out[perm[i]] = in[i]: a level's row order to and from
// the caller's
template <typename T, typename IN>
__global__ void __launch_bounds__(kThreads)
    k_permute_in(i64 n, const u32 *perm, const IN *in, const double *scale_ptr, T *out) {
    const IN scale = scale_ptr ? static_cast<IN>(*scale_ptr) : IN(1);
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        out[i] = static_cast<T>(in[perm[i]] * scale);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_permute_out(i64 n, const u32 *perm, const T *in, T *out) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        out[perm[i]] = in[i];
}

// ---- ILU(0) triangular sweeps over level-scheduled row groups (rows of one group are not
// neighbours, so they update independently). ONE_WG: one workgroup runs groups [g0, g1) with a
// barrier after each; else the grid runs the single group g0.
template <typename T, bool BACK, bool ONE_WG>
__global__ void __launch_bounds__(ONE_WG ? kTailThreads : kThreads)
    k_ilu_group(csr_view<T> A, const u32 *rows, const u32 *gptr, int g0, int g1, const T *pinv,
                const T *r, T *u, T *x) {
    const u32 first = ONE_WG ? threadIdx.x : blockIdx.x * static_cast<u32>(kThreads) + threadIdx.x;
    const u32 step = ONE_WG ? blockDim.x : gridDim.x * static_cast<u32>(kThreads);
    for (int g = g0; g < g1; ++g) {
        const u32 q1 = gptr[g + 1];
        for (u32 q = gptr[g] + first; q < q1; q += step) {
            const u32 i = rows[q];
            // forward: r_i + sum_{j < i} rate u_j; backward: sum_{j > i} rate e_j (columns ascending)
            T sum = BACK ? T(0) : r[i];
            if (A.coded) {
                const u32 len = A.rlen[i];
                const u32 base = A.soff[i >> 6] + (i & 63u);
                for (u32 k = 0; k < len; ++k) {
                    const u32 e = A.ent[base + k * kSlice];
                    const u32 c = e & kColMask;
                    if (BACK ? c > i : c < i) sum += A.vdict[e >> kColBits] * u[c];
                }
            } else {
                for (u32 p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
                    const u32 c = A.col[p];
                    if (BACK ? c > i : c < i) sum += A.val[p] * u[c];
                }
            }
            if constexpr (BACK) {
                const T e = u[i] + sum * pinv[i];
                u[i] = e;
                x[i] += e;
            } else {
                u[i] = sum * pinv[i];
            }
        }
        if constexpr (ONE_WG) __syncthreads();
    }
}

// one Gauss-Seidel half-sweep of the rows of one colour (listed in rows), in place, by one
// workgroup
template <typename T>
__device__ __forceinline__ void tail_gs_half(const csr_view<T> &A, const u32 *rows, i64 nrows, T shift,
                                             const T *rhs, T *x) {
    for (i64 q = threadIdx.x; q < nrows; q += blockDim.x) {
        const u32 i = rows[q];
        T ex = 0;
        const T sum = gather_row(A, i, x, ex);
        x[i] = (rhs[i] + sum) / (shift + ex);
    }
}

// dense terminal blocks: x[block rows] (+)= inv_b r[block rows]
template <bool ASSIGN, typename T>
__device__ __forceinline__ void tail_terminal(const level_desc<T> &lv, const T *r, T *x) {
    for (i64 g = threadIdx.x; g < lv.n_term * lv.term_max; g += blockDim.x) {
        const i64 b = g / lv.term_max, i = g % lv.term_max;
        const u32 m = lv.tsize[b];
        if (i >= m) continue;
        const u32 st = lv.tstart[b];
        const T *row = lv.tinv + lv.toff[b] + static_cast<u64>(i) * m;
        const T *rr = r + st;
        T acc = 0;
#pragma unroll 8
        for (u32 j = 0; j < m; ++j) acc += row[j] * rr[j];
        if constexpr (ASSIGN) {
            x[st + i] = acc;
        } else {
            x[st + i] += acc;
        }
    }
}

// The V-cycle of levels L[0 .. levels) by one workgroup; the top level's result goes to out (in
// the caller's row order) or, without out, to top_x. rhs_in, rhs64, out and top_x are in HBM;
// the top level's right-hand side is first copied to L[0].rhs.
template <typename T>
__device__ __forceinline__ void tail_body(const level_desc<T> *L, int levels, T shift, T omega,
                                          const T *rhs_in, const double *rhs64, const double *scale_ptr,
                                          T *out, T *top_x, int cycles) {
    T *top = nullptr; // the top level's result
    const int tid = threadIdx.x, nt = blockDim.x;
    const double scale = scale_ptr ? *scale_ptr : 1.0;
    const u32 *perm0 = L[0].perm; // the top level's row order differs from the caller's
    T *acc = L[0].acc;            // cycles > 1: the top level's accumulated iterate
    for (int cyc = 0; cyc < cycles; ++cyc) {
        // the top level's last Jacobi sweep writes out directly in a single cycle
        T *vout = cycles == 1 ? out : nullptr;
        // bit k: level k's current iterate is in xt (the Jacobi sweeps alternate x and xt); levels
        // are at most 32
        u32 flip = 0;
        for (int k = 0; k < levels; ++k) {
            const level_desc<T> &lv = L[k];
            const T *rhs = lv.rhs;
            T *x = lv.x, *xt = lv.xt;
            const u8 *colour = lv.colour;
            if (lv.direct && k > 0) {
                // the cycle from zero on a level of terminal blocks only is their dense solve
                tail_terminal<true>(lv, rhs, x);
                __syncthreads();
                continue;
            }
            // first sweep from x = 0 (the FP64 input converted on the top level); red-black: colour 0
            // only, colour 1 follows
            for (i64 i = tid; i < lv.n; i += nt) {
                T b;
                if (k == 0 && cyc == 0) {
                    const i64 src = perm0 ? perm0[i] : i;
                    b = rhs64 ? static_cast<T>(rhs64[src] * scale) : rhs_in[src];
                    lv.rhs[i] = b;
                } else {
                    b = rhs[i];
                }
                if (!colour) {
                    x[i] = lv.pre > 0 ? omega * b / (shift + row_exit(lv.A, i)) : T(0);
                } else if (lv.pre == 0) {
                    x[i] = T(0);
                } else if (colour[i] == 0) {
                    x[i] = b / (shift + row_exit(lv.A, i));
                }
            }
            __syncthreads();
            if (colour) {
                for (int half = 1; half < 2 * lv.pre; ++half) {
                    const int c = half & 1;
                    tail_gs_half(lv.A, lv.crows[c], lv.ncrows[c], shift, rhs, x);
                    __syncthreads();
                }
                if (lv.pre > 0) {
                    // after the last colour-1 half-sweep only colour-0 rows have a residual
                    for (i64 q = tid; q < lv.ncrows[0]; q += nt) {
                        const u32 i = lv.crows[0][q];
                        T ex = 0;
                        const T sum = gather_row(lv.A, i, x, ex);
                        const T d = shift + ex;
                        lv.res[i] = rhs[i] - (d * x[i] - sum);
                    }
                    for (i64 q = tid; q < lv.ncrows[1]; q += nt) lv.res[lv.crows[1][q]] = T(0);
                } else {
                    for (i64 i = tid; i < lv.n; i += nt) {
                        T ex = 0;
                        const T sum = gather_row(lv.A, static_cast<u32>(i), x, ex);
                        const T d = shift + ex;
                        lv.res[i] = rhs[i] - (d * x[i] - sum);
                    }
                }
                __syncthreads();
            }
            for (int p = 1; !colour && p < lv.pre; ++p) {
                for (i64 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = gather_row(lv.A, static_cast<u32>(i), x, ex);
                    const T xr = x[i];
                    const T d = shift + ex;
                    const T ax = d * xr - sum;
                    const T res = rhs[i] - ax;
                    xt[i] = xr + omega * res / d;
                }
                __syncthreads();
                T *t = x;
                x = xt;
                xt = t;
            }
            for (i64 i = tid; !colour && i < lv.n; i += nt) {
                T ex = 0;
                const T sum = gather_row(lv.A, static_cast<u32>(i), x, ex);
                const T d = shift + ex;
                lv.res[i] = rhs[i] - (d * x[i] - sum);
            }
            __syncthreads();
            if (lv.n_term > 0) {
                tail_terminal<false>(lv, lv.res, x);
                __syncthreads();
            }
            if (lv.n_next > 0) {
                T *crhs = L[k + 1].rhs;
                for (i64 c = tid; c < lv.n_next; c += nt) {
                    T ex = 0;
                    crhs[c] = gather_row(lv.R, static_cast<u32>(c), lv.res, ex);
                }
                __syncthreads();
            }
            if (x == lv.xt) flip |= 1u << k;
        }
        for (int k = levels - 1; k >= 0; --k) {
            const level_desc<T> &lv = L[k];
            const T *rhs = lv.rhs;
            const bool f = (flip >> k) & 1u;
            T *x = f ? lv.xt : lv.x, *xt = f ? lv.x : lv.xt;
            if (lv.direct && k > 0) continue;
            if (lv.n_next > 0) {
                const T *cx = ((flip >> (k + 1)) & 1u) ? L[k + 1].xt : L[k + 1].x;
                // red-black with post-smoothing: colour-1 rows are rewritten without being read
                const bool red_only = lv.colour && lv.post > 0;
                const i64 np = red_only ? lv.ncrows[0] : lv.n;
                for (i64 q = tid; q < np; q += nt) {
                    const i64 i = red_only ? lv.crows[0][q] : q;
                    const u32 c = lv.coarse_of[i];
                    if (c != 0xFFFFFFFFu) x[i] += lv.weight[i] * cx[c];
                }
                __syncthreads();
            }
            for (int half = 0; lv.colour && half < 2 * lv.post; ++half) {
                // reverse colour order: 1, 0, 1, 0, ...
                const int c = 1 - (half & 1);
                tail_gs_half(lv.A, lv.crows[c], lv.ncrows[c], shift, rhs, x);
                __syncthreads();
            }
            for (int p = 0; !lv.colour && p < lv.post; ++p) {
                // the top level's last sweep may write out (HBM): dst is accessed as generic memory
                T *dst = (k == 0 && p == lv.post - 1 && vout) ? vout : xt;
                for (i64 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = gather_row(lv.A, static_cast<u32>(i), x, ex);
                    const T xr = x[i];
                    const T d = shift + ex;
                    const T ax = d * xr - sum;
                    const T res = rhs[i] - ax;
                    dst[i] = xr + omega * res / d;
                }
                __syncthreads();
                if (dst == xt) {
                    xt = x;
                    x = dst;
                } else {
                    x = dst;
                }
            }
            if (k > 0) {
                flip = (flip & ~(1u << k)) | (x == lv.xt ? 1u << k : 0u);
            } else {
                top = x;
            }
        }
        if (cycles > 1) {
            // coarse cycles of the caller's W visit (no permutation: the tail starts below the top
            // level): acc = the first result, acc += each later correction (as an axpy), and the
            // last correction added on the way out; every later cycle starts from zero on the
            // top level's residual rhs_in - A acc
            const level_desc<T> &lv = L[0];
            const bool last = cyc + 1 == cycles;
            T *fin = out ? out : top_x;
            for (i64 i = tid; i < lv.n; i += nt) {
                const T t = top[i];
                if (cyc == 0) {
                    acc[i] = t;
                } else if (!last) {
                    acc[i] += t;
                } else {
                    fin[i] = acc[i] + t;
                }
            }
            __syncthreads();
            if (!last) {
                for (i64 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = gather_row(lv.A, static_cast<u32>(i), acc, ex);
                    const T d = shift + ex;
                    lv.rhs[i] = rhs_in[i] - (d * acc[i] - sum);
                }
                __syncthreads();
            }
        }
    }
    if (cycles > 1) return;
    // the top level's result in out (in the caller's row order), else in top_x
    T *dst = out ? out : top_x;
    if (out && perm0) {
        for (i64 i = tid; i < L[0].n; i += nt) dst[perm0[i]] = top[i];
    } else if (top != dst) {
        for (i64 i = tid; i < L[0].n; i += nt) dst[i] = top[i];
    }
}

template <typename T>
__global__ void __launch_bounds__(kTailThreads)
    k_vcycle_tail(const level_desc<T> *L, int levels, T shift, T omega, const T *rhs_in,
                  const double *rhs64, const double *scale_ptr, T *out, int cycles) {
    tail_body(L, levels, shift, omega, rhs_in, rhs64, scale_ptr, out, L[0].x, cycles);
}

constexpr int kTailLdsLevels = 12;
constexpr u32 kLdsMax = 160 * 1024; // LDS of one workgroup on gfx950

// ---- The tail with every level array in LDS. Each level's arrays are addressed by 32-bit LDS
// addresses held in scalar registers (read once per level visit), so a phase's chain of
// dependent reads starts at the row data: no descriptor reloads after barriers, no generic
// pointers.
template <typename P>
using lptr = __attribute__((address_space(3))) P *;
constexpr u32 kNoLds = 0xFFFFFFFFu; // an absent array

template <typename P>
__device__ __forceinline__ lptr<P> at(u32 a) {
    return (lptr<P>)(a);
}

// a coded operator in LDS; exit != kNoLds: per-row exit rates, else edict[rcode[row]]
struct lds_op {
    u32 rlen, soff, ent, vdict, edict, rcode, exit;
};

template <typename T>
struct lds_level {
    lds_op A, R;
    u32 coarse_of, weight, tstart, tsize, toff, tinv, rhs, x, xt, res, colour, crows[2], perm, acc;
    u32 n, n_next, ncrows[2], n_term, term_max;
    int pre, post, direct, pad;
};

// a level's addresses in scalar registers (every field is the same in all lanes)
template <typename T>
__device__ __forceinline__ lds_level<T> uniform_level(const lds_level<T> *p) {
    auto u = [](u32 v) { return static_cast<u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(v))); };
    lds_level<T> s;
    s.A.rlen = u(p->A.rlen);
    s.A.soff = u(p->A.soff);
    s.A.ent = u(p->A.ent);
    s.A.vdict = u(p->A.vdict);
    s.A.edict = u(p->A.edict);
    s.A.rcode = u(p->A.rcode);
    s.A.exit = u(p->A.exit);
    s.R.rlen = u(p->R.rlen);
    s.R.soff = u(p->R.soff);
    s.R.ent = u(p->R.ent);
    s.R.vdict = u(p->R.vdict);
    s.R.edict = u(p->R.edict);
    s.R.rcode = u(p->R.rcode);
    s.R.exit = u(p->R.exit);
    s.coarse_of = u(p->coarse_of);
    s.weight = u(p->weight);
    s.tstart = u(p->tstart);
    s.tsize = u(p->tsize);
    s.toff = u(p->toff);
    s.tinv = u(p->tinv);
    s.rhs = u(p->rhs);
    s.x = u(p->x);
    s.xt = u(p->xt);
    s.res = u(p->res);
    s.colour = u(p->colour);
    s.crows[0] = u(p->crows[0]);
    s.crows[1] = u(p->crows[1]);
    s.perm = u(p->perm);
    s.acc = u(p->acc);
    s.n = u(p->n);
    s.n_next = u(p->n_next);
    s.ncrows[0] = u(p->ncrows[0]);
    s.ncrows[1] = u(p->ncrows[1]);
    s.n_term = u(p->n_term);
    s.term_max = u(p->term_max);
    s.pre = static_cast<int>(u(static_cast<u32>(p->pre)));
    s.post = static_cast<int>(u(static_cast<u32>(p->post)));
    s.direct = static_cast<int>(u(static_cast<u32>(p->direct)));
    s.pad = 0;
    return s;
}

template <typename T>
__device__ __forceinline__ T lds_exit(const lds_op &A, u32 i) {
    return A.exit != kNoLds ? at<T>(A.exit)[i] : at<T>(A.edict)[at<u8>(A.rcode)[i]];
}

// tail_gather's sum over row i's entries in entry order, eight entries then their eight
// gathers in flight; past the row's end the gathers read entry 0 of both arrays (the row has
// entries, so both exist) and are not added
template <typename T>
__device__ __forceinline__ T lds_gather(const lds_op &A, u32 i, u32 x, T &exit_i) {
    const u32 len = at<u8>(A.rlen)[i];
    const u32 base = at<u32>(A.soff)[i >> 6] + (i & 63u);
    exit_i = lds_exit<T>(A, i);
    const lptr<u32> ent = at<u32>(A.ent);
    const lptr<T> vdict = at<T>(A.vdict), xs = at<T>(x);
    T sum = 0;
    for (u32 k0 = 0; k0 < len; k0 += 8) {
        u32 e[8];
        T r[8], v[8];
#pragma unroll
        for (u32 u = 0; u < 8; ++u) e[u] = k0 + u < len ? ent[base + (k0 + u) * kSlice] : 0u;
#pragma unroll
        for (u32 u = 0; u < 8; ++u) {
            r[u] = vdict[e[u] >> kColBits];
            v[u] = xs[e[u] & kColMask];
        }
#pragma unroll
        for (u32 u = 0; u < 8; ++u) sum = k0 + u < len ? fma(r[u], v[u], sum) : sum;
    }
    return sum;
}

template <typename T>
__device__ __forceinline__ void lds_gs_half(const lds_op &A, u32 rows, u32 nrows, T shift, u32 rhs, u32 x) {
    for (u32 q = threadIdx.x; q < nrows; q += blockDim.x) {
        const u32 i = at<u32>(rows)[q];
        T ex = 0;
        const T sum = lds_gather<T>(A, i, x, ex);
        at<T>(x)[i] = (at<T>(rhs)[i] + sum) / (shift + ex);
    }
}

template <bool ASSIGN, typename T>
__device__ __forceinline__ void lds_terminal(const lds_level<T> &lv, u32 r, u32 x) {
    for (u32 g = threadIdx.x; g < lv.n_term * lv.term_max; g += blockDim.x) {
        const u32 b = g / lv.term_max, i = g % lv.term_max;
        const u32 m = at<u32>(lv.tsize)[b];
        if (i >= m) continue;
        const u32 st = at<u32>(lv.tstart)[b];
        const lptr<T> row = at<T>(lv.tinv) + at<u64>(lv.toff)[b] + static_cast<u64>(i) * m;
        const lptr<T> rr = at<T>(r) + st;
        T acc = 0;
#pragma unroll 8
        for (u32 j = 0; j < m; ++j) acc += row[j] * rr[j];
        if constexpr (ASSIGN) {
            at<T>(x)[st + i] = acc;
        } else {
            at<T>(x)[st + i] += acc;
        }
    }
}

// Copies the image into the arena and writes each level's LDS addresses over the start of the
// descriptors; returns them (every thread of the workgroup calls it)
template <typename T>
__device__ __forceinline__ const lds_level<T> *stage_tail(const uint4 *image, u32 image_bytes, int levels,
                                                          uint4 *arena) {
    const int tid = threadIdx.x;
    const u32 nt = blockDim.x;
    // stage the image: eight 16-byte loads per thread in flight before the LDS stores
    const u32 n16 = image_bytes / 16;
    const uint4 zero{0u, 0u, 0u, 0u};
    auto ld = [&](u32 k) { return k < n16 ? image[k] : zero; };
    auto st = [&](u32 k, const uint4 &v) {
        if (k < n16) arena[k] = v;
    };
    for (u32 b0 = tid; b0 < n16; b0 += 8 * nt) {
        const uint4 v0 = ld(b0), v1 = ld(b0 + nt), v2 = ld(b0 + 2 * nt), v3 = ld(b0 + 3 * nt);
        const uint4 v4 = ld(b0 + 4 * nt), v5 = ld(b0 + 5 * nt), v6 = ld(b0 + 6 * nt), v7 = ld(b0 + 7 * nt);
        st(b0, v0);
        st(b0 + nt, v1);
        st(b0 + 2 * nt, v2);
        st(b0 + 3 * nt, v3);
        st(b0 + 4 * nt, v4);
        st(b0 + 5 * nt, v5);
        st(b0 + 6 * nt, v6);
        st(b0 + 7 * nt, v7);
    }
    __syncthreads();
    // the image starts with the level descriptors, pointer fields holding LDS offset + 1
    const u32 a0 = static_cast<u32>(reinterpret_cast<uintptr_t>((lptr<uint4>)arena));
    auto adr = [&](const void *field) {
        const auto off = reinterpret_cast<uintptr_t>(field);
        return off ? a0 + static_cast<u32>(off - 1) : kNoLds;
    };
    auto op = [&](const csr_view<T> &v) {
        return lds_op{adr(v.rlen), adr(v.soff), adr(v.ent), adr(v.vdict), adr(v.edict), adr(v.rcode),
                      adr(v.exit)};
    };
    lds_level<T> s{};
    if (tid < levels) {
        const level_desc<T> &lv = reinterpret_cast<const level_desc<T> *>(arena)[tid];
        s.A = op(lv.A);
        s.R = op(lv.R);
        s.coarse_of = adr(lv.coarse_of);
        s.weight = adr(lv.weight);
        s.tstart = adr(lv.tstart);
        s.tsize = adr(lv.tsize);
        s.toff = adr(lv.toff);
        s.tinv = adr(lv.tinv);
        s.rhs = adr(lv.rhs);
        s.x = adr(lv.x);
        s.xt = adr(lv.xt);
        s.res = adr(lv.res);
        s.colour = adr(lv.colour);
        s.crows[0] = adr(lv.crows[0]);
        s.crows[1] = adr(lv.crows[1]);
        s.perm = adr(lv.perm);
        s.acc = adr(lv.acc);
        s.n = static_cast<u32>(lv.n);
        s.n_next = static_cast<u32>(lv.n_next);
        s.ncrows[0] = static_cast<u32>(lv.ncrows[0]);
        s.ncrows[1] = static_cast<u32>(lv.ncrows[1]);
        s.n_term = static_cast<u32>(lv.n_term);
        s.term_max = static_cast<u32>(lv.term_max);
        s.pre = lv.pre;
        s.post = lv.post;
        s.direct = lv.direct;
    }
    __syncthreads();
    lds_level<T> *out = reinterpret_cast<lds_level<T> *>(arena);
    if (tid < levels) out[tid] = s;
    __syncthreads();
    return out;
}

// tail_body with the levels in LDS: the same phases, formulas and summation orders. With
// lds_out (one cycle, an output in out), the result is also stored there, in out's row order.
template <typename T>
__device__ __forceinline__ void tail_body_lds(const lds_level<T> *L, int levels, T shift, T omega,
                                              const T *rhs_in, const double *rhs64, const double *scale_ptr,
                                              T *out, T *top_x, int cycles, u32 lds_out = kNoLds) {
    const u32 tid = threadIdx.x, nt = blockDim.x;
    const double scale = scale_ptr ? *scale_ptr : 1.0;
    const lds_level<T> l0 = uniform_level(L);
    const u32 perm0 = l0.perm; // the top level's row order differs from the caller's
    u32 top = kNoLds;          // the top level's result (LDS), unless top_out
    bool top_out = false;      // the top level's last Jacobi sweep wrote out
    for (int cyc = 0; cyc < cycles; ++cyc) {
        T *vout = cycles == 1 ? out : nullptr;
        u32 flip = 0;
        for (int k = 0; k < levels; ++k) {
            const lds_level<T> lv = uniform_level(L + k);
            const u32 rhs = lv.rhs;
            u32 x = lv.x, xt = lv.xt;
            const bool colour = lv.colour != kNoLds;
            if (lv.direct && k > 0) {
                lds_terminal<true>(lv, rhs, x);
                __syncthreads();
                continue;
            }
            for (u32 i = tid; i < lv.n; i += nt) {
                T b;
                if (k == 0 && cyc == 0) {
                    const u32 src = perm0 != kNoLds ? at<u32>(perm0)[i] : i;
                    b = rhs64 ? static_cast<T>(rhs64[src] * scale) : rhs_in[src];
                    at<T>(lv.rhs)[i] = b;
                } else {
                    b = at<T>(rhs)[i];
                }
                if (!colour) {
                    at<T>(x)[i] = lv.pre > 0 ? omega * b / (shift + lds_exit<T>(lv.A, i)) : T(0);
                } else if (lv.pre == 0) {
                    at<T>(x)[i] = T(0);
                } else if (at<u8>(lv.colour)[i] == 0) {
                    at<T>(x)[i] = b / (shift + lds_exit<T>(lv.A, i));
                }
            }
            __syncthreads();
            if (colour) {
                for (int half = 1; half < 2 * lv.pre; ++half) {
                    const int c = half & 1;
                    lds_gs_half<T>(lv.A, c ? lv.crows[1] : lv.crows[0], c ? lv.ncrows[1] : lv.ncrows[0], shift, rhs, x);
                    __syncthreads();
                }
                if (lv.pre > 0) {
                    for (u32 q = tid; q < lv.ncrows[0]; q += nt) {
                        const u32 i = at<u32>(lv.crows[0])[q];
                        T ex = 0;
                        const T sum = lds_gather<T>(lv.A, i, x, ex);
                        const T d = shift + ex;
                        at<T>(lv.res)[i] = at<T>(rhs)[i] - (d * at<T>(x)[i] - sum);
                    }
                    for (u32 q = tid; q < lv.ncrows[1]; q += nt) at<T>(lv.res)[at<u32>(lv.crows[1])[q]] = T(0);
                } else {
                    for (u32 i = tid; i < lv.n; i += nt) {
                        T ex = 0;
                        const T sum = lds_gather<T>(lv.A, i, x, ex);
                        const T d = shift + ex;
                        at<T>(lv.res)[i] = at<T>(rhs)[i] - (d * at<T>(x)[i] - sum);
                    }
                }
                __syncthreads();
            }
            for (int p = 1; !colour && p < lv.pre; ++p) {
                for (u32 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = lds_gather<T>(lv.A, i, x, ex);
                    const T xr = at<T>(x)[i];
                    const T d = shift + ex;
                    const T ax = d * xr - sum;
                    const T res = at<T>(rhs)[i] - ax;
                    at<T>(xt)[i] = xr + omega * res / d;
                }
                __syncthreads();
                const u32 t = x;
                x = xt;
                xt = t;
            }
            if (!colour) {
                for (u32 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = lds_gather<T>(lv.A, i, x, ex);
                    const T d = shift + ex;
                    at<T>(lv.res)[i] = at<T>(rhs)[i] - (d * at<T>(x)[i] - sum);
                }
            }
            __syncthreads();
            if (lv.n_term > 0) {
                lds_terminal<false>(lv, lv.res, x);
                __syncthreads();
            }
            if (lv.n_next > 0) {
                const u32 crhs = L[k + 1].rhs;
                const u32 ch = __builtin_amdgcn_readfirstlane(crhs);
                for (u32 c = tid; c < lv.n_next; c += nt) {
                    T ex = 0;
                    at<T>(ch)[c] = lds_gather<T>(lv.R, c, lv.res, ex);
                }
                __syncthreads();
            }
            if (x == lv.xt) flip |= 1u << k;
        }
        for (int k = levels - 1; k >= 0; --k) {
            const lds_level<T> lv = uniform_level(L + k);
            const u32 rhs = lv.rhs;
            const bool f = (flip >> k) & 1u;
            u32 x = f ? lv.xt : lv.x, xt = f ? lv.x : lv.xt;
            bool x_out = false;
            if (lv.direct && k > 0) continue;
            const bool colour = lv.colour != kNoLds;
            if (lv.n_next > 0) {
                const bool fc = (flip >> (k + 1)) & 1u;
                const u32 cx = __builtin_amdgcn_readfirstlane(fc ? L[k + 1].xt : L[k + 1].x);
                const bool red_only = colour && lv.post > 0;
                const u32 np = red_only ? lv.ncrows[0] : lv.n;
                for (u32 q = tid; q < np; q += nt) {
                    const u32 i = red_only ? at<u32>(lv.crows[0])[q] : q;
                    const u32 c = at<u32>(lv.coarse_of)[i];
                    if (c != 0xFFFFFFFFu) at<T>(x)[i] += at<T>(lv.weight)[i] * at<T>(cx)[c];
                }
                __syncthreads();
            }
            for (int half = 0; colour && half < 2 * lv.post; ++half) {
                const int c = 1 - (half & 1);
                lds_gs_half<T>(lv.A, c ? lv.crows[1] : lv.crows[0], c ? lv.ncrows[1] : lv.ncrows[0], shift, rhs, x);
                __syncthreads();
            }
            for (int p = 0; !colour && p < lv.post; ++p) {
                // the top level's last sweep may write out (HBM)
                const bool to_out = k == 0 && p == lv.post - 1 && vout;
                for (u32 i = tid; i < lv.n; i += nt) {
                    T ex = 0;
                    const T sum = lds_gather<T>(lv.A, i, x, ex);
                    const T xr = at<T>(x)[i];
                    const T d = shift + ex;
                    const T ax = d * xr - sum;
                    const T res = at<T>(rhs)[i] - ax;
                    const T v = xr + omega * res / d;
                    if (to_out) {
                        vout[i] = v;
                        if (lds_out != kNoLds) at<T>(lds_out)[i] = v;
                    } else {
                        at<T>(xt)[i] = v;
                    }
                }
                __syncthreads();
                if (to_out) {
                    x_out = true;
                } else {
                    const u32 t = x;
                    x = xt;
                    xt = t;
                }
            }
            if (k > 0) {
                flip = (flip & ~(1u << k)) | (x == lv.xt ? 1u << k : 0u);
            } else {
                top = x;
                top_out = x_out;
            }
        }
        if (cycles > 1) {
            // coarse cycles of the caller's W visit (no permutation: the tail starts below the top
            // level): acc = the first result, acc += each later correction, the last correction
            // added on the way out; every later cycle starts from zero on the top level's residual
            // rhs_in - A acc
            const bool last = cyc + 1 == cycles;
            T *fin = out ? out : top_x;
            const u32 acc = l0.acc;
            for (u32 i = tid; i < l0.n; i += nt) {
                const T t = at<T>(top)[i];
                if (cyc == 0) {
                    at<T>(acc)[i] = t;
                } else if (!last) {
                    at<T>(acc)[i] += t;
                } else {
                    fin[i] = at<T>(acc)[i] + t;
                }
            }
            __syncthreads();
            if (!last) {
                for (u32 i = tid; i < l0.n; i += nt) {
                    T ex = 0;
                    const T sum = lds_gather<T>(l0.A, i, acc, ex);
                    const T d = shift + ex;
                    at<T>(l0.rhs)[i] = rhs_in[i] - (d * at<T>(acc)[i] - sum);
                }
                __syncthreads();
            }
        }
    }
    if (cycles > 1) return;
    // the top level's result in out (in the caller's row order), else in top_x
    T *dst = out ? out : top_x;
    if (top_out) return;
    if (out && perm0 != kNoLds) {
        for (u32 i = tid; i < l0.n; i += nt) {
            const u32 r = at<u32>(perm0)[i];
            const T v = at<T>(top)[i];
            dst[r] = v;
            if (lds_out != kNoLds) at<T>(lds_out)[r] = v;
        }
    } else {
        for (u32 i = tid; i < l0.n; i += nt) {
            const T v = at<T>(top)[i];
            dst[i] = v;
            if (lds_out != kNoLds) at<T>(lds_out)[i] = v;
        }
    }
}

// The same V-cycle with every array of the levels staged in LDS first (one contiguous read of
// the image of the constant arrays); work vectors live in LDS too.
template <typename T>
__global__ void __launch_bounds__(kTailThreads)
    k_vcycle_tail_lds(const uint4 *image, u32 image_bytes, int levels,
                      T shift, T omega, const T *rhs_in, const double *rhs64, const double *scale_ptr,
                      T *out, T *top_x, int cycles) {
    extern __shared__ uint4 arena[];
    const lds_level<T> *d = stage_tail<T>(image, image_bytes, levels, arena);
    tail_body_lds<T>(d, levels, shift, omega, rhs_in, rhs64, scale_ptr, out, top_x, cycles);
}

// ---- FGMRES in one workgroup: the Krylov loop of a problem of at most kTinyRows states whose
// whole hierarchy fits the LDS tail (one row per thread; the basis in HBM, read back from L2)
constexpr int kTinyWaves = kTailThreads / 64;
constexpr int kTinyStride = kTinyBasis + 2;
struct tiny_scratch {
    double red[kTinyWaves][kTinyStride]; // one partial sum per wave and value
    double h1[kTinyStride], h2[kTinyStride];
    double g[kTinyBasis + 1], cs[kTinyBasis], sn[kTinyBasis], y[kTinyBasis];
    double R[kTinyBasis * (kTinyBasis + 1) / 2]; // column j of the triangular factor at j (j + 1) / 2
    double hn, beta;
    int stop;
};

// sums[k] = the workgroup's sum of term(k) for k < cnt (cnt uniform), each in a fixed order: the
// butterfly within a wave, then the waves in index order. Eight values are reduced at a time so
// that their loads and shuffles overlap.
template <typename F>
__device__ __forceinline__ void tiny_sums(int cnt, F term, tiny_scratch &sc, double *sums) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k0 = 0; k0 < cnt; k0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = k0 + u < cnt ? term(k0 + u) : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (k0 + u < cnt) {
                const double t = wave_sum(v[u]);
                if (lane == 0) sc.red[wave][k0 + u] = t;
            }
        }
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kTinyWaves; ++w) t += sc.red[w][threadIdx.x];
        sums[threadIdx.x] = t;
    }
    __syncthreads();
}

// tiny_row from the generator's coded rows and z staged in LDS (the same sum in the same order)
template <typename ZT>
__device__ __forceinline__ double tiny_row_lds(const lds_op &G, u32 i, u32 z, double shift) {
    const u32 len = at<u8>(G.rlen)[i];
    const u32 base = at<u32>(G.soff)[i >> 6] + (i & 63u);
    const double ex = lds_exit<double>(G, i);
    const lptr<u32> ent = at<u32>(G.ent);
    const lptr<double> vdict = at<double>(G.vdict);
    const lptr<ZT> zs = at<ZT>(z);
    double sum = 0.0;
    for (u32 k0 = 0; k0 < len; k0 += 8) {
        u32 e[8];
        double r[8], v[8];
#pragma unroll
        for (u32 u = 0; u < 8; ++u) e[u] = k0 + u < len ? ent[base + (k0 + u) * kSlice] : 0u;
#pragma unroll
        for (u32 u = 0; u < 8; ++u) {
            r[u] = vdict[e[u] >> kColBits];
            v[u] = static_cast<double>(zs[e[u] & kColMask]);
        }
#pragma unroll
        for (u32 u = 0; u < 8; ++u) sum = k0 + u < len ? fma(r[u], v[u], sum) : sum;
    }
    return (shift + ex) * static_cast<double>(zs[i]) - sum;
}

// ((shift I - Q) z)_i from the generator's gather row i (entry order), z in precision ZT
template <typename ZT>
__device__ __forceinline__ double tiny_row(const csr_view<double> &A, u32 i, const ZT *z, double shift) {
    double sum = 0.0, ex;
    if (A.coded) {
        const u32 len = A.rlen[i];
        const u32 base = A.soff[i >> 6] + (i & 63u);
        ex = A.exit ? A.exit[i] : A.edict[A.rcode[i]];
        // as the LDS tail's gather: eight entries and eight gathers in flight, entry order kept
        for (u32 k0 = 0; k0 < len; k0 += 8) {
            u32 e[8];
            double r[8], v[8];
#pragma unroll
            for (u32 u = 0; u < 8; ++u) e[u] = k0 + u < len ? A.ent[base + (k0 + u) * kSlice] : 0u;
#pragma unroll
            for (u32 u = 0; u < 8; ++u) {
                r[u] = A.vdict[e[u] >> kColBits];
                v[u] = static_cast<double>(z[e[u] & kColMask]);
            }
#pragma unroll
            for (u32 u = 0; u < 8; ++u) sum = k0 + u < len ? fma(r[u], v[u], sum) : sum;
        }
    } else {
        ex = A.exit[i];
        for (u32 p = A.ptr[i]; p < A.ptr[i + 1]; ++p) sum += A.val[p] * static_cast<double>(z[A.col[p]]);
    }
    return (shift + ex) * static_cast<double>(z[i]) - sum;
}

// Restarted flexible GMRES(m) (classical Gram-Schmidt applied twice, Givens rotations) with the
// LDS tail's V-cycle as the preconditioner, every phase separated by workgroup barriers; the same
// loop, stopping tests and result as the host-driven solve. Every control value is computed by
// thread 0 into LDS and read by all threads after a barrier, so the branches are uniform.
template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_gmres_tiny(tiny_gmres_args a) {
    extern __shared__ uint4 arena[];
    const lds_level<T> *d = stage_tail<T>(static_cast<const uint4 *>(a.image), a.image_bytes, a.levels, arena);
    tiny_scratch &sc = *reinterpret_cast<tiny_scratch *>(reinterpret_cast<char *>(arena) +
                                                         ((a.lds_bytes + 15u) & ~15u));
    const int tid = threadIdx.x;
    const bool active = tid < a.n;
    const u32 i = static_cast<u32>(tid);
    const i64 ld = a.ld;
    // this workgroup's system
    const i64 sys = blockIdx.x;
    a.b += sys * a.sys_x;
    a.x += sys * a.sys_x;
    a.result += 3 * sys;
    double *V = a.V + sys * a.sys_v;
    T *Z = static_cast<T *>(a.Z) + sys * a.sys_z;
    const T shift_t = static_cast<T>(a.shift), omega_t = static_cast<T>(a.omega);
    const double tol = a.tols ? a.tols[sys] : a.tolerance;
    // r = b - (shift I - Q) x (kept in r) and |r|
    // the generator's coded rows, and z_j, in LDS when the launcher placed them after the scratch
    const bool gen_lds = a.g_ent != 0;
    const u32 a0 = static_cast<u32>(reinterpret_cast<uintptr_t>((lptr<uint4>)arena));
    lds_op G{kNoLds, kNoLds, kNoLds, kNoLds, kNoLds, kNoLds, kNoLds};
    u32 zl = kNoLds;
    if (gen_lds) {
        auto place = [&](u32 off) { return off ? a0 + off : kNoLds; };
        G = lds_op{place(a.g_rlen), place(a.g_soff), place(a.g_ent), place(a.g_vdict), place(a.g_edict),
                   place(a.g_rcode), place(a.g_exit)};
        zl = place(a.g_z);
        const csr_view<double> &A = a.A;
        for (u32 k = tid; k <= A.n_slices; k += blockDim.x) at<u32>(G.soff)[k] = A.soff[k];
        for (u64 k = tid; k < a.n_ent; k += blockDim.x) at<u32>(G.ent)[k] = A.ent[k];
        for (u32 k = tid; k < static_cast<u32>(a.n); k += blockDim.x) {
            at<u8>(G.rlen)[k] = A.rlen[k];
            if (G.exit != kNoLds) {
                at<double>(G.exit)[k] = A.exit[k];
            } else {
                at<u8>(G.rcode)[k] = A.rcode[k];
            }
        }
        for (int k = tid; k < A.n_vdict; k += blockDim.x) at<double>(G.vdict)[k] = A.vdict[k];
        for (int k = tid; G.exit == kNoLds && k < A.n_edict; k += blockDim.x) at<double>(G.edict)[k] = A.edict[k];
        __syncthreads();
    }
    double r = 0.0;
    auto residual = [&]() {
        r = active ? a.b[i] - tiny_row(a.A, i, a.x, a.shift) : 0.0;
        tiny_sums(1, [&](int) { return r * r; }, sc, &sc.beta);
        return sqrt(sc.beta);
    };
    int it = 0, converged = 0;
    double beta = residual();
    if (!(beta > tol) || !__builtin_isfinite(beta)) {
        converged = __builtin_isfinite(beta) ? 1 : 0;
    } else {
        while (it < a.max_iterations) {
            if (active) V[i] = r / beta;
            if (tid == 0) {
                sc.g[0] = beta;
                for (int k = 1; k <= a.m; ++k) sc.g[k] = 0.0;
            }
            int j = 0;
            while (j < a.m && it < a.max_iterations) {
                ++it;
                __syncthreads();
                // z_j = M v_j (the tail reads v_j in its own row order), w = (shift I - Q) z_j
                T *zj = Z + j * ld;
                tail_body_lds<T>(d, a.levels, shift_t, omega_t, nullptr, V + j * ld, nullptr, zj, nullptr, 1, zl);
                __syncthreads();
                double w = 0.0;
                if (active) w = gen_lds ? tiny_row_lds<T>(G, i, zl, a.shift) : tiny_row(a.A, i, zj, a.shift);
                const int cnt = j + 1;
                auto proj = [&](int k) { return active ? V[k * ld + i] * w : 0.0; };
                tiny_sums(cnt, proj, sc, sc.h1);
                if (active)
                    for (int k = 0; k < cnt; ++k) w -= V[k * ld + i] * sc.h1[k];
                tiny_sums(cnt, proj, sc, sc.h2);
                if (active)
                    for (int k = 0; k < cnt; ++k) w -= V[k * ld + i] * sc.h2[k];
                tiny_sums(1, [&](int) { return w * w; }, sc, &sc.hn);
                const double hn = sqrt(sc.hn);
                if (active && hn > 0.0) V[(j + 1) * ld + i] = w / hn;
                if (tid == 0) {
                    // column j = h1 + h2 and hn, rotated by the previous rotations; a new rotation
                    // zeroes hn
                    double *c = sc.h1;
                    for (int k = 0; k <= j; ++k) c[k] += sc.h2[k];
                    for (int k = 0; k < j; ++k) {
                        const double x0 = c[k], x1 = c[k + 1];
                        c[k] = sc.cs[k] * x0 + sc.sn[k] * x1;
                        c[k + 1] = -sc.sn[k] * x0 + sc.cs[k] * x1;
                    }
                    const double x0 = c[j], dd = hypot(x0, hn);
                    sc.cs[j] = dd > 0.0 ? x0 / dd : 1.0;
                    sc.sn[j] = dd > 0.0 ? hn / dd : 0.0;
                    c[j] = dd;
                    double *Rj = sc.R + j * (j + 1) / 2;
                    for (int k = 0; k <= j; ++k) Rj[k] = c[k];
                    sc.g[j + 1] = -sc.sn[j] * sc.g[j];
                    sc.g[j] = sc.cs[j] * sc.g[j];
                    const double est = fabs(sc.g[j + 1]);
                    sc.stop = !(est > tol) || !(hn > 0.0) || !__builtin_isfinite(hn);
                }
                __syncthreads();
                ++j;
                if (sc.stop) break;
            }
            // y = R^-1 g over the cycle's j columns; x += Z y
            if (tid == 0) {
                for (int q = j; q-- > 0;) {
                    double s = sc.g[q];
                    for (int k = q + 1; k < j; ++k) s -= sc.R[k * (k + 1) / 2 + q] * sc.y[k];
                    sc.y[q] = s / sc.R[q * (q + 1) / 2 + q];
                }
            }
            __syncthreads();
            if (active) {
                double s = a.x[i];
                for (int k = 0; k < j; ++k) s += sc.y[k] * static_cast<double>(Z[k * ld + i]);
                a.x[i] = s;
            }
            __syncthreads();
            beta = residual();
            if (!__builtin_isfinite(beta)) break;
            if (!(beta > tol)) {
                converged = 1;
                break;
            }
        }
    }
    if (tid == 0) {
        a.result[0] = it;
        a.result[1] = beta;
        a.result[2] = converged;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill(i64 n, T v, T *x) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        x[i] = v;
}

template <typename A, typename B>
__global__ void __launch_bounds__(kThreads)
    k_convert(i64 n, const A *in, const double *scale_ptr, B *out) {
    const A scale = scale_ptr ? static_cast<A>(*scale_ptr) : A(1);
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        out[i] = static_cast<B>(in[i] * scale);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_axpy(i64 n, T alpha, const T *x, T *y) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        y[i] += alpha * x[i];
}

// four 16-byte loads in flight per thread, the grid's blocks 4 * 256 elements apart
__global__ void __launch_bounds__(kThreads) k_copy16(i64 n, const uint4 *src, uint4 *dst) {
    const i64 stride = static_cast<i64>(gridDim.x) * kThreads * 4;
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) * 4 + threadIdx.x; i < n; i += stride) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * kThreads < n) v[u] = src[i + u * kThreads];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * kThreads < n) dst[i + u * kThreads] = v[u];
    }
}

// ---- Krylov kernels ----
__global__ void __launch_bounds__(kThreads)
    k_multi_dot(const double *V, i64 ld, int count, const double *w, i64 n, double *partial,
                int nb) {
    __shared__ double red[kWaves];
    const int k0 = blockIdx.y * kDotChunk;
    const int kc = min(kDotChunk, count - k0);
    double acc[kDotChunk];
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) acc[k] = 0.0;
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(nb) * kThreads) {
        const double wi = w[i];
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k)
            if (k < kc) acc[k] += V[(k0 + k) * ld + i] * wi;
    }
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) {
        if (k < kc) {
            const double s = block_sum(acc[k], red);
            if (threadIdx.x == 0) partial[static_cast<i64>(k0 + k) * nb + blockIdx.x] = s;
        }
    }
}

constexpr int kMaxBasis = 512;

__global__ void __launch_bounds__(kThreads)
    k_basis_update(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                   double *partial, int nb) {
    __shared__ double hs[kMaxBasis];
    __shared__ double red[kWaves];
    for (int k = threadIdx.x; k < count; k += kThreads) hs[k] = h[k];
    __syncthreads();
    double sq = 0.0;
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(nb) * kThreads) {
        double wi = w[i];
        for (int k = 0; k < count; ++k) wi -= hs[k] * V[k * ld + i];
        w[i] = wi;
        sq += wi * wi;
    }
    if (partial) {
        const double s = block_sum(sq, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kThreads)
    k_reduce(const double *partial, int nb, double *out) {
    __shared__ double red[kWaves];
    const double *p = partial + static_cast<i64>(blockIdx.x) * nb;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) s += p[b];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads)
    k_scale(i64 n, const double *w, const double *norm2, double *v) {
    const double nrm = sqrt(norm2[0]);
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads)
        v[i] = w[i] / nrm;
}

template <typename ZT>
__global__ void __launch_bounds__(kThreads)
    k_combine(double *x, const ZT *Z, i64 ld, int count, const double *y, i64 n) {
    __shared__ double ys[kMaxBasis];
    for (int k = threadIdx.x; k < count; k += kThreads) ys[k] = y[k];
    __syncthreads();
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads) {
        double s = x[i];
        for (int k = 0; k < count; ++k) s += ys[k] * static_cast<double>(Z[k * ld + i]);
        x[i] = s;
    }
}

constexpr int kTileRows = 64;
constexpr int kPerWave = kFusedBasis / kWaves; // basis vectors per wave

__global__ void __launch_bounds__(kThreads)
    k_update_dot(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                 double *partial, int nb) {
    __shared__ double part[kWaves][kTileRows];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double hk[kPerWave], acc[kPerWave];
#pragma unroll
    for (int kk = 0; kk < kPerWave; ++kk) {
        const int k = kk * kWaves + wave;
        hk[kk] = k < count ? h[k] : 0.0;
        acc[kk] = 0.0;
    }
    double nrm = 0.0;
    for (i64 r0 = static_cast<i64>(blockIdx.x) * kTileRows; r0 < n;
         r0 += static_cast<i64>(nb) * kTileRows) {
        const i64 r = r0 + lane;
        const bool ok = r < n;
        double v[kPerWave];
        double u = 0.0;
#pragma unroll
        for (int kk = 0; kk < kPerWave; ++kk) {
            const int k = kk * kWaves + wave;
            v[kk] = (ok && k < count) ? V[k * ld + r] : 0.0;
            u += hk[kk] * v[kk];
        }
        part[wave][lane] = u;
        __syncthreads();
        double wr = 0.0;
        if (ok) wr = w[r] - (((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane]);
        __syncthreads();
        if (ok) {
            if (wave == 0) {
                w[r] = wr;
                nrm += wr * wr;
            }
#pragma unroll
            for (int kk = 0; kk < kPerWave; ++kk) acc[kk] += v[kk] * wr;
        }
    }
#pragma unroll
    for (int kk = 0; kk < kPerWave; ++kk) {
        const int k = kk * kWaves + wave;
        if (k < count) {
            const double t = wave_sum(acc[kk]);
            if (lane == 0) partial[static_cast<i64>(k) * nb + blockIdx.x] = t;
        }
    }
    if (wave == 0) {
        const double t = wave_sum(nrm);
        if (lane == 0) partial[static_cast<i64>(count) * nb + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kThreads)
    k_update_scale(const double *w, const double *V, i64 ld, int count, const double *cw, i64 n,
                   double *v, double *beta_flag) {
    __shared__ double cs[kFusedBasis];
    __shared__ double beta_s;
    for (int k = threadIdx.x; k < count; k += kThreads) cs[k] = cw[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double c2 = 0.0;
        for (int k = 0; k < count; ++k) c2 += cs[k] * cs[k];
        const double w2 = cw[count];
        const bool trusted = c2 <= 0.25 * w2;
        const double beta = trusted ? sqrt(w2 - c2) : 1.0;
        beta_s = beta;
        if (blockIdx.x == 0) {
            beta_flag[0] = beta;
            beta_flag[1] = trusted ? 0.0 : 1.0;
        }
    }
    __syncthreads();
    const double beta = beta_s;
    for (i64 i = blockIdx.x * static_cast<i64>(kThreads) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreads) {
        double wi = w[i];
        for (int k = 0; k < count; ++k) wi -= cs[k] * V[k * ld + i];
        v[i] = wi / beta;
    }
}

// One step of delayed CGS2. V[0 .. count) is the final basis Q, V[count] the candidate u
// (orthogonalised once), w = A z. s = Q^T u, tg[0 .. count) = Q^T w, tg[count] = u^T w,
// beta = |u - Q s|. Rows own lanes of 64-row tiles; waves split the basis, wave wv holding the
// PW vectors k = 4 kk + wv (PW = ceil(count / 4); slots past count load vector count - 1 with
// zero coefficients, so the loads need no branch) in registers for the projections of the new
// candidate:
//   q   = (u - Q s) / beta                          -> V[count]
//   eta = (u^T w - s^T Q^T w) / beta   (= q^T w)     -> eta_out (block 0)
//   u'  = w - Q (Q^T w) - eta q                     -> V[count + 1]
//   partial[k nb + block]: Q_k . u' (k < count), q . u' (k = count), u' . u' (k = count + 1)
template <int PW>
__global__ void __launch_bounds__(kThreads)
    k_dcgs_update(double *V, i64 ld, int count, const double *sc, const double *tg,
                  const double *beta_ptr, const double *w, i64 n, double *partial, int nb,
                  double *eta_out) {
    __shared__ double part[2][kWaves][kTileRows];
    __shared__ double uw[2][kTileRows];
    __shared__ double2 coef[kFusedBasis]; // (s_k, (Q^T w)_k), zero from count on
    __shared__ double eta_s;
    constexpr int kSlots = PW > 0 ? PW : 1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int k = threadIdx.x; k < kFusedBasis; k += kThreads)
        coef[k] = k < count ? make_double2(sc[k], tg[k]) : make_double2(0.0, 0.0);
    const double beta = *beta_ptr;
    if (threadIdx.x == 0) {
        double st = 0.0;
        for (int k = 0; k < count; ++k) st += sc[k] * tg[k];
        const double eta = (tg[count] - st) / beta;
        eta_s = eta;
        if (blockIdx.x == 0) eta_out[0] = eta;
    }
    __syncthreads();
    const double eta = eta_s;
    const double ib = 1.0 / beta;
    double *const q_out = V + static_cast<i64>(count) * ld;
    double *const u_out = V + static_cast<i64>(count + 1) * ld;
    const double *src[kSlots];
#pragma unroll
    for (int kk = 0; kk < PW; ++kk)
        src[kk] = V + static_cast<i64>(min(kk * kWaves + wave, count - 1)) * ld;
    const double *xsrc = wave == 0 ? q_out : w; // waves 0 and 1 stage u and w
    double acc[kSlots], v[kSlots];
#pragma unroll
    for (int kk = 0; kk < PW; ++kk) acc[kk] = 0.0;
    double own = 0.0; // wave 2: q . u', wave 3: u' . u'
    for (i64 r0 = static_cast<i64>(blockIdx.x) * kTileRows; r0 < n;
         r0 += static_cast<i64>(nb) * kTileRows) {
        const i64 r = r0 + lane;
        const bool ok = r < n;
        const i64 rl = ok ? r : n - 1;
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) v[kk] = src[kk][rl];
        const double x = wave < 2 ? xsrc[rl] : 0.0;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) {
            const double2 c = coef[kk * kWaves + wave];
            a += c.x * v[kk];
            b += c.y * v[kk];
        }
        part[0][wave][lane] = a;
        part[1][wave][lane] = b;
        if (wave < 2) uw[wave][lane] = x;
        __syncthreads();
        const double S = ((part[0][0][lane] + part[0][1][lane]) + part[0][2][lane]) + part[0][3][lane];
        const double T = ((part[1][0][lane] + part[1][1][lane]) + part[1][2][lane]) + part[1][3][lane];
        const double q = (uw[0][lane] - S) * ib;
        const double up = ok ? (uw[1][lane] - T) - eta * q : 0.0;
        __syncthreads();
        if (ok) {
            if (wave == 0) q_out[r] = q;
            if (wave == 1) u_out[r] = up;
        }
        if (wave == 2) own += q * up;
        if (wave == 3) own += up * up;
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) acc[kk] += v[kk] * up;
    }
#pragma unroll
    for (int kk = 0; kk < PW; ++kk) {
        const int k = kk * kWaves + wave;
        if (k < count) {
            const double t = wave_sum(acc[kk]);
            if (lane == 0) partial[static_cast<i64>(k) * nb + blockIdx.x] = t;
        }
    }
    if (wave >= 2) {
        const double t = wave_sum(own);
        if (lane == 0) partial[static_cast<i64>(count + wave - 2) * nb + blockIdx.x] = t;
    }
}

int grid_for(i64 n) {
    const i64 g = (n + kThreads - 1) / kThreads;
    return static_cast<int>(std::max<i64>(1, std::min<i64>(g, 8192)));
}

} // namespace

void set_xcd_chunk(int chunk) { g_xcd_chunk = chunk < 0 ? 0 : chunk; }
int xcd_chunk() { return g_xcd_chunk; }

int reduction_blocks(i64 n) {
    const i64 g = (n + 4 * kThreads - 1) / (4 * kThreads);
    return static_cast<int>(std::max<i64>(1, std::min<i64>(g, 2048)));
}

template <typename XT>
void shifted_apply(const csr_view<double> &A, double shift, const XT *x, const double *b,
                   double *y, double *partial, hipStream_t s) {
    const csr_args<double, XT, double> a{A, shift, 0.0, x, b, y, partial, nullptr, nullptr, nullptr, 0u};
    if (b) {
        launch_csr<csr_mode::shifted_residual>(a, s);
    } else {
        launch_csr<csr_mode::shifted>(a, s);
    }
}

template <typename T, typename IN>
void jacobi_first(const csr_view<T> &A, double shift, double omega, const IN *rhs_in, T *rhs_out,
                  T *x, hipStream_t s, const double *scale) {
    if (A.n == 0) return;
    hipLaunchKernelGGL((k_jacobi_first<T, IN>), dim3(grid_for(A.n)), dim3(kThreads), 0, s, A,
                       static_cast<T>(shift), static_cast<T>(omega), rhs_in, scale, rhs_out, x);
}

template <typename T>
void jacobi_sweep(const csr_view<T> &A, double shift, double omega, const T *rhs, const T *x,
                  T *x_out, hipStream_t s) {
    launch_csr<csr_mode::sweep>(csr_args<T, T, T>{A, shift, static_cast<T>(omega), x, rhs, x_out,
                                                  nullptr, nullptr, nullptr, nullptr, 0u}, s);
}

template <typename T>
void level_residual(const csr_view<T> &A, double shift, const T *rhs, const T *x, T *r,
                    hipStream_t s) {
    launch_csr<csr_mode::residual>(
        csr_args<T, T, T>{A, shift, T(0), x, rhs, r, nullptr, nullptr, nullptr, nullptr, 0u}, s);
}

template <typename T, typename IN>
void rb_first_scatter(i64 n, const T *exit_box, double shift, const IN *rhs_in, T *rhs_out, T *x,
                      hipStream_t s, const double *scale, const u32 *pos) {
    if (n == 0) return;
    const i64 blocks = std::max<i64>(1, std::min<i64>((n + kThreads - 1) / kThreads, 65536));
    hipLaunchKernelGGL((k_rb_first_scatter<T, IN>), dim3(static_cast<u32>(blocks)), dim3(kThreads), 0, s, n,
                       exit_box, static_cast<T>(shift), rhs_in, scale, pos, rhs_out, x);
}

template <typename T>
void gs_half(const csr_view<T> &Ac, double shift, const T *rhs, T *x, T *res_zero, hipStream_t s,
             T *nat, const u32 *perm) {
    launch_csr<csr_mode::gs>(csr_args<T, T, T>{Ac, shift, T(0), x, rhs, x, nullptr, res_zero, nat, perm, 0u},
                             s);
}

template <typename T, typename IN>
void permute_in(i64 n, const u32 *perm, const IN *in, T *out, hipStream_t s, const double *scale) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_permute_in<T, IN>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, perm, in, scale,
                       out);
}

template <typename T>
void permute_out(i64 n, const u32 *perm, const T *in, T *out, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_permute_out<T>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, perm, in, out);
}

template <typename T, typename IN>
void rb_first(const csr_view<T> &A, const u8 *colour, double shift, const IN *rhs_in, T *rhs_out,
              T *x, hipStream_t s, const double *scale, const u32 *perm) {
    if (A.n == 0) return;
    hipLaunchKernelGGL((k_rb_first<T, IN>), dim3(grid_for(A.n)), dim3(kThreads), 0, s, A, colour,
                       static_cast<T>(shift), rhs_in, scale, perm, rhs_out, x);
}

template <typename T>
void ilu_forward(const csr_view<T> &A, const u32 *rows, const u32 *gptr, int g0, int g1,
                 i64 wide_rows, const T *pinv, const T *r, T *u, hipStream_t s) {
    if (g1 <= g0) return;
    if (wide_rows > 0) {
        hipLaunchKernelGGL((k_ilu_group<T, false, false>), dim3(grid_for(wide_rows)), dim3(kThreads), 0, s,
                           A, rows, gptr, g0, g0 + 1, pinv, r, u, nullptr);
    } else {
        hipLaunchKernelGGL((k_ilu_group<T, false, true>), dim3(1), dim3(kTailThreads), 0, s, A, rows,
                           gptr, g0, g1, pinv, r, u, nullptr);
    }
}

template <typename T>
void ilu_backward(const csr_view<T> &A, const u32 *rows, const u32 *gptr, int g0, int g1,
                  i64 wide_rows, const T *pinv, T *u, T *x, hipStream_t s) {
    if (g1 <= g0) return;
    if (wide_rows > 0) {
        hipLaunchKernelGGL((k_ilu_group<T, true, false>), dim3(grid_for(wide_rows)), dim3(kThreads), 0, s,
                           A, rows, gptr, g0, g0 + 1, pinv, nullptr, u, x);
    } else {
        hipLaunchKernelGGL((k_ilu_group<T, true, true>), dim3(1), dim3(kTailThreads), 0, s, A, rows,
                           gptr, g0, g1, pinv, nullptr, u, x);
    }
}

template <typename T>
void gather_sum(const csr_view<T> &R, const T *x, T *y, hipStream_t s) {
    launch_csr<csr_mode::gather_sum>(
        csr_args<T, T, T>{R, 0.0, T(0), x, nullptr, y, nullptr, nullptr, nullptr, nullptr, 0u}, s);
}

template <typename T>
void gather_sum_first(const csr_view<T> &R, const T *x, T *y, const csr_view<T> &Ac,
                      const std::uint8_t *colour, double shift, T *cx, hipStream_t s) {
    csr_args<T, T, T> a{R, shift, T(0), x, nullptr, y, nullptr, nullptr, nullptr, nullptr, 0u};
    a.first_colour = colour;
    a.first_A = Ac;
    a.first_x = cx;
    launch_csr<csr_mode::gather_sum>(a, s);
}

template <typename T>
void level_residual_first(const csr_view<T> &A, double shift, const T *rhs, const T *x, T *r,
                          const csr_view<T> &Ac, const std::uint8_t *colour, T *cx, hipStream_t s) {
    csr_args<T, T, T> a{A, shift, T(0), x, rhs, r, nullptr, nullptr, nullptr, nullptr, 0u};
    a.first_colour = colour;
    a.first_A = Ac;
    a.first_x = cx;
    launch_csr<csr_mode::residual>(a, s);
}

template <typename T>
void restrict_children(i64 n_coarse, const u32 *cptr, const u32 *cidx, const T *fine, T *coarse,
                       hipStream_t s) {
    if (n_coarse == 0) return;
    hipLaunchKernelGGL((k_restrict<T>), dim3(grid_for(n_coarse)), dim3(kThreads), 0, s, n_coarse,
                       cptr, cidx, fine, coarse);
}

template <typename T>
void prolong_add(i64 n, const u32 *coarse_of, const T *weight, const T *coarse, T *x,
                 hipStream_t s, const u32 *rows, const T *coarse2) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_prolong_add<T>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, rows, coarse_of,
                       weight, coarse, coarse2, x);
}

template <typename T>
void terminal_solve(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                    const T *inverse, i64 max_size, const T *r, T *x, hipStream_t s, bool assign) {
    if (n_blocks == 0) return;
    if (max_size > kWideTerminal) {
        const i64 threads = n_blocks * max_size * 64;
        const int grid = static_cast<int>((threads + kThreads - 1) / kThreads);
        hipLaunchKernelGGL((k_terminal_wide<T>), dim3(grid), dim3(kThreads), 0, s, n_blocks, start,
                           size, offset, inverse, max_size, r, x, static_cast<int>(assign));
        return;
    }
    const i64 threads = n_blocks * max_size;
    const int grid = static_cast<int>((threads + kThreads - 1) / kThreads);
    hipLaunchKernelGGL((k_terminal<T>), dim3(grid), dim3(kThreads), 0, s, n_blocks, start, size,
                       offset, inverse, max_size, r, x, static_cast<int>(assign));
}

template <typename T>
void terminal_prolong(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                      const T *inverse, i64 max_size, const T *r, T *xc, const csr_view<T> &R,
                      const u8 *fine_colour, const T *weight, T *x_fine, hipStream_t s) {
    if (n_blocks == 0) return;
    const i64 threads = n_blocks * max_size * 64;
    const int grid = static_cast<int>((threads + kThreads - 1) / kThreads);
    hipLaunchKernelGGL((k_terminal_prolong<T>), dim3(grid), dim3(kThreads), 0, s, n_blocks, start, size,
                       offset, inverse, max_size, r, xc, R, fine_colour, weight, x_fine);
}

namespace {
u32 blocks_for(u64 n) { return static_cast<u32>((n + kThreads - 1) / kThreads); }
} // namespace

template <typename T>
void block_solve(const u32 *members, const u32 *member_block, const u32 *block_first,
                 const u64 *inv_off, const T *inv, u64 m0, u64 m1, const T *r, T *x, T *escr,
                 hipStream_t s) {
    if (m1 <= m0) return;
    hipLaunchKernelGGL((k_block_solve<T>), dim3(blocks_for(m1 - m0)), dim3(kThreads), 0, s, members,
                       member_block, block_first, inv_off, inv, m0, m1, r, x, escr);
}

template <typename T>
void block_residual(const csr_view<T> &A, double shift, const u32 *rows, u64 q0, u64 q1,
                    const T *escr, T *r, hipStream_t s) {
    if (q1 <= q0) return;
    hipLaunchKernelGGL((k_block_residual<T>), dim3(blocks_for(q1 - q0)), dim3(kThreads), 0, s, A,
                       static_cast<T>(shift), rows, q0, q1, escr, r);
}

template <typename T>
void block_zero(const u32 *members, u64 m0, u64 m1, T *escr, hipStream_t s) {
    if (m1 <= m0) return;
    hipLaunchKernelGGL((k_block_zero<T>), dim3(blocks_for(m1 - m0)), dim3(kThreads), 0, s, members,
                       m0, m1, escr);
}

template <typename T>
void sweep_rows(const csr_view<T> &A, double shift, double omega, const T *rhs, const T *x, T *out,
                const u32 *rows, u64 nrows, hipStream_t s) {
    if (nrows == 0) return;
    hipLaunchKernelGGL((k_sweep_rows<T>), dim3(blocks_for(nrows)), dim3(kThreads), 0, s, A,
                       static_cast<T>(shift), static_cast<T>(omega), rhs, x, out, rows, nrows);
}

template <typename T>
void copy_rows(const u32 *rows, u64 nrows, const T *src, T *dst, hipStream_t s) {
    if (nrows == 0) return;
    hipLaunchKernelGGL((k_copy_rows<T>), dim3(blocks_for(nrows)), dim3(kThreads), 0, s, rows, nrows,
                       src, dst);
}

template <typename T>
void vcycle_tail(const level_desc<T> *L, int levels, double shift, double omega, const T *rhs,
                 const double *rhs64, const double *scale, T *out, hipStream_t s, int cycles) {
    if (levels <= 0) return;
    hipLaunchKernelGGL((k_vcycle_tail<T>), dim3(1), dim3(kTailThreads), 0, s, L, levels,
                       static_cast<T>(shift), static_cast<T>(omega), rhs, rhs64, scale, out, cycles);
}

template <typename T>
bool vcycle_tail_lds_fits(u32 lds_bytes, int levels) {
    return levels <= kTailLdsLevels && lds_bytes <= kTailLdsMax;
}

template <typename T>
void vcycle_tail_lds(const void *image, u32 image_bytes, u32 lds_bytes,
                     int levels, double shift, double omega, const T *rhs, const double *rhs64,
                     const double *scale, T *out, T *top_x, hipStream_t s, int cycles) {
    if (levels <= 0) return;
    static const bool attr = [] {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vcycle_tail_lds<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(kTailLdsMax)) == hipSuccess;
    }();
    (void)attr;
    hipLaunchKernelGGL((k_vcycle_tail_lds<T>), dim3(1), dim3(kTailThreads), lds_bytes, s,
                       static_cast<const uint4 *>(image), image_bytes, levels, static_cast<T>(shift),
                       static_cast<T>(omega), rhs, rhs64, scale, out, top_x, cycles);
}

u32 gmres_tiny_lds(u32 tail_lds_bytes) {
    const u64 total = ((static_cast<u64>(tail_lds_bytes) + 15u) & ~u64(15)) + sizeof(tiny_scratch);
    return total <= kLdsMax ? static_cast<u32>(total) : 0u;
}

template <typename T>
void gmres_tiny(const tiny_gmres_args &args, hipStream_t s) {
    u32 lds = gmres_tiny_lds(args.lds_bytes);
    if (lds == 0 || args.n > kTinyRows || args.m > kTinyBasis || args.m < 1 || args.levels > kTailLdsLevels)
        throw std::invalid_argument("gmres_tiny: the problem does not fit one workgroup");
    if (args.systems < 1 || (args.systems > 1 && (args.sys_x < args.n || args.sys_v < (args.m + 1) * args.ld ||
                                                  args.sys_z < args.m * args.ld)))
        throw std::invalid_argument("gmres_tiny: the systems' vectors overlap");
    // the generator's coded rows and z after the scratch, where they fit: the product w = A z_j
    // then gathers from LDS instead of L2 (measured on the 969-state case: 5.7% more solves per
    // second with an FP32 hierarchy, 1.3% fewer with an FP64 one, whose z gathers are 8 bytes)
    tiny_gmres_args a = args;
    a.g_soff = a.g_ent = a.g_rlen = a.g_rcode = a.g_vdict = a.g_edict = a.g_exit = a.g_z = 0;
    if (a.A.coded && sizeof(T) == 4) {
        u64 off = lds;
        auto take = [&](u64 bytes) {
            off = (off + 15) & ~u64(15);
            const u64 at_ = off;
            off += bytes;
            return static_cast<u32>(at_);
        };
        const u64 n = static_cast<u64>(a.n);
        tiny_gmres_args g = a;
        g.g_soff = take((a.A.n_slices + 1) * 4);
        g.g_ent = take(a.n_ent * 4 + 4);
        g.g_rlen = take(n);
        if (a.A.exit) {
            g.g_exit = take(n * 8);
        } else {
            g.g_rcode = take(n);
            g.g_edict = take(static_cast<u64>(a.A.n_edict) * 8);
        }
        g.g_vdict = take(static_cast<u64>(a.A.n_vdict) * 8);
        g.g_z = take(n * sizeof(T));
        if (off <= kLdsMax) {
            a = g;
            lds = static_cast<u32>(off);
        }
    }
    static const bool attr = [] {
        return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gmres_tiny<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(kLdsMax)) == hipSuccess;
    }();
    if (!attr) throw std::runtime_error("gmres_tiny: cannot raise the kernel's LDS limit");
    hipLaunchKernelGGL((k_gmres_tiny<T>), dim3(std::max(a.systems, 1)), dim3(kTailThreads), lds, s, a);
}

template <typename T>
void fill(i64 n, T value, T *x, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, value, x);
}

template <typename A, typename B>
void convert(i64 n, const A *in, B *out, hipStream_t s, const double *scale) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_convert<A, B>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, in, scale,
                       out);
}

template <typename T>
void axpy(i64 n, T alpha, const T *x, T *y, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_axpy<T>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, alpha, x, y);
}

void copy16(i64 n, const void *src, void *dst, hipStream_t s) {
    if (n == 0) return;
    const i64 blocks = std::max<i64>(1, std::min<i64>((n + 4 * kThreads - 1) / (4 * kThreads), 65536));
    hipLaunchKernelGGL(k_copy16, dim3(static_cast<u32>(blocks)), dim3(kThreads), 0, s, n,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst));
}

void multi_dot(const double *V, i64 ld, int count, const double *w, i64 n, double *partial,
               int nb, hipStream_t s) {
    if (count <= 0) return;
    const int chunks = (count + kDotChunk - 1) / kDotChunk;
    hipLaunchKernelGGL(k_multi_dot, dim3(nb, chunks), dim3(kThreads), 0, s, V, ld, count, w, n,
                       partial, nb);
}

void basis_update(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                  double *partial, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_basis_update, dim3(nb), dim3(kThreads), 0, s, w, V, ld, count, h, n,
                       partial, nb);
}

void reduce_partials(const double *partial, int nb, int count, double *out, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_reduce, dim3(count), dim3(kThreads), 0, s, partial, nb, out);
}

void scale_by_norm(i64 n, const double *w, const double *norm2, double *v, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kThreads), 0, s, n, w, norm2, v);
}

void update_dot(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                double *partial, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_update_dot, dim3(nb), dim3(kThreads), 0, s, w, V, ld, count, h, n,
                       partial, nb);
}

void update_scale(const double *w, const double *V, i64 ld, int count, const double *cw, i64 n,
                  double *v, double *beta_flag, hipStream_t s) {
    const i64 g = std::max<i64>(1, (n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_update_scale, dim3(static_cast<u32>(std::min<i64>(g, 1 << 30))), dim3(kThreads), 0, s,
                       w, V, ld, count, cw, n, v, beta_flag);
}

void dcgs_update(double *V, i64 ld, int count, const double *sc, const double *tg,
                 const double *beta, const double *w, i64 n, double *partial, int nb,
                 double *eta_out, hipStream_t s) {
    if (count < 0 || count >= kFusedBasis || n <= 0) return;
    switch ((count + kWaves - 1) / kWaves) {
#define MK_DCGS(P)                                                                               \
    case P:                                                                                      \
        hipLaunchKernelGGL(k_dcgs_update<P>, dim3(nb), dim3(kThreads), 0, s, V, ld, count, sc,   \
                           tg, beta, w, n, partial, nb, eta_out);                                \
        break;
        MK_DCGS(0) MK_DCGS(1) MK_DCGS(2) MK_DCGS(3) MK_DCGS(4) MK_DCGS(5) MK_DCGS(6) MK_DCGS(7)
        MK_DCGS(8) MK_DCGS(9) MK_DCGS(10) MK_DCGS(11) MK_DCGS(12) MK_DCGS(13) MK_DCGS(14)
        MK_DCGS(15) MK_DCGS(16)
#undef MK_DCGS
    }
}

template <typename ZT>
void combine(double *x, const ZT *Z, i64 ld, int count, const double *y, i64 n, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_combine<ZT>), dim3(grid_for(n)), dim3(kThreads), 0, s, x, Z, ld, count,
                       y, n);
}

// explicit instantiations
template void shifted_apply<double>(const csr_view<double> &, double, const double *,
                                    const double *, double *, double *, hipStream_t);
template void shifted_apply<float>(const csr_view<double> &, double, const float *,
                                   const double *, double *, double *, hipStream_t);
#define MK_LEVEL(T)                                                                             \
    template void jacobi_sweep<T>(const csr_view<T> &, double, double, const T *, const T *,    \
                                  T *, hipStream_t);                                            \
    template void level_residual<T>(const csr_view<T> &, double, const T *, const T *, T *,     \
                                    hipStream_t);                                               \
    template void restrict_children<T>(i64, const u32 *, const u32 *, const T *, T *,           \
                                       hipStream_t);                                            \
    template void gather_sum<T>(const csr_view<T> &, const T *, T *, hipStream_t);              \
    template void gather_sum_first<T>(const csr_view<T> &, const T *, T *, const csr_view<T> &,  \
                                      const std::uint8_t *, double, T *, hipStream_t);          \
    template void level_residual_first<T>(const csr_view<T> &, double, const T *, const T *, T *, \
                                          const csr_view<T> &, const std::uint8_t *, T *,       \
                                          hipStream_t);                                         \
    template void prolong_add<T>(i64, const u32 *, const T *, const T *, T *, hipStream_t,       \
                                 const u32 *, const T *);                                         \
    template void terminal_solve<T>(i64, const u32 *, const u32 *, const u64 *, const T *, i64, \
                                    const T *, T *, hipStream_t, bool);                         \
    template void terminal_prolong<T>(i64, const u32 *, const u32 *, const u64 *, const T *, i64, \
                                      const T *, T *, const csr_view<T> &, const u8 *, const T *, \
                                      T *, hipStream_t);                                         \
    template void fill<T>(i64, T, T *, hipStream_t);                                            \
    template void block_solve<T>(const u32 *, const u32 *, const u32 *, const u64 *, const T *,   \
                                 u64, u64, const T *, T *, T *, hipStream_t);                     \
    template void block_residual<T>(const csr_view<T> &, double, const u32 *, u64, u64,         \
                                    const T *, T *, hipStream_t);                                 \
    template void block_zero<T>(const u32 *, u64, u64, T *, hipStream_t);                       \
    template void sweep_rows<T>(const csr_view<T> &, double, double, const T *, const T *, T *,   \
                                const u32 *, u64, hipStream_t);                                   \
    template void copy_rows<T>(const u32 *, u64, const T *, T *, hipStream_t);                  \
    template void vcycle_tail<T>(const level_desc<T> *, int, double, double, const T *,         \
                                 const double *, const double *, T *, hipStream_t, int);         \
    template bool vcycle_tail_lds_fits<T>(u32, int);                                              \
    template void gmres_tiny<T>(const tiny_gmres_args &, hipStream_t);                           \
    template void vcycle_tail_lds<T>(const void *, u32, u32, int,      \
                                     double, double, const T *, const double *, const double *,  \
                                     T *, T *, hipStream_t, int);                                  \
    template void axpy<T>(i64, T, const T *, T *, hipStream_t);                                 \
    template void combine<T>(double *, const T *, i64, int, const double *, i64, hipStream_t);     \
    template void gs_half<T>(const csr_view<T> &, double, const T *, T *, T *, hipStream_t, T *,   \
                             const u32 *);                                                        \
    template void permute_out<T>(i64, const u32 *, const T *, T *, hipStream_t);                 \
    template void permute_in<T, T>(i64, const u32 *, const T *, T *, hipStream_t, const double *); \
    template void ilu_forward<T>(const csr_view<T> &, const u32 *, const u32 *, int, int, i64,    \
                                 const T *, const T *, T *, hipStream_t);                         \
    template void ilu_backward<T>(const csr_view<T> &, const u32 *, const u32 *, int, int, i64,   \
                                  const T *, T *, T *, hipStream_t);
MK_LEVEL(float)
MK_LEVEL(double)
#undef MK_LEVEL
template void jacobi_first<double, double>(const csr_view<double> &, double, double,
                                           const double *, double *, double *, hipStream_t,
                                           const double *);
template void jacobi_first<float, double>(const csr_view<float> &, double, double,
                                          const double *, float *, float *, hipStream_t,
                                          const double *);
template void jacobi_first<float, float>(const csr_view<float> &, double, double, const float *,
                                         float *, float *, hipStream_t, const double *);
template void rb_first<double, double>(const csr_view<double> &, const u8 *, double, const double *,
                                       double *, double *, hipStream_t, const double *, const u32 *);
template void rb_first_scatter<double, double>(i64, const double *, double, const double *, double *,
                                               double *, hipStream_t, const double *, const u32 *);
template void rb_first_scatter<float, double>(i64, const float *, double, const double *, float *, float *,
                                              hipStream_t, const double *, const u32 *);
template void rb_first<float, double>(const csr_view<float> &, const u8 *, double, const double *,
                                      float *, float *, hipStream_t, const double *, const u32 *);
template void rb_first<float, float>(const csr_view<float> &, const u8 *, double, const float *,
                                     float *, float *, hipStream_t, const double *, const u32 *);
template void permute_in<float, double>(i64, const u32 *, const double *, float *, hipStream_t,
                                        const double *);
template void convert<double, float>(i64, const double *, float *, hipStream_t, const double *);
template void convert<float, double>(i64, const float *, double *, hipStream_t, const double *);
template void convert<double, double>(i64, const double *, double *, hipStream_t, const double *);
template void convert<float, float>(i64, const float *, float *, hipStream_t, const double *);

} // namespace mk::gpu
