// Device-side assembly of the RDME generator over a population_box: one thread per state
// builds its gather-form row (the jumps into it) and its exit and leak rates, then the rows
// are encoded in the coded sliced layout in place. The result equals mk::assemble_generator
// (setup.cpp) bit for bit: the same enumeration and rank (population_space.hpp:34-100), the
// same transitions and rates (reaction_diffusion.hpp:82-118; generator.hpp:27-44), the same
// summation orders, and no contraction of products into sums.
//
// A state is held as the sorted voxel positions of its molecules, species by species, which
// is at most kMaxMolecules entries where a count vector would be voxels x species.
#include "runtime.hpp"

#pragma clang fp contract(off)

namespace mk::rt {

namespace {

using gpu::u8;

constexpr int kMaxMolecules = 32;
constexpr int kMaxSpecies = 8;
constexpr int kHashSlots = 4096;
constexpr u64 kEmpty = ~u64{0};
constexpr int kThreadsA = 256;

struct box_view {
    int V = 0, S = 0, K = 0;
    i64 n = 0;
    u32 cstride = 0;                 // compositions table: comp(rest, rem) = comp[rest * cstride + rem]
    const u64 *comp = nullptr;
    const i64 *sbegin = nullptr;     // K + 1
    const int *stotals = nullptr;    // K x S
    const u64 *ssizes = nullptr;     // K x S
    int caps[kMaxSpecies] = {};
    // voxel graph
    i64 E = 0;
    const i64 *in_off = nullptr, *in_edge = nullptr, *edge_from = nullptr, *out_off = nullptr;
    const double *factor = nullptr;  // S x E: D_s w_e / vol_from
    // local reactions
    int R = 0;
    const int *changes = nullptr;    // R x S
    i64 rstride[kMaxSpecies] = {};
    i64 T = 1;
    const double *rates = nullptr;   // (r * V + v) * T + tuple
    int absorbing = 0;
};

struct state {
    int tot[kMaxSpecies];
    int off[kMaxSpecies + 1];
    unsigned short pos[kMaxMolecules]; // species s: pos[off[s] .. off[s + 1]), ascending
    u64 digit[kMaxSpecies];            // rank of each species' positions
};

__device__ __forceinline__ u64 comp(const box_view &b, i64 rest, int rem) {
    return b.comp[static_cast<u64>(rest) * b.cstride + static_cast<u32>(rem)];
}

// population_space::state_in_sector
__device__ void unrank(const box_view &b, int k, u64 local, state &st) {
    const u64 *sizes = b.ssizes + static_cast<i64>(k) * b.S;
    u64 r = local;
    for (int s = b.S; s-- > 0;) {
        st.digit[s] = r % sizes[s];
        r /= sizes[s];
    }
    int idx = 0;
    for (int s = 0; s < b.S; ++s) {
        st.off[s] = idx;
        st.tot[s] = b.stotals[static_cast<i64>(k) * b.S + s];
        u64 L = st.digit[s];
        int rem = st.tot[s];
        for (int v = 0; v + 1 < b.V && rem > 0; ++v) {
            const i64 rest = b.V - v - 1;
            int c = 0;
            while (true) {
                const u64 count = comp(b, rest - 1, rem - c);
                if (L < count) break;
                L -= count;
                ++c;
            }
            for (int q = 0; q < c; ++q) st.pos[idx++] = static_cast<unsigned short>(v);
            rem -= c;
        }
        for (int q = 0; q < rem; ++q) st.pos[idx++] = static_cast<unsigned short>(b.V - 1);
    }
    st.off[b.S] = idx;
}

// rank of one species' sorted positions among compositions of tot into V voxels
// (population_space::rank_in_sector, per species)
__device__ u64 rank_positions(const box_view &b, const unsigned short *p, int tot) {
    u64 local = 0;
    int rem = tot;
    for (int q = 0; q < tot;) {
        const int v = p[q];
        int cnt = 1;
        while (q + cnt < tot && p[q + cnt] == v) ++cnt;
        if (v + 1 < b.V) {
            const i64 rest = b.V - v - 1;
            local += comp(b, rest, rem) - comp(b, rest, rem - cnt);
        }
        rem -= cnt;
        q += cnt;
    }
    return local;
}

__device__ __forceinline__ int count_at(const state &st, int s, int v) {
    int c = 0;
    for (int q = st.off[s]; q < st.off[s + 1]; ++q) c += st.pos[q] == v;
    return c;
}

__device__ __forceinline__ i64 sector_of_totals(const box_view &b, const int *tot) {
    i64 k = 0;
    for (int s = 0; s < b.S; ++s) {
        if (tot[s] < 0 || tot[s] > b.caps[s]) return -1;
        k = k * (b.caps[s] + 1) + tot[s];
    }
    return k;
}

__device__ __forceinline__ u64 mixed(const box_view &b, int k, const u64 *digit) {
    const u64 *sizes = b.ssizes + static_cast<i64>(k) * b.S;
    u64 r = 0;
    for (int s = 0; s < b.S; ++s) r = r * sizes[s] + digit[s];
    return r;
}

// species s's positions with one molecule moved from voxel `from` to voxel `to` (or, with
// from < 0, added at `to`; with to < 0, removed from `from`), kept ascending, into out;
// returns the new count
__device__ int move_one(const unsigned short *p, int n, int from, int to, unsigned short *out) {
    int m = 0;
    bool removed = from < 0, inserted = to < 0;
    for (int q = 0; q < n; ++q) {
        if (!removed && p[q] == from) {
            removed = true;
            continue;
        }
        if (!inserted && p[q] >= to) {
            out[m++] = static_cast<unsigned short>(to);
            inserted = true;
        }
        out[m++] = p[q];
    }
    if (!inserted) out[m++] = static_cast<unsigned short>(to);
    return m;
}

__device__ __forceinline__ int find_sector(const box_view &b, i64 i) {
    int lo = 0, hi = b.K; // sbegin[lo] <= i < sbegin[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (b.sbegin[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Row i's jumps in the host's generation order (diffusion: target voxel b ascending, species,
// in-edges; then reactions: voxel, reaction). FILL = false counts them; FILL = true writes
// them at col/val, sorts the row by column (stable), merges equal columns (sums in that
// order) and writes the merged length, exit, leak and the in-box jump count.
template <bool FILL>
__global__ void __launch_bounds__(kThreadsA)
    k_rows(box_view b, const u64 *raw_ptr, u32 *raw_len, u32 *col, double *val, u32 *len,
           double *exit_out, double *leak_out, u32 *jumps_out, u64 *norm_bits) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; i < b.n;
         i += static_cast<i64>(gridDim.x) * kThreadsA) {
        const int k = find_sector(b, i);
        state st;
        unrank(b, k, static_cast<u64>(i - b.sbegin[k]), st);
        u64 m = 0;
        const u64 base = FILL ? raw_ptr[i] : 0;
        auto emit = [&](u64 c, double r) {
            if constexpr (FILL) {
                col[base + m] = static_cast<u32>(c);
                val[base + m] = r;
            }
            ++m;
        };
        unsigned short tmp[kMaxMolecules];
        u64 dg[kMaxSpecies];
        // diffusion into i: a molecule of species s hopped a -> b
        for (int bv = 0; bv < b.V; ++bv) {
            for (int s = 0; s < b.S; ++s) {
                const int nb = count_at(st, s, bv);
                if (nb == 0) continue;
                for (i64 q = b.in_off[bv]; q < b.in_off[bv + 1]; ++q) {
                    const i64 e = b.in_edge[q];
                    const int a = static_cast<int>(b.edge_from[e]);
                    const int cnt = move_one(st.pos + st.off[s], st.tot[s], bv, a, tmp);
                    for (int t = 0; t < b.S; ++t) dg[t] = st.digit[t];
                    dg[s] = rank_positions(b, tmp, cnt);
                    const int na = count_at(st, s, a) + (a == bv ? 0 : 1);
                    emit(static_cast<u64>(b.sbegin[k]) + mixed(b, k, dg),
                         static_cast<double>(na) * b.factor[static_cast<i64>(s) * b.E + e]);
                }
            }
        }
        // local reactions into i: the source held x - change in voxel v
        int totj[kMaxSpecies];
        for (int v = 0; v < b.V && b.R > 0; ++v) {
            for (int r = 0; r < b.R; ++r) {
                const int *ch = b.changes + static_cast<i64>(r) * b.S;
                bool ok = true;
                i64 tuple = 0;
                for (int s = 0; s < b.S && ok; ++s) {
                    const int c = count_at(st, s, v) - ch[s];
                    ok = c >= 0;
                    totj[s] = st.tot[s] - ch[s];
                    tuple += static_cast<i64>(c) * b.rstride[s];
                }
                if (!ok) continue;
                const i64 kj = sector_of_totals(b, totj);
                if (kj < 0) continue;
                const double rate = b.rates[(static_cast<i64>(r) * b.V + v) * b.T + tuple];
                if (!(rate > 0.0)) continue;
                for (int s = 0; s < b.S; ++s) {
                    if (ch[s] == 0) {
                        dg[s] = st.digit[s]; // a species' rank depends on its own total only
                        continue;
                    }
                    int cnt = st.tot[s];
                    unsigned short w[kMaxMolecules];
                    for (int q = 0; q < cnt; ++q) w[q] = st.pos[st.off[s] + q];
                    for (int d = 0; d < ch[s]; ++d) { // remove ch[s] molecules at v
                        cnt = move_one(w, cnt, v, -1, tmp);
                        for (int q = 0; q < cnt; ++q) w[q] = tmp[q];
                    }
                    for (int d = 0; d < -ch[s]; ++d) { // add -ch[s] molecules at v
                        cnt = move_one(w, cnt, -1, v, tmp);
                        for (int q = 0; q < cnt; ++q) w[q] = tmp[q];
                    }
                    dg[s] = rank_positions(b, w, cnt);
                }
                emit(static_cast<u64>(b.sbegin[kj]) + mixed(b, static_cast<int>(kj), dg), rate);
            }
        }
        if constexpr (!FILL) {
            raw_len[i] = static_cast<u32>(m);
        } else {
            // stable insertion sort by column, then merge equal columns in place
            for (u64 p = 1; p < m; ++p) {
                const u32 c = col[base + p];
                const double r = val[base + p];
                u64 q = p;
                while (q > 0 && col[base + q - 1] > c) {
                    col[base + q] = col[base + q - 1];
                    val[base + q] = val[base + q - 1];
                    --q;
                }
                col[base + q] = c;
                val[base + q] = r;
            }
            u64 out = 0;
            for (u64 p = 0; p < m; ++p) {
                if (out > 0 && col[base + out - 1] == col[base + p]) {
                    val[base + out - 1] += val[base + p];
                } else {
                    col[base + out] = col[base + p];
                    val[base + out] = val[base + p];
                    ++out;
                }
            }
            len[i] = static_cast<u32>(out);
            // row i as a source: exit and leak in the order of reaction_diffusion::transitions;
            // `moved` is its in-box out-rate (for the 1-norm)
            double ex = 0.0, lk = 0.0, moved = 0.0;
            u32 jumps = 0;
            for (int v = 0; v < b.V; ++v)
                for (i64 e = b.out_off[v]; e < b.out_off[v + 1]; ++e)
                    for (int s = 0; s < b.S; ++s) {
                        const int cnt = count_at(st, s, v);
                        if (cnt == 0) continue;
                        const double rt = cnt * b.factor[static_cast<i64>(s) * b.E + e];
                        ex += rt;
                        moved += rt;
                        ++jumps;
                    }
            for (int v = 0; v < b.V && b.R > 0; ++v) {
                i64 tuple = 0;
                for (int s = 0; s < b.S; ++s) tuple += static_cast<i64>(count_at(st, s, v)) * b.rstride[s];
                for (int r = 0; r < b.R; ++r) {
                    const double rate = b.rates[(static_cast<i64>(r) * b.V + v) * b.T + tuple];
                    if (!(rate > 0.0)) continue;
                    const int *ch = b.changes + static_cast<i64>(r) * b.S;
                    bool valid = true;
                    for (int s = 0; s < b.S; ++s) {
                        valid = valid && count_at(st, s, v) + ch[s] >= 0;
                        totj[s] = st.tot[s] + ch[s];
                    }
                    if (!valid) continue;
                    if (sector_of_totals(b, totj) >= 0) {
                        ex += rate;
                        moved += rate;
                        ++jumps;
                    } else {
                        lk += rate;
                    }
                }
            }
            if (b.absorbing) ex += lk;
            exit_out[i] = ex;
            leak_out[i] = lk;
            jumps_out[i] = jumps;
            // ||Q||_1 candidate of column i: exit plus the in-box rates out of i (non-negative,
            // so the bit patterns order like the values)
            const double cn = ex + moved;
            atomicMax(reinterpret_cast<unsigned long long *>(norm_bits),
                      static_cast<unsigned long long>(__double_as_longlong(cn)));
        }
    }
}

// distinct values into an open-addressing set of kHashSlots bit patterns; sets *overflow when
// a probe runs through the whole table
__device__ void hash_insert(u64 *table, u64 key, int *overflow) {
    u32 h = static_cast<u32>((key * 0x9E3779B97F4A7C15ull) >> 52) & (kHashSlots - 1);
    for (int probe = 0; probe < kHashSlots; ++probe) {
        const u64 cur = __atomic_load_n(table + h, __ATOMIC_RELAXED);
        if (cur == key) return;
        if (cur == kEmpty) {
            const u64 prev = atomicCAS(reinterpret_cast<unsigned long long *>(table + h),
                                       static_cast<unsigned long long>(kEmpty),
                                       static_cast<unsigned long long>(key));
            if (prev == kEmpty || prev == key) return;
        }
        h = (h + 1) & (kHashSlots - 1);
    }
    atomicExch(overflow, 1);
}

__global__ void __launch_bounds__(kThreadsA)
    k_distinct(i64 n, const u64 *raw_ptr, const u32 *len, const double *val, const double *exit,
               u64 *vtable, u64 *etable, int *overflow) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreadsA) {
        hash_insert(etable, static_cast<u64>(__double_as_longlong(exit[i])), overflow + 1);
        const u64 b = raw_ptr[i];
        for (u32 p = 0; p < len[i]; ++p)
            hash_insert(vtable, static_cast<u64>(__double_as_longlong(val[b + p])), overflow);
    }
}

// widest row of every 64-row slice, and the longest row and largest column overall
__global__ void __launch_bounds__(kThreadsA)
    k_widths(i64 n, const u64 *raw_ptr, const u32 *len, const u32 *col, u32 *width, u32 *maxima) {
    const i64 n_slices = (n + gpu::kSlice - 1) / gpu::kSlice;
    for (i64 sl = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; sl < n_slices;
         sl += static_cast<i64>(gridDim.x) * kThreadsA) {
        u32 w = 0, mc = 0;
        const i64 end = (sl + 1) * gpu::kSlice < n ? (sl + 1) * gpu::kSlice : n;
        for (i64 i = sl * gpu::kSlice; i < end; ++i) {
            w = max(w, len[i]);
            for (u32 p = 0; p < len[i]; ++p) mc = max(mc, col[raw_ptr[i] + p]);
        }
        width[sl] = w;
        atomicMax(maxima, w);
        atomicMax(maxima + 1, mc);
    }
}

__device__ __forceinline__ u32 code_of(const double *dict, int nd, double v) {
    int lo = 0, hi = nd; // first entry >= v
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (dict[mid] < v) lo = mid + 1; else hi = mid;
    }
    return static_cast<u32>(lo);
}

__global__ void __launch_bounds__(kThreadsA)
    k_encode(i64 n, const u64 *raw_ptr, const u32 *len, const u32 *col, const double *val,
             const double *exit, const double *vdict, int nv, const double *edict, int ne,
             const u32 *soff, u32 *ent, u8 *rlen, u8 *rcode) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreadsA) {
        const i64 sl = i / gpu::kSlice, lane = i % gpu::kSlice;
        const u64 b = raw_ptr[i];
        const u32 l = len[i];
        for (u32 p = 0; p < l; ++p)
            ent[soff[sl] + static_cast<u64>(p) * gpu::kSlice + lane] =
                col[b + p] | (code_of(vdict, nv, val[b + p]) << gpu::kColBits);
        rlen[i] = static_cast<u8>(l);
        rcode[i] = static_cast<u8>(code_of(edict, ne, exit[i]));
    }
}

int grid_of(i64 n) { return static_cast<int>(std::max<i64>(1, std::min<i64>((n + kThreadsA - 1) / kThreadsA, 16384))); }

template <typename T>
std::vector<T> download(const dbuf &b, std::size_t count) {
    std::vector<T> v(count);
    if (count) HIPCHK(hipMemcpy(v.data(), b.p, count * sizeof(T), hipMemcpyDeviceToHost));
    return v;
}

std::vector<double> table_values(const std::vector<u64> &table) {
    std::vector<double> out;
    for (u64 key : table) {
        if (key == kEmpty) continue;
        double d;
        std::memcpy(&d, &key, sizeof d);
        out.push_back(d);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

} // namespace

namespace {

// the merged rows of a generator, on the device: row i holds len[i] entries at raw_ptr[i]
struct device_rows {
    i64 n = 0;
    u64 raw_total = 0;
    std::vector<u64> raw_ptr_h;
    dbuf raw_ptr, col, val, len, exit, leak, jumps, norm;
};

// passes 1 and 2 of the assembly; false for spaces outside the kernel's limits (a sector
// space, as the multigrid levels use, is allowed without reactions)
bool assemble_rows(const mk::population_space &space, const mk::voxel_domain &domain,
                   const mk::reaction_tables &reactions, bool absorbing, device_rows &out) {
    const int S = space.species();
    const i64 V = space.voxels();
    const i64 n = space.size();
    if ((!space.is_box() && reactions.n > 0) || S > kMaxSpecies || V >= 65536 || n >= (i64{1} << 32) ||
        n == 0)
        return false;
    int total_cap = 0;
    for (int c : space.caps()) total_cap += c;
    if (total_cap > kMaxMolecules) return false;
    if (domain.voxels != V || domain.species() != S)
        throw std::invalid_argument("generator: the domain does not match the state space");
    const hipStream_t s = stream();

    // tables: sectors, compositions, graph, reactions
    const i64 K = space.n_sectors();
    std::vector<i64> sbegin(K + 1);
    std::vector<int> stot(K * S);
    std::vector<u64> ssz(K * S);
    for (i64 k = 0; k < K; ++k) {
        sbegin[k] = space.begin(k);
        for (int q = 0; q < S; ++q) {
            stot[k * S + q] = space.totals(k)[q];
            ssz[k * S + q] = space.species_sizes(k)[q];
        }
    }
    sbegin[K] = space.size();
    const i64 E = domain.offsets.back();
    std::vector<double> factor(static_cast<std::size_t>(S) * E);
    std::vector<i64> edge_from(E), in_off(V + 1, 0), in_edge(E);
    for (i64 a = 0; a < V; ++a)
        for (i64 e = domain.offsets[a]; e < domain.offsets[a + 1]; ++e) {
            edge_from[e] = a;
            for (int q = 0; q < S; ++q)
                factor[static_cast<std::size_t>(q) * E + e] =
                    domain.diffusivities[q] * domain.weight[e] / domain.volumes[a];
        }
    for (i64 e = 0; e < E; ++e) ++in_off[domain.to[e] + 1];
    for (i64 b = 0; b < V; ++b) in_off[b + 1] += in_off[b];
    {
        std::vector<i64> fill(in_off.begin(), in_off.end() - 1);
        for (i64 e = 0; e < E; ++e) in_edge[fill[domain.to[e]]++] = e;
    }
    dbuf d_comp = upload(space.compositions().data()), d_sbegin = upload(sbegin), d_stot = upload(stot),
         d_ssz = upload(ssz), d_in_off = upload(in_off), d_in_edge = upload(in_edge),
         d_edge_from = upload(edge_from), d_out_off = upload(domain.offsets), d_factor = upload(factor),
         d_changes = upload(reactions.changes), d_rates = upload(reactions.rates);
    box_view b;
    b.V = static_cast<int>(V);
    b.S = S;
    b.K = static_cast<int>(K);
    b.n = n;
    b.cstride = static_cast<u32>(space.compositions().stride());
    b.comp = dp<u64>(d_comp);
    b.sbegin = dp<i64>(d_sbegin);
    b.stotals = dp<int>(d_stot);
    b.ssizes = dp<u64>(d_ssz);
    for (int q = 0; q < S; ++q) b.caps[q] = space.caps()[q];
    b.E = E;
    b.in_off = dp<i64>(d_in_off);
    b.in_edge = dp<i64>(d_in_edge);
    b.edge_from = dp<i64>(d_edge_from);
    b.out_off = dp<i64>(d_out_off);
    b.factor = dp<double>(d_factor);
    b.R = reactions.n;
    b.changes = dp<int>(d_changes);
    for (int q = 0; q < S && reactions.n > 0; ++q) b.rstride[q] = reactions.stride[q];
    b.T = reactions.T;
    b.rates = dp<double>(d_rates);
    b.absorbing = absorbing ? 1 : 0;

    // pass 1: raw jump counts; their running sum on the host
    dbuf d_raw_len(static_cast<std::size_t>(n) * sizeof(u32));
    hipLaunchKernelGGL(k_rows<false>, dim3(grid_of(n)), dim3(kThreadsA), 0, s, b, nullptr,
                       dp<u32>(d_raw_len), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    const auto raw_len = download<u32>(d_raw_len, n);
    d_raw_len = dbuf();
    std::vector<u64> raw_ptr(n + 1, 0);
    for (i64 i = 0; i < n; ++i) raw_ptr[i + 1] = raw_ptr[i] + raw_len[i];
    const u64 raw_total = raw_ptr[n];
    if (raw_total >= (u64{1} << 32)) throw std::overflow_error("operator has more than 2^32 entries");
    dbuf d_raw_ptr = upload(raw_ptr);
    dbuf d_col(std::max<u64>(raw_total, 1) * sizeof(u32)), d_val(std::max<u64>(raw_total, 1) * sizeof(double));
    dbuf d_len(n * sizeof(u32)), d_exit(n * sizeof(double)), d_leak(n * sizeof(double)),
        d_jumps(n * sizeof(u32)), d_norm(sizeof(u64));
    HIPCHK(hipMemsetAsync(d_norm.p, 0, sizeof(u64), s));
    // pass 2: rows, merged; exit, leak, jumps, 1-norm
    hipLaunchKernelGGL(k_rows<true>, dim3(grid_of(n)), dim3(kThreadsA), 0, s, b, dp<u64>(d_raw_ptr),
                       nullptr, dp<u32>(d_col), dp<double>(d_val), dp<u32>(d_len), dp<double>(d_exit),
                       dp<double>(d_leak), dp<u32>(d_jumps), dp<u64>(d_norm));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    out.n = n;
    out.raw_total = raw_total;
    out.raw_ptr_h = std::move(raw_ptr);
    out.raw_ptr = std::move(d_raw_ptr);
    out.col = std::move(d_col);
    out.val = std::move(d_val);
    out.len = std::move(d_len);
    out.exit = std::move(d_exit);
    out.leak = std::move(d_leak);
    out.jumps = std::move(d_jumps);
    out.norm = std::move(d_norm);
    return true;
}

__global__ void __launch_bounds__(kThreadsA)
    k_compact(i64 n, const u64 *raw_ptr, const u64 *ptr, const u32 *col, const double *val,
              u32 *col2, double *val2) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * kThreadsA) {
        const u64 a = raw_ptr[i], b = ptr[i], l = ptr[i + 1] - ptr[i];
        for (u64 p = 0; p < l; ++p) {
            col2[b + p] = col[a + p];
            val2[b + p] = val[a + p];
        }
    }
}

} // namespace

namespace {

// population_transfer of one sector level (multigrid.hpp:50-92), per fine state: the rank of its
// aggregate (counts summed over each coarse voxel's children) in the coarse sector, the log of
// its unnormalised multinomial weight, summed over occupied voxels in ascending order as the
// host does, and its red-black colour
struct transfer_view {
    const u64 *ccomp = nullptr; // coarse compositions table
    u32 ccstride = 0;
    int Vc = 0;
    const u64 *csizes = nullptr; // coarse sector's per-species sizes
    const int *parent = nullptr; // V
    const u8 *parity = nullptr;  // V: coordinate-sum parity
    const double *log_fact = nullptr, *log_children = nullptr;
};

__device__ u64 rank_coarse(const transfer_view &t, const unsigned short *p, int tot) {
    u64 local = 0;
    int rem = tot;
    for (int q = 0; q < tot;) {
        const int v = p[q];
        int cnt = 1;
        while (q + cnt < tot && p[q + cnt] == v) ++cnt;
        if (v + 1 < t.Vc) {
            const i64 rest = t.Vc - v - 1;
            local += t.ccomp[static_cast<u64>(rest) * t.ccstride + rem] -
                     t.ccomp[static_cast<u64>(rest) * t.ccstride + (rem - cnt)];
        }
        rem -= cnt;
        q += cnt;
    }
    return local;
}

// sum of term(v, s, n) over the occupied (v, s) of per-species sorted position lists, v
// ascending, then s ascending
template <typename F>
__device__ void visit_occupied(const unsigned short *pos, const int *off, int S, F &&term) {
    int head[kMaxSpecies];
    for (int s = 0; s < S; ++s) head[s] = off[s];
    while (true) {
        int best = -1, bv = 0;
        for (int s = 0; s < S; ++s)
            if (head[s] < off[s + 1] && (best < 0 || pos[head[s]] < bv)) {
                best = s;
                bv = pos[head[s]];
            }
        if (best < 0) return;
        int n = 0;
        while (head[best] < off[best + 1] && pos[head[best]] == bv) {
            ++head[best];
            ++n;
        }
        term(bv, best, n);
    }
}

__global__ void __launch_bounds__(kThreadsA)
    k_transfer(box_view b, transfer_view t, i64 *coarse_of, double *log_weight, u8 *colour) {
    for (i64 i = blockIdx.x * static_cast<i64>(kThreadsA) + threadIdx.x; i < b.n;
         i += static_cast<i64>(gridDim.x) * kThreadsA) {
        state st;
        unrank(b, 0, static_cast<u64>(i), st);
        double lw = 0.0;
        int par = 0;
        visit_occupied(st.pos, st.off, b.S, [&](int v, int, int n) {
            lw -= t.log_fact[n];
            par ^= (n & 1) & t.parity[v];
        });
        unsigned short cpos[kMaxMolecules];
        for (int s = 0; s < b.S; ++s) { // coarse positions, sorted per species
            int m = st.off[s];
            for (int q = st.off[s]; q < st.off[s + 1]; ++q) {
                const unsigned short a = static_cast<unsigned short>(t.parent[st.pos[q]]);
                int r = m++;
                while (r > st.off[s] && cpos[r - 1] > a) {
                    cpos[r] = cpos[r - 1];
                    --r;
                }
                cpos[r] = a;
            }
        }
        visit_occupied(cpos, st.off, b.S, [&](int a, int, int n) {
            lw += t.log_fact[n] - (n * t.log_children[a]);
        });
        u64 r = 0;
        for (int s = 0; s < b.S; ++s)
            r = r * t.csizes[s] + rank_coarse(t, cpos + st.off[s], st.tot[s]);
        coarse_of[i] = static_cast<i64>(r);
        log_weight[i] = lw;
        if (colour) colour[i] = static_cast<u8>(par);
    }
}

} // namespace

bool transfer_device(const mk::population_space &fine, const mk::population_space &coarse,
                     const mk::voxel_domain &dom, const std::vector<i64> &parent,
                     const std::vector<double> &log_fact, const std::vector<double> &log_children,
                     std::vector<i64> &coarse_of, std::vector<double> &log_weight,
                     std::vector<std::uint8_t> *colour) {
    const int S = fine.species();
    const i64 V = fine.voxels(), n = fine.size();
    int total = 0;
    for (int c : fine.totals(0)) total += c;
    if (fine.is_box() || fine.n_sectors() != 1 || S > kMaxSpecies || V >= 65536 || total > kMaxMolecules ||
        n >= (i64{1} << 32) || n < 4096 || dom.dim == 0)
        return false;
    const hipStream_t s = stream();
    std::vector<i64> sbegin{0, n};
    std::vector<int> stot(fine.totals(0).begin(), fine.totals(0).end());
    std::vector<u64> ssz(fine.species_sizes(0).begin(), fine.species_sizes(0).end());
    std::vector<int> par32(parent.begin(), parent.end());
    std::vector<u8> parity(V);
    for (i64 v = 0; v < V; ++v) {
        i64 c = 0;
        for (i64 a : dom.coordinate(v)) c += a;
        parity[v] = static_cast<u8>(c & 1);
    }
    dbuf d_comp = upload(fine.compositions().data()), d_sbegin = upload(sbegin), d_stot = upload(stot),
         d_ssz = upload(ssz), d_ccomp = upload(coarse.compositions().data()),
         d_csz = upload(coarse.species_sizes(0)), d_parent = upload(par32), d_parity = upload(parity),
         d_lf = upload(log_fact), d_lc = upload(log_children);
    box_view b;
    b.V = static_cast<int>(V);
    b.S = S;
    b.K = 1;
    b.n = n;
    b.cstride = static_cast<u32>(fine.compositions().stride());
    b.comp = dp<u64>(d_comp);
    b.sbegin = dp<i64>(d_sbegin);
    b.stotals = dp<int>(d_stot);
    b.ssizes = dp<u64>(d_ssz);
    transfer_view t;
    t.ccomp = dp<u64>(d_ccomp);
    t.ccstride = static_cast<u32>(coarse.compositions().stride());
    t.Vc = static_cast<int>(coarse.voxels());
    t.csizes = dp<u64>(d_csz);
    t.parent = dp<int>(d_parent);
    t.parity = dp<u8>(d_parity);
    t.log_fact = dp<double>(d_lf);
    t.log_children = dp<double>(d_lc);
    dbuf d_co(n * sizeof(i64)), d_lw(n * sizeof(double)), d_col(colour ? n : 1);
    hipLaunchKernelGGL(k_transfer, dim3(grid_of(n)), dim3(kThreadsA), 0, s, b, t, dp<i64>(d_co),
                       dp<double>(d_lw), colour ? dp<u8>(d_col) : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    coarse_of = download<i64>(d_co, n);
    log_weight = download<double>(d_lw, n);
    if (colour) *colour = download<u8>(d_col, n);
    return true;
}

bool assemble_csr_device(const mk::population_space &space, const mk::voxel_domain &domain,
                         const mk::reaction_tables &reactions, bool absorbing, mk::generator_csr &g) {
    device_rows r;
    if (!assemble_rows(space, domain, reactions, absorbing, r)) return false;
    const hipStream_t s = stream();
    const i64 n = r.n;
    const auto len = download<u32>(r.len, n);
    g = mk::generator_csr();
    g.n = n;
    g.ptr.assign(n + 1, 0);
    for (i64 i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + len[i];
    const u64 nnz = g.ptr[n];
    if (nnz == r.raw_total) { // no merged entries: the raw arrays are the rows
        g.col = download<u32>(r.col, nnz);
        g.val = download<double>(r.val, nnz);
    } else {
        dbuf d_ptr = upload(g.ptr), c2(std::max<u64>(nnz, 1) * sizeof(u32)), v2(std::max<u64>(nnz, 1) * sizeof(double));
        hipLaunchKernelGGL(k_compact, dim3(grid_of(n)), dim3(kThreadsA), 0, s, n, dp<u64>(r.raw_ptr),
                           dp<u64>(d_ptr), dp<u32>(r.col), dp<double>(r.val), dp<u32>(c2), dp<double>(v2));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        g.col = download<u32>(c2, nnz);
        g.val = download<double>(v2, nnz);
    }
    g.exit = download<double>(r.exit, n);
    g.leak = download<double>(r.leak, n);
    const auto jumps = download<u32>(r.jumps, n);
    for (u32 j : jumps) g.n_jumps += j;
    return true;
}

bool assemble_generator_device(const mk::population_space &space, const mk::voxel_domain &domain,
                               const mk::reaction_tables &reactions, bool absorbing,
                               mk_generator_s &G) {
    device_rows r;
    if (!space.is_box() || !assemble_rows(space, domain, reactions, absorbing, r)) return false;
    const hipStream_t s = stream();
    const i64 n = r.n;
    const u64 raw_total = r.raw_total;
    const auto &raw_ptr = r.raw_ptr_h;
    dbuf &d_raw_ptr = r.raw_ptr, &d_col = r.col, &d_val = r.val, &d_len = r.len, &d_exit = r.exit,
         &d_leak = r.leak, &d_jumps = r.jumps, &d_norm = r.norm;

    // distinct rates and exit rates, slice widths, longest row and largest column
    dbuf d_vt(kHashSlots * sizeof(u64)), d_et(kHashSlots * sizeof(u64)), d_over(2 * sizeof(int));
    HIPCHK(hipMemsetAsync(d_vt.p, 0xFF, d_vt.bytes, s));
    HIPCHK(hipMemsetAsync(d_et.p, 0xFF, d_et.bytes, s));
    HIPCHK(hipMemsetAsync(d_over.p, 0, d_over.bytes, s));
    hipLaunchKernelGGL(k_distinct, dim3(grid_of(n)), dim3(kThreadsA), 0, s, n, dp<u64>(d_raw_ptr),
                       dp<u32>(d_len), dp<double>(d_val), dp<double>(d_exit), dp<u64>(d_vt),
                       dp<u64>(d_et), dp<int>(d_over));
    HIPCHK(hipGetLastError());
    const i64 n_slices = (n + gpu::kSlice - 1) / gpu::kSlice;
    dbuf d_width(n_slices * sizeof(u32)), d_max(2 * sizeof(u32));
    HIPCHK(hipMemsetAsync(d_max.p, 0, d_max.bytes, s));
    hipLaunchKernelGGL(k_widths, dim3(grid_of(n_slices)), dim3(kThreadsA), 0, s, n, dp<u64>(d_raw_ptr),
                       dp<u32>(d_len), dp<u32>(d_col), dp<u32>(d_width), dp<u32>(d_max));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    const auto over = download<int>(d_over, 2);
    const auto maxima = download<u32>(d_max, 2);
    const auto jumps = download<u32>(d_jumps, n);
    G.n = n;
    G.n_jumps = 0;
    for (u32 j : jumps) G.n_jumps += j;
    G.exit = download<double>(d_exit, n);
    G.leak = download<double>(d_leak, n);
    {
        const auto nb = download<u64>(d_norm, 1);
        std::memcpy(&G.norm1, nb.data(), sizeof(double));
    }
    auto vdict = over[0] ? std::vector<double>{} : table_values(download<u64>(d_vt, kHashSlots));
    auto edict = over[1] ? std::vector<double>{} : table_values(download<u64>(d_et, kHashSlots));
    if (raw_total == 0 && vdict.empty()) vdict.push_back(0.0); // the host's dictionary of nothing
    const auto width = download<u32>(d_width, n_slices);
    std::vector<u32> soff(n_slices + 1, 0);
    u64 total = 0;
    bool fits = true;
    for (i64 sl = 0; sl < n_slices && fits; ++sl) {
        total += static_cast<u64>(width[sl]) * gpu::kSlice;
        fits = total < (u64{1} << 32);
        soff[sl + 1] = static_cast<u32>(total);
    }
    const bool coded = fits && maxima[0] <= 255 && (raw_total == 0 || maxima[1] < (1u << gpu::kColBits)) &&
                       !vdict.empty() && vdict.size() <= 256 && !edict.empty() && edict.size() <= 256;
    if (!coded) {
        // outside the coded layout: the host's plain encoding of the same rows
        const auto len = download<u32>(d_len, n);
        const auto col = download<u32>(d_col, raw_total);
        const auto val = download<double>(d_val, raw_total);
        mk::generator_csr g;
        g.n = n;
        g.n_jumps = G.n_jumps;
        g.ptr.assign(n + 1, 0);
        for (i64 i = 0; i < n; ++i) {
            g.ptr[i + 1] = g.ptr[i] + len[i];
            g.col.insert(g.col.end(), col.begin() + raw_ptr[i], col.begin() + raw_ptr[i] + len[i]);
            g.val.insert(g.val.end(), val.begin() + raw_ptr[i], val.begin() + raw_ptr[i] + len[i]);
        }
        g.exit = G.exit;
        g.leak = G.leak;
        G.A = upload_csr<double>(g);
        return true;
    }
    // encode in place on the device
    dev_csr<double> &d = G.A;
    d = dev_csr<double>();
    d.n = n;
    d.nnz = 0;
    {
        const auto len = download<u32>(d_len, n);
        for (u32 l : len) d.nnz += l;
    }
    d.coded = true;
    d.padded = total;
    d.soff = upload(soff);
    d.ent = dbuf(std::max<u64>(total, 1) * sizeof(u32));
    HIPCHK(hipMemsetAsync(d.ent.p, 0, d.ent.bytes, s));
    d.rlen = dbuf(n);
    d.rcode = dbuf(n);
    d.vdict = upload(vdict);
    d.edict = upload(edict);
    hipLaunchKernelGGL(k_encode, dim3(grid_of(n)), dim3(kThreadsA), 0, s, n, dp<u64>(d_raw_ptr),
                       dp<u32>(d_len), dp<u32>(d_col), dp<double>(d_val), dp<double>(d_exit),
                       dp<double>(d.vdict), static_cast<int>(vdict.size()), dp<double>(d.edict),
                       static_cast<int>(edict.size()), dp<u32>(d.soff), dp<u32>(d.ent),
                       dp<u8>(d.rlen), dp<u8>(d.rcode));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    d.view.n = n;
    d.view.n_slices = n_slices;
    d.view.n_blocks = (n_slices + (gpu::kThreads / gpu::kSlice) - 1) / (gpu::kThreads / gpu::kSlice);
    d.view.soff = dp<u32>(d.soff);
    d.view.ent = dp<u32>(d.ent);
    d.view.rlen = dp<u8>(d.rlen);
    d.view.rcode = dp<u8>(d.rcode);
    d.view.vdict = dp<double>(d.vdict);
    d.view.edict = dp<double>(d.edict);
    d.view.n_vdict = static_cast<int>(vdict.size());
    d.view.n_edict = static_cast<int>(edict.size());
    d.view.max_width = static_cast<int>(maxima[0]);
    d.view.coded = 1;
    check_view(d.view, d.nnz);
    return true;
}

} // namespace mk::rt
