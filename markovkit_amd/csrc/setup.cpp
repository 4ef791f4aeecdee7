// Host setup: population spaces, gather-form generator assembly, multigrid hierarchy.
// Reference semantics are cited in setup.hpp; the loops here are written for throughput
// (per-row work on a pool of threads, tabulated compositions) and produce the same operator
// as the reference's per-source jump lists, stored by destination.
#include "setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <thread>

namespace mk {

unsigned worker_threads() {
    static const unsigned n = [] {
        unsigned hw = std::thread::hardware_concurrency();
        if (hw == 0) hw = 1;
        if (const char *env = std::getenv("OMP_NUM_THREADS")) {
            const long v = std::strtol(env, nullptr, 10);
            if (v > 0) hw = std::min<unsigned>(hw, static_cast<unsigned>(v));
        }
        return std::min(hw, 64U);
    }();
    return n;
}

namespace {

// run task(t) for t in [0, n_tasks) on the worker pool, dynamically scheduled
template <typename F>
void run_tasks(i64 n_tasks, F &&task) {
    const unsigned threads = static_cast<unsigned>(std::min<i64>(worker_threads(), n_tasks));
    if (threads <= 1) {
        for (i64 t = 0; t < n_tasks; ++t) task(t);
        return;
    }
    std::atomic<i64> next{0};
    std::vector<std::thread> pool;
    std::exception_ptr failure;
    std::atomic<bool> failed{false};
    for (unsigned w = 0; w < threads; ++w) {
        pool.emplace_back([&] {
            try {
                for (i64 t; (t = next.fetch_add(1)) < n_tasks && !failed.load();) task(t);
            } catch (...) {
                if (!failed.exchange(true)) failure = std::current_exception();
            }
        });
    }
    for (auto &th : pool) th.join();
    if (failure) std::rethrow_exception(failure);
}

} // namespace

void parallel_tasks(i64 n_tasks, const std::function<void(i64)> &task) { run_tasks(n_tasks, task); }

namespace {

u64 sat_add(u64 a, u64 b) { return a > std::numeric_limits<u64>::max() - b ? std::numeric_limits<u64>::max() : a + b; }

} // namespace

// ------------------------------------------------------------------ compositions table
compositions_table::compositions_table(i64 voxels, int max_total)
    : table_(static_cast<std::size_t>(voxels) * (max_total + 1)), stride_(max_total + 1) {
    // comp(0, rem) = 1; comp(rest, rem) = sum_{c <= rem} comp(rest - 1, c)
    for (int rem = 0; rem <= max_total; ++rem) table_[rem] = 1;
    for (i64 rest = 1; rest < voxels; ++rest) {
        u64 run = 0;
        for (int rem = 0; rem <= max_total; ++rem) {
            run = sat_add(run, (*this)(rest - 1, rem));
            table_[static_cast<std::size_t>(rest) * stride_ + rem] = run;
        }
    }
}

// ------------------------------------------------------------------ population spaces
population_space population_space::box(i64 voxels, std::vector<int> caps) {
    if (voxels <= 0 || caps.empty())
        throw std::invalid_argument("population state space needs a voxel and a species");
    for (int c : caps)
        if (c < 0) throw std::invalid_argument("species totals must be nonnegative");
    population_space p;
    p.voxels_ = voxels;
    p.caps_ = std::move(caps);
    p.box_ = true;
    p.comp_ = compositions_table(voxels, *std::max_element(p.caps_.begin(), p.caps_.end()));
    // add_sectors (population_space.hpp:178-189): species 0 most significant
    std::vector<int> totals(p.caps_.size(), 0);
    const int S = p.species();
    while (true) {
        p.add_sector(totals);
        int s = S - 1;
        while (s >= 0 && totals[s] == p.caps_[s]) totals[s--] = 0;
        if (s < 0) break;
        ++totals[s];
    }
    return p;
}

population_space population_space::sector(i64 voxels, std::vector<int> totals) {
    if (voxels <= 0 || totals.empty())
        throw std::invalid_argument("population state space needs a voxel and a species");
    for (int c : totals)
        if (c < 0) throw std::invalid_argument("species totals must be nonnegative");
    population_space p;
    p.voxels_ = voxels;
    p.caps_ = totals;
    p.box_ = false;
    p.comp_ = compositions_table(voxels, *std::max_element(totals.begin(), totals.end()));
    p.add_sector(totals);
    return p;
}

void population_space::add_sector(const std::vector<int> &totals) {
    std::vector<u64> sizes;
    u64 size = 1;
    for (int t : totals) {
        const u64 f = comp_(voxels_ - 1, t);
        if (f == std::numeric_limits<u64>::max() || (f != 0 && size > std::numeric_limits<u64>::max() / f))
            throw std::overflow_error("population state space is too large");
        sizes.push_back(f);
        size *= f;
    }
    if (static_cast<u64>(begin_.back()) > static_cast<u64>(std::numeric_limits<i64>::max()) - size)
        throw std::overflow_error("population state space is too large");
    totals_.push_back(totals);
    species_sizes_.push_back(std::move(sizes));
    begin_.push_back(begin_.back() + static_cast<i64>(size));
}

i64 population_space::sector_of_totals(const int *totals) const {
    const int S = species();
    if (box_) {
        i64 k = 0;
        for (int s = 0; s < S; ++s) {
            if (totals[s] < 0 || totals[s] > caps_[s]) return -1;
            k = k * (caps_[s] + 1) + totals[s];
        }
        return k;
    }
    for (int s = 0; s < S; ++s)
        if (totals[s] != caps_[s]) return -1;
    return 0;
}

i64 population_space::sector_of_index(i64 i) const {
    return static_cast<i64>(std::upper_bound(begin_.begin(), begin_.end(), i) - begin_.begin()) - 1;
}

i64 population_space::rank_in_sector(const int *x, i64 k) const {
    const int S = species();
    const auto &totals = totals_[k];
    const auto &sizes = species_sizes_[k];
    u64 result = 0;
    for (int s = 0; s < S; ++s) {
        int remaining = totals[s];
        u64 local = 0;
        for (i64 v = 0; v + 1 < voxels_; ++v) {
            const int count = x[v * S + s];
            if (count < 0 || count > remaining) return -1;
            const i64 rest = voxels_ - v - 1;
            local += comp_(rest, remaining) - comp_(rest, remaining - count);
            remaining -= count;
        }
        if (x[(voxels_ - 1) * S + s] != remaining) return -1;
        result = result * sizes[s] + local;
    }
    return static_cast<i64>(result);
}

i64 population_space::find(const int *x) const {
    const int S = species();
    std::vector<int> totals(S, 0);
    for (i64 i = 0; i < width(); ++i) totals[i % S] += x[i];
    const i64 k = sector_of_totals(totals.data());
    if (k < 0) return -1;
    const i64 local = rank_in_sector(x, k);
    return local < 0 ? -1 : begin_[k] + local;
}

void population_space::state_in_sector(i64 k, i64 local, int *x) const {
    const int S = species();
    const auto &totals = totals_[k];
    const auto &sizes = species_sizes_[k];
    // split the sector rank into per-species ranks (species S-1 least significant)
    std::vector<u64> digit(S);
    u64 r = static_cast<u64>(local);
    for (int s = S; s-- > 0;) {
        digit[s] = r % sizes[s];
        r /= sizes[s];
    }
    for (int s = 0; s < S; ++s) {
        u64 L = digit[s];
        int rem = totals[s];
        for (i64 v = 0; v + 1 < voxels_; ++v) {
            const i64 rest = voxels_ - v - 1;
            int c = 0;
            // states with c molecules in voxel v: compositions of rem - c into `rest` voxels
            while (true) {
                const u64 count = comp_(rest - 1, rem - c);
                if (L < count) break;
                L -= count;
                ++c;
            }
            x[v * S + s] = c;
            rem -= c;
        }
        x[(voxels_ - 1) * S + s] = rem;
    }
}

void population_space::state(i64 i, int *x) const {
    const i64 k = sector_of_index(i);
    state_in_sector(k, i - begin_[k], x);
}

// ------------------------------------------------------------------ voxel domains
voxel_domain voxel_domain::grid(std::vector<i64> dims, std::vector<double> spacing,
                                const double *volumes, std::vector<double> diffusivities) {
    if (dims.empty() || dims.size() > 3 || spacing.size() != dims.size())
        throw std::invalid_argument("grid: 1 to 3 axes, one spacing per axis");
    voxel_domain d;
    d.dim = static_cast<int>(dims.size());
    d.dims = std::move(dims);
    d.spacing = std::move(spacing);
    d.voxels = 1;
    for (i64 n : d.dims) {
        if (n < 1) throw std::invalid_argument("grid: every axis needs a cell");
        d.voxels *= n;
    }
    d.diffusivities = std::move(diffusivities);
    d.offsets.push_back(0);
    for (i64 v = 0; v < d.voxels; ++v) {
        const auto c = d.coordinate(v);
        i64 stride = 1;
        for (int a = 0; a < d.dim; ++a) {
            double area = 1.0;
            for (int b = 0; b < d.dim; ++b)
                if (b != a) area *= d.spacing[b];
            const double w = area / d.spacing[a];
            if (c[a] > 0) {
                d.to.push_back(v - stride);
                d.weight.push_back(w);
            }
            if (c[a] + 1 < d.dims[a]) {
                d.to.push_back(v + stride);
                d.weight.push_back(w);
            }
            stride *= d.dims[a];
        }
        d.offsets.push_back(static_cast<i64>(d.to.size()));
    }
    if (volumes) {
        d.volumes.assign(volumes, volumes + d.voxels);
    } else {
        d.volumes.assign(d.voxels, d.cell_volume());
    }
    for (double v : d.volumes)
        if (!(v > 0.0) || !std::isfinite(v))
            throw std::invalid_argument("voxel_domain: volumes must be positive");
    return d;
}

voxel_domain voxel_domain::graph(i64 voxels, const i64 *offsets, const i64 *to,
                                 const double *weight, const double *volumes,
                                 std::vector<double> diffusivities) {
    if (voxels <= 0 || !offsets || !volumes)
        throw std::invalid_argument("voxel_domain: a graph needs voxels, edges and volumes");
    voxel_domain d;
    d.voxels = voxels;
    d.offsets.assign(offsets, offsets + voxels + 1);
    const i64 E = d.offsets.back();
    d.to.assign(to, to + E);
    d.weight.assign(weight, weight + E);
    for (i64 t : d.to)
        if (t < 0 || t >= voxels) throw std::invalid_argument("voxel graph: edge out of range");
    d.volumes.assign(volumes, volumes + voxels);
    for (double v : d.volumes)
        if (!(v > 0.0) || !std::isfinite(v))
            throw std::invalid_argument("voxel_domain: volumes must be positive");
    d.diffusivities = std::move(diffusivities);
    return d;
}

double voxel_domain::cell_volume() const {
    double v = 1.0;
    for (double h : spacing) v *= h;
    return v;
}

std::vector<i64> voxel_domain::coordinate(i64 v) const {
    std::vector<i64> c(dim);
    for (int a = 0; a < dim; ++a) {
        c[a] = v % dims[a];
        v /= dims[a];
    }
    return c;
}

i64 voxel_domain::vertex(const std::vector<i64> &c) const {
    i64 v = 0;
    for (int a = dim; a-- > 0;) v = v * dims[a] + c[a];
    return v;
}

std::optional<voxel_domain> coarsen(const voxel_domain &fine, std::vector<double> diffusivities) {
    if (fine.dim == 0) return std::nullopt;
    auto dims = fine.dims;
    auto h = fine.spacing;
    bool reduced = false;
    for (int a = 0; a < fine.dim; ++a) {
        if (dims[a] == 1) continue;
        if (dims[a] % 2 != 0) return std::nullopt;
        dims[a] /= 2;
        h[a] *= 2.0;
        reduced = true;
    }
    if (!reduced) return std::nullopt;
    return voxel_domain::grid(dims, h, nullptr, std::move(diffusivities));
}

// ------------------------------------------------------------------ generator assembly
namespace {

struct row_task {
    i64 sector;
    i64 first; // global row
    i64 last;
};

std::vector<row_task> split_rows(const population_space &space, i64 chunk) {
    std::vector<row_task> tasks;
    for (i64 k = 0; k < space.n_sectors(); ++k)
        for (i64 r = space.begin(k); r < space.end(k); r += chunk)
            tasks.push_back({k, r, std::min(space.end(k), r + chunk)});
    return tasks;
}

struct task_out {
    std::vector<u32> len;
    std::vector<u32> col;
    std::vector<double> val;
    i64 jumps = 0;
};

} // namespace

generator_csr assemble_generator(const population_space &space, const voxel_domain &domain,
                                 const reaction_tables &reactions, bool absorbing) {
    const i64 V = space.voxels();
    const int S = space.species();
    const i64 n = space.size();
    if (domain.voxels != V || domain.species() != S)
        throw std::invalid_argument("generator: the domain does not match the state space");
    if (n >= (i64{1} << 32)) throw std::overflow_error("generator: more than 2^32 states");

    // per species and out-edge: D_s * w_e / vol_from (reaction_diffusion.hpp:52-55)
    const i64 E = domain.offsets.back();
    std::vector<double> factor(static_cast<std::size_t>(S) * E);
    std::vector<i64> edge_from(E);
    for (i64 a = 0; a < V; ++a)
        for (i64 e = domain.offsets[a]; e < domain.offsets[a + 1]; ++e) {
            edge_from[e] = a;
            for (int s = 0; s < S; ++s)
                factor[static_cast<std::size_t>(s) * E + e] =
                    domain.diffusivities[s] * domain.weight[e] / domain.volumes[a];
        }
    // in-edges of every voxel: the out-edges (a -> b) grouped by b
    std::vector<i64> in_off(V + 1, 0), in_edge(E);
    for (i64 e = 0; e < E; ++e) ++in_off[domain.to[e] + 1];
    for (i64 b = 0; b < V; ++b) in_off[b + 1] += in_off[b];
    {
        std::vector<i64> fill(in_off.begin(), in_off.end() - 1);
        for (i64 e = 0; e < E; ++e) in_edge[fill[domain.to[e]]++] = e;
    }

    const int R = reactions.n;
    const auto tasks = split_rows(space, 16384);
    std::vector<task_out> outs(tasks.size());
    generator_csr g;
    g.n = n;
    g.exit.assign(n, 0.0);
    g.leak.assign(n, 0.0);

    run_tasks(static_cast<i64>(tasks.size()), [&](i64 t) {
        const row_task &task = tasks[t];
        task_out &out = outs[t];
        const auto &totals = space.totals(task.sector);
        const i64 W = V * S;
        std::vector<int> x(W), y(W), tot_j(S);
        std::vector<std::pair<u32, double>> entries;
        out.len.reserve(task.last - task.first);
        for (i64 i = task.first; i < task.last; ++i) {
            space.state_in_sector(task.sector, i - space.begin(task.sector), x.data());
            entries.clear();
            // diffusion into i: a molecule of species s hopped a -> b
            for (i64 b = 0; b < V; ++b) {
                for (int s = 0; s < S; ++s) {
                    if (x[b * S + s] == 0) continue;
                    for (i64 q = in_off[b]; q < in_off[b + 1]; ++q) {
                        const i64 e = in_edge[q];
                        const i64 a = edge_from[e];
                        std::copy(x.begin(), x.end(), y.begin());
                        --y[b * S + s];
                        ++y[a * S + s];
                        const i64 j = space.begin(task.sector) + space.rank_in_sector(y.data(), task.sector);
                        const int na = y[a * S + s];
                        entries.push_back({static_cast<u32>(j), na * factor[static_cast<std::size_t>(s) * E + e]});
                    }
                }
            }
            // local reactions into i: the source held x - change in voxel v
            for (i64 v = 0; v < V; ++v) {
                for (int r = 0; r < R; ++r) {
                    const int *ch = &reactions.changes[static_cast<std::size_t>(r) * S];
                    bool ok = true;
                    i64 tuple = 0;
                    for (int s = 0; s < S && ok; ++s) {
                        const int c = x[v * S + s] - ch[s];
                        ok = c >= 0;
                        tot_j[s] = totals[s] - ch[s];
                        tuple += static_cast<i64>(c) * reactions.stride[s];
                    }
                    if (!ok) continue;
                    const i64 kj = space.sector_of_totals(tot_j.data());
                    if (kj < 0) continue;
                    const double rate = reactions.rates[(static_cast<std::size_t>(r) * V + v) * reactions.T + tuple];
                    if (!(rate > 0.0)) continue;
                    std::copy(x.begin(), x.end(), y.begin());
                    for (int s = 0; s < S; ++s) y[v * S + s] -= ch[s];
                    const i64 j = space.begin(kj) + space.rank_in_sector(y.data(), kj);
                    entries.push_back({static_cast<u32>(j), rate});
                }
            }
            std::sort(entries.begin(), entries.end(),
                      [](const auto &p, const auto &q) { return p.first < q.first; });
            u32 len = 0;
            for (std::size_t e = 0; e < entries.size(); ++e) {
                if (len > 0 && out.col.back() == entries[e].first) {
                    out.val.back() += entries[e].second;
                } else {
                    out.col.push_back(entries[e].first);
                    out.val.push_back(entries[e].second);
                    ++len;
                }
            }
            out.len.push_back(len);

            // row i as a source: exit and leak in the order of reaction_diffusion::transitions
            double exit = 0.0, leak = 0.0;
            i64 jumps = 0;
            for (i64 v = 0; v < V; ++v)
                for (i64 e = domain.offsets[v]; e < domain.offsets[v + 1]; ++e)
                    for (int s = 0; s < S; ++s) {
                        const int cnt = x[v * S + s];
                        if (cnt == 0) continue;
                        exit += cnt * factor[static_cast<std::size_t>(s) * E + e];
                        ++jumps;
                    }
            for (i64 v = 0; v < V && R > 0; ++v) {
                i64 tuple = 0;
                for (int s = 0; s < S; ++s) tuple += static_cast<i64>(x[v * S + s]) * reactions.stride[s];
                for (int r = 0; r < R; ++r) {
                    const double rate = reactions.rates[(static_cast<std::size_t>(r) * V + v) * reactions.T + tuple];
                    if (!(rate > 0.0)) continue;
                    const int *ch = &reactions.changes[static_cast<std::size_t>(r) * S];
                    bool valid = true;
                    for (int s = 0; s < S; ++s) {
                        valid = valid && x[v * S + s] + ch[s] >= 0;
                        tot_j[s] = totals[s] + ch[s];
                    }
                    if (!valid) continue;
                    if (space.sector_of_totals(tot_j.data()) >= 0) {
                        exit += rate;
                        ++jumps;
                    } else {
                        leak += rate;
                    }
                }
            }
            if (absorbing) exit += leak;
            g.exit[i] = exit;
            g.leak[i] = leak;
            out.jumps += jumps;
        }
    });

    g.ptr.assign(n + 1, 0);
    i64 row = 0;
    u64 total = 0;
    for (const auto &o : outs) {
        for (u32 l : o.len) {
            total += l;
            g.ptr[++row] = total;
        }
        g.n_jumps += o.jumps;
    }
    g.col.resize(total);
    g.val.resize(total);
    std::vector<u64> start(outs.size() + 1, 0);
    for (std::size_t t = 0; t < outs.size(); ++t) start[t + 1] = start[t] + outs[t].col.size();
    run_tasks(static_cast<i64>(outs.size()), [&](i64 t) {
        std::copy(outs[t].col.begin(), outs[t].col.end(), g.col.begin() + start[t]);
        std::copy(outs[t].val.begin(), outs[t].val.end(), g.val.begin() + start[t]);
        std::vector<u32>().swap(outs[t].col);
        std::vector<double>().swap(outs[t].val);
    });
    return g;
}

generator_csr generator_from_jumps(i64 n, const i64 *first, const i64 *to, const double *rate,
                                   const double *exit, const double *leak) {
    if (n < 0 || n >= (i64{1} << 32)) throw std::invalid_argument("generator: bad size");
    generator_csr g;
    g.n = n;
    g.exit.assign(exit, exit + n);
    g.leak.assign(leak, leak + n);
    const i64 m = first[n];
    g.n_jumps = m;
    std::vector<u64> count(n + 1, 0);
    for (i64 k = 0; k < m; ++k) {
        if (to[k] < 0 || to[k] >= n) throw std::invalid_argument("generator: jump target out of range");
        ++count[to[k] + 1];
    }
    for (i64 i = 0; i < n; ++i) count[i + 1] += count[i];
    std::vector<u32> col(m);
    std::vector<double> val(m);
    std::vector<u64> fill(count.begin(), count.end() - 1);
    for (i64 j = 0; j < n; ++j)
        for (i64 k = first[j]; k < first[j + 1]; ++k) {
            const u64 p = fill[to[k]]++;
            col[p] = static_cast<u32>(j);
            val[p] = rate[k];
        }
    g.ptr.assign(n + 1, 0);
    for (i64 i = 0; i < n; ++i) {
        for (u64 p = count[i]; p < count[i + 1]; ++p) {
            if (p > count[i] && col[p] == g.col.back()) {
                g.val.back() += val[p];
            } else {
                g.col.push_back(col[p]);
                g.val.push_back(val[p]);
            }
        }
        g.ptr[i + 1] = g.col.size();
    }
    return g;
}

// ------------------------------------------------------------------ multigrid hierarchy
bool invert_dense(std::vector<double> a, i64 n, std::vector<double> &inverse) {
    inverse.assign(static_cast<std::size_t>(n * n), 0.0);
    for (i64 i = 0; i < n; ++i) inverse[i * n + i] = 1.0;
    for (i64 k = 0; k < n; ++k) {
        i64 p = k;
        for (i64 i = k + 1; i < n; ++i)
            if (std::abs(a[i * n + k]) > std::abs(a[p * n + k])) p = i;
        if (a[p * n + k] == 0.0 || !std::isfinite(a[p * n + k])) return false;
        if (p != k)
            for (i64 c = 0; c < n; ++c) {
                std::swap(a[k * n + c], a[p * n + c]);
                std::swap(inverse[k * n + c], inverse[p * n + c]);
            }
        const double d = a[k * n + k];
        for (i64 c = 0; c < n; ++c) {
            a[k * n + c] /= d;
            inverse[k * n + c] /= d;
        }
        for (i64 i = 0; i < n; ++i) {
            if (i == k) continue;
            const double f = a[i * n + k];
            if (f == 0.0) continue;
            for (i64 c = 0; c < n; ++c) {
                a[i * n + c] -= f * a[k * n + c];
                inverse[i * n + c] -= f * inverse[k * n + c];
            }
        }
    }
    return true;
}

namespace {

// multigrid.hpp:268-300 on one sector level; see block_smoother
block_smoother build_blocks(const population_space &space, const std::vector<i64> &parent,
                            i64 n_agg, const generator_csr &A, double shift) {
    block_smoother bs;
    bs.aggregates = n_agg;
    const i64 n = space.size();
    const int S = space.species();
    const i64 V = space.voxels();
    std::vector<std::vector<i64>> children(n_agg);
    for (i64 v = 0; v < V; ++v) children[parent[v]].push_back(v);
    i64 molecules = 0;
    for (int s = 0; s < S; ++s) molecules += space.totals(0)[s];
    const i64 W = std::max<i64>(1, std::min<i64>(molecules, n_agg)); // occupied aggregates per state
    constexpr u32 none = 0xFFFFFFFFu;
    std::vector<u32> occ(static_cast<std::size_t>(n * W), none);
    const i64 chunk = 16384;
    const i64 tasks = (n + chunk - 1) / chunk;
    // canonical key of every (aggregate, state with molecules in it), per task and aggregate
    std::vector<std::vector<std::vector<std::pair<i64, u32>>>> keyed(tasks);
    run_tasks(tasks, [&](i64 t) {
        auto &kt = keyed[t];
        kt.resize(n_agg);
        std::vector<int> x(V * S), y(V * S), inside(n_agg * S);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i) {
            space.state_in_sector(0, i, x.data());
            std::fill(inside.begin(), inside.end(), 0);
            for (i64 v = 0; v < V; ++v)
                for (int s = 0; s < S; ++s) inside[parent[v] * S + s] += x[v * S + s];
            i64 w = 0;
            for (i64 a = 0; a < n_agg; ++a) {
                bool any = false;
                for (int s = 0; s < S; ++s) any = any || inside[a * S + s] > 0;
                if (!any) continue;
                occ[i * W + w++] = static_cast<u32>(a);
                y = x;
                for (i64 v : children[a])
                    for (int s = 0; s < S; ++s) y[v * S + s] = 0;
                for (int s = 0; s < S; ++s) y[children[a][0] * S + s] = inside[a * S + s];
                kt[a].push_back({space.rank_in_sector(y.data(), 0), static_cast<u32>(i)});
            }
        }
    });
    // blocks of every aggregate: equal keys, members ascending (stable sort of row-ordered lists)
    struct agg_part {
        std::vector<u64> first{0};
        std::vector<u32> members;
        std::vector<u64> inv_offset;
        std::vector<double> inverse;
    };
    std::vector<agg_part> parts(n_agg);
    run_tasks(n_agg, [&](i64 a) {
        std::vector<std::pair<i64, u32>> all;
        for (i64 t = 0; t < tasks; ++t) all.insert(all.end(), keyed[t][a].begin(), keyed[t][a].end());
        std::stable_sort(all.begin(), all.end(),
                         [](const auto &l, const auto &r) { return l.first < r.first; });
        auto &P = parts[a];
        std::size_t b0 = 0;
        std::vector<double> dense, inv;
        while (b0 < all.size()) {
            std::size_t b1 = b0 + 1;
            while (b1 < all.size() && all[b1].first == all[b0].first) ++b1;
            const i64 m = static_cast<i64>(b1 - b0);
            if (m > 1) {
                const std::size_t m0 = P.members.size();
                for (std::size_t q = b0; q < b1; ++q) P.members.push_back(all[q].second);
                const u32 *mem = P.members.data() + m0;
                dense.assign(static_cast<std::size_t>(m * m), 0.0);
                for (i64 k = 0; k < m; ++k) {
                    const u32 i = mem[k];
                    dense[k * m + k] += shift + A.exit[i];
                    for (u64 p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
                        const u32 *hit = std::lower_bound(mem, mem + m, A.col[p]);
                        if (hit != mem + m && *hit == A.col[p]) dense[k * m + (hit - mem)] -= A.val[p];
                    }
                }
                if (!invert_dense(dense, m, inv))
                    throw std::runtime_error("population_multigrid: singular block");
                P.inv_offset.push_back(P.inverse.size());
                P.inverse.insert(P.inverse.end(), inv.begin(), inv.end());
                P.first.push_back(P.members.size());
            }
            b0 = b1;
        }
        std::vector<std::pair<i64, u32>>().swap(all);
    });
    keyed.clear();
    // rows whose residual an aggregate's correction changes: the members and every row with a
    // member among its sources
    std::vector<std::vector<u64>> count(tasks, std::vector<u64>(n_agg, 0));
    auto affected = [&](i64 i, std::vector<i64> &stamp, auto &&emit) {
        auto mark = [&](u32 row) {
            for (i64 w = 0; w < W; ++w) {
                const u32 a = occ[row * W + w];
                if (a == none) break;
                if (stamp[a] != i) {
                    stamp[a] = i;
                    emit(a);
                }
            }
        };
        mark(static_cast<u32>(i));
        for (u64 p = A.ptr[i]; p < A.ptr[i + 1]; ++p) mark(A.col[p]);
    };
    run_tasks(tasks, [&](i64 t) {
        std::vector<i64> stamp(n_agg, -1);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i)
            affected(i, stamp, [&](u32 a) { ++count[t][a]; });
    });
    bs.agg_rows.assign(n_agg + 1, 0);
    std::vector<std::vector<u64>> at(tasks, std::vector<u64>(n_agg, 0));
    for (i64 a = 0; a < n_agg; ++a) {
        u64 off = bs.agg_rows[a];
        for (i64 t = 0; t < tasks; ++t) {
            at[t][a] = off;
            off += count[t][a];
        }
        bs.agg_rows[a + 1] = off;
    }
    bs.rows.resize(bs.agg_rows[n_agg]);
    run_tasks(tasks, [&](i64 t) {
        std::vector<i64> stamp(n_agg, -1);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i)
            affected(i, stamp, [&](u32 a) { bs.rows[at[t][a]++] = static_cast<u32>(i); });
    });
    // concatenate the aggregates
    bs.agg_blocks.assign(1, 0);
    bs.block_first.assign(1, 0);
    for (i64 a = 0; a < n_agg; ++a) {
        const auto &P = parts[a];
        const u64 mbase = bs.members.size(), ibase = bs.inverse.size();
        bs.members.insert(bs.members.end(), P.members.begin(), P.members.end());
        for (std::size_t b = 1; b < P.first.size(); ++b) bs.block_first.push_back(mbase + P.first[b]);
        for (u64 o : P.inv_offset) bs.inv_offset.push_back(ibase + o);
        bs.inverse.insert(bs.inverse.end(), P.inverse.begin(), P.inverse.end());
        bs.agg_blocks.push_back(bs.inv_offset.size());
    }
    return bs;
}

// every entry of A joins the two colours (the red-black smoother's precondition)
void check_colours(const std::vector<std::uint8_t> &colour, const generator_csr &A) {
    const i64 n = A.n, chunk = 1 << 16;
    std::atomic<bool> bad{false};
    run_tasks((n + chunk - 1) / chunk, [&](i64 t) {
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk) && !bad.load(std::memory_order_relaxed); ++i)
            for (u64 p = A.ptr[i]; p < A.ptr[i + 1]; ++p)
                if (colour[A.col[p]] == colour[i]) bad = true;
    });
    if (bad) throw std::logic_error("red-black smoother: the diffusion graph is not bipartite");
}

// red-black colouring of a sector's diffusion graph (a hop changes one molecule's voxel
// coordinate sum by one, so it flips the parity); checked on every entry
std::vector<std::uint8_t> state_colours(const population_space &space, const voxel_domain &dom,
                                        const generator_csr &A) {
    const int S = space.species();
    std::vector<int> parity(dom.voxels);
    for (i64 v = 0; v < dom.voxels; ++v) {
        i64 c = 0;
        for (i64 a : dom.coordinate(v)) c += a;
        parity[v] = static_cast<int>(c & 1);
    }
    const i64 n = space.size();
    std::vector<std::uint8_t> colour(n);
    const i64 chunk = 32768;
    run_tasks((n + chunk - 1) / chunk, [&](i64 t) {
        std::vector<int> x(dom.voxels * S);
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i) {
            space.state_in_sector(0, i, x.data());
            int p = 0;
            for (i64 v = 0; v < dom.voxels; ++v)
                for (int s = 0; s < S; ++s) p ^= (x[v * S + s] & 1) & parity[v];
            colour[i] = static_cast<std::uint8_t>(p);
        }
    });
    check_colours(colour, A);
    return colour;
}

} // namespace

sector_hierarchy build_sector_hierarchy(const voxel_domain &domain, const std::vector<int> &totals,
                                        double shift, const multigrid_options &options,
                                        const level_hooks &hooks) {
    if (!(shift > 0.0)) throw std::invalid_argument("population_multigrid: shift must be positive");
    if (domain.dim == 0) throw std::invalid_argument("population_multigrid needs a grid domain");
    sector_hierarchy h;
    voxel_domain dom = domain;
    multigrid_options opt = options;
    const reaction_tables none;
    while (true) {
        const auto space = population_space::sector(dom.voxels, totals);
        sector_level level;
        level.rows = space.size();
        level.pre_sweeps = opt.pre_sweeps;
        level.post_sweeps = opt.post_sweeps;
        level.A = hooks.assemble ? hooks.assemble(space, dom) : assemble_generator(space, dom, none, true);
        std::vector<double> doubled = dom.diffusivities;
        for (double &d : doubled) d *= 2.0;
        auto coarse = coarsen(dom, doubled);
        if (level.rows <= opt.coarsest_size || opt.max_levels <= 1 || !coarse) {
            // terminal: dense (shift I - Q) on the whole level (multigrid.hpp:247-264)
            const i64 n = level.rows;
            if (n > 8192) throw std::runtime_error("population_multigrid: terminal level too large for a dense solve");
            std::vector<double> A(static_cast<std::size_t>(n * n), 0.0);
            for (i64 i = 0; i < n; ++i) {
                A[i * n + i] += shift + level.A.exit[i];
                for (u64 p = level.A.ptr[i]; p < level.A.ptr[i + 1]; ++p)
                    A[i * n + level.A.col[p]] -= level.A.val[p];
            }
            if (!invert_dense(std::move(A), n, level.inverse))
                throw std::runtime_error("population_multigrid: singular block");
            if (opt.smoother == 2) level.colour = state_colours(space, dom, level.A);
            level.terminal = true;
            h.levels.push_back(std::move(level));
            break;
        }
        // population_transfer (multigrid.hpp:50-92)
        const auto cspace = population_space::sector(coarse->voxels, totals);
        const int S = space.species();
        std::vector<i64> parent(dom.voxels);
        std::vector<i64> children(coarse->voxels, 0);
        for (i64 v = 0; v < dom.voxels; ++v) {
            auto c = dom.coordinate(v);
            for (int a = 0; a < dom.dim; ++a)
                if (dom.dims[a] > 1) c[a] /= 2;
            parent[v] = coarse->vertex(c);
            ++children[parent[v]];
        }
        std::vector<double> log_children(coarse->voxels);
        for (i64 a = 0; a < coarse->voxels; ++a) log_children[a] = std::log(double(children[a]));
        // lgamma(n + 1) for every count a state can hold
        int max_total = 0;
        for (int t : totals) max_total = std::max(max_total, t);
        std::vector<double> log_fact(max_total + 1);
        {
            int sign = 0;
            for (int c = 0; c <= max_total; ++c) log_fact[c] = lgamma_r(c + 1.0, &sign);
        }
        level.coarse_of.resize(level.rows);
        level.weight.resize(level.rows);
        const i64 chunk = 32768;
        std::vector<double> log_weight;
        const bool by_hook =
            hooks.transfer && hooks.transfer(space, cspace, dom, parent, log_fact, log_children, level.coarse_of,
                                             log_weight, opt.smoother == 2 ? &level.colour : nullptr);
        if (by_hook) {
            run_tasks((level.rows + chunk - 1) / chunk, [&](i64 t) {
                for (i64 i = t * chunk; i < std::min(level.rows, (t + 1) * chunk); ++i)
                    level.weight[i] = std::exp(log_weight[i]);
            });
            if (opt.smoother == 2) check_colours(level.colour, level.A);
        } else {
            if (opt.smoother == 2) level.colour = state_colours(space, dom, level.A);
        }
        if (!by_hook) run_tasks((level.rows + chunk - 1) / chunk, [&](i64 t) {
            std::vector<int> x(dom.voxels * S), c(coarse->voxels * S);
            for (i64 i = t * chunk; i < std::min(level.rows, (t + 1) * chunk); ++i) {
                space.state_in_sector(0, i, x.data());
                std::fill(c.begin(), c.end(), 0);
                double log_weight = 0.0;
                // empty voxels add lgamma(1) = 0 and 0 * log(children) = 0, which leave the sums
                // unchanged, so only occupied ones are visited (the same sums as visiting all)
                for (i64 v = 0; v < dom.voxels; ++v)
                    for (int s = 0; s < S; ++s) {
                        const int n = x[v * S + s];
                        if (n == 0) continue;
                        c[parent[v] * S + s] += n;
                        log_weight -= log_fact[n];
                    }
                for (i64 a = 0; a < coarse->voxels; ++a)
                    for (int s = 0; s < S; ++s) {
                        const int n = c[a * S + s];
                        if (n == 0) continue;
                        log_weight += log_fact[n] - (n * log_children[a]);
                    }
                level.coarse_of[i] = cspace.rank_in_sector(c.data(), 0);
                level.weight[i] = std::exp(log_weight);
            }
        });
        std::vector<double> total(cspace.size(), 0.0);
        for (i64 i = 0; i < level.rows; ++i) total[level.coarse_of[i]] += level.weight[i];
        for (i64 i = 0; i < level.rows; ++i) level.weight[i] /= total[level.coarse_of[i]];
        if (opt.smoother == 1) {
            level.blocks = build_blocks(space, parent, coarse->voxels, level.A, shift);
            level.has_blocks = true;
        }
        h.levels.push_back(std::move(level));
        dom = std::move(*coarse);
        opt.max_levels -= 1;
        opt.pre_sweeps += options.sweep_growth;
        opt.post_sweeps += options.sweep_growth;
    }
    return h;
}

// ------------------------------------------------------------------ small dense exponential
std::vector<double> expm_pade(std::vector<double> X, i64 n) {
    double norm = 0.0;
    for (i64 i = 0; i < n; ++i) {
        double r = 0.0;
        for (i64 j = 0; j < n; ++j) r += std::abs(X[i * n + j]);
        norm = std::max(norm, r);
    }
    const int squarings = norm > 0.5 ? static_cast<int>(std::ceil(std::log2(norm / 0.5))) : 0;
    const double scale = std::ldexp(1.0, -squarings);
    for (double &x : X) x *= scale;
    auto product = [n](const std::vector<double> &A, const std::vector<double> &B) {
        std::vector<double> C(static_cast<std::size_t>(n * n), 0.0);
        for (i64 i = 0; i < n; ++i)
            for (i64 k = 0; k < n; ++k) {
                const double a = A[i * n + k];
                if (a == 0.0) continue;
                for (i64 j = 0; j < n; ++j) C[i * n + j] += a * B[k * n + j];
            }
        return C;
    };
    // N = sum c_k X^k, D = sum (-1)^k c_k X^k, c_0 = 1, c_k = c_{k-1} (q - k + 1) / ((2q - k + 1) k)
    constexpr int q = 6;
    std::vector<double> N(static_cast<std::size_t>(n * n), 0.0), D(N), P(N);
    for (i64 i = 0; i < n; ++i) P[i * n + i] = 1.0;
    double c = 1.0;
    for (int k = 0; k <= q; ++k) {
        if (k > 0) {
            c *= double(q - k + 1) / double((2 * q - k + 1) * k);
            P = product(P, X);
        }
        const double sign = (k % 2) ? -1.0 : 1.0;
        for (std::size_t e = 0; e < N.size(); ++e) {
            N[e] += c * P[e];
            D[e] += sign * c * P[e];
        }
    }
    std::vector<double> Dinv;
    if (!invert_dense(std::move(D), n, Dinv)) throw std::runtime_error("expm: singular Pade denominator");
    std::vector<double> F = product(Dinv, N);
    for (int k = 0; k < squarings; ++k) F = product(F, F);
    return F;
}

// ------------------------------------------------------------------ coded sliced layout
namespace {

// the distinct values of v converted to VT, sorted; empty when there are more than 256
template <typename VT>
std::vector<VT> dictionary(const std::vector<double> &v) {
    const i64 n = static_cast<i64>(v.size());
    const i64 chunk = 1 << 20;
    const i64 tasks = (n + chunk - 1) / chunk;
    std::vector<std::vector<VT>> parts(tasks);
    std::vector<char> over(tasks, 0);
    run_tasks(tasks, [&](i64 t) {
        auto &d = parts[t];
        for (i64 i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i) {
            const VT x = static_cast<VT>(v[i]);
            auto it = std::lower_bound(d.begin(), d.end(), x);
            if (it == d.end() || *it != x) {
                if (d.size() >= 256) {
                    over[t] = 1;
                    return;
                }
                d.insert(it, x);
            }
        }
    });
    std::vector<VT> all;
    for (i64 t = 0; t < tasks; ++t) {
        if (over[t]) return {};
        for (VT x : parts[t]) {
            auto it = std::lower_bound(all.begin(), all.end(), x);
            if (it == all.end() || *it != x) all.insert(it, x);
        }
        if (all.size() > 256) return {};
    }
    if (all.empty()) all.push_back(VT(0));
    return all;
}

template <typename VT>
u32 code_of(const std::vector<VT> &dict, double v) {
    return static_cast<u32>(std::lower_bound(dict.begin(), dict.end(), static_cast<VT>(v)) -
                            dict.begin());
}

} // namespace

template <typename VT>
bool encode_sliced(const generator_csr &g, sliced_operator<VT> &out, bool exit_array) {
    const i64 n = g.n;
    for (i64 i = 0; i < n; ++i)
        if (g.ptr[i + 1] - g.ptr[i] > 255) return false;
    // every column (rows of a square operator, finer-level rows of a restriction) fits 24 bits
    u32 max_col = 0;
    for (u32 c : g.col) max_col = std::max(max_col, c);
    if (!g.col.empty() && max_col >= (1u << kCodedColumnBits)) return false;
    out.vdict = dictionary<VT>(g.val);
    if (out.vdict.empty()) return false;
    out.edict = dictionary<VT>(g.exit);
    if (out.edict.empty()) {
        if (!exit_array) return false;
        out.exit.assign(g.exit.begin(), g.exit.end());
    }
    const i64 n_slices = (n + kSliceRows - 1) / kSliceRows;
    out.soff.assign(n_slices + 1, 0);
    u64 total = 0;
    for (i64 sl = 0; sl < n_slices; ++sl) {
        u64 w = 0;
        for (i64 r = sl * kSliceRows; r < std::min(n, (sl + 1) * kSliceRows); ++r)
            w = std::max<u64>(w, g.ptr[r + 1] - g.ptr[r]);
        total += w * kSliceRows;
        if (total >= (u64{1} << 32)) return false;
        out.soff[sl + 1] = static_cast<u32>(total);
    }
    out.padded = total;
    out.ent.assign(total, 0u);
    out.rlen.assign(n, 0);
    out.rcode.assign(n, 0);
    const i64 chunk = 1024; // slices per task
    run_tasks((n_slices + chunk - 1) / chunk, [&](i64 t) {
        for (i64 sl = t * chunk; sl < std::min(n_slices, (t + 1) * chunk); ++sl) {
            for (i64 l = 0; l < kSliceRows; ++l) {
                const i64 i = sl * kSliceRows + l;
                if (i >= n) break;
                out.rlen[i] = static_cast<std::uint8_t>(g.ptr[i + 1] - g.ptr[i]);
                if (!out.edict.empty())
                    out.rcode[i] = static_cast<std::uint8_t>(code_of(out.edict, g.exit[i]));
                u64 k = 0;
                for (u64 p = g.ptr[i]; p < g.ptr[i + 1]; ++p, ++k)
                    out.ent[out.soff[sl] + k * kSliceRows + l] =
                        g.col[p] | (code_of(out.vdict, g.val[p]) << kCodedColumnBits);
            }
        }
    });
    return true;
}

template bool encode_sliced<double>(const generator_csr &, sliced_operator<double> &, bool);
template bool encode_sliced<float>(const generator_csr &, sliced_operator<float> &, bool);

} // namespace mk
