// ORACLE — test infrastructure only. This is an FP64, single-threaded-solve CPU restatement
// of MarkovKit's RDME resolvent path, used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the CHECKER. Nothing in markovkit_amd/ may include, link or
// call it; the product path fails loudly without its HIP library instead of falling back here.
//
// What it restates (reference files, relative to the markovkit repository root):
//   grid<D> (num::structured_grid_graph, absent third-party numerics @ dc6f1ff):
//       axis 0 fastest, non-periodic, neighbour weight = face area / spacing,
//       cell volume = prod(spacing)   [call sites include/markovkit/rdme/multigrid.hpp:20-41,
//                                      pinned by tests/test_rdme.cpp:22-33]
//   coarsen                        include/markovkit/rdme/multigrid.hpp:24-42
//   binomial / compositions / enumerate / rank
//                                  include/markovkit/rdme/population_space.hpp:15-100
//   population_sector / population_box   population_space.hpp:105-196
//   voxel_domain::jump_rate, reaction_diffusion::transitions
//                                  include/markovkit/rdme/reaction_diffusion.hpp:23-118
//   generator (jump lists, absorbing / reflecting, scatter apply / apply_shifted)
//                                  include/markovkit/generator.hpp:27-107
//   population_transfer (multinomial P, aggregation R)   multigrid.hpp:47-118
//   population_multigrid (V/W cycle, weighted Jacobi and aggregate-block smoothers)
//                                                         multigrid.hpp:139-361
//       with two more smoothers: red-black Gauss-Seidel and ILU(0) in rank order
//                                  rdme/cpp/rdme_full_multigrid.cpp:178-195, :204-296
//   sector_multigrid                                       multigrid.hpp:371-408
//   num::gmres (absent; restated as restarted right-preconditioned GMRES, Saad's
//       algorithm with modified Gram-Schmidt and Givens rotations, absolute Euclidean
//       tolerance, x in = initial guess)   call sites tests/test_rdme.cpp:238-243,
//       rdme/cpp/rdme_full_multigrid.cpp:1187-1209
//   dense partial-pivot LU (num::lu / num::solve, absent)  multigrid.hpp:221, :247-264
//   num::expv (absent; restated as Sidje's Expokit expv with a degree-6 Pade exponential of
//       the Hessenberg matrix)   call sites include/markovkit/fsp.hpp:152, :225
//
// Pins (tests/test_oracle_pins.py): the exact values in
// tests/test_rdme.cpp:22-223, the reference's recorded transfer weights
// (rdme/paper/data/transfer_sparsity/*.csv), the lumped sector-sum closed form that
// reproduces the recorded masses 0.337423312883 and 0.82128451753, an independent
// brute-force Python generator on tiny boxes, R P = I and R A P = A_c, and a direct
// sparse LU solve of the assembled system.
//
// Parity unpinned: the iterate sequence, iteration counts and the meaning of
// solver_result::residual of num::gmres (its source is absent); only the converged answer
// is pinned (unique solution of a nonsingular M-matrix system).
//
// Setup loops (enumeration, generator rows, transfer weights) may run on several threads;
// they produce exactly what the serial loop does. Solves are single-threaded.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace mk_oracle {

using idx = std::uint64_t;
using real = double;
using state = std::vector<int>;

inline unsigned setup_threads() {
    const unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 1U : std::min(n, 16U);
}

// run body(begin, end) over [0, n) on a few threads; chunks are contiguous and ordered
template <typename F>
void parallel_chunks(idx n, F &&body) {
    const unsigned threads = n < 4096 ? 1U : setup_threads();
    if (threads <= 1) {
        body(idx{0}, n);
        return;
    }
    std::vector<std::thread> pool;
    const idx step = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        const idx b = std::min<idx>(n, t * step), e = std::min<idx>(n, b + step);
        pool.emplace_back([&, b, e] { body(b, e); });
    }
    for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------- voxel graphs
struct edge {
    idx to;
    real weight;
};

// num::structured_grid_graph<D>: vertex v = c0 + dims0*(c1 + dims1*c2); neighbours
// listed axis by axis, the lower neighbour first; weight = (prod h) / h_axis^2.
struct grid {
    std::vector<idx> dims;
    std::vector<real> spacing;

    grid(std::vector<idx> d, std::vector<real> h = {}) : dims(std::move(d)), spacing(std::move(h)) {
        if (spacing.empty()) spacing.assign(dims.size(), 1.0);
        if (spacing.size() != dims.size()) throw std::invalid_argument("grid: one spacing per axis");
    }
    [[nodiscard]] std::size_t D() const { return dims.size(); }
    [[nodiscard]] idx n_vertices() const {
        idx n = 1;
        for (idx d : dims) n *= d;
        return n;
    }
    [[nodiscard]] real cell_volume() const {
        real v = 1.0;
        for (real h : spacing) v *= h;
        return v;
    }
    [[nodiscard]] std::vector<idx> coordinate(idx v) const {
        std::vector<idx> c(D());
        for (std::size_t a = 0; a < D(); ++a) {
            c[a] = v % dims[a];
            v /= dims[a];
        }
        return c;
    }
    [[nodiscard]] idx vertex(const std::vector<idx> &c) const {
        idx v = 0;
        for (std::size_t a = D(); a-- > 0;) v = v * dims[a] + c[a];
        return v;
    }
    [[nodiscard]] real face_weight(std::size_t axis) const {
        real area = 1.0;
        for (std::size_t b = 0; b < D(); ++b)
            if (b != axis) area *= spacing[b];
        return area / spacing[axis];
    }
    [[nodiscard]] std::vector<edge> neighbors(idx v) const {
        std::vector<edge> out;
        const auto c = coordinate(v);
        idx stride = 1;
        for (std::size_t a = 0; a < D(); ++a) {
            const real w = face_weight(a);
            if (c[a] > 0) out.push_back({v - stride, w});
            if (c[a] + 1 < dims[a]) out.push_back({v + stride, w});
            stride *= dims[a];
        }
        return out;
    }
};

// multigrid.hpp:24-42
inline std::optional<grid> coarsen(const grid &fine) {
    auto dims = fine.dims;
    auto h = fine.spacing;
    bool reduced = false;
    for (std::size_t a = 0; a < dims.size(); ++a) {
        if (dims[a] == 1) continue;
        if (dims[a] % 2 != 0) return std::nullopt;
        dims[a] /= 2;
        h[a] *= 2.0;
        reduced = true;
    }
    return reduced ? std::optional<grid>(grid(dims, h)) : std::nullopt;
}

// reaction_diffusion.hpp:23-55; the graph is held as adjacency lists, built from a grid
// or given explicitly (num::graph), and `structured` keeps the grid for the multigrid
struct voxel_domain {
    std::vector<std::vector<edge>> adjacency;
    std::vector<real> volumes;
    std::vector<real> diffusivities;
    std::optional<grid> structured;

    voxel_domain(std::vector<std::vector<edge>> adj, std::vector<real> vols, std::vector<real> diff)
        : adjacency(std::move(adj)), volumes(std::move(vols)), diffusivities(std::move(diff)) {
        if (volumes.size() != adjacency.size())
            throw std::invalid_argument("voxel_domain: one volume per voxel");
        for (real v : volumes)
            if (!(v > 0.0) || !std::isfinite(v))
                throw std::invalid_argument("voxel_domain: volumes must be positive");
    }
    voxel_domain(const grid &g, std::vector<real> diff)
        : voxel_domain(grid_adjacency(g), std::vector<real>(g.n_vertices(), g.cell_volume()),
                       std::move(diff)) {
        structured = g;
    }
    voxel_domain(const grid &g, std::vector<real> vols, std::vector<real> diff)
        : voxel_domain(grid_adjacency(g), std::move(vols), std::move(diff)) {
        structured = g;
    }
    static std::vector<std::vector<edge>> grid_adjacency(const grid &g) {
        std::vector<std::vector<edge>> adj(g.n_vertices());
        for (idx v = 0; v < g.n_vertices(); ++v) adj[v] = g.neighbors(v);
        return adj;
    }
    [[nodiscard]] idx n_voxels() const { return adjacency.size(); }
    [[nodiscard]] idx n_species() const { return diffusivities.size(); }
    [[nodiscard]] real jump_rate(idx species, idx from, const edge &e) const {
        return diffusivities[species] * e.weight / volumes[from];
    }
};

// ---------------------------------------------------------------- population spaces
namespace detail {

// population_space.hpp:15-25
inline idx binomial(idx n, idx k) {
    k = std::min(k, n - k);
    idx result = 1;
    for (idx i = 1; i <= k; ++i) {
        if (result > std::numeric_limits<idx>::max() / (n - k + i))
            throw std::overflow_error("population state space is too large");
        result = result * (n - k + i) / i;
    }
    return result;
}

// population_space.hpp:28-30
inline idx compositions(idx voxels, int total) {
    return binomial(static_cast<idx>(total) + voxels - 1, voxels - 1);
}

// population_space.hpp:34-52: rank order, species 0 most significant, voxel 0 most
// significant within a species, counts ascending. States are stored flat, V*S ints each.
inline void enumerate(idx voxels, const std::vector<int> &totals, idx species, idx voxel,
                      int remaining, state &x, std::vector<int> &out) {
    const idx S = totals.size();
    if (species == S) {
        out.insert(out.end(), x.begin(), x.end());
        return;
    }
    const idx at = voxel * S + species;
    if (voxel + 1 == voxels) {
        x[at] = remaining;
        enumerate(voxels, totals, species + 1, 0, species + 1 < S ? totals[species + 1] : 0, x,
                  out);
        return;
    }
    for (int count = 0; count <= remaining; ++count) {
        x[at] = count;
        enumerate(voxels, totals, species, voxel + 1, remaining - count, x, out);
    }
}

// population_space.hpp:54-72
inline idx enumerate(idx voxels, const std::vector<int> &totals, std::vector<int> &out) {
    if (voxels == 0 || totals.empty())
        throw std::invalid_argument("population state space needs a voxel and a species");
    idx size = 1;
    for (int total : totals) {
        if (total < 0) throw std::invalid_argument("species totals must be nonnegative");
        const idx factor = compositions(voxels, total);
        if (size > std::numeric_limits<idx>::max() / factor)
            throw std::overflow_error("population state space is too large");
        size *= factor;
    }
    out.reserve(out.size() + size * voxels * totals.size());
    state x(voxels * totals.size(), 0);
    enumerate(voxels, totals, 0, 0, totals[0], x, out);
    return size;
}

// C(rem + rest, rest) for rest < voxels and rem <= max_total, filled with binomial() once per
// space: the same integers the reference's rank recomputes per call
struct binomial_table {
    idx stride = 0;
    std::vector<idx> values;
    binomial_table() = default;
    binomial_table(idx voxels, int max_total) : stride(static_cast<idx>(max_total) + 1) {
        values.resize(voxels * stride);
        for (idx rest = 0; rest < voxels; ++rest)
            for (int rem = 0; rem <= max_total; ++rem) {
                idx v;
                try {
                    v = binomial(static_cast<idx>(rem) + rest, rest);
                } catch (const std::overflow_error &) {
                    v = std::numeric_limits<idx>::max();
                }
                values[rest * stride + rem] = v;
            }
    }
    [[nodiscard]] idx operator()(idx rest, int rem) const { return values[rest * stride + rem]; }
};

// population_space.hpp:75-100 (binomials from the table when one is given)
inline std::optional<idx> rank(idx voxels, const std::vector<int> &totals, const int *x,
                               idx size, const binomial_table *table = nullptr) {
    const idx S = totals.size();
    if (size != voxels * S) return std::nullopt;
    idx result = 0;
    for (idx s = 0; s < S; ++s) {
        int remaining = totals[s];
        idx local = 0;
        for (idx v = 0; v + 1 < voxels; ++v) {
            const int count = x[v * S + s];
            if (count < 0 || count > remaining) return std::nullopt;
            const idx rest = voxels - v - 1;
            if (table) {
                local += (*table)(rest, remaining) - (*table)(rest, remaining - count);
            } else {
                local += binomial(remaining + rest, rest) - binomial(remaining - count + rest, rest);
            }
            remaining -= count;
        }
        if (x[(voxels - 1) * S + s] != remaining) return std::nullopt;
        result = result * (table ? (*table)(voxels - 1, totals[s]) : compositions(voxels, totals[s])) + local;
    }
    return result;
}

} // namespace detail

// population_space.hpp:105-127
class population_sector {
  public:
    population_sector(idx n_voxels, std::vector<int> totals)
        : voxels_(n_voxels), totals_(std::move(totals)) {
        size_ = detail::enumerate(voxels_, totals_, states_);
        table_ = detail::binomial_table(voxels_, *std::max_element(totals_.begin(), totals_.end()));
    }
    [[nodiscard]] idx size() const { return size_; }
    [[nodiscard]] const int *operator[](idx i) const { return states_.data() + i * width(); }
    [[nodiscard]] state at(idx i) const { return state((*this)[i], (*this)[i] + width()); }
    [[nodiscard]] std::optional<idx> find(const int *x, idx n) const {
        return detail::rank(voxels_, totals_, x, n, &table_);
    }
    [[nodiscard]] std::optional<idx> find(const state &x) const { return find(x.data(), x.size()); }
    [[nodiscard]] idx n_voxels() const { return voxels_; }
    [[nodiscard]] idx n_species() const { return totals_.size(); }
    [[nodiscard]] idx width() const { return voxels_ * totals_.size(); }
    [[nodiscard]] const std::vector<int> &totals() const { return totals_; }

  private:
    idx voxels_;
    std::vector<int> totals_;
    std::vector<int> states_;
    idx size_ = 0;
    detail::binomial_table table_;
};

// population_space.hpp:131-196
class population_box {
  public:
    population_box(idx n_voxels, std::vector<int> max_totals)
        : voxels_(n_voxels), max_totals_(std::move(max_totals)) {
        begin_.push_back(0);
        std::vector<int> totals(max_totals_.size(), 0);
        add_sectors(0, totals);
        table_ = detail::binomial_table(voxels_, *std::max_element(max_totals_.begin(), max_totals_.end()));
    }
    [[nodiscard]] idx size() const { return begin_.back(); }
    [[nodiscard]] const int *operator[](idx i) const { return states_.data() + i * width(); }
    [[nodiscard]] state at(idx i) const { return state((*this)[i], (*this)[i] + width()); }
    [[nodiscard]] std::optional<idx> find(const int *x, idx n) const {
        if (n != width()) return std::nullopt;
        const idx S = n_species();
        thread_local std::vector<int> totals;
        totals.assign(S, 0);
        for (idx v = 0; v < voxels_; ++v)
            for (idx s = 0; s < S; ++s) totals[s] += x[v * S + s];
        const auto k = find_sector(totals);
        const auto local = k ? detail::rank(voxels_, totals, x, n, &table_) : std::nullopt;
        return local ? std::optional<idx>(begin_[*k] + *local) : std::nullopt;
    }
    [[nodiscard]] std::optional<idx> find(const state &x) const { return find(x.data(), x.size()); }
    [[nodiscard]] idx n_voxels() const { return voxels_; }
    [[nodiscard]] idx n_species() const { return max_totals_.size(); }
    [[nodiscard]] idx width() const { return voxels_ * max_totals_.size(); }
    [[nodiscard]] idx n_sectors() const { return totals_.size(); }
    [[nodiscard]] const std::vector<int> &sector_totals(idx k) const { return totals_[k]; }
    [[nodiscard]] idx sector_begin(idx k) const { return begin_[k]; }
    [[nodiscard]] idx sector_end(idx k) const { return begin_[k + 1]; }
    [[nodiscard]] idx sector_of(idx i) const {
        return static_cast<idx>(std::upper_bound(begin_.begin(), begin_.end(), i) -
                                begin_.begin() - 1);
    }
    [[nodiscard]] std::optional<idx> find_sector(const std::vector<int> &totals) const {
        idx k = 0;
        for (idx s = 0; s < n_species(); ++s) {
            if (totals[s] < 0 || totals[s] > max_totals_[s]) return std::nullopt;
            k = k * (max_totals_[s] + 1) + totals[s];
        }
        return k;
    }
    [[nodiscard]] const std::vector<int> &max_totals() const { return max_totals_; }

  private:
    void add_sectors(idx species, std::vector<int> &totals) {
        if (species == n_species()) {
            totals_.push_back(totals);
            begin_.push_back(begin_.back() + detail::enumerate(voxels_, totals, states_));
            return;
        }
        for (int total = 0; total <= max_totals_[species]; ++total) {
            totals[species] = total;
            add_sectors(species + 1, totals);
        }
    }
    idx voxels_;
    std::vector<int> max_totals_;
    std::vector<std::vector<int>> totals_;
    std::vector<idx> begin_;
    std::vector<int> states_;
    detail::binomial_table table_;
};

// ---------------------------------------------------------------- the RDME model
// reaction_diffusion.hpp:63-67
struct local_reaction {
    std::vector<int> change;
    std::function<real(const int *counts, real volume)> rate;
};

// reaction_diffusion.hpp:71-118. transitions() hands each jump to a visitor as
// (target state, rate) in the reference order; the target buffer is reused between calls.
struct reaction_diffusion {
    voxel_domain domain;
    std::vector<local_reaction> reactions;

    [[nodiscard]] idx n_voxels() const { return domain.n_voxels(); }
    [[nodiscard]] idx n_species() const { return domain.n_species(); }
    [[nodiscard]] idx index(idx voxel, idx species) const { return voxel * n_species() + species; }

    template <typename Visit>
    void transitions(const int *x, state &y, Visit &&visit) const {
        const idx S = n_species(), W = n_voxels() * S;
        y.assign(x, x + W);
        for (idx v = 0; v < n_voxels(); ++v) {
            for (const edge &e : domain.adjacency[v]) {
                for (idx s = 0; s < S; ++s) {
                    const int n = x[index(v, s)];
                    if (n == 0) continue;
                    --y[index(v, s)];
                    ++y[index(e.to, s)];
                    visit(y, n * domain.jump_rate(s, v, e));
                    ++y[index(v, s)];
                    --y[index(e.to, s)];
                }
            }
        }
        for (idx v = 0; v < n_voxels(); ++v) {
            const int *counts = x + index(v, 0);
            for (const local_reaction &r : reactions) {
                const real rate = r.rate(counts, domain.volumes[v]);
                if (!(rate > 0.0)) continue;
                bool valid = true;
                for (idx s = 0; s < S; ++s) {
                    y[index(v, s)] += r.change[s];
                    valid = valid && y[index(v, s)] >= 0;
                }
                if (valid) visit(y, rate);
                for (idx s = 0; s < S; ++s) y[index(v, s)] -= r.change[s];
            }
        }
    }
};

enum class boundary : std::uint8_t { reflecting, absorbing };

// generator.hpp:27-107: per-source jump lists; scatter apply and apply_shifted
class generator {
  public:
    struct jump {
        idx to;
        real rate;
    };

    template <typename Space>
    generator(const Space &space, const reaction_diffusion &model,
              boundary leaks = boundary::absorbing)
        : first_(space.size() + 1, 0), exit_(space.size(), 0.0), leak_(space.size(), 0.0) {
        const idx n = space.size();
        const idx W = space.width();
        // each thread lists the jumps of a contiguous range of sources, then the lists are
        // concatenated in source order: the same arrays as the serial loop
        const unsigned threads = n < 4096 ? 1U : setup_threads();
        std::vector<std::vector<jump>> parts(threads);
        std::vector<idx> count(n, 0);
        std::vector<std::thread> pool;
        const idx step = (n + threads - 1) / threads;
        for (unsigned t = 0; t < threads; ++t) {
            pool.emplace_back([&, t] {
                const idx b = std::min<idx>(n, t * step), e = std::min<idx>(n, b + step);
                state y;
                for (idx j = b; j < e; ++j) {
                    model.transitions(space[j], y, [&](const state &to, real rate) {
                        if (auto i = space.find(to.data(), W)) {
                            parts[t].push_back({*i, rate});
                            exit_[j] += rate;
                            ++count[j];
                        } else {
                            leak_[j] += rate;
                        }
                    });
                    if (leaks == boundary::absorbing) exit_[j] += leak_[j];
                }
            });
            if (threads == 1) pool.back().join();
        }
        if (threads > 1)
            for (auto &th : pool) th.join();
        for (idx j = 0; j < n; ++j) first_[j + 1] = first_[j] + count[j];
        jumps_.reserve(first_[n]);
        for (auto &p : parts) {
            jumps_.insert(jumps_.end(), p.begin(), p.end());
            std::vector<jump>().swap(p);
        }
    }

    [[nodiscard]] idx rows() const { return exit_.size(); }
    [[nodiscard]] idx cols() const { return exit_.size(); }
    [[nodiscard]] idx n_jumps() const { return jumps_.size(); }
    [[nodiscard]] const jump *jumps_begin(idx j) const { return jumps_.data() + first_[j]; }
    [[nodiscard]] const jump *jumps_end(idx j) const { return jumps_.data() + first_[j + 1]; }
    [[nodiscard]] const std::vector<real> &exit_rates() const { return exit_; }
    [[nodiscard]] const std::vector<real> &leak_rates() const { return leak_; }
    [[nodiscard]] const std::vector<idx> &first() const { return first_; }
    [[nodiscard]] const std::vector<jump> &all_jumps() const { return jumps_; }

    // y = Q x   (generator.hpp:62-70)
    void apply(const real *x, real *y) const {
        std::fill(y, y + rows(), 0.0);
        for (idx j = 0; j < rows(); ++j) {
            y[j] -= exit_[j] * x[j];
            for (const jump *e = jumps_begin(j); e != jumps_end(j); ++e) y[e->to] += e->rate * x[j];
        }
    }
    // y = shift x - Q x   (generator.hpp:73-81)
    void apply_shifted(real shift, const real *x, real *y) const {
        std::fill(y, y + rows(), 0.0);
        for (idx j = 0; j < rows(); ++j) {
            y[j] += (shift + exit_[j]) * x[j];
            for (const jump *e = jumps_begin(j); e != jumps_end(j); ++e) y[e->to] -= e->rate * x[j];
        }
    }

  private:
    std::vector<idx> first_;
    std::vector<jump> jumps_;
    std::vector<real> exit_;
    std::vector<real> leak_;
};

// ---------------------------------------------------------------- dense LU (num::lu)
struct lu_result {
    idx n = 0;
    std::vector<real> a; // row-major, L (unit) and U packed
    std::vector<idx> piv;
    bool singular = false;
};

inline lu_result lu(std::vector<real> a, idx n) {
    lu_result r;
    r.n = n;
    r.piv.resize(n);
    for (idx k = 0; k < n; ++k) {
        idx p = k;
        for (idx i = k + 1; i < n; ++i)
            if (std::abs(a[i * n + k]) > std::abs(a[p * n + k])) p = i;
        r.piv[k] = p;
        if (a[p * n + k] == 0.0) {
            r.singular = true;
            continue;
        }
        if (p != k)
            for (idx c = 0; c < n; ++c) std::swap(a[k * n + c], a[p * n + c]);
        for (idx i = k + 1; i < n; ++i) {
            a[i * n + k] /= a[k * n + k];
            for (idx c = k + 1; c < n; ++c) a[i * n + c] -= a[i * n + k] * a[k * n + c];
        }
    }
    r.a = std::move(a);
    return r;
}

inline void solve(const lu_result &f, const real *b, real *x) {
    const idx n = f.n;
    std::vector<real> y(b, b + n);
    for (idx k = 0; k < n; ++k) std::swap(y[k], y[f.piv[k]]);
    for (idx i = 0; i < n; ++i)
        for (idx c = 0; c < i; ++c) y[i] -= f.a[i * n + c] * y[c];
    for (idx i = n; i-- > 0;) {
        for (idx c = i + 1; c < n; ++c) y[i] -= f.a[i * n + c] * y[c];
        y[i] /= f.a[i * n + i];
    }
    std::copy(y.begin(), y.end(), x);
}

// ---------------------------------------------------------------- multigrid
// multigrid.hpp:47-118
class population_transfer {
  public:
    population_transfer(const grid &fine_grid, const grid &coarse_grid,
                        const population_sector &fine, const population_sector &coarse)
        : parent_(fine.n_voxels()), coarse_of_(fine.size()), weight_(fine.size()),
          coarse_size_(coarse.size()) {
        const idx S = fine.n_species();
        std::vector<idx> children(coarse.n_voxels(), 0);
        for (idx v = 0; v < fine.n_voxels(); ++v) {
            auto c = fine_grid.coordinate(v);
            for (std::size_t a = 0; a < fine_grid.D(); ++a)
                if (fine_grid.dims[a] > 1) c[a] /= 2;
            parent_[v] = coarse_grid.vertex(c);
            ++children[parent_[v]];
        }
        parallel_chunks(fine.size(), [&](idx b, idx e) {
            state c(coarse.n_voxels() * S, 0);
            for (idx i = b; i < e; ++i) {
                std::fill(c.begin(), c.end(), 0);
                real log_weight = 0.0;
                for (idx v = 0; v < fine.n_voxels(); ++v) {
                    for (idx s = 0; s < S; ++s) {
                        const int n = fine[i][v * S + s];
                        c[parent_[v] * S + s] += n;
                        log_weight -= std::lgamma(n + 1.0);
                    }
                }
                for (idx a = 0; a < coarse.n_voxels(); ++a) {
                    for (idx s = 0; s < S; ++s) {
                        const int n = c[a * S + s];
                        log_weight += std::lgamma(n + 1.0) - (n * std::log(real(children[a])));
                    }
                }
                coarse_of_[i] = *coarse.find(c);
                weight_[i] = std::exp(log_weight);
            }
        });
        std::vector<real> total(coarse.size(), 0.0);
        for (idx i = 0; i < fine.size(); ++i) total[coarse_of_[i]] += weight_[i];
        for (idx i = 0; i < fine.size(); ++i) weight_[i] /= total[coarse_of_[i]];
    }
    [[nodiscard]] idx fine_size() const { return coarse_of_.size(); }
    [[nodiscard]] idx coarse_size() const { return coarse_size_; }
    [[nodiscard]] const std::vector<idx> &parent() const { return parent_; }
    [[nodiscard]] const std::vector<idx> &coarse_of() const { return coarse_of_; }
    [[nodiscard]] const std::vector<real> &weight() const { return weight_; }

    void restrict(const real *fine, real *coarse) const {
        std::fill(coarse, coarse + coarse_size_, 0.0);
        for (idx i = 0; i < fine_size(); ++i) coarse[coarse_of_[i]] += fine[i];
    }
    void prolong(const real *coarse, real *fine) const {
        for (idx i = 0; i < fine_size(); ++i) fine[i] = weight_[i] * coarse[coarse_of_[i]];
    }

  private:
    std::vector<idx> parent_;
    std::vector<idx> coarse_of_;
    std::vector<real> weight_;
    idx coarse_size_;
};

// multigrid.hpp:120-134, plus two smoothers of the paper driver's family:
//   red_black: Gauss-Seidel in two colours. The fixed-population diffusion graph on a grid is
//     bipartite (a hop changes the parity of the sum over molecules of their voxels'
//     coordinate sums; rdme/cpp/rdme_full_multigrid.cpp:218-221), so updating every state of
//     one colour from the residual taken just before is exact Gauss-Seidel in the order
//     "colour 0 states, then colour 1 states" (lexicographic GS as at
//     rdme_full_multigrid.cpp:178-195 in that order). Colour 0 first; reversed for
//     post-smoothing.
//   ilu0: x += (L U)^-1 (rhs - A x) with the ILU(0) factors of the level operator in its rank
//     order (rdme_full_multigrid.cpp:204-296, structured case). On a bipartite graph ILU(0)
//     changes only the diagonal: pivot_i = a_ii - sum_{j < i} a_ij a_ji / pivot_j.
enum class relaxation : std::uint8_t { jacobi, aggregate_block, red_black, ilu0 };

struct multigrid_options {
    int pre_sweeps = 2;
    int post_sweeps = 2;
    real jacobi_weight = 2.0 / 3.0;
    idx coarsest_size = 4;
    int max_levels = 32;
    int coarse_cycles = 1;
    int sweep_growth = 0;
    relaxation smoother = relaxation::jacobi;
    // the library's extension: only the top cycle_levels levels take coarse_cycles visits of
    // their coarse level, the ones below take one (0: every level, the reference's semantics)
    int cycle_levels = 0;
};

// multigrid.hpp:139-361 (weighted Jacobi and aggregate-block smoothers)
class population_multigrid {
  public:
    population_multigrid(const voxel_domain &domain, const std::vector<int> &totals, real shift,
                         multigrid_options options = {})
        : space_(domain.n_voxels(), totals),
          Q_(space_, reaction_diffusion{domain, {}}), shift_(shift), options_(options),
          rhs_(rows()), x_(rows()), r_(rows()), e_(rows()) {
        if (!(shift > 0.0))
            throw std::invalid_argument("population_multigrid: shift must be positive");
        if (!domain.structured) throw std::invalid_argument("population_multigrid needs a grid");
        if (options.smoother == relaxation::red_black) build_colours(*domain.structured);
        if (options.smoother == relaxation::ilu0) build_ilu();
        auto coarse_grid = coarsen(*domain.structured);
        if (rows() <= options.coarsest_size || options.max_levels <= 1 || !coarse_grid) {
            std::vector<idx> all(rows());
            std::iota(all.begin(), all.end(), 0);
            terminal_ = factor(all, all);
            return;
        }
        std::vector<real> doubled = domain.diffusivities;
        for (real &d : doubled) d *= 2.0;
        multigrid_options coarse_options = options;
        coarse_options.max_levels -= 1;
        coarse_options.pre_sweeps += options.sweep_growth;
        coarse_options.post_sweeps += options.sweep_growth;
        if (options.cycle_levels == 1) coarse_options.coarse_cycles = 1;
        if (options.cycle_levels > 1) coarse_options.cycle_levels = options.cycle_levels - 1;
        coarse_ = std::make_unique<population_multigrid>(voxel_domain(*coarse_grid, doubled),
                                                         totals, shift, coarse_options);
        transfer_ = std::make_unique<population_transfer>(*domain.structured, *coarse_grid,
                                                          space_, coarse_->space_);
        const idx n = coarse_->rows();
        coarse_rhs_.assign(n, 0.0);
        coarse_x_.assign(n, 0.0);
        coarse_r_.assign(n, 0.0);
        coarse_e_.assign(n, 0.0);
        if (options.smoother == relaxation::aggregate_block) build_blocks();
    }

    [[nodiscard]] idx rows() const { return space_.size(); }
    [[nodiscard]] real shift() const { return shift_; }
    [[nodiscard]] const population_sector &space() const { return space_; }
    [[nodiscard]] const generator &fine_generator() const { return Q_; }
    [[nodiscard]] const population_multigrid *coarse_level() const { return coarse_.get(); }
    [[nodiscard]] const population_transfer &transfer() const { return *transfer_; }
    [[nodiscard]] idx n_levels() const { return coarse_ ? 1 + coarse_->n_levels() : 1; }
    [[nodiscard]] real grid_complexity() const { return real(total_states()) / real(rows()); }
    [[nodiscard]] real operator_complexity() const {
        return real(total_nonzeros()) / real(rows() + Q_.n_jumps());
    }

    void apply_system(const real *x, real *y) const { Q_.apply_shifted(shift_, x, y); }

    // multigrid.hpp:201-228
    void cycle(const real *b, real *x) const {
        std::copy(b, b + rows(), rhs_.begin());
        std::fill(x_.begin(), x_.end(), 0.0);
        smooth(options_.pre_sweeps, false);
        residual();
        if (coarse_) {
            transfer_->restrict(r_.data(), coarse_rhs_.data());
            std::fill(coarse_x_.begin(), coarse_x_.end(), 0.0);
            for (int c = 0; c < options_.coarse_cycles; ++c) {
                coarse_->apply_system(coarse_x_.data(), coarse_r_.data());
                for (idx i = 0; i < coarse_r_.size(); ++i) coarse_r_[i] = coarse_rhs_[i] - coarse_r_[i];
                coarse_->cycle(coarse_r_.data(), coarse_e_.data());
                for (idx i = 0; i < coarse_x_.size(); ++i) coarse_x_[i] += coarse_e_[i];
            }
            transfer_->prolong(coarse_x_.data(), e_.data());
        } else {
            solve(terminal_, r_.data(), e_.data());
        }
        for (idx i = 0; i < rows(); ++i) x_[i] += e_[i];
        smooth(options_.post_sweeps, true);
        std::copy(x_.begin(), x_.end(), x);
    }

    // multigrid.hpp:231-236: `sweeps` smoother sweeps on A x = rhs, starting from x
    void relax(const real *rhs, real *x, int sweeps = 1, bool reverse = false) const {
        std::copy(rhs, rhs + rows(), rhs_.begin());
        std::copy(x, x + rows(), x_.begin());
        smooth(sweeps, reverse);
        std::copy(x_.begin(), x_.end(), x);
    }

    // aggregate a's blocks, in the order the smoother visits them (for tests)
    [[nodiscard]] idx n_aggregates() const { return blocks_.size(); }
    [[nodiscard]] idx n_blocks(idx a) const { return blocks_[a].size(); }
    [[nodiscard]] const std::vector<idx> &block_states(idx a, idx b) const {
        return blocks_[a][b].states;
    }
    // red_black: the colour of every state; ilu0: the ILU(0) pivots (for tests)
    [[nodiscard]] const std::vector<std::uint8_t> &colours() const { return colour_; }
    [[nodiscard]] const std::vector<real> &ilu_pivots() const { return pivot_; }

  private:
    struct block {
        std::vector<idx> states;
        lu_result lu;
        mutable std::vector<real> rhs, e;
    };

    // multigrid.hpp:268-300: group the states that agree outside each coarse voxel; each group
    // of two or more states is one block. The groups are visited in ascending order of the
    // outside configuration (the reference's hash-map order is unspecified; only the rounding
    // of the residual updates depends on it).
    void build_blocks() {
        const idx S = space_.n_species();
        const auto &parent = transfer_->parent();
        std::vector<idx> position(rows(), rows());
        blocks_.resize(coarse_->space().n_voxels());
        for (idx a = 0; a < static_cast<idx>(blocks_.size()); ++a) {
            std::map<state, std::vector<idx>> groups;
            for (idx i = 0; i < rows(); ++i) {
                state outside(space_[i], space_[i] + space_.n_voxels() * S);
                for (idx v = 0; v < space_.n_voxels(); ++v)
                    if (parent[v] == a) std::fill_n(outside.begin() + v * S, S, 0);
                groups[std::move(outside)].push_back(i);
            }
            for (auto &[outside, members] : groups) {
                if (members.size() == 1) continue;
                for (idx k = 0; k < static_cast<idx>(members.size()); ++k) position[members[k]] = k;
                auto f = factor(members, position);
                for (idx i : members) position[i] = rows();
                const idx n = members.size();
                blocks_[a].push_back({std::move(members), std::move(f), std::vector<real>(n, 0.0),
                                      std::vector<real>(n, 0.0)});
            }
        }
    }

    [[nodiscard]] lu_result factor(const std::vector<idx> &states,
                                   const std::vector<idx> &position) const {
        const idx n = states.size();
        std::vector<real> A(n * n, 0.0);
        for (idx k = 0; k < n; ++k) {
            A[k * n + k] = shift_ + Q_.exit_rates()[states[k]];
            for (const auto *e = Q_.jumps_begin(states[k]); e != Q_.jumps_end(states[k]); ++e)
                if (position[e->to] < n) A[position[e->to] * n + k] -= e->rate;
        }
        auto f = lu(std::move(A), n);
        if (f.singular) throw std::runtime_error("population_multigrid: singular block");
        return f;
    }
    // colour of x = parity of sum over voxels v and species s of x[v, s] * (coordinate sum of v);
    // every jump must join the two colours
    void build_colours(const grid &g) {
        const idx S = space_.n_species();
        std::vector<int> parity(space_.n_voxels());
        for (idx v = 0; v < space_.n_voxels(); ++v) {
            idx c = 0;
            for (idx a : g.coordinate(v)) c += a;
            parity[v] = static_cast<int>(c & 1);
        }
        colour_.assign(rows(), 0);
        for (idx i = 0; i < rows(); ++i) {
            int p = 0;
            for (idx v = 0; v < space_.n_voxels(); ++v)
                for (idx s = 0; s < S; ++s) p ^= (space_[i][v * S + s] & 1) & parity[v];
            colour_[i] = static_cast<std::uint8_t>(p);
        }
        for (idx j = 0; j < rows(); ++j)
            for (const auto *e = Q_.jumps_begin(j); e != Q_.jumps_end(j); ++e)
                if (colour_[e->to] == colour_[j])
                    throw std::logic_error("red-black smoother: the diffusion graph is not bipartite");
    }
    // gather lists (row i: every j with a jump j -> i, ascending) and the ILU(0) pivots
    void build_ilu() {
        const idx n = rows();
        in_first_.assign(n + 1, 0);
        for (idx j = 0; j < n; ++j)
            for (const auto *e = Q_.jumps_begin(j); e != Q_.jumps_end(j); ++e) ++in_first_[e->to + 1];
        for (idx i = 0; i < n; ++i) in_first_[i + 1] += in_first_[i];
        in_from_.resize(in_first_[n]);
        in_rate_.resize(in_first_[n]);
        std::vector<idx> fill(in_first_.begin(), in_first_.end() - 1);
        for (idx j = 0; j < n; ++j) // sources ascending: each row's list comes out ascending
            for (const auto *e = Q_.jumps_begin(j); e != Q_.jumps_end(j); ++e) {
                in_from_[fill[e->to]] = j;
                in_rate_[fill[e->to]++] = e->rate;
            }
        // a_ij a_ji = rate(j -> i) rate(i -> j); the reverse rate is found in row j's list
        pivot_.assign(n, 0.0);
        for (idx i = 0; i < n; ++i) {
            real d = shift_ + Q_.exit_rates()[i];
            for (idx p = in_first_[i]; p < in_first_[i + 1] && in_from_[p] < i; ++p) {
                const idx j = in_from_[p];
                real back = 0.0;
                for (idx q = in_first_[j]; q < in_first_[j + 1]; ++q)
                    if (in_from_[q] == i) back = in_rate_[q];
                d -= in_rate_[p] * back / pivot_[j];
            }
            if (!(d > 0.0) || !std::isfinite(d)) throw std::runtime_error("population_multigrid: ILU(0) pivot");
            pivot_[i] = d;
        }
    }
    void residual() const {
        apply_system(x_.data(), r_.data());
        for (idx i = 0; i < rows(); ++i) r_[i] = rhs_[i] - r_[i];
    }
    // multigrid.hpp:309-342
    void smooth(int sweeps, bool reverse) const {
        if (!colour_.empty()) {
            for (int sweep = 0; sweep < sweeps; ++sweep)
                for (int half = 0; half < 2; ++half) {
                    const std::uint8_t c = static_cast<std::uint8_t>(reverse ? 1 - half : half);
                    residual();
                    for (idx i = 0; i < rows(); ++i)
                        if (colour_[i] == c) x_[i] += r_[i] / (shift_ + Q_.exit_rates()[i]);
                }
            return;
        }
        if (!pivot_.empty()) {
            const idx n = rows();
            for (int sweep = 0; sweep < sweeps; ++sweep) {
                residual();
                // forward: (D + L) u = r; backward: (D + U) e = D u; e overwrites u in e_
                for (idx i = 0; i < n; ++i) {
                    real s = r_[i];
                    for (idx p = in_first_[i]; p < in_first_[i + 1] && in_from_[p] < i; ++p)
                        s += in_rate_[p] * e_[in_from_[p]];
                    e_[i] = s / pivot_[i];
                }
                for (idx i = n; i-- > 0;) {
                    real s = 0.0;
                    for (idx p = in_first_[i]; p < in_first_[i + 1]; ++p)
                        if (in_from_[p] > i) s += in_rate_[p] * e_[in_from_[p]];
                    e_[i] += s / pivot_[i];
                }
                for (idx i = 0; i < n; ++i) x_[i] += e_[i];
            }
            return;
        }
        if (blocks_.empty()) {
            for (int sweep = 0; sweep < sweeps; ++sweep) {
                residual();
                for (idx i = 0; i < rows(); ++i)
                    x_[i] += options_.jacobi_weight * r_[i] / (shift_ + Q_.exit_rates()[i]);
            }
            return;
        }
        // block Gauss-Seidel over aggregates, block Jacobi within one aggregate
        residual();
        const idx A = blocks_.size();
        for (int sweep = 0; sweep < sweeps; ++sweep) {
            for (idx step = 0; step < A; ++step) {
                const auto &aggregate = blocks_[reverse ? A - 1 - step : step];
                for (const block &b : aggregate) {
                    for (idx k = 0; k < static_cast<idx>(b.states.size()); ++k) b.rhs[k] = r_[b.states[k]];
                    solve(b.lu, b.rhs.data(), b.e.data());
                }
                for (const block &b : aggregate) {
                    for (idx k = 0; k < static_cast<idx>(b.states.size()); ++k) {
                        const idx j = b.states[k];
                        x_[j] += b.e[k];
                        r_[j] -= (shift_ + Q_.exit_rates()[j]) * b.e[k];
                        for (const auto *e = Q_.jumps_begin(j); e != Q_.jumps_end(j); ++e)
                            r_[e->to] += e->rate * b.e[k];
                    }
                }
            }
        }
    }
    [[nodiscard]] idx total_states() const {
        return rows() + (coarse_ ? coarse_->total_states() : 0);
    }
    [[nodiscard]] idx total_nonzeros() const {
        return rows() + Q_.n_jumps() + (coarse_ ? coarse_->total_nonzeros() : 0);
    }

    population_sector space_;
    generator Q_;
    real shift_;
    multigrid_options options_;
    std::unique_ptr<population_multigrid> coarse_;
    std::unique_ptr<population_transfer> transfer_;
    lu_result terminal_;
    std::vector<std::vector<block>> blocks_;
    std::vector<std::uint8_t> colour_;
    std::vector<idx> in_first_, in_from_;
    std::vector<real> in_rate_, pivot_;
    mutable std::vector<real> rhs_, x_, r_, e_;
    mutable std::vector<real> coarse_rhs_, coarse_x_, coarse_r_, coarse_e_;
};

// multigrid.hpp:371-408
class sector_multigrid {
  public:
    sector_multigrid(const voxel_domain &domain, const population_box &box, real shift,
                     multigrid_options options = {}) {
        for (idx k = 0; k < box.n_sectors(); ++k) {
            levels_.push_back(
                std::make_unique<population_multigrid>(domain, box.sector_totals(k), shift, options));
            begin_.push_back(box.sector_begin(k));
        }
        begin_.push_back(box.size());
    }
    [[nodiscard]] idx rows() const { return begin_.back(); }
    [[nodiscard]] idx n_sectors() const { return levels_.size(); }
    [[nodiscard]] idx max_levels() const {
        idx r = 0;
        for (const auto &l : levels_) r = std::max(r, l->n_levels());
        return r;
    }
    [[nodiscard]] const population_multigrid &sector(idx k) const { return *levels_[k]; }
    void apply(const real *b, real *x) const {
        for (idx k = 0; k < levels_.size(); ++k)
            levels_[k]->cycle(b + begin_[k], x + begin_[k]);
    }

  private:
    std::vector<std::unique_ptr<population_multigrid>> levels_;
    std::vector<idx> begin_;
};

// ---------------------------------------------------------------- GMRES (num::gmres)
struct gmres_options {
    real tolerance = 1e-10; // absolute, Euclidean
    idx max_iterations = 1000;
    idx restart = 30;
};

struct solver_result {
    idx iterations = 0;
    real residual = 0.0; // final true residual ||b - A x||_2
    bool converged = false;
};

inline real dot(const real *a, const real *b, idx n) {
    real s = 0.0;
    for (idx i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// restarted GMRES, right preconditioned: A M^-1 u = b, x = x0 + M^-1 u. `apply_a(x, y)` is
// y = A x and `apply_m(v, z)` is z ~ A^-1 v (identity when absent). The preconditioned
// vectors z_j = M v_j are stored, so the update is x += Z y (flexible form, equal to the
// right-preconditioned iterate for a fixed linear M).
template <typename ApplyA, typename ApplyM>
solver_result gmres(idx n, ApplyA &&apply_a, ApplyM &&apply_m, const real *b, real *x,
                    const gmres_options &opt) {
    solver_result res;
    const idx m = std::max<idx>(1, std::min<idx>(opt.restart, n));
    std::vector<real> V((m + 1) * n), Z(m * n), w(n), r(n);
    std::vector<real> H((m + 1) * m, 0.0), cs(m), sn(m), g(m + 1);
    auto true_residual = [&]() {
        apply_a(x, r.data());
        for (idx i = 0; i < n; ++i) r[i] = b[i] - r[i];
        return std::sqrt(dot(r.data(), r.data(), n));
    };
    real beta = true_residual();
    res.residual = beta;
    if (!(beta > opt.tolerance)) {
        res.converged = std::isfinite(beta);
        return res;
    }
    while (res.iterations < opt.max_iterations) {
        for (idx i = 0; i < n; ++i) V[i] = r[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        idx j = 0;
        bool done = false;
        for (; j < m && res.iterations < opt.max_iterations; ++j) {
            ++res.iterations;
            apply_m(&V[j * n], &Z[j * n]);
            apply_a(&Z[j * n], w.data());
            for (idx k = 0; k <= j; ++k) { // modified Gram-Schmidt
                const real h = dot(&V[k * n], w.data(), n);
                H[k * m + j] = h;
                for (idx i = 0; i < n; ++i) w[i] -= h * V[k * n + i];
            }
            const real hn = std::sqrt(dot(w.data(), w.data(), n));
            H[(j + 1) * m + j] = hn;
            if (hn > 0.0)
                for (idx i = 0; i < n; ++i) V[(j + 1) * n + i] = w[i] / hn;
            for (idx k = 0; k < j; ++k) { // previous rotations
                const real a = H[k * m + j], c = H[(k + 1) * m + j];
                H[k * m + j] = cs[k] * a + sn[k] * c;
                H[(k + 1) * m + j] = -sn[k] * a + cs[k] * c;
            }
            const real a = H[j * m + j], c = H[(j + 1) * m + j];
            const real d = std::hypot(a, c);
            cs[j] = d > 0.0 ? a / d : 1.0;
            sn[j] = d > 0.0 ? c / d : 0.0;
            H[j * m + j] = d;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (!(std::abs(g[j + 1]) > opt.tolerance) || !(hn > 0.0)) {
                ++j;
                done = true;
                break;
            }
        }
        std::vector<real> y(j, 0.0);
        for (idx i = j; i-- > 0;) {
            real s = g[i];
            for (idx k = i + 1; k < j; ++k) s -= H[i * m + k] * y[k];
            y[i] = s / H[i * m + i];
        }
        for (idx k = 0; k < j; ++k)
            for (idx i = 0; i < n; ++i) x[i] += y[k] * Z[k * n + i];
        beta = true_residual();
        res.residual = beta;
        if (!std::isfinite(beta)) break;
        if (!(beta > opt.tolerance)) {
            res.converged = true;
            break;
        }
        (void)done;
    }
    return res;
}

// ---------------------------------------------------------------- Krylov expv (num::expv)
// num::expv(t, A, v, m, tol) is called at include/markovkit/fsp.hpp:152 and :225; its source
// (third-party numerics @ dc6f1ff) is absent. Restated as Sidje's expv (Expokit, ACM TOMS 24,
// 1998): Arnoldi (modified Gram-Schmidt) of dimension m on w, the exponential of the
// (m + 2) x (m + 2) augmented Hessenberg matrix, the a-posteriori error estimate from the
// two corrected terms, and step-size control with gamma 0.9, delta 1.2, at most 10 rejections.

// exp(X) for a small dense X (row-major n x n): diagonal Pade approximant of degree 6 with
// scaling and squaring (Golub & Van Loan, Matrix Computations, Algorithm 9.3.1)
inline std::vector<real> expm_pade6(std::vector<real> X, idx n) {
    real norm = 0.0;
    for (idx i = 0; i < n; ++i) {
        real r = 0.0;
        for (idx j = 0; j < n; ++j) r += std::abs(X[i * n + j]);
        norm = std::max(norm, r);
    }
    int sq = 0;
    if (norm > 0.5) sq = static_cast<int>(std::ceil(std::log2(norm / 0.5)));
    const real scale = std::ldexp(1.0, -sq);
    for (real &x : X) x *= scale;
    auto mul = [n](const std::vector<real> &A, const std::vector<real> &B) {
        std::vector<real> C(n * n, 0.0);
        for (idx i = 0; i < n; ++i)
            for (idx k = 0; k < n; ++k) {
                const real a = A[i * n + k];
                for (idx j = 0; j < n; ++j) C[i * n + j] += a * B[k * n + j];
            }
        return C;
    };
    constexpr int q = 6;
    real c = 1.0;
    std::vector<real> N(n * n, 0.0), D(n * n, 0.0), P(n * n, 0.0);
    for (idx i = 0; i < n; ++i) P[i * n + i] = 1.0;
    for (int k = 0; k <= q; ++k) {
        if (k > 0) {
            c *= real(q - k + 1) / real((2 * q - k + 1) * k);
            P = mul(P, X);
        }
        const real sign = (k % 2) ? -1.0 : 1.0;
        for (idx i = 0; i < n * n; ++i) {
            N[i] += c * P[i];
            D[i] += sign * c * P[i];
        }
    }
    // F = D^-1 N
    auto f = lu(D, n);
    if (f.singular) throw std::runtime_error("expm: singular Pade denominator");
    std::vector<real> F(n * n), col(n), out(n);
    for (idx j = 0; j < n; ++j) {
        for (idx i = 0; i < n; ++i) col[i] = N[i * n + j];
        solve(f, col.data(), out.data());
        for (idx i = 0; i < n; ++i) F[i * n + j] = out[i];
    }
    for (int k = 0; k < sq; ++k) F = mul(F, F);
    return F;
}

struct expv_options {
    idx krylov_dim = 30;
    real tolerance = 1e-8;
};

struct expv_result {
    idx steps = 0;
    idx rejected = 0;
    real error = 0.0; // accumulated local error estimates
};

// w = exp(t A) v with y = A x given by apply(x, y); anorm is a norm estimate of A
template <typename Apply>
expv_result expv(real t, idx n, const Apply &apply, real anorm, const real *v, real *w,
                 const expv_options &opt = {}) {
    expv_result res;
    const idx m = std::max<idx>(1, std::min<idx>(opt.krylov_dim, n));
    const real tol = opt.tolerance, btol = 1e-7, gamma = 0.9, delta = 1.2;
    const int max_reject = 10;
    const real t_out = std::abs(t), sgn = t >= 0.0 ? 1.0 : -1.0;
    const real rndoff = anorm * std::numeric_limits<real>::epsilon();
    auto norm2 = [n](const real *x) {
        real s = 0.0;
        for (idx i = 0; i < n; ++i) s += x[i] * x[i];
        return std::sqrt(s);
    };
    std::copy(v, v + n, w);
    real beta = norm2(w);
    if (!(beta > 0.0) || t_out == 0.0) return res;
    auto round_step = [](real step) {
        const real s = std::pow(10.0, std::floor(std::log10(step)) - 1.0);
        return std::ceil(step / s) * s;
    };
    real xm = 1.0 / real(m);
    const real fact = std::pow(real(m + 1) / std::exp(1.0), real(m + 1)) *
                      std::sqrt(2.0 * 3.141592653589793 * real(m + 1));
    real t_new = (1.0 / anorm) * std::pow((fact * tol) / (4.0 * beta * anorm), xm);
    t_new = round_step(t_new);
    real t_now = 0.0;
    std::vector<real> V((m + 1) * n), p(n);
    const idx ld = m + 2;
    while (t_now < t_out) {
        ++res.steps;
        real t_step = std::min(t_out - t_now, t_new);
        std::vector<real> H(ld * ld, 0.0);
        for (idx i = 0; i < n; ++i) V[i] = w[i] / beta;
        int k1 = 2;
        idx mb = m;
        for (idx j = 0; j < m; ++j) {
            apply(V.data() + j * n, p.data());
            for (idx i = 0; i <= j; ++i) {
                real h = 0.0;
                const real *Vi = V.data() + i * n;
                for (idx r = 0; r < n; ++r) h += Vi[r] * p[r];
                H[i * ld + j] = h;
                for (idx r = 0; r < n; ++r) p[r] -= h * Vi[r];
            }
            const real s = norm2(p.data());
            if (s < btol) {
                k1 = 0;
                mb = j + 1;
                t_step = t_out - t_now;
                break;
            }
            H[(j + 1) * ld + j] = s;
            for (idx r = 0; r < n; ++r) V[(j + 1) * n + r] = p[r] / s;
        }
        real avnorm = 0.0;
        if (k1 != 0) {
            H[(m + 1) * ld + m] = 1.0;
            apply(V.data() + m * n, p.data());
            avnorm = norm2(p.data());
        }
        std::vector<real> F;
        real err_loc = btol;
        for (int rejected = 0;; ++rejected) {
            const idx mx = mb + k1;
            std::vector<real> X(mx * mx);
            for (idx i = 0; i < mx; ++i)
                for (idx j = 0; j < mx; ++j) X[i * mx + j] = sgn * t_step * H[i * ld + j];
            F = expm_pade6(std::move(X), mx);
            if (k1 == 0) {
                err_loc = btol;
                // F is mb x mb: widen to the layout below
                std::vector<real> G(ld * ld, 0.0);
                for (idx i = 0; i < mb; ++i)
                    for (idx j = 0; j < mb; ++j) G[i * ld + j] = F[i * mb + j];
                F = std::move(G);
                break;
            }
            const real phi1 = std::abs(beta * F[m * mx + 0]);
            const real phi2 = std::abs(beta * F[(m + 1) * mx + 0] * avnorm);
            if (phi1 > 10.0 * phi2) {
                err_loc = phi2;
                xm = 1.0 / real(m);
            } else if (phi1 > phi2) {
                err_loc = (phi1 * phi2) / (phi1 - phi2);
                xm = 1.0 / real(m);
            } else {
                err_loc = phi1;
                xm = 1.0 / real(std::max<idx>(1, m - 1));
            }
            if (err_loc <= delta * t_step * tol) {
                std::vector<real> G(ld * ld, 0.0);
                for (idx i = 0; i < mx; ++i)
                    for (idx j = 0; j < mx; ++j) G[i * ld + j] = F[i * mx + j];
                F = std::move(G);
                break;
            }
            if (rejected == max_reject) throw std::runtime_error("expv: the requested tolerance is too high");
            t_step = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
            ++res.rejected;
        }
        const idx mx = mb + std::max(0, k1 - 1);
        std::fill(w, w + n, 0.0);
        for (idx k = 0; k < mx; ++k) {
            const real c = beta * F[k * ld + 0];
            for (idx r = 0; r < n; ++r) w[r] += c * V[k * n + r];
        }
        beta = norm2(w);
        t_now += t_step;
        t_new = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
        res.error += std::max(err_loc, rndoff);
        if (!(beta > 0.0)) break;
    }
    return res;
}

} // namespace mk_oracle
