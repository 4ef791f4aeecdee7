// Header-only C++20 wrappers with MarkovKit's names and signatures over the markovkit_amd C
// ABI (include/markovkit_amd.h), so callers written against the reference's resolvent path
// compile unchanged and run on the MI355X:
//
//   markovkit::population_sector     include/markovkit/rdme/population_space.hpp:103-127
//   markovkit::population_box        include/markovkit/rdme/population_space.hpp:131-196
//   markovkit::state_space            include/markovkit/chain.hpp:45-50
//   markovkit::voxel_domain<G>        include/markovkit/rdme/reaction_diffusion.hpp:23-56
//   markovkit::local_reaction         reaction_diffusion.hpp:63-67
//   markovkit::reaction_diffusion<G>  reaction_diffusion.hpp:71-119
//   markovkit::generator              include/markovkit/generator.hpp:19-112
//   markovkit::shifted_generator      generator.hpp:116-133
//   markovkit::multigrid_options      include/markovkit/rdme/multigrid.hpp:125-134
//   markovkit::coarsen<D>             multigrid.hpp:24-42
//   markovkit::population_transfer    multigrid.hpp:47-118 (applied on host vectors)
//   markovkit::population_multigrid<D> multigrid.hpp:139-236
//   markovkit::sector_multigrid<D>    multigrid.hpp:371-408
//   num::structured_grid_graph<D>, num::graph, num::gmres, num::gmres_options,
//   num::solver_result, num::spmat, num::linear_operator
//                                     (numerics library; call sites tests/test_rdme.cpp)
//
// Errors from the library are rethrown as the reference's exception types
// (std::invalid_argument, std::overflow_error, std::runtime_error). Link with
// -lmarkovkit_amd.
#pragma once

#include "../markovkit_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace num {

using idx = std::size_t;
using real = double;
template <typename T>
using vec = std::vector<T>;
template <typename T>
using array = std::vector<T>;
template <typename T>
using view = std::span<T>;
template <typename T, std::size_t N>
using static_array = std::array<T, N>;

struct weighted_edge {
    idx to;
    real weight;
};

// non-periodic structured grid: vertex = c0 + d0 (c1 + d1 c2); neighbours axis by axis, the
// lower one first; weight = face area / spacing; cell volume = product of spacings
template <std::size_t D>
class structured_grid_graph {
  public:
    using index_type = idx;
    structured_grid_graph(static_array<idx, D> dims, static_array<real, D> spacing)
        : dims_(dims), spacing_(spacing) {}
    explicit structured_grid_graph(static_array<idx, D> dims) : dims_(dims) {
        spacing_.fill(1.0);
    }
    [[nodiscard]] const static_array<idx, D> &dimensions() const { return dims_; }
    [[nodiscard]] const static_array<real, D> &spacing() const { return spacing_; }
    [[nodiscard]] idx n_vertices() const {
        idx n = 1;
        for (idx d : dims_) n *= d;
        return n;
    }
    [[nodiscard]] real cell_volume() const {
        real v = 1.0;
        for (real h : spacing_) v *= h;
        return v;
    }
    [[nodiscard]] static_array<idx, D> coordinate(idx v) const {
        static_array<idx, D> c{};
        for (std::size_t a = 0; a < D; ++a) {
            c[a] = v % dims_[a];
            v /= dims_[a];
        }
        return c;
    }
    [[nodiscard]] idx vertex(const static_array<idx, D> &c) const {
        idx v = 0;
        for (std::size_t a = D; a-- > 0;) v = v * dims_[a] + c[a];
        return v;
    }
    [[nodiscard]] std::vector<weighted_edge> neighbors(idx v) const {
        std::vector<weighted_edge> out;
        const auto c = coordinate(v);
        idx stride = 1;
        for (std::size_t a = 0; a < D; ++a) {
            real area = 1.0;
            for (std::size_t b = 0; b < D; ++b)
                if (b != a) area *= spacing_[b];
            const real w = area / spacing_[a];
            if (c[a] > 0) out.push_back({v - stride, w});
            if (c[a] + 1 < dims_[a]) out.push_back({v + stride, w});
            stride *= dims_[a];
        }
        return out;
    }

  private:
    static_array<idx, D> dims_;
    static_array<real, D> spacing_{};
};

// explicit undirected weighted graph (add_edge adds both directions)
template <typename W = real, typename I = idx>
class graph {
  public:
    using index_type = I;
    struct edge {
        I to;
        W weight;
    };
    explicit graph(I n) : adj_(n) {}
    void add_edge(I u, I v, W w) {
        adj_[u].push_back({v, w});
        adj_[v].push_back({u, w});
    }
    [[nodiscard]] idx n_vertices() const { return adj_.size(); }
    [[nodiscard]] const std::vector<edge> &neighbors(I v) const { return adj_[v]; }

  private:
    std::vector<std::vector<edge>> adj_;
};

// an operator y = A x between vectors (numerics; asserted at tests/test_rdme.cpp:104, :163)
template <typename Op, typename In, typename Out>
concept linear_operator = requires(const Op &op, const In &x, Out &y) {
    typename Op::domain_type;
    typename Op::codomain_type;
    { op.rows() } -> std::convertible_to<idx>;
    { op.cols() } -> std::convertible_to<idx>;
    op.apply(x, y);
};

// host sparse matrix in compressed rows (numerics spmat: from_triplets sums duplicates, apply)
class spmat {
  public:
    using domain_type = vec<real>;
    using codomain_type = vec<real>;
    spmat() = default;
    spmat(idx rows, idx cols, array<idx> ptr, array<idx> col, array<real> val)
        : rows_(rows), cols_(cols), ptr_(std::move(ptr)), col_(std::move(col)), val_(std::move(val)) {}
    static spmat from_triplets(idx rows, idx cols, const array<idx> &row, const array<idx> &col,
                               const array<real> &value) {
        if (row.size() != col.size() || row.size() != value.size())
            throw std::invalid_argument("spmat: triplet arrays differ in length");
        array<idx> order(row.size());
        std::iota(order.begin(), order.end(), idx{0});
        std::stable_sort(order.begin(), order.end(), [&](idx a, idx b) {
            return row[a] != row[b] ? row[a] < row[b] : col[a] < col[b];
        });
        array<idx> ptr(rows + 1, 0), c;
        array<real> v;
        for (idx q = 0; q < order.size(); ++q) {
            const idx t = order[q];
            if (row[t] >= rows || col[t] >= cols) throw std::invalid_argument("spmat: triplet out of range");
            if (q > 0 && row[order[q - 1]] == row[t] && col[order[q - 1]] == col[t]) {
                v.back() += value[t];
                continue;
            }
            c.push_back(col[t]);
            v.push_back(value[t]);
            ++ptr[row[t] + 1];
        }
        for (idx i = 0; i < rows; ++i) ptr[i + 1] += ptr[i];
        return spmat(rows, cols, std::move(ptr), std::move(c), std::move(v));
    }
    [[nodiscard]] idx rows() const { return rows_; }
    [[nodiscard]] idx cols() const { return cols_; }
    [[nodiscard]] idx nnz() const { return col_.size(); }
    [[nodiscard]] const array<idx> &row_ptr() const { return ptr_; }
    [[nodiscard]] const array<idx> &col_idx() const { return col_; }
    [[nodiscard]] const array<real> &values() const { return val_; }
    // y = A x
    void apply(view<const real> x, view<real> y) const {
        for (idx i = 0; i < rows_; ++i) {
            real sum = 0.0;
            for (idx p = ptr_[i]; p < ptr_[i + 1]; ++p) sum += val_[p] * x[col_[p]];
            y[i] = sum;
        }
    }

  private:
    idx rows_ = 0, cols_ = 0;
    array<idx> ptr_, col_;
    array<real> val_;
};

struct gmres_options {
    real tolerance = 1e-10;
    idx max_iterations = 1000;
    idx restart = 30;
    // extension: Arnoldi orthogonalisation of the device solver (MK_DCGS2 or MK_CGS2)
    int orthogonalization = MK_DCGS2;
};

struct solver_result {
    idx iterations = 0;
    real residual = 0.0;
    bool converged = false;
};

} // namespace num

namespace markovkit {

using num::array;
using num::idx;
using num::real;
using num::vec;
using num::view;
using state = array<int>;

namespace detail {

inline void check(mk_status st) {
    if (st == MK_OK) return;
    const std::string msg = mk_last_error();
    if (st == MK_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (st == MK_ERR_OVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}

template <typename H, mk_status (*Destroy)(H)>
struct handle {
    H h = nullptr;
    handle() = default;
    explicit handle(H x) : h(x) {}
    handle(const handle &) = delete;
    handle &operator=(const handle &) = delete;
    handle(handle &&o) noexcept : h(std::exchange(o.h, nullptr)) {}
    handle &operator=(handle &&o) noexcept {
        std::swap(h, o.h);
        return *this;
    }
    ~handle() {
        if (h) Destroy(h);
    }
};

} // namespace detail

// a finite state space numbers its states 0, ..., size() - 1 (chain.hpp:45-50)
template <typename Space>
concept state_space = requires(const Space &space, const state &x, idx i) {
    { space.size() } -> std::convertible_to<idx>;
    { space[i] } -> std::convertible_to<const state &>;
    { space.find(x) } -> std::convertible_to<std::optional<idx>>;
};

// population_sector (population_space.hpp:103-127): every state with totals[s] molecules of
// species s, in rank order; states are voxel-major
class population_sector {
  public:
    population_sector(idx n_voxels, array<int> totals) : voxels_(n_voxels), totals_(std::move(totals)) {
        mk_box b = nullptr;
        detail::check(mk_sector_space_create(static_cast<int64_t>(n_voxels),
                                             static_cast<int32_t>(totals_.size()), totals_.data(), &b));
        space_ = std::make_shared<detail::handle<mk_box, mk_box_destroy>>(b);
        int64_t v = 0;
        detail::check(mk_box_size(b, &v));
        size_ = static_cast<idx>(v);
    }
    [[nodiscard]] idx size() const { return size_; }
    [[nodiscard]] state operator[](idx i) const {
        state x(voxels_ * n_species(), 0);
        detail::check(mk_box_state(space_->h, static_cast<int64_t>(i), x.data()));
        return x;
    }
    [[nodiscard]] std::optional<idx> find(const state &x) const {
        if (x.size() != voxels_ * n_species()) return std::nullopt;
        int64_t i = -1;
        detail::check(mk_box_find(space_->h, x.data(), &i));
        return i < 0 ? std::nullopt : std::optional<idx>(static_cast<idx>(i));
    }
    [[nodiscard]] idx n_voxels() const { return voxels_; }
    [[nodiscard]] idx n_species() const { return totals_.size(); }
    [[nodiscard]] const array<int> &totals() const { return totals_; }
    // every state, enumerated on first use
    [[nodiscard]] const array<state> &states() const {
        std::call_once(*states_once_, [&] {
            states_.reserve(size_);
            for (idx i = 0; i < size_; ++i) states_.push_back((*this)[i]);
        });
        return states_;
    }
    // the largest count of each species in one voxel (rate tables of a generator)
    [[nodiscard]] const array<int> &max_totals() const { return totals_; }
    [[nodiscard]] mk_box native() const { return space_->h; }

  private:
    idx voxels_;
    array<int> totals_;
    std::shared_ptr<detail::handle<mk_box, mk_box_destroy>> space_;
    idx size_ = 0;
    mutable array<state> states_;
    std::shared_ptr<std::once_flag> states_once_ = std::make_shared<std::once_flag>();
};

// population_box (population_space.hpp:131-196); states are voxel-major
class population_box {
  public:
    population_box(idx n_voxels, array<int> max_totals)
        : voxels_(n_voxels), max_totals_(std::move(max_totals)) {
        mk_box b = nullptr;
        detail::check(mk_box_create(static_cast<int64_t>(n_voxels),
                                    static_cast<int32_t>(max_totals_.size()), max_totals_.data(), &b));
        box_ = std::make_shared<detail::handle<mk_box, mk_box_destroy>>(b);
        int64_t v = 0;
        detail::check(mk_box_size(b, &v));
        size_ = static_cast<idx>(v);
        detail::check(mk_box_n_sectors(b, &v));
        sectors_ = static_cast<idx>(v);
        for (idx k = 0; k < sectors_; ++k) {
            int64_t begin = 0, end = 0;
            array<int> totals(n_species(), 0);
            detail::check(mk_box_sector(b, static_cast<int64_t>(k), &begin, &end, totals.data()));
            begin_.push_back(static_cast<idx>(begin));
            totals_.push_back(std::move(totals));
        }
        begin_.push_back(size_);
    }
    [[nodiscard]] idx size() const { return size_; }
    [[nodiscard]] state operator[](idx i) const {
        state x(voxels_ * n_species(), 0);
        detail::check(mk_box_state(box_->h, static_cast<int64_t>(i), x.data()));
        return x;
    }
    [[nodiscard]] std::optional<idx> find(const state &x) const {
        if (x.size() != voxels_ * n_species()) return std::nullopt;
        int64_t i = -1;
        detail::check(mk_box_find(box_->h, x.data(), &i));
        return i < 0 ? std::nullopt : std::optional<idx>(static_cast<idx>(i));
    }
    [[nodiscard]] idx n_voxels() const { return voxels_; }
    [[nodiscard]] idx n_species() const { return max_totals_.size(); }
    [[nodiscard]] idx n_sectors() const { return sectors_; }
    [[nodiscard]] const array<int> &sector_totals(idx k) const { return totals_[k]; }
    [[nodiscard]] idx sector_begin(idx k) const { return begin_[k]; }
    [[nodiscard]] idx sector_end(idx k) const { return begin_[k + 1]; }
    [[nodiscard]] idx sector_of(idx i) const {
        return static_cast<idx>(std::upper_bound(begin_.begin(), begin_.end(), i) - begin_.begin() - 1);
    }
    [[nodiscard]] std::optional<idx> find_sector(const array<int> &totals) const {
        idx k = 0;
        for (idx s = 0; s < n_species(); ++s) {
            if (totals[s] < 0 || totals[s] > max_totals_[s]) return std::nullopt;
            k = k * (max_totals_[s] + 1) + totals[s];
        }
        return k;
    }
    [[nodiscard]] const array<int> &max_totals() const { return max_totals_; }
    [[nodiscard]] mk_box native() const { return box_->h; }

  private:
    idx voxels_;
    array<int> max_totals_;
    std::shared_ptr<detail::handle<mk_box, mk_box_destroy>> box_;
    idx size_ = 0, sectors_ = 0;
    array<idx> begin_;
    array<array<int>> totals_;
};

template <std::size_t D>
using grid = num::structured_grid_graph<D>;

// voxel_domain (reaction_diffusion.hpp:23-56)
template <typename Graph>
struct voxel_domain {
    Graph graph;
    array<real> volumes;
    array<real> diffusivities;

    voxel_domain(Graph g, array<real> voxel_volumes, array<real> species_diffusivities)
        : graph(std::move(g)), volumes(std::move(voxel_volumes)),
          diffusivities(std::move(species_diffusivities)) {
        if (volumes.size() != n_voxels())
            throw std::invalid_argument("voxel_domain: one volume per voxel");
        for (real v : volumes)
            if (!(v > 0.0) || !std::isfinite(v))
                throw std::invalid_argument("voxel_domain: volumes must be positive");
    }
    voxel_domain(Graph g, array<real> species_diffusivities)
        requires requires(const Graph &grid) { grid.cell_volume(); }
        : voxel_domain(g, array<real>(g.n_vertices(), static_cast<real>(g.cell_volume())),
                       std::move(species_diffusivities)) {}

    [[nodiscard]] idx n_voxels() const { return graph.n_vertices(); }
    [[nodiscard]] idx n_species() const { return diffusivities.size(); }
    template <typename Edge>
    [[nodiscard]] real jump_rate(idx species, idx from, const Edge &edge) const {
        return diffusivities[species] * static_cast<real>(edge.weight) / volumes[from];
    }

    // the library-side domain handle
    [[nodiscard]] std::shared_ptr<detail::handle<mk_domain, mk_domain_destroy>> native() const {
        mk_domain d = nullptr;
        if constexpr (requires(const Graph &g) { g.dimensions(); g.spacing(); }) {
            const auto &dims = graph.dimensions();
            const auto &h = graph.spacing();
            std::vector<int64_t> d64(dims.begin(), dims.end());
            std::vector<double> sp(h.begin(), h.end());
            detail::check(mk_domain_create_grid(static_cast<int32_t>(d64.size()), d64.data(),
                                                sp.data(), volumes.data(),
                                                static_cast<int32_t>(n_species()),
                                                diffusivities.data(), &d));
        } else {
            std::vector<int64_t> off{0}, to;
            std::vector<double> w;
            for (idx v = 0; v < n_voxels(); ++v) {
                for (const auto &e : graph.neighbors(static_cast<typename Graph::index_type>(v))) {
                    to.push_back(static_cast<int64_t>(e.to));
                    w.push_back(static_cast<double>(e.weight));
                }
                off.push_back(static_cast<int64_t>(to.size()));
            }
            detail::check(mk_domain_create_graph(static_cast<int64_t>(n_voxels()), off.data(),
                                                 to.data(), w.data(), volumes.data(),
                                                 static_cast<int32_t>(n_species()),
                                                 diffusivities.data(), &d));
        }
        return std::make_shared<detail::handle<mk_domain, mk_domain_destroy>>(d);
    }
};

template <typename Graph>
voxel_domain(Graph, array<real>, array<real>) -> voxel_domain<Graph>;
template <typename Graph>
voxel_domain(Graph, array<real>) -> voxel_domain<Graph>;

// local_reaction (reaction_diffusion.hpp:63-67)
struct local_reaction {
    state change;
    std::function<real(view<const int> counts, real volume)> rate;
};

// reaction_diffusion (reaction_diffusion.hpp:71-119)
template <typename Graph>
struct reaction_diffusion {
    voxel_domain<Graph> domain;
    array<local_reaction> reactions;
    [[nodiscard]] idx n_voxels() const { return domain.n_voxels(); }
    [[nodiscard]] idx n_species() const { return domain.n_species(); }
    [[nodiscard]] idx index(idx voxel, idx species) const { return voxel * n_species() + species; }
};

template <typename Graph>
reaction_diffusion(voxel_domain<Graph>, array<local_reaction>) -> reaction_diffusion<Graph>;

enum class boundary : std::uint8_t { reflecting, absorbing };

// generator (generator.hpp:19-112): assembled on the host, resident in HBM
class generator {
  public:
    using domain_type = vec<real>;
    using codomain_type = vec<real>;

    // one jump out of a state that stays in the space (generator.hpp:21-24)
    struct jump {
        idx to;
        real rate;
    };

    // over a population_box or a population_sector
    template <typename Space, typename Graph>
        requires requires(const Space &sp) { sp.native(); sp.max_totals(); }
    generator(const Space &space, const reaction_diffusion<Graph> &model,
              boundary leaks = boundary::absorbing) {
        const idx S = model.n_species(), V = model.n_voxels();
        if (space.n_species() != S || space.n_voxels() != V)
            throw std::invalid_argument("generator: the model does not match the box");
        // every local reaction's rate at every count tuple a voxel of the box can hold
        const auto &caps = space.max_totals();
        idx T = 1;
        for (int c : caps) T *= static_cast<idx>(c + 1);
        std::vector<int32_t> changes;
        std::vector<double> rates(model.reactions.size() * V * T);
        std::vector<int> counts(S);
        for (idx r = 0; r < model.reactions.size(); ++r) {
            const auto &reaction = model.reactions[r];
            if (reaction.change.size() != S)
                throw std::invalid_argument("local_reaction: one change per species");
            changes.insert(changes.end(), reaction.change.begin(), reaction.change.end());
            for (idx v = 0; v < V; ++v)
                for (idx t = 0; t < T; ++t) {
                    idx rem = t;
                    for (idx s = S; s-- > 0;) {
                        counts[s] = static_cast<int>(rem % (caps[s] + 1));
                        rem /= caps[s] + 1;
                    }
                    rates[(r * V + v) * T + t] =
                        reaction.rate(view<const int>(counts.data(), S), model.domain.volumes[v]);
                }
        }
        auto dom = model.domain.native();
        mk_generator g = nullptr;
        detail::check(mk_generator_create_rdme(space.native(), dom->h,
                                               static_cast<int32_t>(model.reactions.size()),
                                               changes.data(), rates.data(),
                                               leaks == boundary::absorbing ? MK_ABSORBING
                                                                            : MK_REFLECTING,
                                               &g));
        gen_ = std::make_shared<detail::handle<mk_generator, mk_generator_destroy>>(g);
        int64_t rows = 0, jumps = 0;
        detail::check(mk_generator_size(g, &rows, &jumps));
        rows_ = static_cast<idx>(rows);
        jumps_count_ = static_cast<idx>(jumps);
        exit_.resize(rows_);
        leak_.resize(rows_);
        detail::check(mk_generator_rates(g, exit_.data(), leak_.data()));
    }

    [[nodiscard]] idx rows() const { return rows_; }
    [[nodiscard]] idx cols() const { return rows_; }
    [[nodiscard]] idx n_jumps() const { return jumps_count_; }
    [[nodiscard]] const array<real> &exit_rates() const { return exit_; }
    [[nodiscard]] const array<real> &leak_rates() const { return leak_; }

    // y = Q x
    void apply(const vec<real> &x, vec<real> &y) const {
        detail::check(mk_generator_apply(gen_->h, x.data(), y.data()));
    }
    // y = shift x - Q x
    void apply_shifted(real shift, view<const real> x, view<real> y) const {
        detail::check(mk_shifted_apply(gen_->h, shift, x.data(), y.data()));
    }

    // the jumps out of state j that stay in the space (generator.hpp:51-53), read back from the
    // device on first use: targets ascending, jumps of j to one target merged (summed), so their
    // count can be below n_jumps() when two transitions of a state share a target
    [[nodiscard]] view<const jump> jumps(idx j) const {
        const auto &J = *host_jumps();
        return {J.list.data() + J.first[j], J.first[j + 1] - J.first[j]};
    }

    // Q, assembled on the host (generator.hpp:84-85)
    [[nodiscard]] num::spmat matrix() const { return assemble(0.0, 1.0); }
    // shift I - Q, assembled on the host (generator.hpp:87-89)
    [[nodiscard]] num::spmat characteristic_matrix(real shift) const { return assemble(shift, -1.0); }

    [[nodiscard]] mk_generator native() const { return gen_->h; }

  private:
    struct jump_lists {
        array<idx> first;
        array<jump> list;
        // gather form: row i lists (source, rate), sources ascending
        array<idx> ptr, col;
        array<real> val;
    };
    // source-major lists from the library's destination-major entries
    const jump_lists *host_jumps() const {
        std::call_once(*jumps_once_, [&] {
            auto J = std::make_shared<jump_lists>();
            int64_t entries = 0;
            detail::check(mk_generator_entries(gen_->h, &entries, nullptr, nullptr, nullptr));
            std::vector<int64_t> ptr(rows_ + 1), col(static_cast<std::size_t>(entries));
            J->val.resize(static_cast<std::size_t>(entries));
            detail::check(mk_generator_entries(gen_->h, &entries, ptr.data(), col.data(), J->val.data()));
            J->ptr.assign(ptr.begin(), ptr.end());
            J->col.assign(col.begin(), col.end());
            J->first.assign(rows_ + 1, 0);
            for (int64_t c : col) ++J->first[static_cast<idx>(c) + 1];
            for (idx j = 0; j < rows_; ++j) J->first[j + 1] += J->first[j];
            J->list.resize(J->col.size());
            array<idx> fill(J->first.begin(), J->first.end() - 1);
            for (idx i = 0; i < rows_; ++i)
                for (idx p = J->ptr[i]; p < J->ptr[i + 1]; ++p) J->list[fill[J->col[p]]++] = {i, J->val[p]};
            jumps_ = J;
        });
        return jumps_.get();
    }
    // sign (Q - shift I) in compressed rows: row i holds the diagonal and the jumps into i
    [[nodiscard]] num::spmat assemble(real shift, real sign) const {
        const auto &J = *host_jumps();
        array<idx> ptr(rows_ + 1, 0), col;
        array<real> val;
        col.reserve(J.col.size() + rows_);
        val.reserve(J.col.size() + rows_);
        for (idx i = 0; i < rows_; ++i) {
            bool diagonal = false;
            for (idx p = J.ptr[i]; p <= J.ptr[i + 1]; ++p) {
                if (!diagonal && (p == J.ptr[i + 1] || J.col[p] > i)) {
                    col.push_back(i);
                    val.push_back(-sign * (exit_[i] + shift));
                    diagonal = true;
                }
                if (p < J.ptr[i + 1]) {
                    col.push_back(J.col[p]);
                    val.push_back(sign * J.val[p]);
                }
            }
            ptr[i + 1] = col.size();
        }
        return num::spmat(rows_, rows_, std::move(ptr), std::move(col), std::move(val));
    }

    std::shared_ptr<detail::handle<mk_generator, mk_generator_destroy>> gen_;
    idx rows_ = 0, jumps_count_ = 0;
    array<real> exit_, leak_;
    mutable std::shared_ptr<const jump_lists> jumps_;
    std::shared_ptr<std::once_flag> jumps_once_ = std::make_shared<std::once_flag>();
};

// shifted_generator (generator.hpp:116-133): shift I - Q; it reads Q by reference
class shifted_generator {
  public:
    using domain_type = vec<real>;
    using codomain_type = vec<real>;
    shifted_generator(const generator &Q, real shift) : Q_(&Q), shift_(shift) {}
    [[nodiscard]] idx rows() const { return Q_->rows(); }
    [[nodiscard]] idx cols() const { return Q_->cols(); }
    [[nodiscard]] real shift() const { return shift_; }
    void apply(const vec<real> &x, vec<real> &y) const { Q_->apply_shifted(shift_, x, y); }
    // shift I - Q, assembled on the host (generator.hpp:128)
    [[nodiscard]] num::spmat matrix() const { return Q_->characteristic_matrix(shift_); }
    [[nodiscard]] const generator &base() const { return *Q_; }

  private:
    const generator *Q_;
    real shift_;
};

// jacobi and aggregate_block as at multigrid.hpp:120-123; red_black (multicolour Gauss-Seidel)
// and ilu0 (level-scheduled ILU(0) in rank order) are the device's extensions
enum class relaxation : std::uint8_t { jacobi, aggregate_block, red_black, ilu0 };

// precision of the device multigrid (no reference counterpart)
enum class precision : std::uint8_t { fp64, fp32 };

// multigrid_options (multigrid.hpp:125-134) plus the device working precision
struct multigrid_options {
    int pre_sweeps = 2;
    int post_sweeps = 2;
    real jacobi_weight = 2.0 / 3.0;
    idx coarsest_size = 4;
    int max_levels = 32;
    int coarse_cycles = 1;
    int sweep_growth = 0;
    relaxation smoother = relaxation::jacobi;
    precision working_precision = precision::fp64;
    int cycle_levels = 0; // extension: coarse_cycles on the top cycle_levels levels only
};

namespace detail {
inline mk_multigrid_options to_c(const multigrid_options &options) {
    mk_multigrid_options o;
    o.pre_sweeps = options.pre_sweeps;
    o.post_sweeps = options.post_sweeps;
    o.jacobi_weight = options.jacobi_weight;
    o.coarsest_size = static_cast<int64_t>(options.coarsest_size);
    o.max_levels = options.max_levels;
    o.coarse_cycles = options.coarse_cycles;
    o.sweep_growth = options.sweep_growth;
    o.smoother = options.smoother == relaxation::jacobi            ? MK_JACOBI
                 : options.smoother == relaxation::aggregate_block ? MK_AGGREGATE_BLOCK
                 : options.smoother == relaxation::red_black       ? MK_RED_BLACK
                                                                   : MK_ILU0;
    o.precision = options.working_precision == precision::fp32 ? MK_FP32 : MK_FP64;
    o.cycle_levels = options.cycle_levels;
    return o;
}
} // namespace detail

// coarsen (multigrid.hpp:24-42): every axis longer than one cell halved (its spacing doubled);
// nothing when an axis is odd or no axis is longer than one cell
template <std::size_t D>
std::optional<grid<D>> coarsen(const grid<D> &fine) {
    auto dims = fine.dimensions();
    auto spacing = fine.spacing();
    bool halved = false;
    for (std::size_t a = 0; a < D; ++a) {
        if (dims[a] == 1) continue;
        if (dims[a] % 2 != 0) return std::nullopt;
        dims[a] /= 2;
        spacing[a] *= 2.0;
        halved = true;
    }
    if (!halved) return std::nullopt;
    return grid<D>(dims, spacing);
}

// population_transfer (multigrid.hpp:47-118): the coarse state of every fine state and its
// multinomial prolongation weight, as the library's hierarchy holds them; restrict and prolong
// act on host vectors
class population_transfer {
  public:
    population_transfer(array<idx> parent, array<idx> coarse_of, array<real> weight, idx coarse_size)
        : parent_(std::move(parent)), coarse_of_(std::move(coarse_of)), weight_(std::move(weight)),
          coarse_size_(coarse_size) {}
    [[nodiscard]] idx fine_size() const { return coarse_of_.size(); }
    [[nodiscard]] idx coarse_size() const { return coarse_size_; }
    // the coarse voxel containing each fine voxel
    [[nodiscard]] const array<idx> &parent() const { return parent_; }
    [[nodiscard]] const array<idx> &coarse_of() const { return coarse_of_; }
    [[nodiscard]] const array<real> &weight() const { return weight_; }
    // coarse[c] = sum of fine[i] over the fine states of c
    void restrict(view<const real> fine, view<real> coarse) const {
        std::fill(coarse.begin(), coarse.end(), 0.0);
        for (idx i = 0; i < fine_size(); ++i) coarse[coarse_of_[i]] += fine[i];
    }
    // fine[i] = weight[i] coarse[coarse_of[i]]
    void prolong(view<const real> coarse, view<real> fine) const {
        for (idx i = 0; i < fine_size(); ++i) fine[i] = weight_[i] * coarse[coarse_of_[i]];
    }

  private:
    array<idx> parent_, coarse_of_;
    array<real> weight_;
    idx coarse_size_;
};

// population_multigrid (multigrid.hpp:139-236): the multigrid cycle of one population_sector,
// a sector multigrid of the library over a one-sector space. The cycle and relax run on the
// device; apply_system is the device shifted-generator product of the sector's diffusion
// generator; coarse_level() is the population_multigrid of the coarsened grid with doubled
// diffusivities, built on first use as the reference builds it.
template <std::size_t D>
class population_multigrid {
  public:
    using domain_type = vec<real>;
    using codomain_type = vec<real>;

    population_multigrid(const voxel_domain<grid<D>> &domain, const array<int> &totals, real shift,
                         multigrid_options options = {})
        : domain_(domain), space_(domain.n_voxels(), totals), shift_(shift), options_(options) {
        const auto o = detail::to_c(options);
        auto dom = domain.native();
        mk_sector_mg m = nullptr;
        detail::check(mk_sector_mg_create(dom->h, space_.native(), shift, &o, &m));
        mg_ = std::make_shared<detail::handle<mk_sector_mg, mk_sector_mg_destroy>>(m);
        int64_t rows = 0, sectors = 0, levels = 0;
        detail::check(mk_sector_mg_info(m, &rows, &sectors, &levels));
        for (int32_t l = 0; l < levels; ++l) {
            int64_t r = 0, j = 0;
            int32_t t = 0;
            detail::check(mk_sector_mg_level(m, l, &r, &j, &t));
            level_rows_.push_back(static_cast<idx>(r));
            level_jumps_.push_back(static_cast<idx>(j));
        }
    }

    [[nodiscard]] idx rows() const { return space_.size(); }
    [[nodiscard]] idx cols() const { return space_.size(); }
    [[nodiscard]] real shift() const { return shift_; }
    [[nodiscard]] const population_sector &space() const { return space_; }
    [[nodiscard]] idx n_levels() const { return level_rows_.size(); }
    [[nodiscard]] real grid_complexity() const {
        return real(std::accumulate(level_rows_.begin(), level_rows_.end(), idx{0})) / real(rows());
    }
    [[nodiscard]] real operator_complexity() const {
        const idx total = std::accumulate(level_rows_.begin(), level_rows_.end(), idx{0}) +
                          std::accumulate(level_jumps_.begin(), level_jumps_.end(), idx{0});
        return real(total) / real(rows() + level_jumps_[0]);
    }

    // the sector's diffusion generator (reactions are not part of the preconditioner)
    [[nodiscard]] const generator &fine_generator() const {
        std::call_once(lazy_->generator_once, [&] {
            lazy_->Q = std::make_unique<generator>(space_, reaction_diffusion<grid<D>>{domain_, {}});
        });
        return *lazy_->Q;
    }
    // y = A x with A = shift I - Q on this level
    void apply_system(view<const real> x, view<real> y) const { fine_generator().apply_shifted(shift_, x, y); }
    // one cycle from a zero initial guess: x ~ A^{-1} b
    void apply(const vec<real> &b, vec<real> &x) const { cycle(b, x); }
    void cycle(view<const real> b, view<real> x) const {
        if (b.size() != rows() || x.size() != rows())
            throw std::invalid_argument("population_multigrid: vector sizes do not match");
        detail::check(mk_sector_mg_apply(mg_->h, b.data(), x.data()));
    }
    // `sweeps` smoother sweeps on A x = rhs, starting from x
    void relax(const vec<real> &rhs, vec<real> &x, int sweeps = 1, bool reverse = false) const {
        if (rhs.size() != rows() || x.size() != rows())
            throw std::invalid_argument("population_multigrid: vector sizes do not match");
        detail::check(mk_sector_mg_relax(mg_->h, rhs.data(), x.data(), sweeps, reverse ? 1 : 0));
    }

    // the transfer to the coarse level (none on the last level)
    [[nodiscard]] const population_transfer &transfer() const {
        if (n_levels() < 2) throw std::logic_error("population_multigrid: the last level has no transfer");
        std::call_once(lazy_->transfer_once, [&] {
            std::vector<int64_t> co(rows());
            array<real> w(rows());
            detail::check(mk_sector_mg_transfer(mg_->h, 0, co.data(), w.data()));
            const auto coarse = coarsen<D>(domain_.graph);
            array<idx> parent(domain_.n_voxels());
            for (idx v = 0; v < parent.size(); ++v) {
                auto c = domain_.graph.coordinate(v);
                for (std::size_t a = 0; a < D; ++a)
                    if (domain_.graph.dimensions()[a] > 1) c[a] /= 2;
                parent[v] = coarse->vertex(c);
            }
            lazy_->transfer = std::make_unique<population_transfer>(
                std::move(parent), array<idx>(co.begin(), co.end()), std::move(w), level_rows_[1]);
        });
        return *lazy_->transfer;
    }
    // the next coarser level, or null on the last level
    [[nodiscard]] const population_multigrid *coarse_level() const {
        if (n_levels() < 2) return nullptr;
        std::call_once(lazy_->coarse_once, [&] {
            array<real> doubled = domain_.diffusivities;
            for (real &d : doubled) d *= 2.0;
            multigrid_options co = options_;
            co.max_levels -= 1;
            co.pre_sweeps += options_.sweep_growth;
            co.post_sweeps += options_.sweep_growth;
            // the top cycle_levels levels take coarse_cycles visits: one fewer below this one,
            // and single visits once none is left
            if (options_.cycle_levels > 0) {
                co.cycle_levels = options_.cycle_levels - 1;
                if (co.cycle_levels == 0) co.coarse_cycles = 1;
            }
            lazy_->coarse = std::make_unique<population_multigrid>(
                voxel_domain<grid<D>>(*coarsen<D>(domain_.graph), std::move(doubled)), space_.totals(),
                shift_, co);
        });
        return lazy_->coarse.get();
    }
    [[nodiscard]] mk_sector_mg native() const { return mg_->h; }

  private:
    struct lazy_parts {
        std::once_flag generator_once, transfer_once, coarse_once;
        std::unique_ptr<generator> Q;
        std::unique_ptr<population_transfer> transfer;
        std::unique_ptr<population_multigrid> coarse;
    };
    voxel_domain<grid<D>> domain_;
    population_sector space_;
    real shift_;
    multigrid_options options_;
    std::shared_ptr<detail::handle<mk_sector_mg, mk_sector_mg_destroy>> mg_;
    array<idx> level_rows_, level_jumps_;
    std::shared_ptr<lazy_parts> lazy_ = std::make_shared<lazy_parts>();
};

template <std::size_t D>
population_multigrid(voxel_domain<grid<D>>, array<int>, real) -> population_multigrid<D>;
template <std::size_t D>
population_multigrid(voxel_domain<grid<D>>, array<int>, real, multigrid_options)
    -> population_multigrid<D>;

// sector_multigrid (multigrid.hpp:371-408)
template <std::size_t D>
class sector_multigrid {
  public:
    using domain_type = vec<real>;
    using codomain_type = vec<real>;

    sector_multigrid(const voxel_domain<grid<D>> &domain, const population_box &box, real shift,
                     multigrid_options options = {})
        : shift_(shift) {
        const auto o = detail::to_c(options);
        auto dom = domain.native();
        mk_sector_mg m = nullptr;
        detail::check(mk_sector_mg_create(dom->h, box.native(), shift, &o, &m));
        mg_ = std::make_shared<detail::handle<mk_sector_mg, mk_sector_mg_destroy>>(m);
        int64_t rows = 0, sectors = 0, levels = 0;
        detail::check(mk_sector_mg_info(m, &rows, &sectors, &levels));
        rows_ = static_cast<idx>(rows);
        sectors_ = static_cast<idx>(sectors);
        levels_ = static_cast<idx>(levels);
    }
    [[nodiscard]] idx rows() const { return rows_; }
    [[nodiscard]] idx cols() const { return rows_; }
    [[nodiscard]] idx n_sectors() const { return sectors_; }
    [[nodiscard]] idx max_levels() const { return levels_; }
    [[nodiscard]] real shift() const { return shift_; }
    void apply(const vec<real> &b, vec<real> &x) const {
        detail::check(mk_sector_mg_apply(mg_->h, b.data(), x.data()));
    }
    [[nodiscard]] mk_sector_mg native() const { return mg_->h; }

  private:
    std::shared_ptr<detail::handle<mk_sector_mg, mk_sector_mg_destroy>> mg_;
    real shift_;
    idx rows_ = 0, sectors_ = 0, levels_ = 0;
};

template <std::size_t D>
sector_multigrid(voxel_domain<grid<D>>, const population_box &, real) -> sector_multigrid<D>;
template <std::size_t D>
sector_multigrid(voxel_domain<grid<D>>, const population_box &, real, multigrid_options)
    -> sector_multigrid<D>;

} // namespace markovkit

namespace num {

namespace detail {
inline solver_result run_gmres(const markovkit::shifted_generator &A, mk_sector_mg M,
                               const vec<real> &b, vec<real> &x, const gmres_options &options) {
    if (b.size() != A.rows() || x.size() != A.rows())
        throw std::invalid_argument("gmres: vector sizes do not match the operator");
    mk_gmres_options o{options.tolerance, static_cast<int64_t>(options.max_iterations),
                       static_cast<int64_t>(options.restart), options.orthogonalization};
    mk_solver_result r{};
    markovkit::detail::check(
        mk_gmres(A.base().native(), A.shift(), M, b.data(), x.data(), &o, &r));
    return {static_cast<idx>(r.iterations), r.residual, r.converged != 0};
}
} // namespace detail

// gmres(A, b, x, options): plain restarted GMRES on shift I - Q; x is the initial guess
inline solver_result gmres(const markovkit::shifted_generator &A, const vec<real> &b, vec<real> &x,
                           const gmres_options &options) {
    return detail::run_gmres(A, nullptr, b, x, options);
}

// gmres(A, M, b, x, options): flexible GMRES preconditioned by the sector multigrid
template <std::size_t D>
solver_result gmres(const markovkit::shifted_generator &A,
                    const markovkit::sector_multigrid<D> &M, const vec<real> &b, vec<real> &x,
                    const gmres_options &options) {
    return detail::run_gmres(A, M.native(), b, x, options);
}

} // namespace num
