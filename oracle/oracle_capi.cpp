// ORACLE C API — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
// cpu_baseline leg). Plain-pointer entry points over mk_oracle.hpp so Python can drive the
// FP64 CPU restatement of the reference path on the same inputs as the HIP library.
//
// Local reaction rates are passed as tables: rates[(r * V + v) * T + t] is reaction r's rate
// in voxel v when that voxel holds the count tuple t = sum_s n_s * stride_s (species 0 most
// significant, radix max_totals[s] + 1). Every state of a population_box holds at most
// max_totals[s] molecules of species s in a voxel, so the table covers every evaluation the
// reference's std::function rates receive (reaction_diffusion.hpp:99-116). A reference rate
// sees only (counts, volume), so the oracle's closures look the table up by volume: the
// first voxel with that volume supplies the row.
#include "mk_oracle.hpp"
#include "box_mg.hpp"

#include <cstring>
#include <exception>
#include <string>

using namespace mk_oracle;

namespace {

thread_local std::string last_error;

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument &e) {
        last_error = e.what();
        return 1;
    } catch (const std::overflow_error &e) {
        last_error = e.what();
        return 2;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 3;
    }
}

struct problem {
    std::unique_ptr<voxel_domain> domain;
    std::unique_ptr<population_box> box;
    std::unique_ptr<reaction_diffusion> model;
    std::unique_ptr<generator> Q;
};

voxel_domain make_domain(int dim, const int64_t *dims, const double *spacing, int64_t n_voxels,
                         const int64_t *adj_offsets, const int64_t *adj_to,
                         const double *adj_weight, const double *volumes, int n_species,
                         const double *diffusivities) {
    std::vector<real> diff(diffusivities, diffusivities + n_species);
    if (dim > 0) {
        std::vector<idx> d(dims, dims + dim);
        std::vector<real> h(spacing, spacing + dim);
        grid g(d, h);
        if (volumes) return voxel_domain(g, std::vector<real>(volumes, volumes + g.n_vertices()), diff);
        return voxel_domain(g, diff);
    }
    std::vector<std::vector<edge>> adj(n_voxels);
    for (int64_t v = 0; v < n_voxels; ++v)
        for (int64_t e = adj_offsets[v]; e < adj_offsets[v + 1]; ++e)
            adj[v].push_back({static_cast<idx>(adj_to[e]), adj_weight[e]});
    return voxel_domain(std::move(adj), std::vector<real>(volumes, volumes + n_voxels), diff);
}

struct rate_table {
    std::vector<real> values;   // [voxel][tuple]
    std::vector<real> distinct; // distinct volumes, in first-seen order
    std::vector<idx> first;     // the first voxel with each distinct volume
    std::vector<idx> stride;
    idx T = 1;
    void set_volumes(const std::vector<real> &volumes) {
        for (idx v = 0; v < volumes.size(); ++v)
            if (std::find(distinct.begin(), distinct.end(), volumes[v]) == distinct.end()) {
                distinct.push_back(volumes[v]);
                first.push_back(v);
            }
    }
    real operator()(const int *counts, real volume) const {
        idx d = 0;
        while (d < distinct.size() && distinct[d] != volume) ++d;
        if (d == distinct.size()) throw std::invalid_argument("oracle: rate asked at an unknown volume");
        const idx v = first[d];
        idx t = 0;
        for (idx s = 0; s < stride.size(); ++s) t += static_cast<idx>(counts[s]) * stride[s];
        return values[v * T + t];
    }
};

} // namespace

extern "C" {

const char *orc_last_error() { return last_error.c_str(); }

// dim > 0: structured grid (dims, spacing, optional volumes); dim == 0: explicit voxel graph
// (CSR adjacency, volumes). Builds the box, the model and its generator.
int orc_create(int dim, const int64_t *dims, const double *spacing, int64_t n_voxels,
               const int64_t *adj_offsets, const int64_t *adj_to, const double *adj_weight,
               const double *volumes, int n_species, const double *diffusivities,
               const int *max_totals, int n_reactions, const int *changes, const double *rates,
               int absorbing, void **out) {
    return guarded([&] {
        auto p = std::make_unique<problem>();
        p->domain = std::make_unique<voxel_domain>(make_domain(dim, dims, spacing, n_voxels,
                                                               adj_offsets, adj_to, adj_weight,
                                                               volumes, n_species, diffusivities));
        const idx V = p->domain->n_voxels();
        std::vector<int> caps(max_totals, max_totals + n_species);
        p->box = std::make_unique<population_box>(V, caps);
        idx T = 1;
        std::vector<idx> stride(n_species, 1);
        for (int s = n_species; s-- > 0;) {
            stride[s] = T;
            T *= static_cast<idx>(caps[s] + 1);
        }
        p->model = std::make_unique<reaction_diffusion>(reaction_diffusion{*p->domain, {}});
        for (int r = 0; r < n_reactions; ++r) {
            rate_table table;
            table.values.assign(rates + static_cast<idx>(r) * V * T,
                                rates + static_cast<idx>(r + 1) * V * T);
            table.set_volumes(p->domain->volumes);
            table.stride = stride;
            table.T = T;
            std::vector<int> change(changes + r * n_species, changes + (r + 1) * n_species);
            p->model->reactions.push_back({change, table});
        }
        p->Q = std::make_unique<generator>(*p->box, *p->model,
                                           absorbing ? boundary::absorbing : boundary::reflecting);
        *out = p.release();
    });
}

void orc_destroy(void *h) { delete static_cast<problem *>(h); }

int64_t orc_size(void *h) { return static_cast<int64_t>(static_cast<problem *>(h)->box->size()); }
int64_t orc_n_sectors(void *h) {
    return static_cast<int64_t>(static_cast<problem *>(h)->box->n_sectors());
}
int64_t orc_n_jumps(void *h) { return static_cast<int64_t>(static_cast<problem *>(h)->Q->n_jumps()); }

void orc_sector_begins(void *h, int64_t *out) {
    const auto &box = *static_cast<problem *>(h)->box;
    for (idx k = 0; k <= box.n_sectors(); ++k)
        out[k] = static_cast<int64_t>(k < box.n_sectors() ? box.sector_begin(k) : box.size());
}

void orc_states(void *h, int *out) {
    const auto &box = *static_cast<problem *>(h)->box;
    for (idx i = 0; i < box.size(); ++i) std::memcpy(out + i * box.width(), box[i], box.width() * sizeof(int));
}

int64_t orc_find(void *h, const int *x) {
    const auto &box = *static_cast<problem *>(h)->box;
    auto i = box.find(x, box.width());
    return i ? static_cast<int64_t>(*i) : -1;
}

void orc_rates(void *h, double *exit, double *leak) {
    const auto &Q = *static_cast<problem *>(h)->Q;
    std::copy(Q.exit_rates().begin(), Q.exit_rates().end(), exit);
    std::copy(Q.leak_rates().begin(), Q.leak_rates().end(), leak);
}

// source-major jump lists: first[n+1], to[n_jumps], rate[n_jumps]
void orc_jumps(void *h, int64_t *first, int64_t *to, double *rate) {
    const auto &Q = *static_cast<problem *>(h)->Q;
    for (idx j = 0; j <= Q.rows(); ++j) first[j] = static_cast<int64_t>(Q.first()[j]);
    const auto &J = Q.all_jumps();
    for (idx k = 0; k < J.size(); ++k) {
        to[k] = static_cast<int64_t>(J[k].to);
        rate[k] = J[k].rate;
    }
}

void orc_apply(void *h, const double *x, double *y) { static_cast<problem *>(h)->Q->apply(x, y); }

void orc_apply_shifted(void *h, double shift, const double *x, double *y) {
    static_cast<problem *>(h)->Q->apply_shifted(shift, x, y);
}

int orc_mg_create(void *h, double shift, int pre, int post, double omega, int64_t coarsest,
                  int max_levels, int coarse_cycles, int sweep_growth, int smoother, int cycle_levels,
                  void **out) {
    return guarded([&] {
        auto *p = static_cast<problem *>(h);
        multigrid_options o;
        if (smoother < 0 || smoother > 3) throw std::invalid_argument("sector_multigrid: unknown smoother");
        o.smoother = static_cast<relaxation>(smoother);
        o.pre_sweeps = pre;
        o.post_sweeps = post;
        o.jacobi_weight = omega;
        o.coarsest_size = static_cast<idx>(coarsest);
        o.max_levels = max_levels;
        o.coarse_cycles = coarse_cycles;
        o.sweep_growth = sweep_growth;
        if (cycle_levels < 0) throw std::invalid_argument("sector_multigrid: negative cycle_levels");
        o.cycle_levels = cycle_levels;
        *out = new sector_multigrid(*p->domain, *p->box, shift, o);
    });
}

void orc_mg_destroy(void *mg) { delete static_cast<sector_multigrid *>(mg); }

int orc_mg_relax(void *mg, int64_t k, const double *rhs, double *x, int sweeps, int reverse) {
    return guarded([&] {
        auto *m = static_cast<sector_multigrid *>(mg);
        if (k < 0 || k >= static_cast<int64_t>(m->n_sectors())) throw std::invalid_argument("sector out of range");
        m->sector(k).relax(rhs, x, sweeps, reverse != 0);
    });
}

// sector k's finest level: red-black colours (red_black) or ILU(0) pivots (ilu0); -1 when the
// hierarchy has no such smoother
int orc_mg_colours(void *mg, int64_t k, uint8_t *out) {
    return guarded([&] {
        const auto &c = static_cast<sector_multigrid *>(mg)->sector(k).colours();
        if (c.empty()) throw std::invalid_argument("no red-black colours");
        std::copy(c.begin(), c.end(), out);
    });
}

int orc_mg_ilu_pivots(void *mg, int64_t k, double *out) {
    return guarded([&] {
        const auto &p = static_cast<sector_multigrid *>(mg)->sector(k).ilu_pivots();
        if (p.empty()) throw std::invalid_argument("no ILU(0) pivots");
        std::copy(p.begin(), p.end(), out);
    });
}

int64_t orc_mg_max_levels(void *mg) {
    return static_cast<int64_t>(static_cast<sector_multigrid *>(mg)->max_levels());
}

void orc_mg_apply(void *mg, const double *b, double *x) {
    static_cast<sector_multigrid *>(mg)->apply(b, x);
}

// per-sector level structure: out[k * max_levels + l] = rows of sector k on level l (0 past
// its last level)
void orc_mg_level_rows(void *mg, int64_t *out) {
    const auto &M = *static_cast<sector_multigrid *>(mg);
    const idx L = M.max_levels();
    for (idx k = 0; k < M.n_sectors(); ++k) {
        const population_multigrid *lvl = &M.sector(k);
        for (idx l = 0; l < L; ++l) {
            out[k * L + l] = lvl ? static_cast<int64_t>(lvl->rows()) : 0;
            lvl = lvl ? lvl->coarse_level() : nullptr;
        }
    }
}

// the fine-to-coarse transfer of sector k on level l: coarse_of[fine], weight[fine]
int orc_mg_transfer(void *mg, int64_t k, int64_t level, int64_t *coarse_of, double *weight) {
    return guarded([&] {
        const population_multigrid *lvl = &static_cast<sector_multigrid *>(mg)->sector(k);
        for (int64_t l = 0; l < level && lvl; ++l) lvl = lvl->coarse_level();
        if (!lvl || !lvl->coarse_level()) throw std::invalid_argument("no transfer on that level");
        const auto &t = lvl->transfer();
        for (idx i = 0; i < t.fine_size(); ++i) {
            coarse_of[i] = static_cast<int64_t>(t.coarse_of()[i]);
            weight[i] = t.weight()[i];
        }
    });
}

// y = A_l x for sector k on level l (A_l = shift I - Q_diffusion on that level)
int orc_mg_level_apply(void *mg, int64_t k, int64_t level, const double *x, double *y) {
    return guarded([&] {
        const population_multigrid *lvl = &static_cast<sector_multigrid *>(mg)->sector(k);
        for (int64_t l = 0; l < level && lvl; ++l) lvl = lvl->coarse_level();
        if (!lvl) throw std::invalid_argument("no such level");
        lvl->apply_system(x, y);
    });
}

// restarted GMRES on (shift I - Q) x = b, preconditioned by mg when not null
int orc_gmres(void *h, void *mg, double shift, const double *b, double *x, double tol,
              int64_t max_it, int64_t restart, int64_t *iterations, double *residual) {
    auto *p = static_cast<problem *>(h);
    auto *M = static_cast<sector_multigrid *>(mg);
    gmres_options o;
    o.tolerance = tol;
    o.max_iterations = static_cast<idx>(max_it);
    o.restart = static_cast<idx>(restart);
    const idx n = p->box->size();
    auto A = [&](const real *in, real *out) { p->Q->apply_shifted(shift, in, out); };
    solver_result r;
    if (M) {
        r = gmres(n, A, [&](const real *v, real *z) { M->apply(v, z); }, b, x, o);
    } else {
        r = gmres(n, A, [&](const real *v, real *z) { std::copy(v, v + n, z); }, b, x, o);
    }
    *iterations = static_cast<int64_t>(r.iterations);
    *residual = r.residual;
    return r.converged ? 1 : 0;
}

// w = exp(t Q) v for the generator given by source-major jump lists (first, to, rate) and
// exit rates: (Q x)_i = -exit_i x_i + sum over jumps j -> i of rate x_j. The norm estimate is
// ||Q||_1 = max_j (exit_j + sum of j's in-space jump rates).
int orc_expv_jumps(int64_t n, const int64_t *first, const int64_t *to, const double *rate,
                   const double *exit, double t, const double *v, double *w, int64_t krylov_dim,
                   double tol, int64_t *steps, int64_t *rejected) {
    return guarded([&] {
        real anorm = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            real c = exit[j];
            for (int64_t p = first[j]; p < first[j + 1]; ++p) c += rate[p];
            anorm = std::max(anorm, c);
        }
        auto apply = [&](const real *x, real *y) {
            for (int64_t i = 0; i < n; ++i) y[i] = -exit[i] * x[i];
            for (int64_t j = 0; j < n; ++j)
                for (int64_t p = first[j]; p < first[j + 1]; ++p) y[to[p]] += rate[p] * x[j];
        };
        expv_options o;
        o.krylov_dim = static_cast<idx>(krylov_dim);
        o.tolerance = tol;
        const auto r = expv(t, static_cast<idx>(n), apply, anorm > 0.0 ? anorm : 1.0, v, w, o);
        if (steps) *steps = static_cast<int64_t>(r.steps);
        if (rejected) *rejected = static_cast<int64_t>(r.rejected);
    });
}

// ---- whole-box multigrid of the paper driver (box_mg.hpp)
int orc_box_create(int dimension, int cells, int cap, double diffusion, double shift, double c1,
                   double c2, double c3, double c4, int smoother, int sweeps, int cycles,
                   int kcycle, int64_t terminal_rows, void **out) {
    return guarded([&] {
        orc_box::params p;
        p.dimension = dimension;
        p.cells = cells;
        p.cap = cap;
        p.diffusion = diffusion;
        p.shift = shift;
        p.c1 = c1;
        p.c2 = c2;
        p.c3 = c3;
        p.c4 = c4;
        if (smoother < 0 || smoother > 3) throw std::invalid_argument("oracle box: unknown smoother");
        p.smoother = static_cast<orc_box::smoother_kind>(smoother);
        p.sweeps = sweeps;
        p.cycles = cycles;
        p.kcycle = kcycle != 0;
        p.terminal_rows = static_cast<long>(terminal_rows);
        if (dimension < 1 || dimension > 3 || cells < 1 || (cells & (cells - 1)) || cap < 0 ||
            !(shift > 0.0) || sweeps < 1 || cycles < 1 || cycles > 2)
            throw std::invalid_argument("oracle box: invalid parameters");
        *out = new orc_box::level(p);
    });
}

void orc_box_destroy(void *h) { delete static_cast<orc_box::level *>(h); }

namespace {
const orc_box::level &box_at(void *h, int64_t l) {
    return *static_cast<orc_box::level *>(h)->at(static_cast<int>(l));
}
} // namespace

int64_t orc_box_levels(void *h) { return static_cast<orc_box::level *>(h)->n_levels(); }
int64_t orc_box_rows(void *h, int64_t l) { return box_at(h, l).rows(); }
int64_t orc_box_nnz(void *h, int64_t l) { return static_cast<int64_t>(box_at(h, l).A.val.size()); }
int64_t orc_box_voxels(void *h, int64_t l) { return box_at(h, l).S.V; }
int orc_box_terminal(void *h, int64_t l) { return box_at(h, l).terminal; }

int64_t orc_box_rank(void *h, const int *n) {
    const auto &L = box_at(h, 0);
    std::vector<int> x(n, n + L.S.V);
    int total = 0;
    for (int v : x) total += v;
    if (total > L.S.cap) return -1;
    return L.S.rank(x, total);
}

// the driver's initial state: `initial` molecules dealt round-robin over the voxels
int64_t orc_box_initial(void *h, int initial) {
    const auto &L = box_at(h, 0);
    std::vector<int> x(L.S.V, 0);
    for (int k = 0; k < initial; ++k) ++x[k % L.S.V];
    if (initial > L.S.cap) return -1;
    return L.S.rank(x, initial);
}

void orc_box_matrix(void *h, int64_t l, int64_t *ptr, int64_t *col, double *val) {
    const auto &A = box_at(h, l).A;
    for (std::size_t i = 0; i < A.ptr.size(); ++i) ptr[i] = A.ptr[i];
    for (std::size_t j = 0; j < A.col.size(); ++j) {
        col[j] = A.col[j];
        val[j] = A.val[j];
    }
}

void orc_box_leak(void *h, int64_t l, double *out) {
    const auto &L = box_at(h, l);
    std::copy(L.leak.begin(), L.leak.end(), out);
}

void orc_box_colours(void *h, int64_t l, uint8_t *out) {
    const auto &L = box_at(h, l);
    std::copy(L.colour.begin(), L.colour.end(), out);
}

int orc_box_transfer(void *h, int64_t l, int64_t *parent, double *weight) {
    return guarded([&] {
        const auto &L = box_at(h, l);
        if (!L.coarse) throw std::invalid_argument("oracle box: the level has no coarse level");
        for (std::size_t i = 0; i < L.parent.size(); ++i) {
            parent[i] = L.parent[i];
            weight[i] = L.weight[i];
        }
    });
}

// x <- the level's smoother (all its sweeps) on A x = b; reverse = post-smoothing order
int orc_box_smooth(void *h, int64_t l, const double *b, double *x, int reverse) {
    return guarded([&] {
        const auto &L = box_at(h, l);
        if (L.terminal) throw std::invalid_argument("oracle box: a terminal level has no smoother");
        orc_box::Vec bb(b, b + L.rows()), xx(x, x + L.rows());
        L.smooth(bb, xx, reverse != 0);
        std::copy(xx.begin(), xx.end(), x);
    });
}

// x <- one cycle of A x = b from the guess x
void orc_box_cycle(void *h, const double *b, double *x) {
    const auto &L = box_at(h, 0);
    orc_box::Vec bb(b, b + L.rows()), xx(x, x + L.rows());
    L.cycle(bb, xx);
    std::copy(xx.begin(), xx.end(), x);
}

// the standalone iteration from the guess x; the error history (one entry per loop test) goes
// to history[0 .. returned count) when history_cap allows
int64_t orc_box_solve(void *h, const double *b, double *x, double tolerance, int max_iterations,
                      double *history, int64_t history_cap, int *iterations, double *error) {
    const auto &L = box_at(h, 0);
    orc_box::Vec bb(b, b + L.rows()), xx(x, x + L.rows()), hist;
    const auto r = orc_box::standalone(L, bb, xx, tolerance, max_iterations, &hist);
    std::copy(xx.begin(), xx.end(), x);
    if (iterations) *iterations = r.iterations;
    if (error) *error = r.error;
    const int64_t m = std::min<int64_t>(history_cap, static_cast<int64_t>(hist.size()));
    for (int64_t k = 0; k < m; ++k) history[k] = hist[k];
    return static_cast<int64_t>(hist.size());
}

} // extern "C"
