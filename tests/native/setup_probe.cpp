// Test-only probe over the HIP library's host setup code (markovkit_amd/csrc/setup.cpp):
// exports the state space, the gather-form generator and the multigrid hierarchy as plain
// arrays so CPU tests can compare them with the oracle without a GPU.
#include "../../markovkit_amd/csrc/setup.hpp"

#include <cstring>
#include <exception>
#include <memory>
#include <string>

using namespace mk;

namespace {
thread_local std::string err;
template <typename F>
int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument &e) {
        err = e.what();
        return 1;
    } catch (const std::overflow_error &e) {
        err = e.what();
        return 2;
    } catch (const std::exception &e) {
        err = e.what();
        return 3;
    }
}

voxel_domain make_domain(int dim, const i64 *dims, const double *spacing, i64 V, const i64 *off,
                         const i64 *to, const double *w, const double *vols, int S,
                         const double *diff) {
    std::vector<double> d(diff, diff + S);
    if (dim > 0)
        return voxel_domain::grid(std::vector<i64>(dims, dims + dim),
                                  std::vector<double>(spacing, spacing + dim), vols, d);
    return voxel_domain::graph(V, off, to, w, vols, d);
}

struct gen_result {
    generator_csr g;
};
struct mg_result {
    std::vector<sector_hierarchy> H;
};
} // namespace

extern "C" {

const char *probe_error() { return err.c_str(); }

int probe_box(i64 V, int S, const int *caps, i64 *size, i64 *n_sectors) {
    return guard([&] {
        auto p = population_space::box(V, std::vector<int>(caps, caps + S));
        *size = p.size();
        *n_sectors = p.n_sectors();
    });
}

// states[i * V * S ...] for every state, plus begin[k] for k <= n_sectors
int probe_box_states(i64 V, int S, const int *caps, int *states, i64 *begin, i64 *found) {
    return guard([&] {
        auto p = population_space::box(V, std::vector<int>(caps, caps + S));
        for (i64 k = 0; k <= p.n_sectors(); ++k) begin[k] = k < p.n_sectors() ? p.begin(k) : p.size();
        for (i64 i = 0; i < p.size(); ++i) {
            p.state(i, states + i * p.width());
            found[i] = p.find(states + i * p.width());
        }
    });
}

int probe_generator_level(int dim, const i64 *dims, const double *spacing, const double *vols,
                          int S, const double *diff, const int *totals, void **out) {
    return guard([&] {
        auto dom = make_domain(dim, dims, spacing, 0, nullptr, nullptr, nullptr, vols, S, diff);
        auto space = population_space::sector(dom.voxels, std::vector<int>(totals, totals + S));
        auto r = std::make_unique<gen_result>();
        r->g = assemble_generator(space, dom, reaction_tables{}, true);
        *out = r.release();
    });
}

int probe_generator(int dim, const i64 *dims, const double *spacing, i64 V, const i64 *off,
                    const i64 *to, const double *w, const double *vols, int S, const double *diff,
                    const int *caps, int R, const int *changes, const double *rates, int absorbing,
                    void **out) {
    return guard([&] {
        auto dom = make_domain(dim, dims, spacing, V, off, to, w, vols, S, diff);
        auto space = population_space::box(dom.voxels, std::vector<int>(caps, caps + S));
        reaction_tables rt;
        rt.n = R;
        rt.stride.assign(S, 1);
        for (int s = S; s-- > 0;) {
            rt.stride[s] = rt.T;
            rt.T *= caps[s] + 1;
        }
        if (R > 0) {
            rt.changes.assign(changes, changes + static_cast<i64>(R) * S);
            rt.rates.assign(rates, rates + static_cast<i64>(R) * dom.voxels * rt.T);
        }
        auto r = std::make_unique<gen_result>();
        r->g = assemble_generator(space, dom, rt, absorbing != 0);
        *out = r.release();
    });
}

void probe_generator_sizes(void *h, i64 *n, i64 *nnz, i64 *n_jumps) {
    auto &g = static_cast<gen_result *>(h)->g;
    *n = g.n;
    *nnz = static_cast<i64>(g.col.size());
    *n_jumps = g.n_jumps;
}

void probe_generator_arrays(void *h, i64 *ptr, i64 *col, double *val, double *exit, double *leak) {
    auto &g = static_cast<gen_result *>(h)->g;
    for (i64 i = 0; i <= g.n; ++i) ptr[i] = static_cast<i64>(g.ptr[i]);
    for (std::size_t k = 0; k < g.col.size(); ++k) {
        col[k] = g.col[k];
        val[k] = g.val[k];
    }
    std::memcpy(exit, g.exit.data(), g.n * sizeof(double));
    std::memcpy(leak, g.leak.data(), g.n * sizeof(double));
}

void probe_generator_free(void *h) { delete static_cast<gen_result *>(h); }

// wrap a CSR given by the caller (rows may differ from the column range)
void *probe_csr(i64 n, const i64 *ptr, const i64 *col, const double *val, const double *exit) {
    auto r = std::make_unique<gen_result>();
    r->g.n = n;
    r->g.ptr.assign(ptr, ptr + n + 1);
    r->g.col.assign(col, col + ptr[n]);
    r->g.val.assign(val, val + ptr[n]);
    r->g.exit.assign(exit, exit + n);
    r->g.leak.assign(n, 0.0);
    return r.release();
}

// the coded sliced layout of the generator (FP64 dictionaries); returns 0 when it does not fit
int probe_sliced_sizes(void *h, i64 *n_slices, i64 *padded, i64 *n_vdict, i64 *n_edict) {
    auto &g = static_cast<gen_result *>(h)->g;
    sliced_operator<double> so;
    if (!encode_sliced<double>(g, so)) return 0;
    *n_slices = static_cast<i64>(so.soff.size()) - 1;
    *padded = static_cast<i64>(so.padded);
    *n_vdict = static_cast<i64>(so.vdict.size());
    *n_edict = static_cast<i64>(so.edict.size());
    return 1;
}

void probe_sliced_arrays(void *h, u32 *soff, u32 *ent, unsigned char *rlen, unsigned char *rcode,
                         double *vdict, double *edict) {
    auto &g = static_cast<gen_result *>(h)->g;
    sliced_operator<double> so;
    encode_sliced<double>(g, so);
    std::memcpy(soff, so.soff.data(), so.soff.size() * sizeof(u32));
    std::memcpy(ent, so.ent.data(), so.ent.size() * sizeof(u32));
    std::memcpy(rlen, so.rlen.data(), so.rlen.size());
    std::memcpy(rcode, so.rcode.data(), so.rcode.size());
    std::memcpy(vdict, so.vdict.data(), so.vdict.size() * sizeof(double));
    std::memcpy(edict, so.edict.data(), so.edict.size() * sizeof(double));
}

int probe_hierarchy(int dim, const i64 *dims, const double *spacing, const double *vols, int S,
                    const double *diff, const int *caps, double shift, int pre, int post,
                    double omega, i64 coarsest, int max_levels, int cycles, int growth,
                    void **out) {
    return guard([&] {
        auto dom = make_domain(dim, dims, spacing, 0, nullptr, nullptr, nullptr, vols, S, diff);
        auto space = population_space::box(dom.voxels, std::vector<int>(caps, caps + S));
        multigrid_options o;
        o.pre_sweeps = pre;
        o.post_sweeps = post;
        o.jacobi_weight = omega;
        o.coarsest_size = coarsest;
        o.max_levels = max_levels;
        o.coarse_cycles = cycles;
        o.sweep_growth = growth;
        auto r = std::make_unique<mg_result>();
        for (i64 k = 0; k < space.n_sectors(); ++k)
            r->H.push_back(build_sector_hierarchy(dom, space.totals(k), shift, o));
        *out = r.release();
    });
}

i64 probe_hierarchy_levels(void *h, i64 k) {
    return static_cast<i64>(static_cast<mg_result *>(h)->H[k].levels.size());
}

// rows, terminal flag, nnz of sector k's level l
void probe_level_info(void *h, i64 k, i64 l, i64 *rows, int *terminal, i64 *nnz) {
    const auto &s = static_cast<mg_result *>(h)->H[k].levels[l];
    *rows = s.rows;
    *terminal = s.terminal ? 1 : 0;
    *nnz = static_cast<i64>(s.A.col.size());
}

void probe_level_transfer(void *h, i64 k, i64 l, i64 *coarse_of, double *weight) {
    const auto &s = static_cast<mg_result *>(h)->H[k].levels[l];
    for (i64 i = 0; i < s.rows; ++i) {
        coarse_of[i] = s.coarse_of[i];
        weight[i] = s.weight[i];
    }
}

void probe_level_operator(void *h, i64 k, i64 l, i64 *ptr, i64 *col, double *val, double *exit) {
    const auto &s = static_cast<mg_result *>(h)->H[k].levels[l];
    for (i64 i = 0; i <= s.rows; ++i) ptr[i] = static_cast<i64>(s.A.ptr[i]);
    for (std::size_t q = 0; q < s.A.col.size(); ++q) {
        col[q] = s.A.col[q];
        val[q] = s.A.val[q];
    }
    std::memcpy(exit, s.A.exit.data(), s.rows * sizeof(double));
}

void probe_level_inverse(void *h, i64 k, i64 l, double *inverse) {
    const auto &s = static_cast<mg_result *>(h)->H[k].levels[l];
    std::memcpy(inverse, s.inverse.data(), s.inverse.size() * sizeof(double));
}

void probe_hierarchy_free(void *h) { delete static_cast<mg_result *>(h); }

} // extern "C"
