// Drop-in check: a caller written against MarkovKit's RDME resolvent interface, compiled against
// include/markovkit_amd/markovkit.hpp and run on the MI355X through the C ABI. Each function makes
// the calls of one test of the reference's tests/test_rdme.cpp, unchanged (the voxel domains of
// :22-43, the population spaces, generators, multigrid and resolvent solves of :75-297; the SSA
// part of :45-73 is outside the resolvent path); the check wording is this file's own. Exit
// status 0 when every check passes.
#include "markovkit_amd/markovkit.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

using namespace markovkit;

namespace {

int failures = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("  FAILED  %s\n", what);
        ++failures;
    }
}

real sum(const vec<real> &x) { return std::accumulate(x.begin(), x.end(), 0.0); }

array<local_reaction> birth_death(real birth, real death) {
    return {{{1}, [=](view<const int>, real volume) { return birth * volume; }},
            {{-1}, [=](view<const int> n, real) { return death * n[0]; }}};
}

// tests/test_rdme.cpp:22-33
void structured_voxel_domain() {
    voxel_domain domain(grid<2>({3, 2}, {0.5, 2.0}), {0.25, 1.0});
    expect(domain.n_voxels() == 6 && domain.n_species() == 2, "a 3 x 2 grid holds six voxels, two species");
    expect(std::abs(domain.volumes[0] - 1.0) < 1e-15, "volume = product of the spacings");
    const auto neighbors = domain.graph.neighbors(0);
    expect(neighbors.size() == 2, "the corner voxel has two neighbours");
    expect(std::abs(domain.jump_rate(0, 0, neighbors.begin()[0]) - 1.0) < 1e-15, "hop rate, short axis");
    expect(std::abs(domain.jump_rate(0, 0, neighbors.begin()[1]) - 0.0625) < 1e-15, "hop rate, long axis");
}

// tests/test_rdme.cpp:35-43
void explicit_graph_voxel_domain() {
    num::graph<double, num::idx> graph(2);
    graph.add_edge(0, 1, 3.0);
    voxel_domain domain(std::move(graph), {2.0, 4.0}, {0.5});
    expect(std::abs(domain.jump_rate(0, 0, domain.graph.neighbors(0)[0]) - 0.75) < 1e-15, "hop out of voxel 0");
    expect(std::abs(domain.jump_rate(0, 1, domain.graph.neighbors(1)[0]) - 0.375) < 1e-15, "hop out of voxel 1");
}

// tests/test_rdme.cpp:75-98
void population_spaces() {
    expect(population_sector(3, {4}).size() == 15, "C(6, 2) states for 4 molecules in 3 voxels");
    expect(population_sector(3, {16}).size() == 153, "C(18, 2) states for 16 molecules in 3 voxels");
    expect(population_sector(3, {64}).size() == 2145, "C(66, 2) states for 64 molecules in 3 voxels");

    population_sector two_species(2, {1, 2});
    static_assert(state_space<population_sector>);
    expect(two_species.size() == 6, "2 x 3 states with totals (1, 2)");
    for (idx i = 0; i < two_species.size(); ++i) {
        const state &x = two_species[i];
        expect(x[0] + x[2] == 1 && x[1] + x[3] == 2, "every state has the sector's totals");
        expect(two_species.find(x) == i, "find inverts the enumeration");
    }
    expect(!two_species.find({2, 0, 0, 2}), "a state with other totals is not found");

    population_box box(3, {2, 2});
    static_assert(state_space<population_box>);
    for (idx i = 0; i < box.size(); ++i) {
        const idx k = box.sector_of(i);
        expect(box.find(box[i]) == i, "box find inverts the enumeration");
        expect(box.sector_begin(k) <= i && i < box.sector_end(k), "sector_of brackets the index");
    }
    expect(!box.find({3, 0, 0, 0, 0, 0}), "a state above the caps is not found");

    bool threw = false;
    try {
        population_box huge(4096, {40});
    } catch (const std::overflow_error &) {
        threw = true;
    }
    expect(threw, "a box whose size overflows throws overflow_error");
}

// tests/test_rdme.cpp:100-121
void diffusion_generator() {
    population_sector space(3, {4});
    reaction_diffusion diffusion{voxel_domain(grid<1>({3}, {0.5}), {0.25}), {}};
    const generator Q(space, diffusion);
    static_assert(num::linear_operator<generator, vec<real>, vec<real>>);
    expect(Q.rows() == 15, "a row per state of the sector");

    const idx source = *space.find({4, 0, 0});
    expect(std::abs(Q.exit_rates()[source] - 4.0) < 1e-15, "all four molecules hop at rate 1");

    vec<real> p(Q.rows(), 0.0), free(Q.rows(), 0.0), assembled(Q.rows(), 0.0);
    p[source] = 1.0;
    Q.apply(p, free);
    Q.matrix().apply(p, assembled);
    expect(std::abs(free[source] + 4.0) < 1e-15, "the diagonal is minus the exit rate");
    expect(std::abs(sum(free)) < 1e-15, "columns of a diffusion generator sum to zero");
    for (idx i = 0; i < Q.rows(); ++i) {
        expect(std::abs(free[i] - assembled[i]) < 1e-15, "the assembled matrix equals the operator");
    }
}

// tests/test_rdme.cpp:123-158
void reactions_couple_neighboring_sectors() {
    reaction_diffusion model{voxel_domain(grid<1>({2}, {1.0}), {0.5}), birth_death(2.0, 0.25)};
    population_box box(2, {3});
    const generator Q(box, model, boundary::absorbing);
    expect(Q.rows() == 10, "sectors of 1 + 2 + 3 + 4 states");

    bool diffusion = false, birth = false, death = false;
    for (idx j = 0; j < Q.rows(); ++j) {
        for (const auto &e : Q.jumps(j)) {
            const auto change = static_cast<long>(box.sector_of(e.to)) - long(box.sector_of(j));
            expect(std::abs(change) <= 1, "a jump moves at most one sector");
            diffusion = diffusion || change == 0;
            birth = birth || change == 1;
            death = death || change == -1;
        }
    }
    expect(diffusion && birth && death, "jumps within, up and down a sector");

    const idx empty = *box.find({0, 0});
    const idx full = *box.find({3, 0});
    expect(std::abs(Q.exit_rates()[empty] - 4.0) < 1e-15, "the empty state leaves by two births");
    expect(std::abs(Q.exit_rates()[full] - 6.25) < 1e-15, "absorbing: the leak adds to the exit rate");

    vec<real> p(Q.rows(), 0.0), free(Q.rows(), 0.0), assembled(Q.rows(), 0.0);
    p[full] = 1.0;
    Q.apply(p, free);
    Q.matrix().apply(p, assembled);
    expect(std::abs(sum(free) + 4.0) < 1e-15, "probability leaves only through births above the cap");
    for (idx i = 0; i < Q.rows(); ++i) {
        expect(std::abs(free[i] - assembled[i]) < 1e-15, "the assembled matrix equals the operator");
    }

    const generator R(box, model, boundary::reflecting);
    expect(std::abs(R.exit_rates()[full] - 2.25) < 1e-15, "reflecting: no leak in the exit rate");
    expect(std::abs(R.leak_rates()[full] - 4.0) < 1e-15, "reflecting: the leak rate is kept");

    // shift I - Q assembled, three ways
    const shifted_generator A(Q, 0.5);
    vec<real> y1(Q.rows(), 0.0), y2(Q.rows(), 0.0), y3(Q.rows(), 0.0);
    A.apply(p, y1);
    A.matrix().apply(p, y2);
    Q.characteristic_matrix(0.5).apply(p, y3);
    for (idx i = 0; i < Q.rows(); ++i) {
        expect(std::abs(y1[i] - y2[i]) < 1e-15 && y2[i] == y3[i], "characteristic matrix equals sI - Q");
    }
}

// tests/test_rdme.cpp:160-223
void population_multigrid_levels() {
    voxel_domain domain(grid<1>({8}, {1.0}), {1.0});
    population_multigrid multigrid(domain, {2}, 1.0);
    static_assert(num::linear_operator<decltype(multigrid), vec<real>, vec<real>>);
    expect(multigrid.rows() == 36, "36 states of 2 molecules in 8 voxels");
    expect(multigrid.transfer().coarse_size() == 10, "10 states on 4 voxels");
    expect(multigrid.n_levels() == 3, "8, 4 and 2 voxels");
    expect(multigrid.grid_complexity() >= 1.0 && multigrid.operator_complexity() >= 1.0,
           "complexities are at least one");

    vec<real> coarse(10, 0.0), fine(36, 0.0), restricted(10, 0.0);
    for (idx i = 0; i < coarse.size(); ++i) {
        coarse[i] = std::sin(real(i));
    }
    multigrid.transfer().prolong(coarse, fine);
    multigrid.transfer().restrict(fine, restricted);
    for (idx i = 0; i < coarse.size(); ++i) {
        expect(std::abs(restricted[i] - coarse[i]) < 1e-14, "restriction after prolongation is the identity");
    }
    expect(std::abs(sum(fine) - sum(coarse)) < 1e-14, "prolongation keeps the total");

    vec<real> fine_image(36, 0.0), projected(10, 0.0), coarse_image(10, 0.0);
    multigrid.apply_system(fine, fine_image);
    multigrid.transfer().restrict(fine_image, projected);
    multigrid.coarse_level()->apply_system(coarse, coarse_image);
    for (idx i = 0; i < coarse.size(); ++i) {
        expect(std::abs(projected[i] - coarse_image[i]) < 1e-12, "R A P equals the coarse operator");
    }

    const auto residual_norm = [&](const population_multigrid<1> &cycle, const vec<real> &b) {
        vec<real> x(b.size(), 0.0), image(b.size(), 0.0);
        cycle.apply(b, x);
        cycle.apply_system(x, image);
        real squared = 0.0;
        for (idx i = 0; i < b.size(); ++i) {
            squared += (b[i] - image[i]) * (b[i] - image[i]);
        }
        return squared;
    };
    vec<real> b(36, 0.0);
    real initial = 0.0;
    for (idx i = 0; i < b.size(); ++i) {
        b[i] = 0.25 + std::sin(real(i));
        initial += b[i] * b[i];
    }
    const real jacobi = residual_norm(multigrid, b);
    expect(jacobi < initial, "one Jacobi V-cycle lowers the residual");

    population_multigrid block(domain, {2}, 1.0, {.smoother = relaxation::aggregate_block});
    expect(residual_norm(block, b) < jacobi, "the aggregate-block cycle lowers it further");

    population_box box(8, {2});
    sector_multigrid sectors(domain, box, 1.0);
    static_assert(num::linear_operator<decltype(sectors), vec<real>, vec<real>>);
    expect(sectors.rows() == 45 && sectors.n_sectors() == 3, "sectors of 1, 8 and 36 states");
    vec<real> ones(45, 1.0), correction(45, 0.0);
    sectors.apply(ones, correction);
    expect(std::all_of(correction.begin(), correction.end(), [](real v) { return std::isfinite(v); }),
           "the sector cycle stays finite");

    // relax (multigrid.hpp:231-236): Jacobi sweeps from a guess lower the residual
    vec<real> x(36, 0.0), image(36, 0.0);
    multigrid.relax(b, x, 3);
    multigrid.apply_system(x, image);
    real relaxed = 0.0;
    for (idx i = 0; i < b.size(); ++i) relaxed += (b[i] - image[i]) * (b[i] - image[i]);
    expect(relaxed < initial, "three Jacobi sweeps lower the residual");
}

// tests/test_rdme.cpp:225-251
void multigrid_preconditions_resolvent() {
    reaction_diffusion model{voxel_domain(grid<1>({4}, {1.0}), {4.0}), birth_death(0.05, 0.02)};
    population_box box(4, {3});
    const generator Q(box, model, boundary::absorbing);
    constexpr real shift = 0.5;
    const shifted_generator A(Q, shift);
    const sector_multigrid preconditioner(model.domain, box, shift);
    static_assert(num::linear_operator<shifted_generator, vec<real>, vec<real>>);

    vec<real> b(Q.rows(), 0.0);
    for (idx i = 0; i < b.size(); ++i) {
        b[i] = 1.0 + 0.1 * std::cos(real(i));
    }
    const num::gmres_options options{.tolerance = 1e-10,
                                     .max_iterations = 200,
                                     .restart = Q.rows()};
    vec<real> x_plain(Q.rows(), 0.0), x_mg(Q.rows(), 0.0);
    const auto plain = num::gmres(A, b, x_plain, options);
    const auto mg = num::gmres(A, preconditioner, b, x_mg, options);
    std::printf("  RDME GMRES iterations: %zu -> %zu with sector multigrid\n", plain.iterations,
                mg.iterations);
    expect(plain.converged && mg.converged, "plain and preconditioned solves converge");
    expect(mg.iterations <= plain.iterations, "the preconditioner needs no more iterations");
    for (idx i = 0; i < Q.rows(); ++i) {
        expect(std::abs(x_mg[i] - x_plain[i]) < 1e-8, "the two solutions agree");
    }
}

// tests/test_rdme.cpp:253-297
void two_particle_annihilation_transform() {
    reaction_diffusion model{
        voxel_domain(grid<1>({4}, {0.25}), {1.0, 1.0}),
        {{{-1, -1}, [](view<const int> n, real volume) { return 2.0 * n[0] * n[1] / volume; }}}};
    population_box box(4, {1, 1});
    const generator Q(box, model);
    expect(Q.rows() == 25, "vacuum, one A, one B, and both");

    constexpr real shift = 0.5;
    const shifted_generator A(Q, shift);
    const sector_multigrid preconditioner(model.domain, box, shift);
    const num::gmres_options options{.tolerance = 1e-12,
                                     .max_iterations = 100,
                                     .restart = Q.rows()};
    const idx joint = *box.find_sector({1, 1});
    const idx vacuum = *box.find(state(8, 0));

    const auto transform = [&](idx a_voxel, idx b_voxel) {
        state initial(8, 0);
        initial[model.index(a_voxel, 0)] = 1;
        initial[model.index(b_voxel, 1)] = 1;
        vec<real> b(Q.rows(), 0.0), x(Q.rows(), 0.0);
        b[*box.find(initial)] = 1.0;
        expect(num::gmres(A, preconditioner, b, x, options).converged, "the resolvent solve converges");
        real survival = 0.0, flux = 0.0;
        for (idx i = box.sector_begin(joint); i < box.sector_end(joint); ++i) {
            survival += x[i];
            for (idx v = 0; v < 4; ++v) {
                flux += 2.0 * box[i][model.index(v, 0)] * box[i][model.index(v, 1)] /
                        model.domain.volumes[v] * x[i];
            }
        }
        const real absorbed = shift * x[vacuum];
        expect(std::abs(absorbed - flux) < 1e-10, "absorbed mass equals the integrated reaction flux");
        expect(std::abs(absorbed + shift * survival - 1.0) < 1e-10, "absorbed plus surviving mass is one");
        return absorbed;
    };
    const real colocated = transform(0, 0);
    const real separated = transform(0, 3);
    expect(colocated > separated, "particles in one voxel annihilate sooner");
    expect(separated > 0.0 && colocated < 1.0, "both transforms lie strictly between 0 and 1");
}

// the build's own additions: the mixed-precision path and the error types
void mixed_precision_lumped_sums() {
    // 1D, 16 voxels, cap 3 (969 states), s = 0.5, b = e_0, FP32 multigrid inside FP64 GMRES:
    // sector sums against the lumped closed form of the linear birth-death chain
    reaction_diffusion model{voxel_domain(grid<1>({16}, {1.0 / 16}), {1.0}), birth_death(0.05, 0.02)};
    population_box box(16, {3});
    const generator Q(box, model);
    const shifted_generator A(Q, 0.5);
    multigrid_options opt;
    opt.working_precision = precision::fp32;
    const sector_multigrid M(model.domain, box, 0.5, opt);
    vec<real> b(Q.rows(), 0.0), x(Q.rows(), 0.0);
    b[0] = 1.0;
    const auto r = num::gmres(A, M, b, x, {.tolerance = 1e-10, .max_iterations = 200, .restart = 50});
    expect(r.converged, "mixed-precision solve converges");
    const real expected[4] = {1.8240352155078126, 0.16096842646485102, 0.013756057739361041,
                              0.001127545716341069};
    real err = 0.0;
    for (idx k = 0; k < 4; ++k) {
        real s = 0.0;
        for (idx i = box.sector_begin(k); i < box.sector_end(k); ++i) s += x[i];
        err += std::abs(s - expected[k]);
    }
    expect(err < 1e-9, "sector sums match the lumped closed form");
    bool threw = false;
    try {
        const sector_multigrid bad(model.domain, box, 0.0);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    expect(threw, "a non-positive shift throws invalid_argument");
}

} // namespace

int main() {
    structured_voxel_domain();
    explicit_graph_voxel_domain();
    population_spaces();
    diffusion_generator();
    reactions_couple_neighboring_sectors();
    population_multigrid_levels();
    multigrid_preconditions_resolvent();
    two_particle_annihilation_transform();
    mixed_precision_lumped_sums();
    if (failures) {
        std::printf("rdme drop-in: %d check(s) FAILED\n", failures);
        return 1;
    }
    std::printf("rdme drop-in: all checks passed\n");
    return 0;
}
