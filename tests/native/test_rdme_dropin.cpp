// Drop-in check: code written against MarkovKit's RDME resolvent interface (the calls and
// checks of tests/test_rdme.cpp, reference repository) compiled against
// include/markovkit_amd/markovkit.hpp and run on the MI355X through the C ABI. Exit status 0
// when every check passes.
#include "markovkit_amd/markovkit.hpp"

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

void population_spaces() {
    population_box box(3, {2, 2});
    for (idx i = 0; i < box.size(); ++i) {
        const idx k = box.sector_of(i);
        expect(box.find(box[i]) == i, "box rank reproduces enumeration");
        expect(box.sector_begin(k) <= i && i < box.sector_end(k), "sector_of");
    }
    expect(!box.find({3, 0, 0, 0, 0, 0}), "a state outside the box is absent");
    population_box two(2, {1, 2});
    const idx k = *two.find_sector({1, 2});
    expect(two.sector_end(k) - two.sector_begin(k) == 6, "two-species sector size");
    bool threw = false;
    try {
        population_box huge(4096, {40});
    } catch (const std::overflow_error &) {
        threw = true;
    }
    expect(threw, "oversized box throws overflow_error");
}

void diffusion_generator() {
    population_box space(3, {4});
    reaction_diffusion diffusion{voxel_domain(grid<1>({3}, {0.5}), {0.25}), {}};
    const generator Q(space, diffusion);
    const idx source = *space.find({4, 0, 0});
    expect(std::abs(Q.exit_rates()[source] - 4.0) < 1e-15, "four molecules can hop");
    vec<real> p(Q.rows(), 0.0), free(Q.rows(), 0.0);
    p[source] = 1.0;
    Q.apply(p, free);
    expect(std::abs(free[source] + 4.0) < 1e-15, "diagonal");
    expect(std::abs(sum(free)) < 1e-15, "diffusion conserves probability");
}

void reactions_couple_neighboring_sectors() {
    reaction_diffusion model{voxel_domain(grid<1>({2}, {1.0}), {0.5}), birth_death(2.0, 0.25)};
    population_box box(2, {3});
    const generator Q(box, model, boundary::absorbing);
    expect(Q.rows() == 10, "sector k holds k + 1 configurations of two voxels");
    const idx empty = *box.find({0, 0}), full = *box.find({3, 0});
    expect(std::abs(Q.exit_rates()[empty] - 4.0) < 1e-15, "two births from empty");
    expect(std::abs(Q.exit_rates()[full] - 6.25) < 1e-15, "absorbing exit rate");
    vec<real> p(Q.rows(), 0.0), free(Q.rows(), 0.0);
    p[full] = 1.0;
    Q.apply(p, free);
    expect(std::abs(sum(free) + 4.0) < 1e-14, "only the leaking births lose probability");
    const generator R(box, model, boundary::reflecting);
    expect(std::abs(R.exit_rates()[full] - 2.25) < 1e-15, "reflecting drops the leak");
    expect(std::abs(R.leak_rates()[full] - 4.0) < 1e-15, "the leak is still recorded");
}

void multigrid_preconditions_resolvent() {
    reaction_diffusion model{voxel_domain(grid<1>({4}, {1.0}), {4.0}), birth_death(0.05, 0.02)};
    population_box box(4, {3});
    const generator Q(box, model, boundary::absorbing);
    constexpr real shift = 0.5;
    const shifted_generator A(Q, shift);
    const sector_multigrid preconditioner(model.domain, box, shift);
    vec<real> b(Q.rows(), 0.0);
    for (idx i = 0; i < b.size(); ++i) b[i] = 1.0 + 0.1 * std::cos(real(i));
    const num::gmres_options options{.tolerance = 1e-10, .max_iterations = 200, .restart = Q.rows()};
    vec<real> x_plain(Q.rows(), 0.0), x_mg(Q.rows(), 0.0);
    const auto plain = num::gmres(A, b, x_plain, options);
    const auto mg = num::gmres(A, preconditioner, b, x_mg, options);
    std::printf("  RDME GMRES iterations: %zu -> %zu with sector multigrid\n", plain.iterations,
                mg.iterations);
    expect(plain.converged && mg.converged, "both solves converge");
    expect(mg.iterations <= plain.iterations, "multigrid does not slow GMRES down");
    for (idx i = 0; i < Q.rows(); ++i) expect(std::abs(x_mg[i] - x_plain[i]) < 1e-8, "both solves agree");
    vec<real> z(Q.rows(), 0.0);
    preconditioner.apply(b, z);
    for (real v : z) expect(std::isfinite(v), "sector multigrid is finite");
}

void two_particle_annihilation_transform() {
    reaction_diffusion model{
        voxel_domain(grid<1>({4}, {0.25}), {1.0, 1.0}),
        {{{-1, -1}, [](view<const int> n, real volume) { return 2.0 * n[0] * n[1] / volume; }}}};
    population_box box(4, {1, 1});
    const generator Q(box, model);
    expect(Q.rows() == 25, "1 + V + V + V^2 states");
    constexpr real shift = 0.5;
    const shifted_generator A(Q, shift);
    const sector_multigrid preconditioner(model.domain, box, shift);
    const num::gmres_options options{.tolerance = 1e-12, .max_iterations = 100, .restart = Q.rows()};
    const idx joint = *box.find_sector({1, 1});
    const idx vacuum = *box.find(state(8, 0));
    const auto transform = [&](idx a_voxel, idx b_voxel) {
        state initial(8, 0);
        initial[model.index(a_voxel, 0)] = 1;
        initial[model.index(b_voxel, 1)] = 1;
        vec<real> b(Q.rows(), 0.0), x(Q.rows(), 0.0);
        b[*box.find(initial)] = 1.0;
        expect(num::gmres(A, preconditioner, b, x, options).converged, "resolvent solve");
        real survival = 0.0, flux = 0.0;
        for (idx i = box.sector_begin(joint); i < box.sector_end(joint); ++i) {
            const state s = box[i];
            survival += x[i];
            for (idx v = 0; v < 4; ++v)
                flux += 2.0 * s[model.index(v, 0)] * s[model.index(v, 1)] /
                        model.domain.volumes[v] * x[i];
        }
        const real absorbed = shift * x[vacuum];
        expect(std::abs(absorbed - flux) < 1e-10, "absorption equals reaction flux");
        expect(std::abs(absorbed + shift * survival - 1.0) < 1e-10, "E[exp(-sT)] + s S_hat(s) = 1");
        return absorbed;
    };
    const real colocated = transform(0, 0);
    const real separated = transform(0, 3);
    expect(colocated > separated, "colocated particles react sooner");
    expect(separated > 0.0 && colocated < 1.0, "transforms lie in (0, 1)");
}

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

void aggregate_blocks_beat_jacobi() {
    // tests/test_rdme.cpp:191-210 on the 36-state sector of box(8, {2}): the aggregate-block
    // V-cycle leaves a smaller residual than the Jacobi V-cycle, which reduces it
    voxel_domain domain(grid<1>({8}, {1.0}), {1.0});
    population_box box(8, {2});
    const generator Q(box, reaction_diffusion{domain, {}});
    vec<real> b(Q.rows(), 0.0);
    real initial = 0.0;
    const idx k = *box.find_sector({2});
    for (idx i = box.sector_begin(k); i < box.sector_end(k); ++i) {
        b[i] = 0.25 + std::sin(real(i - box.sector_begin(k)));
        initial += b[i] * b[i];
    }
    const auto residual_norm = [&](relaxation smoother) {
        multigrid_options opt;
        opt.smoother = smoother;
        const sector_multigrid M(domain, box, 1.0, opt);
        vec<real> x(b.size(), 0.0), image(b.size(), 0.0);
        M.apply(b, x);
        Q.apply_shifted(1.0, x, image);
        real squared = 0.0;
        for (idx i = box.sector_begin(k); i < box.sector_end(k); ++i)
            squared += (b[i] - image[i]) * (b[i] - image[i]);
        return squared;
    };
    const real jacobi = residual_norm(relaxation::jacobi);
    expect(jacobi < initial, "a Jacobi V-cycle reduces the residual");
    expect(residual_norm(relaxation::aggregate_block) < jacobi, "aggregate blocks beat point Jacobi");
}

} // namespace

int main() {
    aggregate_blocks_beat_jacobi();
    population_spaces();
    diffusion_generator();
    reactions_couple_neighboring_sectors();
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
