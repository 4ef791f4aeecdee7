// Host-side setup of the RDME resolvent operators: population spaces with arithmetic rank,
// the generator assembled by destination (gather form for the GPU), and the per-sector voxel
// multigrid hierarchy. Everything here runs once per problem; the solve runs on the device.
//
// Reference semantics followed (paths relative to the markovkit repository root):
//   population_box / population_sector ordering and rank   rdme/population_space.hpp:15-196
//   reaction_diffusion transitions and rates                rdme/reaction_diffusion.hpp:82-118
//   generator exit / leak / absorbing rule                  generator.hpp:27-44
//   coarsen, population_transfer weights                    rdme/multigrid.hpp:24-118
//   population_multigrid level structure and terminal rule  rdme/multigrid.hpp:145-179
#pragma once

#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace mk {

using i64 = std::int64_t;
using u64 = std::uint64_t;
using u32 = std::uint32_t;

// C(rem + rest, rest): the number of ways to put `rem` molecules in rest + 1 voxels,
// tabulated for rest < voxels and rem <= max_total; saturates at UINT64_MAX
class compositions_table {
  public:
    compositions_table() = default;
    compositions_table(i64 voxels, int max_total);
    [[nodiscard]] u64 operator()(i64 rest, int rem) const {
        return table_[static_cast<std::size_t>(rest) * stride_ + static_cast<std::size_t>(rem)];
    }
    // the table, row-major by rest with stride() entries per row (device setup uploads it)
    [[nodiscard]] const std::vector<u64> &data() const { return table_; }
    [[nodiscard]] std::size_t stride() const { return stride_; }

  private:
    std::vector<u64> table_;
    std::size_t stride_ = 0;
};

// A set of population sectors over V voxels and S species, stored one after another.
// As a population_box it holds every sector with 0 <= totals[s] <= caps[s] in mixed-radix
// order (species 0 most significant); as a population_sector it holds one sector.
class population_space {
  public:
    static population_space box(i64 voxels, std::vector<int> caps);
    static population_space sector(i64 voxels, std::vector<int> totals);

    [[nodiscard]] i64 voxels() const { return voxels_; }
    [[nodiscard]] int species() const { return static_cast<int>(caps_.size()); }
    [[nodiscard]] i64 width() const { return voxels_ * species(); }
    [[nodiscard]] i64 size() const { return begin_.back(); }
    [[nodiscard]] i64 n_sectors() const { return static_cast<i64>(totals_.size()); }
    [[nodiscard]] const std::vector<int> &totals(i64 k) const { return totals_[k]; }
    [[nodiscard]] i64 begin(i64 k) const { return begin_[k]; }
    [[nodiscard]] i64 end(i64 k) const { return begin_[k + 1]; }
    [[nodiscard]] const std::vector<int> &caps() const { return caps_; }
    [[nodiscard]] bool is_box() const { return box_; }
    [[nodiscard]] const compositions_table &compositions() const { return comp_; }
    // per-species state counts of sector k (species 0 most significant in the sector rank)
    [[nodiscard]] const std::vector<u64> &species_sizes(i64 k) const { return species_sizes_[k]; }

    // the sector holding these totals, or -1
    [[nodiscard]] i64 sector_of_totals(const int *totals) const;
    [[nodiscard]] i64 sector_of_index(i64 i) const;
    // rank of x among the states of sector k, or -1 when x does not belong to it
    [[nodiscard]] i64 rank_in_sector(const int *x, i64 k) const;
    // global index of x, or -1 (population_box::find)
    [[nodiscard]] i64 find(const int *x) const;
    // x = the state with global index i
    void state(i64 i, int *x) const;
    void state_in_sector(i64 k, i64 local, int *x) const;

  private:
    i64 voxels_ = 0;
    std::vector<int> caps_;
    bool box_ = true;
    std::vector<std::vector<int>> totals_;
    std::vector<std::vector<u64>> species_sizes_;
    std::vector<i64> begin_{0};
    compositions_table comp_;
    void add_sector(const std::vector<int> &totals);
};

// The voxel graph of a voxel_domain: out-edges of every voxel in the reference's neighbour
// order, with the per-voxel volumes and per-species diffusivities. A grid domain also keeps
// its dimensions and spacing for coarsening.
struct voxel_domain {
    i64 voxels = 0;
    std::vector<i64> offsets;  // voxels + 1
    std::vector<i64> to;
    std::vector<double> weight;
    std::vector<double> volumes;
    std::vector<double> diffusivities;
    int dim = 0; // 0: explicit graph
    std::vector<i64> dims;
    std::vector<double> spacing;

    static voxel_domain grid(std::vector<i64> dims, std::vector<double> spacing,
                             const double *volumes, std::vector<double> diffusivities);
    static voxel_domain graph(i64 voxels, const i64 *offsets, const i64 *to, const double *weight,
                              const double *volumes, std::vector<double> diffusivities);
    [[nodiscard]] int species() const { return static_cast<int>(diffusivities.size()); }
    [[nodiscard]] double cell_volume() const;
    [[nodiscard]] std::vector<i64> coordinate(i64 v) const;
    [[nodiscard]] i64 vertex(const std::vector<i64> &c) const;
};

// coarsen (multigrid.hpp:24-42): halve every axis longer than one cell; nothing when an axis
// is odd or none is long. The coarse domain has uniform cell volumes and the given
// diffusivities.
std::optional<voxel_domain> coarsen(const voxel_domain &fine, std::vector<double> diffusivities);

// local reactions as rate tables: rates[(r * voxels + v) * T + tuple]
struct reaction_tables {
    int n = 0;
    std::vector<int> changes; // n * S
    std::vector<double> rates;
    std::vector<i64> stride;  // per species
    i64 T = 1;
};

// The generator in gather form: row i lists every jump j -> i as (j, rate), columns
// ascending, duplicates merged. exit[j] = -Q_jj and leak[j] as in generator.hpp:27-44.
struct generator_csr {
    i64 n = 0;
    i64 n_jumps = 0; // in-space transitions before merging (generator::n_jumps)
    std::vector<u64> ptr;
    std::vector<u32> col;
    std::vector<double> val;
    std::vector<double> exit;
    std::vector<double> leak;
};

generator_csr assemble_generator(const population_space &space, const voxel_domain &domain,
                                 const reaction_tables &reactions, bool absorbing);

// generator_csr from source-major jump lists (first, to, rate) with exit and leak given
generator_csr generator_from_jumps(i64 n, const i64 *first, const i64 *to, const double *rate,
                                   const double *exit, const double *leak);

// multigrid_options (multigrid.hpp:125-134)
struct multigrid_options {
    int pre_sweeps = 2;
    int post_sweeps = 2;
    double jacobi_weight = 2.0 / 3.0;
    i64 coarsest_size = 4;
    int max_levels = 32;
    int coarse_cycles = 1;
    int sweep_growth = 0;
    int smoother = 0; // 0 Jacobi, 1 aggregate blocks, 2 red-black Gauss-Seidel, 3 ILU(0)
    int precision = 0;
    int cycle_levels = 0; // 0: every level takes coarse_cycles visits; k: the top k levels only
};

// Aggregate-block smoother of one sector level (multigrid.hpp:268-342): for every coarse voxel
// a (aggregate), the blocks of two or more states that agree outside a, with the dense inverse
// of (shift I - Q) restricted to each block. Blocks are ordered by their canonical state (the
// aggregate's molecules moved into its first fine voxel), members ascending. rows lists, per
// aggregate, the states whose residual the aggregate's correction changes (its members and the
// targets of their jumps), ascending.
struct block_smoother {
    i64 aggregates = 0;
    std::vector<u64> agg_blocks;  // aggregates + 1
    std::vector<u64> block_first; // blocks + 1: first member of each block
    std::vector<u32> members;     // level row indices
    std::vector<u64> inv_offset;  // blocks
    std::vector<double> inverse;  // row-major per block
    std::vector<u64> agg_rows;    // aggregates + 1
    std::vector<u32> rows;
};

// One level of one sector's population_multigrid
struct sector_level {
    i64 rows = 0;
    generator_csr A;             // diffusion only, by destination
    bool terminal = false;
    std::vector<i64> coarse_of;  // local index in the next level (non-terminal)
    std::vector<double> weight;  // multinomial prolongation weight (non-terminal)
    std::vector<double> inverse; // rows x rows, row-major (terminal): (shift I - Q)^-1
    int pre_sweeps = 2, post_sweeps = 2;
    bool has_blocks = false;     // aggregate-block smoother (non-terminal levels)
    block_smoother blocks;
    // red-black smoother: colour of every state, the parity of sum over voxels v and species
    // s of x[v, s] * (coordinate sum of v); every diffusion jump joins the two colours
    std::vector<std::uint8_t> colour;
};

struct sector_hierarchy {
    std::vector<sector_level> levels;
};

// Optional replacements of the per-state loops of build_sector_hierarchy (the library runs
// them on the device); each gives exactly what the host loop gives. Empty functions, or a
// false return, leave the loop on the host.
struct level_hooks {
    // one level's diffusion operator (sector space, no reactions, absorbing rule)
    std::function<generator_csr(const population_space &, const voxel_domain &)> assemble;
    // for every state i of `fine`: coarse_of[i], the rank of its aggregate in `coarse`;
    // log_weight[i] = -sum_{v, s} log_fact[x_vs] + sum_{a, s} (log_fact[c_as] - c_as
    // log_children[a]) over occupied voxels in ascending order; and, when colour is given, the
    // red-black colour (parity of sum x_vs * (coordinate sum of v))
    std::function<bool(const population_space &fine, const population_space &coarse,
                       const voxel_domain &dom, const std::vector<i64> &parent,
                       const std::vector<double> &log_fact, const std::vector<double> &log_children,
                       std::vector<i64> &coarse_of, std::vector<double> &log_weight,
                       std::vector<std::uint8_t> *colour)>
        transfer;
};

// population_multigrid's levels for one sector (multigrid.hpp:145-179)
sector_hierarchy build_sector_hierarchy(const voxel_domain &domain, const std::vector<int> &totals,
                                        double shift, const multigrid_options &options,
                                        const level_hooks &hooks = {});

// dense inverse by Gauss-Jordan with partial pivoting; false when singular
bool invert_dense(std::vector<double> a, i64 n, std::vector<double> &inverse);

// exp(X) of a small dense matrix (row-major n x n): diagonal degree-6 Pade approximant of
// X / 2^s with ||X / 2^s||_inf <= 1/2, squared s times
std::vector<double> expm_pade(std::vector<double> X, i64 n);

// The coded sliced-ELL encoding of a gather-form operator (the device's primary layout):
// entry k of row 64 s + l is ent[soff[s] + 64 k + l] = column | rate code << 24, with the
// slice padded to its longest row; rlen / rcode hold each row's length and exit-rate code.
constexpr int kSliceRows = 64;
constexpr int kCodedColumnBits = 24;
template <typename VT>
struct sliced_operator {
    std::vector<u32> soff;
    std::vector<u32> ent;
    std::vector<std::uint8_t> rlen, rcode;
    std::vector<VT> vdict, edict;
    std::vector<VT> exit; // per-row exit rates when edict is empty (exit_array encodings)
    u64 padded = 0;
};

// false when the operator does not fit the coded layout (a column at or above 2^24, more than 256
// distinct rates or exit rates, rows longer than 255 entries, or 2^32 stored entries)
// With exit_array, exit rates that do not fit the 256-entry dictionary are kept per row instead.
template <typename VT>
bool encode_sliced(const generator_csr &g, sliced_operator<VT> &out, bool exit_array = false);

unsigned worker_threads();

// task(t) for t in [0, n_tasks) on the worker pool (dynamically scheduled)
void parallel_tasks(i64 n_tasks, const std::function<void(i64)> &task);

} // namespace mk
