/* markovkit_amd — MI355X (gfx950) RDME resolvent solve behind a C ABI.
 *
 * The reference boundary is C++ (templates and concepts, no C ABI). Each entry point below
 * replaces one reference construct; the citation (paths relative to the markovkit repository
 * root) names it. Header-only C++ wrappers with the reference's names sit on top of these
 * entry points in include/markovkit_amd/markovkit.hpp.
 *
 * Conventions
 *  - Every function returns an mk_status; MK_OK is 0. The message of the last failure on
 *    the calling thread is mk_last_error(). Status values map to the reference's exception
 *    types: MK_ERR_INVALID_ARGUMENT -> std::invalid_argument, MK_ERR_OVERFLOW ->
 *    std::overflow_error, MK_ERR_RUNTIME -> std::runtime_error; MK_ERR_DEVICE is a HIP failure.
 *  - Handles are library-owned (create/destroy). Vectors are caller-owned: the plain entry
 *    points take host buffers, the *_device entry points take device (HBM) pointers.
 *  - Non-convergence is not an error: mk_solver_result.converged == 0.
 *  - Calls on one handle are serialised on the library's HIP stream; a handle is not
 *    reentrant (like the reference's mutable multigrid scratch, multigrid.hpp:359-360).
 */
#ifndef MARKOVKIT_AMD_H
#define MARKOVKIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mk_status;
enum {
    MK_OK = 0,
    MK_ERR_INVALID_ARGUMENT = 1,
    MK_ERR_OVERFLOW = 2,
    MK_ERR_RUNTIME = 3,
    MK_ERR_DEVICE = 4
};

typedef struct mk_box_s *mk_box;
typedef struct mk_domain_s *mk_domain;
typedef struct mk_generator_s *mk_generator;
typedef struct mk_sector_mg_s *mk_sector_mg;
typedef struct mk_batch_s *mk_batch;

/* boundary (generator.hpp:11-13) */
enum { MK_REFLECTING = 0, MK_ABSORBING = 1 };
/* relaxation (multigrid.hpp:120-123), plus two smoothers of the paper driver's family
 * (rdme/cpp/rdme_full_multigrid.cpp:178-195, :204-296):
 *   MK_RED_BLACK: Gauss-Seidel in two colours. The fixed-population diffusion graph on a grid is
 *     bipartite; colour 0 states are updated first (colour 1 first when post-smoothing), each
 *     colour in parallel. jacobi_weight is not used.
 *   MK_ILU0: x += (L U)^-1 (rhs - A x) with the ILU(0) factors of the level operator in rank
 *     order (only the pivots differ from A on a bipartite graph); triangular solves are
 *     level-scheduled. The pivots depend on the shift (mk_sector_mg_set_shift refuses it). */
enum { MK_JACOBI = 0, MK_AGGREGATE_BLOCK = 1, MK_RED_BLACK = 2, MK_ILU0 = 3 };
/* working precision of the multigrid preconditioner (no reference counterpart) */
enum { MK_FP64 = 0, MK_FP32 = 1 };

/* multigrid_options (multigrid.hpp:125-134), same fields and defaults */
typedef struct mk_multigrid_options {
    int32_t pre_sweeps;     /* 2 */
    int32_t post_sweeps;    /* 2 */
    double jacobi_weight;   /* 2/3 */
    int64_t coarsest_size;  /* 4 */
    int32_t max_levels;     /* 32 */
    int32_t coarse_cycles;  /* 1 = V-cycle, 2 = W-cycle */
    int32_t sweep_growth;   /* 0 */
    int32_t smoother;       /* MK_JACOBI */
    int32_t precision;      /* MK_FP64 or MK_FP32 (device-only extension) */
    int32_t cycle_levels;   /* 0: every level takes coarse_cycles visits of its coarse level (the
                               reference); k > 0: only the top k levels do, the ones below one
                               (a W-cycle at the top over a V-cycle tail; extension) */
} mk_multigrid_options;

/* Arnoldi orthogonalisation of the device solver (extension; the reference's field set ends at
 * restart). MK_DCGS2: classical Gram-Schmidt twice with the second projection delayed into the
 * next step, one global reduction per step; used for preconditioned solves with restart <= 64,
 * otherwise the solver falls back to MK_CGS2 (two projections per step). */
enum { MK_DCGS2 = 0, MK_CGS2 = 1 };

/* num::gmres_options (numerics, call site tests/test_rdme.cpp:238-240); tolerance is an
 * absolute Euclidean residual norm, x on input is the initial guess */
typedef struct mk_gmres_options {
    double tolerance;
    int64_t max_iterations;
    int64_t restart;
    int32_t orthogonalization; /* MK_DCGS2 (0, default) or MK_CGS2 */
} mk_gmres_options;

/* num::solver_result (examples/rdme_multigrid_benchmark.cpp:55); residual is the final true
 * residual ||b - A x||_2 recomputed in FP64 */
typedef struct mk_solver_result {
    int64_t iterations;
    double residual;
    int32_t converged;
} mk_solver_result;

const char *mk_last_error(void);
void mk_default_multigrid_options(mk_multigrid_options *out);

/* ---- population_box (rdme/population_space.hpp:131-196) ---------------------------- */
/* population_box(n_voxels, max_totals): population_space.hpp:185-190 */
mk_status mk_box_create(int64_t n_voxels, int32_t n_species, const int32_t *max_totals,
                        mk_box *out);
mk_status mk_box_destroy(mk_box box);
/* size(), n_sectors(): population_space.hpp:192, :206 */
mk_status mk_box_size(mk_box box, int64_t *size);
mk_status mk_box_n_sectors(mk_box box, int64_t *n_sectors);
/* sector_begin(k) / sector_totals(k): population_space.hpp:209-211 */
mk_status mk_box_sector(mk_box box, int64_t k, int64_t *begin, int64_t *end,
                        int32_t *totals /* n_species */);
/* operator[](i): population_space.hpp:193; state is voxel-major, n_voxels*n_species ints */
mk_status mk_box_state(mk_box box, int64_t i, int32_t *state);
/* find(x): population_space.hpp:194-202; *index = -1 when x is outside the box */
mk_status mk_box_find(mk_box box, const int32_t *state, int64_t *index);

/* population_sector(n_voxels, totals): population_space.hpp:103-127, every state with totals[s]
 * molecules of species s, in rank order. The handle is a space with one sector: every mk_box_*
 * query, mk_generator_create_rdme and mk_sector_mg_create take it (a sector multigrid over it is
 * the reference's population_multigrid). */
mk_status mk_sector_space_create(int64_t n_voxels, int32_t n_species, const int32_t *totals,
                                 mk_box *out);

/* ---- voxel_domain (rdme/reaction_diffusion.hpp:23-56) over a grid<D> ---------------- */
/* voxel_domain(grid<D>(dims, spacing), [volumes,] diffusivities): reaction_diffusion.hpp:29-47.
 * volumes may be NULL (uniform cell volume). */
mk_status mk_domain_create_grid(int32_t dim, const int64_t *dims, const double *spacing,
                                const double *volumes, int32_t n_species,
                                const double *diffusivities, mk_domain *out);
/* voxel_domain over an explicit graph (num::graph; tests/test_rdme.cpp:35-43):
 * CSR adjacency offsets[n_voxels+1], to[], weight[]; the multigrid needs a grid domain */
mk_status mk_domain_create_graph(int64_t n_voxels, const int64_t *offsets, const int64_t *to,
                                 const double *weight, const double *volumes, int32_t n_species,
                                 const double *diffusivities, mk_domain *out);
mk_status mk_domain_destroy(mk_domain domain);

/* ---- generator (generator.hpp:19-112) ----------------------------------------------- */
/* generator(box, reaction_diffusion{domain, reactions}, boundary): generator.hpp:27-44 with
 * the model of reaction_diffusion.hpp:82-118. changes[r*n_species + s] is reaction r's change;
 * rates[(r*n_voxels + v)*T + t] is its rate in voxel v at the count tuple t (mixed radix
 * max_totals[s]+1, species 0 most significant), T = prod(max_totals[s]+1). */
mk_status mk_generator_create_rdme(mk_box box, mk_domain domain, int32_t n_reactions,
                                   const int32_t *changes, const double *rates,
                                   int32_t boundary, mk_generator *out);
/* generator from source-major jump lists (the layout of generator.hpp:109-112) for models the
 * RDME builder does not cover: first[n+1], to[first[n]], rate[first[n]], exit[n], leak[n] */
mk_status mk_generator_create_jumps(int64_t n, const int64_t *first, const int64_t *to,
                                    const double *rate, const double *exit, const double *leak,
                                    mk_generator *out);
mk_status mk_generator_destroy(mk_generator gen);
/* rows(), n_jumps(): generator.hpp:46-48 */
mk_status mk_generator_size(mk_generator gen, int64_t *rows, int64_t *n_jumps);
/* exit_rates(), leak_rates(): generator.hpp:55-59 */
mk_status mk_generator_rates(mk_generator gen, double *exit, double *leak);
/* The stored jumps, for generator::jumps(j) and matrix() (generator.hpp:51-53, :84-89): row i
 * (a destination) lists (column j, rate of j -> i), columns ascending, with jumps of one source
 * to one destination merged (summed). *entries receives the count; ptr[rows + 1], col[entries]
 * and rate[entries] may be NULL (size query). Decoded on the host from the device layout. */
mk_status mk_generator_entries(mk_generator gen, int64_t *entries, int64_t *ptr, int64_t *col,
                               double *rate);
/* apply(x, y): y = Q x, generator.hpp:62-70 */
mk_status mk_generator_apply(mk_generator gen, const double *x, double *y);
/* shifted_generator::apply -> apply_shifted(shift, x, y): y = shift x - Q x,
 * generator.hpp:73-81 and :127 */
mk_status mk_shifted_apply(mk_generator gen, double shift, const double *x, double *y);
mk_status mk_shifted_apply_device(mk_generator gen, double shift, const double *d_x,
                                  double *d_y);

/* ---- sector_multigrid (rdme/multigrid.hpp:371-408) ---------------------------------- */
/* sector_multigrid(domain, box, shift, options): multigrid.hpp:377-384; options may be NULL */
mk_status mk_sector_mg_create(mk_domain domain, mk_box box, double shift,
                              const mk_multigrid_options *options, mk_sector_mg *out);
/* Re-target a Jacobi hierarchy to another shift: only the dense terminal inverses depend on it
 * (multigrid.hpp:247-264), so they are re-inverted and the level operators stay resident. No
 * reference counterpart (the reference builds a new sector_multigrid per shift). Rejected for
 * the aggregate-block smoother, whose block inverses depend on the shift. */
mk_status mk_sector_mg_set_shift(mk_sector_mg mg, double shift);
mk_status mk_sector_mg_destroy(mk_sector_mg mg);
/* max_levels(), n_sectors(): multigrid.hpp:388-395 */
mk_status mk_sector_mg_info(mk_sector_mg mg, int64_t *rows, int64_t *n_sectors,
                            int64_t *max_levels);
/* population_multigrid (multigrid.hpp:139-236) = a sector multigrid over a space made by
 * mk_sector_space_create (one sector). For such a handle:
 * level l's states, jumps (the diffusion generator's, before merging) and terminal flag
 * (n_levels, grid_complexity, operator_complexity, multigrid.hpp:188-193); */
mk_status mk_sector_mg_level(mk_sector_mg mg, int32_t level, int64_t *rows, int64_t *jumps,
                             int32_t *terminal);
/* the transfer from level l to level l + 1 (population_transfer, multigrid.hpp:47-118): for each
 * level-l state its coarse state and its multinomial prolongation weight (FP64, rank order) */
mk_status mk_sector_mg_transfer(mk_sector_mg mg, int32_t level, int64_t *coarse_of, double *weight);
/* relax(rhs, x, sweeps, reverse) (multigrid.hpp:231-236): `sweeps` sweeps of the hierarchy's
 * smoother on its finest level from the guess x, on the device in the hierarchy's working
 * precision (any sector multigrid handle; the finest level of every sector). */
mk_status mk_sector_mg_relax(mk_sector_mg mg, const double *rhs, double *x, int32_t sweeps,
                             int32_t reverse);
mk_status mk_sector_mg_relax_device(mk_sector_mg mg, const double *d_rhs, double *d_x,
                                    int32_t sweeps, int32_t reverse);
/* apply(b, x): one cycle per sector from a zero guess, multigrid.hpp:397-403 */
mk_status mk_sector_mg_apply(mk_sector_mg mg, const double *b, double *x);
mk_status mk_sector_mg_apply_device(mk_sector_mg mg, const double *d_b, double *d_x);

/* ---- num::gmres (numerics; call sites tests/test_rdme.cpp:242-243) ------------------ */
/* gmres(shifted_generator(gen, shift), [mg,] b, x, options): flexible GMRES(restart) with
 * CGS2 Arnoldi and Givens rotations in FP64; mg may be NULL (plain GMRES). */
mk_status mk_gmres(mk_generator gen, double shift, mk_sector_mg mg, const double *b, double *x,
                   const mk_gmres_options *options, mk_solver_result *result);
mk_status mk_gmres_device(mk_generator gen, double shift, mk_sector_mg mg, const double *d_b,
                          double *d_x, const mk_gmres_options *options,
                          mk_solver_result *result);
/* mk_gmres_device without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under their own HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The Krylov basis, the preconditioned vectors
 * and the work arrays live in one 64 MiB device array of the library (one per GPU, held until
 * the process ends); the generator and the hierarchy are created beforehand, as for any solve.
 * Limits: restart <= 64, and (restart + 1) basis vectors of 8 bytes plus restart preconditioned
 * vectors (4 bytes each with an FP32 hierarchy, else 8) per row, rounded to 64 rows, and the
 * reduction arrays must fit 64 MiB (about 58,000 rows at restart 50 with an FP32 hierarchy);
 * larger requests fail with MK_ERR_INVALID_ARGUMENT. Calls on one GPU share its array, so calls
 * that overlap on one GPU (from several host threads) race: serialise them. Launches are eager
 * (no graph replay); the kernels and the arithmetic are mk_gmres_device's, so the results are
 * the same bit for bit. */
mk_status mk_gmres_device_noalloc(mk_generator gen, double shift, mk_sector_mg mg,
                                  const double *d_b, double *d_x, const mk_gmres_options *options,
                                  mk_solver_result *result);

/* ---- whole-box multigrid (the paper driver, rdme/cpp/rdme_full_multigrid.cpp) ----------
 * The driver's Level hierarchy (:298-560) for a single species on a Cartesian grid of
 * cells^dimension voxels (unit box, cell volume 1 / V) with every state of total population
 * <= cap, in the population box order (total first, voxel 0 most significant; :66-152). Every
 * level holds the complete operator shift I - Q: diffusion at D cells^2 per molecule and
 * neighbour, births c3 vol + c1 n (n - 1) / (2 vol), deaths c4 n + c2 n (n - 1) (n - 2) /
 * (6 vol^2) (:400-402), births at the cap lost (absorbing). Coarse levels halve the cells and
 * double D (:538-541); transfers are the multinomial weights of :555-572. The one-voxel level is
 * solved by the Thomas algorithm (:519-535); with terminal_rows > 0 the first level of at most
 * that many states is solved exactly with a dense inverse instead (no driver counterpart).
 * Smoothers: MK_BOX_GS is Gauss-Seidel in four-colour order (colour = 2 (total mod 2) + (sum of
 * the molecules' voxel coordinate sums mod 2); replaces the driver's lexicographic "gs", :178-195,
 * whose sweep is sequential), MK_BOX_JACOBI the driver's "jacobi" (w = 2/3, :180-186).
 * The driver's sequential smoothers themselves run as level-scheduled triangular solves (rows in
 * the box order, long dependency chains: slower per cycle, the driver's cycle counts):
 * MK_BOX_GS_LEX is its "gs" (x += (D + L)^-1 (b - A x), backward (D + U) after the correction),
 * MK_BOX_BLOCK_ILU its structured "block-ilu" (population blocks in order, each the block's
 * residual and its ILU(0) forward and backward solves, :204-296, :795-815).
 * Cycles (:819-878): cycles = 1 (V) or 2 (W), cycle_type MK_CYCLE_K for the K-cycle. FP64. */
typedef struct mk_box_mg_s *mk_box_mg;
enum { MK_BOX_GS = 0, MK_BOX_JACOBI = 1, MK_BOX_GS_LEX = 2, MK_BOX_BLOCK_ILU = 3 };
enum { MK_CYCLE_STANDARD = 0, MK_CYCLE_K = 1 };

typedef struct mk_box_mg_params {
    int32_t dimension;     /* 1, 2 or 3 (--dimension) */
    int32_t cells;         /* per axis, a power of two (--cells) */
    int32_t cap;           /* total population cap (--cap) */
    double diffusion;      /* --diffusion */
    double shift;          /* --shift */
    double c1, c2, c3, c4; /* Schlogl constants (--c1 .. --c4; --linear s: 0, 0, 2 s, s) */
    int32_t smoother;      /* MK_BOX_GS or MK_BOX_JACOBI */
    int32_t sweeps;        /* per smoothing step (--sweeps) */
    int32_t cycles;        /* 1 or 2 (--cycles) */
    int32_t cycle_type;    /* MK_CYCLE_STANDARD or MK_CYCLE_K (--cycle-type) */
    int64_t terminal_rows; /* 0, or at most 8192 */
} mk_box_mg_params;

typedef struct mk_box_mg_result {
    int32_t iterations; /* cycles run */
    int32_t converged;  /* error < tolerance */
    double error;       /* |b - A x|_1 / shift after the last cycle (:884-887, :1215) */
    int64_t history;    /* number of errors the iteration produced (the first before any cycle) */
} mk_box_mg_result;

/* Level(p) with the hierarchy (rdme_full_multigrid.cpp:298-573) */
mk_status mk_box_mg_create(const mk_box_mg_params *params, mk_box_mg *out);
mk_status mk_box_mg_destroy(mk_box_mg mg);
mk_status mk_box_mg_info(mk_box_mg mg, int64_t *rows, int32_t *levels);
/* level l's states, voxels, stored off-diagonal entries and terminal kind (0 smoothed, 1
 * Thomas, 2 dense) */
mk_status mk_box_mg_level(mk_box_mg mg, int32_t level, int64_t *rows, int64_t *voxels,
                          int64_t *entries, int32_t *terminal);
/* the finest operator as a generator (borrowed: valid until mk_box_mg_destroy; do not destroy
 * it), for mk_gmres_box, mk_expv, mk_shifted_apply */
mk_status mk_box_mg_generator(mk_box_mg mg, mk_generator *gen);
/* Level::cycle(b, x) (:819-848): one cycle from the guess x, in place (vectors in HBM) */
mk_status mk_box_mg_cycle_device(mk_box_mg mg, const double *d_b, double *d_x);
/* Level::apply(b, x) (:883): one cycle from x = 0 */
mk_status mk_box_mg_apply_device(mk_box_mg mg, const double *d_b, double *d_z);
/* run() with --method mg (:1186-1212): cycles from the guess x until |b - A x|_1 / shift <
 * tolerance or max_iterations cycles; history (NULL allowed) receives up to history_cap errors,
 * the first before any cycle */
mk_status mk_box_mg_solve(mk_box_mg mg, const double *b, double *x, double tolerance,
                          int32_t max_iterations, double *history, int64_t history_cap,
                          mk_box_mg_result *result);
mk_status mk_box_mg_solve_device(mk_box_mg mg, const double *d_b, double *d_x, double tolerance,
                                 int32_t max_iterations, double *history, int64_t history_cap,
                                 mk_box_mg_result *result);
/* One time step of run() (:1174-1245) in place on the distribution p (HBM, box order): the
 * solve of (shift I - Q) p_next = shift p (backward Euler, step 1 / shift; also the two half
 * steps of the driver's Rannacher start) or, with crank_nicolson != 0, = 2 shift p -
 * (shift I - Q) p (Crank-Nicolson, step 2 / shift), from the guess p. p is replaced only when
 * the solve converges. */
mk_status mk_box_mg_step_device(mk_box_mg mg, double *d_p, int32_t crank_nicolson,
                                double tolerance, int32_t max_iterations,
                                mk_box_mg_result *result);
/* run() with --method mg-gmres (:1200-1213): flexible GMRES on (shift I - Q) x = b preconditioned
 * by one cycle (standard cycles only; the driver rejects the K-cycle there, :1441-1442). gen is
 * mk_box_mg_generator's; the driver's tolerance is tolerance * shift / sqrt(rows). */
mk_status mk_gmres_box(mk_generator gen, double shift, mk_box_mg mg, const double *b, double *x,
                       const mk_gmres_options *options, mk_solver_result *result);
mk_status mk_gmres_box_device(mk_generator gen, double shift, mk_box_mg mg, const double *d_b,
                              double *d_x, const mk_gmres_options *options,
                              mk_solver_result *result);
/* The standalone iteration replays one captured cycle (HIP graph) instead of launching its
 * kernels one by one (default on; MK_BOX_GRAPHS=0 turns it off at start; the K-cycle, which
 * reads its step lengths on the host, always launches eagerly). Same kernels and arguments:
 * results are identical. No reference counterpart. */
mk_status mk_box_mg_graphs_enable(int32_t on);
/* Levels of at most 2,048 states run a visit's pre-smoothing, residual and restriction in one
 * single-workgroup kernel, and its prolongation and post-smoothing in another, instead of about
 * 20 launches (four-colour Gauss-Seidel and Jacobi; default on, MK_BOX_SMALL=0 turns it off at
 * start). Same formulas and summation orders as the per-operation kernels. No reference
 * counterpart. */
mk_status mk_box_mg_small_enable(int32_t on);

/* ---- Krylov exponential (finite state projection) -------------------------------------
 * num::expv(t, Q.matrix(), p, krylov_dim, krylov_tol), called at include/markovkit/fsp.hpp:152
 * and :225 (numerics source absent): w = exp(t Q) v for the generator, computed as Sidje's
 * expv (Expokit) with an Arnoldi basis of krylov_dim vectors and local error control at
 * `tolerance`. Works on any generator handle (mk_generator_create_rdme or _jumps). */
typedef struct mk_expv_options {
    int64_t krylov_dim; /* 30 (krylov_fsp_options, fsp.hpp:205) */
    double tolerance;   /* 1e-8 */
} mk_expv_options;

typedef struct mk_expv_result {
    int64_t steps;    /* accepted time steps */
    int64_t rejected; /* rejected step sizes */
    double error;     /* sum of the local error estimates */
} mk_expv_result;

/* host vectors */
mk_status mk_expv(mk_generator gen, double t, const double *v, double *w,
                  const mk_expv_options *options, mk_expv_result *result);
/* v and w in HBM (may alias) */
mk_status mk_expv_device(mk_generator gen, double t, const double *d_v, double *d_w,
                         const mk_expv_options *options, mk_expv_result *result);

/* Batched exponentials w_j = exp(t_j Q) v_j of `count` vectors over one generator (v and w
 * host buffers, system-major: vector j at j * rows), each to options->tolerance with
 * options->krylov_dim basis vectors; results: count entries. No single-call counterpart (the
 * reference calls num::expv once per vector). Up to 256 vectors advance together: one Arnoldi
 * process per substep for all of them (the operator read once per row for 64 vectors) and
 * Expokit's step control per vector; a vector's result does not depend on the others in the
 * call (not bit for bit mk_expv's: the kernels differ). */
mk_status mk_expv_batch(mk_generator gen, int64_t count, const double *t, const double *v, double *w,
                        const mk_expv_options *options, mk_expv_result *results);

/* ---- batched resolvent solves ----------------------------------------------------------
 * No single-call counterpart: the reference solves each (shift, rhs) pair with its own
 * num::gmres call and sector_multigrid (examples/rdme_multigrid_benchmark.cpp:92-121). A batch
 * solves (s_j I - Q) x_j = b_j for j < n_systems together: gen is the shared generator, mg a
 * sector multigrid of the same box with the Jacobi or red-black smoother (red-black with at
 * least one pre- and one post-smoothing sweep); its level operators and transfers are shared,
 * its own shift is not used: every system holds the terminal inverses of its s_j, and the
 * hierarchy's precision selects the FP32 or FP64 preconditioner. gen and mg must outlive the
 * batch. Every system runs flexible GMRES(restart <= 64) with delayed CGS2 and stops
 * individually; the steps of an Arnoldi iteration are replayed from captured HIP graphs. */
mk_status mk_batch_create(mk_generator gen, mk_sector_mg mg, int64_t n_systems,
                          const double *shifts, mk_batch *out);
/* Systems of different sizes in one batch: system j is the leading rows[j] x rows[j] block of
 * s_j I - Q, rows[j] the end of a sector of the box (NULL: every system has all rows). For an
 * absorbing generator of a one-species box that block is the generator of the box of a smaller
 * cap (births out of its top sector are absorbed), and the leading sectors of the sector
 * multigrid are that box's hierarchy: systems of caps 3, 4 and 5 share one launch on the cap-5
 * box. B and X keep `rows` values per system (the box's row count); entries past rows[j] are
 * ignored on input and zero on output. */
mk_status mk_batch_create_sized(mk_generator gen, mk_sector_mg mg, int64_t n_systems,
                                const double *shifts, const int64_t *rows, mk_batch *out);
mk_status mk_batch_destroy(mk_batch batch);
/* Block Krylov for right-hand sides sharing a shift (block = 4; 1 restores per-system FGMRES):
 * every 4 consecutive systems (one shift, one size) run their first restart cycle as one block
 * FGMRES: a shared basis of 4-vector blocks, block classical Gram-Schmidt twice whose projections
 * V^T W (the group-diagonal 4 x 4 blocks) run on FP64 matrix cores (v_mfma_f64_16x16x4f64),
 * Cholesky QR of each new block and Givens rotations on the block Hessenberg least-squares
 * problem. A system the block cycle leaves unconverged (or a group whose residual block is rank
 * deficient) continues with its own FGMRES restarts from the block cycle's iterate. */
mk_status mk_batch_set_block(mk_batch batch, int32_t block);
/* graph replay of the batched Arnoldi steps of mk_batch_solve (default on; MK_BATCH_GRAPHS=0
 * turns it off at start); replays run the same kernels with the same arguments. While a step is
 * captured the runtime rejects synchronous (legacy-stream) HIP calls of other host threads:
 * turn replay off when other threads do device work during mk_batch_solve. */
mk_status mk_batch_graphs_enable(int32_t on);
/* B and X are system-major (system j's vector at j * rows), host buffers; X holds the initial
 * guesses on input. tolerances: n_systems absolute residual norms, or NULL for
 * options->tolerance; results: n_systems entries (orthogonalization is ignored). */
mk_status mk_batch_solve(mk_batch batch, const double *B, double *X, const double *tolerances,
                         const mk_gmres_options *options, mk_solver_result *results);
/* the same with B and X in HBM */
mk_status mk_batch_solve_device(mk_batch batch, const double *d_B, double *d_X,
                                const double *tolerances, const mk_gmres_options *options,
                                mk_solver_result *results);

/* ---- serving -----------------------------------------------------------------------------
 * No reference counterpart. A resolvent family is a generator with a pool of sector multigrid
 * handles (Jacobi or red-black) of the same box, kept resident.
 * Continuous batching: a family of at most the batching row limit gets a batch engine (one slot
 * per system; the generator, the first hierarchy and the slots' Krylov bases resident in HBM) run
 * by its own thread and HIP stream: at every FGMRES restart boundary it retires finished systems
 * and admits queued requests of the family into free slots (restart 50; the request's
 * max_iterations applies). Other families, and exponentials, run on `workers` host threads with
 * their own HIP streams, admitted first come first served when a worker is free, their device
 * footprint (Krylov basis and work vectors) fits hbm_budget_bytes beside the running ones and,
 * for a resolvent, a hierarchy of its family is free (re-targeted to the request's shift with
 * mk_sector_mg_set_shift). A request's result does not depend on the other requests running.
 * The server does not own the handles; they must outlive it. */
typedef struct mk_server_s *mk_server;
enum { MK_REQUEST_RESOLVENT = 0, MK_REQUEST_EXPV = 1 };

typedef struct mk_server_result {
    int32_t kind;             /* MK_REQUEST_RESOLVENT or MK_REQUEST_EXPV */
    int32_t status;           /* MK_OK or the failure's status */
    mk_solver_result solve;   /* resolvent */
    mk_expv_result expv;      /* exponential */
    double queued_s;          /* submission to start */
    double run_s;             /* start to finish */
} mk_server_result;

mk_status mk_server_create(int32_t workers, int64_t hbm_budget_bytes, mk_server *out);
/* finishes the queued requests, then joins the workers */
mk_status mk_server_destroy(mk_server server);
mk_status mk_server_add_family(mk_server server, mk_generator gen, int32_t n_mg,
                               const mk_sector_mg *mg, int32_t *family);
/* (s I - Q) x = b from x = 0 to the absolute tolerance; b is copied at submission */
mk_status mk_server_submit_resolvent(mk_server server, int32_t family, double shift,
                                     const double *b, double tolerance,
                                     const mk_gmres_options *options, int64_t *id);
/* the same for the leading `rows` rows of the family's box (a sector boundary; for an absorbing
 * one-species box, the box of a smaller cap); b has `rows` values, and so has the solution.
 * Batched families only (rows < the box's size) */
mk_status mk_server_submit_resolvent_sized(mk_server server, int32_t family, double shift,
                                           const double *b, int64_t rows, double tolerance,
                                           const mk_gmres_options *options, int64_t *id);
/* families added after this call with at most max_rows states (default 524,288) are served by
 * a batch engine of `slots` systems (default 64; 0 disables batching) */
mk_status mk_server_batching(mk_server server, int64_t max_rows, int32_t slots);
/* w = exp(t Q) v; v is copied at submission */
mk_status mk_server_submit_expv(mk_server server, mk_generator gen, double t, const double *v,
                                const mk_expv_options *options, int64_t *id);
/* blocks until request id is done; copies its solution to out (rows of its generator) */
mk_status mk_server_wait(mk_server server, int64_t id, double *out, mk_server_result *result);

/* ---- device memory and timing (no reference counterpart: HBM-resident callers) ------ */
mk_status mk_device_alloc(int64_t bytes, void **d_ptr);
mk_status mk_device_free(void *d_ptr);
mk_status mk_copy_to_device(void *d_dst, const void *src, int64_t bytes);
mk_status mk_copy_to_host(void *dst, const void *d_src, int64_t bytes);
mk_status mk_device_zero(void *d_ptr, int64_t bytes);
/* d_dst = d_src with the library's plain copy kernel (HBM calibration of byte counters) */
mk_status mk_device_copy(void *d_dst, const void *d_src, int64_t bytes);
mk_status mk_synchronize(void);
/* HIP-graph replay of repeated solver iterations and expv steps on operators of at most 262,144
 * rows (default off; MK_GRAPHS=1 in the environment turns it on at start). Replays run the same
 * kernels with the same arguments as eager launches, so results are identical either way. */
mk_status mk_graphs_enable(int32_t on);
/* Run the multigrid levels of at most 2048 states (and all coarser ones) as one workgroup-wide
 * kernel instead of one launch per operation (default on; MK_MG_TAIL=0 turns it off at start).
 * Same formulas and summation orders as the per-level kernels. */
mk_status mk_fused_tail_enable(int32_t on);
/* Where mk_generator_create_rdme assembles the generator (no reference counterpart): on the
 * device (default; MK_DEVICE_SETUP=0 turns it off), one thread per state, or on the host. The
 * operator, exit and leak rates are the same bit for bit. Boxes over 32 molecules per state,
 * 8 species or 65,535 voxels always assemble on the host. */
mk_status mk_device_setup_enable(int32_t on);
/* The fused tail keeps its levels' operators, transfers and vectors in LDS where they fit in
 * 150 KB (default on; MK_TAIL_LDS=0 turns it off). Results are bitwise the same either way. No
 * reference counterpart. */
mk_status mk_tail_lds_enable(int32_t on);
/* A red-black level's first half-sweep (from x = 0) computed by the kernel that makes the level's
 * right-hand side, the restriction or the coarse-cycle residual, instead of a launch of its own
 * (default on; MK_FUSED_FIRST=0 turns it off at start). Results are bitwise the same either way.
 * No reference counterpart. */
mk_status mk_fused_first_enable(int32_t on);
/* A direct coarse level (every row in a dense block of more than 128 states) solved and
 * prolonged into the level above in one launch: each coarse value is added to its colour-0
 * children as soon as its row of the inverse is applied (default on; MK_COARSE_FUSE=0 turns it
 * off at start). Results are bitwise the same either way. No reference counterpart. */
mk_status mk_coarse_fuse_enable(int32_t on);
/* FGMRES run whole in one workgroup (one kernel launch per solve) when the preconditioner is a
 * sector hierarchy whose levels all fit the LDS tail, the problem has at most 1024 states and
 * the restart is at most 64: the same iteration (classical Gram-Schmidt twice) and stopping
 * tests as the host-driven loop, with sums in another order, so not bitwise the same (default on;
 * MK_TINY_GMRES=0 turns it off at start). mk_batch_solve runs a batch of full-size systems at the
 * hierarchy's own shift the same way, one workgroup per system in one launch, each system's
 * result bitwise its single solve. No reference counterpart. */
mk_status mk_tiny_gmres_enable(int32_t on);
/* Workgroup-to-XCD mapping of the operator kernels (no reference counterpart): XCD j takes the
 * chunks j, j + 8, ... of `chunk` consecutive row blocks; 0 keeps workgroup b on row block b.
 * Default 16 (MK_XCD_CHUNK overrides). Process-wide; the autotuner sets it. */
mk_status mk_set_xcd_chunk(int32_t chunk);
mk_status mk_get_xcd_chunk(int32_t *chunk);
/* per-kernel-class timing with HIP events on the library stream. Classes, in bit order:
 * "shifted_spmv", "jacobi_first", "jacobi_sweep", "residual", "restrict", "prolong",
 * "terminal", "gmres_dot", "gmres_update", "gmres_other". mask: 0 = off, -1 = all, else the
 * classes whose bits are set. */
mk_status mk_profile_enable(int32_t mask);
mk_status mk_profile_reset(void);
mk_status mk_profile_read(const char *kernel, int64_t *launches, double *milliseconds,
                          double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* MARKOVKIT_AMD_H */
