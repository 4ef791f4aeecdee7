// Device kernels of the resolvent solve (gfx950, wave64) and their launchers.
//
// Sparse layout: every operator A = shift I - C is stored by destination ("gather" form): row i
// lists the jumps j -> i as (column j, positive rate), so (A x)_i = (shift + exit_i) x_i -
// sum rate x_j, summed over the row's entries in ascending column order.
//
// Two encodings of the same operator:
//  * coded sliced ELL (used whenever it fits): the rates of a Markov generator take few
//    distinct values, so an entry is one u32 (column in the low 24 bits, index into a rate
//    dictionary in the high 8 bits) and a row is a u8 length plus a u8 index into an
//    exit-rate dictionary. Rows are cut into slices of 64 (one wavefront); a slice stores its
//    entries column-major, entry k of row (64 s + l) at soff[s] + 64 k + l, padded to the
//    slice's longest row. Lane l owns one row: the entry loads are coalesced, and entry k of
//    64 neighbouring states is usually the same kind of jump, so the x gathers of one load
//    instruction fall into few cache lines.
//  * plain CSR-stream: u32 row pointers and columns, rates and exit rates in VT; a workgroup
//    streams a row block's entries, stages products in LDS, and each thread sums one row.
// Workgroups are mapped so that each XCD (8 per MI355X, dispatched round-robin) processes a
// contiguous range of rows, which keeps neighbouring rows' gathers in one L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mk::gpu {

using i64 = std::int64_t;
using u64 = std::uint64_t;
using u32 = std::uint32_t;
using u8 = std::uint8_t;

constexpr int kThreads = 256;      // workgroup size (4 wave64)
constexpr int kSlice = 64;         // rows per slice of the coded layout (one wavefront)
constexpr int kRowsPerBlock = 256; // rows per row block of the plain layout (one per thread)
constexpr int kTile = 4096;        // LDS product tile (entries) per plain row block
constexpr int kDotChunk = 8;       // basis vectors per multi-dot pass
constexpr int kDictMax = 256;
constexpr u32 kColBits = 24;
constexpr u32 kColMask = (1u << kColBits) - 1u;

// Workgroup-to-XCD mapping of the operator kernels. The hardware dispatches workgroup b to
// XCD b % 8. With chunk C > 0 the grid is padded to a multiple of 8 C and XCD j processes the
// logical chunks j, j + 8, j + 16, ... of C consecutive row blocks (neighbouring rows share an
// L2, while costly row ranges still spread over all XCDs); C = 0 keeps block b on row block b.
void set_xcd_chunk(int chunk);
int xcd_chunk();

// device view of one operator (shift I - C) without its shift
template <typename VT>
struct csr_view {
    i64 n = 0;
    i64 n_blocks = 0;            // workgroups per pass (row blocks or groups of 4 slices)
    // plain encoding
    const u32 *blocks = nullptr; // n_blocks + 1 row boundaries
    const u32 *ptr = nullptr;    // n + 1
    const u32 *col = nullptr;
    const VT *val = nullptr;
    const VT *exit = nullptr;    // n
    // coded sliced encoding (active when coded != 0; ent may be null when the operator has
    // no entries). A coded operator whose exit rates take more than 256 values may keep them
    // per row in `exit` instead (rcode and edict then unused, n_edict = 0).
    int coded = 0;
    i64 n_slices = 0;
    const u32 *soff = nullptr;   // n_slices + 1: first entry of each slice
    const u32 *ent = nullptr;    // column | rate code << 24, column-major per slice
    const u8 *rlen = nullptr;    // row lengths
    const u8 *rcode = nullptr;   // exit-rate codes
    const VT *vdict = nullptr;
    const VT *edict = nullptr;
    int n_vdict = 0;
    int n_edict = 0;
    int max_width = 0;           // longest row (0: unknown); picks the sliced kernel's chunk
    // a row subset (one colour of a red-black split): local row r is vector entry rowmap[r];
    // null for the identity
    const u32 *rowmap = nullptr;
};

// y = (shift + exit) o x - C x             (FP64 accumulate; x in XT)
// with b: y = b - ((shift + exit) o x - C x) and, if partial != null, one partial sum of y^2
// per workgroup in partial[0 .. n_blocks)
template <typename XT>
void shifted_apply(const csr_view<double> &A, double shift, const XT *x, const double *b,
                   double *y, double *partial, hipStream_t s);

// Jacobi smoothing on one multigrid level, working precision T:
//   first sweep from x = 0:  rhs = (T)(scale * rhs_in), stored to rhs_out if given;
//   x = w * rhs / (shift + exit); scale is read from device memory (1 when null)
template <typename T, typename IN>
void jacobi_first(const csr_view<T> &A, double shift, double omega, const IN *rhs_in, T *rhs_out,
                  T *x, hipStream_t s, const double *scale = nullptr);
//   sweep: x_out = x + w * (rhs - A x) / (shift + exit)
template <typename T>
void jacobi_sweep(const csr_view<T> &A, double shift, double omega, const T *rhs, const T *x,
                  T *x_out, hipStream_t s);
//   residual: r = rhs - A x
template <typename T>
void level_residual(const csr_view<T> &A, double shift, const T *rhs, const T *x, T *r,
                    hipStream_t s);
// y[r] = sum over row r's entries of rate * x[column] (restriction: rates 1, rows = coarse
// states, columns = their fine children in ascending order)
template <typename T>
void gather_sum(const csr_view<T> &R, const T *x, T *y, hipStream_t s);
// restriction by coarse rows: coarse[c] = sum of fine[children of c] in ascending order
template <typename T>
void restrict_children(i64 n_coarse, const u32 *cptr, const u32 *cidx, const T *fine, T *coarse,
                       hipStream_t s);
// prolongation fused with the correction: x[i] += weight[i] * coarse[coarse_of[i]] for the n
// rows i = rows[q] (all n rows when rows is null)
template <typename T>
void prolong_add(i64 n, const u32 *coarse_of, const T *weight, const T *coarse, T *x,
                 hipStream_t s, const u32 *rows = nullptr, const T *coarse2 = nullptr);
// restriction (gather_sum) or level residual with the next red-black level's first half-sweep
// fused in: besides y (r), cx[i] = y[i] / (shift + exit_i of Ac) where colour[i] == 0, the
// values rb_first computes from y
template <typename T>
void gather_sum_first(const csr_view<T> &R, const T *x, T *y, const csr_view<T> &Ac,
                      const std::uint8_t *colour, double shift, T *cx, hipStream_t s);
template <typename T>
void level_residual_first(const csr_view<T> &A, double shift, const T *rhs, const T *x, T *r,
                          const csr_view<T> &Ac, const std::uint8_t *colour, T *cx, hipStream_t s);
// dense terminal solves: x[start_b + i] += sum_j inv_b[i][j] * r[start_b + j] (assign: x[...] =);
// a thread per row, or a wave per row (lanes over columns) when a block exceeds kWideTerminal
constexpr i64 kWideTerminal = 128;
template <typename T>
void terminal_solve(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                    const T *inverse, i64 max_size, const T *r, T *x, hipStream_t s,
                    bool assign = false);
// a direct coarse level's dense solve fused with the prolongation into the level above: wave g
// computes coarse row c = start_b + i as terminal_solve does (assign), stores it to xc[c], and
// adds weight[f] * xc[c] to x_fine[f] for the children f of c (row c of the restriction R, fine
// rows ascending) whose fine colour is 0 -- the values prolong_add gives for the colour-0 rows
template <typename T>
void terminal_prolong(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                      const T *inverse, i64 max_size, const T *r, T *xc, const csr_view<T> &R,
                      const u8 *fine_colour, const T *weight, T *x_fine, hipStream_t s);

// Red-black Gauss-Seidel on one level. The level operator is split by colour into two
// row-mapped operators (rows of one colour, all their entries); a half-sweep updates one colour
// in place from the other colour's values:
// The levels of a red-black hierarchy number each sector's states colour 0 first, so a
// half-sweep streams contiguous row ranges; perm maps the top level's rows to the caller's
// order (row i is the caller's entry perm[i]).
//   gs_half: x[i] = (rhs[i] + sum rate x[j]) / (shift + exit_i) for the operator's rows;
//   res_zero[i] = 0 when res_zero is given (after the last pre-smoothing half-sweep the residual
//   of the rows just solved is zero by construction); nat[perm[i]] = x[i] when nat is given
template <typename T>
void gs_half(const csr_view<T> &Ac, double shift, const T *rhs, T *x, T *res_zero, hipStream_t s,
             T *nat = nullptr, const u32 *perm = nullptr);
//   rb_first: the first half-sweep from x = 0 on the full level: rhs[i] = (T)(scale *
//   rhs_in[perm[i]]) (perm null: identity), stored to rhs_out if given; x[i] = rhs[i] /
//   (shift + exit_i) for colour[i] == 0 (colour-1 entries of x are left for the following
//   half-sweep to set)
template <typename T, typename IN>
void rb_first(const csr_view<T> &A, const u8 *colour, double shift, const IN *rhs_in, T *rhs_out,
              T *x, hipStream_t s, const double *scale = nullptr, const u32 *perm = nullptr);
//   rb_first_scatter: the same values with rhs_in in the caller's order, read in that order:
//   row pos[j] takes rhs_in[j] (pos the inverse of perm), and exit_box[j] is the exit rate of that
//   row where it has colour 0, else negative; every load is in the caller's order, none depends
//   on another
template <typename T, typename IN>
void rb_first_scatter(i64 n, const T *exit_box, double shift, const IN *rhs_in, T *rhs_out, T *x,
                      hipStream_t s, const double *scale, const u32 *pos);
// out[i] = (T)(scale * in[perm[i]]); out[perm[i]] = in[i]
template <typename T, typename IN>
void permute_in(i64 n, const u32 *perm, const IN *in, T *out, hipStream_t s,
                const double *scale = nullptr);
template <typename T>
void permute_out(i64 n, const u32 *perm, const T *in, T *out, hipStream_t s);

// ILU(0) smoothing of one level (rank order; on the bipartite diffusion graph only the pivots
// differ from A). The triangular solves are level-scheduled: rows are grouped by dependency
// depth (a row's lower / upper neighbours lie in earlier groups), one group per launch, or a
// run of small groups in one workgroup with barriers between them. Groups [g0, g1) of rows[]
// (group k is rows[gptr[k] .. gptr[k + 1])):
// wide_rows > 0 launches the single group g0 (g1 = g0 + 1) of that many rows over a grid;
// wide_rows == 0 runs groups [g0, g1) in one workgroup.
//   ilu_forward:  u[i] = (r[i] + sum_{j < i} rate u[j]) * pinv[i]
//   ilu_backward: e[i] = u[i] + (sum_{j > i} rate e[j]) * pinv[i] (in place in u); x[i] += e[i]
template <typename T>
void ilu_forward(const csr_view<T> &A, const u32 *rows, const u32 *gptr, int g0, int g1,
                 i64 wide_rows, const T *pinv, const T *r, T *u, hipStream_t s);
template <typename T>
void ilu_backward(const csr_view<T> &A, const u32 *rows, const u32 *gptr, int g0, int g1,
                  i64 wide_rows, const T *pinv, T *u, T *x, hipStream_t s);

// aggregate-block smoother (multigrid.hpp:309-342), one aggregate's step at a time:
//   block_solve: for members g in [m0, m1): e = inverse_b r[members of b], x += e, escr = e
//   block_residual: r[i] -= (A escr)[i] for listed rows q in [q0, q1) (escr zero off members)
//   block_zero: escr = 0 on members [m0, m1)
template <typename T>
void block_solve(const u32 *members, const u32 *member_block, const u32 *block_first,
                 const u64 *inv_off, const T *inv, u64 m0, u64 m1, const T *r, T *x, T *escr,
                 hipStream_t s);
template <typename T>
void block_residual(const csr_view<T> &A, double shift, const u32 *rows, u64 q0, u64 q1,
                    const T *escr, T *r, hipStream_t s);
template <typename T>
void block_zero(const u32 *members, u64 m0, u64 m1, T *escr, hipStream_t s);
// Jacobi sweep on listed rows only (out = x + w (rhs - A x) / (shift + exit)), and copy back
template <typename T>
void sweep_rows(const csr_view<T> &A, double shift, double omega, const T *rhs, const T *x, T *out,
                const u32 *rows, u64 nrows, hipStream_t s);
template <typename T>
void copy_rows(const u32 *rows, u64 nrows, const T *src, T *dst, hipStream_t s);

// One multigrid level as the fused tail kernel reads it (all pointers in HBM)
template <typename T>
struct level_desc {
    csr_view<T> A;                     // level operator (shift I - C without the shift)
    csr_view<T> R;                     // restriction to the next level (n_next rows)
    const u32 *coarse_of = nullptr;    // prolongation parent (0xFFFFFFFF: none)
    const T *weight = nullptr;
    i64 n = 0, n_next = 0;
    int pre = 0, post = 0;
    i64 n_term = 0, term_max = 0;      // dense terminal blocks
    const u32 *tstart = nullptr, *tsize = nullptr;
    const u64 *toff = nullptr;
    const T *tinv = nullptr;
    T *rhs = nullptr, *x = nullptr, *xt = nullptr, *res = nullptr; // work vectors
    const u8 *colour = nullptr;        // red-black smoother: colour of every row, else null
    const u32 *perm = nullptr;         // top level of a red-black hierarchy: row -> caller's entry
    int direct = 0;                    // every row in a terminal block, no coarser level
    const u32 *crows[2] = {nullptr, nullptr}; // red-black: the rows of each colour, ascending
    i64 ncrows[2] = {0, 0};
    T *acc = nullptr;                  // the tail's top level under a W visit: its iterate
};
constexpr i64 kTailRows = 2048;  // levels of at most this many rows run in the fused tail
constexpr int kTailThreads = 1024;
// The V-cycle below and including level L[0] (Jacobi smoother, or red-black Gauss-Seidel on
// levels with a colour array; one coarse cycle) in one
// workgroup: the same per-row formulas and summation orders as the per-level kernels, phases
// separated by workgroup barriers. L[0]'s right-hand side is rhs, or (T)(rhs64 * scale) when
// rhs64 is given (scale from device memory, 1 when null); the result goes to out, else to
// L[0].x. cycles > 1 (a W visit from the level above; L[0] is then not the hierarchy's top):
// that many cycles on L[0], each later one from zero on the residual of the accumulated
// iterate in L[0].acc, as the caller's cycle loop does with the same formulas.
template <typename T>
void vcycle_tail(const level_desc<T> *L, int levels, double shift, double omega, const T *rhs,
                 const double *rhs64, const double *scale, T *out, hipStream_t s, int cycles = 1);
// The same tail with every array of its levels in LDS. image holds the arena's first
// image_bytes (a multiple of 16): the level descriptors, with every pointer field replaced by
// (LDS offset + 1) (0 for null), then every constant array (operators, transfers, terminal
// blocks, colours and colour row lists, top-level permutation) at its LDS offset; the work
// vectors follow in the arena up to lds_bytes. The result goes to out, else to top_x (HBM).
struct lds_copy { // a constant array and its place in the image
    const void *src = nullptr;
    u32 offset = 0;
    u32 bytes = 0;
};
constexpr u32 kTailLdsMax = 150 * 1024;
template <typename T>
bool vcycle_tail_lds_fits(u32 lds_bytes, int levels);
template <typename T>
void vcycle_tail_lds(const void *image, u32 image_bytes, u32 lds_bytes,
                     int levels, double shift, double omega, const T *rhs, const double *rhs64,
                     const double *scale, T *out, T *top_x, hipStream_t s, int cycles = 1);

// Flexible GMRES(m) on (shift I - Q) x = b in one workgroup, preconditioned by the LDS tail of a
// hierarchy that the tail holds whole (its image as for vcycle_tail_lds, top level included):
// one row per thread (n <= kTinyRows), m <= kTinyBasis, the Krylov basis V (FP64, m + 1 vectors)
// and the preconditioned vectors Z (precision T, m vectors) in HBM with stride ld. Classical
// Gram-Schmidt twice with workgroup sums in a fixed order; restarts, stopping tests and the
// result (iterations, residual, converged as doubles in result[0 .. 2]) as the host-driven loop.
constexpr int kTinyRows = kTailThreads;
constexpr int kTinyBasis = 64;
struct tiny_gmres_args {
    csr_view<double> A;        // the generator's gather rows
    double shift = 0.0;
    const double *b = nullptr;
    double *x = nullptr;       // the initial guess in, the solution out
    double *V = nullptr;
    void *Z = nullptr;
    i64 ld = 0;
    int n = 0, m = 0, max_iterations = 0;
    double tolerance = 0.0;
    double *result = nullptr;
    const void *image = nullptr; // the tail's image, its size and the LDS its levels take
    u32 image_bytes = 0, lds_bytes = 0;
    int levels = 0;
    double omega = 0.0;
    u64 n_ent = 0;             // the generator's stored entries (coded layout, padding included)
    // a batch of independent systems of the same operator and hierarchy, one workgroup each:
    // system k's b and x at k * sys_x, its basis at k * sys_v, its preconditioned vectors at
    // k * sys_z, its result at 3 k and its tolerance tols[k] (tolerance when tols is null)
    int systems = 1;
    i64 sys_x = 0, sys_v = 0, sys_z = 0;
    const double *tols = nullptr;
    // set by the launcher: LDS offsets of the generator's staged arrays and of z (0: not staged)
    u32 g_soff = 0, g_ent = 0, g_rlen = 0, g_rcode = 0, g_vdict = 0, g_edict = 0, g_exit = 0, g_z = 0;
};
// the dynamic LDS of the one-workgroup solve over a tail of tail_lds_bytes; 0 when above 160 KB
u32 gmres_tiny_lds(u32 tail_lds_bytes);
template <typename T>
void gmres_tiny(const tiny_gmres_args &a, hipStream_t s);

template <typename T>
void fill(i64 n, T value, T *x, hipStream_t s);
// out = (B)(in * scale), scale read from device memory (1 when null)
template <typename A, typename B>
void convert(i64 n, const A *in, B *out, hipStream_t s, const double *scale = nullptr);
template <typename T>
void axpy(i64 n, T alpha, const T *x, T *y, hipStream_t s); // y += alpha x
// plain copy of n 16-byte words (the byte-counter calibration kernel)
void copy16(i64 n, const void *src, void *dst, hipStream_t s);

// ---- Krylov vector kernels (FP64) ----
int reduction_blocks(i64 n);
// partial[k * nb + block] = sum over the block's rows of V_k . w, k in [0, count)
void multi_dot(const double *V, i64 ld, int count, const double *w, i64 n, double *partial,
               int nb, hipStream_t s);
// w -= sum_k h[k] V_k; if partial != null, partial[block] = sum of w^2 over the block's rows
void basis_update(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                  double *partial, int nb, hipStream_t s);
// out[k] = sum_b partial[k * nb + b] (fixed order), k in [0, count)
void reduce_partials(const double *partial, int nb, int count, double *out, hipStream_t s);
// v = w / sqrt(norm2[0])
void scale_by_norm(i64 n, const double *w, const double *norm2, double *v, hipStream_t s);
// Fused classical Gram-Schmidt, twice (CGS2), for count <= kFusedBasis basis vectors: pass 1
// projections by multi_dot; then the update fused with the pass-2 projections (waves split the
// basis, k = 4 kk + wave, lanes own the rows of 64-row tiles, V values reused from registers);
// then the pass-2 update fused with the normalisation.
constexpr int kFusedBasis = 64;
//   w <- w - V h, then the pass-2 projections of the new w and its squared norm:
//   partial[k * nb + block] = sum_rows V_k w (k < count), partial[count * nb + block] = sum w^2
void update_dot(double *w, const double *V, i64 ld, int count, const double *h, i64 n,
                double *partial, int nb, hipStream_t s);
//   v = (w - V c) / beta with beta^2 = cw[count] - sum_k c_k^2 (c = cw[0..count)). When the
//   second pass removes more than a quarter of |w|^2 the norm is not trusted: beta = 1 and
//   flag = 1, and the caller normalises explicitly. beta_flag[0] = beta, beta_flag[1] = flag.
void update_scale(const double *w, const double *V, i64 ld, int count, const double *cw, i64 n,
                  double *v, double *beta_flag, hipStream_t s);
// Delayed CGS2 (one reduction per Arnoldi step; count < kFusedBasis). V[0 .. count) is the
// final orthonormal basis Q and V[count] the next candidate u, orthogonalised once against Q;
// sc = Q^T u, tg[0 .. count) = Q^T w and tg[count] = u^T w for the new operator image w, and
// beta = |u - Q sc| (read from device memory). Finalises q = (u - Q sc) / beta in V[count],
// writes eta = q^T w to eta_out[0] and the new candidate u' = w - Q (Q^T w) - eta q to
// V[count + 1], with the partial sums of its projections Q_k . u', q . u' and of |u'|^2 at
// partial[k * nb + block] for k = 0 .. count + 1.
void dcgs_update(double *V, i64 ld, int count, const double *sc, const double *tg,
                 const double *beta, const double *w, i64 n, double *partial, int nb,
                 double *eta_out, hipStream_t s);
// x += sum_k y[k] Z_k (Z stored in ZT)
template <typename ZT>
void combine(double *x, const ZT *Z, i64 ld, int count, const double *y, i64 n, hipStream_t s);

} // namespace mk::gpu
