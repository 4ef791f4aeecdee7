// Device kernels of batched resolvent solves: many systems (shift s_j, right-hand side b_j)
// over one state space, advanced together.
//
// Layout: a batched vector stores state i of system j at X[i * Bp + j], with Bp the number of
// systems padded to a multiple of 64. A wavefront works on one state row for 64 consecutive
// systems (lane = system): the row's operator entries are wave-uniform (scalar loads, read once
// for all systems of the wave), and every gather x[col * Bp + j] is one contiguous 256-512 B
// access. The operators are the single-system ones (coded sliced or plain, see kernels.hpp);
// only the diagonal shift varies with the system.
//
// limit (when not null) holds one int per system: lane j works on the rows r < limit[j] of the
// level the kernel runs on and loads and stores nothing past them (0: an inactive system), with
// zero partial sums there. A system of fewer rows (a leading block of whole sectors) thus shares
// the launch with larger ones. Reductions run in a fixed order.
#pragma once

#include "kernels.hpp"

namespace mk::gpu::batch {

constexpr int kLanes = 64; // systems per wavefront
constexpr int kBlock = 4;  // systems per block of the block Krylov method (one shift, 4 rhs)

// row partitions used by the batched reductions for an n-row operand
int parts(i64 n);
// more partitions for an operand of few rows: 1024 (at least 16 rows each), so that a reduction
// launch has enough workgroups even for one lane group; it depends on n only, so a system's sums
// do not depend on how many systems share the launch
int parts_wide(i64 n);

// partial sums of the residual kernel: one per 4-row block
i64 residual_parts(i64 n);

// y = (s_j + exit) o x - C x                       (mode shifted)
// y = b - ((s_j + exit) o x - C x), partial sums of y^2 per system (mode shifted_residual):
//   partial[p * Bp + j] for p < residual_parts(n)
// (scale, when given and b is null: lane j of y is multiplied by scale[j])
template <typename XT>
void shifted_apply(const csr_view<double> &A, const double *shifts, const XT *x, const double *b,
                   double *y, double *partial, i64 Bp, const int *limit, hipStream_t s,
                   const double *scale = nullptr);

// multigrid level operations in precision T (see kernels.hpp for the single-system forms)
//   first sweep from zero: rhs = (T)(in * scale_j) (stored to rhs_out if given),
//   x = w rhs / (s_j + exit)
template <typename T, typename IN>
void jacobi_first(const csr_view<T> &A, const double *shifts, double omega, const IN *in,
                  const double *scale, T *rhs_out, T *x, i64 Bp, const int *limit,
                  hipStream_t s);
template <typename T>
void jacobi_sweep(const csr_view<T> &A, const double *shifts, double omega, const T *rhs,
                  const T *x, T *x_out, i64 Bp, const int *limit, hipStream_t s);
template <typename T>
void level_residual(const csr_view<T> &A, const double *shifts, const T *rhs, const T *x, T *r,
                    i64 Bp, const int *limit, hipStream_t s);
template <typename T>
void gather_sum(const csr_view<T> &R, const T *x, T *y, i64 Bp, const int *limit, hipStream_t s);
// x[i] += weight[i] coarse[coarse_of[i]] for the rows i = rows[q], q < n (all n rows when rows
// is null)
template <typename T>
void prolong_add(i64 n, const u32 *coarse_of, const T *weight, const T *coarse, T *x, i64 Bp,
                 const int *limit, hipStream_t s, const u32 *rows = nullptr);
// red-black Gauss-Seidel on a level numbered colour 0 first (see gs_half and rb_first in
// kernels.hpp): the first half-sweep from x = 0, rhs taken from in[perm[i]] (scaled per system)
template <typename T, typename IN>
void rb_first(const csr_view<T> &A, const u8 *colour, const double *shifts, const IN *in,
              const double *scale, T *rhs_out, T *x, const u32 *perm, i64 Bp, const int *limit,
              hipStream_t s);
// one colour's half-sweep in place over its row-mapped operator; res_zero rows cleared; nat
// (with perm) receives the rows in the caller's order
template <typename T>
void gs_half(const csr_view<T> &Ac, const double *shifts, const T *rhs, T *x, T *res_zero, i64 Bp,
             const int *limit, hipStream_t s, T *nat = nullptr, const u32 *perm = nullptr);
// dense terminal blocks with one inverse set per distinct shift: system j uses
// inverse + shift_index[j] * per_shift
template <typename T>
void terminal_solve(i64 n_blocks, const u32 *start, const u32 *size, const u64 *offset,
                    const T *inverse, u64 per_shift, const int *shift_index, i64 max_size,
                    const T *r, T *x, i64 Bp, const int *limit, hipStream_t s);

// ---- Krylov (FP64), count basis vectors V_k = V + k * n * Bp
// partial[p * (count * Bp) + k * Bp + j] = sum over partition p's rows of V_k[i][j] w[i][j]
void multi_dot(const double *V, int count, const double *w, i64 n, i64 Bp, double *partial,
               int np, const int *limit, hipStream_t s);
// out[o] = sum_p partial[p * m + o] for o < m (fixed order); with tmp (reduce_tmp_size doubles)
// a long reduction runs in two stages over groups of 256 partitions
void reduce_partials(const double *partial, int np, i64 m, double *out, hipStream_t s,
                     double *tmp = nullptr);
i64 reduce_tmp_size(int np, i64 m);
// delayed CGS2 step per system (count < kFusedBasis): V[count] holds u_j, w = A z_j;
// sc[k * Bp + j] = (Q^T u)_k, tg[k * Bp + j] = (Q^T w)_k for k < count and tg[count * Bp + j] =
// u^T w; beta[j] = |u - Q s|. Writes q to V[count], eta_j to eta[j], u' to V[count + 1], and
// partial[p * ((count + 2) * Bp) + k * Bp + j] for k <= count + 1 as in the single-system form
void dcgs_update(double *V, int count, const double *sc, const double *tg, const double *beta,
                 const double *w, i64 n, i64 Bp, double *partial, int np, double *eta,
                 const int *limit, hipStream_t s);
// delayed CGS2 without a host step: per lane, from the candidate's projections sn[0 .. count]
// and |u|^2 = sn[count + 1], flag[j] = 1 where |s|^2 > |u|^2 / 4 (the candidate needs an explicit
// second projection; lflag[j] = limit[j] there, else 0)
void dcgs_flags(const double *sn, int count, i64 Bp, const int *limit, int *lflag, double *flag,
                hipStream_t s);
// then the normalisation from sn1, or sn2 (the second projection) where flagged (flag not
// null): scur = s, beta_out = beta = sqrt(max(|u|^2 - |s|^2, 0)), dbeta = beta (infinity when 0),
// scale = 1 / beta (0 when 0), beta2 = beta^2; flag_out (when given) receives the flag test on
// sn1 (a step taken without second projections records where one was due)
void dcgs_norms(const double *sn1, const double *sn2, const double *flag, int count, i64 Bp,
                const int *limit, double *scur, double *beta_out, double *dbeta, double *scale,
                double *beta2, hipStream_t s, double *flag_out = nullptr);
// w[i][j] -= sum_{k < count} h[k * Bp + j] V_k[i][j]
void basis_update(double *w, const double *V, int count, const double *h, i64 n, i64 Bp,
                  const int *limit, hipStream_t s);
// x[i][j] += sum_{k < cnt[j]} y[k * Bp + j] Z_k[i][j]
template <typename ZT>
void combine(double *x, const ZT *Z, const int *cnt, const double *y, i64 n, i64 Bp,
             hipStream_t s);
// ---- block Krylov: lanes 4 g .. 4 g + 3 form group g, which shares one basis
// partial[p * (count Bp 4) + (i Bp + j) 4 + a] = sum over partition p's rows of V_i[r][4 g + a]
// W[r][j], i < count, g = j / 4: the group-diagonal blocks of V^T W on FP64 matrix cores
// (v_mfma_f64_16x16x4f64)
void block_dot(const double *V, int count, const double *W, i64 n, i64 Bp, double *partial,
               int np, const int *limit, hipStream_t s);
// out[r][j] = sign sum_{i < count, a < 4} V_i[r][4 g + a] C[(i Bp + j) 4 + a], added to out unless
// assign; rows r < limit[j]; counts (when given) replaces count per lane
template <typename IN>
void block_apply(double *out, const IN *V, int count, const double *C, i64 n, i64 Bp, double sign,
                 bool assign, const int *limit, hipStream_t s, const int *counts = nullptr);
// y[r][j] = x[r][j] / sqrt(norm2[j]) for rows r < limit[j] (0 where norm2[j] is not positive)
void scale_lanes(const double *x, const double *norm2, double *y, i64 n, i64 Bp, const int *limit,
                 hipStream_t s);
// one lane j of a state-major vector: dst[r * Bp + j] = src[r] for r < rows, 0 for rows <= r < n
void put_lane(const double *src, i64 rows, double *dst, i64 n, i64 Bp, i64 j, hipStream_t s);
// dst[r] = src[r * Bp + j] for r < rows
void get_lane(const double *src, i64 rows, i64 Bp, i64 j, double *dst, hipStream_t s);
// X_k[r * Bp + j] = 0 for rows[j] <= r < n and k < count (vectors ld elements apart), on the
// lanes with sel[j] != 0 (every lane when sel is null)
template <typename T>
void clear_tails(T *X, int count, i64 ld, i64 n, i64 Bp, const int *rows, const int *sel,
                 hipStream_t s);
// system-major (src[j * n + i], j < nsys) to state-major (dst[i * Bp + j], zero for j >= nsys)
void pack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp, hipStream_t s);
// state-major to system-major for j < nsys
void unpack(const double *src, i64 n, i64 nsys, double *dst, i64 Bp, hipStream_t s);

} // namespace mk::gpu::batch
