// ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
// leg). An FP64 CPU restatement of the paper driver's whole-box multigrid,
// rdme/cpp/rdme_full_multigrid.cpp (paths relative to the markovkit repository root):
//
//   Space: weak compositions of total <= cap over V voxels, total ascending, then
//       lexicographic with voxel 0 most significant, rank by binomial counts      :66-152
//   Matrix: shift I - Q stored by destination (row i lists the jumps j -> i), the
//       diagonal inside the row in column order; lexicographic Gauss-Seidel
//       (backward order for post-smoothing) and weighted Jacobi (w = 2/3)         :154-199
//   PopulationBlockILU, structured form: diagonal-only ILU(0) pivots of each
//       population block; forward solve fused with the residual, backward solve
//       fused with the update                                                     :204-296
//   Level: the complete reaction-diffusion operator on every level (Schlogl rates
//       birth(n) = c3 vol + c1 n (n - 1) / (2 vol), death(n) = c4 n +
//       c2 n (n - 1) (n - 2) / (6 vol^2), jump = D cells^2, vol = 1 / V);
//       coarse level: cells / 2, D * 2; multinomial transfer weights normalised
//       per parent; Thomas solve of the one-voxel level                          :298-560
//   smooth dispatch (point smoothers, population-block sweeps forward / reverse)  :768-816
//   cycle (V for cycles = 1, W for cycles = 2; the exact bottom solve visited once) :819-848
//   K-cycle coarse solve (at most two GCR steps, the second skipped when the
//       first reduces |g|^2 by 1/16)                                              :849-878
//   the standalone iteration of run(): x0 = the initial distribution,
//       error = |b - A x|_1 / shift, cycles until error < tolerance              :1186-1212
//
// Two options of the GPU library that the driver does not have are restated here as well, so
// that the library's hierarchy has an exact CPU counterpart:
//   colour_gs: Gauss-Seidel in four-colour order. Colour = 2 (total mod 2) + (sum over
//       molecules of the coordinate sum of their voxel, mod 2). Diffusion changes the second
//       bit and a reaction that changes the population by one changes the first, so no
//       stored coupling joins two states of one colour; a sweep updates colours 0, 1, 2, 3
//       (3, 2, 1, 0 after the coarse correction), rows of one colour in any order.
//   terminal_rows > 0: the first level with at most that many states is solved exactly
//       (dense LU with partial pivoting) instead of being coarsened further.
//
// Pins (tests/test_oracle_box_mg.py): the recorded per-iteration residual histories, the
// iteration counts, masses and minimum probabilities of
// rdme/paper/data/full_multigrid/*.csv; the driver's --check invariants (column sums equal
// shift + leak, R P = I, R A P = A_coarse, :923-979); the Space order against the library's
// population box; a dense direct solve.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

namespace orc_box {

using Index = std::uint32_t;
using Vec = std::vector<double>;

enum class smoother_kind { gs = 0, jacobi = 1, block_ilu = 2, colour_gs = 3 };

struct params {
    int dimension = 1, cells = 4, cap = 40, sweeps = 2, cycles = 1;
    double diffusion = 1.0, shift = 0.05;
    double c1 = 0.0133, c2 = 0.0006, c3 = 0.6924125, c4 = 0.114725; // driver defaults :57
    smoother_kind smoother = smoother_kind::gs;
    bool kcycle = false;
    long terminal_rows = 0;
};

// Weak compositions of at most `cap` molecules over V voxels (rdme_full_multigrid.cpp:66-152)
struct space {
    int V = 0, cap = 0;
    std::vector<std::vector<std::uint64_t>> C; // C[n][r] = binomial(n, r), r <= V
    std::uint64_t size = 0;

    space(int voxels, int k) : V(voxels), cap(k), C(k + voxels + 1, std::vector<std::uint64_t>(voxels + 1, 0)) {
        for (int n = 0; n <= k + voxels; ++n) {
            C[n][0] = 1;
            for (int r = 1; r <= std::min(n, voxels); ++r) C[n][r] = C[n - 1][r - 1] + C[n - 1][r];
        }
        size = C[k + voxels][voxels];
        if (size > (std::uint64_t{1} << 31)) throw std::invalid_argument("oracle box: space too large");
    }
    // number of states with total population below `total`
    std::uint64_t offset(int total) const { return total == 0 ? 0 : C[total + V - 1][V]; }
    // lexicographic rank among compositions of the same total: compositions that agree on
    // voxels < v and hold fewer molecules in v come first; with m molecules left for the
    // rest = V - v - 1 later voxels there are C(m + rest, rest) of them
    Index rank(const std::vector<int> &n, int total) const {
        std::uint64_t r = offset(total);
        int left = total;
        for (int v = 0; v + 1 < V; ++v) {
            const int rest = V - v - 1;
            r += C[left + rest][rest] - C[left - n[v] + rest][rest];
            left -= n[v];
        }
        return static_cast<Index>(r);
    }
    // f(index, counts, total) over every state in index order
    template <class F>
    void each(F f) const {
        std::vector<int> n(V, 0);
        Index i = 0;
        for (int total = 0; total <= cap; ++total) {
            std::fill(n.begin(), n.end(), 0);
            n[V - 1] = total;
            while (true) {
                f(i++, n, total);
                // lexicographic successor: move one molecule from the last occupied voxel
                // q >= 1 into voxel q - 1 and the rest of q's molecules to the last voxel
                int q = V - 1;
                while (q >= 1 && n[q] == 0) --q;
                if (q < 1) break;
                const int s = n[q] - 1;
                n[q] = 0;
                ++n[q - 1];
                n[V - 1] += s;
            }
        }
        if (i != size) throw std::logic_error("oracle box: enumeration size");
    }
};

// shift I - Q by destination (rdme_full_multigrid.cpp:154-199)
struct matrix {
    std::vector<Index> ptr{0}, col, diag;
    Vec val;
    Index rows() const { return static_cast<Index>(diag.size()); }
    void apply(const Vec &x, Vec &y) const {
        for (Index r = 0; r < rows(); ++r) {
            double s = 0.0;
            for (Index j = ptr[r]; j < ptr[r + 1]; ++j) s += val[j] * x[col[j]];
            y[r] = s;
        }
    }
    double residual_l1(const Vec &b, const Vec &x) const {
        double norm = 0.0;
        for (Index r = 0; r < rows(); ++r) {
            double s = 0.0;
            for (Index j = ptr[r]; j < ptr[r + 1]; ++j) s += val[j] * x[col[j]];
            norm += std::abs(b[r] - s);
        }
        return norm;
    }
    void gs_row(Index r, const Vec &b, Vec &x) const {
        double s = b[r];
        for (Index j = ptr[r]; j < ptr[r + 1]; ++j)
            if (col[j] != r) s -= val[j] * x[col[j]];
        x[r] = s / val[diag[r]];
    }
    double entry(Index r, Index c) const {
        auto first = col.begin() + ptr[r], last = col.begin() + ptr[r + 1];
        auto it = std::lower_bound(first, last, c);
        return (it == last || *it != c) ? 0.0 : val[it - col.begin()];
    }
};

// structured population-block ILU(0) of rows [first, last) (rdme_full_multigrid.cpp:204-296)
struct block_ilu {
    Index first = 0, last = 0;
    std::vector<Index> lower_begin, upper_end; // offsets into the row: first in-block column,
                                               // one past the last in-block column
    Vec pinv;                                  // 1 / pivot
    block_ilu(const matrix &A, Index f, Index l) : first(f), last(l) {
        const Index n = l - f;
        lower_begin.resize(n);
        upper_end.resize(n);
        Vec pivot(n);
        for (Index r = f; r < l; ++r) {
            const Index i = r - f, d = A.diag[r];
            const auto b = A.col.begin();
            lower_begin[i] = static_cast<Index>(std::lower_bound(b + A.ptr[r], b + d, f) - b - A.ptr[r]);
            upper_end[i] = static_cast<Index>(std::lower_bound(b + d + 1, b + A.ptr[r + 1], l) - b - A.ptr[r]);
            double p = A.val[d];
            // the block graph is bipartite, so elimination only changes the diagonal:
            // p -= (a_rc / pivot_c) a_cr over in-block columns c < r
            for (Index j = A.ptr[r] + lower_begin[i]; j < d; ++j) {
                const Index c = A.col[j];
                const double a_cr = A.entry(c, r);
                p -= A.val[j] / pivot[c - f] * a_cr;
            }
            if (!(p > 0.0) || !std::isfinite(p)) throw std::runtime_error("oracle box: ILU pivot");
            pivot[i] = p;
        }
        pinv.resize(n);
        for (Index i = 0; i < n; ++i) pinv[i] = 1.0 / pivot[i];
    }
    void smooth(const matrix &A, const Vec &b, Vec &x, Vec &w) const {
        const Index n = last - first;
        for (Index i = 0; i < n; ++i) {
            const Index r = first + i, d = A.diag[r], lo = A.ptr[r] + lower_begin[i];
            double s = b[r];
            for (Index j = A.ptr[r]; j < lo; ++j) s -= A.val[j] * x[A.col[j]];
            for (Index j = lo; j < d; ++j) s -= A.val[j] * (x[A.col[j]] + w[A.col[j] - first]);
            for (Index j = d; j < A.ptr[r + 1]; ++j) s -= A.val[j] * x[A.col[j]];
            w[i] = s * pinv[i];
        }
        for (Index i = n; i-- > 0;) {
            const Index r = first + i;
            double s = 0.0;
            for (Index j = A.diag[r] + 1; j < A.ptr[r] + upper_end[i]; ++j) s += A.val[j] * w[A.col[j] - first];
            w[i] -= s * pinv[i];
            x[r] += w[i];
        }
    }
};

inline int voxel_count(const params &p) {
    int v = 1;
    for (int d = 0; d < p.dimension; ++d) v *= p.cells;
    return v;
}

struct level {
    params p;
    space S;
    matrix A;
    Vec leak;
    std::vector<int> population;
    std::vector<std::uint8_t> colour;
    std::vector<Index> by_colour, colour_begin; // rows of colour c: by_colour[colour_begin[c] ..
                                                // colour_begin[c + 1]), ascending
    std::vector<Index> block_offsets;
    std::vector<block_ilu> ilu;
    std::unique_ptr<level> coarse;
    std::vector<Index> parent;
    Vec weight;
    // exact solves: Thomas (one voxel) or dense LU (terminal_rows)
    int terminal = 0; // 0 none, 1 Thomas, 2 dense
    Vec lower, upper, pivot;
    Vec lu;
    std::vector<Index> piv;
    mutable Vec residual, coarse_rhs, coarse_solution, scratch, work;
    mutable Vec kv1, kw1, kv2, kw2, kr;

    explicit level(const params &pp) : p(pp), S(voxel_count(pp), pp.cap) {
        const int V = S.V;
        const Index n = static_cast<Index>(S.size);
        std::vector<std::vector<int>> nbr(V);
        std::vector<int> csum(V, 0);
        for (int v = 0; v < V; ++v) {
            int stride = 1;
            for (int d = 0; d < p.dimension; ++d) {
                const int c = (v / stride) % p.cells;
                csum[v] += c;
                if (c > 0) nbr[v].push_back(v - stride);
                if (c + 1 < p.cells) nbr[v].push_back(v + stride);
                stride *= p.cells;
            }
        }
        const double vol = 1.0 / V, jump = p.diffusion * p.cells * p.cells;
        auto birth = [&](double m) { return p.c3 * vol + p.c1 * m * (m - 1) / (2 * vol); };
        auto death = [&](double m) { return p.c4 * m + p.c2 * m * (m - 1) * (m - 2) / (6 * vol * vol); };
        leak.assign(n, 0.0);
        population.assign(n, 0);
        colour.assign(n, 0);
        A.diag.resize(n);
        std::vector<std::pair<Index, double>> row;
        S.each([&](Index i, std::vector<int> &x, int total) {
            population[i] = total;
            row.clear();
            double d = p.shift;
            int parity = 0;
            for (int v = 0; v < V; ++v) {
                const int m = x[v];
                parity += m * csum[v];
                d += birth(m) + death(m) + jump * m * static_cast<double>(nbr[v].size());
                if (total == p.cap) leak[i] += birth(m);
                if (m > 0) {
                    --x[v];
                    const double in = birth(m - 1);
                    if (in > 0) row.emplace_back(S.rank(x, total - 1), -in);
                    for (int w : nbr[v]) {
                        ++x[w];
                        row.emplace_back(S.rank(x, total), -jump * x[w]);
                        --x[w];
                    }
                    ++x[v];
                }
                if (total < p.cap) {
                    ++x[v];
                    const double in = death(m + 1);
                    if (in > 0) row.emplace_back(S.rank(x, total + 1), -in);
                    --x[v];
                }
            }
            colour[i] = static_cast<std::uint8_t>(((total & 1) << 1) | (parity & 1));
            row.emplace_back(i, d);
            std::sort(row.begin(), row.end());
            for (auto [c, a] : row) {
                if (c == i) A.diag[i] = static_cast<Index>(A.val.size());
                A.col.push_back(c);
                A.val.push_back(a);
            }
            A.ptr.push_back(static_cast<Index>(A.val.size()));
        });
        if (p.smoother == smoother_kind::colour_gs) {
            for (Index r = 0; r < n; ++r)
                for (Index j = A.ptr[r]; j < A.ptr[r + 1]; ++j)
                    if (A.col[j] != r && colour[A.col[j]] == colour[r])
                        throw std::invalid_argument("oracle box: a coupling joins two states of one colour");
            colour_begin.assign(5, 0);
            for (int c = 0; c < 4; ++c) {
                colour_begin[c + 1] = colour_begin[c];
                for (Index r = 0; r < n; ++r)
                    if (colour[r] == c) {
                        by_colour.push_back(r);
                        ++colour_begin[c + 1];
                    }
            }
        }
        residual.assign(n, 0.0);
        scratch.assign(n, 0.0);
        const bool dense = p.terminal_rows > 0 && static_cast<long>(n) <= p.terminal_rows;
        if (p.cells == 1 && !dense) {
            // Thomas coefficients of the tridiagonal one-voxel chain (rows ordered by total)
            terminal = 1;
            pivot.assign(n, 0.0);
            lower.assign(n, 0.0);
            upper.assign(n, 0.0);
            for (Index i = 0; i < n; ++i) {
                pivot[i] = A.val[A.diag[i]];
                for (Index j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
                    if (A.col[j] + 1 == i) lower[i] = A.val[j];
                    if (A.col[j] == i + 1) upper[i] = A.val[j];
                }
                if (i > 0) {
                    lower[i] /= pivot[i - 1];
                    pivot[i] -= lower[i] * upper[i - 1];
                }
            }
            return;
        }
        if (dense) {
            terminal = 2;
            factor_dense();
            return;
        }
        if (p.smoother == smoother_kind::block_ilu) {
            block_offsets.resize(p.cap + 2);
            for (int k = 0; k <= p.cap; ++k) block_offsets[k] = static_cast<Index>(S.offset(k));
            block_offsets[p.cap + 1] = n;
            for (int k = 0; k <= p.cap; ++k) ilu.emplace_back(A, block_offsets[k], block_offsets[k + 1]);
            work.assign(n, 0.0);
        }
        if (p.cells % 2) throw std::invalid_argument("oracle box: cells must be a power of two");
        params cp = p;
        cp.cells /= 2;
        cp.diffusion *= 2;
        coarse = std::make_unique<level>(cp);
        const Index nc = coarse->rows();
        coarse_rhs.assign(nc, 0.0);
        coarse_solution.assign(nc, 0.0);
        if (p.kcycle && !coarse->terminal) kv1 = kw1 = kv2 = kw2 = kr = coarse_rhs;
        // transfer: the coarse state adds the molecule counts of each 2^d block of voxels;
        // weight = multinomial probability of the fine arrangement given the coarse one,
        // normalised over the fine states of each coarse state
        std::vector<int> vparent(V);
        for (int v = 0; v < V; ++v) {
            int rest = v, stride = 1, cv = 0;
            for (int d = 0; d < p.dimension; ++d) {
                cv += ((rest % p.cells) / 2) * stride;
                rest /= p.cells;
                stride *= cp.cells;
            }
            vparent[v] = cv;
        }
        Vec lf(p.cap + 1);
        for (int k = 0; k <= p.cap; ++k) lf[k] = std::lgamma(k + 1.0);
        const double log_children = p.dimension * std::log(2.0);
        parent.assign(n, 0);
        weight.assign(n, 0.0);
        Vec sums(nc, 0.0);
        std::vector<int> cn(coarse->S.V);
        S.each([&](Index i, const std::vector<int> &x, int total) {
            std::fill(cn.begin(), cn.end(), 0);
            double lw = -total * log_children;
            for (int v = 0; v < V; ++v) {
                cn[vparent[v]] += x[v];
                lw -= lf[x[v]];
            }
            for (int k : cn) lw += lf[k];
            parent[i] = coarse->S.rank(cn, total);
            weight[i] = std::exp(lw);
            sums[parent[i]] += weight[i];
        });
        for (Index i = 0; i < n; ++i) weight[i] /= sums[parent[i]];
    }

    Index rows() const { return A.rows(); }

    void factor_dense() {
        const Index n = rows();
        lu.assign(static_cast<std::size_t>(n) * n, 0.0);
        for (Index r = 0; r < n; ++r)
            for (Index j = A.ptr[r]; j < A.ptr[r + 1]; ++j) lu[static_cast<std::size_t>(r) * n + A.col[j]] = A.val[j];
        piv.resize(n);
        for (Index k = 0; k < n; ++k) {
            Index pr = k;
            for (Index r = k + 1; r < n; ++r)
                if (std::abs(lu[std::size_t(r) * n + k]) > std::abs(lu[std::size_t(pr) * n + k])) pr = r;
            piv[k] = pr;
            if (pr != k)
                for (Index c = 0; c < n; ++c) std::swap(lu[std::size_t(k) * n + c], lu[std::size_t(pr) * n + c]);
            const double d = lu[std::size_t(k) * n + k];
            if (d == 0.0) throw std::runtime_error("oracle box: singular terminal level");
            for (Index r = k + 1; r < n; ++r) {
                const double f = lu[std::size_t(r) * n + k] / d;
                lu[std::size_t(r) * n + k] = f;
                if (f != 0.0)
                    for (Index c = k + 1; c < n; ++c) lu[std::size_t(r) * n + c] -= f * lu[std::size_t(k) * n + c];
            }
        }
    }

    void solve_exact(const Vec &b, Vec &x) const {
        const Index n = rows();
        std::copy(b.begin(), b.begin() + n, x.begin());
        if (terminal == 1) {
            for (Index i = 1; i < n; ++i) x[i] -= lower[i] * x[i - 1];
            x[n - 1] /= pivot[n - 1];
            for (Index i = n - 1; i > 0; --i) x[i - 1] = (x[i - 1] - upper[i - 1] * x[i]) / pivot[i - 1];
            return;
        }
        for (Index k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
        for (Index r = 1; r < n; ++r)
            for (Index c = 0; c < r; ++c) x[r] -= lu[std::size_t(r) * n + c] * x[c];
        for (Index r = n; r-- > 0;) {
            for (Index c = r + 1; c < n; ++c) x[r] -= lu[std::size_t(r) * n + c] * x[c];
            x[r] /= lu[std::size_t(r) * n + r];
        }
    }

    void smooth(const Vec &b, Vec &x, bool reverse) const {
        const Index n = rows();
        for (int sweep = 0; sweep < p.sweeps; ++sweep) {
            switch (p.smoother) {
            case smoother_kind::jacobi:
                A.apply(x, scratch);
                for (Index r = 0; r < n; ++r) x[r] += (2. / 3.) * (b[r] - scratch[r]) / A.val[A.diag[r]];
                break;
            case smoother_kind::gs:
                for (Index i = 0; i < n; ++i) A.gs_row(reverse ? n - 1 - i : i, b, x);
                break;
            case smoother_kind::colour_gs:
                for (int k = 0; k < 4; ++k) {
                    const int c = reverse ? 3 - k : k;
                    for (Index q = colour_begin[c]; q < colour_begin[c + 1]; ++q) A.gs_row(by_colour[q], b, x);
                }
                break;
            case smoother_kind::block_ilu:
                for (int visit = 0; visit <= p.cap; ++visit) {
                    const int k = reverse ? p.cap - visit : visit;
                    ilu[k].smooth(A, b, x, work);
                }
                break;
            }
        }
    }

    // one cycle on A x = b from the guess x (rdme_full_multigrid.cpp:819-848)
    void cycle(const Vec &b, Vec &x) const {
        if (terminal) {
            solve_exact(b, x);
            return;
        }
        smooth(b, x, false);
        A.apply(x, residual);
        std::fill(coarse_rhs.begin(), coarse_rhs.end(), 0.0);
        for (Index i = 0; i < rows(); ++i) coarse_rhs[parent[i]] += b[i] - residual[i];
        std::fill(coarse_solution.begin(), coarse_solution.end(), 0.0);
        if (p.kcycle && !coarse->terminal) {
            krylov_coarse_solve();
        } else {
            const int visits = coarse->terminal ? 1 : p.cycles;
            for (int v = 0; v < visits; ++v) coarse->cycle(coarse_rhs, coarse_solution);
        }
        for (Index i = 0; i < rows(); ++i) x[i] += weight[i] * coarse_solution[parent[i]];
        smooth(b, x, true);
    }

    // (rdme_full_multigrid.cpp:849-878)
    void krylov_coarse_solve() const {
        const level &c = *coarse;
        const Index n = c.rows();
        auto dot = [n](const Vec &a, const Vec &b) {
            double s = 0.0;
            for (Index i = 0; i < n; ++i) s += a[i] * b[i];
            return s;
        };
        const double gg = dot(coarse_rhs, coarse_rhs);
        if (!(gg > 0.0)) return;
        c.apply(coarse_rhs, kv1);
        c.A.apply(kv1, kw1);
        const double ww1 = dot(kw1, kw1);
        if (!(ww1 > 0.0)) return;
        const double a1 = dot(kw1, coarse_rhs) / ww1;
        for (Index i = 0; i < n; ++i) {
            coarse_solution[i] = a1 * kv1[i];
            kr[i] = coarse_rhs[i] - a1 * kw1[i];
        }
        if (dot(kr, kr) <= 0.0625 * gg) return;
        c.apply(kr, kv2);
        c.A.apply(kv2, kw2);
        const double beta = dot(kw2, kw1) / ww1;
        for (Index i = 0; i < n; ++i) {
            kv2[i] -= beta * kv1[i];
            kw2[i] -= beta * kw1[i];
        }
        const double ww2 = dot(kw2, kw2);
        if (!(ww2 > 0.0)) return;
        const double a2 = dot(kw2, kr) / ww2;
        for (Index i = 0; i < n; ++i) coarse_solution[i] += a2 * kv2[i];
    }

    // M b: one cycle from x = 0
    void apply(const Vec &b, Vec &x) const {
        std::fill(x.begin(), x.end(), 0.0);
        cycle(b, x);
    }

    const level *at(int l) const {
        const level *L = this;
        for (int k = 0; k < l && L; ++k) L = L->coarse.get();
        if (!L) throw std::invalid_argument("oracle box: no such level");
        return L;
    }
    int n_levels() const { return 1 + (coarse ? coarse->n_levels() : 0); }
};

// run() of the driver with --method mg (rdme_full_multigrid.cpp:1186-1212): x holds the guess;
// history receives the error before every cycle and after the last
struct solve_result {
    int iterations = 0;
    double error = 0.0;
};

inline solve_result standalone(const level &L, const Vec &b, Vec &x, double tolerance, int max_iterations,
                               Vec *history = nullptr) {
    solve_result r;
    r.error = L.A.residual_l1(b, x) / L.p.shift;
    while (r.iterations < max_iterations) {
        if (history) history->push_back(r.error);
        if (r.error < tolerance || !std::isfinite(r.error) || r.error > 1e8) break;
        L.cycle(b, x);
        ++r.iterations;
        r.error = L.A.residual_l1(b, x) / L.p.shift;
    }
    return r;
}

} // namespace orc_box
