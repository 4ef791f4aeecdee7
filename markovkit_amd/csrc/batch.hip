// Batched resolvent solves: (s_j I - Q) x_j = b_j for many systems j over one population box,
// each preconditioned by the sector multigrid of its own shift, all advanced together on the
// device. The reference solves such families one num::gmres call at a time, each with its own
// sector_multigrid (examples/rdme_multigrid_benchmark.cpp:92-121 per shift).
//
// The shift enters the multigrid only through the diagonal of every level operator and the
// dense terminal inverses, so one hierarchy serves all systems: level operators and transfers
// are shared (read once per row for all systems of a wavefront) and every system slot holds the
// terminal inverses of its own shift. Each system runs flexible GMRES(m) with delayed CGS2 and its
// own Givens least-squares problem.
//
// Systems of different sizes share the launch: system j may be the leading rows[j] x rows[j]
// block of s_j I - Q, rows[j] a boundary between whole sectors. For an absorbing generator of a
// one-species box, the leading block of sectors 0..k is exactly the generator of the box of cap
// k (births out of sector k are absorbed in both), and the sector multigrid is block-diagonal by
// sector, so its leading sectors are that box's hierarchy. Lane j works on the rows below its
// limit on every level; the vectors it gathers past them stay zero.
//
// The batch is an engine of slots: systems are admitted to free slots and retired at restart
// boundaries (continuous batching, used by the request server); mk_batch_solve admits all its
// systems and steps until they are done. A lane's arithmetic does not depend on the other lanes,
// so a system's result is the same alone and in any mix. In mk_batch_solve the operations of one
// Arnoldi step (preconditioner cycle, operator, projections, update) are captured in a HIP graph
// per step index and replayed; the server's engines launch them eagerly.
#include "batch_kernels.hpp"
#include "runtime.hpp"

#include <cstring>
#include <limits>
#include <map>
#include <type_traits>

using namespace mk::rt;
namespace bk = mk::gpu::batch;

namespace {

template <typename T>
struct blevel {
    dbuf rhs, x, xt, res, crhs, cx; // n_level x Bp work vectors
    dbuf tinv;                      // terminal inverses: per_shift values for every slot
    u64 per_shift = 0;
};

// the FGMRES state of one slot
struct slot_state {
    i64 rows = 0;        // rows of the system; 0: a free slot
    double shift = 1.0, tol = 0.0;
    i64 max_iterations = 0;
    bool fresh = false;  // admitted; its initial residual is not taken yet
    bool active = false; // iterating
    bool done = false;   // finished (converged, non-finite or out of iterations)
    mk_solver_result res{};
    double beta = 0.0, bj = 0.0;
    i64 cols = 0;
    char flag = 0;
};

// the batched step graphs are replayed by default (MK_BATCH_GRAPHS=0 launches eagerly)
std::atomic<int> &batch_graph_switch() {
    static std::atomic<int> on{[] {
        const char *env = std::getenv("MK_BATCH_GRAPHS");
        return (env && env[0] == '0') ? 0 : 1;
    }()};
    return on;
}

} // namespace

struct mk_batch_s {
    mk::u64 uid = next_uid();
    mk_generator_s *gen = nullptr;
    mk_sector_mg_s *mg = nullptr;
    i64 n = 0, Bp = 0, nsys = 0;
    int levels = 0;
    std::vector<slot_state> slot;
    std::vector<hessenberg_ls> ls;
    std::vector<std::vector<int>> lim;  // [level][slot]: the slot's rows on the level
    std::vector<i64> sector_end;        // level-0 row where each sector ends (admissible sizes)
    dbuf d_shift, d_sidx, d_lim, d_rows, d_sel;
    int *h_lim = nullptr;               // pinned: the levels x Bp limits of the current step
    std::map<mk::u64, std::vector<std::vector<double>>> inverses; // terminal inverses per shift
    std::vector<blevel<float>> B32;
    std::vector<blevel<double>> B64;
    // Krylov workspace for restart m
    i64 m = 0;
    int np = 1;
    dbuf V, Z, w, partial, ptmp, rpart, rtmp, hd, d_b, d_x, d_io, d_scale, d_beta, d_cnt, d_y;
    double *host = nullptr; // pinned: hd mirror, y, scale, beta
    int *ihost = nullptr;   // pinned: cnt, rows, sel
    graph_cache graphs;
    // block Krylov: with block = 4, every group of 4 consecutive slots with one shift and one row
    // count runs its first restart cycle as one block FGMRES (shared basis, block Gram-Schmidt on
    // matrix cores); lanes it leaves unconverged continue with per-slot restarts
    int block = 1;
    std::vector<char> block_pending; // per group: the block cycle is still to run
    int npb = 1;                     // row partitions of the block projections
    dbuf d_bh, bpartial;             // block coefficients (H1, H2, Gram, C) and partial sums
    dbuf bscratch, bscratch2;        // block vectors of the Cholesky QR passes
    double *bh = nullptr;            // pinned mirror of H1, H2 and the Gram blocks
    dbuf d_bcnt;
    // one-workgroup solves (every system in its own workgroup of one launch): bases, results and
    // tolerances
    dbuf tV, tZ, tres, ttol;
    // step graphs are captured only where no other host thread of the process runs device work:
    // during a capture the runtime rejects synchronous (legacy-stream) calls of other threads,
    // so the server's engines, which run beside other threads, launch eagerly
    bool capture = true;
    ~mk_batch_s() {
        graphs.clear();
        if (host) (void)hipHostFree(host);
        if (ihost) (void)hipHostFree(ihost);
        if (h_lim) (void)hipHostFree(h_lim);
        if (bh) (void)hipHostFree(bh);
    }
};

namespace {

// device scalar slots of the workspace (each (m + 2) x Bp): projections t, the two s buffers,
// then eta and the residual norms (Bp each)
struct slots {
    i64 stride;
    i64 t() const { return 0; }
    i64 s(int which) const { return (1 + which) * stride; }
    i64 eta() const { return 3 * stride; }
    i64 rn(i64 Bp) const { return 3 * stride + Bp; }
    i64 total(i64 Bp) const { return 3 * stride + 2 * Bp; }
};

template <typename T>
std::vector<level_dev<T>> &levels_of(mk_batch_s &B);
template <>
std::vector<level_dev<float>> &levels_of<float>(mk_batch_s &B) { return B.mg->L32; }
template <>
std::vector<level_dev<double>> &levels_of<double>(mk_batch_s &B) { return B.mg->L64; }
template <typename T>
std::vector<blevel<T>> &blevels_of(mk_batch_s &B);
template <>
std::vector<blevel<float>> &blevels_of<float>(mk_batch_s &B) { return B.B32; }
template <>
std::vector<blevel<double>> &blevels_of<double>(mk_batch_s &B) { return B.B64; }

template <typename T>
void build_blevels(mk_batch_s &B) {
    auto &L = levels_of<T>(B);
    auto &BL = blevels_of<T>(B);
    BL.clear();
    BL.resize(L.size());
    const bool cycles = B.mg->opt.coarse_cycles > 1;
    for (std::size_t l = 0; l < L.size(); ++l) {
        const auto &lv = L[l];
        auto &bl = BL[l];
        const std::size_t vb = static_cast<std::size_t>(std::max<i64>(lv.n, 1)) * B.Bp * sizeof(T);
        bl.rhs = dbuf(vb);
        bl.x = dbuf(vb);
        bl.xt = dbuf(vb);
        bl.res = dbuf(vb);
        if (cycles) {
            bl.crhs = dbuf(vb);
            bl.cx = dbuf(vb);
        }
        if (lv.n_term > 0) {
            bl.per_shift = lv.h_tq.size();
            bl.tinv = dbuf(bl.per_shift * B.Bp * sizeof(T));
        }
    }
}

// the terminal inverses of every level for one shift (cached by the shift's bits)
const std::vector<std::vector<double>> &inverses_for(mk_batch_s &B, double shift) {
    auto it = B.inverses.find(bits(shift));
    if (it != B.inverses.end()) return it->second;
    std::vector<std::vector<double>> per_level(B.levels);
    auto fill = [&](const auto &L) {
        for (int l = 0; l < B.levels; ++l) {
            const auto &lv = L[l];
            if (lv.n_term == 0) continue;
            auto &inv = per_level[l];
            inv.assign(lv.h_tq.size(), 0.0);
            for (i64 b = 0; b < lv.n_term; ++b) {
                const i64 m = lv.h_tsize[b];
                std::vector<double> a(lv.h_tq.begin() + lv.h_toff[b], lv.h_tq.begin() + lv.h_toff[b] + m * m);
                for (i64 i = 0; i < m; ++i) a[i * m + i] += shift;
                std::vector<double> out;
                if (!mk::invert_dense(std::move(a), m, out))
                    throw std::runtime_error("population_multigrid: singular block");
                std::copy(out.begin(), out.end(), inv.begin() + lv.h_toff[b]);
            }
        }
    };
    if (B.mg->precision == MK_FP32) {
        fill(B.mg->L32);
    } else {
        fill(B.mg->L64);
    }
    if (B.inverses.size() >= 4096) B.inverses.clear();
    return B.inverses.emplace(bits(shift), std::move(per_level)).first->second;
}

// slot j takes shift s and the leading `rows` rows: limits per level and terminal inverses
template <typename T>
void put_slot_inverses(mk_batch_s &B, i64 j, const std::vector<std::vector<double>> &inv) {
    auto &BL = blevels_of<T>(B);
    for (int l = 0; l < B.levels; ++l) {
        if (BL[l].per_shift == 0) continue;
        std::vector<T> v(inv[l].begin(), inv[l].end());
        HIPCHK(hipMemcpy(static_cast<T *>(BL[l].tinv.p) + j * BL[l].per_shift, v.data(),
                         v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
}

void set_slot(mk_batch_s &B, i64 j, double shift, i64 rows) {
    if (!(shift > 0.0) || !std::isfinite(shift))
        throw std::invalid_argument("population_multigrid: shift must be positive");
    const auto k = std::find(B.sector_end.begin(), B.sector_end.end(), rows);
    require(k != B.sector_end.end(),
            "batch: a system's rows must end at a sector boundary of the box");
    const i64 sectors = (k - B.sector_end.begin()) + 1;
    for (int l = 0; l < B.levels; ++l) {
        i64 r = 0;
        for (i64 q = 0; q < sectors; ++q) r += B.mg->sector_rows[l][q];
        B.lim[l][j] = static_cast<int>(r);
    }
    auto &S = B.slot[j];
    S.rows = rows;
    S.shift = shift;
    const auto &inv = inverses_for(B, shift);
    if (B.mg->precision == MK_FP32) {
        put_slot_inverses<float>(B, j, inv);
    } else {
        put_slot_inverses<double>(B, j, inv);
    }
    HIPCHK(hipMemcpy(dp<double>(B.d_shift) + j, &shift, sizeof(double), hipMemcpyHostToDevice));
}

// one batched cycle of level l from a zero guess (population_multigrid::cycle,
// multigrid.hpp:201-228), the per-system form of run_level in api.hip; lims holds the row limits
// of every level (levels x Bp, 0 for masked lanes)
template <typename T>
T *brun_level(mk_batch_s &B, std::size_t l, const T *rhs, const double *rhs64, const double *scale,
              T *out, const int *lims) {
    auto &L = levels_of<T>(B);
    auto &BL = blevels_of<T>(B);
    auto &lv = L[l];
    auto &bl = BL[l];
    auto &P = prof();
    const hipStream_t s = stream();
    const auto &mg = *B.mg;
    const double omega = mg.opt.jacobi_weight;
    const double *sh = dp<double>(B.d_shift);
    const int *sidx = dp<int>(B.d_sidx);
    const i64 n = lv.n, Bp = B.Bp;
    const int *lim = lims + static_cast<i64>(l) * Bp;
    const double ts = double(sizeof(T)), nb = double(n) * double(Bp);
    const double sweep_bytes = op_bytes(lv.A, 0, 0, 0) + nb * 4.0 * ts;
    T *cur = dp<T>(bl.x);
    T *other = dp<T>(bl.xt);
    T *res = dp<T>(bl.res);
    auto terminal = [&](const T *r, T *x) {
        P.run(K_TERMINAL, 0.0, [&] {
            bk::terminal_solve<T>(lv.n_term, dp<u32>(lv.tstart), dp<u32>(lv.tsize), dp<u64>(lv.toff),
                                  dp<T>(bl.tinv), bl.per_shift, sidx, lv.term_max, r, x, Bp, lim, s);
        });
    };
    if (lv.direct) {
        // a coarse level of terminal blocks only: its cycle from zero is their dense solve
        T *dst = out ? out : cur;
        P.run(K_OTHER, nb * ts, [&] { gpu::fill<T>(n * Bp, T(0), dst, s); });
        terminal(rhs, dst);
        return dst;
    }
    const bool rbgs = lv.rb;
    const u32 *perm0 = dp<u32>(lv.perm);
    if (rbgs && out && !perm0) cur = out; // the in-place smoother works in the output
    bool have_residual = false;
    if (rbgs) {
        // red-black Gauss-Seidel: colour 0 from x = 0, then colour 1, colour 0, ...; after the
        // last colour-1 half-sweep only colour-0 rows have a residual
        T *rb = rhs64 || perm0 ? dp<T>(bl.rhs) : nullptr;
        P.run(K_FIRST, nb * (rhs64 ? 8.0 + ts : ts) + nb * ts, [&] {
            if (rhs64) {
                bk::rb_first<T, double>(lv.A.view, dp<std::uint8_t>(lv.colour), sh, rhs64, scale, rb, cur,
                                        perm0, Bp, lim, s);
            } else {
                bk::rb_first<T, T>(lv.A.view, dp<std::uint8_t>(lv.colour), sh, rhs, nullptr, rb, cur,
                                   perm0, Bp, lim, s);
            }
        });
        if (rb) rhs = rb;
        auto half_bytes = [&](int c) {
            return op_bytes(lv.Ac[c], 0, 0, 0) + double(lv.Ac[c].n) * Bp * 3.0 * ts;
        };
        for (int half = 1; half < 2 * lv.pre; ++half) {
            const int c = half & 1;
            T *rz = half == 2 * lv.pre - 1 ? res : nullptr;
            P.run(K_SWEEP, half_bytes(c), [&] { bk::gs_half<T>(lv.Ac[c].view, sh, rhs, cur, rz, Bp, lim, s); });
        }
        P.run(K_RESIDUAL, half_bytes(0), [&] { bk::level_residual<T>(lv.Ac[0].view, sh, rhs, cur, res, Bp, lim, s); });
        have_residual = true;
    } else if (rhs64) {
        T *rb = dp<T>(bl.rhs);
        P.run(K_FIRST, nb * (8 + 3 * ts), [&] {
            bk::jacobi_first<T, double>(lv.A.view, sh, lv.pre > 0 ? omega : 0.0, rhs64, scale, rb,
                                        cur, Bp, lim, s);
        });
        rhs = rb;
    } else {
        P.run(K_FIRST, nb * 3 * ts, [&] {
            bk::jacobi_first<T, T>(lv.A.view, sh, lv.pre > 0 ? omega : 0.0, rhs, nullptr, nullptr,
                                   cur, Bp, lim, s);
        });
    }
    for (int k = 1; !rbgs && k < lv.pre; ++k) {
        P.run(K_SWEEP, sweep_bytes, [&] {
            bk::jacobi_sweep<T>(lv.A.view, sh, omega, rhs, cur, other, Bp, lim, s);
        });
        std::swap(cur, other);
    }
    if (!have_residual) {
        P.run(K_RESIDUAL, sweep_bytes, [&] {
            bk::level_residual<T>(lv.A.view, sh, rhs, cur, res, Bp, lim, s);
        });
    }
    if (lv.n_term > 0) terminal(res, cur);
    if (lv.n_next > 0) {
        auto &c = L[l + 1];
        auto &cb = BL[l + 1];
        const double cn = double(c.n) * double(Bp);
        const double rbytes = op_bytes(lv.R, 0, 0, 0) + double(lv.n_fine_linked) * Bp * ts + cn * ts;
        // red-black with post-smoothing: only colour-0 rows need the correction (the first
        // post-smoothing half-sweep rewrites colour-1 rows without reading them)
        const bool red_only = rbgs && lv.post > 0;
        const i64 np = red_only ? lv.Ac[0].n : n;
        const u32 *prows = red_only ? dp<u32>(lv.rowmap[0]) : nullptr;
        const double pbytes = double(np) * (4 + 2 * Bp * ts) + cn * ts;
        const int *clim = lims + static_cast<i64>(l + 1) * Bp;
        if (mg.opt.coarse_cycles <= 1 || static_cast<i64>(l) >= cycle_levels(mg.opt)) {
            P.run(K_RESTRICT, rbytes, [&] { bk::gather_sum<T>(lv.R.view, res, dp<T>(cb.rhs), Bp, clim, s); });
            T *cx = brun_level<T>(B, l + 1, dp<T>(cb.rhs), nullptr, nullptr, nullptr, lims);
            P.run(K_PROLONG, pbytes, [&] {
                bk::prolong_add<T>(np, dp<u32>(lv.coarse_of), dp<T>(lv.weight), cx, cur, Bp, lim, s, prows);
            });
        } else {
            P.run(K_RESTRICT, rbytes, [&] { bk::gather_sum<T>(lv.R.view, res, dp<T>(cb.crhs), Bp, clim, s); });
            P.run(K_OTHER, cn * ts, [&] { gpu::fill<T>(c.n * Bp, T(0), dp<T>(cb.cx), s); });
            const double cbytes = op_bytes(c.A, 0, 0, 0) + cn * 3 * ts;
            for (int cyc = 0; cyc < mg.opt.coarse_cycles; ++cyc) {
                P.run(K_RESIDUAL, cbytes, [&] {
                    bk::level_residual<T>(c.A.view, sh, dp<T>(cb.crhs), dp<T>(cb.cx), dp<T>(cb.rhs), Bp, clim, s);
                });
                T *ce = brun_level<T>(B, l + 1, dp<T>(cb.rhs), nullptr, nullptr, nullptr, lims);
                P.run(K_OTHER, cn * 3 * ts, [&] { gpu::axpy<T>(c.n * Bp, T(1), ce, dp<T>(cb.cx), s); });
            }
            P.run(K_PROLONG, pbytes, [&] {
                bk::prolong_add<T>(np, dp<u32>(lv.coarse_of), dp<T>(lv.weight), dp<T>(cb.cx), cur, Bp, lim, s, prows);
            });
        }
    }
    if (rbgs) {
        for (int half = 0; half < 2 * lv.post; ++half) {
            const int c = 1 - (half & 1); // reverse colour order
            // the last two half-sweeps also store their rows to the output in the caller's order
            T *nat = perm0 && out && half >= 2 * lv.post - 2 ? out : nullptr;
            P.run(K_SWEEP, op_bytes(lv.Ac[c], 0, 0, 0) + double(lv.Ac[c].n) * Bp * 3.0 * ts, [&] {
                bk::gs_half<T>(lv.Ac[c].view, sh, rhs, cur, nullptr, Bp, lim, s, nat, perm0);
            });
        }
        if (perm0 && out) return out;
    } else {
        for (int k = 0; k < lv.post; ++k) {
            T *dst = (k == lv.post - 1 && out) ? out : other;
            P.run(K_SWEEP, sweep_bytes, [&] {
                bk::jacobi_sweep<T>(lv.A.view, sh, omega, rhs, cur, dst, Bp, lim, s);
            });
            if (dst == other) {
                std::swap(cur, other);
            } else {
                cur = dst;
            }
        }
    }
    if (out && cur != out) {
        P.run(K_OTHER, nb * 2 * ts, [&] { gpu::copy16(n * Bp * sizeof(T) / 16, cur, out, s); });
        cur = out;
    }
    return cur;
}

void ensure_workspace(mk_batch_s &B, i64 m) {
    if (B.m >= m && B.V.p) return;
    const i64 n = B.n, Bp = B.Bp;
    const std::size_t vec = static_cast<std::size_t>(std::max<i64>(n, 1)) * Bp;
    const std::size_t zs = B.mg->precision == MK_FP32 ? sizeof(float) : sizeof(double);
    B.graphs.clear();
    B.V = dbuf();
    B.Z = dbuf();
    B.V = dbuf(vec * (m + 1) * sizeof(double));
    B.Z = dbuf(vec * m * zs);
    B.w = dbuf(vec * sizeof(double));
    B.np = bk::parts_wide(n);
    B.partial = dbuf(static_cast<std::size_t>(B.np) * (m + 2) * Bp * sizeof(double));
    B.ptmp = dbuf(std::max<i64>(1, bk::reduce_tmp_size(B.np, (m + 2) * Bp)) * sizeof(double));
    B.rpart = dbuf(static_cast<std::size_t>(bk::residual_parts(n)) * Bp * sizeof(double));
    B.rtmp = dbuf(std::max<i64>(1, bk::reduce_tmp_size(static_cast<int>(bk::residual_parts(n)), Bp)) *
                  sizeof(double));
    const slots sl{(m + 2) * Bp};
    B.hd = dbuf(sl.total(Bp) * sizeof(double));
    B.d_scale = dbuf(Bp * sizeof(double));
    B.d_beta = dbuf(Bp * sizeof(double));
    B.d_cnt = dbuf(Bp * sizeof(int));
    B.d_y = dbuf(m * Bp * sizeof(double));
    if (B.host) (void)hipHostFree(B.host);
    B.host = nullptr;
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B.host),
                         (sl.total(Bp) + m * Bp + 2 * Bp) * sizeof(double)));
    B.m = m;
}

// the limits of the current step: lane j's rows on every level where mask[j], else 0
void push_limits(mk_batch_s &B, const char *mask) {
    for (int l = 0; l < B.levels; ++l)
        for (i64 j = 0; j < B.Bp; ++j) B.h_lim[l * B.Bp + j] = mask[j] ? B.lim[l][j] : 0;
    HIPCHK(hipMemcpyAsync(B.d_lim.p, B.h_lim, static_cast<std::size_t>(B.levels) * B.Bp * sizeof(int),
                          hipMemcpyHostToDevice, stream()));
}

// V_0 = b - A x and beta_j = |r_j| for the lanes in mask
void true_residual(mk_batch_s &B, const char *mask, double *dst = nullptr) {
    auto &P = prof();
    const hipStream_t s = stream();
    const i64 n = B.n, Bp = B.Bp;
    const slots sl{(B.m + 2) * Bp};
    push_limits(B, mask);
    double *rp = dp<double>(B.rpart), *hd = dp<double>(B.hd);
    const int nrp = static_cast<int>(bk::residual_parts(n));
    const double nb = double(n) * double(Bp);
    P.run(K_SPMV, op_bytes(B.gen->A, 0, 0, 0) + nb * 24.0, [&] {
        bk::shifted_apply<double>(B.gen->A.view, dp<double>(B.d_shift), dp<double>(B.d_x), dp<double>(B.d_b),
                                  dst ? dst : dp<double>(B.V), rp, Bp, dp<int>(B.d_lim), s);
    });
    P.run(K_DOT, double(nrp) * Bp * 8.0,
          [&] { bk::reduce_partials(rp, nrp, Bp, hd + sl.rn(Bp), s, dp<double>(B.rtmp)); });
    HIPCHK(hipMemcpyAsync(B.host + sl.rn(Bp), hd + sl.rn(Bp), Bp * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    for (i64 j = 0; j < Bp; ++j)
        if (mask[j]) B.slot[j].beta = std::sqrt(B.host[sl.rn(Bp) + j]);
}

// after a residual: converged, non-finite or out of iterations ends the slot's solve
void settle(slot_state &S) {
    S.res.residual = S.beta;
    if (!std::isfinite(S.beta)) {
        S.active = false;
    } else if (!(S.beta > S.tol)) {
        S.res.converged = 1;
        S.active = false;
    } else if (S.res.iterations >= S.max_iterations) {
        S.active = false;
    } else {
        S.active = true;
    }
    S.done = !S.active;
}

// the initial residuals of the slots admitted since the last step
void take_initial_residuals(mk_batch_s &B) {
    std::vector<char> mask(B.Bp, 0);
    bool any_fresh = false;
    for (i64 j = 0; j < B.Bp; ++j) {
        mask[j] = B.slot[j].fresh ? 1 : 0;
        any_fresh = any_fresh || mask[j];
    }
    if (!any_fresh) return;
    true_residual(B, mask.data());
    for (i64 j = 0; j < B.Bp; ++j) {
        auto &S = B.slot[j];
        if (!S.fresh) continue;
        S.fresh = false;
        S.res = mk_solver_result{0, 0.0, 0};
        settle(S);
    }
}

// One restart cycle of FGMRES(m) with delayed CGS2 for the active slots (the per-system
// algorithm of fgmres in api.hip), then x_j += Z y_j and the true residual. Returns whether any
// slot is still active.
template <typename ZT>
bool step(mk_batch_s &B) {
    const i64 n = B.n, Bp = B.Bp, m = B.m;
    auto &P = prof();
    const hipStream_t s = stream();
    auto &A = *B.gen;
    const int np = B.np;
    const i64 vec = n * Bp;
    double *V = dp<double>(B.V), *w = dp<double>(B.w), *partial = dp<double>(B.partial);
    double *hd = dp<double>(B.hd);
    ZT *Z = dp<ZT>(B.Z);
    const double *sh = dp<double>(B.d_shift);
    const slots sl{(m + 2) * Bp};
    double *host = B.host;
    double *yh = host + sl.total(Bp), *scale_h = yh + m * Bp, *beta_h = scale_h + Bp;
    int *cnt_h = B.ihost;
    const double nb = double(vec);
    std::vector<char> mask(Bp, 0);
    bool any = false;
    for (i64 j = 0; j < Bp; ++j) {
        mask[j] = B.slot[j].active ? 1 : 0;
        any = any || mask[j];
    }
    if (!any) return false;
    std::vector<char> iterated(mask);
    std::vector<double> col(m + 1);
    for (i64 j = 0; j < Bp; ++j) {
        auto &S = B.slot[j];
        S.cols = 0;
        if (!mask[j]) continue;
        B.ls[j].start(S.beta);
        S.bj = S.beta;
    }
    auto any_mask = [&] {
        for (i64 j = 0; j < Bp; ++j)
            if (mask[j]) return true;
        return false;
    };
    int which = 0; // s buffer the next update kernel reads
    const bool graphs = B.capture && batch_graph_switch().load() != 0;
    for (i64 c = 0; c < m && any_mask(); ++c) {
        const int cnt = static_cast<int>(c);
        for (i64 j = 0; j < Bp; ++j) {
            scale_h[j] = mask[j] ? 1.0 / B.slot[j].bj : 0.0;
            beta_h[j] = mask[j] ? B.slot[j].bj : 1.0;
        }
        HIPCHK(hipMemcpyAsync(B.d_scale.p, scale_h, Bp * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B.d_beta.p, beta_h, Bp * sizeof(double), hipMemcpyHostToDevice, s));
        push_limits(B, mask.data());
        const int *lims = dp<int>(B.d_lim);
        double *sd = hd + sl.s(which), *sd_next = hd + sl.s(1 - which);
        // z = M (u / beta), w = A z, projections, delayed CGS2 update: one captured graph per step
        const graph_key key{B.uid, static_cast<u64>(sizeof(ZT)), c, which, n, m, 0};
        B.graphs.run(key, s, [&] {
            ZT *Zc = Z + c * vec;
            brun_level<ZT>(B, 0, nullptr, V + c * vec, dp<double>(B.d_scale), Zc, lims);
            P.run(K_SPMV, op_bytes(A.A, 0, 0, 0) + nb * (sizeof(ZT) + 8.0), [&] {
                bk::shifted_apply<ZT>(A.A.view, sh, Zc, nullptr, w, nullptr, Bp, lims, s);
            });
            P.run(K_DOT, double(cnt + 2) * nb * 8.0, [&] {
                bk::multi_dot(V, cnt + 1, w, n, Bp, partial, np, lims, s);
                bk::reduce_partials(partial, np, (cnt + 1) * Bp, hd + sl.t(), s, dp<double>(B.ptmp));
            });
            P.run(K_UPDATE, double(cnt + 4) * nb * 8.0, [&] {
                bk::dcgs_update(V, cnt, sd, hd + sl.t(), dp<double>(B.d_beta), w, n, Bp, partial, np,
                                hd + sl.eta(), lims, s);
                bk::reduce_partials(partial, np, (cnt + 2) * Bp, sd_next, s, dp<double>(B.ptmp));
            });
        }, n, graphs);
        HIPCHK(hipMemcpyAsync(host, hd, sl.total(Bp) * sizeof(double), hipMemcpyDeviceToHost, s));
        drain();
        const double *t = host + sl.t(), *eta = host + sl.eta();
        double *sn = host + sl.s(1 - which);
        bool any_flag = false;
        std::vector<char> flag(Bp, 0);
        for (i64 j = 0; j < Bp; ++j) {
            if (!mask[j]) continue;
            double ss = 0.0;
            for (i64 k = 0; k <= c; ++k) ss += sn[k * Bp + j] * sn[k * Bp + j];
            if (!(ss <= 0.25 * sn[(c + 1) * Bp + j])) {
                flag[j] = 1;
                any_flag = true;
            }
        }
        std::vector<double> first;
        if (any_flag) {
            // explicit second projection of the flagged systems' candidates; the other systems
            // keep their coefficients
            first.assign(sn, sn + (c + 2) * Bp);
            push_limits(B, flag.data());
            double *u = V + (c + 1) * vec;
            P.run(K_UPDATE, double(cnt + 3) * nb * 8.0, [&] {
                bk::basis_update(u, V, cnt + 1, sd_next, n, Bp, lims, s);
            });
            P.run(K_DOT, double(cnt + 3) * nb * 8.0, [&] {
                bk::multi_dot(V, cnt + 2, u, n, Bp, partial, np, lims, s);
                bk::reduce_partials(partial, np, (cnt + 2) * Bp, sd_next, s, dp<double>(B.ptmp));
            });
            HIPCHK(hipMemcpyAsync(host, hd, sl.total(Bp) * sizeof(double), hipMemcpyDeviceToHost, s));
            drain();
            for (i64 j = 0; j < Bp; ++j)
                if (!flag[j])
                    for (i64 k = 0; k <= c + 1; ++k) sn[k * Bp + j] = first[k * Bp + j];
            HIPCHK(hipMemcpyAsync(sd_next, sn, (c + 2) * Bp * sizeof(double), hipMemcpyHostToDevice, s));
        }
        for (i64 j = 0; j < Bp; ++j) {
            if (!mask[j]) continue;
            auto &S = B.slot[j];
            ++S.res.iterations;
            for (i64 k = 0; k < c; ++k) col[k] = t[k * Bp + j] + sn[k * Bp + j];
            col[c] = eta[j] + sn[c * Bp + j];
            double ss = 0.0;
            for (i64 k = 0; k <= c; ++k) ss += sn[k * Bp + j] * sn[k * Bp + j];
            if (flag[j])
                for (i64 k = 0; k <= c; ++k) col[k] += first[k * Bp + j];
            const double hn = std::sqrt(std::max(sn[(c + 1) * Bp + j] - ss, 0.0));
            S.bj = hn;
            S.cols = c + 1;
            const double est = B.ls[j].add_column(c, col.data(), hn);
            if (!(est > S.tol) || !(hn > 0.0) || !std::isfinite(hn) || S.res.iterations >= S.max_iterations)
                mask[j] = 0;
        }
        which = 1 - which;
    }
    // x_j += Z y_j over each system's columns
    std::fill(yh, yh + m * Bp, 0.0);
    for (i64 j = 0; j < Bp; ++j) {
        cnt_h[j] = 0;
        auto &S = B.slot[j];
        if (!iterated[j] || S.cols == 0) continue;
        B.ls[j].solve(S.cols);
        for (i64 k = 0; k < S.cols; ++k) yh[k * Bp + j] = B.ls[j].y[k];
        cnt_h[j] = static_cast<int>(S.cols);
    }
    HIPCHK(hipMemcpyAsync(B.d_y.p, yh, m * Bp * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.d_cnt.p, cnt_h, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    P.run(K_OTHER, nb * (m * sizeof(ZT) + 16.0), [&] {
        bk::combine<ZT>(dp<double>(B.d_x), Z, dp<int>(B.d_cnt), dp<double>(B.d_y), n, Bp, s);
    });
    true_residual(B, iterated.data());
    bool still = false;
    for (i64 j = 0; j < Bp; ++j) {
        if (!iterated[j]) continue;
        settle(B.slot[j]);
        still = still || B.slot[j].active;
    }
    return still;
}

// ---- block FGMRES cycle of the pending groups ----------------------------------------------
constexpr int kP = bk::kBlock;

// G = R^T R with R upper triangular (kP x kP, row-major); false when G is not numerically
// positive definite (the block is rank deficient)
bool block_cholesky(const double *G, double *R) {
    for (int i = 0; i < kP * kP; ++i) R[i] = 0.0;
    for (int j = 0; j < kP; ++j) {
        double d = G[j * kP + j];
        for (int k = 0; k < j; ++k) d -= R[k * kP + j] * R[k * kP + j];
        if (!(d > 1e-24 * std::max(G[j * kP + j], 1e-300)) || !std::isfinite(d)) return false;
        R[j * kP + j] = std::sqrt(d);
        for (int i = j + 1; i < kP; ++i) {
            double v = G[j * kP + i];
            for (int k = 0; k < j; ++k) v -= R[k * kP + j] * R[k * kP + i];
            R[j * kP + i] = v / R[j * kP + j];
        }
    }
    return true;
}

// the inverse of an upper triangular kP x kP matrix
void upper_inverse(const double *R, double *Ri) {
    for (int i = 0; i < kP * kP; ++i) Ri[i] = 0.0;
    for (int c = 0; c < kP; ++c) {
        for (int r = c; r >= 0; --r) {
            double v = r == c ? 1.0 : 0.0;
            for (int k = r + 1; k <= c; ++k) v -= R[r * kP + k] * Ri[k * kP + c];
            Ri[r * kP + c] = v / R[r * kP + r];
        }
    }
}

// max |R_ii| / min |R_ii| of an upper triangular kP x kP matrix: a lower bound of its condition
// number, used to end a block cycle before the block's columns become dependent
double diag_ratio(const double *R) {
    double lo = std::abs(R[0]), hi = lo;
    for (int i = 1; i < kP; ++i) {
        lo = std::min(lo, std::abs(R[i * kP + i]));
        hi = std::max(hi, std::abs(R[i * kP + i]));
    }
    return lo > 0.0 ? hi / lo : std::numeric_limits<double>::infinity();
}

// R = A B for upper triangular kP x kP matrices
void upper_product(const double *A, const double *B, double *R) {
    for (int i = 0; i < kP; ++i)
        for (int j = 0; j < kP; ++j) {
            double v = 0.0;
            for (int k = i; k <= j; ++k) v += A[i * kP + k] * B[k * kP + j];
            R[i * kP + j] = i <= j ? v : 0.0;
        }
}

// the block least-squares problem of one group: the block Hessenberg matrix ((m + 1) kP x m kP,
// row-major), its right-hand side (m + 1) kP x kP, and the Givens rotations applied so far
struct block_ls {
    i64 m = 0, cols = 0;
    std::vector<double> H, g;
    struct rot {
        i64 i, j;
        double c, s;
    };
    std::vector<rot> rots;
    void start(i64 m_, const double *S0) {
        m = m_;
        cols = 0;
        H.assign((m + 1) * kP * m * kP, 0.0);
        g.assign((m + 1) * kP * kP, 0.0);
        for (int a = 0; a < kP; ++a)
            for (int b = 0; b < kP; ++b) g[a * kP + b] = S0[a * kP + b];
        rots.clear();
    }
    double &h(i64 r, i64 c) { return H[r * m * kP + c]; }
    // block column c: projections Hc[(i kP + a) kP + b] for i <= c and the new block's R; returns
    // the residual norm estimate of each right-hand side (column of g)
    void add(i64 c, const double *Hc, const double *R, double *est) {
        const i64 c0 = c * kP, nc = (c + 1) * kP;
        for (i64 i = 0; i <= c; ++i)
            for (int a = 0; a < kP; ++a)
                for (int b = 0; b < kP; ++b) h(i * kP + a, c0 + b) = Hc[(i * kP + a) * kP + b];
        for (int a = 0; a < kP; ++a)
            for (int b = 0; b < kP; ++b) h(nc + a, c0 + b) = R[a * kP + b];
        for (const auto &q : rots)
            for (int b = 0; b < kP; ++b) {
                const double x = h(q.i, c0 + b), y = h(q.j, c0 + b);
                h(q.i, c0 + b) = q.c * x + q.s * y;
                h(q.j, c0 + b) = -q.s * x + q.c * y;
            }
        for (int b = 0; b < kP; ++b) {
            const i64 col = c0 + b;
            for (i64 r = col + 1; r <= col + kP; ++r) {
                const double x = h(col, col), y = h(r, col);
                if (y == 0.0) continue;
                const double d = std::hypot(x, y);
                const rot q{col, r, x / d, y / d};
                for (i64 cc = col; cc < nc; ++cc) {
                    const double u = h(col, cc), v = h(r, cc);
                    h(col, cc) = q.c * u + q.s * v;
                    h(r, cc) = -q.s * u + q.c * v;
                }
                for (int k = 0; k < kP; ++k) {
                    const double u = g[col * kP + k], v = g[r * kP + k];
                    g[col * kP + k] = q.c * u + q.s * v;
                    g[r * kP + k] = -q.s * u + q.c * v;
                }
                rots.push_back(q);
            }
        }
        cols = c + 1;
        for (int k = 0; k < kP; ++k) {
            double ss = 0.0;
            for (int a = 0; a < kP; ++a) ss += g[(nc + a) * kP + k] * g[(nc + a) * kP + k];
            est[k] = std::sqrt(ss);
        }
    }
    // Y = R^-1 g over the first cols block columns (row-major, cols kP x kP)
    void solve(std::vector<double> &Y) {
        const i64 N = cols * kP;
        Y.assign(N * kP, 0.0);
        for (int k = 0; k < kP; ++k)
            for (i64 i = N; i-- > 0;) {
                double v = g[i * kP + k];
                for (i64 j = i + 1; j < N; ++j) v -= h(i, j) * Y[j * kP + k];
                Y[i * kP + k] = v / h(i, i);
            }
    }
};

// One block FGMRES(m) cycle for every pending group with an active slot: the residual block is
// orthonormalised (Cholesky QR twice), and each step applies the preconditioner to the kP basis
// vectors (one batched cycle), projects the operator's image on the block basis twice (block CGS2,
// the projections on matrix cores), orthonormalises the remainder by Cholesky QR twice and extends
// the block Hessenberg least-squares problem. A group whose new block is rank deficient or
// ill-conditioned (diagonal ratio of R above 1e7) ends its cycle there. Then x += Z Y and the true
// residual per slot; the slots left unconverged continue with per-slot restarts.
template <typename ZT>
void block_cycle(mk_batch_s &B) {
    const i64 n = B.n, Bp = B.Bp, m = B.m, G = Bp / kP;
    auto &P = prof();
    const hipStream_t s = stream();
    auto &A = *B.gen;
    const i64 vec = n * Bp;
    const double nb = double(vec);
    double *V = dp<double>(B.V), *w = dp<double>(B.w), *part = dp<double>(B.bpartial);
    ZT *Z = dp<ZT>(B.Z);
    const double *sh = dp<double>(B.d_shift);
    const i64 hsz = (m + 1) * Bp * kP;
    double *H1 = dp<double>(B.d_bh), *H2 = H1 + hsz, *Gd = H2 + hsz, *Cd = Gd + Bp * kP;
    const int *lims = dp<int>(B.d_lim);
    std::vector<char> run(G, 0), mask(Bp, 0);
    bool any = false;
    for (i64 g = 0; g < G; ++g) {
        if (!B.block_pending[g]) continue;
        B.block_pending[g] = 0;
        bool on = false;
        for (int b = 0; b < kP; ++b) on = on || B.slot[g * kP + b].active;
        if (!on) continue;
        run[g] = 1;
        any = true;
        for (int b = 0; b < kP; ++b) mask[g * kP + b] = B.slot[g * kP + b].rows > 0 ? 1 : 0;
    }
    if (!any) return;
    // the projections on `count` block vectors of v, reduced into out
    auto project = [&](const double *v, int count, const double *wv, double *out) {
        P.run(K_DOT, double(count + 1) * nb * 8.0, [&] {
            bk::block_dot(v, count, wv, n, Bp, part, B.npb, lims, s);
            bk::reduce_partials(part, B.npb, static_cast<i64>(count) * Bp * kP, out, s);
        });
    };
    auto fetch = [&](const double *src, i64 count, double *dst) {
        HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, s));
    };
    std::vector<double> C(Bp * kP, 0.0);
    auto put_coefficients = [&](i64 count) {
        HIPCHK(hipMemcpy(Cd, C.data(), count * sizeof(double), hipMemcpyHostToDevice));
    };
    // Cholesky QR twice of the block src (src = dst R with dst orthonormal, R upper triangular),
    // for the groups in `on`; dst != src; R per group into Rg; groups whose block is rank deficient
    // or too ill-conditioned come back false in ok
    std::vector<double> Rg(G * kP * kP, 0.0);
    std::vector<char> ok(G, 0);
    double Gm[kP * kP], R1[kP * kP], R2[kP * kP], Ri[kP * kP];
    auto gram_of = [&](i64 g, const double *h) {
        for (int a = 0; a < kP; ++a)
            for (int b = 0; b < kP; ++b) Gm[a * kP + b] = h[(g * kP + b) * kP + a];
    };
    auto scale_by = [&](const std::vector<char> &on, double *dst, const double *src,
                        const std::vector<double> &Rs) {
        std::fill(C.begin(), C.end(), 0.0);
        std::vector<char> lanes(Bp, 0);
        for (i64 g = 0; g < G; ++g) {
            if (!on[g]) continue;
            upper_inverse(&Rs[g * kP * kP], Ri);
            for (int a = 0; a < kP; ++a)
                for (int b = 0; b < kP; ++b) C[(g * kP + b) * kP + a] = Ri[a * kP + b];
            for (int b = 0; b < kP; ++b) lanes[g * kP + b] = B.slot[g * kP + b].rows > 0 ? 1 : 0;
        }
        push_limits(B, lanes.data());
        put_coefficients(Bp * kP);
        P.run(K_OTHER, nb * 40.0, [&] { bk::block_apply<double>(dst, src, 1, Cd, n, Bp, 1.0, true, lims, s); });
    };
    auto cholqr2 = [&](const std::vector<char> &on, const double *src, double *dst, double *tmp) {
        std::vector<char> lanes(Bp, 0);
        for (i64 g = 0; g < G; ++g)
            if (on[g])
                for (int b = 0; b < kP; ++b) lanes[g * kP + b] = B.slot[g * kP + b].rows > 0 ? 1 : 0;
        push_limits(B, lanes.data());
        project(src, 1, src, Gd);
        fetch(Gd, Bp * kP, B.bh + 2 * hsz);
        drain();
        std::vector<char> live(G, 0);
        std::vector<double> Rfirst(G * kP * kP, 0.0);
        for (i64 g = 0; g < G; ++g) {
            ok[g] = 0;
            std::fill(&Rg[g * kP * kP], &Rg[(g + 1) * kP * kP], 0.0);
            if (!on[g]) continue;
            gram_of(g, B.bh + 2 * hsz);
            if (!block_cholesky(Gm, R1)) continue;
            // an ill-conditioned block keeps its R for the least-squares problem, but the group
            // ends its cycle here
            std::copy(R1, R1 + kP * kP, &Rg[g * kP * kP]);
            if (diag_ratio(R1) > 1e7) continue;
            std::copy(R1, R1 + kP * kP, &Rfirst[g * kP * kP]);
            live[g] = 1;
        }
        scale_by(live, tmp, src, Rfirst);
        push_limits(B, lanes.data());
        project(tmp, 1, tmp, Gd);
        fetch(Gd, Bp * kP, B.bh + 2 * hsz);
        drain();
        std::vector<double> Rsecond(G * kP * kP, 0.0);
        for (i64 g = 0; g < G; ++g) {
            if (!live[g]) continue;
            gram_of(g, B.bh + 2 * hsz);
            if (!block_cholesky(Gm, R2)) {
                live[g] = 0;
                continue;
            }
            std::copy(R2, R2 + kP * kP, &Rsecond[g * kP * kP]);
            upper_product(R2, &Rfirst[g * kP * kP], &Rg[g * kP * kP]);
            ok[g] = 1;
        }
        scale_by(live, dst, tmp, Rsecond);
    };
    // the residual block V_0 = R S0^-1 (through w and Z's last column as scratch)
    double *scratch = reinterpret_cast<double *>(B.bscratch.p);
    cholqr2(run, V, w, scratch);
    std::vector<block_ls> ls(G);
    for (i64 g = 0; g < G; ++g) {
        if (!run[g]) continue;
        if (!ok[g]) {
            run[g] = 0; // a rank-deficient residual block: per-slot restarts take over
            for (int b = 0; b < kP; ++b) mask[g * kP + b] = 0;
            continue;
        }
        ls[g].start(m, &Rg[g * kP * kP]);
    }
    std::fill(C.begin(), C.end(), 0.0);
    for (i64 j = 0; j < Bp; ++j) C[j * kP + (j % kP)] = 1.0; // identity: V_0 = w
    push_limits(B, mask.data());
    put_coefficients(Bp * kP);
    P.run(K_OTHER, nb * 40.0, [&] { bk::block_apply<double>(V, w, 1, Cd, n, Bp, 1.0, true, lims, s); });
    std::vector<char> iterated(mask);
    std::vector<double> Hc((m + 1) * kP * kP);
    double est[kP];
    auto any_run = [&] {
        for (i64 g = 0; g < G; ++g)
            if (run[g]) return true;
        return false;
    };
    for (i64 c = 0; c < m && any_run(); ++c) {
        const int cnt = static_cast<int>(c + 1);
        push_limits(B, mask.data());
        ZT *Zc = Z + c * vec;
        brun_level<ZT>(B, 0, nullptr, V + c * vec, nullptr, Zc, lims);
        P.run(K_SPMV, op_bytes(A.A, 0, 0, 0) + nb * (sizeof(ZT) + 8.0), [&] {
            bk::shifted_apply<ZT>(A.A.view, sh, Zc, nullptr, w, nullptr, Bp, lims, s);
        });
        // block CGS2: two projections of w on the basis blocks
        project(V, cnt, w, H1);
        P.run(K_UPDATE, double(cnt + 2) * nb * 8.0, [&] {
            bk::block_apply<double>(w, V, cnt, H1, n, Bp, -1.0, false, lims, s);
        });
        project(V, cnt, w, H2);
        P.run(K_UPDATE, double(cnt + 2) * nb * 8.0, [&] {
            bk::block_apply<double>(w, V, cnt, H2, n, Bp, -1.0, false, lims, s);
        });
        fetch(H1, cnt * Bp * kP, B.bh);
        fetch(H2, cnt * Bp * kP, B.bh + hsz);
        drain();
        // the next block: V_{c+1} R = w by Cholesky QR twice (the last step's block is not kept)
        const bool next = c + 1 < m;
        double *vnext = next ? V + (c + 1) * vec : scratch;
        cholqr2(run, w, vnext, scratch == vnext ? dp<double>(B.bscratch2) : scratch);
        for (i64 g = 0; g < G; ++g) {
            if (!run[g]) continue;
            for (i64 i = 0; i < cnt; ++i)
                for (int a = 0; a < kP; ++a)
                    for (int b = 0; b < kP; ++b) {
                        const i64 o = (i * Bp + g * kP + b) * kP + a;
                        Hc[(i * kP + a) * kP + b] = B.bh[o] + B.bh[hsz + o];
                    }
            ls[g].add(c, Hc.data(), &Rg[g * kP * kP], est);
            bool done = true, out_of_iterations = false;
            for (int b = 0; b < kP; ++b) {
                auto &S = B.slot[g * kP + b];
                if (S.rows == 0) continue;
                ++S.res.iterations;
                done = done && !(est[b] > S.tol);
                out_of_iterations = out_of_iterations || S.res.iterations >= S.max_iterations;
            }
            if (done || !ok[g] || out_of_iterations) {
                run[g] = 0;
                for (int b = 0; b < kP; ++b) mask[g * kP + b] = 0;
            }
        }
    }
    // x += Z Y per group: coefficients (i Bp + j) kP + a = Y[i kP + a][j mod kP]
    std::vector<double> Yc(m * Bp * kP, 0.0), Y;
    std::vector<int> counts(Bp, 0);
    i64 most = 0;
    for (i64 g = 0; g < G; ++g) {
        if (ls[g].cols == 0) continue;
        ls[g].solve(Y);
        for (i64 i = 0; i < ls[g].cols; ++i)
            for (int a = 0; a < kP; ++a)
                for (int b = 0; b < kP; ++b) Yc[(i * Bp + g * kP + b) * kP + a] = Y[(i * kP + a) * kP + b];
        for (int b = 0; b < kP; ++b) counts[g * kP + b] = static_cast<int>(ls[g].cols);
        most = std::max(most, ls[g].cols);
    }
    std::vector<char> updated(Bp, 0);
    for (i64 j = 0; j < Bp; ++j) updated[j] = counts[j] > 0 && B.slot[j].rows > 0 ? 1 : 0;
    HIPCHK(hipMemcpy(Cd, Yc.data(), m * Bp * kP * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B.d_bcnt.p, counts.data(), Bp * sizeof(int), hipMemcpyHostToDevice));
    push_limits(B, updated.data());
    P.run(K_OTHER, nb * (most * sizeof(ZT) * kP + 16.0), [&] {
        bk::block_apply<ZT>(dp<double>(B.d_x), Z, static_cast<int>(most), Cd, n, Bp, 1.0, false, lims, s,
                            dp<int>(B.d_bcnt));
    });
    true_residual(B, iterated.data());
    for (i64 j = 0; j < Bp; ++j)
        if (iterated[j] && B.slot[j].active) settle(B.slot[j]);
}

bool engine_step(mk_batch_s &B) {
    take_initial_residuals(B);
    // groups still owed their block cycle run it before any per-slot restart
    bool pending = false;
    for (char p : B.block_pending) pending = pending || p;
    if (pending) {
        if (B.mg->precision == MK_FP32) {
            block_cycle<float>(B);
        } else {
            block_cycle<double>(B);
        }
        for (const auto &S : B.slot)
            if (S.active) return true;
        return false;
    }
    if (B.mg->precision == MK_FP32) return step<float>(B);
    return step<double>(B);
}

// the block workspace for restart m
void ensure_block(mk_batch_s &B) {
    const i64 m = B.m, Bp = B.Bp;
    const std::size_t coeff = static_cast<std::size_t>((2 * (m + 1) + 1 + m) * Bp * kP);
    if (B.d_bh.bytes >= coeff * sizeof(double)) return;
    B.npb = std::min(bk::parts(B.n), 128);
    B.d_bh = dbuf(coeff * sizeof(double));
    B.bpartial = dbuf(static_cast<std::size_t>(B.npb) * (m + 1) * Bp * kP * sizeof(double));
    B.d_bcnt = dbuf(Bp * sizeof(int));
    B.bscratch = dbuf(static_cast<std::size_t>(std::max<i64>(B.n, 1)) * Bp * sizeof(double));
    B.bscratch2 = dbuf(static_cast<std::size_t>(std::max<i64>(B.n, 1)) * Bp * sizeof(double));
    if (B.bh) (void)hipHostFree(B.bh);
    B.bh = nullptr;
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B.bh), (2 * (m + 1) + 1) * Bp * kP * sizeof(double)));
}

// zero the tails (rows past each lane's rows) of the solution, the right-hand side and every
// preconditioned vector on the lanes with want[j] whose system is shorter than the box, so the
// operator's gathers across a lane's last sector read zeros (the kernels never write there)
void clear_lane_tails(mk_batch_s &B, const std::vector<char> &want) {
    const hipStream_t s = stream();
    const i64 n = B.n, Bp = B.Bp;
    int *rows_h = B.ihost + Bp, *sel_h = B.ihost + 2 * Bp;
    bool any = false;
    for (i64 j = 0; j < Bp; ++j) {
        rows_h[j] = static_cast<int>(B.slot[j].rows);
        sel_h[j] = want[j] && B.slot[j].rows > 0 && B.slot[j].rows < n ? 1 : 0;
        any = any || sel_h[j];
    }
    if (!any) return;
    HIPCHK(hipMemcpyAsync(B.d_rows.p, rows_h, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.d_sel.p, sel_h, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    const int *rows = dp<int>(B.d_rows), *sel = dp<int>(B.d_sel);
    bk::clear_tails<double>(dp<double>(B.d_x), 1, n * Bp, n, Bp, rows, sel, s);
    bk::clear_tails<double>(dp<double>(B.d_b), 1, n * Bp, n, Bp, rows, sel, s);
    if (B.mg->precision == MK_FP32) {
        bk::clear_tails<float>(dp<float>(B.Z), static_cast<int>(B.m), n * Bp, n, Bp, rows, sel, s);
    } else {
        bk::clear_tails<double>(dp<double>(B.Z), static_cast<int>(B.m), n * Bp, n, Bp, rows, sel, s);
    }
    check_launch("batch: lane tails");
    drain();
}

mk_batch_s *create_engine(mk_generator_s *gen, mk_sector_mg_s *mg, i64 slots) {
    require(gen && mg && slots > 0, "batch: bad arguments");
    require(mg->rows == gen->n, "batch: the preconditioner does not match the operator");
    require(mg->opt.smoother == MK_JACOBI || mg->opt.smoother == MK_RED_BLACK,
            "batch: the Jacobi and red-black smoothers are batched");
    const bool fp32 = mg->precision == MK_FP32;
    const auto &L0 = fp32 ? mg->L32[0].rb : mg->L64[0].rb;
    if (L0 && (mg->opt.pre_sweeps < 1 || mg->opt.post_sweeps < 1))
        throw std::invalid_argument("batch: red-black batches need pre- and post-smoothing sweeps");
    auto B = std::make_unique<mk_batch_s>();
    B->gen = gen;
    B->mg = mg;
    B->n = gen->n;
    B->Bp = (slots + bk::kLanes - 1) / bk::kLanes * bk::kLanes;
    B->levels = static_cast<int>(mg->max_levels);
    B->slot.assign(B->Bp, slot_state{});
    B->ls.assign(B->Bp, hessenberg_ls(1));
    B->lim.assign(B->levels, std::vector<int>(B->Bp, 0));
    i64 end = 0;
    for (i64 k = 0; k < static_cast<i64>(mg->sector_rows[0].size()); ++k) {
        end += mg->sector_rows[0][k];
        B->sector_end.push_back(end);
    }
    require(end == B->n, "batch: the hierarchy does not cover the box");
    B->d_shift = upload(std::vector<double>(B->Bp, 1.0));
    std::vector<int> sidx(B->Bp);
    for (i64 j = 0; j < B->Bp; ++j) sidx[j] = static_cast<int>(j);
    B->d_sidx = upload(sidx);
    B->d_lim = dbuf(static_cast<std::size_t>(std::max(B->levels, 1)) * B->Bp * sizeof(int));
    B->d_rows = dbuf(B->Bp * sizeof(int));
    B->d_sel = dbuf(B->Bp * sizeof(int));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B->h_lim),
                         static_cast<std::size_t>(std::max(B->levels, 1)) * B->Bp * sizeof(int)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B->ihost), 3 * B->Bp * sizeof(int)));
    const std::size_t vb = static_cast<std::size_t>(std::max<i64>(B->n, 1)) * B->Bp * sizeof(double);
    B->d_b = dbuf(vb);
    B->d_x = dbuf(vb);
    if (fp32) {
        build_blevels<float>(*B);
    } else {
        build_blevels<double>(*B);
    }
    return B.release();
}

// A batch of full-size systems at the hierarchy's own shift whose hierarchy sits whole in the LDS
// tail: every system is solved by its own workgroup of one launch, each exactly as a single solve
// of it runs (the one-workgroup FGMRES). Returns false where that does not apply.
template <typename ZT>
bool one_workgroup_batch(mk_batch_s &B, const double *d_B, double *d_X, const double *tolerances,
                         const mk_gmres_options &o, i64 m, mk_solver_result *results) {
    const mk_sector_mg_s &mg = *B.mg;
    const i64 n = B.n;
    if (tiny_gmres_switch().load() == 0 || tail_switch().load() == 0 || tail_lds_switch().load() == 0 ||
        B.block > 1 || mg.tail_start != 0 || mg.tail_lds_bytes == 0 || n > gpu::kTinyRows ||
        m > gpu::kTinyBasis || gpu::gmres_tiny_lds(mg.tail_lds_bytes) == 0 || B.nsys < 1)
        return false;
    for (i64 j = 0; j < B.nsys; ++j)
        if (B.slot[j].rows != n || B.slot[j].shift != mg.shift) return false;
    const i64 ld = (n + 63) / 64 * 64, ns = B.nsys;
    const std::size_t vb = static_cast<std::size_t>(ns) * (m + 1) * ld * sizeof(double);
    const std::size_t zb = static_cast<std::size_t>(ns) * m * ld * sizeof(ZT);
    if (B.tV.bytes < vb) B.tV = dbuf(vb);
    if (B.tZ.bytes < zb) B.tZ = dbuf(zb);
    if (B.tres.bytes < ns * 3 * sizeof(double)) B.tres = dbuf(ns * 3 * sizeof(double));
    if (B.ttol.bytes < ns * sizeof(double)) B.ttol = dbuf(ns * sizeof(double));
    std::vector<double> tol(ns, o.tolerance);
    if (tolerances) std::copy(tolerances, tolerances + ns, tol.begin());
    const hipStream_t s = stream();
    HIPCHK(hipMemcpyAsync(B.ttol.p, tol.data(), ns * sizeof(double), hipMemcpyHostToDevice, s));
    gpu::tiny_gmres_args a;
    a.A = B.gen->A.view;
    a.shift = mg.shift;
    a.b = d_B;
    a.x = d_X;
    a.V = dp<double>(B.tV);
    a.Z = B.tZ.p;
    a.ld = ld;
    a.n = static_cast<int>(n);
    a.m = static_cast<int>(m);
    a.max_iterations = static_cast<int>(std::min<i64>(o.max_iterations, i64{1} << 30));
    a.tolerance = o.tolerance;
    a.result = dp<double>(B.tres);
    a.image = mg.tail_image.p;
    a.image_bytes = mg.tail_image_bytes;
    a.lds_bytes = mg.tail_lds_bytes;
    a.levels = static_cast<int>(mg.precision == MK_FP32 ? mg.L32.size() : mg.L64.size());
    a.omega = mg.opt.jacobi_weight;
    a.n_ent = B.gen->A.padded;
    a.systems = static_cast<int>(ns);
    a.sys_x = n;
    a.sys_v = (m + 1) * ld;
    a.sys_z = m * ld;
    a.tols = dp<double>(B.ttol);
    prof().run(K_OTHER, 0.0, [&] { gpu::gmres_tiny<ZT>(a, s); });
    std::vector<double> r(ns * 3);
    HIPCHK(hipMemcpyAsync(r.data(), B.tres.p, ns * 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    for (i64 j = 0; j < ns; ++j)
        results[j] = mk_solver_result{static_cast<int64_t>(r[3 * j]), r[3 * j + 1], static_cast<int32_t>(r[3 * j + 2])};
    return true;
}

void batch_solve_device(mk_batch_s &B, const double *d_B, double *d_X, const double *tolerances,
                        const mk_gmres_options &o, mk_solver_result *results) {
    require(o.max_iterations >= 0 && o.restart >= 1, "batch gmres: bad options");
    const i64 m = std::max<i64>(1, std::min<i64>(o.restart, B.n));
    require(m <= gpu::kFusedBasis, "batch gmres: restart above 64 is not supported");
    if (B.mg->precision == MK_FP32 ? one_workgroup_batch<float>(B, d_B, d_X, tolerances, o, m, results)
                                   : one_workgroup_batch<double>(B, d_B, d_X, tolerances, o, m, results))
        return;
    ensure_workspace(B, m);
    B.block_pending.assign(B.Bp / kP, 0);
    if (B.block > 1) {
        ensure_block(B);
        for (i64 g = 0; g * kP < B.nsys; ++g) B.block_pending[g] = 1;
    }
    for (i64 j = 0; j < B.Bp; ++j) {
        B.ls[j] = hessenberg_ls(m);
        auto &S = B.slot[j];
        S.tol = j < B.nsys ? (tolerances ? tolerances[j] : o.tolerance) : 0.0;
        S.max_iterations = o.max_iterations;
        S.fresh = j < B.nsys;
        S.active = S.done = false;
    }
    auto &P = prof();
    P.run(K_OTHER, double(B.n) * B.nsys * 16.0, [&] {
        bk::pack(d_B, B.n, B.nsys, dp<double>(B.d_b), B.Bp, stream());
        bk::pack(d_X, B.n, B.nsys, dp<double>(B.d_x), B.Bp, stream());
    });
    clear_lane_tails(B, std::vector<char>(B.Bp, 1));
    while (engine_step(B)) {
    }
    for (i64 j = 0; j < B.nsys; ++j) results[j] = B.slot[j].res;
    P.run(K_OTHER, double(B.n) * B.nsys * 16.0,
          [&] { bk::unpack(dp<double>(B.d_x), B.n, B.nsys, d_X, B.Bp, stream()); });
    drain();
}

} // namespace

// ---- continuous batching for the request server (server.hip) -----------------------------
namespace mk::rt {

mk_batch_s *batch_engine_create(mk_generator_s *gen, mk_sector_mg_s *mg, i64 slots, i64 restart) {
    auto *B = create_engine(gen, mg, slots);
    B->capture = false; // the engine runs beside other host threads
    try {
        const i64 m = std::max<i64>(1, std::min<i64>(restart, B->n));
        require(m <= gpu::kFusedBasis, "batch gmres: restart above 64 is not supported");
        ensure_workspace(*B, m);
        for (auto &l : B->ls) l = hessenberg_ls(m);
    } catch (...) {
        delete B;
        throw;
    }
    return B;
}

void batch_engine_destroy(mk_batch_s *B) { delete B; }

i64 batch_engine_slots(const mk_batch_s &B) { return B.Bp; }

// admit a system to a free slot: b (host, `rows` values), x = 0, its shift, tolerance and
// iteration limit; it takes its initial residual at the next step
void batch_engine_admit(mk_batch_s &B, i64 j, const double *b, i64 rows, double shift, double tol,
                        i64 max_iterations) {
    require(j >= 0 && j < B.Bp && B.slot[j].rows == 0, "batch: the slot is not free");
    set_slot(B, j, shift, rows);
    auto &S = B.slot[j];
    S.tol = tol;
    S.max_iterations = max_iterations;
    S.fresh = true;
    S.active = S.done = false;
    S.res = mk_solver_result{0, 0.0, 0};
    const hipStream_t s = stream();
    if (!B.d_io.p || B.d_io.bytes < static_cast<std::size_t>(B.n) * sizeof(double))
        B.d_io = dbuf(static_cast<std::size_t>(std::max<i64>(B.n, 1)) * sizeof(double));
    HIPCHK(hipMemcpyAsync(B.d_io.p, b, rows * sizeof(double), hipMemcpyHostToDevice, s));
    bk::put_lane(dp<double>(B.d_io), rows, dp<double>(B.d_b), B.n, B.Bp, j, s);
    bk::put_lane(dp<double>(B.d_io), 0, dp<double>(B.d_x), B.n, B.Bp, j, s); // x = 0
    std::vector<char> want(B.Bp, 0);
    want[j] = 1;
    clear_lane_tails(B, want);
    drain();
}

bool batch_engine_step(mk_batch_s &B) { return engine_step(B); }

// a finished slot's result and solution (`rows` values, host); the slot becomes free
bool batch_engine_take(mk_batch_s &B, i64 j, double *x, mk_solver_result *res) {
    auto &S = B.slot[j];
    if (S.rows == 0 || !S.done) return false;
    const hipStream_t s = stream();
    bk::get_lane(dp<double>(B.d_x), S.rows, B.Bp, j, dp<double>(B.d_io), s);
    HIPCHK(hipMemcpyAsync(x, B.d_io.p, S.rows * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    *res = S.res;
    S = slot_state{};
    return true;
}

// free a slot whatever its state (after a failure)
void batch_engine_release(mk_batch_s &B, i64 j) {
    if (j >= 0 && j < B.Bp) B.slot[j] = slot_state{};
}

} // namespace mk::rt

extern "C" {

mk_status mk_batch_create_sized(mk_generator gen, mk_sector_mg mg, int64_t n_systems,
                                const double *shifts, const int64_t *rows, mk_batch *out) {
    return guarded([&] {
        require(gen && mg && shifts && out && n_systems > 0, "batch: bad arguments");
        std::unique_ptr<mk_batch_s> B(create_engine(gen, mg, n_systems));
        B->nsys = n_systems;
        for (i64 j = 0; j < n_systems; ++j) set_slot(*B, j, shifts[j], rows ? rows[j] : B->n);
        *out = B.release();
    });
}

mk_status mk_batch_create(mk_generator gen, mk_sector_mg mg, int64_t n_systems,
                          const double *shifts, mk_batch *out) {
    return mk_batch_create_sized(gen, mg, n_systems, shifts, nullptr, out);
}

mk_status mk_batch_destroy(mk_batch b) {
    delete b;
    return MK_OK;
}

mk_status mk_batch_solve_device(mk_batch b, const double *d_B, double *d_X,
                                const double *tolerances, const mk_gmres_options *options,
                                mk_solver_result *results) {
    return guarded([&] {
        require(b && d_B && d_X && options && results, "batch: bad arguments");
        batch_solve_device(*b, d_B, d_X, tolerances, *options, results);
    });
}

mk_status mk_batch_solve(mk_batch b, const double *B, double *X, const double *tolerances,
                         const mk_gmres_options *options, mk_solver_result *results) {
    return guarded([&] {
        require(b && B && X && options && results, "batch: bad arguments");
        const std::size_t bytes = static_cast<std::size_t>(b->n) * b->nsys * sizeof(double);
        if (!b->d_io.p || b->d_io.bytes < 2 * bytes) b->d_io = dbuf(2 * bytes);
        double *dB = dp<double>(b->d_io), *dX = dB + b->n * b->nsys;
        HIPCHK(hipMemcpy(dB, B, bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice));
        batch_solve_device(*b, dB, dX, tolerances, *options, results);
        HIPCHK(hipMemcpy(X, dX, bytes, hipMemcpyDeviceToHost));
    });
}

mk_status mk_batch_set_block(mk_batch b, int32_t block) {
    return guarded([&] {
        require(b && (block == 1 || block == kP), "batch: block must be 1 or 4");
        if (block > 1) {
            require(b->nsys % kP == 0, "batch: block solves take whole groups of 4 systems");
            for (i64 g = 0; g * kP < b->nsys; ++g)
                for (int q = 1; q < kP; ++q) {
                    const auto &a = b->slot[g * kP], &c = b->slot[g * kP + q];
                    require(a.shift == c.shift && a.rows == c.rows,
                            "batch: the 4 systems of a block group need one shift and one size");
                }
        }
        b->block = block;
    });
}

mk_status mk_batch_graphs_enable(int32_t on) {
    batch_graph_switch().store(on ? 1 : 0);
    return MK_OK;
}

} // extern "C"
