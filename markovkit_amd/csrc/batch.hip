// Batched resolvent solves: (s_j I - Q) x_j = b_j for many systems j over one population box,
// each preconditioned by the sector multigrid of its own shift, all advanced together on the
// device. The reference solves such families one num::gmres call at a time, each with its own
// sector_multigrid (examples/rdme_multigrid_benchmark.cpp:92-121 per shift).
//
// The shift enters the multigrid only through the diagonal of every level operator and the
// dense terminal inverses, so one hierarchy serves all systems: level operators and transfers
// are shared (read once per row for all systems of a wavefront) and the terminal blocks are
// re-inverted once per distinct shift. Each system runs flexible GMRES(m) with delayed CGS2
// and its own Givens least-squares problem; a system leaves the lockstep iteration when its
// residual estimate reaches its tolerance (its lanes then load and store nothing).
#include "batch_kernels.hpp"
#include "runtime.hpp"

#include <type_traits>

using namespace mk::rt;
namespace bk = mk::gpu::batch;

namespace {

template <typename T>
struct blevel {
    dbuf rhs, x, xt, res, crhs, cx; // n_level x Bp work vectors
    dbuf tinv;                      // terminal inverses, one set of per_shift values per shift
    u64 per_shift = 0;
};

} // namespace

struct mk_batch_s {
    mk_generator_s *gen = nullptr;
    mk_sector_mg_s *mg = nullptr;
    i64 n = 0, nsys = 0, Bp = 0, n_shift = 0;
    std::vector<double> shifts; // per padded system (1 for padding)
    dbuf d_shift, d_sidx;
    std::vector<blevel<float>> B32;
    std::vector<blevel<double>> B64;
    // Krylov workspace for restart m
    i64 m = 0;
    int np = 1;
    dbuf V, Z, w, partial, rpart, rtmp, hd, d_b, d_x, d_io, d_scale, d_beta, d_mask, d_active, d_cnt, d_y;
    double *host = nullptr; // pinned: hd mirror, y, scale, beta
    int *ihost = nullptr;   // pinned: mask, active, cnt
    ~mk_batch_s() {
        if (host) (void)hipHostFree(host);
        if (ihost) (void)hipHostFree(ihost);
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
void build_blevels(mk_batch_s &B, const std::vector<level_dev<T>> &L, std::vector<blevel<T>> &BL,
                   const std::vector<double> &distinct) {
    BL.clear();
    BL.resize(L.size());
    const int cycles = B.mg->opt.coarse_cycles;
    for (std::size_t l = 0; l < L.size(); ++l) {
        const auto &lv = L[l];
        auto &bl = BL[l];
        const std::size_t vb = static_cast<std::size_t>(std::max<i64>(lv.n, 1)) * B.Bp * sizeof(T);
        bl.rhs = dbuf(vb);
        bl.x = dbuf(vb);
        bl.xt = dbuf(vb);
        bl.res = dbuf(vb);
        if (cycles > 1) {
            bl.crhs = dbuf(vb);
            bl.cx = dbuf(vb);
        }
        if (lv.n_term > 0) {
            bl.per_shift = lv.h_tq.size();
            std::vector<T> inv(bl.per_shift * distinct.size());
            for (std::size_t d = 0; d < distinct.size(); ++d) {
                for (i64 b = 0; b < lv.n_term; ++b) {
                    const i64 m = lv.h_tsize[b];
                    std::vector<double> a(lv.h_tq.begin() + lv.h_toff[b],
                                          lv.h_tq.begin() + lv.h_toff[b] + m * m);
                    for (i64 i = 0; i < m; ++i) a[i * m + i] += distinct[d];
                    std::vector<double> out;
                    if (!mk::invert_dense(std::move(a), m, out))
                        throw std::runtime_error("population_multigrid: singular block");
                    std::copy(out.begin(), out.end(), inv.begin() + d * bl.per_shift + lv.h_toff[b]);
                }
            }
            bl.tinv = upload(inv);
        }
    }
}

// one batched cycle of level l from a zero guess (population_multigrid::cycle,
// multigrid.hpp:201-228), the per-system form of run_level in api.hip
template <typename T>
T *brun_level(mk_batch_s &B, std::vector<level_dev<T>> &L, std::vector<blevel<T>> &BL,
              std::size_t l, const T *rhs, const double *rhs64, const double *scale, T *out,
              const int *act) {
    auto &lv = L[l];
    auto &bl = BL[l];
    auto &P = prof();
    const hipStream_t s = stream();
    const auto &mg = *B.mg;
    const double omega = mg.opt.jacobi_weight;
    const double *sh = dp<double>(B.d_shift);
    const i64 n = lv.n, Bp = B.Bp;
    const double ts = double(sizeof(T)), nb = double(n) * double(Bp);
    const double sweep_bytes = op_bytes(lv.A, 0, 0, 0) + nb * 4.0 * ts;
    T *cur = dp<T>(bl.x);
    T *other = dp<T>(bl.xt);
    if (rhs64) {
        T *rb = dp<T>(bl.rhs);
        P.run(K_FIRST, nb * (8 + 3 * ts), [&] {
            bk::jacobi_first<T, double>(lv.A.view, sh, lv.pre > 0 ? omega : 0.0, rhs64, scale, rb,
                                        cur, Bp, act, s);
        });
        rhs = rb;
    } else {
        P.run(K_FIRST, nb * 3 * ts, [&] {
            bk::jacobi_first<T, T>(lv.A.view, sh, lv.pre > 0 ? omega : 0.0, rhs, nullptr, nullptr,
                                   cur, Bp, act, s);
        });
    }
    for (int k = 1; k < lv.pre; ++k) {
        P.run(K_SWEEP, sweep_bytes, [&] {
            bk::jacobi_sweep<T>(lv.A.view, sh, omega, rhs, cur, other, Bp, act, s);
        });
        std::swap(cur, other);
    }
    T *res = dp<T>(bl.res);
    P.run(K_RESIDUAL, sweep_bytes, [&] {
        bk::level_residual<T>(lv.A.view, sh, rhs, cur, res, Bp, act, s);
    });
    if (lv.n_term > 0) {
        P.run(K_TERMINAL, 0.0, [&] {
            bk::terminal_solve<T>(lv.n_term, dp<u32>(lv.tstart), dp<u32>(lv.tsize), dp<u64>(lv.toff),
                                  dp<T>(bl.tinv), bl.per_shift, dp<int>(B.d_sidx), lv.term_max, res,
                                  cur, Bp, act, s);
        });
    }
    if (lv.n_next > 0) {
        auto &c = L[l + 1];
        auto &cb = BL[l + 1];
        const double cn = double(c.n) * double(Bp);
        const double rbytes = op_bytes(lv.R, 0, 0, 0) + double(lv.n_fine_linked) * Bp * ts + cn * ts;
        const double pbytes = double(lv.n) * 4 + nb * 2 * ts + cn * ts;
        if (mg.opt.coarse_cycles <= 1 || static_cast<i64>(l) >= cycle_levels(mg.opt)) {
            P.run(K_RESTRICT, rbytes, [&] { bk::gather_sum<T>(lv.R.view, res, dp<T>(cb.rhs), Bp, act, s); });
            T *cx = brun_level<T>(B, L, BL, l + 1, dp<T>(cb.rhs), nullptr, nullptr, nullptr, act);
            P.run(K_PROLONG, pbytes, [&] {
                bk::prolong_add<T>(n, dp<u32>(lv.coarse_of), dp<T>(lv.weight), cx, cur, Bp, act, s);
            });
        } else {
            P.run(K_RESTRICT, rbytes, [&] { bk::gather_sum<T>(lv.R.view, res, dp<T>(cb.crhs), Bp, act, s); });
            P.run(K_OTHER, cn * ts, [&] { gpu::fill<T>(c.n * Bp, T(0), dp<T>(cb.cx), s); });
            const double cbytes = op_bytes(c.A, 0, 0, 0) + cn * 3 * ts;
            for (int cyc = 0; cyc < mg.opt.coarse_cycles; ++cyc) {
                P.run(K_RESIDUAL, cbytes, [&] {
                    bk::level_residual<T>(c.A.view, sh, dp<T>(cb.crhs), dp<T>(cb.cx), dp<T>(cb.rhs), Bp, act, s);
                });
                T *ce = brun_level<T>(B, L, BL, l + 1, dp<T>(cb.rhs), nullptr, nullptr, nullptr, act);
                P.run(K_OTHER, cn * 3 * ts, [&] { gpu::axpy<T>(c.n * Bp, T(1), ce, dp<T>(cb.cx), s); });
            }
            P.run(K_PROLONG, pbytes, [&] {
                bk::prolong_add<T>(n, dp<u32>(lv.coarse_of), dp<T>(lv.weight), dp<T>(cb.cx), cur, Bp, act, s);
            });
        }
    }
    for (int k = 0; k < lv.post; ++k) {
        T *dst = (k == lv.post - 1 && out) ? out : other;
        P.run(K_SWEEP, sweep_bytes, [&] {
            bk::jacobi_sweep<T>(lv.A.view, sh, omega, rhs, cur, dst, Bp, act, s);
        });
        if (dst == other) {
            std::swap(cur, other);
        } else {
            cur = dst;
        }
    }
    if (out && cur != out) {
        P.run(K_OTHER, nb * 2 * ts,
              [&] { HIPCHK(hipMemcpyAsync(out, cur, n * Bp * sizeof(T), hipMemcpyDeviceToDevice, s)); });
        cur = out;
    }
    return cur;
}

void ensure_workspace(mk_batch_s &B, i64 m) {
    if (B.m >= m && B.V.p) return;
    const i64 n = B.n, Bp = B.Bp;
    const std::size_t vec = static_cast<std::size_t>(std::max<i64>(n, 1)) * Bp;
    const std::size_t zs = B.mg->precision == MK_FP32 ? sizeof(float) : sizeof(double);
    B.V = dbuf();
    B.Z = dbuf();
    B.V = dbuf(vec * (m + 1) * sizeof(double));
    B.Z = dbuf(vec * m * zs);
    B.w = dbuf(vec * sizeof(double));
    B.np = bk::parts(n);
    B.partial = dbuf(static_cast<std::size_t>(B.np) * (m + 2) * Bp * sizeof(double));
    B.rpart = dbuf(static_cast<std::size_t>(bk::residual_parts(n)) * Bp * sizeof(double));
    B.rtmp = dbuf(std::max<i64>(1, bk::reduce_tmp_size(static_cast<int>(bk::residual_parts(n)), Bp)) *
                  sizeof(double));
    const slots sl{(m + 2) * Bp};
    B.hd = dbuf(sl.total(Bp) * sizeof(double));
    B.d_scale = dbuf(Bp * sizeof(double));
    B.d_beta = dbuf(Bp * sizeof(double));
    B.d_mask = dbuf(Bp * sizeof(int));
    B.d_active = dbuf(Bp * sizeof(int));
    B.d_cnt = dbuf(Bp * sizeof(int));
    B.d_y = dbuf(m * Bp * sizeof(double));
    if (B.host) (void)hipHostFree(B.host);
    if (B.ihost) (void)hipHostFree(B.ihost);
    B.host = nullptr;
    B.ihost = nullptr;
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B.host),
                         (sl.total(Bp) + m * Bp + 2 * Bp) * sizeof(double)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&B.ihost), 3 * Bp * sizeof(int)));
    B.m = m;
}

// the batched FGMRES(m) with delayed CGS2 (the per-system algorithm of fgmres in api.hip)
template <typename ZT>
void bgmres(mk_batch_s &B, const double *b, double *x, const std::vector<double> &tol,
            const mk_gmres_options &o, mk_solver_result *out) {
    const i64 n = B.n, Bp = B.Bp, nsys = B.nsys;
    const i64 m = std::max<i64>(1, std::min<i64>(o.restart, n));
    require(m <= gpu::kFusedBasis, "batch gmres: restart above 64 is not supported");
    ensure_workspace(B, m);
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
    int *mask_h = B.ihost, *active_h = mask_h + Bp, *cnt_h = active_h + Bp;
    int *d_mask = dp<int>(B.d_mask), *d_active = dp<int>(B.d_active), *d_cnt = dp<int>(B.d_cnt);
    const double nb = double(vec);
    const double spmv_bytes = op_bytes(A.A, 0, 0, 0) + nb * 16.0;

    auto push_ints = [&](const int *h, int *d) {
        HIPCHK(hipMemcpyAsync(d, h, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    };
    // V_0 = b - A x and |r_j|^2 for the active systems
    std::vector<double> beta(Bp, 0.0);
    auto true_residual = [&]() {
        push_ints(active_h, d_active);
        double *rp = dp<double>(B.rpart);
        const int nrp = static_cast<int>(bk::residual_parts(n));
        P.run(K_SPMV, spmv_bytes + nb * 8.0, [&] {
            bk::shifted_apply<double>(A.A.view, sh, x, b, V, rp, Bp, d_active, s);
        });
        P.run(K_DOT, double(nrp) * Bp * 8.0,
              [&] { bk::reduce_partials(rp, nrp, Bp, hd + sl.rn(Bp), s, dp<double>(B.rtmp)); });
        HIPCHK(hipMemcpyAsync(host + sl.rn(Bp), hd + sl.rn(Bp), Bp * sizeof(double), hipMemcpyDeviceToHost, s));
        drain();
        for (i64 j = 0; j < nsys; ++j)
            if (active_h[j]) beta[j] = std::sqrt(host[sl.rn(Bp) + j]);
    };

    for (i64 j = 0; j < Bp; ++j) active_h[j] = j < nsys ? 1 : 0;
    for (i64 j = 0; j < nsys; ++j) out[j] = mk_solver_result{0, 0.0, 0};
    true_residual();
    for (i64 j = 0; j < nsys; ++j) {
        out[j].residual = beta[j];
        if (!(beta[j] > tol[j])) {
            out[j].converged = std::isfinite(beta[j]) ? 1 : 0;
            active_h[j] = 0;
        }
        if (o.max_iterations <= 0) active_h[j] = 0;
    }
    std::vector<hessenberg_ls> ls(nsys, hessenberg_ls(m));
    std::vector<i64> cols(nsys, 0);
    std::vector<double> bj(nsys, 0.0), col(m + 1);
    std::vector<char> flag(nsys, 0);
    auto any = [&](const int *a) {
        for (i64 j = 0; j < nsys; ++j)
            if (a[j]) return true;
        return false;
    };
    while (any(active_h)) {
        for (i64 j = 0; j < Bp; ++j) mask_h[j] = j < nsys ? active_h[j] : 0;
        for (i64 j = 0; j < nsys; ++j) {
            cols[j] = 0;
            if (!mask_h[j]) continue;
            ls[j].start(beta[j]);
            bj[j] = beta[j];
        }
        int which = 0; // s buffer the next update kernel reads
        for (i64 c = 0; c < m && any(mask_h); ++c) {
            const int cnt = static_cast<int>(c);
            for (i64 j = 0; j < Bp; ++j) {
                const bool on = j < nsys && mask_h[j];
                scale_h[j] = on ? 1.0 / bj[j] : 0.0;
                beta_h[j] = on ? bj[j] : 1.0;
            }
            HIPCHK(hipMemcpyAsync(B.d_scale.p, scale_h, Bp * sizeof(double), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B.d_beta.p, beta_h, Bp * sizeof(double), hipMemcpyHostToDevice, s));
            push_ints(mask_h, d_mask);
            // z = M (u / beta), w = A z
            ZT *Zc = Z + c * vec;
            if constexpr (std::is_same_v<ZT, float>) {
                brun_level<float>(B, B.mg->L32, B.B32, 0, nullptr, V + c * vec, dp<double>(B.d_scale), Zc, d_mask);
            } else {
                brun_level<double>(B, B.mg->L64, B.B64, 0, nullptr, V + c * vec, dp<double>(B.d_scale), Zc, d_mask);
            }
            P.run(K_SPMV, op_bytes(A.A, 0, 0, 0) + nb * (sizeof(ZT) + 8.0), [&] {
                bk::shifted_apply<ZT>(A.A.view, sh, Zc, nullptr, w, nullptr, Bp, d_mask, s);
            });
            double *sd = hd + sl.s(which), *sd_next = hd + sl.s(1 - which);
            P.run(K_DOT, double(cnt + 2) * nb * 8.0, [&] {
                bk::multi_dot(V, cnt + 1, w, n, Bp, partial, np, d_mask, s);
                bk::reduce_partials(partial, np, (cnt + 1) * Bp, hd + sl.t(), s);
            });
            P.run(K_UPDATE, double(cnt + 4) * nb * 8.0, [&] {
                bk::dcgs_update(V, cnt, sd, hd + sl.t(), dp<double>(B.d_beta), w, n, Bp, partial, np,
                                hd + sl.eta(), d_mask, s);
                bk::reduce_partials(partial, np, (cnt + 2) * Bp, sd_next, s);
            });
            HIPCHK(hipMemcpyAsync(host, hd, sl.total(Bp) * sizeof(double), hipMemcpyDeviceToHost, s));
            drain();
            const double *t = host + sl.t(), *eta = host + sl.eta();
            double *sn = host + sl.s(1 - which);
            bool any_flag = false;
            for (i64 j = 0; j < nsys; ++j) {
                flag[j] = 0;
                if (!mask_h[j]) continue;
                double ss = 0.0;
                for (i64 k = 0; k <= c; ++k) ss += sn[k * Bp + j] * sn[k * Bp + j];
                if (!(ss <= 0.25 * sn[(c + 1) * Bp + j])) {
                    flag[j] = 1;
                    any_flag = true;
                }
            }
            std::vector<double> first;
            if (any_flag) {
                // explicit second projection of the flagged systems' candidates; the other
                // systems keep their coefficients
                first.assign(sn, sn + (c + 2) * Bp);
                for (i64 j = 0; j < Bp; ++j) cnt_h[j] = (j < nsys && flag[j]) ? 1 : 0;
                push_ints(cnt_h, d_cnt);
                double *u = V + (c + 1) * vec;
                P.run(K_UPDATE, double(cnt + 3) * nb * 8.0, [&] {
                    bk::basis_update(u, V, cnt + 1, sd_next, n, Bp, d_cnt, s);
                });
                P.run(K_DOT, double(cnt + 3) * nb * 8.0, [&] {
                    bk::multi_dot(V, cnt + 2, u, n, Bp, partial, np, d_cnt, s);
                    bk::reduce_partials(partial, np, (cnt + 2) * Bp, sd_next, s);
                });
                HIPCHK(hipMemcpyAsync(host, hd, sl.total(Bp) * sizeof(double), hipMemcpyDeviceToHost, s));
                drain();
                // keep the unflagged systems' first coefficients in the buffer the next step
                // reads
                for (i64 j = 0; j < Bp; ++j)
                    if (!(j < nsys && flag[j]))
                        for (i64 k = 0; k <= c + 1; ++k) sn[k * Bp + j] = first[k * Bp + j];
                HIPCHK(hipMemcpyAsync(sd_next, sn, (c + 2) * Bp * sizeof(double), hipMemcpyHostToDevice, s));
            }
            for (i64 j = 0; j < nsys; ++j) {
                if (!mask_h[j]) continue;
                ++out[j].iterations;
                for (i64 k = 0; k < c; ++k) col[k] = t[k * Bp + j] + sn[k * Bp + j];
                col[c] = eta[j] + sn[c * Bp + j];
                double ss = 0.0;
                for (i64 k = 0; k <= c; ++k) ss += sn[k * Bp + j] * sn[k * Bp + j];
                if (flag[j])
                    for (i64 k = 0; k <= c; ++k) col[k] += first[k * Bp + j];
                const double hn = std::sqrt(std::max(sn[(c + 1) * Bp + j] - ss, 0.0));
                bj[j] = hn;
                cols[j] = c + 1;
                const double est = ls[j].add_column(c, col.data(), hn);
                if (!(est > tol[j]) || !(hn > 0.0) || !std::isfinite(hn) ||
                    out[j].iterations >= o.max_iterations)
                    mask_h[j] = 0;
            }
            which = 1 - which;
        }
        // x_j += Z y_j over each system's columns
        std::fill(yh, yh + m * Bp, 0.0);
        for (i64 j = 0; j < Bp; ++j) cnt_h[j] = 0;
        for (i64 j = 0; j < nsys; ++j) {
            if (!active_h[j] || cols[j] == 0) continue;
            ls[j].solve(cols[j]);
            for (i64 k = 0; k < cols[j]; ++k) yh[k * Bp + j] = ls[j].y[k];
            cnt_h[j] = static_cast<int>(cols[j]);
        }
        HIPCHK(hipMemcpyAsync(B.d_y.p, yh, m * Bp * sizeof(double), hipMemcpyHostToDevice, s));
        push_ints(cnt_h, d_cnt);
        P.run(K_OTHER, nb * (m * sizeof(ZT) + 16.0), [&] {
            bk::combine<ZT>(x, Z, d_cnt, dp<double>(B.d_y), n, Bp, s);
        });
        true_residual();
        for (i64 j = 0; j < nsys; ++j) {
            if (!active_h[j]) continue;
            out[j].residual = beta[j];
            if (!std::isfinite(beta[j])) {
                active_h[j] = 0;
            } else if (!(beta[j] > tol[j])) {
                out[j].converged = 1;
                active_h[j] = 0;
            } else if (out[j].iterations >= o.max_iterations) {
                active_h[j] = 0;
            }
        }
    }
}

void batch_solve_device(mk_batch_s &B, const double *d_B, double *d_X, const double *tolerances,
                        const mk_gmres_options &o, mk_solver_result *results) {
    require(o.max_iterations >= 0 && o.restart >= 1, "batch gmres: bad options");
    const i64 n = B.n, Bp = B.Bp;
    const std::size_t vb = static_cast<std::size_t>(std::max<i64>(n, 1)) * Bp * sizeof(double);
    if (!B.d_b.p || B.d_b.bytes < vb) B.d_b = dbuf(vb);
    if (!B.d_x.p || B.d_x.bytes < vb) B.d_x = dbuf(vb);
    std::vector<double> tol(B.nsys, o.tolerance);
    if (tolerances) std::copy(tolerances, tolerances + B.nsys, tol.begin());
    auto &P = prof();
    P.run(K_OTHER, double(n) * B.nsys * 16.0, [&] {
        bk::pack(d_B, n, B.nsys, dp<double>(B.d_b), Bp, stream());
        bk::pack(d_X, n, B.nsys, dp<double>(B.d_x), Bp, stream());
    });
    if (B.mg->precision == MK_FP32) {
        bgmres<float>(B, dp<double>(B.d_b), dp<double>(B.d_x), tol, o, results);
    } else {
        bgmres<double>(B, dp<double>(B.d_b), dp<double>(B.d_x), tol, o, results);
    }
    P.run(K_OTHER, double(n) * B.nsys * 16.0,
          [&] { bk::unpack(dp<double>(B.d_x), n, B.nsys, d_X, Bp, stream()); });
    drain();
}

} // namespace

extern "C" {

mk_status mk_batch_create(mk_generator gen, mk_sector_mg mg, int64_t n_systems,
                          const double *shifts, mk_batch *out) {
    return guarded([&] {
        require(gen && mg && shifts && out && n_systems > 0, "batch: bad arguments");
        require(mg->rows == gen->n, "batch: the preconditioner does not match the operator");
        require(mg->opt.smoother == MK_JACOBI, "batch: only the Jacobi smoother is batched");
        for (i64 j = 0; j < n_systems; ++j)
            if (!(shifts[j] > 0.0)) throw std::invalid_argument("population_multigrid: shift must be positive");
        auto B = std::make_unique<mk_batch_s>();
        B->gen = gen;
        B->mg = mg;
        B->n = gen->n;
        B->nsys = n_systems;
        B->Bp = (n_systems + bk::kLanes - 1) / bk::kLanes * bk::kLanes;
        B->shifts.assign(B->Bp, 1.0);
        std::copy(shifts, shifts + n_systems, B->shifts.begin());
        std::vector<double> distinct;
        std::vector<int> sidx(B->Bp, 0);
        for (i64 j = 0; j < n_systems; ++j) {
            auto it = std::find(distinct.begin(), distinct.end(), shifts[j]);
            if (it == distinct.end()) {
                distinct.push_back(shifts[j]);
                it = distinct.end() - 1;
            }
            sidx[j] = static_cast<int>(it - distinct.begin());
        }
        B->n_shift = static_cast<i64>(distinct.size());
        B->d_shift = upload(B->shifts);
        B->d_sidx = upload(sidx);
        if (mg->precision == MK_FP32) {
            build_blevels<float>(*B, mg->L32, B->B32, distinct);
        } else {
            build_blevels<double>(*B, mg->L64, B->B64, distinct);
        }
        *out = B.release();
    });
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

} // extern "C"
