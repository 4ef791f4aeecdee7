// Krylov exponential w = exp(t Q) v of a generator on the device: the num::expv the reference's
// FSP solvers call (include/markovkit/fsp.hpp:152, :225; numerics source absent), as Sidje's
// expv (Expokit, ACM TOMS 24, 1998). Each step builds an Arnoldi basis of dimension m for
// B = 0 I - Q = -Q with classical Gram-Schmidt twice, whose coefficients and norms stay on the
// device (no host synchronisation inside a step); the host then reads the Hessenberg matrix,
// exponentiates the (m + 2) x (m + 2) augmented matrix (degree-6 Pade, scaling and squaring),
// estimates the local error from the two corrected terms and accepts or shrinks the step
// (gamma 0.9, delta 1.2, at most 10 rejections); w = V (beta F e_1) on the device.
//
// Batched exponentials (mk_expv_batch_s): many vectors over one generator advance together in
// the state-major layout of the batched resolvent solves (vector j's state i at X[i * Bp + j]).
// Every substep runs one Arnoldi process for all running lanes (the operator's entries read
// once per row for 64 lanes, per-lane classical Gram-Schmidt twice with fixed-order reductions),
// then each lane's step control on the host. Lanes keep their own step sizes, time and error;
// the basis does not depend on the step, so a lane's rejections only recompute its small
// exponential. A lane's arithmetic never mixes with another lane's, so its result does not
// depend on what runs beside it. Vectors join at substep boundaries (continuous batching).
#include "batch_kernels.hpp"
#include "runtime.hpp"

using namespace mk::rt;

namespace {

struct expv_workspace {
    i64 n = 0, ld = 0, m = 0;
    dbuf V, p, x, partial, hd;
    double *host = nullptr;
    graph_cache graphs; // one captured Arnoldi per (generator, n, m)
    ~expv_workspace() {
        graphs.clear();
        if (host) (void)hipHostFree(host);
    }
    // device scalars: h1 and h2 projections ((m + 1) x m each, column-major), |p|^2 per step
    // (m + 1), then beta^2 and the coefficients y (m + 1)
    i64 h1() const { return 0; }
    i64 h2() const { return (m + 1) * m; }
    i64 nrm() const { return 2 * (m + 1) * m; }
    i64 beta2() const { return nrm() + m + 1; }
    i64 y() const { return beta2() + 1; }
    i64 total() const { return y() + m + 1; }
    void ensure(i64 n_, i64 m_) {
        if (n_ <= n && m_ <= m && V.p) return;
        n = std::max(n, n_);
        m = std::max(m, m_);
        ld = (n + 63) / 64 * 64;
        V = dbuf();
        V = dbuf(static_cast<std::size_t>(ld) * (m + 1) * sizeof(double));
        p = dbuf(static_cast<std::size_t>(ld) * sizeof(double));
        x = dbuf(static_cast<std::size_t>(ld) * sizeof(double));
        graphs.clear();
        partial = dbuf(static_cast<std::size_t>(gpu::reduction_blocks(n)) * (m + 2) * sizeof(double));
        hd = dbuf(total() * sizeof(double));
        if (host) (void)hipHostFree(host);
        host = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&host), total() * sizeof(double)));
    }
};

expv_workspace &expv_ws() {
    thread_local expv_workspace w;
    return w;
}

double round_step(double step) {
    const double s = std::pow(10.0, std::floor(std::log10(step)) - 1.0);
    return std::ceil(step / s) * s;
}

constexpr double kBreakdown = 1e-7, kGamma = 0.9, kDelta = 1.2;
constexpr int kMaxReject = 10;

// Expokit's step control of one vector (expv.f): the first step size from the norms, and per
// step the Hessenberg matrix from the Arnoldi coefficients, the augmented exponential (degree-6
// Pade), the local error from the two corrected terms, rejections that shrink the step, the
// coefficients y = beta F e_1 and the next step size. The Arnoldi basis does not depend on the
// step size, so a rejected step only recomputes the small exponential.
struct expv_control {
    i64 m = 1;
    double t_out = 0.0, sgn = -1.0, tol = 0.0, anorm = 1.0, rndoff = 0.0;
    double t_now = 0.0, t_new = 0.0, t_step = 0.0, xm = 1.0, beta = 0.0, err_loc = 0.0;
    mk_expv_result res{0, 0, 0.0};

    // false: nothing to integrate (a zero vector or t = 0: w = v)
    bool start(double t, double beta0, const mk_generator_s &G, double tolerance, i64 m_) {
        m = m_;
        t_out = std::abs(t);
        sgn = t >= 0.0 ? -1.0 : 1.0; // exp(t Q) = exp(-t B)
        tol = tolerance;
        anorm = G.norm1 > 0.0 ? G.norm1 : 1.0;
        rndoff = anorm * std::numeric_limits<double>::epsilon();
        beta = beta0;
        if (!(beta > 0.0) || t_out == 0.0) return false;
        xm = 1.0 / double(m);
        const double fact = std::pow(double(m + 1) / std::exp(1.0), double(m + 1)) *
                            std::sqrt(2.0 * 3.141592653589793 * double(m + 1));
        t_new = round_step((1.0 / anorm) * std::pow((fact * tol) / (4.0 * beta * anorm), xm));
        t_now = 0.0;
        return true;
    }
    bool running() const { return t_now < t_out; }

    // one step from the Arnoldi coefficients h(i, j) (both projection passes summed, i <= j < m)
    // and the squared norms nr(j) of the remainders (j <= m); y receives the coefficients of
    // the basis vectors; returns their number
    template <typename HF, typename NF>
    i64 plan(HF &&h, NF &&nr, std::vector<double> &y) {
        ++res.steps;
        t_step = std::min(t_out - t_now, t_new);
        const i64 L = m + 2;
        std::vector<double> H(static_cast<std::size_t>(L * L), 0.0);
        int k1 = 2;
        i64 mb = m;
        for (i64 j = 0; j < m; ++j) {
            for (i64 i = 0; i <= j; ++i) H[i * L + j] = h(i, j);
            const double sj = std::sqrt(nr(j));
            if (!(sj >= kBreakdown)) {
                k1 = 0;
                mb = j + 1;
                t_step = t_out - t_now;
                break;
            }
            H[(j + 1) * L + j] = sj;
        }
        double avnorm = 0.0;
        if (k1 != 0) {
            H[(m + 1) * L + m] = 1.0;
            avnorm = std::sqrt(nr(m));
        }
        std::vector<double> F;
        i64 mx = mb + k1;
        err_loc = kBreakdown;
        for (int rejected = 0;; ++rejected) {
            mx = mb + k1;
            std::vector<double> X(static_cast<std::size_t>(mx * mx));
            for (i64 i = 0; i < mx; ++i)
                for (i64 j = 0; j < mx; ++j) X[i * mx + j] = sgn * t_step * H[i * L + j];
            F = mk::expm_pade(std::move(X), mx);
            if (k1 == 0) {
                err_loc = kBreakdown;
                break;
            }
            const double phi1 = std::abs(beta * F[m * mx]);
            const double phi2 = std::abs(beta * F[(m + 1) * mx] * avnorm);
            if (phi1 > 10.0 * phi2) {
                err_loc = phi2;
                xm = 1.0 / double(m);
            } else if (phi1 > phi2) {
                err_loc = (phi1 * phi2) / (phi1 - phi2);
                xm = 1.0 / double(m);
            } else {
                err_loc = phi1;
                xm = 1.0 / double(std::max<i64>(1, m - 1));
            }
            if (err_loc <= kDelta * t_step * tol) break;
            if (rejected == kMaxReject) throw std::runtime_error("expv: the requested tolerance is too high");
            t_step = round_step(kGamma * t_step * std::pow(t_step * tol / err_loc, xm));
            ++res.rejected;
        }
        const i64 use = mb + std::max(0, k1 - 1);
        y.resize(static_cast<std::size_t>(use));
        for (i64 k = 0; k < use; ++k) y[k] = beta * F[k * mx];
        return use;
    }

    // the step is taken: beta_new = |w|; false when the iterate vanished or is not finite
    bool accept(double beta_new) {
        beta = beta_new;
        t_now += t_step;
        t_new = round_step(kGamma * t_step * std::pow(t_step * tol / err_loc, xm));
        res.error += std::max(err_loc, rndoff);
        return beta > 0.0 && std::isfinite(beta);
    }
};

mk_expv_result expv_device(mk_generator_s &G, double t, const double *v, double *w_out,
                           const mk_expv_options &o) {
    require(o.krylov_dim >= 1 && o.tolerance > 0.0, "expv: bad options");
    const i64 n = G.n;
    const i64 m = std::max<i64>(1, std::min<i64>(o.krylov_dim, n));
    require(m + 1 <= 512, "expv: krylov_dim above 511 is not supported on the device");
    auto &ws = expv_ws();
    ws.ensure(n, m);
    auto &P = prof();
    const hipStream_t s = stream();
    const int nb = gpu::reduction_blocks(n);
    const i64 ld = ws.ld;
    double *V = dp<double>(ws.V), *pv = dp<double>(ws.p), *partial = dp<double>(ws.partial);
    double *hd = dp<double>(ws.hd);
    const double spmv_bytes = op_bytes(G.A, 8, 8, 1);

    // beta^2 = |w|^2 into hd[beta2]
    auto norm2_to = [&](const double *x, double *out) {
        P.run(K_DOT, n * 8.0, [&] {
            gpu::multi_dot(x, n, 1, x, n, partial, nb, s);
            gpu::reduce_partials(partial, nb, 1, out, s);
        });
    };
    auto read = [&](i64 off, i64 count) {
        HIPCHK(hipMemcpyAsync(ws.host + off, hd + off, count * sizeof(double), hipMemcpyDeviceToHost, s));
        drain();
    };
    // the iterate lives in the workspace (fixed pointers for the captured Arnoldi)
    double *w = dp<double>(ws.x);
    HIPCHK(hipMemcpyAsync(w, v, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    expv_control c;
    auto finish = [&] {
        HIPCHK(hipMemcpyAsync(w_out, w, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        return c.res;
    };
    norm2_to(w, hd + ws.beta2());
    read(ws.beta2(), 1);
    if (!c.start(t, std::sqrt(ws.host[ws.beta2()]), G, o.tolerance, m)) return finish();
    std::vector<double> y;
    while (c.running()) {
        ws.graphs.run(graph_key{G.uid, 0, n, m, 0, 0, 0}, s, [&] {
        // V_0 = w / beta (beta^2 still in hd[beta2])
        P.run(K_OTHER, n * 16.0, [&] { gpu::scale_by_norm(n, w, hd + ws.beta2(), V, s); });
        for (i64 j = 0; j < m; ++j) {
            const int cnt = static_cast<int>(j + 1);
            P.run(K_SPMV, spmv_bytes, [&] {
                gpu::shifted_apply<double>(G.A.view, 0.0, V + j * ld, nullptr, pv, nullptr, s);
            });
            for (int pass = 0; pass < 2; ++pass) {
                double *h = hd + (pass == 0 ? ws.h1() : ws.h2()) + j * (m + 1);
                P.run(K_DOT, double(cnt + 1) * n * 8.0, [&] {
                    gpu::multi_dot(V, ld, cnt, pv, n, partial, nb, s);
                    gpu::reduce_partials(partial, nb, cnt, h, s);
                });
                double *np = pass == 1 ? partial + static_cast<i64>(nb) * (m + 1) : nullptr;
                P.run(K_UPDATE, double(cnt + 2) * n * 8.0, [&] {
                    gpu::basis_update(pv, V, ld, cnt, h, n, np, nb, s);
                });
            }
            P.run(K_DOT, nb * 8.0, [&] {
                gpu::reduce_partials(partial + static_cast<i64>(nb) * (m + 1), nb, 1, hd + ws.nrm() + j, s);
            });
            P.run(K_OTHER, n * 16.0, [&] { gpu::scale_by_norm(n, pv, hd + ws.nrm() + j, V + (j + 1) * ld, s); });
        }
        // |B V_m|^2 for the error estimate
        P.run(K_SPMV, spmv_bytes, [&] {
            gpu::shifted_apply<double>(G.A.view, 0.0, V + m * ld, nullptr, pv, nullptr, s);
        });
        norm2_to(pv, hd + ws.nrm() + m);
        }, n);
        read(0, ws.beta2());
        const double *h1 = ws.host + ws.h1(), *h2 = ws.host + ws.h2(), *nr = ws.host + ws.nrm();
        const i64 use = c.plan([&](i64 i, i64 j) { return h1[j * (m + 1) + i] + h2[j * (m + 1) + i]; },
                               [&](i64 j) { return nr[j]; }, y);
        double *yh = ws.host + ws.y();
        std::copy(y.begin(), y.end(), yh);
        HIPCHK(hipMemcpyAsync(hd + ws.y(), yh, use * sizeof(double), hipMemcpyHostToDevice, s));
        P.run(K_OTHER, double(use) * n * 8.0 + n * 16.0, [&] {
            gpu::fill<double>(n, 0.0, w, s);
            gpu::combine<double>(w, V, ld, static_cast<int>(use), hd + ws.y(), n, s);
        });
        norm2_to(w, hd + ws.beta2());
        read(ws.beta2(), 1);
        if (!c.accept(std::sqrt(ws.host[ws.beta2()]))) break;
    }
    return finish();
}

} // namespace

namespace bk = mk::gpu::batch;

struct mk_expv_batch_s {
    struct lane {
        bool active = false, pending = false, done = false;
        double t = 0.0, tol = 0.0;
        expv_control c;
        std::vector<double> y;
        std::string error;
    };
    mk_generator_s *gen = nullptr;
    i64 n = 0, Bp = 0, slots = 0, m = 0;
    int np = 0;
    dbuf W, V, p, partial, tmp, hd, d_lim, d_cnt, d_rows, d_sel, d_shift, stage, d_lflag;
    double *host = nullptr; // pinned mirror of hd
    int *ihost = nullptr;   // pinned: limits | counts | rows | sel
    std::vector<lane> L;
    ~mk_expv_batch_s() {
        if (host) (void)hipHostFree(host);
        if (ihost) (void)hipHostFree(ihost);
    }
    // device scalars, each [.. * Bp + lane]: per Arnoldi step the projections of both passes
    // (m steps of m + 1 slots; delayed CGS2: Q^T w and u^T w, then the candidate's projections
    // and |u|^2), the remainders' squared norms (m + 1; delayed CGS2 uses the last, |B V_m|^2),
    // |w|^2 and the coefficients y (m + 1); for delayed CGS2 then the explicit second
    // projections (m steps of m + 2), eta (m), beta (m + 1), the flags (m), the current s (m + 1)
    // and the next step's scale, beta (for the update) and beta^2
    i64 h1(i64 j) const { return j * (m + 1) * Bp; }
    i64 h2(i64 j) const { return (m * (m + 1) + j * (m + 1)) * Bp; }
    i64 nrm(i64 j) const { return (2 * m * (m + 1) + j) * Bp; }
    i64 beta2() const { return (2 * m * (m + 1) + m + 1) * Bp; }
    i64 y() const { return beta2() + Bp; }
    i64 s2(i64 j) const { return y() + (m + 1) * Bp + j * (m + 2) * Bp; }
    i64 eta(i64 j) const { return s2(m) + j * Bp; }
    i64 bet(i64 j) const { return eta(m) + j * Bp; }
    i64 flg(i64 j) const { return bet(m + 1) + j * Bp; }
    i64 scur() const { return flg(m); }
    i64 dscale() const { return scur() + (m + 1) * Bp; }
    i64 dbeta() const { return dscale() + Bp; }
    i64 dbeta2() const { return dbeta() + Bp; }
    i64 readback() const { return scur(); }
    i64 total() const { return dbeta2() + Bp; }
    double *dv(i64 k) const { return dp<double>(V) + k * n * Bp; }
};

namespace {

using expv_batch = mk_expv_batch_s;

// limits (n for the lanes in `on`, else 0) to the device
void push_lanes(expv_batch &E, const std::vector<char> &on) {
    for (i64 j = 0; j < E.Bp; ++j) E.ihost[j] = on[j] ? static_cast<int>(E.n) : 0;
    HIPCHK(hipMemcpyAsync(E.d_lim.p, E.ihost, E.Bp * sizeof(int), hipMemcpyHostToDevice, stream()));
}

// |W_j|^2 of the lanes in `on` into hd[beta2] and its host mirror
void lane_norms(expv_batch &E, const std::vector<char> &on) {
    const hipStream_t s = stream();
    push_lanes(E, on);
    double *hd = dp<double>(E.hd);
    const int *lim = dp<int>(E.d_lim);
    bk::multi_dot(dp<double>(E.W), 1, dp<double>(E.W), E.n, E.Bp, dp<double>(E.partial), E.np, lim, s);
    bk::reduce_partials(dp<double>(E.partial), E.np, E.Bp, hd + E.beta2(), s, dp<double>(E.tmp));
    HIPCHK(hipMemcpyAsync(E.host + E.beta2(), hd + E.beta2(), E.Bp * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
}

} // namespace

namespace mk::rt {

expv_batch *expv_engine_create(mk_generator_s *gen, i64 slots, i64 krylov_dim) {
    require(gen && slots >= 1 && slots <= 4096 && krylov_dim >= 1, "expv batch: bad arguments");
    auto E = std::make_unique<expv_batch>();
    E->gen = gen;
    E->n = gen->n;
    E->slots = slots;
    E->Bp = (slots + bk::kLanes - 1) / bk::kLanes * bk::kLanes;
    E->m = std::max<i64>(1, std::min<i64>(krylov_dim, E->n));
    require(E->m + 1 <= 512, "expv: krylov_dim above 511 is not supported on the device");
    const i64 n = E->n, Bp = E->Bp, m = E->m;
    const std::size_t vec = static_cast<std::size_t>(std::max<i64>(n, 1)) * Bp * sizeof(double);
    E->W = dbuf(vec);
    E->V = dbuf(vec * (m + 1));
    E->p = dbuf(vec);
    E->np = bk::parts_wide(n);
    E->partial = dbuf(static_cast<std::size_t>(E->np) * (m + 1) * Bp * sizeof(double));
    E->tmp = dbuf(static_cast<std::size_t>(std::max<i64>(1, bk::reduce_tmp_size(E->np, (m + 1) * Bp))) *
                  sizeof(double));
    E->hd = dbuf(E->total() * sizeof(double));
    E->d_lim = dbuf(Bp * sizeof(int));
    E->d_cnt = dbuf(Bp * sizeof(int));
    E->d_rows = dbuf(Bp * sizeof(int));
    E->d_sel = dbuf(Bp * sizeof(int));
    E->d_shift = dbuf(Bp * sizeof(double));
    E->d_lflag = dbuf(Bp * sizeof(int));
    E->stage = dbuf(static_cast<std::size_t>(std::max<i64>(n, 1)) * sizeof(double));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&E->host), E->total() * sizeof(double)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&E->ihost), 4 * Bp * sizeof(int)));
    const hipStream_t s = stream();
    // B = 0 I - Q: every lane's shift is zero; the iterate starts at zero on every lane
    HIPCHK(hipMemsetAsync(E->d_shift.p, 0, Bp * sizeof(double), s));
    HIPCHK(hipMemsetAsync(E->W.p, 0, vec, s));
    HIPCHK(hipMemsetAsync(E->hd.p, 0, E->total() * sizeof(double), s));
    drain();
    E->L.resize(Bp);
    return E.release();
}

void expv_engine_destroy(expv_batch *E) {
    if (!E) return;
    drain();
    delete E;
}

i64 expv_engine_slots(const expv_batch &E) { return E.slots; }
i64 expv_engine_dim(const expv_batch &E) { return E.m; }

// lane j takes v (n host values) for exp(t Q) v to the tolerance; it starts at the next step
void expv_engine_admit(expv_batch &E, i64 j, const double *v, double t, double tol) {
    require(j >= 0 && j < E.slots && !E.L[j].active, "expv batch: the slot is busy");
    require(v && tol > 0.0 && std::isfinite(t), "expv: bad options");
    const hipStream_t s = stream();
    HIPCHK(hipMemcpyAsync(E.stage.p, v, E.n * sizeof(double), hipMemcpyHostToDevice, s));
    bk::put_lane(dp<double>(E.stage), E.n, dp<double>(E.W), E.n, E.Bp, j, s);
    drain();
    auto &l = E.L[j];
    l = expv_batch::lane{};
    l.active = l.pending = true;
    l.t = t;
    l.tol = tol;
}

// one substep of every running lane; false when none is left running
bool expv_engine_step(expv_batch &E) {
    const i64 n = E.n, Bp = E.Bp, m = E.m;
    const hipStream_t s = stream();
    auto &P = prof();
    // lanes admitted since the last step: their norms and first step sizes (the norms of every
    // unfinished lane are recomputed with them, the same values for the lanes already running:
    // a reduction writes every lane's slot)
    std::vector<char> fresh(Bp, 0), live(Bp, 0);
    bool any_fresh = false;
    for (i64 j = 0; j < Bp; ++j) {
        if (E.L[j].active && E.L[j].pending) fresh[j] = any_fresh = true;
        live[j] = E.L[j].active && !E.L[j].done;
    }
    if (any_fresh) {
        lane_norms(E, live);
        for (i64 j = 0; j < Bp; ++j) {
            if (!fresh[j]) continue;
            auto &l = E.L[j];
            l.pending = false;
            // a zero vector or t = 0: w = v, already in the lane
            if (!l.c.start(l.t, std::sqrt(E.host[E.beta2() + j]), *E.gen, l.tol, m)) l.done = true;
        }
    }
    std::vector<char> run(Bp, 0);
    bool any = false;
    for (i64 j = 0; j < Bp; ++j)
        if (E.L[j].active && !E.L[j].done) run[j] = any = true;
    if (!any) return false;
    push_lanes(E, run);
    const int *lim = dp<int>(E.d_lim);
    double *hd = dp<double>(E.hd), *pv = dp<double>(E.p), *part = dp<double>(E.partial);
    double *tmp = dp<double>(E.tmp), *W = dp<double>(E.W);
    const double *sh = dp<double>(E.d_shift);
    const auto &A = E.gen->A;
    const double nb = double(n) * Bp;
    const double spmv_bytes = op_bytes(A, 0, 0, 0) + nb * 16.0;
    const bool delayed = m < gpu::kFusedBasis;
    // the Arnoldi process of every running lane: V_0 = W / beta, then m steps; the Hessenberg
    // coefficients and norms to the host
    auto arnoldi = [&](bool reproject) {
        P.run(K_OTHER, nb * 16.0, [&] { bk::scale_lanes(W, hd + E.beta2(), E.dv(0), n, Bp, lim, s); });
        if (delayed) {
            // delayed CGS2 (the batched FGMRES's Arnoldi step) with the normalisation on the
            // device: step j applies B to the candidate u_j / beta_j (u_0 = q_0, beta_0 = 1), one
            // reduction gives Q^T w and u_j^T w, the update finalises q_j and forms u_{j+1} with
            // its projections. Where the Pythagorean norm of u_{j+1} would cancel, the candidate
            // needs an explicit second projection: without reproject the steps only record the
            // flags (the host then runs the process again with reproject, which launches the
            // second projection for the flagged lanes)
            int *lflag = dp<int>(E.d_lflag);
            gpu::fill<double>(Bp, 1.0, hd + E.dscale(), s);
            gpu::fill<double>(Bp, 1.0, hd + E.dbeta(), s);
            for (i64 j = 0; j < m; ++j) {
                const int cnt = static_cast<int>(j);
                P.run(K_SPMV, spmv_bytes, [&] {
                    bk::shifted_apply<double>(A.view, sh, E.dv(j), nullptr, pv, nullptr, Bp, lim, s,
                                              hd + E.dscale());
                });
                P.run(K_DOT, double(cnt + 2) * nb * 8.0, [&] {
                    bk::multi_dot(E.dv(0), cnt + 1, pv, n, Bp, part, E.np, lim, s);
                    bk::reduce_partials(part, E.np, static_cast<i64>(cnt + 1) * Bp, hd + E.h1(j), s, tmp);
                });
                P.run(K_UPDATE, double(cnt + 4) * nb * 8.0, [&] {
                    bk::dcgs_update(E.dv(0), cnt, hd + E.scur(), hd + E.h1(j), hd + E.dbeta(), pv, n, Bp,
                                    part, E.np, hd + E.eta(j), lim, s);
                    bk::reduce_partials(part, E.np, static_cast<i64>(cnt + 2) * Bp, hd + E.h2(j), s, tmp);
                });
                P.run(K_OTHER, 0.0, [&] {
                    if (reproject) {
                        bk::dcgs_flags(hd + E.h2(j), cnt, Bp, lim, lflag, hd + E.flg(j), s);
                        bk::basis_update(E.dv(j + 1), E.dv(0), cnt + 1, hd + E.h2(j), n, Bp, lflag, s);
                        bk::multi_dot(E.dv(0), cnt + 2, E.dv(j + 1), n, Bp, part, E.np, lflag, s);
                        bk::reduce_partials(part, E.np, static_cast<i64>(cnt + 2) * Bp, hd + E.s2(j), s, tmp);
                        bk::dcgs_norms(hd + E.h2(j), hd + E.s2(j), hd + E.flg(j), cnt, Bp, lim, hd + E.scur(),
                                       hd + E.bet(j + 1), hd + E.dbeta(), hd + E.dscale(), hd + E.dbeta2(), s);
                    } else {
                        bk::dcgs_norms(hd + E.h2(j), nullptr, nullptr, cnt, Bp, lim, hd + E.scur(),
                                       hd + E.bet(j + 1), hd + E.dbeta(), hd + E.dscale(), hd + E.dbeta2(), s,
                                       hd + E.flg(j));
                    }
                });
            }
            // q_m = (u_m - Q s_m) / beta_m
            P.run(K_UPDATE, double(m + 2) * nb * 8.0 + nb * 16.0, [&] {
                bk::basis_update(E.dv(m), E.dv(0), static_cast<int>(m), hd + E.scur(), n, Bp, lim, s);
                bk::scale_lanes(E.dv(m), hd + E.dbeta2(), E.dv(m), n, Bp, lim, s);
            });
        }
        for (i64 j = 0; j < m && !delayed; ++j) {
            const int cnt = static_cast<int>(j + 1);
            P.run(K_SPMV, spmv_bytes, [&] { bk::shifted_apply<double>(A.view, sh, E.dv(j), nullptr, pv, nullptr, Bp, lim, s); });
            for (int pass = 0; pass < 2; ++pass) {
                double *h = hd + (pass == 0 ? E.h1(j) : E.h2(j));
                P.run(K_DOT, double(cnt + 1) * nb * 8.0, [&] {
                    bk::multi_dot(E.dv(0), cnt, pv, n, Bp, part, E.np, lim, s);
                    bk::reduce_partials(part, E.np, static_cast<i64>(cnt) * Bp, h, s, tmp);
                });
                P.run(K_UPDATE, double(cnt + 2) * nb * 8.0, [&] { bk::basis_update(pv, E.dv(0), cnt, h, n, Bp, lim, s); });
            }
            P.run(K_DOT, nb * 8.0, [&] {
                bk::multi_dot(pv, 1, pv, n, Bp, part, E.np, lim, s);
                bk::reduce_partials(part, E.np, Bp, hd + E.nrm(j), s, tmp);
            });
            P.run(K_OTHER, nb * 16.0, [&] { bk::scale_lanes(pv, hd + E.nrm(j), E.dv(j + 1), n, Bp, lim, s); });
        }
        // |B V_m|^2 for the error estimates
        P.run(K_SPMV, spmv_bytes, [&] { bk::shifted_apply<double>(A.view, sh, E.dv(m), nullptr, pv, nullptr, Bp, lim, s); });
        P.run(K_DOT, nb * 8.0, [&] {
            bk::multi_dot(pv, 1, pv, n, Bp, part, E.np, lim, s);
            bk::reduce_partials(part, E.np, Bp, hd + E.nrm(m), s, tmp);
        });
        HIPCHK(hipMemcpyAsync(E.host, hd, (delayed ? E.readback() : E.beta2()) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        drain();
    };
    arnoldi(false);
    if (delayed) {
        bool flagged = false;
        for (i64 j = 0; j < Bp && !flagged; ++j)
            for (i64 k = 0; k < m && run[j] && !flagged; ++k) flagged = E.host[E.flg(k) + j] != 0.0;
        if (flagged) arnoldi(true);
    }
    // every running lane's step control, in parallel on the host
    std::vector<i64> use(Bp, 0);
    mk::parallel_tasks(Bp, [&](i64 j) {
        if (!run[j]) return;
        auto &l = E.L[j];
        const double *H = E.host;
        try {
            if (delayed) {
                // column k of the Hessenberg matrix from step k's coefficients: the expansion
                // of w = B u_k / beta_k is (t + s_{k+1} [+ the first projection where flagged],
                // eta + ..., beta_{k+1}); since u_k / beta_k = q_k + Q s_k / beta_k, the columns
                // of the earlier basis vectors it carries are subtracted: h_k -= sum_i
                // (s_k[i] / beta_k) h_i, which makes B Q_m = Q_{m+1} H hold for the finalised q
                std::vector<double> Hc(static_cast<std::size_t>((m + 1) * m), 0.0); // column-major
                auto flagged = [&](i64 k) { return H[E.flg(k) + j] != 0.0; };
                auto sn = [&](i64 k, i64 i) { return H[(flagged(k) ? E.s2(k) : E.h2(k)) + i * Bp + j]; };
                for (i64 k = 0; k < m; ++k) {
                    double *c = Hc.data() + k * (m + 1);
                    for (i64 i = 0; i <= k; ++i) {
                        c[i] = (i < k ? H[E.h1(k) + i * Bp + j] : H[E.eta(k) + j]) + sn(k, i);
                        if (flagged(k)) c[i] += H[E.h2(k) + i * Bp + j];
                    }
                    c[k + 1] = H[E.bet(k + 1) + j];
                    const double bk_ = k == 0 ? 1.0 : H[E.bet(k) + j];
                    if (k > 0 && bk_ > 0.0) {
                        for (i64 q = 0; q < k; ++q) {
                            const double f = sn(k - 1, q) / bk_;
                            const double *cq = Hc.data() + q * (m + 1);
                            for (i64 i = 0; i <= q + 1; ++i) c[i] -= f * cq[i];
                        }
                    }
                }
                use[j] = l.c.plan([&](i64 i, i64 k) { return Hc[k * (m + 1) + i]; },
                                  [&](i64 k) {
                                      if (k == m) return H[E.nrm(m) + j];
                                      const double b = H[E.bet(k + 1) + j];
                                      return b * b;
                                  },
                                  l.y);
            } else {
                use[j] = l.c.plan([&](i64 i, i64 k) { return H[E.h1(k) + i * Bp + j] + H[E.h2(k) + i * Bp + j]; },
                                  [&](i64 k) { return H[E.nrm(k) + j]; }, l.y);
            }
        } catch (const std::exception &e) {
            l.error = e.what();
            use[j] = -1;
        }
    });
    double *yh = E.host + E.y();
    int *cnt = E.ihost + Bp, *rows = E.ihost + 2 * Bp, *sel = E.ihost + 3 * Bp;
    std::fill(yh, yh + (m + 1) * Bp, 0.0);
    for (i64 j = 0; j < Bp; ++j) {
        cnt[j] = run[j] && use[j] > 0 ? static_cast<int>(use[j]) : 0;
        rows[j] = 0;
        sel[j] = cnt[j] > 0 ? 1 : 0;
        for (i64 k = 0; k < cnt[j]; ++k) yh[k * Bp + j] = E.L[j].y[k];
        if (run[j] && use[j] < 0) {
            E.L[j].done = true; // the step control failed: the lane ends with its error
            run[j] = 0;
        }
    }
    HIPCHK(hipMemcpyAsync(hd + E.y(), yh, (m + 1) * Bp * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(E.d_cnt.p, cnt, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(E.d_rows.p, rows, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(E.d_sel.p, sel, Bp * sizeof(int), hipMemcpyHostToDevice, s));
    // w = V y on the lanes that took a step
    P.run(K_OTHER, nb * 8.0 * (m + 2), [&] {
        bk::clear_tails<double>(W, 1, n * Bp, n, Bp, dp<int>(E.d_rows), dp<int>(E.d_sel), s);
        bk::combine<double>(W, E.dv(0), dp<int>(E.d_cnt), hd + E.y(), n, Bp, s);
    });
    lane_norms(E, run);
    bool left = false;
    for (i64 j = 0; j < Bp; ++j) {
        if (!run[j]) continue;
        auto &l = E.L[j];
        if (!l.c.accept(std::sqrt(E.host[E.beta2() + j])) || !l.c.running()) l.done = true;
        left = left || !l.done;
    }
    return left;
}

// the result of a finished lane (n values) and its statistics; false while it runs. A lane whose
// step control failed throws its error.
bool expv_engine_take(expv_batch &E, i64 j, double *w, mk_expv_result *res) {
    require(j >= 0 && j < E.slots, "expv batch: bad slot");
    auto &l = E.L[j];
    if (!l.active || !l.done) return false;
    if (!l.error.empty()) throw std::runtime_error(l.error);
    const hipStream_t s = stream();
    bk::get_lane(dp<double>(E.W), E.n, E.Bp, j, dp<double>(E.stage), s);
    HIPCHK(hipMemcpyAsync(w, E.stage.p, E.n * sizeof(double), hipMemcpyDeviceToHost, s));
    drain();
    if (res) *res = l.c.res;
    return true;
}

void expv_engine_release(expv_batch &E, i64 j) {
    if (j >= 0 && j < E.slots) E.L[j] = expv_batch::lane{};
}

} // namespace mk::rt

extern "C" {

mk_status mk_expv_device(mk_generator gen, double t, const double *d_v, double *d_w,
                         const mk_expv_options *options, mk_expv_result *result) {
    return guarded([&] {
        require(gen && d_v && d_w && options && result, "expv: bad arguments");
        *result = expv_device(*gen, t, d_v, d_w, *options);
        drain();
    });
}

mk_status mk_expv_batch(mk_generator gen, int64_t count, const double *t, const double *v, double *w,
                        const mk_expv_options *options, mk_expv_result *results) {
    return guarded([&] {
        require(gen && options && count >= 0 && (count == 0 || (t && v && w && results)),
                "expv batch: bad arguments");
        require(options->krylov_dim >= 1 && options->tolerance > 0.0, "expv: bad options");
        if (count == 0) return;
        const i64 n = gen->n, slots = std::min<i64>(count, 256);
        std::unique_ptr<mk_expv_batch_s, void (*)(mk_expv_batch_s *)> E(
            mk::rt::expv_engine_create(gen, slots, options->krylov_dim), mk::rt::expv_engine_destroy);
        std::vector<i64> in_slot(slots, -1);
        i64 next = 0, finished = 0;
        while (finished < count) {
            for (i64 j = 0; j < slots && next < count; ++j) {
                if (in_slot[j] >= 0) continue;
                mk::rt::expv_engine_admit(*E, j, v + next * n, t[next], options->tolerance);
                in_slot[j] = next++;
            }
            mk::rt::expv_engine_step(*E);
            for (i64 j = 0; j < slots; ++j) {
                const i64 q = in_slot[j];
                if (q < 0 || !mk::rt::expv_engine_take(*E, j, w + q * n, results + q)) continue;
                mk::rt::expv_engine_release(*E, j);
                in_slot[j] = -1;
                ++finished;
            }
        }
    });
}

mk_status mk_expv(mk_generator gen, double t, const double *v, double *w,
                  const mk_expv_options *options, mk_expv_result *result) {
    return guarded([&] {
        require(gen && v && w && options && result, "expv: bad arguments");
        const std::size_t bytes = static_cast<std::size_t>(std::max<i64>(gen->n, 1)) * sizeof(double);
        dbuf dv(bytes), dw(bytes);
        HIPCHK(hipMemcpy(dv.p, v, gen->n * sizeof(double), hipMemcpyHostToDevice));
        *result = expv_device(*gen, t, dp<double>(dv), dp<double>(dw), *options);
        drain();
        HIPCHK(hipMemcpy(w, dw.p, gen->n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

} // extern "C"
