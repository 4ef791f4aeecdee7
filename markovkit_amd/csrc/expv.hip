// Krylov exponential w = exp(t Q) v of a generator on the device: the num::expv the reference's
// FSP solvers call (include/markovkit/fsp.hpp:152, :225; numerics source absent), as Sidje's
// expv (Expokit, ACM TOMS 24, 1998). Each step builds an Arnoldi basis of dimension m for
// B = 0 I - Q = -Q with classical Gram-Schmidt twice, whose coefficients and norms stay on the
// device (no host synchronisation inside a step); the host then reads the Hessenberg matrix,
// exponentiates the (m + 2) x (m + 2) augmented matrix (degree-6 Pade, scaling and squaring),
// estimates the local error from the two corrected terms and accepts or shrinks the step
// (gamma 0.9, delta 1.2, at most 10 rejections); w = V (beta F e_1) on the device.
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

mk_expv_result expv_device(mk_generator_s &G, double t, const double *v, double *w_out,
                           const mk_expv_options &o) {
    require(o.krylov_dim >= 1 && o.tolerance > 0.0, "expv: bad options");
    mk_expv_result res{0, 0, 0.0};
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
    const double tol = o.tolerance, btol = 1e-7, gamma = 0.9, delta = 1.2;
    const int max_reject = 10;
    const double t_out = std::abs(t), sgn = t >= 0.0 ? -1.0 : 1.0; // exp(t Q) = exp(-t B)
    const double anorm = G.norm1 > 0.0 ? G.norm1 : 1.0;
    const double rndoff = anorm * std::numeric_limits<double>::epsilon();

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
    auto finish = [&] {
        HIPCHK(hipMemcpyAsync(w_out, w, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        return res;
    };
    norm2_to(w, hd + ws.beta2());
    read(ws.beta2(), 1);
    double beta = std::sqrt(ws.host[ws.beta2()]);
    if (!(beta > 0.0) || t_out == 0.0) return finish();
    double xm = 1.0 / double(m);
    const double fact = std::pow(double(m + 1) / std::exp(1.0), double(m + 1)) *
                        std::sqrt(2.0 * 3.141592653589793 * double(m + 1));
    double t_new = round_step((1.0 / anorm) * std::pow((fact * tol) / (4.0 * beta * anorm), xm));
    double t_now = 0.0;
    const i64 L = m + 2;
    std::vector<double> H;
    while (t_now < t_out) {
        ++res.steps;
        double t_step = std::min(t_out - t_now, t_new);
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
        H.assign(static_cast<std::size_t>(L * L), 0.0);
        int k1 = 2;
        i64 mb = m;
        for (i64 j = 0; j < m; ++j) {
            for (i64 i = 0; i <= j; ++i) H[i * L + j] = h1[j * (m + 1) + i] + h2[j * (m + 1) + i];
            const double sj = std::sqrt(nr[j]);
            if (!(sj >= btol)) {
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
            avnorm = std::sqrt(nr[m]);
        }
        std::vector<double> F;
        i64 mx = mb + k1;
        double err_loc = btol;
        for (int rejected = 0;; ++rejected) {
            mx = mb + k1;
            std::vector<double> X(static_cast<std::size_t>(mx * mx));
            for (i64 i = 0; i < mx; ++i)
                for (i64 j = 0; j < mx; ++j) X[i * mx + j] = sgn * t_step * H[i * L + j];
            F = mk::expm_pade(std::move(X), mx);
            if (k1 == 0) {
                err_loc = btol;
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
            if (err_loc <= delta * t_step * tol) break;
            if (rejected == max_reject) throw std::runtime_error("expv: the requested tolerance is too high");
            t_step = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
            ++res.rejected;
        }
        const i64 use = mb + std::max(0, k1 - 1);
        double *yh = ws.host + ws.y();
        for (i64 k = 0; k < use; ++k) yh[k] = beta * F[k * mx];
        HIPCHK(hipMemcpyAsync(hd + ws.y(), yh, use * sizeof(double), hipMemcpyHostToDevice, s));
        P.run(K_OTHER, double(use) * n * 8.0 + n * 16.0, [&] {
            gpu::fill<double>(n, 0.0, w, s);
            gpu::combine<double>(w, V, ld, static_cast<int>(use), hd + ws.y(), n, s);
        });
        norm2_to(w, hd + ws.beta2());
        read(ws.beta2(), 1);
        beta = std::sqrt(ws.host[ws.beta2()]);
        t_now += t_step;
        t_new = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
        res.error += std::max(err_loc, rndoff);
        if (!(beta > 0.0) || !std::isfinite(beta)) break;
    }
    return finish();
}

} // namespace

extern "C" {

mk_status mk_expv_device(mk_generator gen, double t, const double *d_v, double *d_w,
                         const mk_expv_options *options, mk_expv_result *result) {
    return guarded([&] {
        require(gen && d_v && d_w && options && result, "expv: bad arguments");
        *result = expv_device(*gen, t, d_v, d_w, *options);
        drain();
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
