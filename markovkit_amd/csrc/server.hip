// Request serving: resolvent solves and Krylov exponentials. No reference counterpart (the
// reference is a library called one solve at a time).
//
// Continuous batching: a resolvent family of at most batch_rows states keeps a batch engine
// resident (batch.hip: one slot per system, the generator and one hierarchy shared, the Krylov
// bases of all slots in HBM) driven by its own thread and HIP stream. At every restart boundary
// the engine retires the finished systems and admits queued requests of its family into the
// free slots, so requests that arrive while others run join them on the device as slots free
// up. A request may be the leading block of whole sectors of its family's box (for an absorbing
// one-species box, the box of a smaller cap), so the caps of one grid share an engine.
//
// Exponentials over a generator of at most batch_rows states share a batched expv engine
// (expv.hip) per generator and Krylov dimension, run by its own thread the same way: vectors
// join at substep boundaries.
//
// Larger resolvent families and exponentials run on worker threads, each with its own HIP
// stream and Krylov workspaces, so their solves overlap on the device too. Such a request is
// admitted first come first served once a worker is free, its footprint fits the HBM budget
// beside the running requests and, for a resolvent, a multigrid hierarchy of its family is
// free; the hierarchy is re-targeted to the request's shift (terminal inverses only).
//
// Results do not depend on what else runs: every slot and every request computes with
// fixed-order reductions of its own, so a request's result is the same alone and in any mix.
#include "runtime.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

using namespace mk::rt;

struct mk_expv_batch_s;

namespace mk::rt {
// the continuous-batching engine of batch.hip
mk_batch_s *batch_engine_create(mk_generator_s *gen, mk_sector_mg_s *mg, i64 slots, i64 restart);
void batch_engine_destroy(mk_batch_s *B);
i64 batch_engine_slots(const mk_batch_s &B);
void batch_engine_admit(mk_batch_s &B, i64 j, const double *b, i64 rows, double shift, double tol,
                        i64 max_iterations);
bool batch_engine_step(mk_batch_s &B);
bool batch_engine_take(mk_batch_s &B, i64 j, double *x, mk_solver_result *res);
void batch_engine_release(mk_batch_s &B, i64 j);
// the batched exponentials of expv.hip
mk_expv_batch_s *expv_engine_create(mk_generator_s *gen, i64 slots, i64 krylov_dim);
void expv_engine_destroy(mk_expv_batch_s *E);
i64 expv_engine_slots(const mk_expv_batch_s &E);
void expv_engine_admit(mk_expv_batch_s &E, i64 j, const double *v, double t, double tol);
bool expv_engine_step(mk_expv_batch_s &E);
bool expv_engine_take(mk_expv_batch_s &E, i64 j, double *w, mk_expv_result *res);
void expv_engine_release(mk_expv_batch_s &E, i64 j);
} // namespace mk::rt

namespace {

using clock_type = std::chrono::steady_clock;

struct request {
    i64 id = 0;
    int kind = MK_REQUEST_RESOLVENT;
    int family = -1;      // resolvent: the family
    int expv_family = -1; // exponential: the batch engine of its generator (-1: a worker)
    i64 rows = 0; // resolvent: the leading rows of the family's box the system has
    mk_generator gen = nullptr;
    double shift = 0.0, t = 0.0, tolerance = 0.0;
    mk_gmres_options gopt{};
    mk_expv_options eopt{};
    std::vector<double> in, out;
    i64 footprint = 0;
    bool done = false;
    mk_status status = MK_OK;
    std::string error;
    mk_solver_result solve{};
    mk_expv_result expv{};
    clock_type::time_point submitted, started, finished;
};

struct family {
    mk_generator gen = nullptr;
    std::vector<mk_sector_mg> free;
    i64 n = 0;
    std::size_t zs = 8;
    // continuous batching: the engine, its queue and its thread (null engine: worker path)
    mk_batch_s *engine = nullptr;
    std::deque<request *> queue;
    std::thread thread;
};

// the exponentials of one generator and Krylov dimension, batched: engine, queue, thread
struct expv_family {
    mk_generator gen = nullptr;
    i64 m = 0;
    mk_expv_batch_s *engine = nullptr;
    std::deque<request *> queue;
    std::thread thread;
};

// device footprint of a request beyond the resident operators: the Krylov basis (m + 1
// vectors), Z (m vectors in the preconditioner's precision) or the expv basis, and the work
// vectors
i64 resolvent_footprint(i64 n, i64 m, std::size_t zs) {
    const i64 ld = (n + 63) / 64 * 64;
    return ld * (8 * (m + 1) + static_cast<i64>(zs) * m + 8 * 4);
}
i64 expv_footprint(i64 n, i64 m) {
    const i64 ld = (n + 63) / 64 * 64;
    return ld * (8 * (m + 1) + 8 * 4);
}

// device in/out buffers of one worker thread
struct worker_buffers {
    dbuf a, b;
    void ensure(i64 n) {
        const std::size_t bytes = static_cast<std::size_t>(std::max<i64>(n, 1)) * sizeof(double);
        if (a.bytes < bytes) a = dbuf(bytes);
        if (b.bytes < bytes) b = dbuf(bytes);
    }
};

} // namespace

struct mk_server_s {
    i64 budget = 0;
    i64 batch_rows = i64{1} << 19; // families of at most this many states are batched
    i64 batch_slots = 64;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<request *> queue;
    std::map<i64, std::unique_ptr<request>> requests;
    // a deque keeps every family at its address while mk_server_add_family appends (workers hold
    // no reference into it outside the lock either: run() gets the generator handle by value)
    std::deque<family> families;
    std::deque<expv_family> expv_families;
    i64 reserved = 0;
    i64 next_id = 1;
    bool stopping = false;
    std::vector<std::thread> threads;

    bool admissible(const request &r) const {
        if (reserved + r.footprint > budget && reserved > 0) return false;
        if (r.kind == MK_REQUEST_RESOLVENT && families[r.family].free.empty()) return false;
        return true;
    }

    void run(request &r, worker_buffers &buf, mk_generator gen, mk_sector_mg mg) {
        const hipStream_t s = stream();
        const i64 n = static_cast<i64>(r.in.size());
        buf.ensure(n);
        double *da = dp<double>(buf.a), *db = dp<double>(buf.b);
        HIPCHK(hipMemcpyAsync(da, r.in.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        mk_status st;
        if (r.kind == MK_REQUEST_RESOLVENT) {
            HIPCHK(hipMemsetAsync(db, 0, n * sizeof(double), s));
            st = mk_sector_mg_set_shift(mg, r.shift);
            if (st == MK_OK) {
                mk_gmres_options o = r.gopt;
                o.tolerance = r.tolerance;
                st = mk_gmres_device(gen, r.shift, mg, da, db, &o, &r.solve);
            }
        } else {
            st = mk_expv_device(gen, r.t, da, db, &r.eopt, &r.expv);
        }
        if (st != MK_OK) throw std::runtime_error(mk_last_error());
        r.out.resize(n);
        HIPCHK(hipMemcpyAsync(r.out.data(), db, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }

    // the engine thread of a batched family: admit queued requests into free slots, run one
    // restart cycle of every active slot, hand out the finished ones; repeat
    void engine_loop(family &f) {
        mk_batch_s &B = *f.engine;
        const i64 slots = batch_engine_slots(B);
        std::vector<request *> in_slot(slots, nullptr);
        i64 busy = 0;
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            cv.wait(lock, [&] { return stopping || !f.queue.empty() || busy > 0; });
            if (stopping && f.queue.empty() && busy == 0) return;
            std::vector<std::pair<i64, request *>> admit;
            for (i64 j = 0; j < slots && !f.queue.empty(); ++j) {
                if (in_slot[j]) continue;
                request *r = f.queue.front();
                f.queue.pop_front();
                r->started = clock_type::now();
                in_slot[j] = r;
                ++busy;
                admit.emplace_back(j, r);
            }
            lock.unlock();
            std::vector<std::pair<request *, std::string>> failed;
            for (auto &[j, r] : admit) {
                try {
                    batch_engine_admit(B, j, r->in.data(), r->rows, r->shift, r->tolerance,
                                       r->gopt.max_iterations);
                } catch (const std::exception &e) {
                    failed.emplace_back(r, e.what());
                    in_slot[j] = nullptr;
                    batch_engine_release(B, j);
                }
            }
            std::vector<request *> finished;
            std::string step_error;
            try {
                batch_engine_step(B);
                for (i64 j = 0; j < slots; ++j) {
                    request *r = in_slot[j];
                    if (!r) continue;
                    r->out.resize(r->rows);
                    if (batch_engine_take(B, j, r->out.data(), &r->solve)) {
                        finished.push_back(r);
                        in_slot[j] = nullptr;
                    }
                }
            } catch (const std::exception &e) {
                // a device failure ends every request of the engine
                step_error = e.what();
                for (i64 j = 0; j < slots; ++j)
                    if (in_slot[j]) {
                        failed.emplace_back(in_slot[j], step_error);
                        in_slot[j] = nullptr;
                        batch_engine_release(B, j);
                    }
            }
            lock.lock();
            for (request *r : finished) {
                r->finished = clock_type::now();
                r->status = MK_OK;
                r->done = true;
                --busy;
            }
            for (auto &[r, e] : failed) {
                r->finished = clock_type::now();
                r->status = MK_ERR_RUNTIME;
                r->error = e;
                r->done = true;
                --busy;
            }
            cv.notify_all();
        }
    }

    // the engine thread of batched exponentials: admit queued vectors into free lanes, run one
    // substep of every running lane, hand out the finished ones; repeat
    void expv_loop(expv_family &f) {
        mk_expv_batch_s &E = *f.engine;
        const i64 slots = expv_engine_slots(E);
        std::vector<request *> in_slot(slots, nullptr);
        i64 busy = 0;
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            cv.wait(lock, [&] { return stopping || !f.queue.empty() || busy > 0; });
            if (stopping && f.queue.empty() && busy == 0) return;
            std::vector<std::pair<i64, request *>> admit;
            for (i64 j = 0; j < slots && !f.queue.empty(); ++j) {
                if (in_slot[j]) continue;
                request *r = f.queue.front();
                f.queue.pop_front();
                r->started = clock_type::now();
                in_slot[j] = r;
                ++busy;
                admit.emplace_back(j, r);
            }
            lock.unlock();
            std::vector<std::pair<request *, std::string>> failed;
            for (auto &[j, r] : admit) {
                try {
                    expv_engine_admit(E, j, r->in.data(), r->t, r->eopt.tolerance);
                } catch (const std::exception &e) {
                    failed.emplace_back(r, e.what());
                    in_slot[j] = nullptr;
                    expv_engine_release(E, j);
                }
            }
            std::vector<request *> finished;
            try {
                expv_engine_step(E);
                for (i64 j = 0; j < slots; ++j) {
                    request *r = in_slot[j];
                    if (!r) continue;
                    r->out.resize(r->in.size());
                    bool done = false;
                    try {
                        done = expv_engine_take(E, j, r->out.data(), &r->expv);
                        if (done) finished.push_back(r);
                    } catch (const std::exception &e) {
                        // this lane's step control failed; the others go on
                        failed.emplace_back(r, e.what());
                        done = true;
                    }
                    if (done) {
                        in_slot[j] = nullptr;
                        expv_engine_release(E, j);
                    }
                }
            } catch (const std::exception &e) {
                // a device failure ends every request of the engine
                for (i64 j = 0; j < slots; ++j)
                    if (in_slot[j]) {
                        failed.emplace_back(in_slot[j], e.what());
                        in_slot[j] = nullptr;
                        expv_engine_release(E, j);
                    }
            }
            lock.lock();
            for (request *r : finished) {
                r->finished = clock_type::now();
                r->status = MK_OK;
                r->done = true;
                --busy;
            }
            for (auto &[r, e] : failed) {
                r->finished = clock_type::now();
                r->status = MK_ERR_RUNTIME;
                r->error = e;
                r->done = true;
                --busy;
            }
            cv.notify_all();
        }
    }

    void worker() {
        worker_buffers buf;
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            // a request that is not admissible now becomes so when a running one finishes
            cv.wait(lock, [&] {
                return (stopping && queue.empty()) || (!queue.empty() && admissible(*queue.front()));
            });
            if (queue.empty()) return;
            request *r = queue.front();
            queue.pop_front();
            reserved += r->footprint;
            mk_sector_mg mg = nullptr;
            mk_generator gen = r->gen;
            if (r->kind == MK_REQUEST_RESOLVENT) {
                family &f = families[r->family];
                mg = f.free.back();
                f.free.pop_back();
                gen = f.gen;
            }
            r->started = clock_type::now();
            lock.unlock();
            mk_status status = MK_OK;
            std::string error;
            try {
                run(*r, buf, gen, mg);
            } catch (const std::exception &e) {
                status = MK_ERR_RUNTIME;
                error = e.what();
            }
            lock.lock();
            r->finished = clock_type::now();
            r->status = status;
            r->error = error;
            r->done = true;
            reserved -= r->footprint;
            if (mg) families[r->family].free.push_back(mg);
            cv.notify_all();
        }
    }
};

namespace {

double seconds(clock_type::duration d) { return std::chrono::duration<double>(d).count(); }

i64 submit(mk_server s, std::unique_ptr<request> r) {
    std::lock_guard<std::mutex> lock(s->mu);
    require(!s->stopping, "server: stopping");
    r->id = s->next_id++;
    r->submitted = clock_type::now();
    const i64 id = r->id;
    if (r->kind == MK_REQUEST_RESOLVENT && s->families[r->family].engine) {
        s->families[r->family].queue.push_back(r.get());
    } else if (r->kind == MK_REQUEST_EXPV && r->expv_family >= 0) {
        s->expv_families[r->expv_family].queue.push_back(r.get());
    } else {
        s->queue.push_back(r.get());
    }
    s->requests[id] = std::move(r);
    s->cv.notify_all();
    return id;
}

} // namespace

extern "C" {

mk_status mk_server_create(int32_t workers, int64_t hbm_budget_bytes, mk_server *out) {
    return guarded([&] {
        require(out && workers >= 1 && workers <= 64 && hbm_budget_bytes > 0, "server: bad arguments");
        auto s = std::make_unique<mk_server_s>();
        s->budget = hbm_budget_bytes;
        for (int w = 0; w < workers; ++w) s->threads.emplace_back([p = s.get()] { p->worker(); });
        *out = s.release();
    });
}

mk_status mk_server_destroy(mk_server s) {
    if (!s) return MK_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    for (auto &t : s->threads) t.join();
    for (auto &f : s->families) {
        if (f.thread.joinable()) f.thread.join();
        if (f.engine) batch_engine_destroy(f.engine);
    }
    for (auto &f : s->expv_families) {
        if (f.thread.joinable()) f.thread.join();
        if (f.engine) expv_engine_destroy(f.engine);
    }
    delete s;
    return MK_OK;
}

mk_status mk_server_add_family(mk_server s, mk_generator gen, int32_t n_mg, const mk_sector_mg *mg,
                               int32_t *family_out) {
    return guarded([&] {
        require(s && gen && mg && family_out && n_mg >= 1, "server: bad family");
        for (int k = 0; k < n_mg; ++k) {
            require(mg[k] && mg[k]->rows == gen->n, "server: a hierarchy does not match the generator");
            require(mg[k]->opt.smoother == MK_JACOBI || mg[k]->opt.smoother == MK_RED_BLACK,
                    "server: hierarchies must use the Jacobi or red-black smoother (re-targeted per shift)");
        }
        std::lock_guard<std::mutex> lock(s->mu);
        require(!s->stopping, "server: stopping");
        family &f = s->families.emplace_back();
        f.gen = gen;
        f.n = gen->n;
        for (int k = 0; k < n_mg; ++k) {
            f.zs = mg[k]->precision == MK_FP32 ? sizeof(float) : sizeof(double);
            f.free.push_back(mg[k]);
        }
        const bool rb_ok = mg[0]->opt.smoother != MK_RED_BLACK ||
                           (mg[0]->opt.pre_sweeps >= 1 && mg[0]->opt.post_sweeps >= 1);
        if (gen->n <= s->batch_rows && s->batch_slots > 0 && rb_ok) {
            try {
                f.engine = batch_engine_create(gen, mg[0], s->batch_slots, 50);
            } catch (...) {
                s->families.pop_back();
                throw;
            }
            f.thread = std::thread([s, &f] { s->engine_loop(f); });
        }
        *family_out = static_cast<int32_t>(s->families.size() - 1);
    });
}

mk_status mk_server_submit_resolvent_sized(mk_server s, int32_t family, double shift,
                                           const double *b, int64_t rows, double tolerance,
                                           const mk_gmres_options *options, int64_t *id) {
    return guarded([&] {
        require(s && b && options && id, "server: bad arguments");
        if (!(shift > 0.0)) throw std::invalid_argument("population_multigrid: shift must be positive");
        i64 n;
        std::size_t zs;
        bool batched;
        {
            std::lock_guard<std::mutex> lock(s->mu);
            require(family >= 0 && family < static_cast<int32_t>(s->families.size()), "server: unknown family");
            n = s->families[family].n;
            zs = s->families[family].zs;
            batched = s->families[family].engine != nullptr;
        }
        if (rows < 0) rows = n;
        require(rows >= 1 && rows <= n, "server: bad row count");
        require(rows == n || batched, "server: only batched families take systems of fewer rows");
        auto r = std::make_unique<request>();
        r->kind = MK_REQUEST_RESOLVENT;
        r->family = family;
        r->rows = rows;
        r->shift = shift;
        r->tolerance = tolerance;
        r->gopt = *options;
        r->in.assign(b, b + rows);
        r->footprint = batched ? 0 : resolvent_footprint(n, std::max<i64>(1, options->restart), zs);
        *id = submit(s, std::move(r));
    });
}

mk_status mk_server_submit_resolvent(mk_server s, int32_t family, double shift, const double *b,
                                     double tolerance, const mk_gmres_options *options,
                                     int64_t *id) {
    return mk_server_submit_resolvent_sized(s, family, shift, b, -1, tolerance, options, id);
}

mk_status mk_server_batching(mk_server s, int64_t max_rows, int32_t slots) {
    return guarded([&] {
        require(s && max_rows >= 0 && slots >= 0 && slots <= 4096, "server: bad batching arguments");
        std::lock_guard<std::mutex> lock(s->mu);
        s->batch_rows = max_rows;
        s->batch_slots = slots;
    });
}

mk_status mk_server_submit_expv(mk_server s, mk_generator gen, double t, const double *v,
                                const mk_expv_options *options, int64_t *id) {
    return guarded([&] {
        require(s && gen && v && options && id, "server: bad arguments");
        require(options->krylov_dim >= 1 && options->tolerance > 0.0, "expv: bad options");
        auto r = std::make_unique<request>();
        r->kind = MK_REQUEST_EXPV;
        r->gen = gen;
        r->t = t;
        r->eopt = *options;
        r->in.assign(v, v + gen->n);
        r->footprint = expv_footprint(gen->n, std::max<i64>(1, options->krylov_dim));
        {
            // continuous batching: the exponentials of a generator of at most the batching row
            // limit share a batch engine per Krylov dimension, made at its first request
            std::lock_guard<std::mutex> lock(s->mu);
            require(!s->stopping, "server: stopping");
            const i64 m = std::max<i64>(1, std::min<i64>(options->krylov_dim, gen->n));
            if (gen->n <= s->batch_rows && s->batch_slots > 0) {
                int found = -1;
                for (std::size_t k = 0; k < s->expv_families.size(); ++k)
                    if (s->expv_families[k].gen == gen && s->expv_families[k].m == m) found = static_cast<int>(k);
                if (found < 0) {
                    expv_family &f = s->expv_families.emplace_back();
                    f.gen = gen;
                    f.m = m;
                    try {
                        f.engine = expv_engine_create(gen, s->batch_slots, m);
                    } catch (...) {
                        s->expv_families.pop_back();
                        throw;
                    }
                    f.thread = std::thread([s, &f] { s->expv_loop(f); });
                    found = static_cast<int>(s->expv_families.size() - 1);
                }
                r->expv_family = found;
                r->footprint = 0;
            }
        }
        *id = submit(s, std::move(r));
    });
}

mk_status mk_server_wait(mk_server s, int64_t id, double *out, mk_server_result *result) {
    return guarded([&] {
        require(s && result, "server: bad arguments");
        std::unique_lock<std::mutex> lock(s->mu);
        auto it = s->requests.find(id);
        require(it != s->requests.end(), "server: unknown request");
        request &r = *it->second;
        s->cv.wait(lock, [&] { return r.done; });
        result->kind = r.kind;
        result->status = r.status;
        result->solve = r.solve;
        result->expv = r.expv;
        result->queued_s = seconds(r.started - r.submitted);
        result->run_s = seconds(r.finished - r.started);
        if (out && r.status == MK_OK) std::copy(r.out.begin(), r.out.end(), out);
        const mk_status st = r.status;
        const std::string err = r.error;
        s->requests.erase(it);
        if (st != MK_OK) throw std::runtime_error("server: request failed: " + err);
    });
}

} // extern "C"
