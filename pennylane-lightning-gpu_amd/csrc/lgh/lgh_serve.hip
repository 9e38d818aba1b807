// Serving engine: a request queue in front of the device.
//
// Requests are whole circuits (gates with bound parameters, one Pauli-word
// observable, optionally the gradient with respect to every parametrised
// gate).  Worker threads, each with its own HIP stream, take work from one
// queue:
//   - continuous batching: a worker that finds circuits of at most
//     `batch_max_qubits` qubits waiting takes all of them (up to a limit) into
//     one persistent batched launch (lgh_batch_run).  Several batches run at
//     once on different workers' streams (by default all workers but one), so
//     small circuits that arrive while a batch is on the device start on the
//     next idle worker instead of waiting for that batch.  Gate structures
//     are compiled once into templates, cached by structure, so a stream of
//     requests from a few circuit families costs no host compilation;
//   - larger circuits run one per worker on device states of their own
//     (lgh_apply, Pauli-word expectation, lgh_adjoint_jacobian), their fused
//     passes spread over the whole device, overlapping with the batches and
//     with each other.
// Memory admission: every piece of work reserves the device bytes it will
// hold (states, plus the adjoint's working copy) against a budget before it
// starts; work that does not fit waits, and later work that fits may pass it.
// Per request the engine records queueing delay and run time.
//
// The reference has no serving layer: each PennyLane execution calls the
// device synchronously (pennylane_lightning_gpu/lightning_gpu.py:519-555,
// :638-752); one request here is one such execution.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>

#include "lgh_internal.hpp"

using namespace lgh;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Req {
    uint64_t ticket = 0;
    int n = 0;
    std::vector<lgh_op> ops;
    std::vector<std::vector<int32_t>> wires;
    std::vector<std::vector<double>> mats;
    std::vector<int32_t> obs_paulis, obs_wires;
    bool grad = false;
    int64_t num_par = 0; // parametrised gates
    uint64_t bytes = 0;  // device bytes reserved while running alone
    bool batchable = false;
    std::string key;     // gate structure (batched path)
    Clock::time_point t_submit, t_start, t_end;
    // results
    double expval = 0.0;
    std::vector<double> grad_out;
    lgh_status status = LGH_OK;
    std::string error;
    bool done = false;
};

bool template_gate(int g) {
    switch (g) {
    case LGH_GATE_RX: case LGH_GATE_RY: case LGH_GATE_RZ: case LGH_GATE_PHASESHIFT: case LGH_GATE_CRX:
    case LGH_GATE_CRY: case LGH_GATE_CRZ: case LGH_GATE_CONTROLLEDPHASESHIFT: case LGH_GATE_ISINGZZ:
    case LGH_GATE_MULTIRZ: case LGH_GATE_PAULIX: case LGH_GATE_PAULIY: case LGH_GATE_PAULIZ:
    case LGH_GATE_HADAMARD: case LGH_GATE_S: case LGH_GATE_T: case LGH_GATE_CNOT: case LGH_GATE_CY:
    case LGH_GATE_CZ: case LGH_GATE_TOFFOLI: case LGH_GATE_IDENTITY:
        return true;
    default:
        return false;
    }
}

// Batches hold circuits of one width class: a persistent launch runs each
// circuit on one workgroup and lasts as long as its largest circuit, so a
// batch of 10-qubit circuits should not wait for 20-qubit ones, which get
// batches of their own (on another worker, at the same time).
int width_class(int n) { return n <= 12 ? 0 : n <= 16 ? 1 : 2; }

bool parametric(int g) {
    switch (g) {
    case LGH_GATE_RX: case LGH_GATE_RY: case LGH_GATE_RZ: case LGH_GATE_PHASESHIFT: case LGH_GATE_CRX:
    case LGH_GATE_CRY: case LGH_GATE_CRZ: case LGH_GATE_CONTROLLEDPHASESHIFT: case LGH_GATE_ISINGZZ:
    case LGH_GATE_MULTIRZ:
        return true;
    default:
        return false;
    }
}

void check_ok(lgh_status s) {
    if (s != LGH_OK) throw Error(s, lgh_last_error());
}

} // namespace

struct lgh_server {
    int device = 0;
    uint64_t budget = 0;
    int batch_max_qubits = 20;
    int max_batch = 4096;
    int max_batches = 1; // batches that may run at once
    int max_singles = 1; // large circuits that may run at once (one worker stays free for batches)
    int running_singles = 0;
    std::vector<void *> streams;
    std::vector<std::thread> workers;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_ticket = 1;
    uint64_t reserved = 0;
    int running_batches = 0;
    // Batch formation under load: while other work runs, a batch smaller
    // than min_batch waits up to batch_wait for more circuits of its class
    // (fewer, larger launches); an idle engine launches at once.
    size_t min_batch = 64;
    std::chrono::microseconds batch_wait{1000};
    bool deferred = false;
    Clock::time_point defer_until{};
    uint64_t inject_failures = 0; // the next pieces of work fail before launching (tests)
    std::deque<Req *> small, large;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    // template cache by gate structure
    std::mutex tmpl_mu;
    std::unordered_map<std::string, lgh_tmpl *> tmpls;
    // counters
    uint64_t n_batches = 0, n_batched = 0, n_single = 0;

    ~lgh_server() {
        for (auto &kv : tmpls) lgh_tmpl_destroy(kv.second);
    }

    lgh_tmpl *template_for(const Req &r) {
        std::lock_guard<std::mutex> lk(tmpl_mu);
        auto it = tmpls.find(r.key);
        if (it != tmpls.end()) return it->second;
        std::vector<int32_t> pidx(r.ops.size(), -1);
        int32_t k = 0;
        for (size_t i = 0; i < r.ops.size(); ++i)
            if (r.ops[i].num_params > 0) pidx[i] = k++;
        lgh_tmpl *t = nullptr;
        check_ok(lgh_tmpl_create(&t, r.n, r.ops.data(), int32_t(r.ops.size()), pidx.data(), r.obs_paulis.data(),
                                 r.obs_wires.data(), int32_t(r.obs_paulis.size()), device));
        tmpls.emplace(r.key, t);
        return t;
    }

    void run_batch(std::vector<Req *> &batch, void *stream) {
        std::vector<lgh_tmpl *> ts;
        std::vector<double> params;
        std::vector<int64_t> poff, joff;
        int64_t nj = 0;
        for (Req *r : batch) {
            ts.push_back(template_for(*r));
            poff.push_back(int64_t(params.size()));
            for (const lgh_op &o : r->ops)
                if (o.num_params > 0) params.push_back(o.params[0]);
            joff.push_back(nj);
            nj += r->num_par;
        }
        bool any_grad = false;
        for (Req *r : batch) any_grad = any_grad || r->grad;
        std::vector<double> ev(batch.size()), jac(size_t(std::max<int64_t>(1, nj)));
        check_ok(lgh_batch_run(ts.data(), int32_t(batch.size()), LGH_C128, params.empty() ? nullptr : params.data(),
                               poff.data(), ev.data(), any_grad ? jac.data() : nullptr, joff.data(), stream));
        for (size_t i = 0; i < batch.size(); ++i) {
            Req *r = batch[i];
            r->expval = ev[i];
            if (r->grad) r->grad_out.assign(jac.begin() + joff[i], jac.begin() + joff[i] + r->num_par);
        }
    }

    void run_single(Req *r, void *stream) {
        lgh_sv *psi = nullptr, *bra = nullptr;
        struct Del {
            lgh_sv *&a, *&b;
            ~Del() {
                if (a) lgh_sv_destroy(a);
                if (b) lgh_sv_destroy(b);
            }
        } del{psi, bra};
        check_ok(lgh_sv_create(&psi, r->n, LGH_C128, device, stream));
        check_ok(lgh_sv_reset(psi));
        check_ok(lgh_apply(psi, r->ops.data(), int64_t(r->ops.size())));
        std::string word(static_cast<size_t>(r->n), 'I');
        const char *names = "IXYZ";
        for (size_t k = 0; k < r->obs_paulis.size(); ++k)
            word[size_t(r->obs_wires[k])] = names[r->obs_paulis[k] - LGH_GATE_PAULIX + 1];
        std::vector<int32_t> wl(static_cast<size_t>(r->n));
        std::iota(wl.begin(), wl.end(), 0);
        const int64_t offs[2] = {0, r->n};
        const double coef[2] = {1.0, 0.0};
        check_ok(lgh_expval_pauli_words(psi, 1, word.data(), wl.data(), offs, coef, nullptr, &r->expval));
        if (!r->grad) return;
        check_ok(lgh_sv_create(&bra, r->n, LGH_C128, device, stream));
        check_ok(lgh_sv_copy_d2d(bra, psi));
        std::vector<lgh_op> pops;
        for (size_t k = 0; k < r->obs_paulis.size(); ++k) {
            lgh_op o{};
            o.gate = r->obs_paulis[k];
            o.num_wires = 1;
            o.wires = &r->obs_wires[k];
            pops.push_back(o);
        }
        if (!pops.empty()) check_ok(lgh_apply(bra, pops.data(), int64_t(pops.size())));
        std::vector<int64_t> tp(size_t(r->num_par));
        std::iota(tp.begin(), tp.end(), 0);
        r->grad_out.assign(size_t(r->num_par), 0.0);
        lgh_sv *bl[1] = {bra};
        check_ok(lgh_adjoint_jacobian(psi, bl, 1, r->ops.data(), int64_t(r->ops.size()), tp.data(), r->num_par, 0,
                                      r->grad_out.data()));
    }

    // Under mu: the next piece of work that fits the memory budget, or
    // nothing.  Waiting small circuits go first: those of the oldest one's
    // width class (up to max_batch) into one batch, while fewer than
    // max_batches batches run;
    // otherwise the oldest large circuit that fits, while fewer than
    // max_singles run (with several workers one always stays free for the
    // small circuits, whose latency would otherwise be set by the largest
    // circuits in flight).  Work is always admitted when nothing else holds
    // memory.
    bool pick(std::vector<Req *> &batch, Req *&single, uint64_t &bytes) {
        if (!small.empty() && running_batches < max_batches) {
            // the oldest waiting circuit's width class, its peers in arrival order
            const int cls = width_class(small.front()->n);
            uint64_t b = 0;
            std::deque<Req *> rest;
            for (Req *r : small) {
                const bool fits = batch.empty() ? (reserved + r->bytes <= budget || reserved == 0)
                                                : reserved + b + r->bytes <= budget;
                if (width_class(r->n) == cls && batch.size() < size_t(max_batch) && fits) {
                    batch.push_back(r);
                    b += r->bytes;
                } else {
                    rest.push_back(r);
                }
            }
            const bool busy = running_batches + running_singles > 0;
            const auto ready_at = small.front()->t_submit + batch_wait;
            if (!batch.empty() && batch.size() < min_batch && busy && Clock::now() < ready_at) {
                batch.clear(); // leave the queue as it is; look again when more arrive or at ready_at
                deferred = true;
                defer_until = ready_at;
            } else if (!batch.empty()) {
                small.swap(rest);
                bytes = b;
                ++running_batches;
                return true;
            }
        }
        if (running_singles >= max_singles) return false;
        for (auto it = large.begin(); it != large.end(); ++it)
            if (reserved + (*it)->bytes <= budget || reserved == 0) {
                single = *it;
                large.erase(it);
                bytes = single->bytes;
                ++running_singles;
                return true;
            }
        return false;
    }

    void worker(int idx) {
        void *stream = streams[size_t(idx)];
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            std::vector<Req *> batch;
            Req *single = nullptr;
            uint64_t bytes = 0;
            for (;;) {
                deferred = false;
                if (stop || pick(batch, single, bytes)) break;
                if (deferred)
                    cv_work.wait_until(lk, defer_until);
                else
                    cv_work.wait(lk);
            }
            if (stop) return;
            const auto now = Clock::now();
            if (single) single->t_start = now;
            for (Req *r : batch) r->t_start = now;
            reserved += bytes;
            const bool fail = inject_failures > 0;
            if (fail) --inject_failures;
            lk.unlock();
            lgh_status code = LGH_OK;
            std::string msg;
            try {
                if (fail)
                    throw Error(LGH_ERR_DEVICE, "injected failure (lgh_server_inject_failures)");
                if (single)
                    run_single(single, stream);
                else
                    run_batch(batch, stream);
            } catch (const Error &e) {
                msg = e.what();
                code = e.code;
            } catch (const std::exception &e) {
                msg = e.what();
                code = LGH_ERR_INVALID_ARGUMENT;
            }
            lk.lock();
            reserved -= bytes;
            const auto end = Clock::now();
            if (single) {
                batch.push_back(single);
                --running_singles;
                ++n_single;
            } else {
                --running_batches;
                ++n_batches;
                n_batched += batch.size();
            }
            for (Req *r : batch) {
                r->t_end = end;
                if (!msg.empty()) {
                    r->status = code;
                    r->error = msg;
                }
                r->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

lgh_status lgh_server_create(lgh_server **out, int32_t device, uint64_t memory_budget, int32_t num_streams,
                             int32_t batch_max_qubits) {
    return guard([&] {
        LGH_REQUIRE(out, "null pointer");
        *out = nullptr;
        LGH_REQUIRE(num_streams >= 1 && num_streams <= 16, "num_streams must be in [1, 16]");
        LGH_REQUIRE(batch_max_qubits >= 0 && batch_max_qubits <= 20, "batch_max_qubits must be in [0, 20]");
        auto s = std::make_unique<lgh_server>();
        s->device = device;
        uint64_t free_b = 0, total_b = 0;
        check_ok(lgh_device_memory(device, &free_b, &total_b));
        s->budget = memory_budget ? std::min<uint64_t>(memory_budget, free_b) : free_b / 10 * 9;
        s->batch_max_qubits = batch_max_qubits;
        s->max_batches = std::max(1, num_streams - 1);
        s->max_singles = std::max(1, num_streams - 1);
        for (int i = 0; i < num_streams; ++i) {
            void *st = nullptr;
            check_ok(lgh_stream_create(device, &st));
            s->streams.push_back(st);
        }
        lgh_server *raw = s.get();
        for (int i = 0; i < num_streams; ++i) s->workers.emplace_back([raw, i] { raw->worker(i); });
        *out = s.release();
    });
}

lgh_status lgh_server_destroy(lgh_server *srv) {
    return guard([&] {
        if (!srv) return;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv_work.notify_all();
        for (auto &t : srv->workers) t.join();
        for (void *st : srv->streams) lgh_stream_destroy(st);
        delete srv;
    });
}

lgh_status lgh_server_submit(lgh_server *srv, const lgh_request *req, uint64_t *ticket) {
    return guard([&] {
        LGH_REQUIRE(srv && req && ticket, "null pointer");
        LGH_REQUIRE(req->num_qubits >= 1 && req->num_qubits <= 36, "num_qubits must be in [1, 36]");
        LGH_REQUIRE(req->num_ops >= 0 && (req->num_ops == 0 || req->ops), "bad op list");
        LGH_REQUIRE(req->obs_len >= 0 && (req->obs_len == 0 || (req->obs_paulis && req->obs_wires)),
                    "bad observable");
        auto r = std::make_unique<Req>();
        r->n = req->num_qubits;
        r->grad = req->want_grad != 0;
        bool tmpl_ok = r->n <= srv->batch_max_qubits;
        r->key = std::to_string(r->n);
        for (int32_t i = 0; i < req->num_ops; ++i) {
            const lgh_op &o = req->ops[i];
            validate_op(o, r->n);
            LGH_REQUIRE(o.gate != LGH_GATE_STATEPREP, "state preparation is not part of a served circuit");
            r->ops.push_back(o);
            r->wires.emplace_back(o.wires, o.wires + o.num_wires);
            if (o.gate == LGH_GATE_MATRIX) {
                const size_t dim = size_t(1) << (o.num_wires - o.num_ctrl);
                r->mats.emplace_back(o.matrix, o.matrix + 2 * dim * dim);
            } else {
                r->mats.emplace_back();
            }
            if (o.num_params > 0) ++r->num_par;
            // templates take every parametric gate's angle from the request
            if (!template_gate(o.gate) || (o.num_params > 0) != parametric(o.gate)) tmpl_ok = false;
            r->key += ";" + std::to_string(o.gate) + (o.inverse ? "i" : "") + (o.num_params > 0 ? "p" : "");
            for (int32_t w = 0; w < o.num_wires; ++w) r->key += "," + std::to_string(o.wires[w]);
        }
        for (size_t i = 0; i < r->ops.size(); ++i) {
            r->ops[i].wires = r->wires[i].data();
            if (!r->mats[i].empty()) r->ops[i].matrix = r->mats[i].data();
        }
        r->key += "|";
        for (int32_t k = 0; k < req->obs_len; ++k) {
            const int32_t g = req->obs_paulis[k];
            LGH_REQUIRE(g == LGH_GATE_PAULIX || g == LGH_GATE_PAULIY || g == LGH_GATE_PAULIZ,
                        "observable factors are PauliX, PauliY or PauliZ");
            LGH_REQUIRE(req->obs_wires[k] >= 0 && req->obs_wires[k] < r->n, "observable wire out of range");
            for (int32_t j = 0; j < k; ++j)
                LGH_REQUIRE(req->obs_wires[j] != req->obs_wires[k], "repeated observable wire");
            r->obs_paulis.push_back(g);
            r->obs_wires.push_back(req->obs_wires[k]);
            r->key += std::to_string(g) + "," + std::to_string(req->obs_wires[k]) + ";";
        }
        const uint64_t state = uint64_t(16) << r->n;
        if (tmpl_ok) {
            // a batched circuit holds psi and bra on at least 9 qubits
            r->batchable = true;
            r->bytes = 2 * (uint64_t(16) << std::max(r->n, 9));
            // the template must accept the circuit (e.g. observable in one pass)
            try {
                srv->template_for(*r);
            } catch (const Error &) {
                r->batchable = false;
            }
        }
        if (!r->batchable) r->bytes = r->grad ? 3 * state : state;
        LGH_REQUIRE(r->bytes <= srv->budget, "request needs " + std::to_string(r->bytes) +
                                                 " device bytes, more than the server's budget");
        std::lock_guard<std::mutex> lk(srv->mu);
        r->ticket = srv->next_ticket++;
        r->t_submit = Clock::now();
        Req *raw = r.get();
        srv->reqs.emplace(raw->ticket, std::move(r));
        (raw->batchable ? srv->small : srv->large).push_back(raw);
        *ticket = raw->ticket;
        srv->cv_work.notify_all();
    });
}

lgh_status lgh_server_wait(lgh_server *srv, uint64_t ticket, double *expval, double *grad, int64_t grad_len,
                           double *timing_ms) {
    return guard([&] {
        LGH_REQUIRE(srv, "null server");
        std::unique_ptr<Req> r;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            auto it = srv->reqs.find(ticket);
            LGH_REQUIRE(it != srv->reqs.end(), "unknown ticket");
            Req *raw = it->second.get();
            srv->cv_done.wait(lk, [&] { return raw->done; });
            r = std::move(it->second);
            srv->reqs.erase(it);
        }
        if (r->status != LGH_OK) throw Error(r->status, r->error);
        if (expval) *expval = r->expval;
        if (grad) {
            LGH_REQUIRE(grad_len >= int64_t(r->grad_out.size()), "gradient buffer too small");
            std::copy(r->grad_out.begin(), r->grad_out.end(), grad);
        }
        if (timing_ms) {
            timing_ms[0] = ms_between(r->t_submit, r->t_start);
            timing_ms[1] = ms_between(r->t_start, r->t_end);
            timing_ms[2] = ms_between(r->t_submit, r->t_end);
        }
    });
}

lgh_status lgh_server_set_limits(lgh_server *srv, int32_t max_concurrent_batches, int32_t max_batch_circuits) {
    return guard([&] {
        LGH_REQUIRE(srv, "null server");
        LGH_REQUIRE(max_concurrent_batches >= 1, "max_concurrent_batches must be at least 1");
        LGH_REQUIRE(max_batch_circuits >= 1, "max_batch_circuits must be at least 1");
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->max_batches = max_concurrent_batches;
        srv->max_batch = max_batch_circuits;
        srv->cv_work.notify_all();
    });
}

lgh_status lgh_server_set_batching(lgh_server *srv, int32_t min_batch, int32_t wait_us) {
    return guard([&] {
        LGH_REQUIRE(srv, "null server");
        LGH_REQUIRE(min_batch >= 1 && wait_us >= 0, "min_batch >= 1 and wait_us >= 0");
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->min_batch = size_t(min_batch);
        srv->batch_wait = std::chrono::microseconds(wait_us);
        srv->cv_work.notify_all();
    });
}

lgh_status lgh_server_inject_failures(lgh_server *srv, uint64_t count) {
    return guard([&] {
        LGH_REQUIRE(srv, "null server");
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->inject_failures = count;
    });
}

lgh_status lgh_server_stats(lgh_server *srv, uint64_t *batches, uint64_t *batched_requests,
                            uint64_t *single_requests, uint64_t *budget) {
    return guard([&] {
        LGH_REQUIRE(srv, "null server");
        std::lock_guard<std::mutex> lk(srv->mu);
        if (batches) *batches = srv->n_batches;
        if (batched_requests) *batched_requests = srv->n_batched;
        if (single_requests) *single_requests = srv->n_single;
        if (budget) *budget = srv->budget;
    });
}

} // extern "C"
