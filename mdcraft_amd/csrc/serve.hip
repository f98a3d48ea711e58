// Continuous batching of MD requests on one GPU: the mdc_serve_* entry
// points of include/mdc.h.  No reference counterpart: the reference runs one
// system per process.
//
// A request is one structure (atom records, box) for one registered
// potential, with a task: a single-point energy/force/virial evaluation
// (steps = 0) or an NVE run of `steps` velocity-Verlet steps followed by
// one.  The server
//  * admits submitted requests in arrival order into free slots of running
//    batches: a batch is one engine system whose boxes are slots of one
//    size class (records per slot), for one potential and time step; a
//    request takes a free slot of its class (mdc_sys_refill_box), so
//    arrivals join batches that are already stepping;
//  * opens a new batch for a request that finds no free slot, while the
//    estimated device bytes of all open batches stay within the budget
//    (idle batches are closed first when the budget is short);
//  * drives every batch from its own thread on its own context (stream):
//    batches step concurrently and their kernels overlap on the device;
//  * advances each batch in chunks of at most `chunk` steps, never past
//    the smallest step count left among its running slots, and retires a
//    request when its count is reached (a final E/F/virial evaluation of
//    the batch, its records downloaded, its slot switched off and freed).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "abi.h"

using namespace mdc;

namespace {

using Clock = std::chrono::steady_clock;

// smallest size class (records per slot) holding n atoms: 64 * 2^(k/2)
int size_class(int n) {
    int k = 0;
    while (true) {
        const double c = std::floor(64.0 * std::pow(2.0, 0.5 * k) + 0.5);
        if (c >= n) return (int)c;
        ++k;
    }
}

struct SKind {
    const mdc_pot* pot;
    int precision;
};

struct SBatch;

struct SReq {
    uint64_t id = 0;
    int kind = 0, steps = 0, left = 0;
    double dt = 0.0;
    std::vector<AtomRec> atoms;
    double lo[3], hi[3];
    uint8_t per[3];
    int cls = 0;
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    enum { QUEUED, ASSIGNED, RUNNING, DONE, FAILED } state = QUEUED;
    int code = 0;
    std::string msg;
    std::vector<AtomRec> out;
    double thermo[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    SBatch* batch = nullptr;
    int slot = -1;
};

struct SBatch {
    int kind = 0, cls = 0, nslots = 0;
    double dt = 0.0, bytes = 0.0;
    std::vector<SReq*> slot;                  // request per slot (nullptr: free)
    std::deque<std::pair<int, SReq*>> fills;  // assigned, not yet in the system
    bool stop = false, exited = false;
    std::condition_variable cv;
    std::thread th;
};

}  // namespace

struct mdc_server {
    int device = 0, chunk = 25, slots_per_batch = 8;
    double budget = 200e9;
    std::vector<SKind> kinds;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<SReq*> queue;  // submitted, not admitted, arrival order
    std::map<uint64_t, std::unique_ptr<SReq>> reqs;
    std::vector<std::unique_ptr<SBatch>> batches;
    uint64_t next_id = 1;
    double open_bytes = 0.0, peak_bytes = 0.0;
    long done = 0, failed = 0, batches_opened = 0, joined = 0;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    std::thread sched;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Device bytes of a batch of `records` records for requests like r: the
// per-record state of a system (positions and per-atom fields, double
// buffered, list bookkeeping: about 470 bytes), the neighbour list at the
// request's density (capacity 1.3 x the mean neighbour count + 16, as the
// engine sizes it, 4 bytes a slot) and, for BPNN, the work arrays of an
// evaluation with the virial (descriptors and dE/dG, force and virial
// shares, reverse-slot index).
double batch_bytes(const SReq& r, int type, int precision, int records) {
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= std::max(r.hi[a] - r.lo[a], 1e-9);
    double rns = 0.0;
    for (const AtomRec& x : r.atoms) rns = std::max(rns, x.rns);
    const double dens = r.atoms.size() / vol;
    const double cap = 1.3 * dens * 4.18879 * rns * rns * rns + 16.0;
    double per = 470.0 + 4.0 * cap;
    if (type == PT_BPNN) {
        const double rs = precision == MDC_FP64 ? 8.0 : 4.0;
        per += 2.0 * 64.0 * rs + 2.0 * cap * 3.0 * rs + 6.0 * rs + 4.0 * cap;
    }
    return per * records;
}

void fail(mdc_server* S, SReq* r, int code, const std::string& msg) {
    r->state = SReq::FAILED;
    r->code = code ? code : E_INVAL;
    r->msg = msg;
    r->t_done = S->now();
    ++S->failed;
}

std::string ctx_msg(mdc_ctx* c) {
    const char* m = c ? mdc_ctx_last_error(c) : mdc_last_error();
    return m ? std::string(m) : std::string("engine error");
}

#define SRV_CALL(expr)                                                                \
    do {                                                                              \
        const int st_ = (expr);                                                       \
        if (st_ != 0) throw Error(st_, std::string(#expr ": ") + ctx_msg(ctx));       \
    } while (0)

// The final E/F/virial evaluation of the batch and, for each finishing
// request, its records and the thermo row of its slot; then the slot is
// emptied (all its records void, so the next sort moves them out of the
// kernels' range) and freed.
void finish(mdc_server* S, SBatch* B, mdc_ctx* ctx, mdc_sys* sys, const SKind& K, const std::vector<SReq*>& reqs) {
    SRV_CALL(mdc_compute(sys, K.pot, MDC_FORCE | MDC_ENERGY | MDC_VIRIAL, K.precision));
    std::vector<double> th((size_t)8 * B->nslots);
    SRV_CALL(mdc_thermo(sys, th.data()));
    const double lo0[3] = {0.0, 0.0, 0.0}, hi0[3] = {1.0, 1.0, 1.0};
    const uint8_t per0[3] = {0, 0, 0};
    for (SReq* r : reqs) {
        r->out.resize(r->atoms.size());
        const uint64_t first = (uint64_t)r->slot * (uint64_t)B->cls;
        SRV_CALL(mdc_sys_download_range(sys, first, r->atoms.size(), r->out.data(), MDC_F_ALL));
        std::memcpy(r->thermo, &th[(size_t)8 * r->slot], sizeof(r->thermo));
        SRV_CALL(mdc_sys_refill_box(sys, r->slot, 0, nullptr, lo0, hi0, per0));
    }
    std::lock_guard<std::mutex> lk(S->mu);
    for (SReq* r : reqs) {
        r->state = SReq::DONE;
        r->t_done = S->now();
        B->slot[r->slot] = nullptr;
        r->batch = nullptr;
        ++S->done;
    }
    S->cv_done.notify_all();
    S->cv_sched.notify_all();
}

void batch_loop(mdc_server* S, SBatch* B) {
    SKind K;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        K = S->kinds[B->kind];
    }
    mdc_ctx* ctx = nullptr;
    mdc_sys* sys = nullptr;
    auto fail_all = [&](int code, const std::string& msg) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (int s = 0; s < B->nslots; ++s)
            if (B->slot[s]) {
                fail(S, B->slot[s], code, msg);
                B->slot[s]->batch = nullptr;
                B->slot[s] = nullptr;
            }
        B->fills.clear();
        B->stop = true;
        S->cv_done.notify_all();
        S->cv_sched.notify_all();
    };
    try {
        SRV_CALL(mdc_ctx_create(S->device, &ctx));
        std::vector<uint64_t> natoms(B->nslots, (uint64_t)B->cls);
        std::vector<double> lo(3 * B->nslots, 0.0), hi(3 * B->nslots, 1.0);
        std::vector<uint8_t> per(3 * B->nslots, 0);
        SRV_CALL(mdc_sys_create(ctx, B->nslots, natoms.data(), lo.data(), hi.data(), per.data(), 1, &sys));
        for (int s = 0; s < B->nslots; ++s) SRV_CALL(mdc_sys_refill_box(sys, s, 0, nullptr, &lo[0], &hi[0], &per[0]));
        while (true) {
            std::vector<std::pair<int, SReq*>> fills;
            std::vector<SReq*> running;
            {
                std::unique_lock<std::mutex> lk(S->mu);
                B->cv.wait(lk, [&] {
                    if (B->stop || !B->fills.empty()) return true;
                    for (SReq* r : B->slot)
                        if (r && r->state == SReq::RUNNING) return true;
                    return false;
                });
                fills.assign(B->fills.begin(), B->fills.end());
                B->fills.clear();
                for (auto& f : fills) f.second->state = SReq::RUNNING;
                for (SReq* r : B->slot)
                    if (r && r->state == SReq::RUNNING) running.push_back(r);
                if (B->stop && running.empty()) break;
            }
            if (!fills.empty()) {
                for (auto& f : fills) {
                    SReq* r = f.second;
                    SRV_CALL(mdc_sys_refill_box(sys, f.first, r->atoms.size(), r->atoms.data(), r->lo, r->hi, r->per));
                }
                SRV_CALL(mdc_nlist_build(sys));
                SRV_CALL(mdc_compute(sys, K.pot, MDC_FORCE, K.precision));
                std::vector<SReq*> singles;
                for (auto& f : fills)
                    if (f.second->left == 0) singles.push_back(f.second);
                if (!singles.empty()) finish(S, B, ctx, sys, K, singles);
                continue;  // pick up further fills before stepping
            }
            int k = S->chunk;
            for (SReq* r : running) k = std::min(k, r->left);
            if (k <= 0) continue;
            int reb = 0;
            SRV_CALL(mdc_step(sys, K.pot, B->dt, k, K.precision, &reb));
            std::vector<SReq*> ended;
            for (SReq* r : running) {
                r->left -= k;
                if (r->left == 0) ended.push_back(r);
            }
            if (!ended.empty()) finish(S, B, ctx, sys, K, ended);
        }
    } catch (const Error& e) {
        fail_all(e.code, e.what());
    } catch (const std::exception& e) {
        fail_all(E_INVAL, e.what());
    }
    if (sys) (void)mdc_sys_destroy(sys);
    if (ctx) (void)mdc_ctx_destroy(ctx);
    std::lock_guard<std::mutex> lk(S->mu);
    B->exited = true;
    S->cv_sched.notify_all();
}

// Admission in arrival order (the scheduler's lock held).  Returns when the
// queue's head can neither take a slot nor open a batch.
void admit(mdc_server* S) {
    while (!S->queue.empty()) {
        SReq* r = S->queue.front();
        // a free slot of the request's class in a running batch
        SBatch* best = nullptr;
        int bs = -1;
        for (auto& b : S->batches) {
            // a single point never steps, so it joins a batch of any time
            // step; a batch that has stepped nothing yet takes the first
            // run's time step
            const bool dt_ok = r->steps == 0 || b->dt == r->dt || b->dt == 0.0;
            if (b->stop || b->kind != r->kind || !dt_ok || b->cls != r->cls) continue;
            for (int s = 0; s < b->nslots && bs < 0; ++s)
                if (!b->slot[s]) {
                    best = b.get();
                    bs = s;
                }
            if (best) break;
        }
        if (!best) {
            const SKind& K = S->kinds[r->kind];
            const double need = batch_bytes(*r, K.pot->p.type, K.precision, r->cls * S->slots_per_batch);
            if (need > S->budget) {
                S->queue.pop_front();
                fail(S, r, E_BUDGET, "request needs a batch of about " + std::to_string((long long)need) +
                                         " bytes, above the whole memory budget");
                S->cv_done.notify_all();
                continue;
            }
            if (S->open_bytes + need > S->budget) {
                // close idle batches to make room; the request waits (later
                // arrivals wait behind it)
                for (auto& b : S->batches) {
                    bool idle = b->fills.empty();
                    for (SReq* q : b->slot) idle = idle && !q;
                    if (idle && !b->stop) {
                        b->stop = true;
                        b->cv.notify_all();
                    }
                }
                return;
            }
            auto nb = std::make_unique<SBatch>();
            nb->kind = r->kind;
            nb->dt = r->dt;
            nb->cls = r->cls;
            nb->nslots = S->slots_per_batch;
            nb->slot.assign(nb->nslots, nullptr);
            nb->bytes = need;
            S->open_bytes += need;
            S->peak_bytes = std::max(S->peak_bytes, S->open_bytes);
            ++S->batches_opened;
            best = nb.get();
            bs = 0;
            S->batches.push_back(std::move(nb));
            best->th = std::thread(batch_loop, S, best);
        }
        S->queue.pop_front();
        if (r->steps > 0) best->dt = r->dt;
        best->slot[bs] = r;
        best->fills.emplace_back(bs, r);
        r->batch = best;
        r->slot = bs;
        r->state = SReq::ASSIGNED;
        r->t_admit = S->now();
        best->cv.notify_all();
    }
}

// Joins batch threads that have ended and returns their budget.
void reap(mdc_server* S, std::unique_lock<std::mutex>& lk) {
    for (size_t q = 0; q < S->batches.size();) {
        SBatch* b = S->batches[q].get();
        if (!b->exited) {
            ++q;
            continue;
        }
        std::unique_ptr<SBatch> owned = std::move(S->batches[q]);
        S->batches.erase(S->batches.begin() + q);
        S->open_bytes -= owned->bytes;
        lk.unlock();
        if (owned->th.joinable()) owned->th.join();
        lk.lock();
        ++S->joined;
    }
}

void sched_loop(mdc_server* S) {
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop) {
        reap(S, lk);
        admit(S);
        S->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
    }
    for (auto& b : S->batches) {
        b->stop = true;
        b->cv.notify_all();
    }
    while (!S->batches.empty()) {
        reap(S, lk);
        if (!S->batches.empty()) S->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
    }
}

template <class F>
int sguard(F&& f) {
    try {
        f();
        return 0;
    } catch (const Error& e) {
        return report_error(nullptr, e.code, e.what());
    } catch (const std::exception& e) {
        return report_error(nullptr, E_INVAL, e.what());
    }
}

}  // namespace

extern "C" {

mdc_status mdc_serve_create(int device, double budget_bytes, int chunk_steps, int slots_per_batch, mdc_server** out) {
    return sguard([&] {
        if (!out) throw Error(E_INVAL, "null output pointer");
        if (!(budget_bytes > 0) || chunk_steps < 1 || slots_per_batch < 1 || slots_per_batch > 4096)
            throw Error(E_INVAL, "budget must be positive, chunk and slots per batch at least 1");
        int ndev = 0;
        HIPCHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(E_INVAL, "no such device");
        auto S = std::make_unique<mdc_server>();
        S->device = device;
        S->budget = budget_bytes;
        S->chunk = chunk_steps;
        S->slots_per_batch = slots_per_batch;
        S->sched = std::thread(sched_loop, S.get());
        *out = S.release();
    });
}

mdc_status mdc_serve_add_potential(mdc_server* S, const mdc_pot* pot, int precision, int* kind) {
    return sguard([&] {
        if (!S || !pot || !kind) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        std::lock_guard<std::mutex> lk(S->mu);
        S->kinds.push_back(SKind{pot, precision});
        *kind = (int)S->kinds.size() - 1;
    });
}

mdc_status mdc_serve_submit(mdc_server* S, int kind, const void* atoms, uint64_t natoms, const double* box_lo,
                            const double* box_hi, const uint8_t* periodic, double dt, int steps, uint64_t* id) {
    return sguard([&] {
        if (!S || !id || !box_lo || !box_hi || !periodic || (natoms && !atoms)) throw Error(E_INVAL, "null argument");
        if (natoms == 0 || natoms >= (uint64_t)kMaxAtomsShift) throw Error(E_INVAL, "a request needs 1 to 2^26 - 1 atoms");
        if (steps < 0 || (steps > 0 && !(dt > 0))) throw Error(E_INVAL, "steps must be >= 0 with a positive time step");
        auto r = std::make_unique<SReq>();
        r->atoms.resize(natoms);
        std::memcpy(r->atoms.data(), atoms, natoms * sizeof(AtomRec));
        for (int a = 0; a < 3; ++a) {
            r->lo[a] = box_lo[a];
            r->hi[a] = box_hi[a];
            r->per[a] = periodic[a] ? 1 : 0;
        }
        r->steps = r->left = steps;
        r->dt = steps > 0 ? dt : 0.0;
        r->cls = size_class((int)natoms);
        std::lock_guard<std::mutex> lk(S->mu);
        if (kind < 0 || kind >= (int)S->kinds.size()) throw Error(E_INVAL, "unknown potential kind");
        r->kind = kind;
        r->id = S->next_id++;
        r->t_submit = S->now();
        *id = r->id;
        S->queue.push_back(r.get());
        S->reqs.emplace(r->id, std::move(r));
        S->cv_sched.notify_all();
    });
}

mdc_status mdc_serve_wait(mdc_server* S, uint64_t id, double timeout_s, int* state) {
    return sguard([&] {
        if (!S || !state) throw Error(E_INVAL, "null argument");
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) throw Error(E_INVAL, "unknown request");
        SReq* r = it->second.get();
        auto ended = [&] { return r->state == SReq::DONE || r->state == SReq::FAILED; };
        if (timeout_s > 0) S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
        *state = r->state == SReq::DONE ? 1 : r->state == SReq::FAILED ? -1 : 0;
    });
}

mdc_status mdc_serve_take(mdc_server* S, uint64_t id, void* atoms_out, double* thermo, double* times) {
    std::unique_ptr<SReq> r;
    {
        int code = 0;
        std::string msg;
        if (!S) return report_error(nullptr, E_INVAL, "null argument");
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) return report_error(nullptr, E_INVAL, "unknown request");
        SReq* q = it->second.get();
        if (q->state != SReq::DONE && q->state != SReq::FAILED) return report_error(nullptr, E_STATE, "request not done");
        r = std::move(it->second);
        S->reqs.erase(it);
        if (r->state == SReq::FAILED) {
            code = r->code;
            msg = r->msg;
        }
        if (code) return report_error(nullptr, code, msg);
    }
    if (atoms_out) std::memcpy(atoms_out, r->out.data(), r->out.size() * sizeof(AtomRec));
    if (thermo) std::memcpy(thermo, r->thermo, sizeof(r->thermo));
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    return 0;
}

mdc_status mdc_serve_stats(mdc_server* S, double* out) {
    return sguard([&] {
        if (!S || !out) throw Error(E_INVAL, "null argument");
        std::lock_guard<std::mutex> lk(S->mu);
        long queued = (long)S->queue.size(), running = 0;
        for (auto& kv : S->reqs)
            if (kv.second->state == SReq::ASSIGNED || kv.second->state == SReq::RUNNING) ++running;
        out[0] = (double)S->done;
        out[1] = (double)S->failed;
        out[2] = (double)queued;
        out[3] = (double)running;
        out[4] = (double)S->batches.size();
        out[5] = (double)S->batches_opened;
        out[6] = S->open_bytes;
        out[7] = S->peak_bytes;
    });
}

mdc_status mdc_serve_destroy(mdc_server* S) {
    if (!S) return 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    delete S;
    return 0;
}

}  // extern "C"
