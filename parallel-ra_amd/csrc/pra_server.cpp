// Serving loop for transitive-closure / reachability requests on one GPU,
// with continuous batching.
//
// Requests (an edge list in the tsv_to_bin row layout) are queued in arrival
// order.  One serving thread owns one context whose program instances
// (pra_ctx_set_group) are the requests in flight, up to `slots` of them: each
// lock-step iteration (pra_ctx_step) advances every request in flight by one
// iteration, all of them in the same batched launches, and between iterations
// the thread retires the requests that reached their fixed point (results
// read, instance released) and admits queued ones into the freed slots when
// their memory fits.  Admission counts every request in flight at the larger
// of its estimate and the bytes its relations hold (pra_ctx_group_status),
// the new request's estimate, and the largest amount of context scratch seen
// after any step (the device bytes of the context beyond its requests'
// relations: candidate, sort and staging buffers shared by all instances),
// against the HBM budget; the first request of an idle server is always
// admitted.  Relations stay resident on the device from
// admission to completion.  Estimates can be low, so the loop restarts
// requests rather than fail them when memory runs short: past the budget after
// an iteration, the context's scratch is given back (pra_ctx_trim), then the
// newest admissions in flight are released and queued again at the head of
// the queue (released bytes count at once); an iteration that
// runs the device out of memory with several requests in flight requeues all
// of them on a fresh context and halves the number in flight, which then
// grows back by one per completed request.  A large request (estimate at or
// above `large`, by default that of a 10,000-vertex graph) fills the device
// by itself: two of them in flight together finish later than one after the
// other, their random-access membership bitmaps evicting each other from the
// last-level cache.  So one large request is in flight at a time; a large
// request that waits only for that one lets the requests behind it pass, and
// one that waits for memory holds the queue as any other.  The program solved is the reference's
// tests/transitive_closure.cpp:56-77 (copy seed, then the T(DELTA) x G(FULL)
// join to the fixed point).
#include "../../include/pra.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0, admitted = 0;   // arrival number, admission number (last admission)
    std::vector<uint64_t> edges;
    uint64_t m = 0, est = 0;
    pra_request_result res{};
    std::vector<uint64_t> rows;
    bool done = false;
};

struct Slot {
    Request* r = nullptr;
    pra_rel* T = nullptr;
};

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// device bytes a TC request on m edges is expected to need at its peak:
// random graphs with 1.3 edges per vertex close to about 0.18 n^2 paths, and
// a solve peaks at up to 90 bytes per final path (24-byte rows of the full,
// delta and newt versions in blocks up to 1.5x their size, the growth copy,
// the join candidates and the dense membership bitmap), plus a 64 MiB
// allowance for scratch
uint64_t default_estimate(uint64_t m)
{
    const double n = (double)m / 1.3;
    return (uint64_t)(0.18 * n * n * 90.0) + (64ULL << 20);
}

} // namespace

struct pra_server {
    int device = 0, nb = 1, slots = 1, keep_rows = 0;
    uint64_t large = 0;   // estimates from here up run one at a time (0: no such rule)
    int limit = 1;   // requests in flight at most for now: slots, halved after running out of memory
    uint64_t budget = 0, next_id = 1, admissions = 0;
    uint64_t scratch_hw = 0;   // largest context scratch seen after a step (bytes beyond the relations)
    bool stopping = false;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::thread loop;
    // owned by the serving thread
    pra_ctx* ctx = nullptr;
    std::vector<Slot> slot;

    int in_flight() const
    {
        int k = 0;
        for (const Slot& s : slot) k += s.r != nullptr;
        return k;
    }

    // device bytes the relations of instance g hold
    uint64_t held_by(int g)
    {
        uint64_t held = 0;
        pra_ctx_group_status(ctx, g, nullptr, nullptr, &held);
        return held;
    }

    // after a step: the context's bytes beyond the relations in flight
    void observe_scratch()
    {
        uint64_t live = 0, held = 0;
        if (pra_ctx_device_bytes(ctx, &live, nullptr) != PRA_OK) return;
        for (int g = 0; g < slots; g++)
            if (slot[(size_t)g].r) held += held_by(g);
        if (live > held) scratch_hw = std::max(scratch_hw, live - held);
    }

    bool is_large(const Request* r) const { return large && r->est >= large; }

    bool large_in_flight() const
    {
        for (const Slot& s : slot)
            if (s.r && is_large(s.r)) return true;
        return false;
    }

    // may the head of the queue join the requests in flight and the `pending`
    // bytes of estimates admitted before it in the same pass?
    bool fits(const Request* r, bool first, uint64_t pending)
    {
        if ((first && !in_flight()) || !budget) return true;
        uint64_t need = scratch_hw + pending + r->est;
        for (int g = 0; g < slots; g++)
            if (slot[(size_t)g].r) need += std::max(slot[(size_t)g].r->est, held_by(g));
        return need <= budget;
    }

    // the program of tests/transitive_closure.cpp as instance g
    int build(int g, Request* r)
    {
        pra_rel *T = nullptr, *G = nullptr;
        pra_scc *fix = nullptr, *seed = nullptr;
        const int proj[2] = {3, 1}, reorder[2] = {1, 0};
        int rc = pra_ctx_set_group(ctx, g);
        if (rc == PRA_OK) rc = pra_rel_create(ctx, 2, 1, 257, 1, &T);
        if (rc == PRA_OK) rc = pra_rel_create(ctx, 2, 1, 256, 1, &G);
        if (rc == PRA_OK) rc = pra_rel_load_rows(G, r->edges.data(), r->m, PRA_FULL);
        if (rc == PRA_OK) rc = pra_scc_create(ctx, 1, 13237, &fix);
        if (rc == PRA_OK) rc = pra_scc_add_relation(fix, G, 0);
        if (rc == PRA_OK) rc = pra_scc_add_relation(fix, T, 1);
        if (rc == PRA_OK) rc = pra_scc_add_join(fix, T, T, PRA_DELTA, G, PRA_FULL, proj, 2);
        if (rc == PRA_OK) rc = pra_scc_create(ctx, 0, 13239, &seed);
        if (rc == PRA_OK) rc = pra_scc_add_relation(seed, G, 0);
        if (rc == PRA_OK) rc = pra_scc_add_relation(seed, T, 1);
        if (rc == PRA_OK) rc = pra_scc_add_copy(seed, T, G, PRA_FULL, reorder, 2);
        if (rc == PRA_OK) rc = pra_scc_add_dependency(seed, fix);
        slot[(size_t)g].T = T;
        return rc;
    }

    void finish(Request* r, int status)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            r->res.status = status;
            r->res.t_done = seconds_since(t0);
            r->edges.clear();
            r->edges.shrink_to_fit();
            r->done = true;
        }
        cv_done.notify_all();
    }

    // results of a finished instance, then its release
    void retire(int g)
    {
        Slot& s = slot[(size_t)g];
        Request* r = s.r;
        int rc = pra_rel_count(s.T, PRA_FULL, &r->res.tuples);
        if (rc == PRA_OK) rc = pra_ctx_group_joined(ctx, g, &r->res.joined);
        if (rc == PRA_OK) rc = pra_ctx_group_status(ctx, g, nullptr, &r->res.iterations, nullptr);
        if (rc == PRA_OK && keep_rows) {
            r->rows.resize(r->res.tuples * 3);
            uint64_t got = 0;
            rc = pra_rel_download(s.T, PRA_FULL, r->rows.data(), r->res.tuples, &got);
        }
        if (rc != PRA_OK)
            std::fprintf(stderr, "pra_server: results of request %llu not read (%d): %s\n",
                         (unsigned long long)r->id, rc, pra_last_error(ctx));
        const int rel = pra_ctx_release_group(ctx, g);
        s = Slot{};
        finish(r, rc != PRA_OK ? rc : rel);
    }

    // a request released before its end needs at least what it held: its
    // estimate becomes at least twice that, so that it is not admitted again
    // into the same shortage
    void raise_estimate(Request* r, int g)
    {
        if (ctx) r->est = std::max(r->est, 2 * held_by(g));
    }

    // requests back to the head of the queue, in arrival order
    void requeue(std::vector<Request*> rs)
    {
        std::sort(rs.begin(), rs.end(), [](const Request* a, const Request* b) { return a->id > b->id; });
        std::lock_guard<std::mutex> lk(mu);
        for (Request* r : rs) {
            r->res.restarts++;
            queue.push_front(r);
        }
    }

    // a failed step leaves the context unusable: the requests in flight are
    // queued again after running out of memory with several in flight, else
    // they fail; the next ones start on a fresh context
    void reset_context(int rc)
    {
        const int n = in_flight();
        const bool retry = (rc == PRA_E_NOMEM || rc == PRA_E_BUDGET) && n > 1;
        std::fprintf(stderr, "pra_server: step failed (%d): %s; %d request(s) in flight %s\n", rc,
                     pra_last_error(ctx), n, retry ? "queued again" : "fail");
        std::vector<Request*> back;
        for (size_t g = 0; g < slot.size(); g++)
            if (Request* r = slot[g].r) {
                if (retry) {
                    raise_estimate(r, (int)g);
                    back.push_back(r);
                } else {
                    finish(r, rc);
                }
            }
        if (retry) limit = std::max(1, n / 2);
        slot.assign((size_t)slots, Slot{});
        pra_ctx_destroy(ctx);
        ctx = nullptr;
        if (pra_ctx_create(device, nb, 0, &ctx) == PRA_OK) pra_ctx_set_batch_size(ctx, 1);
        requeue(back);
    }

    // past the budget after an iteration: the context's scratch is given
    // back, then the newest admissions are released until the bytes held fit
    // again (one request always stays)
    void relieve()
    {
        if (!budget) return;
        bool trimmed = false;
        while (in_flight() > 1) {
            uint64_t live = 0;
            if (pra_ctx_device_bytes(ctx, &live, nullptr) != PRA_OK || live <= budget) return;
            if (!trimmed) {
                trimmed = true;
                if (pra_ctx_trim(ctx) == PRA_OK) continue;
            }
            int newest = -1;
            for (int g = 0; g < slots; g++)
                if (slot[(size_t)g].r && (newest < 0 || slot[(size_t)g].r->admitted > slot[(size_t)newest].r->admitted))
                    newest = g;
            Request* r = slot[(size_t)newest].r;
            raise_estimate(r, newest);
            pra_ctx_release_group(ctx, newest);
            slot[(size_t)newest] = Slot{};
            requeue({r});
        }
    }

    void run()
    {
        slot.assign((size_t)slots, Slot{});
        const int created = pra_ctx_create(device, nb, 0, &ctx);
        if (created == PRA_OK) pra_ctx_set_batch_size(ctx, 1);
        for (;;) {
            std::vector<Request*> admit;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !queue.empty() || in_flight(); });
                if (stopping && queue.empty() && !in_flight()) break;
                int free_slots = std::min(slots, limit) - in_flight();
                uint64_t pending = 0;
                bool large_busy = large_in_flight();
                for (auto it = queue.begin(); ctx && free_slots > 0 && it != queue.end();) {
                    Request* r = *it;
                    const bool lg = is_large(r);
                    if (lg && large_busy) {   // waits for the large request in flight
                        ++it;
                        continue;
                    }
                    if (!fits(r, admit.empty(), pending)) break;
                    pending += r->est;
                    admit.push_back(r);
                    it = queue.erase(it);
                    free_slots--;
                    large_busy = large_busy || lg;
                }
            }
            if (!ctx) {   // no context could be created: fail what is queued
                for (Request* r : admit) finish(r, created != PRA_OK ? created : PRA_E_HIP);
                std::lock_guard<std::mutex> lk(mu);
                for (Request* r : queue) {
                    r->res.status = PRA_E_HIP;
                    r->done = true;
                }
                queue.clear();
                cv_done.notify_all();
                continue;
            }
            for (Request* r : admit) {
                int g = 0;
                while (slot[(size_t)g].r) g++;
                slot[(size_t)g].r = r;
                r->admitted = ++admissions;
                r->res.t_start = seconds_since(t0);
                const int rc = build(g, r);
                if (rc != PRA_OK) {
                    std::fprintf(stderr, "pra_server: request %llu (%llu edges) not set up (%d): %s\n",
                                 (unsigned long long)r->id, (unsigned long long)r->m, rc, pra_last_error(ctx));
                    pra_ctx_release_group(ctx, g);
                    slot[(size_t)g] = Slot{};
                    finish(r, rc);
                }
            }
            if (!in_flight()) continue;
            uint64_t active = 0;
            const int rc = pra_ctx_step(ctx, &active);
            if (rc != PRA_OK) {
                reset_context(rc);
                continue;
            }
            observe_scratch();
            for (int g = 0; g < slots; g++) {
                if (!slot[(size_t)g].r) continue;
                int done = 0;
                pra_ctx_group_status(ctx, g, &done, nullptr, nullptr);
                if (done) {
                    retire(g);
                    limit = std::min(slots, limit + 1);
                }
            }
            relieve();
        }
        if (ctx) pra_ctx_destroy(ctx);
        ctx = nullptr;
    }
};

extern "C" {

int pra_server_create(int device, int nbuckets, int slots, uint64_t hbm_budget, int keep_rows, pra_server** out)
{
    if (!out || slots < 1 || slots > 4096 || nbuckets < 1 || nbuckets > 256) return PRA_E_ARG;
    pra_server* s = new pra_server();
    s->device = device;
    s->nb = nbuckets;
    s->slots = slots;
    s->limit = slots;
    s->budget = hbm_budget;
    s->keep_rows = keep_rows;
    s->large = default_estimate((uint64_t)(1.3 * 10000));
    if (const char* v = std::getenv("PRA_SERVE_LARGE_BYTES")) s->large = std::strtoull(v, nullptr, 10);
    s->loop = std::thread([s] { s->run(); });
    *out = s;
    return PRA_OK;
}

int pra_server_submit_tc(pra_server* s, const uint64_t* edge_rows, uint64_t m, uint64_t est_bytes, uint64_t* ticket)
{
    if (!s || !ticket || (m && !edge_rows)) return PRA_E_ARG;
    std::unique_ptr<Request> r(new Request());
    r->edges.assign(edge_rows, edge_rows + m * 3);
    r->m = m;
    r->est = est_bytes ? est_bytes : default_estimate(m);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return PRA_E_STATE;
        r->id = s->next_id++;
        r->res.t_submit = seconds_since(s->t0);
        *ticket = r->id;
        s->queue.push_back(r.get());
        s->all[r->id] = std::move(r);
    }
    s->cv_work.notify_one();
    return PRA_OK;
}

int pra_server_wait(pra_server* s, uint64_t ticket, pra_request_result* out)
{
    if (!s || !out) return PRA_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return PRA_E_ARG;
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    *out = r->res;
    return PRA_OK;
}

int pra_server_rows(pra_server* s, uint64_t ticket, uint64_t* rows, uint64_t cap, uint64_t* nrows)
{
    if (!s || !nrows) return PRA_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end() || !it->second->done) return PRA_E_STATE;
    Request* r = it->second.get();
    uint64_t n = r->rows.size() / 3;
    *nrows = n;
    for (uint64_t i = 0; rows && i < std::min(n, cap) * 3; i++) rows[i] = r->rows[i];
    return PRA_OK;
}

int pra_server_destroy(pra_server* s)
{
    if (!s) return PRA_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->loop.joinable()) s->loop.join();
    delete s;
    return PRA_OK;
}

} // extern "C"
