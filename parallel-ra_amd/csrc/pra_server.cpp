// Serving loop for transitive-closure / reachability requests on one GPU.
//
// Requests (an edge list in the tsv_to_bin row layout) are queued in arrival
// order; `slots` worker threads admit the head of the queue as soon as a slot
// is free and its estimated device memory fits into the HBM budget next to
// the requests already running (a request larger than the whole budget runs
// alone).  Each admitted request is solved on its own context and stream, its
// relations resident on the device for all iterations, so requests that
// arrive while others run overlap with them on the GPU.  The program solved is
// the reference's tests/transitive_closure.cpp:56-77 (copy seed, then the
// T(DELTA) x G(FULL) join to the fixed point).
#include "../../include/pra.h"

#include <algorithm>
#include <chrono>
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
    uint64_t id = 0;
    std::vector<uint64_t> edges;
    uint64_t m = 0, est = 0;
    pra_request_result res{};
    std::vector<uint64_t> rows;
    bool done = false;
};

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

} // namespace

struct pra_server {
    int device = 0, nb = 1, slots = 1, keep_rows = 0;
    uint64_t budget = 0, inflight = 0, next_id = 1;
    bool stopping = false;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::vector<std::thread> workers;

    bool admissible()
    {
        if (queue.empty()) return false;
        return inflight == 0 || budget == 0 || inflight + queue.front()->est <= budget;
    }

    void solve(Request* r)
    {
        pra_ctx* c = nullptr;
        pra_rel *T = nullptr, *G = nullptr;
        pra_scc *fix = nullptr, *seed = nullptr;
        int rc = pra_ctx_create(device, nb, 0, &c);
        const int proj[2] = {3, 1}, reorder[2] = {1, 0};
        if (rc == PRA_OK) rc = pra_ctx_set_batch_size(c, 1);
        if (rc == PRA_OK) rc = pra_rel_create(c, 2, 1, 257, 1, &T);
        if (rc == PRA_OK) rc = pra_rel_create(c, 2, 1, 256, 1, &G);
        if (rc == PRA_OK) rc = pra_rel_load_rows(G, r->edges.data(), r->m, PRA_FULL);
        if (rc == PRA_OK) rc = pra_scc_create(c, 1, 13237, &fix);
        if (rc == PRA_OK) rc = pra_scc_add_relation(fix, G, 0);
        if (rc == PRA_OK) rc = pra_scc_add_relation(fix, T, 1);
        if (rc == PRA_OK) rc = pra_scc_add_join(fix, T, T, PRA_DELTA, G, PRA_FULL, proj, 2);
        if (rc == PRA_OK) rc = pra_scc_create(c, 0, 13239, &seed);
        if (rc == PRA_OK) rc = pra_scc_add_relation(seed, G, 0);
        if (rc == PRA_OK) rc = pra_scc_add_relation(seed, T, 1);
        if (rc == PRA_OK) rc = pra_scc_add_copy(seed, T, G, PRA_FULL, reorder, 2);
        if (rc == PRA_OK) rc = pra_scc_add_dependency(seed, fix);
        if (rc == PRA_OK) rc = pra_execute(c);
        if (rc == PRA_OK) rc = pra_rel_count(T, PRA_FULL, &r->res.tuples);
        if (rc == PRA_OK) rc = pra_ctx_joined_tuples(c, &r->res.joined);
        if (rc == PRA_OK) rc = pra_ctx_iterations(c, &r->res.iterations);
        if (rc == PRA_OK && keep_rows) {
            r->rows.resize(r->res.tuples * 3);
            uint64_t got = 0;
            rc = pra_rel_download(T, PRA_FULL, r->rows.data(), r->res.tuples, &got);
        }
        if (c) pra_ctx_destroy(c);
        r->res.status = rc;
    }

    void worker()
    {
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return (stopping && queue.empty()) || admissible(); });
                if (!admissible()) return;
                r = queue.front();
                queue.pop_front();
                inflight += r->est;
                r->res.t_start = seconds_since(t0);
            }
            solve(r);
            {
                std::lock_guard<std::mutex> lk(mu);
                inflight -= r->est;
                r->res.t_done = seconds_since(t0);
                r->edges.clear();
                r->edges.shrink_to_fit();
                r->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int pra_server_create(int device, int nbuckets, int slots, uint64_t hbm_budget, int keep_rows, pra_server** out)
{
    if (!out || slots < 1 || slots > 64 || nbuckets < 1 || nbuckets > 256) return PRA_E_ARG;
    pra_server* s = new pra_server();
    s->device = device;
    s->nb = nbuckets;
    s->slots = slots;
    s->budget = hbm_budget;
    s->keep_rows = keep_rows;
    for (int i = 0; i < slots; i++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return PRA_OK;
}

int pra_server_submit_tc(pra_server* s, const uint64_t* edge_rows, uint64_t m, uint64_t est_bytes, uint64_t* ticket)
{
    if (!s || !ticket || (m && !edge_rows)) return PRA_E_ARG;
    std::unique_ptr<Request> r(new Request());
    r->edges.assign(edge_rows, edge_rows + m * 3);
    r->m = m;
    r->est = est_bytes;
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
    for (auto& t : s->workers) t.join();
    delete s;
    return PRA_OK;
}

} // extern "C"
