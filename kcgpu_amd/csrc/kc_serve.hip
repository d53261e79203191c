// kc_serve.hip -- serving loop: oriented graphs resident in HBM, (graph, k)
// queries admitted by memory footprint and counted by continuous batching.
//
// No reference counterpart (the reference counts one graph per process).
// Design:
//   * kc_server_add_graph orients a graph on the device and keeps it; it is
//     refused (KC_ERR_BUDGET) when the resident CSRs plus the orientation's
//     transient buffers would exceed the HBM budget.
//   * kc_server_submit only queues the query and returns a ticket.
//   * `workers` host threads, each with its own HIP stream (default priority)
//     and grow-only scratch, take every pending query that fits the free
//     budget (FIFO, at most max_batch) and count them in one run_queries call:
//     persistent launches draining a queue of (query, task).  Queries that
//     arrive while a batch is on the GPU go into the next batch, which the
//     other worker launches on its stream, so consecutive batches overlap on
//     the device.  Each worker's persistent launches take 1/workers of the
//     device's resident block slots, so a long batch (a dense graph at a large
//     k) leaves room for the short batches behind it instead of filling every
//     CU until it ends.
//   * Pivoting queries on graphs whose tasks all fit 64 vertices (every
//     query of most serving mixes) skip the workers: the engine thread
//     appends their tasks to the open queues of one resident launch
//     (kc::OpenEngine) while it runs, so a query that arrives while others
//     are on the GPU is drained by the waves already resident, next to them.
//     Queries of at most EXPRESS_TASKS tasks go to the queue the waves try
//     first, so a short query is not held behind a long one.  A query is
//     admitted when a query slot and queue room are free; each query's
//     result comes back through its slot's mailbox in host memory.
//   * Latency is measured from submit to the moment the result is on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kc_device.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Resident {
    kc_oriented *o = nullptr;
    uint64_t epoch = 0;  // order of arrival among the server's graphs
    uint64_t bytes = 0;
    uint32_t pending = 0;
    bool removed = false;
};

struct Ticket {
    uint64_t id;
    uint32_t gid;
    kc_query q;
    uint64_t scratch;  // estimated HBM scratch bytes of this query
    Clock::time_point t_submit;
};

struct Done {
    kc_result r{};
    double latency_ms = 0;
    int rc = KC_OK;
};

}  // namespace

struct kc_server {
    kc_server_config cfg{};
    std::unique_ptr<kc::OpenEngine> eng;  // resident launch with open queues (null: workers only)
    std::deque<Ticket> open_pending;
    std::thread engine;
    uint64_t admissions = 0, admitted = 0, joined_running = 0, launches = 0;
    // graphs added so far, and how many there were when the resident launch
    // started: a graph uploaded after that start is counted by the workers
    // until the launch restarts (a launch sees memory written before it began)
    uint64_t graph_epoch = 0, launch_epoch = 0;
    int dev = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::unordered_map<uint32_t, Resident> graphs;
    uint32_t next_gid = 1;
    uint64_t next_ticket = 1;
    std::deque<Ticket> pending;
    std::unordered_map<uint64_t, Done> done;
    uint64_t resident_bytes = 0, inflight_scratch = 0, held_scratch = 0;
    // removed graphs (and their bytes) waiting to be freed by a worker outside
    // the lock: hipFree waits for every stream of the device, a resident
    // launch included, and only the engine thread -- which takes this lock
    // and must not wait on the device -- can end that launch
    std::vector<std::pair<kc_oriented *, uint64_t>> free_later;
    uint64_t submitted = 0, completed = 0;             // lifetime, for drain
    uint64_t st_completed = 0, batches = 0, batched_queries = 0;  // since the last stats reset
    std::vector<double> lat;
    Clock::time_point t_first{}, t_last{};
    bool any = false;
    std::vector<std::thread> workers;
};

namespace {

constexpr uint64_t EXPRESS_TASKS = 1u << 16;

uint64_t task_count(const kc_oriented *o, const kc_query &q) { return q.gran == KC_GRAN_EDGE ? o->m : o->n; }

// can the resident launch take this query
bool open_eligible(const kc_server *s, const kc_oriented *o, const kc_query &q)
{
    return s->eng && q.algo == KC_ALGO_PIVOT && q.k >= 3 && o->max_out <= kc::OPEN_MAX_ROWS;
}

// record a finished query (caller holds s->mu)
void complete(kc_server *s, const Ticket &t, const Done &d)
{
    s->lat.push_back(d.latency_ms);
    s->done[t.id] = d;
    Resident &g = s->graphs[t.gid];
    if (--g.pending == 0 && g.removed) {
        s->free_later.emplace_back(g.o, g.bytes);
        s->graphs.erase(t.gid);
        s->cv_work.notify_all();
    }
    s->completed++;
    s->st_completed++;
}

uint64_t oriented_bytes(uint32_t n, uint64_t m) { return 4ull * (n + 1) + 8ull * m + 4ull * n + 4096; }

// free removed graphs (caller holds no lock and is not the engine thread);
// their bytes leave the budget once they are freed
void free_graphs(kc_server *s, std::vector<std::pair<kc_oriented *, uint64_t>> &frees)
{
    uint64_t bytes = 0;
    for (auto &f : frees) {
        kc_oriented_free(f.first);
        bytes += f.second;
    }
    frees.clear();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->resident_bytes -= bytes;
    }
    s->cv_work.notify_all();
}

// transient device memory of kc_orient: full CSR, out-degrees, peel buffers
uint64_t orient_transient(const kc_graph *g)
{
    const uint64_t n = g->csr.n, nnz = g->csr.col.size();
    return 4 * (n + 1) + 4 * nnz + 4 * (n + 1) + 13 * n + 4096;
}

// scratch of one query in a launch: task sizes and codes, plus the HBM bitmap
// slices when out-degrees exceed the LDS classes
uint64_t scratch_bytes(const kc::Scratch &c)
{
    return c.tasks.bytes + c.meta.bytes + c.binom.bytes + c.gscr.bytes + c.size.bytes + c.qdesc.bytes;
}

uint64_t query_scratch(const kc_oriented *o, const kc_query &q)
{
    const uint64_t ntask = q.gran == KC_GRAN_EDGE ? o->m : o->n;
    uint64_t b = 12 * ntask + 64 * 1024;
    if (o->max_out > 1024) b += 2ull << 30;
    return b;
}

void worker_main(kc_server *s)
{
    hipSetDevice(s->dev);
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess) return;
    kc::Scratch scr;
    uint64_t held = 0;  // device bytes this worker's scratch holds between batches
    for (;;) {
        std::vector<Ticket> batch;
        std::vector<std::pair<kc_oriented *, uint64_t>> frees;
        uint64_t batch_scratch = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->free_later.empty(); });
            frees.swap(s->free_later);
            if (frees.empty() && s->stop && s->pending.empty()) break;
            // admit FIFO while the batch fits next to the resident graphs
            const uint64_t budget = s->cfg.budget_bytes;
            while (frees.empty() && !s->pending.empty() && batch.size() < s->cfg.max_batch) {
                const Ticket &t = s->pending.front();
                const uint64_t need =
                    s->resident_bytes + s->held_scratch + s->inflight_scratch + batch_scratch + t.scratch;
                if (!batch.empty() && need > budget) break;
                batch_scratch += t.scratch;
                batch.push_back(t);
                s->pending.pop_front();
            }
            s->inflight_scratch += batch_scratch;
        }
        if (!frees.empty()) {
            free_graphs(s, frees);
            continue;
        }
        std::vector<kc::QueryRun> runs;
        runs.reserve(batch.size());
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (const Ticket &t : batch) runs.push_back(kc::QueryRun{s->graphs[t.gid].o, t.q});
        }
        std::vector<kc_result> res;
        int rc = kc::run_queries(scr, st, runs, res, nullptr, nullptr, s->cfg.workers);
        const auto t_end = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->inflight_scratch -= batch_scratch;
            const uint64_t now_held = scratch_bytes(scr);
            s->held_scratch = s->held_scratch - held + now_held;
            held = now_held;
            s->batches++;
            s->batched_queries += batch.size();
            for (size_t i = 0; i < batch.size(); ++i) {
                Done d;
                d.rc = rc;
                if (rc == KC_OK) {
                    d.r = res[i];
                    if (d.r.overflow) d.rc = KC_ERR_OVERFLOW;
                }
                d.latency_ms = std::chrono::duration<double, std::milli>(t_end - batch[i].t_submit).count();
                complete(s, batch[i], d);
            }
            s->t_last = t_end;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    hipStreamDestroy(st);
}


// The engine thread: admits queued queries into the resident launches' queues
// (express: at most EXPRESS_TASKS tasks), starts a queue's launch while it has
// queries in flight, ends it after STOP_IDLE without any, and collects results
// from the mailboxes.
void engine_main(kc_server *s)
{
    hipSetDevice(s->dev);
    kc::OpenEngine &e = *s->eng;
    // an idle launch ends after this long, so its waves free their CUs for the
    // workers' launches: a worker's persistent launch is sized for the
    // device's free slots and completes only when its last block has run, so
    // an express launch kept up for a second between queries held the heavy
    // queries on the workers back by that second (serving mix at 8
    // queries/s: p99 107 ms -> 12.6 s)
    const std::chrono::microseconds STOP_IDLE[kc::OpenEngine::NQ] = {std::chrono::microseconds(2000),
                                                                     std::chrono::microseconds(2000)};
    struct Flight {
        Ticket t;
        uint32_t seq;
        int queue;
    };
    std::unordered_map<uint32_t, Flight> inflight;  // by slot
    Clock::time_point last_active[kc::OpenEngine::NQ] = {Clock::now(), Clock::now()};
    uint64_t launch_epoch[kc::OpenEngine::NQ] = {0, 0};
    for (;;) {
        std::vector<Ticket> take[kc::OpenEngine::NQ];
        std::vector<kc::OpenQueryIn> in[kc::OpenEngine::NQ];
        bool any_running = false, stopping = false;
        for (int j = 0; j < kc::OpenEngine::NQ; ++j) any_running |= e.running(j);
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto has_work = [&] { return s->stop || !s->open_pending.empty(); };
            if (inflight.empty() && !any_running) s->cv_work.wait(lk, has_work);
            else s->cv_work.wait_for(lk, std::chrono::microseconds(20), has_work);
            if (s->stop && s->open_pending.empty() && inflight.empty()) break;
            stopping = s->stop;
            for (int j = 0; j < kc::OpenEngine::NQ; ++j)  // indices restart at 0 when a queue is idle
                if (!e.qs[j].inflight && e.qs[j].nseg && !e.running(j)) e.reset(j);
            size_t slots = e.free_slots.size();
            while (!s->open_pending.empty() && slots > 0) {
                const Ticket &t = s->open_pending.front();
                const Resident &r = s->graphs[t.gid];
                const int j = task_count(r.o, t.q) <= EXPRESS_TASKS ? 0 : 1;
                // a launch sees memory written before it began: a graph uploaded
                // since goes to the workers while that launch runs
                if (e.running(j) && !e.qs[j].stopped && r.epoch > launch_epoch[j]) {
                    s->pending.push_back(t);
                    s->open_pending.pop_front();
                    s->cv_work.notify_all();
                    continue;
                }
                if (e.room(j) <= take[j].size()) break;  // waits for the queue to drain and restart
                in[j].push_back(kc::OpenQueryIn{r.o, t.q, (uint32_t)(t.id % 0xFFFFFFFEull) + 1});
                take[j].push_back(t);
                s->open_pending.pop_front();
                --slots;
            }
        }
        for (int j = 0; j < kc::OpenEngine::NQ; ++j) {
            if (take[j].empty()) continue;
            // other queries on the GPU in a resident launch: this one joins them
            bool was_running = false;
            for (int i = 0; i < kc::OpenEngine::NQ; ++i) was_running |= e.running(i) && !e.qs[i].stopped;
            std::vector<uint32_t> slot;
            e.admit(j, in[j], slot);
            const auto now = Clock::now();
            last_active[j] = now;
            std::lock_guard<std::mutex> lk(s->mu);
            s->admissions++;
            s->admitted += take[j].size();
            if (was_running) s->joined_running += take[j].size();
            for (size_t i = 0; i < take[j].size(); ++i) {
                if (slot[i] == ~0u) {  // no task: the count is 0
                    Done d;
                    d.r.algo = take[j][i].q.algo;
                    d.r.gran = take[j][i].q.gran;
                    d.r.part_size = take[j][i].q.part_size;
                    d.latency_ms = std::chrono::duration<double, std::milli>(now - take[j][i].t_submit).count();
                    complete(s, take[j][i], d);
                    s->t_last = now;
                } else {
                    inflight[slot[i]] = Flight{take[j][i], in[j][i].seq, j};
                }
            }
            s->cv_done.notify_all();
        }
        // launches: start a queue's launch while it has work, end it when idle
        for (int j = 0; j < kc::OpenEngine::NQ; ++j) {
            const bool run = e.running(j);
            if (e.qs[j].inflight && !run) {
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    launch_epoch[j] = s->graph_epoch;
                }
                if (e.launch(j) == KC_OK) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->launches++;
                }
                last_active[j] = Clock::now();
            } else if (run && !e.qs[j].stopped && !e.qs[j].inflight &&
                       (stopping || Clock::now() - last_active[j] > STOP_IDLE[j])) {
                e.stop(j);
            }
        }
        // results
        std::vector<std::pair<Flight, kc::OpenMail>> fin;
        for (auto it = inflight.begin(); it != inflight.end();) {
            volatile kc::OpenMail *m = e.mail + it->first;
            if (m->seq == it->second.seq) {
                std::atomic_thread_fence(std::memory_order_acquire);
                kc::OpenMail r{m->count, m->ovf, m->seq};
                fin.emplace_back(it->second, r);
                e.release(it->first);
                e.qs[it->second.queue].inflight--;
                last_active[it->second.queue] = Clock::now();
                it = inflight.erase(it);
            } else {
                ++it;
            }
        }
        if (!fin.empty()) {
            const auto now = Clock::now();
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &f : fin) {
                Done d;
                d.r.count = f.second.count;
                d.r.overflow = f.second.ovf ? 1 : 0;
                d.r.algo = f.first.t.q.algo;
                d.r.gran = f.first.t.q.gran;
                d.r.part_size = f.first.t.q.part_size;
                d.rc = d.r.overflow ? KC_ERR_OVERFLOW : KC_OK;
                d.latency_ms = std::chrono::duration<double, std::milli>(now - f.first.t.t_submit).count();
                complete(s, f.first.t, d);
            }
            s->t_last = now;
            s->cv_done.notify_all();
        }
    }
    for (int j = 0; j < kc::OpenEngine::NQ; ++j)
        if (e.running(j) && !e.qs[j].stopped) e.stop(j);
}

}  // namespace

using namespace kc;

extern "C" {

int kc_server_create(const kc_server_config *cfg, kc_server **out)
{
    if (!out) return fail(KC_ERR_ARG, "kc_server_create: null argument");
    *out = nullptr;
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    kc_server *s = new kc_server;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.workers == 0) s->cfg.workers = 4;
    if (s->cfg.workers > 8) s->cfg.workers = 8;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 256;
    if (s->cfg.budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        hipMemGetInfo(&fr, &tot);
        s->cfg.budget_bytes = (uint64_t)(fr * 0.9);
    }
    s->dev = ctx->dev;
    // the resident launch's slot sums and binomial table come out of the budget
    const uint64_t eng_bytes = 1ull << 20;
    if (s->cfg.budget_bytes > (64ull << 20)) {
        auto e = std::make_unique<OpenEngine>();
        if ((rc = e->init(s->dev))) {
            delete s;
            return rc;
        }
        s->eng = std::move(e);
        s->resident_bytes += eng_bytes;
        s->engine = std::thread(engine_main, s);
    }
    for (uint32_t i = 0; i < s->cfg.workers; ++i) s->workers.emplace_back(worker_main, s);
    *out = s;
    return KC_OK;
}

int kc_server_add_graph(kc_server *s, const kc_graph *g, int orient_mode, uint32_t *gid)
{
    if (!s || !g || !gid) return fail(KC_ERR_ARG, "kc_server_add_graph: null argument");
    const uint64_t foot = oriented_bytes(g->csr.n, g->csr.col.size() / 2);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const uint64_t need =
            s->resident_bytes + s->held_scratch + s->inflight_scratch + foot + orient_transient(g);
        if (need > s->cfg.budget_bytes)
            return fail(KC_ERR_BUDGET, "graph needs %llu bytes, %llu of %llu in use",
                        (unsigned long long)(foot + orient_transient(g)),
                        (unsigned long long)(s->resident_bytes + s->inflight_scratch),
                        (unsigned long long)s->cfg.budget_bytes);
        s->resident_bytes += foot;  // reserve before the upload
    }
    kc_oriented *o = nullptr;
    int rc = kc_orient(g, orient_mode, &o);
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
        s->resident_bytes -= foot;
        return rc;
    }
    Resident r;
    r.o = o;
    r.bytes = foot;
    r.epoch = ++s->graph_epoch;
    *gid = s->next_gid++;
    s->graphs[*gid] = r;
    return KC_OK;
}

int kc_server_remove_graph(kc_server *s, uint32_t gid)
{
    if (!s) return fail(KC_ERR_ARG, "kc_server_remove_graph: null server");
    std::vector<std::pair<kc_oriented *, uint64_t>> frees;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->graphs.find(gid);
        if (it == s->graphs.end() || it->second.removed) return fail(KC_ERR_NOTFOUND, "no graph %u", gid);
        if (it->second.pending == 0) {
            frees.emplace_back(it->second.o, it->second.bytes);
            s->graphs.erase(it);
        } else {
            it->second.removed = true;  // freed by a worker after its last query
        }
    }
    // outside the lock: the free waits for the device, resident launches
    // included, which the engine thread ends while this thread waits
    if (!frees.empty()) free_graphs(s, frees);
    return KC_OK;
}

int kc_server_submit(kc_server *s, uint32_t gid, const kc_query *q, uint64_t *ticket)
{
    if (!s || !q || !ticket) return fail(KC_ERR_ARG, "kc_server_submit: null argument");
    int rc = check_query(q);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->graphs.find(gid);
    if (it == s->graphs.end() || it->second.removed) return fail(KC_ERR_NOTFOUND, "no graph %u", gid);
    Ticket t;
    t.id = s->next_ticket++;
    t.gid = gid;
    t.q = tune_query(it->second.o, q);
    t.t_submit = Clock::now();
    if (t.q.k <= 2) {  // closed forms, no device work
        Done d;
        d.r.count = t.q.k == 1 ? it->second.o->n : it->second.o->m;
        d.r.algo = t.q.algo;
        d.r.gran = t.q.gran;
        d.r.part_size = t.q.part_size;
        s->done[t.id] = d;
        s->lat.push_back(0.0);
        s->submitted++;
        s->completed++;
        s->st_completed++;
        *ticket = t.id;
        return KC_OK;
    }
    t.scratch = query_scratch(it->second.o, t.q);
    if (s->resident_bytes + t.scratch > s->cfg.budget_bytes)
        return fail(KC_ERR_BUDGET, "query scratch %llu bytes does not fit next to the resident graphs",
                    (unsigned long long)t.scratch);
    it->second.pending++;
    if (!s->any) {
        s->t_first = t.t_submit;
        s->any = true;
    }
    if (open_eligible(s, it->second.o, t.q)) s->open_pending.push_back(t);
    else s->pending.push_back(t);
    s->submitted++;
    *ticket = t.id;
    s->cv_work.notify_all();
    return KC_OK;
}

int kc_server_poll(kc_server *s, uint64_t ticket, kc_result *r, int *done)
{
    if (!s || !r || !done) return fail(KC_ERR_ARG, "kc_server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(ticket);
    if (it == s->done.end()) {
        if (ticket == 0 || ticket >= s->next_ticket) return fail(KC_ERR_NOTFOUND, "no ticket %llu", (unsigned long long)ticket);
        *done = 0;
        return KC_OK;
    }
    *done = 1;
    *r = it->second.r;
    const int rc = it->second.rc;
    s->done.erase(it);
    return rc == KC_OK ? KC_OK : fail(rc, "query %llu failed", (unsigned long long)ticket);
}

int kc_server_wait(kc_server *s, uint64_t ticket, kc_result *r)
{
    if (!s || !r) return fail(KC_ERR_ARG, "kc_server_wait: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket == 0 || ticket >= s->next_ticket) return fail(KC_ERR_NOTFOUND, "no ticket %llu", (unsigned long long)ticket);
    s->cv_done.wait(lk, [&] { return s->done.count(ticket) != 0; });
    auto it = s->done.find(ticket);
    *r = it->second.r;
    const int rc = it->second.rc;
    s->done.erase(it);
    return rc == KC_OK ? KC_OK : fail(rc, "query %llu failed", (unsigned long long)ticket);
}

int kc_server_drain(kc_server *s)
{
    if (!s) return fail(KC_ERR_ARG, "kc_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return s->pending.empty() && s->open_pending.empty() && s->completed == s->submitted;
    });
    return KC_OK;
}

int kc_server_get_stats(kc_server *s, kc_server_stats *st)
{
    if (!s || !st) return fail(KC_ERR_ARG, "kc_server_get_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->st_completed;
    st->batches = s->batches + s->admissions;
    st->mean_batch = st->batches ? (double)(s->batched_queries + s->admitted) / st->batches : 0.0;
    st->resident_admitted = s->admitted;
    st->joined_running = s->joined_running;
    st->resident_launches = s->launches;
    st->graphs = s->graphs.size();
    st->resident_bytes = s->resident_bytes;
    std::vector<double> l = s->lat;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : l) sum += x;
        st->mean_ms = sum / l.size();
    }
    if (s->any && s->st_completed) {
        const double span = std::chrono::duration<double>(s->t_last - s->t_first).count();
        st->queries_per_s = span > 0 ? s->st_completed / span : 0.0;
    }
    st->held_scratch_bytes = s->held_scratch;
    return KC_OK;
}

int kc_server_reset_stats(kc_server *s)
{
    if (!s) return fail(KC_ERR_ARG, "kc_server_reset_stats: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->st_completed = 0;
    s->batches = 0;
    s->batched_queries = 0;
    s->admissions = s->admitted = s->joined_running = s->launches = 0;
    s->any = false;
    return KC_OK;
}

void kc_server_destroy(kc_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    if (s->engine.joinable()) s->engine.join();
    s->eng.reset();  // waits for the resident launch to end
    for (auto &f : s->free_later) kc_oriented_free(f.first);
    for (auto &kv : s->graphs) kc_oriented_free(kv.second.o);
    delete s;
}

}  // extern "C"
