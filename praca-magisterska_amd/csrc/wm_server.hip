// Request server with continuous batching and device overlap.
//
// The reference spreads frequencies over MPI ranks that time-share one GPU
// ("wersja gpu/Wire_MoM.cu":331-339).  Here one process keeps a queue of
// requests (a geometry, its frequencies and feeds) and serves them with
// `lanes`: each lane is a worker thread with its own device context (stream
// and workspaces), so the batches of two lanes run on the device at the same
// time -- a small request that arrives while a large batch is factored is
// solved beside it instead of after it.  Each wm_server_step (non-blocking):
//   1. retires the batches the lanes have finished: each frequency's result is
//      already in its request; a batch that failed (negative status) completes
//      every request it belonged to with that status;
//   2. admits queued requests in arrival order while their footprint fits the
//      HBM budget (a request that can never fit is refused at submit);
//   3. hands a batch to every idle lane: pending frequencies of admitted
//      requests whose padded size is that of the oldest one with pending work
//      or down to half of it (uneven sizes share a batch, the smaller matrices
//      padded with the identity), oldest request first (each request's
//      frequencies highest first),
//      as many as the lane's share of the budget and the in-flight cap allow;
//      in mixed precision, frequencies that the refinement rates already seen
//      on their structure predict not to refine are factored in FP64 directly;
//   4. completes the requests whose last frequency is done.
// With wait set it blocks until a request completes, a request is submitted,
// or nothing is pending.  Every item's arithmetic depends only on that item,
// so a result is the same alone and in a mix, whichever lane solves it.
//
// The lanes' panel launches share the device's resident workgroups
// (wm_ctx::panel_share), so every panel cluster of every concurrent launch is
// co-resident and their cross-workgroup waits always complete.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "wm_internal.h"

namespace {

// Frequencies per lane batch at most: a batch is retired as a whole, so the
// latency of its requests grows with it (32 frequencies of N = 8192 in mixed
// precision hold about 52 GB and take about half a second).
constexpr int kLaneBatch = 32;

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    int64_t id = 0;
    const wm_geom *g = nullptr;
    std::vector<float> freqs;
    std::vector<int> gaps;
    float ampl = 1.f, phas = 0.f;
    int nrhs = 0;
    std::vector<wm_c128> I, zin;
    std::vector<wm_freq_stats> st;
    std::vector<int> order;  // frequencies highest first (refinement rates then route lower ones)
    int next = 0;            // first entry of order not yet scheduled
    int done = 0;            // frequencies finished (or failed)
    int status = WM_OK;
    std::string err;         // message of the first failed batch
    double t_submit = 0, t_done = 0;
};

// A worker thread with its own context; it solves one batch at a time.
struct Lane {
    wm_ctx *ctx = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    bool has_job = false, finished = false, quit = false;
    std::vector<wm::WorkItem> items;
    std::vector<Req *> owner;
    wm_solve_opts o{};
    int rc = WM_OK;
    std::string err;
    bool inject = false;    // fail the current batch before any launch (tests of the error path)
    double t0 = 0, t1 = 0;  // host time the current batch started and finished
};

}  // namespace

struct wm_server {
    wm_ctx *ctx = nullptr;
    wm_server_opts o;
    size_t budget = 0;
    std::mutex qmu;  // guards queue, completed, next_id and `event`
    std::condition_variable qcv;
    bool event = false;  // a submission or a finished batch since the stepping thread last looked
    int64_t next_id = 0;
    std::deque<std::unique_ptr<Req>> queue;
    std::vector<std::unique_ptr<Req>> running;  // driven only by wm_server_step
    std::map<int64_t, std::unique_ptr<Req>> completed;
    size_t hbm_used = 0;
    std::vector<std::unique_ptr<Lane>> lanes;
    // busy intervals of the lanes' batches (host time), for wm_server_overlap
    std::vector<std::pair<double, double>> busy;
    int64_t batches = 0;
    int inject_failures = 0;  // batches still to fail on purpose (wmx_server_inject_failure)
};

using namespace wm;

namespace {

// Solves a lane's batch; when the device cannot hold its workspace (the other
// lanes' batches hold theirs) the workspace is released and the batch is
// solved in halves.
int solve_split(wm_ctx *c, const wm_solve_opts &o, WorkItem *items, int count)
{
    int rc = solve_items(c, o, items, count, c->stream);
    if (rc == WM_ERR_NOMEM && count > 1) {
        c->ws_sweep.release();
        const int h = count / 2;
        rc = solve_split(c, o, items, h);
        if (rc >= 0) {
            const int r2 = solve_split(c, o, items + h, count - h);
            if (r2 != WM_OK && (rc == WM_OK || r2 < 0)) rc = r2;
        }
        return rc;
    }
    if (rc >= 0 && o.prec != WM_PREC_F64) record_rates(items, count);
    return rc;
}

void lane_main(wm_server *s, Lane *L)
{
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(L->mu);
            L->cv.wait(lk, [&] { return L->has_job || L->quit; });
            if (L->quit) return;
        }
        int rc;
        if (L->inject) {
            set_error("injected batch failure");
            rc = WM_ERR_ARG;
        } else {
            std::lock_guard<std::recursive_mutex> lk(L->ctx->mu);
            (void)hipSetDevice(L->ctx->device);
            rc = solve_split(L->ctx, L->o, L->items.data(), (int)L->items.size());
        }
        {
            std::lock_guard<std::mutex> lk(L->mu);
            L->rc = rc;
            L->err = rc < 0 ? wm_last_error() : "";
            L->t1 = now_s();
            L->has_job = false;
            L->finished = true;
        }
        {
            std::lock_guard<std::mutex> lk(s->qmu);
            s->event = true;
        }
        s->qcv.notify_all();
    }
}

// Retires finished batches; returns the number of lanes that finished.
int retire(wm_server *s)
{
    int n = 0;
    for (auto &lp : s->lanes) {
        Lane *L = lp.get();
        std::lock_guard<std::mutex> lk(L->mu);
        if (!L->finished) continue;
        L->finished = false;
        ++n;
        s->busy.emplace_back(L->t0, L->t1);
        for (size_t i = 0; i < L->items.size(); ++i) {
            Req *r = L->owner[i];
            r->done++;
            if (L->rc < 0) {
                if (r->status == WM_OK) {
                    r->status = L->rc;
                    r->err = L->err;
                }
            } else if (L->items[i].st && L->items[i].st->info != 0 && r->status == WM_OK) {
                r->status = L->items[i].st->info;
            }
        }
        if (L->rc < 0) set_error("server batch failed: %s", L->err.c_str());
        L->items.clear();
        L->owner.clear();
    }
    return n;
}

// Moves requests whose frequencies are all done to `completed`.
int complete(wm_server *s)
{
    const double t = now_s();
    int ndone = 0;
    std::lock_guard<std::mutex> lk(s->qmu);
    for (auto it = s->running.begin(); it != s->running.end();) {
        Req *r = it->get();
        if (r->done == (int)r->freqs.size()) {
            r->t_done = t;
            s->hbm_used -= item_bytes(r->g->ld, r->nrhs, s->o.prec);
            s->completed[r->id] = std::move(*it);
            it = s->running.erase(it);
            ++ndone;
        } else {
            ++it;
        }
    }
    return ndone;
}

void admit(wm_server *s)
{
    std::lock_guard<std::mutex> lk(s->qmu);
    while (!s->queue.empty()) {
        Req *r = s->queue.front().get();
        const size_t need = item_bytes(r->g->ld, r->nrhs, s->o.prec);
        if (s->hbm_used + need > s->budget && !s->running.empty()) break;
        s->hbm_used += need;
        s->running.push_back(std::move(s->queue.front()));
        s->queue.pop_front();
    }
}

// Forms a batch for an idle lane; false if no frequency is pending.
bool dispatch(wm_server *s, Lane *L)
{
    // the batch's padded size is the oldest pending request's; requests down
    // to half that size join it (their matrices padded with the identity, the
    // factorization skipping the padding), larger ones wait for another batch
    int ld = -1;
    for (auto &r : s->running)
        if (r->next < (int)r->freqs.size()) { ld = r->g->ld; break; }
    if (ld < 0) return false;
    auto joins = [&](const Req *r) { return r->g->ld <= ld && 2 * r->g->ld > ld; };
    int nr = 1;
    for (auto &r : s->running)
        if (joins(r.get()) && r->next < (int)r->freqs.size()) nr = std::max(nr, r->nrhs);
    const size_t per = item_bytes(ld, nr, s->o.prec);
    int cap = (int)std::max<size_t>(1, s->budget / s->lanes.size() / per);
    cap = std::min(cap, kLaneBatch);
    if (s->o.max_inflight > 0) cap = std::min(cap, s->o.max_inflight);
    std::vector<WorkItem> items;
    std::vector<Req *> owner;
    for (auto &r : s->running) {
        if (!joins(r.get())) continue;
        while (r->next < (int)r->freqs.size() && (int)items.size() < cap) {
            const int f = r->order[r->next++];
            WorkItem w;
            w.g = r->g;
            w.freq = r->freqs[f];
            w.nrhs = r->nrhs;
            w.gaps = r->gaps.data();
            w.ampl = r->ampl;
            w.phas = r->phas;
            w.I = r->I.data() + (size_t)f * r->nrhs * r->g->npls;
            w.zin = r->zin.data() + (size_t)f * r->nrhs;
            w.st = r->st.data() + f;
            items.push_back(w);
            owner.push_back(r.get());
        }
        if ((int)items.size() >= cap) break;
    }
    // group items of one geometry so their fills share launches (stable: keeps age order)
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return items[a].g < items[b].g; });
    wm_solve_opts o;
    wm_solve_opts_default(&o);
    o.prec = s->o.prec;
    std::lock_guard<std::mutex> lk(L->mu);
    L->items.clear();
    L->owner.clear();
    for (int i : order) {
        L->items.push_back(items[i]);
        L->owner.push_back(owner[i]);
    }
    // mixed precision: structures solved before carry refinement rates that
    // route frequencies which cannot refine to FP64 factors at once
    if (o.prec != WM_PREC_F64) {
        if (s->o.size_policy && prefer_f64(L->ctx, o.prec, ld, (int)L->items.size()))
            for (auto &w : L->items) w.fprec = WM_PREC_F64;
        else
            route_items(o, L->items.data(), (int)L->items.size());
    }
    L->o = o;
    L->inject = s->inject_failures > 0;
    if (L->inject) --s->inject_failures;
    L->t0 = now_s();
    L->has_job = true;
    ++s->batches;
    L->cv.notify_one();
    return true;
}

}  // namespace

extern "C" {

void wm_server_opts_default(wm_server_opts *o)
{
    if (!o) return;
    o->hbm_budget = 0;
    o->max_inflight = 0;
    o->prec = WM_PREC_F32;
    o->streams = 2;
    o->size_policy = 1;
}

int wm_server_destroy(wm_server *s)
{
    if (!s) return WM_OK;
    for (auto &L : s->lanes) {
        {
            std::lock_guard<std::mutex> lk(L->mu);
            L->quit = true;
        }
        L->cv.notify_one();
    }
    for (auto &L : s->lanes) {
        if (L->th.joinable()) L->th.join();
        if (L->ctx) (void)wm_shutdown(L->ctx);
    }
    delete s;
    return WM_OK;
}

int wm_server_create(wm_ctx *c, const wm_server_opts *o, wm_server **out)
{
    if (!c || !out) {
        set_error("wm_server_create: bad arguments");
        return WM_ERR_ARG;
    }
    auto *s = new wm_server();
    s->ctx = c;
    wm_server_opts_default(&s->o);
    if (o) s->o = *o;
    if (s->o.prec != WM_PREC_F64 && s->o.prec != WM_PREC_F32 && s->o.prec != WM_PREC_BF16X3) {
        delete s;
        set_error("wm_server_create: precision %d not available", o ? o->prec : -1);
        return WM_ERR_ARG;
    }
    if (s->o.streams <= 0) s->o.streams = 2;
    s->o.streams = std::min(s->o.streams, 4);
    if (s->o.hbm_budget == 0) {
        size_t freeb = 0, total = 0;
        (void)hipSetDevice(c->device);
        if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = c->hbm_bytes / 2;
        s->budget = (size_t)(0.8 * (double)freeb);
    } else {
        s->budget = s->o.hbm_budget;
    }
    for (int i = 0; i < s->o.streams; ++i) {
        auto L = std::make_unique<Lane>();
        const int rc = wm_init(c->device, &L->ctx);
        if (rc != WM_OK) {
            wm_server_destroy(s);
            return rc;
        }
        L->ctx->panel_share = s->o.streams;
        // Lanes launch their factorizations directly.  A stream capture in one
        // lane is invalidated by the other lanes' device-synchronizing calls
        // (allocation growth, frees, synchronous copies) made while it is open.
        L->ctx->use_graphs = false;
        for (int p = 0; p < 3; ++p)
            for (int k = 0; k < wm_ctx::kTuneSizes; ++k) L->ctx->getrf_block[p][k] = c->getrf_block[p][k];
        s->lanes.push_back(std::move(L));
    }
    for (auto &L : s->lanes) L->th = std::thread(lane_main, s, L.get());
    *out = s;
    return WM_OK;
}

int wm_server_submit(wm_server *s, const wm_geom *g, int nfreq, const float *freq_mhz, int nrhs, const int *gaps,
                     float ampl, float phas, int64_t *id)
{
    if (!s || !g || nfreq <= 0 || !freq_mhz || nrhs <= 0 || nrhs > 64 || !gaps || !id) {
        set_error("wm_server_submit: bad arguments");
        return WM_ERR_ARG;
    }
    for (int r = 0; r < nrhs; ++r)
        if (gaps[r] < 1 || gaps[r] > g->npls) {
            set_error("wm_server_submit: feed basis %d outside 1..%d", gaps[r], g->npls);
            return WM_ERR_ARG;
        }
    if (item_bytes(g->ld, nrhs, s->o.prec) > s->budget / s->lanes.size()) {
        set_error("wm_server_submit: one frequency of N=%d needs %zu bytes, a lane's budget is %zu", g->npls,
                  item_bytes(g->ld, nrhs, s->o.prec), s->budget / s->lanes.size());
        return WM_ERR_NOMEM;
    }
    auto r = std::make_unique<Req>();
    r->g = g;
    r->freqs.assign(freq_mhz, freq_mhz + nfreq);
    r->gaps.assign(gaps, gaps + nrhs);
    r->ampl = ampl;
    r->phas = phas;
    r->nrhs = nrhs;
    r->I.assign((size_t)nfreq * nrhs * g->npls, wm_c128{0, 0});
    r->zin.assign((size_t)nfreq * nrhs, wm_c128{0, 0});
    r->st.assign(nfreq, wm_freq_stats{});
    r->order.resize(nfreq);
    for (int f = 0; f < nfreq; ++f) r->order[f] = f;
    std::stable_sort(r->order.begin(), r->order.end(), [&](int a, int b) { return r->freqs[a] > r->freqs[b]; });
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->qmu);
        r->id = s->next_id++;
        *id = r->id;
        s->queue.push_back(std::move(r));
        s->event = true;
    }
    s->qcv.notify_all();
    return WM_OK;
}

int wm_server_step(wm_server *s, int wait, int *done)
{
    if (!s) return WM_ERR_ARG;
    if (done) *done = 0;
    int ndone = 0, status = WM_OK;
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(s->qmu);
            s->event = false;
        }
        retire(s);
        ndone += complete(s);
        admit(s);
        for (auto &L : s->lanes) {
            bool idle;
            {
                std::lock_guard<std::mutex> lk(L->mu);
                idle = !L->has_job && !L->finished;
            }
            if (idle && !dispatch(s, L.get())) break;
        }
        if (!wait || ndone > 0) break;
        // block until a lane finishes or a request arrives, unless nothing is pending
        bool busy = false;
        for (auto &L : s->lanes) {
            std::lock_guard<std::mutex> lk(L->mu);
            busy = busy || L->has_job || L->finished;
        }
        std::unique_lock<std::mutex> lk(s->qmu);
        if (!busy && s->queue.empty()) break;
        s->qcv.wait(lk, [&] { return s->event; });
    }
    if (done) *done = ndone;
    return status;
}

int wm_server_overlap(wm_server *s, int64_t *batches, double *busy_s, double *overlap_s)
{
    if (!s) return WM_ERR_ARG;
    // union of the busy intervals, and the time covered by two or more of them
    std::vector<std::pair<double, int>> ev;
    for (const auto &b : s->busy) {
        ev.emplace_back(b.first, +1);
        ev.emplace_back(b.second, -1);
    }
    std::sort(ev.begin(), ev.end());
    double busy = 0.0, over = 0.0, prev = 0.0;
    int depth = 0;
    for (const auto &e : ev) {
        if (depth >= 1) busy += e.first - prev;
        if (depth >= 2) over += e.first - prev;
        depth += e.second;
        prev = e.first;
    }
    if (batches) *batches = (int64_t)s->busy.size();
    if (busy_s) *busy_s = busy;
    if (overlap_s) *overlap_s = over;
    return WM_OK;
}

int wm_server_result(wm_server *s, int64_t id, wm_c128 *I, wm_c128 *zin, wm_freq_stats *st, double *latency_s)
{
    if (!s) return WM_ERR_ARG;
    std::unique_ptr<Req> r;
    {
        std::lock_guard<std::mutex> lk(s->qmu);
        auto it = s->completed.find(id);
        if (it == s->completed.end()) {
            set_error("wm_server_result: request %lld is not complete", (long long)id);
            return WM_ERR_STATE;
        }
        r = std::move(it->second);
        s->completed.erase(it);
    }
    if (I) std::copy(r->I.begin(), r->I.end(), I);
    if (zin) std::copy(r->zin.begin(), r->zin.end(), zin);
    if (st) std::copy(r->st.begin(), r->st.end(), st);
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    if (r->status < 0) set_error("request %lld failed: %s", (long long)id, r->err.c_str());
    return r->status;
}

int wm_server_done(wm_server *s, int64_t *ids, int cap, int *count)
{
    if (!s || !count || cap < 0 || (cap > 0 && !ids)) return WM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->qmu);
    int k = 0;
    for (auto it = s->completed.begin(); it != s->completed.end() && k < cap; ++it) ids[k++] = it->first;
    *count = k;
    return WM_OK;
}

int wm_server_status(wm_server *s, int *queued, int *running, int *completed, size_t *hbm_used)
{
    if (!s) return WM_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->qmu);
    if (queued) *queued = (int)s->queue.size();
    if (running) *running = (int)s->running.size();
    if (completed) *completed = (int)s->completed.size();
    if (hbm_used) *hbm_used = s->hbm_used;
    return WM_OK;
}

}  // extern "C"

// Internal test hook (not part of include/wiremom.h): the next `n` batches
// fail on the host before any launch, as a batch whose solve returned an
// error would, so the error path of the server can be tested.
extern "C" int wmx_server_inject_failure(wm_server *s, int n)
{
    if (!s || n < 0) return WM_ERR_ARG;
    s->inject_failures = n;
    return WM_OK;
}
