// Request server with continuous batching.
//
// The reference spreads frequencies over MPI ranks that time-share one GPU
// ("wersja gpu/Wire_MoM.cu":331-339).  Here one process keeps a queue of
// requests (a geometry, its frequencies and feeds) and drives the device in
// steps.  Each step:
//   1. admits queued requests in arrival order while their footprint fits
//      the HBM budget (a request that can never fit is failed at submit);
//   2. packs pending frequencies of admitted requests that share the padded
//      size of the oldest one into one batch, oldest request first (and each
//      request's frequencies highest first), as many as the budget and the
//      in-flight cap allow; in mixed precision, frequencies that the refinement
//      rates already seen on their structure predict not to refine are
//      factored in FP64 directly;
//   3. fills, factors and solves that batch (solve_items) and retires the
//      requests whose last frequency finished.
// A request that arrives during a step joins at the next one, next to the
// frequencies still pending from earlier requests.  Every item's arithmetic
// depends only on that item, so a result is the same alone and in a mix.
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "wm_internal.h"

namespace {

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
    int done = 0;
    int status = WM_OK;
    double t_submit = 0, t_done = 0;
};

}  // namespace

struct wm_server {
    wm_ctx *ctx = nullptr;
    wm_server_opts o;
    size_t budget = 0;
    std::mutex qmu;  // guards queue, completed and next_id
    int64_t next_id = 0;
    std::deque<std::unique_ptr<Req>> queue;
    std::vector<std::unique_ptr<Req>> running;  // driven only by wm_server_step
    std::map<int64_t, std::unique_ptr<Req>> completed;
    size_t hbm_used = 0;
};

using namespace wm;

extern "C" {

void wm_server_opts_default(wm_server_opts *o)
{
    if (!o) return;
    o->hbm_budget = 0;
    o->max_inflight = 0;
    o->prec = WM_PREC_F32;
    o->streams = 1;
    o->size_policy = 1;
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
    if (s->o.hbm_budget == 0) {
        size_t freeb = 0, total = 0;
        (void)hipSetDevice(c->device);
        if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = c->hbm_bytes / 2;
        s->budget = (size_t)(0.8 * (double)freeb);
    } else {
        s->budget = s->o.hbm_budget;
    }
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
    if (item_bytes(g->ld, nrhs, s->o.prec) > s->budget) {
        set_error("wm_server_submit: one frequency of N=%d needs %zu bytes, budget is %zu", g->npls,
                  item_bytes(g->ld, nrhs, s->o.prec), s->budget);
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
    std::lock_guard<std::mutex> lk(s->qmu);
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(std::move(r));
    return WM_OK;
}

int wm_server_step(wm_server *s, int wait, int *done)
{
    if (!s) return WM_ERR_ARG;
    if (done) *done = 0;
    wm_ctx *c = s->ctx;
    // 1. admission: a request is admitted while the budget holds one of its
    //    frequencies next to those of the requests already admitted
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
    if (s->running.empty()) return WM_OK;
    // 2. batch: pending frequencies of requests with the oldest request's size
    int ld = -1;
    for (auto &r : s->running)
        if (r->next < (int)r->freqs.size()) { ld = r->g->ld; break; }
    if (ld < 0) return WM_OK;
    int nr = 1;
    for (auto &r : s->running)
        if (r->g->ld == ld && r->next < (int)r->freqs.size()) nr = std::max(nr, r->nrhs);
    const size_t per = item_bytes(ld, nr, s->o.prec);
    int cap = (int)std::max<size_t>(1, s->budget / per);
    cap = std::min(cap, 64);
    if (s->o.max_inflight > 0) cap = std::min(cap, s->o.max_inflight);
    std::vector<WorkItem> items;
    std::vector<Req *> owner;
    for (auto &r : s->running) {
        if (r->g->ld != ld) continue;
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
    std::vector<WorkItem> sorted;
    std::vector<Req *> sorted_owner;
    for (int i : order) { sorted.push_back(items[i]); sorted_owner.push_back(owner[i]); }
    wm_solve_opts o;
    wm_solve_opts_default(&o);
    o.prec = s->o.prec;
    int rc;
    {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        // mixed precision: structures solved before carry refinement rates that
        // route frequencies which cannot refine to FP64 factors at once
        if (o.prec != WM_PREC_F64) {
            if (s->o.size_policy && prefer_f64(c, o.prec, ld, (int)sorted.size()))
                for (auto &w : sorted) w.fprec = WM_PREC_F64;
            else
                route_items(o, sorted.data(), (int)sorted.size());
        }
        rc = solve_items(c, o, sorted.data(), (int)sorted.size(), c->stream);
        if (rc >= 0 && o.prec != WM_PREC_F64) record_rates(sorted.data(), (int)sorted.size());
    }
    if (rc < 0) return rc;
    // 3. retire
    const double t = now_s();
    for (size_t i = 0; i < sorted.size(); ++i) {
        Req *r = sorted_owner[i];
        r->done++;
        if (sorted[i].st && sorted[i].st->info != 0 && r->status == WM_OK) r->status = sorted[i].st->info;
    }
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
    if (done) *done = ndone;
    (void)wait;
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

int wm_server_destroy(wm_server *s)
{
    delete s;
    return WM_OK;
}

}  // extern "C"
