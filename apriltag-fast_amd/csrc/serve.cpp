// The serving loop: frames from many camera streams, admitted as they
// arrive into batches that overlap on the device.
//
// Each stream has a bounded queue of waiting frames (host memory). A
// worker thread owns the device: it releases fed arrivals at their times,
// collects finished batches, and whenever a lane (a kernel chain on its
// own stream, with its own fixed arena) is idle, forms the next batch from
// the oldest waiting frames while the lane's arena holds them (and, if
// the configuration caps them, while a stream has fewer frames in flight
// than its cap) and launches the chain, which copies each frame into the
// batch's arena (by DMA from page-locked memory). A frame's device memory
// is its batch's: frames in flight are bounded by the lanes' arenas only.
// With several lanes, frames that arrive while batches run are admitted as
// soon as a lane frees up, and their work overlaps the running batches on
// the device: continuous batching. The budget covers the lanes' arenas
// and result buffers.
//
// A frame's detections do not depend on the batch it runs in: each frame
// has its own arena region and every kernel indexes by frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "atg.h"
#include "chain.h"

using namespace atg;

namespace {

constexpr int kMaxLanes = 8;
constexpr int kQueuePerStream = 64;  // waiting frames per stream before drops

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Req {
    int32_t stream;
    const uint8_t *px;
    int64_t stride;
    uint64_t tag;
    int64_t arrival, admit;
};

struct Lane {
    atg_ctx *ctx = nullptr;
    bool busy = false;
    std::vector<Req> batch;
    size_t cap = 0;  // arena bytes
};

struct StreamS {
    int32_t w, h;
    size_t frame_bytes;  // arena bytes of one frame (its pixels included)
    int queued = 0;      // waiting or in flight
    int inflight = 0;    // in a launched batch
};

struct Done {
    atg_served rec;
    std::vector<at_detection> dets;
};

}  // namespace

struct atg_server {
    atg_server_config cfg;
    atg_params params;
    std::vector<const at_family *> fams;
    std::vector<int32_t> bits;
    std::string err;
    int64_t t0 = 0;
    std::vector<StreamS> streams;
    std::vector<Lane> lanes;
    atg_ctx *spill = nullptr;  // frames that need more point capacity, one at a time
    bool started = false;
    // shared with the worker
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req> pending;                 // arrived, waiting for a lane
    std::vector<atg_arrival> feed;           // not yet arrived, by time
    size_t feed_next = 0;
    int64_t feed_t0 = 0;
    std::deque<Done> done;
    std::vector<Done> polled;                // the last poll's records (their dets)
    int64_t submitted = 0, dropped = 0, completed = 0, batches = 0, max_batch_seen = 0, inflight = 0;
    std::atomic<bool> stop{false};
    std::thread worker;
};

namespace {

int sfail(atg_server *s, int code, const std::string &m) {
    s->err = m;
    return code;
}

void finish(atg_server *s, const Req &r, int status, int32_t nb, std::vector<at_detection> &&dets, int64_t nquads) {
    Done d;
    d.rec.stream = r.stream;
    d.rec.status = status;
    d.rec.tag = r.tag;
    d.rec.arrival_ns = r.arrival - s->t0;
    d.rec.admit_ns = r.admit - s->t0;
    d.rec.done_ns = now_ns() - s->t0;
    d.rec.nquads = nquads;
    d.rec.batch_frames = nb;
    d.dets = std::move(dets);
    std::lock_guard<std::mutex> g(s->mu);
    s->streams[(size_t)r.stream].inflight--;
    s->streams[(size_t)r.stream].queued--;
    s->done.push_back(std::move(d));
    s->completed++;
    s->inflight--;
    s->cv.notify_all();
}

void collect(atg_server *s, Lane &L) {
    int rc = chain_collect(L.ctx);
    const int32_t nb = (int32_t)L.batch.size();
    for (int32_t i = 0; i < nb; i++) {
        const Req &r = L.batch[(size_t)i];
        int st = rc < 0 ? rc : ctx_frame_status(L.ctx, i);
        if (st == 1) {
            // more points than the lane's capacity: this frame alone, with
            // the worst-case capacity
            const StreamS &S = s->streams[(size_t)r.stream];
            atg_result res;
            st = atg_detect(s->spill, &s->params, r.px, S.w, S.h, r.stride, 0, &res);
            std::vector<at_detection> v;
            if (st == ATG_OK) v.assign(res.dets, res.dets + res.ndets);
            finish(s, r, st, nb, std::move(v), st == ATG_OK ? res.nquads : 0);
            continue;
        }
        std::vector<at_detection> v;
        int64_t nq = 0;
        if (st == ATG_OK) {
            v = ctx_dets(L.ctx, i);
            nq = ctx_nquads(L.ctx, i);
        }
        finish(s, r, st, nb, std::move(v), nq);
    }
    L.batch.clear();
    L.busy = false;
}

// The next batch for lane L: the oldest waiting frames while the lane's
// arena holds them (and their stream is under its in-flight cap, if any).
bool launch(atg_server *s, Lane &L) {
    std::vector<Req> take;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->pending.empty()) return false;
        size_t bytes = 0;
        const int cap = s->cfg.slots_per_stream;
        for (auto it = s->pending.begin(); it != s->pending.end() && (int32_t)take.size() < s->cfg.max_batch;) {
            StreamS &S = s->streams[(size_t)it->stream];
            if ((cap > 0 && S.inflight >= cap) || bytes + S.frame_bytes > L.cap) {
                ++it;
                continue;
            }
            S.inflight++;
            bytes += S.frame_bytes;
            Req r = *it;
            r.admit = now_ns();
            take.push_back(r);
            it = s->pending.erase(it);
        }
        s->inflight += (int64_t)take.size();
    }
    if (take.empty()) return false;
    std::vector<atg_frame> fr(take.size());
    for (size_t i = 0; i < take.size(); i++) {
        const StreamS &S = s->streams[(size_t)take[i].stream];
        fr[i] = atg_frame{take[i].px, S.w, S.h, take[i].stride};
    }
    const int rc = chain_enqueue(L.ctx, &s->params, fr.data(), (int32_t)fr.size(), 0);
    L.batch = std::move(take);
    if (rc != ATG_OK) {
        for (const Req &r : L.batch) finish(s, r, rc, 0, {}, 0);
        L.batch.clear();
        return true;
    }
    L.busy = true;
    std::lock_guard<std::mutex> g(s->mu);
    s->batches++;
    s->max_batch_seen = std::max<int64_t>(s->max_batch_seen, (int64_t)L.batch.size());
    return true;
}

void release_arrivals(atg_server *s) {
    std::lock_guard<std::mutex> g(s->mu);
    int64_t t = now_ns();
    while (s->feed_next < s->feed.size() && s->feed_t0 + s->feed[s->feed_next].at_ns <= t) {
        const atg_arrival &a = s->feed[s->feed_next++];
        StreamS &S = s->streams[(size_t)a.stream];
        s->submitted++;
        if (S.queued >= kQueuePerStream) {
            // dropped: its record comes back from poll with ATG_ERR_LIMIT
            s->dropped++;
            Done d;
            d.rec = atg_served{};
            d.rec.stream = a.stream;
            d.rec.status = ATG_ERR_LIMIT;
            d.rec.tag = a.tag;
            d.rec.arrival_ns = s->feed_t0 + a.at_ns - s->t0;
            d.rec.admit_ns = d.rec.done_ns = t - s->t0;
            s->done.push_back(std::move(d));
            continue;
        }
        S.queued++;
        s->pending.push_back(Req{a.stream, a.pixels, a.stride, a.tag, s->feed_t0 + a.at_ns, 0});
    }
}

void work(atg_server *s) {
    hipSetDevice(s->cfg.device);
    for (;;) {
        bool did = false;
        release_arrivals(s);
        for (auto &L : s->lanes)
            if (L.busy && chain_done(L.ctx)) {
                collect(s, L);
                did = true;
            }
        for (auto &L : s->lanes)
            if (!L.busy && launch(s, L)) did = true;
        bool idle;
        {
            std::unique_lock<std::mutex> g(s->mu);
            idle = s->pending.empty() && s->feed_next == s->feed.size() && s->inflight == 0;
            if (s->stop && idle) break;
            if (!did) {
                // nothing to do now: wake on a submit, or soon for a
                // finishing batch or the next fed arrival
                int64_t wait_us = 20;
                if (idle && s->feed_next == s->feed.size()) wait_us = 2000;
                s->cv.wait_for(g, std::chrono::microseconds(wait_us));
            }
        }
    }
}

int start(atg_server *s) {
    if (s->started) return ATG_OK;
    if (s->streams.empty()) return sfail(s, ATG_ERR_ARG, "no streams");
    for (int i = 0; i < s->cfg.lanes; i++) {
        Lane L;
        int rc = atg_ctx_create(s->cfg.device, &L.ctx);
        if (rc) return sfail(s, rc, "lane context");
        s->lanes.push_back(L);
        atg_set_graphs(L.ctx, 0);  // batches differ from one launch to the next
        rc = atg_set_families(L.ctx, s->fams.data(), s->bits.data(), (int32_t)s->fams.size());
        if (rc) return sfail(s, rc, std::string("lane families: ") + atg_last_error(L.ctx));
    }
    // a frame's arena bytes with its pixels, by stream (the same for every
    // frame of it; the sizing reads no pixel)
    static const uint8_t kSizing = 0;
    for (auto &S : s->streams) {
        atg_frame f{&kSizing, S.w, S.h, S.w};
        S.frame_bytes = batch_arena_bytes(s->lanes[0].ctx, &s->params, &f, 1, 0);
    }
    size_t fixed = (size_t)s->cfg.lanes * ctx_result_bytes(s->cfg.max_batch);
    if (fixed >= s->cfg.budget_bytes) return sfail(s, ATG_ERR_LIMIT, "the lanes' result buffers fill the budget");
    size_t cap = ((s->cfg.budget_bytes - fixed) / (size_t)s->cfg.lanes) & ~(size_t)255;
    for (auto &S : s->streams)
        if (S.frame_bytes > cap) return sfail(s, ATG_ERR_LIMIT, "a stream's frame does not fit a lane's budget");
    for (auto &L : s->lanes) {
        int rc = ctx_reserve(L.ctx, cap, s->cfg.max_batch);
        if (rc) return sfail(s, rc, std::string("lane arena: ") + atg_last_error(L.ctx));
        L.cap = cap;
    }
    int rc = atg_ctx_create(s->cfg.device, &s->spill);
    if (!rc) rc = atg_set_families(s->spill, s->fams.data(), s->bits.data(), (int32_t)s->fams.size());
    if (rc) return sfail(s, rc, "spill context");
    s->started = true;
    s->worker = std::thread(work, s);
    return ATG_OK;
}

}  // namespace

extern "C" void atg_server_config_default(atg_server_config *c) {
    c->device = 0;
    c->budget_bytes = (size_t)16 << 30;
    c->lanes = 3;
    c->max_batch = 128;
    c->slots_per_stream = 0;  // no cap on a stream's frames in flight
}

extern "C" int atg_server_create(const atg_server_config *c, const atg_params *p, const at_family *const *fams,
                                 const int32_t *bits, int32_t nfam, atg_server **out) {
    *out = nullptr;
    if (!c || !p || nfam < 1 || nfam > 16 || c->lanes < 1 || c->lanes > kMaxLanes || c->max_batch < 1 ||
        c->slots_per_stream < 0)
        return ATG_ERR_ARG;
    atg_server *s = new atg_server();
    s->cfg = *c;
    s->params = *p;
    s->fams.assign(fams, fams + nfam);
    s->bits.assign(bits, bits + nfam);
    s->t0 = now_ns();
    *out = s;
    return ATG_OK;
}

extern "C" void atg_server_destroy(atg_server *s) {
    if (!s) return;
    if (s->started) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
            s->feed_next = s->feed.size();  // arrivals not yet due are cancelled
            s->cv.notify_all();
        }
        s->worker.join();
    }
    for (auto &L : s->lanes) atg_ctx_destroy(L.ctx);
    atg_ctx_destroy(s->spill);
    delete s;
}

extern "C" const char *atg_server_error(const atg_server *s) { return s ? s->err.c_str() : "no server"; }

extern "C" int64_t atg_server_now_ns(const atg_server *s) { return now_ns() - s->t0; }

extern "C" int atg_server_add_stream(atg_server *s, int32_t width, int32_t height, int32_t *stream_id) {
    if (!s || width < 1 || height < 1 || !stream_id) return ATG_ERR_ARG;
    if (s->started) return sfail(s, ATG_ERR_ARG, "streams are added before the first frame");
    StreamS S;
    S.w = width;
    S.h = height;
    S.frame_bytes = 0;  // known once the lanes exist
    *stream_id = (int32_t)s->streams.size();
    s->streams.push_back(std::move(S));
    return ATG_OK;
}

extern "C" int atg_server_submit(atg_server *s, int32_t stream, const uint8_t *pixels, int64_t stride, uint64_t tag) {
    if (!s || stream < 0 || (size_t)stream >= s->streams.size() || !pixels || stride < s->streams[(size_t)stream].w)
        return ATG_ERR_ARG;
    int rc = start(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    StreamS &S = s->streams[(size_t)stream];
    s->submitted++;
    if (S.queued >= kQueuePerStream) {
        s->dropped++;
        return ATG_ERR_LIMIT;
    }
    S.queued++;
    s->pending.push_back(Req{stream, pixels, stride, tag, now_ns(), 0});
    s->cv.notify_all();
    return ATG_OK;
}

extern "C" int atg_server_feed(atg_server *s, const atg_arrival *arrivals, int64_t n) {
    if (!s || n < 0 || (n > 0 && !arrivals)) return ATG_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        const atg_arrival &a = arrivals[i];
        if (a.stream < 0 || (size_t)a.stream >= s->streams.size() || !a.pixels ||
            a.stride < s->streams[(size_t)a.stream].w || a.at_ns < 0)
            return ATG_ERR_ARG;
    }
    int rc = start(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->feed_next < s->feed.size()) return sfail(s, ATG_ERR_ARG, "a feed is still running");
    s->feed.assign(arrivals, arrivals + n);
    std::stable_sort(s->feed.begin(), s->feed.end(),
                     [](const atg_arrival &a, const atg_arrival &b) { return a.at_ns < b.at_ns; });
    s->feed_next = 0;
    s->feed_t0 = now_ns();
    s->cv.notify_all();
    return ATG_OK;
}

extern "C" int atg_server_poll(atg_server *s, atg_served *out, int32_t cap) {
    if (!s || cap < 0 || (cap > 0 && !out)) return ATG_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    s->polled.clear();
    while (!s->done.empty() && (int32_t)s->polled.size() < cap) {
        s->polled.push_back(std::move(s->done.front()));
        s->done.pop_front();
    }
    for (size_t i = 0; i < s->polled.size(); i++) {
        Done &d = s->polled[i];
        d.rec.dets = d.dets.data();
        d.rec.ndets = (int64_t)d.dets.size();
        out[i] = d.rec;
    }
    return (int)s->polled.size();
}

extern "C" int atg_server_drain(atg_server *s, int64_t timeout_ms) {
    if (!s) return ATG_ERR_ARG;
    if (!s->started) return ATG_OK;
    std::unique_lock<std::mutex> g(s->mu);
    bool ok = s->cv.wait_for(g, std::chrono::milliseconds(timeout_ms), [s] {
        return s->pending.empty() && s->feed_next == s->feed.size() && s->inflight == 0;
    });
    return ok ? ATG_OK : sfail(s, ATG_ERR_LIMIT, "drain timed out");
}

extern "C" int atg_server_stats(atg_server *s, atg_server_counts *out) {
    if (!s || !out) return ATG_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    out->submitted = s->submitted;
    out->dropped = s->dropped;
    out->completed = s->completed;
    out->batches = s->batches;
    out->max_batch_seen = s->max_batch_seen;
    out->lanes = (int32_t)s->lanes.size();
    out->lane_arena_bytes = s->lanes.empty() ? 0 : s->lanes[0].cap;
    out->slot_bytes = 0;  // (frames live in the lanes' arenas)
    return ATG_OK;
}

extern "C" int atg_host_alloc(size_t bytes, void **out) {
    if (!out) return ATG_ERR_ARG;
    return hipHostMalloc(out, bytes) == hipSuccess ? ATG_OK : ATG_ERR_HIP;
}

extern "C" int atg_host_free(void *p) { return hipHostFree(p) == hipSuccess ? ATG_OK : ATG_ERR_HIP; }
