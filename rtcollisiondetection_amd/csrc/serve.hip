// Serving loop: continuous batching of discrete / linear / quadratic requests over resident scenes.
//
// RTCD has no serving layer; its closest analogue is the NSTREAM ThreadPool in
// CollisionDetector::check (RTCD/CollisionDetector/CollisionDetector.h:396-433, threadPool.h:10-92),
// which launches fixed batches on round-robin streams and joins every NSTREAM tasks.  Here a worker
// thread owns `slots` workspaces, each on its own stream.  Requests are admitted while a slot is
// free and their estimated device bytes fit the memory budget; a request that does not fit may be
// overtaken by smaller ones for at most kBypassMs, so small requests are not blocked behind large
// ones.  Admitted requests run concurrently on the device; completion is detected with events, the
// flags are copied to the caller's buffer, and submit->done latency is recorded.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "ctx.hpp"

namespace rtcd {

namespace {
constexpr double kBypassMs = 5.0;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n) return;
        if (p) (void) hipHostFree(p);
        p = nullptr;
        size_t c = count + count / 4 + 64;
        RTCD_HIP(hipHostMalloc((void**) &p, c * sizeof(T), hipHostMallocDefault));
        n = c;
    }
    ~PinnedBuf() {
        if (p) (void) hipHostFree(p);
    }
};

struct Req {
    uint64_t id = 0;
    rtcd_request r{};
    std::shared_ptr<SceneRes> scene;
    size_t bytes = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    rtcd_status status = RTCD_OK;
    std::string err;
    bool done = false;
    uint32_t n_wp_total = 0;
};

struct Slot {
    Workspace ws;
    hipEvent_t done_ev = nullptr;
    Req* cur = nullptr;
    PinnedBuf<float> in;
    PinnedBuf<uint32_t> offs;
    PinnedBuf<uint32_t> out;
};
}  // namespace

struct ServeState {
    rtcd_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread worker;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    size_t budget = 0, in_use = 0;
    std::vector<double> lat;
    uint64_t completed = 0, failed = 0;
    double t_first = 0, t_last = 0;
    size_t queued_max = 0;
};

namespace {

size_t request_bytes(const rtcd_ctx* c, const rtcd_request& r, uint32_t nw) {
    const size_t staging = (size_t) nw * c->robot.dof * 4 + (size_t) r.n * 4 * (r.kind == RTCD_REQ_QUADRATIC ? 1 : 1);
    if (r.kind == RTCD_REQ_DISCRETE) return discrete_bytes(c, r.n) + staging;
    if (r.kind == RTCD_REQ_LINEAR) return sweep_bytes(c, r.n, nw, 0) + staging;
    return sweep_bytes(c, r.n, nw, r.n_ctrl) + staging;
}

// Launch request q on slot sl (worker thread, no lock held).
void launch(ServeState* S, Slot& sl, Req* q) {
    rtcd_ctx* c = S->ctx;
    const rtcd_request& r = q->r;
    Workspace& ws = sl.ws;
    const hipStream_t s = ws.stream;
    const uint32_t dof = (uint32_t) c->robot.dof;
    const size_t nin = (size_t) q->n_wp_total * dof;
    sl.in.reserve(nin);
    std::memcpy(sl.in.p, r.input, nin * sizeof(float));
    ws.h_poses.reserve(nin);
    RTCD_HIP(hipMemcpyAsync(ws.h_poses.p, sl.in.p, nin * sizeof(float), hipMemcpyHostToDevice, s));
    ws.h_out.reserve(r.n);
    if (r.kind == RTCD_REQ_DISCRETE) {
        enqueue_discrete(c, ws, *q->scene, r.direction, ws.h_poses.p, r.n, ws.h_out.p, 0);
    } else if (r.kind == RTCD_REQ_LINEAR) {
        sl.offs.reserve(r.n + 1);
        std::memcpy(sl.offs.p, r.offsets, sizeof(uint32_t) * (r.n + 1));
        ws.h_offsets.reserve(r.n + 1);
        RTCD_HIP(hipMemcpyAsync(ws.h_offsets.p, sl.offs.p, sizeof(uint32_t) * (r.n + 1), hipMemcpyHostToDevice, s));
        enqueue_linear(c, ws, *q->scene, ws.h_poses.p, ws.h_offsets.p, r.n, q->n_wp_total, ws.h_out.p, nullptr, 0);
    } else {
        enqueue_quadratic(c, ws, *q->scene, ws.h_poses.p, r.n, r.n_wp, r.n_ctrl, nullptr, ws.h_out.p, nullptr, 0);
    }
    sl.out.reserve(r.n);
    RTCD_HIP(hipMemcpyAsync(sl.out.p, ws.h_out.p, sizeof(uint32_t) * r.n, hipMemcpyDeviceToHost, s));
    RTCD_HIP(hipEventRecord(sl.done_ev, s));
}

// Completion of the request on slot sl (its done event has fired).
void complete(ServeState* S, Slot& sl) {
    Req* q = sl.cur;
    rtcd_ctx* c = S->ctx;
    const rtcd_request& r = q->r;
    if (r.kind == RTCD_REQ_DISCRETE) {
        // re-runs synchronously on the slot stream if the candidate buffer overflowed
        std::vector<uint32_t> sv = sl.ws.scalars.download(2, sl.ws.stream);
        if (sv[1]) {
            finish_discrete(c, sl.ws, *q->scene, r.direction, sl.ws.h_poses.p, r.n, sl.ws.h_out.p, 0, nullptr);
            RTCD_HIP(hipMemcpyAsync(sl.out.p, sl.ws.h_out.p, sizeof(uint32_t) * r.n, hipMemcpyDeviceToHost,
                                    sl.ws.stream));
            RTCD_HIP(hipStreamSynchronize(sl.ws.stream));
        }
    } else {
        // re-runs synchronously if a candidate segment or the FP64 queue overflowed
        std::vector<uint32_t> sv = sl.ws.scalars.download(3, sl.ws.stream);
        if (sv[1] & 7u) {
            auto rerun = [&]() {
                if (r.kind == RTCD_REQ_LINEAR)
                    enqueue_linear(c, sl.ws, *q->scene, sl.ws.h_poses.p, sl.ws.h_offsets.p, r.n, q->n_wp_total,
                                   sl.ws.h_out.p, nullptr, 0);
                else
                    enqueue_quadratic(c, sl.ws, *q->scene, sl.ws.h_poses.p, r.n, r.n_wp, r.n_ctrl, nullptr, sl.ws.h_out.p,
                                      nullptr, 0);
            };
            finish_sweep(c, sl.ws, nullptr, rerun);
            RTCD_HIP(hipMemcpyAsync(sl.out.p, sl.ws.h_out.p, sizeof(uint32_t) * r.n, hipMemcpyDeviceToHost,
                                    sl.ws.stream));
            RTCD_HIP(hipStreamSynchronize(sl.ws.stream));
        }
    }
    std::memcpy(r.out, sl.out.p, sizeof(uint32_t) * r.n);
}

void worker_main(ServeState* S) {
    (void) hipSetDevice(S->ctx->cfg.device);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        bool progressed = false;
        // 1. completions
        for (auto& up : S->slots) {
            Slot& sl = *up;
            if (!sl.cur) continue;
            if (hipEventQuery(sl.done_ev) == hipErrorNotReady) continue;
            Req* q = sl.cur;
            lk.unlock();
            rtcd_status st = RTCD_OK;
            std::string err;
            try {
                RTCD_HIP(hipEventSynchronize(sl.done_ev));
                complete(S, sl);
            } catch (const HipError& e) {
                st = e.code == hipErrorOutOfMemory ? RTCD_ERR_OOM : RTCD_ERR_HIP;
                err = e.what();
            } catch (const std::exception& e) {
                st = RTCD_ERR_INTERNAL;
                err = e.what();
            }
            lk.lock();
            q->t_done = now_ms();
            q->status = st;
            q->err = err;
            q->done = true;
            q->scene.reset();
            S->in_use -= q->bytes;
            sl.cur = nullptr;
            if (st == RTCD_OK) {
                S->lat.push_back(q->t_done - q->t_submit);
                S->completed++;
            } else {
                S->failed++;
            }
            S->t_last = q->t_done;
            progressed = true;
            S->cv_done.notify_all();
        }
        // 2. admissions
        for (auto& up : S->slots) {
            Slot& sl = *up;
            if (sl.cur || S->queue.empty()) continue;
            Req* pick = nullptr;
            const double tnow = now_ms();
            for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
                if (S->in_use + (*it)->bytes <= S->budget) {
                    pick = *it;
                    S->queue.erase(it);
                    break;
                }
                // the head has waited too long: stop letting smaller requests overtake it
                if (tnow - (*it)->t_submit > kBypassMs) break;
            }
            if (!pick) break;
            S->in_use += pick->bytes;
            pick->t_start = tnow;
            sl.cur = pick;
            lk.unlock();
            try {
                launch(S, sl, pick);
            } catch (const std::exception& e) {
                lk.lock();
                pick->status = RTCD_ERR_HIP;
                pick->err = e.what();
                pick->done = true;
                pick->t_done = now_ms();
                pick->scene.reset();
                S->in_use -= pick->bytes;
                S->failed++;
                sl.cur = nullptr;
                S->cv_done.notify_all();
                lk.unlock();
            }
            lk.lock();
            progressed = true;
        }
        bool busy = false;
        for (auto& up : S->slots) busy |= up->cur != nullptr;
        if (S->stop && !busy && S->queue.empty()) break;
        if (!progressed) {
            if (busy) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

}  // namespace

void serve_shutdown(rtcd_ctx* ctx) {
    ServeState* S = ctx->serve;
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->m);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (auto& sl : S->slots) {
        if (sl->done_ev) (void) hipEventDestroy(sl->done_ev);
        sl->ws.destroy();
    }
    ctx->serve_bytes = 0;
    delete S;
    ctx->serve = nullptr;
}

}  // namespace rtcd

extern "C" {

rtcd_status rtcd_serve_start(rtcd_ctx* ctx, uint32_t slots, uint64_t mem_budget) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    if (slots < 1 || slots > 16) return fail(RTCD_ERR_INVALID_ARG, "slots must be 1..16");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    if (ctx->serve) return fail(RTCD_ERR_INVALID_ARG, "serving loop already running");
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        auto S = std::make_unique<ServeState>();
        S->ctx = ctx;
        S->budget = mem_budget ? mem_budget : ((size_t) 8 << 30);
        for (uint32_t i = 0; i < slots; ++i) {
            auto sl = std::make_unique<Slot>();
            hipStream_t st = nullptr;
            RTCD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            sl->ws.init(st, true);
            RTCD_HIP(hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming));
            S->slots.push_back(std::move(sl));
        }
        ctx->serve = S.release();
        ctx->serve->worker = std::thread(worker_main, ctx->serve);
        return RTCD_OK;
    });
}

rtcd_status rtcd_serve_stop(rtcd_ctx* ctx) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->serve) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    serve_shutdown(ctx);
    return RTCD_OK;
}

rtcd_status rtcd_submit(rtcd_ctx* ctx, const rtcd_request* r, uint64_t* id) {
    if (!ctx || !r || !id) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    ServeState* S = ctx->serve;
    if (!S) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    if (r->kind < RTCD_REQ_DISCRETE || r->kind > RTCD_REQ_QUADRATIC) return fail(RTCD_ERR_INVALID_ARG, "bad kind");
    if (r->n == 0 || !r->input || !r->out) return fail(RTCD_ERR_INVALID_ARG, "empty request or null buffer");
    if (r->kind == RTCD_REQ_DISCRETE && (r->direction < 1 || r->direction > 3))
        return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (r->kind == RTCD_REQ_QUADRATIC && (r->n_ctrl < 3 || r->n_wp < r->n_ctrl || r->n_ctrl > 64 || r->n_wp > 4096))
        return fail(RTCD_ERR_INVALID_ARG, "need 3 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    uint32_t nw = r->n;
    if (r->kind == RTCD_REQ_LINEAR) {
        if (!r->offsets || r->offsets[0] != 0) return fail(RTCD_ERR_INVALID_ARG, "offsets must start at 0");
        for (uint32_t i = 0; i < r->n; ++i)
            if (r->offsets[i + 1] <= r->offsets[i])
                return fail(RTCD_ERR_INVALID_ARG, "every trajectory needs at least one waypoint");
        nw = r->offsets[r->n];
    } else if (r->kind == RTCD_REQ_QUADRATIC) {
        nw = r->n * r->n_wp;
    }
    if (!finite_all(r->input, (size_t) nw * ctx->robot.dof)) return fail(RTCD_ERR_INVALID_ARG, "non-finite joint value");
    std::shared_ptr<SceneRes> sc;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        sc = find_scene(ctx, r->scene_id);
    }
    if (!sc) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    auto q = std::make_unique<Req>();
    q->r = *r;
    q->scene = sc;
    q->n_wp_total = nw;
    q->bytes = request_bytes(ctx, *r, nw);
    if (q->bytes > S->budget) return fail(RTCD_ERR_OOM, "request larger than the serving memory budget");
    q->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> g(S->m);
        q->id = S->next_id++;
        *id = q->id;
        if (S->completed + S->failed == 0 && S->reqs.empty()) S->t_first = q->t_submit;
        S->queue.push_back(q.get());
        S->queued_max = std::max(S->queued_max, S->queue.size());
        S->reqs[q->id] = std::move(q);
    }
    S->cv_work.notify_one();
    return RTCD_OK;
}

rtcd_status rtcd_poll(rtcd_ctx* ctx, uint64_t id, int32_t* done) {
    if (!ctx || !done) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    ServeState* S = ctx->serve;
    if (!S) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    std::lock_guard<std::mutex> g(S->m);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(RTCD_ERR_UNKNOWN_ID, "unknown request id");
    *done = it->second->done ? 1 : 0;
    return RTCD_OK;
}

rtcd_status rtcd_wait(rtcd_ctx* ctx, uint64_t id, double* latency_ms) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    ServeState* S = ctx->serve;
    if (!S) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    std::unique_lock<std::mutex> g(S->m);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(RTCD_ERR_UNKNOWN_ID, "unknown request id");
    Req* q = it->second.get();
    S->cv_done.wait(g, [&] { return q->done; });
    if (latency_ms) *latency_ms = q->t_done - q->t_submit;
    rtcd_status st = q->status;
    std::string err = q->err;
    S->reqs.erase(it);
    if (st != RTCD_OK) return fail(st, err);
    return RTCD_OK;
}

rtcd_status rtcd_serve_get_stats(rtcd_ctx* ctx, rtcd_serve_stats* out, int32_t reset) {
    if (!ctx || !out) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    ServeState* S = ctx->serve;
    if (!S) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    std::lock_guard<std::mutex> g(S->m);
    std::vector<double> l = S->lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t) (p * (l.size() - 1) + 0.5))]; };
    out->completed = S->completed;
    out->failed = S->failed;
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = l.empty() ? 0.0 : l.back();
    double sum = 0;
    for (double x : l) sum += x;
    out->mean_ms = l.empty() ? 0.0 : sum / l.size();
    out->span_ms = S->t_last - S->t_first;
    out->queued_max = (uint32_t) S->queued_max;
    if (reset) {
        S->lat.clear();
        S->completed = S->failed = 0;
        S->queued_max = 0;
        S->t_first = now_ms();
    }
    return RTCD_OK;
}

}  // extern "C"
