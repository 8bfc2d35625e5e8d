// Serving loop: continuous batching of discrete / linear / quadratic requests over resident scenes.
//
// RTCD has no serving layer; its closest analogue is the NSTREAM ThreadPool in
// CollisionDetector::check (RTCD/CollisionDetector/CollisionDetector.h:396-433, threadPool.h:10-92),
// which launches fixed batches on round-robin streams and joins every NSTREAM tasks.  Here each of
// `slots` slots has its own host thread, workspace and stream, and runs work units:
//  * a request is split into units of at most chunk_discrete poses or chunk_sweep trajectories
//    (queries are independent), so no unit holds a slot for long;
//  * a free slot takes the queued unit of the smallest size class (log2 of the request size,
//    FIFO within a class), unless some unit has waited longer than kAgingMs, which then goes
//    first, and batches with it the queued units of the same kind, scene and shape (FIFO) up to the
//    chunk size: many small requests become one launch (continuous batching);
//  * with more than one slot, slot 0 is a lane for small requests (at most lane_max queries, of
//    the kinds in lane_kinds): it never ages, never takes a larger request and batches at most
//    lane_max queries, so a small request finds a slot that no bulk unit holds; with lane CUs set,
//    stream CU masks give slot 0 its own CUs and the other slots the rest, so bulk kernels cannot
//    fill the lane's CUs either;
//  * a batch is admitted only while its estimated device bytes fit the memory budget;
//  * admitted batches run concurrently on the device; completion is detected with events, each
//    unit's flags are copied into its range of the caller's buffer, and a request completes (its
//    submit->done latency recorded) when its last unit does.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include <hip/hip_ext.h>

#include "ctx.hpp"

namespace rtcd {

namespace {
constexpr double kAgingMs = 20.0;

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
    double t_submit = 0, t_start = 0, t_done = 0;  // t_start: the first unit's launch
    rtcd_status status = RTCD_OK;
    std::string err;
    bool done = false;
    uint32_t units_left = 0;
};

// A contiguous range [i0, i0 + n) of a request's queries.
struct Unit {
    Req* q = nullptr;
    uint32_t i0 = 0, n = 0;
    uint32_t nw = 0;         // waypoints of the range (linear: sum of its lengths; quadratic: n * n_wp)
    size_t bytes = 0;
    int cls = 0;             // size class of the request: admission order
};

// Units launched together: same kind, scene and shape; unit k covers batch queries
// [first[k], first[k] + units[k].n).
struct Batch {
    std::vector<Unit> units;
    std::vector<uint32_t> first;
    uint32_t n = 0, nw = 0;
    size_t bytes = 0;
};

struct Slot {
    Workspace ws;
    hipEvent_t done_ev = nullptr;
    Batch cur;
    PinnedBuf<float> in;
    PinnedBuf<uint32_t> offs;
    PinnedBuf<uint32_t> out;
    PinnedBuf<uint32_t> sv;  // the workspace's status scalars, copied back with the flags
    std::thread th;
};
}  // namespace

struct ServeState {
    rtcd_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Unit> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    size_t budget = 0, in_use = 0;
    uint32_t chunk_discrete = 4096, chunk_sweep = 1024;
    uint32_t lane_max = 0;   // slot 0 takes only requests of at most lane_max queries (0: any)
    uint32_t lane_kinds = 15; // ... and of these kinds (bit kind - 1)
    uint32_t lane_cus = 0;   // CUs of slot 0's stream mask (0: no masks)
    std::vector<double> lat, qlat;
    uint64_t completed = 0, failed = 0;
    double t_first = 0, t_last = 0;
    size_t queued_max = 0;
};

namespace {

size_t unit_bytes(const rtcd_ctx* c, const rtcd_request& r, uint32_t n, uint32_t nw) {
    const size_t staging = (size_t) nw * c->robot.dof * 4 + (size_t) n * 4;
    if (r.kind == RTCD_REQ_DISCRETE) return discrete_bytes(c, n) + staging;
    if (r.kind == RTCD_REQ_LINEAR) return sweep_bytes(c, n, nw, 0) + staging;
    return sweep_bytes(c, n, nw, r.n_ctrl) + staging;
}

// First input waypoint of a unit's range in its request's input.
size_t first_waypoint(const Unit& u) {
    const rtcd_request& r = u.q->r;
    if (r.kind == RTCD_REQ_DISCRETE) return u.i0;
    if (r.kind == RTCD_REQ_LINEAR) return r.offsets[u.i0];
    return (size_t) u.i0 * r.n_wp;
}

// Launch batch b on slot sl (the slot's thread, no lock held): stage the units' inputs one after the
// other, enqueue one check over the whole batch.
void launch(ServeState* S, Slot& sl, const Batch& b) {
    rtcd_ctx* c = S->ctx;
    const rtcd_request& r = b.units[0].q->r;  // kind, scene and shape are common to the batch
    Workspace& ws = sl.ws;
    const hipStream_t s = ws.stream;
    const uint32_t dof = (uint32_t) c->robot.dof;
    const size_t nin = (size_t) b.nw * dof;
    sl.in.reserve(nin);
    size_t at = 0;
    for (const Unit& u : b.units) {
        std::memcpy(sl.in.p + at, u.q->r.input + first_waypoint(u) * dof, (size_t) u.nw * dof * sizeof(float));
        at += (size_t) u.nw * dof;
    }
    ws.h_poses.reserve(nin);
    RTCD_HIP(hipMemcpyAsync(ws.h_poses.p, sl.in.p, nin * sizeof(float), hipMemcpyHostToDevice, s));
    ws.h_out.reserve(b.n);
    SceneRes& sc = *b.units[0].q->scene;
    if (r.kind == RTCD_REQ_DISCRETE) {
        enqueue_discrete(c, ws, sc, r.direction, ws.h_poses.p, b.n, ws.h_out.p, 0);
    } else if (r.kind == RTCD_REQ_LINEAR) {
        sl.offs.reserve(b.n + 1);
        uint32_t k = 0, base = 0;
        for (const Unit& u : b.units) {
            const uint32_t* o = u.q->r.offsets + u.i0;
            for (uint32_t i = 0; i < u.n; ++i) sl.offs.p[k++] = base + (o[i] - o[0]);
            base += u.nw;
        }
        sl.offs.p[k] = base;
        ws.h_offsets.reserve(b.n + 1);
        RTCD_HIP(hipMemcpyAsync(ws.h_offsets.p, sl.offs.p, sizeof(uint32_t) * (b.n + 1), hipMemcpyHostToDevice, s));
        enqueue_linear(c, ws, sc, ws.h_poses.p, ws.h_offsets.p, b.n, b.nw, ws.h_out.p, nullptr, 0);
    } else {
        enqueue_curve(c, ws, sc, ws.h_poses.p, b.n, r.n_wp, r.n_ctrl, r.kind == RTCD_REQ_CUBIC ? 3 : 2, nullptr,
                      ws.h_out.p, nullptr, 0);
    }
    sl.out.reserve(b.n);
    sl.sv.reserve(4);
    RTCD_HIP(hipMemcpyAsync(sl.sv.p, ws.scalars.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RTCD_HIP(hipMemcpyAsync(sl.out.p, ws.h_out.p, sizeof(uint32_t) * b.n, hipMemcpyDeviceToHost, s));
    RTCD_HIP(hipEventRecord(sl.done_ev, s));
}

// Completion of the batch on slot sl (its done event has fired): re-run on overflow, scatter the
// flags into each unit's range of its caller's buffer.
void complete(ServeState* S, Slot& sl) {
    const Batch& b = sl.cur;
    rtcd_ctx* c = S->ctx;
    const rtcd_request& r = b.units[0].q->r;
    SceneRes& sc = *b.units[0].q->scene;
    if (r.kind == RTCD_REQ_DISCRETE) {
        // re-runs synchronously on the slot stream if a candidate segment or the FP64 queue overflowed
        if (sl.sv.p[1]) {
            finish_discrete(c, sl.ws, sc, r.direction, sl.ws.h_poses.p, b.n, sl.ws.h_out.p, 0, nullptr);
            RTCD_HIP(hipMemcpyAsync(sl.out.p, sl.ws.h_out.p, sizeof(uint32_t) * b.n, hipMemcpyDeviceToHost,
                                    sl.ws.stream));
            RTCD_HIP(hipStreamSynchronize(sl.ws.stream));
        }
    } else {
        if (sl.sv.p[1] & 7u) {
            auto rerun = [&]() {
                if (r.kind == RTCD_REQ_LINEAR)
                    enqueue_linear(c, sl.ws, sc, sl.ws.h_poses.p, sl.ws.h_offsets.p, b.n, b.nw, sl.ws.h_out.p, nullptr,
                                   0);
                else
                    enqueue_curve(c, sl.ws, sc, sl.ws.h_poses.p, b.n, r.n_wp, r.n_ctrl,
                                  r.kind == RTCD_REQ_CUBIC ? 3 : 2, nullptr, sl.ws.h_out.p, nullptr, 0);
            };
            finish_sweep(c, sl.ws, nullptr, rerun);
            RTCD_HIP(hipMemcpyAsync(sl.out.p, sl.ws.h_out.p, sizeof(uint32_t) * b.n, hipMemcpyDeviceToHost,
                                    sl.ws.stream));
            RTCD_HIP(hipStreamSynchronize(sl.ws.stream));
        }
    }
    for (size_t k = 0; k < b.units.size(); ++k)
        std::memcpy(b.units[k].q->r.out + b.units[k].i0, sl.out.p + b.first[k], sizeof(uint32_t) * b.units[k].n);
}

// Units that can share one launch: same kind, scene, direction and B-spline shape.
bool same_shape(const Unit& a, const Unit& b) {
    const rtcd_request &x = a.q->r, &y = b.q->r;
    return x.kind == y.kind && a.q->scene == b.q->scene &&
           (x.kind != RTCD_REQ_DISCRETE || x.direction == y.direction) &&
           ((x.kind != RTCD_REQ_QUADRATIC && x.kind != RTCD_REQ_CUBIC) || (x.n_wp == y.n_wp && x.n_ctrl == y.n_ctrl));
}

// Whether the small-request lane (slot 0) may take unit u.
bool lane_ok(const ServeState* S, const Unit& u, uint32_t lane) {
    return u.q->r.n <= lane && ((S->lane_kinds >> (u.q->r.kind - 1)) & 1u);
}

// Next batch to admit (lock held).  Its lead unit is the smallest size class, FIFO within it,
// unless `aging` and the oldest unit has waited longer than kAgingMs; then queued units of the same
// shape join it in submit order while the batch stays within the chunk size and the memory budget.
// With `lane` > 0 only units of requests of at most `lane` queries and of the lane's kinds are taken,
// and the batch holds at most `lane` queries.  The batch's units are removed from the queue.  Empty if nothing fits.
Batch pick_batch(ServeState* S, double tnow, bool aging, uint32_t lane) {
    Batch b;
    auto lead = S->queue.end();
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
        if (S->in_use + it->bytes > S->budget) continue;
        if (lane && !lane_ok(S, *it, lane)) continue;
        if (aging && tnow - it->q->t_submit > kAgingMs) {  // the queue is in submit order
            lead = it;
            break;
        }
        if (lead == S->queue.end() || it->cls < lead->cls) lead = it;
    }
    if (lead == S->queue.end()) return b;
    const Unit head = *lead;
    const rtcd_ctx* c = S->ctx;
    uint32_t chunk = head.q->r.kind == RTCD_REQ_DISCRETE ? S->chunk_discrete : S->chunk_sweep;
    if (lane) chunk = std::min(chunk, lane);
    auto add = [&](const Unit& u) {
        b.first.push_back(b.n);
        b.units.push_back(u);
        b.n += u.n;
        b.nw += u.nw;
    };
    add(head);
    b.bytes = head.bytes;
    S->queue.erase(lead);
    for (auto it = S->queue.begin(); it != S->queue.end() && b.n < chunk;) {
        if (!same_shape(*it, head) || b.n + it->n > chunk || (lane && !lane_ok(S, *it, lane))) {
            ++it;
            continue;
        }
        const size_t nb = unit_bytes(c, head.q->r, b.n + it->n, b.nw + it->nw);
        if (S->in_use + nb > S->budget) break;
        add(*it);
        b.bytes = nb;
        it = S->queue.erase(it);
    }
    return b;
}

// Records the end of unit u (lock held): the request completes with its last unit.
void unit_done(ServeState* S, const Unit& u, rtcd_status st, const std::string& err) {
    Req* q = u.q;
    if (st != RTCD_OK && q->status == RTCD_OK) {
        q->status = st;
        q->err = err;
    }
    if (--q->units_left > 0) return;
    q->t_done = now_ms();
    q->done = true;
    q->scene.reset();
    if (q->status == RTCD_OK) {
        S->lat.push_back(q->t_done - q->t_submit);
        S->qlat.push_back(q->t_start - q->t_submit);
        S->completed++;
    } else {
        S->failed++;
    }
    S->t_last = q->t_done;
    S->cv_done.notify_all();
}

// One host thread per slot: take a batch (lock held), launch it, wait for its event, scatter the
// flags, record the units' completion.  Slots never wait on each other's host work, so a small
// request on the lane is not held up by a bulk launch being enqueued on another slot.
void slot_main(ServeState* S, size_t si) {
    (void) hipSetDevice(S->ctx->cfg.device);
    Slot& sl = *S->slots[si];
    // slot 0, with other slots beside it, is the small-request lane: smallest request first (no
    // aging), only requests of at most lane_max queries of the lane's kinds
    const bool lane_slot = si == 0 && S->slots.size() > 1;
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        Batch b;
        const double tnow = now_ms();
        if (!S->queue.empty()) b = pick_batch(S, tnow, !lane_slot, lane_slot ? S->lane_max : 0);
        if (b.units.empty()) {
            if (S->stop && S->queue.empty()) break;
            // woken by submits, completions (they free budget) and stop
            S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            continue;
        }
        for (const Unit& u : b.units)
            if (u.q->t_start == 0) u.q->t_start = tnow;
        S->in_use += b.bytes;
        sl.cur = std::move(b);
        lk.unlock();
        rtcd_status st = RTCD_OK;
        std::string err;
        try {
            launch(S, sl, sl.cur);
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
        for (const Unit& u : sl.cur.units) unit_done(S, u, st, err);
        S->in_use -= sl.cur.bytes;
        sl.cur = Batch{};
        S->cv_work.notify_all();
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
    for (auto& sl : S->slots)
        if (sl->th.joinable()) sl->th.join();
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
        S->chunk_discrete = ctx->serve_chunk_discrete;
        S->chunk_sweep = ctx->serve_chunk_sweep;
        S->lane_max = ctx->serve_lane_max;
        S->lane_kinds = ctx->serve_lane_kinds;
        // CU masks: on the 256-CU part (8 XCDs of 32 CUs), CU i goes to slot 0 iff
        // ((i mod 8) - floor(i / 32)) mod 8 < lane_cus / 32.  Whether the driver maps mask bit i to XCD
        // i mod 8 (interleaved) or to XCD i / 32 (blocked), both masks then hold CUs on every XCD.
        std::vector<uint32_t> lane_mask, bulk_mask;
        int n_cu = 0;
        RTCD_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->cfg.device));
        if (slots > 1 && ctx->serve_lane_cus && n_cu == 256) {
            lane_mask.assign(8, 0u);
            bulk_mask.assign(8, 0u);
            const int m = (int) ctx->serve_lane_cus / 32;
            for (int i = 0; i < 256; ++i) {
                const bool lane = (((i % 8) - i / 32) % 8 + 8) % 8 < m;
                (lane ? lane_mask : bulk_mask)[i / 32] |= 1u << (i % 32);
            }
        }
        try {
            for (uint32_t i = 0; i < slots; ++i) {
                S->slots.push_back(std::make_unique<Slot>());
                Slot& sl = *S->slots.back();
                hipStream_t st = nullptr;
                if (!lane_mask.empty()) {
                    std::vector<uint32_t>& mk = i == 0 ? lane_mask : bulk_mask;
                    if (hipExtStreamCreateWithCUMask(&st, (uint32_t) mk.size(), mk.data()) != hipSuccess) {
                        (void) hipGetLastError();
                        st = nullptr;
                        if (i > 0) throw std::runtime_error("a CU-masked serving stream could not be created");
                        lane_mask.clear();  // no masks on any slot
                    }
                }
                if (!st) RTCD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                sl.ws.stream = st;  // owned from here on
                sl.ws.own_stream = true;
                sl.ws.init(st, true);
                RTCD_HIP(hipEventCreateWithFlags(&sl.done_ev, hipEventDisableTiming));
            }
        } catch (...) {
            for (auto& sl : S->slots) {
                if (sl->done_ev) (void) hipEventDestroy(sl->done_ev);
                sl->ws.destroy();
            }
            throw;
        }
        S->lane_cus = lane_mask.empty() ? 0 : ctx->serve_lane_cus;
        ctx->serve = S.release();
        for (size_t i = 0; i < ctx->serve->slots.size(); ++i)
            ctx->serve->slots[i]->th = std::thread(slot_main, ctx->serve, i);
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
    if (r->kind < RTCD_REQ_DISCRETE || r->kind > RTCD_REQ_CUBIC) return fail(RTCD_ERR_INVALID_ARG, "bad kind");
    if (r->n == 0 || !r->input || !r->out) return fail(RTCD_ERR_INVALID_ARG, "empty request or null buffer");
    if (r->kind == RTCD_REQ_DISCRETE && (r->direction < 1 || r->direction > 3))
        return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (r->kind == RTCD_REQ_QUADRATIC && (r->n_ctrl < 3 || r->n_wp < r->n_ctrl || r->n_ctrl > 64 || r->n_wp > 4096))
        return fail(RTCD_ERR_INVALID_ARG, "need 3 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    if (r->kind == RTCD_REQ_CUBIC && (r->n_ctrl < 4 || r->n_wp < r->n_ctrl || r->n_ctrl > 64 || r->n_wp > 4096))
        return fail(RTCD_ERR_INVALID_ARG, "need 4 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    uint32_t nw = r->n;
    if (r->kind == RTCD_REQ_LINEAR) {
        if (!r->offsets || r->offsets[0] != 0) return fail(RTCD_ERR_INVALID_ARG, "offsets must start at 0");
        for (uint32_t i = 0; i < r->n; ++i)
            if (r->offsets[i + 1] <= r->offsets[i])
                return fail(RTCD_ERR_INVALID_ARG, "every trajectory needs at least one waypoint");
        nw = r->offsets[r->n];
    } else if (r->kind == RTCD_REQ_QUADRATIC || r->kind == RTCD_REQ_CUBIC) {
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
    // units of at most `chunk` queries; the size class orders admission (small requests first)
    const uint32_t chunk = r->kind == RTCD_REQ_DISCRETE ? S->chunk_discrete : S->chunk_sweep;
    int cls = 0;
    while ((1u << cls) < r->n && cls < 31) ++cls;
    std::vector<Unit> units;
    for (uint32_t i0 = 0; i0 < r->n; i0 += chunk) {
        Unit u;
        u.q = q.get();
        u.i0 = i0;
        u.n = std::min(chunk, r->n - i0);
        u.nw = r->kind == RTCD_REQ_DISCRETE ? u.n
               : r->kind == RTCD_REQ_LINEAR ? r->offsets[i0 + u.n] - r->offsets[i0]
                                            : u.n * r->n_wp;
        u.bytes = unit_bytes(ctx, *r, u.n, u.nw);
        u.cls = cls;
        if (u.bytes > S->budget) return fail(RTCD_ERR_OOM, "request unit larger than the serving memory budget");
        units.push_back(u);
    }
    q->units_left = (uint32_t) units.size();
    q->t_submit = now_ms();
    {
        std::lock_guard<std::mutex> g(S->m);
        q->id = S->next_id++;
        *id = q->id;
        if (S->completed + S->failed == 0 && S->reqs.empty()) S->t_first = q->t_submit;
        for (const Unit& u : units) S->queue.push_back(u);
        S->queued_max = std::max(S->queued_max, S->queue.size());
        S->reqs[q->id] = std::move(q);
    }
    S->cv_work.notify_all();  // the lane slot may not take this request: wake every slot
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

rtcd_status rtcd_wait(rtcd_ctx* ctx, uint64_t id, double* latency_ms) { return rtcd_wait_timed(ctx, id, latency_ms, nullptr); }

rtcd_status rtcd_wait_timed(rtcd_ctx* ctx, uint64_t id, double* latency_ms, double* queued_ms) {
    if (!ctx) return fail(RTCD_ERR_INVALID_ARG, "null ctx");
    ServeState* S = ctx->serve;
    if (!S) return fail(RTCD_ERR_NOT_READY, "serving loop not running");
    std::unique_lock<std::mutex> g(S->m);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(RTCD_ERR_UNKNOWN_ID, "unknown request id");
    Req* q = it->second.get();
    S->cv_done.wait(g, [&] { return q->done; });
    if (latency_ms) *latency_ms = q->t_done - q->t_submit;
    if (queued_ms) *queued_ms = q->t_start > 0 ? q->t_start - q->t_submit : 0.0;
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
    std::vector<double> l = S->lat, ql = S->qlat;
    std::sort(l.begin(), l.end());
    std::sort(ql.begin(), ql.end());
    auto pct_of = [](const std::vector<double>& v, double p) {
        return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t) (p * (v.size() - 1) + 0.5))];
    };
    auto pct = [&](double p) { return pct_of(l, p); };
    out->queue_p50_ms = pct_of(ql, 0.50);
    out->queue_p99_ms = pct_of(ql, 0.99);
    out->lane_cus = S->lane_cus;
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
        S->qlat.clear();
        S->completed = S->failed = 0;
        S->queued_max = 0;
        S->t_first = now_ms();
    }
    return RTCD_OK;
}

}  // extern "C"
