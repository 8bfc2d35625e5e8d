// Autotuner: times the caller's own sample request on the device under the launch-shape choices of
// the context and keeps the fastest.
//
// RTCD fixes its launch shapes at compile time (BATCHSIZE / NSTREAM template arguments,
// RTCD/CollisionDetector/CollisionDetector.h:396-433) and leaves BVH shapes to OptiX.  Here every
// choice is a context parameter, and rtcd_autotune searches them by coordinate descent: for each knob
// in turn it times every candidate value (one warm-up, median of three device-timed runs) with the
// others fixed, and keeps the best.  Results never depend on these choices; the tuner checks that:
// every configuration must reproduce the flags of the first one, or the call fails.
#include "ctx.hpp"

namespace {
using namespace rtcd;

struct Knob {
    uint32_t* field;
    std::vector<uint32_t> values;
};

struct Runner {
    rtcd_ctx* c;
    rtcd_request r;
    DevBuf<float> in;
    DevBuf<uint32_t> offs, out;
    uint32_t nw = 0;
    std::vector<uint32_t> ref;
    bool have_ref = false;
    uint32_t timed = 0;

    double run_once() {
        Workspace& ws = c->main;
        std::shared_ptr<SceneRes> sc = find_scene(c, r.scene_id);  // a rebuild replaces the SceneRes
        if (!sc) throw std::invalid_argument("unknown scene id");
        rtcd_stats st{};
        if (r.kind == RTCD_REQ_DISCRETE) {
            enqueue_discrete(c, ws, *sc, r.direction, in.p, r.n, out.p, 0);
            RTCD_HIP(hipStreamSynchronize(ws.stream));
            finish_discrete(c, ws, *sc, r.direction, in.p, r.n, out.p, 0, &st);
        } else if (r.kind == RTCD_REQ_LINEAR) {
            auto run = [&]() { enqueue_linear(c, ws, *sc, in.p, offs.p, r.n, nw, out.p, nullptr, 0); };
            run();
            finish_sweep(c, ws, &st, run);
        } else {
            const int deg = r.kind == RTCD_REQ_CUBIC ? 3 : 2;
            auto run = [&]() { enqueue_curve(c, ws, *sc, in.p, r.n, r.n_wp, r.n_ctrl, deg, nullptr, out.p, nullptr, 0); };
            run();
            finish_sweep(c, ws, &st, run);
        }
        std::vector<uint32_t> f = out.download(r.n, ws.stream);
        if (!have_ref) {
            ref = std::move(f);
            have_ref = true;
        } else if (f != ref) {
            throw std::runtime_error("autotune: a configuration changed the collision flags");
        }
        return st.last_ms;
    }

    double time() {
        run_once();
        double t[3];
        for (double& x : t) x = run_once();
        std::sort(t, t + 3);
        ++timed;
        return t[1];
    }
};

// Tries every value of one knob (the others fixed) and keeps the fastest.
double descend(Runner& R, Knob& k, double best) {
    const uint32_t keep = *k.field;
    uint32_t arg = keep;
    for (uint32_t v : k.values) {
        if (v == keep) continue;
        *k.field = v;
        const double t = R.time();
        if (t < best) {
            best = t;
            arg = v;
        }
    }
    *k.field = arg;
    return best;
}

// The same for one of the BVH leaf sizes (2, 4, 8): every candidate rebuilds the BVHs.
double descend_leaf(Runner& R, uint32_t* field, double best) {
    const uint32_t keep = *field;
    uint32_t arg = keep;
    for (uint32_t v : {2u, 4u, 8u}) {
        if (v == keep) continue;
        *field = v;
        rebuild_bvhs(R.c);
        const double t = R.time();
        if (t < best) {
            best = t;
            arg = v;
        }
    }
    if (*field != arg) {
        *field = arg;
        rebuild_bvhs(R.c);
    }
    return best;
}

}  // namespace

extern "C" {

rtcd_status rtcd_autotune(rtcd_ctx* ctx, const rtcd_request* r, uint32_t flags, rtcd_tune_report* rep) {
    if (!ctx || !r || !rep) return fail(RTCD_ERR_INVALID_ARG, "null argument");
    if (r->kind < RTCD_REQ_DISCRETE || r->kind > RTCD_REQ_CUBIC) return fail(RTCD_ERR_INVALID_ARG, "bad kind");
    if (r->n == 0 || !r->input) return fail(RTCD_ERR_INVALID_ARG, "empty sample");
    if (r->kind == RTCD_REQ_DISCRETE && (r->direction < 1 || r->direction > 3))
        return fail(RTCD_ERR_INVALID_ARG, "direction must be 1, 2 or 3");
    if (r->kind == RTCD_REQ_LINEAR && !r->offsets) return fail(RTCD_ERR_INVALID_ARG, "linear sample needs offsets");
    if (r->kind == RTCD_REQ_QUADRATIC && (r->n_ctrl < 3 || r->n_wp < r->n_ctrl || r->n_ctrl > 64 || r->n_wp > 4096))
        return fail(RTCD_ERR_INVALID_ARG, "need 3 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    if (r->kind == RTCD_REQ_CUBIC && (r->n_ctrl < 4 || r->n_wp < r->n_ctrl || r->n_ctrl > 64 || r->n_wp > 4096))
        return fail(RTCD_ERR_INVALID_ARG, "need 4 <= n_ctrl <= 64 and n_ctrl <= n_wp <= 4096");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (!ctx->robot_loaded) return fail(RTCD_ERR_NOT_READY, "robot not loaded");
    if (ctx->serve) return fail(RTCD_ERR_NOT_READY, "stop the serving loop before tuning");
    if (!find_scene(ctx, r->scene_id)) return fail(RTCD_ERR_UNKNOWN_ID, "unknown scene id");
    // on any failure the context keeps the settings it had on entry (BVHs rebuilt if needed)
    struct Restore {
        rtcd_ctx* c;
        bool counting;
        uint32_t leaf_r, leaf_t, leaf_e, nbs, passes, passes_small, bud, bud_small, bud_next, sw_bud, sw_passes, n_passes,
            n_bud, o_passes;
        bool done = false;
        ~Restore() {
            c->counting = counting;
            if (done) return;
            c->narrow_bs = nbs;
            c->broad_passes = passes;
            c->broad_passes_small = passes_small;
            c->broad_budget = bud;
            c->broad_budget_small = bud_small;
            c->broad_budget_next = bud_next;
            c->sweep_budget = sw_bud;
            c->sweep_passes = sw_passes;
            c->narrow_passes = n_passes;
            c->narrow_budget = n_bud;
            c->o2r_passes = o_passes;
            if (c->leaf_robot != leaf_r || c->leaf_scene_tri != leaf_t || c->leaf_scene_edge != leaf_e) {
                c->leaf_robot = leaf_r;
                c->leaf_scene_tri = leaf_t;
                c->leaf_scene_edge = leaf_e;
                try {
                    rebuild_bvhs(c);
                } catch (...) {
                }
            }
        }
    } restore{ctx,
              ctx->counting,
              ctx->leaf_robot,
              ctx->leaf_scene_tri,
              ctx->leaf_scene_edge,
              ctx->narrow_bs,
              ctx->broad_passes,
              ctx->broad_passes_small,
              ctx->broad_budget,
              ctx->broad_budget_small,
              ctx->broad_budget_next,
              ctx->sweep_budget,
              ctx->sweep_passes,
              ctx->narrow_passes,
              ctx->narrow_budget,
              ctx->o2r_passes};
    return guarded([&]() -> rtcd_status {
        RTCD_HIP(hipSetDevice(ctx->cfg.device));
        ctx->counting = false;
        ctx->main.counting = false;
        drop_graph(ctx);
        Runner R{ctx, *r};
        const uint32_t dof = (uint32_t) ctx->robot.dof;
        const hipStream_t s = ctx->main.stream;
        size_t n_in;
        if (r->kind == RTCD_REQ_LINEAR) {
            const uint32_t* o = r->offsets;
            if (o[0] != 0) return fail(RTCD_ERR_INVALID_ARG, "offsets[0] must be 0");
            for (uint32_t i = 0; i < r->n; ++i)
                if (o[i + 1] <= o[i]) return fail(RTCD_ERR_INVALID_ARG, "every trajectory needs a waypoint");
            R.nw = o[r->n];
            n_in = (size_t) R.nw * dof;
            R.offs.upload(o, r->n + 1, s);
        } else {
            n_in = (size_t) r->n * (r->kind >= RTCD_REQ_QUADRATIC ? r->n_wp : 1) * dof;
        }
        if (!finite_all(r->input, n_in)) return fail(RTCD_ERR_INVALID_ARG, "non-finite joint value");
        R.in.upload(r->input, n_in, s);
        R.out.reserve(r->n);
        const double before = R.time();
        double best = before;
        bool small_batch = false;
        if (r->kind == RTCD_REQ_DISCRETE) {
            const size_t tasks = (size_t) r->n * (ctx->robot.n_links - (ctx->cfg.skip_base_link ? 1 : 0));
            const bool small = tasks < ctx->frontier_max_tasks;
            small_batch = small;
            if (flags & RTCD_TUNE_LEAF_SIZE) {
                best = descend_leaf(R, &ctx->leaf_scene_tri, best);
                best = descend_leaf(R, &ctx->leaf_robot, best);
            }
            std::vector<Knob> knobs = {
                {&ctx->narrow_bs, {64, 128, 256}},
                {&ctx->narrow_passes, {1, 4, 6, 8}},
                {&ctx->narrow_budget, {8, 16, 32, 64}},
                {&ctx->o2r_passes, {1, 2, 4}},
                small ? Knob{&ctx->broad_passes_small, {2, 3, 4, 6}} : Knob{&ctx->broad_passes, {3, 4, 6, 8, 12}},
                {&ctx->broad_budget_next, {8, 16, 32}},
                small ? Knob{&ctx->broad_budget_small, {8, 16, 32, 64}} : Knob{&ctx->broad_budget, {16, 32, 64, 128}},
            };
            for (Knob& k : knobs) best = descend(R, k, best);
        } else {
            if (flags & RTCD_TUNE_LEAF_SIZE) best = descend_leaf(R, &ctx->leaf_scene_edge, best);
            std::vector<Knob> knobs = {{&ctx->sweep_passes, {4, 6, 8, 12}}, {&ctx->sweep_budget, {8, 16, 32, 64}}};
            for (Knob& k : knobs) best = descend(R, k, best);
        }
        rep->ms_before = before;
        rep->ms_after = R.time();
        rep->configs_timed = R.timed;
        rep->leaf_robot = ctx->leaf_robot;
        rep->leaf_scene_tri = ctx->leaf_scene_tri;
        rep->leaf_scene_edge = ctx->leaf_scene_edge;
        rep->narrow_block = ctx->narrow_bs;
        rep->broad_passes = small_batch ? ctx->broad_passes_small : ctx->broad_passes;
        rep->broad_budget = ctx->broad_budget;
        rep->broad_budget_small = ctx->broad_budget_small;
        rep->broad_budget_next = ctx->broad_budget_next;
        rep->sweep_budget = ctx->sweep_budget;
        rep->sweep_passes = ctx->sweep_passes;
        restore.done = true;
        return RTCD_OK;
    });
}

}  // extern "C"
