// Discrete checks: batched FK, broad phase (link boxes vs scene LBVH) and the any-hit narrow phase.
//
// Reference path (RTCD/CollisionDetector/CollisionDetector.h:919-983 detectWithOBB):
//   fkineKernelAdj (RTCD/Robot/robot.cu:202-241) -> per-batch IAS rebuild -> optixLaunch of
//   __raygen__oneside/__anyhit__IAS (CCCuda.cu:31-77, obstacle edges vs link instances) or
//   __raygen__robotRayCompact (CCRay.cu:74-116, link edges moved to the world vs the scene).
// Here no acceleration structure is rebuilt per batch: for each (pose, link) the static scene LBVH
// is queried with the link's boxes, and each candidate scene triangle is moved into the link frame,
// where its owned edges trace the static link-triangle LBVH (obstacle->robot) and the triangle
// itself queries the static link-edge LBVH (robot->obstacle).
#include <type_traits>

#include "band32.hpp"
#include "cand.hpp"
#include "exact64.hpp"
#include "kernels.hpp"

namespace rtcd {
namespace {

constexpr int kBS = 256;
constexpr int kSeg = kCandSegments;

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
// Rows r_j = (R_0j, R_1j, R_2j, t_j): rot() applies R^T, to_link() applies R^T (p - t).
__device__ __forceinline__ float3 rot(const float4& r0, const float4& r1, const float4& r2, float3 p) {
    return make_float3(r0.x * p.x + r0.y * p.y + r0.z * p.z, r1.x * p.x + r1.y * p.y + r1.z * p.z,
                       r2.x * p.x + r2.y * p.y + r2.z * p.z);
}
__device__ __forceinline__ float3 to_link(const float4& r0, const float4& r1, const float4& r2, float3 p) {
    return rot(r0, r1, r2, make_float3(p.x - r0.w, p.y - r1.w, p.z - r2.w));
}
__device__ __forceinline__ float3 add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 neg(float3 a) { return make_float3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ uint32_t load_flag(const uint32_t* f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Scene triangle t moved into the frame of a link with rows (r0, r1, r2) = [R^T | t]: vertex A as
// R^T (A - t) and the edges rotated from world differences, so that their error is relative to
// their own length.  delta bounds the abs. error of A per component, |A' - A'_exact| <= eps |t| +
// 5 eps |A - t|_1, plus 8 eps |e| for B' = A' + e1', C' = A' + e2'.
struct LinkTri {
    float3 A, e1, e2;
    float delta;
};
__device__ __forceinline__ LinkTri link_tri(const Tri& t, const float4& r0, const float4& r1, const float4& r2) {
    LinkTri o;
    const float tinf = fmaxf(fmaxf(fabsf(r0.w), fabsf(r1.w)), fabsf(r2.w));
    const float3 wA = f3(t.a), wB = f3(t.b), wC = f3(t.c);
    const float3 rel = make_float3(wA.x - r0.w, wA.y - r1.w, wA.z - r2.w);
    o.A = rot(r0, r1, r2, rel);
    o.e1 = rot(r0, r1, r2, sub(wB, wA));
    o.e2 = rot(r0, r1, r2, sub(wC, wA));
    const float en = fmaxf(l1(o.e1.x, o.e1.y, o.e1.z), l1(o.e2.x, o.e2.y, o.e2.z));
    o.delta = 1.05f * kEps32 * (tinf + 5.f * l1(rel.x, rel.y, rel.z) + 8.f * en) + 1e-30f;
    return o;
}

// ------------------------------------------------------------------------------------------ FK
// Row r of the world AABB of a link box (b = centre[3], lo[3], hi[3] of LinkInfo) under the link
// transform T (rounded outwards).
__device__ __forceinline__ void link_box_row(const double* T, const float* b, int r, float& lo, float& hi) {
    const double c0 = b[0], c1 = b[1], c2 = b[2];
    const double h0 = 0.5 * ((double) b[6] - b[3]), h1 = 0.5 * ((double) b[7] - b[4]),
                 h2 = 0.5 * ((double) b[8] - b[5]);
    const double wc = T[r * 4 + 0] * c0 + T[r * 4 + 1] * c1 + T[r * 4 + 2] * c2 + T[r * 4 + 3];
    const double wh = fabs(T[r * 4 + 0]) * h0 + fabs(T[r * 4 + 1]) * h1 + fabs(T[r * 4 + 2]) * h2;
    const double pad = 1e-6 * (1.0 + fabs(wc) + wh);
    lo = (float) (wc - wh - pad);
    hi = (float) (wc + wh + pad);
}

// The per-call clears the pose kernel does on the side (it runs first, so nothing reads these
// before it ends): the caller's flags, the status words and the candidate-segment counters.
__device__ __forceinline__ void fk_clear(const FkClear& z, int p, bool pose_lane) {
    if (pose_lane && p < z.n) z.flags[p] = 0u;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < kScalars; i += kBS) z.scalars[i] = 0u;
        for (int b = threadIdx.x; b < kSeg; b += kBS) z.seg_count[(size_t) b * kSegStride] = 0u;
    }
}

// One thread per pose: T64, Ti and wbox of every link.
__global__ __launch_bounds__(kBS) void k_fk(const RobotDev* __restrict__ R, const float* __restrict__ poses, int n,
                                           double* __restrict__ T64, float4* __restrict__ Ti,
                                           float4* __restrict__ wbox, FkClear z) {
    const int p = blockIdx.x * kBS + threadIdx.x;
    fk_clear(z, p, true);
    if (p >= n) return;
    const int dof = R->dof, NL = dof + 1;
    double T[12];
    for (int k = 0; k < 12; ++k) T[k] = (double) R->base[k];
    for (int l = 0; l < NL; ++l) {
        if (l > 0) {
            const int i = l - 1;
            double s, c, L[12], Tn[12];
            x64::sincos((double) poses[(size_t) p * dof + i], &s, &c);
            {
#pragma clang fp contract(off)
                for (int k = 0; k < 12; ++k)
                    L[k] = (c * (double) R->cmask[i * 12 + k] + s * (double) R->smask[i * 12 + k]) +
                           (double) R->omask[i * 12 + k];
            }
            x64::affine_mul(T, L, Tn);
            for (int k = 0; k < 12; ++k) T[k] = Tn[k];
        }
        const size_t pl = (size_t) p * NL + l;
        for (int k = 0; k < 12; ++k) T64[pl * 12 + k] = T[k];
        // rows of R^T with the world translation in w: p_link = R^T (p - t)
        for (int j = 0; j < 3; ++j)
            Ti[pl * 3 + j] = make_float4((float) T[0 * 4 + j], (float) T[1 * 4 + j], (float) T[2 * 4 + j],
                                         (float) T[j * 4 + 3]);
        float lo[3], hi[3];
        const LinkInfo& L = R->link[l];
        const float b[9] = {L.c[0], L.c[1], L.c[2], L.lo[0], L.lo[1], L.lo[2], L.hi[0], L.hi[1], L.hi[2]};
        for (int r = 0; r < 3; ++r) link_box_row(T, b, r, lo[r], hi[r]);
        wbox[pl * 2] = make_float4(lo[0], lo[1], lo[2], 0.f);
        wbox[pl * 2 + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
}

// One wave per pose, for small batches, where the serial FP64 chain of a single thread is the whole
// latency: lane i < dof evaluates the sine and cosine of joint i, lane k < 12 owns entry k of the
// running transform (the products of the chain), lanes 0..2 own the rows of Ti and of the world box.
// Every value comes from the same expressions in the same order as in k_fk, so both kernels write
// identical bits.
__global__ __launch_bounds__(kBS) void k_fk_wave(const RobotDev* __restrict__ R, const float* __restrict__ poses,
                                                int n, double* __restrict__ T64, float4* __restrict__ Ti,
                                                float4* __restrict__ wbox, FkClear z) {
    // the joint masks, the base and the link boxes in LDS: the chain then waits on LDS, not on a global
    // load per joint and link
    __shared__ float cm[12 * kMaxDof], sm[12 * kMaxDof], om[12 * kMaxDof], bs[12], lb[kMaxLinks][9];
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * (kBS / 64) + (threadIdx.x >> 6);  // uniform per wave
    fk_clear(z, p, lane == 0);
    const int dof = R->dof, NL = dof + 1;
    for (int i = threadIdx.x; i < 12 * dof; i += kBS) {
        cm[i] = R->cmask[i];
        sm[i] = R->smask[i];
        om[i] = R->omask[i];
    }
    if (threadIdx.x < 12) bs[threadIdx.x] = R->base[threadIdx.x];
    for (int i = threadIdx.x; i < 9 * NL; i += kBS) {
        const LinkInfo& L = R->link[i / 9];
        const int k = i % 9;
        lb[i / 9][k] = k < 3 ? L.c[k] : k < 6 ? L.lo[k - 3] : L.hi[k - 6];
    }
    __syncthreads();
    if (p >= n) return;
    double sj = 0.0, cj = 0.0;
    if (lane < dof) x64::sincos((double) poses[(size_t) p * dof + lane], &sj, &cj);
    const int k = lane < 12 ? lane : 0;  // lanes 12.. shadow entry 0
    const int r = k >> 2, cc = k & 3;
    double T = (double) bs[k];
    for (int l = 0; l < NL; ++l) {
        if (l > 0) {
            const int i = l - 1;
            const double s = __shfl(sj, i, 64), c = __shfl(cj, i, 64);
            const double a0 = __shfl(T, r * 4 + 0, 64), a1 = __shfl(T, r * 4 + 1, 64), a2 = __shfl(T, r * 4 + 2, 64),
                         a3 = __shfl(T, r * 4 + 3, 64);
            {
#pragma clang fp contract(off)
                // column cc of the joint transform (rows 0..2), then entry (r, cc) of T * L
                const int m = i * 12 + cc;
                const double L0 = (c * (double) cm[m] + s * (double) sm[m]) + (double) om[m];
                const double L1 = (c * (double) cm[m + 4] + s * (double) sm[m + 4]) + (double) om[m + 4];
                const double L2 = (c * (double) cm[m + 8] + s * (double) sm[m + 8]) + (double) om[m + 8];
                double v = (a0 * L0 + a1 * L1) + a2 * L2;
                if (cc == 3) v = v + a3;
                T = v;
            }
        }
        double Tf[12];
        for (int j = 0; j < 12; ++j) Tf[j] = __shfl(T, j, 64);
        const size_t pl = (size_t) p * NL + l;
        if (lane < 12) T64[pl * 12 + lane] = T;
        float lo = 0.f, hi = 0.f;
        if (lane < 3) {
            Ti[pl * 3 + lane] = make_float4((float) Tf[0 * 4 + lane], (float) Tf[1 * 4 + lane], (float) Tf[2 * 4 + lane],
                                            (float) Tf[lane * 4 + 3]);
            link_box_row(Tf, lb[l], lane, lo, hi);
        }
        const float lo1 = __shfl(lo, 1, 64), lo2 = __shfl(lo, 2, 64), hi1 = __shfl(hi, 1, 64), hi2 = __shfl(hi, 2, 64);
        if (lane == 0) {
            wbox[pl * 2] = make_float4(lo, lo1, lo2, 0.f);
            wbox[pl * 2 + 1] = make_float4(hi, hi1, hi2, 0.f);
        }
    }
}

// ------------------------------------------------------------------------------------------ broad
// Child box (lo, hi) of a scene node against link (p, l): world AABB test and link-frame test.
__device__ __forceinline__ bool broad_test(float lx, float ly, float lz, float hx, float hy, float hz, float4 wlo,
                                           float4 whi, const float4& r0, const float4& r1, const float4& r2,
                                           const float3& lc, const float3& lh, float pad) {
    if (lx > whi.x || hx < wlo.x || ly > whi.y || hy < wlo.y || lz > whi.z || hz < wlo.z) return false;
    float cx = 0.5f * (lx + hx), cy = 0.5f * (ly + hy), cz = 0.5f * (lz + hz);
    float ex = 0.5f * (hx - lx), ey = 0.5f * (hy - ly), ez = 0.5f * (hz - lz);
    float3 c = to_link(r0, r1, r2, make_float3(cx, cy, cz));
    float hx2 = fabsf(r0.x) * ex + fabsf(r0.y) * ey + fabsf(r0.z) * ez;
    float hy2 = fabsf(r1.x) * ex + fabsf(r1.y) * ey + fabsf(r1.z) * ez;
    float hz2 = fabsf(r2.x) * ex + fabsf(r2.y) * ey + fabsf(r2.z) * ez;
    return fabsf(c.x - lc.x) <= hx2 + lh.x + pad && fabsf(c.y - lc.y) <= hy2 + lh.y + pad &&
           fabsf(c.z - lc.z) <= hz2 + lh.z + pad;
}

// Broad phase, pass `pass` of a.n_passes.  Pass 0: every (pose, link) task, split over F frontier
// entries, walks the scene triangle LBVH and appends each leaf that passes both box tests to the
// candidate list.  A walk that has visited a.budget nodes in a pass that may hand off work appends
// its pending subtrees (the next node and its stack) as (pose * NL + link, node) to the queue of the
// next pass and ends; passes 1.. run those subtrees, so that a link deep inside dense geometry is
// walked by many threads instead of holding one wave.  The last pass never hands off.
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kBS) void k_broad(DiscreteArgs a, int pass) {
    __shared__ int stk[STACK * kBS];
    __shared__ uint32_t wcnt[kBS / 64];
    __shared__ uint32_t tris_block;
    int* st = stk + threadIdx.x;
    if (threadIdx.x < kBS / 64) wcnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) tris_block = 0;
    __syncthreads();
    uint32_t* my_wcnt = wcnt + (threadIdx.x >> 6);
    // Block b owns candidate segment b.  In pass 0 it walks one contiguous chunk of the (task,
    // frontier entry) items, so the segments in order keep the link-major task order (poses flagged
    // through an earlier link skip the later ones in the narrow phase); later passes stride.
    // Later passes rotate the block -> segment map by 167 segments per pass: a short hand-off queue runs
    // on the first few blocks, and without the rotation their segments collected the records of every
    // pass (a 4,096-pose dense check needed 8,192 records in one segment for 765,710 in all).
    const size_t seg = pass == 0 ? blockIdx.x : (blockIdx.x + (size_t) pass * 167u) % kSeg;
    uint32_t* seg_count = a.seg_count + seg * kSegStride;
    uint2* seg_cand = a.cand + seg * a.seg_cap;
    const uint2* in_q = pass > 0 ? ((pass - 1) & 1 ? a.spill[1] : a.spill[0]) : nullptr;
    uint2* out_q = (pass & 1 ? a.spill[1] : a.spill[0]);
    uint32_t* out_cnt = a.spill_count + pass;
    const bool may_spill = pass < a.n_passes - 1;
    const uint32_t budget = pass == 0 ? a.budget : a.budget_next;
    size_t beg, end, step;
    if (pass == 0) {
        const size_t n_items = (size_t) a.n_active * a.n_poses * a.F;
        const size_t chunk = ((n_items + gridDim.x - 1) / gridDim.x + kBS - 1) / kBS * kBS;
        beg = (size_t) blockIdx.x * chunk;
        end = min(beg + chunk, n_items);
        step = kBS;
    } else {
        beg = (size_t) blockIdx.x * kBS;
        end = min(a.spill_count[pass - 1], a.spill_cap);
        step = (size_t) gridDim.x * kBS;
    }
    const float pad = a.broad_pad;
    uint32_t ntris = 0, c_nodes = 0;
    auto emit = [&](uint32_t key, int start, int n, uint32_t o) {
        if (o < a.seg_cap) seg_cand[o] = make_uint2(key, pack_leaf(start, n));
        else atomicOr(a.overflow, 1u);
        ntris += (uint32_t) n;
    };
    for (size_t id = beg + threadIdx.x; id < end; id += step) {
        uint32_t key;
        int f = 0, ni = 0;
        if (pass == 0) {
            // item = (task, frontier entry); task = (link-major) (link, pose), distal links first: a
            // serial arm's outer links sweep the most space, so their records flag the most poses
            // early and the narrow passes skip the later records of those poses (dense 65,536-pose
            // check: obstacle->robot 0.96 -> 0.94 ms)
            const int task = (int) (id / (size_t) a.F);
            f = (int) (id - (size_t) task * a.F);
            const int li = task / a.n_poses, p = task - li * a.n_poses;
            key = (uint32_t) ((size_t) p * a.n_links + a.l0 + (a.n_active - 1 - li));
        } else {
            const uint2 r = in_q[id];
            if (r.x == 0xffffffffu) continue;  // slot of a hand-off that did not fit
            key = r.x;
            ni = (int) r.y;
            // a split check's early obstacle->robot pass may have flagged the pose: its walks end
            if (load_flag(a.flags + key / (uint32_t) a.n_links)) continue;
        }
        const int l = (int) (key % (uint32_t) a.n_links);
        const size_t pl = key;
        const float4 wlo = a.wbox[pl * 2], whi = a.wbox[pl * 2 + 1];
        const float4 r0 = a.Ti[pl * 3], r1 = a.Ti[pl * 3 + 1], r2 = a.Ti[pl * 3 + 2];
        const LinkInfo& L = a.robot->link[l];
        const float3 lc = make_float3(L.c[0], L.c[1], L.c[2]);
        const float3 lh =
            make_float3(0.5f * (L.hi[0] - L.lo[0]), 0.5f * (L.hi[1] - L.lo[1]), 0.5f * (L.hi[2] - L.lo[2]));
        if (pass == 0) {
            const FEntry fe = a.front[f];
            const bool go = a.sc.tri_has_root && (L.n_tri > 0 || L.n_edge > 0) &&
                            broad_test(fe.lo[0], fe.lo[1], fe.lo[2], fe.hi[0], fe.hi[1], fe.hi[2], wlo, whi, r0, r1,
                                       r2, lc, lh, pad);
            if (!go) continue;
            if (fe.n > 0) {
                emit(key, fe.c, fe.n, wave_alloc(my_wcnt, seg_count, 1u));
                continue;
            }
            ni = fe.c;
        }
        bool spill_ok = may_spill;
        uint32_t work = 0;
        int sp = 0;
        for (;;) {
            if (spill_ok && work >= budget) {
                const uint32_t want = (uint32_t) sp + 1u;
                const uint32_t o = atomicAdd(out_cnt, want);
                if ((uint64_t) o + want <= a.spill_cap) {
                    out_q[o] = make_uint2(key, (uint32_t) ni);
                    for (int j = 0; j < sp; ++j) out_q[o + 1 + j] = make_uint2(key, (uint32_t) st[j * kBS]);
                    break;
                }
                // queue full: mark the reserved slots empty and finish this subtree here (bit 16 lets
                // the host grow the queue for later calls)
                for (uint64_t j = o; j < a.spill_cap && j < (uint64_t) o + want; ++j)
                    out_q[j] = make_uint2(0xffffffffu, 0u);
                atomicOr(a.overflow, 16u);
                spill_ok = false;
            }
            const BNode nd = a.sc.tri_nodes[ni];
            ++work;
            if (COUNT) c_nodes++;
            bool h0 = nd.c0 >= 0 && broad_test(nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z, wlo, whi, r0, r1,
                                               r2, lc, lh, pad);
            bool h1 = nd.c1 >= 0 && broad_test(nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z, wlo, whi, r0, r1,
                                               r2, lc, lh, pad);
            const bool e0 = h0 && nd.n0 > 0, e1 = h1 && nd.n1 > 0;
            if (e0 || e1) {
                uint32_t o = wave_alloc(my_wcnt, seg_count, (uint32_t) e0 + (uint32_t) e1);
                if (e0) {
                    emit(key, nd.c0, nd.n0, o++);
                    h0 = false;
                }
                if (e1) {
                    emit(key, nd.c1, nd.n1, o);
                    h1 = false;
                }
            }
            if (h0 && h1) {
                if (sp >= STACK) { atomicOr(a.overflow, 2u); break; }
                st[sp * kBS] = nd.c1;
                ++sp;
                ni = nd.c0;
            } else if (h0) {
                ni = nd.c0;
            } else if (h1) {
                ni = nd.c1;
            } else {
                if (sp == 0) break;
                --sp;
                ni = st[sp * kBS];
            }
        }
    }
    if (ntris) atomicAdd(&tris_block, ntris);
    if (COUNT) atomicAdd(&a.stats->broad_nodes, (unsigned long long) c_nodes);
    __syncthreads();
    if (threadIdx.x == 0 && tris_block) atomicAdd(a.tri_candidates, (unsigned long long) tris_block);
}

// Last broad pass of a check with hand-offs (a.broad_share): the few subtrees still pending (36,084
// of them in pass 5 of a dense 65,536-pose check, 51 in pass 9) are long walks deep inside dense
// geometry, and one thread per subtree in hand-off passes of 8 nodes left most of the GPU idle for
// five or six launches of about 20 us each.  Here persistent lanes walk them to the end instead: wave
// w takes subtrees w, w + W, w + 2W, ... (W waves), so a wave holds a few of them, every lane walks
// its own and, when the wave has no subtree left to start, idle lanes take the bottom pending subtree
// of busy lanes (the one nearest the walk's start, so the largest) through the wave's LDS slots.
// Every subtree is still walked once: the records are those of the one-thread walks, in another order.
#ifndef RTCD_BROAD_STEPS
#define RTCD_BROAD_STEPS 1
#endif
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kBS) void k_broad_share(DiscreteArgs a, int pass) {
    constexpr int kSteps = RTCD_BROAD_STEPS;  // node visits between refill checks
    __shared__ int stk[STACK * kBS];
    __shared__ uint2 bq[kBS / 64][128];
    __shared__ uint2 bsh[kBS / 64][64];
    __shared__ uint32_t wcnt[kBS / 64];
    __shared__ uint32_t tris_block;
    int* st = stk + threadIdx.x;
    const int wv = (int) (threadIdx.x >> 6), lane = (int) (threadIdx.x & 63);
    const uint64_t below = (1ull << lane) - 1ull;
    if (threadIdx.x < kBS / 64) wcnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) tris_block = 0;
    __syncthreads();
    uint32_t* my_wcnt = wcnt + wv;
    uint2* q = bq[wv];
    uint2* sh = bsh[wv];
    // segments as in k_broad's later passes
    const size_t seg = (blockIdx.x + (size_t) pass * 167u) % kSeg;
    uint32_t* seg_count = a.seg_count + seg * kSegStride;
    uint2* seg_cand = a.cand + seg * a.seg_cap;
    const uint2* in_q = (pass - 1) & 1 ? a.spill[1] : a.spill[0];
    const uint32_t n_items = min(a.spill_count[pass - 1], a.spill_cap);
    const uint32_t n_waves = gridDim.x * (kBS / 64), wave = blockIdx.x * (kBS / 64) + (uint32_t) wv;
    uint32_t k_next = 0;  // the wave's next subtree is wave + k_next * n_waves
    const uint32_t NL = (uint32_t) a.n_links;
    const float pad = a.broad_pad;
    uint32_t ntris = 0, c_nodes = 0;
    auto emit = [&](uint32_t k, int start, int n, uint32_t o) {
        if (o < a.seg_cap) seg_cand[o] = make_uint2(k, pack_leaf(start, n));
        else atomicOr(a.overflow, 1u);
        ntris += (uint32_t) n;
    };
    // lane state: the walk, its pending subtrees st[sb .. sp) and the boxes of its (pose, link) task
    bool act = false;
    uint32_t key = 0;
    int ni = 0, sp = 0, sb = 0;
    float4 wlo = {}, whi = {}, r0 = {}, r1 = {}, r2 = {};
    float3 lc = {}, lh = {};
    auto start = [&](uint2 e) {
        key = e.x;
        ni = (int) e.y;
        wlo = a.wbox[(size_t) key * 2];
        whi = a.wbox[(size_t) key * 2 + 1];
        r0 = a.Ti[(size_t) key * 3];
        r1 = a.Ti[(size_t) key * 3 + 1];
        r2 = a.Ti[(size_t) key * 3 + 2];
        const LinkInfo& L = a.robot->link[key % NL];
        lc = make_float3(L.c[0], L.c[1], L.c[2]);
        lh = make_float3(0.5f * (L.hi[0] - L.lo[0]), 0.5f * (L.hi[1] - L.lo[1]), 0.5f * (L.hi[2] - L.lo[2]));
        sp = sb = 0;
        act = true;
    };
    int qn = 0;            // queued subtrees (wave-uniform)
    bool drained = false;  // the wave's subtrees are all taken (wave-uniform)
    for (;;) {
        const uint64_t idle = __ballot(!act);
        const int n_idle = __popcll(idle);
        if (n_idle > 0) {
            while (qn < n_idle && !drained) {
                const uint64_t i = (uint64_t) wave + ((uint64_t) k_next + (uint64_t) lane) * n_waves;
                if ((uint64_t) wave + (uint64_t) k_next * n_waves >= n_items) {
                    drained = true;
                    break;
                }
                k_next += 64;
                bool live = false;
                uint2 e = make_uint2(0u, 0u);
                if (i < n_items) {
                    e = in_q[i];
                    // a slot of a hand-off that did not fit, or a pose an obstacle->robot part flagged
                    live = e.x != 0xffffffffu && !load_flag(a.flags + e.x / NL);
                }
                const uint64_t m = __ballot(live);
                if (live) q[qn + __popcll(m & below)] = e;
                qn += __popcll(m);
            }
            __builtin_amdgcn_wave_barrier();
            // idle lane number r (in lane order) takes queue entry qn - 1 - r
            const int r = __popcll(idle & below);
            const int take = min(n_idle, qn);
            const bool mine = !act && r < take;
            uint2 e = make_uint2(0u, 0u);
            if (mine) e = q[qn - 1 - r];
            qn -= take;
            __builtin_amdgcn_wave_barrier();
            if (mine) start(e);
            if (qn == 0 && drained) {
                // no subtree left to start: the idle lanes take the bottom pending subtree of busy lanes
                const uint64_t idle2 = __ballot(!act), donors = __ballot(act && sp > sb);
                const int n_give = min(__popcll(idle2), __popcll(donors));
                if (n_give > 0) {
                    const int dr = __popcll(donors & below);
                    if (act && sp > sb && dr < n_give) {
                        sh[dr] = make_uint2(key, (uint32_t) st[sb * kBS]);
                        if (++sb == sp) sb = sp = 0;
                    }
                    __builtin_amdgcn_wave_barrier();
                    const int ir = __popcll(idle2 & below);
                    const bool taker = !act && ir < n_give;
                    uint2 g = make_uint2(0u, 0u);
                    if (taker) g = sh[ir];
                    __builtin_amdgcn_wave_barrier();
                    if (taker) start(g);
                }
                if (__ballot(act) == 0) break;
            }
        }
        for (int s = 0; s < kSteps && act; ++s) {
            const BNode nd = a.sc.tri_nodes[ni];
            if (COUNT) c_nodes++;
            bool h0 = nd.c0 >= 0 && broad_test(nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z, wlo, whi, r0, r1,
                                               r2, lc, lh, pad);
            bool h1 = nd.c1 >= 0 && broad_test(nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z, wlo, whi, r0, r1,
                                               r2, lc, lh, pad);
            const bool e0 = h0 && nd.n0 > 0, e1 = h1 && nd.n1 > 0;
            if (e0 || e1) {
                uint32_t o = wave_alloc(my_wcnt, seg_count, (uint32_t) e0 + (uint32_t) e1);
                if (e0) {
                    emit(key, nd.c0, nd.n0, o++);
                    h0 = false;
                }
                if (e1) {
                    emit(key, nd.c1, nd.n1, o);
                    h1 = false;
                }
            }
            if (h0 && h1) {
                if (sp >= STACK) {
                    atomicOr(a.overflow, 2u);
                    act = false;
                    break;
                }
                st[sp * kBS] = nd.c1;
                ++sp;
                ni = nd.c0;
            } else if (h0) {
                ni = nd.c0;
            } else if (h1) {
                ni = nd.c1;
            } else if (sp == sb) {
                act = false;
            } else {
                ni = st[--sp * kBS];
                if (sp == sb) sb = sp = 0;
            }
        }
    }
    if (ntris) atomicAdd(&tris_block, ntris);
    if (COUNT) atomicAdd(&a.stats->broad_nodes, (unsigned long long) c_nodes);
    __syncthreads();
    if (threadIdx.x == 0 && tris_block) atomicAdd(a.tri_candidates, (unsigned long long) tris_block);
}

// ------------------------------------------------------------------------------------------ gather
// Exclusive prefix of the segment counts (each clamped to seg_cap): one block, seg_pre[kSeg] = total.
// Split checks gather in parts: seg_base[b] holds the records of segment b gathered so far and
// seg_base[kSeg] their total.  `gpart` 1 (first part) sets them; `gpart` 2 scans only the records
// added since and places them after the earlier parts (seg_pre[0] = first new record), then updates
// them.  Each word of seg_base is read and written by one thread only.
__global__ __launch_bounds__(kBS) void k_seg_scan(const uint32_t* __restrict__ seg_count, uint32_t seg_cap,
                                                 uint32_t* __restrict__ seg_pre, uint32_t* __restrict__ seg_base,
                                                 int gpart) {
    constexpr int per = kSeg / kBS;
    __shared__ uint32_t part[kBS / 64];
    __shared__ uint32_t start_s;
    if (threadIdx.x == 0) start_s = gpart == 2 ? seg_base[kSeg] : 0u;
    uint32_t loc[per], acc = 0;
    for (int k = 0; k < per; ++k) {
        const int b = threadIdx.x * per + k;
        const uint32_t cnt = min(seg_count[(size_t) b * kSegStride], seg_cap);
        loc[k] = gpart == 2 ? cnt - seg_base[b] : cnt;
        if (gpart) seg_base[b] = cnt;
        acc += loc[k];
    }
    // inclusive scan of the per-thread sums within each wave, then across the waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = acc;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) part[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kBS / 64; ++w) {
        before += w < wave ? part[w] : 0u;
        all += part[w];
    }
    const uint32_t start = start_s;
    if (threadIdx.x == 0) {
        seg_pre[kSeg] = start + all;
        if (gpart) seg_base[kSeg] = start + all;
    }
    uint32_t run = start + before + inc - acc;
    for (int k = 0; k < per; ++k) {
        seg_pre[threadIdx.x * per + k] = run;
        run += loc[k];
    }
}

// Small batches: both steps in one launch.  Every block builds the whole prefix in LDS (block 0 also
// stores it), then the blocks copy the records grid-stride, each finding its segment by binary search.
__global__ __launch_bounds__(kBS) void k_seg_scan_gather(const uint2* __restrict__ cand,
                                                        const uint32_t* __restrict__ seg_count, uint32_t seg_cap,
                                                        uint32_t* __restrict__ seg_pre, uint2* __restrict__ flat) {
    constexpr int per = kSeg / kBS;
    __shared__ uint32_t pre[kSeg + 1];
    __shared__ uint32_t part[kBS / 64];
    uint32_t loc[per], acc = 0;
    for (int k = 0; k < per; ++k) {
        loc[k] = min(seg_count[(size_t) (threadIdx.x * per + k) * kSegStride], seg_cap);
        acc += loc[k];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = acc;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) part[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < kBS / 64; ++w) {
        before += w < wave ? part[w] : 0u;
        all += part[w];
    }
    uint32_t run = before + inc - acc;
    for (int k = 0; k < per; ++k) {
        pre[threadIdx.x * per + k] = run;
        run += loc[k];
    }
    if (threadIdx.x == 0) pre[kSeg] = all;
    __syncthreads();
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b <= kSeg; b += kBS) seg_pre[b] = pre[b];
    for (uint32_t i = blockIdx.x * kBS + threadIdx.x; i < all; i += gridDim.x * kBS) {
        int lo = 0, hi = kSeg;  // the segment b with pre[b] <= i < pre[b + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[mid] <= i) lo = mid;
            else hi = mid;
        }
        flat[i] = cand[(size_t) lo * seg_cap + (i - pre[lo])];
    }
}

// Block b copies the newest seg_pre[b + 1] - seg_pre[b] records of segment b (all of them, unless a
// split check gathered some before) to flat[seg_pre[b] ..): the narrow phase reads records by index.
__global__ __launch_bounds__(kBS) void k_seg_gather(const uint2* __restrict__ cand, const uint32_t* __restrict__ seg_count,
                                                   uint32_t seg_cap, const uint32_t* __restrict__ seg_pre,
                                                   uint2* __restrict__ flat) {
    const uint32_t b = blockIdx.x;
    const uint32_t base = seg_pre[b], n = seg_pre[b + 1] - base;
    const uint2* src = cand + (size_t) b * seg_cap + (min(seg_count[(size_t) b * kSegStride], seg_cap) - n);
    for (uint32_t j = threadIdx.x; j < n; j += kBS) flat[(size_t) base + j] = src[j];
}

// After the obstacle->robot pass: the records of poses it left unflagged into a.r2o_list (block b
// compacts the b-th contiguous chunk of the gathered list: one sweep counts, one global atomic per
// block reserves, a second sweep writes with block-wide ranks).  Each block's records keep their order;
// the blocks' ranges follow the order of their reservations, which the robot->obstacle pass does not
// depend on (it reads the list as a set).  KEEP_FLAGGED keeps the records of flagged poses instead
// (the other polarity, used by the test entry point only).
template <bool KEEP_FLAGGED>
__global__ __launch_bounds__(kBS) void k_compact(DiscreteArgs a) {
    __shared__ uint32_t wsum[kBS / 64];
    __shared__ uint32_t base;
    const uint32_t total = a.seg_pre[kSeg];
    const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t beg = min(total, blockIdx.x * chunk), end = min(total, beg + chunk);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto keep = [&](uint32_t i) {
        return i < end && (load_flag(a.flags + a.flat[i].x / (uint32_t) a.n_links) != 0u) == KEEP_FLAGGED;
    };
    uint32_t mine = 0;
    for (uint32_t i = beg + threadIdx.x; i < end; i += kBS) mine += keep(i) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) mine += __shfl_down(mine, o, 64);
    if (lane == 0) wsum[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBS / 64; ++w) t += wsum[w];
        base = t ? atomicAdd(a.r2o_count, t) : 0u;
    }
    __syncthreads();
    uint32_t run = base;
    for (uint32_t it = beg; it < end; it += kBS) {
        const uint32_t i = it + threadIdx.x;
        const bool k = keep(i);
        const uint64_t m = __ballot(k);
        const uint32_t rank = (uint32_t) __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wsum[wave] = (uint32_t) __popcll(m);
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int w = 0; w < kBS / 64; ++w) {
            before += w < (int) wave ? wsum[w] : 0u;
            all += wsum[w];
        }
        if (k) a.r2o_list[run + before + rank] = a.flat[i];
        run += all;
    }
}

// ------------------------------------------------------------------------------------------ narrow
struct Ctr {
    uint32_t nodes = 0, prims = 0, fp64 = 0, edges = 0, tris = 0;
};

// Work hand-off of a narrow traversal: once a walk has visited `budget` nodes in a pass that may
// hand off, its pending subtrees (the next node and its stack) are appended to `out` as
// (pose * NL + link, scene triangle, link-BVH node, edge j of the triangle or 3) and the walk ends;
// the next pass resumes them.  A full queue only stops hand-offs.
struct NarrowSpill {
    uint4* out;
    uint32_t* count;
    uint32_t cap, budget;
    bool may;
    uint32_t* full;  // gets bit 8 when a hand-off did not fit (the host then grows the queue)
};

// Appends the pending subtrees (ni and the stack) of a walk; false if the queue is full (the slots
// it reserved are marked empty and the walk goes on by itself).
template <int BS>
__device__ __forceinline__ bool spill_walk(const NarrowSpill& sp_, uint32_t key, uint32_t tri, uint32_t j, int ni,
                                           const int* st, int sp) {
    const uint32_t want = (uint32_t) sp + 1u;
    const uint32_t o = atomicAdd(sp_.count, want);
    if ((uint64_t) o + want <= sp_.cap) {
        sp_.out[o] = make_uint4(key, tri, (uint32_t) ni, j);
        for (int k = 0; k < sp; ++k) sp_.out[o + 1 + k] = make_uint4(key, tri, (uint32_t) st[k * BS], j);
        return true;
    }
    for (uint64_t k = o; k < sp_.cap && k < (uint64_t) o + want; ++k) sp_.out[k] = make_uint4(0xffffffffu, 0u, 0u, 0u);
    atomicOr(sp_.full, 8u);
    return false;
}

// A test whose FP32 signs fall inside their band is queued for the FP64 kernel: (pose*NL + link,
// scene triangle, robot primitive, kind) with kind 0..2 = obstacle edge j of the triangle (O2R),
// 3 = robot edge (R2O).
__device__ __forceinline__ void defer(const DiscreteArgs& a, uint32_t key, uint32_t tri, uint32_t prim, uint32_t kind) {
    const uint32_t slot = atomicAdd(a.unsure_count, 1u);
    if (slot < a.unsure_cap) a.unsure[slot] = make_uint4(key, tri, prim, kind);
    else atomicOr(a.overflow, 4u);
}

enum { kVisitDone = 0, kVisitMore = 1, kVisitHit = 2 };

// Persistent narrow lanes: a wave takes new candidates once this many lanes are idle, and checks
// after this many node visits of its active lanes (measured on MI355X; a build may override them).
// Robot->obstacle walks are short (2.3 link-BVH nodes per candidate on the dense scene), so its waves
// check after every visit: dense two-way check, robot->obstacle stage 0.50 -> 0.38 ms at 65,536 poses,
// 0.23 -> 0.13 ms at 8,192, against 4 visits (2: 0.41 / 0.17; 8: 0.64 / 0.32).  The refill threshold
// (8, 16 or 32 lanes) changed neither direction; obstacle->robot walks (3.5 nodes) keep 4 visits
// (2, 3 and 6 were no faster).
#ifndef RTCD_NARROW_REFILL
#define RTCD_NARROW_REFILL 16
#endif
#ifndef RTCD_NARROW_STEPS
#define RTCD_NARROW_STEPS 4
#endif
#ifndef RTCD_R2O_REFILL
#define RTCD_R2O_REFILL RTCD_NARROW_REFILL
#endif
#ifndef RTCD_R2O_STEPS
#define RTCD_R2O_STEPS 1
#endif

// Walk of obstacle edge e of a scene triangle (link-frame origin P with abs. error <= delta per
// component, direction d = R^T (Q - P) with rel. error <= 6 eps) through the link-triangle LBVH.
struct O2RWalk {
    float3 P, d, inv;
    float pad;
    MTBands bd;
    int node0, prim0;
    uint32_t e;
};

// Sets up the walk; false if the segment misses the link box.
__device__ __forceinline__ bool o2r_begin(const LinkInfo& L, float3 P, float3 d, float delta, uint32_t e, O2RWalk& w) {
    w.P = P;
    w.d = d;
    w.e = e;
    w.inv = safe_inv(d);
    const float D = l1(d.x, d.y, d.z) * (1.f + 8.f * kEps32);
    const float ad = 6.1f * kEps32 * D;
    w.pad = 2.f * delta + 2.f * ad + 1e-6f * L.scale;
    if (!slab(P, w.inv, L.lo[0], L.lo[1], L.lo[2], L.hi[0], L.hi[1], L.hi[2], w.pad)) return false;
    const float S = l1(P.x - L.c[0], P.y - L.c[1], P.z - L.c[2]) * (1.f + 2.f * kEps32) + 3.f * delta + L.rv;
    w.bd = make_bands(D, ad, S, delta + kEps32 * S, L.e1max, kEps32 * L.e1max, L.e2max, kEps32 * L.e2max);
    w.node0 = L.tri_node0;
    w.prim0 = L.tri_prim0;
    return true;
}

// One node visit: both child boxes, the link triangles of the children that are leaves (banded
// FP32 test, unsure ones queued for FP64), then the next node from the children or the LDS stack.
template <int STACK, int BS, bool COUNT>
__device__ __forceinline__ int o2r_visit(const DiscreteArgs& a, const O2RWalk& w, uint32_t key, uint32_t tri, int& ni,
                                         int* st, int& sp, Ctr& ctr) {
    const BNode nd = a.rb.tri_nodes[w.node0 + ni];
    if (COUNT) ctr.nodes++;
    bool h0 = nd.c0 >= 0 && slab(w.P, w.inv, nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z, w.pad);
    bool h1 = nd.c1 >= 0 && slab(w.P, w.inv, nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z, w.pad);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const bool hs = side ? h1 : h0;
        const int cnt = side ? nd.n1 : nd.n0;
        const int first = side ? nd.c1 : nd.c0;
        if (hs && cnt > 0) {
            for (int k = 0; k < cnt; ++k) {
                const Tri t = a.rb.tri_prims[w.prim0 + first + k];
                if (COUNT) ctr.prims++;
                const float3 ta = f3(t.a);
                const int r = mt32(w.P, w.d, ta, sub(f3(t.b), ta), sub(f3(t.c), ta), w.bd);
                if (r == kHit) return kVisitHit;
                if (r == kUnsure) {
                    if (COUNT) ctr.fp64++;
                    defer(a, key, tri, (uint32_t) (w.prim0 + first + k), w.e);
                }
            }
            if (side) h1 = false; else h0 = false;
        }
    }
    if (h0 && h1) {
        if (sp >= STACK) {
            atomicOr(a.overflow, 2u);
            return kVisitDone;
        }
        st[sp * BS] = nd.c1;
        ++sp;
        ni = nd.c0;
    } else if (h0) {
        ni = nd.c0;
    } else if (h1) {
        ni = nd.c1;
    } else {
        if (sp == 0) return kVisitDone;
        --sp;
        ni = st[sp * BS];
    }
    return kVisitMore;
}

// Walk of a scene triangle (link frame A with abs. error <= delta, e1 = R^T (B - A), e2 = R^T (C - A)
// with rel. error <= 6 eps) through the link-edge LBVH.
struct R2OWalk {
    float3 A, e1, e2, nrm;
    float delta, lx, ly, lz, hx, hy, hz, nl, ne;
    float sg;  // orientation of the triangle's shadow on the plane of the axes other than ax (0: none)
    MTBands bd;
    int node0, prim0, ax;
};

__device__ __forceinline__ float pick(const float3& p, int k) { return k == 0 ? p.x : k == 1 ? p.y : p.z; }

// Sets up the walk; false if the triangle's box misses the link box.
__device__ __forceinline__ bool r2o_begin(const LinkInfo& L, float3 A, float3 e1, float3 e2, float delta, R2OWalk& w) {
    const float3 B = add(A, e1), C = add(A, e2);
    const float pad = 2.f * delta + 12.f * kEps32 * (l1(e1.x, e1.y, e1.z) + l1(e2.x, e2.y, e2.z)) + 1e-6f * L.scale;
    w.lx = fminf(fminf(A.x, B.x), C.x) - pad;
    w.ly = fminf(fminf(A.y, B.y), C.y) - pad;
    w.lz = fminf(fminf(A.z, B.z), C.z) - pad;
    w.hx = fmaxf(fmaxf(A.x, B.x), C.x) + pad;
    w.hy = fmaxf(fmaxf(A.y, B.y), C.y) + pad;
    w.hz = fmaxf(fmaxf(A.z, B.z), C.z) + pad;
    if (w.lx > L.hi[0] || w.hx < L.lo[0] || w.ly > L.hi[1] || w.hy < L.lo[1] || w.lz > L.hi[2] || w.hz < L.lo[2])
        return false;
    const float E1 = l1(e1.x, e1.y, e1.z) * (1.f + 8.f * kEps32);
    const float E2 = l1(e2.x, e2.y, e2.z) * (1.f + 8.f * kEps32);
    const float S = l1(A.x - L.c[0], A.y - L.c[1], A.z - L.c[2]) * (1.f + 2.f * kEps32) + 3.f * delta + L.rv;
    w.bd = make_bands(L.dmax, kEps32 * L.dmax, S, delta + kEps32 * S, E1, 6.1f * kEps32 * E1, E2, 6.1f * kEps32 * E2);
    // plane of the triangle: a node box that lies strictly on one side cannot hold a crossing edge.
    // n = e1 x e2 and d0 = n.A carry errors bounded by 14 eps E1 E2 (n) and |n| delta (A); the
    // margin in r2o_visit covers them for any box centre c: |Delta(n.(c - A))| <= 14 eps E1 E2 |c - A|_1 + |n|_1 delta.
    w.nrm = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    w.nl = l1(w.nrm.x, w.nrm.y, w.nrm.z);
    w.ne = 32.f * kEps32 * E1 * E2;
    // shadow of the triangle on the plane of the two axes other than its normal's largest one; its
    // orientation counts only where the rounded 2D area is larger than its error bound
    const float ax_ = fabsf(w.nrm.x), ay_ = fabsf(w.nrm.y), az_ = fabsf(w.nrm.z);
    w.ax = ax_ >= ay_ && ax_ >= az_ ? 0 : ay_ >= az_ ? 1 : 2;
    {
        const int iu = w.ax == 2 ? 0 : w.ax + 1, iv = w.ax == 0 ? 2 : w.ax - 1;
        const float Au = pick(A, iu), Av = pick(A, iv);
        const float p1 = (pick(B, iu) - Au) * (pick(C, iv) - Av), p2 = (pick(B, iv) - Av) * (pick(C, iu) - Au);
        const float area = p1 - p2, err = 5.f * kEps32 * (fabsf(p1) + fabsf(p2));
        w.sg = area > err ? 1.f : area < -err ? -1.f : 0.f;
    }
    w.A = A;
    w.e1 = e1;
    w.e2 = e2;
    w.delta = delta;
    w.node0 = L.edge_node0;
    w.prim0 = L.edge_prim0;
    return true;
}

template <int STACK, int BS, bool COUNT>
__device__ __forceinline__ int r2o_visit(const DiscreteArgs& a, const R2OWalk& w, uint32_t key, uint32_t tri, int& ni,
                                         int* st, int& sp, Ctr& ctr) {
    auto straddles = [&](float bx0, float by0, float bz0, float bx1, float by1, float bz1) {
        const float cx = 0.5f * (bx0 + bx1) - w.A.x, cy = 0.5f * (by0 + by1) - w.A.y, cz = 0.5f * (bz0 + bz1) - w.A.z;
        const float hx = 0.5f * (bx1 - bx0), hy = 0.5f * (by1 - by0), hz = 0.5f * (bz1 - bz0);
        const float sd = w.nrm.x * cx + w.nrm.y * cy + w.nrm.z * cz;
        const float rr = fabsf(w.nrm.x) * hx + fabsf(w.nrm.y) * hy + fabsf(w.nrm.z) * hz;
        const float m = w.ne * (fabsf(cx) + fabsf(cy) + fabsf(cz) + hx + hy + hz) + (w.nl + w.ne) * 2.f * w.delta +
                        4.f * kEps32 * (fabsf(sd) + rr);
        return fabsf(sd) <= rr + m;
    };
    // A robot edge inside a node box crosses the triangle at a point of the box, so the box meets the
    // triangle's shadow on the (u, v) plane.  The box, grown by delta (the error of the triangle's
    // vertices), entirely outside one edge line of the shadow holds no crossing.  The box corner that
    // maximises the edge function and the signs are exact; m covers its rounding and the growth.
    auto shadow_misses = [&](float bx0, float by0, float bz0, float bx1, float by1, float bz1) {
        if (w.sg == 0.f) return false;
        const int iu = w.ax == 2 ? 0 : w.ax + 1, iv = w.ax == 0 ? 2 : w.ax - 1;
        const float3 lo = make_float3(bx0, by0, bz0), hi = make_float3(bx1, by1, bz1);
        const float lu = pick(lo, iu), lv = pick(lo, iv), hu = pick(hi, iu), hv = pick(hi, iv);
        const float Au = pick(w.A, iu), Av = pick(w.A, iv);
        const float Bu = Au + pick(w.e1, iu), Bv = Av + pick(w.e1, iv);
        const float Cu = Au + pick(w.e2, iu), Cv = Av + pick(w.e2, iv);
        auto outside = [&](float Pu, float Pv, float Qu, float Qv) {
            const float du = (Qu - Pu) * w.sg, dv = (Qv - Pv) * w.sg;
            // sg * E(x) = du (x_v - P_v) - dv (x_u - P_u), largest at this corner
            const float xu = dv > 0.f ? lu : hu, xv = du > 0.f ? hv : lv;
            const float t1 = du * (xv - Pv), t2 = dv * (xu - Pu);
            const float m = 5.f * kEps32 * (fabsf(t1) + fabsf(t2)) + 1.01f * (fabsf(du) + fabsf(dv)) * w.delta;
            return (t1 - t2) + m < 0.f;
        };
        return outside(Au, Av, Bu, Bv) || outside(Bu, Bv, Cu, Cv) || outside(Cu, Cv, Au, Av);
    };
    const BNode nd = a.rb.edge_nodes[w.node0 + ni];
    if (COUNT) ctr.nodes++;
    bool h0 = nd.c0 >= 0 &&
              !(w.lx > nd.hi0x || w.hx < nd.lo0x || w.ly > nd.hi0y || w.hy < nd.lo0y || w.lz > nd.hi0z || w.hz < nd.lo0z) &&
              straddles(nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z) &&
              !shadow_misses(nd.lo0x, nd.lo0y, nd.lo0z, nd.hi0x, nd.hi0y, nd.hi0z);
    bool h1 = nd.c1 >= 0 &&
              !(w.lx > nd.hi1x || w.hx < nd.lo1x || w.ly > nd.hi1y || w.hy < nd.lo1y || w.lz > nd.hi1z || w.hz < nd.lo1z) &&
              straddles(nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z) &&
              !shadow_misses(nd.lo1x, nd.lo1y, nd.lo1z, nd.hi1x, nd.hi1y, nd.hi1z);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const bool hs = side ? h1 : h0;
        const int cnt = side ? nd.n1 : nd.n0;
        const int first = side ? nd.c1 : nd.c0;
        if (hs && cnt > 0) {
            for (int k = 0; k < cnt; ++k) {
                const Edge e = a.rb.edge_prims[w.prim0 + first + k];
                if (COUNT) {
                    ctr.prims++;
                    ctr.edges++;
                }
                const float3 ep = f3(e.p);
                const int r = mt32(ep, sub(f3(e.q), ep), w.A, w.e1, w.e2, w.bd);
                if (r == kHit) return kVisitHit;
                if (r == kUnsure) {
                    if (COUNT) ctr.fp64++;
                    defer(a, key, tri, (uint32_t) (w.prim0 + first + k), 3u);
                }
            }
            if (side) h1 = false; else h0 = false;
        }
    }
    if (h0 && h1) {
        if (sp >= STACK) {
            atomicOr(a.overflow, 2u);
            return kVisitDone;
        }
        st[sp * BS] = nd.c1;
        ++sp;
        ni = nd.c0;
    } else if (h0) {
        ni = nd.c0;
    } else if (h1) {
        ni = nd.c1;
    } else {
        if (sp == 0) return kVisitDone;
        --sp;
        ni = st[sp * BS];
    }
    return kVisitMore;
}

// A walk from node `start` until it hits, ends, or hands its pending subtrees off (budget `hs`):
// 1 on a hit.  `j` is the walk's kind in the hand-off and FP64 records (obstacle edge 0..2, or 3).
template <int DIR, int STACK, int BS, bool COUNT, typename W>
__device__ __forceinline__ int walk(const DiscreteArgs& a, const W& w, uint32_t key, uint32_t tri, uint32_t j, int start,
                                   int* st, const uint32_t* flag, Ctr& ctr, const NarrowSpill& hs) {
    int sp = 0, ni = start, iter = 0;
    bool spill_ok = hs.may;
    for (;;) {
        if ((++iter & 15) == 0 && load_flag(flag)) return 0;
        if (spill_ok && (uint32_t) iter > hs.budget) {
            if (spill_walk<BS>(hs, key, tri, j, ni, st, sp)) return 0;
            spill_ok = false;
        }
        int r;
        if constexpr (DIR == 1) r = o2r_visit<STACK, BS, COUNT>(a, w, key, tri, ni, st, sp, ctr);
        else r = r2o_visit<STACK, BS, COUNT>(a, w, key, tri, ni, st, sp, ctr);
        if (r == kVisitHit) return 1;
        if (r == kVisitDone) return 0;
    }
}

// Candidate (key, scene triangle) resumed from a hand-off: edge j of the triangle (obstacle->robot)
// or the triangle (robot->obstacle) against the link BVH from node `start`.
template <int DIR, int STACK, int BS, bool COUNT>
__device__ __forceinline__ int resume_candidate(const DiscreteArgs& a, uint32_t key, uint32_t tri, uint32_t j,
                                                int start, int* st, uint32_t* flag, Ctr& ctr,
                                                const NarrowSpill& hs) {
    const LinkInfo& L = a.robot->link[key % (uint32_t) a.n_links];
    const float4 r0 = a.Ti[(size_t) key * 3], r1 = a.Ti[(size_t) key * 3 + 1], r2 = a.Ti[(size_t) key * 3 + 2];
    const Tri t = a.sc.tri_prims[tri];
    const LinkTri m = link_tri(t, r0, r1, r2);
    if constexpr (DIR == 2) {
        R2OWalk w;
        if (!L.edge_has_root || !r2o_begin(L, m.A, m.e1, m.e2, m.delta, w)) return 0;
        return walk<DIR, STACK, BS, COUNT>(a, w, key, tri, 3u, start, st, flag, ctr, hs);
    } else {
        // the edges A->B, B->C, C->A
        const float3 P = j == 0 ? m.A : j == 1 ? add(m.A, m.e1) : add(m.A, m.e2);
        const float3 d = j == 0 ? m.e1 : j == 1 ? rot(r0, r1, r2, sub(f3(t.c), f3(t.b))) : neg(m.e2);
        O2RWalk w;
        if (!L.tri_has_root || !o2r_begin(L, P, d, m.delta, j, w)) return 0;
        return walk<DIR, STACK, BS, COUNT>(a, w, key, tri, j, start, st, flag, ctr, hs);
    }
}

// Pass 0 of one direction with persistent lanes: every lane holds its own candidate (scene
// triangle, pose-link) and, once that candidate is decided or handed off, takes the next one from
// the wave's LDS queue, which the wave tops up with the live candidates of its share of the leaf
// records (a leaf record spreads over `lanes` adjacent lanes, one scene triangle each; candidates of
// poses already flagged never enter the queue).  A wave that ran 64 candidates to completion
// together would keep every lane waiting for the longest walk: walks are short on average (a few
// link-BVH nodes) but vary widely.  Obstacle->robot traces the owned edges of the triangle one after
// the other; robot->obstacle walks the triangle once.
//
// Once the wave's candidates are used up, idle lanes share the walks still running in the wave
// (a.narrow_share): every busy lane with a pending subtree on its stack gives the top one to an idle
// lane through the wave's LDS slots `sh`, as the record a hand-off would write (pose-link, triangle,
// node, edge), and the idle lane walks that subtree with its own stack.  Each subtree is still walked
// once, so the flags and FP64 tests are those of a single walk; a long walk (a large scene triangle
// whose plane cuts a link) then spreads over the wave instead of holding one lane while 63 wait.
template <int DIR, int STACK, int BS, bool COUNT, bool SPILL>
__device__ __forceinline__ void narrow_lanes(const DiscreteArgs& a, const uint2* list, uint32_t total, uint2* q,
                                             uint4* sh, int* st, Ctr& ctr, const NarrowSpill& hs) {
    constexpr int kRefill = DIR == 1 ? RTCD_NARROW_REFILL : RTCD_R2O_REFILL;  // idle lanes that make the wave refill
    constexpr int kSteps = DIR == 1 ? RTCD_NARROW_STEPS : RTCD_R2O_STEPS;     // node visits between refill checks
    const int lane = (int) (threadIdx.x & 63);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t lanes = (uint32_t) a.leaf_lanes;
    const uint32_t gpw = 64u / lanes, g = (uint32_t) lane / lanes, sub_k = (uint32_t) lane - g * lanes;
    const uint32_t wave = blockIdx.x * (BS / 64) + (threadIdx.x >> 6), n_waves = gridDim.x * (BS / 64);
    const uint32_t NL = (uint32_t) a.n_links;
    uint32_t base = wave * gpw;  // next leaf-record group of this wave
    int qn = 0;                  // queued candidates (wave-uniform)
    // lane state: the candidate and its current walk
    bool act = false, spill_ok = false;
    uint32_t key = 0, tri = 0;
    uint32_t* flag = a.flags;
    int sp = 0, ni = 0, iter = 0;
    using W = std::conditional_t<DIR == 1, O2RWalk, R2OWalk>;
    W w = {};
    // obstacle->robot: the triangle in the link frame, its edge B->C and the edges not yet traced
    float3 A = {}, e1 = {}, e2 = {}, d1 = {};
    float delta = 0.f;
    uint32_t own = 0;
    auto begin_walk = [&]() {
        sp = 0;
        ni = 0;
        iter = 0;
        spill_ok = hs.may;
    };
    // obstacle->robot: the next owned edge whose segment meets the link box; false if none is left
    auto next_edge = [&](const LinkInfo& L) -> bool {
        if constexpr (DIR == 1) {
            while (own) {
                const uint32_t e = (uint32_t) __builtin_ctz(own);
                own &= own - 1u;
                const float3 P = e == 0 ? A : e == 1 ? add(A, e1) : add(A, e2);
                const float3 d = e == 0 ? e1 : e == 1 ? d1 : neg(e2);
                if (o2r_begin(L, P, d, delta, e, w)) {
                    begin_walk();
                    return true;
                }
            }
        }
        return false;
    };
    auto start = [&](uint2 c) -> bool {
        key = c.x;
        tri = c.y;
        flag = a.flags + key / NL;
        if (load_flag(flag)) return false;
        if (COUNT) ctr.tris++;
        const LinkInfo& L = a.robot->link[key % NL];
        if (DIR == 1 ? !L.tri_has_root : !L.edge_has_root) return false;
        const float4 r0 = a.Ti[(size_t) key * 3], r1 = a.Ti[(size_t) key * 3 + 1], r2 = a.Ti[(size_t) key * 3 + 2];
        const Tri t = a.sc.tri_prims[tri];
        const LinkTri m = link_tri(t, r0, r1, r2);
        if constexpr (DIR == 1) {
            A = m.A;
            e1 = m.e1;
            e2 = m.e2;
            delta = m.delta;
            d1 = rot(r0, r1, r2, sub(f3(t.c), f3(t.b)));
            own = __float_as_uint(t.b.w) & 7u;
            return next_edge(L);
        } else {
            if (!r2o_begin(L, m.A, m.e1, m.e2, m.delta, w)) return false;
            begin_walk();
            return true;
        }
    };
    // a subtree shared by another lane of the wave: the same walk set-up (a pure function of the
    // candidate and the edge), an empty stack, no further owned edges
    auto start_shared = [&](uint4 e) -> bool {
        key = e.x;
        tri = e.y;
        flag = a.flags + key / NL;
        const LinkInfo& L = a.robot->link[key % NL];
        const float4 r0 = a.Ti[(size_t) key * 3], r1 = a.Ti[(size_t) key * 3 + 1], r2 = a.Ti[(size_t) key * 3 + 2];
        const Tri t = a.sc.tri_prims[tri];
        const LinkTri m = link_tri(t, r0, r1, r2);
        bool ok;
        if constexpr (DIR == 1) {
            own = 0u;
            const float3 P = e.w == 0 ? m.A : e.w == 1 ? add(m.A, m.e1) : add(m.A, m.e2);
            const float3 d = e.w == 0 ? m.e1 : e.w == 1 ? rot(r0, r1, r2, sub(f3(t.c), f3(t.b))) : neg(m.e2);
            ok = o2r_begin(L, P, d, m.delta, e.w, w);
        } else {
            ok = r2o_begin(L, m.A, m.e1, m.e2, m.delta, w);
        }
        begin_walk();
        ni = (int) e.z;
        return ok;
    };
    // the lane's walk has ended without a hit: obstacle->robot goes on with the next owned edge
    auto walk_ended = [&]() -> bool {
        if constexpr (DIR == 1) return next_edge(a.robot->link[key % NL]);
        return false;
    };
    for (;;) {
        const uint64_t idle = __ballot(!act);
        const int n_idle = __popcll(idle);
        if (n_idle >= kRefill) {
            // top the queue up with the live candidates of further leaf records
            while (qn < n_idle && base < total) {
                const uint32_t i = base + g;
                bool live = false;
                uint2 cand = make_uint2(0u, 0u);
                if (g < gpw && i < total) {
                    const uint2 c = list[i];
                    if (sub_k <= leaf_last(c.y) && !load_flag(a.flags + c.x / NL)) {
                        live = true;
                        cand = make_uint2(c.x, leaf_start(c.y) + sub_k);
                    }
                }
                const uint64_t m = __ballot(live);
                if (live) q[qn + __popcll(m & below)] = cand;
                qn += __popcll(m);
                base += n_waves * gpw;
            }
            __builtin_amdgcn_wave_barrier();
            // idle lane number r (in lane order) takes queue entry qn - 1 - r
            const int r = __popcll(idle & below);
            const int take = min(n_idle, qn);
            const bool mine = !act && r < take;
            uint2 c = make_uint2(0u, 0u);
            if (mine) c = q[qn - 1 - r];
            qn -= take;
            __builtin_amdgcn_wave_barrier();
            if (mine) act = start(c);
            if (a.narrow_share && qn == 0 && base >= total) {
                // no candidates left: the idle lanes take the top pending subtree of busy lanes
                const uint64_t idle2 = __ballot(!act), donors = __ballot(act && sp > 0);
                const int n_give = min(__popcll(idle2), __popcll(donors));
                if (n_give > 0) {
                    const int dr = __popcll(donors & below);
                    if (act && sp > 0 && dr < n_give) {
                        --sp;
                        uint32_t j = 3u;
                        if constexpr (DIR == 1) j = w.e;
                        sh[dr] = make_uint4(key, tri, (uint32_t) st[sp * BS], j);
                    }
                    __builtin_amdgcn_wave_barrier();
                    const int ir = __popcll(idle2 & below);
                    uint4 e = make_uint4(0u, 0u, 0u, 0u);
                    const bool taker = !act && ir < n_give;
                    if (taker) e = sh[ir];
                    __builtin_amdgcn_wave_barrier();
                    if (taker) act = start_shared(e);
                }
            }
            if (qn == 0 && base >= total && __ballot(act) == 0) break;
        }
        for (int s = 0; s < kSteps && act; ++s) {
            // The pose flag is read before every node visit: in cluttered scenes most poses are
            // flagged while other lanes still walk them, and stopping those walks a few visits
            // sooner outweighs the extra (L2-resident) load: 0.93 -> 0.84 ms obstacle->robot on
            // the dense scene, against checking every 16th visit.
            ++iter;
            if (load_flag(flag)) {
                act = false;
                break;
            }
            if (SPILL && spill_ok && (uint32_t) iter > hs.budget) {
                uint32_t j = 3u;
                if constexpr (DIR == 1) j = w.e;
                if (spill_walk<BS>(hs, key, tri, j, ni, st, sp)) {
                    act = walk_ended();
                    continue;
                }
                spill_ok = false;
            }
            int v;
            if constexpr (DIR == 1) v = o2r_visit<STACK, BS, COUNT>(a, w, key, tri, ni, st, sp, ctr);
            else v = r2o_visit<STACK, BS, COUNT>(a, w, key, tri, ni, st, sp, ctr);
            if (v == kVisitHit) {
                atomicOr(flag, 1u);
                act = false;
            } else if (v == kVisitDone) {
                act = walk_ended();
            }
        }
    }
}

// One direction (DIR 1: obstacle->robot, 2: robot->obstacle), pass `pass` of a.narrow_passes.  Pass
// 0 runs over the candidates with persistent lanes (robot->obstacle over the records of the poses
// obstacle->robot left unflagged, when that ran first); later passes resume the walks the previous
// pass handed off.
template <int DIR, int STACK, int BS, bool COUNT>
__global__ __launch_bounds__(BS) void k_narrow(DiscreteArgs a, int pass) {
    __shared__ int stk[STACK * BS];
    int* st = stk + threadIdx.x;
    const NarrowSpill hs{(pass & 1 ? a.nspill[1] : a.nspill[0]), a.nspill_count + pass, a.nspill_cap,
                         pass == 0 ? a.narrow_budget : a.narrow_budget_next, pass < a.narrow_passes - 1, a.overflow};
    const uint32_t stride = gridDim.x * BS;
    Ctr ctr;
    if (pass == 0) {
        __shared__ uint2 lq[BS / 64][128];
        __shared__ uint4 sh[BS / 64][64];
        // the gathered list: the records of the latest gather (a split check gathers in parts)
        const uint32_t first = DIR == 1 || !a.r2o_list ? a.seg_pre[0] : 0u;
        const uint2* list = DIR == 2 && a.r2o_list ? a.r2o_list : a.flat + first;
        const uint32_t total = DIR == 2 && a.r2o_list ? *a.r2o_count : a.seg_pre[kSeg] - first;
        // a pass that may hand off and one that never does (the walks of the latter keep fewer registers)
        if (hs.may)
            narrow_lanes<DIR, STACK, BS, COUNT, true>(a, list, total, lq[threadIdx.x >> 6], sh[threadIdx.x >> 6], st, ctr, hs);
        else
            narrow_lanes<DIR, STACK, BS, COUNT, false>(a, list, total, lq[threadIdx.x >> 6], sh[threadIdx.x >> 6], st, ctr, hs);
    } else {
        const uint4* in = ((pass - 1) & 1 ? a.nspill[1] : a.nspill[0]);
        const uint32_t total = min(a.nspill_count[pass - 1], a.nspill_cap);
        for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < total; i += stride) {
            const uint4 e = in[i];
            if (e.x == 0xffffffffu) continue;  // slot of a hand-off that did not fit
            uint32_t* flag = a.flags + e.x / (uint32_t) a.n_links;
            if (load_flag(flag)) continue;
            if (resume_candidate<DIR, STACK, BS, COUNT>(a, e.x, e.y, e.w, (int) e.z, st, flag, ctr, hs))
                atomicOr(flag, 1u);
        }
    }
    if (COUNT) {
        atomicAdd(&a.stats->node_fetches, (unsigned long long) ctr.nodes);
        atomicAdd(&a.stats->prim_tests, (unsigned long long) ctr.prims);
        atomicAdd(&a.stats->fp64, (unsigned long long) ctr.fp64);
        atomicAdd(&a.stats->edge_tests, (unsigned long long) ctr.edges);
        atomicAdd(&a.stats->narrow_tris, (unsigned long long) ctr.tris);
        atomicAdd(&a.stats->dir_nodes[DIR - 1], (unsigned long long) ctr.nodes);
        atomicAdd(&a.stats->dir_prims[DIR - 1], (unsigned long long) ctr.prims);
        atomicAdd(&a.stats->dir_tris[DIR - 1], (unsigned long long) ctr.tris);
    }
}

// FP64 re-decision of the queued tests, exactly as the oracle computes them.
__device__ __forceinline__ void redecide_range(const DiscreteArgs& a, uint32_t first, uint32_t stride) {
    const uint32_t total = min(*a.unsure_count, a.unsure_cap);
    for (uint32_t i = first; i < total; i += stride) {
        const uint4 u = a.unsure[i];
        const uint32_t p = u.x / (uint32_t) a.n_links;
        uint32_t* flag = a.flags + p;
        if (load_flag(flag)) continue;
        const double* T64 = a.T64 + (size_t) u.x * 12;
        const Tri t = a.sc.tri_prims[u.y];
        const float4 w[3] = {t.a, t.b, t.c};
        double P[3], Q[3], A[3], B[3], C[3];
        if (u.w < 3u) {
            const float4 wa = w[u.w], wb = w[(u.w + 1) % 3];
            double pa[3] = {wa.x, wa.y, wa.z}, pb[3] = {wb.x, wb.y, wb.z};
            x64::inv_xform(T64, pa, P);
            x64::inv_xform(T64, pb, Q);
            const Tri lt = a.rb.tri_prims[u.z];
            A[0] = lt.a.x; A[1] = lt.a.y; A[2] = lt.a.z;
            B[0] = lt.b.x; B[1] = lt.b.y; B[2] = lt.b.z;
            C[0] = lt.c.x; C[1] = lt.c.y; C[2] = lt.c.z;
        } else {
            const Edge e = a.rb.edge_prims[u.z];
            double p0[3] = {e.p.x, e.p.y, e.p.z}, q0[3] = {e.q.x, e.q.y, e.q.z};
            x64::xform(T64, p0, P);
            x64::xform(T64, q0, Q);
            A[0] = t.a.x; A[1] = t.a.y; A[2] = t.a.z;
            B[0] = t.b.x; B[1] = t.b.y; B[2] = t.b.z;
            C[0] = t.c.x; C[1] = t.c.y; C[2] = t.c.z;
        }
        if (x64::seg_tri(P, Q, A, B, C)) atomicOr(flag, 1u);
    }
}
__global__ __launch_bounds__(kBS) void k_redecide(DiscreteArgs a) {
    redecide_range(a, blockIdx.x * kBS + threadIdx.x, gridDim.x * kBS);
}

// Status words and small flag arrays into pinned host memory (see launch_redecide).
__global__ __launch_bounds__(kBS) void k_publish(const uint32_t* __restrict__ scalars, uint32_t* status_out,
                                                const uint32_t* __restrict__ flags, uint32_t* flags_out, uint32_t n) {
    if (flags_out)
        for (uint32_t i = blockIdx.x * kBS + threadIdx.x; i < n; i += gridDim.x * kBS) flags_out[i] = flags[i];
    if (blockIdx.x == 0 && threadIdx.x < kPublish) status_out[threadIdx.x] = scalars[threadIdx.x];
}

// Small batches: the whole re-decision queue in one block, which then publishes (its own flag
// updates are visible to it after the barrier; everything else was written by earlier kernels).
__global__ __launch_bounds__(kBS) void k_redecide_publish(DiscreteArgs a, const uint32_t* __restrict__ scalars,
                                                         uint32_t* status_out, uint32_t* flags_out) {
    redecide_range(a, threadIdx.x, kBS);
    __syncthreads();
    if (flags_out)
        for (uint32_t i = threadIdx.x; i < (uint32_t) a.n_poses; i += kBS) flags_out[i] = load_flag(a.flags + i);
    if (threadIdx.x < kPublish) status_out[threadIdx.x] = scalars[threadIdx.x];
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void) hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    }
    return cus;
}

}  // namespace

void launch_fk(const RobotDev* robot, int, int, const float* poses, int n, double* T64, float4* Ti, float4* wbox,
               const FkClear& z, hipStream_t s) {
    if (n <= 0) return;
    // a wave per pose while the batch leaves most of the device idle (one thread per pose would run
    // the whole FP64 chain serially), a thread per pose beyond
    if ((size_t) n * 64 <= (size_t) device_cus() * kBS * 4)
        hipLaunchKernelGGL(k_fk_wave, dim3(blocks_for((size_t) n * 64, kBS)), dim3(kBS), 0, s, robot, poses, n, T64, Ti,
                           wbox, z);
    else
        hipLaunchKernelGGL(k_fk, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, robot, poses, n, T64, Ti, wbox, z);
    RTCD_LAUNCH_CHECK();
}

template <int STACK>
static void broad_t(const DiscreteArgs& a, bool count, hipStream_t s, int p0, int p1) {
    const unsigned nb0 = (unsigned) std::min<size_t>(blocks_for((size_t) a.n_active * a.n_poses * a.F, kBS), kSeg);
    if (!nb0) return;
    // later passes stride over the hand-off queue; they need at most kSeg blocks (one per segment)
    const unsigned nb1 = std::min<unsigned>((unsigned) device_cus() * 8u, kSeg);
    for (int pass = p0; pass < (p1 < 0 ? a.n_passes : p1); ++pass) {
        const unsigned nb = pass == 0 ? nb0 : nb1;
        if (a.broad_share && pass > 0 && pass == a.n_passes - 1) {
            // persistent: the blocks the CUs hold at once (LDS: 4 per CU)
            const unsigned nbs = std::min<unsigned>((unsigned) device_cus() * 4u, kSeg);
            if (count) hipLaunchKernelGGL((k_broad_share<STACK, true>), dim3(nbs), dim3(kBS), 0, s, a, pass);
            else hipLaunchKernelGGL((k_broad_share<STACK, false>), dim3(nbs), dim3(kBS), 0, s, a, pass);
        } else if (count) {
            hipLaunchKernelGGL((k_broad<STACK, true>), dim3(nb), dim3(kBS), 0, s, a, pass);
        } else {
            hipLaunchKernelGGL((k_broad<STACK, false>), dim3(nb), dim3(kBS), 0, s, a, pass);
        }
        RTCD_LAUNCH_CHECK();
    }
}

void launch_broad(const DiscreteArgs& a, int stack, bool count, hipStream_t s, int p0, int p1) {
    if (stack <= 20) broad_t<20>(a, count, s, p0, p1);
    else if (stack <= 26) broad_t<26>(a, count, s, p0, p1);
    else if (stack <= 32) broad_t<32>(a, count, s, p0, p1);
    else if (stack <= 40) broad_t<40>(a, count, s, p0, p1);
    else if (stack <= 64) broad_t<64>(a, count, s, p0, p1);
    else throw std::runtime_error("BVH deeper than 64 levels");
}

template <int DIR, int STACK, int BS>
static void narrow_bs(const DiscreteArgs& a, bool count, hipStream_t s) {
    // as many threads as 8 blocks of 256 per CU, whatever the block size
    const unsigned nb = (unsigned) device_cus() * 8u * (256u / BS);
    // resume passes carry only the handed-off walks: a quarter of the grid (dense 65,536-pose two-way
    // check: 1.93 -> 1.91 ms; a sixteenth: 1.94)
    for (int pass = 0; pass < a.narrow_passes; ++pass) {
        const unsigned g = pass ? std::max(1u, nb / 4u) : nb;
        if (count) hipLaunchKernelGGL((k_narrow<DIR, STACK, BS, true>), dim3(g), dim3(BS), 0, s, a, pass);
        else hipLaunchKernelGGL((k_narrow<DIR, STACK, BS, false>), dim3(g), dim3(BS), 0, s, a, pass);
        RTCD_LAUNCH_CHECK();
    }
}

template <int DIR, int STACK>
static void narrow_t(const DiscreteArgs& a, bool count, hipStream_t s) {
    if (a.narrow_bs == 64) narrow_bs<DIR, STACK, 64>(a, count, s);
    else if (a.narrow_bs == 128) narrow_bs<DIR, STACK, 128>(a, count, s);
    else narrow_bs<DIR, STACK, 256>(a, count, s);
}

template <int DIR>
static void narrow_dir(const DiscreteArgs& a, int stack, bool count, hipStream_t s) {
    if (stack <= 16) narrow_t<DIR, 16>(a, count, s);
    else if (stack <= 24) narrow_t<DIR, 24>(a, count, s);
    else if (stack <= 32) narrow_t<DIR, 32>(a, count, s);
    else if (stack <= 64) narrow_t<DIR, 64>(a, count, s);
    else throw std::runtime_error("BVH deeper than 64 levels");
}

void launch_gather(const DiscreteArgs& a, bool small, hipStream_t s, int part) {
    if (small) {
        hipLaunchKernelGGL(k_seg_scan_gather, dim3(16), dim3(kBS), 0, s, a.cand, a.seg_count, a.seg_cap, a.seg_pre,
                           a.flat);
        RTCD_LAUNCH_CHECK();
        return;
    }
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(kBS), 0, s, a.seg_count, a.seg_cap, a.seg_pre, a.seg_base, part);
    RTCD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_seg_gather, dim3(kSeg), dim3(kBS), 0, s, a.cand, a.seg_count, a.seg_cap, a.seg_pre, a.flat);
    RTCD_LAUNCH_CHECK();
}

void launch_narrow(const DiscreteArgs& a, int dir, int stack, bool count, hipStream_t s) {
    if (dir == 1) narrow_dir<1>(a, stack, count, s);
    else narrow_dir<2>(a, stack, count, s);
}

void launch_compact_unflagged(const DiscreteArgs& a, hipStream_t s, bool keep_flagged) {
    if (keep_flagged) hipLaunchKernelGGL(k_compact<true>, dim3(kSeg), dim3(kBS), 0, s, a);
    else hipLaunchKernelGGL(k_compact<false>, dim3(kSeg), dim3(kBS), 0, s, a);
    RTCD_LAUNCH_CHECK();
}

void launch_redecide(const DiscreteArgs& a, bool one_block, const uint32_t* scalars, uint32_t* status_out,
                     uint32_t* flags_out, hipStream_t s) {
    if (one_block) {
        hipLaunchKernelGGL(k_redecide_publish, dim3(1), dim3(kBS), 0, s, a, scalars, status_out, flags_out);
        RTCD_LAUNCH_CHECK();
        return;
    }
    hipLaunchKernelGGL(k_redecide, dim3((unsigned) device_cus() * 2u), dim3(kBS), 0, s, a);
    RTCD_LAUNCH_CHECK();
    const uint32_t n = (uint32_t) a.n_poses;
    const unsigned nb = flags_out ? std::max(1u, std::min(blocks_for(n, kBS), 64u)) : 1u;
    hipLaunchKernelGGL(k_publish, dim3(nb), dim3(kBS), 0, s, scalars, status_out, a.flags, flags_out, n);
    RTCD_LAUNCH_CHECK();
}

}  // namespace rtcd
