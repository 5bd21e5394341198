// CDNA4 kernels of the XPBD substep pipeline.
//
// Data layout (all FP32, structure of arrays, 16-byte rows for coalesced and
// single-transaction gathers):
//   bodies     x4 (position, inverse mass)  q4 (w,x,y,z)  v4, w4 (velocities)
//              px4, pq4 (previous pose)  f4 (force, mass)  t4 (torque)
//              I0/I1 (local inertia xx yy zz xy | xz yz)
//              Ii0/Ii1 (local inverse inertia xx yy zz xy | xz yz type -)
//   colliders  meta (kind, body, is_static, bp slot)  dims, lpos, lrot, mat
//   SAP        fat AABB per slot, endpoint keys/values (stable across frames),
//              endpoint positions, sorted sweep rows
//   pairs      frame pair list (canonical), colour, solve slots; contact rows
//              live in the pair's solve slot, four points per slot.
#pragma once
#include "rbgpu_collide.hpp"
#include "rbgpu_sort.hpp"

namespace rb {

constexpr uint32_t BP_INVALID = 0xffffffffu;
constexpr int MAX_COLOURS = 256;
constexpr int NBINS = 10;
enum Bin { BIN_SS = 0, BIN_SB, BIN_BB, BIN_SC, BIN_CC, BIN_CB, BIN_SP, BIN_BP, BIN_CP, BIN_OTHER };

struct Bodies {
    float4 *x, *q, *v, *w, *px, *pq, *f, *t, *I0, *I1, *Ii0, *Ii1;
};
struct Colliders {
    int4 *meta;      // kind, body, is_static, bp slot
    float4 *dims, *lpos, *lrot, *mat;   // mat: restitution, static friction, dynamic friction
};
struct Contacts {
    int *cnt;        // points per slot this substep
    uint2 *cc;       // colliders (a, b) per slot
    uint2 *bb;       // bodies (a, b) per slot
    float4 *nr;      // normal, restitution
    float2 *fr;      // static, dynamic friction
    float4 *ra, *rb; // body-local lever arms (rb.w = depth), 4 per slot
    float4 *pa, *pb; // world points (pa.w = pre-solve normal velocity, pb.w = collision flag)
    float2 *lam;     // normal, tangent lambda
};

__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ Q quat(float4 a) { return Q{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float4 f4q(Q q) { return make_float4(q.w, q.x, q.y, q.z); }
__device__ __forceinline__ SM sm_from(float4 a, float4 b) { return SM{a.x, a.y, a.z, a.w, b.x, b.y}; }
__device__ __forceinline__ Shape shape_of(int kind, float4 d) { return Shape{kind, d.x, d.y, d.z, d.w}; }
__device__ __forceinline__ TF collider_tf(const Colliders &C, const Bodies &B, uint32_t c, uint32_t body) {
    V3 bp = xyz(B.x[body]);
    Q bq = quat(B.q[body]);
    return TF{bp + rotate(bq, xyz(C.lpos[c])), qmul(bq, quat(C.lrot[c]))};   // Collider.hpp:73-81
}

// ------------------------------------------------------------------ broadphase
// Refit of one collider's swept AABB with fat-AABB hysteresis
// (CollisionPipeline.cpp:14-53, BroadPhase.hpp:228-249, BroadPhase.cpp:211-232).
__global__ void k_refit(Colliders C, Bodies B, uint32_t M, float dt, float margin, float reinsert, float4 *fmin,
                        float4 *fmax, uint32_t *ep_key, const uint2 *ep_pos, int *dirty) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int4 m = C.meta[i];
    if (m.z != 0 || (uint32_t)m.w == BP_INVALID) return;
    const uint32_t body = (uint32_t)m.y, slot = (uint32_t)m.w;
    Box3 t = shape_aabb(shape_of(m.x, C.dims[i]), collider_tf(C, B, i, body));
    V3 dv = xyz(B.v[body]) * dt;
    V3 emn{margin + (dv.x < 0.0f ? -dv.x : 0.0f), margin + (dv.y < 0.0f ? -dv.y : 0.0f), margin + (dv.z < 0.0f ? -dv.z : 0.0f)};
    V3 emx{margin + (dv.x > 0.0f ? dv.x : 0.0f), margin + (dv.y > 0.0f ? dv.y : 0.0f), margin + (dv.z > 0.0f ? dv.z : 0.0f)};
    V3 smn = t.mn - emn, smx = t.mx + emx;
    float4 fa = fmin[slot], fb = fmax[slot];
    if (smn.x >= fa.x && smx.x <= fb.x && smn.y >= fa.y && smx.y <= fb.y && smn.z >= fa.z && smx.z <= fb.z) return;
    float g = margin + reinsert;
    V3 nmn = smn - V3{g, g, g}, nmx = smx + V3{g, g, g};
    fmin[slot] = f4(nmn, 0.0f);
    fmax[slot] = f4(nmx, 0.0f);
    uint2 p = ep_pos[slot];
    ep_key[p.x] = float_key(nmn.x);
    ep_key[p.y] = float_key(nmx.x);
    *dirty = 1;
}

// Sorted sweep rows and endpoint positions after a sort.
// row (float4): min endpoint -> (x_max, y_min, y_max, z_min) and zmx separately.
__global__ void k_sweep_prep(const uint32_t *ep_val, uint32_t nep, const float4 *fmin, const float4 *fmax, uint2 *ep_pos,
                             uint32_t *sw_meta, float4 *sw_a, float2 *sw_b, const int *dirty) {
    if (*dirty == 0) return;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nep) return;
    uint32_t v = ep_val[j];
    uint32_t slot = v >> 1, is_max = v & 1u;
    if (is_max) ep_pos[slot].y = j; else ep_pos[slot].x = j;
    sw_meta[j] = v;
    float4 a = fmin[slot], b = fmax[slot];
    sw_a[j] = make_float4(a.x, b.x, a.y, b.y);
    sw_b[j] = make_float2(a.z, b.z);
}

// Pair filter of the narrowphase driver (CollisionPipeline.cpp:121-133).
__device__ __forceinline__ bool pair_ok(const Colliders &C, uint32_t ca, uint32_t cb) {
    int4 a = C.meta[ca], b = C.meta[cb];
    if (a.z != 0 && b.z != 0) return false;
    return a.y != b.y;
}

// Forward scan from each min endpoint to its own max endpoint: every pair is
// found once, from the object that opens first (same set as the active-set sweep
// of BroadPhase.cpp:104-152, inclusive 3-D fat-AABB test AABB.hpp:54-59).
template <bool EMIT>
__global__ void k_sweep(const uint32_t *sw_meta, const float4 *sw_a, const float2 *sw_b, const uint2 *ep_pos,
                        const uint32_t *uid, Colliders C, uint32_t nep, uint32_t *cnt, const uint32_t *off,
                        uint2 *pairs, uint32_t cap, int *overflow) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nep) return;
    uint32_t v = sw_meta[i];
    if (v & 1u) { if (!EMIT) cnt[i] = 0; return; }
    const uint32_t slot = v >> 1;
    const uint32_t end = ep_pos[slot].y;
    const float4 A = sw_a[i];
    const float2 Az = sw_b[i];
    const uint32_t ua = uid[slot];
    uint32_t n = 0, o = EMIT ? off[i] : 0;
    for (uint32_t j = i + 1; j < end; ++j) {
        uint32_t w = sw_meta[j];
        if (w & 1u) continue;
        float4 Bv = sw_a[j];
        float2 Bz = sw_b[j];
        bool ov = (A.x <= Bv.y && A.y >= Bv.x) && (A.z <= Bv.w && A.w >= Bv.z) && (Az.x <= Bz.y && Az.y >= Bz.x);
        if (!ov) continue;
        uint32_t ub = uid[w >> 1];
        uint32_t lo = ua < ub ? ua : ub, hi = ua < ub ? ub : ua;
        if (!pair_ok(C, lo, hi)) continue;
        if (EMIT) {
            uint32_t k = o + n;
            if (k < cap) pairs[k] = make_uint2(lo, hi);
            else *overflow = 1;
        }
        ++n;
    }
    if (!EMIT) cnt[i] = n;
}

__global__ void k_append(uint2 *pairs, const uint32_t *n_sap, const uint2 *extra, uint32_t n_extra, uint32_t cap,
                         uint32_t *n_total, int *overflow) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t base = *n_sap;
    if (t == 0) {
        *n_total = base + n_extra;
        if (base + n_extra > cap) *overflow = 1;
    }
    if (t < n_extra && base + t < cap) pairs[base + t] = extra[t];
}

// ------------------------------------------------------------------ colouring
__device__ __forceinline__ uint32_t mix32(uint32_t x) {   // lowbias32 (C. Wellons, public domain)
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Pair -> bodies, dynamic-body degree counts, colour reset.
__global__ void k_pair_bodies(const uint2 *pairs, const uint32_t *n_pairs, Colliders C, Bodies B, uint2 *pbod,
                              uint32_t *deg, int *colour, uint32_t *uncoloured) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t P = *n_pairs;
    if (p == 0) *uncoloured = P;
    if (p >= P) return;
    uint2 pr = pairs[p];
    uint32_t ba = (uint32_t)C.meta[pr.x].y, bb = (uint32_t)C.meta[pr.y].y;
    pbod[p] = make_uint2(ba, bb);
    colour[p] = -1;
    if (B.Ii1[ba].z != 0.0f) atomicAdd(&deg[ba], 1u);
    if (B.Ii1[bb].z != 0.0f) atomicAdd(&deg[bb], 1u);
}

__global__ void k_fill_adj(const uint2 *pbod, const uint32_t *n_pairs, Bodies B, const uint32_t *aoff, uint32_t *cursor,
                           uint32_t *adj) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *n_pairs) return;
    uint2 b = pbod[p];
    if (B.Ii1[b.x].z != 0.0f) adj[aoff[b.x] + atomicAdd(&cursor[b.x], 1u)] = p;
    if (B.Ii1[b.y].z != 0.0f) adj[aoff[b.y] + atomicAdd(&cursor[b.y], 1u)] = p;
}

// One Jones-Plassmann round: a pair whose uncoloured neighbours all have lower
// priority takes the smallest colour unused by its coloured neighbours (which are
// then exactly its higher-priority neighbours, so the result equals sequential
// first-fit in descending priority whatever the timing).
__global__ void k_jp_round(const uint2 *pbod, const uint32_t *n_pairs, Bodies B, const uint32_t *aoff,
                           const uint32_t *adeg, const uint32_t *adj, int *colour, uint32_t *uncoloured, int *err) {
    if (__atomic_load_n(uncoloured, __ATOMIC_RELAXED) == 0) return;
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *n_pairs) return;
    if (__atomic_load_n(&colour[p], __ATOMIC_RELAXED) >= 0) return;
    const uint32_t hp = mix32(p);
    uint64_t used[MAX_COLOURS / 64] = {0, 0, 0, 0};
    uint2 bb = pbod[p];
    uint32_t bodies[2] = {bb.x, bb.y};
    for (int s = 0; s < 2; ++s) {
        uint32_t b = bodies[s];
        if (B.Ii1[b].z == 0.0f) continue;
        uint32_t o = aoff[b], n = adeg[b];
        for (uint32_t k = 0; k < n; ++k) {
            uint32_t q = adj[o + k];
            if (q == p) continue;
            int c = __atomic_load_n(&colour[q], __ATOMIC_RELAXED);
            if (c < 0) {
                uint32_t hq = mix32(q);
                if (hq > hp || (hq == hp && q > p)) return;   // a higher-priority neighbour is uncoloured
            } else {
                used[c >> 6] |= 1ull << (c & 63);
            }
        }
    }
    int c = -1;
    for (int w = 0; w < MAX_COLOURS / 64; ++w)
        if (~used[w]) { c = w * 64 + __builtin_ctzll(~used[w]); break; }
    if (c < 0) { *err = 1; c = MAX_COLOURS - 1; }
    __atomic_store_n(&colour[p], c, __ATOMIC_RELAXED);
    atomicSub(uncoloured, 1u);
}

__device__ __forceinline__ int bin_of(int ka, int kb) {
    if (kb == SH_PLANE) return ka == SH_SPHERE ? BIN_SP : ka == SH_BOX ? BIN_BP : ka == SH_CAPSULE ? BIN_CP : BIN_OTHER;
    if (ka > kb) { int t = ka; ka = kb; kb = t; }
    if (ka == SH_SPHERE && kb == SH_SPHERE) return BIN_SS;
    if (ka == SH_SPHERE && kb == SH_BOX) return BIN_SB;
    if (ka == SH_BOX && kb == SH_BOX) return BIN_BB;
    if (ka == SH_SPHERE && kb == SH_CAPSULE) return BIN_SC;
    if (ka == SH_CAPSULE && kb == SH_CAPSULE) return BIN_CC;
    if (ka == SH_BOX && kb == SH_CAPSULE) return BIN_CB;
    return BIN_OTHER;
}

// Per-frame pair data indexed by pair id: colliders, bodies, mixed materials, bin key.
__global__ void k_pair_gather(const uint2 *pairs, const uint32_t *n_pairs, Colliders C, Contacts K, uint32_t *bin_key,
                              uint32_t *bin_val) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *n_pairs) return;
    uint2 cc = pairs[p];
    int4 ma_ = C.meta[cc.x], mb_ = C.meta[cc.y];
    K.cc[p] = cc;
    K.bb[p] = make_uint2((uint32_t)ma_.y, (uint32_t)mb_.y);
    K.cnt[p] = 0;
    float4 ma = C.mat[cc.x], mb = C.mat[cc.y];
    float rest = fmin_ref(ma.x, mb.x);                                     // CollisionPipeline.cpp:75-77
    K.nr[p] = make_float4(0.0f, 0.0f, 0.0f, rest);
    K.fr[p] = make_float2((ma.y + mb.y) * 0.5f, (ma.z + mb.z) * 0.5f);
    bin_key[p] = (uint32_t)bin_of(ma_.x, mb_.x);
    bin_val[p] = p;
}

// ------------------------------------------------------------------ grid barrier
// Barrier for cooperative (co-resident) grids.  The last block to arrive evaluates
// `*probe == 0` (when probe is given) after every block has finished the phase, and
// publishes it with the new generation, so every block takes the same decision.
// Every wait is bounded: after BARRIER_TIMEOUT_TICKS of the 100 MHz wall clock a
// waiting block raises *fault and leaves; once *fault is set no block waits again,
// so the kernel always drains (the host then reports the fault as an error).
constexpr uint64_t BARRIER_TIMEOUT_TICKS = 200000000ull;   // 2 s
struct GridBar {
    uint32_t *count, *gen;
    int *decision;
    int *fault;
};
__device__ __forceinline__ int grid_barrier(GridBar b, const uint32_t *probe) {
    __shared__ int s_dec;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t g = __hip_atomic_load(b.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(b.count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            int d = probe ? (__hip_atomic_load(probe, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0u) : 0;
            __hip_atomic_store(b.decision, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(b.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(b.gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(b.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(b.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (wall_clock64() - t0 > BARRIER_TIMEOUT_TICKS) {
                    __hip_atomic_store(b.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        s_dec = __hip_atomic_load(b.decision, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_load(b.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) s_dec = 1;   // leave any loop
        __threadfence();
    }
    __syncthreads();
    return s_dec;
}

// ------------------------------------------------------------------ substep colouring of active pairs
__global__ void k_active_flags(const int *cnt, uint32_t P, uint32_t *flag) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) flag[p] = cnt[p] > 0 ? 1u : 0u;
}
// Compacted active list (ascending pair id); colour reset; counters reset.
__global__ void k_compact(const uint32_t *flag, const uint32_t *off, uint32_t P, const uint32_t *n_act, uint32_t *act,
                          int *colour, uint32_t *cround, uint32_t *uncoloured, uint32_t *ncol, uint32_t *deg, uint32_t *cur,
                          uint32_t N) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) { *uncoloured = *n_act; *ncol = 0; }
    if (p <= N) { deg[p] = 0; cur[p] = 0; }
    if (p < P && flag[p]) { act[off[p]] = p; colour[off[p]] = -1; cround[off[p]] = 0xffffffffu; }
}
__global__ void k_act_deg(const uint32_t *act, const uint32_t *n_act, const uint2 *pbod, Bodies B, uint32_t *deg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_act) return;
    uint2 b = pbod[act[i]];
    if (B.Ii1[b.x].z != 0.0f) atomicAdd(&deg[b.x], 1u);
    if (B.Ii1[b.y].z != 0.0f) atomicAdd(&deg[b.y], 1u);
}
__global__ void k_act_adj(const uint32_t *act, const uint32_t *n_act, const uint2 *pbod, Bodies B, const uint32_t *aoff,
                          uint32_t *cur, uint32_t *adj) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_act) return;
    uint2 b = pbod[act[i]];
    if (B.Ii1[b.x].z != 0.0f) adj[aoff[b.x] + atomicAdd(&cur[b.x], 1u)] = i;
    if (B.Ii1[b.y].z != 0.0f) adj[aoff[b.y] + atomicAdd(&cur[b.y], 1u)] = i;
}

__device__ __forceinline__ bool higher_priority(uint32_t q, uint32_t p) {
    const uint32_t hq = mix32(q), hp = mix32(p);
    return hq > hp || (hq == hp && q > p);
}
// The idx-th colour not set in `used`.
__device__ __forceinline__ int nth_free(const uint64_t used[MAX_COLOURS / 64], uint32_t idx) {
    for (int w = 0; w < MAX_COLOURS / 64; ++w) {
        uint64_t fr = ~used[w];
        const uint32_t n = (uint32_t)__popcll(fr);
        if (idx < n) {
            for (; idx; --idx) fr &= fr - 1;
            return w * 64 + __builtin_ctzll(fr);
        }
        idx -= n;
    }
    return -1;
}

// Colouring of the active pairs by synchronous randomized trials, all rounds in one
// cooperative launch.  Round r: (A) every uncoloured vertex i (pair p = act[i]) picks
// t = the (mix32(p ^ r*0x9e3779b9) mod k)-th colour unused by neighbours coloured in
// earlier rounds, k = uncoloured neighbour entries + 1; (B) it keeps t unless an
// uncoloured neighbour of higher priority (mix32(q), q) picked t too.  cround[i] holds
// the round a vertex was coloured in (0xffffffff = not yet), so "uncoloured at the start
// of round r" is cround >= r in both phases.  Every round colours at least the highest-
// priority uncoloured vertex; the expected number of rounds is O(log degree).
__global__ void __launch_bounds__(256) k_colour_coop(const uint32_t *act, const uint32_t *n_act, const uint2 *pbod, Bodies B,
                                                     const uint32_t *aoff, const uint32_t *adeg, const uint32_t *adj,
                                                     int *colour, int *tent, uint32_t *cround, uint32_t *uncoloured,
                                                     uint32_t *ncol, int *err, GridBar bar, uint32_t max_rounds,
                                                     uint32_t *rounds_out) {
    const uint32_t n = *n_act;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = 0;; ++r) {
        // phase A: tentative colours
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            if (cround[i] < r) continue;
            const uint32_t p = act[i];
            uint64_t used[MAX_COLOURS / 64] = {0, 0, 0, 0};
            uint32_t k = 1;
            const uint2 bb = pbod[p];
            for (int sd = 0; sd < 2; ++sd) {
                const uint32_t b = sd ? bb.y : bb.x;
                if (B.Ii1[b].z == 0.0f) continue;
                const uint32_t o = aoff[b], nn = adeg[b];
                for (uint32_t e = 0; e < nn; ++e) {
                    const uint32_t j = adj[o + e];
                    if (j == i) continue;
                    if (cround[j] < r) {
                        const int c = colour[j];
                        used[c >> 6] |= 1ull << (c & 63);
                    } else {
                        ++k;
                    }
                }
            }
            int t = nth_free(used, mix32(p ^ (r * 0x9e3779b9u)) % k);
            if (t < 0) { *err = 1; t = MAX_COLOURS - 1; }
            tent[i] = t;
        }
        grid_barrier(bar, nullptr);
        // phase B: keep the tentative colour unless a higher-priority neighbour chose it too
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            if (cround[i] < r) continue;
            const uint32_t p = act[i];
            const int t = tent[i];
            const uint2 bb = pbod[p];
            bool ok = true;
            for (int sd = 0; sd < 2 && ok; ++sd) {
                const uint32_t b = sd ? bb.y : bb.x;
                if (B.Ii1[b].z == 0.0f) continue;
                const uint32_t o = aoff[b], nn = adeg[b];
                for (uint32_t e = 0; e < nn; ++e) {
                    const uint32_t j = adj[o + e];
                    if (j == i) continue;
                    if (__hip_atomic_load(&cround[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= r && tent[j] == t &&
                        higher_priority(act[j], p)) { ok = false; break; }
                }
            }
            if (ok) {
                colour[i] = t;
                __hip_atomic_store(&cround[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicSub(uncoloured, 1u);
                atomicMax(ncol, (uint32_t)(t + 1));
            }
        }
        const int done = grid_barrier(bar, uncoloured);
        if (done) {
            if (blockIdx.x == 0 && threadIdx.x == 0) *rounds_out = r + 1;
            break;
        }
        if (r + 1 >= max_rounds) {
            if (blockIdx.x == 0 && threadIdx.x == 0) { *err = 2; *rounds_out = r + 1; }
            break;
        }
    }
}

__global__ void k_colour_keys(const int *colour, const uint32_t *act, const uint32_t *n_act, uint32_t *key, uint32_t *val) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_act) return;
    key[i] = (uint32_t)colour[i];
    val[i] = act[i];
}

// ------------------------------------------------------------------ narrowphase
// Contact emission (CollisionPipeline.cpp:59-104): split the depth across the two
// surfaces and store body-local lever arms.
__device__ __forceinline__ void emit(const Manifold &m, const Bodies &B, Contacts &K, uint32_t s, uint2 bb) {
    V3 xa = xyz(B.x[bb.x]), xb = xyz(B.x[bb.y]);
    Q qa = conj(quat(B.q[bb.x])), qb = conj(quat(B.q[bb.y]));
    float4 nr = K.nr[s];
    K.nr[s] = f4(m.n, nr.w);
    for (int i = 0; i < m.np; ++i) {
        float half = m.dep[i] * 0.5f;
        V3 pa = m.p[i] + m.n * half;
        V3 pb = m.p[i] - m.n * half;
        uint32_t k = s * 4 + i;
        K.ra[k] = f4(rotate(qa, pa - xa), 0.0f);
        K.rb[k] = f4(rotate(qb, pb - xb), m.dep[i]);
        K.pa[k] = f4(pa, 0.0f);
        K.pb[k] = f4(pb, 0.0f);
        K.lam[k] = make_float2(0.0f, 0.0f);
    }
    K.cnt[s] = m.np;
}

template <int BIN>
__device__ __forceinline__ bool collide_bin(const Shape &a, const TF &ta, const Shape &b, const TF &tb, Manifold &m) {
    // Analytic table with (B,A) = (A,B) and the normal flipped (Dispatcher.hpp:104-116).
    const bool sw = (BIN != BIN_SP && BIN != BIN_BP && BIN != BIN_CP) && (a.kind > b.kind);
    const Shape &A = sw ? b : a;
    const Shape &Bs = sw ? a : b;
    const TF &TA = sw ? tb : ta;
    const TF &TB = sw ? ta : tb;
    bool hit = false;
    if (BIN == BIN_SS) hit = sphere_sphere(A, TA, Bs, TB, m);
    else if (BIN == BIN_SB) hit = sphere_box(A, TA, Bs, TB, m);
    else if (BIN == BIN_BB) hit = box_box(A, TA, Bs, TB, m);
    else if (BIN == BIN_SC) hit = sphere_capsule(A, TA, Bs, TB, m);
    else if (BIN == BIN_CC) hit = capsule_capsule(A, TA, Bs, TB, m);
    else if (BIN == BIN_CB) {
        // capsule (kind 3) vs box (kind 1): the routine takes the capsule first
        const bool cap_first = a.kind == SH_CAPSULE;
        hit = cap_first ? capsule_box(a, ta, b, tb, m) : capsule_box(b, tb, a, ta, m);
        if (hit && !cap_first) m.n = -m.n;
        return hit;
    } else if (BIN == BIN_SP) hit = sphere_plane(a, ta, b, tb, m);
    else if (BIN == BIN_BP) hit = box_plane(a, ta, b, tb, m);
    else if (BIN == BIN_CP) hit = capsule_plane(a, ta, b, tb, m);
    if (hit && sw) m.n = -m.n;
    return hit;
}

template <int BIN>
__global__ void __launch_bounds__(128) k_narrow(const uint32_t *bin_slot, uint32_t off, uint32_t n, Colliders C, Bodies B,
                                                Contacts K) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t s = bin_slot[off + t];
    uint2 cc = K.cc[s];
    uint2 bb = K.bb[s];
    int4 ma = C.meta[cc.x], mb = C.meta[cc.y];
    Shape a = shape_of(ma.x, C.dims[cc.x]), b = shape_of(mb.x, C.dims[cc.y]);
    TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
    Manifold m;
    m.np = 0;
    if (collide_bin<BIN>(a, ta, b, tb, m) && m.np > 0) emit(m, B, K, s, bb);
    else K.cnt[s] = 0;
}

// Single-pair narrowphase over a batch (the analytic dispatch of Dispatcher.hpp:212-217
// without a world): out row = n, normal xyz, then (point xyz, depth) x 4.
__global__ void k_collide_pairs(uint32_t n, const int *kind_a, const float4 *dims_a, const float4 *pos_a,
                                const float4 *rot_a, const int *kind_b, const float4 *dims_b, const float4 *pos_b,
                                const float4 *rot_b, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Shape a = shape_of(kind_a[i], dims_a[i]), b = shape_of(kind_b[i], dims_b[i]);
    TF ta{xyz(pos_a[i]), quat(rot_a[i])}, tb{xyz(pos_b[i]), quat(rot_b[i])};
    Manifold m;
    m.np = 0;
    bool hit = false;
    switch (bin_of(a.kind, b.kind)) {
        case BIN_SS: hit = collide_bin<BIN_SS>(a, ta, b, tb, m); break;
        case BIN_SB: hit = collide_bin<BIN_SB>(a, ta, b, tb, m); break;
        case BIN_BB: hit = collide_bin<BIN_BB>(a, ta, b, tb, m); break;
        case BIN_SC: hit = collide_bin<BIN_SC>(a, ta, b, tb, m); break;
        case BIN_CC: hit = collide_bin<BIN_CC>(a, ta, b, tb, m); break;
        case BIN_CB: hit = collide_bin<BIN_CB>(a, ta, b, tb, m); break;
        case BIN_SP: hit = collide_bin<BIN_SP>(a, ta, b, tb, m); break;
        case BIN_BP: hit = collide_bin<BIN_BP>(a, ta, b, tb, m); break;
        case BIN_CP: hit = collide_bin<BIN_CP>(a, ta, b, tb, m); break;
        default: break;
    }
    float *o = out + (size_t)i * 20;
    for (int k = 0; k < 20; ++k) o[k] = 0.0f;
    if (!hit) return;
    o[0] = (float)m.np;
    o[1] = m.n.x; o[2] = m.n.y; o[3] = m.n.z;
    for (int k = 0; k < m.np; ++k) {
        o[4 + 4 * k] = m.p[k].x; o[5 + 4 * k] = m.p[k].y; o[6 + 4 * k] = m.p[k].z; o[7 + 4 * k] = m.dep[k];
    }
}

// ------------------------------------------------------------------ bodies (Body.cpp)
__global__ void k_integrate(Bodies B, uint32_t N, float dt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 ii1 = B.Ii1[i];
    if (ii1.z == 0.0f) return;   // static
    float4 x4 = B.x[i], q4 = B.q[i], v4 = B.v[i], w4 = B.w[i], f = B.f[i], tq = B.t[i];
    Q q = quat(q4);
    M3 R = mat3_cast(q);
    SM Iw = world_tensor(R, sm_from(B.I0[i], B.I1[i]));
    SM Iwi = world_tensor(R, sm_from(B.Ii0[i], ii1));
    B.px[i] = x4;
    B.pq[i] = q4;
    V3 v = xyz(v4) + xyz(f) * x4.w * dt;
    V3 x = xyz(x4) + v * dt;
    V3 w = xyz(w4);
    w = w + smv(sscale(Iwi, dt), xyz(tq) - cross(w, smv(Iw, w)));
    V3 ws = (0.5f * dt) * w;
    q = qstep(q, ws);
    B.x[i] = f4(x, x4.w);
    B.q[i] = f4q(q);
    B.v[i] = f4(v, 0.0f);
    B.w[i] = f4(w, 0.0f);
}

__global__ void k_update_vel(Bodies B, uint32_t N, float inv_dt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (B.Ii1[i].z == 0.0f) return;
    V3 x = xyz(B.x[i]), px = xyz(B.px[i]);
    Q dq = qmul(quat(B.q[i]), conj(quat(B.pq[i])));
    V3 v = (x - px) * inv_dt;
    V3 w = (2.0f * inv_dt) * V3{dq.x, dq.y, dq.z};
    B.v[i] = f4(v, 0.0f);
    B.w[i] = f4(dq.w >= 0.0f ? w : -w, 0.0f);
}

// ------------------------------------------------------------------ contact solve (Contact.cpp)
struct BodyReg {
    V3 x;
    Q q;
    float im;
    SM Ii;       // local inverse inertia
    bool dyn;
    SM Iwi;      // world inverse inertia at the last generalized-mass evaluation
};

__device__ __forceinline__ void load_body(const Bodies &B, uint32_t i, BodyReg &r) {
    float4 x4 = B.x[i];
    float4 a = B.Ii0[i], b = B.Ii1[i];
    r.x = xyz(x4);
    r.im = x4.w;
    r.q = quat(B.q[i]);
    r.Ii = sm_from(a, b);
    r.dyn = b.z != 0.0f;
}
// get_positional_generalized_inverse_mass (Body.cpp:94-101): refreshes the world
// tensor from the current orientation.
__device__ __forceinline__ float w_pos(BodyReg &r, V3 rr, V3 n) {
    if (!r.dyn) return 0.0f;
    r.Iwi = world_tensor(mat3_cast(r.q), r.Ii);
    V3 c = cross(rr, n);
    return r.im + dot(c, smv(r.Iwi, c));
}
__device__ __forceinline__ void apply_pos(BodyReg &r, V3 imp, V3 rr) {   // Body.cpp:59-68
    if (!r.dyn) return;
    r.x = r.x + imp * r.im;
    V3 rot = smv(r.Iwi, cross(rr, imp)) * 0.5f;
    r.q = qstep(r.q, rot);
}

__device__ __forceinline__ void solve_pos_pair(const Bodies &B, Contacts &K, uint32_t s) {
    const int np = K.cnt[s];
    if (np == 0) return;
    const uint2 bb = K.bb[s];
    const float4 nr = K.nr[s];
    const V3 n = xyz(nr);
    const float sf = K.fr[s].x;
    BodyReg A, Bd;
    load_body(B, bb.x, A);
    load_body(B, bb.y, Bd);
    const V3 pxa = xyz(B.px[bb.x]), pxb = xyz(B.px[bb.y]);
    const Q pqa = quat(B.pq[bb.x]), pqb = quat(B.pq[bb.y]);
    const V3 va = xyz(B.v[bb.x]), vb = xyz(B.v[bb.y]);
    const V3 wa = xyz(B.w[bb.x]), wb = xyz(B.w[bb.y]);
    for (int i = 0; i < np; ++i) {
        const uint32_t k = s * 4 + i;
        const V3 rla = xyz(K.ra[k]), rlb = xyz(K.rb[k]);
        float2 lam = K.lam[k];
        // Contact.cpp:35-75
        V3 pa = A.x + rotate(A.q, rla);
        V3 pb = Bd.x + rotate(Bd.q, rlb);
        float d = dot(pa - pb, n);
        if (d <= 0.0f) {
            K.pa[k] = f4(pa, 0.0f);
            K.pb[k] = f4(pb, 0.0f);
            continue;
        }
        V3 raw = pa - A.x, rbw = pb - Bd.x;
        V3 vr = (va + cross(wa, raw)) - (vb + cross(wb, rbw));
        const float relv = dot(vr, n);
        // Contact.cpp:81-115
        {
            float wA = w_pos(A, raw, n), wB = w_pos(Bd, rbw, n);
            float den = wA + wB;
            if (!(den < EPS)) {
                float dl = -d / den;
                lam.x += dl;
                if (lam.x > 0.0f) lam.x = 0.0f;
                V3 imp = dl * n;
                apply_pos(A, imp, raw);
                apply_pos(Bd, -imp, rbw);
            }
        }
        // Contact.cpp:121-195
        V3 ra = rotate(A.q, rla), rb = rotate(Bd.q, rlb);
        V3 pac = A.x + ra, pbc = Bd.x + rb;
        K.pa[k] = f4(pac, relv);
        K.pb[k] = f4(pbc, 1.0f);
        V3 ppa = pxa + rotate(pqa, rla);
        V3 ppb = pxb + rotate(pqb, rlb);
        V3 dp = (pac - ppa) - (pbc - ppb);
        V3 dpt = dp - dot(dp, n) * n;
        float slip = len(dpt);
        if (!(slip < EPS)) {
            V3 tv = normalize(dpt);
            float wA = w_pos(A, ra, tv), wB = w_pos(Bd, rb, tv);
            float den = wA + wB;
            if (!(den < EPS)) {
                float dl = -slip / den;
                float old = lam.y;
                lam.y += dl;
                float cone = sf * lam.x;
                if (lam.y < cone) lam.y = cone;
                float act = lam.y - old;
                if (!(absf(act) < EPS)) {
                    V3 imp = act * tv;
                    apply_pos(A, imp, ra);
                    apply_pos(Bd, -imp, rb);
                }
            }
        }
        K.lam[k] = lam;
    }
    if (A.dyn) { B.x[bb.x] = f4(A.x, A.im); B.q[bb.x] = f4q(A.q); }
    if (Bd.dyn) { B.x[bb.y] = f4(Bd.x, Bd.im); B.q[bb.y] = f4q(Bd.q); }
}

struct VelReg {
    V3 x, v, w;
    float im;
    SM Ii, Iwi;
    Q q;
    bool dyn;
};
__device__ __forceinline__ void load_vel(const Bodies &B, uint32_t i, VelReg &r) {
    float4 x4 = B.x[i];
    float4 a = B.Ii0[i], b = B.Ii1[i];
    r.x = xyz(x4);
    r.im = x4.w;
    r.q = quat(B.q[i]);
    r.v = xyz(B.v[i]);
    r.w = xyz(B.w[i]);
    r.Ii = sm_from(a, b);
    r.dyn = b.z != 0.0f;
}
__device__ __forceinline__ float w_vel(VelReg &r, V3 rr, V3 n) {
    if (!r.dyn) return 0.0f;
    r.Iwi = world_tensor(mat3_cast(r.q), r.Ii);
    V3 c = cross(rr, n);
    return r.im + dot(c, smv(r.Iwi, c));
}
__device__ __forceinline__ void apply_vel(VelReg &r, V3 imp, V3 rr) {   // Body.cpp:79-85
    if (!r.dyn) return;
    r.v = r.v + imp * r.im;
    r.w = r.w + smv(r.Iwi, cross(rr, imp));
}

__device__ __forceinline__ void solve_vel_pair(const Bodies &B, Contacts &K, uint32_t s, float h) {
    const int np = K.cnt[s];
    if (np == 0) return;
    const uint2 bb = K.bb[s];
    const float4 nr = K.nr[s];
    const V3 n = xyz(nr);
    const float rest = nr.w, df = K.fr[s].y;
    VelReg A, Bd;
    load_vel(B, bb.x, A);
    load_vel(B, bb.y, Bd);
    bool any = false;
    for (int i = 0; i < np; ++i) {
        const uint32_t k = s * 4 + i;
        const float4 pb4 = K.pb[k];
        if (pb4.w == 0.0f) continue;   // no collision at the position level
        any = true;
        const float4 pa4 = K.pa[k];
        const float relv = pa4.w;
        const float lamn = K.lam[k].x;
        // Contact.cpp:206-289
        V3 raw = xyz(pa4) - A.x, rbw = xyz(pb4) - Bd.x;
        V3 vr = (A.v + cross(A.w, raw)) - (Bd.v + cross(Bd.w, rbw));
        float vn = dot(vr, n);
        V3 vt = vr - n * vn;
        float vtm = len(vt);
        float e = rest;
        if (absf(relv) <= 2.0f * 9.8f * h) e = 0.0f;
        float vtarget = fmin_ref(-e * relv, 0.0f);
        float dvn = vtarget - vn;
        float dvt = 0.0f;
        if (vtm > EPS) {
            float lim = df * absf(lamn) / h;
            dvt = -fmin_ref(lim, vtm);
        }
        float wan = w_vel(A, raw, n), wbn = w_vel(Bd, rbw, n);
        float denn = wan + wbn;
        if (denn > EPS && absf(dvn) > EPS) {
            V3 imp = (dvn / denn) * n;
            apply_vel(A, imp, raw);
            apply_vel(Bd, -imp, rbw);
        }
        if (vtm > EPS && absf(dvt) > EPS) {
            V3 tv = vdiv(vt, vtm);
            float wat = w_vel(A, raw, tv), wbt = w_vel(Bd, rbw, tv);
            float dent = wat + wbt;
            if (dent > EPS) {
                V3 imp = (dvt / dent) * tv;
                apply_vel(A, imp, raw);
                apply_vel(Bd, -imp, rbw);
            }
        }
    }
    if (!any) return;
    if (A.dyn) { B.v[bb.x] = f4(A.v, 0.0f); B.w[bb.x] = f4(A.w, 0.0f); }
    if (Bd.dyn) { B.v[bb.y] = f4(Bd.v, 0.0f); B.w[bb.y] = f4(Bd.w, 0.0f); }
}

// Position / velocity solve of one substep: colour after colour, all in one cooperative
// launch; pairs of one colour touch disjoint dynamic bodies (grid barrier between colours).
__global__ void __launch_bounds__(256) k_solve_pos_coop(Bodies B, Contacts K, const uint32_t *order, const uint32_t *col_off,
                                                        const uint32_t *ncol, GridBar bar) {
    const uint32_t nc = *ncol;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t o = col_off[c], e = col_off[c + 1];
        for (uint32_t t = o + blockIdx.x * blockDim.x + threadIdx.x; t < e; t += stride) solve_pos_pair(B, K, order[t]);
        if (c + 1 < nc) grid_barrier(bar, nullptr);
    }
}
__global__ void __launch_bounds__(256) k_solve_vel_coop(Bodies B, Contacts K, const uint32_t *order, const uint32_t *col_off,
                                                        const uint32_t *ncol, float h, GridBar bar) {
    const uint32_t nc = *ncol;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t o = col_off[c], e = col_off[c + 1];
        for (uint32_t t = o + blockIdx.x * blockDim.x + threadIdx.x; t < e; t += stride) solve_vel_pair(B, K, order[t], h);
        if (c + 1 < nc) grid_barrier(bar, nullptr);
    }
}

}  // namespace rb
