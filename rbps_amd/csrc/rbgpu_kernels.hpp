// CDNA4 kernels of the XPBD substep pipeline.
//
// Data layout (all FP32):
//   bodies     one 128-byte row (a cache line) per body: x (position, inverse mass),
//              q (w,x,y,z), v, w (velocities), px, pq (pose at the substep start),
//              ii0/ii1 (local inverse inertia xx yy zz xy | xz yz dynamic -); beside it
//              f (force, mass), t (torque), I0/I1 (local inertia), read by integrate only
//   colliders  meta (kind, body, is_static, bp slot)  dims, lpos, lrot, mat
//   SAP        fat AABB per slot, endpoint keys/values (stable across frames),
//              endpoint positions, sorted sweep rows
//   pairs      frame pair list (canonical), colour, solve slots; a slot's contact
//              header (32 B) and its four points' blocks (64 B each) are contiguous.
#pragma once
#include "rbgpu_collide.hpp"
#include "rbgpu_sort.hpp"

namespace rb {

constexpr uint32_t BP_INVALID = 0xffffffffu;
constexpr int MAX_COLOURS = 1024;
constexpr int NBINS = 13;
// pair-type bins: analytic pairs, shape-plane, hull-hull, hull with a sphere / box /
// capsule (either order), hull-plane
enum Bin { BIN_SS = 0, BIN_SB, BIN_BB, BIN_SC, BIN_CC, BIN_CB, BIN_SP, BIN_BP, BIN_CP, BIN_HH, BIN_HX, BIN_HP, BIN_OTHER };

// One 128-byte row per body (one cache line): everything the contact and joint solves
// gather for a body, so a gather touches one line.
struct BodyRow {
    float4 x;     // position, inverse mass
    float4 q;     // orientation (w, x, y, z)
    float4 v;     // linear velocity
    float4 w;     // angular velocity
    float4 px;    // position at the start of the substep
    float4 pq;    // orientation at the start of the substep
    float4 ii0;   // local inverse inertia xx yy zz xy
    float4 ii1;   // local inverse inertia xz yz, dynamic flag (1 dynamic, 0 static)
};
static_assert(sizeof(BodyRow) == 128, "a body row is one cache line");
struct Bodies {
    BodyRow *r;
    float4 *f, *t, *I0, *I1;   // force (w: mass), torque, local inertia (integrate only)
};
struct Colliders {
    int4 *meta;      // kind, body, is_static, bp slot
    float4 *dims, *lpos, *lrot, *mat;   // mat: restitution, static friction, dynamic friction
    int *hull;       // hull index of a hull collider (-1 otherwise)
};

// Convex hull tables of a world (built on the host from the caller's vertices and
// triangles; see rbgpu_hull.hpp).
struct HullMeta {   // 64 bytes; table offsets index HullSet::tab
    uint32_t v0, nv, pn0, np;
    uint32_t poff, pidx, v2p_o, v2p_i;
    uint32_t v2n_o, v2n_i, p2p_o, p2p_i;
    uint32_t adj, pad0, pad1, pad2;
};
struct HullSet {
    const float4 *v;       // local vertices (x y z 0), all hulls
    const float4 *pn;      // polygon normals
    const uint32_t *tab;   // polygon loops and adjacency lists (CSR, hull-local indices)
    const HullMeta *meta;
    const float4 *amin, *amax;   // local AABB
    const float4 *ell;           // per hull: bounding ellipsoid centre, then its semi-axes (local frame)
};
struct Contacts {
    int *cnt;        // points per slot this substep
    uint2 *cc;       // colliders (a, b) per slot
    uint2 *bb;       // bodies (a, b) per slot (read by the colouring)
    // per slot, one 32-byte header: (normal, restitution), (bodies a and b as bits, static
    // friction, dynamic friction)
    float4 *hd;
    // per point (4 per slot), one 64-byte block: (r_a body-local, normal lambda),
    // (r_b body-local, tangent lambda), (p_a world, pre-solve normal velocity),
    // (p_b world, collision flag)
    float4 *pt;
    float *dep;      // per point: penetration depth
};
// Rows of a slot's header and of a point's block (point k = slot * 4 + i).
__device__ __forceinline__ float4 *hd_row(const Contacts &K, uint32_t s) { return K.hd + 2 * (size_t)s; }
__device__ __forceinline__ float4 *pt_row(const Contacts &K, uint32_t k) { return K.pt + 4 * (size_t)k; }

__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ Q quat(float4 a) { return Q{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float4 f4q(Q q) { return make_float4(q.w, q.x, q.y, q.z); }
__device__ __forceinline__ SM sm_from(float4 a, float4 b) { return SM{a.x, a.y, a.z, a.w, b.x, b.y}; }
__device__ __forceinline__ Shape shape_of(int kind, float4 d) { return Shape{kind, d.x, d.y, d.z, d.w}; }
__device__ __forceinline__ TF collider_tf(const Colliders &C, const Bodies &B, uint32_t c, uint32_t body) {
    V3 bp = xyz(B.r[body].x);
    Q bq = quat(B.r[body].q);
    return TF{bp + rotate(bq, xyz(C.lpos[c])), qmul(bq, quat(C.lrot[c]))};   // Collider.hpp:73-81
}

// ------------------------------------------------------------------ broadphase
// Refit of one collider's swept AABB with fat-AABB hysteresis
// (CollisionPipeline.cpp:14-53, BroadPhase.hpp:228-249, BroadPhase.cpp:211-232).
__global__ void k_refit(Colliders C, Bodies B, HullSet hs, uint32_t M, float dt, float margin, float reinsert,
                        float4 *fmin, float4 *fmax, uint32_t *ep_key, const uint2 *ep_pos, int *dirty) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int4 m = C.meta[i];
    if (m.z != 0 || (uint32_t)m.w == BP_INVALID) return;
    const uint32_t body = (uint32_t)m.y, slot = (uint32_t)m.w;
    const TF tf = collider_tf(C, B, i, body);
    Box3 t;
    if (m.x == SH_HULL) {
        const float4 a = hs.amin[C.hull[i]], b = hs.amax[C.hull[i]];
        t = hull_aabb(V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, tf);
    } else {
        t = shape_aabb(shape_of(m.x, C.dims[i]), tf);
    }
    V3 dv = xyz(B.r[body].v) * dt;
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

// Scene of each endpoint (key of the batch sort passes that follow the value passes).
__global__ void k_ep_scene_key(const uint32_t *ep_val, const uint32_t *slot_scene, uint32_t nep, uint32_t *key,
                               const int *dirty) {
    if (*dirty == 0) return;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nep) key[j] = slot_scene[ep_val[j] >> 1];
}

// Sorted sweep rows, endpoint positions and endpoint keys after a sort (the batch sort
// reorders values without their keys; a key is float_key of its fat-box bound).
// Row of endpoint j, two float4: (x_min, x_max, y_min, y_max) and (z_min, z_max, filter
// word 0, filter word 1) of its slot's fat box, the words as bits: collider | static << 31
// and body | is_max << 31 (everything the sweep's pair filter reads).
__global__ void k_sweep_prep(const uint32_t *ep_val, uint32_t nep, const float4 *fmin, const float4 *fmax, uint2 *ep_pos,
                             const uint32_t *uid, Colliders C, float4 *sw_a, float4 *sw_b, uint32_t *ep_key,
                             const int *dirty) {
    if (*dirty == 0) return;
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nep) return;
    uint32_t v = ep_val[j];
    uint32_t slot = v >> 1, is_max = v & 1u;
    if (is_max) ep_pos[slot].y = j; else ep_pos[slot].x = j;
    ep_key[j] = float_key(is_max ? fmax[slot].x : fmin[slot].x);
    float4 a = fmin[slot], b = fmax[slot];
    const uint32_t c = uid[slot];
    const int4 m = C.meta[c];
    const uint32_t w0 = c | (m.z != 0 ? 0x80000000u : 0u), w1 = ((uint32_t)m.y & 0x7fffffffu) | (is_max << 31);
    sw_a[j] = make_float4(a.x, b.x, a.y, b.y);
    sw_b[j] = make_float4(a.z, b.z, __uint_as_float(w0), __uint_as_float(w1));
}

// Sweep: the range of every broadphase slot, from its min endpoint to its own max
// endpoint, is scanned for the min endpoints it contains, so every pair is found once,
// from the object that opens first (same set as the active-set sweep of
// BroadPhase.cpp:104-152, inclusive 3-D fat-AABB test AABB.hpp:54-59).
// One 16-lane row (a quarter wavefront) per slot: a step tests the 16 endpoints that
// follow, loaded together, and the row's ballot ranks the hits, so each slot's pairs
// are written in endpoint order at the offset its count gave (the count pass, EMIT
// false, writes the slot's count at its min endpoint and 0 at its max endpoint).
constexpr int SWEEP_ROW = 16;
template <bool EMIT, int W = SWEEP_ROW>
__global__ void __launch_bounds__(256) k_sweep(const float4 *sw_a, const float4 *sw_b, const uint2 *ep_pos,
                                               uint32_t nslot, uint32_t *cnt, const uint32_t *off, uint2 *pairs,
                                               uint32_t cap, int *overflow) {
    static_assert(W == 4 || W == 8 || W == 16 || W == 32, "sweep row width");
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t / W, r = t % W;
    const uint32_t sh = (threadIdx.x & 63u) & ~(uint32_t)(W - 1);   // the row's first lane
    if (s >= nslot) return;   // whole rows: nslot rows of W threads
    const uint2 p = ep_pos[s];
    const uint32_t i = p.x, end = p.y;
    const float4 A = sw_a[i], Ab = sw_b[i];
    const uint32_t a0 = __float_as_uint(Ab.z), a1 = __float_as_uint(Ab.w);
    const uint32_t ua = a0 & 0x7fffffffu;
    uint32_t n = 0;
    const uint32_t o = EMIT ? off[i] : 0u;
    for (uint32_t base = i + 1; base < end; base += W) {   // row-uniform trip count
        const uint32_t j = base + r;
        bool hit = false;
        uint32_t lo = 0, hi = 0;
        if (j < end) {
            const float4 Bv = sw_a[j], Bb = sw_b[j];
            const uint32_t b0 = __float_as_uint(Bb.z), b1 = __float_as_uint(Bb.w);
            // a min endpoint whose box overlaps; pair filter (CollisionPipeline.cpp:121-133):
            // not both static, not the same body
            if (!(b1 >> 31) && (A.x <= Bv.y && A.y >= Bv.x) && (A.z <= Bv.w && A.w >= Bv.z) &&
                (Ab.x <= Bb.y && Ab.y >= Bb.x) && !((a0 & b0) >> 31) && ((a1 ^ b1) & 0x7fffffffu) != 0u) {
                const uint32_t ub = b0 & 0x7fffffffu;
                lo = ua < ub ? ua : ub;
                hi = ua < ub ? ub : ua;
                hit = true;
            }
        }
        const uint32_t m = (uint32_t)(__ballot(hit) >> sh) & (uint32_t)((1ull << W) - 1ull);
        if (EMIT && hit) {
            const uint32_t k = o + n + (uint32_t)__popc(m & ((1u << r) - 1u));
            if (k < cap) pairs[k] = make_uint2(lo, hi);
            else *overflow = 1;
        }
        n += (uint32_t)__popc(m);
    }
    if (!EMIT && r == 0) {
        cnt[i] = n;
        cnt[end] = 0u;
    }
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

__device__ __forceinline__ int bin_of(int ka, int kb) {
    if (ka == SH_HULL || kb == SH_HULL) {
        if (ka == SH_PLANE || kb == SH_PLANE) return kb == SH_PLANE ? BIN_HP : BIN_OTHER;   // plane always B
        return (ka == SH_HULL && kb == SH_HULL) ? BIN_HH : BIN_HX;
    }
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
    float4 *h = hd_row(K, p);
    h[0] = make_float4(0.0f, 0.0f, 0.0f, rest);
    h[1] = make_float4(__uint_as_float((uint32_t)ma_.y), __uint_as_float((uint32_t)mb_.y), (ma.y + mb.y) * 0.5f,
                       (ma.z + mb.z) * 0.5f);
    bin_key[p] = (uint32_t)bin_of(ma_.x, mb_.x);
    bin_val[p] = p;
}

// Local pair ids (a pair's index in its scene's own list: SAP pairs, then plane pairs).
// SAP pairs are scene-major (the batch sort makes every scene's endpoints contiguous), so
// a scene's SAP pairs form one run [start, end).
__global__ void k_scene_bounds(const uint2 *pairs, const uint32_t *n_sap_dev, const uint32_t *cscene, uint32_t *sap_start,
                               uint32_t *sap_end) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *n_sap_dev;
    if (p >= n) return;
    const uint32_t s = cscene[pairs[p].x];
    if (p == 0 || cscene[pairs[p - 1].x] != s) sap_start[s] = p;
    if (p + 1 == n || cscene[pairs[p + 1].x] != s) sap_end[s] = p + 1;
}
__global__ void k_pair_lid(const uint2 *pairs, const uint32_t *n_sap_dev, const uint32_t *n_pairs_dev, const uint32_t *cscene,
                           const uint32_t *sap_start, const uint32_t *sap_end, const uint32_t *plane_lid, uint32_t *lid) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ns = *n_sap_dev;
    if (p >= *n_pairs_dev) return;
    const uint32_t s = cscene[pairs[p].x];
    lid[p] = p < ns ? p - sap_start[s] : (sap_end[s] - sap_start[s]) + plane_lid[p - ns];
}

// ------------------------------------------------------------------ substep colouring of active pairs
// Pairs with contact points are coloured every substep; two pairs conflict when they share a
// dynamic body.  Synchronous randomized rounds whose cost does not depend on body degree.
// Per dynamic body v: used(v) = colours taken by its pairs (bit mask), u(v) = its uncoloured
// pairs, claim(v, slot) = the largest key that targeted the slot.  Round r, uncoloured pair p:
//   claim:  t = the (tent_hash(p, r) mod min(max(ceil(k/4), 1), pal_max))-th colour outside the
//           union of used(v), k = sum of u(v); key = (r + 1) << 32 | rank_hash(p, r);
//           atomicMax(claim(v, t mod COLOUR_SLOTS), key) for each dynamic body v;
//   settle: p takes t if its key is the claim at each of its dynamic bodies; then it sets
//           t in used(v) and decrements u(v).
// One key per slot means two pairs of one body never settle on one colour in one round, and
// used(v) keeps later rounds off taken colours: the colouring is valid.  The largest key
// always settles, so every round makes progress.  Keys carry the round, so claims of earlier
// rounds lose without clearing.  Outcomes depend only on the keys, not on thread or list
// order, so work lists may be built with atomics.  Arrays are indexed by pair id; the
// hashes take the pair's local id (its index in its own scene's list), so every scene of
// a batch is coloured exactly as it would be alone.
constexpr uint32_t COLOUR_SLOTS = 16;
constexpr uint32_t COLOUR_WORDS = MAX_COLOURS / 32;
constexpr uint32_t MAX_HEAD_ROUNDS = 256;
constexpr uint32_t TAIL_THREADS = 1024;
constexpr uint32_t LEFT_STRIPES = 64;   // the uncoloured count is striped over 64 words
// ctr[]: counters of the colouring (cleared every substep)
enum {
    CTR_ACT = 0,      // active pairs (the scan total of the per-block counts)
    CTR_REST = 1,     // rest list length
    CTR_ROUNDS = 2,   // rounds used
    CTR_NCOL = 3,     // colours used
    CTR_ERR = 4,      // 1: colour capacity exceeded, 2: no convergence
    CTR_LEFT = 64,    // LEFT_STRIPES words; their sum (mod 2^32) = uncoloured pairs
    CTR_FLAGS = CTR_LEFT + LEFT_STRIPES,   // one word per grid-wide round: had pairs
    CTR_SIZE = CTR_FLAGS + MAX_HEAD_ROUNDS
};
constexpr int32_t COLOUR_NONE = -2;     // a pair without points this substep
constexpr int32_t COLOUR_OPEN = -1;     // an active pair not coloured yet

struct ColourBufs {
    uint32_t *act;              // active pair ids, ascending (CTR_ACT of them)
    uint32_t *blk;              // per-block active counts, then their exclusive scan
    int *acol;                  // colour per pair id: >= 0, COLOUR_OPEN or COLOUR_NONE
    uint32_t *tent;             // tentative colour per pair id
    uint32_t *rest;             // pair ids left to the single-workgroup rounds
    uint32_t *used;             // N x COLOUR_WORDS colour masks
    uint32_t *ucnt;             // N uncoloured-pair counts
    unsigned long long *claim;  // N x COLOUR_SLOTS
    uint32_t *hist, *off, *cur; // per colour: size, start (MAX_COLOURS + 1), scatter cursor
    uint32_t *ctr;              // CTR_*
    const uint32_t *lid;        // local pair id per pair (the hashes' input)
    uint32_t pal_max;           // largest palette a pair draws its tentative colour from
    uint32_t *epoch;            // per-scene path: substeps since the last full clear (claim keys)
};
// High word of a round's slot-claim key.  The grid-wide path clears the claims every substep
// and uses r + 1; the per-scene path puts a per-substep epoch above it (stale claims of earlier
// substeps are then smaller than any key of the running round, so only the used masks need a
// clear, which each scene does for its own bodies).  Only keys of one round are ever compared,
// so the outcome is the same either way.
constexpr uint32_t COLOUR_EPOCH_SHIFT = 17;   // r + 1 <= 2^16 (the rounds cap)
constexpr uint32_t COLOUR_EPOCH_MAX = (1u << (32 - COLOUR_EPOCH_SHIFT)) - 1u;
__device__ __forceinline__ unsigned long long claim_key(uint32_t kh, uint32_t lo) {
    return ((unsigned long long)kh << 32) | lo;
}

__device__ __forceinline__ uint32_t tent_hash(uint32_t p, uint32_t r) { return mix32(p ^ (r * 0x9e3779b9u)); }
__device__ __forceinline__ uint32_t rank_hash(uint32_t p, uint32_t r) { return mix32(p ^ (r * 0x85ebca6bu) ^ 0xc2b2ae35u); }
__device__ __forceinline__ uint32_t palette(uint32_t k, uint32_t pal_max) {
    const uint32_t m = (k + 3u) / 4u;
    return m < 1u ? 1u : (m > pal_max ? pal_max : m);
}
__device__ __forceinline__ bool is_dynamic(const Bodies &B, uint32_t b) { return B.r[b].ii1.z != 0.0f; }
// Loads of words that other threads update with atomics in the same kernel.
__device__ __forceinline__ uint32_t ld_shared(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_shared(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Rank of this lane among the set lanes of a wave-wide ballot.
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Wave-aggregated append: one atomic per wave; every lane of the wave must call it.
__device__ __forceinline__ void wave_append(bool take, uint32_t val, uint32_t *counter, uint32_t *list) {
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (take) list[base + lane_rank(m)] = val;
}
// Sum of the stripes, by one wave (every lane gets it).
__device__ __forceinline__ uint32_t uncoloured_left(const uint32_t *ctr) {
    uint32_t v = ld_shared(&ctr[CTR_LEFT + (threadIdx.x & 63)]);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Clears the per-body colouring state and the counters.  Grid covers N * COLOUR_WORDS / 4.
__global__ void k_colour_reset(uint32_t N, ColourBufs cb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    if (i < N * (COLOUR_WORDS / 4)) reinterpret_cast<uint4 *>(cb.used)[i] = z;
    if (i < N * (COLOUR_SLOTS / 2)) reinterpret_cast<uint4 *>(cb.claim)[i] = z;
    if (i < N) cb.ucnt[i] = 0u;
    if (i < (uint32_t)MAX_COLOURS) cb.hist[i] = 0u;
    if (i < (uint32_t)CTR_SIZE) cb.ctr[i] = 0u;
    if (i == 0 && cb.epoch) *cb.epoch = 1u;
}
// Per-scene path, the substeps after a full clear: the counters only, the next epoch.
__global__ void __launch_bounds__(512) k_colour_ctr_reset(ColourBufs cb) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)CTR_SIZE; i += blockDim.x) cb.ctr[i] = 0u;
    if (threadIdx.x == 0) *cb.epoch += 1u;
}

// Marks the pairs with points as open, counts them per dynamic body and per block.
__global__ void __launch_bounds__(256) k_colour_activate(const int *cnt, uint32_t P, const uint2 *pbod, Bodies B,
                                                         ColourBufs cb) {
    __shared__ uint32_t sh;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool a = p < P && cnt[p] > 0;
    if (p < P) cb.acol[p] = a ? COLOUR_OPEN : COLOUR_NONE;
    if (a) {
        const uint2 bb = pbod[p];
        if (is_dynamic(B, bb.x)) atomicAdd(&cb.ucnt[bb.x], 1u);
        if (is_dynamic(B, bb.y)) atomicAdd(&cb.ucnt[bb.y], 1u);
    }
    if (threadIdx.x == 0) sh = 0u;
    __syncthreads();
    const uint64_t m = __ballot(a);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&sh, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) cb.blk[blockIdx.x] = sh;
}
// Active list in ascending pair order from the scanned block counts; the uncoloured
// count starts at the total.
__global__ void __launch_bounds__(256) k_colour_compact(uint32_t P, ColourBufs cb) {
    __shared__ uint32_t wc[4];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool a = p < P && cb.acol[p] == COLOUR_OPEN;
    const uint64_t m = __ballot(a);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wc[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = cb.blk[blockIdx.x];
    for (int w = 0; w < wave; ++w) base += wc[w];
    if (a) cb.act[base + lane_rank(m)] = p;
    if (p == 0) cb.ctr[CTR_LEFT] = cb.ctr[CTR_ACT];
}

__device__ __forceinline__ void colour_claim(const ColourBufs &cb, const uint2 *pbod, const Bodies &B, uint32_t p,
                                             uint32_t r, uint32_t kh) {
    const uint2 bb = pbod[p];
    const bool da = is_dynamic(B, bb.x), db = is_dynamic(B, bb.y);
    const uint32_t k = (da ? ld_shared(&cb.ucnt[bb.x]) : 0u) + (db ? ld_shared(&cb.ucnt[bb.y]) : 0u);
    const uint32_t q = cb.lid[p];
    uint32_t idx = tent_hash(q, r) % palette(k, cb.pal_max);
    uint32_t t = MAX_COLOURS;
    for (uint32_t w = 0; w < COLOUR_WORDS; ++w) {
        uint32_t fr = ~((da ? ld_shared(&cb.used[bb.x * COLOUR_WORDS + w]) : 0u) |
                        (db ? ld_shared(&cb.used[bb.y * COLOUR_WORDS + w]) : 0u));
        const uint32_t n = (uint32_t)__popc(fr);
        if (idx < n) {
            for (; idx; --idx) fr &= fr - 1u;
            t = w * 32u + (uint32_t)__builtin_ctz(fr);
            break;
        }
        idx -= n;
    }
    if (t >= (uint32_t)MAX_COLOURS) { atomicOr(&cb.ctr[CTR_ERR], 1u); t = MAX_COLOURS - 1; }
    cb.tent[p] = t;
    const unsigned long long key = claim_key(kh, rank_hash(q, r));
    if (da) atomicMax(&cb.claim[bb.x * COLOUR_SLOTS + t % COLOUR_SLOTS], key);
    if (db) atomicMax(&cb.claim[bb.y * COLOUR_SLOTS + t % COLOUR_SLOTS], key);
}

__device__ __forceinline__ bool colour_settle(const ColourBufs &cb, const uint2 *pbod, const Bodies &B, uint32_t p,
                                              uint32_t r, uint32_t kh) {
    const uint2 bb = pbod[p];
    const bool da = is_dynamic(B, bb.x), db = is_dynamic(B, bb.y);
    const uint32_t t = cb.tent[p];
    const unsigned long long key = claim_key(kh, rank_hash(cb.lid[p], r));
    if (da && ld_shared(&cb.claim[bb.x * COLOUR_SLOTS + t % COLOUR_SLOTS]) != key) return false;
    if (db && ld_shared(&cb.claim[bb.y * COLOUR_SLOTS + t % COLOUR_SLOTS]) != key) return false;
    cb.acol[p] = (int)t;
    const uint32_t bit = 1u << (t % 32u);
    if (da) { atomicOr(&cb.used[bb.x * COLOUR_WORDS + t / 32u], bit); atomicSub(&cb.ucnt[bb.x], 1u); }
    if (db) { atomicOr(&cb.used[bb.y * COLOUR_WORDS + t / 32u], bit); atomicSub(&cb.ucnt[bb.y], 1u); }
    return true;
}

// Grid-wide round, claim phase, over the active list (grid-stride).  Skipped once at most
// `tail_min` pairs are uncoloured (the uncoloured count changes only in settle launches,
// so every block of one claim launch decides alike).  Marks round r as run.
__global__ void __launch_bounds__(256) k_colour_claim(const uint2 *pbod, Bodies B, ColourBufs cb, uint32_t r,
                                                      uint32_t tail_min) {
    if (r > 0 && uncoloured_left(cb.ctr) <= tail_min) return;
    const uint32_t n = cb.ctr[CTR_ACT];
    bool any = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = cb.act[i];
        if (cb.acol[p] == COLOUR_OPEN) {
            any = true;
            colour_claim(cb, pbod, B, p, r, r + 1u);
        }
    }
    if (__ballot(any) && (threadIdx.x & 63) == 0) cb.ctr[CTR_FLAGS + r] = 1u;
}
// Grid-wide round, settle phase (runs exactly when its claim phase had pairs).
__global__ void __launch_bounds__(256) k_colour_settle(const uint2 *pbod, Bodies B, ColourBufs cb, uint32_t r) {
    __shared__ uint32_t sh;
    if (!cb.ctr[CTR_FLAGS + r]) return;   // the claim phase of round r was skipped or empty
    const uint32_t n = cb.ctr[CTR_ACT];
    uint32_t won = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = cb.act[i];
        if (cb.acol[p] == COLOUR_OPEN && colour_settle(cb, pbod, B, p, r, r + 1u)) ++won;
    }
    for (int o = 32; o > 0; o >>= 1) won += __shfl_xor(won, o, 64);
    if (threadIdx.x == 0) sh = 0u;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && won) atomicAdd(&sh, won);
    __syncthreads();
    if (threadIdx.x == 0 && sh) atomicSub(&cb.ctr[CTR_LEFT + blockIdx.x % LEFT_STRIPES], sh);
}
// Collects the uncoloured pairs into the rest list (after the grid-wide rounds).
__global__ void k_colour_collect(ColourBufs cb) {
    const uint32_t n = cb.ctr[CTR_ACT];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {   // every lane of a wave iterates alike
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t p = i < n ? cb.act[i] : 0u;
        wave_append(i < n && cb.acol[p] == COLOUR_OPEN, p, &cb.ctr[CTR_REST], cb.rest);
    }
}

// The remaining rounds in one workgroup: the rest list is small after the grid-wide rounds,
// and a workgroup barrier costs far less than a kernel boundary.  Starts at the first round
// the grid-wide launches skipped (r_max when none did) and runs until every pair is
// coloured (or `cap` rounds: error 2).  Records the round count.
__global__ void __launch_bounds__(TAIL_THREADS) k_colour_tail(const uint2 *pbod, Bodies B, ColourBufs cb, uint32_t r_max,
                                                                uint32_t cap) {
    __shared__ uint32_t won;
    const uint32_t n = cb.ctr[CTR_REST];
    uint32_t r0 = r_max;
    for (uint32_t q = 0; q < r_max; ++q)
        if (!cb.ctr[CTR_FLAGS + q]) { r0 = q; break; }
    uint32_t left = n, r = r0;
    for (; left > 0 && r < cap; ++r) {
        if (threadIdx.x == 0) won = 0u;
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
            const uint32_t p = cb.rest[j];
            if (cb.acol[p] == COLOUR_OPEN) colour_claim(cb, pbod, B, p, r, r + 1u);
        }
        __syncthreads();
        uint32_t w = 0;
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
            const uint32_t p = cb.rest[j];
            if (cb.acol[p] == COLOUR_OPEN && colour_settle(cb, pbod, B, p, r, r + 1u)) ++w;
        }
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
        if ((threadIdx.x & 63) == 0 && w) atomicAdd(&won, w);
        __syncthreads();
        left -= won;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint32_t rounds = r;
        if (n == 0) {   // finished in the grid-wide rounds: count the rounds that had work
            rounds = 0;
            for (uint32_t q = 0; q < r0; ++q)
                if (cb.ctr[CTR_FLAGS + q]) rounds = q + 1;
        }
        cb.ctr[CTR_ROUNDS] = rounds;
        if (left) atomicOr(&cb.ctr[CTR_ERR], 2u);
    }
}

// Colour sizes (LDS histogram per block, grid-stride over the pairs).
__global__ void __launch_bounds__(256) k_colour_hist(ColourBufs cb) {
    __shared__ uint32_t h[MAX_COLOURS];
    for (uint32_t c = threadIdx.x; c < (uint32_t)MAX_COLOURS; c += blockDim.x) h[c] = 0u;
    __syncthreads();
    const uint32_t n = cb.ctr[CTR_ACT];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = cb.acol[cb.act[i]];
        if (c >= 0) atomicAdd(&h[c], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < (uint32_t)MAX_COLOURS; c += blockDim.x)
        if (h[c]) atomicAdd(&cb.hist[c], h[c]);
}
// Colour offsets (one block of MAX_COLOURS threads): off = exclusive scan, cur = off, ncol,
// active pairs.
__global__ void __launch_bounds__(MAX_COLOURS) k_colour_offsets(ColourBufs cb) {
    __shared__ uint32_t sh[MAX_COLOURS / 64 + 1];
    __shared__ uint32_t top;
    const uint32_t c = threadIdx.x;
    if (c == 0) top = 0u;
    const uint32_t v = cb.hist[c];
    uint32_t tot;
    const uint32_t o = block_exclusive_scan(v, sh, &tot);
    cb.off[c] = o;
    cb.cur[c] = o;
    if (c == 0) cb.off[MAX_COLOURS] = tot;
    if (v) atomicMax(&top, c + 1u);
    __syncthreads();
    if (c == 0) cb.ctr[CTR_NCOL] = top;
}
// Solve list ordered by colour: each block counts its chunk per colour in LDS, reserves
// ranges with one atomic per (block, colour), then scatters.
__global__ void __launch_bounds__(256) k_colour_scatter(ColourBufs cb, uint32_t *solve_list, uint32_t chunk) {
    __shared__ uint32_t h[MAX_COLOURS];
    __shared__ uint32_t base[MAX_COLOURS];
    for (uint32_t c = threadIdx.x; c < (uint32_t)MAX_COLOURS; c += blockDim.x) h[c] = 0u;
    __syncthreads();
    const uint32_t n = cb.ctr[CTR_ACT];
    const uint32_t i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int c = cb.acol[cb.act[i]];
        if (c >= 0) atomicAdd(&h[c], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < (uint32_t)MAX_COLOURS; c += blockDim.x) {
        base[c] = h[c] ? atomicAdd(&cb.cur[c], h[c]) : 0u;
        h[c] = 0u;
    }
    __syncthreads();
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const uint32_t p = cb.act[i];
        const int c = cb.acol[p];
        if (c >= 0) solve_list[base[c] + atomicAdd(&h[c], 1u)] = p;
    }
}

// ------------------------------------------------------------------ narrowphase
// Contact emission (CollisionPipeline.cpp:59-104): split the depth across the two
// surfaces and store body-local lever arms.
__device__ __forceinline__ void emit(const Manifold &m, const Bodies &B, Contacts &K, uint32_t s, uint2 bb) {
    V3 xa = xyz(B.r[bb.x].x), xb = xyz(B.r[bb.y].x);
    Q qa = conj(quat(B.r[bb.x].q)), qb = conj(quat(B.r[bb.y].q));
    float4 *h = hd_row(K, s);
    h[0] = f4(m.n, h[0].w);
#pragma unroll
    for (int i = 0; i < 4; ++i) {   // constant indices keep the manifold in registers
        if (i >= m.np) break;
        float half = m.dep[i] * 0.5f;
        V3 pa = m.p[i] + m.n * half;
        V3 pb = m.p[i] - m.n * half;
        uint32_t k = s * 4 + i;
        float4 *r = pt_row(K, k);
        r[0] = f4(rotate(qa, pa - xa), 0.0f);
        r[1] = f4(rotate(qb, pb - xb), 0.0f);
        r[2] = f4(pa, 0.0f);
        r[3] = f4(pb, 0.0f);
        K.dep[k] = m.dep[i];
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

// Bounding-sphere rejection ahead of the analytic routines whose separation test is exact
// (box-box: the 15-axis SAT is complete for boxes; sphere-box, sphere-capsule, capsule-
// capsule: closest points): when the bounding spheres are more than SEP_MARGIN apart the
// shapes are disjoint, these routines report no contact, and the margin keeps the decision
// far above rounding (and above box-box's 0.001 edge bias), so contact sets are unchanged.
// Not for capsule-box: the reference tests only 6 SAT axes there (CapsuleBox.hpp:51-97) and
// reports contacts for shapes several centimetres apart near box edges, which must be kept.
constexpr float SEP_MARGIN = 1e-2f;
__device__ __forceinline__ float bound_radius(const Shape &s) {
    return s.kind == SH_SPHERE ? s.a : (s.kind == SH_BOX ? sqrtf(s.a * s.a + s.b * s.b + s.c * s.c) : s.a + s.b);
}
template <int BIN>
__device__ __forceinline__ bool far_apart(const Shape &a, const TF &ta, const Shape &b, const TF &tb) {
    if (BIN != BIN_BB && BIN != BIN_SB && BIN != BIN_SC && BIN != BIN_CC) return false;
    const float r = bound_radius(a) + bound_radius(b) + SEP_MARGIN;
    return len2(ta.pos - tb.pos) > r * r;
}

// flip_cc (contact mode 1, capsule-capsule bin only): the reference's capsule-capsule routine
// returns a normal from B to A (CapsuleCapsule.hpp:68, delta = closest_a - closest_b) where
// every other routine returns one from A to B, so the solver pushes such a pair into each
// other; mode 1 turns that normal around (points and depths unchanged).
template <int BIN>
__global__ void __launch_bounds__(128) k_narrow(const uint32_t *bin_slot, uint32_t off, uint32_t n, Colliders C, Bodies B,
                                                Contacts K, int flip_cc = 0) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t s = bin_slot[off + t];
    uint2 cc = K.cc[s];
    uint2 bb = K.bb[s];
    int4 ma = C.meta[cc.x], mb = C.meta[cc.y];
    Shape a = shape_of(ma.x, C.dims[cc.x]), b = shape_of(mb.x, C.dims[cc.y]);
    TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
    if (far_apart<BIN>(a, ta, b, tb)) {
        K.cnt[s] = 0;
        return;
    }
    Manifold m;
    m.np = 0;
    if (collide_bin<BIN>(a, ta, b, tb, m) && m.np > 0) {
        if (BIN == BIN_CC && flip_cc) m.n = -m.n;
        emit(m, B, K, s, bb);
    } else {
        K.cnt[s] = 0;
    }
}

// Box-box in two launches: the separating-axis test for every pair of the bin (registers
// only, full occupancy), then the full routine, whose clipping buffers live in scratch
// memory, for the overlapping pairs alone (it recomputes the same test, so results are
// unchanged).  Each block of the first launch lists its overlapping pairs in its own
// 128-entry segment; block b of the second launch takes segment b (no atomics).
__global__ void __launch_bounds__(128) k_narrow_bb_sat(const uint32_t *bin_slot, uint32_t off, uint32_t n, Colliders C,
                                                       Bodies B, Contacts K, uint32_t *hits, uint32_t *blk_hits) {
    __shared__ uint32_t wc[2];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    uint32_t s = 0;
    if (t < n) {
        s = bin_slot[off + t];
        const uint2 cc = K.cc[s], bb = K.bb[s];
        const Shape a = shape_of(SH_BOX, C.dims[cc.x]), b = shape_of(SH_BOX, C.dims[cc.y]);
        const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
        hit = !far_apart<BIN_BB>(a, ta, b, tb) && box_box_overlap(a, ta, b, tb);
        if (!hit) K.cnt[s] = 0;
    }
    const uint64_t m = __ballot(hit);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wc[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (hit) hits[blockIdx.x * 128 + (wave ? wc[0] : 0u) + lane_rank(m)] = s;
    if (threadIdx.x == 0) blk_hits[blockIdx.x] = wc[0] + wc[1];
}
__global__ void __launch_bounds__(128) k_narrow_bb_full(const uint32_t *hits, const uint32_t *blk_hits, Colliders C,
                                                        Bodies B, Contacts K) {
    if (threadIdx.x >= blk_hits[blockIdx.x]) return;
    const uint32_t s = hits[blockIdx.x * 128 + threadIdx.x];
    const uint2 cc = K.cc[s], bb = K.bb[s];
    const Shape a = shape_of(SH_BOX, C.dims[cc.x]), b = shape_of(SH_BOX, C.dims[cc.y]);
    const TF ta = collider_tf(C, B, cc.x, bb.x), tb = collider_tf(C, B, cc.y, bb.y);
    Manifold m;
    m.np = 0;
    if (collide_bin<BIN_BB>(a, ta, b, tb, m) && m.np > 0) emit(m, B, K, s, bb);
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
__device__ __forceinline__ void integrate_body(const Bodies &B, uint32_t i, float dt) {
    BodyRow &b = B.r[i];
    float4 ii1 = b.ii1;
    if (ii1.z == 0.0f) return;   // static
    float4 x4 = b.x, q4 = b.q, v4 = b.v, w4 = b.w, f = B.f[i], tq = B.t[i];
    Q q = quat(q4);
    M3 R = mat3_cast(q);
    SM Iw = world_tensor(R, sm_from(B.I0[i], B.I1[i]));
    SM Iwi = world_tensor(R, sm_from(b.ii0, ii1));
    b.px = x4;
    b.pq = q4;
    V3 v = xyz(v4) + xyz(f) * x4.w * dt;
    V3 x = xyz(x4) + v * dt;
    V3 w = xyz(w4);
    w = w + smv(sscale(Iwi, dt), xyz(tq) - cross(w, smv(Iw, w)));
    V3 ws = (0.5f * dt) * w;
    q = qstep(q, ws);
    b.x = f4(x, x4.w);
    b.q = f4q(q);
    b.v = f4(v, 0.0f);
    b.w = f4(w, 0.0f);
}

__global__ void k_integrate(Bodies B, uint32_t N, float dt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) integrate_body(B, i, dt);
}

__device__ __forceinline__ void update_vel_body(const Bodies &B, uint32_t i, float inv_dt) {
    BodyRow &b = B.r[i];
    if (b.ii1.z == 0.0f) return;
    V3 x = xyz(b.x), px = xyz(b.px);
    Q dq = qmul(quat(b.q), conj(quat(b.pq)));
    V3 v = (x - px) * inv_dt;
    V3 w = (2.0f * inv_dt) * V3{dq.x, dq.y, dq.z};
    b.v = f4(v, 0.0f);
    b.w = f4(dq.w >= 0.0f ? w : -w, 0.0f);
}

__global__ void k_update_vel(Bodies B, uint32_t N, float inv_dt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) update_vel_body(B, i, inv_dt);
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
    const BodyRow &row = B.r[i];
    float4 x4 = row.x;
    float4 a = row.ii0, b = row.ii1;
    r.x = xyz(x4);
    r.im = x4.w;
    r.q = quat(row.q);
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

// KNOWN: the point count and bodies come with the call (a caller that staged them), so the
// body loads need not wait for the pair's own rows.
template <bool KNOWN = false>
__device__ __forceinline__ void solve_pos_pair(const Bodies &B, Contacts &K, uint32_t s, int npk = 0, uint2 bbk = {}) {
    const int np = KNOWN ? npk : K.cnt[s];
    if (np == 0) return;
    const float4 *hd = hd_row(K, s);
    const float4 h0 = hd[0], h1 = hd[1];
    const uint2 bb = KNOWN ? bbk : make_uint2(__float_as_uint(h1.x), __float_as_uint(h1.y));
    const V3 n = xyz(h0);
    const float sf = h1.z;
    BodyReg A, Bd;
    load_body(B, bb.x, A);
    load_body(B, bb.y, Bd);
    const BodyRow &ra_ = B.r[bb.x], &rb_ = B.r[bb.y];
    const V3 pxa = xyz(ra_.px), pxb = xyz(rb_.px);
    const Q pqa = quat(ra_.pq), pqb = quat(rb_.pq);
    const V3 va = xyz(ra_.v), vb = xyz(rb_.v);
    const V3 wa = xyz(ra_.w), wb = xyz(rb_.w);
    for (int i = 0; i < np; ++i) {
        float4 *row = pt_row(K, s * 4 + i);
        const float4 r0 = row[0], r1 = row[1];
        const V3 rla = xyz(r0), rlb = xyz(r1);
        float2 lam = make_float2(r0.w, r1.w);
        // Contact.cpp:35-75
        V3 pa = A.x + rotate(A.q, rla);
        V3 pb = Bd.x + rotate(Bd.q, rlb);
        float d = dot(pa - pb, n);
        if (d <= 0.0f) {
            row[2] = f4(pa, 0.0f);
            row[3] = f4(pb, 0.0f);
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
        row[2] = f4(pac, relv);
        row[3] = f4(pbc, 1.0f);
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
        row[0] = f4(rla, lam.x);
        row[1] = f4(rlb, lam.y);
    }
    if (A.dyn) { B.r[bb.x].x = f4(A.x, A.im); B.r[bb.x].q = f4q(A.q); }
    if (Bd.dyn) { B.r[bb.y].x = f4(Bd.x, Bd.im); B.r[bb.y].q = f4q(Bd.q); }
}

struct VelReg {
    V3 x, v, w;
    float im;
    SM Ii, Iwi;
    Q q;
    bool dyn;
};
__device__ __forceinline__ void load_vel(const Bodies &B, uint32_t i, VelReg &r) {
    const BodyRow &row = B.r[i];
    float4 x4 = row.x;
    float4 a = row.ii0, b = row.ii1;
    r.x = xyz(x4);
    r.im = x4.w;
    r.q = quat(row.q);
    r.v = xyz(row.v);
    r.w = xyz(row.w);
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

template <bool KNOWN = false>
__device__ __forceinline__ void solve_vel_pair(const Bodies &B, Contacts &K, uint32_t s, float h, int npk = 0,
                                               uint2 bbk = {}) {
    const int np = KNOWN ? npk : K.cnt[s];
    if (np == 0) return;
    const float4 *hd = hd_row(K, s);
    const float4 h0 = hd[0], h1 = hd[1];
    const uint2 bb = KNOWN ? bbk : make_uint2(__float_as_uint(h1.x), __float_as_uint(h1.y));
    const V3 n = xyz(h0);
    const float rest = h0.w, df = h1.w;
    VelReg A, Bd;
    load_vel(B, bb.x, A);
    load_vel(B, bb.y, Bd);
    bool any = false;
    for (int i = 0; i < np; ++i) {
        const float4 *row = pt_row(K, s * 4 + i);
        const float4 pb4 = row[3];
        if (pb4.w == 0.0f) continue;   // no collision at the position level
        any = true;
        const float4 pa4 = row[2];
        const float relv = pa4.w;
        const float lamn = row[0].w;
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
    if (A.dyn) { B.r[bb.x].v = f4(A.v, 0.0f); B.r[bb.x].w = f4(A.w, 0.0f); }
    if (Bd.dyn) { B.r[bb.y].v = f4(Bd.v, 0.0f); B.r[bb.y].w = f4(Bd.w, 0.0f); }
}

// Position / velocity solve of one colour: its pairs touch disjoint dynamic bodies.
// One colour per launch.  With `pts`, also counts the contact points solved (algorithmic
// bytes of the timed launches): one atomic per wave.
__device__ __forceinline__ void count_points(const Contacts &K, uint32_t s, bool valid, unsigned long long *pts) {
    uint32_t v = valid ? (uint32_t)K.cnt[s] : 0u;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(pts, (unsigned long long)v);
}
__global__ void __launch_bounds__(256) k_solve_pos(Bodies B, Contacts K, const uint32_t *order, uint32_t off, uint32_t count,
                                                   unsigned long long *pts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t < count ? order[off + t] : 0u;
    if (pts) count_points(K, s, t < count, pts);
    if (t < count) solve_pos_pair(B, K, s);
}
__global__ void __launch_bounds__(256) k_solve_vel(Bodies B, Contacts K, const uint32_t *order, uint32_t off, uint32_t count,
                                                   float h, unsigned long long *pts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t < count ? order[off + t] : 0u;
    if (pts) count_points(K, s, t < count, pts);
    if (t < count) solve_vel_pair(B, K, s, h);
}
// A run of small colours [c0, c1) in one workgroup: colours are solved in order with a
// workgroup barrier between them (pairs of one colour share no dynamic body).
__global__ void __launch_bounds__(TAIL_THREADS) k_solve_pos_run(Bodies B, Contacts K, const uint32_t *order,
                                                                  const uint32_t *coff, uint32_t c0, uint32_t c1) {
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t e = coff[c + 1];
        for (uint32_t i = coff[c] + threadIdx.x; i < e; i += blockDim.x) solve_pos_pair(B, K, order[i]);
        __syncthreads();
    }
}
__global__ void __launch_bounds__(TAIL_THREADS) k_solve_vel_run(Bodies B, Contacts K, const uint32_t *order,
                                                                  const uint32_t *coff, uint32_t c0, uint32_t c1, float h) {
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t e = coff[c + 1];
        for (uint32_t i = coff[c] + threadIdx.x; i < e; i += blockDim.x) solve_vel_pair(B, K, order[i], h);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ batches: one workgroup per scene
// Scenes share no bodies, so a batch's colouring and colour-ordered solve can run scene by
// scene, each in one workgroup with workgroup barriers instead of kernel boundaries and host
// round trips.  k_scene_colour runs exactly the rounds of the grid-wide path (rounds 0, 1, ...
// with the same keys, palettes and masks), so every pair gets the same colour; the solve
// kernels walk a scene's colours in order, and pairs of one colour share no dynamic body, so
// the states are bit-identical to the per-colour launches.
constexpr uint32_t SCENE_THREADS = 256;      // k_scene_colour's workgroup (the bucketing assumes it)
constexpr uint32_t SCENE_LDS_PAIRS = 1024;   // active pairs of a scene kept in LDS (more: the global slice)
struct SceneTabs {
    const uint32_t *sap_start, *sap_end;     // the scene's SAP pairs [start, end) (both 0: none)
    const uint32_t *plane_off, *plane_idx;   // its plane pairs: plane list entries plane_idx[plane_off[s] ..)
    const uint32_t *base;                    // exclusive scan of the scenes' pair counts
    uint32_t *work;                          // per scene at base[s]: active pairs (build order)
    uint32_t *list;                          // per scene at base[s]: active pairs by colour
    uint32_t *coff;                          // per scene MAX_COLOURS + 1 offsets into its list
    uint32_t *ncol;                          // per scene: colours used
    uint32_t *err;                           // sticky: 1 colour capacity exceeded, 2 no convergence
    // joints solved inside the per-scene solves (null: the global joint launches run instead):
    // jorder[jbase[s] ..) the scene's joints by joint colour, jcoff[s * (njc + 1) + c] offsets
    const uint32_t *jorder, *jbase, *jcoff;
    uint32_t njc;
    // the scene's bodies: body_idx[body_off[s] ..) in ascending order; body_loc: a body's
    // index in its scene's list
    const uint32_t *body_off, *body_idx, *body_loc;
    const uint32_t *corder;   // the colouring's scene per workgroup (largest scenes first)
};

// Pairs per scene (SAP run + plane pairs), for the slice offsets.
__global__ void k_scene_sizes(SceneTabs T, uint32_t nsc, uint32_t *size) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsc) size[s] = (T.sap_end[s] - T.sap_start[s]) + (T.plane_off[s + 1] - T.plane_off[s]);
}

// The rounds keep a scene's active pairs in LDS when they fit (SCENE_LDS_PAIRS): pair id,
// bodies with their dynamic flags, local id, colour and tentative colour, so that a round
// reads only the bodies' counts, masks and slot claims from global memory (a chain of one or
// two loads in place of six); larger scenes read the same fields from the global arrays.
// Either way the rounds, keys, palettes and masks are those of colour_claim / colour_settle.
constexpr uint32_t SCENE_BDYN = 0x80000000u;   // body word: index | dynamic flag
constexpr uint32_t SCENE_LDYN = 0x8000u;       // local body half-word: local id | dynamic flag
// LDS colouring state: scenes of up to SCENE_LDS_BODIES bodies of which at most
// SCENE_LDS_ACTIVE dynamic ones have active pairs (numbered in scene order for the tables)
constexpr uint32_t SCENE_LDS_BODIES = 512, SCENE_LDS_ACTIVE = 256, SCENE_LDS_UWORDS = 2;
static_assert(SCENE_LDS_PAIRS >= (uint32_t)MAX_COLOURS, "the histogram lives in the hash-input array");
static_assert(SCENE_LDS_BODIES < SCENE_LDYN, "local ids fit in 15 bits");
static_assert(SCENE_LDS_BODIES == 2 * SCENE_THREADS, "the body numbering takes two bodies per thread");
static_assert(SCENE_LDS_ACTIVE * COLOUR_SLOTS * 4 >= SCENE_LDS_PAIRS * 8, "the claim table overlays the body pairs");
__global__ void __launch_bounds__(SCENE_THREADS) k_scene_colour(const int *cnt, const uint2 *pbod, const uint32_t *n_sap_dev,
                                                                Bodies B, ColourBufs cb, SceneTabs T, uint32_t cap_rounds,
                                                                uint32_t n_int, float dt) {
    __shared__ uint32_t L[SCENE_LDS_PAIRS];    // pair id
    __shared__ uint32_t LQ[SCENE_LDS_PAIRS];   // local pair id (the hashes' input); the histogram after the rounds
    __shared__ int LA[SCENE_LDS_PAIRS];        // colour, COLOUR_OPEN
    __shared__ uint16_t LT[SCENE_LDS_PAIRS];   // tentative colour of the round
    __shared__ uint32_t LL[SCENE_LDS_PAIRS];   // bodies' scene-local, then active ids | SCENE_LDYN, a in the low half
    // the pairs' bodies | SCENE_BDYN; with the LDS state, once its bodies are numbered, the
    // rounds' slot claims per active body (the keys' low words) in the same place
    __shared__ union {
        uint2 b[SCENE_LDS_PAIRS];
        uint32_t c[SCENE_LDS_ACTIVE * COLOUR_SLOTS];
    } LBC;
    uint2 *const LB = LBC.b;
    uint32_t *const LC = LBC.c;
    // per active body: uncoloured counts, colours 0 .. 32 SCENE_LDS_UWORDS - 1 of the used
    // masks (higher colours stay global), body index; per scene body: its active id
    __shared__ uint32_t LU[SCENE_LDS_ACTIVE];
    __shared__ uint32_t LW[SCENE_LDS_ACTIVE * SCENE_LDS_UWORDS];
    __shared__ uint32_t LG[SCENE_LDS_ACTIVE];
    __shared__ uint16_t LM[SCENE_LDS_BODIES];
    uint32_t *const hist = LQ;
    __shared__ uint32_t sh[SCENE_THREADS / 64 + 1];
    __shared__ uint32_t nact, won, top, serr;
    uint32_t nbs_active = 0u;
    const uint32_t s = T.corder ? T.corder[blockIdx.x] : blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const uint32_t ns = *n_sap_dev;
    const uint32_t a0 = T.sap_start[s], na = T.sap_end[s] - a0;
    const uint32_t q0 = T.plane_off[s], total = na + (T.plane_off[s + 1] - q0);
    uint32_t *G = T.work + T.base[s];
    const uint32_t nbs = T.body_loc ? T.body_off[s + 1] - T.body_off[s] : 0u;
    const bool lfit = T.body_loc && nbs <= SCENE_LDS_BODIES;   // bodies numbered in LDS
    if (tid == 0) { nact = 0u; won = 0u; top = 0u; serr = 0u; }
    for (uint32_t i = tid; lfit && i < nbs; i += bd) LM[i] = 0u;
    __syncthreads();
    // activation, in two parts.  (1) The scan: the point counts of ACT_AHEAD pairs per thread
    // at once, the pairs without points marked, the others listed (no atomics in this loop: a
    // global atomic between loads made every step wait a round trip).  (2) The listed pairs:
    // their bodies, dynamic flags and uncoloured counts.
    constexpr uint32_t ACT_AHEAD = 8;
    for (uint32_t c0 = 0; c0 < total; c0 += ACT_AHEAD * bd) {
        uint32_t pid[ACT_AHEAD];
        bool act[ACT_AHEAD];
#pragma unroll
        for (uint32_t u = 0; u < ACT_AHEAD; ++u) {
            const uint32_t j = c0 + u * bd + tid;
            pid[u] = j < total ? (j < na ? a0 + j : ns + T.plane_idx[q0 + j - na]) : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < ACT_AHEAD; ++u) act[u] = c0 + u * bd + tid < total && cnt[pid[u]] > 0;
#pragma unroll
        for (uint32_t u = 0; u < ACT_AHEAD; ++u) {
            if (c0 + u * bd >= total) break;   // uniform
            const uint32_t j = c0 + u * bd + tid;
            const bool a = act[u];
            if (j < total && !a) cb.acol[pid[u]] = COLOUR_NONE;
            const uint64_t m = __ballot(a);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                uint32_t b0 = 0u;
                if ((int)(tid & 63) == leader) b0 = atomicAdd(&nact, (uint32_t)__popcll(m));
                b0 = __shfl(b0, leader, 64);
                if (a) {
                    const uint32_t pos = b0 + lane_rank(m);
                    G[pos] = pid[u];
                    if (pos < SCENE_LDS_PAIRS) L[pos] = pid[u];
                }
            }
        }
    }
    __syncthreads();
    const uint32_t n = nact;
    const bool in_lds = n <= SCENE_LDS_PAIRS;
    // the bodies' colouring state in LDS when the pairs are there and the scene's active
    // dynamic bodies fit: mark them (scene-local ids), number them in scene order, then count
    const bool lmark = in_lds && lfit;
    for (uint32_t j = tid; j < n; j += bd) {
        const uint32_t p = j < SCENE_LDS_PAIRS ? L[j] : G[j];
        const uint2 bb = pbod[p];
        const uint32_t q = cb.lid[p];
        const bool da = is_dynamic(B, bb.x), db = is_dynamic(B, bb.y);
        cb.acol[p] = COLOUR_OPEN;
        if (lmark) {   // every body of a scene's pair is in the scene
            const uint32_t la = T.body_loc[bb.x], lb = T.body_loc[bb.y];
            if (da) LM[la] = 1u;
            if (db) LM[lb] = 1u;
            LL[j] = (la | (da ? SCENE_LDYN : 0u)) | ((lb | (db ? SCENE_LDYN : 0u)) << 16);
        } else {
            if (da) atomicAdd(&cb.ucnt[bb.x], 1u);
            if (db) atomicAdd(&cb.ucnt[bb.y], 1u);
        }
        if (j < SCENE_LDS_PAIRS) {
            LB[j] = make_uint2(bb.x | (da ? SCENE_BDYN : 0u), bb.y | (db ? SCENE_BDYN : 0u));
            LQ[j] = q;
            LA[j] = COLOUR_OPEN;
        }
    }
    __syncthreads();
    bool lbod = false;
    if (lmark) {   // uniform
        const uint32_t i0 = 2u * tid;
        const uint32_t m0 = i0 < nbs ? LM[i0] : 0u, m1 = i0 + 1u < nbs ? LM[i0 + 1u] : 0u;
        uint32_t nact_b;
        const uint32_t base = block_exclusive_scan(m0 + m1, sh, &nact_b);
        lbod = nact_b <= SCENE_LDS_ACTIVE;
        if (i0 < nbs) LM[i0] = (uint16_t)base;
        if (i0 + 1u < nbs) LM[i0 + 1u] = (uint16_t)(base + m0);
        for (uint32_t i = tid; lbod && i < nact_b; i += bd) LU[i] = 0u;
        for (uint32_t i = tid; lbod && i < nact_b * SCENE_LDS_UWORDS; i += bd) LW[i] = 0u;
        __syncthreads();
        for (uint32_t j = tid; j < n; j += bd) {
            const uint32_t lw = LL[j];
            const bool da = lw & SCENE_LDYN, db = (lw >> 16) & SCENE_LDYN;
            const uint2 bw = LB[j];
            if (lbod) {   // active ids, their bodies and uncoloured counts
                const uint32_t ca = da ? LM[lw & (SCENE_LDYN - 1u)] : 0u;
                const uint32_t cc = db ? LM[(lw >> 16) & (SCENE_LDYN - 1u)] : 0u;
                LL[j] = (ca | (da ? SCENE_LDYN : 0u)) | ((cc | (db ? SCENE_LDYN : 0u)) << 16);
                if (da) { LG[ca] = bw.x & ~SCENE_BDYN; atomicAdd(&LU[ca], 1u); }
                if (db) { LG[cc] = bw.y & ~SCENE_BDYN; atomicAdd(&LU[cc], 1u); }
            } else {   // too many active bodies: the global state
                if (da) atomicAdd(&cb.ucnt[bw.x & ~SCENE_BDYN], 1u);
                if (db) atomicAdd(&cb.ucnt[bw.y & ~SCENE_BDYN], 1u);
            }
        }
        __syncthreads();
        if (lbod) {   // the claim table over the body pairs (read above, not after)
            uint4 *c4 = reinterpret_cast<uint4 *>(LC);
            for (uint32_t i = tid; i < nact_b * (COLOUR_SLOTS / 4); i += bd) c4[i] = make_uint4(0u, 0u, 0u, 0u);
            nbs_active = nact_b;
        }
        __syncthreads();
    }
    const uint32_t ep = cb.epoch ? *cb.epoch : 0u;   // 0: the claims were cleared this substep
    // rounds 0, 1, ... as the grid-wide path runs them
    uint32_t left = n, r = 0;
    for (; left > 0 && r < cap_rounds; ++r) {
        const uint32_t kh = (ep << COLOUR_EPOCH_SHIFT) | (r + 1u);
        if (lbod) {
            for (uint32_t j = tid; j < n; j += bd) {   // claim (colour_claim on the LDS state)
                if (LA[j] != COLOUR_OPEN) continue;
                const uint32_t lw = LL[j];
                const bool da = lw & SCENE_LDYN, db = (lw >> 16) & SCENE_LDYN;
                const uint32_t la = lw & (SCENE_LDYN - 1u), lb = (lw >> 16) & (SCENE_LDYN - 1u);
                const uint32_t k = (da ? LU[la] : 0u) + (db ? LU[lb] : 0u);
                const uint32_t q = LQ[j];
                uint32_t idx = tent_hash(q, r) % palette(k, cb.pal_max);
                uint32_t t = MAX_COLOURS;
                for (uint32_t w = 0; w < COLOUR_WORDS; ++w) {
                    uint32_t ua = 0u, ub = 0u;
                    if (w < SCENE_LDS_UWORDS) {
                        ua = da ? LW[la * SCENE_LDS_UWORDS + w] : 0u;
                        ub = db ? LW[lb * SCENE_LDS_UWORDS + w] : 0u;
                    } else {
                        ua = da ? ld_shared(&cb.used[LG[la] * COLOUR_WORDS + w]) : 0u;
                        ub = db ? ld_shared(&cb.used[LG[lb] * COLOUR_WORDS + w]) : 0u;
                    }
                    uint32_t fr = ~(ua | ub);
                    const uint32_t nf = (uint32_t)__popc(fr);
                    if (idx < nf) {
                        for (; idx; --idx) fr &= fr - 1u;
                        t = w * 32u + (uint32_t)__builtin_ctz(fr);
                        break;
                    }
                    idx -= nf;
                }
                if (t >= (uint32_t)MAX_COLOURS) { atomicOr(&serr, 1u); t = MAX_COLOURS - 1; }
                LT[j] = (uint16_t)t;
                // only this round's claims are in the table, so the keys' low words decide
                const uint32_t lo = rank_hash(q, r);
                if (da) atomicMax(&LC[la * COLOUR_SLOTS + t % COLOUR_SLOTS], lo);
                if (db) atomicMax(&LC[lb * COLOUR_SLOTS + t % COLOUR_SLOTS], lo);
            }
        } else if (in_lds) {
            for (uint32_t j = tid; j < n; j += bd) {   // claim (colour_claim)
                if (LA[j] != COLOUR_OPEN) continue;
                const uint2 bw = LB[j];
                const bool da = bw.x & SCENE_BDYN, db = bw.y & SCENE_BDYN;
                const uint32_t ba = bw.x & ~SCENE_BDYN, bb = bw.y & ~SCENE_BDYN;
                const uint32_t k = (da ? ld_shared(&cb.ucnt[ba]) : 0u) + (db ? ld_shared(&cb.ucnt[bb]) : 0u);
                const uint32_t q = LQ[j];
                uint32_t idx = tent_hash(q, r) % palette(k, cb.pal_max);
                uint32_t t = MAX_COLOURS;
                for (uint32_t w = 0; w < COLOUR_WORDS; ++w) {
                    uint32_t fr = ~((da ? ld_shared(&cb.used[ba * COLOUR_WORDS + w]) : 0u) |
                                    (db ? ld_shared(&cb.used[bb * COLOUR_WORDS + w]) : 0u));
                    const uint32_t nf = (uint32_t)__popc(fr);
                    if (idx < nf) {
                        for (; idx; --idx) fr &= fr - 1u;
                        t = w * 32u + (uint32_t)__builtin_ctz(fr);
                        break;
                    }
                    idx -= nf;
                }
                if (t >= (uint32_t)MAX_COLOURS) { atomicOr(&serr, 1u); t = MAX_COLOURS - 1; }
                LT[j] = (uint16_t)t;
                const unsigned long long key = claim_key(kh, rank_hash(q, r));
                if (da) atomicMax(&cb.claim[ba * COLOUR_SLOTS + t % COLOUR_SLOTS], key);
                if (db) atomicMax(&cb.claim[bb * COLOUR_SLOTS + t % COLOUR_SLOTS], key);
            }
        } else {
            for (uint32_t j = tid; j < n; j += bd) {
                const uint32_t p = G[j];
                if (cb.acol[p] == COLOUR_OPEN) colour_claim(cb, pbod, B, p, r, kh);
            }
        }
        __syncthreads();
        uint32_t w = 0;
        if (lbod) {
            for (uint32_t j = tid; j < n; j += bd) {   // settle (colour_settle on the LDS state)
                if (LA[j] != COLOUR_OPEN) continue;
                const uint32_t lw = LL[j];
                const bool da = lw & SCENE_LDYN, db = (lw >> 16) & SCENE_LDYN;
                const uint32_t la = lw & (SCENE_LDYN - 1u), lb = (lw >> 16) & (SCENE_LDYN - 1u);
                const uint32_t t = LT[j];
                const uint32_t lo = rank_hash(LQ[j], r);
                if (da && LC[la * COLOUR_SLOTS + t % COLOUR_SLOTS] != lo) continue;
                if (db && LC[lb * COLOUR_SLOTS + t % COLOUR_SLOTS] != lo) continue;
                LA[j] = (int)t;
                const uint32_t bit = 1u << (t % 32u), tw = t / 32u;
                if (tw < SCENE_LDS_UWORDS) {
                    if (da) atomicOr(&LW[la * SCENE_LDS_UWORDS + tw], bit);
                    if (db) atomicOr(&LW[lb * SCENE_LDS_UWORDS + tw], bit);
                } else {
                    if (da) atomicOr(&cb.used[LG[la] * COLOUR_WORDS + tw], bit);
                    if (db) atomicOr(&cb.used[LG[lb] * COLOUR_WORDS + tw], bit);
                }
                if (da) atomicSub(&LU[la], 1u);
                if (db) atomicSub(&LU[lb], 1u);
                ++w;
            }
        } else if (in_lds) {
            for (uint32_t j = tid; j < n; j += bd) {   // settle (colour_settle)
                if (LA[j] != COLOUR_OPEN) continue;
                const uint2 bw = LB[j];
                const bool da = bw.x & SCENE_BDYN, db = bw.y & SCENE_BDYN;
                const uint32_t ba = bw.x & ~SCENE_BDYN, bb = bw.y & ~SCENE_BDYN;
                const uint32_t t = LT[j];
                const unsigned long long key = claim_key(kh, rank_hash(LQ[j], r));
                if (da && ld_shared(&cb.claim[ba * COLOUR_SLOTS + t % COLOUR_SLOTS]) != key) continue;
                if (db && ld_shared(&cb.claim[bb * COLOUR_SLOTS + t % COLOUR_SLOTS]) != key) continue;
                LA[j] = (int)t;
                const uint32_t bit = 1u << (t % 32u);
                if (da) { atomicOr(&cb.used[ba * COLOUR_WORDS + t / 32u], bit); atomicSub(&cb.ucnt[ba], 1u); }
                if (db) { atomicOr(&cb.used[bb * COLOUR_WORDS + t / 32u], bit); atomicSub(&cb.ucnt[bb], 1u); }
                ++w;
            }
        } else {
            for (uint32_t j = tid; j < n; j += bd) {
                const uint32_t p = G[j];
                if (cb.acol[p] == COLOUR_OPEN && colour_settle(cb, pbod, B, p, r, kh)) ++w;
            }
        }
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
        if ((tid & 63) == 0 && w) atomicAdd(&won, w);
        __syncthreads();
        left = n - won;
        if (lbod) {   // the next round's claim table
            uint4 *c4 = reinterpret_cast<uint4 *>(LC);
            for (uint32_t i = tid; i < nbs_active * (COLOUR_SLOTS / 4); i += bd) c4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
    }
    if (in_lds)   // the colours to the global array (pair colours are read from it)
        for (uint32_t j = tid; j < n; j += bd) cb.acol[L[j]] = LA[j];
    auto colour_of = [&](uint32_t j) { return in_lds ? LA[j] : cb.acol[G[j]]; };
    // bucket by colour: sizes, offsets (four colours per thread), scatter
    for (uint32_t c = tid; c < (uint32_t)MAX_COLOURS; c += bd) hist[c] = 0u;
    __syncthreads();
    for (uint32_t j = tid; j < n; j += bd) {
        const int c = colour_of(j);
        if (c >= 0) atomicAdd(&hist[c], 1u);
    }
    __syncthreads();
    constexpr uint32_t PER = MAX_COLOURS / SCENE_THREADS;
    uint32_t v[PER], sum = 0u;
    for (uint32_t k = 0; k < PER; ++k) {
        v[k] = hist[tid * PER + k];
        sum += v[k];
        if (v[k]) atomicMax(&top, tid * PER + k + 1u);
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(sum, sh, &tot);
    const uint32_t tp = top;   // final: the scan ends with a barrier
    // offsets of colours 0 .. top (the solve reads no others; colours from top on are empty, so
    // the offset at top is the total)
    uint32_t *co = T.coff + (size_t)s * (MAX_COLOURS + 1);
    for (uint32_t k = 0; k < PER; ++k) {
        if (tid * PER + k <= tp) co[tid * PER + k] = run;
        hist[tid * PER + k] = run;   // scatter cursor
        run += v[k];
    }
    if (tid == 0 && tp == (uint32_t)MAX_COLOURS) co[MAX_COLOURS] = tot;
    __syncthreads();
    uint32_t *out = T.list + T.base[s];
    for (uint32_t j = tid; j < n; j += bd) {
        const int c = colour_of(j);
        if (c >= 0) out[atomicAdd(&hist[c], 1u)] = in_lds ? L[j] : G[j];
    }
    if (ep) {   // what this scene set: its bodies' used-mask words of its colours and their counts
        const uint32_t nw = (top + 31u) / 32u;   // final since the barrier before the offsets
        for (uint32_t i = tid; i < nbs_active && lbod; i += bd)   // LDS state: only the global words
            for (uint32_t w = SCENE_LDS_UWORDS; w < nw; ++w) cb.used[LG[i] * COLOUR_WORDS + w] = 0u;
        for (uint32_t j = tid; j < n && !lbod; j += bd) {
            uint32_t ba, bb;
            bool da, db;
            if (in_lds) {
                const uint2 bw = LB[j];
                da = bw.x & SCENE_BDYN; db = bw.y & SCENE_BDYN;
                ba = bw.x & ~SCENE_BDYN; bb = bw.y & ~SCENE_BDYN;
            } else {
                const uint2 b2 = pbod[G[j]];
                ba = b2.x; bb = b2.y;
                da = is_dynamic(B, ba); db = is_dynamic(B, bb);
            }
            for (uint32_t w = 0; w < nw; ++w) {
                if (da) cb.used[ba * COLOUR_WORDS + w] = 0u;
                if (db) cb.used[bb * COLOUR_WORDS + w] = 0u;
            }
            if (da) cb.ucnt[ba] = 0u;
            if (db) cb.ucnt[bb] = 0u;
        }
    }
    if (tid == 0) {
        T.ncol[s] = top;
        // colour capacity: this scene's LDS flag (the global word only where colour_claim ran)
        uint32_t e = serr | (in_lds ? 0u : ld_shared(&cb.ctr[CTR_ERR]) & 1u);
        if (left) e |= 2u;
        if (e) atomicOr(T.err, e);
    }
    // the substep's integration (k_integrate) of n_int bodies, spread over the workgroups: the
    // colouring reads only the bodies' dynamic flags, which the integration does not write
    for (uint32_t i = s * bd + tid; i < n_int; i += gridDim.x * bd) integrate_body(B, i, dt);
    if (tid == 0 && n) {   // statistics (no wait on these atomics)
        atomicAdd(&cb.ctr[CTR_ACT], n);
        atomicMax(&cb.ctr[CTR_NCOL], top);
        atomicMax(&cb.ctr[CTR_ROUNDS], r);
    }
}

}  // namespace rb
