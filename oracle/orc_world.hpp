// Oracle: CPU restatement of rbps::World::step (broadphase, narrowphase, contact
// emission, XPBD position/velocity solve), templated on the scalar type.
//
// TEST INFRASTRUCTURE ONLY (see orc_math.hpp).
//
// Two solve orders:
//   ORDER_REFERENCE — the reference's sequential Gauss-Seidel in contact emission
//                     order (pairs in SAP active-set order, then the plane loop).
//                     With T=double this is the reference step bit-for-bit.
//   ORDER_COLOUR    — the order the GPU pipeline uses: the frame's pairs in
//                     canonical sweep order; every substep the pairs that have
//                     contact points are coloured by synchronous randomized trials
//                     (colour_active) and solved colour by colour; a pair's points
//                     are solved in sequence.
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>
#include "orc_collide.hpp"
#include "orc_trig.hpp"

namespace orc {

enum { ORDER_REFERENCE = 0, ORDER_COLOUR = 1 };
static constexpr uint32_t BP_INVALID = 0xffffffffu;

// lowbias32 integer mix (public-domain hash by C. Wellons); JP priorities.
inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <class T> struct Contact {
    uint32_t ba, bb, ca, cb;
    V3<T> ra, rb, n;            // body-local lever arms, world normal A->B
    T rest, sf, df;
    T lam_n, lam_t;
    V3<T> pa, pb;               // world points (rebuilt in the position solve)
    T depth;
    bool collision;
    T relv;
};

template <class T> struct World {
    T timestep = T(1.0 / 60.0);
    int substeps = 20;
    int order = ORDER_REFERENCE;

    // ---- bodies (Body.hpp:52-69)
    std::vector<int> btype;   // 0 STATIC, 1 DYNAMIC
    std::vector<uint32_t> bscene;   // independent scene of a batch (0 for a single world)
    std::vector<T> mass, inv_mass;
    std::vector<V3<T>> force, torque, pos, lv, av, ppos;
    std::vector<Q<T>> q, pq;
    std::vector<SM<T>> I, Iinv, Iw, Iwinv;

    // ---- colliders (Collider.hpp:43-63)
    std::vector<ShapeP<T>> shape;
    std::vector<V3<T>> lpos;
    std::vector<Q<T>> lrot;
    std::vector<uint32_t> cbody, bph;
    std::vector<char> cstatic;
    std::vector<T> rest, sfr, dfr;

    // ---- broadphase (BroadPhase.hpp:78-125)
    struct Obj { AABB<T> fat, tight; uint32_t uid; bool alive; uint32_t scene; };
    bool batched = false;   // some body has scene > 0
    struct EP { float v; uint32_t slot; uint32_t is_max; };
    std::vector<Obj> objs;
    std::vector<EP> eps;
    std::vector<uint32_t> ep_at[2];   // slot -> endpoint index (min, max); replaces the O(n) scan
    bool dirty = false;
    T fat_margin = T(0.1), reinsert = T(0.05);
    std::vector<std::pair<uint32_t, uint32_t>> sap_pairs;   // user ids, (lo, hi)

    // ---- contacts
    std::vector<Contact<T>> contacts;             // ORDER_REFERENCE: emission order
    struct SPair { uint32_t ca, cb, ba, bb; };
    std::vector<SPair> spairs;                     // ORDER_COLOUR: frame pair list
    std::vector<int> colour;
    int ncolours = 0;
    uint32_t colour_rounds = 0;
    std::vector<int> pcount;                       // points per pair this substep
    std::vector<Contact<T>> pcont;                 // 4 slots per pair

    // ---- joints (Joint.hpp:82-106) with their constraint rows (Constraint.hpp:31-53)
    enum { J_PRISMATIC = 0, J_REVOLUTE = 1, J_FIXED = 2, J_SPHERICAL = 3, J_ROW = 4 };
    enum { ACT_FREE = 0, ACT_POSITION = 1, ACT_SPEED = 2 };
    struct Joint {
        int type = 0, act = 0;
        uint32_t b1 = 0, b2 = 0;
        V3<T> r1, r2, axis, laxis;
        bool limited = false;
        T lower = 0, upper = 0, target = 0, speed = 0, current = 0, damping = 0;
        int nrows = 0;
        bool rot[4] = {false, false, false, false};   // ROTATIONAL row (else POSITIONAL)
        T comp[4] = {0, 0, 0, 0};                       // row compliance
        V3<T> dir[4];
        T mag[4] = {0, 0, 0, 0}, lam[4] = {0, 0, 0, 0};
    };
    std::vector<Joint> joints;
    std::vector<int> jcolour;
    std::vector<uint32_t> joint_order;   // colour mode: joints by colour (stable)
    int njcolours = 0;
    bool joints_dirty = false;
    // Joint solver mode.
    // 0 (the reference): errors of every joint evaluated once per substep before any row is
    //   solved (World.hpp:168-170); lambda accumulates over the whole run (never reset);
    //   row direction = normalize(v) (Constraint.cpp:5-12), which is zero for errors in
    //   (1e-9, ~3.2e-5) although the row is kept, so a ROTATIONAL row with compliance 0
    //   then divides 0 by 0; the revolute alignment row is a1 x a2 (Joint.cpp:101), whose
    //   correction turns the axes apart.  Chains diverge and hinges turn NaN.
    // 1 (XPBD, Mueller et al. 2020): each row's error evaluated from the current state just
    //   before the row is solved, lambda reset every substep, direction v/|v| whenever
    //   |v| > EPS, revolute alignment a2 x a1.  Equal to mode 0 in everything else.
    int joint_mode = 0;
    // 1: turn the capsule-capsule normal around (the reference's routine returns it from B to
    // A, CapsuleCapsule.hpp:68, every other routine from A to B)
    int contact_mode = 0;
    // EPA guard (gjk_epa_manifold), colour order only: 1 the pair's contact normal of the
    // previous substep of the same frame; 2 that and the face normals of both shapes
    int epa_guard = 0;
    uint64_t guard_hits = 0;   // EPA results the guard replaced, and the largest depth it cut
    double guard_max_cut = 0.0;
    bool narrow(const ShapeP<T> &a, const TF<T> &ta, const ShapeP<T> &b, const TF<T> &tb, Manifold<T> &m,
                const EpaGuard<T> *guard = nullptr, T *guard_cut = nullptr) const {
        if (!dispatch(a, ta, b, tb, m, guard, guard_cut)) return false;
        if (contact_mode == 1 && a.kind == SH_CAPSULE && b.kind == SH_CAPSULE) m.normal = -m.normal;
        return true;
    }

    // ------------------------------------------------------------ creation
    // BodyAPI.hpp:53-80
    uint32_t add_body(int type, T m, const SM<T> &inertia, V3<T> p, Q<T> o, V3<T> v, V3<T> w, uint32_t scene = 0) {
        bscene.push_back(scene);
        if (scene) batched = true;
        uint32_t id = (uint32_t)btype.size();
        T im = (type == 0) ? T(0) : T(1) / m;
        SM<T> Ii = (type == 0) ? SM<T>() : inertia.inverse();
        btype.push_back(type);
        force.push_back(V3<T>(T(0), T(-9.8f) * m, T(0)));
        torque.push_back(V3<T>(T(0)));
        mass.push_back(m);
        inv_mass.push_back(im);
        pos.push_back(p); q.push_back(o); lv.push_back(v); av.push_back(w);
        ppos.push_back(p); pq.push_back(o);
        I.push_back(inertia); Iinv.push_back(Ii); Iw.push_back(inertia); Iwinv.push_back(Ii);
        return id;
    }
    TF<T> collider_tf(uint32_t c) const {   // Collider.hpp:73-81
        uint32_t b = cbody[c];
        return {pos[b] + rotate(q[b], lpos[c]), q[b] * lrot[c]};
    }
    // ColliderAPI.hpp:61-91 + BroadPhase.cpp:169-197
    uint32_t add_collider(const ShapeP<T> &s, uint32_t body, bool is_static, V3<T> lp, Q<T> lr, T e, T fs, T fd) {
        uint32_t id = (uint32_t)shape.size();
        shape.push_back(s); lpos.push_back(lp); lrot.push_back(lr); cbody.push_back(body);
        cstatic.push_back(is_static ? 1 : 0); rest.push_back(e); sfr.push_back(fs); dfr.push_back(fd);
        uint32_t h = BP_INVALID;
        if (s.kind != SH_PLANE) {
            TF<T> tf = collider_tf(id);
            AABB<T> tight = shape_aabb(s, tf);
            h = (uint32_t)objs.size();
            Obj o;
            o.tight = tight;
            V3<T> m(fat_margin);
            o.fat = {tight.mn - m, tight.mx + m};
            o.uid = id;
            o.scene = bscene[body];
            o.alive = true;
            objs.push_back(o);
            ep_at[0].push_back((uint32_t)eps.size());
            eps.push_back({(float)o.fat.mn.x, h, 0});
            ep_at[1].push_back((uint32_t)eps.size());
            eps.push_back({(float)o.fat.mx.x, h, 1});
            dirty = true;
        }
        bph.push_back(h);
        return id;
    }

    // JointAPI.hpp:160-265 (rows per type); the spherical joint is this build's extension:
    // one POSITIONAL attachment row carrying the compliance, plus a ROTATIONAL swing-limit
    // row when limited.
    uint32_t add_joint(const Joint &in, T compliance) {
        Joint j = in;
        for (int k = 0; k < 4; ++k) { j.rot[k] = false; j.comp[k] = T(0); j.dir[k] = V3<T>(T(1), T(0), T(0)); }
        switch (j.type) {
            case J_PRISMATIC: j.nrows = 3; j.rot[0] = true; j.comp[2] = compliance; j.laxis = V3<T>(); break;
            case J_REVOLUTE: j.nrows = 4; j.rot[0] = j.rot[2] = j.rot[3] = true; j.comp[3] = compliance; break;
            case J_FIXED:
                j.nrows = 2; j.rot[0] = true; j.act = ACT_FREE; j.limited = false; j.lower = j.upper = T(0);
                j.axis = j.laxis = V3<T>(); j.damping = T(0);
                break;
            case J_ROW:   // a constraint row (ConstraintAPI.hpp:37-64): act = its type, direction (1,0,0), magnitude 0
                j.nrows = 1; j.rot[0] = j.act == 1; j.comp[0] = compliance; j.damping = T(0);
                break;
            default:
                j.nrows = j.limited ? 2 : 1; j.rot[1] = true; j.comp[0] = compliance; break;
        }
        j.current = T(0);   // target and speed: the reference's public fields, 0 unless set
        joints.push_back(j);
        joints_dirty = true;
        return (uint32_t)joints.size() - 1;
    }

    // ------------------------------------------------------------ broadphase
    // BroadPhase.cpp:211-232 with the swept box of BroadPhase.hpp:228-249
    void bp_move(uint32_t h, const AABB<T> &nt) {
        Obj &o = objs[h];
        o.tight = nt;
        const AABB<T> &f = o.fat;
        if (nt.mn.x >= f.mn.x && nt.mx.x <= f.mx.x && nt.mn.y >= f.mn.y && nt.mx.y <= f.mx.y &&
            nt.mn.z >= f.mn.z && nt.mx.z <= f.mx.z)
            return;
        V3<T> m(fat_margin + reinsert);
        o.fat = {nt.mn - m, nt.mx + m};
        eps[ep_at[0][h]].v = (float)o.fat.mn.x;
        eps[ep_at[1][h]].v = (float)o.fat.mx.x;
        dirty = true;
    }
    // CollisionPipeline.cpp:14-53
    void refit(T dt) {
        for (uint32_t i = 0; i < shape.size(); ++i) {
            if (cstatic[i]) continue;
            if (bph[i] == BP_INVALID) continue;
            uint32_t b = cbody[i];
            AABB<T> tight = shape_aabb(shape[i], collider_tf(i));
            V3<T> dv = lv[b] * dt;
            V3<T> emn(fat_margin + (dv.x < 0 ? -dv.x : T(0)), fat_margin + (dv.y < 0 ? -dv.y : T(0)),
                      fat_margin + (dv.z < 0 ? -dv.z : T(0)));
            V3<T> emx(fat_margin + (dv.x > 0 ? dv.x : T(0)), fat_margin + (dv.y > 0 ? dv.y : T(0)),
                      fat_margin + (dv.z > 0 ? dv.z : T(0)));
            bp_move(bph[i], AABB<T>{tight.mn - emn, tight.mx + emx});
        }
    }
    // BroadPhase.cpp:80-94 (a stable sort on the FP32 key; equal keys keep their previous order)
    // Batches sort by (scene, value): a scene's endpoints are contiguous, so the canonical
    // sweep (min endpoint to its own max) never leaves the scene.  Ties keep the previous
    // frame's order in both keys.
    void bp_sort() {
        if (!batched) {
            std::stable_sort(eps.begin(), eps.end(), [](const EP &a, const EP &b) { return a.v < b.v; });
        } else {
            std::stable_sort(eps.begin(), eps.end(), [&](const EP &a, const EP &b) {
                const uint32_t sa = objs[a.slot].scene, sb = objs[b.slot].scene;
                return sa < sb || (sa == sb && a.v < b.v);
            });
        }
        for (uint32_t k = 0; k < eps.size(); ++k) ep_at[eps[k].is_max][eps[k].slot] = k;
    }
    static std::pair<uint32_t, uint32_t> ordered(uint32_t a, uint32_t b) {
        return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    }
    // BroadPhase.cpp:104-152: active-set sweep, pairs in the reference's emission order.
    void sweep_reference() {
        sap_pairs.clear();
        std::vector<uint32_t> active;
        for (const EP &e : eps) {
            const Obj &o = objs[e.slot];
            if (!o.alive) continue;
            if (!e.is_max) {
                for (uint32_t s : active) {
                    const Obj &ot = objs[s];
                    if (!ot.alive) continue;
                    if (!overlap(o.fat, ot.fat)) continue;
                    if (o.scene == ot.scene) sap_pairs.push_back(ordered(o.uid, ot.uid));
                }
                active.push_back(e.slot);
            } else {
                auto it = std::find(active.begin(), active.end(), e.slot);
                if (it != active.end()) active.erase(it);
            }
        }
    }
    // Same pair set, canonical order: each min endpoint scans forward to its own max.
    void sweep_canonical() {
        sap_pairs.clear();
        for (uint32_t i = 0; i < eps.size(); ++i) {
            const EP &e = eps[i];
            if (e.is_max || !objs[e.slot].alive) continue;
            const Obj &a = objs[e.slot];
            for (uint32_t j = i + 1; j < eps.size(); ++j) {
                const EP &f = eps[j];
                if (f.slot == e.slot) break;   // own max endpoint
                if (f.is_max || !objs[f.slot].alive) continue;
                if (overlap(a.fat, objs[f.slot].fat)) sap_pairs.push_back(ordered(a.uid, objs[f.slot].uid));
            }
        }
    }
    void bp_update() {   // BroadPhase.cpp:234-245
        if (dirty) { bp_sort(); dirty = false; }
        if (order == ORDER_REFERENCE) sweep_reference(); else sweep_canonical();
    }

    // ------------------------------------------------------------ narrowphase + emission
    // CollisionPipeline.cpp:59-104
    void make_contacts(const Manifold<T> &m, uint32_t ba, uint32_t bb, uint32_t ca, uint32_t cb, Contact<T> *out) const {
        for (int i = 0; i < m.n; ++i) {
            Contact<T> &k = out[i];
            T half = m.depth[i] * T(0.5);
            V3<T> pa = m.pos[i] + m.normal * half;
            V3<T> pb = m.pos[i] - m.normal * half;
            k.ba = ba; k.bb = bb; k.ca = ca; k.cb = cb;
            k.ra = rotate(conj(q[ba]), pa - pos[ba]);
            k.rb = rotate(conj(q[bb]), pb - pos[bb]);
            k.n = m.normal;
            k.rest = smin(rest[ca], rest[cb]);
            k.sf = (sfr[ca] + sfr[cb]) * T(0.5);
            k.df = (dfr[ca] + dfr[cb]) * T(0.5);
            k.lam_n = T(0); k.lam_t = T(0);
            k.pa = pa; k.pb = pb;
            k.depth = m.depth[i];
            k.collision = false;
            k.relv = T(0);
        }
    }
    bool pair_ok(uint32_t ca, uint32_t cb) const {   // CollisionPipeline.cpp:121-133
        if (cstatic[ca] && cstatic[cb]) return false;
        return cbody[ca] != cbody[cb];
    }
    // CollisionPipeline.cpp:110-198
    void narrowphase_reference() {
        contacts.clear();
        Contact<T> buf[4];
        for (auto &pr : sap_pairs) {
            uint32_t ca = pr.first, cb = pr.second;
            if (ca >= shape.size() || cb >= shape.size()) continue;
            if (!pair_ok(ca, cb)) continue;
            Manifold<T> m;
            if (narrow(shape[ca], collider_tf(ca), shape[cb], collider_tf(cb), m)) {
                make_contacts(m, cbody[ca], cbody[cb], ca, cb, buf);
                for (int i = 0; i < m.n; ++i) contacts.push_back(buf[i]);
            }
        }
        for (uint32_t pi = 0; pi < shape.size(); ++pi) {
            if (shape[pi].kind != SH_PLANE) continue;
            for (uint32_t di = 0; di < shape.size(); ++di) {
                if (di == pi || shape[di].kind == SH_PLANE) continue;
                if (!pair_ok(pi, di)) continue;
                Manifold<T> m;
                if (narrow(shape[di], collider_tf(di), shape[pi], collider_tf(pi), m)) {
                    make_contacts(m, cbody[di], cbody[pi], di, pi, buf);
                    for (int i = 0; i < m.n; ++i) contacts.push_back(buf[i]);
                }
            }
        }
    }
    // Frame pair list for the colour order: canonical SAP pairs then the plane loop (a
    // plane pairs with the colliders of its own scene).  lid = the pair's index in its
    // scene's own list (its SAP pairs, then its plane pairs): the colouring hashes lid, so
    // each scene of a batch is coloured exactly as it would be alone.
    std::vector<uint32_t> lid;
    void build_pairs() {
        spairs.clear();
        lid.clear();
        std::vector<uint32_t> nsap(bscene.empty() ? 1 : *std::max_element(bscene.begin(), bscene.end()) + 1, 0);
        std::vector<uint32_t> nplane(nsap.size(), 0);
        for (auto &pr : sap_pairs)
            if (pair_ok(pr.first, pr.second)) {
                const uint32_t sc = bscene[cbody[pr.first]];
                spairs.push_back({pr.first, pr.second, cbody[pr.first], cbody[pr.second]});
                lid.push_back(nsap[sc]++);
            }
        std::vector<uint32_t> plane_sc;
        std::vector<std::vector<uint32_t>> by_scene(nsap.size());
        for (uint32_t c = 0; c < shape.size(); ++c) by_scene[bscene[cbody[c]]].push_back(c);
        for (uint32_t pi = 0; pi < shape.size(); ++pi) {
            if (shape[pi].kind != SH_PLANE) continue;
            for (uint32_t di : by_scene[bscene[cbody[pi]]]) {
                if (di == pi || shape[di].kind == SH_PLANE) continue;
                if (!pair_ok(pi, di)) continue;
                const uint32_t sc = bscene[cbody[di]];
                spairs.push_back({di, pi, cbody[di], cbody[pi]});
                lid.push_back(nplane[sc]++);
                plane_sc.push_back(sc);
            }
        }
        for (size_t k = 0; k < plane_sc.size(); ++k) lid[lid.size() - plane_sc.size() + k] += nsap[plane_sc[k]];
        const uint32_t P = (uint32_t)spairs.size();
        colour.assign(P, -1);
        pcount.assign(P, 0);
        pcont.resize((size_t)P * 4);
    }
    // Colouring of the substep's active pairs (pairs with contact points); two pairs
    // conflict when they share a dynamic body.  Synchronous randomized rounds whose cost is
    // independent of body degree, replayed exactly as the GPU runs them.  Per dynamic body
    // v: used(v) = colours taken by its pairs, u(v) = its uncoloured pairs.  Round r, for
    // every uncoloured pair p with dynamic bodies D(p):
    //   claim: t(p) = the (tent_hash(p, r) mod min(max(ceil(k/4), 1), palette_max))-th colour outside the
    //          union of used(v), k = sum of u(v) over D(p); key(p) = rank_hash(p, r); each v
    //          in D(p) records the largest key that targets its slot t mod COLOUR_SLOTS;
    //   settle: p takes t(p) if its key is the recorded one at every v in D(p).
    // Two pairs of one body never settle on one colour in one round (one key per slot), nor
    // on a colour used earlier (used(v) is excluded), so the colouring is valid; the pair
    // with the largest key overall always settles, so every round makes progress.
    static constexpr uint32_t COLOUR_SLOTS = 16, MAX_COLOURS = 1024;
    uint32_t palette_max = 8;
    static uint32_t tent_hash(uint32_t p, uint32_t r) { return mix32(p ^ (r * 0x9e3779b9u)); }
    static uint32_t rank_hash(uint32_t p, uint32_t r) { return mix32(p ^ (r * 0x85ebca6bu) ^ 0xc2b2ae35u); }
    uint32_t palette(uint32_t k) const {
        uint32_t m = (k + 3u) / 4u;
        return m < 1u ? 1u : (m > palette_max ? palette_max : m);
    }
    void colour_active() {
        const uint32_t P = (uint32_t)spairs.size(), N = (uint32_t)btype.size();
        constexpr uint32_t W = MAX_COLOURS / 32;
        std::vector<uint32_t> act;
        for (uint32_t p = 0; p < P; ++p)
            if (pcount[p] > 0) act.push_back(p);
        std::vector<uint32_t> ucnt(N, 0), used((size_t)N * W, 0);
        std::vector<uint64_t> claim((size_t)N * COLOUR_SLOTS, 0);
        auto dyn = [&](uint32_t b) { return btype[b] == 1; };
        for (uint32_t p : act) {
            if (dyn(spairs[p].ba)) ++ucnt[spairs[p].ba];
            if (dyn(spairs[p].bb)) ++ucnt[spairs[p].bb];
        }
        colour.assign(P, -1);
        ncolours = 0;
        colour_rounds = 0;
        std::vector<uint32_t> tent(P, 0);
        std::vector<uint64_t> key(P, 0);
        size_t left = act.size();
        for (uint32_t r = 0; left > 0; ++r) {
            for (uint32_t p : act) {
                if (colour[p] >= 0) continue;
                const uint32_t bs[2] = {spairs[p].ba, spairs[p].bb};
                uint32_t k = 0;
                for (uint32_t b : bs)
                    if (dyn(b)) k += ucnt[b];
                uint32_t idx = tent_hash(lid[p], r) % palette(k), t = MAX_COLOURS;
                for (uint32_t w = 0; w < W && t == MAX_COLOURS; ++w) {
                    uint32_t fr = ~0u;
                    for (uint32_t b : bs)
                        if (dyn(b)) fr &= ~used[(size_t)b * W + w];
                    for (uint32_t j = 0; j < 32; ++j)
                        if ((fr >> j) & 1u) {
                            if (idx == 0) { t = w * 32 + j; break; }
                            --idx;
                        }
                }
                if (t == MAX_COLOURS) throw std::runtime_error("more than 1024 colours needed");
                tent[p] = t;
                key[p] = ((uint64_t)(r + 1) << 32) | rank_hash(lid[p], r);
                for (uint32_t b : bs)
                    if (dyn(b)) {
                        uint64_t &c = claim[(size_t)b * COLOUR_SLOTS + t % COLOUR_SLOTS];
                        if (key[p] > c) c = key[p];
                    }
            }
            std::vector<uint32_t> won;
            for (uint32_t p : act) {
                if (colour[p] >= 0) continue;
                bool ok = true;
                for (uint32_t b : {spairs[p].ba, spairs[p].bb})
                    if (dyn(b) && claim[(size_t)b * COLOUR_SLOTS + tent[p] % COLOUR_SLOTS] != key[p]) ok = false;
                if (ok) won.push_back(p);
            }
            for (uint32_t p : won) {
                const uint32_t t = tent[p];
                colour[p] = (int)t;
                for (uint32_t b : {spairs[p].ba, spairs[p].bb})
                    if (dyn(b)) {
                        used[(size_t)b * W + t / 32] |= 1u << (t % 32);
                        --ucnt[b];
                    }
                if ((int)t + 1 > ncolours) ncolours = (int)t + 1;
            }
            left -= won.size();
            colour_rounds = r + 1;
        }
    }
    void narrowphase_colour() {
        for (uint32_t p = 0; p < spairs.size(); ++p) {
            const SPair &s = spairs[p];
            Manifold<T> m;
            EpaGuard<T> g;
            g.has_prev = epa_guard && pcount[p] > 0;   // none in a frame's first substep
            if (g.has_prev) g.prev = pcont[(size_t)p * 4].n;
            g.faces = epa_guard == 2;
            pcount[p] = 0;
            T cut = T(0);
            const bool hit = narrow(shape[s.ca], collider_tf(s.ca), shape[s.cb], collider_tf(s.cb), m,
                                    (g.has_prev || g.faces) ? &g : nullptr, &cut);
            if (cut > T(0)) {
                ++guard_hits;
                if (double(cut) > guard_max_cut) guard_max_cut = double(cut);
            }
            if (hit) {
                make_contacts(m, s.ba, s.bb, s.ca, s.cb, &pcont[(size_t)p * 4]);
                pcount[p] = m.n;
            }
        }
    }

    // ------------------------------------------------------------ body kernels (Body.cpp)
    void update_Iw(uint32_t i) {   // Body.cpp:11-16
        M3<T> R = mat3_cast(q[i]);
        Iw[i] = world_tensor(R, I[i]);
        Iwinv[i] = world_tensor(R, Iinv[i]);
    }
    void integrate(T dt) {   // Body.cpp:18-39
        for (uint32_t i = 0; i < btype.size(); ++i) {
            if (btype[i] == 0) continue;
            update_Iw(i);
            ppos[i] = pos[i];
            lv[i] = lv[i] + force[i] * inv_mass[i] * dt;
            pos[i] = pos[i] + lv[i] * dt;
            pq[i] = q[i];
            av[i] = av[i] + (dt * Iwinv[i]) * (torque[i] - cross(av[i], Iw[i] * av[i]));
            V3<T> ws = (T(0.5) * dt) * av[i];
            Q<T> r(T(0.0), ws.x, ws.y, ws.z);
            q[i] = normalize(q[i] + r * q[i]);
        }
    }
    void update_velocities(T inv_dt) {   // Body.cpp:41-57
        for (uint32_t i = 0; i < btype.size(); ++i) {
            if (btype[i] == 0) continue;
            lv[i] = (pos[i] - ppos[i]) * inv_dt;
            Q<T> dq = q[i] * conj(pq[i]);
            V3<T> w = (T(2) * inv_dt) * V3<T>(dq.x, dq.y, dq.z);
            av[i] = dq.w >= T(0) ? w : -w;
        }
    }
    void apply_pos_impulse(uint32_t i, V3<T> imp, V3<T> r) {   // Body.cpp:59-68
        if (btype[i] == 0) return;
        pos[i] = pos[i] + imp * inv_mass[i];
        V3<T> rot = (Iwinv[i] * cross(r, imp)) * T(0.5);
        Q<T> rq(T(0.0), rot.x, rot.y, rot.z);
        q[i] = normalize(q[i] + rq * q[i]);
    }
    void apply_vel_impulse(uint32_t i, V3<T> imp, V3<T> r) {   // Body.cpp:79-85
        if (btype[i] == 0) return;
        lv[i] = lv[i] + imp * inv_mass[i];
        av[i] = av[i] + Iwinv[i] * cross(r, imp);
    }
    T w_pos(uint32_t i, V3<T> r, V3<T> n) {   // Body.cpp:94-101 (refreshes the world tensors)
        if (btype[i] == 0) return T(0.0);
        update_Iw(i);
        V3<T> c = cross(r, n);
        return inv_mass[i] + dot(c, Iwinv[i] * c);
    }

    // ------------------------------------------------------------ joints (Joint.cpp, Constraint.cpp)
    T w_rot(uint32_t i, V3<T> n) {   // Body.cpp:103-109
        if (btype[i] == 0) return T(0.0);
        update_Iw(i);
        return dot(n, Iwinv[i] * n);
    }
    void apply_rot_impulse(uint32_t i, V3<T> imp) {   // Body.cpp:70-77
        if (btype[i] == 0) return;
        V3<T> rot = (Iwinv[i] * imp) * T(0.5);
        Q<T> rq(T(0.0), rot.x, rot.y, rot.z);
        q[i] = normalize(q[i] + rq * q[i]);
    }
    void apply_rot_vel_impulse(uint32_t i, V3<T> imp) {   // Body.cpp:87-92
        if (btype[i] == 0) return;
        av[i] = av[i] + Iwinv[i] * imp;
    }
    static Q<T> from_axis_angle(V3<T> axis, T angle) {   // quat.hpp:52-58
        T half = angle * T(0.5);
        T sn = t_sin(half);
        V3<T> n = normalize(axis);
        return Q<T>(t_cos(half), n.x * sn, n.y * sn, n.z * sn);
    }
    void set_row(Joint &j, int k, V3<T> v) const {   // Constraint.cpp:5-12
        j.mag[k] = length(v);
        if (!(j.mag[k] > C<T>::EPS)) j.dir[k] = V3<T>(T(1.0), T(0.0), T(0.0));
        else if (joint_mode == 0) j.dir[k] = normalize(v);
        else j.dir[k] = v * (T(1.0) / std::sqrt(length_sq(v)));
    }
    // Joint.cpp:18-71 (prismatic), 92-159 (revolute), 161-182 (fixed); spherical: attachment
    // row, and a swing limit built like the revolute angle limit (Joint.cpp:73-90).
    void joint_errors(Joint &j, T h, bool advance_target = true) {
        if (j.type == J_ROW) return;   // a constraint row keeps the direction and magnitude its caller set
        const uint32_t b1 = j.b1, b2 = j.b2;
        const V3<T> r1w = rotate(q[b1], j.r1), r2w = rotate(q[b2], j.r2);
        const V3<T> p1 = pos[b1] + r1w, p2 = pos[b2] + r2w;
        if (j.type == J_PRISMATIC || j.type == J_FIXED) {
            Q<T> dq = q[b1] * conj(q[b2]);
            set_row(j, 0, V3<T>(dq.x, dq.y, dq.z) * T(2.0));
        }
        if (j.type == J_FIXED) { set_row(j, 1, p1 - p2); return; }
        if (j.type == J_PRISMATIC) {
            V3<T> dp = p1 - p2;
            V3<T> axis = rotate(q[b1], j.axis);
            T ps = dot(dp, axis);
            T over = T(0);
            if (j.limited) {
                T cl = sclamp(ps, j.lower, j.upper);
                over = ps - cl;
            }
            set_row(j, 1, (dp - axis * ps) + axis * over);
            V3<T> drive(T(0.0), T(0.0), T(0.0));
            if (j.act == ACT_SPEED && advance_target) j.target = j.target + j.speed * h;
            if (j.act == ACT_SPEED || j.act == ACT_POSITION) {
                V3<T> ddp = p2 - p1;
                T curr = dot(ddp, axis);
                drive = axis * (j.target - curr);
            }
            set_row(j, 2, drive);
            j.current = ps;
            return;
        }
        if (j.type == J_REVOLUTE) {
            V3<T> a1 = rotate(q[b1], j.axis), a2 = rotate(q[b2], j.axis);
            set_row(j, 0, joint_mode == 0 ? cross(a1, a2) : cross(a2, a1));
            set_row(j, 1, p1 - p2);
            V3<T> l1 = rotate(q[b1], j.laxis), l2 = rotate(q[b2], j.laxis);
            T phi = t_atan2(dot(cross(l1, l2), a1), dot(l1, l2));
            V3<T> lim(T(0.0), T(0.0), T(0.0));
            if (j.limited && (phi < j.lower || phi > j.upper)) {
                phi = sclamp(phi, j.lower, j.upper);
                V3<T> n1 = rotate(from_axis_angle(a1, phi), l1);
                lim = cross(n1, l2);
            }
            set_row(j, 2, -lim);
            V3<T> drive(T(0.0), T(0.0), T(0.0));
            if (j.act == ACT_SPEED && advance_target) j.target = j.target + j.speed * h;
            if (j.act == ACT_SPEED || j.act == ACT_POSITION) {
                T theta = j.target + T(3.14159265358979323846);
                V3<T> bt = (l1 * t_cos(theta) + cross(a1, l1) * t_sin(theta)) + a1 * dot(a1, l1) * (T(1) - t_cos(theta));
                drive = cross(bt, l2);
            }
            set_row(j, 3, drive);
            j.current = phi;
            return;
        }
        // spherical
        set_row(j, 0, p1 - p2);
        if (j.limited) {
            V3<T> a1 = rotate(q[b1], j.axis), a2 = rotate(q[b2], j.axis);
            V3<T> c = cross(a1, a2);
            T th = t_atan2(length(c), dot(a1, a2));
            V3<T> lim(T(0.0), T(0.0), T(0.0));
            if (th > j.upper) {
                V3<T> n1 = rotate(from_axis_angle(c, j.upper), a1);
                lim = cross(n1, a2);
            }
            set_row(j, 1, -lim);
            j.current = th;
        }
    }
    T joint_delta_lambda(const Joint &j, int k, T w1, T w2, T inv_h) const {   // Constraint.cpp:14-18
        T ap = j.comp[k] * inv_h * inv_h;
        return (-j.mag[k] - ap * j.lam[k]) / (w1 + w2 + ap);
    }
    // Constraint.cpp:20-46 for the joint's rows in order (mode 1: each row's error first
    // re-evaluated from the current state).
    void joint_solve(Joint &j, T inv_h, T h) {
        for (int k = 0; k < j.nrows; ++k) {
            if (joint_mode == 1) joint_errors(j, h, false);
            const V3<T> n = j.dir[k];
            if (j.rot[k]) {
                T w1 = w_rot(j.b1, n), w2 = w_rot(j.b2, n);
                T dl = joint_delta_lambda(j, k, w1, w2, inv_h);
                j.lam[k] = j.lam[k] + dl;
                V3<T> imp = dl * n;
                apply_rot_impulse(j.b1, imp);
                apply_rot_impulse(j.b2, -imp);
            } else {
                V3<T> r1w = rotate(q[j.b1], j.r1), r2w = rotate(q[j.b2], j.r2);
                T w1 = w_pos(j.b1, r1w, n), w2 = w_pos(j.b2, r2w, n);
                T dl = joint_delta_lambda(j, k, w1, w2, inv_h);
                j.lam[k] = j.lam[k] + dl;
                V3<T> imp = dl * n;
                apply_pos_impulse(j.b1, imp, r1w);
                apply_pos_impulse(j.b2, -imp, r2w);
            }
        }
    }
    // Joint.cpp:184-220 (prismatic and revolute only)
    void joint_damp(Joint &j, T h) {
        const uint32_t b1 = j.b1, b2 = j.b2;
        if (j.type == J_PRISMATIC) {
            V3<T> dv = (lv[b2] - lv[b1]) * smin(j.damping * h, T(1.0));
            if (length_sq(dv) < C<T>::EPS) return;
            V3<T> r1w = rotate(q[b1], j.r1), r2w = rotate(q[b2], j.r2);
            V3<T> n = normalize(dv);
            T w1 = w_pos(b1, r1w, n), w2 = w_pos(b2, r2w, n);
            V3<T> imp = dv / (w1 + w2);
            apply_vel_impulse(b1, imp, r1w);
            apply_vel_impulse(b2, -imp, r2w);
        } else if (j.type == J_REVOLUTE) {
            V3<T> dw = (av[b2] - av[b1]) * smin(j.damping * h, T(1.0));
            if (length_sq(dw) < C<T>::EPS) return;
            V3<T> n = normalize(dw);
            T w1 = w_rot(b1, n), w2 = w_rot(b2, n);
            V3<T> imp = dw / (w1 + w2);
            apply_rot_vel_impulse(b1, imp);
            apply_rot_vel_impulse(b2, -imp);
        }
    }
    // Colour mode: joints coloured once (first fit in joint order; conflict = a shared
    // dynamic body), solved colour by colour.
    void colour_joints() {
        const uint32_t J = (uint32_t)joints.size();
        jcolour.assign(J, 0);
        njcolours = 0;
        std::vector<std::vector<int>> body_cols(btype.size());
        for (uint32_t i = 0; i < J; ++i) {
            const uint32_t bs[2] = {joints[i].b1, joints[i].b2};
            int c = 0;
            for (;; ++c) {
                bool taken = false;
                for (uint32_t b : bs)
                    if (btype[b] == 1)
                        for (int u : body_cols[b]) taken = taken || u == c;
                if (!taken) break;
            }
            jcolour[i] = c;
            for (uint32_t b : bs)
                if (btype[b] == 1) body_cols[b].push_back(c);
            if (c + 1 > njcolours) njcolours = c + 1;
        }
        joint_order.resize(J);
        for (uint32_t i = 0; i < J; ++i) joint_order[i] = i;
        std::stable_sort(joint_order.begin(), joint_order.end(),
                         [&](uint32_t a, uint32_t b) { return jcolour[a] < jcolour[b]; });
        joints_dirty = false;
    }

    // ------------------------------------------------------------ contact solve (Contact.cpp)
    void contact_position(Contact<T> &k, T inv_h) {   // Contact.cpp:35-75
        (void)inv_h;
        const uint32_t a = k.ba, b = k.bb;
        V3<T> pa = pos[a] + rotate(q[a], k.ra);
        V3<T> pb = pos[b] + rotate(q[b], k.rb);
        k.pa = pa; k.pb = pb;
        T d = dot(pa - pb, k.n);
        if (d <= T(0.0)) { k.collision = false; return; }
        k.collision = true;
        V3<T> raw = pa - pos[a], rbw = pb - pos[b];
        {
            V3<T> vr = (lv[a] + cross(av[a], raw)) - (lv[b] + cross(av[b], rbw));
            k.relv = dot(vr, k.n);
        }
        // normal (Contact.cpp:81-115)
        {
            T wa = w_pos(a, raw, k.n), wb = w_pos(b, rbw, k.n);
            T den = wa + wb;
            if (!(den < C<T>::EPS)) {
                T dl = -d / den;
                k.lam_n += dl;
                if (k.lam_n > T(0.0)) k.lam_n = T(0.0);
                V3<T> imp = dl * k.n;
                apply_pos_impulse(a, imp, raw);
                apply_pos_impulse(b, -imp, rbw);
            }
        }
        // tangent (Contact.cpp:121-195)
        V3<T> ra = rotate(q[a], k.ra), rb = rotate(q[b], k.rb);
        V3<T> pac = pos[a] + ra, pbc = pos[b] + rb;
        k.pa = pac; k.pb = pbc;
        V3<T> ppa = ppos[a] + rotate(pq[a], k.ra);
        V3<T> ppb = ppos[b] + rotate(pq[b], k.rb);
        V3<T> dp = (pac - ppa) - (pbc - ppb);
        V3<T> dpt = dp - dot(dp, k.n) * k.n;
        T slip = length(dpt);
        if (slip < C<T>::EPS) return;
        V3<T> t = normalize(dpt);
        T wa = w_pos(a, ra, t), wb = w_pos(b, rb, t);
        T den = wa + wb;
        if (den < C<T>::EPS) return;
        T dl = -slip / den;
        T old = k.lam_t;
        k.lam_t += dl;
        T cone = k.sf * k.lam_n;
        if (k.lam_t < cone) k.lam_t = cone;
        T act = k.lam_t - old;
        if (std::abs(act) < C<T>::EPS) return;
        V3<T> imp = act * t;
        apply_pos_impulse(a, imp, ra);
        apply_pos_impulse(b, -imp, rb);
    }
    void contact_velocity(Contact<T> &k, T h) {   // Contact.cpp:206-289
        if (!k.collision) return;
        const uint32_t a = k.ba, b = k.bb;
        V3<T> raw = k.pa - pos[a], rbw = k.pb - pos[b];
        V3<T> vr = (lv[a] + cross(av[a], raw)) - (lv[b] + cross(av[b], rbw));
        T vn = dot(vr, k.n);
        V3<T> vt = vr - k.n * vn;
        T vtm = length(vt);
        T e = k.rest;
        if (std::abs(k.relv) <= T(2.0) * T(9.8) * h) e = T(0.0);
        T vtarget = smin(-e * k.relv, T(0.0));
        T dvn = vtarget - vn;
        T dvt = T(0.0);
        if (vtm > C<T>::EPS) {
            T lim = k.df * std::abs(k.lam_n) / h;
            dvt = -smin(lim, vtm);
        }
        T wan = w_pos(a, raw, k.n), wbn = w_pos(b, rbw, k.n);
        T denn = wan + wbn;
        if (denn > C<T>::EPS && std::abs(dvn) > C<T>::EPS) {
            V3<T> imp = (dvn / denn) * k.n;
            apply_vel_impulse(a, imp, raw);
            apply_vel_impulse(b, -imp, rbw);
        }
        if (vtm > C<T>::EPS && std::abs(dvt) > C<T>::EPS) {
            V3<T> t = vt / vtm;
            T wat = w_pos(a, raw, t), wbt = w_pos(b, rbw, t);
            T dent = wat + wbt;
            if (dent > C<T>::EPS) {
                V3<T> imp = (dvt / dent) * t;
                apply_vel_impulse(a, imp, raw);
                apply_vel_impulse(b, -imp, rbw);
            }
        }
    }
    // Solve order of the colour mode: active pairs stably sorted by colour.
    std::vector<uint32_t> solve_order;
    void build_solve_order() {
        const uint32_t P = (uint32_t)spairs.size();
        std::vector<std::pair<int, uint32_t>> keyed;
        for (uint32_t p = 0; p < P; ++p)
            if (pcount[p] > 0) keyed.push_back({colour[p], p});
        std::stable_sort(keyed.begin(), keyed.end(), [](auto &a, auto &b) { return a.first < b.first; });
        solve_order.clear();
        for (auto &k : keyed) solve_order.push_back(k.second);
    }
    // Planted schedule mistakes, for tests that show the colour-order pins catch them:
    // 1 skips the pairs of the last colour, 2 solves the first pair twice.  0 (default): none.
    int planted_fault = 0;
    template <class F> void for_each_colour_contact(F &&f) {
        const int last = solve_order.empty() ? -1 : colour[solve_order.back()];
        for (uint32_t p : solve_order) {
            if (planted_fault == 1 && colour[p] == last) continue;
            const int reps = planted_fault == 2 && p == solve_order.front() ? 2 : 1;
            for (int r = 0; r < reps; ++r)
                for (int k = 0; k < pcount[p]; ++k) f(pcont[(size_t)p * 4 + k]);
        }
    }

    // ------------------------------------------------------------ World::step (World.hpp:192-212)
    void step() {
        const T h = timestep / T(substeps);
        const T inv_h = T(1.0) / h;
        refit(timestep);
        bp_update();
        if (order == ORDER_COLOUR) build_pairs();
        if (joints_dirty) colour_joints();
        for (int s = 0; s < substeps; ++s) {
            if (order == ORDER_REFERENCE) narrowphase_reference();
            else { narrowphase_colour(); colour_active(); build_solve_order(); }
            integrate(h);
            if (order == ORDER_REFERENCE)
                for (auto &k : contacts) contact_position(k, inv_h);
            else
                for_each_colour_contact([&](Contact<T> &k) { contact_position(k, inv_h); });
            // World.hpp:168-170: joint errors from the state after the contact solve, then the rows
            for (auto &j : joints) {
                joint_errors(j, h);   // mode 1: targets advance here; rows re-evaluated in the solve
                if (joint_mode == 1)
                    for (int k = 0; k < 4; ++k) j.lam[k] = T(0);
            }
            if (order == ORDER_REFERENCE)
                for (auto &j : joints) joint_solve(j, inv_h, h);
            else
                for (uint32_t i : joint_order) joint_solve(joints[i], inv_h, h);
            update_velocities(inv_h);
            if (order == ORDER_REFERENCE)
                for (auto &k : contacts) contact_velocity(k, h);
            else
                for_each_colour_contact([&](Contact<T> &k) { contact_velocity(k, h); });
            // World.hpp:174-178
            if (order == ORDER_REFERENCE)
                for (auto &j : joints) joint_damp(j, h);
            else
                for (uint32_t i : joint_order) joint_damp(joints[i], h);
        }
    }
};

}  // namespace orc
