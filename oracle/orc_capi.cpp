// Oracle C-ABI (ctypes entry points for tests/ and bench.py's cpu_baseline leg).
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU pipeline, never the product.
// The descriptor layouts match the product's C-ABI descriptors byte for byte so
// one numpy structured array feeds both.
#include <exception>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>
#include <stdexcept>
#include "orc_world.hpp"

extern "C" {

typedef struct {
    int32_t type;
    uint32_t scene;   // independent scene of a batch
    double mass;
    double inertia[6];   // xx yy zz xy xz yz
    double position[3];
    double orientation[4];   // w x y z
    double linear_velocity[3];
    double angular_velocity[3];
} orc_body_desc;

typedef struct {
    int32_t shape;
    uint32_t body_id;
    int32_t is_static;
    int32_t hull_index;
    double dims[4];
    double local_pos[3];
    double local_rot[4];
    double restitution, static_friction, dynamic_friction;
} orc_collider_desc;

typedef struct {
    int32_t type, actuation;   // JointType / JointActuationType order (Joint.hpp), 3 = spherical
    uint32_t body_1, body_2;
    double r_1[3], r_2[3];
    double axis[3], limit_axis[3];
    double compliance, damping;
    int32_t limited, _pad;
    double lower_limit, upper_limit;
    double target_position, target_speed;
} orc_joint_desc;

}  // extern "C"

namespace {

// Registered hulls (process lifetime): tables in double, and the hull in each precision.
std::deque<orc::HullTables> g_tables;
std::deque<orc::Hull<float>> g_hull_f;
std::deque<orc::Hull<double>> g_hull_d;
template <class T> const orc::Hull<T> *hull_ref(int i);
template <> const orc::Hull<float> *hull_ref<float>(int i) {
    if (i < 0 || (size_t)i >= g_hull_f.size()) throw std::out_of_range("hull index");
    return &g_hull_f[i];
}
template <> const orc::Hull<double> *hull_ref<double>(int i) {
    if (i < 0 || (size_t)i >= g_hull_d.size()) throw std::out_of_range("hull index");
    return &g_hull_d[i];
}
template <class T> orc::ShapeP<T> shape_from(const orc_collider_desc &d) {
    orc::ShapeP<T> s;
    s.kind = d.shape;
    s.a = T(d.dims[0]); s.b = T(d.dims[1]); s.c = T(d.dims[2]); s.d = T(d.dims[3]);
    if (d.shape == orc::SH_HULL) s.hull = hull_ref<T>(d.hull_index);
    return s;
}

struct Any {
    virtual ~Any() {}
    virtual void add_body(const orc_body_desc &d) = 0;
    virtual void add_collider(const orc_collider_desc &d) = 0;
    virtual void step(int frames) = 0;
    virtual int nbodies() const = 0;
    virtual void bodies(double *p, double *q, double *v, double *w) const = 0;
    virtual int npairs() const = 0;
    virtual void pairs(uint32_t *out) const = 0;
    virtual int ncontacts() const = 0;
    virtual void contacts(uint32_t *ids4, double *n3, double *ra3, double *rb3, double *dep, double *lam2) const = 0;
    virtual int nspairs() const = 0;
    virtual void spairs(uint32_t *ca_cb, int32_t *col, int32_t *cnt, int32_t *kind) const = 0;
    virtual int ncolours() const = 0;
    virtual int colour_rounds() const = 0;
    virtual void set_palette_max(uint32_t m) = 0;
    virtual void set_planted_fault(int f) = 0;
    virtual int solve_order(uint32_t *out) const = 0;
    virtual void set_joint_mode(int m) = 0;
    virtual void set_contact_mode(int m) = 0;
    virtual void set_epa_guard(int g) = 0;
    virtual void guard_stats(double *out2) const = 0;
    virtual void set_state(const double *p, const double *q, const double *v, const double *w) = 0;
    virtual void add_joint(const orc_joint_desc &d) = 0;
    virtual int njoints() const = 0;
    virtual void joint_state(double *cur, double *tgt, double *lam4, int32_t *colour) const = 0;
    virtual void set_joint_targets(int n, const uint32_t *ids, const double *tgt, const double *speed) = 0;
    virtual void set_rows(int n, const uint32_t *ids, const double *dir3, const double *mag) = 0;
    virtual void set_forces(const double *f3, const double *t3) = 0;
};

template <class T> typename orc::World<T>::Joint joint_from(const orc_joint_desc &d) {
    typename orc::World<T>::Joint j;
    auto v3 = [](const double *a) { return orc::V3<T>(T(a[0]), T(a[1]), T(a[2])); };
    j.type = d.type; j.act = d.actuation; j.b1 = d.body_1; j.b2 = d.body_2;
    j.r1 = v3(d.r_1); j.r2 = v3(d.r_2); j.axis = v3(d.axis); j.laxis = v3(d.limit_axis);
    j.limited = d.limited != 0; j.lower = T(d.lower_limit); j.upper = T(d.upper_limit);
    j.target = T(d.target_position); j.speed = T(d.target_speed); j.damping = T(d.damping);
    return j;
}

template <class T> struct Impl : Any {
    orc::World<T> w;
    void add_body(const orc_body_desc &d) override {
        orc::SM<T> I(T(d.inertia[0]), T(d.inertia[1]), T(d.inertia[2]), T(d.inertia[3]), T(d.inertia[4]), T(d.inertia[5]));
        w.add_body(d.type, T(d.mass), I, orc::V3<T>(T(d.position[0]), T(d.position[1]), T(d.position[2])),
                   orc::Q<T>(T(d.orientation[0]), T(d.orientation[1]), T(d.orientation[2]), T(d.orientation[3])),
                   orc::V3<T>(T(d.linear_velocity[0]), T(d.linear_velocity[1]), T(d.linear_velocity[2])),
                   orc::V3<T>(T(d.angular_velocity[0]), T(d.angular_velocity[1]), T(d.angular_velocity[2])), d.scene);
    }
    void add_collider(const orc_collider_desc &d) override {
        const orc::ShapeP<T> s = shape_from<T>(d);
        w.add_collider(s, d.body_id, d.is_static != 0,
                       orc::V3<T>(T(d.local_pos[0]), T(d.local_pos[1]), T(d.local_pos[2])),
                       orc::Q<T>(T(d.local_rot[0]), T(d.local_rot[1]), T(d.local_rot[2]), T(d.local_rot[3])),
                       T(d.restitution), T(d.static_friction), T(d.dynamic_friction));
    }
    void step(int frames) override { for (int f = 0; f < frames; ++f) w.step(); }
    int nbodies() const override { return (int)w.btype.size(); }
    void bodies(double *p, double *q, double *v, double *av) const override {
        for (size_t i = 0; i < w.btype.size(); ++i) {
            if (p) { p[3 * i] = w.pos[i].x; p[3 * i + 1] = w.pos[i].y; p[3 * i + 2] = w.pos[i].z; }
            if (q) { q[4 * i] = w.q[i].w; q[4 * i + 1] = w.q[i].x; q[4 * i + 2] = w.q[i].y; q[4 * i + 3] = w.q[i].z; }
            if (v) { v[3 * i] = w.lv[i].x; v[3 * i + 1] = w.lv[i].y; v[3 * i + 2] = w.lv[i].z; }
            if (av) { av[3 * i] = w.av[i].x; av[3 * i + 1] = w.av[i].y; av[3 * i + 2] = w.av[i].z; }
        }
    }
    void set_state(const double *p, const double *q, const double *v, const double *av) override {
        for (size_t i = 0; i < w.btype.size(); ++i) {
            if (p) w.pos[i] = orc::V3<T>(T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2]));
            if (q) w.q[i] = orc::Q<T>(T(q[4 * i]), T(q[4 * i + 1]), T(q[4 * i + 2]), T(q[4 * i + 3]));
            if (v) w.lv[i] = orc::V3<T>(T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2]));
            if (av) w.av[i] = orc::V3<T>(T(av[3 * i]), T(av[3 * i + 1]), T(av[3 * i + 2]));
        }
    }
    int npairs() const override { return (int)w.sap_pairs.size(); }
    void pairs(uint32_t *out) const override {
        for (size_t i = 0; i < w.sap_pairs.size(); ++i) { out[2 * i] = w.sap_pairs[i].first; out[2 * i + 1] = w.sap_pairs[i].second; }
    }
    // Contacts of the last substep, in solve-list order: reference mode = emission
    // order; colour mode = pair order, points in manifold order.
    template <class F> void each_contact(F &&f) const {
        if (w.order == orc::ORDER_REFERENCE) {
            for (auto &k : w.contacts) f(k);
        } else {
            for (size_t p = 0; p < w.spairs.size(); ++p)
                for (int k = 0; k < w.pcount[p]; ++k) f(w.pcont[p * 4 + k]);
        }
    }
    int ncontacts() const override { int n = 0; each_contact([&](const orc::Contact<T> &) { ++n; }); return n; }
    void contacts(uint32_t *ids4, double *n3, double *ra3, double *rb3, double *dep, double *lam2) const override {
        size_t i = 0;
        each_contact([&](const orc::Contact<T> &k) {
            if (ids4) { ids4[4 * i] = k.ba; ids4[4 * i + 1] = k.bb; ids4[4 * i + 2] = k.ca; ids4[4 * i + 3] = k.cb; }
            if (n3) { n3[3 * i] = k.n.x; n3[3 * i + 1] = k.n.y; n3[3 * i + 2] = k.n.z; }
            if (ra3) { ra3[3 * i] = k.ra.x; ra3[3 * i + 1] = k.ra.y; ra3[3 * i + 2] = k.ra.z; }
            if (rb3) { rb3[3 * i] = k.rb.x; rb3[3 * i + 1] = k.rb.y; rb3[3 * i + 2] = k.rb.z; }
            if (dep) dep[i] = k.depth;
            if (lam2) { lam2[2 * i] = k.lam_n; lam2[2 * i + 1] = k.lam_t; }
            ++i;
        });
    }
    int nspairs() const override { return (int)w.spairs.size(); }
    // colour: active pairs -> colour of the last substep, inactive -> -1;
    // kind: active -> -1, inactive -> -2.
    void spairs(uint32_t *ca_cb, int32_t *col, int32_t *cnt, int32_t *kind) const override {
        for (size_t p = 0; p < w.spairs.size(); ++p) {
            if (ca_cb) { ca_cb[2 * p] = w.spairs[p].ca; ca_cb[2 * p + 1] = w.spairs[p].cb; }
            const bool act = w.pcount[p] > 0;
            if (col) col[p] = act ? w.colour[p] : -1;
            if (cnt) cnt[p] = w.pcount[p];
            if (kind) kind[p] = act ? -1 : -2;
        }
    }
    int ncolours() const override { return w.ncolours; }
    int colour_rounds() const override { return (int)w.colour_rounds; }
    void set_palette_max(uint32_t m) override { w.palette_max = m; }
    void set_planted_fault(int f) override { w.planted_fault = f; }
    int solve_order(uint32_t *out) const override {
        if (out) std::copy(w.solve_order.begin(), w.solve_order.end(), out);
        return (int)w.solve_order.size();
    }
    void set_joint_mode(int m) override { w.joint_mode = m; }
    void set_contact_mode(int m) override { w.contact_mode = m; }
    void set_epa_guard(int g) override { w.epa_guard = g; }
    void guard_stats(double *out2) const override { out2[0] = double(w.guard_hits); out2[1] = w.guard_max_cut; }
    void add_joint(const orc_joint_desc &d) override { w.add_joint(joint_from<T>(d), T(d.compliance)); }
    int njoints() const override { return (int)w.joints.size(); }
    void joint_state(double *cur, double *tgt, double *lam4, int32_t *colour) const override {
        for (size_t i = 0; i < w.joints.size(); ++i) {
            const auto &j = w.joints[i];
            if (cur) cur[i] = j.current;
            if (tgt) tgt[i] = j.target;
            if (lam4) for (int k = 0; k < 4; ++k) lam4[4 * i + k] = k < j.nrows ? double(j.lam[k]) : 0.0;
            if (colour) colour[i] = i < w.jcolour.size() ? w.jcolour[i] : -1;
        }
    }
    void set_joint_targets(int n, const uint32_t *ids, const double *tgt, const double *speed) override {
        for (int k = 0; k < n; ++k) {
            auto &j = w.joints[ids[k]];
            if (tgt) j.target = T(tgt[k]);
            if (speed) j.speed = T(speed[k]);
        }
    }
    void set_rows(int n, const uint32_t *ids, const double *dir3, const double *mag) override {
        for (int k = 0; k < n; ++k) {
            auto &j = w.joints[ids[k]];
            if (j.type != orc::World<T>::J_ROW) continue;
            if (dir3) j.dir[0] = orc::V3<T>(T(dir3[3 * k]), T(dir3[3 * k + 1]), T(dir3[3 * k + 2]));
            if (mag) j.mag[0] = T(mag[k]);
        }
    }
    void set_forces(const double *f3, const double *t3) override {
        for (size_t i = 0; i < w.force.size(); ++i) {
            if (f3) w.force[i] = orc::V3<T>(T(f3[3 * i]), T(f3[3 * i + 1]), T(f3[3 * i + 2]));
            if (t3) w.torque[i] = orc::V3<T>(T(t3[3 * i]), T(t3[3 * i + 1]), T(t3[3 * i + 2]));
        }
    }
};

// Errors of one joint between two bodies at given poses (Joint.cpp:18-182): out =
// 4 rows x (direction xyz, magnitude), current position, target position.
template <class T> void joint_errors_t(const orc_joint_desc *d, const double *pose1, const double *pose2, double h,
                                       double *out) {
    orc::World<T> w;
    orc::SM<T> I(T(1), T(1), T(1), T(0), T(0), T(0));
    for (const double *p : {pose1, pose2})
        w.add_body(1, T(1), I, orc::V3<T>(T(p[0]), T(p[1]), T(p[2])), orc::Q<T>(T(p[3]), T(p[4]), T(p[5]), T(p[6])),
                   orc::V3<T>(), orc::V3<T>());
    orc_joint_desc dd = *d;
    dd.body_1 = 0; dd.body_2 = 1;
    w.add_joint(joint_from<T>(dd), T(d->compliance));
    auto &j = w.joints[0];
    w.joint_errors(j, T(h));
    for (int k = 0; k < 4; ++k) {
        out[4 * k] = j.dir[k].x; out[4 * k + 1] = j.dir[k].y; out[4 * k + 2] = j.dir[k].z;
        out[4 * k + 3] = k < j.nrows ? double(j.mag[k]) : 0.0;
    }
    out[16] = j.current;
    out[17] = j.target;
}

template <class T> orc::TF<T> tf7(const double *p) {
    orc::TF<T> t;
    t.pos = orc::V3<T>(T(p[0]), T(p[1]), T(p[2]));
    t.rot = orc::Q<T>(T(p[3]), T(p[4]), T(p[5]), T(p[6]));
    return t;
}

// GJK + EPA of one pair with statistics: out = gjk status, gjk iterations, epa status,
// epa iterations, normal xyz, depth, EPA witness xyz, GJK exit direction xyz, fallback
// (15 values).  guess6 (optional): warm start in A's frame, GJK direction xyz and EPA
// seed normal xyz (a zero seed normal: none).
template <class T> int gjk_epa_t(const orc_collider_desc *a, const double *ta, const orc_collider_desc *b,
                                 const double *tb, const double *guess6, double *out, const double *guard3 = nullptr) {
    const orc::ShapeP<T> sa = shape_from<T>(*a), sb = shape_from<T>(*b);
    const orc::TF<T> A = tf7<T>(ta), Bt = tf7<T>(tb);
    orc::Manifold<T> m;
    orc::GjkEpaStats st{};
    orc::EpaOut<T> e;
    orc::V3<T> g, sn, gd;
    bool seed = false;
    if (guess6) {
        g = orc::V3<T>(T(guess6[0]), T(guess6[1]), T(guess6[2]));
        sn = orc::V3<T>(T(guess6[3]), T(guess6[4]), T(guess6[5]));
        seed = !(guess6[3] == 0.0 && guess6[4] == 0.0 && guess6[5] == 0.0);
    }
    orc::EpaGuard<T> eg;
    if (guard3) {   // guard3: a direction (has_prev) and, in guard3[3], the face-normal flag
        eg.has_prev = !(guard3[0] == 0.0 && guard3[1] == 0.0 && guard3[2] == 0.0);
        eg.prev = orc::V3<T>(T(guard3[0]), T(guard3[1]), T(guard3[2]));
        eg.faces = guard3[3] != 0.0;
    }
    const bool hit = orc::gjk_epa_manifold(sa, A, sb, Bt, m, guess6 ? &g : nullptr, &st, &e, seed ? &sn : nullptr, &gd,
                                           guard3 ? &eg : nullptr);
    out[0] = st.gjk_status; out[1] = st.gjk_iters; out[2] = st.epa_status; out[3] = st.epa_iters;
    out[4] = e.n.x; out[5] = e.n.y; out[6] = e.n.z; out[7] = e.depth;
    out[8] = e.p.x; out[9] = e.p.y; out[10] = e.p.z;
    out[11] = gd.x; out[12] = gd.y; out[13] = gd.z; out[14] = st.fallback;
    return hit ? 1 : 0;
}

// One narrowphase call (same layout as the reference driver's ref_collide).  guess6
// (optional): warm start of a GJK pair in A's frame (GJK direction, EPA seed normal; a
// zero seed normal: none); other pairs ignore it.
template <class T> int collide_t(const orc_collider_desc *a, const double *ta, const orc_collider_desc *b,
                                        const double *tb, double *out, const double *guess6 = nullptr) {
    orc::Manifold<T> m;
    const orc::ShapeP<T> sa = shape_from<T>(*a), sb = shape_from<T>(*b);
    const bool gjk_pair = (sa.kind == orc::SH_HULL || sb.kind == orc::SH_HULL) && sa.kind != orc::SH_PLANE &&
                          sb.kind != orc::SH_PLANE;
    bool hit;
    if (guess6 && gjk_pair) {
        const orc::V3<T> g{T(guess6[0]), T(guess6[1]), T(guess6[2])}, sn{T(guess6[3]), T(guess6[4]), T(guess6[5])};
        const bool seed = !(guess6[3] == 0.0 && guess6[4] == 0.0 && guess6[5] == 0.0);
        // a zero direction means the plain seed (the GPU's convention)
        const bool warm = !(guess6[0] == 0.0 && guess6[1] == 0.0 && guess6[2] == 0.0);
        hit = orc::gjk_epa_manifold<T>(sa, tf7<T>(ta), sb, tf7<T>(tb), m, warm ? &g : nullptr, nullptr, nullptr,
                                    warm && seed ? &sn : nullptr);
    } else {
        hit = orc::dispatch(sa, tf7<T>(ta), sb, tf7<T>(tb), m);
    }
    out[0] = hit ? m.n : 0;
    out[1] = m.normal.x; out[2] = m.normal.y; out[3] = m.normal.z;
    for (int i = 0; i < 4; ++i) {
        out[4 + 4 * i] = i < m.n ? (double)m.pos[i].x : 0.0;
        out[5 + 4 * i] = i < m.n ? (double)m.pos[i].y : 0.0;
        out[6 + 4 * i] = i < m.n ? (double)m.pos[i].z : 0.0;
        out[7 + 4 * i] = i < m.n ? (double)m.depth[i] : 0.0;
    }
    return hit ? 1 : 0;
}
}  // namespace

// Test infrastructure: GJK (plain seed tf_b.pos - tf_a.pos, GJK.cpp:42-96) of one pair with
// its states recorded; out4 = status, iterations, the first iteration whose state (rank,
// simplex points w, search direction) equals the state p <= 16 iterations earlier (-1 if
// none), and that p.  A repeated state repeats forever, since the iteration is a function
// of that state alone.
template <class T> int gjk_trace_t(const orc_collider_desc *a, const double *ta, const orc_collider_desc *b,
                                   const double *tb, int *out4) {
    const orc::ShapeP<T> sa = shape_from<T>(*a), sb = shape_from<T>(*b);
    orc::MinkDiff<T> md{&sa, &sb, tf7<T>(ta), tf7<T>(tb)};
    orc::V3<T> guess = md.tb.pos - md.ta.pos;
    if (orc::length_sq(guess) < orc::C<T>::EPS) guess = orc::V3<T>(1, 0, 0);
    struct St { int rank; T w[12]; T d[3]; };
    std::vector<St> hist;
    orc::Simplex<T> s;
    orc::V3<T> dir = guess;
    const orc::SVert<T> v0 = md.sample(dir);
    s.push(v0);
    dir = -v0.w;
    out4[0] = orc::GJK_FAILED; out4[1] = 0; out4[2] = -1; out4[3] = 0;
    for (int it = 0; it < 128; ++it) {
        out4[1] = it + 1;
        const orc::SVert<T> nx = md.sample(dir);
        if (orc::dot(nx.w, dir) < T(0)) { out4[0] = orc::GJK_VALID; return 0; }
        s.push(nx);
        bool inside = false;
        if (s.rank == 2) orc::simplex2(s, dir);
        else if (s.rank == 3) orc::simplex3(s, dir);
        else if (s.rank == 4) inside = orc::simplex4(s, dir);
        if (inside) { out4[0] = orc::GJK_INSIDE; return 0; }
        St st{};
        st.rank = s.rank;
        for (int k = 0; k < s.rank; ++k) { st.w[3 * k] = s.v[k].w.x; st.w[3 * k + 1] = s.v[k].w.y; st.w[3 * k + 2] = s.v[k].w.z; }
        st.d[0] = dir.x; st.d[1] = dir.y; st.d[2] = dir.z;
        if (out4[2] < 0)
            for (int p = 1; p <= 16 && p <= (int)hist.size(); ++p) {
                const St &o = hist[hist.size() - p];
                if (o.rank == st.rank && std::memcmp(o.w, st.w, sizeof(T) * 3 * st.rank) == 0 && std::memcmp(o.d, st.d, sizeof(T) * 3) == 0) {
                    out4[2] = it + 1;
                    out4[3] = p;
                    break;
                }
            }
        hist.push_back(st);
    }
    return 0;
}
extern "C" {

void *orc_create(int fp32, int order, double timestep, int substeps) {
    Any *a;
    if (fp32) {
        auto *i = new Impl<float>();
        i->w.timestep = (float)timestep; i->w.substeps = substeps; i->w.order = order;
        a = i;
    } else {
        auto *i = new Impl<double>();
        i->w.timestep = timestep; i->w.substeps = substeps; i->w.order = order;
        a = i;
    }
    return a;
}
void orc_destroy(void *h) { delete static_cast<Any *>(h); }
int orc_add_bodies(void *h, int n, const orc_body_desc *d) {
    for (int i = 0; i < n; ++i) static_cast<Any *>(h)->add_body(d[i]);
    return 0;
}
int orc_add_colliders(void *h, int n, const orc_collider_desc *d) {
    try {
        for (int i = 0; i < n; ++i) static_cast<Any *>(h)->add_collider(d[i]);
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}
int orc_step(void *h, int frames) {
    try {
        static_cast<Any *>(h)->step(frames);
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}
int orc_num_bodies(void *h) { return static_cast<Any *>(h)->nbodies(); }
void orc_get_bodies(void *h, double *p, double *q, double *v, double *w) { static_cast<Any *>(h)->bodies(p, q, v, w); }
void orc_set_bodies(void *h, const double *p, const double *q, const double *v, const double *w) {
    static_cast<Any *>(h)->set_state(p, q, v, w);
}
int orc_num_pairs(void *h) { return static_cast<Any *>(h)->npairs(); }
void orc_get_pairs(void *h, uint32_t *out) { static_cast<Any *>(h)->pairs(out); }
int orc_num_contacts(void *h) { return static_cast<Any *>(h)->ncontacts(); }
void orc_get_contacts(void *h, uint32_t *ids4, double *n3, double *ra3, double *rb3, double *dep, double *lam2) {
    static_cast<Any *>(h)->contacts(ids4, n3, ra3, rb3, dep, lam2);
}
int orc_num_solve_pairs(void *h) { return static_cast<Any *>(h)->nspairs(); }
void orc_get_solve_pairs(void *h, uint32_t *ca_cb, int32_t *col, int32_t *cnt, int32_t *kind) {
    static_cast<Any *>(h)->spairs(ca_cb, col, cnt, kind);
}
int orc_num_colours(void *h) { return static_cast<Any *>(h)->ncolours(); }
int orc_colour_rounds(void *h) { return static_cast<Any *>(h)->colour_rounds(); }
int orc_add_joints(void *h, int n, const orc_joint_desc *d) {
    for (int i = 0; i < n; ++i) {
        if (d[i].type < 0 || d[i].type > 4) return -1;
        static_cast<Any *>(h)->add_joint(d[i]);
    }
    return 0;
}
int orc_num_joints(void *h) { return static_cast<Any *>(h)->njoints(); }
void orc_get_joints(void *h, double *cur, double *tgt, double *lam4, int32_t *colour) {
    static_cast<Any *>(h)->joint_state(cur, tgt, lam4, colour);
}
void orc_set_joint_targets(void *h, int n, const uint32_t *ids, const double *tgt, const double *speed) {
    static_cast<Any *>(h)->set_joint_targets(n, ids, tgt, speed);
}
void orc_set_joint_mode(void *h, int m) { static_cast<Any *>(h)->set_joint_mode(m); }
void orc_set_contact_mode(void *h, int m) { static_cast<Any *>(h)->set_contact_mode(m); }
void orc_set_epa_guard(void *h, int g) { static_cast<Any *>(h)->set_epa_guard(g); }
// EPA results the guard replaced so far and the largest depth it cut (m).
void orc_guard_stats(void *h, double *out2) { static_cast<const Any *>(h)->guard_stats(out2); }
void orc_joint_errors(int fp32, const orc_joint_desc *d, const double *pose1, const double *pose2, double h, double *out) {
    if (fp32) joint_errors_t<float>(d, pose1, pose2, h, out);
    else joint_errors_t<double>(d, pose1, pose2, h, out);
}
float orc_sinf(float x) { return orc::rb_sin(x); }
float orc_cosf(float x) { return orc::rb_cos(x); }
float orc_atan2f(float y, float x) { return orc::rb_atan2(y, x); }
int orc_set_palette_max(void *h, int m) {
    if (m < 1) return -1;
    static_cast<Any *>(h)->set_palette_max((uint32_t)m);
    return 0;
}
uint32_t orc_mix32(uint32_t x) { return orc::mix32(x); }
// Direction and magnitude of constraint rows (joints of type 4), as callers write them.
void orc_set_rows(void *h, int n, const uint32_t *ids, const double *dir3, const double *mag) {
    static_cast<Any *>(h)->set_rows(n, ids, dir3, mag);
}
// BodyCollection::force / torque of every body (public fields the reference's callers write).
void orc_set_forces(void *h, const double *f3, const double *t3) { static_cast<Any *>(h)->set_forces(f3, t3); }
// Test infrastructure: plants a colour-schedule mistake (see World::planted_fault).
void orc_set_planted_fault(void *h, int f) { static_cast<Any *>(h)->set_planted_fault(f); }
// The last substep's colour-order solve list (pair indices of the frame's pair list).
int orc_get_solve_order(void *h, uint32_t *out) { return static_cast<Any *>(h)->solve_order(out); }

int orc_collide(int fp32, const orc_collider_desc *a, const double *ta, const orc_collider_desc *b, const double *tb,
                double *out) {
    try {
        return fp32 ? collide_t<float>(a, ta, b, tb, out) : collide_t<double>(a, ta, b, tb, out);
    } catch (const std::exception &) {
        return -1;
    }
}
int orc_collide_warm(int fp32, const orc_collider_desc *a, const double *ta, const orc_collider_desc *b,
                     const double *tb, const double *guess6, double *out) {
    try {
        return fp32 ? collide_t<float>(a, ta, b, tb, out, guess6) : collide_t<double>(a, ta, b, tb, out, guess6);
    } catch (const std::exception &) {
        return -1;
    }
}
int orc_gjk_epa(int fp32, const orc_collider_desc *a, const double *ta, const orc_collider_desc *b, const double *tb,
                const double *guess6, double *out) {
    try {
        return fp32 ? gjk_epa_t<float>(a, ta, b, tb, guess6, out) : gjk_epa_t<double>(a, ta, b, tb, guess6, out);
    } catch (const std::exception &) {
        return -1;
    }
}

// orc_gjk_epa with the EPA guard (see gjk_epa_manifold): guard4 = a direction (world; zero:
// none) and a face-normal flag.
int orc_gjk_epa_guarded(int fp32, const orc_collider_desc *a, const double *ta, const orc_collider_desc *b,
                        const double *tb, const double *guard3, double *out) {
    try {
        return fp32 ? gjk_epa_t<float>(a, ta, b, tb, nullptr, out, guard3) : gjk_epa_t<double>(a, ta, b, tb, nullptr, out, guard3);
    } catch (const std::exception &) {
        return -1;
    }
}

int orc_gjk_trace(int fp32, const orc_collider_desc *a, const double *ta, const orc_collider_desc *b, const double *tb,
                  int *out4) {
    return fp32 ? gjk_trace_t<float>(a, ta, b, tb, out4) : gjk_trace_t<double>(a, ta, b, tb, out4);
}

// Registers a hull (ConvexHull.hpp:149-420 tables); returns its index for
// orc_collider_desc.hull_index.  faces may be null (vertex-only hull).
int orc_hull_register(const double *verts, int nv, const uint32_t *faces, int nf) {
    if (nv <= 0) return -1;
    for (int i = 0; i < 3 * nf; ++i)
        if (faces[i] >= (uint32_t)nv) return -1;
    g_tables.push_back(orc::build_hull_tables(verts, (uint32_t)nv, faces, faces ? (uint32_t)nf : 0u));
    g_hull_f.push_back(orc::Hull<float>::from(g_tables.back()));
    g_hull_d.push_back(orc::Hull<double>::from(g_tables.back()));
    return (int)g_tables.size() - 1;
}
// Table sizes: polygons, polygon indices, v2p, v2n, p2p entries; adjacency flag.
int orc_hull_sizes(int i, int32_t *out6) {
    if (i < 0 || (size_t)i >= g_tables.size()) return -1;
    const auto &t = g_tables[i];
    out6[0] = (int32_t)t.pn.size(); out6[1] = (int32_t)t.pidx.size(); out6[2] = (int32_t)t.v2p_i.size();
    out6[3] = (int32_t)t.v2n_i.size(); out6[4] = (int32_t)t.p2p_i.size(); out6[5] = g_hull_d[i].adj ? 1 : 0;
    return 0;
}
// Tables: poly normals (3 per poly), poly offsets (P+1), indices, and the three adjacency
// lists as offsets (nv+1 / nv+1 / P+1) plus indices.
int orc_hull_tables(int i, double *pn, uint32_t *poff, uint32_t *pidx, uint32_t *v2p_o, uint32_t *v2p_i,
                    uint32_t *v2n_o, uint32_t *v2n_i, uint32_t *p2p_o, uint32_t *p2p_i) {
    if (i < 0 || (size_t)i >= g_tables.size()) return -1;
    const auto &t = g_tables[i];
    for (size_t k = 0; k < t.pn.size(); ++k) { pn[3 * k] = t.pn[k].x; pn[3 * k + 1] = t.pn[k].y; pn[3 * k + 2] = t.pn[k].z; }
    auto cp = [](const std::vector<uint32_t> &v, uint32_t *o) { if (o) for (size_t k = 0; k < v.size(); ++k) o[k] = v[k]; };
    cp(t.poff, poff); cp(t.pidx, pidx); cp(t.v2p_o, v2p_o); cp(t.v2p_i, v2p_i);
    cp(t.v2n_o, v2n_o); cp(t.v2n_i, v2n_i); cp(t.p2p_o, p2p_o); cp(t.p2p_i, p2p_i);
    return 0;
}
// Volume and unit-density inertia (xx yy zz xy xz yz) of a registered hull.
int orc_hull_mass(int i, double *out7) {
    if (i < 0 || (size_t)i >= g_tables.size()) return -1;
    orc::hull_mass_properties(g_tables[i], out7);
    return 0;
}

}  // extern "C"
