// Host orchestration of the XPBD substep pipeline and the rbgpu C-ABI.
//
// One world = one HIP stream.  Per frame (World::step, include/rbps/API/World.hpp:192-212):
//   refit + stable endpoint sort + sweep (pairs), plane pairs appended, narrowphase bins
//   (one host synchronisation: pair count and bin offsets);
//   then per substep: narrowphase -> colouring of the pairs with points (grid-wide rounds,
//   then one workgroup for the rest; one host synchronisation for the colour offsets) ->
//   integrate -> position solve -> velocity update -> velocity solve.  Each solve runs a
//   large colour as its own launch and a run of small colours in one workgroup.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/rbgpu.h"
#include "rbgpu_kernels.hpp"
#include "rbgpu_joints.hpp"
#include "rbgpu_gjk.hpp"
#include "rbgpu_hull.hpp"
#include "rbgpu_mlp.hpp"

using namespace rb;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(RBGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int TPB = 256;
inline uint32_t nblk(uint32_t n, uint32_t t = TPB) { return (n + t - 1) / t; }

template <class T> int dalloc(T *&p, size_t n) {
    p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)&p, n * sizeof(T));
    if (e != hipSuccess) return fail(RBGPU_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return RBGPU_OK;
}
template <class T> void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Readback block (pinned host memory).
struct Readback {
    uint32_t n_sap;
    uint32_t n_pairs;
    uint32_t pad0_;
    int32_t dirty, overflow, pad1_, pad2_;
    uint32_t ctr[CTR_SIZE];
    uint32_t col_off[MAX_COLOURS + 1];
    uint32_t bin_off[NBINS + 1];
    uint64_t points;
    uint32_t log_rows;
    uint32_t scene_err;
    uint32_t scene_err_tail;   // sc_err as queued after the last frame of an rbgpu_step call
};

__global__ void k_set_int(int *p, int v) { *p = v; }

// Every body row, in one contiguous copy to the host (a strided copy of single fields moves
// 16 bytes per row and is far slower).
hipError_t body_rows(std::vector<BodyRow> &dst, const BodyRow *rows, uint32_t n) {
    dst.resize(n);
    return n ? hipMemcpy(dst.data(), rows, (size_t)n * sizeof(BodyRow), hipMemcpyDeviceToHost) : hipSuccess;
}

__global__ void k_offsets(const uint32_t *hoff, uint32_t ntiles, uint32_t nd, const uint32_t *n, uint32_t *out) {
    uint32_t d = threadIdx.x + blockIdx.x * blockDim.x;
    if (d < nd) out[d] = hoff[d * ntiles];
    if (d == 0) out[nd] = *n;
}

}  // namespace

struct rbgpu_world {
    double timestep = 1.0 / 60.0;
    int substeps = 20;
    float margin = 0.1f, reinsert = 0.05f;
    hipStream_t st = nullptr;

    // host-side scene (descriptors) and broadphase bookkeeping for (re)uploads
    std::vector<rbgpu_body_desc> hb;
    std::vector<rbgpu_collider_desc> hc;
    std::vector<rbgpu_joint_desc> hj;
    struct HostHull {
        rbh::HullTables t;
        float amin[3], amax[3];   // local AABB of the FP32 vertices
        // bounding ellipsoid of the FP32 vertices (local frame, axes along the local axes): centre
        // and semi-axes, every vertex v satisfying sum_k ((v_k - c_k) / a_k)^2 <= 1 with room for
        // FP32 rounding; extent along a unit u at most c . u + |diag(a) u|
        float ec[3], ea[3];
    };
    std::vector<HostHull> hh;           // convex hull data (rbgpu_add_hull)
    std::vector<float4> h_jtg;          // joint target / speed / current (survives rebuilds)
    std::vector<float> h_jlam;          // joint row lambdas (never reset)
    std::vector<float4> h_jrow_user;    // a constraint row's (direction, magnitude) as its caller wrote it
    std::vector<float4> h_force, h_torque;   // caller-written force / torque (w = 1), else gravity / zero
    // BodyCollection::force / torque rows as the integrator reads them: xyz and, in force.w, the mass
    void body_forces(std::vector<float4> &f, std::vector<float4> &t) const {
        const size_t n = hb.size();
        f.resize(n);
        t.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const float mass = (float)hb[i].mass;
            const bool uf = i < h_force.size() && h_force[i].w != 0.0f, ut = i < h_torque.size() && h_torque[i].w != 0.0f;
            f[i] = uf ? make_float4(h_force[i].x, h_force[i].y, h_force[i].z, mass)
                      : make_float4(0.0f, -9.8f * mass, 0.0f, mass);   // BodyAPI.hpp:61
            t[i] = ut ? make_float4(h_torque[i].x, h_torque[i].y, h_torque[i].z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    std::vector<float4> h_fmin, h_fmax;
    std::vector<uint32_t> h_uid, h_key, h_val;
    std::vector<uint2> h_eppos;
    bool built = false, sorted_once = false;

    uint32_t N = 0, M = 0, S = 0, NEP = 0, NJ = 0;
    uint32_t nscenes = 1;               // independent scenes (batch) in this world
    // one workgroup per scene for colouring and solve (option scene_kernels: -1 auto, 0 off, 1 on)
    int scene_mode = -1;
    bool scenes_on = false;
    uint32_t max_scene_bodies = 0, scene_threads = 64;
    uint32_t *sc_plane_off = nullptr, *sc_plane_idx = nullptr, *sc_size = nullptr, *sc_base = nullptr;
    uint32_t *sc_work = nullptr, *sc_list = nullptr, *sc_coff = nullptr, *sc_ncol = nullptr, *sc_err = nullptr;
    uint32_t *sj_order = nullptr, *sj_base = nullptr, *sj_coff = nullptr, sj_njc = 0;   // per-scene joint tables
    // per-scene body lists, each body's index in its scene's list, the colouring's scene order
    uint32_t *sb_off = nullptr, *sb_idx = nullptr, *sb_loc = nullptr, *sc_order = nullptr;
    SceneTabs scene_tabs() const {
        return SceneTabs{sap_start, sap_end, sc_plane_off, sc_plane_idx, sc_base, sc_work, sc_list, sc_coff, sc_ncol, sc_err,
                         sj_order, sj_base, sj_coff, sj_njc, sb_off, sb_idx, sb_loc, sc_order};
    }
    bool scene_path() const {
        return nscenes > 1 && (scene_mode == 1 || (scene_mode == -1 && nscenes >= 16 && max_scene_bodies <= 4096));
    }
    uint32_t *cscene = nullptr, *slot_scene = nullptr, *sap_start = nullptr, *sap_end = nullptr;
    uint32_t *plane_lid = nullptr, *pair_lid = nullptr;
    uint32_t *bb_hits = nullptr, *bb_blk = nullptr;   // box-box overlap lists (per 128-pair block)
    // hull tables on the device and GJK/EPA statistics per pair slot
    float4 *hv = nullptr, *hpn = nullptr, *hamin = nullptr, *hamax = nullptr, *hell = nullptr;
    uint32_t *htab = nullptr;
    HullMeta *hmeta = nullptr;
    HullSet hs{};
    int hull_vpl = 1;                   // hull vertices per lane in the wave kernels (1: all hulls <= 64)
    uint32_t *gstat = nullptr;
    uint32_t *gjk_keep = nullptr, *gjk_pos = nullptr, *gjk_cand = nullptr, *gjk_ncand = nullptr;   // cull -> candidates
    uint4 *gjk_rec = nullptr;   // per GJK-bin position, this frame: slot, colliders, body a (the cull's first load)
    uint32_t *gjk_recb = nullptr;   // and body b
    uint32_t *gjk_dn = nullptr;   // pairs the fused GJK/EPA pass hands to the full-size EPA pass
    uint32_t *gjk_ckey = nullptr, *gjk_ckey2 = nullptr, *gjk_cand2 = nullptr;   // candidates sorted by last iterations
    int gjk_sort = 0;   // candidate order: 0 bin order, 1 iterations, 2 overlap, 3 both
    float4 *gjk_axis = nullptr;   // per pair slot: the axis that last separated it (rejection hint)
    // the hints carried across a frame's new pair list (table keyed by collider pair)
    unsigned long long *ax_keys = nullptr;
    float4 *ax_vals = nullptr;
    uint32_t ax_cap = 0;
    bool axis_carry = true;
    int64_t ax_restore = -1;   // >= 0: the next frame's hint table came from a snapshot (its size; 0 empty)
    int contact_mode = 0;                               // 1: capsule-capsule normal from A to B
    int epa_guard = 0;                                  // EPA depth bounded along the previous normal (1), and faces (2)
    bool gjk_cull = true;                               // bounding-sphere rejection in front of GJK
    uint32_t gjk_grid = 4096, epa_grid = 8192;          // workgroups of the two GJK/EPA passes
    int gjk_layout = 4;                                 // GJK pass: 0 a pair per wavefront, 1 per 16-lane row (hulls <= 64
                                                        // vertices), 2 per lane, 3 refilled rows, 4 GJK + EPA per row
                                                        // (5: with row buffers most pairs outgrow, for tests)
    uint32_t epa_sub = 16;                              // EPA pass: listed pairs per work item
    int gjk_occ = 3;                                    // GJK row pass: waves per EU the registers are budgeted for
    int epa_occ = 0, solve_occ = 0;                     // EPA pass / per-scene position solve: 4 = budget for 4 waves
    // the narrowphase bins other than GJK run on a second stream beside the GJK chain (default
    // priority, joined back by an event before the colouring)
    // 0: off; where the side stream forks: 1 before the GJK chain, 2 after its rejection tests
    // and compaction, 3 after the GJK pass (beside the EPA pass)
    // Colouring state between substeps (per-scene path): clean after a frame whose scenes cleared
    // what they set; colour_fast: this frame clears only the counters; colour_epochs: the claim
    // epoch reached on the device since its last full clear.
    bool colour_clean = false, colour_fast = false;
    uint32_t colour_epochs = 0;
    int side_stream = 2;
    // The fork point: the main stream records the fork event there (the side stream waits on
    // it); the side kernels themselves are queued once the whole GJK chain is, so that the
    // host's launches of the side bins never hold back the main stream's next kernel.
    bool fork_armed = false, fork_recorded = false;   // set by launch_narrowphase while launch_gjk runs
    int fork_at(int stage) {
        if (!fork_armed || stage < side_stream) return RBGPU_OK;
        fork_armed = false;
        HIPCHK(hipEventRecord(ev_fork, st));
        HIPCHK(hipStreamWaitEvent(st2, ev_fork, 0));
        fork_recorded = true;
        return RBGPU_OK;
    }
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint32_t *gjk_hits = nullptr, *gjk_blk = nullptr;   // overlapping GJK pairs per 4-pair workgroup
    float4 *gjk_sx = nullptr;                          // their final simplices (8 float4 each)
    // learned warm start: parameters (FP32 master, BF16 copy), gradient, Adam moments, the
    // logged training set (features, targets) and the per-frame inference buffers
    float *mlp_wf = nullptr, *mlp_grad = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_loss = nullptr;
    bf16 *mlp_wb = nullptr;
    uint64_t mlp_step = 0;
    int mlp_mode = 0;                   // 0 plain, 1 GJK direction, 2 GJK direction + EPA seed
    bf16 *ds_feat = nullptr;
    float4 *ds_tgt = nullptr;
    uint64_t ds_cap = 0, ds_n = 0;   // log capacity; rows logged (host copy of *ds_cnt after sync_log_count)
    uint64_t ds_hi = 0;               // upper bound of the rows logged (ds_n + GJK pairs since the last sync)
    uint32_t *ds_cnt = nullptr;       // rows logged (device)
    bool ds_logging = false;
    bf16 *inf_feat = nullptr;
    float4 *inf_guess = nullptr;
    uint32_t inf_cap = 0;
    unsigned long long *gjk_tot = nullptr;   // iterations and counts of the GJK pairs (rbgpu_get_gjk_totals)
    Joints Jt{};
    uint32_t *jorder = nullptr;         // joints by colour
    std::vector<uint32_t> jcol_off;     // joint colour offsets
    std::vector<int32_t> h_jcolour;
    Bodies B{};
    Colliders C{};
    Contacts K{};
    float4 *fmin = nullptr, *fmax = nullptr;
    uint32_t *uid = nullptr;
    uint32_t *ep_key[2] = {nullptr, nullptr}, *ep_val[2] = {nullptr, nullptr};
    uint2 *ep_pos = nullptr;
    float4 *sw_a = nullptr;
    float4 *sw_b = nullptr;
    uint32_t *ep_cnt = nullptr, *ep_off = nullptr;
    uint32_t *ep_hist = nullptr;
    int *flags = nullptr;            // [0] dirty [1] overflow
    uint32_t *scal = nullptr;        // [0] n_sap [1] n_pairs
    uint32_t *tile_sums = nullptr;   // scan scratch
    ColourBufs cb{};                 // substep colouring (per-body state sized N, lists sized pcap)
    uint2 *plane_pairs = nullptr;
    uint32_t n_plane = 0;
    unsigned long long *d_points = nullptr;

    // pair stage (capacity pcap); contact rows are indexed by pair id
    uint32_t pcap = 0;
    uint2 *pairs = nullptr;
    uint32_t *ka = nullptr, *va = nullptr, *kb = nullptr, *solve_list = nullptr, *bin_list = nullptr;
    uint32_t *p_hist = nullptr, *offs = nullptr;   // offs: bin offsets

    Readback *rb = nullptr;          // pinned
    uint32_t n_pairs = 0, n_sap = 0, ncol = 0;
    std::vector<uint32_t> col_off, bin_off;
    struct Run { uint32_t c0, c1, group; };
    std::vector<Run> runs;          // this substep's solve launches
    uint32_t frames = 0;
    // tuning: grid-wide colouring rounds before the single-workgroup rounds; colours with
    // fewer pairs than run_min are solved in single-workgroup runs
    uint32_t head_rounds = 96, round_batch = 12, tail_min = 2048, run_min = 256, palette_max = 8;
    int joint_mode = 0;             // joint solver: 0 the reference's, 1 XPBD rows (rbgpu_joints.hpp)

    // timing
    int timing_kernel = 0;
    uint32_t timing_stride = 1;     // time every timing_stride-th frame
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;
    double timing_bytes = 0;
    bool timing_points_pending = false;

    ~rbgpu_world() {
        release();
        release_mlp();
    }

    void release_pairs() {
        dfree(pairs); dfree(ka); dfree(va); dfree(kb);
        dfree(solve_list); dfree(bin_list); dfree(p_hist); dfree(pair_lid); dfree(bb_hits); dfree(bb_blk); dfree(gstat);
        dfree(gjk_hits); dfree(gjk_blk); dfree(gjk_sx); dfree(gjk_keep); dfree(gjk_pos); dfree(gjk_cand);
        dfree(gjk_rec); dfree(gjk_recb);
        dfree(gjk_ckey); dfree(gjk_ckey2); dfree(gjk_cand2);
        dfree(gjk_ncand); dfree(gjk_dn); dfree(gjk_axis); dfree(sc_work); dfree(sc_list);
        dfree(cb.acol); dfree(cb.tent); dfree(cb.rest); dfree(cb.act); dfree(cb.blk);
        dfree(K.cnt); dfree(K.cc); dfree(K.bb); dfree(K.hd); dfree(K.pt); dfree(K.dep);
        pcap = 0;
    }
    void release_mlp() {
        dfree(mlp_wf); dfree(mlp_grad); dfree(mlp_m); dfree(mlp_v); dfree(mlp_loss); dfree(mlp_wb);
        dfree(ds_feat); dfree(ds_tgt); dfree(ds_cnt); dfree(inf_feat); dfree(inf_guess); dfree(gjk_tot);
        ds_cap = ds_n = ds_hi = 0;
        inf_cap = 0;
    }
    void release() {
        if (st) (void)hipStreamSynchronize(st);
        release_pairs();
        dfree(B.r); dfree(B.f); dfree(B.t); dfree(B.I0); dfree(B.I1);
        dfree(C.meta); dfree(C.dims); dfree(C.lpos); dfree(C.lrot); dfree(C.mat); dfree(C.hull);
        dfree(hv); dfree(hpn); dfree(hamin); dfree(hamax); dfree(hell); dfree(htab); dfree(hmeta);
        hs = HullSet{};
        dfree(Jt.jm); dfree(Jt.jr1); dfree(Jt.jr2); dfree(Jt.jax); dfree(Jt.jlx); dfree(Jt.jtg); dfree(Jt.jrow);
        dfree(Jt.jlam); dfree(jorder); dfree(sj_order); dfree(sj_base); dfree(sj_coff);
        sj_njc = 0;
        dfree(fmin); dfree(fmax); dfree(uid); dfree(ep_key[0]); dfree(ep_key[1]); dfree(ep_val[0]); dfree(ep_val[1]);
        dfree(ep_pos); dfree(sw_a); dfree(sw_b); dfree(ep_cnt); dfree(ep_off); dfree(ep_hist);
        dfree(flags); dfree(scal); dfree(tile_sums);
        dfree(cb.used); dfree(cb.ucnt); dfree(cb.claim); dfree(cb.hist); dfree(cb.off); dfree(cb.cur); dfree(cb.ctr); dfree(cb.epoch);
        colour_clean = false;
        dfree(plane_pairs); dfree(offs); dfree(d_points);
        dfree(ax_keys); dfree(ax_vals);
        ax_cap = 0;
        n_pairs = n_sap = 0;   // no pair list survives the buffers
        bin_off.clear();
        dfree(cscene); dfree(slot_scene); dfree(sap_start); dfree(sap_end); dfree(plane_lid);
        dfree(sc_plane_off); dfree(sc_plane_idx); dfree(sc_size); dfree(sc_base); dfree(sc_coff); dfree(sc_ncol);
        dfree(sc_err); dfree(sb_off); dfree(sb_idx); dfree(sb_loc); dfree(sc_order);
        for (auto &e : ev_pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        ev_pool.clear();
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (rb) (void)hipHostFree(rb);
        rb = nullptr;
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        ev_fork = ev_join = nullptr;
        if (st2) (void)hipStreamDestroy(st2);
        st2 = nullptr;
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
        built = false;
    }

    // ------------------------------------------------------------ primitives
    int scan(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total) {
        uint32_t nt = std::max<uint32_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
        if (nt > (uint32_t)(SCAN_TOP_THREADS * SCAN_TOP_ITEMS)) return fail(RBGPU_ERR_CAPACITY, "scan too large");
        k_scan_reduce<<<nt, SCAN_THREADS, 0, st>>>(in, nullptr, n, tile_sums);
        k_scan_top<<<1, SCAN_TOP_THREADS, 0, st>>>(tile_sums, nt, total);
        k_scan_apply<<<nt, SCAN_THREADS, 0, st>>>(in, nullptr, n, tile_sums, out);
        return RBGPU_OK;
    }
    // Scan of a device-counted array (n_max bounds the launch).
    int scan_dev(const uint32_t *in, uint32_t *out, const uint32_t *n_dev, uint32_t n_max, uint32_t *total) {
        uint32_t nt = std::max<uint32_t>(1, (n_max + SCAN_TILE - 1) / SCAN_TILE);
        if (nt > (uint32_t)(SCAN_TOP_THREADS * SCAN_TOP_ITEMS)) return fail(RBGPU_ERR_CAPACITY, "scan too large");
        k_scan_reduce<<<nt, SCAN_THREADS, 0, st>>>(in, n_dev, n_max, tile_sums);
        k_scan_top<<<1, SCAN_TOP_THREADS, 0, st>>>(tile_sums, nt, total);
        k_scan_apply<<<nt, SCAN_THREADS, 0, st>>>(in, n_dev, n_max, tile_sums, out);
        return RBGPU_OK;
    }
    // One stable 8-bit radix pass (k0,v0) -> (k1,v1) on bits [shift, shift+8); the element
    // count is n (or *n_dev when given, bounded by n).
    int radix_pass(const uint32_t *k0, const uint32_t *v0, uint32_t *k1, uint32_t *v1, uint32_t n, int shift,
                   uint32_t *hist, const int *skip, const uint32_t *n_dev = nullptr) {
        uint32_t nt = std::max<uint32_t>(1, (n + RS_TILE - 1) / RS_TILE);
        k_rs_hist<<<nt, RS_THREADS, 0, st>>>(k0, n_dev, n, shift, nt, hist, skip);
        int rc = scan(hist, hist, nt * RS_BINS, nullptr);
        if (rc) return rc;
        k_rs_scatter<<<nt, RS_THREADS, 0, st>>>(k0, v0, k1, v1, n_dev, n, shift, nt, hist, skip);
        return RBGPU_OK;
    }

    // The GJK bins [g0, g0 + ng) (hull-hull and hull-primitive, adjacent in the bin order)
    // in one go: bounding-sphere rejection and compaction (world steps), features for the
    // training log (every bin position) and for inference (the candidates), the network's
    // warm starts, then the two GJK/EPA passes over the candidates.
    bool gjk_log = false;
    bool gjk_explicit = false;   // warm starts given by the caller (rbgpu_collide_world_pairs)
    int launch_gjk(uint32_t g0, uint32_t ng) {
        const int guard = epa_guard;
        gjk_log = false;
        if (ds_logging && !gjk_explicit && ds_feat) {
            if (ds_hi + ng > ds_cap) {   // near capacity: the exact count decides
                int rc = sync_log_count();
                if (rc) return rc;
            }
            gjk_log = ds_hi + ng <= ds_cap;
            if (!gjk_log) ds_logging = false;   // the log is full
        }
        const bool infer = mlp_mode && mlp_wb && !gjk_explicit && inf_guess && ng <= inf_cap;
        float4 *tl = gjk_log ? ds_tgt : nullptr;   // rows after *ds_cnt
        const bool timed = timing_kernel == 11 && frames % timing_stride == 0;
        const uint32_t *cand = nullptr, *ncand = nullptr;
        if (int rcf = fork_at(1)) return rcf;
        if (gjk_cull && !gjk_explicit) {
            static_assert(TPB == 256, "k_gjk_compact_blk takes the cull's 256-pair workgroups");
            // bin order from the cull's per-workgroup counts (their scan in one workgroup)
            const bool blk = !gjk_sort && nblk(ng) <= (uint32_t)(SCAN_TOP_THREADS * SCAN_TOP_ITEMS);
            // records of this frame's list (the GJK bins start at g0 every substep)
            const bool rec = gjk_rec && g0 == bin_off[BIN_HH] && ng == bin_off[BIN_HP] - bin_off[BIN_HH];
            k_gjk_cull<<<nblk(ng), TPB, 0, st>>>(bin_list, g0, ng, C, B, K, hs, gjk_axis, gstat, gjk_keep,
                                                 blk ? gjk_pos : nullptr, rec ? gjk_rec : nullptr, gjk_recb);
            if (blk) {
                k_scan_top<<<1, SCAN_TOP_THREADS, 0, st>>>(gjk_pos, nblk(ng), gjk_ncand);
                k_gjk_compact_blk<<<nblk(ng), TPB, 0, st>>>(gjk_keep, gjk_pos, ng, gjk_cand);
                cand = gjk_cand;
            } else {
                int rc = scan(gjk_keep, gjk_pos, ng, gjk_ncand);
                if (rc) return rc;
                if (gjk_sort) {   // candidates by a key from their slot's previous substep (stable)
                    k_gjk_compact_keyed<<<nblk(ng), TPB, 0, st>>>(gjk_keep, gjk_pos, ng, bin_list, g0, gstat, gjk_cand,
                                                                   gjk_ckey, gjk_sort);
                    rc = radix_pass(gjk_ckey, gjk_cand, gjk_ckey2, gjk_cand2, ng, 0, p_hist, nullptr, gjk_ncand);
                    if (rc) return rc;
                    cand = gjk_cand2;
                } else {
                    k_gjk_compact<<<nblk(ng), TPB, 0, st>>>(gjk_keep, gjk_pos, ng, gjk_cand);
                    cand = gjk_cand;
                }
            }
            ncand = gjk_ncand;
            if (timed) k_add_count<<<1, 1, 0, st>>>(gjk_ncand, d_points);
        }
        if (int rcf = fork_at(2)) return rcf;
        if (gjk_log) k_gjk_features<<<nblk(ng), TPB, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, ds_feat, ds_cnt);
        if (infer) {
            k_gjk_features<<<nblk(ng), TPB, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, inf_feat, nullptr);
            const uint32_t tiles = (ng + MLP_TILE - 1) / MLP_TILE;
            k_mlp_infer<<<std::min<uint32_t>(tiles, 1024), MLP_INFER_THREADS, 0, st>>>(inf_feat, ng, ncand,
                                                                                    MlpParams{mlp_wb, mlp_wf},
                                                                         inf_guess);
        }
        const float4 *gp = (gjk_explicit || infer) ? inf_guess : nullptr;
        const int seed = mlp_mode == 2 ? 1 : 0;
        // four pairs per wavefront (one per DPP row) when every hull has at most 64 vertices
        const bool lanes = gjk_layout == 2;
        const bool refill = gjk_layout == 3 && hull_vpl == 1;
        const bool fused = gjk_layout >= 4 && hull_vpl == 1;
        const bool rows = !lanes && gjk_layout == 1 && hull_vpl == 1;
        const uint32_t ppb = lanes ? GJK_LANE_SEG
                             : refill ? GJK_REFILL_SEG
                             : rows ? (uint32_t)gjk_pass_pairs<16>() : (uint32_t)gjk_pass_pairs<64>();
        const uint32_t nvb = (ng + ppb - 1) / ppb;
        const uint32_t sub = std::min<uint32_t>(epa_sub, ppb);
        const uint32_t g1 = std::min<uint32_t>(nvb, gjk_grid), g2 = std::min<uint32_t>(nvb * (ppb / sub), epa_grid);
        hipEvent_t *e = timed_begin(11);   // the GJK pass alone
        if (fused) {
            // GJK, EPA and the manifold per 16-lane row (four pairs per one-wave workgroup);
            // the pairs whose polytope outgrew the row's go to the full-size EPA pass
            k_set_int<<<1, 1, 0, st>>>((int *)gjk_dn, 0);
            // one-wave workgroups: twice the GJK grid (8,192 by default) measured as fast as four times
            // and half the register-spill prologue writes of every wave
            const uint32_t gf = std::min<uint32_t>((ng + 3) / 4, 2 * gjk_grid);
            if (gjk_layout == 5)   // row buffers that most pairs outgrow: exercises the hand-over
                k_gjk_epa_rows<8, 16, 8, 3><<<gf, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp, seed, tl,
                                                               ds_cnt, gjk_dn, gjk_hits, gjk_sx, gjk_axis, guard);
            else if (gjk_occ == 4)
                k_gjk_epa_rows<FUSED_NV, FUSED_NF, FUSED_HC, 4><<<gf, 64, 0, st>>>(
                    bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp, seed, tl, ds_cnt, gjk_dn, gjk_hits, gjk_sx, gjk_axis,
                    guard);
            else if (gjk_occ == 2)
                k_gjk_epa_rows<FUSED_NV, FUSED_NF, FUSED_HC, 2><<<gf, 64, 0, st>>>(
                    bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp, seed, tl, ds_cnt, gjk_dn, gjk_hits, gjk_sx, gjk_axis,
                    guard);
            else
                k_gjk_epa_rows<FUSED_NV, FUSED_NF, FUSED_HC, 3><<<gf, 64, 0, st>>>(
                    bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp, seed, tl, ds_cnt, gjk_dn, gjk_hits, gjk_sx, gjk_axis,
                    guard);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            k_epa_pass<1><<<FUSED_DEFER_GRID, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, 1, 1, C, B, K, hs, gstat, gp, seed, tl,
                                                          ds_cnt, gjk_hits, nullptr, gjk_sx, guard, gjk_dn);
        } else if (refill) {
            // a wavefront per range of GJK_REFILL_SEG candidates, its rows refilled as they finish
            const uint32_t gr = std::min<uint32_t>((nvb + GJK_PASS_WAVES - 1) / GJK_PASS_WAVES, gjk_grid);
            if (gjk_occ == 2)
                k_gjk_refill<2><<<gr, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp,
                                                                   tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            else if (gjk_occ == 3)
                k_gjk_refill<3><<<gr, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp,
                                                                   tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            else
                k_gjk_refill<4><<<gr, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp,
                                                                   tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            if (epa_occ == 4)
                k_epa_pass<1, 4><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                    ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
            else
                k_epa_pass<1><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                 ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
        } else if (lanes) {
            // one pair per lane, 64-thread workgroups; the EPA pass reads 16-candidate segments
            const uint32_t gl = std::min<uint32_t>(nblk(ng, 64), gjk_grid);
            k_gjk_lanes<<<gl, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp, tl, ds_cnt, gjk_hits,
                                           gjk_blk, gjk_sx, gjk_axis);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            if (hull_vpl == 1)
                k_epa_pass<1><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                 ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
            else
                k_epa_pass<HULL_VPL><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp,
                                                        seed, tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
        } else if (rows) {
            // waves per EU the compiler budgets registers for: 4 caps them at 128 (some spill), 3 and 2 do not
            if (gjk_occ == 2)
                k_gjk_pass<4, 16, 2><<<g1, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat,
                                                                         gp, tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            else if (gjk_occ == 3)
                k_gjk_pass<4, 16, 3><<<g1, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat,
                                                                         gp, tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            else
                k_gjk_pass<4, 16><<<g1, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp,
                                                                      tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            if (epa_occ == 4)
                k_epa_pass<1, 4><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                    ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
            else
                k_epa_pass<1><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                 ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
        } else if (hull_vpl == 1) {
            k_gjk_pass<1, 64><<<g1, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat, gp,
                                                                  tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx, gjk_axis);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            k_epa_pass<1><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl, ds_cnt,
                                             gjk_hits, gjk_blk, gjk_sx, guard);
        } else {
            k_gjk_pass<HULL_VPL, 64><<<g1, 64 * GJK_PASS_WAVES, 0, st>>>(bin_list, g0, ng, cand, ncand, C, B, K, hs, gstat,
                                                                         gp, tl, ds_cnt, gjk_hits, gjk_blk, gjk_sx,
                                                                         gjk_axis);
            timed_end(e);
            if (int rcf = fork_at(3)) return rcf;
            k_epa_pass<HULL_VPL><<<g2, 64, 0, st>>>(bin_list, g0, ng, cand, ncand, ppb, sub, C, B, K, hs, gstat, gp, seed, tl,
                                                    ds_cnt, gjk_hits, gjk_blk, gjk_sx, guard);
        }
        if (e && !cand) timing_bytes += (double)ng * gjk_pass_bytes();
        if (gjk_tot)
            k_gjk_totals<<<std::min<uint32_t>(nblk(ng, 256), GJK_TOTALS_BLOCKS), 256, 0, st>>>(bin_list, g0, ng, gstat, gjk_tot);
        if (gjk_tot && fused) k_add_count<<<1, 1, 0, st>>>(gjk_dn, gjk_tot + 6);
        if (gjk_log) {
            k_log_advance<<<1, 1, 0, st>>>(ds_cnt, ncand, ng);
            ds_hi += ng;
        }
        gjk_log = false;
        return RBGPU_OK;
    }

    // Substep graph: captured from `body` (one substep), then launched `substeps` times.  The
    // executable graph is kept and updated in place by the next capture when the update
    // succeeds (same launch shapes), else instantiated again.
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    template <class F> int run_graph(F &&body) {
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = body();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(RBGPU_ERR_HIP, "substep graph capture failed");
        }
        if (gexec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(gexec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
            }
        }
        if (!gexec && hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0) != hipSuccess) {
            gexec = nullptr;
            (void)hipGraphDestroy(g);
            return fail(RBGPU_ERR_HIP, "substep graph instantiation failed");
        }
        (void)hipGraphDestroy(g);
        for (int k = 0; k < substeps; ++k) HIPCHK(hipGraphLaunch(gexec, st));
        ++graph_frames;
        return RBGPU_OK;
    }
    uint64_t graph_frames = 0;

    // Per-scene colouring errors of the last frames of an rbgpu_step call: the call queues a
    // copy of the sticky flag behind its work; calls that wait for the stream report it.
    bool err_tail_queued = false;
    int queue_error_check() {
        if (!sc_err || !rb || !scene_path()) return RBGPU_OK;
        HIPCHK(hipMemcpyAsync(&rb->scene_err_tail, sc_err, 4, hipMemcpyDeviceToHost, st));
        err_tail_queued = true;
        return RBGPU_OK;
    }
    // After a synchronisation of the world's stream.
    int pending_error() {
        if (!err_tail_queued) return RBGPU_OK;
        err_tail_queued = false;
        const uint32_t e = rb->scene_err_tail;
        rb->scene_err_tail = 0;
        if (!e) return RBGPU_OK;
        colour_clean = false;
        HIPCHK(hipMemsetAsync(sc_err, 0, 4, st));
        HIPCHK(hipStreamSynchronize(st));
        if (e & 1u) return fail(RBGPU_ERR_CAPACITY, "more than 1024 colours needed (a frame of the last step)");
        return fail(RBGPU_ERR_CAPACITY, "colouring did not converge (a frame of the last step)");
    }
    // Waits for the world's stream and reports a pending colouring error.
    int sync() {
        if (!st) return RBGPU_OK;
        HIPCHK(hipStreamSynchronize(st));
        return pending_error();
    }

    // Rows in the training log (device count -> ds_n, ds_hi).
    int sync_log_count() {
        uint32_t c = 0;
        if (ds_cnt) {
            HIPCHK(hipMemcpyAsync(&rb->log_rows, ds_cnt, 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            c = rb->log_rows;
        }
        ds_n = ds_hi = c;
        return RBGPU_OK;
    }

    // Narrowphase of every bin of the current pair list (bin_off): one launch per bin.  The
    // GJK chain (both GJK bins) goes first on the main stream; with side_stream and pairs in
    // both, the other bins run meanwhile on st2 (their slots and scratch are disjoint from the
    // chain's) and the main stream waits for them at the end.  colour_reset: also clear the
    // colouring state there (it depends on no narrowphase result).
    int launch_narrowphase(bool colour_reset = false) {
        const uint32_t g0 = bin_off[BIN_HH], ng = bin_off[BIN_HP] - bin_off[BIN_HH];
        const bool fork = side_stream && st2 && ng && bin_off[NBINS] - ng > 0;
        if (ng) {
            fork_armed = fork;
            fork_recorded = false;
            int rc = launch_gjk(g0, ng);
            const bool recorded = fork_recorded;   // the chain reached the fork point
            fork_armed = fork_recorded = false;
            if (fork && (!rc || recorded)) {
                // the side bins once the whole chain is queued (behind the fork event, recorded at
                // the fork point or, when the chain ended before it, now), also when the chain
                // failed after the fork point; then the join
                const int rs = launch_side(true, colour_reset, recorded);
                if (!rc) rc = rs;
                const hipError_t e1 = hipEventRecord(ev_join, st2);
                const hipError_t e2 = hipStreamWaitEvent(st, ev_join, 0);
                if (!rc && (e1 != hipSuccess || e2 != hipSuccess)) rc = fail(RBGPU_ERR_HIP, "side stream join failed");
            }
            if (rc || fork) return rc;
        }
        return launch_side(false, colour_reset);
    }
    // The narrowphase bins other than GJK (and the colouring state's reset): on st2, forked
    // from the main stream (launch_narrowphase joins it back), or on the main stream.
    int launch_side(bool fork, bool colour_reset, bool recorded = false) {
        if (fork && !recorded) {
            HIPCHK(hipEventRecord(ev_fork, st));
            HIPCHK(hipStreamWaitEvent(st2, ev_fork, 0));
        }
        const hipStream_t so = fork ? st2 : st;
        if (colour_reset && colour_fast)
            k_colour_ctr_reset<<<1, 512, 0, so>>>(cb);
        else if (colour_reset)
            k_colour_reset<<<nblk(std::max<uint32_t>(N * (COLOUR_WORDS / 4), MAX_COLOURS)), TPB, 0, so>>>(N, cb);
        for (int b = 0; b < NBINS; ++b) {
            uint32_t o = bin_off[b], n = bin_off[b + 1] - bin_off[b];
            if (!n || b == BIN_HH || b == BIN_HX) continue;
            const uint32_t g = nblk(n, 128);
            switch (b) {
                case BIN_SS: k_narrow<BIN_SS><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_SB: k_narrow<BIN_SB><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_BB:
                    k_narrow_bb_sat<<<g, 128, 0, so>>>(bin_list, o, n, C, B, K, bb_hits, bb_blk);
                    k_narrow_bb_full<<<g, 128, 0, so>>>(bb_hits, bb_blk, C, B, K);
                    break;
                case BIN_SC: k_narrow<BIN_SC><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_CC: k_narrow<BIN_CC><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K, contact_mode); break;
                case BIN_CB: k_narrow<BIN_CB><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_SP: k_narrow<BIN_SP><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_BP: k_narrow<BIN_BP><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_CP: k_narrow<BIN_CP><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
                case BIN_HP:
                    if (hull_vpl == 1) k_narrow_hull_plane<1><<<n, 64, 0, so>>>(bin_list, o, C, B, K, hs);
                    else k_narrow_hull_plane<HULL_VPL><<<n, 64, 0, so>>>(bin_list, o, C, B, K, hs);
                    break;
                default: k_narrow<BIN_OTHER><<<g, 128, 0, so>>>(bin_list, o, n, C, B, K); break;
            }
        }
        return RBGPU_OK;
    }

    // Algorithmic bytes of one pair in the GJK pass: per side the hull vertices (16 B each, the
    // world's mean count), hull meta 64 B, collider rows 68 B, body pose 32 B; the pair's slot,
    // colliders and bodies 20 B; out: status 4 B and the separating-axis hint 16 B (an
    // overlapping pair's simplex, 128 B, is not counted).
    double gjk_pass_bytes() const {
        double nv = 0.0;
        for (const HostHull &h : hh) nv += (double)h.t.v.size();
        nv = hh.empty() ? 0.0 : nv / (double)hh.size();
        return 2.0 * (16.0 * nv + 64.0 + 68.0 + 32.0) + 20.0 + 20.0;
    }

    hipEvent_t *timed_begin(int kernel) {
        if (timing_kernel != kernel || frames % timing_stride) return nullptr;
        if (ev_used == ev_pool.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
            ev_pool.push_back({a, b});
        }
        auto &e = ev_pool[ev_used++];
        (void)hipEventRecord(e.first, st);
        return &e.second;
    }
    void timed_end(hipEvent_t *e) {
        if (e) (void)hipEventRecord(*e, st);
    }

    // ------------------------------------------------------------ upload
    int build();
    float bounding_radius(const rbgpu_collider_desc &d) const;
    int build_joints();
    int download_state();
    int ensure_pair_capacity(uint32_t need);
    int frame();
};

// ----------------------------------------------------------------- build / upload
int rbgpu_world::ensure_pair_capacity(uint32_t need) {
    if (need <= pcap && pairs) return RBGPU_OK;
    HIPCHK(hipStreamSynchronize(st));
    release_pairs();
    uint32_t cap = std::max<uint32_t>(need + need / 2, 4096);
    int rc = 0;
    rc |= dalloc(pairs, cap);
    rc |= dalloc(ka, cap); rc |= dalloc(va, cap); rc |= dalloc(kb, cap); rc |= dalloc(solve_list, cap);
    rc |= dalloc(bin_list, cap); rc |= dalloc(pair_lid, cap); rc |= dalloc(sc_work, cap); rc |= dalloc(sc_list, cap);
    rc |= dalloc(bb_hits, nblk(cap, 128) * 128); rc |= dalloc(bb_blk, nblk(cap, 128)); rc |= dalloc(gstat, cap);
    if (!hh.empty()) {
        // listed pairs: segments of up to gjk_pass_pairs<16>() = 16 per virtual block of 4 or 16 candidates
        const size_t nb1 = (cap + GJK_PASS_WAVES - 1) / GJK_PASS_WAVES;
        const size_t nh = ((size_t)cap + 15) / 16 * 16 + 16;
        rc |= dalloc(gjk_hits, nh); rc |= dalloc(gjk_blk, nb1);
        rc |= dalloc(gjk_sx, nh * 8);
        rc |= dalloc(gjk_keep, cap); rc |= dalloc(gjk_pos, cap); rc |= dalloc(gjk_cand, cap); rc |= dalloc(gjk_ncand, 1);
        rc |= dalloc(gjk_rec, cap); rc |= dalloc(gjk_recb, cap);
        rc |= dalloc(gjk_dn, 1);
        rc |= dalloc(gjk_ckey, cap); rc |= dalloc(gjk_ckey2, cap); rc |= dalloc(gjk_cand2, cap);
        rc |= dalloc(gjk_axis, cap);
        if (!rc) HIPCHK(hipMemsetAsync(gjk_axis, 0, (size_t)cap * 16, st));   // no hints yet
    }
    if (mlp_wb) {
        dfree(inf_feat); dfree(inf_guess);
        rc |= dalloc(inf_feat, (size_t)cap * MLP_IN); rc |= dalloc(inf_guess, 2 * (size_t)cap);
        inf_cap = cap;
    }
    rc |= dalloc(cb.acol, cap); rc |= dalloc(cb.tent, cap); rc |= dalloc(cb.rest, cap); rc |= dalloc(cb.act, cap);
    rc |= dalloc(cb.blk, nblk(cap) + 1);
    uint32_t nt = (cap + RS_TILE - 1) / RS_TILE;
    rc |= dalloc(p_hist, (size_t)nt * RS_BINS);
    rc |= dalloc(K.cnt, cap); rc |= dalloc(K.cc, cap); rc |= dalloc(K.bb, cap);
    rc |= dalloc(K.hd, 2 * (size_t)cap);    // 32-byte header per slot
    rc |= dalloc(K.pt, 16 * (size_t)cap);   // 64-byte block per point, 4 points per slot
    rc |= dalloc(K.dep, 4 * (size_t)cap);
    if (rc) return RBGPU_ERR_OOM;
    HIPCHK(hipMemsetAsync(K.cnt, 0, cap * sizeof(int), st));
    pcap = cap;
    return RBGPU_OK;
}

// Radius of a sphere about a collider's origin that holds its shape (FP32, rounded up):
// sphere r, box |half extents|, capsule half-height + r, hull max |vertex| (0 for planes).
float rbgpu_world::bounding_radius(const rbgpu_collider_desc &d) const {
    double r = 0.0;
    switch (d.shape) {
        case RBGPU_SPHERE: r = d.dims[0]; break;
        case RBGPU_BOX: r = std::sqrt(d.dims[0] * d.dims[0] + d.dims[1] * d.dims[1] + d.dims[2] * d.dims[2]); break;
        case RBGPU_CAPSULE: r = d.dims[0] + d.dims[1]; break;
        case RBGPU_HULL:
            for (const auto &v : hh[d.hull_index].t.v) r = std::max(r, std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z));
            break;
        default: break;
    }
    return std::nextafter(std::nextafter((float)r, INFINITY), INFINITY);
}

int rbgpu_world::build() {
    // Keep broadphase history; bodies come from the descriptors (or the downloaded state).
    const uint32_t n = (uint32_t)hb.size(), m = (uint32_t)hc.size();
    std::vector<float4> x(n), q(n), v(n), w(n), f, t, I0(n), I1(n), Ii0(n), Ii1(n);
    body_forces(f, t);
    for (uint32_t i = 0; i < n; ++i) {
        const rbgpu_body_desc &d = hb[i];
        const bool dyn = d.type != RBGPU_STATIC;
        const float mass = (float)d.mass;
        const float im = dyn ? 1.0f / mass : 0.0f;
        SM I{(float)d.inertia[0], (float)d.inertia[1], (float)d.inertia[2], (float)d.inertia[3], (float)d.inertia[4],
             (float)d.inertia[5]};
        SM Ii = dyn ? sinverse(I) : SM{0, 0, 0, 0, 0, 0};
        x[i] = make_float4((float)d.position[0], (float)d.position[1], (float)d.position[2], im);
        q[i] = make_float4((float)d.orientation[0], (float)d.orientation[1], (float)d.orientation[2], (float)d.orientation[3]);
        v[i] = make_float4((float)d.linear_velocity[0], (float)d.linear_velocity[1], (float)d.linear_velocity[2], 0.0f);
        w[i] = make_float4((float)d.angular_velocity[0], (float)d.angular_velocity[1], (float)d.angular_velocity[2], 0.0f);
        I0[i] = make_float4(I.xx, I.yy, I.zz, I.xy);
        I1[i] = make_float4(I.xz, I.yz, 0.0f, 0.0f);
        Ii0[i] = make_float4(Ii.xx, Ii.yy, Ii.zz, Ii.xy);
        Ii1[i] = make_float4(Ii.xz, Ii.yz, dyn ? 1.0f : 0.0f, 0.0f);
    }
    std::vector<int4> meta(m);
    std::vector<float4> dims(m), lpos(m), lrot(m), mat(m);
    std::vector<int> chull(m);
    std::vector<uint2> planes;
    uint32_t slot = 0;
    for (uint32_t c = 0; c < m; ++c) {
        const rbgpu_collider_desc &d = hc[c];
        const bool plane = d.shape == RBGPU_PLANE;
        meta[c] = make_int4(d.shape, (int)d.body_id, d.is_static ? 1 : 0, plane ? (int)BP_INVALID : (int)slot);
        if (!plane) ++slot;
        dims[c] = make_float4((float)d.dims[0], (float)d.dims[1], (float)d.dims[2], (float)d.dims[3]);
        // w: radius of a sphere about the collider's origin that holds the shape (rounded up),
        // for the exact bounding-sphere rejection in front of GJK
        lpos[c] = make_float4((float)d.local_pos[0], (float)d.local_pos[1], (float)d.local_pos[2],
                              bounding_radius(d));
        lrot[c] = make_float4((float)d.local_rot[0], (float)d.local_rot[1], (float)d.local_rot[2], (float)d.local_rot[3]);
        mat[c] = make_float4((float)d.restitution, (float)d.static_friction, (float)d.dynamic_friction, 0.0f);
        chull[c] = d.shape == RBGPU_HULL ? d.hull_index : -1;
    }
    // hull tables, concatenated: vertices, polygon normals (rounded from double), index tables
    std::vector<float4> hvv, hpnv, hminv, hmaxv, hellv;
    std::vector<uint32_t> htabv;
    std::vector<HullMeta> hmv;
    for (const HostHull &h : hh) {
        HullMeta mt{};
        mt.v0 = (uint32_t)hvv.size();
        mt.nv = (uint32_t)h.t.v.size();
        mt.pn0 = (uint32_t)hpnv.size();
        mt.np = (uint32_t)h.t.pn.size();
        for (const auto &p : h.t.v) hvv.push_back(make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f));
        for (const auto &p : h.t.pn) hpnv.push_back(make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f));
        auto put = [&](const std::vector<uint32_t> &a) {
            const uint32_t o = (uint32_t)htabv.size();
            htabv.insert(htabv.end(), a.begin(), a.end());
            return o;
        };
        mt.poff = put(h.t.poff); mt.pidx = put(h.t.pidx); mt.v2p_o = put(h.t.v2p_o); mt.v2p_i = put(h.t.v2p_i);
        mt.v2n_o = put(h.t.v2n_o); mt.v2n_i = put(h.t.v2n_i); mt.p2p_o = put(h.t.p2p_o); mt.p2p_i = put(h.t.p2p_i);
        mt.adj = h.t.adjacency() ? 1u : 0u;
        hmv.push_back(mt);
        hminv.push_back(make_float4(h.amin[0], h.amin[1], h.amin[2], 0.0f));
        hmaxv.push_back(make_float4(h.amax[0], h.amax[1], h.amax[2], 0.0f));
        hellv.push_back(make_float4(h.ec[0], h.ec[1], h.ec[2], 0.0f));
        hellv.push_back(make_float4(h.ea[0], h.ea[1], h.ea[2], 0.0f));
    }
    // Plane loop of the narrowphase driver (CollisionPipeline.cpp:156-197): constant per scene;
    // a plane pairs with the colliders of its own scene.  plane_lid: index of the pair among
    // its scene's plane pairs.
    std::vector<uint32_t> csc(m), ssc(S), plids;
    uint32_t nsc = 1;
    for (uint32_t c = 0; c < m; ++c) {
        csc[c] = hb[hc[c].body_id].scene;
        nsc = std::max(nsc, csc[c] + 1);
        if ((uint32_t)meta[c].w != BP_INVALID) ssc[(uint32_t)meta[c].w] = csc[c];
    }
    for (const rbgpu_body_desc &d : hb) nsc = std::max(nsc, d.scene + 1);
    if (nsc > 65536) return fail(RBGPU_ERR_CAPACITY, "more than 65536 scenes in one world");
    {
        std::vector<std::vector<uint32_t>> by_scene(nsc);
        for (uint32_t c = 0; c < m; ++c) by_scene[csc[c]].push_back(c);
        std::vector<uint32_t> nplane(nsc, 0);
        for (uint32_t pi = 0; pi < m; ++pi) {
            if (hc[pi].shape != RBGPU_PLANE) continue;
            for (uint32_t di : by_scene[csc[pi]]) {
                if (di == pi || hc[di].shape == RBGPU_PLANE) continue;
                if (hc[pi].is_static && hc[di].is_static) continue;
                if (hc[pi].body_id == hc[di].body_id) continue;
                planes.push_back(make_uint2(di, pi));
                plids.push_back(nplane[csc[pi]]++);
            }
        }
    }
    if (slot != S) return fail(RBGPU_ERR_INVALID, "broadphase slot bookkeeping mismatch");
    if (S > (1u << 26)) return fail(RBGPU_ERR_INVALID, "more than 2^26 broadphase colliders");   // sweep rows: 16 threads each

    release();
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void **)&rb, sizeof(Readback), hipHostMallocDefault));
    N = n; M = m; NEP = 2 * S;
    int rc = 0;
    rc |= dalloc(B.r, n); rc |= dalloc(B.f, n); rc |= dalloc(B.t, n); rc |= dalloc(B.I0, n); rc |= dalloc(B.I1, n);
    rc |= dalloc(C.meta, m); rc |= dalloc(C.dims, m); rc |= dalloc(C.lpos, m); rc |= dalloc(C.lrot, m); rc |= dalloc(C.mat, m);
    rc |= dalloc(C.hull, m);
    rc |= dalloc(hv, hvv.size()); rc |= dalloc(hpn, hpnv.size()); rc |= dalloc(htab, htabv.size());
    rc |= dalloc(hmeta, hmv.size()); rc |= dalloc(hamin, hminv.size()); rc |= dalloc(hamax, hmaxv.size());
    rc |= dalloc(hell, hellv.size());
    rc |= dalloc(fmin, S); rc |= dalloc(fmax, S); rc |= dalloc(uid, S);
    rc |= dalloc(ep_key[0], NEP); rc |= dalloc(ep_key[1], NEP); rc |= dalloc(ep_val[0], NEP); rc |= dalloc(ep_val[1], NEP);
    rc |= dalloc(ep_pos, S); rc |= dalloc(sw_a, NEP); rc |= dalloc(sw_b, NEP);
    rc |= dalloc(ep_cnt, NEP + 1); rc |= dalloc(ep_off, NEP + 1);
    rc |= dalloc(ep_hist, (size_t)((NEP + RS_TILE - 1) / RS_TILE) * RS_BINS);
    rc |= dalloc(flags, 4); rc |= dalloc(scal, 8);
    rc |= dalloc(tile_sums, SCAN_TOP_THREADS * SCAN_TOP_ITEMS);
    rc |= dalloc(cb.used, (size_t)n * COLOUR_WORDS); rc |= dalloc(cb.ucnt, n);
    rc |= dalloc(cb.claim, (size_t)n * COLOUR_SLOTS); rc |= dalloc(cb.hist, MAX_COLOURS);
    rc |= dalloc(cb.off, MAX_COLOURS + 1); rc |= dalloc(cb.cur, MAX_COLOURS); rc |= dalloc(cb.ctr, CTR_SIZE);
    rc |= dalloc(cb.epoch, 1);
    rc |= dalloc(plane_pairs, planes.size()); rc |= dalloc(offs, NBINS + 4);
    rc |= dalloc(cscene, m); rc |= dalloc(slot_scene, S); rc |= dalloc(sap_start, nsc); rc |= dalloc(sap_end, nsc);
    rc |= dalloc(plane_lid, planes.size());
    rc |= dalloc(d_points, 1);
    if (rc) return RBGPU_ERR_OOM;
    auto up = [&](void *dst, const void *src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st); };
    if (n) {
        std::vector<BodyRow> rows(n);
        for (uint32_t i = 0; i < n; ++i) rows[i] = BodyRow{x[i], q[i], v[i], w[i], x[i], q[i], Ii0[i], Ii1[i]};
        HIPCHK(up(B.r, rows.data(), n * sizeof(BodyRow)));
        HIPCHK(hipStreamSynchronize(st));   // `rows` leaves scope
        HIPCHK(up(B.f, f.data(), n * 16)); HIPCHK(up(B.t, t.data(), n * 16)); HIPCHK(up(B.I0, I0.data(), n * 16));
        HIPCHK(up(B.I1, I1.data(), n * 16));
    }
    if (m) {
        HIPCHK(up(C.meta, meta.data(), m * 16)); HIPCHK(up(C.dims, dims.data(), m * 16));
        HIPCHK(up(C.lpos, lpos.data(), m * 16)); HIPCHK(up(C.lrot, lrot.data(), m * 16)); HIPCHK(up(C.mat, mat.data(), m * 16));
        HIPCHK(up(C.hull, chull.data(), m * 4));
    }
    if (!hmv.empty()) {
        HIPCHK(up(hv, hvv.data(), hvv.size() * 16));
        if (!hpnv.empty()) HIPCHK(up(hpn, hpnv.data(), hpnv.size() * 16));
        if (!htabv.empty()) HIPCHK(up(htab, htabv.data(), htabv.size() * 4));
        HIPCHK(up(hmeta, hmv.data(), hmv.size() * sizeof(HullMeta)));
        HIPCHK(up(hamin, hminv.data(), hminv.size() * 16));
        HIPCHK(up(hamax, hmaxv.data(), hmaxv.size() * 16));
        HIPCHK(up(hell, hellv.data(), hellv.size() * 16));
    }
    hs = HullSet{hv, hpn, htab, hmeta, hamin, hamax, hell};
    hull_vpl = 1;
    for (const HostHull &h : hh)
        if (h.t.v.size() > 64) hull_vpl = HULL_VPL;
    if (S) {
        HIPCHK(up(fmin, h_fmin.data(), S * 16)); HIPCHK(up(fmax, h_fmax.data(), S * 16)); HIPCHK(up(uid, h_uid.data(), S * 4));
        HIPCHK(up(ep_key[0], h_key.data(), NEP * 4)); HIPCHK(up(ep_val[0], h_val.data(), NEP * 4));
        HIPCHK(up(ep_pos, h_eppos.data(), S * 8));
    }
    n_plane = (uint32_t)planes.size();
    if (n_plane) HIPCHK(up(plane_pairs, planes.data(), n_plane * 8));
    if (n_plane) HIPCHK(up(plane_lid, plids.data(), n_plane * 4));
    nscenes = nsc;
    {   // per-scene tables: plane-pair lists (constant per build), slice sizes and offsets, colours
        std::vector<uint32_t> poff(nsc + 1, 0), pidx(n_plane), nb(nsc, 0);
        for (uint32_t j = 0; j < n_plane; ++j) ++poff[csc[planes[j].y] + 1];
        for (uint32_t k = 0; k < nsc; ++k) poff[k + 1] += poff[k];
        std::vector<uint32_t> cur(poff.begin(), poff.end() - 1);
        for (uint32_t j = 0; j < n_plane; ++j) pidx[cur[csc[planes[j].y]]++] = j;
        for (const rbgpu_body_desc &d : hb) ++nb[d.scene];
        max_scene_bodies = nsc ? *std::max_element(nb.begin(), nb.end()) : 0;
        int rs = 0;
        rs |= dalloc(sc_plane_off, nsc + 1); rs |= dalloc(sc_plane_idx, std::max<uint32_t>(n_plane, 1));
        rs |= dalloc(sc_size, nsc + 1); rs |= dalloc(sc_base, nsc + 1); rs |= dalloc(sc_ncol, nsc);
        rs |= dalloc(sc_coff, (size_t)nsc * (MAX_COLOURS + 1)); rs |= dalloc(sc_err, 1);
        // the scenes' bodies in ascending order (the per-scene solve's velocity update, the
        // colouring's local body ids); the colouring takes the scenes largest first, so that
        // the longest workgroups start early
        std::vector<uint32_t> boff(nsc + 1, 0), bidx(N), bloc(N), sord(nsc);
        for (uint32_t k = 0; k < nsc; ++k) boff[k + 1] = boff[k] + nb[k];
        {
            std::vector<uint32_t> bcur(boff.begin(), boff.end() - 1);
            for (uint32_t i = 0; i < N; ++i) {
                bloc[i] = bcur[hb[i].scene] - boff[hb[i].scene];
                bidx[bcur[hb[i].scene]++] = i;
            }
        }
        for (uint32_t k = 0; k < nsc; ++k) sord[k] = k;
        std::stable_sort(sord.begin(), sord.end(), [&](uint32_t a, uint32_t b) { return nb[a] > nb[b]; });
        rs |= dalloc(sb_off, nsc + 1); rs |= dalloc(sb_idx, std::max<uint32_t>(N, 1));
        rs |= dalloc(sb_loc, std::max<uint32_t>(N, 1)); rs |= dalloc(sc_order, nsc);
        if (rs) return RBGPU_ERR_OOM;
        HIPCHK(up(sc_plane_off, poff.data(), (nsc + 1) * 4));
        if (n_plane) HIPCHK(up(sc_plane_idx, pidx.data(), n_plane * 4));
        HIPCHK(up(sb_off, boff.data(), (nsc + 1) * 4));
        if (N) HIPCHK(up(sb_idx, bidx.data(), N * 4));
        if (N) HIPCHK(up(sb_loc, bloc.data(), N * 4));
        HIPCHK(up(sc_order, sord.data(), nsc * 4));
        HIPCHK(hipMemsetAsync(sc_err, 0, 4, st));
    }
    if (m) HIPCHK(up(cscene, csc.data(), m * 4));
    if (S) HIPCHK(up(slot_scene, ssc.data(), S * 4));
    int rcj = build_joints();
    if (rcj) return rcj;
    int init_flags[4] = {1, 0, 0, 0};   // dirty: the endpoint array has never been sorted
    HIPCHK(up(flags, init_flags, sizeof(init_flags)));
    HIPCHK(hipMemsetAsync(scal, 0, 32, st));
    HIPCHK(hipMemsetAsync(cb.ctr, 0, CTR_SIZE * 4, st));
    HIPCHK(hipMemsetAsync(d_points, 0, 8, st));
    HIPCHK(hipStreamSynchronize(st));
    int rc2 = ensure_pair_capacity(std::max<uint32_t>(8 * S + n_plane, 4096));
    if (rc2) return rc2;
    built = true;
    return RBGPU_OK;
}

// Joint upload and colouring.  Colouring: first fit in joint order, conflict = a shared
// dynamic body (computed once per upload; the joint graph does not change while stepping).
int rbgpu_world::build_joints() {
    NJ = (uint32_t)hj.size();
    h_jtg.resize(NJ);
    h_jlam.resize((size_t)NJ * 4, 0.0f);
    std::vector<int4> jm(NJ);
    std::vector<float4> jr1(NJ), jr2(NJ), jax(NJ), jlx(NJ), jrow((size_t)NJ * 4, make_float4(1.0f, 0.0f, 0.0f, 0.0f));
    h_jcolour.assign(NJ, 0);
    std::vector<std::vector<int32_t>> body_cols(hb.size());
    int ncolj = 0;
    for (uint32_t i = 0; i < NJ; ++i) {
        const rbgpu_joint_desc &d = hj[i];
        const bool row = d.type == RBGPU_JOINT_ROW;
        int rows = d.type == RBGPU_JOINT_PRISMATIC ? 3 : d.type == RBGPU_JOINT_REVOLUTE ? 4
                 : d.type == RBGPU_JOINT_FIXED ? 2 : row ? 1 : (d.limited ? 2 : 1);
        const bool fixed = d.type == RBGPU_JOINT_FIXED || row;   // a row has no axes, limits or damping
        const int act = d.type == RBGPU_JOINT_FIXED ? RBGPU_ACT_FREE : d.actuation;   // a row: its kind
        const int lim = fixed ? 0 : (d.limited ? 1 : 0);
        if (row) jrow[(size_t)i * 4] = h_jrow_user[i];
        jm[i] = make_int4(d.type | (act << 8) | (lim << 16) | (rows << 24), (int)d.body_1, (int)d.body_2, 0);
        jr1[i] = make_float4((float)d.r_1[0], (float)d.r_1[1], (float)d.r_1[2], fixed ? 0.0f : (float)d.damping);
        jr2[i] = make_float4((float)d.r_2[0], (float)d.r_2[1], (float)d.r_2[2], (float)d.compliance);
        const bool lax = d.type == RBGPU_JOINT_REVOLUTE;   // the others carry a zero limit axis (JointAPI.hpp)
        jax[i] = make_float4(fixed ? 0.0f : (float)d.axis[0], fixed ? 0.0f : (float)d.axis[1], fixed ? 0.0f : (float)d.axis[2],
                             fixed ? 0.0f : (float)d.lower_limit);
        jlx[i] = make_float4(lax ? (float)d.limit_axis[0] : 0.0f, lax ? (float)d.limit_axis[1] : 0.0f,
                             lax ? (float)d.limit_axis[2] : 0.0f, fixed ? 0.0f : (float)d.upper_limit);
        const uint32_t bs[2] = {d.body_1, d.body_2};
        int c = 0;
        for (;; ++c) {
            bool taken = false;
            for (uint32_t b : bs)
                if (hb[b].type == RBGPU_DYNAMIC)
                    for (int32_t u : body_cols[b]) taken = taken || u == c;
            if (!taken) break;
        }
        h_jcolour[i] = c;
        for (uint32_t b : bs)
            if (hb[b].type == RBGPU_DYNAMIC) body_cols[b].push_back(c);
        ncolj = std::max(ncolj, c + 1);
    }
    std::vector<uint32_t> order(NJ);
    for (uint32_t i = 0; i < NJ; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return h_jcolour[a] < h_jcolour[b]; });
    jcol_off.assign(ncolj + 1, 0);
    for (uint32_t i = 0; i < NJ; ++i) ++jcol_off[h_jcolour[i] + 1];
    for (int c = 0; c < ncolj; ++c) jcol_off[c + 1] += jcol_off[c];
    int rc = 0;
    rc |= dalloc(Jt.jm, NJ); rc |= dalloc(Jt.jr1, NJ); rc |= dalloc(Jt.jr2, NJ); rc |= dalloc(Jt.jax, NJ);
    rc |= dalloc(Jt.jlx, NJ); rc |= dalloc(Jt.jtg, NJ); rc |= dalloc(Jt.jrow, (size_t)NJ * 4);
    rc |= dalloc(Jt.jlam, (size_t)NJ * 4); rc |= dalloc(jorder, NJ);
    if (rc) return RBGPU_ERR_OOM;
    if (!NJ) return RBGPU_OK;
    auto up = [&](void *dst, const void *src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st); };
    HIPCHK(up(Jt.jm, jm.data(), NJ * 16)); HIPCHK(up(Jt.jr1, jr1.data(), NJ * 16)); HIPCHK(up(Jt.jr2, jr2.data(), NJ * 16));
    HIPCHK(up(Jt.jax, jax.data(), NJ * 16)); HIPCHK(up(Jt.jlx, jlx.data(), NJ * 16));
    HIPCHK(up(Jt.jtg, h_jtg.data(), NJ * 16)); HIPCHK(up(Jt.jrow, jrow.data(), (size_t)NJ * 64));
    HIPCHK(up(Jt.jlam, h_jlam.data(), (size_t)NJ * 16)); HIPCHK(up(jorder, order.data(), NJ * 4));
    // per-scene joint tables for the per-scene solves: the scene's joints by colour (the global
    // order restricted to the scene); only when every joint stays inside one scene
    bool inside = nscenes > 1;
    std::vector<uint32_t> jsc(NJ);
    for (uint32_t i = 0; i < NJ && inside; ++i) {
        jsc[i] = hb[hj[i].body_1].scene;
        inside = hb[hj[i].body_2].scene == jsc[i];
    }
    if (inside) {
        std::vector<uint32_t> so(order);
        std::stable_sort(so.begin(), so.end(), [&](uint32_t a, uint32_t b) { return jsc[a] < jsc[b]; });
        std::vector<uint32_t> base(nscenes + 1, 0), coff((size_t)nscenes * (ncolj + 1), 0);
        for (uint32_t i = 0; i < NJ; ++i) {
            ++base[jsc[i] + 1];
            ++coff[(size_t)jsc[i] * (ncolj + 1) + h_jcolour[i] + 1];
        }
        for (uint32_t k = 0; k < nscenes; ++k) {
            base[k + 1] += base[k];
            for (int c = 0; c < ncolj; ++c) coff[(size_t)k * (ncolj + 1) + c + 1] += coff[(size_t)k * (ncolj + 1) + c];
        }
        rc = 0;
        rc |= dalloc(sj_order, NJ); rc |= dalloc(sj_base, nscenes + 1); rc |= dalloc(sj_coff, coff.size());
        if (rc) return RBGPU_ERR_OOM;
        HIPCHK(up(sj_order, so.data(), NJ * 4)); HIPCHK(up(sj_base, base.data(), (nscenes + 1) * 4));
        HIPCHK(up(sj_coff, coff.data(), coff.size() * 4));
        sj_njc = (uint32_t)ncolj;
    }
    HIPCHK(hipStreamSynchronize(st));
    return RBGPU_OK;
}

// Pull the evolving state back into the host descriptors and broadphase mirrors so that
// a later build() (after bodies or colliders were added) continues from it.
int rbgpu_world::download_state() {
    if (!built) return RBGPU_OK;
    HIPCHK(hipStreamSynchronize(st));
    std::vector<BodyRow> rows;
    HIPCHK(body_rows(rows, B.r, N));
    for (uint32_t i = 0; i < N; ++i) {
        rbgpu_body_desc &d = hb[i];
        const BodyRow &r = rows[i];
        d.position[0] = r.x.x; d.position[1] = r.x.y; d.position[2] = r.x.z;
        d.orientation[0] = r.q.x; d.orientation[1] = r.q.y; d.orientation[2] = r.q.z; d.orientation[3] = r.q.w;
        d.linear_velocity[0] = r.v.x; d.linear_velocity[1] = r.v.y; d.linear_velocity[2] = r.v.z;
        d.angular_velocity[0] = r.w.x; d.angular_velocity[1] = r.w.y; d.angular_velocity[2] = r.w.z;
    }
    if (S) {
        HIPCHK(hipMemcpy(h_fmin.data(), fmin, S * 16, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_fmax.data(), fmax, S * 16, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_key.data(), ep_key[0], NEP * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_val.data(), ep_val[0], NEP * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_eppos.data(), ep_pos, S * 8, hipMemcpyDeviceToHost));
    }
    if (NJ) {
        HIPCHK(hipMemcpy(h_jtg.data(), Jt.jtg, NJ * 16, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h_jlam.data(), Jt.jlam, (size_t)NJ * 16, hipMemcpyDeviceToHost));
    }
    // a pending dirty flag is re-established by build()
    return RBGPU_OK;
}

// ----------------------------------------------------------------- one frame
int rbgpu_world::frame() {
    const float dt = (float)timestep;
    const float h = dt / (float)substeps;
    const float inv_h = 1.0f / h;
    int *dirty = flags, *overflow = flags + 1;

    // 1. refit (CollisionPipeline.cpp:14-53)
    hipEvent_t *ebp = timed_begin(6);
    if (M) k_refit<<<nblk(M), TPB, 0, st>>>(C, B, hs, M, dt, margin, reinsert, fmin, fmax, ep_key[0], ep_pos, dirty);
    // 2. stable sort of the endpoints if any fat box changed (BroadPhase.cpp:234-245)
    if (NEP) {
        for (int pass = 0; pass < 4; ++pass) {
            int s = pass & 1;
            int rc = radix_pass(ep_key[s], ep_val[s], ep_key[s ^ 1], ep_val[s ^ 1], NEP, 8 * pass, ep_hist, dirty);
            if (rc) return rc;
        }
        if (nscenes > 1) {   // then by scene (stable): (scene, value) order, scenes contiguous
            k_ep_scene_key<<<nblk(NEP), TPB, 0, st>>>(ep_val[0], slot_scene, NEP, ep_cnt, dirty);
            int rc = radix_pass(ep_cnt, ep_val[0], ep_key[1], ep_val[1], NEP, 0, ep_hist, dirty);
            if (!rc) rc = radix_pass(ep_key[1], ep_val[1], ep_cnt, ep_val[0], NEP, 8, ep_hist, dirty);
            if (rc) return rc;
        }
        k_sweep_prep<<<nblk(NEP), TPB, 0, st>>>(ep_val[0], NEP, fmin, fmax, ep_pos, uid, C, sw_a, sw_b, ep_key[0],
                                                dirty);
        // 3. sweep: count, scan, readback of the pair count
        k_sweep<false><<<nblk(S * SWEEP_ROW), TPB, 0, st>>>(sw_a, sw_b, ep_pos, S, ep_cnt, nullptr, nullptr, 0,
                                                            overflow);
        int rc = scan(ep_cnt, ep_off, NEP, scal + 0);
        if (rc) return rc;
        k_set_int<<<1, 1, 0, st>>>(dirty, 0);
    } else {
        HIPCHK(hipMemsetAsync(scal, 0, 4, st));
    }
    // 3. pair list: SAP pairs (canonical order) then the plane pairs; per-pair data and
    //    narrowphase bins.  One host synchronisation per frame: the pair count (capacity
    //    check) and the bin offsets.  The GJK slots' rejection hints are saved by collider
    //    pair first and restored into the new list's slots after it.
    //    After a snapshot restore the table comes from the snapshot instead.
    const bool restored = ax_restore >= 0;
    const uint32_t ng_old = !restored && n_pairs && !hh.empty() && bin_off.size() == NBINS + 1
                                ? bin_off[BIN_HP] - bin_off[BIN_HH] : 0u;
    bool carry = axis_carry && gjk_cull && gjk_axis && ng_old;
    uint32_t amask = 0;
    if (restored) {
        carry = axis_carry && gjk_cull && ax_restore > 0;
        amask = ax_restore > 0 ? (uint32_t)ax_restore - 1 : 0u;
        ax_restore = -1;
    } else if (carry) {
        uint32_t H = 1;
        while (H < 2 * ng_old) H <<= 1;
        if (H > ax_cap) {
            dfree(ax_keys); dfree(ax_vals);
            ax_cap = 0;
            if (dalloc(ax_keys, H) | dalloc(ax_vals, H)) return RBGPU_ERR_OOM;
            ax_cap = H;
        }
        amask = H - 1;
        k_axis_clear<<<nblk(H), TPB, 0, st>>>(ax_keys, H);
        k_axis_save<<<nblk(ng_old), TPB, 0, st>>>(bin_list + bin_off[BIN_HH], ng_old, pcap, K, gjk_axis, ax_keys, ax_vals,
                                                   amask);
    }
    for (int attempt = 0;; ++attempt) {
        HIPCHK(hipMemsetAsync(overflow, 0, 4, st));
        if (NEP)
            k_sweep<true><<<nblk(S * SWEEP_ROW), TPB, 0, st>>>(sw_a, sw_b, ep_pos, S, nullptr, ep_off, pairs, pcap,
                                                               overflow);
        k_append<<<nblk(std::max<uint32_t>(n_plane, 1)), TPB, 0, st>>>(pairs, scal + 0, plane_pairs, n_plane, pcap,
                                                                        scal + 1, overflow);
        k_pair_gather<<<nblk(pcap), TPB, 0, st>>>(pairs, scal + 1, C, K, ka, va);
        // local pair ids (colouring hash input): a scene's SAP pairs are one run
        HIPCHK(hipMemsetAsync(sap_start, 0, nscenes * 4, st));
        HIPCHK(hipMemsetAsync(sap_end, 0, nscenes * 4, st));
        k_scene_bounds<<<nblk(pcap), TPB, 0, st>>>(pairs, scal + 0, cscene, sap_start, sap_end);
        k_pair_lid<<<nblk(pcap), TPB, 0, st>>>(pairs, scal + 0, scal + 1, cscene, sap_start, sap_end, plane_lid, pair_lid);
        int rc = radix_pass(ka, va, kb, bin_list, pcap, 0, p_hist, nullptr, scal + 1);
        if (rc) return rc;
        k_offsets<<<1, 64, 0, st>>>(p_hist, (pcap + RS_TILE - 1) / RS_TILE, NBINS, scal + 1, offs);
        HIPCHK(hipMemcpyAsync(&rb->n_sap, scal, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&rb->overflow, overflow, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(rb->bin_off, offs, (NBINS + 1) * 4, hipMemcpyDeviceToHost, st));
        if (sc_err) HIPCHK(hipMemcpyAsync(&rb->scene_err, sc_err, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        err_tail_queued = false;   // the flag read just now covers the earlier frames
        // errors of the per-scene colouring (sticky until reported here, one frame after the fact)
        if (sc_err && rb->scene_err) {
            const uint32_t e = rb->scene_err;
            rb->scene_err = 0;
            colour_clean = false;
            HIPCHK(hipMemsetAsync(sc_err, 0, 4, st));
            HIPCHK(hipStreamSynchronize(st));
            if (e & 1u) return fail(RBGPU_ERR_CAPACITY, "more than 1024 colours needed");
            return fail(RBGPU_ERR_CAPACITY, "colouring did not converge");
        }
        if (!rb->overflow) break;
        if (attempt > 0) return fail(RBGPU_ERR_CAPACITY, "pair buffer overflow");
        int rc2 = ensure_pair_capacity(rb->n_pairs);
        if (rc2) return rc2;
    }
    timed_end(ebp);
    n_sap = rb->n_sap;
    n_pairs = rb->n_pairs;
    bin_off.assign(rb->bin_off, rb->bin_off + NBINS + 1);
    if (gjk_rec && gjk_cull && bin_off[BIN_HP] > bin_off[BIN_HH])   // the frame's GJK-bin records
        k_gjk_records<<<nblk(bin_off[BIN_HP] - bin_off[BIN_HH]), TPB, 0, st>>>(bin_list + bin_off[BIN_HH],
                                                                                 bin_off[BIN_HP] - bin_off[BIN_HH], K,
                                                                                 gjk_rec, gjk_recb);
    if (gjk_axis && bin_off[BIN_HP] > bin_off[BIN_HH])
        k_axis_load<<<nblk(bin_off[BIN_HP] - bin_off[BIN_HH]), TPB, 0, st>>>(bin_list + bin_off[BIN_HH],
                                                                               bin_off[BIN_HP] - bin_off[BIN_HH], K,
                                                                               carry ? ax_keys : nullptr, ax_vals, amask,
                                                                               gjk_axis);
    const uint32_t P = n_pairs;
    const uint32_t colour_rounds_cap = 1u << 16;
    const bool by_scene = scene_path();
    const SceneTabs T = scene_tabs();
    if (by_scene && P) {   // slice offsets of the scenes' pair lists
        k_scene_sizes<<<nblk(nscenes), TPB, 0, st>>>(T, nscenes, sc_size);
        int rcs = scan(sc_size, sc_base, nscenes, nullptr);
        if (rcs) return rcs;
    }

    // 4. substeps (World.hpp:203-211).  One substep's launches; with the per-scene kernels
    //    nothing in it waits for the host, so it is captured once per frame into a hipGraph
    //    and the graph is launched for every substep.
    cb.pal_max = palette_max;
    cb.lid = pair_lid;
    // joints inside the per-scene solves (which run only when the frame has pairs)
    const bool scene_joints = by_scene && P && NJ && sj_order;
    SceneTabs TJ = T;
    if (!scene_joints) TJ.jorder = nullptr;
    colour_fast = by_scene && colour_clean;   // this frame's colouring-state clear: counters only
    // On the per-scene path the body passes ride in the per-scene kernels: the integration in
    // the colouring's workgroups (it is independent of the colouring), the velocity update and
    // the velocity solve in the position solve's (a scene's bodies are touched by its own
    // pairs and joints only) -- unless one of those launches is timed this frame.
    const bool timed_frame = timing_kernel != 0 && frames % timing_stride == 0;
    auto timed_k = [&](std::initializer_list<int> ks) {
        for (int k : ks)
            if (timed_frame && timing_kernel == k) return true;
        return false;
    };
    const bool fold_int = by_scene && P && N && !timed_k({4, 7});
    const bool fold_vel = by_scene && P && N && (NJ == 0 || scene_joints) && solve_occ != 4 && !timed_k({1, 2, 5});
    auto substep = [&]() -> int {
        runs.clear();
        if (P) {
            hipEvent_t *e = timed_begin(3);
            int rcn = launch_narrowphase(true);   // with the colouring state cleared beside it
            if (rcn) return rcn;
            timed_end(e);
            // colouring of the pairs with points (conflict = a shared dynamic body)
            hipEvent_t *ec = timed_begin(7);
            if (by_scene) {   // per scene: rounds, then the colour-ordered lists (no host round trip)
                k_scene_colour<<<nscenes, SCENE_THREADS, 0, st>>>(K.cnt, K.bb, scal + 0, B, cb, T, colour_rounds_cap,
                                                                  fold_int ? N : 0u, h);
                timed_end(ec);
            } else {
            k_colour_activate<<<nblk(P), TPB, 0, st>>>(K.cnt, P, K.bb, B, cb);
            int rcs = scan(cb.blk, cb.blk, nblk(P), cb.ctr + CTR_ACT);
            if (rcs) return rcs;
            k_colour_compact<<<nblk(P), TPB, 0, st>>>(P, cb);
            const uint32_t G = std::min<uint32_t>(nblk(P), 1024);
            // grid-wide rounds in batches (they skip themselves once at most tail_min pairs are
            // left); a host check between batches
            uint32_t launched = 0;
            for (;;) {
                const uint32_t end = std::min(head_rounds, launched + round_batch);
                for (; launched < end; ++launched) {
                    k_colour_claim<<<G, TPB, 0, st>>>(K.bb, B, cb, launched, tail_min);
                    k_colour_settle<<<G, TPB, 0, st>>>(K.bb, B, cb, launched);
                }
                if (launched >= head_rounds) break;
                HIPCHK(hipMemcpyAsync(rb->ctr, cb.ctr, CTR_SIZE * 4, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                uint32_t left = 0;
                for (uint32_t q = 0; q < LEFT_STRIPES; ++q) left += rb->ctr[CTR_LEFT + q];
                if (left <= tail_min || !rb->ctr[CTR_FLAGS + launched - 1]) break;
            }
            k_colour_collect<<<G, TPB, 0, st>>>(cb);
            k_colour_tail<<<1, TAIL_THREADS, 0, st>>>(K.bb, B, cb, launched, colour_rounds_cap);
            const uint32_t hist_blocks = std::min<uint32_t>(nblk(P), 512);
            k_colour_hist<<<hist_blocks, 256, 0, st>>>(cb);
            k_colour_offsets<<<1, MAX_COLOURS, 0, st>>>(cb);
            const uint32_t chunk = 4096;
            k_colour_scatter<<<nblk(P, chunk), 256, 0, st>>>(cb, solve_list, chunk);
            HIPCHK(hipMemcpyAsync(rb->ctr, cb.ctr, CTR_SIZE * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(rb->col_off, cb.off, (MAX_COLOURS + 1) * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (rb->ctr[CTR_ERR] & 1u) return fail(RBGPU_ERR_CAPACITY, "more than 1024 colours needed");
            if (rb->ctr[CTR_ERR] & 2u) return fail(RBGPU_ERR_CAPACITY, "colouring did not converge");
            ncol = rb->ctr[CTR_NCOL];
            col_off.assign(rb->col_off, rb->col_off + ncol + 1);
            runs.clear();   // [c0, c1, single-workgroup?]
            for (uint32_t c = 0; c < ncol;) {
                if (col_off[c + 1] - col_off[c] >= run_min) { runs.push_back({c, c + 1, 0}); ++c; continue; }
                uint32_t c1 = c + 1;
                while (c1 < ncol && col_off[c1 + 1] - col_off[c1] < run_min) ++c1;
                runs.push_back({c, c1, 1});
                c = c1;
            }
            timed_end(ec);
            }
        }
        if (N && !fold_int) {
            hipEvent_t *e = timed_begin(4);
            k_integrate<<<nblk(N), TPB, 0, st>>>(B, N, h);
            timed_end(e);
        }
        if (fold_vel) {   // positions, velocities, velocity solve: one launch
            k_scene_solve<true, 1, true><<<nscenes, scene_threads, 0, st>>>(B, K, TJ, h, nullptr, Jt, inv_h, joint_mode);
            return RBGPU_OK;
        }
        if (by_scene && P) {
            hipEvent_t *e = timed_begin(1);
            if (solve_occ == 4)
                k_scene_solve<true, 4><<<nscenes, scene_threads, 0, st>>>(B, K, TJ, h, e ? d_points : nullptr, Jt, inv_h,
                                                                         joint_mode);
            else
                k_scene_solve<true><<<nscenes, scene_threads, 0, st>>>(B, K, TJ, h, e ? d_points : nullptr, Jt, inv_h,
                                                                      joint_mode);
            timed_end(e);
        }
        for (const Run &u : runs) {
            if (u.group) {
                hipEvent_t *e = timed_begin(8);
                k_solve_pos_run<<<1, TAIL_THREADS, 0, st>>>(B, K, solve_list, cb.off, u.c0, u.c1);
                timed_end(e);
            } else {
                const uint32_t o = col_off[u.c0], n = col_off[u.c1] - o;
                hipEvent_t *e = timed_begin(1);
                k_solve_pos<<<nblk(n), TPB, 0, st>>>(B, K, solve_list, o, n, e ? d_points : nullptr);
                timed_end(e);
            }
        }
        // joints: errors from the state after the contact solve, then rows per colour
        // (World.hpp:166-171); on the per-scene path they run inside k_scene_solve
        if (NJ && !scene_joints) {
            hipEvent_t *e = timed_begin(10);
            k_joint_errors<<<nblk(NJ), TPB, 0, st>>>(B, Jt, NJ, h, joint_mode);
            for (size_t c = 0; c + 1 < jcol_off.size(); ++c) {
                const uint32_t o = jcol_off[c], n = jcol_off[c + 1] - o;
                if (n) k_joint_solve<<<nblk(n), TPB, 0, st>>>(B, Jt, jorder, o, n, inv_h, h, joint_mode);
            }
            timed_end(e);
        }
        if (N) {
            hipEvent_t *e = timed_begin(5);
            k_update_vel<<<nblk(N), TPB, 0, st>>>(B, N, inv_h);
            timed_end(e);
        }
        if (by_scene && P) {
            hipEvent_t *e = timed_begin(2);
            k_scene_solve<false><<<nscenes, scene_threads, 0, st>>>(B, K, TJ, h, e ? d_points : nullptr, Jt, inv_h,
                                                                   joint_mode);
            timed_end(e);
        }
        for (const Run &u : runs) {
            if (u.group) {
                hipEvent_t *e = timed_begin(9);
                k_solve_vel_run<<<1, TAIL_THREADS, 0, st>>>(B, K, solve_list, cb.off, u.c0, u.c1, h);
                timed_end(e);
            } else {
                const uint32_t o = col_off[u.c0], n = col_off[u.c1] - o;
                hipEvent_t *e = timed_begin(2);
                k_solve_vel<<<nblk(n), TPB, 0, st>>>(B, K, solve_list, o, n, h, e ? d_points : nullptr);
                timed_end(e);
            }
        }
        // joint damping (World.hpp:174-178)
        if (NJ && !scene_joints) {
            hipEvent_t *e = timed_begin(10);
            for (size_t c = 0; c + 1 < jcol_off.size(); ++c) {
                const uint32_t o = jcol_off[c], n = jcol_off[c + 1] - o;
                if (n) k_joint_damp<<<nblk(n), TPB, 0, st>>>(B, Jt, jorder, o, n, h);
            }
            timed_end(e);
        }
        return RBGPU_OK;
    };
    const bool timed = timing_kernel != 0 && frames % timing_stride == 0;   // per-launch events
    if (use_graph && by_scene && !timed && !ds_logging && substeps > 1) {
        int rc = run_graph(substep);
        if (rc) return rc;
    } else {
        for (int s = 0; s < substeps; ++s) {
            int rc = substep();
            if (rc) return rc;
        }
    }
    // colouring state for the next frame: after per-scene frames only the counters need a
    // clear (the claims' epoch rises by one per substep); a full clear before the epoch wraps
    // and after the grid-wide path (its claims carry no epoch, its masks are not cleared)
    if (P && by_scene) {
        colour_epochs = colour_fast ? colour_epochs + (uint32_t)substeps : 1u;
        colour_clean = colour_epochs + 2u * (uint32_t)substeps < COLOUR_EPOCH_MAX;
    } else if (P) {
        colour_clean = false;
    }
    ++frames;
    return RBGPU_OK;
}

// ----------------------------------------------------------------- C-ABI
extern "C" {

int rbgpu_world_create(double timestep, int substeps, rbgpu_world **out) {
    if (!out || substeps <= 0 || !(timestep > 0.0)) return fail(RBGPU_ERR_INVALID, "invalid timestep/substeps");
    auto *w = new rbgpu_world();
    w->timestep = timestep;
    w->substeps = substeps;
    *out = w;
    return RBGPU_OK;
}

int rbgpu_world_destroy(rbgpu_world *w) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    delete w;
    return RBGPU_OK;
}

int rbgpu_add_bodies(rbgpu_world *w, int n, const rbgpu_body_desc *desc, uint32_t *ids_out) {
    if (!w || n < 0 || (n && !desc)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (w->hb.size() + (size_t)n >= (1u << 24)) return fail(RBGPU_ERR_CAPACITY, "more than 2^24 bodies");
    int rc = w->download_state();
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        if (desc[i].type == RBGPU_DYNAMIC && !(desc[i].mass > 0.0)) return fail(RBGPU_ERR_INVALID, "dynamic body needs mass > 0");
        if (ids_out) ids_out[i] = (uint32_t)w->hb.size();
        w->hb.push_back(desc[i]);
    }
    w->built = false;
    return RBGPU_OK;
}

int rbgpu_add_colliders(rbgpu_world *w, int n, const rbgpu_collider_desc *desc, uint32_t *ids_out) {
    if (!w || n < 0 || (n && !desc)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    int rc = w->download_state();
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        const rbgpu_collider_desc &d = desc[i];
        if (d.body_id >= w->hb.size()) return fail(RBGPU_ERR_INVALID, "collider body_id out of range");
        if (d.shape != RBGPU_SPHERE && d.shape != RBGPU_BOX && d.shape != RBGPU_CAPSULE && d.shape != RBGPU_PLANE &&
            d.shape != RBGPU_HULL)
            return fail(RBGPU_ERR_UNSUPPORTED, "shape kind not supported by this build");
        if (d.shape == RBGPU_HULL && (d.hull_index < 0 || (size_t)d.hull_index >= w->hh.size()))
            return fail(RBGPU_ERR_INVALID, "hull collider: hull_index does not name a hull of this world");
        const uint32_t id = (uint32_t)w->hc.size();
        w->hc.push_back(d);
        if (ids_out) ids_out[i] = id;
        if (d.shape == RBGPU_PLANE) continue;   // planes bypass the broadphase (ColliderAPI.hpp:29-33)
        // initial tight AABB at the body's current pose, fattened by the margin (BroadPhase.cpp:169-197)
        const rbgpu_body_desc &b = w->hb[d.body_id];
        TF bt{V3{(float)b.position[0], (float)b.position[1], (float)b.position[2]},
              Q{(float)b.orientation[0], (float)b.orientation[1], (float)b.orientation[2], (float)b.orientation[3]}};
        V3 lp{(float)d.local_pos[0], (float)d.local_pos[1], (float)d.local_pos[2]};
        Q lr{(float)d.local_rot[0], (float)d.local_rot[1], (float)d.local_rot[2], (float)d.local_rot[3]};
        TF tf{bt.pos + rotate(bt.rot, lp), qmul(bt.rot, lr)};
        Shape s{d.shape, (float)d.dims[0], (float)d.dims[1], (float)d.dims[2], (float)d.dims[3]};
        Box3 t;
        if (d.shape == RBGPU_HULL) {
            const auto &h = w->hh[d.hull_index];
            t = hull_aabb(V3{h.amin[0], h.amin[1], h.amin[2]}, V3{h.amax[0], h.amax[1], h.amax[2]}, tf);
        } else {
            t = shape_aabb(s, tf);
        }
        const float mg = w->margin;
        V3 mn = t.mn - V3{mg, mg, mg}, mx = t.mx + V3{mg, mg, mg};
        const uint32_t slot = w->S++;
        w->h_fmin.push_back(make_float4(mn.x, mn.y, mn.z, 0.0f));
        w->h_fmax.push_back(make_float4(mx.x, mx.y, mx.z, 0.0f));
        w->h_uid.push_back(id);
        w->h_eppos.push_back(make_uint2((uint32_t)w->h_key.size(), (uint32_t)w->h_key.size() + 1));
        w->h_key.push_back(float_key(mn.x));
        w->h_val.push_back(slot << 1);
        w->h_key.push_back(float_key(mx.x));
        w->h_val.push_back((slot << 1) | 1u);
    }
    w->built = false;
    return RBGPU_OK;
}

int rbgpu_add_joints(rbgpu_world *w, int n, const rbgpu_joint_desc *desc, uint32_t *ids_out) {
    if (!w || n < 0 || (n && !desc)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    for (int i = 0; i < n; ++i) {
        const rbgpu_joint_desc &d = desc[i];
        if (d.type < RBGPU_JOINT_PRISMATIC || d.type > RBGPU_JOINT_ROW) return fail(RBGPU_ERR_INVALID, "unknown joint type");
        if (d.type == RBGPU_JOINT_ROW && d.actuation != 0 && d.actuation != 1)
            return fail(RBGPU_ERR_INVALID, "constraint row type is 0 (positional) or 1 (rotational)");
        if (d.actuation < RBGPU_ACT_FREE || d.actuation > RBGPU_ACT_SPEED) return fail(RBGPU_ERR_INVALID, "unknown actuation");
        if (d.body_1 >= w->hb.size() || d.body_2 >= w->hb.size()) return fail(RBGPU_ERR_INVALID, "joint body out of range");
        if (d.body_1 == d.body_2) return fail(RBGPU_ERR_INVALID, "joint needs two different bodies");
        if (w->hb[d.body_1].scene != w->hb[d.body_2].scene) return fail(RBGPU_ERR_INVALID, "joint across two scenes");
    }
    int rc = w->download_state();
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        if (ids_out) ids_out[i] = (uint32_t)w->hj.size();
        w->hj.push_back(desc[i]);
        w->h_jtg.push_back(make_float4((float)desc[i].target_position, (float)desc[i].target_speed, 0.0f, 0.0f));
        for (int k = 0; k < 4; ++k) w->h_jlam.push_back(0.0f);
        w->h_jrow_user.push_back(make_float4(1.0f, 0.0f, 0.0f, 0.0f));   // ConstraintAPI.hpp:52-54
    }
    w->built = false;
    return RBGPU_OK;
}

int rbgpu_get_joints(rbgpu_world *w, double *cur, double *tgt, double *lam4, int32_t *colour) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    const size_t nj = w->hj.size();
    std::vector<float4> tg(w->h_jtg.begin(), w->h_jtg.end());
    std::vector<float> lam(w->h_jlam.begin(), w->h_jlam.end());
    if (w->built && w->NJ) {
        HIPCHK(hipStreamSynchronize(w->st));
        HIPCHK(hipMemcpy(tg.data(), w->Jt.jtg, w->NJ * 16, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lam.data(), w->Jt.jlam, (size_t)w->NJ * 16, hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < nj; ++i) {
        if (cur) cur[i] = tg[i].z;
        if (tgt) tgt[i] = tg[i].x;
        if (lam4) for (int k = 0; k < 4; ++k) lam4[4 * i + k] = lam[4 * i + k];
        if (colour) colour[i] = (w->built && i < w->h_jcolour.size()) ? w->h_jcolour[i] : -1;
    }
    return RBGPU_OK;
}

int rbgpu_set_joint_targets(rbgpu_world *w, int n, const uint32_t *ids, const double *tgt, const double *speed) {
    if (!w || n < 0 || (n && !ids)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    for (int k = 0; k < n; ++k)
        if (ids[k] >= w->hj.size()) return fail(RBGPU_ERR_INVALID, "joint id out of range");
    int rc = w->download_state();
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        float4 &t = w->h_jtg[ids[k]];
        if (tgt) t.x = (float)tgt[k];
        if (speed) t.y = (float)speed[k];
    }
    if (w->built && w->NJ) {
        HIPCHK(hipMemcpy(w->Jt.jtg, w->h_jtg.data(), w->NJ * 16, hipMemcpyHostToDevice));
    }
    return RBGPU_OK;
}

int rbgpu_add_constraints(rbgpu_world *w, int n, const rbgpu_constraint_desc *desc, uint32_t *ids_out) {
    if (!w || n < 0 || (n && !desc)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    std::vector<rbgpu_joint_desc> d(n);
    for (int i = 0; i < n; ++i) {
        if (desc[i].type != 0 && desc[i].type != 1)
            return fail(RBGPU_ERR_INVALID, "constraint row type is 0 (positional) or 1 (rotational)");
        d[i] = rbgpu_joint_desc{};
        d[i].type = RBGPU_JOINT_ROW;
        d[i].actuation = desc[i].type;
        d[i].body_1 = desc[i].body_1;
        d[i].body_2 = desc[i].body_2;
        for (int k = 0; k < 3; ++k) { d[i].r_1[k] = desc[i].r_1[k]; d[i].r_2[k] = desc[i].r_2[k]; }
        d[i].compliance = desc[i].compliance;
    }
    return rbgpu_add_joints(w, n, d.data(), ids_out);
}

int rbgpu_set_constraints(rbgpu_world *w, int n, const uint32_t *ids, const double *direction3, const double *magnitude) {
    if (!w || n < 0 || (n && !ids)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    for (int k = 0; k < n; ++k) {
        if (ids[k] >= w->hj.size()) return fail(RBGPU_ERR_INVALID, "constraint id out of range");
        if (w->hj[ids[k]].type != RBGPU_JOINT_ROW) return fail(RBGPU_ERR_INVALID, "not a constraint row (a joint's rows follow its errors)");
    }
    if (w->built && w->NJ) HIPCHK(hipStreamSynchronize(w->st));
    for (int k = 0; k < n; ++k) {
        float4 &r = w->h_jrow_user[ids[k]];
        if (direction3) { r.x = (float)direction3[3 * k]; r.y = (float)direction3[3 * k + 1]; r.z = (float)direction3[3 * k + 2]; }
        if (magnitude) r.w = (float)magnitude[k];
        if (w->built && ids[k] < w->NJ) HIPCHK(hipMemcpy(w->Jt.jrow + (size_t)ids[k] * 4, &r, 16, hipMemcpyHostToDevice));
    }
    return RBGPU_OK;
}

int rbgpu_get_constraint_rows(rbgpu_world *w, double *direction3, double *magnitude) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    const size_t nj = w->hj.size();
    std::vector<float4> rows(nj * 4, make_float4(1.0f, 0.0f, 0.0f, 0.0f));
    for (size_t i = 0; i < nj; ++i)
        if (w->hj[i].type == RBGPU_JOINT_ROW) rows[i * 4] = w->h_jrow_user[i];
    if (w->built && w->NJ) {
        if (int rcs = w->sync()) return rcs;
        HIPCHK(hipMemcpy(rows.data(), w->Jt.jrow, (size_t)w->NJ * 64, hipMemcpyDeviceToHost));
    }
    for (size_t j = 0; j < nj * 4; ++j) {
        if (direction3) { direction3[3 * j] = rows[j].x; direction3[3 * j + 1] = rows[j].y; direction3[3 * j + 2] = rows[j].z; }
        if (magnitude) magnitude[j] = rows[j].w;
    }
    return RBGPU_OK;
}

int rbgpu_step(rbgpu_world *w, int n_frames) {
    if (!w || n_frames < 0) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (!w->built) {
        int rc = w->build();
        if (rc) return rc;
    }
    for (int f = 0; f < n_frames; ++f) {
        int rc = w->frame();
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return w->queue_error_check();
}

int rbgpu_autotune(rbgpu_world *w, int frames_per_trial, rbgpu_tune_report *out) {
    if (!w || frames_per_trial < 1) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (!w->built) {
        int rc = w->build();
        if (rc) return rc;
    }
    if (w->timing_kernel) return fail(RBGPU_ERR_INVALID, "autotune with kernel timing on (rbgpu_set_timing 0 first)");
    // wall time of k frames between two synchronisations of the world's stream
    auto time_frames = [&](int k, double *ms) -> int {
        HIPCHK(hipStreamSynchronize(w->st));
        const auto t0 = std::chrono::steady_clock::now();
        for (int f = 0; f < k; ++f) {
            int rc = w->frame();
            if (rc) return rc;
        }
        HIPCHK(hipStreamSynchronize(w->st));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return RBGPU_OK;
    };
    struct Knob {
        int opt;
        std::vector<int> vals;
    };
    std::vector<Knob> knobs;
    if (w->scene_path()) {
        knobs.push_back({RBGPU_OPT_SCENE_THREADS, {64, 128, 256}});
        knobs.push_back({RBGPU_OPT_GRAPH, {1, 0}});
    } else {
        knobs.push_back({RBGPU_OPT_RUN_MIN, {128, 256, 512, 1024}});
        knobs.push_back({RBGPU_OPT_ROUND_BATCH, {6, 12, 24}});
    }
    // Not tuned: the GJK row layout, the EPA work item size and the GJK/EPA grids, whose best
    // values flip between the sparse early frames the trials see and settled piles (their
    // defaults were measured on settled piles).
    rbgpu_tune_report rep{};
    double ms = 0.0;
    int rc = time_frames(1, &ms);   // settles allocations and the first capture
    if (rc) return rc;
    rep.frames += 1;
    double best_all = 0.0;
    for (const Knob &k : knobs) {
        int cur = 0;
        rc = rbgpu_get_option(w, k.opt, &cur);
        if (rc) return rc;
        // two passes over the values, the second in reverse order, times summed per value: the
        // scene moves between trials, so a single trial per value mostly measures which frames
        // it got
        int best = cur;
        double best_ms = 1e300;
        std::vector<double> acc(k.vals.size(), 0.0);
        const size_t nv = k.vals.size();
        for (size_t t = 0; t < 2 * nv; ++t) {
            const size_t j = t < nv ? t : 2 * nv - 1 - t;
            rc = rbgpu_set_option(w, k.opt, k.vals[j]);
            if (rc) return rc;
            rc = time_frames(frames_per_trial, &ms);
            if (rc) return rc;
            rep.frames += frames_per_trial;
            ++rep.trials;
            acc[j] += ms;
        }
        for (size_t j = 0; j < nv; ++j)
            if (acc[j] < best_ms) { best_ms = acc[j]; best = k.vals[j]; }
        (void)cur;
        rc = rbgpu_set_option(w, k.opt, best);
        if (rc) return rc;
        best_all = best_ms / (2.0 * frames_per_trial);
    }
    rbgpu_get_option(w, RBGPU_OPT_SCENE_THREADS, &rep.scene_threads);
    rbgpu_get_option(w, RBGPU_OPT_GRAPH, &rep.graph);
    rbgpu_get_option(w, RBGPU_OPT_RUN_MIN, &rep.run_min);
    rbgpu_get_option(w, RBGPU_OPT_ROUND_BATCH, &rep.round_batch);
    rbgpu_get_option(w, RBGPU_OPT_GJK_GRID, &rep.gjk_grid);
    rbgpu_get_option(w, RBGPU_OPT_EPA_GRID, &rep.epa_grid);
    rep.scene_kernels = w->scene_path() ? 1 : 0;
    rep.best_ms_per_frame = best_all;
    if (out) *out = rep;
    return RBGPU_OK;
}

int rbgpu_set_step(rbgpu_world *w, double timestep, int substeps) {
    if (!w || substeps <= 0 || !(timestep > 0.0)) return fail(RBGPU_ERR_INVALID, "invalid timestep/substeps");
    w->timestep = timestep;
    w->substeps = substeps;
    return RBGPU_OK;
}

int rbgpu_synchronize(rbgpu_world *w) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    return w->sync();
}

int rbgpu_get_bodies(rbgpu_world *w, double *p, double *q, double *v, double *av) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    const uint32_t n = (uint32_t)w->hb.size();
    if (!w->built) {
        for (uint32_t i = 0; i < n; ++i) {
            const rbgpu_body_desc &d = w->hb[i];
            for (int k = 0; k < 3; ++k) {
                if (p) p[3 * i + k] = (float)d.position[k];
                if (v) v[3 * i + k] = (float)d.linear_velocity[k];
                if (av) av[3 * i + k] = (float)d.angular_velocity[k];
            }
            for (int k = 0; k < 4; ++k)
                if (q) q[4 * i + k] = (float)d.orientation[k];
        }
        return RBGPU_OK;
    }
    if (int rcs = w->sync()) return rcs;
    std::vector<BodyRow> rows;
    HIPCHK(body_rows(rows, w->B.r, n));
    for (uint32_t i = 0; i < n; ++i) {
        const BodyRow &r = rows[i];
        if (p) { p[3 * i] = r.x.x; p[3 * i + 1] = r.x.y; p[3 * i + 2] = r.x.z; }
        if (q) { q[4 * i] = r.q.x; q[4 * i + 1] = r.q.y; q[4 * i + 2] = r.q.z; q[4 * i + 3] = r.q.w; }
        if (v) { v[3 * i] = r.v.x; v[3 * i + 1] = r.v.y; v[3 * i + 2] = r.v.z; }
        if (av) { av[3 * i] = r.w.x; av[3 * i + 1] = r.w.y; av[3 * i + 2] = r.w.z; }
    }
    return RBGPU_OK;
}

int rbgpu_set_bodies(rbgpu_world *w, const double *p, const double *q, const double *v, const double *av) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    int rc = w->download_state();
    if (rc) return rc;
    for (size_t i = 0; i < w->hb.size(); ++i) {
        rbgpu_body_desc &d = w->hb[i];
        for (int k = 0; k < 3; ++k) {
            if (p) d.position[k] = p[3 * i + k];
            if (v) d.linear_velocity[k] = v[3 * i + k];
            if (av) d.angular_velocity[k] = av[3 * i + k];
        }
        for (int k = 0; k < 4; ++k)
            if (q) d.orientation[k] = q[4 * i + k];
    }
    w->built = false;
    return RBGPU_OK;
}

int rbgpu_set_body_forces(rbgpu_world *w, const double *force3, const double *torque3) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    const size_t n = w->hb.size();
    w->h_force.resize(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    w->h_torque.resize(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (size_t i = 0; i < n; ++i) {
        if (force3) w->h_force[i] = make_float4((float)force3[3 * i], (float)force3[3 * i + 1], (float)force3[3 * i + 2], 1.0f);
        if (torque3)
            w->h_torque[i] = make_float4((float)torque3[3 * i], (float)torque3[3 * i + 1], (float)torque3[3 * i + 2], 1.0f);
    }
    if (!w->built || !n) return RBGPU_OK;
    if (int rcs = w->sync()) return rcs;
    std::vector<float4> f, t;
    w->body_forces(f, t);
    HIPCHK(hipMemcpy(w->B.f, f.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(w->B.t, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    return RBGPU_OK;
}

int rbgpu_get_body_forces(rbgpu_world *w, double *force3, double *torque3) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    std::vector<float4> f, t;
    w->body_forces(f, t);
    for (size_t i = 0; i < f.size(); ++i) {
        if (force3) { force3[3 * i] = f[i].x; force3[3 * i + 1] = f[i].y; force3[3 * i + 2] = f[i].z; }
        if (torque3) { torque3[3 * i] = t[i].x; torque3[3 * i + 1] = t[i].y; torque3[3 * i + 2] = t[i].z; }
    }
    return RBGPU_OK;
}

int rbgpu_get_sap_pairs(rbgpu_world *w, uint32_t *pairs2, uint32_t cap, uint32_t *n) {
    if (!w || !n) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *n = w->built ? w->n_sap : 0;
    if (!w->built || !pairs2) return RBGPU_OK;
    if (int rcs = w->sync()) return rcs;
    uint32_t k = std::min(cap, w->n_sap);
    if (k) HIPCHK(hipMemcpy(pairs2, w->pairs, (size_t)k * 8, hipMemcpyDeviceToHost));
    return RBGPU_OK;
}

int rbgpu_get_pair_colours(rbgpu_world *w, uint32_t *pairs2, int32_t *colour, uint32_t cap, uint32_t *n) {
    if (!w || !n) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *n = w->built ? w->n_pairs : 0;
    if (!w->built) return RBGPU_OK;
    if (int rcs = w->sync()) return rcs;
    const uint32_t k = std::min(cap, w->n_pairs);
    if (k && pairs2) HIPCHK(hipMemcpy(pairs2, w->pairs, (size_t)k * 8, hipMemcpyDeviceToHost));
    if (k && colour) {
        // colour of each pair in the last substep's colouring; -1 = no points
        std::vector<int> col(w->n_pairs);
        HIPCHK(hipMemcpy(col.data(), w->cb.acol, (size_t)w->n_pairs * 4, hipMemcpyDeviceToHost));
        for (uint32_t p = 0; p < k; ++p) colour[p] = col[p] >= 0 ? col[p] : -1;
    }
    return RBGPU_OK;
}

int rbgpu_get_contacts(rbgpu_world *w, uint32_t *ids4, double *n3, double *ra3, double *rb3, double *dep, double *lam2,
                       uint32_t cap, uint32_t *n) {
    if (!w || !n) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *n = 0;
    if (!w->built || !w->n_pairs) return RBGPU_OK;
    if (int rcs = w->sync()) return rcs;
    const uint32_t P = w->n_pairs;
    std::vector<int> cnt(P);
    std::vector<uint2> cc(P), bb(P);
    std::vector<float4> hd(2 * (size_t)P), pt(16 * (size_t)P);
    std::vector<float> dp(4 * (size_t)P);
    HIPCHK(hipMemcpy(cnt.data(), w->K.cnt, P * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cc.data(), w->K.cc, P * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(bb.data(), w->K.bb, P * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hd.data(), w->K.hd, (size_t)P * 32, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pt.data(), w->K.pt, (size_t)P * 256, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dp.data(), w->K.dep, (size_t)P * 16, hipMemcpyDeviceToHost));
    uint32_t k = 0;
    for (uint32_t s = 0; s < P; ++s) {
        for (int i = 0; i < cnt[s]; ++i, ++k) {
            if (k >= cap) continue;
            const size_t j = (size_t)s * 4 + i;
            const float4 r0 = pt[4 * j], r1 = pt[4 * j + 1], nr = hd[2 * (size_t)s];
            if (ids4) { ids4[4 * k] = bb[s].x; ids4[4 * k + 1] = bb[s].y; ids4[4 * k + 2] = cc[s].x; ids4[4 * k + 3] = cc[s].y; }
            if (n3) { n3[3 * k] = nr.x; n3[3 * k + 1] = nr.y; n3[3 * k + 2] = nr.z; }
            if (ra3) { ra3[3 * k] = r0.x; ra3[3 * k + 1] = r0.y; ra3[3 * k + 2] = r0.z; }
            if (rb3) { rb3[3 * k] = r1.x; rb3[3 * k + 1] = r1.y; rb3[3 * k + 2] = r1.z; }
            if (dep) dep[k] = dp[j];
            if (lam2) { lam2[2 * k] = r0.w; lam2[2 * k + 1] = r1.w; }
        }
    }
    *n = k;
    return RBGPU_OK;
}

int rbgpu_get_contact_points(rbgpu_world *w, double *point_on_a3, double *point_on_b3, double *relative_velocity,
                             int32_t *collision, uint32_t cap, uint32_t *n) {
    if (!w || !n) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *n = 0;
    if (!w->built || !w->n_pairs) return RBGPU_OK;
    if (int rcs = w->sync()) return rcs;
    const uint32_t P = w->n_pairs;
    std::vector<int> cnt(P);
    std::vector<float4> pt(16 * (size_t)P);
    HIPCHK(hipMemcpy(cnt.data(), w->K.cnt, P * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pt.data(), w->K.pt, (size_t)P * 256, hipMemcpyDeviceToHost));
    uint32_t k = 0;
    for (uint32_t s = 0; s < P; ++s)
        for (int i = 0; i < cnt[s]; ++i, ++k) {
            if (k >= cap) continue;
            const float4 a = pt[4 * ((size_t)s * 4 + i) + 2], b = pt[4 * ((size_t)s * 4 + i) + 3];
            if (point_on_a3) { point_on_a3[3 * k] = a.x; point_on_a3[3 * k + 1] = a.y; point_on_a3[3 * k + 2] = a.z; }
            if (point_on_b3) { point_on_b3[3 * k] = b.x; point_on_b3[3 * k + 1] = b.y; point_on_b3[3 * k + 2] = b.z; }
            if (relative_velocity) relative_velocity[k] = a.w;
            if (collision) collision[k] = b.w != 0.0f ? 1 : 0;
        }
    *n = k;
    return RBGPU_OK;
}

int rbgpu_get_broadphase_objects(rbgpu_world *w, double *fat_min3, double *fat_max3, uint32_t *user_id, uint32_t cap,
                                 uint32_t *n) {
    if (!w || !n) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    const uint32_t S = (uint32_t)w->h_uid.size();
    *n = S;
    std::vector<float4> mn(w->h_fmin), mx(w->h_fmax);
    if (w->built && w->S) {
        if (int rcs = w->sync()) return rcs;
        HIPCHK(hipMemcpy(mn.data(), w->fmin, (size_t)w->S * 16, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mx.data(), w->fmax, (size_t)w->S * 16, hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < S && i < cap; ++i) {
        if (fat_min3) { fat_min3[3 * i] = mn[i].x; fat_min3[3 * i + 1] = mn[i].y; fat_min3[3 * i + 2] = mn[i].z; }
        if (fat_max3) { fat_max3[3 * i] = mx[i].x; fat_max3[3 * i + 1] = mx[i].y; fat_max3[3 * i + 2] = mx[i].z; }
        if (user_id) user_id[i] = w->h_uid[i];
    }
    return RBGPU_OK;
}

int rbgpu_get_stats(rbgpu_world *w, rbgpu_stats *out) {
    if (!w || !out) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    std::memset(out, 0, sizeof(*out));
    out->frames = w->frames;
    out->sap_pairs = w->n_sap;
    out->pairs = w->n_pairs;
    out->graph_frames = (uint32_t)w->graph_frames;
    out->scene_kernels = w->built && w->scene_path() ? 1u : 0u;
    if (w->built) {
        if (int rcs = w->sync()) return rcs;
        uint32_t sc[CTR_SIZE];
        HIPCHK(hipMemcpy(sc, w->cb.ctr, sizeof(sc), hipMemcpyDeviceToHost));
        out->colours = sc[CTR_NCOL];
        out->colour_rounds = sc[CTR_ROUNDS];
        out->active_pairs = sc[CTR_ACT];
    }
    for (int b = 0; b < NBINS && (size_t)b + 1 < w->bin_off.size(); ++b) out->bins[b] = w->bin_off[b + 1] - w->bin_off[b];
    if (w->built && w->n_pairs) {
        HIPCHK(hipStreamSynchronize(w->st));
        std::vector<int> cnt(w->n_pairs);
        HIPCHK(hipMemcpy(cnt.data(), w->K.cnt, w->n_pairs * 4, hipMemcpyDeviceToHost));
        uint32_t s = 0;
        for (int c : cnt) s += (uint32_t)c;
        out->contacts = s;
    }
    return RBGPU_OK;
}

int rbgpu_set_option(rbgpu_world *w, int option, int value) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    switch (option) {
        case RBGPU_OPT_PALETTE_MAX:
            if (value < 1 || value > 64) return fail(RBGPU_ERR_INVALID, "palette maximum must be in [1, 64]");
            w->palette_max = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_HEAD_ROUNDS:
            if (value < 1 || value > (int)MAX_HEAD_ROUNDS) return fail(RBGPU_ERR_INVALID, "head rounds must be in [1, 256]");
            w->head_rounds = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_ROUND_BATCH:
            if (value < 1) return fail(RBGPU_ERR_INVALID, "round batch must be >= 1");
            w->round_batch = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_TAIL_MIN:
            if (value < 0) return fail(RBGPU_ERR_INVALID, "tail minimum must be >= 0");
            w->tail_min = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_RUN_MIN:
            if (value < 0) return fail(RBGPU_ERR_INVALID, "run minimum must be >= 0");
            w->run_min = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_TIMING_STRIDE:
            if (value < 1) return fail(RBGPU_ERR_INVALID, "timing stride must be >= 1");
            w->timing_stride = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_JOINT_MODE:
            if (value != 0 && value != 1) return fail(RBGPU_ERR_INVALID, "joint mode is 0 or 1");
            w->joint_mode = value; return RBGPU_OK;
        case RBGPU_OPT_SCENE_KERNELS:
            if (value < -1 || value > 1) return fail(RBGPU_ERR_INVALID, "scene kernels: -1 auto, 0 off, 1 on");
            w->scene_mode = value; return RBGPU_OK;
        case RBGPU_OPT_SCENE_THREADS:
            if (value != 64 && value != 128 && value != 256) return fail(RBGPU_ERR_INVALID, "scene threads: 64, 128 or 256");
            w->scene_threads = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_GJK_CULL:
            if (value != 0 && value != 1) return fail(RBGPU_ERR_INVALID, "GJK cull is 0 or 1");
            w->gjk_cull = value != 0; return RBGPU_OK;
        case RBGPU_OPT_GJK_GRID:
            if (value < 64 || value > (1 << 20)) return fail(RBGPU_ERR_INVALID, "GJK grid must be in [64, 2^20]");
            w->gjk_grid = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_EPA_GRID:
            if (value < 64 || value > (1 << 20)) return fail(RBGPU_ERR_INVALID, "EPA grid must be in [64, 2^20]");
            w->epa_grid = (uint32_t)value; return RBGPU_OK;
        case RBGPU_OPT_GRAPH:
            if (value != 0 && value != 1) return fail(RBGPU_ERR_INVALID, "graph is 0 or 1");
            w->use_graph = value != 0; return RBGPU_OK;
        case RBGPU_OPT_CONTACT_MODE:
            if (value != 0 && value != 1) return fail(RBGPU_ERR_INVALID, "contact mode is 0 or 1");
            w->contact_mode = value; return RBGPU_OK;
        case RBGPU_OPT_GJK_ROWS:
            if (value < 0 || value > 5)
                return fail(RBGPU_ERR_INVALID, "GJK layout: 0 wavefront, 1 row, 2 lane per pair, 3 refilled rows, "
                                               "4 GJK + EPA per row, 5 the same with row buffers most pairs outgrow");
            w->gjk_layout = value; return RBGPU_OK;
        case RBGPU_OPT_GJK_OCCUPANCY:
            if (value < 2 || value > 4) return fail(RBGPU_ERR_INVALID, "GJK occupancy: 2, 3 or 4 waves per EU");
            w->gjk_occ = value; return RBGPU_OK;
        case RBGPU_OPT_EPA_OCCUPANCY:
            if (value != 0 && value != 4) return fail(RBGPU_ERR_INVALID, "EPA occupancy: 0 (no budget) or 4");
            w->epa_occ = value; return RBGPU_OK;
        case RBGPU_OPT_SOLVE_OCCUPANCY:
            if (value != 0 && value != 4) return fail(RBGPU_ERR_INVALID, "solve occupancy: 0 (no budget) or 4");
            w->solve_occ = value; return RBGPU_OK;
        case RBGPU_OPT_SIDE_STREAM:
            if (value < 0 || value > 3) return fail(RBGPU_ERR_INVALID, "side stream: 0 (off), 1, 2 or 3 (fork point)");
            w->side_stream = value; return RBGPU_OK;
        case RBGPU_OPT_AXIS_CARRY:
            if (value != 0 && value != 1) return fail(RBGPU_ERR_INVALID, "axis carry is 0 or 1");
            w->axis_carry = value != 0; return RBGPU_OK;
        case RBGPU_OPT_EPA_GUARD:
            if (value < 0 || value > 2) return fail(RBGPU_ERR_INVALID, "EPA guard is 0, 1 or 2");
            w->epa_guard = value; return RBGPU_OK;
        case RBGPU_OPT_GJK_SORT:
            if (value < 0 || value > 3) return fail(RBGPU_ERR_INVALID, "GJK sort is 0, 1, 2 or 3");
            w->gjk_sort = value; return RBGPU_OK;
        case RBGPU_OPT_EPA_SUB:
            if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
                return fail(RBGPU_ERR_INVALID, "EPA work item: 1, 2, 4, 8 or 16 pairs");
            w->epa_sub = (uint32_t)value; return RBGPU_OK;
        default: return fail(RBGPU_ERR_INVALID, "unknown option");
    }
}

int rbgpu_get_option(rbgpu_world *w, int option, int *value) {
    if (!w || !value) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    switch (option) {
        case RBGPU_OPT_PALETTE_MAX: *value = (int)w->palette_max; return RBGPU_OK;
        case RBGPU_OPT_HEAD_ROUNDS: *value = (int)w->head_rounds; return RBGPU_OK;
        case RBGPU_OPT_ROUND_BATCH: *value = (int)w->round_batch; return RBGPU_OK;
        case RBGPU_OPT_TAIL_MIN: *value = (int)w->tail_min; return RBGPU_OK;
        case RBGPU_OPT_RUN_MIN: *value = (int)w->run_min; return RBGPU_OK;
        case RBGPU_OPT_TIMING_STRIDE: *value = (int)w->timing_stride; return RBGPU_OK;
        case RBGPU_OPT_JOINT_MODE: *value = w->joint_mode; return RBGPU_OK;
        case RBGPU_OPT_SCENE_KERNELS: *value = w->scene_mode; return RBGPU_OK;
        case RBGPU_OPT_SCENE_THREADS: *value = (int)w->scene_threads; return RBGPU_OK;
        case RBGPU_OPT_GJK_CULL: *value = w->gjk_cull ? 1 : 0; return RBGPU_OK;
        case RBGPU_OPT_GJK_GRID: *value = (int)w->gjk_grid; return RBGPU_OK;
        case RBGPU_OPT_EPA_GRID: *value = (int)w->epa_grid; return RBGPU_OK;
        case RBGPU_OPT_GRAPH: *value = w->use_graph ? 1 : 0; return RBGPU_OK;
        case RBGPU_OPT_CONTACT_MODE: *value = w->contact_mode; return RBGPU_OK;
        case RBGPU_OPT_GJK_ROWS: *value = w->gjk_layout; return RBGPU_OK;
        case RBGPU_OPT_EPA_SUB: *value = (int)w->epa_sub; return RBGPU_OK;
        case RBGPU_OPT_GJK_OCCUPANCY: *value = w->gjk_occ; return RBGPU_OK;
        case RBGPU_OPT_EPA_OCCUPANCY: *value = w->epa_occ; return RBGPU_OK;
        case RBGPU_OPT_SOLVE_OCCUPANCY: *value = w->solve_occ; return RBGPU_OK;
        case RBGPU_OPT_SIDE_STREAM: *value = w->side_stream; return RBGPU_OK;
        case RBGPU_OPT_AXIS_CARRY: *value = w->axis_carry ? 1 : 0; return RBGPU_OK;
        case RBGPU_OPT_EPA_GUARD: *value = w->epa_guard; return RBGPU_OK;
        case RBGPU_OPT_GJK_SORT: *value = w->gjk_sort; return RBGPU_OK;
        default: return fail(RBGPU_ERR_INVALID, "unknown option");
    }
}

int rbgpu_set_timing(rbgpu_world *w, int kernel) {
    if (!w || kernel < 0 || kernel > 11) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (w->st) HIPCHK(hipStreamSynchronize(w->st));
    w->timing_kernel = kernel;
    w->ev_used = 0;
    w->timing_bytes = 0.0;
    if (w->d_points) HIPCHK(hipMemset(w->d_points, 0, 8));
    return RBGPU_OK;
}

int rbgpu_get_timing(rbgpu_world *w, double *total_ms, uint64_t *launches, double *bytes) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (w->st) HIPCHK(hipStreamSynchronize(w->st));
    double ms = 0.0;
    for (size_t i = 0; i < w->ev_used; ++i) {
        float e = 0.0f;
        HIPCHK(hipEventElapsedTime(&e, w->ev_pool[i].first, w->ev_pool[i].second));
        ms += e;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = w->ev_used;
    if (bytes) {
        unsigned long long pts = 0;
        if (w->d_points) HIPCHK(hipMemcpy(&pts, w->d_points, 8, hipMemcpyDeviceToHost));
        // algorithmic bytes per contact point of the timed per-colour launches: position
        // solve 296 B, velocity solve 280 B
        // GJK pass: bytes per candidate (counted on the device) or per pair without the rejection
        *bytes = w->timing_kernel == 11 ? w->timing_bytes + (double)pts * w->gjk_pass_bytes()
                                        : (double)pts * (w->timing_kernel == 1 ? 296.0 : w->timing_kernel == 2 ? 280.0 : 0.0);
    }
    return RBGPU_OK;
}

int rbgpu_collide_pairs(int n, const int32_t *kind_a, const double *dims_a, const double *tf_a, const int32_t *kind_b,
                        const double *dims_b, const double *tf_b, double *out) {
    if (n < 0 || (n && (!kind_a || !dims_a || !tf_a || !kind_b || !dims_b || !tf_b || !out)))
        return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (n == 0) return RBGPU_OK;
    std::vector<float4> da(n), db(n), pa(n), pb(n), ra(n), rbq(n);
    for (int i = 0; i < n; ++i) {
        da[i] = make_float4((float)dims_a[4 * i], (float)dims_a[4 * i + 1], (float)dims_a[4 * i + 2], (float)dims_a[4 * i + 3]);
        db[i] = make_float4((float)dims_b[4 * i], (float)dims_b[4 * i + 1], (float)dims_b[4 * i + 2], (float)dims_b[4 * i + 3]);
        const double *a = tf_a + 7 * i, *b = tf_b + 7 * i;
        pa[i] = make_float4((float)a[0], (float)a[1], (float)a[2], 0.0f);
        ra[i] = make_float4((float)a[3], (float)a[4], (float)a[5], (float)a[6]);
        pb[i] = make_float4((float)b[0], (float)b[1], (float)b[2], 0.0f);
        rbq[i] = make_float4((float)b[3], (float)b[4], (float)b[5], (float)b[6]);
    }
    int *dka = nullptr, *dkb = nullptr;
    float4 *dda = nullptr, *ddb = nullptr, *dpa = nullptr, *dpb = nullptr, *dra = nullptr, *drb = nullptr;
    float *dout = nullptr;
    int rc = 0;
    rc |= dalloc(dka, n); rc |= dalloc(dkb, n); rc |= dalloc(dda, n); rc |= dalloc(ddb, n); rc |= dalloc(dpa, n);
    rc |= dalloc(dpb, n); rc |= dalloc(dra, n); rc |= dalloc(drb, n); rc |= dalloc(dout, 20 * (size_t)n);
    std::vector<float> fo(20 * (size_t)n);
    auto cleanup = [&]() {
        dfree(dka); dfree(dkb); dfree(dda); dfree(ddb); dfree(dpa); dfree(dpb); dfree(dra); dfree(drb); dfree(dout);
    };
    hipError_t e = hipSuccess;
    if (!rc) {
        if (e == hipSuccess) e = hipMemcpy(dka, kind_a, n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dkb, kind_b, n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dda, da.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ddb, db.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dpa, pa.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dpb, pb.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dra, ra.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(drb, rbq.data(), n * 16, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            k_collide_pairs<<<nblk(n), TPB>>>(n, dka, dda, dpa, dra, dkb, ddb, dpb, drb, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(fo.data(), dout, 20 * (size_t)n * 4, hipMemcpyDeviceToHost);
    }
    cleanup();
    if (rc) return rc;
    if (e != hipSuccess) return fail(RBGPU_ERR_HIP, std::string("rbgpu_collide_pairs: ") + hipGetErrorString(e));
    for (size_t k = 0; k < fo.size(); ++k) out[k] = fo[k];
    return RBGPU_OK;
}

// ---------------------------------------------------------------- narrowphase without allocation
// rbgpu_collide_pairs_noalloc stages its inputs and results in `scratch`, a file-scope device
// array of fixed size, instead of allocating per call.  Host code never passes `scratch`
// itself to a kernel or a copy (its name is a host-side stand-in there): each device's
// address of it comes from hipGetSymbolAddress, once per device.
namespace {
constexpr uint32_t NOALLOC_MAX_PAIRS = 16384;
// per pair: kinds 2 x 4 B, dims 2 x 32 B and poses 2 x 56 B as passed in (doubles), six staged
// float4 rows 96 B, the float result 80 B and its double copy 160 B
constexpr size_t NA_KA = 0, NA_KB = 4, NA_DA = 8, NA_DB = 40, NA_TA = 72, NA_TB = 128, NA_STAGE = 184, NA_FOUT = 280,
                 NA_DOUT = 360, NA_PER_PAIR = 520;
}  // namespace
static __device__ __align__(256) unsigned char scratch[(size_t)NOALLOC_MAX_PAIRS * NA_PER_PAIR];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoAllocDevice {
    std::once_flag once;
    unsigned char *ptr = nullptr;
    hipError_t err = hipSuccess;
};
NoAllocDevice g_noalloc[NOALLOC_MAX_DEVICES];

// The current device's address of `scratch`, looked up on the first call on that device.
int noalloc_scratch(unsigned char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(RBGPU_ERR_UNSUPPORTED, "device index beyond 64");
    NoAllocDevice &d = g_noalloc[dev];
    std::call_once(d.once, [&d]() { d.err = hipGetSymbolAddress((void **)&d.ptr, HIP_SYMBOL(scratch)); });
    if (d.err != hipSuccess || !d.ptr)
        return fail(RBGPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(d.err));
    *out = d.ptr;
    return RBGPU_OK;
}

// doubles as passed in -> the float4 rows k_collide_pairs reads (round to nearest, as the
// host conversion of rbgpu_collide_pairs)
__global__ void k_stage_pairs(uint32_t n, const double *dims_a, const double *tf_a, const double *dims_b,
                              const double *tf_b, float4 *da, float4 *db, float4 *pa, float4 *ra, float4 *pb, float4 *rb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *x = dims_a + 4 * (size_t)i, *y = dims_b + 4 * (size_t)i;
    da[i] = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
    db[i] = make_float4((float)y[0], (float)y[1], (float)y[2], (float)y[3]);
    const double *a = tf_a + 7 * (size_t)i, *b = tf_b + 7 * (size_t)i;
    pa[i] = make_float4((float)a[0], (float)a[1], (float)a[2], 0.0f);
    ra[i] = make_float4((float)a[3], (float)a[4], (float)a[5], (float)a[6]);
    pb[i] = make_float4((float)b[0], (float)b[1], (float)b[2], 0.0f);
    rb[i] = make_float4((float)b[3], (float)b[4], (float)b[5], (float)b[6]);
}

__global__ void k_widen(uint32_t n, const float *in, double *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}
}  // namespace

extern "C" int rbgpu_collide_pairs_noalloc(int n, const int32_t *kind_a, const double *dims_a, const double *tf_a,
                                           const int32_t *kind_b, const double *dims_b, const double *tf_b,
                                           double *out) {
    if (n < 0 || (n && (!kind_a || !dims_a || !tf_a || !kind_b || !dims_b || !tf_b || !out)))
        return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if ((uint32_t)n > NOALLOC_MAX_PAIRS)
        return fail(RBGPU_ERR_CAPACITY, "rbgpu_collide_pairs_noalloc: more than 16384 pairs in one call");
    if (n == 0) return RBGPU_OK;
    unsigned char *p = nullptr;
    int rc = noalloc_scratch(&p);
    if (rc) return rc;
    const size_t M = NOALLOC_MAX_PAIRS, un = (size_t)n;
    int *ka = (int *)(p + NA_KA * M), *kb = (int *)(p + NA_KB * M);
    double *da = (double *)(p + NA_DA * M), *db = (double *)(p + NA_DB * M);
    double *ta = (double *)(p + NA_TA * M), *tb = (double *)(p + NA_TB * M);
    float4 *st = (float4 *)(p + NA_STAGE * M);   // six rows of M float4 each
    float *fo = (float *)(p + NA_FOUT * M);
    double *dout = (double *)(p + NA_DOUT * M);
    HIPCHK(hipMemcpy(ka, kind_a, un * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(kb, kind_b, un * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(da, dims_a, un * 32, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, dims_b, un * 32, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ta, tf_a, un * 56, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tb, tf_b, un * 56, hipMemcpyHostToDevice));
    k_stage_pairs<<<nblk(n), TPB>>>(n, da, ta, db, tb, st, st + M, st + 2 * M, st + 3 * M, st + 4 * M, st + 5 * M);
    k_collide_pairs<<<nblk(n), TPB>>>(n, ka, st, st + 2 * M, st + 3 * M, kb, st + M, st + 4 * M, st + 5 * M, fo);
    k_widen<<<nblk(20 * n), TPB>>>(20 * n, fo, dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dout, un * 160, hipMemcpyDeviceToHost));
    return RBGPU_OK;
}

// ---------------------------------------------------------------- convex hulls
namespace {
int check_hull_input(const double *v, uint32_t nv, const uint32_t *tri, uint32_t nt) {
    if (!v || nv == 0) return fail(RBGPU_ERR_INVALID, "hull needs at least one vertex");
    if (nv > (uint32_t)HULL_MAX_VERTS) return fail(RBGPU_ERR_CAPACITY, "hull has more than 256 vertices");
    if (nt && !tri) return fail(RBGPU_ERR_INVALID, "hull triangles missing");
    for (uint32_t i = 0; i < 3 * nt; ++i)
        if (tri[i] >= nv) return fail(RBGPU_ERR_INVALID, "hull triangle index out of range");
    return RBGPU_OK;
}
}  // namespace

int rbgpu_add_hull(rbgpu_world *w, const double *vertices, uint32_t n_vertices, const uint32_t *triangles,
                   uint32_t n_triangles, uint32_t *hull_id_out) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    int rc = check_hull_input(vertices, n_vertices, triangles, n_triangles);
    if (rc) return rc;
    rbgpu_world::HostHull h;
    h.t = rbh::prepare_hull(vertices, n_vertices, triangles, n_triangles);
    const uint32_t P = (uint32_t)h.t.pn.size();
    for (uint32_t p = 0; p < P; ++p) {
        if (h.t.poff[p + 1] - h.t.poff[p] > (uint32_t)HULL_MAX_POLY)
            return fail(RBGPU_ERR_CAPACITY, "hull polygon with more than 64 vertices");
        if (h.t.p2p_o[p + 1] - h.t.p2p_o[p] > (uint32_t)HULL_MAX_POLY)
            return fail(RBGPU_ERR_CAPACITY, "hull polygon with more than 64 neighbours");
    }
    for (int k = 0; k < 3; ++k) { h.amin[k] = 1e30f; h.amax[k] = -1e30f; }
    for (const auto &p : h.t.v) {
        const float c[3] = {(float)p.x, (float)p.y, (float)p.z};
        for (int k = 0; k < 3; ++k) {
            h.amin[k] = c[k] < h.amin[k] ? c[k] : h.amin[k];
            h.amax[k] = h.amax[k] < c[k] ? c[k] : h.amax[k];
        }
    }
    {   // bounding ellipsoid: the AABB's centre and half extents scaled to hold every vertex,
        // widened by 1e-4 relative and 1e-6 absolute (far above the FP32 rounding of the bound)
        double c[3], e[3], emax = 0.0;
        for (int k = 0; k < 3; ++k) {
            c[k] = (double)(float)(0.5 * ((double)h.amin[k] + (double)h.amax[k]));   // the stored centre
            e[k] = 0.5 * ((double)h.amax[k] - (double)h.amin[k]);
            emax = std::max(emax, e[k]);
        }
        for (int k = 0; k < 3; ++k) e[k] = std::max(e[k], 1e-3 * emax + 1e-9);
        double s2 = 0.0;
        for (const auto &p : h.t.v) {
            const double v[3] = {(double)(float)p.x, (double)(float)p.y, (double)(float)p.z};
            double q = 0.0;
            for (int k = 0; k < 3; ++k) q += ((v[k] - c[k]) / e[k]) * ((v[k] - c[k]) / e[k]);
            s2 = std::max(s2, q);
        }
        const double sc = std::sqrt(s2);
        for (int k = 0; k < 3; ++k) {
            h.ec[k] = (float)c[k];
            const double a = sc * e[k] * (1.0 + 1e-4) + 1e-6;
            h.ea[k] = std::nextafter((float)a, 3.0e38f);
        }
    }
    rc = w->download_state();
    if (rc) return rc;
    if (hull_id_out) *hull_id_out = (uint32_t)w->hh.size();
    w->hh.push_back(std::move(h));
    w->built = false;
    return RBGPU_OK;
}

int rbgpu_hull_mass_properties(const double *vertices, uint32_t n_vertices, const uint32_t *triangles,
                               uint32_t n_triangles, double *volume, double *inertia6) {
    int rc = check_hull_input(vertices, n_vertices, triangles, n_triangles);
    if (rc) return rc;
    const rbh::HullTables t = rbh::prepare_hull(vertices, n_vertices, nullptr, 0);
    double vol, I[6];
    rbh::hull_mass(t, triangles, n_triangles, vol, I);
    if (volume) *volume = vol;
    if (inertia6) for (int k = 0; k < 6; ++k) inertia6[k] = I[k];
    return RBGPU_OK;
}

int rbgpu_hull_tables(const double *vertices, uint32_t n_vertices, const uint32_t *triangles, uint32_t n_triangles,
                      uint32_t *sizes6, double *poly_normals, uint32_t *poly_offsets, uint32_t *poly_indices,
                      uint32_t *v2p_offsets, uint32_t *v2p_indices, uint32_t *v2n_offsets, uint32_t *v2n_indices,
                      uint32_t *p2p_offsets, uint32_t *p2p_indices) {
    int rc = check_hull_input(vertices, n_vertices, triangles, n_triangles);
    if (rc) return rc;
    const rbh::HullTables t = rbh::prepare_hull(vertices, n_vertices, triangles, n_triangles);
    if (sizes6) {
        sizes6[0] = (uint32_t)t.pn.size(); sizes6[1] = (uint32_t)t.pidx.size(); sizes6[2] = (uint32_t)t.v2p_i.size();
        sizes6[3] = (uint32_t)t.v2n_i.size(); sizes6[4] = (uint32_t)t.p2p_i.size(); sizes6[5] = t.adjacency() ? 1u : 0u;
    }
    if (poly_normals)
        for (size_t k = 0; k < t.pn.size(); ++k) {
            poly_normals[3 * k] = t.pn[k].x; poly_normals[3 * k + 1] = t.pn[k].y; poly_normals[3 * k + 2] = t.pn[k].z;
        }
    auto cp = [](const std::vector<uint32_t> &a, uint32_t *o) { if (o) std::copy(a.begin(), a.end(), o); };
    cp(t.poff, poly_offsets); cp(t.pidx, poly_indices); cp(t.v2p_o, v2p_offsets); cp(t.v2p_i, v2p_indices);
    cp(t.v2n_o, v2n_offsets); cp(t.v2n_i, v2n_indices); cp(t.p2p_o, p2p_offsets); cp(t.p2p_i, p2p_indices);
    return RBGPU_OK;
}

int rbgpu_collide_world_pairs(rbgpu_world *w, uint32_t n, const uint32_t *pairs2, double *out32, uint32_t *gjk_stats,
                              const float *guess6, int seed_epa) {
    if (!w || (n && (!pairs2 || !out32))) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (!w->built) {
        int rc = w->build();
        if (rc) return rc;
    }
    for (uint32_t i = 0; i < 2 * n; ++i)
        if (pairs2[i] >= w->M) return fail(RBGPU_ERR_INVALID, "pair collider out of range");
    if (!n) return RBGPU_OK;
    int rc = w->ensure_pair_capacity(n);
    if (rc) return rc;
    hipStream_t st = w->st;
    HIPCHK(hipMemcpyAsync(w->pairs, pairs2, (size_t)n * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(w->scal + 1, &n, 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(w->gstat, 0xff, (size_t)n * 4, st));
    k_pair_gather<<<nblk(n), TPB, 0, st>>>(w->pairs, w->scal + 1, w->C, w->K, w->ka, w->va);
    rc = w->radix_pass(w->ka, w->va, w->kb, w->bin_list, w->pcap, 0, w->p_hist, nullptr, w->scal + 1);
    if (rc) return rc;
    k_offsets<<<1, 64, 0, st>>>(w->p_hist, (w->pcap + RS_TILE - 1) / RS_TILE, NBINS, w->scal + 1, w->offs);
    std::vector<uint32_t> off(NBINS + 1);
    HIPCHK(hipMemcpyAsync(off.data(), w->offs, (NBINS + 1) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    w->bin_off = off;
    // explicit warm starts (per input pair, A's frame) are permuted into bin order
    const int mode_saved = w->mlp_mode;
    float4 *gin = nullptr, *gsave = w->inf_guess;
    if (guess6) {
        std::vector<float4> g(2 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            g[2 * i] = make_float4(guess6[6 * i], guess6[6 * i + 1], guess6[6 * i + 2], 0.0f);
            g[2 * i + 1] = make_float4(guess6[6 * i + 3], guess6[6 * i + 4], guess6[6 * i + 5], 0.0f);
        }
        rc = dalloc(gin, 4 * (size_t)n);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(gin + 2 * (size_t)n, g.data(), (size_t)n * 32, hipMemcpyHostToDevice, st));
        const uint32_t g0 = off[BIN_HH], ng = off[BIN_HP] - off[BIN_HH];
        if (ng) k_gather_guess<<<nblk(ng), TPB, 0, st>>>(w->bin_list + g0, ng, gin + 2 * (size_t)n, gin);
        w->inf_guess = gin;
        w->mlp_mode = seed_epa ? 2 : 1;
    }
    w->gjk_explicit = guess6 != nullptr;
    const bool cull_saved = w->gjk_cull;
    const int guard_saved = w->epa_guard;
    w->gjk_cull = false;   // every pair runs GJK here: the entry reports its iterations and statuses
    w->epa_guard = 0;      // the listed pairs have no previous substep
    const int rcn = w->launch_narrowphase();
    w->gjk_cull = cull_saved;
    w->epa_guard = guard_saved;
    w->gjk_explicit = false;
    w->mlp_mode = mode_saved;
    w->inf_guess = gsave;
    if (rcn) return rcn;
    HIPCHK(hipGetLastError());
    std::vector<int> cnt(n);
    std::vector<float4> hd(2 * (size_t)n), pt(16 * (size_t)n);
    std::vector<float> dp(4 * (size_t)n);
    std::vector<uint32_t> gs(n);
    HIPCHK(hipMemcpyAsync(cnt.data(), w->K.cnt, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hd.data(), w->K.hd, (size_t)n * 32, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pt.data(), w->K.pt, (size_t)n * 256, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dp.data(), w->K.dep, (size_t)n * 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(gs.data(), w->gstat, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(gin);
    for (uint32_t i = 0; i < n; ++i) {
        double *o = out32 + (size_t)i * 32;
        for (int k = 0; k < 32; ++k) o[k] = 0.0;
        o[0] = cnt[i];
        if (!cnt[i]) continue;
        const float4 nr = hd[2 * (size_t)i];
        o[1] = nr.x; o[2] = nr.y; o[3] = nr.z;
        for (int k = 0; k < cnt[i]; ++k) {
            const size_t j = 4 * (size_t)i + k;
            const float4 a = pt[4 * j + 2], b = pt[4 * j + 3];
            double *q = o + 4 + 7 * k;
            q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = b.x; q[4] = b.y; q[5] = b.z; q[6] = dp[j];
        }
    }
    if (gjk_stats) std::copy(gs.begin(), gs.end(), gjk_stats);
    return RBGPU_OK;
}

// ---------------------------------------------------------------- learned warm start
namespace {
int need_mlp(rbgpu_world *w) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (!w->mlp_wb) return fail(RBGPU_ERR_INVALID, "warm-start network not initialised (rbgpu_mlp_init)");
    if (!w->st) HIPCHK(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking));
    return RBGPU_OK;
}
uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
}  // namespace

int rbgpu_mlp_init(rbgpu_world *w, uint64_t seed) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (!w->st) HIPCHK(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking));
    HIPCHK(hipStreamSynchronize(w->st));
    w->release_mlp();
    int rc = 0;
    rc |= dalloc(w->mlp_wf, MLP_NPARAM); rc |= dalloc(w->mlp_wb, MLP_NPARAM); rc |= dalloc(w->mlp_grad, MLP_NPARAM);
    rc |= dalloc(w->mlp_m, MLP_NPARAM); rc |= dalloc(w->mlp_v, MLP_NPARAM); rc |= dalloc(w->mlp_loss, 1);
    if (w->pcap) {
        rc |= dalloc(w->inf_feat, (size_t)w->pcap * MLP_IN); rc |= dalloc(w->inf_guess, 2 * (size_t)w->pcap);
        w->inf_cap = w->pcap;
    }
    if (rc) return RBGPU_ERR_OOM;
    // He-uniform weights, zero biases
    std::vector<float> p(MLP_NPARAM, 0.0f);
    uint64_t x = seed;
    const int wo[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4}, fin[4] = {MLP_IN, MLP_H, MLP_H, MLP_H};
    const int fout[4] = {MLP_H, MLP_H, MLP_H, MLP_OUT};
    for (int L = 0; L < 4; ++L) {
        const double lim = std::sqrt(6.0 / fin[L]);
        for (int i = 0; i < fin[L] * fout[L]; ++i)
            p[wo[L] + i] = (float)(lim * (2.0 * (double)(splitmix(x) >> 11) / 9007199254740992.0 - 1.0));
    }
    w->mlp_step = 0;
    return rbgpu_mlp_set_weights(w, p.data(), MLP_NPARAM);
}

int rbgpu_mlp_set_weights(rbgpu_world *w, const float *params, uint32_t n) {
    int rc = need_mlp(w);
    if (rc) return rc;
    if (!params || n != (uint32_t)MLP_NPARAM) return fail(RBGPU_ERR_INVALID, "parameter count must be 39312");
    hipStream_t st = w->st;
    HIPCHK(hipMemcpyAsync(w->mlp_wf, params, MLP_NPARAM * 4, hipMemcpyHostToDevice, st));
    k_f32_to_bf16<<<nblk(MLP_NPARAM), TPB, 0, st>>>(w->mlp_wf, w->mlp_wb, MLP_NPARAM);
    HIPCHK(hipMemsetAsync(w->mlp_grad, 0, MLP_NPARAM * 4, st));
    HIPCHK(hipMemsetAsync(w->mlp_m, 0, MLP_NPARAM * 4, st));
    HIPCHK(hipMemsetAsync(w->mlp_v, 0, MLP_NPARAM * 4, st));
    HIPCHK(hipStreamSynchronize(st));
    w->mlp_step = 0;
    return RBGPU_OK;
}

int rbgpu_mlp_get_weights(rbgpu_world *w, float *params, uint32_t n) {
    int rc = need_mlp(w);
    if (rc) return rc;
    if (!params || n != (uint32_t)MLP_NPARAM) return fail(RBGPU_ERR_INVALID, "parameter count must be 39312");
    HIPCHK(hipStreamSynchronize(w->st));
    HIPCHK(hipMemcpy(params, w->mlp_wf, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    return RBGPU_OK;
}

int rbgpu_mlp_set_mode(rbgpu_world *w, int mode) {
    if (!w || mode < 0 || mode > 2) return fail(RBGPU_ERR_INVALID, "mode: 0 plain, 1 GJK direction, 2 + EPA seed");
    if (mode) {
        int rc = need_mlp(w);
        if (rc) return rc;
    }
    w->mlp_mode = mode;
    return RBGPU_OK;
}

int rbgpu_mlp_log(rbgpu_world *w, uint64_t capacity) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (w->st) HIPCHK(hipStreamSynchronize(w->st));
    dfree(w->ds_feat); dfree(w->ds_tgt);
    w->ds_cap = w->ds_n = w->ds_hi = 0;
    w->ds_logging = false;
    if (!capacity) return RBGPU_OK;
    if (capacity >= (1ull << 31)) return fail(RBGPU_ERR_INVALID, "log capacity must be below 2^31 samples");
    int rc = 0;
    rc |= dalloc(w->ds_feat, (size_t)capacity * MLP_IN); rc |= dalloc(w->ds_tgt, 2 * (size_t)capacity);
    if (!w->ds_cnt) rc |= dalloc(w->ds_cnt, 1);
    if (rc) return RBGPU_ERR_OOM;
    HIPCHK(hipMemset(w->ds_cnt, 0, 4));
    w->ds_cap = capacity;
    w->ds_logging = true;
    return RBGPU_OK;
}

int rbgpu_mlp_logged(rbgpu_world *w, uint64_t *n, int stop) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (w->st) {
        int rc = w->sync_log_count();
        if (rc) return rc;
    }
    if (n) *n = w->ds_n;
    if (stop) w->ds_logging = false;
    return RBGPU_OK;
}

int rbgpu_mlp_get_log(rbgpu_world *w, uint64_t start, uint32_t n, float *features32, float *targets8) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (w->st) {
        int rc = w->sync_log_count();
        if (rc) return rc;
    }
    if (start + n > w->ds_n) return fail(RBGPU_ERR_INVALID, "range beyond the logged samples");
    if (!n) return RBGPU_OK;
    HIPCHK(hipStreamSynchronize(w->st));
    float *f = nullptr;
    if (dalloc(f, (size_t)n * MLP_IN)) return RBGPU_ERR_OOM;
    k_bf16_to_f32<<<nblk(n * MLP_IN), TPB, 0, w->st>>>(w->ds_feat + start * MLP_IN, f, n * MLP_IN);
    int rc = RBGPU_OK;
    if (features32 && hipMemcpyAsync(features32, f, (size_t)n * MLP_IN * 4, hipMemcpyDeviceToHost, w->st) != hipSuccess)
        rc = fail(RBGPU_ERR_HIP, "copy failed");
    if (targets8 && hipMemcpyAsync(targets8, w->ds_tgt + 2 * start, (size_t)n * 32, hipMemcpyDeviceToHost, w->st) != hipSuccess)
        rc = fail(RBGPU_ERR_HIP, "copy failed");
    if (hipStreamSynchronize(w->st) != hipSuccess) rc = fail(RBGPU_ERR_HIP, "rbgpu_mlp_get_log failed");
    dfree(f);
    return rc;
}

int rbgpu_mlp_train(rbgpu_world *w, int steps, uint32_t batch, double lr, double *mean_loss) {
    int rc = need_mlp(w);
    if (rc) return rc;
    if (steps < 0 || batch == 0 || batch > (1u << 20) || !(lr > 0.0)) return fail(RBGPU_ERR_INVALID, "invalid training arguments");
    if (w->st) {
        rc = w->sync_log_count();
        if (rc) return rc;
    }
    if (!w->ds_n) return fail(RBGPU_ERR_INVALID, "no logged samples (rbgpu_mlp_log, then step)");
    hipStream_t st = w->st;
    bf16 *bf = nullptr;
    float4 *bt = nullptr;
    rc = dalloc(bf, (size_t)batch * MLP_IN);
    rc |= dalloc(bt, 2 * (size_t)batch);
    if (rc) { dfree(bf); dfree(bt); return RBGPU_ERR_OOM; }
    // sample order: a multiplier coprime to the log size walks it as a permutation
    const uint64_t N = w->ds_n;
    uint64_t mult = 2654435761ull % N;
    auto gcd = [](uint64_t a, uint64_t b) { while (b) { uint64_t t = a % b; a = b; b = t; } return a; };
    while (N > 1 && (mult == 0 || gcd(mult, N) != 1)) ++mult;
    if (N == 1) mult = 1;
    HIPCHK(hipMemsetAsync(w->mlp_loss, 0, 4, st));
    const uint32_t tiles = (batch + MLP_TILE - 1) / MLP_TILE;
    for (int k = 0; k < steps; ++k) {
        const uint64_t start = (w->mlp_step * (uint64_t)batch) % N;
        k_mlp_gather<<<nblk(batch), TPB, 0, st>>>(w->ds_feat, w->ds_tgt, N, start, mult, batch, bf, bt);
        k_mlp_grad<<<std::min<uint32_t>(tiles, 512), 256, 0, st>>>(bf, bt, batch, MlpParams{w->mlp_wb, w->mlp_wf},
                                                                  1.0f / (float)batch, w->mlp_grad, w->mlp_loss);
        ++w->mlp_step;
        const float b1c = 1.0f - std::pow(0.9f, (float)w->mlp_step), b2c = 1.0f - std::pow(0.999f, (float)w->mlp_step);
        k_mlp_adam<<<nblk(MLP_NPARAM), TPB, 0, st>>>(w->mlp_wf, w->mlp_wb, w->mlp_grad, w->mlp_m, w->mlp_v, MLP_NPARAM,
                                                     (float)lr, b1c, b2c);
    }
    float loss = 0.0f;
    HIPCHK(hipMemcpyAsync(&loss, w->mlp_loss, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    dfree(bf);
    dfree(bt);
    if (mean_loss) *mean_loss = steps ? (double)loss / steps : 0.0;
    return RBGPU_OK;
}

int rbgpu_mlp_eval(rbgpu_world *w, uint32_t n, const float *feat32, float *out6) {
    int rc = need_mlp(w);
    if (rc) return rc;
    if (n && (!feat32 || !out6)) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    if (!n) return RBGPU_OK;
    hipStream_t st = w->st;
    float *f = nullptr;
    bf16 *fb = nullptr;
    float4 *o = nullptr;
    rc = dalloc(f, (size_t)n * MLP_IN);
    rc |= dalloc(fb, (size_t)n * MLP_IN);
    rc |= dalloc(o, 2 * (size_t)n);
    if (!rc) {
        (void)hipMemcpyAsync(f, feat32, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, st);
        k_f32_to_bf16<<<nblk(n * MLP_IN), TPB, 0, st>>>(f, fb, n * MLP_IN);
        k_mlp_infer<<<std::min<uint32_t>((n + MLP_TILE - 1) / MLP_TILE, 1024), MLP_INFER_THREADS, 0, st>>>(
            fb, n, nullptr, MlpParams{w->mlp_wb, w->mlp_wf}, o);
        std::vector<float4> h(2 * (size_t)n);
        (void)hipMemcpyAsync(h.data(), o, 2 * (size_t)n * 16, hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(RBGPU_ERR_HIP, "rbgpu_mlp_eval failed");
        for (uint32_t i = 0; i < n && !rc; ++i) {
            const float4 a = h[2 * i], b = h[2 * i + 1];
            float *q = out6 + 6 * (size_t)i;
            q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = b.x; q[4] = b.y; q[5] = b.z;
        }
    }
    dfree(f); dfree(fb); dfree(o);
    return rc;
}

int rbgpu_mlp_gradients(rbgpu_world *w, uint32_t n, const float *feat32, const float *tgt8, float *grad, double *loss) {
    int rc = need_mlp(w);
    if (rc) return rc;
    if (!n || !feat32 || !tgt8 || !grad) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    hipStream_t st = w->st;
    float *f = nullptr, *g = nullptr, *l = nullptr;
    bf16 *fb = nullptr;
    float4 *t = nullptr;
    rc = dalloc(f, (size_t)n * MLP_IN);
    rc |= dalloc(fb, (size_t)n * MLP_IN);
    rc |= dalloc(t, 2 * (size_t)n);
    rc |= dalloc(g, MLP_NPARAM);
    rc |= dalloc(l, 1);
    if (!rc) {
        (void)hipMemcpyAsync(f, feat32, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(t, tgt8, (size_t)n * 32, hipMemcpyHostToDevice, st);
        (void)hipMemsetAsync(g, 0, MLP_NPARAM * 4, st);
        (void)hipMemsetAsync(l, 0, 4, st);
        k_f32_to_bf16<<<nblk(n * MLP_IN), TPB, 0, st>>>(f, fb, n * MLP_IN);
        k_mlp_grad<<<std::min<uint32_t>((n + MLP_TILE - 1) / MLP_TILE, 512), 256, 0, st>>>(
            fb, t, n, MlpParams{w->mlp_wb, w->mlp_wf}, 1.0f / (float)n, g, l);
        float lh = 0.0f;
        (void)hipMemcpyAsync(grad, g, MLP_NPARAM * 4, hipMemcpyDeviceToHost, st);
        (void)hipMemcpyAsync(&lh, l, 4, hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(RBGPU_ERR_HIP, "rbgpu_mlp_gradients failed");
        if (loss) *loss = lh;
    }
    dfree(f); dfree(fb); dfree(t); dfree(g); dfree(l);
    return rc;
}

int rbgpu_get_gjk_counts(rbgpu_world *w, uint64_t *out, uint32_t n, int reset) {
    if (!w) return fail(RBGPU_ERR_INVALID, "null world");
    if (n > RBGPU_GJK_COUNTS) return fail(RBGPU_ERR_INVALID, "at most RBGPU_GJK_COUNTS counters");
    if (!w->st) HIPCHK(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking));
    constexpr size_t bytes = sizeof(unsigned long long) * RBGPU_GJK_COUNTS;
    if (!w->gjk_tot) {
        if (dalloc(w->gjk_tot, RBGPU_GJK_COUNTS)) return RBGPU_ERR_OOM;
        HIPCHK(hipMemsetAsync(w->gjk_tot, 0, bytes, w->st));
    }
    unsigned long long h[RBGPU_GJK_COUNTS] = {};
    HIPCHK(hipMemcpyAsync(h, w->gjk_tot, bytes, hipMemcpyDeviceToHost, w->st));
    if (reset) HIPCHK(hipMemsetAsync(w->gjk_tot, 0, bytes, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    if (out) for (uint32_t k = 0; k < n; ++k) out[k] = h[k];
    return RBGPU_OK;
}

int rbgpu_get_gjk_totals(rbgpu_world *w, uint64_t *out6, int reset) { return rbgpu_get_gjk_counts(w, out6, 6, reset); }

}  // extern "C"

// ------------------------------------------------------------------ snapshot / restore
// The state that carries from one frame to the next, copied device to device on the
// world's stream: body rows (pose, velocity, previous pose), force and torque, fat AABBs,
// the endpoint order the next frame's stable sort starts from, endpoint positions, joint
// targets / current positions, joint lambdas (never reset in joint mode 0), constraint
// rows, and the GJK rejection hints keyed by collider pair.  Everything else a frame
// rebuilds from these.  The header holds the counts (the layout check), the frame count,
// the step and every option.
struct rbgpu_snapshot {
    struct Header {
        char magic[8];
        uint32_t version;
        uint32_t N, M, S, NEP, NJ, nscenes, nhulls;
        uint32_t frames;
        uint32_t ax_H;            // hint table entries (a power of two; 0: none)
        int32_t substeps;
        int32_t n_options;
        int32_t options[32];
        double timestep;
        uint64_t payload;         // bytes of the section block
        uint64_t checksum;        // FNV-1a 64 of the section block (serialised form)
    } hd{};
    unsigned char *dev = nullptr;  // the section block (device)
    size_t cap = 0;
    ~rbgpu_snapshot() { dfree(dev); }
};

namespace {

constexpr char SNAP_MAGIC[8] = {'R', 'B', 'G', 'P', 'U', 'S', 'N', 'P'};
constexpr uint32_t SNAP_VERSION = 1;
constexpr int SNAP_OPTIONS = RBGPU_OPT_GJK_SORT + 1;
constexpr size_t SNAP_ALIGN = 256;

struct SnapSec {
    void *world;   // the world's buffer (nullptr: the hint table, handled apart)
    size_t bytes;
};
// Sections in block order; the last two are the hint table's keys and values.
std::vector<SnapSec> snap_sections(rbgpu_world *w, const rbgpu_snapshot::Header &h) {
    return {{w->B.r, (size_t)h.N * sizeof(BodyRow)}, {w->B.f, (size_t)h.N * 16}, {w->B.t, (size_t)h.N * 16},
            {w->fmin, (size_t)h.S * 16}, {w->fmax, (size_t)h.S * 16}, {w->ep_key[0], (size_t)h.NEP * 4},
            {w->ep_val[0], (size_t)h.NEP * 4}, {w->ep_pos, (size_t)h.S * 8}, {w->Jt.jtg, (size_t)h.NJ * 16},
            {w->Jt.jlam, (size_t)h.NJ * 16}, {w->Jt.jrow, (size_t)h.NJ * 64}, {nullptr, (size_t)h.ax_H * 8},
            {nullptr, (size_t)h.ax_H * 16}};
}
std::vector<size_t> snap_offsets(const std::vector<SnapSec> &sec) {
    std::vector<size_t> off(sec.size() + 1, 0);
    for (size_t k = 0; k < sec.size(); ++k) off[k + 1] = off[k] + (sec[k].bytes + SNAP_ALIGN - 1) / SNAP_ALIGN * SNAP_ALIGN;
    return off;
}
uint64_t fnv1a(const unsigned char *p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}
int snap_reserve(rbgpu_snapshot *s, size_t bytes) {
    if (bytes <= s->cap && s->dev) return RBGPU_OK;
    dfree(s->dev);
    s->cap = 0;
    if (dalloc(s->dev, bytes)) return RBGPU_ERR_OOM;
    s->cap = bytes;
    return RBGPU_OK;
}

int snap_capture(rbgpu_world *w, rbgpu_snapshot *s) {
    if (!w->built) {
        int rc = w->build();
        if (rc) return rc;
    }
    if (int rc = w->sync()) return rc;
    rbgpu_snapshot::Header h{};
    std::memcpy(h.magic, SNAP_MAGIC, 8);
    h.version = SNAP_VERSION;
    h.N = w->N; h.M = w->M; h.S = w->S; h.NEP = w->NEP; h.NJ = w->NJ;
    h.nscenes = w->nscenes; h.nhulls = (uint32_t)w->hh.size(); h.frames = w->frames;
    h.timestep = w->timestep; h.substeps = w->substeps;
    h.n_options = SNAP_OPTIONS;
    for (int k = 0; k < SNAP_OPTIONS; ++k)
        if (int rc = rbgpu_get_option(w, k, &h.options[k])) return rc;
    // the hint table the next frame would carry: the one restored and not yet used, or the
    // one saved from the last frame's GJK slots
    const uint32_t ng = w->n_pairs && !w->hh.empty() && w->bin_off.size() == NBINS + 1
                            ? w->bin_off[BIN_HP] - w->bin_off[BIN_HH] : 0u;
    const bool from_restore = w->ax_restore >= 0;
    if (from_restore) {
        h.ax_H = (uint32_t)w->ax_restore;
    } else if (w->axis_carry && w->gjk_cull && w->gjk_axis && ng) {
        h.ax_H = 1;
        while (h.ax_H < 2 * ng) h.ax_H <<= 1;
    }
    const std::vector<SnapSec> sec = snap_sections(w, h);
    const std::vector<size_t> off = snap_offsets(sec);
    h.payload = off.back();
    if (int rc = snap_reserve(s, h.payload)) return rc;
    for (size_t k = 0; k + 2 < sec.size(); ++k)
        if (sec[k].bytes)
            HIPCHK(hipMemcpyAsync(s->dev + off[k], sec[k].world, sec[k].bytes, hipMemcpyDeviceToDevice, w->st));
    if (h.ax_H) {
        auto *keys = (unsigned long long *)(s->dev + off[sec.size() - 2]);
        auto *vals = (float4 *)(s->dev + off[sec.size() - 1]);
        if (from_restore) {
            HIPCHK(hipMemcpyAsync(keys, w->ax_keys, (size_t)h.ax_H * 8, hipMemcpyDeviceToDevice, w->st));
            HIPCHK(hipMemcpyAsync(vals, w->ax_vals, (size_t)h.ax_H * 16, hipMemcpyDeviceToDevice, w->st));
        } else {
            k_axis_clear<<<nblk(h.ax_H), TPB, 0, w->st>>>(keys, h.ax_H);
            k_axis_save<<<nblk(ng), TPB, 0, w->st>>>(w->bin_list + w->bin_off[BIN_HH], ng, w->pcap, w->K, w->gjk_axis,
                                                      keys, vals, h.ax_H - 1);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipStreamSynchronize(w->st));
    s->hd = h;
    return RBGPU_OK;
}

int snap_check(const rbgpu_snapshot::Header &h) {
    if (std::memcmp(h.magic, SNAP_MAGIC, 8) != 0) return fail(RBGPU_ERR_INVALID, "not a world snapshot");
    if (h.version != SNAP_VERSION) return fail(RBGPU_ERR_INVALID, "snapshot version mismatch");
    if (h.n_options != SNAP_OPTIONS) return fail(RBGPU_ERR_INVALID, "snapshot option count mismatch");
    if (h.ax_H & (h.ax_H - 1)) return fail(RBGPU_ERR_INVALID, "corrupt snapshot (hint table size)");
    return RBGPU_OK;
}

}  // namespace

extern "C" {

int rbgpu_snapshot_create(rbgpu_world *w, rbgpu_snapshot **out) {
    if (!w || !out) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *out = nullptr;
    auto *s = new rbgpu_snapshot();
    int rc = snap_capture(w, s);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return RBGPU_OK;
}

int rbgpu_snapshot_capture(rbgpu_world *w, rbgpu_snapshot *s) {
    if (!w || !s) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    return snap_capture(w, s);
}

int rbgpu_snapshot_restore(rbgpu_world *w, const rbgpu_snapshot *s) {
    if (!w || !s || !s->dev) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    const rbgpu_snapshot::Header &h = s->hd;
    if (int rc = snap_check(h)) return rc;
    if (!w->built) {
        int rc = w->build();
        if (rc) return rc;
    }
    if (h.N != w->N || h.M != w->M || h.S != w->S || h.NEP != w->NEP || h.NJ != w->NJ || h.nscenes != w->nscenes ||
        h.nhulls != (uint32_t)w->hh.size())
        return fail(RBGPU_ERR_INVALID, "snapshot of a world with other bodies, colliders, joints, scenes or hulls");
    if (int rc = w->sync()) return rc;
    for (int k = 0; k < SNAP_OPTIONS; ++k)
        if (int rc = rbgpu_set_option(w, k, h.options[k])) return rc;
    if (int rc = rbgpu_set_step(w, h.timestep, h.substeps)) return rc;
    const std::vector<SnapSec> sec = snap_sections(w, h);
    const std::vector<size_t> off = snap_offsets(sec);
    if (off.back() != h.payload || h.payload > s->cap) return fail(RBGPU_ERR_INVALID, "corrupt snapshot (layout)");
    for (size_t k = 0; k + 2 < sec.size(); ++k)
        if (sec[k].bytes)
            HIPCHK(hipMemcpyAsync(sec[k].world, s->dev + off[k], sec[k].bytes, hipMemcpyDeviceToDevice, w->st));
    if (h.ax_H > w->ax_cap) {
        HIPCHK(hipStreamSynchronize(w->st));
        dfree(w->ax_keys); dfree(w->ax_vals);
        w->ax_cap = 0;
        if (dalloc(w->ax_keys, h.ax_H) | dalloc(w->ax_vals, h.ax_H)) return RBGPU_ERR_OOM;
        w->ax_cap = h.ax_H;
    }
    if (h.ax_H) {
        HIPCHK(hipMemcpyAsync(w->ax_keys, s->dev + off[sec.size() - 2], (size_t)h.ax_H * 8, hipMemcpyDeviceToDevice, w->st));
        HIPCHK(hipMemcpyAsync(w->ax_vals, s->dev + off[sec.size() - 1], (size_t)h.ax_H * 16, hipMemcpyDeviceToDevice,
                              w->st));
    }
    w->ax_restore = h.ax_H;
    w->frames = h.frames;
    // the sweep's per-endpoint arrays derive from the restored boxes and order only when the
    // frame marks them dirty: mark it (the stable sort of the restored, sorted order keeps it)
    k_set_int<<<1, 1, 0, w->st>>>(w->flags, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return RBGPU_OK;
}

int rbgpu_snapshot_size(const rbgpu_snapshot *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *bytes = sizeof(rbgpu_snapshot::Header) + s->hd.payload;
    return RBGPU_OK;
}

int rbgpu_snapshot_save(const rbgpu_snapshot *s, void *host, uint64_t cap) {
    if (!s || !host || !s->dev) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    const uint64_t need = sizeof(rbgpu_snapshot::Header) + s->hd.payload;
    if (cap < need) return fail(RBGPU_ERR_CAPACITY, "host buffer smaller than rbgpu_snapshot_size");
    auto *p = (unsigned char *)host;
    HIPCHK(hipMemcpy(p + sizeof(rbgpu_snapshot::Header), s->dev, s->hd.payload, hipMemcpyDeviceToHost));
    rbgpu_snapshot::Header h = s->hd;
    h.checksum = fnv1a(p + sizeof(h), h.payload);
    std::memcpy(p, &h, sizeof(h));
    return RBGPU_OK;
}

int rbgpu_snapshot_load(const void *host, uint64_t n, rbgpu_snapshot **out) {
    if (!host || !out) return fail(RBGPU_ERR_INVALID, "invalid arguments");
    *out = nullptr;
    rbgpu_snapshot::Header h;
    if (n < sizeof(h)) return fail(RBGPU_ERR_INVALID, "snapshot bytes shorter than the header");
    std::memcpy(&h, host, sizeof(h));
    if (int rc = snap_check(h)) return rc;
    if (n != sizeof(h) + h.payload) return fail(RBGPU_ERR_INVALID, "snapshot bytes of the wrong length");
    const auto *p = (const unsigned char *)host + sizeof(h);
    if (fnv1a(p, h.payload) != h.checksum) return fail(RBGPU_ERR_INVALID, "snapshot checksum mismatch");
    auto *s = new rbgpu_snapshot();
    if (int rc = snap_reserve(s, h.payload)) {
        delete s;
        return rc;
    }
    hipError_t e = hipMemcpy(s->dev, p, h.payload, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete s;
        return fail(RBGPU_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    s->hd = h;
    *out = s;
    return RBGPU_OK;
}

int rbgpu_snapshot_destroy(rbgpu_snapshot *s) {
    delete s;
    return RBGPU_OK;
}

const char *rbgpu_last_error(void) { return g_err.c_str(); }

}  // extern "C"
