"""Shared helpers for the parity tests."""
import numpy as np

from rbps_amd import scenes

COLLIDE_DIMS = {scenes.SPHERE: 1, scenes.BOX: 3, scenes.CAPSULE: 2, scenes.PLANE: 4}


def shape(kind, *dims):
    rec = np.zeros((), scenes.COLLIDER_DTYPE)
    rec["shape"] = kind
    rec["dims"] = list(dims) + [0.0] * (4 - len(dims))
    rec["local_rot"] = (1, 0, 0, 0)
    return rec


def tf_at(x, y, z, q=(1.0, 0.0, 0.0, 0.0)):
    return np.array([x, y, z, *q], dtype=np.float64)


def state_close(a, b, rtol):
    """max |a-b| / max(1, |b|) over all state fields."""
    worst = 0.0
    for k in a:
        d = np.abs(a[k] - b[k]) / np.maximum(1.0, np.abs(b[k]))
        worst = max(worst, float(d.max()) if d.size else 0.0)
    return worst <= rtol, worst


def colouring_valid(pairs_bodies, colour, dynamic):
    """No two pairs of one colour share a dynamic body (colour < 0 = not coloured)."""
    seen = {}
    for (ba, bb), c in zip(pairs_bodies, colour):
        if c < 0:
            continue
        for b in (int(ba), int(bb)):
            if not dynamic[b]:
                continue
            key = (b, int(c))
            if key in seen:
                return False
            seen[key] = True
    return True


def greedy_by_priority(pairs_bodies, dynamic, mix32, active=None):
    """Sequential first-fit colouring in descending (mix32(p), p) priority of the active
    pairs (all pairs when active is None); inactive pairs get -1."""
    P = len(pairs_bodies)
    act = [p for p in range(P) if active is None or active[p]]
    order = sorted(act, key=lambda p: (mix32(p), p), reverse=True)
    by_body = {}
    for p in act:
        ba, bb = pairs_bodies[p]
        for b in (int(ba), int(bb)):
            if dynamic[b]:
                by_body.setdefault(b, []).append(p)
    col = [-1] * P
    for p in order:
        used = set()
        for b in (int(pairs_bodies[p][0]), int(pairs_bodies[p][1])):
            if dynamic[b]:
                used.update(col[q] for q in by_body[b] if q != p and col[q] >= 0)
        c = 0
        while c in used:
            c += 1
        col[p] = c
    return np.array(col, np.int32)


def filtered_ref_pairs(pairs, colliders):
    """The narrowphase driver's filters (CollisionPipeline.cpp:121-133) on a pair list."""
    keep = []
    for a, b in pairs:
        ca, cb = colliders[a], colliders[b]
        if ca["is_static"] and cb["is_static"]:
            continue
        if ca["body_id"] == cb["body_id"]:
            continue
        keep.append((a, b))
    return np.array(keep, np.uint32).reshape(-1, 2)


def sorted_pairs(p):
    p = np.asarray(p, np.uint32).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def trial_colouring(pairs_bodies, dynamic, active, mix32, slots=16, palette_max=8):
    """Independent replay of the active-pair colouring (synchronous randomized rounds with
    per-body slot claims).  Round r, every uncoloured pair p with dynamic bodies D(p):
    t(p) = the (tent mod min(max(ceil(k/4), 1), 8))-th colour used by no pair of D(p) so far,
    k = uncoloured pairs at D(p) summed; key(p) = (r + 1, rank); p keeps t(p) iff its key is
    the largest among this round's uncoloured pairs of each v in D(p) whose t has the
    same residue mod `slots`."""
    M = 0xffffffff
    tent_h = lambda p, r: mix32(p ^ ((r * 0x9e3779b9) & M))
    rank_h = lambda p, r: mix32(p ^ ((r * 0x85ebca6b) & M) ^ 0xc2b2ae35)
    P = len(pairs_bodies)
    act = [p for p in range(P) if active[p]]
    dyn_of = {p: [int(b) for b in pairs_bodies[p] if dynamic[int(b)]] for p in act}
    used, unc = {}, {}
    for p in act:
        for b in dyn_of[p]:
            used.setdefault(b, set())
            unc[b] = unc.get(b, 0) + 1
    col = {p: -1 for p in act}
    r = 0
    while any(c < 0 for c in col.values()):
        tent, key, best = {}, {}, {}
        for p in act:
            if col[p] >= 0:
                continue
            taken = set().union(*[used[b] for b in dyn_of[p]]) if dyn_of[p] else set()
            k = sum(unc[b] for b in dyn_of[p])
            idx = tent_h(p, r) % min(max((k + 3) // 4, 1), palette_max)
            c = 0
            while c in taken or idx > 0:
                if c not in taken:
                    idx -= 1
                c += 1
            tent[p], key[p] = c, ((r + 1) << 32) | rank_h(p, r)
            for b in dyn_of[p]:
                best[(b, c % slots)] = max(best.get((b, c % slots), 0), key[p])
        won = [p for p in tent if all(best[(b, tent[p] % slots)] == key[p] for b in dyn_of[p])]
        for p in won:
            col[p] = tent[p]
            for b in dyn_of[p]:
                used[b].add(tent[p])
                unc[b] -= 1
        r += 1
    out = np.full(P, -1, np.int32)
    for p, c in col.items():
        out[p] = c
    return out, r


def unpack_hulls(g, prefix="hull_"):
    """Hull list of a fixture written by make_golden.pack_hulls."""
    v, voff, f, foff, has = (g[prefix + k] for k in ("v", "voff", "f", "foff", "has_faces"))
    return [(v[voff[i]:voff[i + 1]], f[foff[i]:foff[i + 1]] if has[i] else None) for i in range(len(has))]


def scene_from_golden(g):
    """The scene of a world fixture (bodies, colliders, joints and hulls when present)."""
    sc = scenes.Scene("g", g["bodies"], g["colliders"], float(g["timestep"]), int(g["substeps"]))
    if "joints" in g:
        sc.joints = g["joints"]
    if "hull_v" in g:
        sc.hulls = unpack_hulls(g)
    return sc


def sat_depth(va, fa, vb, fb):
    """Penetration depth of two convex polytopes (world vertices, triangles) by brute
    force: the minimum over the Minkowski-difference face normals (the triangle normals of
    both hulls and the cross products of their edge directions, both signs) of
    max(a.n) - min(b.n).  Negative: separated (then a separating axis exists among them)."""
    def tri_normals(v, f):
        n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
        return n / np.linalg.norm(n, axis=1, keepdims=True)

    def edges(v, f):
        e = np.concatenate([v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 1]], v[f[:, 0]] - v[f[:, 2]]])
        return e / np.linalg.norm(e, axis=1, keepdims=True)
    ax = [tri_normals(va, fa), tri_normals(vb, fb)]
    ea, eb = edges(va, fa), edges(vb, fb)
    c = np.cross(ea[:, None, :], eb[None, :, :]).reshape(-1, 3)
    ln = np.linalg.norm(c, axis=1)
    ax.append(c[ln > 1e-9] / ln[ln > 1e-9, None])
    n = np.concatenate(ax)
    n = np.concatenate([n, -n])
    d = (va @ n.T).max(0) - (vb @ n.T).min(0)
    k = int(np.argmin(d))
    return float(d[k]), n[k]


def colouring_valid_np(pairs_bodies, colour, dynamic):
    """colouring_valid for large pair lists: no (dynamic body, colour) key twice."""
    pb = np.asarray(pairs_bodies, np.int64).reshape(-1, 2)
    col = np.asarray(colour, np.int64)
    keys = []
    for side in (0, 1):
        sel = (col >= 0) & dynamic[pb[:, side]]
        keys.append(pb[sel, side] * (1 << 20) + col[sel])
    k = np.concatenate(keys)
    return len(np.unique(k)) == len(k)
