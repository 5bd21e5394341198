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


def trial_colouring(pairs_bodies, dynamic, active, mix32):
    """Independent replay of the active-pair colouring (synchronous randomized trials):
    round r, uncoloured pair p picks the (mix32(p ^ r*0x9e3779b9) mod k)-th colour unused by
    neighbours coloured in earlier rounds (k = uncoloured neighbour entries + 1) and keeps it
    unless an uncoloured neighbour of higher priority (mix32(q), q) picked the same."""
    P = len(pairs_bodies)
    act = [p for p in range(P) if active[p]]
    by_body = {}
    for p in act:
        for b in (int(pairs_bodies[p][0]), int(pairs_bodies[p][1])):
            if dynamic[b]:
                by_body.setdefault(b, []).append(p)
    nbrs = {p: [q for b in (int(pairs_bodies[p][0]), int(pairs_bodies[p][1])) if dynamic[b]
                for q in by_body[b] if q != p] for p in act}
    col = {p: -1 for p in act}
    r = 0
    while any(c < 0 for c in col.values()):
        tent = {}
        for p in act:
            if col[p] >= 0:
                continue
            used = {col[q] for q in nbrs[p] if col[q] >= 0}
            k = 1 + sum(1 for q in nbrs[p] if col[q] < 0)
            idx = mix32(p ^ ((r * 0x9e3779b9) & 0xffffffff)) % k
            c = 0
            while True:
                if c not in used:
                    if idx == 0:
                        break
                    idx -= 1
                c += 1
            tent[p] = c
        acc = [p for p in tent if not any(col[q] < 0 and tent[q] == tent[p] and (mix32(q), q) > (mix32(p), p)
                                          for q in nbrs[p])]
        for p in acc:
            col[p] = tent[p]
        r += 1
    out = np.full(P, -1, np.int32)
    for p, c in col.items():
        out[p] = c
    return out, r
