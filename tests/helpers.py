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
    """No two pairs of one colour share a dynamic body."""
    seen = {}
    for (ba, bb), c in zip(pairs_bodies, colour):
        for b in (int(ba), int(bb)):
            if not dynamic[b]:
                continue
            key = (b, int(c))
            if key in seen:
                return False
            seen[key] = True
    return True


def greedy_by_priority(pairs_bodies, dynamic, mix32):
    """Sequential first-fit colouring in descending (mix32(p), p) priority."""
    P = len(pairs_bodies)
    order = sorted(range(P), key=lambda p: (mix32(p), p), reverse=True)
    by_body = {}
    for p, (ba, bb) in enumerate(pairs_bodies):
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
