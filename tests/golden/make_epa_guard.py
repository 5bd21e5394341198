"""Writes tests/golden/epa_guard_pairs.npz: two convex hulls of the library (scenes.hull_library)
put face to face and touching, where the FP32 restatement of EPA settles on a face that is not
the minimum (depth of centimetres at zero penetration) while FP64 finds the contact; and the
same pair pressed 1 mm together, where FP32 is right.  With the face normal as the guard's
direction the guarded FP32 result is the contact.  Searches seeded face-to-face placements
(touching found by bisection on the FP64 GJK); uses the oracle only.

    PYTHONPATH=. python3 tests/golden/make_epa_guard.py
"""
import os

import numpy as np

from oracle import orc
from rbps_amd import scenes as S

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "epa_guard_pairs.npz")


def desc(i):
    r = np.zeros((), S.COLLIDER_DTYPE)
    r["shape"] = S.HULL
    r["hull_index"] = i
    r["local_rot"] = (1, 0, 0, 0)
    return r


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def qrot(q, v):
    u = q[1:]
    return v + 2 * q[0] * np.cross(u, v) + 2 * np.cross(u, np.cross(u, v))


def q_from_to(a, b):
    a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
    c, d = np.cross(a, b), np.dot(a, b)
    if d < -0.999999:
        ax = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(ax) < 1e-6:
            ax = np.cross(a, [0.0, 1.0, 0.0])
        return np.array([0.0, *(ax / np.linalg.norm(ax))])
    q = np.array([1 + d, *c])
    return q / np.linalg.norm(q)


def main(seed=7, trials=20000):
    rng = np.random.default_rng(seed)
    hulls = [(v, f) for v, f in S.hull_library()]
    pns = [orc.hull_tables("orc_", v, f)["pn"] for v, f in hulls]
    for _ in range(trials):
        ia, ib = (int(x) for x in rng.integers(0, len(hulls), 2))
        qa = rng.normal(size=4)
        qa /= np.linalg.norm(qa)
        fa, fb = pns[ia][rng.integers(len(pns[ia]))], pns[ib][rng.integers(len(pns[ib]))]
        na = qrot(qa, fa)
        ang = rng.uniform(0, 2 * np.pi)
        qb = qmul(np.array([np.cos(ang / 2), *(np.sin(ang / 2) * na)]), q_from_to(fb, -na))
        off = rng.normal(size=3) * 0.05
        off -= np.dot(off, na) * na
        tfa = np.array([0.0, 0.0, 0.0, *qa])
        lo, hi = 0.0, 4.0
        for _ in range(34):
            mid = 0.5 * (lo + hi)
            r = orc.oracle_gjk_epa(desc(ia), tfa, desc(ib), np.array([*(na * mid + off), *qb]), hulls=hulls)
            lo, hi = (mid, hi) if r["gjk_status"] == 1 else (lo, mid)
        pen = rng.choice([-1e-6, -1e-7, 0.0, 1e-7, 1e-6, 1e-5])
        tfb = np.array([*(na * (lo - pen) + off), *qb])
        r64 = orc.oracle_gjk_epa(desc(ia), tfa, desc(ib), tfb, hulls=hulls)
        r32 = orc.oracle_gjk_epa(desc(ia), tfa, desc(ib), tfb, fp32=True, hulls=hulls)
        if r64["hit"] and r64["depth"] < 1e-6 and r32["hit"] and r32["depth"] > 0.01:
            pressed = np.array([*(na * (lo - 1e-3) + off), *qb])
            np.savez_compressed(OUT, va=hulls[ia][0], fa=hulls[ia][1], vb=hulls[ib][0], fb=hulls[ib][1], tfa=tfa,
                                tfb=tfb, tfb_pressed=pressed, guard=na, seed=seed)
            print("written", OUT, "fp64 depth", r64["depth"], "fp32 depth", r32["depth"])
            return
    raise SystemExit("no failing pair found")


if __name__ == "__main__":
    main()
