"""Writes tests/golden/gjk_cycling_pairs.npz: hull pairs on which FP32 GJK cycles.

The poses come from the combined workload (1,024 scenes, 262,144 bodies) stepped 120
frames on the GPU (whose trajectory equals the FP32 oracle's bit for bit); its GJK pairs
are run through the pair entry, those ending GJK_FAILED are kept when the FP32 oracle's
trace shows a repeated loop state.  Needs an MI355X for the first step:

    python tests/golden/make_gjk_cycling.py dump     # on the GPU box -> gjk_failed_pairs.npz
    python tests/golden/make_gjk_cycling.py write    # here -> gjk_cycling_pairs.npz
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from rbps_amd import scenes  # noqa: E402

DUMP = os.path.join(ROOT, "gpurun_out", "gjk_failed_pairs.npz")


def dump():
    from rbps_amd import world
    sc = scenes.combined_batch()
    w = world.World(sc.timestep, sc.substeps).load(sc)
    w.set_option("joint_mode", 1)
    w.set_option("contact_mode", 1)
    w.step(120)
    pairs = w.sap_pairs()
    k = sc.colliders["shape"][pairs]
    hp = pairs[(k == scenes.HULL).any(1) & ~(k == scenes.PLANE).any(1)]
    r = w.collide_pairs(hp)
    st = w.state()
    os.makedirs(os.path.dirname(DUMP), exist_ok=True)
    np.savez_compressed(DUMP, pairs=hp[r["gjk_status"] == 2], position=st["position"], orientation=st["orientation"])


def write(limit=48):
    from oracle import orc
    import make_golden as mg
    sc = scenes.combined_batch()
    d = np.load(DUMP, allow_pickle=False)
    A, B, TA, TB, TR = [], [], [], [], []
    for a, b in d["pairs"]:
        ba, bb = sc.colliders["body_id"][a], sc.colliders["body_id"][b]
        ta = np.concatenate([d["position"][ba], d["orientation"][ba]])
        tb = np.concatenate([d["position"][bb], d["orientation"][bb]])
        t = orc.oracle_gjk_trace(sc.colliders[a], ta, sc.colliders[b], tb, fp32=True, hulls=sc.hulls)
        if t["gjk_status"] == orc.GJK_FAILED and t["first_repeat"] >= 0:
            A.append(sc.colliders[a]); B.append(sc.colliders[b]); TA.append(ta); TB.append(tb)
            TR.append([t["gjk_status"], t["gjk_iters"], t["first_repeat"], t["period"]])
        if len(A) >= limit:
            break
    A, B = np.array(A), np.array(B)
    used = sorted({int(h) for arr in (A, B) for h in arr["hull_index"][arr["shape"] == scenes.HULL]})
    remap = {h: i for i, h in enumerate(used)}
    for arr in (A, B):
        sel = arr["shape"] == scenes.HULL
        arr["hull_index"][sel] = [remap[int(h)] for h in arr["hull_index"][sel]]
    out = mg.pack_hulls([sc.hulls[h] for h in used])
    out.update(a=A, b=B, tfa=np.array(TA), tfb=np.array(TB), trace=np.array(TR, np.int32))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "gjk_cycling_pairs.npz"), **out)
    print(len(A), "pairs")


if __name__ == "__main__":
    {"dump": dump, "write": write}[sys.argv[1]]()
