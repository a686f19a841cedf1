"""Per-ijk energies of the oracle at the full restricted size (libpt_rpt,
o=45, v=477, seed 2024), written to tests/golden/ijk_rpt_45_477.json.

The whole input is 74 GB of FP64; an ijk item only reads the ovvv slabs of
its own three occupied indices (src/pt.c:236-254, :361-392), so this script
builds t2, oovo, oovv and the ov arrays whole and ovvv only for the slabs
the chosen items use: each ovvv block is allocated at full size with
np.zeros (untouched pages cost no memory) and only those slabs are filled
from the generator.  An item that read any other slab would see zeros and
disagree with the GPU, which reads the complete arrays; the GPU test would
then fail rather than pass.

Items: the first three and the last ijk triple of each spin case in
work-list order (the last ones read slabs 42-44, at byte offsets above
2^35 in the ovvv array).  Energies are per item, without the restricted
factor 2, in FP64 (oracle_pt_*_d, the restatement of src/pt.c) and in FP32
(oracle_pt_*_f, of src/ptmp.c, on float inputs).

    python tests/golden/make_ijk_goldens.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "ijk_rpt_45_477.json")

O, V, SEED = 45, 477, 2024


def items(same):
    """First three and last ijk triple of a spin case in work-list order
    (i<j<k for the same-spin case, i<j and any k for the mixed case)."""
    if same:
        first = [(0, 1, 2), (0, 1, 3), (0, 1, 4)]
        last = (O - 3, O - 2, O - 1)
        n = O * (O - 1) * (O - 2) // 6
    else:
        first = [(0, 1, 0), (0, 1, 1), (0, 1, 2)]
        last = (O - 2, O - 1, O - 1)
        n = O * (O - 1) // 2 * O
    return first + [last], [0, 1, 2, n - 1]


def build(fp32):
    from libpt_amd import api
    spec = api.make_spec("rpt", (O, V), seed=SEED)
    dt = np.float32 if fp32 else np.float64
    ids = api.tensor_ids("rpt")
    arr = []
    for t in ids[:6]:
        a = np.empty(api.synth_size(spec, t), dtype=dt)
        api.synth_range(spec, t, a, 0)
        arr.append(a)
    P = V * (V - 1) // 2
    slab_a, slab_b = V * P, V * V * V
    n_a = O * slab_a
    va = np.zeros(n_a, dtype=dt)
    vb = np.zeros(O * slab_b, dtype=dt)
    used = sorted({x for s in (True, False) for it in items(s)[0] for x in it})
    for k in used:
        api.synth_range(spec, ids[6], va[k * slab_a:(k + 1) * slab_a],
                        k * slab_a)
        api.synth_range(spec, ids[6], vb[k * slab_b:(k + 1) * slab_b],
                        n_a + k * slab_b)
    return arr, va, vb, used


def energies(fp32):
    import oracle
    arr, va, vb, used = build(fp32)
    d, f, t1, t2, oovo, oovv = arr
    n2, n3 = O * O * V * V, O * O * O * V
    tA, tB = t2[:n2], t2[n2:]
    oA, oB = oovo[:n3], oovo[n3:]
    gA, gB = oovv[:n2], oovv[n2:]
    out = {}
    for same in (True, False):
        its, idx = items(same)
        t0 = time.perf_counter()
        if same:
            _, e = oracle.spin_case("pt_aaa", (O, V),
                                    [d, f, t1, tA, oA, gA, va],
                                    np.array(its), mp=fp32)
        else:
            _, e = oracle.spin_case("pt_aab", (O, V, O, V),
                                    [d, d, f, f, t1, t1, tA, tB, tB, oA, oB,
                                     oB, gA, gB, va, vb, vb],
                                    np.array(its), mp=fp32)
        name = "aaa" if same else "aab"
        out[name] = {"ijk": [list(x) for x in its], "index": idx,
                     "energy": [float(x) for x in e]}
        print(f"[ijk goldens] {'fp32' if fp32 else 'fp64'} {name}: "
              f"{time.perf_counter() - t0:.0f} s {list(e)}", flush=True)
    return out, used


def main():
    data = {"call": "rpt", "dims": [O, V], "seed": SEED, "omega": 0.0,
            "sig_f": 0.0,
            "about": "oracle per-ijk energies (no factor 2) at the full "
                     "restricted size; written by "
                     "tests/golden/make_ijk_goldens.py"}
    data["fp64"], used = energies(False)
    data["fp32"], _ = energies(True)
    data["ovvv_slabs_filled"] = used
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
