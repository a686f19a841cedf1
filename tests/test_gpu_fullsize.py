"""The full-size configurations -- rpt o=45 v=477 (the throughput
measurement: 74 GB of FP64 inputs in HBM, packed operands beside them), upt
24/200/20/204, rft 30/300 and uft 16/160/14/160 -- through size-independent
properties on the
first ijk triples of each spin case:
- repeated runs give bitwise the same per-ijk energies;
- doubling the denominators (d_ov -> 2 d_ov, exact in binary) halves every
  per-ijk energy bit for bit: W and Z do not depend on d_ov, every element's
  term is divided by exactly twice its denominator, and compensated sums of
  halved terms are the halved sums;
- doubling the amplitudes multiplies every per-ijk energy by exactly 4 for
  (T) (t1, t2 and f_ov: W and Z both double) and by exactly 2 for (fT) (t2
  alone: only W_R depends on it);
- a split form agrees with FP64 on the same triples (split-BF16 to 1e-6
  relative for (T); for (fT), whose signed products cancel, F16X3 to 1e-4
  and split-BF16 to 5e-4).  Per-ijk energies at o=45 v=477 against the
  oracle itself are in test_gpu_goldens.py."""

import numpy as np
import pytest

from libpt_amd import api

pytestmark = pytest.mark.gpu

NIJK = 3   # triples per spin case


def _slice(plan):
    out = []
    for sc in range(plan.ncases):
        m = min(NIJK, plan.nijk(sc))
        plan.run(sc, 0, m)
        plan.sync()
        out.append(plan.ijk_energies(sc)[:m].copy())
    return out


@pytest.mark.parametrize("kind,dims,omega,split,bound", [
    ("rpt", (45, 477), 0.0, "bf16x3", 1e-6),
    ("upt", (24, 200, 20, 204), 0.0, "bf16x3", 1e-6),
    ("rft", (30, 300), 0.3, "f16x3", 1e-4),
    ("uft", (16, 160, 14, 160), 0.3, "f16x3", 1e-4),
    ("rft", (30, 300), 0.3, "bf16x3", 5e-4),
    ("uft", (16, 160, 14, 160), 0.3, "bf16x3", 5e-4),
])
def test_full_size_properties(kind, dims, omega, split, bound):
    spec = api.make_spec(kind, dims, seed=2024, omega=omega)
    dev = api.synth_device(kind, spec)
    d2 = a2 = None
    try:
        with api.Plan(kind, "fp64", dims, dev) as pl:
            e1 = _slice(pl)
            e1b = _slice(pl)
        print(kind, "fp64", [a.tolist() for a in e1])
        for a, b in zip(e1, e1b):
            assert np.array_equal(a, b)
            assert np.all(np.isfinite(a)) and np.all(a != 0)
        # d_ov is the first pt.h array
        dov = api.from_device(dev.ptrs[0], dev.sizes[0])
        d2 = api.to_device(kind, [2.0 * dov])
        dev2 = api.DeviceArrays(kind, spec, [d2.ptrs[0]] + dev.ptrs[1:],
                                dev.sizes, False)
        with api.Plan(kind, "fp64", dims, dev2) as pl:
            e2 = _slice(pl)
        for a, b in zip(e1, e2):
            assert np.array_equal(b, 0.5 * a)
        # amplitudes doubled: pt.h arrays 1-3 (f_ov, t1, t2) for (T), array
        # 3 (t2) for (fT)
        ft = kind.endswith("ft")
        idx = [3] if ft else [1, 2, 3]
        a2 = api.to_device(kind, [2.0 * api.from_device(dev.ptrs[t],
                                                        dev.sizes[t])
                                  for t in idx])
        ptrs = list(dev.ptrs)
        for t, p in zip(idx, a2.ptrs):
            ptrs[t] = p
        dev3 = api.DeviceArrays(kind, spec, ptrs, dev.sizes, False)
        with api.Plan(kind, "fp64", dims, dev3) as pl:
            e3 = _slice(pl)
        for a, b in zip(e1, e3):
            assert np.array_equal(b, (2.0 if ft else 4.0) * a)
        with api.Plan(kind, split, dims, dev) as pl:
            es = _slice(pl)
        for a, b in zip(e1, es):
            rel = abs(b.sum() - a.sum()) / abs(a.sum())
            print(kind, split, "rel", rel)
            assert rel <= bound
    finally:
        for x in (d2, a2):
            if x is not None:
                x.close()
        dev.close()
