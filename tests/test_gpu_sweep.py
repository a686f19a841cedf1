"""A seeded sweep of small and degenerate sizes through the drop-in, all
four calls, FP64 and FP32, against the oracle: sizes below one tile, odd
edges, and spin blocks of zero size (the reference's rule, src/pt.c:195-196
and :318-319: an empty spin case contributes 0)."""

import numpy as np
import pytest

import oracle
from libpt_amd import api

pytestmark = pytest.mark.gpu


def _cases(n=24, seed=77):
    rng = np.random.default_rng(seed)
    kinds = ["rpt", "upt", "rft", "uft"]
    out = [("upt", (3, 5, 0, 0)), ("upt", (0, 0, 3, 5)), ("rpt", (1, 1)),
           ("uft", (2, 3, 0, 4)), ("rpt", (0, 4)), ("rft", (4, 0))]
    for q in range(n):
        kind = kinds[q % 4]
        oa, va = int(rng.integers(0, 7)), int(rng.integers(0, 41))
        if kind in ("rpt", "rft"):
            out.append((kind, (oa, va)))
        else:
            out.append((kind, (oa, va, int(rng.integers(0, 7)),
                               int(rng.integers(0, 41)))))
    return out


CASES = _cases()


@pytest.mark.parametrize("kind,dims", CASES,
                         ids=[f"{k}-{'x'.join(map(str, d))}" for k, d in CASES])
@pytest.mark.parametrize("mp", [False, True], ids=["fp64", "fp32"])
def test_sweep(kind, dims, mp):
    spec = api.make_spec(kind, dims, seed=sum(dims) + 3,
                         omega=0.3 if kind.endswith("ft") else 0.0, sig_f=1e-2)
    arrays = api.synth_host(kind, spec)
    e = api.dropin(kind, dims, arrays, mp=mp)
    eo = oracle.energy(kind, dims, arrays, mp=mp)
    print(kind, dims, "mp" if mp else "fp64", "gpu", e, "oracle", eo)
    assert np.isfinite(e)
    if eo == 0.0:
        assert e == 0.0
    elif mp:
        assert abs(e - eo) <= 1e-12 + 1e-5 * abs(eo)
    else:
        assert abs(e - eo) <= 1e-16 + 1e-12 * abs(eo)
