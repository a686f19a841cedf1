"""The epilogue's operand staging in every kernel instance against the oracle.

Which epilogue operands a kernel stages in LDS, and how many b it takes per
pass over W, depends on the form, the number of W sets and the wave count.
Every combination runs here at sizes whose last a, b and c tiles are partial
(the staged rows and columns beyond the arrays repeat the last one and are
masked), with more than one tile per ijk and enough occupied orbitals for
both spin cases."""

import os

import numpy as np
import pytest

import oracle
from libpt_amd import api

pytestmark = pytest.mark.gpu

# (kind, dims): v not a multiple of 32 and above one tile; unrestricted
# cases with different extents for the two spins
CASES = [("rpt", (5, 45)), ("upt", (5, 41, 4, 37)), ("rft", (4, 39)),
         ("uft", (4, 35, 5, 43))]
TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4, "f16x3": 1e-4}


class _waves:
    def __init__(self, w):
        self.w = str(w)

    def __enter__(self):
        self.old = os.environ.get("LIBPT_AMD_WAVES")
        os.environ["LIBPT_AMD_WAVES"] = self.w

    def __exit__(self, *a):
        if self.old is None:
            os.environ.pop("LIBPT_AMD_WAVES", None)
        else:
            os.environ["LIBPT_AMD_WAVES"] = self.old


@pytest.mark.parametrize("waves", [4, 8])
@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x3", "f16x3"])
@pytest.mark.parametrize("kind,dims", CASES,
                         ids=[f"{k}-{'x'.join(map(str, d))}" for k, d in CASES])
def test_staged_epilogue_matches_oracle(kind, dims, form, waves):
    ft = kind.endswith("ft")
    spec = api.make_spec(kind, dims, seed=sum(dims) + waves,
                         omega=0.3 if ft else 0.0, sig_f=1e-2)
    arr = api.synth_host(kind, spec)
    eo = oracle.energy(kind, dims, arr, mp=form == "fp32")
    tol = TOL[form] * (5 if ft and form in ("bf16x3", "f16x3") else 1)
    with _waves(waves):
        dev = api.to_device(kind, arr, fp32=form == "fp32")
        try:
            with api.Plan(kind, form, dims, dev) as pl:
                assert pl.geometry()[0] == waves
                pl.run_all()
                e = pl.energy()
        finally:
            dev.close()
    print(kind, dims, form, "waves", waves, e, eo, abs(e - eo) / abs(eo))
    assert np.isfinite(e)
    assert abs(e - eo) <= tol * abs(eo)
