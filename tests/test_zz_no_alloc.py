"""libpt_amd_energy_noalloc (every buffer from the library's static
scratch, no device allocation) against libpt_amd_energy on the same inputs:
the same kernels over the same work list, so the energies agree bit for
bit.  Two calls in a row per case, the second smaller, so the second runs
on scratch that still holds the first call's contents."""

import numpy as np
import pytest

from libpt_amd import api

pytestmark = pytest.mark.gpu

# (call, form, first dims, second (smaller) dims, generator args)
CASES = [
    ("rpt", "fp64", (8, 64), (6, 37), dict(sig_f=1e-2)),
    ("rpt", "fp32", (7, 50), (5, 33), dict(sig_f=1e-2)),
    ("upt", "bf16x3", (5, 40, 4, 44), (4, 20, 3, 23), dict(sig_f=1e-2)),
    ("rft", "fp64", (6, 48), (4, 30), dict(sig_f=1e-2, omega=0.3)),
    ("uft", "f16x3", (4, 30, 3, 33), (3, 17, 2, 19), dict(sig_f=1e-2,
                                                          omega=0.3)),
]


def _one(kind, form, dims, kw, seed):
    fp32 = form == "fp32"
    spec = api.make_spec(kind, dims, seed=seed, **kw)
    dev = api.synth_device(kind, spec, fp32=fp32)
    try:
        host = [api.from_device(p, n, fp32) for p, n in zip(dev.ptrs,
                                                            dev.sizes)]
        print(f"noalloc {kind} {form} {dims}", flush=True)
        e_na = api.energy_noalloc(kind, form, dims, dev)
        e_ref = api.energy(kind, form, dims, host)
    finally:
        dev.close()
    print(f"  noalloc {e_na!r} energy {e_ref!r}", flush=True)
    return e_na, e_ref


@pytest.mark.parametrize("kind,form,d1,d2,kw", CASES)
def test_noalloc_matches_energy(kind, form, d1, d2, kw):
    for dims, seed in ((d1, 5), (d2, 6)):
        e_na, e_ref = _one(kind, form, dims, kw, seed)
        assert np.isfinite(e_na)
        assert e_na == e_ref


def test_noalloc_rejects_what_does_not_fit():
    dims = (24, 200)
    spec = api.make_spec("rpt", dims, seed=1)
    dev = api.synth_device("rpt", spec)
    try:
        with pytest.raises(api.LibptError, match="static scratch"):
            api.energy_noalloc("rpt", "fp64", dims, dev)
    finally:
        dev.close()
    # the scratch is still usable afterwards
    e_na, e_ref = _one("rpt", "fp64", (5, 21), dict(sig_f=1e-2), 7)
    assert e_na == e_ref
