"""Tile-kernel geometry: the wave count (4-wave 32x32x16 / 32x32x32 tiles,
8-wave 32x32x32 tiles) and the grouping of consecutive ijk in a launch.

- Grouping only reorders the items of a launch; each item writes its energy
  at its own place in the work list, so per-ijk energies are bitwise equal
  for every group size (including groups cut short at the end of a range
  and ranges that are not multiples of the group).
- The two wave counts tile (a, b, c) differently, so the per-item sums
  differ in order: both agree with the oracle to the FP64 bound."""

import os

import numpy as np
import pytest

import oracle
from libpt_amd import api

pytestmark = pytest.mark.gpu


class _env:
    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kw}
        os.environ.update({k: str(v) for k, v in self.kw.items()})

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _ijk_energies(kind, form, dims, dev, ranges):
    with api.Plan(kind, form, dims, dev) as pl:
        geo = pl.geometry()
        for sc in range(pl.ncases):
            n = pl.nijk(sc)
            for lo, hi in ranges:
                pl.run(sc, min(lo, n), min(hi, n))
        pl.sync()
        return geo, [pl.ijk_energies(sc) for sc in range(pl.ncases)]


@pytest.mark.parametrize("kind,form,dims,kw", [
    ("rpt", "fp64", (7, 45), dict(sig_f=1e-2)),
    ("upt", "bf16x3", (6, 33, 5, 40), dict(sig_f=1e-2)),
    ("rft", "fp64", (6, 30), dict(sig_f=1e-2, omega=0.3)),
    ("rpt", "fp32", (7, 45), dict(sig_f=1e-2)),
])
def test_ijk_groups_do_not_change_energies(kind, form, dims, kw):
    spec = api.make_spec(kind, dims, seed=17, **kw)
    dev = api.synth_device(kind, spec, fp32=form == "fp32")
    # ranges cut across groups: [0, 5), [5, 13), [13, end)
    ranges = [(0, 5), (5, 13), (13, 10 ** 6)]
    try:
        ref = None
        for g in (1, 3, 8, 64):
            with _env(LIBPT_AMD_KGROUP=g):
                geo, e = _ijk_energies(kind, form, dims, dev, ranges)
            assert geo[2] == g
            print(kind, form, "group", g, [x.sum() for x in e])
            if ref is None:
                ref = e
            else:
                for a, b in zip(ref, e):
                    assert np.array_equal(a, b)
    finally:
        dev.close()


@pytest.mark.parametrize("form", ["fp64", "bf16x3", "fp32"])
@pytest.mark.parametrize("kind,dims,kw", [
    ("rpt", (8, 64), dict(sig_f=1e-2)),
    ("uft", (4, 18, 3, 21), dict(sig_f=1e-2, omega=0.3)),
])
def test_both_wave_counts_match_the_oracle(form, kind, dims, kw):
    spec = api.make_spec(kind, dims, seed=19, **kw)
    arr = api.synth_host(kind, spec)
    eo = oracle.energy(kind, dims, arr, mp=form == "fp32")
    tol = {"fp64": 1e-12, "bf16x3": 5e-4 if "ft" in kind else 1e-4,
           "fp32": 1e-5}[form]
    for w in (4, 8):
        with _env(LIBPT_AMD_WAVES=w):
            dev = api.to_device(kind, arr, fp32=form == "fp32")
            try:
                with api.Plan(kind, form, dims, dev) as pl:
                    assert pl.geometry()[0] == w
                    pl.run_all()
                    e = pl.energy()
            finally:
                dev.close()
        print(kind, form, "waves", w, e, eo, abs(e - eo) / abs(eo))
        assert abs(e - eo) <= tol * abs(eo)
