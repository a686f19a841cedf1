"""The persistent grid (LIBPT_AMD_PERSISTENT=1): a grid of every CU's share
of workgroups takes the work items from one counter per XCD instead of one
workgroup per item.  Each item runs the same code on the same operands, so
per-ijk energies equal those of the default launch bit for bit, for both
geometries, with and without ijk groups, for (T) and for (fT) (where each
workgroup keeps one save area, so a spin case of more items than the
default launches' save areas runs in one launch), and inside batches."""

import os

import numpy as np
import pytest

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


def _ijk(kind, form, dims, dev):
    with api.Plan(kind, form, dims, dev) as pl:
        pl.run_all()
        e = pl.energy()
        return e, [pl.ijk_energies(sc) for sc in range(pl.ncases)]


@pytest.mark.parametrize("kind,form,dims,kw", [
    ("rpt", "fp64", (8, 64), dict(sig_f=1e-2)),
    ("upt", "bf16x3", (5, 36, 4, 41), dict(sig_f=1e-2)),
    ("rft", "fp64", (6, 30), dict(sig_f=1e-2, omega=0.3)),
    ("uft", "f16x3", (4, 18, 3, 21), dict(sig_f=1e-2, omega=0.3)),
    ("rpt", "fp32", (7, 45), dict(sig_f=1e-2)),
])
@pytest.mark.parametrize("waves,kgroup", [(4, 4), (4, 1), (8, 4)])
def test_persistent_equals_default(kind, form, dims, kw, waves, kgroup):
    spec = api.make_spec(kind, dims, seed=31, **kw)
    dev = api.synth_device(kind, spec, fp32=form == "fp32")
    try:
        with _env(LIBPT_AMD_WAVES=waves, LIBPT_AMD_KGROUP=kgroup,
                  LIBPT_AMD_PERSISTENT=0):
            e0, ijk0 = _ijk(kind, form, dims, dev)
        print(f"persistent {kind} {form} {dims} waves {waves} kgroup "
              f"{kgroup}", flush=True)
        with _env(LIBPT_AMD_WAVES=waves, LIBPT_AMD_KGROUP=kgroup,
                  LIBPT_AMD_PERSISTENT=1):
            e1, ijk1 = _ijk(kind, form, dims, dev)
    finally:
        dev.close()
    print(kind, form, waves, kgroup, e0, e1)
    assert e1 == e0
    for a, b in zip(ijk0, ijk1):
        assert np.array_equal(a, b)


def test_persistent_ft_beyond_one_default_launch():
    """rft o=12 v=160 in FP64: the mixed-spin case has more items than
    the default launches' 4 GB of save areas hold (several launches); the
    persistent grid runs it in one launch with one area per workgroup."""
    kind, dims = "rft", (12, 160)
    spec = api.make_spec(kind, dims, seed=33, sig_f=1e-2, omega=0.3)
    dev = api.synth_device(kind, spec)
    try:
        with _env(LIBPT_AMD_PERSISTENT=0):
            with api.Plan(kind, "fp64", dims, dev) as pl:
                assert max(pl.items(sc) for sc in range(pl.ncases)) > 32768
                pl.run_all()
                e0 = pl.energy()
        print("persistent rft (12, 160)", flush=True)
        with _env(LIBPT_AMD_PERSISTENT=1):
            e1, _ = _ijk(kind, "fp64", dims, dev)
    finally:
        dev.close()
    print(e0, e1)
    assert e1 == e0


def test_persistent_batch_equals_alone():
    dims_list = [(6, 37), (3, 50), (1, 9), (5, 12)]
    systems = []
    try:
        for n, d in enumerate(dims_list):
            systems.append((d, api.synth_device("rft", api.make_spec(
                "rft", d, seed=41 + n, sig_f=1e-2, omega=0.3))))
        with _env(LIBPT_AMD_PERSISTENT=1):
            print("persistent batch", flush=True)
            with api.Batch("rft", "fp64", systems) as b:
                b.run()
                eb = b.energies()
        with _env(LIBPT_AMD_PERSISTENT=0):
            for (d, dev), e in zip(systems, eb):
                ea, _ = _ijk("rft", "fp64", d, dev)
                print(d, e, ea)
                assert e == ea
    finally:
        for _, dev in systems:
            dev.close()
