"""Tile geometry, tile-order blocking and their autotuner
(libpt_amd_plan_set_geometry / libpt_amd_plan_set_block /
libpt_amd_plan_tune): the order of the tiles changes only the summation
order of the per-tile energies, so every blocking gives the same energy to
rounding; the two geometries (4 waves, 16-wide c tiles in FP64 and
32-wide in the other forms; 8 waves, 32-wide c tiles) tile (a, b, c)
differently and both match the oracle; the
tuned plan matches the oracle, and a new plan of the same sizes takes the
tuned geometry and blocking from the cache."""

import os

import numpy as np
import pytest

import oracle
from libpt_amd import api

pytestmark = pytest.mark.gpu

CANDIDATES = {(1, 1, 64), (2, 2, 8), (4, 4, 2), (4, 4, 4), (8, 4, 2),
              (8, 8, 1), (8, 8, 2), (16, 16, 1), (4, 2, 8), (2, 2, 16)}


@pytest.mark.parametrize("kind,dims,form", [
    ("rpt", (7, 70), "fp64"), ("upt", (5, 36, 4, 41), "fp64"),
    ("rpt", (7, 70), "bf16x3")])
def test_blockings_agree(kind, dims, form):
    spec = api.make_spec(kind, dims, seed=21, sig_f=1e-2)
    dev = api.synth_device(kind, spec)
    try:
        with api.Plan(kind, form, dims, dev) as pl:
            es = []
            for blk in [(1, 1, 64), (3, 2, 5), (16, 16, 1), (1, 1, 1)]:
                pl.set_block(*blk)
                pl.run_all()
                es.append(pl.energy())
            print(kind, form, es)
            for e in es[1:]:
                assert abs(e - es[0]) <= 1e-13 * abs(es[0])
    finally:
        dev.close()


def test_tune_picks_candidate_and_keeps_parity():
    kind, dims = "rpt", (8, 64)
    spec = api.make_spec(kind, dims, seed=5, sig_f=1e-2)
    dev = api.synth_device(kind, spec)
    try:
        with api.Plan(kind, "fp64", dims, dev) as pl:
            blk = pl.tune(sample_ijk=4)
            assert blk in CANDIDATES
            waves = pl.geometry()[0]
            assert waves in (4, 8)
            pl.run_all()
            e = pl.energy()
        # a new plan of the same sizes takes the tuned geometry and
        # blocking from the cache and gives the same energy bit for bit
        with api.Plan(kind, "fp64", dims, dev) as pl2:
            assert pl2.geometry()[0] == waves
            pl2.run_all()
            assert pl2.energy() == e
    finally:
        dev.close()
    eo = oracle.energy(kind, dims, api.synth_host(kind, spec))
    print("tuned", blk, e, "oracle", eo)
    assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))


@pytest.mark.parametrize("kind,dims,form,kw,tol", [
    ("rpt", (6, 40), "fp64", dict(sig_f=1e-2), 1e-12),
    ("upt", (5, 36, 4, 41), "bf16x3", dict(sig_f=1e-2), 1e-4),
    ("rft", (6, 30), "fp64", dict(sig_f=1e-2, omega=0.3), 1e-12),
    ("uft", (4, 18, 3, 21), "f16x3", dict(sig_f=1e-2, omega=0.3), 1e-4),
    ("rpt", (6, 40), "fp32", dict(sig_f=1e-2), 1e-5),
])
def test_geometry_switch_in_place(kind, dims, form, kw, tol):
    """One plan run in one geometry, switched to the other (tile lists,
    item slots and (fT) save areas rebuilt) and back: every run matches
    the oracle, and switching back gives the first energy bit for bit."""
    spec = api.make_spec(kind, dims, seed=23, **kw)
    arr = api.synth_host(kind, spec)
    eo = oracle.energy(kind, dims, arr, mp=form == "fp32")
    dev = api.to_device(kind, arr, fp32=form == "fp32")
    try:
        with api.Plan(kind, form, dims, dev) as pl:
            w0 = pl.geometry()[0]
            es = []
            for w in (w0, 12 - w0, w0):
                pl.set_geometry(w)
                assert pl.geometry()[0] == w
                pl.run_all()
                es.append(pl.energy())
            with pytest.raises(api.LibptError):
                pl.set_geometry(6)
    finally:
        dev.close()
    print(kind, form, "waves", w0, 12 - w0, es, "oracle", eo)
    assert es[2] == es[0]
    for e in es:
        assert abs(e - eo) <= tol * abs(eo)


def test_plans_read_records_from_the_cache_file(tmp_path):
    """A record in LIBPT_AMD_TUNE_CACHE (written by an earlier process)
    sets the geometry and blocking of new plans of its sizes in a fresh
    process; older 9-field lines (no geometry) keep the default geometry."""
    import subprocess
    import sys
    import textwrap
    from conftest import ROOT
    cache = tmp_path / "tune.txt"
    cache.write_text("0 0 6 45 6 45 4 4 4\n0 0 7 45 7 45 2 2 8 8\n")
    code = textwrap.dedent("""
        from libpt_amd import api
        for dims in ((6, 45), (7, 45)):
            spec = api.make_spec("rpt", dims, seed=3, sig_f=1e-2)
            dev = api.synth_device("rpt", spec)
            with api.Plan("rpt", "fp64", dims, dev) as pl:
                print(dims, pl.geometry()[0], flush=True)
            dev.close()
    """)
    env = dict(os.environ, LIBPT_AMD_TUNE_CACHE=str(cache))
    env.pop("LIBPT_AMD_WAVES", None)
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "(6, 45) 4" in r.stdout
    assert "(7, 45) 8" in r.stdout
