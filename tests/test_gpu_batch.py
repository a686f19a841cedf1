"""Batched launch of many independent systems (libpt_amd_batch_*): every
system's energy must equal the energy of that system run alone through a
plan, bit for bit (same per-item arithmetic, same fixed-order reduction),
and match the oracle; the recorded graph must give the same energies on
every replay."""

import numpy as np
import pytest

import oracle
from libpt_amd import api

pytestmark = pytest.mark.gpu


def _systems(kind, dims_list, seed0=11, fp32=False, **kw):
    """(dims, spec, device arrays) per system; FP32 arrays for the FP32
    form (pt.h _mp element type)"""
    out = []
    for n, d in enumerate(dims_list):
        spec = api.make_spec(kind, d, seed=seed0 + n, **kw)
        out.append((d, spec, api.synth_device(kind, spec, fp32=fp32)))
    return out


def _close(systems):
    for _, _, dev in systems:
        dev.close()


def _alone(kind, form, d, dev):
    with api.Plan(kind, form, d, dev) as pl:
        pl.run_all()
        return pl.energy()


# uneven sizes: tile edges, o < 3 (no same-spin triples), o = 1 (no
# triples at all), v below one tile
RPT_DIMS = [(6, 37), (3, 50), (8, 64), (2, 20), (1, 9), (5, 12), (4, 33)]


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x3", "f16x3"])
def test_batch_rpt_matches_alone(form):
    systems = _systems("rpt", RPT_DIMS, fp32=form == "fp32", sig_f=1e-2)
    try:
        with api.Batch("rpt", form, [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            eb = b.energies()
        for (d, spec, dev), e in zip(systems, eb):
            ea = _alone("rpt", form, d, dev)
            print(form, d, "batch", e, "alone", ea)
            assert e == ea
            if d[0] == 1:
                assert e == 0.0
    finally:
        _close(systems)


def test_batch_rpt_oracle():
    systems = _systems("rpt", RPT_DIMS, sig_f=1e-2)
    try:
        with api.Batch("rpt", "fp64", [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            eb = b.energies()
        for (d, spec, _), e in zip(systems, eb):
            eo = oracle.energy("rpt", d, api.synth_host("rpt", spec))
            print(d, "batch", e, "oracle", eo)
            assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))
    finally:
        _close(systems)


@pytest.mark.parametrize("kind,dims_list,kw", [
    ("upt", [(5, 20, 4, 23), (4, 33, 6, 17), (3, 9, 2, 11)], dict(sig_f=1e-2)),
    ("rft", [(5, 30), (4, 17), (3, 40)], dict(sig_f=1e-2, omega=0.3)),
    ("uft", [(4, 18, 3, 21), (3, 12, 4, 10)], dict(sig_f=1e-2, omega=0.3)),
])
@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_batch_other_calls(kind, dims_list, kw, form):
    systems = _systems(kind, dims_list, fp32=form == "fp32", **kw)
    try:
        with api.Batch(kind, form, [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            eb = b.energies()
        for (d, spec, dev), e in zip(systems, eb):
            ea = _alone(kind, form, d, dev)
            eo = oracle.energy(kind, d, api.synth_host(kind, spec),
                               mp=form == "fp32")
            print(kind, form, d, "batch", e, "alone", ea, "oracle", eo)
            assert e == ea
            tol = max(1e-10, 1e-13 * abs(eo)) if form == "fp64" else \
                1e-5 * abs(eo)
            assert abs(e - eo) <= tol
    finally:
        _close(systems)


# relative bounds of the split forms against the FP64 oracle (as in the
# single-system parity tests; 5x for (fT), whose signed products cancel)
SPLIT_TOL = {"bf16x3": 1e-4, "f16x3": 2e-6}


@pytest.mark.parametrize("kind,dims_list,kw", [
    ("upt", [(5, 20, 4, 23), (3, 9, 2, 11)], dict(sig_f=1e-2)),
    ("rft", [(5, 30), (3, 40)], dict(sig_f=1e-2, omega=0.3)),
    ("uft", [(4, 18, 3, 21), (3, 12, 4, 10)], dict(sig_f=1e-2, omega=0.3)),
])
@pytest.mark.parametrize("form", ["bf16x3", "f16x3"])
def test_batch_split_forms(kind, dims_list, kw, form):
    systems = _systems(kind, dims_list, **kw)
    try:
        with api.Batch(kind, form, [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            eb = b.energies()
        tol = SPLIT_TOL[form] * (5 if "ft" in kind else 1)
        for (d, spec, dev), e in zip(systems, eb):
            ea = _alone(kind, form, d, dev)
            eo = oracle.energy(kind, d, api.synth_host(kind, spec))
            rel = abs(e - eo) / abs(eo)
            print(kind, form, d, "batch", e, "alone", ea, "oracle", eo,
                  "rel", rel)
            assert e == ea
            assert rel <= tol
    finally:
        _close(systems)


@pytest.mark.parametrize("kind,dims_list,kw", [
    ("rpt", [(7, 41), (5, 26), (7, 41), (4, 19)], dict(sig_f=1e-2)),
    ("rft", [(6, 35), (4, 22), (6, 35)], dict(sig_f=1e-2, omega=0.3)),
])
def test_batch_tune(kind, dims_list, kw):
    """Tuned tile orders: each distinct size is timed once (a second tune
    times none), and the tuned batch's energies equal, bit for bit, those of
    plans of the same sizes created afterwards (which take the recorded
    order), and match the oracle.  (fT) tuning runs in the batch's own save
    areas."""
    systems = _systems(kind, dims_list, seed0=40, **kw)
    try:
        with api.Batch(kind, "fp64", [(d, dev) for d, _, dev in systems]) as b:
            timed = b.tune(2)
            print("systems timed", timed)
            assert 0 <= timed <= len(set(dims_list))
            assert b.tune(2) == 0
            b.run()
            eb = b.energies()
        for (d, spec, dev), e in zip(systems, eb):
            ea = _alone(kind, "fp64", d, dev)
            eo = oracle.energy(kind, d, api.synth_host(kind, spec))
            print(d, "batch", e, "alone", ea, "oracle", eo)
            assert e == ea
            assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))
    finally:
        _close(systems)


def test_batch_graph_replays():
    """Replays of the recorded graph, with other eager work and
    synchronisation between them, all give the first replay's energies."""
    systems = _systems("rpt", RPT_DIMS[:4], sig_f=1e-2)
    other = _systems("upt", [(4, 21, 3, 19)], seed0=90)
    try:
        with api.Batch("rpt", "fp64", [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            first = b.energies()
            for rep in range(3):
                eo = _alone("upt", "fp64", other[0][0], other[0][2])
                assert np.isfinite(eo)
                b.run()
                again = b.energies()
                print("replay", rep, again)
                assert np.array_equal(again, first)
    finally:
        _close(systems)
        _close(other)


@pytest.mark.parametrize("form", ["fp64", "bf16x3"])
def test_batch_mixed_geometry(form):
    """Systems recorded with different tile geometries (4 and 8 waves) in
    one (fT) batch: the batch's save areas fit the larger geometry, and
    every system's energy equals its plan's alone (which reads the same
    records) and matches the oracle."""
    dims_list = [(5, 31), (4, 19), (3, 41)]
    geo = [4, 8, 4] if form == "fp64" else [8, 4, 8]
    for d, w, blk in zip(dims_list, geo, [(2, 2, 8), (8, 8, 1), (4, 4, 2)]):
        api.tune_record("rft", form, d, blk, w)
    systems = _systems("rft", dims_list, seed0=41, sig_f=1e-2, omega=0.3)
    try:
        with api.Batch("rft", form, [(d, dev) for d, _, dev in systems]) as b:
            b.run()
            eb = b.energies()
        for (d, spec, dev), e, w in zip(systems, eb, geo):
            with api.Plan("rft", form, d, dev) as pl:
                assert pl.geometry()[0] == w
                pl.run_all()
                ea = pl.energy()
            eo = oracle.energy("rft", d, api.synth_host("rft", spec))
            print(form, d, "waves", w, "batch", e, "alone", ea, "oracle", eo)
            assert e == ea
            tol = 1e-12 if form == "fp64" else 5e-4
            assert abs(e - eo) <= max(1e-10, tol * abs(eo))
    finally:
        _close(systems)
