"""The GPU against committed golden energies at the workload sizes:

- ref_energies.json: libpt compiled from its own sources (MKL sequential,
  OpenMP threads as recorded) on the synthetic generator's inputs, written
  by tests/golden/make_ref_goldens.py: whole calls of rpt 8/64, rpt 45/140,
  upt 24/200/20/204, rft 30/100 and uft 16/160/14/160 (FP64, some also
  _mp), and the 64 systems of the batch workload;
- ijk_rpt_45_477.json: the oracle's per-ijk energies at rpt o=45 v=477
  (74 GB of FP64 inputs) for the first three and the last ijk triple of
  each spin case, FP64 and FP32 (_mp), written by
  tests/golden/make_ijk_goldens.py.

Inputs are regenerated in HBM from (call, dims, seed, omega, sig_f): the
device generator is bitwise the host generator the goldens were made from.
"""

import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from libpt_amd import api

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "ref_energies.json")) as _fh:
    REF = json.load(_fh)["entries"]
with open(os.path.join(GOLDEN, "ijk_rpt_45_477.json")) as _fh:
    IJK = json.load(_fh)

WHOLE = sorted(n for n in REF if n != "batch64")


def _spec(e):
    return api.make_spec(e["call"], tuple(e["dims"]), seed=e["seed"],
                         omega=e["omega"], sig_f=e["sig_f"])


def _energy(kind, form, dims, dev):
    with api.Plan(kind, form, dims, dev) as pl:
        pl.run_all()
        return pl.energy()


@pytest.mark.parametrize("name", WHOLE)
def test_whole_call_against_reference(name):
    """FP64: within max(1e-10 Eh, 1e-12 |E|) of libpt at the recorded
    thread count (1e-10 Eh at rpt o=8 v=64); libpt at 1 thread differs from
    itself at 8 threads by its uncompensated per-thread sums, and the GPU
    must lie within that spread plus the same bound.  _mp: libpt's _mp
    rounds W, Z and Dn in float in its own order, so 1e-5 relative."""
    e = REF[name]
    kind, dims = e["call"], tuple(e["dims"])
    dev = api.synth_device(kind, _spec(e), fp32=e["mp"])
    try:
        got = _energy(kind, "fp32" if e["mp"] else "fp64", dims, dev)
    finally:
        dev.close()
    ref = e["E_ref"]
    print(name, "gpu", repr(got), "libpt", repr(ref), "threads", e["threads"],
          "dE", got - ref, "rel", abs(got - ref) / abs(ref))
    if e["mp"]:
        assert abs(got - ref) <= 1e-5 * abs(ref)
    elif e["threads"] == 1:
        spread = abs(ref - REF[name[:-3]]["E_ref"])
        assert abs(got - ref) <= spread + 1e-10
    else:
        assert abs(got - ref) <= max(1e-10, 1e-12 * abs(ref))


@pytest.mark.parametrize("form,bound", [("fp64", 1e-12), ("bf16x3", 1e-6)])
def test_batch64_against_reference(form, bound):
    """The 64-system batch workload (o in [4, 20], v in [40, 160]) as one
    graph: every system's energy against libpt on the same arrays; FP64 to
    1e-12 relative, split-BF16 to 1e-6 relative."""
    g = REF["batch64"]
    dims = [tuple(d) for d in g["dims"]]
    assert dims == api.batch_dims(g["systems"], seed=g["seed"])
    systems = [(d, api.synth_device("rpt", api.make_spec(
        "rpt", d, seed=g["seed"] + 1 + k))) for k, d in enumerate(dims)]
    try:
        with api.Batch("rpt", form, systems) as b:
            b.run()
            e = np.asarray(b.energies())
    finally:
        for _, dev in systems:
            dev.close()
    ref = np.asarray(g["E_ref"])
    rel = np.abs(e - ref) / np.abs(ref)
    k = int(np.argmax(rel))
    print(form, "worst system", k, dims[k], "rel", rel[k], "max |dE|",
          np.max(np.abs(e - ref)))
    assert np.all(rel <= bound)


def _items(pl, sc, index):
    """per-ijk energies of the listed work-list positions"""
    for i in index:
        pl.run(sc, i, i + 1)
    pl.sync()
    return pl.ijk_energies(sc)[index]


def _check_ijk(pl, golden, bound, label):
    worst = 0.0
    for sc, name in enumerate(("aaa", "aab")):
        g = golden[name]
        lst = pl.ijk_list(sc)
        for i, ijk in zip(g["index"], g["ijk"]):
            assert list(lst[i]) == ijk
        got = _items(pl, sc, g["index"])
        ref = np.asarray(g["energy"])
        rel = np.abs(got - ref) / np.abs(ref)
        worst = max(worst, float(rel.max()))
        print(label, name, got.tolist(), "rel", rel.tolist())
        assert np.all(rel <= bound)
    return worst


def test_full_size_per_ijk_against_oracle():
    """rpt o=45 v=477 with the whole 74 GB input in HBM: per-ijk energies
    of the first three and the last triple of each spin case against the
    oracle.  FP64 to 1e-12 relative; split-BF16 and F16X3 against the FP64
    oracle to 1e-6 relative; FP32 (_mp, float inputs) against the oracle's
    _mp to 1e-5 relative."""
    kind, dims = "rpt", tuple(IJK["dims"])
    spec = api.make_spec(kind, dims, seed=IJK["seed"])
    dev = api.synth_device(kind, spec)
    try:
        for form, bound in (("fp64", 1e-12), ("bf16x3", 1e-6),
                            ("f16x3", 1e-6)):
            print("full size", form, flush=True)
            with api.Plan(kind, form, dims, dev) as pl:
                _check_ijk(pl, IJK["fp64"], bound, form)
    finally:
        dev.close()
    dev = api.synth_device(kind, spec, fp32=True)
    try:
        print("full size fp32", flush=True)
        with api.Plan(kind, "fp32", dims, dev) as pl:
            _check_ijk(pl, IJK["fp32"], 1e-5, "fp32")
    finally:
        dev.close()
