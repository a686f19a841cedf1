"""Tile-order blocking and its autotuner (libpt_amd_plan_set_block /
libpt_amd_plan_tune): the order of the tiles changes only the summation
order of the per-tile energies, so every blocking gives the same energy to
rounding, and the tuned plan matches the oracle."""

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
            pl.run_all()
            e = pl.energy()
        # a new plan of the same sizes takes the tuned blocking from the
        # cache and gives the same energy bit for bit
        with api.Plan(kind, "fp64", dims, dev) as pl2:
            pl2.run_all()
            assert pl2.energy() == e
    finally:
        dev.close()
    eo = oracle.energy(kind, dims, api.synth_host(kind, spec))
    print("tuned", blk, e, "oracle", eo)
    assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))
