"""Launch chunking of the (fT) calls.  A launch with two W sets holds at
most 4 GB of save areas (16384 to 32768 items by form and tile geometry),
because every item of a launch keeps W of set 0 in its own save area;
larger spin cases run in several launches.  The energy of
every ijk triple must not depend on how the work list is cut into
launches: the same items run one ijk per launch give the same bits."""

import numpy as np
import pytest

from libpt_amd import api

pytestmark = pytest.mark.gpu

FT_CHUNK = 32768


@pytest.mark.parametrize("form", ["fp64", "bf16x3"])
def test_ft_chunked_launches_match_single_ijk_runs(form):
    kind, dims = "rft", (12, 160)
    spec = api.make_spec(kind, dims, seed=3, omega=0.3, sig_f=1e-2)
    dev = api.synth_device(kind, spec)
    try:
        with api.Plan(kind, form, dims, dev) as full:
            items = [full.items(sc) for sc in range(full.ncases)]
            print(form, "items per spin case", items)
            assert max(items) > FT_CHUNK
            full.run_all()
            e_full = full.energy()
            ref = [full.ijk_energies(sc) for sc in range(full.ncases)]
        with api.Plan(kind, form, dims, dev) as one:
            for sc in range(one.ncases):
                for b in range(one.nijk(sc)):
                    one.run(sc, b, b + 1)
            e_one = one.energy()
            for sc in range(one.ncases):
                np.testing.assert_array_equal(one.ijk_energies(sc), ref[sc])
        print(form, "energy", e_full, e_one)
        assert np.isfinite(e_full) and e_full != 0.0
        assert e_one == e_full
    finally:
        dev.close()


def test_ft_batch_chunks_match_alone():
    """A batch of (fT) systems with more than 32768 items in all runs as
    several grids; each system's energy still equals its own plan's."""
    kind = "rft"
    dims_list = [(12, 160), (5, 30), (4, 37)]
    systems = []
    try:
        for n, d in enumerate(dims_list):
            spec = api.make_spec(kind, d, seed=20 + n, omega=0.3, sig_f=1e-2)
            systems.append((d, api.synth_device(kind, spec)))
        with api.Batch(kind, "fp64", systems) as b:
            print("batch items", b.items)
            assert b.items > FT_CHUNK
            b.run()
            eb = b.energies()
        for (d, dev), e in zip(systems, eb):
            with api.Plan(kind, "fp64", d, dev) as pl:
                pl.run_all()
                ea = pl.energy()
            print(d, "batch", e, "alone", ea)
            assert e == ea
    finally:
        for _, dev in systems:
            dev.close()
