import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = ["rpt02", "rpt09", "upt01", "upt03", "uft02"]


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line(
        "markers", "slow: CPU test that takes more than a few seconds")


def fixture_path(name):
    return os.path.join(GOLDEN, name + ".dat.gz")


@pytest.fixture(scope="session")
def golden():
    from libpt_amd.fixtures import read_fixture
    return {n: read_fixture(fixture_path(n)) for n in FIXTURES}


def dims_of(fx):
    if fx.kind in ("rpt", "rft"):
        return (fx.oa, fx.va)
    return (fx.oa, fx.va, fx.ob, fx.vb)


def embed_restricted(kind, o, v, arrays):
    """Restricted inputs -> unrestricted layout with beta = alpha
    (bbbb = aaaa, baba = abab), the identity behind src/pt.c:469-473 vs
    :515-529 and :958-966 vs :1024-1044."""
    P = v * (v - 1) // 2

    def ov(a):
        return np.concatenate([a, a])

    def blk2(a, n0):
        return np.concatenate([a[:n0], a[n0:], a[:n0], a[n0:]])

    n_oovv, n_oovo, n_ovvv = o * o * v * v, o * o * o * v, o * v * P
    if kind == "rpt":
        d, f, t1, t2, oovo, oovv, ovvv = arrays
        return [ov(d), ov(f), ov(t1), blk2(t2, n_oovv), blk2(oovo, n_oovo),
                blk2(oovv, n_oovv), blk2(ovvv, n_ovvv)]
    d, f2, l1, t2, l2, oovv, i2, i3, i6, i7 = arrays
    return [ov(d), ov(f2), ov(l1), blk2(t2, n_oovv), blk2(l2, n_oovv),
            blk2(oovv, n_oovv), blk2(i2, n_oovo), blk2(i3, n_ovvv),
            blk2(i6, n_oovo), blk2(i7, n_ovvv)]
