"""Deterministic generator: layouts, closed-shell relations, determinism."""

import numpy as np
import pytest

from libpt_amd import api


def test_restricted_relations():
    o, v = 4, 7
    spec = api.make_spec("rpt", (o, v), seed=11, sig_f=1e-2)
    d, f, t1, t2, oovo, oovv, ovvv = api.synth_host("rpt", spec)
    T = t2.reshape(2, o, o, v, v)
    G = oovv.reshape(2, o, o, v, v)
    for X in (T, G):
        assert np.array_equal(X[1], X[1].transpose(1, 0, 3, 2))
        assert np.array_equal(X[0], X[1] - X[1].transpose(0, 1, 3, 2))
    n3 = o * o * o * v
    Oa = oovo[:n3].reshape(o, o, v, o)
    Ob = oovo[n3:].reshape(o, o, v, o)
    assert np.array_equal(Oa, Ob - Ob.transpose(1, 0, 2, 3))
    P = v * (v - 1) // 2
    Va = ovvv[:o * v * P].reshape(o, v, P)
    Vb = ovvv[o * v * P:].reshape(o, v, v, v)
    b, c = np.tril_indices(v, -1)
    assert np.array_equal(Va, Vb[:, :, b, c] - Vb[:, :, c, b])
    D = d.reshape(o, v)
    assert np.all(D < 0)
    assert np.allclose(D - D[:, :1], D[:1] - D[:1, :1], atol=1e-14)


def test_unrestricted_antisymmetry():
    oa, va, ob, vb = 3, 4, 2, 5
    spec = api.make_spec("upt", (oa, va, ob, vb), seed=5)
    d, f, t1, t2, oovo, oovv, ovvv = api.synth_host("upt", spec)
    n0, n1 = oa * oa * va * va, oa * ob * va * vb
    n2 = ob * ob * vb * vb
    Taa = t2[:n0].reshape(oa, oa, va, va)
    assert np.array_equal(Taa, -Taa.transpose(1, 0, 2, 3))
    assert np.array_equal(Taa, -Taa.transpose(0, 1, 3, 2))
    Tab = t2[n0:n0 + n1].reshape(oa, ob, va, vb)
    Tba = t2[n0 + n1 + n2:].reshape(ob, oa, vb, va)
    assert np.array_equal(Tba, Tab.transpose(1, 0, 3, 2))
    assert len(d) == oa * va + ob * vb


def test_deterministic_and_seeded():
    s1 = api.make_spec("rpt", (3, 5), seed=1)
    s2 = api.make_spec("rpt", (3, 5), seed=2)
    a = api.synth_host("rpt", s1)
    b = api.synth_host("rpt", s1)
    c = api.synth_host("rpt", s2)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[3], c[3])


def test_moments():
    spec = api.make_spec("rpt", (6, 24), seed=9)
    d, f, t1, t2, oovo, oovv, ovvv = api.synth_host("rpt", spec)
    n = 6 * 6 * 24 * 24
    assert abs(t2[n:].std() - 1e-2) < 1e-3
    assert abs(oovo[6 * 6 * 6 * 24:].std() - 5e-2) < 3e-3
    assert np.all(f == 0)


@pytest.mark.parametrize("kind,dims", [("rpt", (4, 6)), ("upt", (3, 4, 2, 5)),
                                       ("rft", (3, 5)), ("uft", (2, 4, 3, 3))])
def test_sizes_match_api(kind, dims):
    from libpt_amd.sizes import api_sizes
    spec = api.make_spec(kind, dims)
    arr = api.synth_host(kind, spec)
    assert [len(a) for a in arr] == api_sizes(kind, dims)


def test_range_is_a_slice_of_the_whole_tensor():
    spec = api.make_spec("rft", (5, 9), seed=3, sig_f=1e-2, omega=0.3)
    for t, whole in zip(api.tensor_ids("rft"), api.synth_host("rft", spec)):
        for b, n in ((0, whole.size), (7, 13), (whole.size - 5, 5),
                     (whole.size, 0)):
            part = np.empty(n)
            api.synth_range(spec, t, part, b)
            assert np.array_equal(part, whole[b:b + n])
        part = np.empty(4, dtype=np.float32)
        api.synth_range(spec, t, part, 2)
        assert np.array_equal(part, whole[2:6].astype(np.float32))
        with pytest.raises(api.LibptError):
            api.synth_range(spec, t, np.empty(2), whole.size - 1)
