"""Parity of the HIP engine (through its C ABI) with the CPU oracle and
libpt's known answers.  Sizes are ones the oracle finishes in seconds;
size-independent properties cover o=45, v=477 (test_gpu_fullsize.py)."""

import numpy as np
import pytest

import oracle
from libpt_amd import api
from conftest import FIXTURES, dims_of, embed_restricted

pytestmark = pytest.mark.gpu


def _synth(kind, dims, seed=3, **kw):
    spec = api.make_spec(kind, dims, seed=seed, **kw)
    return api.synth_host(kind, spec)


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_fp64(golden, name):
    fx = golden[name]
    e = api.dropin(fx.kind, dims_of(fx), fx.args())
    eo = oracle.energy(fx.kind, dims_of(fx), fx.args())
    print(name, "gpu", e, "oracle", eo, "golden", fx.e_ref)
    assert abs(e - fx.e_ref) <= 1e-8          # test.cpp:29,990
    assert abs(e - eo) <= 1e-14 + 1e-12 * abs(eo)


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_fp32(golden, name):
    fx = golden[name]
    e = api.dropin(fx.kind, dims_of(fx), fx.args(np.float32), mp=True)
    eo = oracle.energy(fx.kind, dims_of(fx), fx.args(), mp=True)
    print(name, "gpu_mp", e, "oracle_mp", eo, "golden", fx.e_ref)
    assert abs(e - fx.e_ref) <= 1e-8
    assert abs(e - eo) <= 1e-5 * abs(eo)


# (kind, dims, extra generator args): tile edges (v not a multiple of
# 16/32), odd/even spin sizes, (fT) with omega folded into d_ov
CASES = [
    ("rpt", (8, 64), dict(sig_f=1e-2)),          # the small rpt workload
    ("rpt", (6, 37), dict(sig_f=1e-2)),
    ("rpt", (3, 50), dict()),
    ("upt", (5, 20, 4, 23), dict(sig_f=1e-2)),
    ("upt", (4, 33, 6, 17), dict()),
    ("rft", (5, 30), dict(sig_f=1e-2, omega=0.3)),
    ("uft", (4, 18, 3, 21), dict(sig_f=1e-2, omega=0.3)),
]


@pytest.mark.parametrize("kind,dims,kw", CASES)
def test_synthetic_fp64(kind, dims, kw):
    arr = _synth(kind, dims, **kw)
    e = api.dropin(kind, dims, arr)
    eo = oracle.energy(kind, dims, arr)
    print(kind, dims, "gpu", e, "oracle", eo, "dE", e - eo)
    # FP64 parity: 1e-10 Eh absolute at the small rpt size (o=8, v=64), and a
    # relative bound that scales with |E| (summation order only)
    assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))


@pytest.mark.parametrize("kind,dims,kw", CASES[:1] + CASES[3:4] + CASES[5:6])
def test_synthetic_fp32(kind, dims, kw):
    arr = _synth(kind, dims, **kw)
    e = api.dropin(kind, dims, [a.astype(np.float32) for a in arr], mp=True)
    eo = oracle.energy(kind, dims, arr, mp=True)
    print(kind, dims, "gpu_mp", e, "oracle_mp", eo)
    assert abs(e - eo) <= 1e-5 * abs(eo)


@pytest.mark.parametrize("form,tol", [("bf16x3", 1e-4), ("f16x3", 2e-6)])
@pytest.mark.parametrize("kind,dims,kw", [CASES[0], CASES[3], CASES[5],
                                          CASES[6]])
def test_split_forms(form, tol, kind, dims, kw):
    """hi+lo products with FP32 accumulation: relative error vs the FP64
    oracle, bound stated per form.  (fT) energies sum signed products
    (W_L - Z) W_R / Dn, whose cancellation loosens the relative bound 5x."""
    if "ft" in kind:
        tol *= 5
    arr = _synth(kind, dims, **kw)
    e = api.energy(kind, form, dims, arr)
    eo = oracle.energy(kind, dims, arr)
    rel = abs(e - eo) / abs(eo)
    print(kind, dims, form, e, eo, "rel", rel)
    assert rel <= tol


@pytest.mark.parametrize("form", ["bf16x3", "f16x3"])
@pytest.mark.parametrize("name", FIXTURES)
def test_split_forms_on_fixtures(golden, name, form):
    """The split forms on libpt's own molecular inputs (physical
    amplitudes and integrals, |E| of order 1e-3 Eh): within 1e-6 Eh of the
    known answer and of the FP64 oracle."""
    fx = golden[name]
    e = api.energy(fx.kind, form, dims_of(fx), fx.args())
    eo = oracle.energy(fx.kind, dims_of(fx), fx.args())
    print(name, form, "gpu", e, "oracle", eo, "golden", fx.e_ref,
          "|dE| vs oracle", abs(e - eo))
    assert abs(e - eo) <= 1e-6
    assert abs(e - fx.e_ref) <= 1e-6


def test_restricted_equals_embedded_unrestricted():
    o, v = 5, 21
    arr = _synth("rpt", (o, v), sig_f=1e-2)
    e_r = api.dropin("rpt", (o, v), arr)
    e_u = api.dropin("upt", (o, v, o, v), embed_restricted("rpt", o, v, arr))
    assert abs(e_r - e_u) <= 1e-13 * abs(e_r)


def test_per_ijk_energies_match_oracle():
    """Per-item energies of each spin case (the per-ijk partials the
    reference sums in its omp-for) against the oracle on the same items."""
    kind, dims = "upt", (5, 19, 4, 22)
    arr = _synth(kind, dims, sig_f=1e-2)
    dev = api.to_device(kind, arr)
    try:
        with api.Plan(kind, "fp64", dims, dev) as pl:
            pl.run_all()
            oa, va, ob, vb = dims
            n = [oa * oa * va * va, oa * ob * va * vb, ob * ob * vb * vb]
            d, f, t1, t2, oovo, oovv, ovvv = arr
            for sc in range(pl.ncases):
                got = pl.ijk_energies(sc)
                items = pl.ijk_list(sc)
                if sc == 0:   # aaa alpha
                    _, ref = oracle.spin_case(
                        "pt_aaa", (oa, va),
                        [d[:oa * va], f[:oa * va], t1[:oa * va], t2[:n[0]],
                         oovo[:oa * oa * va * oa], oovv[:n[0]],
                         ovvv[:oa * va * va * (va - 1) // 2]], items)
                    np.testing.assert_allclose(got, ref, rtol=1e-12,
                                               atol=1e-18)
            e = pl.energy()
        assert abs(e - oracle.energy(kind, dims, arr)) <= 1e-13 * abs(e)
    finally:
        dev.close()


def test_device_generator_is_bitwise_host_generator():
    for kind, dims in (("rpt", (4, 9)), ("uft", (3, 5, 2, 6))):
        spec = api.make_spec(kind, dims, seed=42, sig_f=1e-2, omega=0.3)
        host = api.synth_host(kind, spec)
        dev = api.synth_device(kind, spec)
        try:
            for h, p, n in zip(host, dev.ptrs, dev.sizes):
                g = api.from_device(p, n)
                assert np.array_equal(h.view(np.uint64), g.view(np.uint64))
        finally:
            dev.close()


def test_deterministic():
    arr = _synth("rpt", (7, 45), sig_f=1e-2)
    e1 = api.dropin("rpt", (7, 45), arr)
    e2 = api.dropin("rpt", (7, 45), arr)
    assert e1 == e2


@pytest.mark.parametrize("kind,dims,kw", [
    ("rpt", (8, 64), dict(sig_f=1e-2)),
    ("upt", (5, 20, 4, 23), dict(sig_f=1e-2)),
    ("rft", (5, 30), dict(sig_f=1e-2, omega=0.3)),
    ("uft", (4, 18, 3, 21), dict(sig_f=1e-2, omega=0.3)),
    # sizes above the known-answer files (seconds of reference CPU time)
    ("rpt", (12, 100), dict(sig_f=1e-2)),
    ("upt", (9, 70, 8, 74), dict(sig_f=1e-2)),
    ("rft", (10, 90), dict(sig_f=1e-2, omega=0.3)),
])
def test_fp64_against_compiled_reference(kind, dims, kw):
    """The GPU against libpt itself (oracle/_ref, built from the reference's
    sources).  libpt sums each thread's energies without compensation, so
    its own result moves with the thread count: 6.4e-10 Eh between 1 and 2
    threads at rpt o=8 v=64 (|E| ~ 2e3 Eh).  Bound: that spread plus the
    oracle bound, 1e-9 Eh or 1e-12 relative."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    arr = _synth(kind, dims, seed=11, **kw)
    e = api.dropin(kind, dims, arr)
    er = oracle.ref_energy(kind, dims, arr)
    print(kind, dims, "gpu", e, "reference", er, "dE", e - er)
    assert abs(e - er) <= max(1e-9, 1e-12 * abs(er))


@pytest.mark.parametrize("kind,dims,kw", [
    ("rpt", (8, 64), dict(sig_f=1e-2)),
    ("upt", (5, 20, 4, 23), dict(sig_f=1e-2)),
    ("rft", (5, 30), dict(sig_f=1e-2, omega=0.3)),
    ("uft", (4, 18, 3, 21), dict(sig_f=1e-2, omega=0.3)),
    ("rpt", (12, 100), dict(sig_f=1e-2)),
    ("uft", (8, 60, 7, 64), dict(sig_f=1e-2, omega=0.3)),
])
def test_fp32_against_compiled_reference(kind, dims, kw):
    """_mp against libpt's _mp (sgemm, float W, Z and Dn, double sum):
    both round W in float in different orders, so the bound is relative,
    1e-5, as against the oracle."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    arr = [a.astype(np.float32) for a in _synth(kind, dims, seed=11, **kw)]
    e = api.dropin(kind, dims, arr, mp=True)
    er = oracle.ref_energy(kind, dims, arr, mp=True)
    print(kind, dims, "gpu _mp", e, "reference _mp", er, "rel",
          abs(e - er) / abs(er))
    assert abs(e - er) <= 1e-5 * abs(er)


def test_dropin_stages_large_inputs_through_pinned_buffers():
    """Inputs above 256 MB go to HBM through the pinned, multi-threaded
    staging path of the drop-in: the energy equals that of a plan on the
    same arrays copied with hipMemcpy, bit for bit."""
    kind, dims = "rpt", (10, 180)
    arr = _synth(kind, dims, seed=7)
    assert sum(a.nbytes for a in arr) > 256 << 20
    e = api.rpt(*dims, *arr)
    dev = api.to_device(kind, arr)
    try:
        with api.Plan(kind, "fp64", dims, dev) as pl:
            pl.run_all()
            ep = pl.energy()
    finally:
        dev.close()
    print("drop-in", e, "plan", ep)
    assert e == ep
