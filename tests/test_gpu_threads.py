"""Calls from several host threads at once: each drop-in call and each
plan owns its stream and buffers, the device is made current on every
calling thread, and the last-error text is per thread.  Energies computed
concurrently must equal the same calls made one after the other, bit for
bit."""

import threading

import numpy as np
import pytest

from libpt_amd import api

pytestmark = pytest.mark.gpu

SYSTEMS = [("rpt", (6, 37)), ("upt", (4, 21, 3, 19)), ("rft", (4, 30)),
           ("uft", (3, 17, 4, 13)), ("rpt", (5, 48)), ("upt", (5, 16, 5, 24))]


def _arrays(kind, dims, seed):
    spec = api.make_spec(kind, dims, seed=seed, sig_f=1e-2,
                         omega=0.3 if kind.endswith("ft") else 0.0)
    return api.synth_host(kind, spec)


def test_concurrent_dropin_calls_match_sequential():
    inputs = [(k, d, _arrays(k, d, 40 + n)) for n, (k, d) in
              enumerate(SYSTEMS)]
    seq = [api.dropin(k, d, a) for k, d, a in inputs]
    out = [None] * len(inputs)
    errs = []

    def work(n):
        try:
            k, d, a = inputs[n]
            for _ in range(3):
                e = api.dropin(k, d, a)
                assert out[n] is None or out[n] == e
                out[n] = e
        except Exception as exc:  # reported by the main thread
            errs.append(repr(exc))

    threads = [threading.Thread(target=work, args=(n,))
               for n in range(len(inputs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errs, errs
    print("sequential", seq, "concurrent", out)
    assert out == seq


def test_error_text_is_per_thread():
    """A failing call on one thread leaves another thread's error text
    alone."""
    got = {}
    arrays = _arrays("rpt", (4, 10), 1)
    # a successful call clears this thread's error text
    e = api.energy("rpt", "fp64", (4, 10), arrays)
    assert np.isfinite(e) and api.last_error() == ""

    def bad():
        try:
            api.energy("rpt", 99, (4, 10), arrays)
        except api.LibptError as exc:
            got["bad"] = str(exc)

    t = threading.Thread(target=bad)
    t.start()
    t.join(timeout=60)
    assert "form" in got.get("bad", "")
    assert api.last_error() == ""
