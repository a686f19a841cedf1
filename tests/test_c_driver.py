"""The pt.h drop-in called from C.  tests/c/pt_driver.c is compiled with gcc
against include/pt.h and linked to libpt_amd.so, the way a C caller (the
reference's test driver test.cpp, Q-Chem) links libpt; it installs counting
scratch hooks (libpt_set_malloc / libpt_set_free).  The CPU tests build and
link it and compile both headers as C++; the GPU tests run libpt's
known-answer files and a synthetic (fT) system through it."""

import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import FIXTURES, ROOT, dims_of

LIBDIR = os.path.join(ROOT, "libpt_amd")
INCLUDE = os.path.join(ROOT, "include")
SRC = os.path.join(ROOT, "tests", "c", "pt_driver.c")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libpt_amd.so")):
        pytest.skip("libpt_amd.so is not built (python -m libpt_amd.build)")
    exe = str(tmp_path_factory.mktemp("cdriver") / "pt_driver")
    cmd = ["gcc", "-O2", "-std=c99", "-Wall", "-Wextra", "-Werror",
           "-I" + INCLUDE, SRC, "-L" + LIBDIR, "-lpt_amd",
           "-Wl,-rpath," + LIBDIR, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c_driver_builds_and_links(driver):
    """Every pt.h symbol the driver calls resolves in libpt_amd.so."""
    assert os.access(driver, os.X_OK)


def test_headers_compile_as_cxx(tmp_path):
    src = tmp_path / "both.cpp"
    src.write_text('#include "pt.h"\n#include "libpt_amd.h"\n'
                   "int main() {\n"
                   "  double (*f)(size_t, size_t, const double *,\n"
                   "      const double *, const double *, const double *,\n"
                   "      const double *, const double *, const double *) =\n"
                   "      libpt_rpt;\n"
                   "  return f != libpt_rpt;\n}\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror",
                        "-fsyntax-only", "-I" + INCLUDE, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def _run(driver, kind, dims, mp, arrays, tmp_path):
    files = []
    for n, a in enumerate(arrays):
        p = tmp_path / f"a{n}.bin"
        np.ascontiguousarray(a, dtype=np.float64).tofile(p)
        files.append(str(p))
    print("C driver:", kind, dims, "mp" if mp else "fp64", flush=True)
    r = subprocess.run([driver, kind, str(int(mp)), *map(str, dims), *files],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = dict(ln.split(None, 1) for ln in r.stdout.splitlines())
    n_malloc, n_free = (int(x) for x in out["hooks"].split())
    return float(out["E"]), n_malloc, n_free


@pytest.mark.gpu
@pytest.mark.parametrize("mp", [False, True])
@pytest.mark.parametrize("name", FIXTURES)
def test_fixtures_through_c(driver, golden, name, mp, tmp_path):
    from libpt_amd import api
    fx = golden[name]
    e, n_malloc, n_free = _run(driver, fx.kind, dims_of(fx), mp, fx.args(),
                               tmp_path)
    e_py = api.dropin(fx.kind, dims_of(fx), fx.args(), mp=mp)
    print(name, "C", e, "ctypes", e_py, "golden", fx.e_ref,
          "hooks", n_malloc, n_free)
    assert abs(e - fx.e_ref) <= 1e-8          # test.cpp:29,990
    assert e == e_py
    assert n_malloc > 0 and n_malloc == n_free


@pytest.mark.gpu
def test_rft_through_c(driver, tmp_path):
    """rft has no known-answer file: a synthetic system against the oracle
    and the ctypes call."""
    from libpt_amd import api
    dims = (5, 30)
    spec = api.make_spec("rft", dims, seed=8, omega=0.3, sig_f=1e-2)
    arrays = api.synth_host("rft", spec)
    e, n_malloc, n_free = _run(driver, "rft", dims, False, arrays, tmp_path)
    eo = oracle.energy("rft", dims, arrays)
    print("rft C", e, "oracle", eo)
    assert e == api.dropin("rft", dims, arrays)
    assert abs(e - eo) <= max(1e-10, 1e-13 * abs(eo))
    assert n_malloc == n_free


def _gpu_present():
    from libpt_amd import api
    try:
        api.energy("rpt", "fp64", (3, 4), [np.zeros(n) for n in
                                             (12, 12, 12, 288, 216, 288,
                                              72 + 192)])
        return True
    except api.LibptError:
        return False


def test_dropin_without_gpu_exits_like_err(driver, tmp_path):
    """No device: a pt.h call prints 'libpt: <reason>' and exits with status
    1, as the reference's err(1, ...) does (src/pt.c:208-209)."""
    if _gpu_present():
        pytest.skip("a GPU is present")
    arrays = [np.zeros(n) for n in (12, 12, 12, 288, 216, 288, 72 + 192)]
    files = []
    for n, a in enumerate(arrays):
        p = tmp_path / f"a{n}.bin"
        a.tofile(p)
        files.append(str(p))
    r = subprocess.run([driver, "rpt", "0", "3", "4", *files],
                       capture_output=True, text=True, timeout=120)
    print(r.returncode, r.stderr)
    assert r.returncode == 1
    assert r.stderr.startswith("libpt: ")


def test_extension_api_without_gpu_returns_nan():
    from libpt_amd import api
    if _gpu_present():
        pytest.skip("a GPU is present")
    with pytest.raises(api.LibptError) as exc:
        api.energy("rpt", "fp64", (3, 4), [np.zeros(n) for n in
                                           (12, 12, 12, 288, 216, 288,
                                            72 + 192)])
    assert "device" in str(exc.value).lower()
