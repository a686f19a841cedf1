"""The tuning cache file (LIBPT_AMD_TUNE_CACHE) written by
libpt_amd_tune_record: one line per record, `form call oa va ob vb ga gb
gc waves`, restricted calls with ob, vb = oa, va; records already in the
file are read before the first write and kept; bad records are refused
with an error.  No GPU needed: recording touches no device.  Each case runs
in a fresh process, since the library reads the file once per process."""

import os
import subprocess
import sys
import textwrap

from conftest import ROOT


def _run(code, cache):
    env = dict(os.environ, LIBPT_AMD_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(code)],
                       cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_records_are_appended_in_the_documented_format(tmp_path):
    cache = tmp_path / "tune.txt"
    # an older 9-field line (no geometry) must survive
    cache.write_text("0 0 8 64 8 64 8 8 2\n")
    _run("""
        from libpt_amd import api
        api.tune_record("rpt", "fp64", (45, 477), (8, 4, 2), 4)
        api.tune_record("upt", "bf16x3", (24, 200, 20, 204), (16, 16, 1), 8)
        api.tune_record("rft", "f16x3", (30, 300), (4, 4, 4), 0)
    """, cache)
    lines = cache.read_text().split("\n")
    assert lines[0] == "0 0 8 64 8 64 8 8 2"
    assert lines[1:4] == ["0 0 45 477 45 477 8 4 2 4",
                          "2 1 24 200 20 204 16 16 1 8",
                          "3 2 30 300 30 300 4 4 4 0"]


def test_bad_records_are_refused(tmp_path):
    cache = tmp_path / "tune.txt"
    out = _run("""
        from libpt_amd import api
        for blk, w in (((8, 4, 2), 6), ((0, 4, 2), 4)):
            try:
                api.tune_record("rpt", "fp64", (8, 64), blk, w)
                print("accepted")
            except api.LibptError as e:
                print("refused:", e)
    """, cache)
    assert out.count("refused: bad tuning record") == 2
    assert not cache.exists() or cache.read_text() == ""
