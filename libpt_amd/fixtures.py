"""Reader for libpt's ``.dat`` known-answer files.

The format and the spin-orbital -> spin-block slicing follow the reference
test driver:

* header ``o v E`` (restricted) or ``oa va ob vb E`` (unrestricted),
  ``test.cpp:68-85``;
* then sections, each a title line followed by whitespace-separated values,
  ``test.cpp:87-112``;
* restricted (T) files store the API blocks directly, ``test.cpp:114-151``;
* unrestricted (T) files store spin-orbital tensors over o = oa+ob,
  v = va+vb, sliced into aaaa/abab/bbbb/baba blocks, ``test.cpp:153-315``;
  ovvv is stored as a c<b packed list of the antisymmetric tensor
  (``test.cpp:279-312``);
* (fT) files: ``read_test_data_rft`` ``test.cpp:317-485`` (alpha blocks
  only, with oa=ob, va=vb) and ``read_test_data_uft`` ``test.cpp:487-731``.

Every function returns the arrays in exactly the layout the pt.h API takes
(``src/pt.h:43-57,79-93,123-124,156-157``), as float64 numpy arrays.
"""

from __future__ import annotations

import gzip
import os
import re
from dataclasses import dataclass, field

import numpy as np

_TITLE = re.compile(r"^[A-Za-z]")


@dataclass
class Fixture:
    name: str
    kind: str  # rpt / upt / rft / uft
    oa: int
    va: int
    ob: int
    vb: int
    e_ref: float
    arrays: dict = field(default_factory=dict)

    def args(self, dtype=np.float64):
        """API arguments in pt.h order, cast to ``dtype`` (C-contiguous)."""
        order = _ARG_ORDER[self.kind]
        return [np.ascontiguousarray(self.arrays[n], dtype=dtype) for n in order]


_ARG_ORDER = {
    "rpt": ["d_ov", "f_ov", "t1", "t2", "i_oovo", "i_oovv", "i_ovvv"],
    "upt": ["d_ov", "f_ov", "t1", "t2", "i_oovo", "i_oovv", "i_ovvv"],
    "rft": ["d_ov", "f2_ov", "l1", "t2", "l2", "i_oovv", "i2_t2f2_oovo",
            "i3_ovvv", "i6_oovo", "i7_ovvv"],
    "uft": ["d_ov", "f2_ov", "l1", "t2", "l2", "i_oovv", "i2_t2f2_oovo",
            "i3_ovvv", "i6_oovo", "i7_ovvv"],
}


def _sections(text: str):
    lines = text.splitlines()
    header = lines[0].split()
    secs = []
    cur_title, cur = None, []
    for ln in lines[1:]:
        if _TITLE.match(ln):
            if cur_title is not None:
                secs.append((cur_title, cur))
            cur_title, cur = ln.split()[0], []
        else:
            cur.append(ln)
    if cur_title is not None:
        secs.append((cur_title, cur))
    out = []
    for title, body in secs:
        vals = np.array(" ".join(body).split(), dtype=np.float64)
        out.append((title, vals))
    return header, out


def _read_text(path: str) -> str:
    if path.endswith(".gz"):
        with gzip.open(path, "rt") as fh:
            return fh.read()
    with open(path, "r") as fh:
        return fh.read()


def _unpack_ovvv(vals, o, v):
    """c<b packed spin-orbital ovvv -> full antisymmetric [o][v][v][v]
    (test.cpp:279-292)."""
    full = np.zeros((o, v, v, v))
    b_idx, c_idx = np.tril_indices(v, -1)  # b > c, row-major (b, c)
    packed = vals.reshape(o, v, -1)
    full[:, :, b_idx, c_idx] = packed
    full[:, :, c_idx, b_idx] = -packed
    return full


def _tri_pack(full_oovv_like):
    """[o][v][v][v] -> packed [o][v][P(v)] over b > c (src/pt.h:51-52)."""
    o, v = full_oovv_like.shape[0], full_oovv_like.shape[1]
    b_idx, c_idx = np.tril_indices(full_oovv_like.shape[2], -1)
    return full_oovv_like[:, :, b_idx, c_idx].reshape(o, v, -1)


def _ov_blocks(vals, oa, va, ob, vb):
    """spin-orbital [o][v] -> [alpha oa*va | beta ob*vb] (test.cpp:167-178)."""
    m = vals.reshape(oa + ob, va + vb)
    return np.concatenate([m[:oa, :va].ravel(), m[oa:, va:].ravel()])


def _oovv_blocks(vals, oa, va, ob, vb, nblk=4):
    o, v = oa + ob, va + vb
    t = vals.reshape(o, o, v, v)
    blocks = [t[:oa, :oa, :va, :va], t[:oa, oa:, :va, va:],
              t[oa:, oa:, va:, va:], t[oa:, :oa, va:, :va]]
    return np.concatenate([b.ravel() for b in blocks[:nblk]])


def _oovo_blocks(vals, oa, va, ob, vb, nblk=4):
    o, v = oa + ob, va + vb
    t = vals.reshape(o, o, v, o)
    blocks = [t[:oa, :oa, :va, :oa], t[:oa, oa:, :va, oa:],
              t[oa:, oa:, va:, oa:], t[oa:, :oa, va:, :oa]]
    return np.concatenate([b.ravel() for b in blocks[:nblk]])


def _ovvv_blocks(vals, oa, va, ob, vb, nblk=4):
    o, v = oa + ob, va + vb
    full = _unpack_ovvv(vals, o, v)
    blocks = [_tri_pack(full[:oa, :va, :va, :va]),
              full[:oa, va:, :va, va:],
              _tri_pack(full[oa:, va:, va:, va:]),
              full[oa:, :va, va:, :va]]
    return np.concatenate([b.ravel() for b in blocks[:nblk]])


def read_fixture(path: str) -> Fixture:
    """Parse one ``.dat`` (optionally gzipped) known-answer file."""
    name = os.path.basename(path).split(".")[0]
    kind = name[:3]
    header, secs = _sections(_read_text(path))
    if kind in ("rpt", "rft"):
        oa, va = int(header[0]), int(header[1])
        ob, vb = oa, va
        e_ref = float(header[2])
    else:
        oa, va, ob, vb = (int(x) for x in header[:4])
        e_ref = float(header[4])
    fx = Fixture(name, kind, oa, va, ob, vb, e_ref)
    s = dict(secs) if kind != "rpt" else None
    if kind == "rpt":
        # test.cpp:114-151: blocks stored as the API takes them, in order
        vals = [v for _, v in secs]
        fx.arrays = {
            "d_ov": vals[0], "f_ov": vals[1], "t1": vals[2],
            "t2": np.concatenate([vals[3], vals[4]]),
            "i_oovo": np.concatenate([vals[5], vals[6]]),
            "i_oovv": np.concatenate([vals[7], vals[8]]),
            "i_ovvv": np.concatenate([vals[9], vals[10]]),
        }
    elif kind == "upt":
        fx.arrays = {
            "d_ov": _ov_blocks(s["d_ov"], oa, va, ob, vb),
            "f_ov": _ov_blocks(s["f_ov"], oa, va, ob, vb),
            "t1": _ov_blocks(s["t1"], oa, va, ob, vb),
            "t2": _oovv_blocks(s["t2"], oa, va, ob, vb),
            "i_oovo": _oovo_blocks(s["i_oovo"], oa, va, ob, vb),
            "i_oovv": _oovv_blocks(s["i_oovv"], oa, va, ob, vb),
            "i_ovvv": _ovvv_blocks(s["i_ovvv"], oa, va, ob, vb),
        }
    elif kind in ("uft", "rft"):
        nb = 4 if kind == "uft" else 2
        ovs = (lambda x: _ov_blocks(x, oa, va, ob, vb)) if kind == "uft" else \
            (lambda x: x.reshape(oa + ob, va + vb)[:oa, :va].ravel())
        fx.arrays = {
            "d_ov": ovs(s["d_ov"]),
            "f2_ov": ovs(s["f2_ov"]),
            "l1": ovs(s["l1"]),
            "t2": _oovv_blocks(s["t2"], oa, va, ob, vb, nb),
            "l2": _oovv_blocks(s["l2"], oa, va, ob, vb, nb),
            "i_oovv": _oovv_blocks(s["i_oovv"], oa, va, ob, vb, nb),
            "i2_t2f2_oovo": _oovo_blocks(s["i2_t2f2_oovo"], oa, va, ob, vb, nb),
            "i3_ovvv": _ovvv_blocks(s["i3_ovvv"], oa, va, ob, vb, nb),
            "i6_oovo": _oovo_blocks(s["i6_oovo"], oa, va, ob, vb, nb),
            "i7_ovvv": _ovvv_blocks(s["i7_ovvv"], oa, va, ob, vb, nb),
        }
    else:
        raise ValueError(f"unknown fixture kind {kind!r} in {path}")
    return fx
