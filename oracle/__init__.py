"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product
package ``libpt_amd`` never imports it.

Two libraries may be loaded:

* ``liboracle.so`` -- this repository's plain-C restatement of the
  reference algorithm (``oracle/pt_oracle.c``), always built by ``make``;
* ``_ref/libpt_ref.so`` -- the reference libpt compiled unmodified from its
  own sources (``make ref``, needs the reference tree and MKL); optional.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORA = None
_REF = None

_sz = C.c_size_t
_p = C.c_void_p


def build(ref: bool = False) -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _lib():
    global _ORA
    if _ORA is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _ORA = C.CDLL(path)
        for suf in ("d", "f"):
            for name, nsz, narr in (("rpt", 2, 7), ("upt", 4, 7),
                                    ("rft", 2, 10), ("uft", 4, 10)):
                fn = getattr(_ORA, f"oracle_{name}_{suf}")
                fn.restype = C.c_double
                fn.argtypes = [_sz] * nsz + [_p] * narr
            fn = getattr(_ORA, f"oracle_pt_aaa_{suf}")
            fn.restype = C.c_double
            fn.argtypes = [_sz, _sz] + [_p] * 7 + [_sz, _p, _p]
            fn = getattr(_ORA, f"oracle_pt_aab_{suf}")
            fn.restype = C.c_double
            fn.argtypes = [_sz] * 4 + [_p] * 17 + [_sz, _p, _p]
            fn = getattr(_ORA, f"oracle_ft_aaa_{suf}")
            fn.restype = C.c_double
            fn.argtypes = [_sz, _sz] + [_p] * 10 + [_sz, _p, _p]
            fn = getattr(_ORA, f"oracle_ft_aab_{suf}")
            fn.restype = C.c_double
            fn.argtypes = [_sz] * 4 + [_p] * 26 + [_sz, _p, _p]
    return _ORA


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libpt_ref.so"))


def _ref():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libpt_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run `make -C oracle ref`)")
        _REF = C.CDLL(path)
        for name, nsz, narr in (("rpt", 2, 7), ("upt", 4, 7), ("rft", 2, 10),
                                ("uft", 4, 10)):
            for suf in ("", "_mp"):
                fn = getattr(_REF, f"libpt_{name}{suf}")
                fn.restype = C.c_double
                fn.argtypes = [_sz] * nsz + [_p] * narr
    return _REF


def _ptrs(arrays, dtype):
    keep = [np.ascontiguousarray(a, dtype=dtype) for a in arrays]
    return keep, [a.ctypes.data for a in keep]


def energy(kind: str, dims, arrays, mp: bool = False) -> float:
    """Oracle energy of one API call: kind in rpt/upt/rft/uft, dims the size
    arguments, arrays the pt.h arrays in API order."""
    dt = np.float32 if mp else np.float64
    keep, ptrs = _ptrs(arrays, dt)
    fn = getattr(_lib(), f"oracle_{kind}_{'f' if mp else 'd'}")
    return fn(*[int(d) for d in dims], *ptrs)


def ref_energy(kind: str, dims, arrays, mp: bool = False) -> float:
    """Energy from the reference libpt compiled from its own sources."""
    dt = np.float32 if mp else np.float64
    keep, ptrs = _ptrs(arrays, dt)
    fn = getattr(_ref(), f"libpt_{kind}{'_mp' if mp else ''}")
    return fn(*[int(d) for d in dims], *ptrs)


def spin_case(case: str, dims, arrays, items=None, mp: bool = False):
    """Energy of one spin case (pt_aaa, pt_aab, ft_aaa, ft_aab) over a list
    of ijk items (int array [n,3]); returns (total, per-item energies)."""
    dt = np.float32 if mp else np.float64
    keep, ptrs = _ptrs(arrays, dt)
    fn = getattr(_lib(), f"oracle_{case}_{'f' if mp else 'd'}")
    if items is None:
        return fn(*[int(d) for d in dims], *ptrs, 0, None, None), None
    it = np.ascontiguousarray(items, dtype=np.int32).reshape(-1, 3)
    out = np.zeros(len(it))
    tot = fn(*[int(d) for d in dims], *ptrs, len(it), it.ctypes.data,
             out.ctypes.data)
    return tot, out
