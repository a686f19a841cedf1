"""Python mirror of the libpt C API, calling the HIP engine through its C ABI.

``rpt/upt/rft/uft`` (and ``*_mp``) take numpy arrays in exactly the pt.h
layouts (reference ``src/pt.h:37-178``) and return the energy, like
``libpt_rpt`` & co.  ``energy`` exposes the arithmetic forms of
``libpt_amd.h``; ``Plan`` wraps device-resident plans; ``synth`` the
deterministic input generator.

There is no fallback: if ``libpt_amd.so`` is missing or no GPU is present,
the calls raise.
"""

from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpt_amd.so")

FP64, FP32, BF16X3, F16X3 = 0, 1, 2, 3
FORMS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "f16x3": F16X3}
RPT, UPT, RFT, UFT = 0, 1, 2, 3
CALLS = {"rpt": RPT, "upt": UPT, "rft": RFT, "uft": UFT}

# synthetic tensor ids (libpt_amd/csrc/synth.h)
(T_D_OV, T_F_OV, T_T1, T_T2, T_OOVO, T_OOVV, T_OVVV, T_L1, T_L2, T_I2, T_I3,
 T_I6, T_I7) = range(13)
PT_ARGS = [T_D_OV, T_F_OV, T_T1, T_T2, T_OOVO, T_OOVV, T_OVVV]
FT_ARGS = [T_D_OV, T_F_OV, T_L1, T_T2, T_L2, T_OOVV, T_I2, T_I3, T_I6, T_I7]

_LIB = None
_sz, _p, _i, _i64, _d = C.c_size_t, C.c_void_p, C.c_int, C.c_int64, C.c_double


class LibptError(RuntimeError):
    pass


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("oa", C.c_int64), ("va", C.c_int64),
                ("ob", C.c_int64), ("vb", C.c_int64),
                ("restricted", C.c_int32), ("pad_", C.c_int32),
                ("omega", C.c_double), ("sig_t1", C.c_double),
                ("sig_t2", C.c_double), ("sig_int", C.c_double),
                ("sig_f", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("h2d_s", C.c_double),
                ("prep_s", C.c_double), ("kernel_s", C.c_double),
                ("flop", C.c_double), ("scale", C.c_double * 4)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise LibptError(f"{LIBPATH} not built (python -m libpt_amd.build)")
        L = C.CDLL(LIBPATH)
        for name, nsz, narr in (("rpt", 2, 7), ("upt", 4, 7), ("rft", 2, 10),
                                ("uft", 4, 10)):
            for suf in ("", "_mp"):
                fn = getattr(L, f"libpt_{name}{suf}")
                fn.restype = _d
                fn.argtypes = [_sz] * nsz + [_p] * narr
        L.libpt_amd_energy.restype = _d
        L.libpt_amd_energy.argtypes = [_i, _i, _sz, _sz, _sz, _sz,
                                       C.POINTER(_p), C.POINTER(Stats)]
        L.libpt_amd_energy_noalloc.restype = _d
        L.libpt_amd_energy_noalloc.argtypes = [_i, _i, _sz, _sz, _sz, _sz,
                                               _p, _p]
        L.libpt_amd_plan_create.restype = _p
        L.libpt_amd_plan_create.argtypes = [_i, _i, _sz, _sz, _sz, _sz,
                                            C.POINTER(_p)]
        L.libpt_amd_plan_destroy.argtypes = [_p]
        L.libpt_amd_plan_ncases.restype = _i
        L.libpt_amd_plan_ncases.argtypes = [_p]
        for f in ("nijk", "items"):
            fn = getattr(L, f"libpt_amd_plan_{f}")
            fn.restype = _i64
            fn.argtypes = [_p, _i]
        for f in ("flop_per_ijk", "factor"):
            fn = getattr(L, f"libpt_amd_plan_{f}")
            fn.restype = _d
            fn.argtypes = [_p, _i]
        L.libpt_amd_plan_run.restype = _i
        L.libpt_amd_plan_geometry.argtypes = [_p, _p]
        L.libpt_amd_plan_run.argtypes = [_p, _i, _i64, _i64]
        L.libpt_amd_plan_sync.restype = _i
        L.libpt_amd_plan_sync.argtypes = [_p]
        L.libpt_amd_plan_energy.restype = _d
        L.libpt_amd_plan_energy.argtypes = [_p]
        L.libpt_amd_plan_ijk_energies.restype = _i
        L.libpt_amd_plan_ijk_energies.argtypes = [_p, _i, _p]
        L.libpt_amd_plan_ijk_list.restype = _i
        L.libpt_amd_plan_ijk_list.argtypes = [_p, _i, _p]
        L.libpt_amd_plan_set_block.restype = _i
        L.libpt_amd_plan_set_block.argtypes = [_p, _i, _i, _i]
        L.libpt_amd_tune_record.restype = _i
        L.libpt_amd_tune_record.argtypes = [_i, _i, _sz, _sz, _sz, _sz, _i,
                                            _i, _i, _i]
        L.libpt_amd_plan_set_geometry.restype = _i
        L.libpt_amd_plan_set_geometry.argtypes = [_p, _i]
        L.libpt_amd_plan_tune.restype = _i
        L.libpt_amd_plan_tune.argtypes = [_p, _i, _p]
        L.libpt_amd_plan_timer_begin.restype = _i
        L.libpt_amd_plan_timer_begin.argtypes = [_p]
        L.libpt_amd_plan_timer_end_ms.restype = _d
        L.libpt_amd_plan_timer_end_ms.argtypes = [_p]
        L.libpt_amd_batch_create.restype = _p
        L.libpt_amd_batch_create.argtypes = [_i, _i, _i, C.POINTER(_i64),
                                             C.POINTER(_p)]
        L.libpt_amd_batch_destroy.argtypes = [_p]
        L.libpt_amd_batch_run.restype = _i
        L.libpt_amd_batch_run.argtypes = [_p]
        L.libpt_amd_batch_energies.restype = _i
        L.libpt_amd_batch_energies.argtypes = [_p, _p]
        L.libpt_amd_batch_last_ms.restype = _d
        L.libpt_amd_batch_last_ms.argtypes = [_p]
        L.libpt_amd_batch_flop.restype = _d
        L.libpt_amd_batch_flop.argtypes = [_p]
        L.libpt_amd_batch_items.restype = _i64
        L.libpt_amd_batch_items.argtypes = [_p]
        L.libpt_amd_batch_tune.restype = _i
        L.libpt_amd_batch_tune.argtypes = [_p, _i]
        L.libpt_amd_dev_alloc.restype = _p
        L.libpt_amd_dev_alloc.argtypes = [_sz]
        L.libpt_amd_dev_free.argtypes = [_p]
        L.libpt_amd_memcpy.restype = _i
        L.libpt_amd_memcpy.argtypes = [_p, _p, _sz, _i]
        L.libpt_amd_synth.restype = _i
        L.libpt_amd_synth.argtypes = [C.POINTER(SynthSpec), _i, _p, _i, _i]
        L.libpt_amd_synth_size.restype = _i64
        L.libpt_amd_synth_size.argtypes = [C.POINTER(SynthSpec), _i]
        L.libpt_amd_synth_range.argtypes = [C.POINTER(SynthSpec), _i, _p,
                                            _i64, _i64, _i]
        L.libpt_amd_copy_kernel.restype = _i
        L.libpt_amd_copy_kernel.argtypes = [_p, _p, _i64]
        L.libpt_amd_last_error.restype = C.c_char_p
        L.libpt_print_banner.argtypes = []
        L.libpt_set_malloc.argtypes = [_p]
        L.libpt_set_free.argtypes = [_p]
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().libpt_amd_last_error().decode()


def _arrs(arrays, dtype):
    keep = [np.ascontiguousarray(a, dtype=dtype) for a in arrays]
    return keep, [a.ctypes.data for a in keep]


def _drop(name, dims, arrays, mp):
    keep, ptrs = _arrs(arrays, np.float32 if mp else np.float64)
    fn = getattr(lib(), f"libpt_{name}{'_mp' if mp else ''}")
    return fn(*[int(d) for d in dims], *ptrs)


# pt.h mirror ---------------------------------------------------------------

def rpt(oa, va, d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv):
    return _drop("rpt", (oa, va), (d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv),
                 False)


def rpt_mp(oa, va, d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv):
    return _drop("rpt", (oa, va), (d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv),
                 True)


def upt(oa, va, ob, vb, d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv):
    return _drop("upt", (oa, va, ob, vb),
                 (d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv), False)


def upt_mp(oa, va, ob, vb, d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv):
    return _drop("upt", (oa, va, ob, vb),
                 (d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv), True)


def rft(oa, va, *arrays):
    return _drop("rft", (oa, va), arrays, False)


def rft_mp(oa, va, *arrays):
    return _drop("rft", (oa, va), arrays, True)


def uft(oa, va, ob, vb, *arrays):
    return _drop("uft", (oa, va, ob, vb), arrays, False)


def uft_mp(oa, va, ob, vb, *arrays):
    return _drop("uft", (oa, va, ob, vb), arrays, True)


def dropin(kind: str, dims, arrays, mp: bool = False) -> float:
    """pt.h call by name: kind in rpt/upt/rft/uft, dims the size args."""
    return _drop(kind, dims, arrays, mp)


# extension API --------------------------------------------------------------

def _dims4(kind, dims):
    if kind in ("rpt", "rft"):
        return (dims[0], dims[1], dims[0], dims[1])
    return tuple(dims)


def energy(kind: str, form, dims, arrays, with_stats=False):
    """libpt_amd_energy: host arrays (float64; float32 for FP32)."""
    form = FORMS.get(form, form)
    keep, ptrs = _arrs(arrays, np.float32 if form == FP32 else np.float64)
    arr = (_p * len(ptrs))(*ptrs)
    st = Stats()
    e = lib().libpt_amd_energy(CALLS[kind], form, *_dims4(kind, dims), arr,
                               C.byref(st))
    if math.isnan(e):
        raise LibptError(last_error())
    return (e, st) if with_stats else e


def energy_noalloc(kind: str, form, dims, dev: "DeviceArrays",
                   stream=None) -> float:
    """libpt_amd_energy_noalloc: device-resident inputs, every buffer from
    the library's static scratch (no device allocation); raises on error."""
    form = FORMS.get(form, form)
    arr = (_p * len(dev.ptrs))(*dev.ptrs)
    e = lib().libpt_amd_energy_noalloc(CALLS[kind], form, *_dims4(kind, dims),
                                       arr, stream)
    if math.isnan(e):
        raise LibptError(last_error())
    return e


def tune_record(kind: str, form, dims, block, waves=0):
    """libpt_amd_tune_record: geometry (waves 4 or 8, 0 for the default)
    and blocking (ga, gb, gc) for new plans of this call, form and size."""
    form = FORMS.get(form, form)
    if lib().libpt_amd_tune_record(CALLS[kind], form, *_dims4(kind, dims),
                                   *block, waves) != 0:
        raise LibptError(last_error())


@dataclass
class DeviceArrays:
    """Inputs of one call in HBM (owned; freed by close())."""
    kind: str
    spec: SynthSpec
    ptrs: list
    sizes: list
    fp32: bool

    def close(self):
        for p in self.ptrs:
            lib().libpt_amd_dev_free(p)
        self.ptrs = []


def make_spec(kind, dims, seed=1, omega=0.0, sig_t1=1e-2, sig_t2=1e-2,
              sig_int=5e-2, sig_f=0.0) -> SynthSpec:
    oa, va, ob, vb = _dims4(kind, dims)
    return SynthSpec(seed, oa, va, ob, vb, 1 if kind in ("rpt", "rft") else 0,
                     0, omega, sig_t1, sig_t2, sig_int, sig_f)


def tensor_ids(kind):
    return PT_ARGS if kind in ("rpt", "upt") else FT_ARGS


def synth_host(kind, spec: SynthSpec, fp32=False):
    """All input arrays of `kind` on the host, pt.h order."""
    out = []
    for t in tensor_ids(kind):
        n = lib().libpt_amd_synth_size(C.byref(spec), t)
        a = np.empty(n, dtype=np.float32 if fp32 else np.float64)
        if lib().libpt_amd_synth(C.byref(spec), t, a.ctypes.data, 0,
                                 int(fp32)) != 0:
            raise LibptError(last_error())
        out.append(a)
    return out


def synth_size(spec: SynthSpec, tensor: int) -> int:
    return int(lib().libpt_amd_synth_size(C.byref(spec), tensor))


def synth_range(spec: SynthSpec, tensor: int, out: np.ndarray,
                begin: int) -> None:
    """Elements [begin, begin + out.size) of one synthetic tensor into the
    contiguous host array out (float64 or float32)."""
    assert out.flags.c_contiguous and out.dtype in (np.float64, np.float32)
    if lib().libpt_amd_synth_range(C.byref(spec), tensor, out.ctypes.data,
                                   begin, out.size,
                                   int(out.dtype == np.float32)) != 0:
        raise LibptError(last_error())


def synth_device(kind, spec: SynthSpec, fp32=False) -> DeviceArrays:
    ptrs, sizes = [], []
    el = 4 if fp32 else 8
    for t in tensor_ids(kind):
        n = lib().libpt_amd_synth_size(C.byref(spec), t)
        p = lib().libpt_amd_dev_alloc(max(n, 1) * el)
        if not p:
            for q in ptrs:
                lib().libpt_amd_dev_free(q)
            raise LibptError(last_error())
        if lib().libpt_amd_synth(C.byref(spec), t, p, 1, int(fp32)) != 0:
            raise LibptError(last_error())
        ptrs.append(p)
        sizes.append(n)
    return DeviceArrays(kind, spec, ptrs, sizes, fp32)


def to_device(kind, arrays, fp32=False) -> DeviceArrays:
    ptrs, sizes = [], []
    for a in arrays:
        a = np.ascontiguousarray(a, dtype=np.float32 if fp32 else np.float64)
        p = lib().libpt_amd_dev_alloc(max(a.nbytes, 8))
        if not p:
            raise LibptError(last_error())
        if a.nbytes and lib().libpt_amd_memcpy(p, a.ctypes.data, a.nbytes,
                                               1) != 0:
            raise LibptError(last_error())
        ptrs.append(p)
        sizes.append(a.size)
    return DeviceArrays(kind, None, ptrs, sizes, fp32)


def from_device(ptr, n, fp32=False):
    a = np.empty(n, dtype=np.float32 if fp32 else np.float64)
    if n and lib().libpt_amd_memcpy(a.ctypes.data, ptr, a.nbytes, 2) != 0:
        raise LibptError(last_error())
    return a


class Plan:
    """Device-resident plan (libpt_amd_plan_*)."""

    def __init__(self, kind, form, dims, dev: DeviceArrays):
        form = FORMS.get(form, form)
        self.kind, self.form = kind, form
        arr = (_p * len(dev.ptrs))(*dev.ptrs)
        self.h = lib().libpt_amd_plan_create(CALLS[kind], form,
                                             *_dims4(kind, dims), arr)
        if not self.h:
            raise LibptError(last_error())
        self.ncases = lib().libpt_amd_plan_ncases(self.h)

    def nijk(self, sc):
        return lib().libpt_amd_plan_nijk(self.h, sc)

    def items(self, sc):
        return lib().libpt_amd_plan_items(self.h, sc)

    def flop_per_ijk(self, sc):
        return lib().libpt_amd_plan_flop_per_ijk(self.h, sc)

    def factor(self, sc):
        return lib().libpt_amd_plan_factor(self.h, sc)

    def geometry(self):
        """(waves per workgroup, c tile width, ijk per item group)"""
        out = (C.c_int * 3)()
        if lib().libpt_amd_plan_geometry(self.h, out) != 0:
            raise LibptError(last_error())
        return tuple(out)

    def run(self, sc, b=0, e=None):
        if e is None:
            e = self.nijk(sc)
        if lib().libpt_amd_plan_run(self.h, sc, b, e) != 0:
            raise LibptError(last_error())

    def run_all(self):
        for sc in range(self.ncases):
            self.run(sc)

    def sync(self):
        if lib().libpt_amd_plan_sync(self.h) != 0:
            raise LibptError(last_error())

    def energy(self):
        e = lib().libpt_amd_plan_energy(self.h)
        if math.isnan(e):
            raise LibptError(last_error())
        return e

    def ijk_energies(self, sc):
        out = np.zeros(self.nijk(sc))
        if lib().libpt_amd_plan_ijk_energies(self.h, sc, out.ctypes.data) != 0:
            raise LibptError(last_error())
        return out

    def ijk_list(self, sc):
        out = np.zeros((self.nijk(sc), 3), dtype=np.int32)
        if lib().libpt_amd_plan_ijk_list(self.h, sc, out.ctypes.data) != 0:
            raise LibptError(last_error())
        return out

    def set_block(self, ga, gb, gc):
        if lib().libpt_amd_plan_set_block(self.h, ga, gb, gc) != 0:
            raise LibptError(last_error())

    def set_geometry(self, waves):
        if lib().libpt_amd_plan_set_geometry(self.h, waves) != 0:
            raise LibptError(last_error())

    def tune(self, sample_ijk=8):
        """Autotune the tile geometry and order (see libpt_amd_plan_tune);
        returns the chosen blocking (ga, gb, gc); geometry() gives the
        chosen waves."""
        out = (C.c_int * 4)()
        if lib().libpt_amd_plan_tune(self.h, sample_ijk, out) != 0:
            raise LibptError(last_error())
        return tuple(out[:3])

    def timer_begin(self):
        lib().libpt_amd_plan_timer_begin(self.h)

    def timer_end_ms(self):
        return lib().libpt_amd_plan_timer_end_ms(self.h)

    def close(self):
        if self.h:
            lib().libpt_amd_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many independent systems of one call kind in one graph launch
    (libpt_amd_batch_*).  `systems` is a list of (dims, DeviceArrays)."""

    def __init__(self, kind, form, systems):
        form = FORMS.get(form, form)
        self.kind, self.form, self.n = kind, form, len(systems)
        dims = []
        ptrs = []
        for d, dev in systems:
            dims.extend(_dims4(kind, d))
            ptrs.extend(dev.ptrs)
        cd = (_i64 * len(dims))(*dims)
        cp = (_p * len(ptrs))(*ptrs)
        self.h = lib().libpt_amd_batch_create(CALLS[kind], form, self.n, cd,
                                              cp)
        if not self.h:
            raise LibptError(last_error())

    def run(self):
        if lib().libpt_amd_batch_run(self.h) != 0:
            raise LibptError(last_error())

    def energies(self):
        out = np.zeros(self.n)
        if lib().libpt_amd_batch_energies(self.h, out.ctypes.data) != 0:
            raise LibptError(last_error())
        return out

    def last_ms(self):
        return lib().libpt_amd_batch_last_ms(self.h)

    def tune(self, sample_ijk=8):
        """Autotune every system's tile order (libpt_amd_batch_tune);
        returns the number of systems timed."""
        r = lib().libpt_amd_batch_tune(self.h, sample_ijk)
        if r < 0:
            raise LibptError(last_error())
        return r

    @property
    def flop(self):
        return lib().libpt_amd_batch_flop(self.h)

    @property
    def items(self):
        return lib().libpt_amd_batch_items(self.h)

    def close(self):
        if self.h:
            lib().libpt_amd_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def batch_dims(n=64, seed=64, o=(4, 20), v=(40, 160)):
    """Sizes of a batch of restricted systems: o and v uniform integers in
    the given closed ranges, from a seeded generator."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(o[0], o[1] + 1)), int(rng.integers(v[0], v[1] + 1)))
            for _ in range(n)]


def device_copy_calibration(n_doubles: int) -> None:
    """Copy n doubles device-to-device with the engine's own copy kernel
    (reads 8n bytes, writes 8n bytes): the reference point for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    L = lib()
    a = L.libpt_amd_dev_alloc(8 * n_doubles)
    b = L.libpt_amd_dev_alloc(8 * n_doubles)
    if not a or not b:
        raise LibptError(last_error())
    try:
        if L.libpt_amd_copy_kernel(b, a, n_doubles) != 0:
            raise LibptError(last_error())
    finally:
        L.libpt_amd_dev_free(a)
        L.libpt_amd_dev_free(b)


def flop(kind, dims):
    """Reference-formulation GEMM flops of one call (closed form)."""
    def same(o, v):
        return (o * (o - 1) * (o - 2) // 6) * (3.0 * v**3 * (v - 1) + 6.0 * o * v**3)

    def mixed(oa, va, ob, vb):
        per = (2 * va**3 * vb + 4 * va**2 * vb**2 + 2 * vb * va**2 * (va - 1)
               + 6 * oa * va**2 * vb + 4 * ob * va**2 * vb)
        return (oa * (oa - 1) // 2) * ob * float(per)

    oa, va, ob, vb = _dims4(kind, dims)
    if kind in ("rpt", "rft"):
        f = same(oa, va) + mixed(oa, va, oa, va)
    else:
        f = same(oa, va) + same(ob, vb) + mixed(oa, va, ob, vb) + \
            mixed(ob, vb, oa, va)
    return f * (2 if kind in ("rft", "uft") else 1)
