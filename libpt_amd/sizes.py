"""Element counts of the pt.h arrays, as the reference test driver sizes
them (test.cpp:743-749 rpt, :783-791 upt, :825-834 rft, :874-887 uft)."""


def api_sizes(kind, dims):
    if kind in ("rpt", "rft"):
        oa, va = dims
        ob, vb = oa, va
    else:
        oa, va, ob, vb = dims
    ov = oa * va if kind in ("rpt", "rft") else oa * va + ob * vb
    if kind in ("rpt", "rft"):
        t2 = oa * oa * va * va + oa * ob * va * vb
        oovo = oa * oa * va * oa + oa * ob * va * ob
        ovvv = oa * va * va * (va - 1) // 2 + oa * vb * va * vb
    else:
        t2 = oa * oa * va * va + 2 * oa * ob * va * vb + ob * ob * vb * vb
        oovo = (oa * oa * va * oa + oa * ob * va * ob + ob * oa * vb * oa +
                ob * ob * vb * ob)
        ovvv = (oa * va * va * (va - 1) // 2 + oa * vb * va * vb +
                ob * va * vb * va + ob * vb * vb * (vb - 1) // 2)
    if kind in ("rpt", "upt"):
        return [ov, ov, ov, t2, oovo, t2, ovvv]
    return [ov, ov, ov, t2, t2, t2, oovo, ovvv, oovo, ovvv]
