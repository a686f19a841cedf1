"""Benchmark: restricted CCSD(T) (libpt_rpt) at o=45, v=477 on one MI355X.

The workload is the README system of the reference (9 H2O / cc-pVTZ,
522 functions: o=45, v=477) with synthetic amplitudes and integrals from
the deterministic generator, generated directly in HBM (74 GB FP64).

A *step* is one batch of the ijk work list at full shape: 1/NBATCH of the
same-spin and of the mixed-spin ijk triples (every triple costs the same,
so each batch is 1/NBATCH of the whole (T) correction).  The timed region
holds K steps bracketed by device synchronisation.  `value` is the
algorithmic throughput of the FP64 MFMA path (the libpt_rpt drop-in),
counted with the reference formulation's GEMM flops; the split-BF16 form
runs the same steps and is reported beside it with its error against FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("(T) wall s + TFLOP/s at o=45 v=477: split-BF16 vs FP64 MFMA, "
          "|dE|<=1e-6 Eh")
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "f16x3": 2500.0}
EXEC_FACTOR = {"fp64": 1, "fp32": 1, "bf16x3": 3, "f16x3": 3}
# fastest published run of this system (README.md:52): 440 s on 4096 cores
PUBLISHED_WALL_S = 440.0
# HBM traffic of the dominant launch at the default configuration, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel
TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "k_tiles_fp64_rpt_o45_v477_traffic.json")


def traffic_of(form, nbatch, o, v):
    """HBM bytes per launch of the dominant kernel, if a counter summary
    for this configuration is committed; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if (t.get("form"), t.get("nbatch"), t.get("o"), t.get("v")) != \
            (form, nbatch, o, v):
        return None
    return t.get("bytes_per_launch")


def host_info():
    info = {"nproc": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    info["cpu"] = ln.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            info["mem_gb"] = round(int(f.readline().split()[1]) / 2**20, 1)
    except OSError:
        pass
    return info


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(api, form, dims, dev, args, nbatch):
    """Warm up, then time K steps; returns a dict of measurements."""
    t0 = time.perf_counter()
    plan = api.Plan("rpt", form, dims, dev)
    prep = time.perf_counter() - t0
    blk, tune_s = "default 8x8x2", 0.0
    if args.tune:
        t0 = time.perf_counter()
        blk = plan.tune(args.tune_sample)
        tune_s = time.perf_counter() - t0
        log(f"[bench] {form}: tile-order blocking {blk} "
            f"(tuned in {tune_s:.1f} s)")
    ncase = plan.ncases
    nijk = [plan.nijk(sc) for sc in range(ncase)]
    fpi = [plan.flop_per_ijk(sc) for sc in range(ncase)]

    def rng(sc, b):
        n = nijk[sc]
        return b * n // nbatch, (b + 1) * n // nbatch

    def step(b, times):
        for sc in range(ncase):
            lo, hi = rng(sc, b)
            plan.timer_begin()
            plan.run(sc, lo, hi)
            times[sc] += plan.timer_end_ms()

    for w in range(args.warmup):
        step(w % nbatch, [0.0] * ncase)
        log(f"[bench] {form} warmup step {w} done")
    plan.sync()
    kern_ms = [0.0] * ncase
    flop = 0.0
    batches = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        b = (args.warmup + s) % nbatch
        batches.append(b)
        step(b, kern_ms)
        for sc in range(ncase):
            lo, hi = rng(sc, b)
            flop += (hi - lo) * fpi[sc]
        log(f"[bench] {form} step {s} (batch {b}) done")
    plan.sync()
    wall = time.perf_counter() - t0
    # batch energies (with the call's factor) for the cross-form check
    e_batches = 0.0
    for sc in range(ncase):
        ee = plan.ijk_energies(sc)
        for b in set(batches):
            lo, hi = rng(sc, b)
            e_batches += plan.factor(sc) * float(ee[lo:hi].sum())
    dom = max(range(ncase), key=lambda sc: nijk[sc] * fpi[sc])
    dom_flop = sum((rng(dom, b)[1] - rng(dom, b)[0]) * fpi[dom]
                   for b in batches)
    waves, tcw, kgrp = plan.geometry()
    # per spin case: launch time per step and its own rate
    cases = []
    for sc in range(ncase):
        f = sum((rng(sc, b)[1] - rng(sc, b)[0]) * fpi[sc] for b in batches)
        if kern_ms[sc] > 0:
            cases.append({"case": "aaa" if sc == 0 else "aab",
                          "ijk_per_step": rng(sc, batches[0])[1] -
                          rng(sc, batches[0])[0],
                          "launch_ms": round(kern_ms[sc] / args.steps, 2),
                          "tflops": round(f / kern_ms[sc] / 1e9, 2)})
    res = dict(form=form, prep_s=prep, tile_block=blk, tune_s=tune_s,
               geometry={"waves": waves, "tile": f"32x32x{tcw}",
                         "ijk_group": kgrp},
               wall_s=wall, flop=flop,
               tflops=flop / wall / 1e12, ms_per_step=1e3 * wall / args.steps,
               kern_ms=kern_ms, dom_case=dom,
               dom_launch_ms=kern_ms[dom] / args.steps,
               dom_flop_per_launch=dom_flop / args.steps,
               e_batches=e_batches, cases=cases,
               total_flop=sum(n * f for n, f in zip(nijk, fpi)))
    plan.close()
    return res


def run_batch(api, form, n, seed, tune=0, sample=8):
    """n independent restricted systems (o in [4, 20], v in [40, 160]) in
    one graph launch; one untimed replay first.  tune: autotune every
    system's tile geometry and order first (each distinct size timed once);
    the plans timed one after another then take the same choices."""
    dims = api.batch_dims(n, seed=seed)
    systems = [(d, api.synth_device("rpt", api.make_spec("rpt", d,
                                                         seed=seed + 1 + k),
                                    fp32=form == "fp32"))
               for k, d in enumerate(dims)]
    try:
        timed, tune_s = 0, 0.0
        if tune:
            # the geometry of a batch's plans is fixed when its graph is
            # recorded: tune each distinct size on a plan of its own first
            # (geometry and order go to the tuning cache, which the
            # batch's plans read)
            t0 = time.perf_counter()
            seen = set()
            for d, dev in systems:
                if d in seen:
                    continue
                seen.add(d)
                with api.Plan("rpt", form, d, dev) as pl:
                    pl.tune(sample)
                timed += 1
            tune_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = api.Batch("rpt", form, systems)
        prep = time.perf_counter() - t0
        with b:
            b.run()
            b.energies()
            t0 = time.perf_counter()
            b.run()
            e = b.energies()
            wall = time.perf_counter() - t0
            ms = b.last_ms()
            flop, items = b.flop, b.items
        # the same systems one after another, one plan each (plans made
        # beforehand; the time covers their runs and energies only)
        plans = [api.Plan("rpt", form, d, dev) for d, dev in systems]
        try:
            for pl in plans[:2]:
                pl.run_all()
                pl.energy()
            t0 = time.perf_counter()
            e_seq = [(pl.run_all(), pl.energy())[1] for pl in plans]
            seq = time.perf_counter() - t0
        finally:
            for pl in plans:
                pl.close()
    finally:
        for _, dev in systems:
            dev.close()
    return {"systems": n, "o_range": [4, 20], "v_range": [40, 160],
            "form": form, "items": items, "flop": flop,
            "prep_s": round(prep, 3), "tuned_sizes": timed,
            "tune_s": round(tune_s, 2), "graph_ms": round(ms, 3),
            "wall_s": round(wall, 4), "tflops": round(flop / ms / 1e9, 3),
            "sequential_plans_s": round(seq, 4),
            "speedup_vs_sequential": round(seq / wall, 3),
            "same_energies_as_sequential": bool(np.array_equal(e, e_seq)),
            "E_sum": float(np.sum(e))}


# the other configurations of the metric's definition, at their real shapes:
# (name, call, dims, fraction of each spin case's ijk list, generator args)
CONFIGS = [
    ("rpt o=8 v=64", "rpt", (8, 64), 1, dict(sig_f=1e-2)),
    ("upt 24/200/20/204", "upt", (24, 200, 20, 204), 1, dict()),
    ("rft o=30 v=300 omega=0.30", "rft", (30, 300), 16, dict(omega=0.3)),
    ("uft 16/160/14/160 omega=0.30", "uft", (16, 160, 14, 160), 1,
     dict(omega=0.3)),
]


def run_config(api, name, kind, dims, frac, kw, forms, seed):
    """One configuration in each form: the first 1/frac of every spin
    case's ijk list (all of it for frac = 1), GPU-timed, with the energy of
    the part run and its difference from the first form."""
    spec = api.make_spec(kind, dims, seed=seed, **kw)
    dev = api.synth_device(kind, spec)
    # the FP32 (_mp) form reads float arrays of the same generator
    dev32 = api.synth_device(kind, spec, fp32=True) if "fp32" in forms \
        else None
    out = {"config": name, "call": kind, "dims": list(dims),
           "fraction_of_ijk": f"1/{frac}"}
    try:
        e0 = None
        for form in forms:
            t0 = time.perf_counter()
            with api.Plan(kind, form, dims,
                          dev32 if form == "fp32" else dev) as pl:
                prep = time.perf_counter() - t0
                ms, flop, e = 0.0, 0.0, 0.0
                for sc in range(pl.ncases):
                    n = pl.nijk(sc)
                    m = n if frac == 1 else max(1, n // frac)
                    if n == 0:
                        continue
                    pl.timer_begin()
                    pl.run(sc, 0, m)
                    ms += pl.timer_end_ms()
                    flop += m * pl.flop_per_ijk(sc)
                    e += pl.factor(sc) * float(pl.ijk_energies(sc)[:m].sum())
            r = {"tflops": round(flop / ms / 1e9, 3), "kernel_s": round(
                ms / 1e3, 3), "prep_s": round(prep, 3), "flop": flop,
                 "E": e}
            if e0 is None:
                e0 = e
            else:
                r["rel_dE_vs_" + forms[0]] = abs(e - e0) / abs(e0)
            out[form] = r
            log(f"[bench] {name} {form}: {r}")
    finally:
        dev.close()
        if dev32 is not None:
            dev32.close()
    return out


def omp_set_threads(n):
    """Set the OpenMP thread count of later parallel regions started from
    this thread (libgomp, shared by the reference library)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(api, o, vs, seed, full, check=False):
    """The reference libpt compiled from its sources (OpenMP threads spread
    over the cores, MKL in sequential mode) on a bounded sample: whole calls
    over the full o=45 ijk list at reduced v (each v of `vs` in turn, one
    progress line each), the rate of the last (largest) one extrapolated to
    the full v by the reference formulation's flop ratio; falls back to the
    oracle.  Also the reference's own thread spread: one small call (rpt
    o=8 v=64) on all threads and on one thread."""
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle
    threads = ref_threads()
    kind = "reference" if oracle.ref_available() else "port"
    fn = oracle.ref_energy if kind == "reference" else oracle.energy
    runs = []
    for v in vs:
        arr = api.synth_host("rpt", api.make_spec("rpt", (o, v), seed=seed))
        t0 = time.perf_counter()
        e = fn("rpt", (o, v), arr)
        dt = time.perf_counter() - t0
        f = api.flop("rpt", (o, v))
        runs.append({"v": v, "flop": f, "s": round(dt, 2),
                     "tflops": round(f / dt / 1e12, 4), "E": e})
        log(f"[bench] cpu baseline: libpt_rpt o={o} v={v} on {threads} "
            f"threads, {dt:.1f} s, {f / dt / 1e12:.3f} TFLOP/s")
    v, f, dt, e = runs[-1]["v"], runs[-1]["flop"], runs[-1]["s"], \
        runs[-1]["E"]
    f_full = api.flop("rpt", full)
    rate = f / dt
    out = {"value": round(rate / 1e12, 4), "unit": "TFLOP/s",
           "cores": threads, "kind": kind, "host": host_info(),
           "blas": "MKL (MKL_THREADING_LAYER=SEQUENTIAL)" if kind ==
           "reference" else "none (restatement)",
           "omp": {k: os.environ.get(k) for k in
                   ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
           "sample": f"libpt_rpt o={o} v={v} FP64 (full call: every ijk "
                     f"triple of o={o} at v={v}, {f:.3e} flop, {dt:.1f} s, "
                     f"E={e:.6e})",
           "samples": runs,
           "extrapolated_full_s": {
               "dims": list(full), "s": round(f_full / rate, 1),
               "how": "flop ratio of the reference formulation "
                      f"({f_full:.4e} / {f:.4e}) at the sample's rate"}}
    small = api.synth_host("rpt", api.make_spec("rpt", (8, 64), seed=seed,
                                                sig_f=1e-2))
    e_n = fn("rpt", (8, 64), small)
    if kind == "reference":
        omp_set_threads(1)
        e_1 = fn("rpt", (8, 64), small)
        omp_set_threads(threads)
        out["thread_spread"] = {
            "call": "libpt_rpt o=8 v=64", f"E_{threads}_threads": e_n,
            "E_1_thread": e_1, "abs_dE": abs(e_n - e_1)}
        log(f"[bench] cpu baseline: thread spread at rpt o=8 v=64, "
            f"|dE| {abs(e_n - e_1):.2e} Eh")
    if check:
        # the same sample through the GPU drop-in (host arrays in), checked
        # against the CPU energy; opt-in, since its launches share the tile
        # kernels and would mix into a profiled default run's statistics
        arr = api.synth_host("rpt", api.make_spec("rpt", (o, v), seed=seed))
        t0 = time.perf_counter()
        e_gpu = api.rpt(o, v, *arr)
        gdt = time.perf_counter() - t0
        out["gpu_same_sample"] = {"E_gpu": e_gpu, "E_cpu": e,
                                  "rel_dE": abs(e_gpu - e) / abs(e),
                                  "wall_s": round(gdt, 3)}
    return out


FIXTURES = ["rpt02", "rpt09", "upt01", "upt03", "uft02"]


def split_on_fixtures(api, forms):
    """The split forms on libpt's known-answer inputs (molecular amplitudes
    and integrals, |E| ~ 1e-3 Eh): |dE| in Eh against each file's golden
    energy (8 decimals), next to the synthetic relative error."""
    from libpt_amd.fixtures import read_fixture
    out = {}
    for form in forms:
        rows = {}
        for name in FIXTURES:
            fx = read_fixture(os.path.join(ROOT, "tests", "golden",
                                           name + ".dat.gz"))
            dims = (fx.oa, fx.va) if fx.kind in ("rpt", "rft") else \
                (fx.oa, fx.va, fx.ob, fx.vb)
            e = api.energy(fx.kind, form, dims, fx.args())
            rows[name] = {"E": e, "E_golden": fx.e_ref,
                          "abs_dE_Eh": abs(e - fx.e_ref)}
        out[form] = {"max_abs_dE_Eh": max(r["abs_dE_Eh"]
                                          for r in rows.values()),
                     "per_fixture": rows}
        log(f"[bench] {form} on the known-answer inputs: max |dE| "
            f"{out[form]['max_abs_dE_Eh']:.2e} Eh")
    return out


def ref_threads():
    """OpenMP threads of the CPU reference: OMP_NUM_THREADS, else the host
    cores this process may use, at most 16"""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if threads <= 0:
        threads = min(16, len(os.sched_getaffinity(0)))
        os.environ["OMP_NUM_THREADS"] = str(threads)
    return threads


def ref_check(kind, dims, seed, omega):
    """One FP64 drop-in call on the GPU and the same call of libpt compiled
    from its sources (OpenMP on the host cores, MKL sequential) on the same
    synthetic host arrays: both energies, their difference and both wall
    times."""
    from libpt_amd import api
    import oracle
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    threads = ref_threads()
    if not oracle.ref_available():
        print("--ref-check needs oracle/_ref (libpt built from its sources)",
              file=sys.stderr)
        sys.exit(2)
    spec = api.make_spec(kind, dims, seed=seed, omega=omega)
    dev = api.synth_device(kind, spec)
    host = [api.from_device(p, n) for p, n in zip(dev.ptrs, dev.sizes)]
    dev.close()
    f = api.flop(kind, dims)
    t0 = time.perf_counter()
    e_gpu = api.dropin(kind, dims, host)
    tg = time.perf_counter() - t0
    log(f"[bench] ref check: GPU libpt_{kind} done in {tg:.2f} s; "
        f"reference on {threads} threads ({f / 0.4e12:.0f} s at 0.4 TFLOP/s)")
    t0 = time.perf_counter()
    e_ref = oracle.ref_energy(kind, dims, host)
    tr = time.perf_counter() - t0
    print(json.dumps({"ref_check": {
        "call": f"libpt_{kind}", "dims": list(dims), "omega": omega,
        "seed": seed, "flop": f, "E_gpu": e_gpu, "E_ref": e_ref,
        "abs_dE": abs(e_gpu - e_ref),
        "rel_dE": abs(e_gpu - e_ref) / abs(e_ref) if e_ref else None,
        "gpu_wall_s": round(tg, 3), "gpu_tflops": round(f / tg / 1e12, 3),
        "ref_wall_s": round(tr, 2), "ref_tflops": round(f / tr / 1e12, 4),
        "ref_cores": threads, "host": host_info(),
        "blas": "MKL (MKL_THREADING_LAYER=SEQUENTIAL)"}}), flush=True)


def ref_check_batch(n, seed):
    """The n-system restricted batch (o in [4, 20], v in [40, 160]) as one
    FP64 graph on the GPU, and every system in libpt compiled from its
    sources on the same arrays: per-system energies and the largest
    difference."""
    from libpt_amd import api
    import oracle
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    threads = ref_threads()
    if not oracle.ref_available():
        print("--ref-check needs oracle/_ref (libpt built from its sources)",
              file=sys.stderr)
        sys.exit(2)
    dims = api.batch_dims(n, seed=seed)
    specs = [api.make_spec("rpt", d, seed=seed + 1 + k)
             for k, d in enumerate(dims)]
    systems = [(d, api.synth_device("rpt", s)) for d, s in zip(dims, specs)]
    try:
        with api.Batch("rpt", "fp64", systems) as b:
            b.run()
            e_gpu = b.energies()
            ms = b.last_ms()
    finally:
        for _, dev in systems:
            dev.close()
    rows, tr = [], 0.0
    for k, (d, s) in enumerate(zip(dims, specs)):
        t0 = time.perf_counter()
        e_ref = oracle.ref_energy("rpt", d, api.synth_host("rpt", s))
        tr += time.perf_counter() - t0
        rows.append({"dims": list(d), "E_gpu": float(e_gpu[k]),
                     "E_ref": e_ref, "abs_dE": abs(float(e_gpu[k]) - e_ref)})
        log(f"[bench] ref check batch: system {k} {d} |dE| "
            f"{rows[-1]['abs_dE']:.2e}")
    worst = max(rows, key=lambda r: r["abs_dE"] / max(abs(r["E_ref"]), 1e-300))
    print(json.dumps({"ref_check_batch": {
        "systems": n, "seed": seed, "graph_ms": round(ms, 2),
        "ref_wall_s": round(tr, 2), "ref_cores": threads,
        "max_abs_dE": max(r["abs_dE"] for r in rows),
        "max_rel_dE": worst["abs_dE"] / abs(worst["E_ref"]),
        "worst": worst, "per_system": rows}}), flush=True)


def full_call(form, kind, dims, seed, omega):
    """One pt.h-style call with host arrays: the inputs are generated in
    HBM, copied to pageable host memory and freed on the device, then the call
    copies them back, packs its operands and runs every ijk triple.  Progress
    lines (one per 1/16 of a spin case) come from the library."""
    from libpt_amd import api
    spec = api.make_spec(kind, dims, seed=seed, omega=omega)
    t0 = time.perf_counter()
    dev = api.synth_device(kind, spec)
    host = [api.from_device(p, n) for p, n in zip(dev.ptrs, dev.sizes)]
    dev.close()
    if form == "fp32":
        # the _mp drop-in takes float arrays: round them before the timer
        host = [a.astype(np.float32) for a in host]
    gb = sum(a.nbytes for a in host) / 1e9
    log(f"[bench] full call: {gb:.1f} GB of host inputs ready in "
        f"{time.perf_counter() - t0:.1f} s")
    os.environ["LIBPT_AMD_PROGRESS"] = "1"
    t0 = time.perf_counter()
    e, st = api.energy(kind, form, dims, host, with_stats=True)
    wall = time.perf_counter() - t0
    print(json.dumps({"full_call": {
        "call": f"libpt_{kind}" if form == "fp64" else f"libpt_{kind}_mp"
        if form == "fp32" else f"libpt_amd_energy({kind}, {form})",
        "dims": list(dims), "omega": omega, "form": form,
        "host_input_gb": round(gb, 2),
        "wall_s": round(wall, 2), "h2d_s": round(st.h2d_s, 2),
        "prep_s": round(st.prep_s, 2), "kernel_s": round(st.kernel_s, 2),
        "flop": st.flop, "tflops_wall": round(st.flop / wall / 1e12, 3),
        "tflops_kernel": round(st.flop / st.kernel_s / 1e12, 3),
        "E": e}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--o", type=int, default=45)
    ap.add_argument("--v", type=int, default=477)
    ap.add_argument("--nbatch", type=int, default=128)
    ap.add_argument("--forms", default="fp64,bf16x3",
                    help="comma list of fp64, fp32 (the _mp form, float "
                    "inputs), bf16x3, f16x3; the first is the headline and "
                    "the others report their energy difference against it")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tune", type=int, default=0,
                    help="autotune the tile order before timing (1), also "
                         "of every batch system with --batch; the "
                         "default blocking is the tuner's choice for this "
                         "size to within 1 %%, and leaving the tuner's sample "
                         "launches out keeps the kernel statistics of a "
                         "profiled run equal to the timed launches")
    ap.add_argument("--tune-sample", type=int, default=8,
                    help="ijk triples per spin case per tuning candidate")
    ap.add_argument("--configs", type=int, default=0,
                    help="also run the other configurations (small rpt, upt, "
                    "rft, uft) in each form (1); off by default so that the "
                    "kernel statistics of a profiled default run cover the "
                    "metric's launches only")
    ap.add_argument("--batch", type=int, default=0,
                    help="also time a batch of this many independent small "
                    "systems (64: o in [4, 20], v in [40, 160]) as one graph "
                    "against their plans run one after another; off by "
                    "default, since its launches share the tile kernels and "
                    "would mix into a profiled default run's statistics")
    ap.add_argument("--cpu", default="45,32,48",
                    help="o,v1[,v2...] of the CPU-baseline sample (whole "
                    "calls over every ijk triple of o at each reduced v; "
                    "the last one's rate is extrapolated to --o/--v by the "
                    "flop ratio); 'none' to skip")
    ap.add_argument("--fixtures", type=int, default=1,
                    help="run the split forms of --forms on libpt's five "
                    "known-answer inputs and report |dE| in Eh (1)")
    ap.add_argument("--cpu-check", type=int, default=0,
                    help="also run the CPU-baseline sample through the GPU "
                    "drop-in and report the energy difference (1)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a device copy of this many GB (for scaling "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--full-call", default="",
                    help="only time one whole host-array call (host to device "
                         "copy included) in this form and exit; fp64 is the "
                         "pt.h drop-in path")
    ap.add_argument("--ref-check", type=int, default=0,
                    help="only run one FP64 drop-in call of --call/--dims on "
                    "the GPU and in libpt compiled from its sources, on the "
                    "same host arrays, and exit (1); with --batch n, the "
                    "n-system batch instead, every system checked")
    ap.add_argument("--call", default="rpt",
                    help="--full-call, --ref-check: rpt, upt, rft or uft")
    ap.add_argument("--dims", default="",
                    help="--full-call: o,v (restricted) or oa,va,ob,vb; "
                         "default --o,--v")
    ap.add_argument("--omega", type=float, default=0.3,
                    help="--full-call of rft/uft: excitation energy (Eh)")
    args = ap.parse_args()
    if args.ref_check and args.batch > 0:
        ref_check_batch(args.batch, args.seed)
        return
    if args.ref_check:
        dims = tuple(int(x) for x in args.dims.split(",")) if args.dims \
            else (args.o, args.v)
        ref_check(args.call, dims, args.seed,
                  args.omega if args.call.endswith("ft") else 0.0)
        return
    if args.full_call:
        dims = tuple(int(x) for x in args.dims.split(",")) if args.dims \
            else (args.o, args.v)
        full_call(args.full_call, args.call, dims, args.seed,
                  args.omega if args.call.endswith("ft") else 0.0)
        return
    if args.copy_calibration > 0:
        from libpt_amd import api
        n = int(args.copy_calibration * 1e9) // 8
        t0 = time.perf_counter()
        api.device_copy_calibration(n)
        print(json.dumps({"copy_calibration_bytes": 8 * n,
                          "s": time.perf_counter() - t0}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is "
              "not supported (the ijk work list would be split across GPUs "
              "with one 8-byte sum)", file=sys.stderr)
        sys.exit(2)

    from libpt_amd import api
    dims = (args.o, args.v)
    spec = api.make_spec("rpt", dims, seed=args.seed)
    t0 = time.perf_counter()
    dev = api.synth_device("rpt", spec)
    gen_s = time.perf_counter() - t0
    gb = sum(n * 8 for n in dev.sizes) / 1e9
    log(f"[bench] generated {gb:.1f} GB of FP64 inputs in HBM in {gen_s:.1f} s")

    forms = [f for f in args.forms.split(",") if f]
    res = {}
    for form in forms:
        if form == "fp32":
            # the _mp form reads float inputs: the same generator, rounded
            dev32 = api.synth_device("rpt", spec, fp32=True)
            try:
                res[form] = run_form(api, form, dims, dev32, args,
                                     args.nbatch)
            finally:
                dev32.close()
        else:
            res[form] = run_form(api, form, dims, dev, args, args.nbatch)
        r = res[form]
        log(f"[bench] {form}: {r['tflops']:.2f} TFLOP/s, "
            f"{r['ms_per_step']:.1f} ms/step, dominant launch "
            f"{r['dom_launch_ms']:.1f} ms")
    dev.close()

    main_form = forms[0]
    r = res[main_form]
    dom_tflops = r["dom_flop_per_launch"] / (r["dom_launch_ms"] * 1e-3) / 1e12
    peak = PEAK[main_form]
    achieved = dom_tflops * EXEC_FACTOR[main_form]
    proj_wall = r["total_flop"] / (r["tflops"] * 1e12)
    out = {
        "metric": METRIC,
        "value": round(r["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(PUBLISHED_WALL_S / proj_wall, 4),
        "dtype": "f64" if main_form == "fp64" else main_form,
        "data": "synthetic (deterministic counter-based generator, in HBM)",
        "config": {
            "workload": f"libpt_rpt restricted CCSD(T) o={args.o} v={args.v}, "
                        f"step = 1/{args.nbatch} of the ijk work list",
            "o": args.o, "v": args.v, "form": main_form,
            "flop_per_step": r["flop"] / args.steps,
            "total_flop": r["total_flop"],
            "projected_full_T_wall_s": round(proj_wall, 1),
            "tile_block": r["tile_block"],
            "geometry": r["geometry"],
            "spin_cases": r["cases"],
        },
        "roofline": {
            "bound": "mfma",
            "kernel": f"k_tiles<{main_form}, mixed-spin (T)>",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic_of(main_form, args.nbatch, args.o, args.v),
        },
    }
    for form in forms[1:]:
        q = res[form]
        out[form] = {
            "tflops": round(q["tflops"], 3),
            "ms_per_step": round(q["ms_per_step"], 2),
            "speedup_vs_" + main_form: round(q["tflops"] / r["tflops"], 3),
            "tile_block": q["tile_block"],
            "geometry": q["geometry"],
            "spin_cases": q["cases"],
            "projected_full_T_wall_s": round(q["total_flop"] /
                                             (q["tflops"] * 1e12), 1),
            "mfma_frac": round(q["dom_flop_per_launch"] * EXEC_FACTOR[form] /
                               (q["dom_launch_ms"] * 1e-3) / 1e12 /
                               PEAK[form], 4),
            "E_batches": q["e_batches"],
            "E_batches_" + main_form: r["e_batches"],
            "abs_dE": abs(q["e_batches"] - r["e_batches"]),
            "rel_dE": abs(q["e_batches"] - r["e_batches"]) /
                      abs(r["e_batches"]),
        }
    if args.fixtures:
        out["split_on_fixtures"] = split_on_fixtures(
            api, [f for f in forms if f in ("bf16x3", "f16x3")])
    if args.configs:
        out["configs"] = [run_config(api, *c, forms, args.seed)
                          for c in CONFIGS]
    if args.batch > 0:
        out["batch"] = {}
        for form in forms:
            out["batch"][form] = run_batch(api, form, args.batch, args.seed,
                                           args.tune, args.tune_sample)
            log(f"[bench] batch {form}: {out['batch'][form]}")
    if args.cpu != "none":
        cs = [int(x) for x in args.cpu.split(",")]
        out["cpu_baseline"] = cpu_baseline(api, cs[0], cs[1:], args.seed,
                                           (args.o, args.v), args.cpu_check)
        log(f"[bench] cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
