"""Energies of libpt compiled from its own sources on the synthetic
workloads, written to tests/golden/ref_energies.json.

The reference is oracle/_ref/libpt_ref.so (``make -C oracle ref``: the
reference's src/pt.c, src/ptmp.c and src/misc.c compiled unmodified, MKL in
sequential mode).  Inputs come from the engine's host generator
(libpt_amd_synth, bitwise the device generator), so a GPU test regenerates
the same arrays in HBM from (call, dims, seed, omega, sig_f) alone.

Every entry runs in a process of its own with OMP_NUM_THREADS fixed and
OMP_PROC_BIND=spread; the reference sums each thread's energies without
compensation, so its result depends on the thread count, which the file
records (one entry also at 1 thread, to show that spread).

    python tests/golden/make_ref_goldens.py            # every entry
    python tests/golden/make_ref_goldens.py NAME ...   # some entries

Entries already in the file are kept unless named.
"""
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "ref_energies.json")
THREADS = 8
SEED = 2024

# name: (call, dims, generator args, _mp, threads)
ENTRIES = {
    "rpt_8_64": ("rpt", (8, 64), dict(sig_f=1e-2), False, THREADS),
    "rpt_8_64_t1": ("rpt", (8, 64), dict(sig_f=1e-2), False, 1),
    "rpt_8_64_mp": ("rpt", (8, 64), dict(sig_f=1e-2), True, THREADS),
    "uft_16_160_14_160": ("uft", (16, 160, 14, 160), dict(omega=0.3), False,
                          THREADS),
    "uft_16_160_14_160_mp": ("uft", (16, 160, 14, 160), dict(omega=0.3),
                             True, THREADS),
    "rft_30_100": ("rft", (30, 100), dict(omega=0.3), False, THREADS),
    "rpt_45_140": ("rpt", (45, 140), dict(), False, THREADS),
    "rpt_45_140_mp": ("rpt", (45, 140), dict(), True, THREADS),
    "upt_24_200_20_204": ("upt", (24, 200, 20, 204), dict(), False, THREADS),
    # the 64 restricted systems of the batch workload (o in [4, 20], v in
    # [40, 160]; sizes from api.batch_dims(64, seed=2024), system k from
    # seed 2024 + 1 + k)
    "batch64": ("batch", 64, dict(), False, THREADS),
}


def one(name):
    """Run one entry in this process and print its JSON record."""
    sys.path.insert(0, ROOT)
    import numpy as np
    import oracle
    from libpt_amd import api
    call, dims, kw, mp, threads = ENTRIES[name]
    if call == "batch":
        sizes = api.batch_dims(dims, seed=SEED)
        es, wall = [], 0.0
        for k, d in enumerate(sizes):
            arr = api.synth_host("rpt", api.make_spec("rpt", d,
                                                      seed=SEED + 1 + k))
            t0 = time.perf_counter()
            es.append(oracle.ref_energy("rpt", d, arr))
            wall += time.perf_counter() - t0
            print(f"[goldens] {name} system {k} {d} {es[-1]!r}",
                  file=sys.stderr, flush=True)
        rec = {"call": "rpt", "systems": dims, "seed": SEED,
               "system_seed": "seed + 1 + k", "dims": [list(d) for d in sizes],
               "E_ref": es, "wall_s": round(wall, 1)}
    else:
        spec = api.make_spec(call, dims, seed=SEED, **kw)
        arr = api.synth_host(call, spec, fp32=mp)
        t0 = time.perf_counter()
        e = oracle.ref_energy(call, dims, arr, mp=mp)
        wall = time.perf_counter() - t0
        rec = {"call": call, "dims": list(dims), "seed": SEED,
               "omega": kw.get("omega", 0.0), "sig_f": kw.get("sig_f", 0.0),
               "mp": mp, "E_ref": e, "wall_s": round(wall, 1),
               "flop": api.flop(call, dims)}
    rec["threads"] = threads
    rec["blas"] = "MKL (MKL_THREADING_LAYER=SEQUENTIAL)"
    print(json.dumps(rec), flush=True)


def main(names):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as fh:
            data = json.load(fh)
    data.setdefault("about", (
        "libpt compiled from its sources (oracle/_ref, MKL sequential, "
        "OpenMP threads as recorded) on inputs of the engine's synthetic "
        "generator; written by tests/golden/make_ref_goldens.py"))
    ents = data.setdefault("entries", {})
    for name in names or list(ENTRIES):
        if not names and name in ents:
            continue
        threads = ENTRIES[name][4]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads),
                   OMP_PROC_BIND="spread", MKL_THREADING_LAYER="SEQUENTIAL")
        t0 = time.perf_counter()
        r = subprocess.run([sys.executable, __file__, "--one", name], env=env,
                           stdout=subprocess.PIPE, text=True, check=True)
        ents[name] = json.loads(r.stdout.strip().splitlines()[-1])
        with open(OUT + ".tmp", "w") as fh:
            json.dump(data, fh, indent=1)
        os.replace(OUT + ".tmp", OUT)
        print(f"[goldens] {name}: {time.perf_counter() - t0:.0f} s",
              flush=True)


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--one":
        one(sys.argv[2])
    else:
        main(sys.argv[1:])
