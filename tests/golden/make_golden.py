"""Writes tests/golden/oracle_runs.json: results of the CPU oracle on the
parity configurations (verify numbers, cycle count, final time and SHA-256
of the final energy and x fields).  Run from the repo root after
`make -C oracle`:  python tests/golden/make_golden.py

With --large it writes tests/golden/oracle_runs_large.json instead: the
full-size workloads (the s=192 headline's first 500 cycles, s=96 runs) and
the s=30 reference run into its stall, SHA-256 of all eleven state fields;
minutes of CPU time on 8 threads, so only the GPU tests read them back.
Case names after --large regenerate only those cases and keep the others
already in the file:  python tests/golden/make_golden.py --large s30_reference_1700"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

CASES = [  # (name, nx, quirks, num_iters)
    ("s8_to_stoptime", 8, 7, 10 ** 9),
    ("s12_to_stoptime", 12, 7, 10 ** 9),
    ("s16_to_stoptime", 16, 7, 10 ** 9),
    ("s8_smoke_10", 8, 7, 10),
    ("s30_first_200", 30, 7, 200),
    ("s30_quirk_free_to_stoptime", 30, 0, 10 ** 9),
]


LARGE_CASES = [
    # examples/benchmark.jl's loop at s=30 with the reference arithmetic: from
    # cycle ~1350 dt decays geometrically and time stalls at 9.6672e-3, below
    # stoptime, so the reference's `while time < stoptime` never ends there;
    # the case pins the state 1700 cycles in, deep in the stall
    ("s30_reference_1700", 30, 7, 1700),
    ("s192_first_500", 192, 7, 500),
    ("s96_first_2000", 96, 7, 2000),
    ("s96_quirk_free_to_stoptime", 96, 0, 10 ** 9),
]
STATE = ("e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd")


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def run_case(nx, quirks, num_iters, nthreads=8):
    od = O.OracleDomain(nx, quirks, nthreads)
    rc, _ = od.run(num_iters)
    v = od.verify()
    return {"rc": rc, "cycle": od.get("cycle"), "time": od.get("time").hex(),
            "deltatime": od.get("deltatime").hex(),
            "verify": [float(x).hex() for x in v],
            "verify_printed": ["%.6e" % x for x in v],
            "sha256_e": digest(od.field("e")), "sha256_x": digest(od.field("x"))}


def run_large(nx, quirks, num_iters, nthreads):
    od = O.OracleDomain(nx, quirks, nthreads)
    rc, _ = od.run(num_iters)
    v = od.verify()
    r = {"rc": rc, "cycle": od.get("cycle"), "time": od.get("time").hex(),
         "deltatime": od.get("deltatime").hex(),
         "dtcourant": od.get("dtcourant").hex(), "dthydro": od.get("dthydro").hex(),
         "verify": [float(x).hex() for x in v],
         "verify_printed": ["%.6e" % x for x in v]}
    for f in STATE:
        r["sha256_" + f] = digest(od.field(f))
    return r


def main():
    large = "--large" in sys.argv
    only = [a for a in sys.argv[1:] if not a.startswith("--")]
    fn = "oracle_runs_large.json" if large else "oracle_runs.json"
    out = {}
    if only and os.path.exists(os.path.join(HERE, fn)):
        with open(os.path.join(HERE, fn)) as f:
            out = json.load(f)
    for name, nx, q, it in (LARGE_CASES if large else CASES):
        if only and name not in only:
            continue
        r = run_large(nx, q, it, os.cpu_count() or 8) if large else run_case(nx, q, it)
        r.update({"nx": nx, "quirks": q, "num_iters": it})
        out[name] = r
        print(name, r["cycle"], r["verify_printed"], flush=True)
    with open(os.path.join(HERE, fn), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
