"""The CPU code built with the address and undefined-behaviour sanitizers
(-fsanitize=address,undefined, no recovery: the first report aborts):

* `make -C oracle sanitize` drives the oracle through every exported entry
  point (domain construction with several region sets, runs to stop time
  serial and with OpenMP threads, each phase, field and scalar access, two
  forced reference errors, every element function on random hexes); its
  s=8 run-to-stop-time results must still equal the golden fixture;
* `make -C tests/host sanitize` runs tests/host/plan_check.cpp over the
  library's host planning code (k_nodal tiles, the depth model, the XCD
  dispatch table, the revolve schedule) and its region builder, each
  against the property it must keep."""
import json
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def make(path):
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, path), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    lines = dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l)
    assert lines.get("failures") == "0", r.stdout[-3000:]
    return lines


def test_oracle_under_sanitizers():
    out = make("oracle")
    g = json.load(open(os.path.join(REPO, "tests", "golden", "oracle_runs.json")))["s8_to_stoptime"]
    for t in (1, 4):
        assert int(out[f"s8_q7_t{t}_cycle"]) == g["cycle"]
        assert float.fromhex(out[f"s8_q7_t{t}_time"]).hex() == g["time"]
        for k in range(4):
            assert float.fromhex(out[f"s8_q7_t{t}_verify{k}"]).hex() == g["verify"][k]
    # serial and threaded quirk-free runs agree
    for key in ("cycle", "time", "verify0", "verify3"):
        assert out[f"s8_q0_t1_{key}"] == out[f"s8_q0_t4_{key}"]


def test_host_planning_code_under_sanitizers():
    make("tests/host")
