"""Golden fixtures (tests/golden/oracle_runs.json, written by
tests/golden/make_golden.py from the oracle): the oracle must keep
reproducing them, and the GPU path must match them bit for bit.  The
full-size ones (oracle_runs_large.json: the s=192 headline workload's first
500 cycles and s=96 runs, minutes of oracle time each) are checked on the
GPU only, every state field by SHA-256."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "oracle_runs.json")))
_LARGE = os.path.join(HERE, "golden", "oracle_runs_large.json")
GOLD_LARGE = json.load(open(_LARGE)) if os.path.exists(_LARGE) else {}
STATE = ("e", "p", "q", "v", "ss", "x", "y", "z", "xd", "yd", "zd")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_reproduces_golden(name):
    import oracle as O
    g = GOLD[name]
    od = O.OracleDomain(g["nx"], g["quirks"], 4)
    rc, _ = od.run(g["num_iters"])
    assert rc == g["rc"] and od.get("cycle") == g["cycle"]
    assert od.get("time").hex() == g["time"]
    assert [float(x).hex() for x in od.verify()] == g["verify"]
    assert sha(od.field("e")) == g["sha256_e"] and sha(od.field("x")) == g["sha256_x"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLD))
def test_gpu_matches_golden(name):
    from lulesh_amd import Domain, LuleshProblem, run, verify
    g = GOLD[name]
    print(f"[golden] {name}", flush=True)
    dom = Domain(LuleshProblem(g["num_iters"], True, g["nx"], 11, 1, 1), quirks=g["quirks"])
    run(dom, g["num_iters"])
    assert dom.cycle == g["cycle"]
    assert dom.time.hex() == g["time"] and dom.deltatime.hex() == g["deltatime"]
    assert [float(x).hex() for x in verify(dom)] == g["verify"]
    assert sha(dom.e) == g["sha256_e"] and sha(dom.x) == g["sha256_x"]
    dom.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLD_LARGE))
def test_gpu_matches_golden_full_size(name):
    from lulesh_amd import Domain, LuleshProblem, run, verify
    g = GOLD_LARGE[name]
    print(f"[golden-large] {name}", flush=True)
    dom = Domain(LuleshProblem(g["num_iters"], True, g["nx"], 11, 1, 1), quirks=g["quirks"])
    run(dom, g["num_iters"])
    assert dom.cycle == g["cycle"]
    assert dom.time.hex() == g["time"] and dom.deltatime.hex() == g["deltatime"]
    assert dom.dtcourant.hex() == g["dtcourant"] and dom.dthydro.hex() == g["dthydro"]
    assert [float(x).hex() for x in verify(dom)] == g["verify"]
    for f in STATE:
        assert sha(dom.get(f)) == g["sha256_" + f], f
    dom.close()
