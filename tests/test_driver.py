"""The C benchmark driver (lulesh.jl_amd/lib/lulesh_benchmark, built from
lulesh.jl_amd/csrc/lulesh_benchmark.c): the reference's examples/benchmark.jl
loop and report over the C ABI alone.  CPU tests cover its command line
(the paths that end before any device call); GPU tests compare its printed
report with the oracle's run of the same problem."""
import os
import re
import subprocess

import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "lulesh.jl_amd", "lib", "lulesh_benchmark")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built (run build())")


def drive(*args, timeout=120):
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_help_and_usage():
    r = drive("--help")
    assert r.returncode == 0 and "usage:" in r.stderr
    r = drive("-s", "8", "10", "12")
    assert r.returncode == 2 and "usage:" in r.stderr
    r = drive("-s", "-3")
    assert r.returncode == 2


def test_unstructured_is_the_reference_error():
    """Without -s the reference's Domain(prob) raises (src/domain.jl:612)."""
    r = drive("8", "10")
    assert r.returncode == 1
    assert "Reading unstructured mesh is currently missing in the Julia version of LULESH." in r.stderr
    assert r.stdout == ""


def test_mpi_refused_and_size_checked():
    r = drive("--mpi", "-s", "8")
    assert r.returncode == 2 and "one rank" in r.stderr
    r = drive("-s", "801")
    assert r.returncode == 2 and "1..800" in r.stderr


def parse_report(text):
    """The numbers of verifyAndWriteFinalOutput's report (src/utils.jl:131-163)."""
    pat = {"size": r"Problem size\s+=\s+(\d+)", "tasks": r"MPI tasks\s+=\s+(\d+)",
           "cycle": r"Iteration count\s+=\s+(\d+)", "e1": r"Final Origin Energy =\s+(\S+)",
           "maxabs": r"MaxAbsDiff\s+=\s+(\S+)", "totabs": r"TotalAbsDiff = (\S+)",
           "maxrel": r"MaxRelDiff\s+=\s+(\S+)"}
    out = {}
    for k, p in pat.items():
        m = re.search(p, text)
        assert m, (k, text)
        out[k] = m.group(1)
    assert re.search(r"Grind time \(us/z/c\)\s+=\s+\S+\s+\(per dom\)", text)
    assert re.search(r"FOM\s+=\s+\S+ \(z/s\)", text)
    return out


def oracle_report(nx, iters):
    od = O.OracleDomain(nx, O.QUIRKS_REFERENCE, 4)
    rc, _ = od.run(iters)
    assert rc == 0
    v = od.verify()
    out = {"size": str(nx), "tasks": "1", "cycle": str(int(od.get("cycle"))), "e1": "%.6e" % v[0],
           "maxabs": "%.6e" % v[1], "totabs": "%.6e" % v[2], "maxrel": "%.6e" % v[3]}
    od.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("nx,iters", [(8, 10 ** 9), (12, 60), (30, 25)])
def test_driver_report_matches_oracle(nx, iters):
    print(f"[driver] -s {nx} {iters}", flush=True)
    r = drive("-s", nx, iters)
    assert r.returncode == 0, r.stderr
    assert parse_report(r.stdout) == oracle_report(nx, iters)


@pytest.mark.gpu
def test_driver_per_cycle_and_enzyme_same_report():
    """Host-stepped loop (one "Completed" line per cycle, as the reference
    logs) and the Enzyme-call loop (lulesh_vjp_cycle advances the state
    exactly as lagrangeLeapFrog) print the device loop's numbers."""
    print("[driver] per-cycle / enzyme", flush=True)
    base = drive("-s", 12, 40)
    assert base.returncode == 0, base.stderr
    want = parse_report(base.stdout)
    step = drive("-s", "--per-cycle", 12, 40)
    assert step.returncode == 0, step.stderr
    assert parse_report(step.stdout) == want
    assert len(re.findall(r"Completed cycle=", step.stderr)) == 40
    enz = drive("--enzyme", "-s", 12, 40)
    assert enz.returncode == 0, enz.stderr
    assert parse_report(enz.stdout) == want
    assert len(re.findall(r"Completed cycle=", enz.stderr)) == 40


def test_report_numbers_written_as_julia_does():
    """@printf writes non-finite numbers as Inf, -Inf, NaN (the reference's
    report shows MaxRelDiff = Inf once an element's mirror still has e = 0)."""
    from lulesh_amd.domain import _julia_num
    assert _julia_num("%.6e", float("inf")) == "Inf"
    assert _julia_num("%.6e", float("-inf")) == "-Inf"
    assert _julia_num("%.8f", float("nan")) == "NaN"
    assert _julia_num("%.6e", 202507.5) == "2.025075e+05"
    assert _julia_num("%.8f", 0.5) == "0.50000000"


@pytest.mark.gpu
@pytest.mark.parametrize("per_cycle", [False, True])
def test_python_main_report_equals_c_driver(per_cycle):
    """lulesh_amd.main (the Python mirror of the same driver) prints the C
    driver's report, timing lines aside."""
    import io

    from lulesh_amd import main
    print(f"[driver] python main per_cycle={per_cycle}", flush=True)
    buf = io.StringIO()
    rep = main(12, True, 60, per_cycle=per_cycle, file=buf)
    assert rep["cycle"] == 60
    c = drive("-s", 12, 60)
    assert c.returncode == 0, c.stderr
    assert parse_report(buf.getvalue()) == parse_report(c.stdout)
