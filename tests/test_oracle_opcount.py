"""Floating-point operations per cell and substep of the reference formulation (oracle/opcount.cc:
the oracle's right-hand side and RK3 update run on a scalar type that records its arithmetic).

Pinned by closed forms: a 6th-order first derivative is 3 differences, 3 products by the
coefficients, 2 sums and a product by 1/ds (5 additions, 4 multiplications); a second derivative
adds a centre product and the 1/ds^2 scaling (6, 6); a cross derivative takes 3 additions per
offset pair, 3 products, 2 sums and 2 scalings (11, 5); the substep-0 update of a field is
cur + beta dt F (1 addition, 2 multiplications), the later ones
cur + beta (a (cur - prev) + dt F) (3, 3). The right-hand side's live count must lie below its
full count by at least the 12 cross derivatives nothing reads (12 x 16 operations). bench.py's
FLOP constants are these counts.
"""
import os
import shutil
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def counts(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = str(tmp_path_factory.mktemp("opcount") / "opcount")
    cmd = [cxx, "-O1", "-std=c++17", "-Wall", "-Werror", "-Wno-unknown-pragmas", "-ffp-contract=off",
           "-I", os.path.join(REPO, "include"), "-I", os.path.join(REPO, "oracle"),
           "-DAC_DOUBLE_PRECISION=1", os.path.join(REPO, "oracle", "opcount.cc"), "-o", exe]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.splitlines():
        what, kind, *v = line.split()
        res[(what, kind)] = dict(zip(("add", "mul", "div", "exp", "neg"), map(int, v)))
    return res


def test_counter_closed_forms(counts):
    for kind in ("all", "live"):
        assert (counts[("first", kind)]["add"], counts[("first", kind)]["mul"]) == (5, 4)
        assert (counts[("second", kind)]["add"], counts[("second", kind)]["mul"]) == (6, 6)
        assert (counts[("cross", kind)]["add"], counts[("cross", kind)]["mul"]) == (11, 5)
        assert (counts[("update0", kind)]["add"], counts[("update0", kind)]["mul"]) == (8, 16)
        assert (counts[("update12", kind)]["add"], counts[("update12", kind)]["mul"]) == (24, 24)


def test_rhs_counts_and_bench_constants(counts):
    full, live = counts[("rhs", "all")], counts[("rhs", "live")]
    cross = 11 + 5
    assert (full["add"] + full["mul"]) - (live["add"] + live["mul"]) >= 12 * cross
    assert live["exp"] == full["exp"] and live["div"] == full["div"]
    sys.path.insert(0, REPO)
    import bench
    assert bench.FLOP_RHS == live["add"] + live["mul"] + live["div"]
    assert bench.EXP_RHS == live["exp"]
    assert bench.FLOP_UPDATE == (8 + 16, 24 + 24, 24 + 24)
