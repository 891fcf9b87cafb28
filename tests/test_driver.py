"""The stdin driver (tests/driver/faces_main.cpp): the reference's input
contract `lx ly lz mx my mz n nk nj ni` (README.md:61, gpu/main.cc:136-161)
and its output lines (gpu/main.cc:104, 138-142, 194), with the oracle as the
correctness check."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "driver", "_build", "faces_main")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER),
                                  reason="driver not built (make -C tests/driver)")


def run(stdin, *args, timeout=120):
    print("case driver", repr(stdin), args, flush=True)
    p = subprocess.run([DRIVER, *args], input=stdin, capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout, p.stderr


@needs_driver
@pytest.mark.parametrize("stdin,msg", [("1 1 1 4 4", "expected 10 integers"),
                                       ("0 1 1 4 4 4 5 1 1 1", "invalid sizes"),
                                       ("1 1 1 4 4 4 1 1 1 1", "invalid sizes")])
def test_driver_rejects_bad_input(stdin, msg):
    rc, out, err = run(stdin)
    assert rc == 2 and msg in err


@needs_driver
def test_driver_rejects_bad_options():
    rc, _, err = run("1 1 1 4 4 4 5 1 1 1", "--boundary", "torus")
    assert rc == 2 and "unknown boundary" in err
    rc, _, err = run("2 1 1 4 4 4 5 1 1 1", "--rank", "2")
    assert rc == 2 and "outside the rank grid" in err


def check_output(out, lgrid, m, n, nk, nj, ni, npass):
    lines = out.splitlines()
    assert lines[0] == "%d %d %d tasks" % lgrid
    assert lines[1] == "%d %d %d local elements of size %d" % (*m, n)
    tail = ("%d shares without compute kernels" % -ni) if ni < 0 else ("%d shares" % ni)
    assert lines[2] == "%d face inits x %d element inits x %s" % (nk, nj, tail)
    prefix = "".join("%d: " % (k + 1) for k in range(nk))
    assert lines[3] == prefix + "%d tasks passed correctness check" % npass, lines[3]
    assert re.fullmatch(r"time \S+ avg \S+ min \S+ max", lines[4]), lines[4]


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("stdin,args,npass", [
    ("1 1 1 4 4 4 5 1 1 1", (), 1),                                  # 4^3 n=5 periodic, integer
    ("1 1 1 4 4 4 5 2 2 3", (), 1),
    ("1 1 1 3 2 2 6 1 2 -2", (), 1),                                 # compute = 0
    ("1 1 1 3 3 3 8 1 1 2", ("--init", "rand"), 1),
    ("1 1 1 4 3 2 5 1 1 2", ("--reference",), 1),                    # reference mode
    ("1 1 1 4 3 2 5 1 1 -2", ("--reference",), 1),
    ("2 2 1 2 2 2 4 1 1 2", ("--boundary", "open", "--all-ranks"), 4),
    ("3 3 3 2 2 1 3 1 1 1", ("--boundary", "open", "--rank", "13"), 1),  # interior rank
    ("2 2 2 2 2 2 4 1 1 1", ("--all-ranks", "--init", "rand"), 8),
])
def test_driver_passes(stdin, args, npass):
    rc, out, err = run(stdin, *args)
    assert rc == 0, (out, err)
    v = [int(x) for x in stdin.split()]
    check_output(out, tuple(v[:3]), tuple(v[3:6]), v[6], v[7], v[8], v[9], npass)
    assert "bit-exact" in err
    m = re.search(r"(\d+) of (\d+) ranks bit-exact", err)
    assert m and m.group(1) == m.group(2), err


@pytest.mark.gpu
@needs_driver
def test_reference_run_line_a():
    """The reference's run line A (gpu/frontier-run-all-1.sh:15) on one rank in
    the reference's own mode: open boundary, positional init, 11,000 shares."""
    rc, out, err = run("1 1 1 15 14 13 12 10 10 100", "--reference", timeout=300)
    assert rc == 0, (out, err)
    check_output(out, (1, 1, 1), (15, 14, 13), 12, 10, 10, 100, 1)
    assert "1 of 1 ranks bit-exact" in err


@pytest.mark.gpu
@pytest.mark.slow
@needs_driver
def test_reference_run_line_b():
    """The reference's run line B (gpu/frontier-run-all-1.sh:17: 95x94x93
    elements, n=12, an 11.5 GB field, 3 plans x 2 inits x 10 shares) on one
    rank in the reference's own mode."""
    rc, out, err = run("1 1 1 95 94 93 12 3 1 10", "--reference", timeout=600)
    assert rc == 0, (out, err)
    check_output(out, (1, 1, 1), (95, 94, 93), 12, 3, 1, 10, 1)
    assert "1 of 1 ranks bit-exact" in err
