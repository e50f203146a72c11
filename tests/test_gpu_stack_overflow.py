"""GPU: the traversal stack's private overflow path against the oracle.

The path kernels keep the first 20 traversal-stack entries of each lane in
LDS and deeper ones in a per-lane private array.  A Karras tree over the
benchmark scenes never needs more than 18 entries in the binary walk, so the
normal build rarely reaches the overflow; the library built with a 2-entry LDS stack
(lisa_amd/_build/stack2, made by __graft_entry__.build) sends nearly every
push there.  tests/overflow_check.py runs against that library in a child
process (one library per process) and checks hits and per-sample path
records bit for bit against the oracle, and the 4-wide walk's hits and a
render.
"""
import os
import subprocess
import sys

import pytest

from lisa_amd import build as lb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_overflow_stack_matches_oracle():
    lib = os.path.join(lb.SMALL_STACK, "liblisa.so")
    assert os.path.exists(lib), "build the small-stack library first (__graft_entry__.build)"
    env = dict(os.environ, LISA_LIBRARY=lib)
    res = subprocess.run([sys.executable, "-u", os.path.join(HERE, "overflow_check.py"), lib],
                         env=env, capture_output=True, text=True, timeout=600)
    print(res.stdout, res.stderr[-3000:], flush=True)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK")
