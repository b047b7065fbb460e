"""Mutation check of the 3D oracle's parity pins: the oracle is rebuilt with
each planted mistake of oracle/orc_3d.cpp (ORC_MUTANT 1, 2, 3, 5, 6: MIN/MAX swap
in the per-tet combination, double cofactor sums, '>' for '>=' in the
unpredictable test, cbrtf for acbrt1, a tetrahedron outside the Kuhn split),
and every such
build must fail at least one pin of tests/_pins3d.py. Builds go to a temporary
directory; nothing else loads them."""
import os
import subprocess

import pytest

from _libs import Reference, have_reference, load_oracle
from _pins3d import PINS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC = os.path.join(ROOT, "oracle")
SRC = ["orc_huffman.cpp", "orc_2d.cpp", "orc_3d.cpp", "orc_cp.cpp", "orc_learn.cpp", "orc_capi.cpp"]
MUTANTS = {1: "MIN/MAX swap", 2: "double cofactor sums", 3: "'>' for '>='", 5: "cbrtf for acbrt1",
           6: "non-Kuhn tetrahedron"}
# ORC_MUTANT 4 (two vertices of a tetrahedron swapped) is an equivalent
# mutant: the bound is a symmetric function of the four vertices, and the
# double cofactor sums are exact unless the triple products span more than
# 2^29 and then round to the same float; test_vertex_order_is_not_observable
# shows it.

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference build absent (oracle/_ref)")


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("orc_mutants")
    procs = {}
    for m in list(MUTANTS) + [4]:
        out = str(d / f"liborc_m{m}.so")
        cmd = ["g++", "-O2", "-ffp-contract=off", "-std=gnu++17", "-fPIC", "-shared", f"-DORC_MUTANT={m}",
               "-o", out] + [os.path.join(ORC, s) for s in SRC]
        procs[m] = (out, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    libs = {}
    for m, (out, p) in procs.items():
        log = p.communicate()[0].decode()
        assert p.returncode == 0, log
        libs[m] = out
    return libs


@pytest.mark.parametrize("mutant", sorted(MUTANTS), ids=lambda m: MUTANTS[m])
def test_mutant_fails_a_pin(mutant_libs, mutant):
    orc, ref = load_oracle(mutant_libs[mutant]), Reference()
    failed = []
    for pin in PINS:
        try:
            pin(orc, ref)
        except AssertionError:
            failed.append(pin.__name__)
    print(f"mutant {mutant} ({MUTANTS[mutant]}) fails: {failed}")
    assert failed, f"no pin catches mutant {mutant} ({MUTANTS[mutant]})"


def test_vertex_order_is_not_observable(mutant_libs):
    """A permuted vertex order inside a tetrahedron (ORC_MUTANT 4) gives the
    same per-vertex bounds, bit for bit, on normal, small-integer and
    wide-range (1e-8..1e8) fields: the table's order within a tetrahedron is not
    something the pins could or need to fix."""
    import numpy as np
    from _pins3d import _fields
    a, b = load_oracle(), load_oracle(mutant_libs[4])
    for U, V, W in _fields(np.random.default_rng(15)):
        assert np.array_equal(a.eb_3d(U, V, W, 1.0).view(np.uint64), b.eb_3d(U, V, W, 1.0).view(np.uint64))
