"""The split vector launches (an interior launch without UPML code plus a
launch over the UPML shell, wipems_amd/csrc/fdtd_shell.h) cover every
vector the update bodies process exactly once, keep the interior launch off
slab cells, and index inside the field arrays; also with the shell in
vectors of a half or a quarter of the interior's width, and with the
interior's z range aligned to 4 or 8 vectors.  Checked on the host with the
engine's own planning code (tests/native/shell_partition.cpp) for the slab
sets the operator builds, for partial shells, for a z extent so short that
one vector touches both z slabs, and for slabs that form no shell (one
UPML launch over everything)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from wipems_amd.operator import BC_PEC, BC_PML, Operator

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("shell") / "shell_partition")
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "wipems_amd", "csrc"),
                    os.path.join(HERE, "native", "shell_partition.cpp"), "-o", out], check=True, capture_output=True)
    return out


def operator_slabs(N, bc, pml):
    op = Operator()
    op.set_mesh([np.arange(n, dtype=float) for n in N], 1e-3)
    op.set_boundary(bc, pml)
    op.set_gauss_excite(1e9, 1e9)
    d = op.build(0, nr_ts=2)
    return [(tuple(int(v) for v in u.start), tuple(int(v) for v in u.num_lines)) for u in d.upml]


def run(harness, cases):
    text = ""
    for N, slabs in cases:
        text += "%d %d %d %d\n" % (N[0], N[1], N[2], len(slabs))
        for s, n in slabs:
            text += "%d %d %d %d %d %d\n" % (s + n)
    r = subprocess.run([harness], input=text, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout.splitlines()


def test_operator_shells(harness):
    cases = []
    for N, bc, pml in [((41, 37, 33), [BC_PML] * 6, [4] * 6),
                       ((30, 26, 50), [BC_PML] * 6, [8] * 6),
                       ((20, 21, 20), [BC_PML] * 6, [8] * 6),          # z slabs 0-8 and 11-19 share vectors
                       ((25, 19, 47), [BC_PEC] * 4 + [BC_PML] * 2, [0] * 4 + [6, 6]),
                       ((33, 18, 29), [BC_PML, BC_PEC, BC_PEC, BC_PML, BC_PML, BC_PEC], [5, 0, 0, 3, 7, 0]),
                       ((12, 40, 17), [BC_PML] * 6, [3, 3, 8, 8, 2, 2])]:
        slabs = operator_slabs(N, bc, pml)
        assert slabs
        cases.append((N, slabs))
    lines = run(harness, cases)
    assert len(lines) == 48 * len(cases)
    assert all(" shell 1 " in l for l in lines)


def test_non_shell_slabs_use_one_launch(harness):
    # a slab in the middle of the domain forms no shell
    lines = run(harness, [((20, 20, 24), [((5, 5, 5), (4, 6, 7))]),
                          ((16, 16, 16), [((0, 0, 0), (4, 16, 16)), ((0, 0, 0), (1, 1, 1))][:1])])
    assert " shell 0 " in lines[0]
    assert " shell 1 " in lines[-1]


def test_no_upml_single_launch(harness):
    # no slabs: one launch over everything, XCD-contiguous block order
    lines = run(harness, [((37, 29, 45), []), ((9, 300, 17), [])])
    assert len(lines) == 96 and all(" shell 0 " in l for l in lines)
