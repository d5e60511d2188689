"""Parity of the HIP engine against the CPU oracle (same operator in).

FP32 handles must reproduce the oracle's float fields bit for bit (the
reference engines agree bitwise, TESTSUITE/enginetests/cavity.m:155); FP64
handles must agree with the double oracle bit for bit as well.  Voltage probes
are double sums in the same order (exact); current probes accumulate in FP64
on the device and in float in the reference (Common/processcurrent.cpp:95),
so they are compared at rel-L2 <= 1e-6.
"""
import numpy as np
import pytest

from wipems_amd import scenarios
from wipems_amd.engine import engine_from_operator
from oracle.oracle import oracle_from_operator

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def compare(eng, orc, exact=True):
    for which in (0, 1):
        g = eng.get_field(which)
        o = orc.get_field(which)
        if exact:
            bad = np.argwhere(g != o)
            assert bad.size == 0, "field %d differs at %d places, first %s: gpu %r oracle %r" % (
                which, len(bad), tuple(bad[0]), g[tuple(bad[0])], o[tuple(bad[0])])
        else:
            assert rel_l2(g, o) <= 1e-5
    for pid, p in enumerate(eng.probe_ids):
        tg, vg = eng.read_probe(p)
        to, vo = orc.read_probe(pid)
        assert len(tg) == len(to)
        np.testing.assert_array_equal(tg, to)
        kind = eng_probe_kind(eng, pid)
        if kind == 0 and exact:
            np.testing.assert_array_equal(vg, vo)
        else:
            assert rel_l2(vg, vo) <= 1e-6


_kinds = {}


def eng_probe_kind(eng, pid):
    return _kinds[id(eng)][pid]


def run_pair(sc, steps, precision="fp32", compressed=True, chunks=(1,)):
    d = sc.data
    eng = engine_from_operator(d, precision=precision, compressed=compressed)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    done = 0
    for c in chunks:
        eng.iterate(c)
        orc.iterate(c)
        done += c
    eng.iterate(steps - done)
    orc.iterate(steps - done)
    eng.synchronize()
    return eng, orc


@pytest.mark.parametrize("compressed", [True, False])
def test_cavity_fp32_bitwise(gpu, compressed):
    sc = scenarios.cavity_point_source(24, 300, prec=0)
    eng, orc = run_pair(sc, 300, compressed=compressed, chunks=(1, 2, 17))
    compare(eng, orc)
    assert abs(eng.energy() - orc.energy()) <= 1e-12 * abs(orc.energy())


def test_cavity_fp64_bitwise(gpu):
    sc = scenarios.cavity_point_source(20, 200, prec=1)
    eng, orc = run_pair(sc, 200, precision="fp64", compressed=False)
    compare(eng, orc)


def test_graded_upml_port_fp32_bitwise(gpu):
    sc = scenarios.graded_port_scenario(n_cells=40, steps=400, prec=0, n_boxes=6, pml=8)
    assert len(sc.data.upml) == 6
    eng, orc = run_pair(sc, 400, compressed=False, chunks=(3, 16))
    compare(eng, orc)


def test_graded_upml_port_fp32_compressed(gpu):
    sc = scenarios.graded_port_scenario(n_cells=32, steps=200, prec=0, n_boxes=4, pml=6)
    eng, orc = run_pair(sc, 200, compressed=True)
    assert eng.compressed or sc.data.index is None
    compare(eng, orc)


@pytest.mark.parametrize("split", [0, 1])
def test_kernel_shapes_bitwise(gpu, split):
    """One UPML launch and the interior/shell split give the oracle's bits
    (graded mesh, UPML, port)."""
    sc = scenarios.graded_port_scenario(n_cells=36, steps=150, prec=0, n_boxes=5, pml=8)
    d = sc.data
    eng = engine_from_operator(d, precision="fp32", compressed=False)
    eng.set_pml_split(split)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(150)
    orc.iterate(150)
    compare(eng, orc)


def test_cavity64_matches_golden_fixture(gpu):
    """Full-size 64^3-cell / 2000-step case on the GPU against the committed
    oracle fixture (tests/golden/cavity64.npz)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cavity64.npz"), allow_pickle=False)
    sc = scenarios.cavity_point_source(64, 2000, prec=0)
    eng = engine_from_operator(sc.data, precision="fp32", compressed=True)
    eng.iterate(2000)
    tu, u = eng.read_probe(0)
    ti, i = eng.read_probe(1)
    np.testing.assert_array_equal(tu, g["t_u"])
    np.testing.assert_array_equal(u, g["u"])
    assert rel_l2(i, g["i"]) <= 1e-6
    assert eng.energy() == pytest.approx(float(g["energy"]), rel=1e-12)


@pytest.mark.parametrize("n_cells", [21, 30, 33])
def test_vector_kernel_odd_sizes(gpu, n_cells):
    """Vector kernels with Nz not a multiple of the vector width, compressed
    and full coefficients, against the oracle."""
    sc = scenarios.graded_port_scenario(n_cells=n_cells, steps=120, prec=0, n_boxes=4, pml=5)
    d = sc.data
    for compressed in (False, True):
        eng = engine_from_operator(d, precision="fp32", compressed=compressed)
        _kinds[id(eng)] = [p.kind for p in d.probes]
        orc = oracle_from_operator(d, ftz=True)
        eng.iterate(120)
        orc.iterate(120)
        compare(eng, orc)


def test_vector_kernel_fp64(gpu):
    sc = scenarios.graded_port_scenario(n_cells=27, steps=100, prec=1, n_boxes=4, pml=5)
    eng = engine_from_operator(sc.data, precision="fp64", compressed=False)
    _kinds[id(eng)] = [p.kind for p in sc.data.probes]
    orc = oracle_from_operator(sc.data)
    eng.iterate(100)
    orc.iterate(100)
    compare(eng, orc)


@pytest.mark.parametrize("prec,precision,kw", [
    (0, "fp32", dict(n_cells=30, margin=4)),
    (0, "fp32", dict(n_cells=28, margin=3, prop_dir=(1.0, 1.0, 1.0), e_vec=(1.0, -1.0, 0.0), graded=True, phase_freq=5e9)),
    (1, "fp64", dict(n_cells=26, margin=3, prop_dir=(0.0, -1.0, 0.0), e_vec=(1.0, 0.0, 0.0))),
])
def test_plane_wave_tfsf_bitwise(gpu, prec, precision, kw):
    """TFSF injection (repeated edge locations at the box edges applied in
    list order) against the oracle, bit for bit."""
    sc = scenarios.plane_wave_box(steps=260, prec=prec, pml=6, **kw)
    print("case: plane wave %s %s" % (precision, kw), flush=True)
    eng, orc = run_pair(sc, 260, precision=precision, compressed=None, chunks=(5, 31))
    compare(eng, orc)


@pytest.mark.parametrize("prec,precision", [(0, "fp32"), (1, "fp64")])
def test_dispersive_ade_tfsf_bitwise(gpu, prec, precision):
    """Drude slab + Lorentz spheres (ADE pre/apply hooks) lit by a TFSF plane
    wave, UPML on all faces, against the oracle bit for bit."""
    sc = scenarios.dispersive_scenario(num_lines=(41, 37, 45), steps=300, prec=prec, n_spheres=5,
                                       material_scale=1e-3, ts_factor=0.9)
    assert sc.data.ade and sc.data.ade[0].pos.shape[1] > 0
    print("case: dispersive 41x37x45 %s" % precision, flush=True)
    eng, orc = run_pair(sc, 300, precision=precision, compressed=None, chunks=(1, 16, 33))
    compare(eng, orc)


@pytest.mark.parametrize("mode", [0, 1])
def test_dispersive_ade_coefficient_forms(gpu, mode):
    """ADE coefficients from per-entry lists (mode 0) and from tables of
    distinct tuples (mode 1) give the oracle's bits, fused hooks and one
    launch per hook."""
    sc = scenarios.dispersive_scenario(num_lines=(41, 37, 45), steps=240, prec=0, n_spheres=5,
                                       material_scale=1e-3, ts_factor=0.9)
    d = sc.data
    print("case: dispersive 41x37x45 ADE coefficient mode %d" % mode, flush=True)
    for fuse in (1, 0):
        eng = engine_from_operator(d, precision="fp32", compressed=None, ade_compression=mode)
        eng.set_fuse_hooks(fuse)
        for o in range(len(d.ade)):
            v, c = eng.ade_compressed(o)
            assert v == (mode == 1 and bool(d.ade[o].flags & 1)) and c == (mode == 1 and bool(d.ade[o].flags & 2))
        _kinds[id(eng)] = [p.kind for p in d.probes]
        orc = oracle_from_operator(d, ftz=True)
        eng.iterate(240)
        orc.iterate(240)
        compare(eng, orc)
        eng.close()


def test_fused_hooks_early_ade_pre_update(gpu):
    """The fused hooks run the ADE pre-update of an unshared component right
    after its apply; shared components (TFSF locations) keep theirs in the
    other field's hook launch, and a launch sequence's first step redoes all
    of them.  Electric and magnetic poles (the magnetic lists copied from the
    electric ones), TFSF, uneven chunks that cross graph and eager steps,
    and fields set between calls: the oracle's bits, fused and unfused."""
    import dataclasses
    sc = scenarios.dispersive_scenario(num_lines=(41, 37, 45), steps=200, prec=0, n_spheres=5, material_scale=1e-3, ts_factor=0.9)
    d = sc.data
    a = d.ade[0]
    lists = list(a.lists)
    lists[3], lists[4], lists[5] = (np.ascontiguousarray(lists[k] * np.float32(0.5)) if k != 1 else np.ascontiguousarray(lists[k])
                                    for k in (0, 1, 2))
    d = dataclasses.replace(d, ade=[dataclasses.replace(a, lists=lists, flags=a.flags | 2 | (8 if a.flags & 4 else 0))])
    print("case: dispersive 41x37x45, electric + magnetic ADE", flush=True)
    for fuse in (1, 0):
        eng = engine_from_operator(d, precision="fp32", compressed=None)
        eng.set_fuse_hooks(fuse)
        _kinds[id(eng)] = [p.kind for p in d.probes]
        orc = oracle_from_operator(d, ftz=True)
        for c in (1, 19, 2, 33, 17):
            eng.iterate(c)
            orc.iterate(c)
        V = orc.get_field(0) * np.float32(0.75)
        eng.set_field(0, V)
        orc.set_field(0, V)
        for c in (16, 5):
            eng.iterate(c)
            orc.iterate(c)
        compare(eng, orc)
        eng.close()


def test_lorentz_filled_cavity_bitwise(gpu):
    """Undamped Lorentz pole filling a PEC cavity (the ADE dispersion pin)
    on the GPU, 3000 steps, against the oracle."""
    from tests.test_ade import filled_cavity
    d = filled_cavity(2e9, 2e9, steps=3000)
    sc = scenarios.Scenario("lorentz_cavity", d, 3000, None)
    eng = engine_from_operator(d, precision="fp32", compressed=None)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(3000)
    orc.iterate(3000)
    compare(eng, orc)


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1), ("fp16", 0)])
def test_pml_coefficient_rows_match_full_arrays(gpu, precision, prec):
    """UPML coefficients read through the per-cell row index give the same
    bits as the full slab arrays (and as the oracle for FP32/FP64)."""
    sc = scenarios.graded_port_scenario(n_cells=38, steps=150, prec=prec, n_boxes=5, pml=7)
    d = sc.data
    print("case: pml rows %s" % precision, flush=True)
    fields = []
    for use in (True, False):
        eng = engine_from_operator(d, precision=precision, compressed=False)
        assert eng.pml_compressed()          # the default with full main coefficients
        eng.set_pml_compression(1 if use else 0)
        assert eng.pml_compressed() == use
        eng.iterate(3)
        eng.iterate(147)
        fields.append((eng.get_field(0), eng.get_field(1)))
        if precision != "fp16" and use:
            _kinds[id(eng)] = [p.kind for p in d.probes]
            orc = oracle_from_operator(d, ftz=True)
            orc.iterate(150)
            compare(eng, orc)
        eng.close()
    for k in range(2):
        np.testing.assert_array_equal(fields[0][k], fields[1][k])


def _h_source_cavity(n=(30, 26, 40), steps=200, prec=0):
    """PEC box with a soft H (current) excitation edge and a soft E edge."""
    from wipems_amd.operator import BC_PEC, Operator
    lines = [np.arange(k + 1, dtype=np.float64) for k in n]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6)
    op.set_gauss_excite(0.0, 20e9)
    c = [lines[k][n[k] // 2] for k in range(3)]
    e = op.add_excitation(2, (0.0, 1.0, 0.0))
    op.add_box(e, (c[0], c[1], c[2]), (c[0], lines[1][n[1] // 2 + 1], c[2]), priority=5)
    e2 = op.add_excitation(0, (1.0, 0.0, 0.0))
    op.add_box(e2, (lines[0][3], lines[1][4], lines[2][5]), (lines[0][4], lines[1][4], lines[2][5]), priority=5)
    data = op.build(prec, nr_ts=steps)
    data.probes = [op.voltage_probe((c[0], c[1], c[2]), (c[0], c[1], lines[2][n[2] // 2 + 1]), td_interval=1, name="u")]
    return scenarios.Scenario("hsrc", data, steps, op)


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1), ("fp16", 0), ("bf16", 0)])
@pytest.mark.parametrize("case", ["cavity_odd", "cavity_wide", "port_pec", "hsource"])
def test_graph_chunks_match_eager(gpu, precision, prec, case):
    """Uneven iterate() chunks (crossing the 16-step graph, the 1-step graph
    and the eager path) give the bits of one eager step at a time, and of
    the oracle for FP32/FP64."""
    if case == "cavity_odd":
        sc = scenarios.cavity_point_source(n_cells=36, steps=120, prec=prec)
    elif case == "cavity_wide":
        from wipems_amd.operator import BC_PEC, Operator
        sc = scenarios.cavity_point_source(n_cells=20, steps=120, prec=prec)
        lines = [np.arange(k + 1, dtype=np.float64) for k in (23, 19, 300)]
        op = Operator()
        op.set_mesh(lines, 1e-3)
        op.set_boundary([BC_PEC] * 6)
        op.set_gauss_excite(0.0, 10e9)
        c = [lines[k][len(lines[k]) // 2] for k in range(3)]
        e = op.add_excitation(0, (0.0, 0.0, 1.0))
        op.add_box(e, tuple(c), (c[0], c[1], lines[2][len(lines[2]) // 2 + 1]), priority=5)
        d = op.build(prec, nr_ts=120)
        d.probes = [op.voltage_probe(tuple(c), (c[0], c[1], lines[2][len(lines[2]) // 2 + 1]), td_interval=3, name="u")]
        sc = scenarios.Scenario("wide", d, 120, op)
    elif case == "port_pec":
        sc = scenarios.graded_port_scenario(n_cells=34, steps=120, prec=prec, n_boxes=5, pml=0)
    else:
        sc = _h_source_cavity(prec=prec)
    d = sc.data
    assert not d.upml
    print("case: graph chunks %s %s" % (precision, case), flush=True)
    chunks = (1, 16, 3, 16, 33, 51)
    out = []
    for graph in (True, False):
        eng = engine_from_operator(d, precision=precision, compressed=None)
        eng.set_use_graph(graph)
        for c in chunks:
            eng.iterate(c)
        out.append((eng.get_field(0), eng.get_field(1), [eng.read_probe(p)[1] for p in eng.probe_ids]))
        if graph and precision in ("fp32", "fp64"):
            _kinds[id(eng)] = [p.kind for p in d.probes]
            orc = oracle_from_operator(d, ftz=True)
            orc.iterate(sum(chunks))
            compare(eng, orc)
        eng.close()
    for k in range(2):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    for a, b in zip(out[0][2], out[1][2]):
        np.testing.assert_array_equal(a, b)


def _short_z_upml(prec, nz_lines=20, pml=8):
    """Graded mesh whose z extent is so short that one vector of cells holds
    cells of both z slabs (lines 0-8 and 11-19 with 8-cell UPML) plus a
    dielectric box and a soft source: the UPML shell launch must apply both
    slabs to that vector."""
    from wipems_amd.operator import BC_PML, Operator
    rng = np.random.default_rng(21)
    lines = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.5, n - 1))]) for n in (30, 28, nz_lines)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PML] * 6, [pml] * 6)
    op.set_gauss_excite(4e9, 4e9)
    m = op.add_material(eps_r=3.5, kappa=0.02)
    op.add_box(m, (lines[0][10], lines[1][9], lines[2][2]), (lines[0][18], lines[1][17], lines[2][15]), priority=1)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (lines[0][14], lines[1][13], lines[2][9]), (lines[0][14], lines[1][13], lines[2][11]), priority=5)
    d = op.build(prec, nr_ts=150)
    d.probes = [op.voltage_probe((lines[0][14], lines[1][13], lines[2][9]), (lines[0][14], lines[1][13], lines[2][11]),
                                 td_interval=1, name="u")]
    return d


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1)])
@pytest.mark.parametrize("compressed", [False, True])
@pytest.mark.parametrize("split", [0, 1])
def test_upml_shell_launch_short_z_bitwise(gpu, precision, prec, compressed, split):
    d = _short_z_upml(prec)
    assert len(d.upml) == 6
    print("case: short-z UPML shell, %s, compressed %s, split %d" % (precision, compressed, split), flush=True)
    eng = engine_from_operator(d, precision=precision, compressed=compressed)
    eng.set_pml_split(split)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(150)
    orc.iterate(150)
    compare(eng, orc)


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1), ("fp16", 0), ("bf16", 0)])
@pytest.mark.parametrize("case", ["short_z", "graded_port"])
def test_upml_split_launches_equal_one_launch(gpu, precision, prec, case):
    """Interior + shell launches give the same bits as one launch with the
    UPML branch, for every storage type (FP16/BF16 have no bitwise oracle,
    so this is their check of the shell launch)."""
    if case == "short_z":
        d = _short_z_upml(prec)
    else:
        d = scenarios.graded_port_scenario(n_cells=34, steps=150, prec=prec, n_boxes=5, pml=7).data
    print("case: UPML split vs one launch, %s, %s" % (case, precision), flush=True)
    out = []
    for mode in (0, 1):
        eng = engine_from_operator(d, precision=precision)
        eng.set_pml_split(mode)
        assert eng.pml_split() == bool(mode)
        eng.iterate(150)
        eng.synchronize()
        out.append([eng.get_field(w) for w in (0, 1)] + [eng.read_probe(eng.probe_ids[0])[1]])
        eng.close()
    for other in out[1:]:
        for a, b in zip(out[0], other):
            np.testing.assert_array_equal(a, b)
    assert np.abs(out[0][0]).max() > 0


def test_three_slabs_in_one_vector_are_rejected(gpu):
    """The vector kernels apply at most two UPML slabs to one vector of z
    cells; a slab set that puts three into one vector is refused when
    stepping is prepared, before any kernel runs."""
    from wipems_amd._native import FdtdError
    from wipems_amd.engine import Engine
    from wipems_amd.operator import UpmlSlab
    N = (6, 6, 12)
    eng = Engine(N, 1e-12, "fp32")
    z = np.zeros((3, 6, 6, 12), np.float32)
    eng.set_coeffs(z, z, z, z)
    for z0 in (0, 1, 2):           # three 1-cell slabs, z 0..2: one 4-cell vector
        n = (6, 6, 1)
        arrs = [np.ones((3,) + n, np.float32)] * 6
        eng.add_upml(UpmlSlab(start=(0, 0, z0), num_lines=n, arrays=arrs))
    with pytest.raises(FdtdError, match="share a vector"):
        eng.iterate(1)
    eng.close()


def test_upml_split_auto_policy(gpu):
    """Default: split launches when the interior holds at least half of the
    cells, for 2-byte storage and for FP32 (compressed or full coefficient
    arrays); one launch for FP64."""
    from wipems_amd.operator import BC_PML, Operator

    def domain(pml, n=129, prec=0):
        lines = np.arange(n, dtype=np.float64)
        op = Operator()
        op.set_mesh([lines, lines, lines], 1e-3)
        op.set_boundary([BC_PML] * 6, [pml] * 6)
        op.set_gauss_excite(2e9, 2e9)
        return op.build(prec, nr_ts=4)

    thin, thick, fat = domain(1), domain(8), domain(24)     # interior 91%, 76%, 37% of the cells
    thin64 = domain(1, 129, 1)
    cases = [(thin, "fp16", None, True), (thin, "bf16", None, True), (thin, "fp32", True, True), (thick, "fp32", True, True),
             (thick, "fp16", None, True), (thin, "fp32", False, True), (fat, "fp32", True, False), (fat, "fp16", None, False),
             (thin64, "fp64", False, False)]
    for d, precision, compressed, want in cases:
        eng = engine_from_operator(d, precision=precision, compressed=compressed)
        assert eng.pml_split() == want, (precision, compressed, want)
        eng.close()


def _mur_case(prec, bc, n=(26, 22, 24), steps=400):
    from wipems_amd.operator import Operator
    rng = np.random.default_rng(31)
    lines = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.5, k - 1))]) for k in n]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary(bc, [5] * 6)
    op.set_gauss_excite(5e9, 10e9)
    m = op.add_material(eps_r=3.0, kappa=0.01)
    op.add_box(m, (lines[0][0], lines[1][4], lines[2][5]), (lines[0][6], lines[1][12], lines[2][14]), priority=1)
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    c = [lines[k][n[k] // 2] for k in range(3)]
    op.add_box(e, (c[0], c[1], lines[2][10]), (c[0], c[1], lines[2][13]), priority=5)
    # an excitation on the x-lo plane: that face starts after the pulse
    e2 = op.add_excitation(0, (0.0, 1.0, 0.0), delay=2e-11)
    op.add_box(e2, (lines[0][0], lines[1][8], lines[2][9]), (lines[0][0], lines[1][10], lines[2][9]), priority=5)
    d = op.build(prec, nr_ts=steps)
    d.probes = [op.voltage_probe((c[0], c[1], lines[2][10]), (c[0], c[1], lines[2][13]), td_interval=1, name="u")]
    return d


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1)])
@pytest.mark.parametrize("bc", [[2] * 6, [2, 2, 3, 3, 2, 0]])
def test_mur_abc_bitwise(gpu, precision, prec, bc):
    d = _mur_case(prec, bc)
    assert len(d.mur) == bc.count(2)
    assert d.mur[0].face == 0 and d.mur[0].start_ts > 0
    print("case: Mur ABC, bc %s, %s" % (bc, precision), flush=True)
    eng = engine_from_operator(d, precision=precision)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    assert d.mur[0].start_ts < 380
    for chunk in (1, 37, 362):          # the x-lo face starts after the pulse on its plane
        eng.iterate(chunk)
        orc.iterate(chunk)
    compare(eng, orc)
    assert np.abs(eng.get_field(0)[2][0]).max() > 0     # V_z on the PEC-coefficient x-lo plane: set by Mur


def test_mur_abc_fp16_close_to_fp32(gpu):
    from wipems_amd.engine import pilot_scales
    d = _mur_case(0, [2] * 6)
    out = []
    for precision in ("fp32", "fp16"):
        eng = engine_from_operator(d, precision=precision, scales=pilot_scales(d) if precision == "fp16" else None)
        eng.iterate(400)
        out.append(eng.read_probe(eng.probe_ids[0])[1])
        eng.close()
    assert rel_l2(out[1], out[0]) < 1e-2


def _rlc_case(prec, bc=(0,) * 6, steps=300):
    from wipems_amd.operator import Operator
    rng = np.random.default_rng(41)
    n = (24, 22, 26)
    lines = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.6, 1.5, k - 1))]) for k in n]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary(list(bc), [5] * 6)
    op.set_gauss_excite(4e9, 4e9)
    c = [lines[k][n[k] // 2] for k in range(3)]
    e = op.add_excitation(0, (0.0, 0.0, 1.0))
    op.add_box(e, (c[0], c[1], lines[2][9]), (c[0], c[1], lines[2][11]), priority=5)
    par = op.add_lumped_rlc(2, 0, R=0.0, L=2e-9, C=0.5e-12, caps=False)
    op.add_box(par, (lines[0][8], lines[1][8], lines[2][12]), (lines[0][8], lines[1][8], lines[2][14]), priority=6)
    ser = op.add_lumped_rlc(0, 1, R=25.0, L=1e-9, C=1e-12, caps=True)
    op.add_box(ser, (lines[0][14], lines[1][13], lines[2][10]), (lines[0][16], lines[1][14], lines[2][11]), priority=6)
    d = op.build(prec, nr_ts=steps)
    d.probes = [op.voltage_probe((lines[0][8], lines[1][8], lines[2][12]), (lines[0][8], lines[1][8], lines[2][14]),
                                 td_interval=1, name="u_par")]
    return d


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1)])
@pytest.mark.parametrize("bc", [(0,) * 6, (2, 2, 3, 3, 2, 1)])
def test_lumped_rlc_bitwise(gpu, precision, prec, bc):
    d = _rlc_case(prec, bc)
    assert d.rlc is not None and len(d.rlc.dir) == 2 + 2 * 2 * 2
    print("case: lumped RLC, bc %s, %s" % (bc, precision), flush=True)
    eng = engine_from_operator(d, precision=precision)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    for chunk in (1, 2, 297):
        eng.iterate(chunk)
        orc.iterate(chunk)
    compare(eng, orc)
    assert np.abs(eng.read_probe(eng.probe_ids[0])[1]).max() > 0


@pytest.mark.parametrize("precision,prec", [("fp32", 0), ("fp64", 1), ("fp16", 0), ("bf16", 0)])
@pytest.mark.parametrize("case", ["dispersive_tfsf", "graded_port"])
def test_fused_hooks_equal_separate_launches(gpu, precision, prec, case):
    """The fused hook launch (TFSF rounds + ADE apply + excitation per
    location, with the next half step's ADE pre-update) gives the bits of one
    launch per hook, for every storage type, across uneven iterate() chunks
    (graph runs of 16 and 1 steps, eager steps) and a field write between
    chunks (the ADE pre-update must see it)."""
    if case == "dispersive_tfsf":
        d = scenarios.dispersive_scenario(num_lines=(41, 37, 29), steps=150, prec=prec, n_spheres=4,
                                          material_scale=1e-3, ts_factor=0.9).data
    else:
        d = scenarios.graded_port_scenario(n_cells=30, steps=150, prec=prec, n_boxes=4, pml=6).data
    print("case: fused hooks %s %s" % (case, precision), flush=True)
    scales = ([2.0 ** -4] * 3, [2.0 ** -12] * 3) if precision in ("fp16", "bf16") else None
    out = []
    for mode in (0, 1):
        eng = engine_from_operator(d, precision=precision, scales=scales)
        eng.set_fuse_hooks(mode)
        assert eng.hooks_fused() == bool(mode)
        for c in (1, 17, 3):
            eng.iterate(c)
        V = eng.get_field(0)
        V[0, 20, 18, 14] += 1e-3
        eng.set_field(0, V)
        for c in (16, 40, 5):
            eng.iterate(c)
        out.append([eng.get_field(w) for w in (0, 1)] + [eng.read_probe(p)[1] for p in eng.probe_ids])
        eng.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out[0][0]).max() > 0


@pytest.mark.parametrize("precision,compressed,split,band,xcd,want_xcd", [
    ("fp32", False, 0, 0, 1, 1), ("fp32", True, 1, 0, 1, 1), ("fp32", True, 0, 3, 1, 1), ("fp32", False, 0, 5, 0, 0),
    ("fp32", True, 1, 2, 0, 0), ("fp64", False, 0, 3, 1, 1), ("fp64", False, 0, 0, 1, 1), ("fp64", False, 1, 0, 8, 2),
])
def test_tile_orders_bitwise(gpu, precision, compressed, split, band, xcd, want_xcd):
    """Banded tile orders (y bands of `band` tiles swept plane by plane, the
    last band partial; with xcd, chunks of xcd tiles of each band-plane to
    the eight XCDs in turn, capped at one sub-band per XCD, padding blocks
    idle) give the oracle's bits, with one UPML launch and with the
    interior/shell split, for the ascending voltage and the descending
    current sweep; 64 cells per axis give planes of 6 tiles (FP32; 11 for
    FP64), no multiple of 8 or of the band."""
    sc = scenarios.graded_port_scenario(n_cells=64, steps=160, prec=1 if precision == "fp64" else 0, n_boxes=5, pml=6)
    d = sc.data
    eng = engine_from_operator(d, precision=precision, compressed=compressed)
    eng.set_tile_order(band, xcd)
    eng.set_pml_split(split)
    got_band, got_xcd = eng.tile_order()
    assert (got_band, got_xcd) == (band, want_xcd)
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(160)
    orc.iterate(160)
    compare(eng, orc)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_tile_orders_reduced_precision(gpu, precision):
    """2-byte storage: banded and whole-plane tile orders give identical fields
    (planes of 3 tiles in the one-launch form)."""
    sc = scenarios.graded_port_scenario(n_cells=64, steps=120, prec=0, n_boxes=5, pml=6)
    d = sc.data
    from wipems_amd.engine import pilot_scales
    scales = pilot_scales(d, steps=64)
    out = []
    for band, xcd in ((0, 0), (2, 1), (1, 0)):
        eng = engine_from_operator(d, precision=precision, scales=scales)
        eng.set_pml_split(0)
        eng.set_tile_order(band, xcd)
        eng.iterate(120)
        out.append((eng.get_field(0), eng.get_field(1)))
        eng.close()
    for k in (1, 2):
        for which in (0, 1):
            np.testing.assert_array_equal(out[0][which], out[k][which])


@pytest.mark.parametrize("precision,compressed,align", [
    ("fp32", True, 4), ("fp32", True, 8), ("fp32", False, 8), ("fp64", False, 4),
])
def test_shell_align_bitwise(gpu, precision, compressed, align):
    """The interior/shell split with the interior's z range aligned to 4 or 8
    vectors (the shell launch taking the non-slab cells between) gives the
    oracle's bits; 64 cells per axis with 6-cell UPML."""
    sc = scenarios.graded_port_scenario(n_cells=64, steps=140, prec=1 if precision == "fp64" else 0, n_boxes=5, pml=6)
    d = sc.data
    eng = engine_from_operator(d, precision=precision, compressed=compressed)
    eng.set_pml_split(1)
    eng.set_shell_align(align)
    assert eng.pml_split()
    _kinds[id(eng)] = [p.kind for p in d.probes]
    orc = oracle_from_operator(d, ftz=True)
    eng.iterate(140)
    orc.iterate(140)
    compare(eng, orc)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_shell_align_reduced_precision(gpu, precision):
    """2-byte storage (half-width shell vectors): aligned and unaligned
    interior ranges give identical fields."""
    from wipems_amd.engine import pilot_scales
    sc = scenarios.graded_port_scenario(n_cells=64, steps=120, prec=0, n_boxes=5, pml=6)
    d = sc.data
    scales = pilot_scales(d, steps=64)
    out = []
    for align in (1, 4, 8):
        eng = engine_from_operator(d, precision=precision, scales=scales)
        eng.set_pml_split(1)
        eng.set_shell_align(align)
        eng.iterate(120)
        out.append((eng.get_field(0), eng.get_field(1)))
        eng.close()
    for k in (1, 2):
        for which in (0, 1):
            np.testing.assert_array_equal(out[0][which], out[k][which])
