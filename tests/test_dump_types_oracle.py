"""Pins of the dump-type restatement (oracle/dft.py dump_value / dump_box) and
of the operator's stored materials, on the CPU.

The restatement covers the dump types E, H, J, rot H, D, B of ProcessFields
(Common/processfields.h:43) with the three interpolations of
Engine_Interface_FDTD (FDTD/engine_interface_fdtd.cpp:60-288).  It is pinned
by closed forms that a plausible mistake breaks: node values of linear
fields, cell averages of linear fields (an offset of the wrong sign moves
them by a whole cell), the discrete curl of fields with a known constant
curl (a wrong sign, a swapped axis or a forward instead of a backward
difference changes it), material scaling with materials that vary along one
axis (a transposed index reads the wrong edge), the boundary rules, and the
node interpolation of E and H against the vectorised restatement that the
nf2ff tests pin.  Runs without a GPU."""
import numpy as np
import pytest

from oracle.dft import dump_box, dump_time, float_ftz, interp_e, interp_h, split_edges
from wipems_amd.dump import box_lines, dump_on_dual_mesh
from wipems_amd.operator import BC_PEC, Operator

EPS0 = 8.85418781762039e-12
MUE0 = 4e-7 * np.pi


def graded(rng, n, lo=0.5, hi=2.0):
    return np.concatenate([[0.0], np.cumsum(rng.uniform(lo, hi, n - 1))]) * 1e-3


def edges_of(lines):
    """Primary and dual edge lengths per axis (Operator::GetDiscDelta, FDTD/operator.cpp:151-207)."""
    P, D = [], []
    for l in lines:
        prim = np.append(np.diff(l), l[-1] - l[-2])
        dl = np.append(0.5 * (l[:-1] + l[1:]), l[-1] + 0.5 * (l[-1] - l[-2]))
        P.append(prim)
        D.append(np.concatenate([[l[1] - l[0]], np.diff(dl)]))
    return np.concatenate(P), np.concatenate(D)


def full(N):
    return (0, 0, 0), tuple(k - 1 for k in N)


def test_node_interpolation_matches_vectorised_e_and_h():
    rng = np.random.default_rng(11)
    N = (7, 6, 8)
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    V = rng.standard_normal((3,) + N)
    I = rng.standard_normal((3,) + N)
    a, b = full(N)
    np.testing.assert_array_equal(dump_box(V, I, ep, ed, {}, 0, 1, a, b), interp_e(V, split_edges(ep, N), a, b))
    np.testing.assert_array_equal(dump_box(V, I, ep, ed, {}, 1, 1, a, b), interp_h(I, split_edges(ed, N), a, b))


def test_no_interpolation_is_the_raw_edge_value():
    rng = np.random.default_rng(12)
    N = (6, 5, 7)
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    P, D = split_edges(ep, N), split_edges(ed, N)
    f = rng.standard_normal((3,) + N)
    V = np.empty_like(f)
    I = np.empty_like(f)
    for n in range(3):
        sh = [1, 1, 1]
        sh[n] = N[n]
        V[n] = f[n] * P[n].reshape(sh)
        I[n] = f[n] * D[n].reshape(sh)
    a, b = full(N)
    np.testing.assert_allclose(dump_box(V, I, ep, ed, {}, 0, 0, a, b), f, rtol=1e-14)
    np.testing.assert_allclose(dump_box(V, I, ep, ed, {}, 1, 0, a, b), f, rtol=1e-14)


def test_cell_interpolation_of_linear_e_is_the_cell_centre_value():
    """E_n = c + g * (index along nP): the cell average over pos..pos+1 in nP
    and nPP is c + g * (pos[nP] + 1/2) inside, 0 on the last lines."""
    rng = np.random.default_rng(13)
    N = (6, 7, 5)
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    P = split_edges(ep, N)
    c, g = [0.3, -1.2, 2.0], [1.5, -0.25, 0.75]
    V = np.empty((3,) + N)
    idx = np.meshgrid(*[np.arange(k) for k in N], indexing="ij")
    for n in range(3):
        nP = (n + 1) % 3
        sh = [1, 1, 1]
        sh[n] = N[n]
        V[n] = (c[n] + g[n] * idx[nP]) * P[n].reshape(sh)
    a, b = full(N)
    out = dump_box(V, V, ep, ed, {}, 0, 2, a, b)
    inner = (idx[0] < N[0] - 1) & (idx[1] < N[1] - 1) & (idx[2] < N[2] - 1)
    for n in range(3):
        want = np.where(inner, c[n] + g[n] * (idx[(n + 1) % 3] + 0.5), 0.0)
        np.testing.assert_allclose(out[n], want, rtol=1e-13, atol=1e-13)


def test_cell_interpolation_of_linear_h_on_a_uniform_mesh():
    """H_n linear along its own axis: the dual-grid cell rule blends pos and
    pos+1 half and half on a uniform mesh, giving the value half a cell up."""
    N = (6, 5, 7)
    h = 1e-3
    lines = [np.arange(k) * h for k in N]
    ep, ed = edges_of(lines)
    D = split_edges(ed, N)
    idx = np.meshgrid(*[np.arange(k) for k in N], indexing="ij")
    I = np.empty((3,) + N)
    c, g = [1.0, -2.0, 0.5], [0.2, 0.7, -1.1]
    for n in range(3):
        sh = [1, 1, 1]
        sh[n] = N[n]
        I[n] = (c[n] + g[n] * idx[n]) * D[n].reshape(sh)
    a, b = full(N)
    out = dump_box(I, I, ep, ed, {}, 1, 2, a, b)
    for n in range(3):
        want = np.where(idx[n] < N[n] - 1, c[n] + g[n] * (idx[n] + 0.5), 0.0)
        np.testing.assert_allclose(out[n], want, rtol=1e-13, atol=1e-13)


def _curl_case():
    """Graded mesh, H_z = gz * x and H_y = gy * z at their dual node
    positions (I = H times the dual edge length along the component): the
    dual edge lengths are the distances between dual nodes, so the discrete
    circulation over each dual face gives curl H = (-gy, -gz, 0) exactly at
    nodes away from the lower faces."""
    rng = np.random.default_rng(31)
    N = (6, 7, 6)
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    D = split_edges(ed, N)
    dual_nodes = [np.append(0.5 * (l[:-1] + l[1:]), l[-1] + 0.5 * (l[-1] - l[-2])) for l in lines]
    gz, gy = 3.0, -1.75
    idx = np.meshgrid(*[np.arange(k) for k in N], indexing="ij")
    I = np.zeros((3,) + N)
    I[2] = gz * dual_nodes[0][idx[0]] * D[2][idx[2]]
    I[1] = gy * dual_nodes[2][idx[2]] * D[1][idx[1]]
    return N, ep, ed, I, idx, (-gy, -gz, 0.0)


def test_roth_of_a_known_curl():
    N, ep, ed, I, idx, curl = _curl_case()
    V = np.zeros_like(I)
    a, b = full(N)
    out = dump_box(V, I, ep, ed, {}, 3, 0, a, b)
    inner = (idx[0] > 0) & (idx[1] > 0) & (idx[2] > 0)
    for n in range(3):
        np.testing.assert_allclose(out[n][inner], curl[n], rtol=1e-12, atol=1e-9)
    # at the lower boundary the backward terms are dropped (pos[nP] or pos[nPP] == 0)
    q = (0, 3, 3)
    assert out[1][q] == pytest.approx(-I[2][q] / (ed[N[0] + N[1] + q[2]] * ed[q[0]]))
    cell = dump_box(V, I, ep, ed, {}, 3, 2, (1, 1, 1), (N[0] - 2, N[1] - 2, N[2] - 2))
    for n in range(3):
        np.testing.assert_allclose(cell[n], curl[n], rtol=1e-12, atol=1e-9)


def test_material_scaling_reads_the_edge_of_the_value():
    """J = kappa E, D = eps E, B = mu H with random materials per edge and
    unit fields: the dump returns the material of each edge."""
    rng = np.random.default_rng(17)
    N = (5, 5, 5)                     # a transposed index still reads a valid, wrong, edge
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    P, D = split_edges(ep, N), split_edges(ed, N)
    V = np.empty((3,) + N)
    I = np.empty((3,) + N)
    for n in range(3):
        sh = [1, 1, 1]
        sh[n] = N[n]
        V[n] = np.broadcast_to(P[n].reshape(sh), N)
        I[n] = np.broadcast_to(D[n].reshape(sh), N)
    mats = {k: rng.uniform(0.5, 3.0, (3,) + N).astype(np.float32) for k in range(3)}
    a, b = full(N)
    for dump_type, kind, F in ((2, 1, V), (4, 0, V), (5, 2, I)):
        out = dump_box(V, I, ep, ed, mats, dump_type, 0, a, b)
        np.testing.assert_allclose(out, mats[kind].astype(np.float64), rtol=1e-14)
    # node interpolation of uniform materials is the material times the E / H node value
    uni = {k: np.broadcast_to(np.array([1.5, 2.5, 4.0], np.float32).reshape(3, 1, 1, 1), (3,) + N).copy() for k in range(3)}
    Vr, Ir = rng.standard_normal((3,) + N), rng.standard_normal((3,) + N)
    e = dump_box(Vr, Ir, ep, ed, {}, 0, 1, a, b)
    hh = dump_box(Vr, Ir, ep, ed, {}, 1, 1, a, b)
    scale = np.array([1.5, 2.5, 4.0]).reshape(3, 1, 1, 1)
    np.testing.assert_allclose(dump_box(Vr, Ir, ep, ed, uni, 2, 1, a, b), e * scale, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(dump_box(Vr, Ir, ep, ed, uni, 4, 1, a, b), e * scale, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(dump_box(Vr, Ir, ep, ed, uni, 5, 1, a, b), hh * scale, rtol=1e-13, atol=1e-12)
    # without the material the reference returns 0
    assert not dump_box(Vr, Ir, ep, ed, {}, 2, 0, a, b).any()


def test_cell_interpolation_of_h_known_answer():
    """Lines 0, 1, 3, 4.5: dual edge lengths 1, 1.5, 1.75, 1.5.  Raw H_x 1, 2,
    4, 8 along x blend with rel = d_i / (d_i + d_{i+1}): 1*0.6 + 2*0.4 = 1.4,
    2*7/13 + 4*6/13 = 38/13, 4*6/13 + 8*7/13 = 80/13, and 0 on the last line."""
    l = np.array([0.0, 1.0, 3.0, 4.5])
    N = (4, 4, 4)
    ep, ed = edges_of([l, l, l])
    D = split_edges(ed, N)
    np.testing.assert_allclose(D[0], [1.0, 1.5, 1.75, 1.5])
    I = np.zeros((3,) + N)
    I[0] = (np.array([1.0, 2.0, 4.0, 8.0]) * D[0]).reshape(4, 1, 1)
    out = dump_box(I, I, ep, ed, {}, 1, 2, (0, 1, 1), (3, 1, 1))
    np.testing.assert_allclose(out[0, :, 0, 0], [1.4, 38 / 13, 80 / 13, 0.0], rtol=1e-14)


def test_node_and_boundary_rules_of_the_dual_grid():
    rng = np.random.default_rng(19)
    N = (5, 5, 5)
    lines = [graded(rng, n) for n in N]
    ep, ed = edges_of(lines)
    I = rng.standard_normal((3,) + N)
    a, b = full(N)
    out = dump_box(I, I, ep, ed, {}, 1, 1, a, b)
    idx = np.meshgrid(*[np.arange(k) for k in N], indexing="ij")
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        zero = (idx[0] == N[0] - 1) | (idx[1] == N[1] - 1) | (idx[2] == N[2] - 1) | (idx[nP] == 0) | (idx[nPP] == 0)
        assert not out[n][zero].any()
    cell = dump_box(I, I, ep, ed, {}, 1, 2, a, b)
    for n in range(3):
        assert not cell[n][idx[n] == N[n] - 1].any()


def test_dump_time_and_file_mesh():
    assert dump_time(1, 10, 2.0) == 21.0
    for t in (0, 2, 3, 4, 5):
        assert dump_time(t, 10, 2.0) == 20.0
    # the file mesh (ProcessFields::SetDumpMode + openems.cpp:595-600)
    assert dump_on_dual_mesh(1, 0) and not dump_on_dual_mesh(0, 0) and not dump_on_dual_mesh(5, 0)
    assert all(dump_on_dual_mesh(t, 2) for t in range(6))
    assert not any(dump_on_dual_mesh(t, 1) for t in range(6))
    l = np.array([0.0, 1.0, 3.0, 4.5])
    got = box_lines([l, l, l], (0, 1, 2), (3, 3, 3), dual=True)
    np.testing.assert_array_equal(got[0], [0.5, 2.0, 3.75, 5.25])
    np.testing.assert_array_equal(got[2], [3.75, 5.25])
    np.testing.assert_array_equal(box_lines([l, l, l], (1, 1, 1), (2, 2, 2))[1], [1.0, 3.0])


# ---------------------------------------------------------------------------
# operator material storage (Operator::Calc_ECPos, FDTD/operator.cpp:1176-1188)
# ---------------------------------------------------------------------------
def _material_op(streamed=False):
    rng = np.random.default_rng(23)
    lines = [graded(rng, n, 0.8, 1.6) * 1e3 for n in (10, 9, 8)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_gauss_excite(0.0, 10e9)
    m = op.add_material(eps_r=4.0, kappa=0.05, mue_r=2.0)
    op.add_box(m, (lines[0][2], lines[1][1], lines[2][2]), (lines[0][6], lines[1][6], lines[2][5]), priority=1)
    for k in range(3):
        op.set_material_store_flags(k)
    data = op.build_streamed(nr_ts=10) if streamed else op.build(0, nr_ts=10)
    return op, data


def test_stored_materials_follow_the_ec_values():
    """C = eps * area / delta and G = kappa * area / delta on the primary
    grid, L = mu * area / delta on the dual grid: the stored materials are
    the EffMat values behind the operator's EC arrays."""
    op, d = _material_op()
    assert set(d.materials) == {0, 1, 2}
    C, G, L, R = op.get_ec()
    N = d.num_lines
    P, D = split_edges(d.edge_primary, N), split_edges(d.edge_dual, N)
    eps, kap, mu = (d.materials[k].astype(np.float64) for k in range(3))
    for n in range(3):
        nP, nPP = (n + 1) % 3, (n + 2) % 3
        sh = lambda ax: [N[ax] if k == ax else 1 for k in range(3)]
        area_p = D[nP].reshape(sh(nP)) * D[nPP].reshape(sh(nPP))
        area_d = P[nP].reshape(sh(nP)) * P[nPP].reshape(sh(nPP))
        np.testing.assert_allclose(C[n], eps[n] * area_p / P[n].reshape(sh(n)), rtol=2e-7)
        np.testing.assert_allclose(G[n], kap[n] * area_p / P[n].reshape(sh(n)), rtol=2e-7)
        np.testing.assert_allclose(L[n], mu[n] * area_d / D[n].reshape(sh(n)), rtol=2e-7)
    # vacuum edges far from the box hold the vacuum values exactly (as float)
    assert eps[0][0, 0, 0] == np.float32(EPS0) and mu[2][0, 0, 0] == np.float32(MUE0) and kap[1][0, 0, 0] == 0.0
    assert eps.max() == pytest.approx(4.0 * EPS0, rel=1e-6) and mu.max() == pytest.approx(2.0 * MUE0, rel=1e-6)
    assert kap.max() == pytest.approx(0.05, rel=1e-6)


def test_streamed_build_stores_the_same_materials():
    _, full_d = _material_op(False)
    _, st = _material_op(True)
    for k in range(3):
        np.testing.assert_array_equal(st.materials[k], full_d.materials[k])


def test_materials_not_stored_unless_flagged():
    rng = np.random.default_rng(29)
    lines = [graded(rng, n) * 1e3 for n in (5, 5, 5)]
    op = Operator()
    op.set_mesh(lines, 1e-3)
    op.set_boundary([BC_PEC] * 6, [0] * 6)
    op.set_gauss_excite(0.0, 10e9)
    op.set_material_store_flags(1)
    d = op.build(0, nr_ts=10)
    assert set(d.materials) == {1}
    with pytest.raises(ValueError):
        op.set_material_store_flags(3)


def test_float_rounding_flushes_denormals():
    """The reference rounds dump values to float with flush-to-zero set: a
    double that rounds to a float denormal is 0, FLT_MIN and larger keep
    their float value, signs stay."""
    tiny = float(np.finfo(np.float32).tiny)
    x = np.array([1e-39, -3e-39, tiny, -tiny, 2.5e-38, 1.0 / 3.0, 0.0])
    y = float_ftz(x)
    assert y.dtype == np.float32
    np.testing.assert_array_equal(y, np.array([0.0, -0.0, tiny, -tiny, 2.5e-38, 1.0 / 3.0, 0.0], dtype=np.float32))
    assert np.signbit(y[1])
