"""Pinning the CPU oracle (oracle/upside_oracle.py) before it checks the GPU.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * reference-computed energies of example/15.SpatialTransformation frames
    (committed fixture) -- the whole ff_2.1 graph on real coordinates;
  * the reference engine compiled from its own sources (oracle/_ref) on seeded
    random chains -- energies, forces, edge counts, BP iterations;
  * mathematics: central differences of the FP64 oracle, zero net force and
    torque, BP on trees against brute-force enumeration, Kabsch best fit,
    spline partition of unity and exact grid interpolation;
  * Random123 known-answer vectors for Threefry-4x32-20.
"""
import ctypes as ct
import itertools
import os

import numpy as np
import pytest

from conftest import GOLDEN, REF_LIB, RefEngine, read_pos, ref_lib, rel_rms

import upside_oracle as O


# ------------------------------------------------------------------ golden
def test_chig_frames_match_reference_run(chig):
    """Potential of the 20 logged frames of the reference's chignolin run
    (example/15.SpatialTransformation/outputs/simple_test/chig.run.up)."""
    path, z = chig
    eng = O.Engine(path)
    got = np.array([eng.energy(p) for p in z["frame_pos"]], np.float64)
    want = z["frame_potential"].ravel().astype(np.float64)
    # values were logged as f32 after a BP solve at tol 1e-3; the measured
    # agreement is ~3e-5 absolute on |E| ~ 10-40
    assert np.abs(got - want).max() < 2e-3, np.abs(got - want).max()


def test_chig_logged_terms(chig):
    """Terms the reference run logged per frame (same fixture): per-site
    hydrogen-bond scores (protein_hbond output column 6), the Rama map energy,
    the sigmoid-coupling energy and the rotamer free energy (per-residue
    splits summed)."""
    path, z = chig
    eng = O.Engine(path)
    for f in range(0, len(z["frame_pos"]), 3):
        eng.energy(z["frame_pos"][f])
        np.testing.assert_allclose(eng.nodes["protein_hbond"].output[:, 6], z["frame_hbond"][f], atol=5e-6)
        assert abs(eng.nodes["rama_map_pot"].potential - z["frame_rama_map_potential"][f].sum()) < 1e-4
        assert abs(eng.nodes["sigmoid_coupling_environment"].potential - z["frame_sigmoid_coupling"][f].sum()) < 1e-4
        assert abs(eng.nodes["rotamer"].potential - z["frame_rotamer_free_energy"][f].sum()) < 2e-3


# -------------------------------------------------------- compiled reference
needs_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (needs /root/reference)")


@needs_ref
@pytest.mark.parametrize("n_res,seed", [(30, 1), (60, 2), (120, 3)])
def test_oracle_matches_compiled_reference(chain_file, n_res, seed):
    path = chain_file(n_res, seed)
    pos = read_pos(path)
    ref = RefEngine(ref_lib(), path, pos.shape[0])
    eng = O.Engine(path)
    e_ref, d_ref = ref.energy(pos), ref.deriv(pos)
    e, d = float(eng.energy(pos)), eng.deriv(pos)
    assert abs(e - e_ref) <= 1e-5 * abs(e_ref) + 1e-3, (e, e_ref)
    assert rel_rms(d, d_ref) < 1e-4
    # BP iteration count of the last solve and the rotamer edge counts by type
    rot = eng.nodes["rotamer"]
    n_type = rot.ig.type1.max() + 1
    cnt = np.zeros((n_type, rot.ig.n_type2), np.float32)
    np.add.at(cnt, (rot.ig.type1[rot.ig.i1], rot.ig.type2[rot.ig.i2]), 1)
    got = ref.value(cnt.shape, "rotamer", "count_edges_by_type")
    np.testing.assert_array_equal(got, cnt)


@needs_ref
def test_oracle_matches_reference_at_trajectory_frames(chain_file):
    """Forces along a short reference trajectory (positions from the reference
    CLI would need its writer; instead perturb the start with seeded noise)."""
    path = chain_file(50, 4)
    pos0 = read_pos(path)
    ref = RefEngine(ref_lib(), path, pos0.shape[0])
    eng = O.Engine(path)
    rng = np.random.default_rng(0)
    for k in range(4):
        pos = (pos0 + rng.normal(scale=0.3, size=pos0.shape)).astype(np.float32)
        assert rel_rms(eng.deriv(pos), ref.deriv(pos)) < 1e-4


# ------------------------------------------------------------- mathematics
def test_forces_central_differences(chain_file):
    """FP64 oracle: analytic dE/dx against central differences, with BP solved
    to 1e-10 so the stationarity the derivative assumes holds.  The start is
    jittered: the exact ideal geometry of the initial chain (omega exactly pi)
    sits on branch points of the dihedral wrap."""
    path = chain_file(12, 5, tol=1e-10, max_iter=20000)
    rng = np.random.default_rng(1)
    pos = read_pos(path).astype(np.float64) + rng.normal(scale=0.1, size=(36, 3))
    eng = O.Engine(path, dtype=np.float64)
    d = eng.deriv(pos)
    h = 1e-5
    idx = rng.choice(pos.size, size=40, replace=False)
    fd = []
    for i in idx:
        p1 = pos.copy().ravel(); p1[i] += h
        p2 = pos.copy().ravel(); p2[i] -= h
        e1 = eng.compute(p1.reshape(pos.shape))
        e2 = eng.compute(p2.reshape(pos.shape))
        fd.append((e1 - e2) / (2 * h))
    fd = np.array(fd)
    # measured 8.5e-6 over all coordinates (the largest single deviations,
    # ~3e-4 relative, are residue 0's z components, in the compiled reference too)
    assert rel_rms(d.ravel()[idx], fd) < 5e-5


def test_net_force_and_torque_vanish(chain_file):
    """Every ff_2.1 term is internal: sum of forces and torques is zero (on a
    jittered start: at the exact ideal geometry the affine frames' eigenvalue
    derivative is degenerate and the torque does not vanish, in the compiled
    reference as well)."""
    path = chain_file(40, 6)
    pos = read_pos(path).astype(np.float64)
    pos += np.random.default_rng(0).normal(scale=0.05, size=pos.shape)
    d = O.Engine(path, dtype=np.float64).deriv(pos)
    scale = np.abs(d).sum()
    assert np.abs(d.sum(0)).max() < 1e-12 * scale
    c = pos - pos.mean(0)
    assert np.abs(np.cross(c, d).sum(0)).max() < 1e-6 * scale


def _tree_rotamer(seed):
    """A Rotamer oracle node on a tree of two 3-state and three 6-state
    residues: 3a-6a, 6a-6b, 6b-3b, 6b-6c."""
    rng = np.random.default_rng(seed)
    T = np.float64
    r = object.__new__(O.Rotamer)
    r.T = T
    r.damping, r.tol, r.max_iter, r.chunk = T(0.4), np.float64(1e-12), 100000, 2
    r.n_elem = {1: 0, 3: 2, 6: 3}
    r.prob = {3: rng.uniform(0.05, 1, (2, 3)), 6: rng.uniform(0.05, 1, (3, 6))}
    r.offset = {1: np.zeros(0), 3: np.zeros(2), 6: np.zeros(3)}
    r.prob[1] = np.ones((0, 1))
    E = {}
    for c in ((1, 1), (1, 3), (1, 6)):
        E[c] = dict(P=np.ones((0, c[0], c[1])), i1=np.zeros(0, int), i2=np.zeros(0, int))
    E[(3, 3)] = dict(P=np.ones((0, 3, 3)), i1=np.zeros(0, int), i2=np.zeros(0, int))
    E[(3, 6)] = dict(P=rng.uniform(0.1, 2, (2, 3, 6)), i1=np.array([0, 1]), i2=np.array([0, 1]))
    E[(6, 6)] = dict(P=rng.uniform(0.1, 2, (2, 6, 6)), i1=np.array([0, 1]), i2=np.array([1, 2]))
    r.E = E
    return r


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bp_exact_on_trees(seed):
    """On a tree BP is exact: marginals equal brute-force enumeration and the
    Bethe free energy equals -log Z."""
    r = _tree_rotamer(seed)
    r.solve()
    F = r.free_energy()
    nodes = [(3, 0), (3, 1), (6, 0), (6, 1), (6, 2)]
    Z = 0.0
    marg = {n: np.zeros(n[0]) for n in nodes}
    for x in itertools.product(range(3), range(3), range(6), range(6), range(6)):
        s = dict(zip(nodes, x))
        w = np.prod([r.prob[k][i, s[(k, i)]] for (k, i) in nodes])
        for c in ((3, 6), (6, 6)):
            e = r.E[c]
            for j in range(len(e["i1"])):
                w *= e["P"][j, s[(c[0], e["i1"][j])], s[(c[1], e["i2"][j])]]
        Z += w
        for n in nodes:
            marg[n][s[n]] += w
    for (k, i) in nodes:
        np.testing.assert_allclose(r.node_marg[k][i], marg[(k, i)] / Z, atol=1e-8)
    assert abs(F - (-np.log(Z))) < 1e-6


def test_bp_damped_fixed_point_is_damping_free():
    """Damping changes the path, not the fixed point: marginals at damping 0.0
    and 0.8 agree (a wrongly placed damping term moves the fixed point)."""
    a, b = _tree_rotamer(3), _tree_rotamer(3)
    a.damping, b.damping = np.float64(0.0), np.float64(0.8)
    a.solve(); b.solve()
    for k in (3, 6):
        np.testing.assert_allclose(a.node_marg[k], b.node_marg[k], atol=1e-8)


def test_symm_eig4_dominant_vector():
    """Grouped Householder + implicit QR (src/eig.cpp:233-273) against LAPACK."""
    rng = np.random.default_rng(7)
    A = rng.normal(size=(64, 4, 4))
    A = (A + A.transpose(0, 2, 1)).astype(np.float64)
    vals, vecs = O.symm_eig4(A, np.float64)  # slot 0 the largest, eigenvectors as rows
    w, v = np.linalg.eigh(A)
    got = vecs[:, 0, :]
    want = v[:, :, -1]
    np.testing.assert_allclose(vals[:, 0], w[:, -1], atol=1e-9)
    np.testing.assert_allclose(np.sort(vals, 1), w, atol=1e-9)
    np.testing.assert_allclose(np.abs((got * want).sum(1)), 1.0, atol=1e-9)


def test_affine_alignment_is_kabsch_best_fit(chain_file):
    """affine_alignment's quaternion rotates the reference triple onto the
    atoms with a least-squares optimal rotation: its residual equals the
    Kabsch (SVD) optimum, and the frame origin is the atoms' centroid."""
    path = chain_file(20, 8)
    pos = read_pos(path).astype(np.float64)
    pos += np.random.default_rng(4).normal(scale=0.05, size=pos.shape)
    eng = O.Engine(path, dtype=np.float64)
    eng.energy(pos)
    aff = eng.nodes["affine_alignment"]
    out = aff.output
    q = out[:, 3:7]
    np.testing.assert_allclose((q * q).sum(1), 1.0, atol=1e-9)
    U = O.quat_to_rot(q)
    for n in range(len(out)):
        x = pos[aff.atoms[n]]
        ref = aff.ref[n]
        xc = x - x.mean(0)
        rc = ref - ref.mean(0)
        u, s, vt = np.linalg.svd(xc.T @ rc)
        dfix = np.sign(np.linalg.det(u @ vt))
        R = u @ np.diag([1, 1, dfix]) @ vt
        best = ((xc - rc @ R.T) ** 2).sum()
        got = ((xc - rc @ U[n].T) ** 2).sum()
        # the QR stops at off-diagonal 1e-5, so the fit is optimal to second order in that
        assert got <= best * (1 + 1e-6) + 1e-8, (n, got, best)
        np.testing.assert_allclose(out[n, :3], x.mean(0), atol=1e-9)


def test_clamped_spline_partition_of_unity():
    """Uniform cubic B-spline weights sum to one, linear coefficients reproduce
    the line, and the clamped ends hold the end value (src/spline.h:278-358)."""
    x = np.linspace(1.0, 9.999, 50)
    v, d = O.clamped_deboor(np.ones((50, 12)), x, 12)
    np.testing.assert_allclose(v, 1.0, atol=1e-12)
    np.testing.assert_allclose(d, 0.0, atol=1e-12)
    # B-spline k is centred at k-1, so coefficients c_k = k-1 give f(x) = x - 2 ... shifted
    xs = np.array([3.25, 4.5, 7.75])
    c = np.tile(np.arange(12.0), (3, 1))
    v, d = O.clamped_deboor(c, xs, 12)
    np.testing.assert_allclose(d, 1.0, atol=1e-12)
    np.testing.assert_allclose(np.diff(v), np.diff(xs), atol=1e-12)
    lo, dlo = O.clamped_deboor(c[:2], np.array([0.2, -3.0]), 12)
    np.testing.assert_allclose(lo, 1.0, atol=1e-12)  # c0/6 + 2 c1/3 + c2/6 with c = 0,1,2
    np.testing.assert_allclose(dlo, 0.0)
    hi, dhi = O.clamped_deboor(c[:2], np.array([10.0, 50.0]), 12)
    np.testing.assert_allclose(hi, 10.0, atol=1e-12)
    np.testing.assert_allclose(dhi, 0.0)


def test_periodic_bicubic_interpolates_grid():
    """The periodic bicubic spline passes through its grid values and its
    derivative matches central differences (src/spline.cpp:262-292)."""
    rng = np.random.default_rng(3)
    data = rng.normal(size=(1, 12, 12))
    coef = O.periodic_spline_2d(data)
    for i in range(12):
        for j in range(12):
            v, dx, dy = O.bicubic_eval(coef, 0, np.array([i + 0.0]), np.array([j + 0.0]))
            assert abs(float(np.ravel(v)[0]) - data[0, i, j]) < 1e-10
    h = 1e-6
    x, y = np.array([3.3]), np.array([7.6])
    v, dx, dy = O.bicubic_eval(coef, 0, x, y)
    vp, _, _ = O.bicubic_eval(coef, 0, x + h, y)
    vm, _, _ = O.bicubic_eval(coef, 0, x - h, y)
    assert abs(float(np.ravel((vp - vm) / (2 * h))[0]) - float(np.ravel(dx)[0])) < 1e-6


def test_compact_sigmoid_ends_and_slope():
    """1 below -1/sharp, 0 above +1/sharp, C1 at both ends, derivative matches
    finite differences (src/vector_math.h:701-720)."""
    sharp = 2.5
    x = np.array([-1.0, -0.5 / sharp, 0.0, 0.3 / sharp, 1.0])
    v, d = O.compact_sigmoid(x, sharp)
    assert v[0] == 1 and v[-1] == 0 and d[0] == 0 and d[-1] == 0
    h = 1e-7
    for xi, di in zip(x[1:4], d[1:4]):
        vp, _ = O.compact_sigmoid(np.array([xi + h]), sharp)
        vm, _ = O.compact_sigmoid(np.array([xi - h]), sharp)
        assert abs((vp[0] - vm[0]) / (2 * h) - di) < 1e-6
    e, _ = O.compact_sigmoid(np.array([-1 / sharp + 1e-9, 1 / sharp - 1e-9]), sharp)
    np.testing.assert_allclose(e, [1, 0], atol=1e-6)


def test_pair_list_is_brute_force_with_reference_dist2(chain_file):
    """The oracle's edge set equals a literal double loop that applies the
    exclusion, i<j for the symmetric graph and the reference's dist2
    expression fma(x,x,fma(z,z,y*y)) compared in float32."""
    path = chain_file(25, 9)
    pos = read_pos(path)
    eng = O.Engine(path)
    eng.energy(pos)
    ig = eng.nodes["rotamer"].ig
    x = eng.nodes[eng.nodes["rotamer"].args[0]].output[ig.loc1, :3].astype(np.float32)
    cut2 = np.float32(ig.cutoff) * np.float32(ig.cutoff)
    want = set()
    for i in range(len(x)):
        for j in range(i + 1, len(x)):
            if (ig.id1[i] >> 4) == (ig.id1[j] >> 4):
                continue
            dx, dy, dz = (x[i] - x[j]).astype(np.float64)
            d2 = np.float32(dz * dz + np.float64(np.float32(dy * dy)))
            d2 = np.float32(dx * dx + np.float64(d2))
            if d2 < cut2:
                want.add((i, j))
    assert set(zip(ig.i1.tolist(), ig.i2.tolist())) == want


# ------------------------------------------------------------------- RNG
# Random123 known-answer vectors for threefry4x32 with 20 rounds
# (kat_vectors of the Random123 distribution): counter, key -> output.
THREEFRY_KAT = [
    ([0, 0, 0, 0], [0, 0, 0, 0], [0x9c6ca96a, 0xe17eae66, 0xfc10ecd4, 0x5256a7d8]),
    ([0xffffffff] * 4, [0xffffffff] * 4, [0x2a881696, 0x57012287, 0xf6c7446e, 0xa16a6732]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0, 0x082efa98, 0xec4e6c89],
     [0x59cd1dbb, 0xb8879579, 0x86b5d00c, 0xac8b6d84]),
]


@pytest.mark.parametrize("ctr,key,out", THREEFRY_KAT)
def test_threefry_known_answers(ctr, key, out):
    got = O.threefry4x32_20(np.array([ctr], np.uint32), key)[0]
    assert [int(v) for v in got] == out


@needs_ref
def test_threefry_and_normals_match_reference_headers():
    """Bitwise Threefry and the thermostat normals against the reference's own
    random.h (vendored Random123) compiled into oracle/_ref."""
    lib = ct.CDLL(REF_LIB)
    lib.ref_threefry4x32_20.argtypes = [ct.c_void_p] * 3
    lib.ref_normal3.argtypes = [ct.c_uint32, ct.c_uint32, ct.c_int, ct.c_uint64, ct.c_void_p]
    rng = np.random.default_rng(11)
    ctr = rng.integers(0, 2 ** 32, size=(50, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=4, dtype=np.uint64).astype(np.uint32)
    want = np.zeros((50, 4), np.uint32)
    for i in range(50):
        c = np.ascontiguousarray(ctr[i])
        lib.ref_threefry4x32_20(c.ctypes.data, key.ctypes.data, want[i:].ctypes.data)
    np.testing.assert_array_equal(O.threefry4x32_20(ctr, key), want)
    for step in (0, 7, 2 ** 33 + 5):
        ref = np.zeros((300, 3), np.float32)
        lib.ref_normal3(3, 0, 300, step, ref.ctypes.data)
        got = O.thermostat_normals(3, 0, 300, step)
        # same bits into Box-Muller; glibc sinf/cosf/logf vs numpy's float32
        # routines may differ in the last place (parity unpinned at bit level)
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)
