"""Host-side tests of libupside.so and the Python host code (no GPU needed).

* the library loads and exports every entry point include/upside_engine.h
  declares;
* the clamped-spline helpers (host code in the reference too,
  src/engine_c_library.cpp:197-277) against the compiled reference and
  against closed forms;
* without a GPU the engine refuses to construct (there is no CPU fallback);
* the .up writer and HDF5 reader round-trip and are deterministic per seed.
"""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from conftest import PKG, REF_LIB, ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "upside_engine.h")
LIB = os.path.join(PKG, "lib", "libupside.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import __graft_entry__
        __graft_entry__.build()
    return ct.CDLL(LIB)


def test_header_declares_reference_abi():
    names = declared_functions()
    for n in ("construct_deriv_engine", "free_deriv_engine", "evaluate_energy", "evaluate_deriv",
              "set_param", "get_param", "get_param_deriv", "get_output_dims", "get_output", "get_sens",
              "get_value_by_name", "clamped_spline_solve", "clamped_spline_value",
              "get_clamped_value_and_deriv", "get_clamped_coeff_deriv"):
        assert n in names, n


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    """The device code object inside the library targets gfx950 only."""
    data = open(LIB, "rb").read()
    assert b"gfx950" in data
    assert b"gfx942" not in data and b"gfx90a" not in data


def _spline_fns(L):
    L.clamped_spline_solve.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
    L.clamped_spline_value.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p]
    L.get_clamped_value_and_deriv.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p]
    L.get_clamped_coeff_deriv.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_float]
    return L


def _solve(L, values):
    N = len(values) + 2
    c = np.zeros(N, np.float32)
    v = np.ascontiguousarray(values, np.float32)
    assert L.clamped_spline_solve(N, c.ctypes.data, v.ctypes.data) == 0
    return c


def _value_deriv(L, c, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(2 * len(x), np.float32)
    assert L.get_clamped_value_and_deriv(len(c), out.ctypes.data, c.ctypes.data, len(x), x.ctypes.data) == 0
    return out[0::2], out[1::2]


def test_clamped_spline_interpolates_and_is_flat_at_ends(lib):
    L = _spline_fns(lib)
    rng = np.random.default_rng(0)
    values = rng.normal(size=10)
    c = _solve(L, values)
    knots = np.arange(1, 11, dtype=np.float32)
    v, d = _value_deriv(L, c, knots[:-1])
    np.testing.assert_allclose(v, values[:-1], atol=2e-6)
    # zero derivative at the clamped ends; constant beyond them
    np.testing.assert_allclose(c[0], c[2], atol=1e-6)
    np.testing.assert_allclose(c[-1], c[-3], atol=1e-6)
    vout, dout = _value_deriv(L, c, np.array([-2.0, 0.5, 40.0], np.float32))
    assert dout.tolist() == [0.0, 0.0, 0.0]
    # clamped_spline_value agrees with the value half of the pair
    x = rng.uniform(0, 12, 37).astype(np.float32)
    res = np.zeros(37, np.float32)
    assert L.clamped_spline_value(len(c), res.ctypes.data, c.ctypes.data, 37, x.ctypes.data) == 0
    np.testing.assert_array_equal(res, _value_deriv(L, c, x)[0])
    # coefficient derivative: d value / d coeff is the B-spline weight row
    row = np.zeros(len(c), np.float32)
    assert L.get_clamped_coeff_deriv(len(c), row.ctypes.data, c.ctypes.data, ct.c_float(4.3)) == 0
    assert abs(row.sum() - 1.0) < 1e-6
    assert abs(float(row @ c) - _value_deriv(L, c, [4.3])[0][0]) < 1e-5


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (needs /root/reference)")
def test_clamped_spline_helpers_match_reference(lib):
    mine = _spline_fns(lib)
    ref = _spline_fns(ct.CDLL(REF_LIB))
    rng = np.random.default_rng(1)
    values = rng.normal(size=14)
    c1, c2 = _solve(mine, values), _solve(ref, values)
    np.testing.assert_allclose(c1, c2, rtol=0, atol=2e-6)
    x = rng.uniform(-1, 17, 64).astype(np.float32)
    v1, d1 = _value_deriv(mine, c2, x)
    v2, d2 = _value_deriv(ref, c2, x)
    np.testing.assert_allclose(v1, v2, atol=1e-5)
    np.testing.assert_allclose(d1, d2, atol=1e-5)
    for xv in (0.3, 1.0, 5.5, 14.9, 20.0):
        r1 = np.zeros(16, np.float32)
        r2 = np.zeros(16, np.float32)
        mine.get_clamped_coeff_deriv(16, r1.ctypes.data, c2.ctypes.data, ct.c_float(xv))
        ref.get_clamped_coeff_deriv(16, r2.ctypes.data, c2.ctypes.data, ct.c_float(xv))
        np.testing.assert_allclose(r1, r2, atol=1e-6)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU behaviour")
def test_engine_refuses_without_gpu(chain_file):
    """No CPU fallback: with no visible GPU, construction fails loudly."""
    import upside_engine as ue
    path = chain_file(10, 1)
    with pytest.raises(RuntimeError):
        ue.Upside(path, quiet=True)


def test_up_writer_round_trip_and_determinism(updir):
    import h5lite
    import upconfig
    a, b = str(updir / "det_a.up"), str(updir / "det_b.up")
    seq_a, pos_a = upconfig.make_chain(a, 30, seed=17)
    seq_b, pos_b = upconfig.make_chain(b, 30, seed=17)
    assert list(seq_a) == list(seq_b)
    np.testing.assert_array_equal(pos_a, pos_b)
    with h5lite.File(a) as fa, h5lite.File(b) as fb:
        names = fa.children("/input/potential")
        assert names == fb.children("/input/potential")
        for n in names:
            g = "/input/potential/" + n
            assert list(fa.attr(g, "arguments")) == list(fb.attr(g, "arguments"))
        np.testing.assert_array_equal(fa.read("/input/pos"), fb.read("/input/pos"))
        assert fa.read("/input/pos").shape == (90, 3, 1)


def test_random_initial_config_geometry():
    """Bond lengths, bond angles and trans omegas of the seeded random-Rama
    backbone built by TAB matrices (upside_config.py:140-201): the chain grows
    from the origin by bonds 1.453, 1.526, 1.300 A with angles 120, 120, 109.5."""
    import upconfig
    pos = upconfig.random_initial_config(25, np.random.default_rng(3))
    full = np.concatenate([np.zeros((1, 3)), pos])
    b = full[1:] - full[:-1]
    d = np.linalg.norm(b, axis=1)
    np.testing.assert_allclose(d[0::3], 1.453, atol=1e-6)
    np.testing.assert_allclose(d[1::3], 1.526, atol=1e-6)
    np.testing.assert_allclose(d[2::3], 1.300, atol=1e-6)
    cosang = -(b[:-1] * b[1:]).sum(1) / (d[:-1] * d[1:])
    ang = np.degrees(np.arccos(cosang))
    np.testing.assert_allclose(sorted(set(np.round(ang, 6))), [109.5, 120.0], atol=1e-6)


@pytest.mark.skipif(not os.path.exists("/root/reference/example/15.SpatialTransformation/inputs/chig.up"),
                    reason="reference example input not present")
def test_config_writer_matches_reference_chig_up(chig):
    """Every /input/potential dataset and attribute of the writer's chignolin
    equals the reference-written example input (same sequence, pos, rama)."""
    import h5lite
    path, _ = chig
    ref = "/root/reference/example/15.SpatialTransformation/inputs/chig.up"
    with h5lite.File(ref) as fr, h5lite.File(path) as fm:
        rn = fr.children("/input/potential")
        mn = fm.children("/input/potential")
        assert sorted(rn) == sorted(mn)

        def walk(f, g, out):
            for c in f.children(g):
                p = g + "/" + c
                if f.is_group(p):
                    walk(f, p, out)
                else:
                    out.append(p)
        dsets = []
        walk(fr, "/input/potential", dsets)
        for p in dsets:
            a, b = fr.read(p), fm.read(p)
            assert a.shape == b.shape, p
            if a.dtype.kind in "fc":
                np.testing.assert_allclose(b, a, rtol=2e-6, atol=2e-6, err_msg=p)
            else:
                np.testing.assert_array_equal(b, a, err_msg=p)


# ---------------------------------------------------- per-system team plans
N_XCD = 8


def _team_map(lib, teams, per_xcd):
    t = np.ascontiguousarray(teams, dtype=np.intc)
    fn = lib.upside_plan_team_map
    fn.restype = ct.c_int
    n = fn(t.ctypes.data_as(ct.c_void_p), ct.c_int(len(t)), ct.c_int(per_xcd), None, ct.c_int(0))
    if n < 0:
        return None
    m = np.zeros(n, dtype=np.intc)
    assert fn(t.ctypes.data_as(ct.c_void_p), ct.c_int(len(t)), ct.c_int(per_xcd),
              m.ctypes.data_as(ct.c_void_p), ct.c_int(n)) == n
    if n > 0:  # a buffer one short is refused
        assert fn(t.ctypes.data_as(ct.c_void_p), ct.c_int(len(t)), ct.c_int(per_xcd),
                  m.ctypes.data_as(ct.c_void_p), ct.c_int(n - 1)) == -1
    return m


def _check_map(m, teams, per_xcd):
    """Every system's ranks 0..G-1 appear once, all on one XCD (b % 8), every
    XCD holds at most per_xcd workgroups, and the rest is padding."""
    seen = {}
    for b, e in enumerate(m):
        if e < 0:
            assert e == -1
            continue
        s, g, r = e >> 14, (e >> 7) & 127, e & 127
        assert g == teams[s] and 0 <= r < g
        assert (s, r) not in seen
        seen[(s, r)] = b % N_XCD
    assert len(seen) == sum(teams)
    for s, g in enumerate(teams):
        assert len({seen[(s, r)] for r in range(g)}) == 1, s
    load = np.bincount([x for x in seen.values()], minlength=N_XCD)
    assert load.max() <= per_xcd
    assert len(m) == N_XCD * load.max()


def test_team_map_places_every_team_on_one_xcd(lib):
    rng = np.random.default_rng(4)
    for trial in range(200):
        n = int(rng.integers(1, 80))
        teams = [int(x) for x in rng.integers(1, 9, n)]
        m = _team_map(lib, teams, 32)
        if m is None:  # must really not fit: more than the device or a failed greedy near the limit
            assert sum(teams) > 32 * N_XCD - 7 * N_XCD
            continue
        _check_map(m, teams, 32)
    # uniform teams fill the XCDs evenly
    m = _team_map(lib, [4] * 64, 32)
    _check_map(m, [4] * 64, 32)
    assert len(m) == 256 and np.all(m >= 0)
    # a team larger than an XCD, or more workgroups than the device, does not fit
    assert _team_map(lib, [33], 32) is None
    assert _team_map(lib, [32] * 9, 32) is None
    assert _team_map(lib, [0, 1], 32) is None


def test_balanced_teams_follow_the_atom_counts(lib):
    fn = lib.upside_plan_balanced_teams
    fn.restype = ct.c_int

    def plan(atoms, budget, per_xcd=32):
        a = np.ascontiguousarray(atoms, dtype=np.intc)
        t = np.zeros(len(a), dtype=np.intc)
        rc = fn(a.ctypes.data_as(ct.c_void_p), ct.c_int(len(a)), ct.c_int(budget), ct.c_int(per_xcd),
                t.ctypes.data_as(ct.c_void_p))
        return None if rc else t

    rng = np.random.default_rng(3)
    atoms = 3 * rng.integers(50, 401, 64)
    t = plan(atoms, 256)
    assert t is not None and t.sum() <= 256 and t.min() >= 1
    order = np.argsort(atoms, kind="stable")
    assert np.all(np.diff(t[order]) >= 0)            # more atoms, no smaller team
    assert (atoms / t).max() <= 2.0 * (atoms / t).min() + 1e-9 or t.min() == 1
    _check_map(_team_map(lib, t, 32), list(t), 32)
    # equal systems get equal teams; one system gets one XCD
    assert np.all(plan([450] * 16, 256) == 16)
    assert list(plan([1500], 256)) == [32]
    # the budget bounds the total; too many systems is an error
    assert plan(atoms, 100).sum() <= 100
    assert plan([10] * 300, 256) is None


# ------------------------------------------------ lowered program invariants
@pytest.mark.parametrize("lanes", ["4", "2", "0"])
def test_lowered_program_invariants(lib, tmp_path, monkeypatch, lanes):
    """The host lowering of the node graph (value-pass order, lane sections,
    the rotamer's overlap range and early deriv ops) keeps the invariants the
    kernel relies on, for chains of several sizes, the chignolin fixture, and
    with sections of four lanes, of two, or none (upside_check_program runs
    on the host only)."""
    import upconfig
    monkeypatch.setenv("UPSIDE_LANES", lanes)
    paths = []
    for L in (20, 150, 500):
        p = str(tmp_path / ("c%d.up" % L))
        upconfig.make_chain(p, L, seed=5)
        paths.append(p)
    z = np.load(os.path.join(ROOT, "tests", "golden", "chig_frames.npz"))
    g = upconfig.build_ff21([s.decode() for s in z["sequence"]], z["initial_pos"], rama_pot=z["rama_pot"],
                            normalize_rama=False)
    chig = str(tmp_path / "chig.up")
    upconfig.write_up(chig, g)
    paths.append(chig)
    fn = lib.upside_check_program
    fn.restype = ct.c_int
    for p in paths:
        buf = ct.create_string_buffer(1024)
        rc = fn(p.encode(), buf, len(buf))
        msg = buf.value.decode()
        assert rc == 0, (p, msg)
        n_sec = int(re.search(r"(\d+) sections", msg).group(1))
        if lanes == "0":
            assert n_sec == 0, msg
        else:
            assert n_sec >= 1 and ("of %s lanes" % lanes) in msg, msg
    buf = ct.create_string_buffer(256)
    assert fn(str(tmp_path / "missing.up").encode(), buf, len(buf)) == -1
