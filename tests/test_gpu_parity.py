"""Parity of the HIP engine with the oracle, through the C ABI (needs a GPU).

Bars (written per test): total energy within 1e-4 relative (|dE|/max(|E|,1):
ff_2.1 totals cancel to near zero on folded frames) and dE/dx within 1e-4
relative RMS (||F-F_ref||/||F_ref||) of the float32 oracle, of the FP64
oracle and, where oracle/_ref is present, of the compiled reference (against
FP64 and the reference: 1e-4 or 1.25 times the reference's own distance from
FP64, whichever is larger, since float32 rounding grows with chain length);
pair lists and BP edge sets bit-exact; integer statistics (edge counts by
type, BP iterations) equal.
"""
import os

import numpy as np
import pytest

from conftest import REF_LIB, RefEngine, read_pos, ref_lib, rel_rms, say

import upside_oracle as O

pytestmark = pytest.mark.gpu

E_TOL = 1e-4
F_TOL = 1e-4

PAIR_NODES = ["rotamer", "protein_hbond", "hbond_coverage", "hbond_coverage_hydrophobe",
              "environment_coverage_sc", "environment_coverage_hb", "hb_environment_coverage_hn",
              "hb_environment_coverage_oc", "backbone_pairs"]


@pytest.fixture(scope="module")
def ue():
    import upside_engine
    upside_engine.load_upside_library()
    return upside_engine


def oracle_edges(eng, name):
    node = eng.nodes[name]
    obj = node.ig if hasattr(node, "ig") else node
    return set(zip(obj.i1.tolist(), obj.i2.tolist()))


def oracle_edges_at(ue_eng, ora, name):
    """The oracle's edge set of a pair graph computed from the engine's own
    coordinates of the graph's producers (its last evaluation): the refine is
    then compared on bit-identical inputs, whatever rounding the upstream
    nodes of the two engines differ by."""
    node = ora.nodes[name]
    ig = node.ig if hasattr(node, "ig") else None
    if ig is None:   # backbone_pairs: residue pairs of the alignment frames
        aff = ora.nodes[node.args[0]]
        saved = aff.output.copy()
        aff.output[:] = ue_eng.get_output(node.args[0])
        try:
            t = aff.output[node.res, :3]
            d2 = O.refine_dist2(t[:, None] - t[None, :], ora.T)
            i1, i2 = np.nonzero(np.triu((d2 < node.dist_cutoff * node.dist_cutoff)
                                        & (np.abs(node.res[:, None] - node.res[None, :]) > 1), 1))
        finally:
            aff.output[:] = saved
        return set(zip(i1.tolist(), i2.tolist()))
    prods = {ig.node1.name: ig.node1, ig.node2.name: ig.node2}
    saved = {k: n.output.copy() for k, n in prods.items()}
    for k, n in prods.items():
        n.output[:] = ue_eng.get_output(k)
    try:
        i1, i2 = ig.edges()
    finally:
        for k, n in prods.items():
            n.output[:] = saved[k]
        ig.edges()
    return set(zip(i1.tolist(), i2.tolist()))


def dump_case(tag, **arrays):
    """Save the inputs of a failing comparison for study on the CPU (into the
    folder UPSIDE_DUMP_DIR names, when it is set and exists)."""
    d = os.environ.get("UPSIDE_DUMP_DIR", "")
    if d and os.path.isdir(d):
        np.savez(os.path.join(d, "case_" + "".join(ch if ch.isalnum() else "_" for ch in tag) + ".npz"), **arrays)


def check_point(ue_eng, ora, ora64, pos, label, ref=None):
    e, d = float(ue_eng.energy(pos)), ue_eng.deriv(pos)
    e32, d32 = float(ora.energy(pos)), ora.deriv(pos)
    e64, d64 = float(ora64.energy(pos.astype(np.float64))), ora64.deriv(pos.astype(np.float64))
    err = dict(e32=abs(e - e32) / max(abs(e32), 1.0), f32=rel_rms(d, d32),
               e64=abs(e - e64) / max(abs(e64), 1.0), f64=rel_rms(d, d64))
    # float32 rounding grows with the chain: where the compiled reference is
    # itself further than F_TOL from float64 (about 2e-4 at 1000 residues),
    # the bar against float64 and the reference is 1.25 times its own distance
    f_wide = F_TOL
    if ref is not None:
        er, dr = ref.energy(pos), ref.deriv(pos)
        err.update(eref=abs(e - er) / max(abs(er), 1.0), fref=rel_rms(d, dr))
        f_wide = max(F_TOL, 1.25 * rel_rms(dr, d64))
    say(label, " ".join("%s=%.2e" % kv for kv in sorted(err.items())), "(force bar vs f64/ref %.2e)" % f_wide)
    for k, v in err.items():
        bar = E_TOL if k.startswith("e") else (F_TOL if k == "f32" else f_wide)
        assert v < bar, (label, k, v, bar)
    return e, d


def test_chig_frames(ue, chig):
    """The 20 frames of the reference's chignolin run: potential against the
    reference-logged values and the oracle, forces against the oracle."""
    path, z = chig
    say("case: chignolin golden frames")
    eng = ue.Upside(path, quiet=True)
    ora, ora64 = O.Engine(path), O.Engine(path, dtype=np.float64)
    got = []
    # per-residue rotamer values of the reference's detailed log
    # (src/rotamer.cpp:708-731): 2e-3 absolute, the BP-dependent bar of the
    # logged potential
    logged = ["rotamer_free_energy", "rotamer_potential", "rotamer_entropy"]
    res_dev = {k: 0.0 for k in logged + ["rotamer_1body_energy"]}
    for f in range(len(z["frame_pos"])):
        pos = np.ascontiguousarray(z["frame_pos"][f], np.float32)
        e, _ = check_point(eng, ora, ora64, pos, "chig frame %d" % f)
        got.append(e)
        nn = z["frame_rotamer_free_energy"].shape[1]
        for name in logged:
            v = eng.get_value_by_name((nn,), "rotamer", name)
            res_dev[name] = max(res_dev[name], float(np.abs(v - z["frame_" + name][f]).max()))
        one = eng.get_value_by_name((nn, 3), "rotamer", "rotamer_1body_energy")
        for q in range(3):
            res_dev["rotamer_1body_energy"] = max(res_dev["rotamer_1body_energy"],
                                                  float(np.abs(one[:, q] - z["frame_rotamer_1body_energy%d" % q][f]).max()))
    dev = np.abs(np.array(got) - z["frame_potential"].ravel())
    say("chig max |E - E_logged| = %.2e" % dev.max())
    say("chig per-residue max |value - logged|:", " ".join("%s=%.2e" % kv for kv in sorted(res_dev.items())))
    assert dev.max() < 2e-3
    for name, d in res_dev.items():
        assert d < 2e-3, (name, d)


@pytest.mark.parametrize("level", [1, 2])
def test_detailed_and_extensive_node_loggers(ue, chig, tmp_path, level):
    """Frames logged at the chignolin golden positions: rama_map_potential and
    sigmoid_coupling (detailed level, src/rama_map_pot.cpp:49-55,
    src/environment.cpp:892-900) against the values the reference logged for
    those frames (1e-4 absolute on values of order 1-10); at the extensive
    level environment_coverage, hbond_environment_coverage and virtual
    (src/environment.cpp:304-310, :355-360, src/hbond.cpp:53-58) equal to
    the node outputs of an evaluation at the same positions."""
    import h5lite
    path, z = chig
    frames = [0, 9, 19]
    say("case: node loggers at level %d on chignolin frames" % level)
    b = ue.Batch([path], [0.8], seed=1)
    out = str(tmp_path / ("log%d.up" % level))
    b.open_output([out], log_level=level)
    for f in frames:
        b.set_pos(0, np.ascontiguousarray(z["frame_pos"][f], np.float32))
        b.run_logged(1, 1)
    b.close_output()
    with h5lite.File(out) as h:
        names = h.children("/output")
        got = {k: h.read("/output/" + k) for k in ("rama_map_potential", "sigmoid_coupling")}
        ext = {k: h.read("/output/" + k) for k in ("environment_coverage", "hbond_environment_coverage", "virtual")
               if k in names}
    for k, v in got.items():
        want = z["frame_" + k][frames]
        say("%s max |logged - reference| %.2e" % (k, np.abs(v - want).max()))
        assert v.shape == want.shape and np.abs(v - want).max() < 1e-4, k
    if level < 2:
        assert not ext
        return
    assert set(ext) == {"environment_coverage", "hbond_environment_coverage", "virtual"}
    eng = ue.Upside(path, quiet=True)
    with h5lite.File(path) as h:
        nodes = h.children("/input/potential")
    env = [n for n in nodes if n.startswith("environment_coverage")][0]
    hbe = [n for n in nodes if n.startswith("hb_environment_coverage")][0]
    for k, f in enumerate(frames):
        eng.energy(np.ascontiguousarray(z["frame_pos"][f], np.float32))
        np.testing.assert_allclose(ext["environment_coverage"][k], eng.get_output(env).ravel(), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(ext["hbond_environment_coverage"][k], eng.get_output(hbe).ravel(), rtol=1e-6,
                                   atol=1e-7)
        np.testing.assert_allclose(ext["virtual"][k], eng.get_output("infer_H_O")[:, :3], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n_res,seed", [(40, 1), (100, 2), (500, 3), (1000, 4)])
def test_random_chain_energy_forces(ue, chain_file, n_res, seed):
    path = chain_file(n_res, seed)
    say("case: random chain L=%d seed=%d" % (n_res, seed))
    pos0 = read_pos(path)
    eng = ue.Upside(path, quiet=True)
    ora, ora64 = O.Engine(path), O.Engine(path, dtype=np.float64)
    ref = RefEngine(ref_lib(), path, len(pos0)) if os.path.exists(REF_LIB) else None
    check_point(eng, ora, ora64, pos0, "L=%d start" % n_res, ref)
    rng = np.random.default_rng(seed)
    for k in range(2):
        pos = (pos0 + rng.normal(scale=0.2, size=pos0.shape)).astype(np.float32)
        check_point(eng, ora, ora64, pos, "L=%d jitter %d" % (n_res, k), ref)


@pytest.mark.parametrize("n_res,seed", [(100, 2), (500, 3)])
def test_pair_lists_bit_exact(ue, chain_file, n_res, seed):
    """Every interaction graph's edge set equals the oracle's brute force on
    the same coordinates, at the start and after the Verlet caches have aged
    through a short trajectory."""
    path = chain_file(n_res, seed)
    pos0 = read_pos(path)
    eng = ue.Upside(path, quiet=True)
    ora = O.Engine(path)
    b = ue.Batch([path], [0.85], seed=seed)
    say("case: pair lists L=%d: run 20 rounds" % n_res)
    b.run(20)
    pos_run = b.pos(0)
    for label, pos in (("start", pos0), ("after 20 rounds", pos_run)):
        eng.energy(pos)
        ora.energy(pos)
        for name in PAIR_NODES:
            got = set(map(tuple, eng.edge_pairs(name).tolist()))
            want = oracle_edges(ora, name)
            say("L=%d %s %-28s edges %6d  xor %d" % (n_res, label, name, len(want), len(got ^ want)))
            assert got == want, (name, label, sorted(got ^ want)[:10])


def test_coordinate_node_outputs(ue, chain_file):
    """Outputs and sensitivities of the coordinate nodes (get_output/get_sens)."""
    path = chain_file(60, 7)
    pos = read_pos(path)
    pos = pos + np.random.default_rng(0).normal(scale=0.1, size=pos.shape).astype(np.float32)
    say("case: coordinate node outputs L=60")
    eng = ue.Upside(path, quiet=True)
    ora = O.Engine(path)
    eng.deriv(pos)
    ora.deriv(pos)
    for name in ora.order:
        node = ora.nodes[name]
        if node.potential_term:
            got = float(eng.get_output(name)[0, 0])
            assert abs(got - float(node.potential)) <= 1e-5 * max(1.0, abs(float(node.potential))), name
            continue
        out = eng.get_output(name)
        want = node.output[:, :out.shape[1]]
        scale = max(1.0, float(np.abs(want).max()))
        assert np.abs(out - want).max() <= 2e-5 * scale, (name, np.abs(out - want).max())
        sens = eng.get_sens(name)
        ws = node.sens[:, :sens.shape[1]]
        sscale = max(1.0, float(np.abs(ws).max()))
        assert np.abs(sens - ws).max() <= 1e-4 * sscale, (name, np.abs(sens - ws).max())


def test_rotamer_bp_statistics(ue, chain_file):
    """BP iteration count, edge counts by type and n_bad_solve equal the
    oracle's on the same coordinates."""
    path = chain_file(100, 2)
    pos = read_pos(path)
    say("case: rotamer BP statistics L=100")
    eng = ue.Upside(path, quiet=True)
    ora = O.Engine(path)
    rng = np.random.default_rng(5)
    for k in range(3):
        p = (pos + rng.normal(scale=0.15, size=pos.shape)).astype(np.float32)
        eng.energy(p)
        ora.energy(p)
        it = int(eng.get_value_by_name((1,), "rotamer", "iterations")[0])
        rot = ora.nodes["rotamer"]
        say("jitter %d: BP iterations gpu %d oracle %d" % (k, it, rot.iterations))
        assert abs(it - rot.iterations) <= 2 * rot.chunk
        ig = rot.ig
        cnt = np.zeros((ig.type1.max() + 1, ig.n_type2), np.float32)
        np.add.at(cnt, (ig.type1[ig.i1], ig.type2[ig.i2]), 1)
        got = eng.get_value_by_name(cnt.shape, "rotamer", "count_edges_by_type")
        np.testing.assert_array_equal(got, cnt)
    assert eng.get_value_by_name((1,), "rotamer", "read n_bad_solve and reset")[0] == 0


def test_rotamer_values_match_reference(ue, chain_file):
    """The rotamer node's value hooks (src/rotamer.cpp:736-834) equal the
    compiled reference's at the same coordinates: per-residue free energies
    and 1-body energies, node energies, edge energies and edge marginals in
    the reference's graph order.  Tolerances: 2e-4 relative to the largest
    magnitude for values fixed before BP (node and edge energies, float32
    exp/log of the same sums); 2e-3 absolute for BP-dependent values (BP stops
    at a belief change of 1e-3, and a float32 rounding difference can end it
    one chunk earlier or later)."""
    import h5lite
    lib = ref_lib()
    if lib is None:
        pytest.skip("reference engine not built")
    path = chain_file(50, 21)
    pos0 = read_pos(path)
    say("case: rotamer value hooks L=50")
    eng = ue.Upside(path, quiet=True)
    ref = RefEngine(lib, path, len(pos0))
    rng = np.random.default_rng(8)
    for k in range(2):
        pos = (pos0 + rng.normal(scale=0.1, size=pos0.shape)).astype(np.float32)
        eng.energy(pos)
        ref.energy(pos)
        nn = int(ref.value((1,), "rotamer", "n_node")[0])
        assert int(eng.get_value_by_name((1,), "rotamer", "n_node")[0]) == nn
        with h5lite.File(path) as f:
            n_prob = len(f.attr("/input/potential/rotamer", "arguments")) - 1
        for name, shape, bp in (("rotamer_free_energy", (nn,), True),
                                ("rotamer_1body_energy", (nn, n_prob), True),
                                ("node_energy", (nn, 6), False),
                                ("edge_energy", (nn, nn, 6, 6), False),
                                ("edge_marginal_in_graph_order", (nn, nn, 6, 6), True)):
            got = eng.get_value_by_name(shape, "rotamer", name)
            want = ref.value(shape, "rotamer", name)
            if name == "node_energy":   # 1e5 marks states past a residue's rotamer count
                pad = want == np.float32(1e5)
                assert pad.sum() > 0 and np.array_equal(got[pad], want[pad])
                got, want = got[~pad], want[~pad]
            err = float(np.abs(got - want).max())
            scale = max(1.0, float(np.abs(want).max()))
            say("jitter %d %s: max abs diff %.3g (max |value| %.3g)" % (k, name, err, scale))
            if bp:
                assert err <= 2e-3, (name, err)
            else:
                assert err <= 2e-4 * scale, (name, err)
            if name == "rotamer_free_energy":   # the split sums to the node's potential
                total = float(eng.get_output("rotamer").reshape(-1)[0])
                assert abs(float(got.sum()) - total) <= 1e-4 * max(1.0, abs(total)), (got.sum(), total)
            if name == "edge_marginal_in_graph_order":   # each residue pair's block is a distribution
                blk = got.sum(axis=(2, 3))
                assert np.abs(blk - 1).max() <= 1e-4, np.abs(blk - 1).max()


def test_batch_matches_single_engine(ue, chain_file):
    """Systems of uneven length in one launch: each system's potential equals
    the single-engine value at its positions."""
    paths = [chain_file(n, s) for n, s in ((30, 11), (90, 12), (150, 13), (90, 12))]
    say("case: uneven batch of 4")
    b = ue.Batch(paths, [0.8, 0.85, 0.9, 0.95], seed=3)
    b.run(5)
    e_b = b.energies()
    for s, p in enumerate(paths):
        eng = ue.Upside(p, quiet=True)
        e = eng.energy(b.pos(s))
        assert abs(e - e_b[s]) <= 1e-6 * abs(e), (s, e, e_b[s])
    # the two copies of one input file started identically but use different seeds
    assert not np.array_equal(b.pos(1), b.pos(3))
    # the per-solve iteration histogram agrees with the running totals
    for s in range(len(paths)):
        st, h = b.bp_stats(s), b.bp_histogram(s)
        assert h[-1] == 0
        assert h.sum() == st["solves"] and (h * np.arange(len(h))).sum() == st["total"], (s, st)
        assert h[st["last"]] >= 1


def test_trajectory_matches_oracle_integrator(ue, chain_file):
    """Thermalisation plus 4 rounds (12 force evaluations and integrator
    stages) against the oracle's restatement of the simulation loop with the
    same Threefry noise."""
    path = chain_file(40, 21)
    say("case: trajectory L=40, 4 rounds")
    pos0 = read_pos(path)
    b = ue.Batch([path], [0.85], seed=7, dt=0.009)
    mom_gpu0 = b.mom(0)
    ora = O.Engine(path)
    mom0, n_inv = O.thermalize(len(pos0), 0.85, 7, 0.009)
    np.testing.assert_allclose(mom_gpu0, mom0, rtol=2e-5, atol=2e-6)
    b.run(4)
    pos, mom, _ = O.run_rounds(ora, pos0, mom0, 4, 0.009, 0.85, 7, n_inv)
    say("trajectory pos rel %.2e mom rel %.2e" % (rel_rms(b.pos(0), pos), rel_rms(b.mom(0), mom)))
    assert rel_rms(b.pos(0), pos) < 1e-5
    assert rel_rms(b.mom(0), mom) < 1e-3


def test_replica_exchange(ue, chain_file):
    """Swap decisions follow the Metropolis test on the batch energies with the
    uniform numbers of Threefry stream 1 (src/main.cpp:243-316)."""
    path = chain_file(40, 31)
    temps = list(0.80 * (0.95 / 0.80) ** (np.arange(8) / 7))
    say("case: replica exchange 8 replicas")
    b = ue.Batch([path] * 8, temps, seed=5)
    b.run(6)
    pairs = np.array([[0, 1], [2, 3], [4, 5], [6, 7]])
    pos_before = [b.pos(s) for s in range(8)]
    e_before = b.energies()
    acc = b.replica_exchange(pairs, 3)
    beta = 1 / np.array(temps, np.float32)
    key = [5, 1, 0, 0]
    ctr = 0
    for (s1, s2), a in zip(pairs, acc):
        diff = np.float32((-beta[s1] * e_before[s2] - beta[s2] * e_before[s1])
                          - (-beta[s1] * e_before[s1] - beta[s2] * e_before[s2]))
        want = 1
        if diff < 0:
            bits = O.threefry4x32_20(np.array([[3, 0, 0, ctr]], np.uint32), key)[0]
            ctr += 1
            u = np.float32(bits[0]) * np.float32(2.0 ** -32) + np.float32(2.0 ** -33)
            want = 0 if np.exp(diff) < u else 1
        assert a == want, (s1, s2, diff)
        np.testing.assert_array_equal(b.pos(s1), pos_before[s2] if a else pos_before[s1])
        np.testing.assert_array_equal(b.pos(s2), pos_before[s1] if a else pos_before[s2])
    # exchange criterion 1 (src/main.cpp:298-311): test on the current weights only
    b.run(3)
    pos_before = [b.pos(s) for s in range(8)]
    e_before = b.energies()
    pairs = np.array([[1, 2], [3, 4], [5, 6]])
    acc = b.replica_exchange(pairs, 4, criterion=1)
    ctr = 0
    n_acc = 0
    for (s1, s2), a in zip(pairs, acc):
        diff = np.float32(-beta[s2] * e_before[s2]) - np.float32(-beta[s1] * e_before[s1])
        want = 1
        if diff < 0:
            bits = O.threefry4x32_20(np.array([[4, 0, 0, ctr]], np.uint32), key)[0]
            ctr += 1
            u = np.float32(bits[0]) * np.float32(2.0 ** -32) + np.float32(2.0 ** -33)
            want = 1 if np.exp(diff) >= u else 0
        assert a == want, (s1, s2, diff)
        n_acc += a
        np.testing.assert_array_equal(b.pos(s1), pos_before[s2] if a else pos_before[s1])
    say("criterion 1: accepted %d of %d" % (n_acc, len(pairs)))
    with pytest.raises(RuntimeError):
        b.replica_exchange(pairs, 5, criterion=2)


def test_thermostat_temperature(ue, chig):
    """Equipartition: from an equilibrated chignolin frame of the reference's
    T=0.8 run with Maxwell-Boltzmann momenta, the kinetic energy per degree of
    freedom averages T/2 (32 replicas each at T=0.7 and 0.9, own thermostat
    seeds).  The fresh-start thermalisation alone is deliberately weak (one OU
    step of 0.027 time units gives ~1% of kT, src/main.cpp:800-803), so the
    momenta are drawn here instead.  Samples 10 rounds apart are correlated
    over the thermostat time (~185 rounds): 64 replicas and 100 samples keep
    the statistical error near 1.3%, a third of the 4% bar."""
    path, z = chig
    temps = [0.7] * 32 + [0.9] * 32
    say("case: thermostat equipartition, 64 chignolin replicas, 1300 rounds")
    b = ue.Batch([path] * 64, temps, seed=9)
    rng = np.random.default_rng(0)
    for s in range(64):
        b.set_pos(s, np.ascontiguousarray(z["frame_pos"][19], np.float32))
        b.set_mom(s, (np.sqrt(temps[s]) * rng.normal(size=(30, 3))).astype(np.float32))
    b.run(300)
    ke = np.zeros(64)
    n = 0
    for _ in range(100):
        b.run(10)
        for s in range(64):
            ke[s] += 0.5 * (b.mom(s) ** 2).sum() / (3 * b.n_atom[s])
        n += 1
    ke /= n
    got = [ke[:32].mean(), ke[32:].mean()]
    say("mean KE/dof", got, "targets", [0.35, 0.45])
    np.testing.assert_allclose(got, [0.35, 0.45], rtol=0.04)


def test_error_behaviour(ue, chain_file, updir):
    """Reference error conventions: NULL engine on a bad file, nonzero return
    codes on bad names and sizes, raised as RuntimeError by the binding."""
    say("case: error behaviour")
    calc = ue.load_upside_library()
    assert not calc.construct_deriv_engine(10, str(updir / "does_not_exist.up").encode(), True)
    with pytest.raises((RuntimeError, OSError)):  # the binding reads /input/pos first
        ue.Upside(str(updir / "does_not_exist.up"), quiet=True)
    path = chain_file(20, 51)
    eng = ue.Upside(path, quiet=True)
    with pytest.raises(RuntimeError):
        eng.get_output("no_such_node")
    with pytest.raises(RuntimeError):
        eng.get_value_by_name((3,), "rotamer", "n_node")
    with pytest.raises(RuntimeError):
        eng.get_param_deriv((4,), "rotamer")
    e = eng.energy(read_pos(path))
    assert np.isfinite(e)


def test_capacity_regrowth_is_transparent(ue, chain_file, monkeypatch):
    """Neighbour-list capacities that start far too small overflow, regrow and
    repeat the run: results are bit-identical to a run that never overflowed
    (slot order is canonical, independent of capacity)."""
    path = chain_file(100, 61)
    pos = read_pos(path)
    say("case: capacity regrowth L=100")
    ref = ue.Upside(path, quiet=True)
    e0, d0 = ref.energy(pos), ref.deriv(pos)
    b0 = ue.Batch([path] * 2, [0.85, 0.9], seed=4)
    b0.run(5)
    monkeypatch.setenv("UPSIDE_CAPACITY_SCALE", "0.02")
    small = ue.Upside(path, quiet=True)
    assert small.energy(pos) == e0
    np.testing.assert_array_equal(small.deriv(pos), d0)
    b1 = ue.Batch([path] * 2, [0.85, 0.9], seed=4)
    early = b1.save_state()       # saved at the small capacities, before any overflow
    b1.run(5)
    for s in range(2):
        np.testing.assert_array_equal(b1.pos(s), b0.pos(s))
    # a batch that grew rolls back to its own earlier (smaller) checkpoint and
    # repeats the run bit for bit
    b1.load_state(early)
    b1.run(5)
    for s in range(2):
        np.testing.assert_array_equal(b1.pos(s), b0.pos(s))


@pytest.mark.parametrize("name", ["rotamer", "hbond_coverage", "hbond_coverage_hydrophobe",
                                  "environment_coverage_sc", "protein_hbond"])
def test_param_deriv_matches_reference(ue, chain_file, name):
    """get_param_deriv of the spline pair graphs against the reference engine
    built from its sources with PARAM_DERIV (oracle/_ref), on the same
    jittered coordinates: relative norm error within 1e-4, and the same
    result on a second call."""
    import ctypes as ct
    import h5lite
    lib = ref_lib()
    if lib is None:
        pytest.skip("reference engine not built")
    lib.get_param_deriv.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p]
    path = chain_file(60, 7)
    pos0 = read_pos(path)
    pos = (pos0 + np.random.default_rng(3).normal(scale=0.1, size=pos0.shape)).astype(np.float32)
    with h5lite.File(path) as f:
        grp = "/input/potential/rotamer/pair_interaction" if name == "rotamer" else "/input/potential/" + name
        shape = tuple(f.read(grp + "/interaction_param").shape)
    say("case: param deriv %s, table %s" % (name, shape))
    eng = ue.Upside(path, quiet=True)
    ref = RefEngine(lib, path, len(pos0))
    eng.deriv(pos)
    ref.deriv(pos)
    want = np.zeros(shape, np.float32)
    rc = lib.get_param_deriv(want.size, want.ctypes.data, ct.c_void_p(ref.h), name.encode())
    if name == "protein_hbond":          # no derivative hook in the reference: an error in both
        assert rc != 0
        with pytest.raises(RuntimeError):
            eng.get_param_deriv(shape, name)
        return
    assert rc == 0
    got = eng.get_param_deriv(shape, name)
    err = float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-30))
    say("%s: |dE/dparam| %.4g, rel err %.2e, nonzero %d of %d" % (name, np.linalg.norm(want), err,
                                                                   np.count_nonzero(want), want.size))
    if name.startswith("environment"):   # the reference has no derivative for these: zeros
        np.testing.assert_array_equal(got, want)
        return
    assert np.linalg.norm(want) > 0
    assert err < 1e-4, err
    np.testing.assert_array_equal(eng.get_param_deriv(shape, name), got)
    # set_param: the same perturbed table in both engines gives the same
    # energy and forces, and get_param returns it
    lib.set_param.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p]
    table = eng.get_param(shape, name)
    noise = np.random.default_rng(4).standard_normal(shape)
    if name == "rotamer":
        # a symmetric table must stay compatible with its transpose (angular
        # halves swapped, radial equal, src/bead_interaction.h:211-220): change
        # the radial splines by a symmetric factor only
        noise = 0.5 * (noise + noise.transpose(1, 0, 2))
        noise[:, :, :30] = 0.0
    new = (table * (1.0 + 0.05 * noise)).astype(np.float32)
    eng.set_param(new, name)
    assert lib.set_param(new.size, new.ctypes.data, ct.c_void_p(ref.h), name.encode()) == 0
    np.testing.assert_array_equal(eng.get_param(shape, name), new)
    e, d = eng.energy(pos), eng.deriv(pos)
    er, dr = ref.energy(pos), ref.deriv(pos)
    say("%s after set_param: E %.5f ref %.5f, force rel %.2e" % (name, e, er, rel_rms(d, dr)))
    assert abs(e - er) / max(abs(er), 1.0) < E_TOL and rel_rms(d, dr) < F_TOL
    with pytest.raises(RuntimeError):
        eng.set_param(new.ravel()[:-1], name)
    if name == "rotamer":
        bad = new.copy()
        bad[0, 1, 40] += 1.0          # radial part no longer equal to its transpose
        with pytest.raises(RuntimeError):
            eng.set_param(bad, name)
        np.testing.assert_array_equal(eng.get_param(shape, name), new)


def test_mc_pivot_matches_reference_cli(ue, chain_file, tmp_path):
    """Pivot moves every 2 rounds of a 6-round run against the reference CLI
    built from its sources (--monte-carlo-interval, detailed log level): the
    same moves are accepted (pivot_stats), every logged frame (positions,
    potential) matches within 1e-4 relative, and the detailed hbond and
    bad-solve datasets agree."""
    import shutil
    import subprocess
    import h5lite
    from conftest import REF_BIN
    if not os.path.exists(REF_BIN):
        pytest.skip("reference CLI not built")
    src = chain_file(40, 21)
    ref_up, our_up = str(tmp_path / "ref.up"), str(tmp_path / "ours.up")
    shutil.copy(src, ref_up)
    shutil.copy(src, our_up)
    say("case: MC pivots L=40, 6 rounds, pivots every 2, against the reference CLI")
    r = subprocess.run([REF_BIN, "--duration", "0.162", "--frame-interval", "0.027", "--monte-carlo-interval",
                        "0.0675", "--temperature", "0.85", "--seed", "7", "--log-level", "detailed",
                        "--disable-recentering", ref_up], env=dict(os.environ, OMP_NUM_THREADS="1"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    b = ue.Batch([our_up], [0.85], seed=7, dt=0.009)
    b.set_monte_carlo(2)
    b.open_output(None, log_level=1)
    b.run_logged(6, 1)
    b.close_output()
    assert b.mc_stats(0)[1] == 2
    keys = ("pos", "potential", "pivot_stats", "time", "hbond", "rotamer_bad_solves_cumulative")
    with h5lite.File(ref_up) as f:
        want = {k: f.read("/output/" + k) for k in keys}
    with h5lite.File(our_up) as f:
        got = {k: f.read("/output/" + k) for k in keys}
    np.testing.assert_array_equal(got["rotamer_bad_solves_cumulative"], want["rotamer_bad_solves_cumulative"])
    assert got["hbond"].shape == want["hbond"].shape
    np.testing.assert_allclose(got["hbond"], want["hbond"], rtol=1e-3, atol=1e-5)
    say("pivot_stats ours %s reference %s" % (got["pivot_stats"].tolist(), want["pivot_stats"].tolist()))
    np.testing.assert_array_equal(got["pivot_stats"], want["pivot_stats"])
    np.testing.assert_array_equal(got["time"], want["time"])
    assert got["pos"].shape == want["pos"].shape == (6, 1, 120, 3)
    for k in range(6):
        e = rel_rms(got["pos"][k, 0], want["pos"][k, 0])
        de = abs(got["potential"][k, 0] - want["potential"][k, 0]) / max(abs(want["potential"][k, 0]), 1.0)
        say("frame %d: pos rel %.2e potential rel %.2e" % (k, e, de))
        assert e < 1e-4 and de < 1e-4, (k, e, de)


@pytest.mark.parametrize("name,size", [("sigmoid_coupling_environment", 3 * 20 + 400),
                                       ("bb_sigmoid_coupling_environment", 4), ("hbond_energy", 12)])
def test_coupling_param_deriv_matches_reference(ue, chain_file, name, size):
    """get_param_deriv of the sigmoid coupling nodes (scale, center, sharpness
    and weights) and of hbond_energy against the reference's PARAM_DERIV
    build."""
    import ctypes as ct
    lib = ref_lib()
    if lib is None:
        pytest.skip("reference engine not built")
    lib.get_param_deriv.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p]
    path = chain_file(60, 7)
    pos0 = read_pos(path)
    pos = (pos0 + np.random.default_rng(5).normal(scale=0.1, size=pos0.shape)).astype(np.float32)
    say("case: coupling param deriv %s" % name)
    eng = ue.Upside(path, quiet=True)
    ref = RefEngine(lib, path, len(pos0))
    eng.deriv(pos)
    ref.deriv(pos)
    want = np.zeros(size, np.float32)
    assert lib.get_param_deriv(size, want.ctypes.data, ct.c_void_p(ref.h), name.encode()) == 0
    got = eng.get_param_deriv((size,), name)
    err = float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-30))
    say("%s: |dE/dparam| %.4g rel err %.2e nonzero %d" % (name, np.linalg.norm(want), err, np.count_nonzero(want)))
    assert np.linalg.norm(want) > 0 and err < 1e-4, err
    with pytest.raises(RuntimeError):
        eng.get_param_deriv((size + 1,), name)
    # get_param equals the reference's; after the same set_param in both
    # engines, energies and forces agree
    lib.get_param.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p]
    lib.set_param.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p]
    p_ref = np.zeros(size, np.float32)
    assert lib.get_param(size, p_ref.ctypes.data, ct.c_void_p(ref.h), name.encode()) == 0
    p_ours = eng.get_param((size,), name)
    np.testing.assert_array_equal(p_ours, p_ref)
    new = (p_ref * (1.0 + 0.05 * np.random.default_rng(6).standard_normal(size))).astype(np.float32)
    eng.set_param(new, name)
    assert lib.set_param(size, new.ctypes.data, ct.c_void_p(ref.h), name.encode()) == 0
    np.testing.assert_array_equal(eng.get_param((size,), name), new)
    e, d = eng.energy(pos), eng.deriv(pos)
    er, dr = ref.energy(pos), ref.deriv(pos)
    say("%s after set_param: E %.5f ref %.5f, force rel %.2e" % (name, e, er, rel_rms(d, dr)))
    assert abs(e - er) / max(abs(er), 1.0) < E_TOL and rel_rms(d, dr) < F_TOL


def test_replica_exchange_loop_matches_reference_cli(ue, chain_file, tmp_path):
    """Four replicas, swap sets 0-1,2-3 and 1-2 tried in turn every 2 rounds,
    6 rounds with a frame every round, against the reference CLI
    (--replica-interval, --swap-set): the same replica indices and cumulative
    swap counts in every frame, the same swap partners, and positions and
    potentials within 1e-4 relative."""
    import shutil
    import subprocess
    import h5lite
    from conftest import REF_BIN
    if not os.path.exists(REF_BIN):
        pytest.skip("reference CLI not built")
    src = chain_file(40, 21)
    temps = [0.85, 0.88, 0.91, 0.95]
    refs = [str(tmp_path / ("ref%d.up" % k)) for k in range(4)]
    ours = [str(tmp_path / ("ours%d.up" % k)) for k in range(4)]
    for a, c in zip(refs, ours):
        shutil.copy(src, a)
        shutil.copy(src, c)
    say("case: replica exchange loop, 4 x L=40, 6 rounds, against the reference CLI")
    r = subprocess.run([REF_BIN, "--duration", "0.162", "--frame-interval", "0.027", "--replica-interval", "0.0675",
                        "--swap-set", "0-1,2-3", "--swap-set", "1-2", "--temperature", ",".join(map(str, temps)),
                        "--seed", "7", "--log-level", "basic", "--disable-recentering"] + refs,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    b = ue.Batch(ours, temps, seed=7, dt=0.009)
    b.set_replica_exchange(2, [[(0, 1), (2, 3)], [(1, 2)]])
    b.open_output(None)
    b.run_logged(6, 1)
    b.close_output()
    idx, suc, att = b.replica_stats()
    say("after 6 rounds: replica index %s, swaps %s of %s" % (idx.tolist(), suc.tolist(), att.tolist()))
    np.testing.assert_array_equal(att, [2, 2, 1])
    keys = ("pos", "potential", "replica_index", "replica_cumulative_swaps", "replica_swap_partner")
    for k in range(4):
        with h5lite.File(refs[k]) as f:
            want = {q: f.read("/output/" + q) for q in keys}
        with h5lite.File(ours[k]) as f:
            got = {q: f.read("/output/" + q) for q in keys}
        say("system %d replica_index %s reference %s" % (k, got["replica_index"].ravel().tolist(),
                                                         want["replica_index"].ravel().tolist()))
        for q in ("replica_index", "replica_cumulative_swaps", "replica_swap_partner"):
            np.testing.assert_array_equal(got[q], want[q], err_msg=q)
        assert got["pos"].shape == want["pos"].shape
        for fr in range(got["pos"].shape[0]):
            e = rel_rms(got["pos"][fr, 0], want["pos"][fr, 0])
            de = abs(got["potential"][fr, 0] - want["potential"][fr, 0]) / max(abs(want["potential"][fr, 0]), 1.0)
            assert e < 1e-4 and de < 1e-4, (k, fr, e, de)


def test_recentering_matches_reference_cli(ue, chain_file, tmp_path):
    """Recentering before every frame (the reference CLI's default): frame 0
    equals the CLI's recentered start; every frame has its centroid at the
    origin, a potential equal to the oracle's at the logged positions, and
    positions equal to an unrecentered run's minus their centroid.  (After
    frame 0 the CLI's own trajectory is not a reference: its affine alignment
    of residue 0 of this chain changes under translation, which moves that
    residue's side chain, so its energy takes two values under rigid shifts.)"""
    import shutil
    import subprocess
    import h5lite
    from conftest import REF_BIN
    if not os.path.exists(REF_BIN):
        pytest.skip("reference CLI not built")
    src = chain_file(40, 21)
    ref_up, our_up, plain_up = str(tmp_path / "ref.up"), str(tmp_path / "ours.up"), str(tmp_path / "plain.up")
    for q in (ref_up, our_up, plain_up):
        shutil.copy(src, q)
    say("case: recentering, L=40, 6 rounds")
    r = subprocess.run([REF_BIN, "--duration", "0.162", "--frame-interval", "0.027", "--temperature", "0.85",
                        "--seed", "7", "--log-level", "basic", ref_up], env=dict(os.environ, OMP_NUM_THREADS="1"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    b = ue.Batch([our_up], [0.85], seed=7, dt=0.009)
    b.set_recentering(1)
    b.open_output(None)
    b.run_logged(6, 1)
    b.close_output()
    plain = ue.Batch([plain_up], [0.85], seed=7, dt=0.009)
    plain.open_output(None)
    plain.run_logged(6, 1)
    plain.close_output()
    with h5lite.File(ref_up) as f:
        ref_pos = f.read("/output/pos")
    with h5lite.File(our_up) as f:
        got_pos, got_pot = f.read("/output/pos"), f.read("/output/potential")
    with h5lite.File(plain_up) as f:
        plain_pos = f.read("/output/pos")
    assert got_pos.shape == ref_pos.shape == plain_pos.shape
    assert rel_rms(got_pos[0, 0], ref_pos[0, 0]) < 1e-6
    ora = O.Engine(src)
    for fr in range(got_pos.shape[0]):
        p = got_pos[fr, 0]
        np.testing.assert_allclose(p.mean(axis=0), 0.0, atol=1e-4)
        e_ora = float(ora.energy(p))
        de = abs(got_pot[fr, 0] - e_ora) / max(abs(e_ora), 1.0)
        q = plain_pos[fr, 0] - plain_pos[fr, 0].mean(axis=0)
        e = rel_rms(p, q)
        say("frame %d: potential vs oracle %.2e, pos vs unrecentered run %.2e" % (fr, de, e))
        assert de < 1e-4 and e < 1e-4, (fr, de, e)
    with pytest.raises(RuntimeError):
        b.set_recentering(3)


@pytest.mark.parametrize("interval", [1, 2])
def test_annealing_matches_reference_cli(ue, chain_file, tmp_path, interval):
    """--anneal-factor 0.5 --anneal-start 0.03 --anneal-end 0.13 over 6 rounds
    (the schedule holds T0, then blends, then reaches T1 inside the run)
    against the reference CLI: logged temperatures equal, positions and
    potentials within 1e-4 relative.  With a thermostat interval of 2 rounds
    the temperature changes only at thermostat rounds (src/main.cpp:974-977),
    so frames between them log the last one set."""
    import shutil
    import subprocess
    import h5lite
    from conftest import REF_BIN
    if not os.path.exists(REF_BIN):
        pytest.skip("reference CLI not built")
    src = chain_file(40, 21)
    ref_up, our_up = str(tmp_path / "ref.up"), str(tmp_path / "ours.up")
    shutil.copy(src, ref_up)
    shutil.copy(src, our_up)
    say("case: annealing, L=40, 6 rounds, thermostat interval %d, against the reference CLI" % interval)
    r = subprocess.run([REF_BIN, "--duration", "0.162", "--frame-interval", "0.027", "--temperature", "0.85",
                        "--anneal-factor", "0.5", "--anneal-start", "0.03", "--anneal-end", "0.13",
                        "--thermostat-interval", "%.3f" % (0.027 * interval),
                        "--seed", "7", "--log-level", "basic", "--disable-recentering", ref_up],
                       env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    b = ue.Batch([our_up], [0.85], seed=7, dt=0.009, thermostat_interval=interval)
    b.set_annealing(0.5, 0.03, 0.13)
    b.open_output(None)
    b.run_logged(6, 1)
    b.close_output()
    keys = ("pos", "potential", "temperature")
    with h5lite.File(ref_up) as f:
        want = {k: f.read("/output/" + k) for k in keys}
    with h5lite.File(our_up) as f:
        got = {k: f.read("/output/" + k) for k in keys}
    say("temperatures ours %s reference %s" % (got["temperature"].ravel().tolist(),
                                               want["temperature"].ravel().tolist()))
    np.testing.assert_array_equal(got["temperature"], want["temperature"])
    for fr in range(got["pos"].shape[0]):
        e = rel_rms(got["pos"][fr, 0], want["pos"][fr, 0])
        de = abs(got["potential"][fr, 0] - want["potential"][fr, 0]) / max(abs(want["potential"][fr, 0]), 1.0)
        say("frame %d: pos rel %.2e potential rel %.2e" % (fr, e, de))
        assert e < 1e-4 and de < 1e-4, (fr, e, de)


def test_thermostat_timescale_matches_reference_cli(ue, chain_file, tmp_path):
    """--thermostat-timescale 2: 6 rounds with frames every round against the
    reference CLI, positions and potentials within 1e-4 relative."""
    import shutil
    import subprocess
    import h5lite
    from conftest import REF_BIN
    if not os.path.exists(REF_BIN):
        pytest.skip("reference CLI not built")
    src = chain_file(40, 21)
    ref_up, our_up = str(tmp_path / "ref.up"), str(tmp_path / "ours.up")
    shutil.copy(src, ref_up)
    shutil.copy(src, our_up)
    say("case: thermostat timescale 2, L=40, 6 rounds, against the reference CLI")
    r = subprocess.run([REF_BIN, "--duration", "0.162", "--frame-interval", "0.027", "--temperature", "0.85",
                        "--thermostat-timescale", "2", "--seed", "7", "--log-level", "basic",
                        "--disable-recentering", ref_up], env=dict(os.environ, OMP_NUM_THREADS="1"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    b = ue.Batch([our_up], [0.85], seed=7, dt=0.009)
    b.set_thermostat_timescale(2.0)
    b.open_output(None)
    b.run_logged(6, 1)
    b.close_output()
    with h5lite.File(ref_up) as f:
        want_pos, want_pot = f.read("/output/pos"), f.read("/output/potential")
    with h5lite.File(our_up) as f:
        got_pos, got_pot = f.read("/output/pos"), f.read("/output/potential")
    for fr in range(got_pos.shape[0]):
        e = rel_rms(got_pos[fr, 0], want_pos[fr, 0])
        de = abs(got_pot[fr, 0] - want_pot[fr, 0]) / max(abs(want_pot[fr, 0]), 1.0)
        say("frame %d: pos rel %.2e potential rel %.2e" % (fr, e, de))
        assert e < 1e-4 and de < 1e-4, (fr, e, de)
    with pytest.raises(RuntimeError):
        b.set_thermostat_timescale(3.0)       # only before the first round


def test_output_logger(ue, chain_file, tmp_path):
    """/output datasets as the reference's basic log level writes them: frames
    at rounds 0, 3, 6 of a 7-round run hold the state before that round's
    thermostat and integration; the logged run follows an unlogged launch
    with in-kernel frames at the same rounds bit for bit."""
    import h5lite
    path = chain_file(40, 21)
    say("case: /output logger, 2 x L=40, 7 rounds, frames every 3")
    outs = [str(tmp_path / "out0.up"), str(tmp_path / "out1.up")]
    temps, dt = [0.85, 0.9], 0.009
    b = ue.Batch([path] * 2, temps, seed=8, dt=dt)
    with pytest.raises(RuntimeError):
        b.open_output(None)          # both systems read one input file
    b.open_output(outs, record_momentum=True, log_level=1)
    b.run_logged(7, 3)
    b.close_output()
    ref = ue.Batch([path] * 2, temps, seed=8, dt=dt)
    pots = ref.run(7, frame_interval=3, n_frames=3)
    pos0 = read_pos(path)
    n_atom = len(pos0)
    for s, o in enumerate(outs):
        np.testing.assert_array_equal(b.pos(s), ref.pos(s))
        with h5lite.File(o) as f:
            pos, mom = f.read("/output/pos"), f.read("/output/mom")
            kin, pot = f.read("/output/kinetic"), f.read("/output/potential")
            t, temp = f.read("/output/time"), f.read("/output/temperature")
            hb, bad = f.read("/output/hbond"), f.read("/output/rotamer_bad_solves_cumulative")
        assert pos.shape == (3, 1, n_atom, 3) and mom.shape == (3, 1, n_atom, 3)
        assert kin.shape == (3, 1) and pot.shape == (3, 1) and temp.shape == (3, 1) and t.shape == (3,)
        assert pos.dtype == np.float32 and kin.dtype == np.float64 and t.dtype == np.float64
        np.testing.assert_array_equal(pos[0, 0], pos0)
        np.testing.assert_allclose(pot[:, 0], pots[s], rtol=1e-6, atol=1e-5)
        m2 = (mom[:, 0].astype(np.float64) ** 2).sum(axis=(1, 2))
        np.testing.assert_allclose(kin[:, 0], 0.5 * m2 / n_atom, rtol=1e-6)
        want_t = [float(np.float32(3 * np.float32(dt)) * np.float32(r)) for r in (0, 3, 6)]
        np.testing.assert_array_equal(t, want_t)
        np.testing.assert_array_equal(temp[:, 0], np.float32(temps[s]))
        # detailed level: hbond of every donor and acceptor at the frame's
        # positions, cumulative bad solves
        eng = ue.Upside(path, quiet=True)
        eng.energy(pos[0, 0])
        want_hb = eng.get_output("protein_hbond")[:, 6]
        assert hb.shape == (3, len(want_hb)) and hb.dtype == np.float32
        np.testing.assert_allclose(hb[0], want_hb, rtol=1e-6, atol=1e-7)
        assert bad.shape == (3, 1) and (bad == 0).all()
        # per-residue rotamer values (src/rotamer.cpp:711-731) of every frame
        # equal the single engine's at the frame's positions
        nn = int(eng.get_value_by_name((1,), "rotamer", "n_node")[0])
        names = ["rotamer_free_energy", "rotamer_potential", "rotamer_entropy"]
        with h5lite.File(outs[s]) as f:
            logged = {k: f.read("/output/" + k) for k in names + ["rotamer_1body_energy%d" % q for q in range(3)]}
        for fr in range(3):
            eng.energy(pos[fr, 0])
            one = eng.get_value_by_name((nn, 3), "rotamer", "rotamer_1body_energy")
            want = {k: eng.get_value_by_name((nn,), "rotamer", k) for k in names}
            want.update({"rotamer_1body_energy%d" % q: one[:, q] for q in range(3)})
            for k, v in want.items():
                assert logged[k].shape == (3, nn) and logged[k].dtype == np.float32, k
                np.testing.assert_allclose(logged[k][fr], v, rtol=1e-5, atol=1e-5, err_msg=k)


def test_checkpoint_resume_is_bit_identical(ue, chain_file):
    """A saved batch state loaded into a fresh batch of the same files (other
    seed, other team size) continues bit for bit, also through a replica
    exchange; a state of other files is refused."""
    path = chain_file(90, 12)
    say("case: checkpoint and resume, 2 x L=90")
    b1 = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    b1.run(3)
    state = b1.save_state()
    b1.run(4)
    acc1 = b1.replica_exchange([[0, 1]], round_num=7)
    b1.run(2)
    b2 = ue.Batch([path] * 2, [0.85, 0.9], seed=99)
    b2.set_team(min(4, b2.max_team()))
    b2.load_state(state)
    b2.run(4)
    acc2 = b2.replica_exchange([[0, 1]], round_num=7)
    b2.run(2)
    np.testing.assert_array_equal(acc1, acc2)
    for s in range(2):
        np.testing.assert_array_equal(b2.pos(s), b1.pos(s))
        np.testing.assert_array_equal(b2.mom(s), b1.mom(s))
    np.testing.assert_array_equal(b2.energies(), b1.energies())
    # the loop state travels too: pivot moves and replica exchange resume
    c1 = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    c1.set_monte_carlo(3)
    c1.set_replica_exchange(2, [[(0, 1)]])
    c1.run_logged(3, 0)
    loop_state = c1.save_state()
    c1.run_logged(5, 0)
    c2 = ue.Batch([path] * 2, [0.85, 0.9], seed=17)
    c2.load_state(loop_state)
    c2.run_logged(5, 0)
    for s in range(2):
        np.testing.assert_array_equal(c2.pos(s), c1.pos(s))
        assert c2.mc_stats(s) == c1.mc_stats(s)
    for x, y in zip(c2.replica_stats(1), c1.replica_stats()):
        np.testing.assert_array_equal(x, y)
    other = ue.Batch([chain_file(40, 21)] * 2, [0.85, 0.9], seed=6)
    with pytest.raises(RuntimeError):
        other.load_state(state)
    with pytest.raises(RuntimeError):
        ue.Batch([path], [0.85], seed=6).load_state(state)


@pytest.mark.parametrize("team", [2, 8])
def test_team_execution_matches_single_workgroup(ue, chain_file, team):
    """A system evaluated by a team of workgroups (grid barriers between
    phases) follows the single-workgroup trajectory bit for bit: forces are
    sums in a canonical order that does not depend on the team size."""
    path = chain_file(120, 71)
    say("case: team %d vs single workgroup, L=120" % team)
    b1 = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    b1.run(4)
    bt = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    if team > bt.max_team():
        pytest.skip("team %d exceeds resident limit %d" % (team, bt.max_team()))
    bt.set_team(team)
    bt.run(4)
    for s in range(2):
        np.testing.assert_array_equal(bt.pos(s), b1.pos(s))
    np.testing.assert_allclose(bt.energies(), b1.energies(), rtol=1e-5)


def test_section_lane_split_matches_single_workgroup(ue, chain_file):
    """The four lanes of a section run on quarters of a team of eight
    (split 4) or two by two on its halves (split 2): both follow the
    single-workgroup trajectory bit for bit, the lanes being independent."""
    path = chain_file(150, 72)
    b1 = ue.Batch([path] * 2, [0.85, 0.9], seed=7)
    b1.run(4)
    for split in (2, 4):
        bt = ue.Batch([path] * 2, [0.85, 0.9], seed=7)
        if bt.max_team() < 8:
            pytest.skip("no team of eight fits")
        bt.set_team(8)
        bt.set_lane_split(split)
        assert bt.lane_split() == split
        say("case: team 8, lane split %d, L=150" % split)
        bt.run(4)
        for s in range(2):
            np.testing.assert_array_equal(bt.pos(s), b1.pos(s))
            np.testing.assert_array_equal(bt.mom(s), b1.mom(s))
    with pytest.raises(RuntimeError):
        bt.set_lane_split(3)


def test_per_system_teams_match_single_workgroup(ue, chain_file):
    """Every system with its own team size (sizes 1, 3, 6, 5 and 2, so odd
    and uneven splits into halves too; each team placed on one XCD) follows
    the single-workgroup trajectories bit for bit, and so do the balanced
    teams (sizes in proportion to the atom counts, within the resident
    workgroups) for the rounds after."""
    lens = [40, 150, 90, 120, 60]
    paths = [chain_file(L, 80 + k) for k, L in enumerate(lens)]
    temps = [0.8, 0.85, 0.9, 0.85, 0.8]
    b1 = ue.Batch(paths, temps, seed=9)
    b1.run(4)
    bt = ue.Batch(paths, temps, seed=9)
    teams = [1, 3, 6, 5, 2]
    bt.set_teams(teams)
    assert list(bt.teams()) == teams and bt.team() == 6
    say("case: per-system teams %s, L=%s" % (teams, lens))
    bt.run(4)
    for s in range(len(lens)):
        np.testing.assert_array_equal(bt.pos(s), b1.pos(s))
        np.testing.assert_array_equal(bt.mom(s), b1.mom(s))
    bal = bt.balanced_teams()
    say("balanced teams %s (atoms %s)" % (list(bal), bt.n_atom))
    order = np.argsort(bt.n_atom, kind="stable")
    assert np.all(np.diff(bal[order]) >= 0) and bal.min() >= 1
    bt.set_teams(bal)
    bt.run(2)
    b1.run(2)
    for s in range(len(lens)):
        np.testing.assert_array_equal(bt.pos(s), b1.pos(s))
    np.testing.assert_array_equal(bt.energies(), b1.energies())
    with pytest.raises(RuntimeError):
        bt.set_teams([1, 2])
    bt.set_team(1)
    assert list(bt.teams()) == [1] * len(lens)
    # replica exchange inside the launch (grid barriers over every team) with
    # per-system teams: the same swaps and trajectories as one workgroup each
    path = chain_file(90, 88)
    temps = [0.80, 0.85, 0.90, 0.95]
    sets = [[(0, 1), (2, 3)], [(1, 2)]]
    runs = []
    for teams in ([1, 1, 1, 1], [2, 5, 1, 3]):
        b = ue.Batch([path] * 4, temps, seed=5)
        b.set_teams(teams)
        b.set_replica_exchange(3, sets)
        say("case: in-launch replica exchange with teams %s" % teams)
        b.run_logged(24, 0)
        runs.append(([b.pos(s) for s in range(4)], b.replica_stats()))
    for s in range(4):
        np.testing.assert_array_equal(runs[1][0][s], runs[0][0][s])
    for x, y in zip(runs[1][1], runs[0][1]):
        np.testing.assert_array_equal(x, y)
    assert runs[0][1][2].sum() > 0  # swaps were attempted


@pytest.mark.parametrize("n_res,layout", [(120, 1), (120, 3), (400, 0), (400, 2)])
def test_rotamer_solve_layouts_match(ue, chain_file, n_res, layout):
    """The rotamer BP solve on the whole team (state in memory) or on one
    workgroup (state in LDS, or in memory when it does not fit) computes the
    same arithmetic in the same order: the trajectory is bit for bit the
    single-workgroup one.  At 400 residues the random start has more BP edges
    than one workgroup's LDS holds, so layout 0 solves on the team there."""
    path = chain_file(n_res, 73)
    b1 = ue.Batch([path], [0.85], seed=6)
    b1.run(3)
    bt = ue.Batch([path], [0.85], seed=6)
    team = min(8, bt.max_team())
    if team < 2:
        pytest.skip("no team of two workgroups fits")
    bt.set_team(team)
    bt.set_bp_layout(layout)
    say("case: L=%d team %d bp layout %d vs single workgroup" % (n_res, team, layout))
    bt.run(3)
    st1, stt = b1.bp_stats(0), bt.bp_stats(0)
    assert st1["total"] == stt["total"] and st1["solves"] == stt["solves"]
    np.testing.assert_array_equal(bt.pos(0), b1.pos(0))
    np.testing.assert_array_equal(bt.mom(0), b1.mom(0))


def test_team_evaluations_are_fresh(ue, chain_file):
    """Every evaluation launch of a team recomputes every term, whatever ran
    before it: right after a trajectory launch, after positions set by the
    host, and with the team size switched in between.  Compared per term
    against single-workgroup evaluations of the same positions.  (A barrier
    that took its generation from a value read at launch start once let the
    first op run of a launch go stale.)"""
    path = chain_file(120, 71)
    b1 = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    b1.run(1)
    want_e = b1.energies()
    want_t = [b1.potential_terms(s) for s in range(2)]
    team = min(16, ue.Batch([path] * 2, [0.85, 0.9], seed=6).max_team())
    if team < 2:
        pytest.skip("no team of two workgroups fits")

    def check(b, tag):
        say("case: %s, team %d" % (tag, b.team()))
        np.testing.assert_allclose(b.energies(), want_e, rtol=1e-5, err_msg=tag)
        for s in range(2):
            got = b.potential_terms(s)
            for k, v in want_t[s].items():
                assert abs(got[k] - v) <= 1e-4 * max(1.0, abs(v)), (tag, s, k, got[k], v)

    bt = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    bt.set_team(team)
    bt.run(1)
    check(bt, "after a trajectory launch")
    check(bt, "second evaluation")
    bt.set_team(2)
    check(bt, "team size switched")
    bf = ue.Batch([path] * 2, [0.85, 0.9], seed=6)
    bf.set_team(team)
    for s in range(2):
        bf.set_pos(s, b1.pos(s))
    check(bf, "fresh batch, positions from the host")


def _threefry_uniform(seed, done, k):
    """Draw k of RandomGenerator(seed, 1, 0, done).uniform_open_closed()
    (src/main.cpp:266, src/random.h): Threefry-4x32-20 of counter
    (done lo, done hi, 0, k), first word to (0, 1] in float32."""
    bits = O.threefry4x32_20(np.array([[done & 0xFFFFFFFF, done >> 32, 0, k]], np.uint32), [seed, 1, 0, 0])[0]
    return np.float32(bits[0]) * np.float32(2.0 ** -32) + np.float32(2.0 ** -33)


def oracle_deriv_given(ora, pos, given):
    """The oracle's dE/dx with the outputs of the nodes in `given` (name ->
    array, e.g. the engine's affine-alignment frames) substituted after each
    such node computes: the comparison then covers the arithmetic downstream
    of those inputs alone."""
    saved = {}
    for name, arr in given.items():
        node = ora.nodes[name]
        saved[name] = node.compute

        def patched(mode, node=node, arr=arr, orig=node.compute):
            orig(mode)
            node.output[:] = arr
        node.compute = patched
    try:
        return ora.deriv(pos)
    finally:
        for name, f in saved.items():
            ora.nodes[name].compute = f


def f32_conditioning_bar(err_f32_vs_f64):
    """Force bar against float32 results where float32 itself is far from
    float64: the steep backbone-pair repulsion (width 0.3 A^2 in r^2) turns
    the float32 rounding of the residue frames (~4e-6 A at 30-60 A
    coordinates) into force differences up to ~1e-4 of the total at some
    configurations (measured per term against float64: backbone_pairs carries
    all but ~1e-5 of it).  Two float32 engines then differ by about as much as
    either differs from float64: 1.25 times the checker's own distance."""
    return max(F_TOL, 1.25 * err_f32_vs_f64)


# Force bar where the rotamer solves of two engines stopped after different
# iteration counts.  The reference stops its BP when the one-sided change
# over a chunk is at most tol = 1e-3 (src/rotamer.cpp:1139-1152), which leaves
# the marginals up to ~1e-3 from the fixed point; two float32 engines whose
# roundings differ can stop a chunk apart, and the forces then differ by up
# to a few 1e-4 (measured: 1.0e-4 at worst over 64 chains after 20 rounds,
# 3e-4 at the ideal start of a 150-residue chain).  Compared with the
# iteration count matched, the bar is F_TOL.
F_TOL_STOP = 5e-4


def _check_pairs_and_edges(ue, path, positions, label, ref_lib_obj, f64_every=1):
    """Energy and forces of every position set against the float32 oracle,
    the compiled reference and (every f64_every-th) the float64 oracle; every
    pair graph's edge set against the float32 oracle's on the engine's own
    coordinates (XOR 0).  Strict bars (F_TOL): the oracles with their BP run
    for exactly the engine's iteration count and the engine's affine frames
    substituted (the arithmetic of every node downstream of the frames).  The
    plain comparisons: F_TOL where the BP iteration counts agree and the
    float32 oracle is within F_TOL of float64, else the documented bars for
    the BP stopping point (F_TOL_STOP) and float32 conditioning
    (f32_conditioning_bar); with equal iteration counts the plain float32
    comparison also allows the oracle's own difference between its frames and
    the engine's (the strict check bounds the rest); the compiled reference's
    iteration count is not readable, so its bar allows both."""
    n_atom = len(positions[0])
    eng = ue.Upside(path, quiet=True)
    ora = O.Engine(path)
    ora64 = O.Engine(path, dtype=np.float64)
    ref = RefEngine(ref_lib_obj, path, n_atom) if ref_lib_obj is not None else None
    worst = {}
    for k, pos in enumerate(positions):
        pos = np.ascontiguousarray(pos, np.float32)
        e, d = float(eng.energy(pos)), eng.deriv(pos)
        it = int(eng.get_value_by_name((1,), "rotamer", "iterations")[0])
        frames = {"affine_alignment": eng.get_output("affine_alignment")}
        p64 = pos.astype(np.float64)
        e32, d32 = float(ora.energy(pos)), ora.deriv(pos)
        same32 = ora.nodes["rotamer"].iterations == it
        d64_own = ora64.deriv(p64)
        cond = f32_conditioning_bar(rel_rms(d32, d64_own))
        err = dict(e32=abs(e - e32) / max(abs(e32), 1.0), f32=rel_rms(d, d32))
        ora.nodes["rotamer"].force_iters = it
        d32_matched = oracle_deriv_given(ora, pos, frames)
        ora.nodes["rotamer"].force_iters = None
        err["f32_matched"] = rel_rms(d, d32_matched)
        # with the iteration counts equal, the plain comparison differs from
        # the strict one only by the frames: by the triangle inequality it is
        # within F_TOL plus the oracle's own frame difference
        frame_term = rel_rms(d32_matched, d32) if same32 else 0.0
        bars = dict(f32=max(cond, F_TOL + frame_term if same32 else F_TOL_STOP))
        if k % f64_every == 0:
            e64, d64 = float(ora64.energy(p64)), ora64.deriv(p64)
            same64 = ora64.nodes["rotamer"].iterations == it
            err.update(e64=abs(e - e64) / max(abs(e64), 1.0), f64=rel_rms(d, d64))
            ora64.nodes["rotamer"].force_iters = it
            d64_matched = oracle_deriv_given(ora64, p64, frames)
            ora64.nodes["rotamer"].force_iters = None
            err["f64_matched"] = rel_rms(d, d64_matched)
            frame64 = rel_rms(d64_matched, d64) if same64 else 0.0
            bars["f64"] = max(cond, F_TOL + frame64 if same64 else F_TOL_STOP)
        if ref is not None:
            er, dr = ref.energy(pos), ref.deriv(pos)
            err.update(eref=abs(e - er) / max(abs(er), 1.0), fref=rel_rms(d, dr))
            bars["fref"] = max(cond, F_TOL_STOP)
        xor_own = 0
        for name in PAIR_NODES:
            got = set(map(tuple, eng.edge_pairs(name).tolist()))
            assert got == oracle_edges_at(eng, ora, name), (label, k, name)
            xor_own += len(got ^ oracle_edges(ora, name))
        err["xor_vs_oracle_coordinates"] = xor_own
        bad = [(key, v) for key, v in err.items() if key != "xor_vs_oracle_coordinates" and
               not v < (E_TOL if key.startswith("e") else bars.get(key, F_TOL))]
        if bad:
            dump_case("%s_%d" % (label, k), pos=pos)
        for key, v in err.items():
            worst[key] = max(worst.get(key, 0.0), v)
        assert not bad, (label, k, bad, {key: bars.get(key, F_TOL) for key, _ in bad})
    say(label, "worst over %d:" % len(positions), " ".join("%s=%.2e" % kv for kv in sorted(worst.items())))
    return worst


def test_replica_exchange_16x150_in_one_launch(ue, chain_file):
    """Sixteen replicas of one 150-residue chain at temperatures geometric
    over 0.80-0.95, neighbour swap sets alternating every 7 rounds, the
    exchange inside the persistent launch (no host round trip per swap).
    Every attempt's decisions replayed from Threefry stream 1 on the
    energies the replicas held (src/main.cpp:243-316, criterion 0); replica
    indices permuted by exactly the accepted swaps; after 105 rounds every
    replica's energy and forces against the oracles and the compiled
    reference, and every pair graph's edge set against the oracle's."""
    path = chain_file(150, 2)
    temps = [0.80 * (0.95 / 0.80) ** (i / 15.0) for i in range(16)]
    beta = (1.0 / np.array(temps, np.float32)).astype(np.float32)
    sets = [[(i, i + 1) for i in range(0, 15, 2)], [(i, i + 1) for i in range(1, 14, 2)]]
    b = ue.Batch([path] * 16, temps, seed=2)
    b.set_team(min(4, b.max_team()))
    b.set_replica_exchange(7, sets, criterion=0)
    say("case: 16x150 replica exchange in the launch, team %d" % b.team())
    offs = [0, len(sets[0]), len(sets[0]) + len(sets[1])]
    rep_prev = np.arange(16)
    suc_prev = np.zeros(offs[-1], np.int64)
    att_prev = np.zeros(offs[-1], np.int64)
    n_acc = near = 0
    for seg in range(15):
        b.run_logged(7, 0)
        done = 7 * (seg + 1)
        rep, suc, att = b.replica_stats()
        k = seg % 2
        lo, hi = offs[k], offs[k + 1]
        assert np.all(att[lo:hi] - att_prev[lo:hi] == 1) and np.all(np.delete(att - att_prev, np.s_[lo:hi]) == 0)
        acc = (suc - suc_prev)[lo:hi]
        e_now = b.energies()
        want_rep = rep_prev.copy()
        draw = 0
        for (s1, s2), a in zip(sets[k], acc):
            # energies the pair held at the attempt (an accepted pair traded positions)
            e1, e2 = (e_now[s2], e_now[s1]) if a else (e_now[s1], e_now[s2])
            diff = np.float32((np.float32(-beta[s1] * e2) + np.float32(-beta[s2] * e1))
                              - (np.float32(-beta[s1] * e1) + np.float32(-beta[s2] * e2)))
            want = 1
            if diff < 0:
                u = _threefry_uniform(2, done, draw)
                draw += 1
                want = 0 if np.exp(diff) < u else 1
                if abs(float(np.log(u)) - float(diff)) < 1e-4:
                    near += 1
                    continue
            assert a == want, (seg, s1, s2, diff)
            if a:
                want_rep[[s1, s2]] = want_rep[[s2, s1]]
            n_acc += a
        np.testing.assert_array_equal(rep, want_rep)
        rep_prev, suc_prev, att_prev = rep, suc, att
    say("accepted %d swaps over 15 attempts; %d decisions within 1e-4 of the threshold" % (n_acc, near))
    assert 0 < n_acc and near == 0
    _check_pairs_and_edges(ue, path, [b.pos(s) for s in range(16)], "16x150 after 105 rounds", ref_lib())


def test_uneven_batch_64_chains(ue, chain_file):
    """Sixty-four independent chains of lengths uniform in [50, 400] in one
    launch (team autotuned as the bench does): after 20 rounds every chain's
    energy and forces against the float32 oracle and the compiled reference
    (float64 for every fourth), and every pair graph's edge set against the
    oracle's."""
    lens = [int(x) for x in np.random.default_rng(3).integers(50, 401, 64)]
    paths = [chain_file(L, 100 + i) for i, L in enumerate(lens)]
    b = ue.Batch(paths, [0.85] * 64, seed=3)
    cand = [t for t in (1, 2, 4) if t <= b.max_team()]
    b.autotune(2, cand)
    say("case: 64 uneven chains (%d residues), team %d" % (sum(lens), b.team()))
    b.run(20)
    lib = ref_lib()
    for s, p in enumerate(paths):
        _check_pairs_and_edges(ue, p, [b.pos(s)], "chain %d (L=%d)" % (s, lens[s]), lib, f64_every=1 if s % 4 == 0 else 2)


def test_long_run_500_residues_sampled(ue, chain_file):
    """The plain-method larger workload over a long run: 3,333 rounds (9,999
    steps) of the 500-residue chain in 10 launches; at the end of each, the
    engine's energy and forces against the float32 and float64 oracles and
    the compiled reference, and edge sets against the oracle's."""
    path = chain_file(500, 1)
    b = ue.Batch([path], [0.85], seed=1)
    b.set_team(min(16, b.max_team()))
    samples = []
    for k in range(10):
        say("case: 500-residue long run, segment %d" % k)
        b.run(333 + (3 if k == 9 else 0))
        samples.append(b.pos(0))
    assert np.all(np.isfinite(samples[-1]))
    _check_pairs_and_edges(ue, path, samples, "500 residues, 10 samples over 3333 rounds", ref_lib())
