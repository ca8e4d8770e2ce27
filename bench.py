"""Benchmark: residue-steps/s of the HIP Upside engine on one GPU.

Workload (the plain-method larger configuration): one 500-residue chain,
sequence drawn uniformly from the 20 amino acids with seed 1, backbone from
random Ramachandran angles, ff_2.1 potential (hbond, dynamic-rotamer 1-body,
rotamer placement and interaction, environment, rama), FP32, T=0.85,
thermostat every round.  A step is one force evaluation plus one integrator
stage (dt=0.009), as the reference's us/systems/step counts it
(src/main.cpp:1007-1012); rounds are 3 steps, so K is rounded up to a
multiple of 3 and the line reports the steps actually timed.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  The timed region is one persistent launch of K steps
with positions, momenta and tables already resident in HBM, bracketed by
device synchronisation.  `roofline` prices the algorithmic bytes of the
launch (pair graphs, BP messages, integrator; formula in `algorithmic_bytes`)
against the measured kernel duration from HIP events; `cpu_baseline` is the
reference engine compiled from its own sources (oracle/_ref/upside_ref),
timed on this host with one thread on a bounded sample of the same workload.
"""
import argparse
import ctypes as ct
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "upside2-md_amd")
sys.path.insert(0, PKG)

import h5lite  # noqa: E402
import upconfig  # noqa: E402

METRIC = "residue-steps/s vs plain GPU & CPU at ≤1e-4 rel force err; BP iters/step"
N_RES, SEED, TEMP, DT = 500, 1, 0.85, 0.009
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "engine_step_kernel_chain500_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hip_sync():
    lib = ct.CDLL("libamdhip64.so")
    if lib.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


# ------------------------------------------------------------- CPU baseline
def cpu_baseline(path, rounds, use_ref=True):
    """The reference CLI (compiled from the reference's sources into
    oracle/_ref by oracle/Makefile) on a copy of the same .up, one thread;
    falls back to timing the numpy oracle when the reference build is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", "upside_ref")
    if use_ref and os.path.exists(exe):
        d = tempfile.mkdtemp()
        try:
            p = os.path.join(d, "cpu.up")
            shutil.copy(path, p)
            dur = rounds * 3 * DT
            env = dict(os.environ, OMP_NUM_THREADS="1")
            t0 = time.perf_counter()
            r = subprocess.run([exe, "--duration", "%.6f" % dur, "--frame-interval", "%.6f" % dur,
                                "--temperature", str(TEMP), "--seed", str(SEED), "--log-level", "basic",
                                "--disable-recentering", p], env=env, capture_output=True, text=True, timeout=600)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError("reference CLI failed: " + r.stderr[-500:])
            m = re.search(r"\(([0-9.eE+-]+) us/systems/step", r.stdout)
            us = float(m.group(1)) if m else wall * 1e6 / (rounds * 3)
            return dict(value=N_RES / (us * 1e-6), unit="residue-steps/s", cores=1, kind="reference",
                        sample="reference upside CLI built from its sources (oracle/_ref), OMP_NUM_THREADS=1, "
                               "same %d-residue chain, T=%.2f, %d rounds = %d steps; %.1f us/step as it reports, "
                               "%.1f s wall" % (N_RES, TEMP, rounds, rounds * 3, us, wall))
        finally:
            shutil.rmtree(d, ignore_errors=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import upside_oracle  # CPU checker restatement, timed as the baseline of last resort
    eng = upside_oracle.Engine(path)
    with h5lite.File(path) as f:
        pos = f.read("/input/pos")[:, :, 0].astype(np.float32)
    mom, n_inv = upside_oracle.thermalize(len(pos), TEMP, SEED, DT)
    n = max(1, rounds // 20)
    t0 = time.perf_counter()
    upside_oracle.run_rounds(eng, pos, mom, n, DT, TEMP, SEED, n_inv)
    wall = time.perf_counter() - t0
    return dict(value=N_RES * 3 * n / wall, unit="residue-steps/s", cores=1, kind="port",
                sample="numpy oracle restatement, %d rounds of the same chain" % n)


def reference_errors(path, pos, energy_gpu, deriv_gpu):
    """Relative energy and force errors (||F_gpu - F_ref|| / ||F_ref||) at the
    final positions against the compiled reference engine (part of the CPU
    checker leg)."""
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libupside_ref.so")
    if not os.path.exists(lib_path):
        return None, None
    lib = ct.CDLL(lib_path)
    lib.construct_deriv_engine.restype = ct.c_void_p
    lib.construct_deriv_engine.argtypes = [ct.c_int, ct.c_char_p, ct.c_bool]
    lib.evaluate_energy.argtypes = [ct.c_void_p] * 3
    lib.evaluate_deriv.argtypes = [ct.c_void_p] * 3
    lib.free_deriv_engine.argtypes = [ct.c_void_p]
    h = lib.construct_deriv_engine(len(pos), path.encode(), True)
    d = np.zeros_like(pos)
    e = np.zeros(1, np.float32)
    p = np.ascontiguousarray(pos, np.float32)
    lib.evaluate_energy(e.ctypes.data, ct.c_void_p(h), p.ctypes.data)
    lib.evaluate_deriv(d.ctypes.data, ct.c_void_p(h), p.ctypes.data)
    lib.free_deriv_engine(ct.c_void_p(h))
    return (abs(energy_gpu - float(e[0])) / max(abs(float(e[0])), 1.0),
            float(np.sqrt(((deriv_gpu - d) ** 2).sum() / (d ** 2).sum())))


def oracle_edges_on(probe, ora, name):
    """The float32 oracle's edge set of a pair graph refined from the engine's
    own coordinates of the graph's producers (probe's last evaluation), so
    that the refine is compared on identical inputs; and the same set from the
    oracle's own producer coordinates."""
    import upside_oracle
    node = ora.nodes[name]
    own = None
    if hasattr(node, "ig"):
        ig = node.ig
        own = set(zip(ig.i1.tolist(), ig.i2.tolist()))
        prods = {ig.node1.name: ig.node1, ig.node2.name: ig.node2}
        saved = {k: n.output.copy() for k, n in prods.items()}
        for k, n in prods.items():
            n.output[:] = probe.get_output(k)
        try:
            i1, i2 = ig.edges()
        finally:
            for k, n in prods.items():
                n.output[:] = saved[k]
            ig.edges()
        return set(zip(i1.tolist(), i2.tolist())), own
    if not (hasattr(node, "i1") and hasattr(node, "res") and hasattr(node, "dist_cutoff")):
        return None, None
    own = set(zip(node.i1.tolist(), node.i2.tolist()))
    # backbone pairs: residue pairs of the alignment frames' centres
    t = probe.get_output(node.args[0])[node.res, :3].astype(np.float32)
    d2 = upside_oracle.refine_dist2(t[:, None] - t[None, :], np.float32)
    cut2 = np.float32(node.dist_cutoff) * np.float32(node.dist_cutoff)
    i1, i2 = np.nonzero(np.triu((d2 < cut2) & (np.abs(node.res[:, None] - node.res[None, :]) > 1), 1))
    return set(zip(i1.tolist(), i2.tolist())), own


def oracle_checks(path, pos, energy_gpu, deriv_gpu, probe, graph_names):
    """At the final positions: relative energy and force errors against the
    float64 oracle, and the XOR count of every pair graph's edge set against
    the float32 oracle's refine of the engine's own producer coordinates
    (must be 0).  The XOR against the oracle's refine of its own producer
    coordinates is reported beside it: a pair whose distance lies within the
    two engines' rounding of the bead positions from the cutoff can differ
    there.  Part of the CPU checker leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import upside_oracle
    ora = upside_oracle.Engine(path, dtype=np.float64)
    p64 = pos.astype(np.float64)
    e64 = float(ora.energy(p64))
    d64 = ora.deriv(p64)
    ora32 = upside_oracle.Engine(path)      # edge sets are defined in float32
    ora32.energy(np.ascontiguousarray(pos, np.float32))
    xor, xor_own = {}, {}
    for name in graph_names:
        if name not in ora32.nodes:
            continue
        want, own = oracle_edges_on(probe, ora32, name)
        if want is None:
            continue
        got = set(map(tuple, probe.edge_pairs(name).tolist()))
        xor[name] = len(got ^ want)
        xor_own[name] = len(got ^ own)
    return (abs(energy_gpu - e64) / max(abs(e64), 1.0),
            float(np.sqrt(((deriv_gpu - d64) ** 2).sum() / (d64 ** 2).sum())), xor, xor_own)


def percentile_from_hist(h, q):
    """Smallest k with cumulative count >= q * total."""
    c = np.cumsum(h)
    if c[-1] == 0:
        return None
    return int(np.searchsorted(c, q * c[-1]))


# ------------------------------------------------------- algorithmic bytes
# pair-graph nodes of the ff_2.1 graph (InteractionGraph users, src/interaction_graph.h)
PAIR_NODES = ("rotamer", "protein_hbond", "hbond_coverage", "hbond_coverage_hydrophobe",
              "environment_coverage_sc", "environment_coverage_hb", "hb_environment_coverage_hn",
              "hb_environment_coverage_oc", "backbone_pairs")


def graph_counts(eng, node_names):
    """Edge counts of every interaction graph and BP class counts at the
    engine's last evaluation."""
    out = {}
    for name in node_names:
        if name not in PAIR_NODES:
            continue
        shape = eng.get_value_by_name((4,), name, "graph_shape")
        n_edge = float(eng.get_value_by_name((1,), name, "n_edge")[0])
        out[name] = dict(shape=[float(x) for x in shape], n_edge=n_edge)
    bp = eng.get_value_by_name((6,), "rotamer", "bp_edge_classes")
    nodes = eng.get_value_by_name((3,), "rotamer", "node_classes")
    return out, [float(x) for x in bp], [float(x) for x in nodes]


def small_node_bytes(eng, node_names, graph_names):
    b = 0.0
    for name in node_names:
        if name in graph_names:
            continue
        n_elem, width = eng._dims(name)
        if n_elem * width > 1:
            b += 16.0 * n_elem * width   # value written+read, sensitivity written+read
    return b


def algorithmic_bytes(graphs, bp, nodes, n_atom, small, steps, passes):
    """Bytes a launch of `steps` steps must move at minimum, by the convention
    that every global dependency stage round-trips HBM:
      per step: sum_g [8 (n1 d1 + n2 d2) + 8 E_g]  (coordinates in,
                sensitivities out, edge list) + 8 E_rot (edge energies out and
                sensitivities in) + 2*4 a b M_ab over BP classes (fill and
                marginals) + 60 n_atom (integrator) + small nodes
      per BP pass: sum_ab M_ab 4 [a b + 3 (a + b)] (matrix, old and new
                messages, gathered beliefs) + 8 n_rot over 3- and 6-nodes."""
    per_step = 0.0
    for name, g in graphs.items():
        n1, d1, n2, d2 = g["shape"]
        per_step += 8.0 * (n1 * d1 + n2 * d2) + 8.0 * g["n_edge"]
    per_step += 8.0 * graphs.get("rotamer", {"n_edge": 0})["n_edge"]
    classes = [(3, 3, bp[3]), (3, 6, bp[4]), (6, 6, bp[5])]
    per_step += sum(2 * 4.0 * a * b * m for a, b, m in classes)
    per_step += 60.0 * n_atom + small
    per_pass = sum(m * 4.0 * (a * b + 3 * (a + b)) for a, b, m in classes) + 8.0 * (3 * nodes[1] + 6 * nodes[2])
    return per_step * steps + per_pass * passes, per_step, per_pass


def traffic_per_step():
    """Measured HBM bytes per step of this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary, if present."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"traffic_bytes_per_step:\s*([0-9.eE+-]+)", open(TRAFFIC_FILE).read())
    return float(m.group(1)) if m else None


# ------------------------------------------------------ batched workloads
def host_info():
    """Host cores this process may use and the CPU model (for every CPU baseline)."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return nproc, model


def cpu_reference_batch(paths, temps, rounds, seed, extra=()):
    """The reference CLI (oracle/_ref/upside_ref) over copies of the systems'
    .up files in one invocation, OMP_NUM_THREADS = min(systems, host cores):
    the reference parallelises across systems only (src/main.cpp:917)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "upside_ref")
    if not os.path.exists(exe):
        return None
    nproc, model = host_info()
    threads = max(1, min(len(paths), nproc))
    d = tempfile.mkdtemp()
    try:
        files = []
        for k, q in enumerate(paths):
            f = os.path.join(d, "sys%03d.up" % k)
            shutil.copy(q, f)
            files.append(f)
        dur = rounds * 3 * DT
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        cmd = [exe, "--duration", "%.6f" % dur, "--frame-interval", "%.6f" % dur, "--temperature",
               ",".join("%.6f" % t for t in temps), "--seed", str(seed), "--log-level", "basic",
               "--disable-recentering"] + list(extra) + files
        t0 = time.perf_counter()
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("reference CLI failed: " + r.stderr[-500:])
        m = re.search(r"\(([0-9.eE+-]+) us/systems/step", r.stdout)
        n_res = sum(h5lite_n_res(q) for q in paths)
        us = float(m.group(1)) if m else wall * 1e6 / (rounds * 3 * len(paths))
        value = n_res / (us * 1e-6 * len(paths))
        return dict(value=value, unit="residue-steps/s", cores=threads, kind="reference", nproc=nproc,
                    cpu_model=model,
                    sample="reference upside CLI built from its sources (oracle/_ref), one invocation over the "
                           "%d systems, OMP_NUM_THREADS=%d, %d rounds = %d steps; %.2f us/systems/step as it "
                           "reports, %.1f s wall" % (len(paths), threads, rounds, rounds * 3, us, wall))
    finally:
        shutil.rmtree(d, ignore_errors=True)


def h5lite_n_res(path):
    with h5lite.File(path) as f:
        return f.read("/input/pos").shape[0] // 3


def batch_roofline(b, paths, st0, st1, steps, kernel_ms):
    """Algorithmic bytes of a batch launch (sum over systems of the per-step
    and per-BP-pass bytes at the end state, section `algorithmic_bytes`) over
    the launch's kernel time."""
    import upside_engine
    probes = {}
    total = 0.0
    for s, q in enumerate(paths):
        if q not in probes:
            with h5lite.File(q) as f:
                names = f.children("/input/potential")
            probes[q] = (upside_engine.Upside(q, quiet=True), names)
        eng, names = probes[q]
        eng.energy(b.pos(s))
        g, bp, nodes = graph_counts(eng, names)
        small = small_node_bytes(eng, names, set(g))
        passes = (st1[s]["total"] - st0[s]["total"]) + (st1[s]["solves"] - st0[s]["solves"])
        n_atom = b.pos(s).shape[0]
        alg, _, _ = algorithmic_bytes(g, bp, nodes, n_atom, small, steps, passes)
        total += alg
    achieved = total / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": total,
            "kernel_ms": kernel_ms, "kernel": "upside::engine_step_kernel"}


def bp_iters(b, st0, st1):
    it = sum(a["total"] - c["total"] for a, c in zip(st1, st0))
    so = sum(a["solves"] - c["solves"] for a, c in zip(st1, st0))
    return it / max(so, 1)


def batched_workloads(work, rounds=60, cpu=True):
    """The batch workloads reported beside the single-chain line (not
    `value`), each with its own roofline and reference CPU baseline:
    16 replicas of one 150-residue chain at temperatures geometric over
    0.80-0.95 with alternating neighbour swap sets every 7 rounds (21 steps),
    exchanged inside the persistent launch; 64 independent chains with
    lengths uniform in [50, 400].  Each is autotuned (team size and rotamer
    solve layout), warmed up, then timed over `rounds` rounds in one launch.
    Returns the lines and the 64-chain batch's paths (training chains of the
    warm start)."""
    import upside_engine
    out = []

    def tune(b):
        # team sizes shared by every system, and (0) per-system teams in
        # proportion to the systems' sizes
        cand = [t for t in (1, 2, 4, 8, 16) if t <= b.max_team()] + [0]
        b.run(3)
        ms = b.autotune_layout(3, cand, (0, 1, 3))
        print("batch of %d: team / BP layout autotune (ms per 3 rounds; team 0: per-system teams %s): %s"
              % (b.n_sys, list(map(int, b.balanced_teams())), ms), file=sys.stderr)

    def teams(b):
        t = [int(x) for x in b.teams()]
        return t[0] if len(set(t)) == 1 else t

    p = os.path.join(work, "rex150.up")
    upconfig.make_chain(p, 150, seed=2)
    temps = [0.80 * (0.95 / 0.80) ** (i / 15.0) for i in range(16)]
    sets = [[(i, i + 1) for i in range(0, 15, 2)], [(i, i + 1) for i in range(1, 14, 2)]]
    b = upside_engine.Batch([p] * 16, temps, seed=2, dt=DT)
    tune(b)
    b.set_replica_exchange(7, sets)
    st0 = [b.bp_stats(s) for s in range(16)]
    hip_sync()
    t0 = time.perf_counter()
    b.run_logged(rounds, 0)          # one launch: the swaps happen inside it
    hip_sync()
    wall = time.perf_counter() - t0
    st1 = [b.bp_stats(s) for s in range(16)]
    _, suc, att = b.replica_stats()
    line = {"workload": "16 replicas x 150 residues, T geometric 0.80-0.95, neighbour swap sets alternating "
                        "every 21 steps, exchanged inside the persistent launch",
            "value": 16 * 150 * 3 * rounds / wall, "unit": "residue-steps/s", "steps": 3 * rounds,
            "seconds": wall, "workgroups_per_system": teams(b), "bp_layout": b.bp_layout(), "lane_split": b.lane_split(),
            "swaps_accepted_attempted": [int(suc.sum()), int(att.sum())],
            "bp_iters_per_solve": bp_iters(b, st0, st1),
            "roofline": batch_roofline(b, [p] * 16, st0, st1, 3 * rounds, b.last_run_ms())}
    if cpu:
        line["cpu_baseline"] = cpu_reference_batch(
            [p] * 16, temps, rounds, 2,
            ["--replica-interval", "%.4f" % (7 * 3 * DT + 1e-4)] +
            sum((["--swap-set", ",".join("%d-%d" % q for q in st)] for st in sets), []))
    out.append(line)
    del b
    rng = np.random.default_rng(3)
    lens = [int(x) for x in rng.integers(50, 401, 64)]
    paths = []
    for i, L in enumerate(lens):
        q = os.path.join(work, "uneven%02d.up" % i)
        upconfig.make_chain(q, L, seed=100 + i)
        paths.append(q)
    b = upside_engine.Batch(paths, [TEMP] * 64, seed=3, dt=DT)
    tune(b)
    st0 = [b.bp_stats(s) for s in range(64)]
    hip_sync()
    t0 = time.perf_counter()
    b.run(rounds)
    hip_sync()
    wall = time.perf_counter() - t0
    st1 = [b.bp_stats(s) for s in range(64)]
    line = {"workload": "64 independent chains, lengths uniform in [50, 400] (%d residues)" % sum(lens),
            "value": sum(lens) * 3 * rounds / wall, "unit": "residue-steps/s", "steps": 3 * rounds,
            "seconds": wall, "workgroups_per_system": teams(b), "bp_layout": b.bp_layout(), "lane_split": b.lane_split(),
            "bp_iters_per_solve": bp_iters(b, st0, st1),
            "roofline": batch_roofline(b, paths, st0, st1, 3 * rounds, b.last_run_ms())}
    if cpu:
        line["cpu_baseline"] = cpu_reference_batch(paths, [TEMP] * 64, rounds, 3)
    del b
    # the same engine one system at a time: every 8th chain alone on the GPU
    # (largest team, autotuned layout, the same warmup), timed over the same
    # rounds; the rate is the sample's residue-steps over its summed times
    res_steps, secs = 0, 0.0
    for q, L in list(zip(paths, lens))[::8]:
        one = upside_engine.Batch([q], [TEMP], seed=3, dt=DT)
        one.set_team(one.max_team())
        one.run(3)
        one.autotune_layout(3, [one.max_team()], (0, 1, 3))
        hip_sync()
        t0 = time.perf_counter()
        one.run(rounds)
        hip_sync()
        secs += time.perf_counter() - t0
        res_steps += L * 3 * rounds
        del one
    line["one_system_at_a_time"] = {
        "value": res_steps / secs, "unit": "residue-steps/s",
        "sample": "every 8th chain of the batch (8 chains, %d residues) alone on the GPU with a team of 32, "
                  "%d rounds each; residue-steps over the summed times" % (sum(lens[::8]), rounds)}
    line["batched_vs_one_at_a_time"] = line["value"] / (res_steps / secs)
    out.append(line)
    return out, paths


def warm_start_workloads(work, train_paths, rounds=20, train_steps=2000):
    """The learned BP warm start: rows harvested from converged plain solves
    (schedule 1, residual 1e-5) of the 64-chain batch, the network trained on
    the GPU, then a 1000-residue chain and 32 replicas of a 150-residue chain,
    each run from the same state under the reference's schedule (0), the
    plain start with the residual rule (1), the warm start with the
    residual rule (2) and the warm start with the reference's rule (3).
    Each run's end state is evaluated again with the plain start and the
    residual rule at 1e-5 (converged solves; the result does not depend on
    the start), and the force difference to that is reported."""
    import upside_engine
    b = upside_engine.Batch(train_paths, [TEMP] * len(train_paths), seed=5, dt=DT)
    b.set_team(min(2, b.max_team()))
    b.set_bp_schedule(1, 1e-5)
    F, T = [], []
    t0 = time.perf_counter()
    for frame in range(8):
        b.run(2)
        for s in range(len(train_paths)):
            f, t = b.bp_training_rows(s)
            F.append(f)
            T.append(t)
    del b
    F, T = np.concatenate(F), np.concatenate(T)
    net = upside_engine.WarmStartNet(seed=1)
    net.set_data(F, T)
    l0, _ = net.loss_grad()
    hip_sync()
    t1 = time.perf_counter()
    net.train(train_steps, batch_rows=8192, lr=1e-3)
    hip_sync()
    t_train = time.perf_counter() - t1
    l1, _ = net.loss_grad()
    params = net.params()
    info = {"training_rows": int(len(F)), "harvest_seconds": t1 - t0, "train_steps": train_steps,
            "train_seconds": t_train, "loss_before_after": [l0, l1]}
    p1000 = os.path.join(work, "warm1000.up")
    upconfig.make_chain(p1000, 1000, seed=4)
    p150 = os.path.join(work, "warm150.up")
    upconfig.make_chain(p150, 150, seed=6)
    lines = []
    for label, paths in (("one 1000-residue chain", [p1000]), ("32 replicas x 150 residues", [p150] * 32)):
        n_res = sum(h5lite_n_res(q) for q in paths)
        res = {}
        for sched in (0, 1, 2, 3):
            bb = upside_engine.Batch(paths, [TEMP] * len(paths), seed=6, dt=DT)
            bb.set_team(bb.max_team())  # 32 for the single chain, 8 for the replicas
            bb.run(2)
            if sched >= 2:
                bb.set_warm_start(params)
            if sched:
                bb.set_bp_schedule(sched, 1e-5 if sched < 3 else 0.0)  # 3: the node's own tolerance
            st0 = [bb.bp_stats(s) for s in range(len(paths))]
            hip_sync()
            t0 = time.perf_counter()
            bb.run(rounds)
            hip_sync()
            wall = time.perf_counter() - t0
            st1 = [bb.bp_stats(s) for s in range(len(paths))]
            ws = [bb.bp_warm_stats(s) for s in range(len(paths))]
            res[sched] = {"value": n_res * 3 * rounds / wall, "ms_per_step": wall * 1e3 / (3 * rounds),
                          "bp_iters_per_solve": bp_iters(bb, st0, st1),
                          "fallbacks": int(sum(w["fallbacks"] for w in ws))}
            if sched != 1:
                # the converged result does not depend on the start: evaluate the
                # run's end state again with the plain start and the residual rule
                bb.energies()
                d_run = [bb.deriv(s) for s in range(min(len(paths), 4))]
                bb.set_bp_schedule(1, 1e-5)
                bb.energies()
                d_conv = [bb.deriv(s) for s in range(min(len(paths), 4))]
                res[sched]["force_rel_diff_vs_converged"] = max(
                    float(np.sqrt(((a - c) ** 2).sum() / (c ** 2).sum())) for a, c in zip(d_run, d_conv))
        lines.append({"workload": label + ", warm start vs plain", "unit": "residue-steps/s", "steps": 3 * rounds,
                      "reference_schedule": res[0], "plain_start_residual_1e-5": res[1],
                      "warm_start_residual_1e-5": res[2], "warm_start_reference_rule": res[3],
                      "warm_vs_plain_iters": res[2]["bp_iters_per_solve"] / res[1]["bp_iters_per_solve"],
                      "warm_vs_reference_iters": res[3]["bp_iters_per_solve"] / res[0]["bp_iters_per_solve"],
                      "warm_vs_reference_speed": res[3]["value"] / res[0]["value"]})
    return info, lines


# ----------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cpu-rounds", type=int, default=400, help="rounds of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batch and warm-start workloads")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the warm-start workloads")
    ap.add_argument("--team", type=int, default=0,
                    help="workgroups per system (0: time 1..32 on the warmed state and keep the fastest)")
    ap.add_argument("--lane-split", type=int, default=2, choices=(2, 4),
                    help="sections on quarters (4) or halves (2) of the team, with --team (default: autotuned)")
    ap.add_argument("--bp-layout", type=int, default=-1,
                    help="rotamer solve layout with --team (0 auto, 1 team, 2 one workgroup, 3 split; "
                         "default: autotuned with the team size, or 0 with --team)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batch workloads and print their lines (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only an elementwise copy of 1 GiB (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine runs every system of a job on a single "
              "device and has no multi-GPU path (systems would shard across GPUs by whole replicas).",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 0.0, out=y)
        torch.cuda.synchronize()
        log("copy calibration: 3 launches, each reads and writes %d bytes" % (x.numel() * 4))
        return

    rounds = max(1, math.ceil(args.steps / 3))
    warm_rounds = max(1, math.ceil(args.warmup / 3)) if args.warmup > 0 else 0
    work = tempfile.mkdtemp()
    if args.batched_only:
        lines, _ = batched_workloads(work, cpu=not args.no_cpu_baseline)
        shutil.rmtree(work, ignore_errors=True)
        print(json.dumps({"batched": lines}), flush=True)
        return
    path = os.path.join(work, "chain%d_seed%d.up" % (N_RES, SEED))
    upconfig.make_chain(path, N_RES, seed=SEED)
    with h5lite.File(path) as f:
        node_names = f.children("/input/potential")
        n_atom = f.read("/input/pos").shape[0]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(path, args.cpu_rounds)
        log("cpu baseline: %.4g residue-steps/s (%s)" % (cpu["value"], cpu["sample"]))

    import upside_engine
    b = upside_engine.Batch([path], [TEMP], seed=SEED, dt=DT)
    probe = upside_engine.Upside(path, quiet=True)
    if warm_rounds:
        b.run(warm_rounds)
    tune = None
    if args.team > 0:
        b.set_team(args.team)
        b.set_bp_layout(max(args.bp_layout, 0))
        b.set_lane_split(args.lane_split)
    else:
        cand = [t for t in (1, 2, 4, 8, 16, 24, 28, 32) if t <= b.max_team()]
        tune = b.autotune_layout(6, cand, (0, 1, 2, 3))  # restores the warmed state after each candidate
        log("team / BP layout autotune (ms per 6 rounds): %s -> team %d layout %d lane split %d"
            % (tune, b.team(), b.bp_layout(), b.lane_split()))
    probe.energy(b.pos(0))
    g0, bp0, nodes = graph_counts(probe, node_names)
    small = small_node_bytes(probe, node_names, set(g0))
    st0 = b.bp_stats(0)
    h0 = b.bp_histogram(0)

    hip_sync()
    t0 = time.perf_counter()
    b.run(rounds)
    hip_sync()
    wall = time.perf_counter() - t0
    kernel_ms = b.last_run_ms()

    st1 = b.bp_stats(0)
    hist = b.bp_histogram(0) - h0
    pos_end = b.pos(0)
    probe.energy(pos_end)
    g1, bp1, _ = graph_counts(probe, node_names)
    graphs = {k: dict(shape=g0[k]["shape"], n_edge=0.5 * (g0[k]["n_edge"] + g1[k]["n_edge"])) for k in g0}
    bp = [0.5 * (x + y) for x, y in zip(bp0, bp1)]
    steps = 3 * rounds
    solves = st1["solves"] - st0["solves"]
    iters = st1["total"] - st0["total"]
    passes = iters + solves          # the first undamped pass plus chunked passes
    alg, per_step, per_pass = algorithmic_bytes(graphs, bp, nodes, n_atom, small, steps, passes)
    achieved = alg / (kernel_ms * 1e-3) / 1e9
    tps = traffic_per_step()
    value = N_RES * steps / wall

    checks = None
    if not args.no_cpu_baseline:
        e_end, d_end = float(probe.energy(pos_end)), probe.deriv(pos_end)
        er, fr = reference_errors(path, pos_end, e_end, d_end)
        eo, fo, xor, xor_own = oracle_checks(path, pos_end, e_end, d_end, probe, list(g1))
        checks = {"energy_rel_err_vs_reference": er, "force_rel_err_vs_reference": fr,
                  "energy_rel_err_vs_oracle_f64": eo, "force_rel_err_vs_oracle_f64": fo,
                  "edge_set_xor_vs_oracle": xor, "edge_set_xor_vs_oracle_own_bead_coordinates": xor_own}

    line = {
        "metric": METRIC, "value": value, "unit": "residue-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 3 * warm_rounds, "ms_per_step": wall * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "one %d-residue seeded random-sequence chain (seed %d), random-Rama start, "
                               "ff_2.1, FP32, T=%.2f, dt=%.3f, thermostat every round" % (N_RES, SEED, TEMP, DT),
                   "n_res": N_RES, "n_atom": n_atom, "launch": "one persistent launch of all steps",
                   "workgroups_per_system": b.team(), "bp_layout": b.bp_layout(),
                   "lane_split": b.lane_split()},
        "bp_iters_per_step": iters / max(solves, 1),
        "bp_passes_per_solve": passes / max(solves, 1),
        "bp_iters_per_solve_p50_p99": [percentile_from_hist(hist, 0.5), percentile_from_hist(hist, 0.99)],
        "bp_solves_in_lds_fraction": (st1["lds"] - st0["lds"]) / max(solves, 1),
        "bp_edges_3x3_3x6_6x6": [bp0[3:6], bp1[3:6]],
        "force_rel_err_vs_reference_at_end": checks["force_rel_err_vs_reference"] if checks else None,
        "checks_at_end": checks,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (tps * steps) if tps is not None else None,
                     "algorithmic_bytes_per_step": per_step, "algorithmic_bytes_per_bp_pass": per_pass,
                     "kernel_ms": kernel_ms, "kernel": "upside::engine_step_kernel"},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_baseline"] = value / cpu["value"]
    if not args.no_batched:
        line["batched"], uneven = batched_workloads(work, cpu=not args.no_cpu_baseline)
        if not args.no_warm_start:
            info, wl = warm_start_workloads(work, uneven)
            line["warm_start"] = dict(info, workloads=wl)
    if cpu:
        cpu["nproc"], cpu["cpu_model"] = host_info()
    shutil.rmtree(work, ignore_errors=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
