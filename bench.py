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


def oracle_checks(path, pos, energy_gpu, deriv_gpu, probe, graph_names):
    """At the final positions: relative energy and force errors against the
    float64 oracle, and the XOR count of every pair graph's edge set (GPU list
    against the float32 oracle's refined list; must be 0).  Part of the CPU
    checker leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import upside_oracle
    ora = upside_oracle.Engine(path, dtype=np.float64)
    p64 = pos.astype(np.float64)
    e64 = float(ora.energy(p64))
    d64 = ora.deriv(p64)
    ora32 = upside_oracle.Engine(path)      # edge sets are defined in float32
    ora32.energy(np.ascontiguousarray(pos, np.float32))
    xor = {}
    for name in graph_names:
        node = ora32.nodes.get(name)
        if node is None:
            continue
        obj = node.ig if hasattr(node, "ig") else node
        if not hasattr(obj, "i1"):
            continue
        want = set(zip(obj.i1.tolist(), obj.i2.tolist()))
        got = set(map(tuple, probe.edge_pairs(name).tolist()))
        xor[name] = len(got ^ want)
    return (abs(energy_gpu - e64) / max(abs(e64), 1.0),
            float(np.sqrt(((deriv_gpu - d64) ** 2).sum() / (d64 ** 2).sum())), xor)


def percentile_from_hist(h, q):
    """Smallest k with cumulative count >= q * total."""
    c = np.cumsum(h)
    if c[-1] == 0:
        return None
    return int(np.searchsorted(c, q * c[-1]))


# ------------------------------------------------------- algorithmic bytes
def graph_counts(eng, node_names):
    """Edge counts of every interaction graph and BP class counts at the
    engine's last evaluation."""
    out = {}
    for name in node_names:
        try:
            shape = eng.get_value_by_name((4,), name, "graph_shape")
        except RuntimeError:
            continue
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
def batched_workloads(work, rounds=60):
    """Two batch workloads reported beside the single-chain line (not `value`):
    16 replicas of one 150-residue chain at temperatures geometric over
    0.80-0.95 with alternating neighbour swap sets every 7 rounds (21 steps),
    and 64 independent chains with lengths uniform in [50, 400].  Each is
    autotuned, warmed up, then timed over `rounds` rounds."""
    import upside_engine
    out = []

    def timed(b, n_res_total, segment=None, sets=None):
        cand = [t for t in (1, 2, 4, 8, 16) if t <= b.max_team()]
        b.run(3)
        b.autotune(3, cand)
        hip_sync()
        t0 = time.perf_counter()
        done, k = 0, 0
        while done < rounds:
            n = min(segment or rounds, rounds - done)
            b.run(n)
            done += n
            if sets:
                b.replica_exchange(sets[k % 2], done)
                k += 1
        hip_sync()
        wall = time.perf_counter() - t0
        return n_res_total * 3 * rounds / wall, wall, b.team()

    p = os.path.join(work, "rex150.up")
    upconfig.make_chain(p, 150, seed=2)
    temps = [0.80 * (0.95 / 0.80) ** (i / 15.0) for i in range(16)]
    b = upside_engine.Batch([p] * 16, temps, seed=2, dt=DT)
    sets = [[[i, i + 1] for i in range(0, 15, 2)], [[i, i + 1] for i in range(1, 14, 2)]]
    v, wall, team = timed(b, 16 * 150, segment=7, sets=sets)
    out.append({"workload": "16 replicas x 150 residues, T geometric 0.80-0.95, neighbour swaps every 21 steps",
                "value": v, "unit": "residue-steps/s", "steps": 3 * rounds, "seconds": wall,
                "workgroups_per_system": team})
    del b
    rng = np.random.default_rng(3)
    lens = [int(x) for x in rng.integers(50, 401, 64)]
    paths = []
    for i, L in enumerate(lens):
        q = os.path.join(work, "uneven%02d.up" % i)
        upconfig.make_chain(q, L, seed=100 + i)
        paths.append(q)
    b = upside_engine.Batch(paths, [TEMP] * 64, seed=3, dt=DT)
    v, wall, team = timed(b, sum(lens))
    out.append({"workload": "64 independent chains, lengths uniform in [50, 400] (%d residues)" % sum(lens),
                "value": v, "unit": "residue-steps/s", "steps": 3 * rounds, "seconds": wall,
                "workgroups_per_system": team})
    return out


# ----------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cpu-rounds", type=int, default=400, help="rounds of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the two batch workloads")
    ap.add_argument("--team", type=int, default=0,
                    help="workgroups per system (0: time 1..32 on the warmed state and keep the fastest)")
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
    else:
        cand = [t for t in (1, 2, 4, 8, 16, 32) if t <= b.max_team()]
        tune = b.autotune(6, cand)       # restores the warmed state after each candidate
        log("team autotune (ms per 6 rounds): %s -> team %d" % (tune, b.team()))
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
        eo, fo, xor = oracle_checks(path, pos_end, e_end, d_end, probe, list(g1))
        checks = {"energy_rel_err_vs_reference": er, "force_rel_err_vs_reference": fr,
                  "energy_rel_err_vs_oracle_f64": eo, "force_rel_err_vs_oracle_f64": fo,
                  "edge_set_xor_vs_oracle": xor}

    line = {
        "metric": METRIC, "value": value, "unit": "residue-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": 3 * warm_rounds, "ms_per_step": wall * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "one %d-residue seeded random-sequence chain (seed %d), random-Rama start, "
                               "ff_2.1, FP32, T=%.2f, dt=%.3f, thermostat every round" % (N_RES, SEED, TEMP, DT),
                   "n_res": N_RES, "n_atom": n_atom, "launch": "one persistent launch of all steps",
                   "workgroups_per_system": b.team()},
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
        line["batched"] = batched_workloads(work)
    shutil.rmtree(work, ignore_errors=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
