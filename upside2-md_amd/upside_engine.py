"""ctypes binding of libupside.so, mirroring the reference's py/upside_engine.py.

The class `Upside` keeps the reference's method names, argument meanings and
error behaviour (py/upside_engine.py:186-276): it reads /input/pos and
/input/sequence of a `.up` file, constructs an engine through
construct_deriv_engine and raises RuntimeError when a call returns nonzero.
`Batch` wraps the batched extension of include/upside_engine.h.

The library is the HIP engine; there is no CPU fallback.  If it is missing or
no GPU is visible, constructing an engine raises.
"""
import ctypes as ct
import os

import numpy as np

try:
    from . import h5lite
except ImportError:
    import h5lite

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libupside.so")

_calc = None


def load_upside_library(path=LIB_PATH):
    """Load libupside.so and declare the C ABI (py/upside_engine.py:21-91)."""
    global _calc
    if _calc is not None:
        return _calc
    if not os.path.exists(path):
        raise OSError("libupside.so not built at %s (run build() / make -C upside2-md_amd)" % path)
    c = ct.cdll.LoadLibrary(path)
    vp, i, f, cp = ct.c_void_p, ct.c_int, ct.c_float, ct.c_char_p
    sig = {
        "construct_deriv_engine": (vp, [i, cp, ct.c_bool]),
        "free_deriv_engine": (None, [vp]),
        "evaluate_energy": (i, [vp, vp, vp]),
        "evaluate_deriv": (i, [vp, vp, vp]),
        "set_param": (i, [i, vp, vp, cp]),
        "get_param_deriv": (i, [i, vp, vp, cp]),
        "get_param": (i, [i, vp, vp, cp]),
        "get_output_dims": (i, [vp, vp, vp, cp]),
        "get_sens": (i, [i, vp, vp, cp]),
        "get_output": (i, [i, vp, vp, cp]),
        "get_value_by_name": (i, [i, vp, vp, cp, cp]),
        "get_clamped_value_and_deriv": (i, [i, vp, vp, i, vp]),
        "clamped_spline_value": (i, [i, vp, vp, i, vp]),
        "clamped_spline_solve": (i, [i, vp, vp]),
        "get_clamped_coeff_deriv": (i, [i, vp, vp, f]),
        "upside_batch_create": (vp, [i, vp, vp, ct.c_uint32, f, i]),
        "upside_batch_free": (None, [vp]),
        "upside_batch_n_atom": (i, [vp, i]),
        "upside_batch_set_pos": (i, [vp, i, vp]),
        "upside_batch_get_pos": (i, [vp, i, vp]),
        "upside_batch_set_mom": (i, [vp, i, vp]),
        "upside_batch_get_mom": (i, [vp, i, vp]),
        "upside_batch_run": (i, [vp, i, i, vp, i]),
        "upside_batch_energies": (i, [vp, vp]),
        "upside_batch_get_deriv": (i, [vp, i, vp]),
        "upside_batch_replica_exchange": (i, [vp, i, vp, ct.c_uint64, vp]),
        "upside_batch_bp_stats": (i, [vp, i, vp, vp, vp, vp]),
        "upside_batch_potential_terms": (i, [vp, i, vp, i, ct.c_char_p, i]),
        "upside_batch_replica_exchange_criterion": (i, [vp, i, vp, ct.c_uint64, i, vp]),
        "upside_batch_bp_histogram": (i, [vp, i, vp, i]),
        "upside_batch_bp_lds_solves": (i, [vp, i, vp]),
        "upside_batch_open_output": (i, [vp, vp, i, i]),
        "upside_batch_run_logged": (i, [vp, i, i]),
        "upside_batch_close_output": (i, [vp]),
        "upside_batch_set_monte_carlo": (i, [vp, i]),
        "upside_batch_set_recentering": (i, [vp, i]),
        "upside_batch_set_thermostat_timescale": (i, [vp, f]),
        "upside_batch_set_annealing": (i, [vp, ct.c_double, ct.c_double, ct.c_double, ct.c_double]),
        "upside_batch_set_replica_exchange": (i, [vp, i, i, i, vp, vp]),
        "upside_batch_replica_stats": (i, [vp, vp, vp, vp]),
        "upside_batch_mc_stats": (i, [vp, i, vp, vp]),
        "upside_batch_state_size": (ct.c_longlong, [vp]),
        "upside_batch_save_state": (i, [vp, vp, ct.c_longlong]),
        "upside_batch_load_state": (i, [vp, vp, ct.c_longlong]),
        "upside_batch_last_run_ms": (f, [vp]),
        "upside_batch_enable_profile": (i, [vp, i]),
        "upside_batch_max_team": (i, [vp]),
        "upside_batch_set_team": (i, [vp, i]),
        "upside_batch_set_teams": (i, [vp, vp, i]),
        "upside_batch_set_lane_split": (i, [vp, i]),
        "upside_batch_get_lane_split": (i, [vp]),
        "upside_batch_get_teams": (i, [vp, vp, i]),
        "upside_batch_balanced_teams": (i, [vp, i, vp, i]),
        "upside_batch_get_team": (i, [vp]),
        "upside_batch_autotune": (i, [vp, i, vp, i, vp]),
        "upside_batch_autotune_layout": (i, [vp, i, vp, i, vp, i, vp]),
        "upside_batch_set_bp_layout": (i, [vp, i]),
        "upside_batch_get_bp_layout": (i, [vp]),
        "upside_batch_profile_report": (i, [vp, i, ct.c_char_p, i]),
        "upside_batch_set_bp_schedule": (i, [vp, i, f, i, i]),
        "upside_batch_set_warm_start": (i, [vp, vp, i]),
        "upside_batch_bp_training_rows": (i, [vp, i, vp, vp, i]),
        "upside_batch_bp_warm_stats": (i, [vp, i, vp, vp, vp]),
        "upside_mlp_create": (vp, [ct.c_uint32]),
        "upside_mlp_free": (None, [vp]),
        "upside_mlp_n_param": (i, []),
        "upside_mlp_get_params": (i, [vp, vp, i]),
        "upside_mlp_set_params": (i, [vp, vp, i]),
        "upside_mlp_set_data": (i, [vp, vp, vp, i]),
        "upside_mlp_loss_grad": (i, [vp, vp, vp]),
        "upside_mlp_train": (i, [vp, i, i, f, vp]),
        "upside_mlp_forward": (i, [vp, vp, i, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(c, name)
        fn.restype = res
        fn.argtypes = args
    _calc = c
    return c


def _f4(x, shape=None):
    x = np.require(x, dtype="f4", requirements="C")
    if shape is not None:
        assert x.shape == shape, (x.shape, shape)
    return x


class Upside:
    """Force engine over one `.up` file (py/upside_engine.py:186-276)."""

    def __init__(self, config_file_path, quiet=False):
        self.calc = load_upside_library()
        self.config_file_path = str(config_file_path)
        with h5lite.File(self.config_file_path) as t:
            self.initial_pos = t.read("/input/pos")[:, :, 0]
            self.n_atom = self.initial_pos.shape[0]
            self.sequence = t.read("/input/sequence") if t.exists("/input/sequence") else None
        self.engine = self.calc.construct_deriv_engine(self.n_atom, self.config_file_path.encode(), bool(quiet))
        if not self.engine:
            raise RuntimeError("Unable to initialize upside engine for %s" % (config_file_path,))

    def __repr__(self):
        return "Upside(%r, %r)" % (self.n_atom, self.config_file_path)

    def energy(self, pos):
        pos = _f4(pos, (self.n_atom, 3))
        energy = np.zeros(1, dtype="f4")
        if self.calc.evaluate_energy(energy.ctypes.data, self.engine, pos.ctypes.data):
            raise RuntimeError("Unable to evaluate energy")
        return energy[0]

    def deriv(self, pos):
        pos = _f4(pos, (self.n_atom, 3))
        deriv = np.zeros_like(pos)
        if self.calc.evaluate_deriv(deriv.ctypes.data, self.engine, pos.ctypes.data):
            raise RuntimeError("Unable to evaluate derivative")
        return deriv

    def set_param(self, param, node_name):
        p = _f4(np.asarray(param).ravel())
        if self.calc.set_param(int(p.shape[0]), p.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to set param with size %s for node %s" % (np.shape(param), node_name))

    def get_param_deriv(self, param_shape, node_name):
        deriv = np.zeros(param_shape, dtype="f4")
        if self.calc.get_param_deriv(int(np.prod(param_shape)), deriv.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to get param deriv")
        return deriv

    def get_param(self, param_shape, node_name):
        param = np.zeros(param_shape, dtype="f4")
        if self.calc.get_param(int(np.prod(param_shape)), param.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to get param")
        return param

    def _dims(self, node_name):
        n_elem = np.zeros(1, dtype=np.intc)
        width = np.zeros(1, dtype=np.intc)
        if self.calc.get_output_dims(n_elem.ctypes.data, width.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to get output dims")
        return int(n_elem[0]), int(width[0])

    def get_sens(self, node_name):
        shape = self._dims(node_name)
        out = np.zeros(shape, dtype="f4")
        if self.calc.get_sens(out.size, out.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to get output")
        return out

    def get_output(self, node_name):
        shape = self._dims(node_name)
        out = np.zeros(shape, dtype="f4")
        if self.calc.get_output(out.size, out.ctypes.data, self.engine, node_name.encode()):
            raise RuntimeError("Unable to get output")
        return out

    def get_value_by_name(self, value_shape, node_name, log_name):
        value = np.zeros(value_shape, dtype="f4")
        if self.calc.get_value_by_name(int(np.prod(value_shape)), value.ctypes.data, self.engine,
                                       node_name.encode(), log_name.encode()):
            raise RuntimeError("Unable to get value by name")
        return value

    def edge_pairs(self, node_name):
        """Sorted (i1, i2) element pairs of a pair-graph node's current edge set."""
        n = int(self.get_value_by_name((1,), node_name, "n_edge")[0])
        if n == 0:
            return np.zeros((0, 2), np.int64)
        return self.get_value_by_name((n, 2), node_name, "edge_pairs").astype(np.int64)

    def __del__(self):
        if getattr(self, "engine", None):
            self.calc.free_deriv_engine(self.engine)
            self.engine = None


class Batch:
    """Many systems integrated in one persistent launch (extension ABI)."""

    def __init__(self, up_paths, temperatures, seed=1, dt=0.009, thermostat_interval=1):
        self.calc = load_upside_library()
        self.n_sys = len(up_paths)
        arr = (ct.c_char_p * self.n_sys)(*[str(p).encode() for p in up_paths])
        temps = _f4(np.asarray(temperatures, dtype="f4").reshape(self.n_sys))
        self._keep = (arr, temps)
        self.h = self.calc.upside_batch_create(self.n_sys, ct.cast(arr, ct.c_void_p), temps.ctypes.data,
                                               int(seed), float(dt), int(thermostat_interval))
        if not self.h:
            raise RuntimeError("Unable to create batch")
        self.n_atom = [self.calc.upside_batch_n_atom(self.h, s) for s in range(self.n_sys)]

    def pos(self, s):
        out = np.zeros((self.n_atom[s], 3), "f4")
        if self.calc.upside_batch_get_pos(self.h, s, out.ctypes.data):
            raise RuntimeError("get_pos failed")
        return out

    def set_pos(self, s, pos):
        pos = _f4(pos, (self.n_atom[s], 3))
        if self.calc.upside_batch_set_pos(self.h, s, pos.ctypes.data):
            raise RuntimeError("set_pos failed")

    def mom(self, s):
        out = np.zeros((self.n_atom[s], 3), "f4")
        if self.calc.upside_batch_get_mom(self.h, s, out.ctypes.data):
            raise RuntimeError("get_mom failed")
        return out

    def set_mom(self, s, mom):
        mom = _f4(mom, (self.n_atom[s], 3))
        if self.calc.upside_batch_set_mom(self.h, s, mom.ctypes.data):
            raise RuntimeError("set_mom failed")

    def run(self, n_round, frame_interval=0, n_frames=0):
        frames = np.zeros((self.n_sys, max(n_frames, 0)), "f4")
        ptr = frames.ctypes.data if n_frames > 0 else None
        if self.calc.upside_batch_run(self.h, int(n_round), int(frame_interval), ptr, int(n_frames)):
            raise RuntimeError("batch run failed")
        return frames

    def deriv(self, s):
        """dE/dx of system s at the last evaluation (call energies() first)."""
        d = np.zeros((self.calc.upside_batch_n_atom(self.h, s), 3), "f4")
        if self.calc.upside_batch_get_deriv(self.h, s, d.ctypes.data):
            raise RuntimeError("get_deriv failed")
        return d

    def energies(self):
        e = np.zeros(self.n_sys, "f4")
        if self.calc.upside_batch_energies(self.h, e.ctypes.data):
            raise RuntimeError("energies failed")
        return e

    def replica_exchange(self, pairs, round_num, criterion=0):
        """One swap attempt per pair (exchange criterion 0 or 1); returns accepted flags."""
        pairs = np.require(np.asarray(pairs, dtype=np.intc).reshape(-1, 2), requirements="C")
        acc = np.zeros(len(pairs), np.intc)
        if self.calc.upside_batch_replica_exchange_criterion(self.h, len(pairs), pairs.ctypes.data, int(round_num),
                                                             int(criterion), acc.ctypes.data):
            raise RuntimeError("replica exchange failed")
        return acc

    def bp_stats(self, s):
        last = ct.c_int(0)
        tot, sol, bad = ct.c_longlong(0), ct.c_longlong(0), ct.c_longlong(0)
        if self.calc.upside_batch_bp_stats(self.h, s, ct.byref(last), ct.byref(tot), ct.byref(sol), ct.byref(bad)):
            raise RuntimeError("bp stats failed")
        lds = ct.c_longlong(0)
        if self.calc.upside_batch_bp_lds_solves(self.h, s, ct.byref(lds)):
            raise RuntimeError("bp stats failed")
        return dict(last=last.value, total=tot.value, solves=sol.value, bad=bad.value, lds=lds.value)

    def open_output(self, paths=None, record_momentum=False, log_level=0):
        """Start /output logging: into each system's input .up (paths None) or
        into new files paths[s]; log_level 1 adds the detailed datasets."""
        arr = None
        if paths is not None:
            enc = [p.encode() for p in paths]
            arr = (ct.c_char_p * len(enc))(*enc)
        if self.calc.upside_batch_open_output(self.h, arr, int(bool(record_momentum)), int(log_level)):
            raise RuntimeError("open_output failed (see stderr)")

    def run_logged(self, n_round, frame_interval):
        """Advance n_round rounds, logging a frame at rounds that are multiples of frame_interval."""
        if self.calc.upside_batch_run_logged(self.h, int(n_round), int(frame_interval)):
            raise RuntimeError("run_logged failed (see stderr)")

    def set_monte_carlo(self, mc_interval):
        """Pivot moves every mc_interval rounds in run_logged (0: none)."""
        if self.calc.upside_batch_set_monte_carlo(self.h, int(mc_interval)):
            raise RuntimeError("set_monte_carlo failed (see stderr)")

    def set_replica_exchange(self, replica_interval, swap_sets, criterion=0):
        """Replica exchange in run_logged every replica_interval rounds over
        swap_sets (a list of lists of (i, j) pairs, tried in turn)."""
        sizes = np.array([len(s) for s in swap_sets], np.intc)
        pairs = np.ascontiguousarray(np.array([p for s in swap_sets for p in s], np.intc).reshape(-1, 2))
        if self.calc.upside_batch_set_replica_exchange(self.h, int(replica_interval), int(criterion), len(sizes),
                                                       sizes.ctypes.data, pairs.ctypes.data):
            raise RuntimeError("set_replica_exchange failed (see stderr)")
        self._n_pairs = int(sizes.sum())

    def replica_stats(self, n_pairs=None):
        """(replica index of each system, successes and attempts of every pair);
        n_pairs defaults to the pairs given to set_replica_exchange."""
        n = getattr(self, "_n_pairs", 0) if n_pairs is None else int(n_pairs)
        idx = np.zeros(self.n_sys, np.intc)
        suc, att = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
        if self.calc.upside_batch_replica_stats(self.h, idx.ctypes.data, suc.ctypes.data, att.ctypes.data):
            raise RuntimeError("replica_stats failed")
        return idx, suc[:n], att[:n]

    def set_thermostat_timescale(self, timescale):
        """Thermostat timescale (before the first round; re-thermalises)."""
        if self.calc.upside_batch_set_thermostat_timescale(self.h, float(timescale)):
            raise RuntimeError("set_thermostat_timescale failed (see stderr)")

    def set_annealing(self, factor, start=-1.0, end=-1.0, duration=-1.0):
        """Simulated annealing with the CLI's --anneal-* meanings (times in time units)."""
        if self.calc.upside_batch_set_annealing(self.h, float(factor), float(start), float(end), float(duration)):
            raise RuntimeError("set_annealing failed (see stderr)")

    def set_recentering(self, mode):
        """Recenter before each logged frame: 0 none, 1 x/y/z, 2 x/y only."""
        if self.calc.upside_batch_set_recentering(self.h, int(mode)):
            raise RuntimeError("set_recentering failed")

    def mc_stats(self, s):
        """(accepted, attempted) pivot moves of system s."""
        a, t = ct.c_longlong(0), ct.c_longlong(0)
        if self.calc.upside_batch_mc_stats(self.h, s, ct.byref(a), ct.byref(t)):
            raise RuntimeError("mc_stats failed")
        return a.value, t.value

    def close_output(self):
        if self.calc.upside_batch_close_output(self.h):
            raise RuntimeError("close_output failed")

    def save_state(self):
        """The whole device state of the batch as bytes (see load_state)."""
        n = self.calc.upside_batch_state_size(self.h)
        if n < 0:
            raise RuntimeError("state size failed")
        buf = ct.create_string_buffer(n)
        if self.calc.upside_batch_save_state(self.h, buf, n):
            raise RuntimeError("save_state failed")
        return buf.raw

    def load_state(self, state):
        """Resume from save_state() bytes of a batch made from the same .up files."""
        buf = ct.create_string_buffer(bytes(state), len(state))
        if self.calc.upside_batch_load_state(self.h, buf, len(state)):
            raise RuntimeError("load_state failed (see stderr)")

    def bp_histogram(self, s):
        """counts[k] = solves of system s that took k BP iterations (last bin: the rest)."""
        h = np.zeros(256, np.int64)
        if self.calc.upside_batch_bp_histogram(self.h, s, h.ctypes.data, len(h)) < 0:
            raise RuntimeError("bp histogram failed")
        return h

    def potential_terms(self, s):
        """{node name: potential} of system s at the last evaluation."""
        vals = np.zeros(256, "f4")
        names = ct.create_string_buffer(1 << 14)
        n = self.calc.upside_batch_potential_terms(self.h, s, vals.ctypes.data, len(vals), names, len(names))
        if n < 0:
            raise RuntimeError("potential terms failed")
        return dict(zip(names.value.decode().split("\n")[:n], vals[:n].tolist()))

    def last_run_ms(self):
        return float(self.calc.upside_batch_last_run_ms(self.h))

    def max_team(self):
        return int(self.calc.upside_batch_max_team(self.h))

    def team(self):
        return int(self.calc.upside_batch_get_team(self.h))

    def set_team(self, team):
        if self.calc.upside_batch_set_team(self.h, int(team)):
            raise RuntimeError("team size %d not possible" % team)

    def set_lane_split(self, split):
        """Sections on teams of eight or more: 4 quarters or 2 halves."""
        if self.calc.upside_batch_set_lane_split(self.h, int(split)):
            raise RuntimeError("lane split %d not possible" % split)

    def lane_split(self):
        return int(self.calc.upside_batch_get_lane_split(self.h))

    def set_teams(self, teams):
        """One team size per system (each team on one XCD, all resident)."""
        t = np.ascontiguousarray(teams, dtype=np.intc)
        if self.calc.upside_batch_set_teams(self.h, t.ctypes.data, len(t)):
            raise RuntimeError("per-system teams %s not possible" % list(t))

    def teams(self):
        t = np.zeros(self.n_sys, dtype=np.intc)
        if self.calc.upside_batch_get_teams(self.h, t.ctypes.data, len(t)):
            raise RuntimeError("get_teams failed")
        return t

    def balanced_teams(self, budget=0):
        """Team sizes in proportion to the systems' atom counts (not applied)."""
        t = np.zeros(self.n_sys, dtype=np.intc)
        if self.calc.upside_batch_balanced_teams(self.h, int(budget), t.ctypes.data, len(t)):
            raise RuntimeError("balanced_teams failed")
        return t

    def autotune(self, n_round=3, candidates=(1, 2, 4, 8, 16, 32)):
        """Time n_round rounds per team size from the current state, keep the
        fastest; returns {team: kernel ms} of the candidates that ran (0: the
        balanced per-system teams)."""
        cand = np.asarray(candidates, dtype=np.intc)
        ms = np.zeros(len(cand), dtype="f4")
        if self.calc.upside_batch_autotune(self.h, int(n_round), cand.ctypes.data, len(cand), ms.ctypes.data):
            raise RuntimeError("autotune failed")
        return {int(t): float(m) for t, m in zip(cand, ms) if m >= 0}

    def autotune_layout(self, n_round=3, teams=(1, 2, 4, 8, 16, 32), layouts=(0, 1)):
        """Time n_round rounds per (team size, rotamer-solve layout) pair from
        the current state, keep the fastest; returns {(team, layout): ms}."""
        t = np.asarray(teams, dtype=np.intc)
        lay = np.asarray(layouts, dtype=np.intc)
        ms = np.zeros(len(t) * len(lay), dtype="f4")
        if self.calc.upside_batch_autotune_layout(self.h, int(n_round), t.ctypes.data, len(t), lay.ctypes.data,
                                                  len(lay), ms.ctypes.data):
            raise RuntimeError("autotune failed")
        return {(int(a), int(b)): float(ms[k * len(lay) + l]) for k, a in enumerate(t) for l, b in enumerate(lay)
                if ms[k * len(lay) + l] >= 0}

    def bp_layout(self):
        return int(self.calc.upside_batch_get_bp_layout(self.h))

    def set_bp_layout(self, layout):
        if self.calc.upside_batch_set_bp_layout(self.h, int(layout)):
            raise RuntimeError("rotamer solve layout %d not possible" % layout)

    def set_bp_schedule(self, mode, residual=None, budget=200, klocal=4):
        """Rotamer BP schedule: 0 the reference's, 1 plain start with the
        residual stopping rule, 2 warm start (needs set_warm_start) with the
        residual rule and a plain restart after `budget` iterations, 3 the
        warm start stopped by the reference's rule.  `residual` is the
        residual rule's bound (default 1e-5) for 1 and 2, and for 3 the
        one-sided tolerance (default the rotamer node's own)."""
        if residual is None:
            residual = 0.0 if mode == 3 else 1e-5
        if self.calc.upside_batch_set_bp_schedule(self.h, int(mode), float(residual), int(budget), int(klocal)):
            raise RuntimeError("set_bp_schedule failed")

    def set_warm_start(self, params):
        p = _f4(params)
        if self.calc.upside_batch_set_warm_start(self.h, p.ctypes.data, len(p)):
            raise RuntimeError("set_warm_start failed")

    def bp_training_rows(self, s):
        """Features and converged marginals of system s's last solve (schedule
        1 or 2): (features (n, 32), targets (n, 8))."""
        n = self.calc.upside_batch_bp_training_rows(self.h, s, None, None, 0)
        if n < 0:
            raise RuntimeError("bp_training_rows failed")
        f = np.zeros((n, 32), "f4")
        t = np.zeros((n, 8), "f4")
        if self.calc.upside_batch_bp_training_rows(self.h, s, f.ctypes.data, t.ctypes.data, n) != n:
            raise RuntimeError("bp_training_rows failed")
        return f, t

    def bp_warm_stats(self, s):
        w, fb, r = ct.c_longlong(0), ct.c_longlong(0), ct.c_longlong(0)
        if self.calc.upside_batch_bp_warm_stats(self.h, s, ct.byref(w), ct.byref(fb), ct.byref(r)):
            raise RuntimeError("bp_warm_stats failed")
        return dict(warm=w.value, fallbacks=fb.value, residual=r.value)

    def enable_profile(self, on=True):
        if self.calc.upside_batch_enable_profile(self.h, int(on)):
            raise RuntimeError("enable_profile failed")

    def profile_report(self, s=0):
        buf = ct.create_string_buffer(1 << 16)
        r = self.calc.upside_batch_profile_report(self.h, s, buf, len(buf))
        if r != 0:
            raise RuntimeError("profile report failed")
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "h", None):
            self.calc.upside_batch_free(self.h)
            self.h = None


class WarmStartNet:
    """The BP warm-start network (no reference counterpart): parameters,
    training data and Adam state on the GPU; every product on the matrix
    cores (bf16 operands, float32 accumulation)."""

    N_IN, N_OUT = 32, 16

    def __init__(self, seed=1):
        self.calc = load_upside_library()
        self.h = self.calc.upside_mlp_create(int(seed))
        if not self.h:
            raise RuntimeError("upside_mlp_create failed")
        self.n_param = int(self.calc.upside_mlp_n_param())

    def params(self):
        p = np.zeros(self.n_param, "f4")
        if self.calc.upside_mlp_get_params(self.h, p.ctypes.data, self.n_param):
            raise RuntimeError("get_params failed")
        return p

    def set_params(self, p):
        p = _f4(p, (self.n_param,))
        if self.calc.upside_mlp_set_params(self.h, p.ctypes.data, self.n_param):
            raise RuntimeError("set_params failed")

    def set_data(self, feat, target):
        feat = _f4(feat)
        target = _f4(target)
        if feat.ndim != 2 or feat.shape[1] != self.N_IN or target.shape != (len(feat), 8):
            raise ValueError("features (n, 32) and targets (n, 8) expected")
        if self.calc.upside_mlp_set_data(self.h, feat.ctypes.data, target.ctypes.data, len(feat)):
            raise RuntimeError("set_data failed")

    def loss_grad(self):
        loss = np.zeros(1, "f4")
        g = np.zeros(self.n_param, "f4")
        if self.calc.upside_mlp_loss_grad(self.h, loss.ctypes.data, g.ctypes.data):
            raise RuntimeError("loss_grad failed")
        return float(loss[0]), g

    def train(self, n_steps, batch_rows=4096, lr=1e-3):
        losses = np.zeros(n_steps, "f4")
        if self.calc.upside_mlp_train(self.h, int(n_steps), int(batch_rows), float(lr), losses.ctypes.data):
            raise RuntimeError("train failed")
        return losses

    def forward(self, feat):
        feat = _f4(feat)
        out = np.zeros((len(feat), self.N_OUT), "f4")
        if self.calc.upside_mlp_forward(self.h, feat.ctypes.data, len(feat), out.ctypes.data):
            raise RuntimeError("forward failed")
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.calc.upside_mlp_free(self.h)
            self.h = None
