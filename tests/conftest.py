"""Shared test set-up: import paths, the `gpu` marker and seeded inputs.

Tests marked `gpu` need a visible GPU and call the HIP engine through its C
ABI; every other test runs on the CPU (oracle, host logic, library exports).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "upside2-md_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libupside_ref.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "upside_ref")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def updir(tmp_path_factory):
    return tmp_path_factory.mktemp("up")


@pytest.fixture(scope="session")
def chain_file(updir):
    """chain_file(n_res, seed, **build_kw) -> path of a seeded random chain."""
    import upconfig
    cache = {}

    def make(n_res, seed=1, **kw):
        key = (n_res, seed, tuple(sorted(kw.items())))
        if key not in cache:
            tag = "_".join("%s%s" % (k, v) for k, v in sorted(kw.items()))
            path = str(updir / ("chain%d_s%d%s.up" % (n_res, seed, ("_" + tag) if tag else "")))
            upconfig.make_chain(path, n_res, seed=seed, **kw)
            cache[key] = path
        return cache[key]
    return make


@pytest.fixture(scope="session")
def chig(updir):
    """Chignolin (GYDPETGTWG) built from the committed fixture: (path, fixture)."""
    import upconfig
    z = np.load(os.path.join(GOLDEN, "chig_frames.npz"))
    seq = [s.decode() for s in z["sequence"]]
    g = upconfig.build_ff21(seq, z["initial_pos"], rama_pot=z["rama_pot"], normalize_rama=False)
    path = str(updir / "chig.up")
    upconfig.write_up(path, g)
    return path, z


def read_pos(path):
    import h5lite
    with h5lite.File(path) as f:
        return np.ascontiguousarray(f.read("/input/pos")[:, :, 0], dtype=np.float32)


def rel_rms(a, b):
    """||a-b|| / ||b|| (the reference's relative_rms_deviation, src/deriv_engine.h:377-389)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def ref_lib():
    """ctypes handle of the reference engine built from its own sources, or None."""
    import ctypes as ct
    if not os.path.exists(REF_LIB):
        return None
    lib = ct.CDLL(REF_LIB)
    lib.construct_deriv_engine.restype = ct.c_void_p
    lib.construct_deriv_engine.argtypes = [ct.c_int, ct.c_char_p, ct.c_bool]
    for fn in ("evaluate_energy", "evaluate_deriv"):
        getattr(lib, fn).argtypes = [ct.c_void_p] * 3
    lib.get_value_by_name.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_char_p, ct.c_char_p]
    lib.free_deriv_engine.argtypes = [ct.c_void_p]
    return lib


class RefEngine:
    """The reference engine (oracle/_ref) over one .up file."""

    def __init__(self, lib, path, n_atom):
        import ctypes as ct
        self.ct = ct
        self.lib = lib
        self.n_atom = n_atom
        self.h = lib.construct_deriv_engine(n_atom, path.encode(), True)
        assert self.h

    def energy(self, pos):
        e = np.zeros(1, np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        assert self.lib.evaluate_energy(e.ctypes.data, self.ct.c_void_p(self.h), pos.ctypes.data) == 0
        return float(e[0])

    def deriv(self, pos):
        d = np.zeros((self.n_atom, 3), np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        assert self.lib.evaluate_deriv(d.ctypes.data, self.ct.c_void_p(self.h), pos.ctypes.data) == 0
        return d

    def value(self, shape, node, name):
        out = np.zeros(shape, np.float32)
        rc = self.lib.get_value_by_name(int(np.prod(shape)), out.ctypes.data, self.ct.c_void_p(self.h),
                                        node.encode(), name.encode())
        assert rc == 0
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.free_deriv_engine(self.ct.c_void_p(self.h))
            self.h = None


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def say(*a):
    """Name the case before a GPU step, flushed, so a failing run says which."""
    print(*a, flush=True)
