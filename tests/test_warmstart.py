"""The BP warm-start network: the CPU oracle pinned by closed forms, central
differences and torch's bfloat16 rounding (CPU tests), and the GPU kernels
(matrix-core forward, backward and training) against the oracle (gpu tests).
The network has no counterpart in the reference; what it starts is the
reference's BP solve (src/rotamer.cpp:1106-1162)."""
import numpy as np
import pytest

import mlp_oracle as M
from conftest import say


def synthetic_rows(n, seed):
    """Feature rows shaped as the engine writes them: log probabilities and
    first-pass log beliefs (<= 0) over 3 or 6 states, their difference, the
    state mask, degree and class columns; targets are a distribution over the
    row's states."""
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 32), np.float32)
    t = np.zeros((n, 8), np.float32)
    for k in range(n):
        ns = 3 if rng.random() < 0.45 else 6
        lp = np.zeros(6)
        lb = np.zeros(6)
        lp[:ns] = -rng.exponential(1.5, ns)
        lb[:ns] = -rng.exponential(2.0, ns)
        lp[:ns] -= lp[:ns].max()
        lb[:ns] -= lb[:ns].max()
        f[k, 0:6] = lp
        f[k, 6:12] = lb
        f[k, 12:18] = lb - lp
        f[k, 18:18 + ns] = 1.0
        f[k, 24] = rng.integers(1, 20) / 8.0
        f[k, 25] = ns == 6
        q = np.exp(lb[:ns] + rng.normal(0, 0.7, ns))
        t[k, :ns] = q / q.sum()
    return f, t


def random_params(seed, scale_out=0.3):
    rng = np.random.default_rng(seed)
    d = dict(W1=rng.normal(0, np.sqrt(2 / 32), (128, 32)), b1=rng.normal(0, 0.1, 128),
             W2=rng.normal(0, np.sqrt(2 / 128), (128, 128)), b2=rng.normal(0, 0.1, 128),
             W3=rng.normal(0, scale_out / np.sqrt(128), (16, 128)), b3=rng.normal(0, 0.1, 16))
    return M.pack(d).astype(np.float32)


# ------------------------------------------------------------------ CPU

def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-6, 1000), rng.normal(0, 1e4, 1000)]).astype(np.float32)
    # exact ties: float32 values with the low 16 bits 0x8000, both parities
    ties = (np.arange(1000, dtype=np.uint32) << 16 | 0x8000).astype(np.uint32) + 0x3F800000
    x = np.concatenate([x, ties.view(np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(M.bf16(x), want)


def test_oracle_loss_closed_form_with_zero_output_layer():
    """With W3 = b3 = 0 the network leaves the first-pass beliefs unchanged:
    the loss is the cross-entropy of softmax(f[6:12]) over each row's states."""
    f, t = synthetic_rows(50, 1)
    p = random_params(2)
    d = M.unpack(p)
    d["W3"][:] = 0
    d["b3"][:] = 0
    loss = M.loss_only(M.pack(d), f, t)
    want = 0.0
    for k in range(len(f)):
        ns = int(f[k, 18:24].sum())
        z = f[k, 6:6 + ns].astype(np.float64)
        want -= (t[k, :ns] * (z - np.log(np.exp(z).sum()))).sum()
    assert abs(loss - want / len(f)) < 1e-12


def test_oracle_gradient_matches_central_differences():
    """The backward formulas (quantize=False) against central differences of
    the float64 loss, for entries of every weight and bias block."""
    f, t = synthetic_rows(24, 3)
    p = random_params(4, scale_out=1.0).astype(np.float64)
    loss, g = M.loss_grad(p, f, t, quantize=False)
    rng = np.random.default_rng(5)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), 12, replace=False) for k in range(6)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (M.loss_only(pp, f, t) - M.loss_only(pm, f, t)) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1e-3, abs(num)) + 1e-9, (i, num, g[i])


# ------------------------------------------------------------------ GPU

@pytest.fixture(scope="module")
def ue():
    import upside_engine
    upside_engine.load_upside_library()
    return upside_engine


def block_errors(got, want):
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    out = {}
    for k, (name, _) in enumerate(M.SHAPES):
        a, b = got[offs[k]:offs[k + 1]], want[offs[k]:offs[k + 1]]
        out[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 1000, 4133])
def test_gpu_forward_matches_oracle(ue, n):
    """Matrix-core forward pass against the oracle's float64 pass over the
    same bf16-rounded operands: 99.5% of the outputs within 1e-5 of the
    largest output.  The rest are rows where a hidden activation lies within
    float32 rounding of a bf16 rounding boundary, so that the two sides round
    it one bf16 step apart; such a step moves an output by ~|W3| 2^-8 |h|,
    which bounds the largest error (1e-3 of the largest output)."""
    say("case: warm-start forward, %d rows" % n)
    f, _ = synthetic_rows(n, 10 + n)
    p = random_params(11)
    net = ue.WarmStartNet(seed=1)
    net.set_params(p)
    got = net.forward(f)
    want = M.forward(p.astype(np.float64), f)["y"]
    scale = np.abs(want).max()
    err = np.abs(got - want) / scale
    say("forward error: max %.2e, 99.5%% %.2e (of max |y| %.3f)" % (err.max(), np.quantile(err, 0.995), scale))
    assert np.quantile(err, 0.995) < 1e-5
    assert err.max() < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 3000])
def test_gpu_gradient_matches_oracle(ue, n):
    """Loss and gradient of the GPU backward kernels (matrix-core products,
    fixed-order reductions) against the oracle's float64 restatement of the
    same bf16 arithmetic: 1e-4 relative per block (rare bf16 rounding ties of
    intermediates that float32 and float64 place on different sides account
    for what is left)."""
    say("case: warm-start gradient, %d rows" % n)
    f, t = synthetic_rows(n, 20 + n)
    p = random_params(21, scale_out=1.0)
    net = ue.WarmStartNet(seed=1)
    net.set_params(p)
    net.set_data(f, t)
    loss, g = net.loss_grad()
    loss_o, g_o = M.loss_grad(p.astype(np.float64), f, t)
    errs = block_errors(g.astype(np.float64), g_o)
    say("loss %.6f oracle %.6f; gradient errors per block %s" % (loss, loss_o, errs))
    assert abs(loss - loss_o) <= 1e-6 * abs(loss_o)
    for k, v in errs.items():
        assert v < 1e-4, (k, v)
    # twice in a row: the fixed-order reductions give the same bits
    loss2, g2 = net.loss_grad()
    assert loss2 == loss and np.array_equal(g2, g)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(ue):
    """Adam on the GPU learns a correction that the data holds: targets are
    softmax(f[6:12] + 0.8 f[0:6]) (the 1-body term under-weighted in the
    first pass); from the zero output layer the loss falls to near the
    targets' own entropy."""
    say("case: warm-start training on a learnable correction")
    f, _ = synthetic_rows(6000, 30)
    t = np.zeros((len(f), 8), np.float32)
    for k in range(len(f)):
        ns = int(f[k, 18:24].sum())
        z = f[k, 6:6 + ns] + 0.8 * f[k, 0:ns]
        q = np.exp(z - z.max())
        t[k, :ns] = q / q.sum()
    ent = -np.where(t > 0, t * np.log(np.where(t > 0, t, 1)), 0).sum() / len(f)
    net = ue.WarmStartNet(seed=3)
    net.set_data(f, t)
    loss0, _ = net.loss_grad()
    losses = net.train(400, batch_rows=2048, lr=2e-3)
    loss1, _ = net.loss_grad()
    say("loss %.4f -> %.4f (target entropy %.4f); first/last batch %.4f %.4f" % (loss0, loss1, ent, losses[0], losses[-1]))
    assert np.all(np.isfinite(losses))
    assert loss1 - ent < 0.25 * (loss0 - ent)


# ------------------------------------------------------- engine (GPU)

def _chains(chain_file, lens, seed0):
    return [chain_file(int(L), seed0 + i) for i, L in enumerate(lens)]


@pytest.mark.gpu
def test_warm_start_trained_on_gpu_reduces_iterations(ue, chain_file):
    """The whole warm-start loop on the GPU: training rows harvested from
    converged plain solves (schedule 1) of uneven chains, the network
    trained by the matrix-core kernels, then held-out chains solved three
    ways from identical states: plain start, untrained network (first-pass
    beliefs) and trained network, all stopped by the residual rule (1e-5).
    The trained warm start takes fewer iterations per solve than the plain
    start and than the untrained network, and falls back nowhere.  Stopped
    by the reference's rule instead (schedule 3 against schedule 0, the
    node's tolerance 1e-3), the trained start also takes fewer iterations,
    and at one state its forces stay as close to the converged solve's as
    the plain start's do (within twice its difference, or 1e-4)."""
    rng = np.random.default_rng(7)
    train = _chains(chain_file, rng.integers(50, 401, 12), 300)
    b = ue.Batch(train, [0.85] * len(train), seed=3)
    b.set_team(min(4, b.max_team()))
    b.set_bp_schedule(1, 1e-5)
    F, T = [], []
    say("case: warm start, harvest from %d chains" % len(train))
    for frame in range(12):
        b.run(2)
        for s in range(len(train)):
            f, t = b.bp_training_rows(s)
            F.append(f)
            T.append(t)
    F, T = np.concatenate(F), np.concatenate(T)
    assert np.all(np.isfinite(F)) and np.allclose(T.sum(1), 1, atol=1e-5)
    net = ue.WarmStartNet(seed=1)
    net.set_data(F, T)
    l0, _ = net.loss_grad()
    net.train(2000, batch_rows=8192, lr=1e-3)
    l1, _ = net.loss_grad()
    say("harvested %d rows; training loss %.5f -> %.5f" % (len(F), l0, l1))
    assert l1 < l0
    held = _chains(chain_file, np.random.default_rng(8).integers(50, 401, 6), 400)
    runs = {}
    schedules = {"plain": 1, "untrained": 2, "trained": 2, "reference": 0, "trained, reference rule": 3}
    for name, sched in schedules.items():
        hb = ue.Batch(held, [0.85] * len(held), seed=4)
        hb.set_team(min(4, hb.max_team()))
        hb.set_bp_schedule(1, 1e-5)
        hb.run(1)  # identical first round: previous marginals exist for every schedule
        if sched >= 2:
            hb.set_warm_start(ue.WarmStartNet(seed=1).params() if name == "untrained" else net.params())
        if sched != 1:
            hb.set_bp_schedule(sched)
        st0 = [hb.bp_stats(s) for s in range(len(held))]
        h0 = sum(hb.bp_histogram(s) for s in range(len(held)))
        hb.run(6)
        st1 = [hb.bp_stats(s) for s in range(len(held))]
        hist = sum(hb.bp_histogram(s) for s in range(len(held))) - h0
        it = sum(a["total"] - c["total"] for a, c in zip(st1, st0))
        so = sum(a["solves"] - c["solves"] for a, c in zip(st1, st0))
        ws = [hb.bp_warm_stats(s) for s in range(len(held))]
        p99 = int(np.searchsorted(np.cumsum(hist), 0.99 * hist.sum()))
        runs[name] = (it / so, sum(w["fallbacks"] for w in ws), p99)
        say("%-23s iterations per solve %.2f, fallbacks %d, p99 %d" % (name, *runs[name]))
    assert runs["trained"][1] == 0 and runs["trained, reference rule"][1] == 0
    assert runs["trained"][0] < runs["plain"][0]
    assert runs["trained"][0] <= runs["untrained"][0]
    assert runs["trained, reference rule"][0] < runs["reference"][0]
    # accuracy at one state (after an identical first round) against converged solves
    from conftest import rel_rms
    forces = {}
    for sched in (1, 0, 3):
        hb = ue.Batch(held, [0.85] * len(held), seed=4)
        hb.set_team(min(4, hb.max_team()))
        hb.set_bp_schedule(1, 1e-5)
        hb.run(1)
        if sched == 3:
            hb.set_warm_start(net.params())
        hb.set_bp_schedule(sched)
        hb.energies()
        forces[sched] = [hb.deriv(s) for s in range(len(held))]
    e0 = max(rel_rms(a, c) for a, c in zip(forces[0], forces[1]))
    e3 = max(rel_rms(a, c) for a, c in zip(forces[3], forces[1]))
    say("forces against converged solves: reference rule plain start %.2e, warm start %.2e" % (e0, e3))
    assert e3 <= max(2 * e0, 1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("n_res", [60, 300])
def test_warm_start_free_energy_and_forces_match_converged_oracle(ue, chain_file, n_res):
    """A warm-started solve (a network with random output weights, so that
    its start is far from the first pass) converges to the BP fixed point:
    energy and forces within 1e-4 of the oracle's plain BP run to the same
    residual (float32 and float64), and of the engine's own plain start."""
    import upside_oracle as O
    from conftest import read_pos, rel_rms
    path = chain_file(n_res, 77)
    pos = read_pos(path)
    rng = np.random.default_rng(5)
    pos = (pos + rng.normal(0, 0.2, pos.shape)).astype(np.float32)
    p = random_params(12, scale_out=2.0)
    got = {}
    for name in ("plain", "warm"):
        b = ue.Batch([path], [0.85], seed=1)
        b.set_pos(0, pos)
        if name == "warm":
            b.set_warm_start(p)
            b.set_bp_schedule(2, 1e-5)
        else:
            b.set_bp_schedule(1, 1e-5)
        say("case: converged BP, %s start, L=%d" % (name, n_res))
        e = float(b.energies()[0])
        d = b.deriv(0)
        got[name] = (e, d, b.bp_stats(0)["last"], b.bp_warm_stats(0))
    e32 = O.Engine(path, bp_residual=1e-5)
    e64 = O.Engine(path, dtype=np.float64, bp_residual=1e-5)
    ref32 = (float(e32.energy(pos)), e32.deriv(pos))
    ref64 = (float(e64.energy(pos.astype(np.float64))), e64.deriv(pos.astype(np.float64)))
    for name, (e, d, it, ws) in got.items():
        for tag, (er, dr) in (("f32", ref32), ("f64", ref64)):
            de = abs(e - er) / max(abs(er), 1.0)
            df = rel_rms(d, dr)
            say("%s start (%d iterations, %s) vs oracle %s: energy %.2e force %.2e" % (name, it, ws, tag, de, df))
            assert de < 1e-4 and df < 1e-4, (name, tag, de, df)
    assert rel_rms(got["warm"][1], got["plain"][1]) < 1e-4
    assert got["warm"][3]["warm"] >= 1
