"""The learned proposal on the GPU: gradients of the FP32 training kernels and
logits of the BF16 MFMA inference against numpy, training, and the proposal's
effect on the chain (it only removes objects; the exact cuts decide)."""
import numpy as np
import pytest

import oracle
from sdl_amd import gen, hip, proposal as P, runtime

pytestmark = pytest.mark.gpu


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, P.N_IN)).astype(np.float32)
    Y = (X[:, 0] + 0.5 * X[:, 3] - 0.3 * X[:, 7] > 0.2).astype(np.float32)
    return X, Y


def test_gradient_matches_cpu():
    print("case mlp gradient", flush=True)
    X, Y = _rows(1000, 1)
    flat = P.init_params(seed=2)
    m = P.MLP(flat)
    dX, dY = hip.DeviceArray(X), hip.DeviceArray(Y)
    g, loss = m.gradient(dX.address, dY.address, len(Y), 4.0)
    ref_loss, ref_g = P.loss_grad_np(flat, X, Y, 4.0)
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss) + 1e-6
    err = np.abs(g - ref_g)
    assert np.max(err) <= 1e-5 + 2e-3 * np.max(np.abs(ref_g)), np.max(err)
    assert np.linalg.norm(g - ref_g) <= 1e-3 * np.linalg.norm(ref_g)


def test_bf16_mfma_forward_matches_cpu():
    print("case mlp forward", flush=True)
    X, _ = _rows(3001, 4)  # not a multiple of 16
    flat = P.init_params(seed=6)
    m = P.MLP(flat)
    dX = hip.DeviceArray(X)
    out = hip.DeviceArray.empty(len(X), np.float32)
    m.forward(dX.address, len(X), out.address)
    z = out.to_host()
    z_bf = P.forward_np(flat, X, bf16=True)
    z_fp = P.forward_np(flat, X, bf16=False)
    scale = np.max(np.abs(z_fp))
    # the bf16 emulation differs only by accumulation order and the bf16 rounding of the output layer's input
    assert np.max(np.abs(z - z_bf)) <= 0.02 * scale
    assert np.max(np.abs(z - z_fp)) <= 0.05 * scale
    assert np.corrcoef(z, z_fp)[0, 1] > 0.999


def test_training_learns_a_separable_rule():
    print("case mlp training", flush=True)
    X, Y = _rows(8192, 7)
    m = P.MLP(P.init_params(seed=8))
    dX, dY = hip.DeviceArray(X), hip.DeviceArray(Y)
    perm = hip.DeviceArray(np.random.default_rng(9).permutation(len(Y)).astype(np.int32))
    _, loss0 = m.gradient(dX.address, dY.address, len(Y), 1.0)
    for epoch in range(6):
        m.train(dX.address, dY.address, perm.address, 0, 512, len(Y) // 512, 3e-3, 1.0)
    _, loss1 = m.gradient(dX.address, dY.address, len(Y), 1.0)
    assert loss1 < 0.5 * loss0
    out = hip.DeviceArray.empty(len(X), np.float32)
    m.forward(dX.address, len(X), out.address)
    acc = np.mean((out.to_host() > 0) == (Y > 0.5))
    assert acc > 0.95, acc


def _chain(ev, e):
    ev.reset()
    ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    ev.run_chain()
    ev.sync()
    return ev.export_all(0)


def test_proposal_only_removes_objects():
    print("case proposal subset", flush=True)
    wl = 12
    e = gen.event(wl, 0)
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1])
    ev = runtime.Event(g)
    plain = _chain(ev, e)
    ref, _ = oracle.run(gen.geometry(gen.geom_kind(wl))[1], e, oracle.CR)
    X, Y, n = P.pair_rows(ev)
    y = Y.to_host()[:n]
    assert n == ev.tl_work()["pairs"]
    # a pair (inner SG, outer SG's inner MD) is positive iff some tracklet starts with it
    positive_pairs = {tuple(t[:6]) for t in plain["TL"]}
    assert int((y == 1).sum()) == len(positive_pairs)
    # rows of pairs that fail the first stage are marked -1 (never scored)
    assert set(np.unique(y).tolist()) <= {-1.0, 0.0, 1.0}
    # keep everything: identical to the plain chain
    keep_all = np.zeros(P.N_PARAMS, np.float32)
    keep_all[-1] = 10.0
    m = P.MLP(keep_all)
    P.set_proposal(ev, m, 0.5)
    got = _chain(ev, e)
    for s in oracle.STAGES:
        assert np.array_equal(oracle.as_multiset(got[s]), oracle.as_multiset(ref[s])), s
    # drop everything: no tracklets (and no track candidates), earlier stages unchanged
    drop_all = keep_all.copy()
    drop_all[-1] = -10.0
    m2 = P.MLP(drop_all)
    P.set_proposal(ev, m2, 0.5)
    got = _chain(ev, e)
    assert len(got["TL"]) == 0 and len(got["TCP"]) == 0 and len(got["TCL"]) == 0
    for s in ("MD", "SG", "TP"):
        assert np.array_equal(oracle.as_multiset(got[s]), oracle.as_multiset(ref[s])), s
    # a random network: every object kept is an object of the plain chain
    m3 = P.MLP(P.init_params(seed=11))
    P.set_proposal(ev, m3, 0.5)
    got = _chain(ev, e)
    for s in ("TL", "TCP", "TCL"):
        a = {tuple(r) for r in got[s]}
        b = {tuple(r) for r in ref[s]}
        assert a <= b, s
    P.set_proposal(ev, None, 0.5)


def test_proposal_result_independent_of_launch_shape():
    print("case proposal launch shapes", flush=True)
    wl = 12
    e = gen.event(wl, 1)
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1])
    ev = runtime.Event(g)
    m = P.MLP(P.init_params(seed=13))
    P.set_proposal(ev, m, 0.4)
    base = ev.tuning()
    got = []
    for waves, b1 in ((4, 64), (2, 256), (8, 128)):
        ev.set_tuning(dict(base, scored_waves=waves, tl1_block=b1))
        got.append(_chain(ev, e))
    for s in ("TL", "TCP", "TCL"):
        assert np.array_equal(oracle.as_multiset(got[0][s]), oracle.as_multiset(got[1][s])), s
        assert np.array_equal(oracle.as_multiset(got[0][s]), oracle.as_multiset(got[2][s])), s
    P.set_proposal(ev, None, 0.5)
