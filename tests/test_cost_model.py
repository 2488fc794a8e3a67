"""The learned cost model: its kernels against a numpy float64 reference of the same network (bf16 rounding at the same
points), the reference's gradient against finite differences, and regrouping's exactness (a regrouped decision is
word-identical to the plain one)."""
import sys

import numpy as np
import pytest

import rsim.cost as cm

# ---- the float64 reference --------------------------------------------------------------------------------------------------


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[cm.W1:cm.B1].reshape(48, 64), p[cm.B1:cm.W2], p[cm.W2:cm.B2].reshape(64, 64), p[cm.B2:cm.W3],
            p[cm.W3:cm.B3].reshape(64, 2), p[cm.B3:])


def ref_forward(p, feat, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = r(feat)
    h1 = r(np.maximum(x @ r(w1) + b1, 0.0))
    h2 = r(np.maximum(h1 @ r(w2) + b2, 0.0))
    y = h2 @ r(w3) + b3
    return x, h1, h2, y


def ref_loss_grad(p, feat, target, contact, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x, h1, h2, y = ref_forward(p, feat, rounding)
    b = feat.shape[0]
    e = y[:, 0] - target
    sig = 1.0 / (1.0 + np.exp(-y[:, 1]))
    loss = e * e + (np.logaddexp(0.0, y[:, 1]) - contact * y[:, 1])
    dy = np.stack([2.0 * e / b, (sig - contact) / b], 1)
    dh2 = (dy @ r(w3).T) * (h2 > 0)
    dh1 = (dh2 @ r(w2).T) * (h1 > 0)
    g = np.concatenate([(x.T @ dh1).ravel(), dh1.sum(0), (h1.T @ dh2).ravel(), dh2.sum(0), (h2.T @ dy).ravel(), dy.sum(0)])
    return loss, g


def test_reference_gradient_matches_finite_differences():
    """The reference's backward pass (without rounding) against central differences of its mean loss: a wrong index, sign
    or transposition in the formulas fails this."""
    rng = np.random.default_rng(0)
    p = cm.init_params(1).astype(np.float64)
    feat = rng.standard_normal((16, 48)) * 0.5
    target = rng.standard_normal(16)
    contact = (rng.random(16) < 0.5).astype(np.float64)
    _, g = ref_loss_grad(p, feat, target, contact, rounding=False)
    for k in list(rng.choice(cm.PARAMS, 60, replace=False)) + [cm.B1, cm.B3, cm.B3 + 1, cm.W3 + 7]:
        h = 1e-6
        q1, q2 = p.copy(), p.copy()
        q1[k] += h
        q2[k] -= h
        l1 = ref_loss_grad(q1, feat, target, contact, rounding=False)[0].mean()
        l2 = ref_loss_grad(q2, feat, target, contact, rounding=False)[0].mean()
        assert abs((l1 - l2) / (2 * h) - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), k


def test_order_by_cost():
    rng = np.random.default_rng(2)
    cost = rng.standard_normal(1000).astype(np.float32)
    s = cm.order_by_cost(cost, "sorted")
    assert sorted(s.tolist()) == list(range(1000)) and (np.diff(cost[s]) <= 0).all()
    b = cm.order_by_cost(cost, "balanced")
    assert sorted(b.tolist()) == list(range(1000))
    waves = [cost[b[i:i + 64]] for i in range(0, 1000, 64)]
    assert max(w.max() for w in waves) - min(w.max() for w in waves[:-1]) < max(cost) - np.median(cost)  # spread out


# ---- the kernels --------------------------------------------------------------------------------------------------------


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
def test_mfma_inference_matches_reference(rsim, device):
    note("cost model MFMA inference vs float64 reference, 1000 rows")
    rng = np.random.default_rng(3)
    n = 1000
    p = cm.init_params(4)
    p[cm.B1:cm.W2] = rng.standard_normal(64) * 0.1
    feat = np.clip(rng.standard_normal((n, 48)), -4, 4).astype(np.float32)
    model = cm.CostModel(device, p)
    f = device.upload_words(feat.view(np.uint32).ravel())
    out = model.predict(f, n)
    device.sync()
    got = device.read_words(out, 2 * n).view(np.float32).reshape(n, 2)
    want = ref_forward(p, feat)[3]
    # sums in another order (MFMA) can move a hidden activation by one bf16 step; the outputs stay within 2%
    assert np.max(np.abs(got - want)) <= 0.02 * max(1.0, np.max(np.abs(want))), np.max(np.abs(got - want))
    assert np.mean(np.abs(got - want)) < 2e-3


@pytest.mark.gpu
def test_training_gradients_match_reference(rsim, device):
    note("cost model gradients vs float64 reference, batch 512")
    rng = np.random.default_rng(5)
    b = 512
    p = cm.init_params(6)
    feat = np.clip(rng.standard_normal((b, 48)), -4, 4).astype(np.float32)
    target = rng.standard_normal(b).astype(np.float32)
    contact = (rng.random(b) < 0.3).astype(np.float32)
    model = cm.CostModel(device, p)
    d = [device.upload_words(a.view(np.uint32).ravel()) for a in (feat, target, contact)]
    _, loss = model.gradients(d[0], d[1], d[2], b)
    device.sync()
    got_g = device.read_words(model.grad).view(np.float32).astype(np.float64)
    got_l = device.read_words(loss).view(np.float32)
    want_l, want_g = ref_loss_grad(p, feat, target.astype(np.float64), contact.astype(np.float64))
    assert np.allclose(got_l, want_l, rtol=1e-4, atol=1e-5)
    assert np.allclose(got_g, want_g, rtol=2e-3, atol=2e-6), np.max(np.abs(got_g - want_g))


@pytest.mark.gpu
def test_training_learns(rsim, device):
    note("cost model training on a learnable target")
    rng = np.random.default_rng(7)
    n = 8192
    feat = np.clip(rng.standard_normal((n, 48)), -4, 4).astype(np.float32)
    target = (0.8 * feat[:, 0] - 0.5 * feat[:, 3] * feat[:, 4] + 0.3 * np.abs(feat[:, 10])).astype(np.float32)
    contact = (feat[:, 1] > 0.2).astype(np.float32)
    model = cm.CostModel(device, seed=8)
    curve = model.train(feat, target, contact, epochs=25, batch=1024, lr=3e-3)
    assert curve[-1] < 0.4 * curve[0], curve


@pytest.mark.gpu
def test_regrouped_decisions_are_identical(rsim, oracle, world, device, dev_world):
    note("regrouped decisions (sorted and balanced) vs plain, 4096 arenas x 3")
    n = 4096
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 131, n)
    ctl = [rsim.sample_controls(132, k, n) for k in range(3)]
    plain = device.upload_states(states)
    r = device.alloc_words(n)
    plain_res = []
    for c in ctl:
        device.sim8(plain, dev_world, device.upload_controls(c), r)
        device.sync()
        plain_res.append(device.read_words(r))
    want = device.download_states(plain)
    model = cm.CostModel(device, seed=9)
    for mode in ("sorted", "balanced"):
        g = cm.Regrouper(device, model, dev_world, device.upload_states(states), mode)
        for k, c in enumerate(ctl):
            g.regroup()
            g.step(device.upload_controls(c), r)
            device.sync()
            assert np.array_equal(device.read_words(r), plain_res[k]), "%s decision %d" % (mode, k)
        got = device.download_states(g.states_in_arena_order())
        assert np.array_equal(got[:, :2567], want[:, :2567]), mode


@pytest.mark.gpu
def test_features_and_labels(rsim, world, device):
    note("features and per-arena cost labels, 512 arenas")
    n = 512
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 141, n)
    dw = device.upload_static(world)
    feat, target, contact = cm.label_costs(device, dw, states, [rsim.sample_controls(142, 0, n)], 1)
    assert feat.shape == (n, 48) and np.isfinite(feat).all() and (np.abs(feat) <= 4).all()
    bp = states[:, 4:7].view(np.float32)
    assert np.allclose(feat[:, 0], np.clip(bp[:, 0] / 4096.0, -4, 4), rtol=1e-6)
    assert np.isfinite(target).all() and target.std() > 0  # arenas differ in cost
    assert set(np.unique(contact)) <= {0.0, 1.0}
