"""The learned cost model: its kernels against a numpy float64 reference of the same network (bf16 rounding at the same
points), the reference's gradient against finite differences, and regrouping's exactness (a regrouped decision is
word-identical to the plain one)."""
import sys

import numpy as np
import pytest

import rsim.cost as cm
from conftest import first_diff

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
    """The host statement of the device order, against brute force: a stable sort by (bucket, slot), and the balanced deal
    puts one arena of every cost rank in each wave."""
    rng = np.random.default_rng(2)
    for n in (1, 5, 63, 64, 65, 1000, 4099):
        cost = (rng.integers(-40, 240, n) / 64.0).astype(np.float32)  # many ties, values past both ends of the range
        cost[rng.random(n) < 0.02] = np.nan
        b = cm.cost_buckets(cost)
        for i in rng.choice(n, min(n, 50), replace=False):  # the bucket from its definition
            c = float(cost[i])
            want = 255 if np.isnan(c) else int(min(255, max(0, np.floor((3.0 - c) * 64.0))))
            assert b[i] == want, (c, b[i], want)
        s = cm.order_by_cost(cost, "sorted")
        assert s.tolist() == sorted(range(n), key=lambda i: (int(b[i]), i))
        bal = cm.order_by_cost(cost, "balanced")
        assert sorted(bal.tolist()) == list(range(n))
        rank = np.empty(n, np.int64)
        rank[s] = np.arange(n)
        W = (n + 63) // 64
        for w in range(W):  # wave w holds ranks w, w + W, w + 2W, ... (the partial last wave: its first rounds only)
            r = np.sort(rank[bal[w * 64:(w + 1) * 64]])
            assert r[0] == w and (n <= 64 or (np.diff(r) <= W).all()), (n, w, r[:4])


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


@pytest.mark.gpu
def test_device_order_matches_host(rsim, device):
    note("rsim_cost_order vs its host statement, sizes 1..70001, both modes")
    rng = np.random.default_rng(11)
    for n in (1, 63, 65, 1000, 4099, 70001):
        pred = np.zeros((n, 2), np.float32)
        pred[:, 0] = (rng.integers(-40, 240, n) / 64.0).astype(np.float32) + rng.random(n).astype(np.float32) * 0.01
        pred[rng.random(n) < 0.01, 0] = np.nan
        d_pred = device.upload_words(pred.view(np.uint32).ravel())
        work = device.alloc_words(int(rsim.lib().rsim_cost_order_work_words(n)))
        perm = device.alloc_words(n)
        for mode, code in cm.ORDER_MODES.items():
            rsim._check(rsim.lib().rsim_cost_order(device.handle, d_pred.handle, n, code, work.handle, perm.handle), "order")
            device.sync()
            got = device.read_words(perm)
            want = cm.order_by_cost(pred[:, 0], mode)
            assert np.array_equal(got, want), (n, mode, np.flatnonzero(got != want)[:5])


@pytest.mark.gpu
def test_graph_replay_matches_eager(rsim, world, device, dev_world):
    note("3 sim8 decisions captured in a HIP graph, replayed 3 times vs eager, 2048 arenas")
    n = 2048
    states = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 151, n)
    ctl = [device.upload_controls(rsim.sample_controls(152, k, n)) for k in range(3)]
    init = device.upload_states(states)
    work = device.alloc_states(n)
    res = [device.alloc_words(n) for _ in range(3)]
    device.copy_states(work, init)
    for k in range(3):
        device.sim8(work, dev_world, ctl[k], res[k])
    device.sync()
    want_res = [device.read_words(r) for r in res]
    want = device.download_states(work)
    device.capture_begin()
    for k in range(3):
        device.sim8(work, dev_world, ctl[k], res[k])
    graph = device.capture_end()
    for rep in range(3):
        for r in res:
            device.write_words(r, np.zeros(n, np.uint32))
        device.copy_states(work, init)  # eager work between replays
        device.sync()
        graph.launch()
        device.sync()
        for k in range(3):
            assert np.array_equal(device.read_words(res[k]), want_res[k]), (rep, k)
        assert np.array_equal(device.download_states(work)[:, :2567], want[:, :2567]), rep
    graph.free()


@pytest.mark.gpu
def test_regrouper_in_a_graph(rsim, world, device, dev_world):
    note("two regrouped decisions captured in a graph, replayed twice (4 decisions) vs plain, 4096 arenas")
    n = 4096
    states = rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 161, n)
    ctl = [rsim.sample_controls(162, k, n) for k in range(4)]
    plain = device.upload_states(states)
    r = device.alloc_words(n)
    want_res = []
    for c in ctl:
        device.sim8(plain, dev_world, device.upload_controls(c), r)
        device.sync()
        want_res.append(device.read_words(r))
    want = device.download_states(plain)
    model = cm.CostModel(device, seed=12)
    for mode in ("sorted", "balanced"):
        g = cm.Regrouper(device, model, dev_world, device.upload_states(states), mode)
        stage = [device.alloc_words(n * rsim.CONTROL_WORDS) for _ in range(2)]
        out = [device.alloc_words(n) for _ in range(2)]
        device.sync()
        device.capture_begin()
        for k in range(2):
            g.regroup()
            g.step(stage[k], out[k])
        graph = g.dev.capture_end()
        for rep in range(2):
            for k in range(2):
                device.write_words(stage[k], ctl[2 * rep + k])
            graph.launch()
            device.sync()
            for k in range(2):
                assert np.array_equal(device.read_words(out[k]), want_res[2 * rep + k]), (mode, rep, k)
        got = device.download_states(g.states_in_arena_order())
        assert np.array_equal(got[:, :2567], want[:, :2567]), mode
        graph.free()


@pytest.mark.gpu
def test_trained_model_predicts_cost_better_than_untrained(rsim, world, device, dev_world):
    """What training gains: on held-out arenas, the correlation of predicted and measured log cost (lane clocks of the
    simulator itself, one arena per wave) rises well above the untrained network's."""
    note("cost model trained on simulator labels vs untrained, held-out correlation")
    n = 4096
    f, t, c = cm.label_costs(device, dev_world, rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 171, n),
                             [rsim.sample_controls(172, d, n) for d in range(4)], decisions=4)
    hf, ht, _ = cm.label_costs(device, dev_world, rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 173, 2048),
                               [rsim.sample_controls(174, d, 2048) for d in range(2)], decisions=2)
    d_hf = device.upload_words(hf.view(np.uint32).ravel())

    def corr(model):
        out = model.predict(d_hf, hf.shape[0])
        device.sync()
        p = device.read_words(out, 2 * hf.shape[0]).view(np.float32).reshape(-1, 2)[:, 0]
        return float(np.corrcoef(p, ht)[0, 1])

    untrained = cm.CostModel(device, seed=13)
    before = corr(untrained)
    model = cm.CostModel(device, seed=13)
    curve = model.train(f, t, c, epochs=15, batch=2048, lr=3e-3)
    after = corr(model)
    print("[case] held-out pearson: untrained %.3f trained %.3f, loss %.3f -> %.3f" % (before, after, curve[0], curve[-1]))
    assert after > 0.35 and after > before + 0.2, (before, after)


@pytest.mark.gpu
def test_device_batched_training_equals_host_batches(rsim, device):
    """CostModel.train shuffles on the device and trains on views into the shuffled copy; the same epochs with the
    batches gathered on the host and uploaded one by one give bit-identical parameters."""
    note("training on device-side batches (rsim_words_view) vs host-gathered batches")
    rng = np.random.default_rng(31)
    n, batch, epochs = 5000, 512, 2  # 9 full batches per epoch, the remainder dropped
    feat = rng.standard_normal((n, cm.FEATURES)).astype(np.float32)
    target = rng.standard_normal(n).astype(np.float32)
    contact = (rng.random(n) < 0.3).astype(np.float32)
    a = cm.CostModel(device, seed=4)
    a.train(feat, target, contact, epochs=epochs, batch=batch, lr=3e-3, seed=9)
    b = cm.CostModel(device, seed=4)
    order_rng = np.random.default_rng(9)
    d_f, d_t, d_c = device.alloc_words(batch * cm.FEATURES), device.alloc_words(batch), device.alloc_words(batch)
    scratch, loss = device.alloc_words(batch * cm.ACT_WORDS), device.alloc_words(batch)
    for _ in range(epochs):
        order = order_rng.permutation(n)
        for s in range(0, n - batch + 1, batch):
            idx = order[s:s + batch]
            device.write_words(d_f, np.ascontiguousarray(feat[idx]).view(np.uint32).ravel())
            device.write_words(d_t, np.ascontiguousarray(target[idx]).view(np.uint32))
            device.write_words(d_c, np.ascontiguousarray(contact[idx]).view(np.uint32))
            b.gradients(d_f, d_t, d_c, batch, scratch, loss)
            b.step += 1
            rsim._check(rsim.lib().rsim_cost_adam(device.handle, b.params.handle, b.m.handle, b.v.handle, b.grad.handle, 3e-3,
                                                  b.step), "rsim_cost_adam")
    device.sync()
    assert a.step == b.step == epochs * (n // batch)
    assert np.array_equal(a.host_params().view(np.uint32), b.host_params().view(np.uint32))


@pytest.mark.gpu
def test_words_view(rsim, device):
    note("rsim_words_view: a range of a buffer as a buffer of its own")
    base = device.upload_words(np.arange(100, dtype=np.uint32))
    v = base.view(10, 20)
    assert v.n == 20 and np.array_equal(device.read_words(v), np.arange(10, 30, dtype=np.uint32))
    device.write_words(v, np.full(20, 7, np.uint32))
    got = device.read_words(base)
    assert np.array_equal(got[10:30], np.full(20, 7)) and np.array_equal(got[:10], np.arange(10)) and got[30] == 30
    perm = device.upload_words(np.arange(5, dtype=np.uint32)[::-1].copy())
    dst = base.view(50, 5)  # held: a view's handle is freed with its Python object
    rsim._check(rsim.lib().rsim_permute_words(device.handle, base.handle, perm.handle, 5, 1, 0, dst.handle),
                "rsim_permute_words")
    assert np.array_equal(device.read_words(base)[50:55], np.array([4, 3, 2, 1, 0], np.uint32))
    del v
    assert np.array_equal(device.read_words(base)[:10], np.arange(10))  # freeing a view leaves the memory alone
    with pytest.raises(rsim.RsimError):
        base.view(90, 11)


def _regrouped_eager(rsim, device, model, dev_world, states, ctl, mode):
    g = cm.Regrouper(device, model, dev_world, device.upload_states(states), mode)
    r = device.alloc_words(states.shape[0])
    res = []
    for c in ctl:
        d_c = device.upload_controls(c)
        g.regroup()
        g.step(d_c, r)
        device.sync()
        res.append(device.read_words(r))
    return device.download_states(g.states_in_arena_order()), res


@pytest.mark.gpu
def test_regrouped_decisions_equal_the_oracle(rsim, oracle, world, device, dev_world):
    """The learned-regrouping workload's mix (1:1 open play and contact) stepped in cost-regrouped order, against the
    oracle directly: sorted and balanced, eager, and one regroup + decision captured in a graph and replayed three times
    (an odd count: every buffer keeps its role across replays). Every result word of every decision and every persistent
    word of the final states."""
    n, D = 4096, 3
    note("regrouped decisions vs the oracle, 1:1 open play and contact, %d arenas x %d" % (n, D))
    states = rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 181, n)
    ctl = [rsim.sample_controls(182, k, n) for k in range(D)]
    cpu = states.copy()
    want_res = [oracle.sim8(cpu, world, c, threads=8) for c in ctl]
    model = cm.CostModel(device, seed=14)
    for mode in ("sorted", "balanced"):
        got, res = _regrouped_eager(rsim, device, model, dev_world, states, ctl, mode)
        for k in range(D):
            assert np.array_equal(res[k], want_res[k]), "eager %s decision %d" % (mode, k)
        fd = first_diff(got[:, :2567], cpu[:, :2567])
        assert fd is None, "eager %s: arena %d word %d" % ((mode,) + fd)
        g = cm.Regrouper(device, model, dev_world, device.upload_states(states), mode)
        stage, out = device.alloc_words(n * rsim.CONTROL_WORDS), device.alloc_words(n)
        device.sync()
        device.capture_begin()
        g.regroup()
        g.step(stage, out)
        graph = device.capture_end()
        for k in range(D):
            device.write_words(stage, ctl[k])
            graph.launch()
            device.sync()
            assert np.array_equal(device.read_words(out), want_res[k]), "graph %s replay %d" % (mode, k)
        got = device.download_states(g.states_in_arena_order())
        fd = first_diff(got[:, :2567], cpu[:, :2567])
        assert fd is None, "graph %s: arena %d word %d" % ((mode,) + fd)
        graph.free()


@pytest.mark.gpu
def test_regrouped_full_size(rsim, oracle, world, device, dev_world):
    """At the learned-regrouping workload's full size (65,536 arenas, 1:1 open play and contact), two decisions regrouped
    by a trained-shape model in sorted order equal the plain order and the oracle, word for word."""
    n, D = 65536, 2
    note("regrouped vs plain vs oracle, %d arenas x %d decisions" % (n, D))
    states = rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 191, n)
    ctl = [rsim.sample_controls(192, k, n) for k in range(D)]
    model = cm.CostModel(device, seed=15)
    got, res = _regrouped_eager(rsim, device, model, dev_world, states, ctl, "sorted")
    plain = device.upload_states(states)
    r = device.alloc_words(n)
    for k, c in enumerate(ctl):
        device.sim8(plain, dev_world, device.upload_controls(c), r)
        device.sync()
        assert np.array_equal(device.read_words(r), res[k]), "decision %d" % k
    assert np.array_equal(device.download_states(plain)[:, :2567], got[:, :2567])
    cpu = states.copy()
    for k, c in enumerate(ctl):
        assert np.array_equal(oracle.sim8(cpu, world, c, threads=16), res[k]), "oracle decision %d" % k
    fd = first_diff(got[:, :2567], cpu[:, :2567])
    assert fd is None, "arena %d word %d" % fd
