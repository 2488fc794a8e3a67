"""The learned cost model and arena regrouping (no counterpart in the reference).

A 48 -> 64 -> 64 -> 2 MLP predicts, from an arena's state, the log of its decision's cost and a contact logit. Its labels
come from the simulator itself: `label_costs` steps a batch with one arena per wave (block 1), so each lane's wall clock
(`rsim_sim8_timed`) is that arena's own cost, and reads back whether the arena ended the decision in contact. Training
runs the library's own kernels (forward/backward in fp32 with the inference kernel's bf16 rounding points, a batch
reduction, Adam); inference runs on bf16 MFMA.

Regrouping reorders arenas into waves by predicted cost, physically (the AoSoA-64 states are gathered into a work buffer,
controls gathered, result words scattered back), so every lane still reads its arena with coalesced accesses. Results
do not depend on arena order, so a regrouped decision is word-identical to the plain one.
"""
import numpy as np

from . import (CONTROL_WORDS, SIM8_BLOCK, _check, _ptr, _vp, _words, lib)

FEATURES = 48
PARAMS = 7426
ACT_WORDS = 306
W1, B1, W2, B2, W3, B3 = 0, 3072, 3136, 7232, 7296, 7424
LOG_SCALE = (10.0, 2.0)  # normalised target = (log(cycles) - 10) / 2


def init_params(seed=0):
    """He-initialised parameters (fp32 words)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[W1:B1] = rng.standard_normal(48 * 64) * np.sqrt(2.0 / 48)
    p[W2:B2] = rng.standard_normal(64 * 64) * np.sqrt(2.0 / 64)
    p[W3:B3] = rng.standard_normal(64 * 2) * np.sqrt(1.0 / 64)
    return p


def normalise_cycles(cycles):
    return ((np.log(np.maximum(cycles.astype(np.float64), 1.0)) - LOG_SCALE[0]) / LOG_SCALE[1]).astype(np.float32)


class CostModel:
    def __init__(self, dev, params=None, seed=0):
        self.dev = dev
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = dev.upload_words(p.view(np.uint32))
        self.m = dev.alloc_words(PARAMS)
        self.v = dev.alloc_words(PARAMS)
        self.grad = dev.alloc_words(PARAMS)
        self.step = 0

    def host_params(self):
        return self.dev.read_words(self.params).view(np.float32).copy()

    def features(self, states, out=None):
        """Features of a device state buffer (a DeviceWords of n x 48 fp32)."""
        n = states.n
        out = out if out is not None else self.dev.alloc_words(max(n, 1) * FEATURES)
        _check(lib().rsim_cost_features(self.dev.handle, states.handle, out.handle), "rsim_cost_features")
        return out

    def predict(self, features, n, out=None):
        out = out if out is not None else self.dev.alloc_words(max(n, 1) * 2)
        _check(lib().rsim_cost_predict(self.dev.handle, self.params.handle, features.handle, n, out.handle), "rsim_cost_predict")
        return out

    def gradients(self, feat, target, contact, batch, scratch=None, loss=None):
        """Per-row losses and the parameter gradient of the mean loss for one batch (device buffers)."""
        scratch = scratch if scratch is not None else self.dev.alloc_words(batch * ACT_WORDS)
        loss = loss if loss is not None else self.dev.alloc_words(batch)
        _check(lib().rsim_cost_gradients(self.dev.handle, self.params.handle, feat.handle, target.handle, contact.handle, batch,
                                         scratch.handle, loss.handle, self.grad.handle), "rsim_cost_gradients")
        return scratch, loss

    def train(self, feat, target, contact, epochs=30, batch=4096, lr=3e-3, seed=0, log=None):
        """Adam on host arrays (features n x 48, normalised log-cost targets n, contact flags n). Returns the loss curve."""
        rng = np.random.default_rng(seed)
        n = feat.shape[0]
        batch = min(batch, n)
        d_f, d_t, d_c = (self.dev.alloc_words(batch * FEATURES), self.dev.alloc_words(batch), self.dev.alloc_words(batch))
        scratch, loss = self.dev.alloc_words(batch * ACT_WORDS), self.dev.alloc_words(batch)
        curve = []
        for ep in range(epochs):
            order = rng.permutation(n)
            tot, cnt = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                idx = order[s:s + batch]
                self.dev.write_words(d_f, np.ascontiguousarray(feat[idx], np.float32).view(np.uint32))
                self.dev.write_words(d_t, np.ascontiguousarray(target[idx], np.float32).view(np.uint32))
                self.dev.write_words(d_c, np.ascontiguousarray(contact[idx], np.float32).view(np.uint32))
                self.gradients(d_f, d_t, d_c, batch, scratch, loss)
                self.step += 1
                _check(lib().rsim_cost_adam(self.dev.handle, self.params.handle, self.m.handle, self.v.handle, self.grad.handle,
                                            lr, self.step), "rsim_cost_adam")
                tot += float(self.dev.read_words(loss).view(np.float32).sum())
                cnt += batch
            curve.append(tot / max(cnt, 1))
            if log:
                log("epoch %d loss %.4f" % (ep, curve[-1]))
        return curve


def label_costs(dev, world, states, controls, decisions=1):
    """Per-arena cost labels from the simulator: each arena stepped in a wave of its own (block 1) with its lane's wall
    clock recorded, for `decisions` decisions from `states` (host array). Returns (features n*decisions x 48, normalised
    log cycles, contact flags) with features taken before each decision."""
    n = states.shape[0]
    d = dev.upload_states(states)
    r = dev.alloc_words(n)
    cyc = dev.alloc_words(2 * n)
    model = CostModel.__new__(CostModel)
    model.dev = dev
    feats, targets, contacts = [], [], []
    for k in range(decisions):
        f = model.features(d)
        dev.sync()
        feats.append(dev.read_words(f, n * FEATURES).view(np.float32).reshape(n, FEATURES).copy())
        c = dev.upload_controls(controls[k])
        _check(lib().rsim_sim8_timed(dev.handle, d.handle, world.handle, None, c.handle, r.handle, 1, None, cyc.handle),
               "rsim_sim8_timed")
        dev.sync()
        targets.append(normalise_cycles(dev.read_words(cyc).view(np.uint64)))
        after = model.features(d)
        dev.sync()
        fa = dev.read_words(after, n * FEATURES).view(np.float32).reshape(n, FEATURES)
        contacts.append((fa[:, 46] > 0).astype(np.float32))  # a manifold retained after the decision
    return np.concatenate(feats), np.concatenate(targets), np.concatenate(contacts)


def order_by_cost(pred_cost, mode, wave=64):
    """Arena order from predicted costs: 'sorted' (cost-homogeneous waves, most expensive first) or 'balanced' (the sorted
    arenas dealt round-robin over the waves, so every wave gets one of each cost rank). Returns perm: slot i <- arena perm[i]."""
    n = pred_cost.shape[0]
    srt = np.argsort(-pred_cost, kind="stable").astype(np.uint32)
    if mode == "sorted":
        return srt
    if mode == "balanced":
        waves = (n + wave - 1) // wave
        slots = np.arange(n)
        # rank r goes to wave r % waves, lane r // waves
        dest = (slots % waves) * wave + slots // waves
        perm = np.empty(n, np.uint32)
        order = np.argsort(dest, kind="stable")
        perm[:] = srt[order]
        return perm
    raise ValueError(mode)


class Regrouper:
    """Steps decisions of n arenas in a cost-regrouped order: the arenas live in a work buffer in the current order; each
    decision re-orders them by the model's prediction (gather), gathers the controls, steps, and scatters the result words
    back to arena order. `states_in_arena_order` un-permutes."""

    def __init__(self, dev, model, world, states, mode="sorted", block=SIM8_BLOCK):
        self.dev, self.model, self.world, self.mode, self.block = dev, model, world, mode, block
        self.n = n = states.n
        self.cur = states  # work buffer, slot i holds arena self.where[i]
        self.spare = dev.alloc_states(n)
        self.where = np.arange(n, dtype=np.uint32)
        self.perm = dev.alloc_words(n)
        self.feat = dev.alloc_words(n * FEATURES)
        self.pred = dev.alloc_words(2 * n)
        self.ctl = dev.alloc_words(n * CONTROL_WORDS)
        self.res = dev.alloc_words(n)

    def regroup(self):
        n = self.n
        self.model.features(self.cur, self.feat)
        self.model.predict(self.feat, n, self.pred)
        cost = self.dev.read_words(self.pred).view(np.float32).reshape(n, 2)[:, 0]
        rel = order_by_cost(cost, self.mode)  # new slot i <- current slot rel[i]
        _check(lib().rsim_set_permutation(self.dev.handle, _ptr(np.ascontiguousarray(rel)), n, self.perm.handle),
               "rsim_set_permutation")
        _check(lib().rsim_gather_states(self.dev.handle, self.cur.handle, self.perm.handle, self.spare.handle), "rsim_gather_states")
        self.cur, self.spare = self.spare, self.cur
        self.where = self.where[rel]
        _check(lib().rsim_set_permutation(self.dev.handle, _ptr(self.where), n, self.perm.handle), "rsim_set_permutation")

    def step(self, controls, result):
        """One decision: controls (DeviceWords, arena order) and result (DeviceWords, arena order)."""
        n = self.n
        _check(lib().rsim_permute_words(self.dev.handle, controls.handle, self.perm.handle, n, CONTROL_WORDS, 0, self.ctl.handle),
               "rsim_permute_words")
        self.dev.sim8(self.cur, self.world, self.ctl, self.res, self.block)
        _check(lib().rsim_permute_words(self.dev.handle, self.res.handle, self.perm.handle, n, 1, 1, result.handle),
               "rsim_permute_words")

    def states_in_arena_order(self):
        inv = np.empty(self.n, np.uint32)
        inv[self.where] = np.arange(self.n, dtype=np.uint32)
        _check(lib().rsim_set_permutation(self.dev.handle, _ptr(inv), self.n, self.perm.handle), "rsim_set_permutation")
        out = self.dev.alloc_states(self.n)
        _check(lib().rsim_gather_states(self.dev.handle, self.cur.handle, self.perm.handle, out.handle), "rsim_gather_states")
        return out
