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
        """Adam on host arrays (features n x 48, normalised log-cost targets n, contact flags n), reshuffled every epoch,
        full batches only. The data go to the device once; each epoch gathers them into a shuffled copy there
        (rsim_permute_words) and the batches are views into it (rsim_words_view), so an epoch costs one permutation
        upload and one read of its losses. Returns the loss curve (the mean row loss of each epoch)."""
        rng = np.random.default_rng(seed)
        n = feat.shape[0]
        batch = min(batch, n)
        nb = n // batch if batch else 0
        if nb == 0:
            return []
        dev = self.dev
        widths = (FEATURES, 1, 1)
        src = [dev.upload_words(np.ascontiguousarray(x, np.float32).reshape(-1).view(np.uint32)) for x in (feat, target, contact)]
        shuffled = [dev.alloc_words(n * w) for w in widths]
        perm = dev.alloc_words(n)
        scratch, loss = dev.alloc_words(batch * ACT_WORDS), dev.alloc_words(nb * batch)
        views = [[buf.view(b * batch * w, batch * w) for b in range(nb)] for buf, w in zip(shuffled, widths)]
        loss_views = [loss.view(b * batch, batch) for b in range(nb)]
        curve = []
        for ep in range(epochs):
            order = np.ascontiguousarray(rng.permutation(n), np.uint32)
            _check(lib().rsim_set_permutation(dev.handle, _ptr(order), n, perm.handle), "rsim_set_permutation")
            for x, y, w in zip(src, shuffled, widths):
                _check(lib().rsim_permute_words(dev.handle, x.handle, perm.handle, n, w, 0, y.handle), "rsim_permute_words")
            for b in range(nb):
                self.gradients(views[0][b], views[1][b], views[2][b], batch, scratch, loss_views[b])
                self.step += 1
                _check(lib().rsim_cost_adam(dev.handle, self.params.handle, self.m.handle, self.v.handle, self.grad.handle,
                                            lr, self.step), "rsim_cost_adam")
            curve.append(float(dev.read_words(loss).view(np.float32).astype(np.float64).sum()) / (nb * batch))
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


ORDER_BUCKETS, ORDER_HI, ORDER_SCALE = 256, np.float32(3.0), np.float32(64.0)
ORDER_MODES = {"sorted": 0, "balanced": 1}


def cost_buckets(pred_cost):
    """The device's bucket of each predicted normalised log cost (0 = most expensive; NaN in the last)."""
    p = np.asarray(pred_cost, np.float32)
    q = (ORDER_HI - p) * ORDER_SCALE  # float32 arithmetic, as on the device
    b = np.where(np.isnan(q), ORDER_BUCKETS - 1, 0).astype(np.int64)
    ok = ~np.isnan(q)
    b[ok] = np.where(q[ok] <= 0, 0, np.where(q[ok] >= ORDER_BUCKETS - 1, ORDER_BUCKETS - 1, np.floor(np.maximum(q[ok], 0))))
    return b.astype(np.uint32)


def deal_slots(n, wave=64):
    """Slot of each rank in 'balanced' order: ranks dealt round-robin over the ceil(n / wave) waves; the last (partial)
    wave of r lanes takes part in the first r rounds only."""
    W = (n + wave - 1) // wave
    r = n - wave * (W - 1)
    s = np.arange(n, dtype=np.int64)
    slot = (s % W) * wave + s // W
    if W > 1:
        late = s >= r * W
        t = s[late] - r * W
        slot[late] = (t % (W - 1)) * wave + r + t // (W - 1)
    return slot


def order_by_cost(pred_cost, mode, wave=64):
    """The host statement of rsim_cost_order: perm[i] = the current slot that moves to slot i, by bucketed predicted cost
    (most expensive first, stable) in 'sorted' mode, or those ranks dealt over the waves in 'balanced' mode."""
    if mode not in ORDER_MODES:
        raise ValueError(mode)
    srt = np.argsort(cost_buckets(pred_cost), kind="stable").astype(np.uint32)
    if mode == "sorted":
        return srt
    perm = np.empty_like(srt)
    perm[deal_slots(srt.shape[0], wave)] = srt
    return perm


class Regrouper:
    """Steps decisions of n arenas in a cost-regrouped order, entirely on the device (no host synchronisation, so a
    decision can be captured in a HIP graph). The arenas live in a work buffer in the current order, slot i holding arena
    where[i]. Each regroup: features -> MLP -> rsim_cost_order -> gather the states into the new order (through a spare
    buffer, copied back) and compose `where`. Each step gathers the controls by `where`, runs sim8 and scatters the result words back to arena order."""

    def __init__(self, dev, model, world, states, mode="sorted", block=SIM8_BLOCK, variant=None, max_decisions=1):
        """world: a one-world buffer, or a world table with `variant` (DeviceWords, arena order) naming each arena's
        world; max_decisions: the most decisions one `run` call takes (its regrouped controls and results are held)."""
        if mode not in ORDER_MODES:
            raise ValueError(mode)
        self.dev, self.model, self.world, self.mode, self.block = dev, model, world, mode, block
        self.n = n = states.n
        self.cur = states  # the work buffer (the caller's states buffer is used and reordered)
        self.spare = dev.alloc_states(n)
        iota = np.arange(n, dtype=np.uint32)
        self.iota = dev.upload_words(iota)
        self.where = dev.upload_words(iota)
        self.where2 = dev.alloc_words(n)
        self.rel = dev.alloc_words(n)
        self.work = dev.alloc_words(int(lib().rsim_cost_order_work_words(n)))
        self.feat = dev.alloc_words(n * FEATURES)
        self.pred = dev.alloc_words(2 * n)
        self.ctl = dev.alloc_words(n * CONTROL_WORDS)
        self.res = dev.alloc_words(n)
        self.variant = variant  # arena order
        self.var_cur = dev.alloc_words(n) if variant is not None else None  # slot order
        if variant is not None:
            self._permute(variant, self.where, 1, False, self.var_cur)
        self.max_decisions = max_decisions
        if max_decisions > 1:
            self.ctl_run = dev.alloc_words(max_decisions * n * CONTROL_WORDS)
            self.res_run = dev.alloc_words(max_decisions * n)
            self.ctl_run_v = [self.ctl_run.view(k * n * CONTROL_WORDS, n * CONTROL_WORDS) for k in range(max_decisions)]
            self.res_run_v = [self.res_run.view(k * n, n) for k in range(max_decisions)]
        self._views = {}  # per-decision views of the caller's buffers, made once and kept (views are not made in a capture)

    def _permute(self, src, perm, words, scatter, dst):
        _check(lib().rsim_permute_words(self.dev.handle, src.handle, perm.handle, self.n, words, 1 if scatter else 0, dst.handle),
               "rsim_permute_words")
    def reset(self, states):
        """Back to arena order with the states of `states` (a buffer in arena order): copies them into the work buffer
        and resets `where` (asynchronous and capturable, so a recorded sequence can start with it and be replayed)."""
        self.dev.copy_states(self.cur, states)
        self._permute(self.iota, self.iota, 1, False, self.where)
        if self.variant is not None:
            self._permute(self.variant, self.iota, 1, False, self.var_cur)

    def regroup(self):
        n = self.n
        self.model.features(self.cur, self.feat)
        self.model.predict(self.feat, n, self.pred)
        _check(lib().rsim_cost_order(self.dev.handle, self.pred.handle, n, ORDER_MODES[self.mode], self.work.handle,
                                     self.rel.handle), "rsim_cost_order")
        # gather into the spare buffers, then copy back: every buffer keeps its role, so a graph that records any number
        # of regroups replays correctly any number of times (swapping the buffers on the host would bake one direction
        # into the graph)
        _check(lib().rsim_gather_states(self.dev.handle, self.cur.handle, self.rel.handle, self.spare.handle), "rsim_gather_states")
        self.dev.copy_states(self.cur, self.spare)
        _check(lib().rsim_permute_words(self.dev.handle, self.where.handle, self.rel.handle, n, 1, 0, self.where2.handle),
               "rsim_permute_words")
        _check(lib().rsim_permute_words(self.dev.handle, self.where2.handle, self.iota.handle, n, 1, 0, self.where.handle),
               "rsim_permute_words")
        if self.variant is not None:
            self._permute(self.variant, self.where, 1, False, self.var_cur)

    def step(self, controls, result):
        """One decision: controls (DeviceWords, arena order) and result (DeviceWords, arena order)."""
        n = self.n
        _check(lib().rsim_permute_words(self.dev.handle, controls.handle, self.where.handle, n, CONTROL_WORDS, 0, self.ctl.handle),
               "rsim_permute_words")
        self.dev.sim8_variants(self.cur, self.world, self.var_cur, self.ctl, self.res, self.block)
        _check(lib().rsim_permute_words(self.dev.handle, self.res.handle, self.where.handle, n, 1, 1, result.handle),
               "rsim_permute_words")

    def views(self, controls, results, decisions):
        """Per-decision views of a decisions x n x 16 controls buffer and a decisions x n results buffer (made once per
        buffer and kept: call this before a capture that records `run` on them)."""
        key = (id(controls), id(results), decisions)
        if key not in self._views:
            n = self.n
            self._views[key] = ([controls.view(k * n * CONTROL_WORDS, n * CONTROL_WORDS) for k in range(decisions)],
                                [results.view(k * n, n) for k in range(decisions)], controls, results)
        return self._views[key][:2]

    def run(self, controls, results, decisions, queued=True):
        """`decisions` (<= max_decisions) decisions in one launch in the current order: controls decisions x n x 16 words
        and results decisions x n words, both in arena order. The launch is rsim_run_queued, whose queue takes the blocks
        in index order (after a sorted regroup: heaviest first), or rsim_run."""
        if decisions > self.max_decisions:
            raise ValueError("run: %d decisions, the regrouper holds %d" % (decisions, self.max_decisions))
        cv, rv = self.views(controls, results, decisions)
        for k in range(decisions):
            self._permute(cv[k], self.where, CONTROL_WORDS, False, self.ctl_run_v[k])
        ctl, res = (self.ctl_run, self.res_run) if decisions == self.max_decisions else self._run_views(decisions)
        if queued:
            self.dev.run_queued(self.cur, self.world, ctl, res, decisions, variant=self.var_cur, block=self.block)
        else:
            self.dev.run(self.cur, self.world, ctl, res, decisions, variant=self.var_cur, block=self.block)
        for k in range(decisions):
            self._permute(self.res_run_v[k], self.where, 1, True, rv[k])

    def _run_views(self, decisions):
        key = ("run", decisions)
        if key not in self._views:
            n = self.n
            self._views[key] = (self.ctl_run.view(0, decisions * n * CONTROL_WORDS), self.res_run.view(0, decisions * n))
        return self._views[key]

    def slots_of(self, arenas, out):
        """out[i] = the slot that holds arena arenas[i] now, for the n entries of `arenas` (DeviceWords of n arena indices,
        e.g. a sample padded with zeros) into `out` (n words); asynchronous and capturable. Gathering the first m slots
        with rsim_gather_states then reads a sample of m arenas from the regrouped buffer. Call once before a capture
        (it allocates its inverse map on the first call)."""
        if not hasattr(self, "_inv"):
            self._inv = self.dev.alloc_words(self.n)
        self._permute(self.iota, self.where, 1, True, self._inv)  # inv[where[i]] = i
        self._permute(self._inv, arenas, 1, False, out)

    def states_in_arena_order(self, out=None):
        """A states buffer with arena a at position a (asynchronous)."""
        inv = self.dev.alloc_words(self.n)
        _check(lib().rsim_permute_words(self.dev.handle, self.iota.handle, self.where.handle, self.n, 1, 1, inv.handle),
               "rsim_permute_words")
        out = out if out is not None else self.dev.alloc_states(self.n)
        _check(lib().rsim_gather_states(self.dev.handle, self.cur.handle, inv.handle, out.handle), "rsim_gather_states")
        self.dev.sync()
        return out


class BlockOrder:
    """The heaviest-first order of the AoSoA blocks of a state buffer by the cost model (features -> MLP ->
    rsim_cost_block_order), for rsim_run_queued's queue: scheduling by predicted cost without moving any arena.
    Asynchronous and capturable; the buffers are allocated here, before any capture."""

    def __init__(self, dev, model, states):
        self.dev, self.model, self.states, self.n = dev, model, states, states.n
        self.feat = dev.alloc_words(self.n * FEATURES)
        self.pred = dev.alloc_words(2 * self.n)
        self.work = dev.alloc_words(int(lib().rsim_cost_block_order_work_words(self.n)))
        self.order = dev.alloc_words((self.n + 63) // 64)

    def update(self):
        self.model.features(self.states, self.feat)
        self.model.predict(self.feat, self.n, self.pred)
        _check(lib().rsim_cost_block_order(self.dev.handle, self.pred.handle, self.n, self.work.handle, self.order.handle),
               "rsim_cost_block_order")
        return self.order
