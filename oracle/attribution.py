"""TEST INFRASTRUCTURE ONLY: attributes the words in which a device result differs from a target (V3, or a stand-in for
it) to the libm call that causes them, with the oracle's libm attribution build (liboracle_libm.so). It restates the
reference's rule (crates/rocketsim-cuda/tests/common/mod.rs:458-503): from the same input state, a differing word is
charged to function f when making f alone the platform library (the others correctly rounded) changes that word; to
`combo` when no single function does but all seven together do; otherwise to `other` (the difference is not the libm's).
"""
import numpy as np

from oracle import LIBM_FUNCS, PLATFORM_ALL


class Attribution:
    def __init__(self):
        self.ran = 0
        self.states_diff = 0
        self.words = 0
        self.by_func = [0] * len(LIBM_FUNCS)
        self.combo = 0
        self.other = 0
        self.examples = []

    def add(self, label, diff, dev, target, cr, plat, single):
        """One state whose device result `dev` differs from `target` in the words `diff`; `cr`, `plat` and `single[f]`
        are the correctly rounded, platform and single-function builds' results from the same input (None for a run
        that did not reach this state, which counts as a change)."""
        self.states_diff += 1
        for w in diff:
            self.words += 1
            hits = [f for f in range(len(LIBM_FUNCS)) if single[f] is None or single[f][w] != cr[w]]
            if hits:
                for f in hits:
                    self.by_func[f] += 1
            elif plat[w] != cr[w]:
                self.combo += 1
            else:
                self.other += 1
                if len(self.examples) < 8:
                    self.examples.append("%s: word %d device %08x target %08x platform %08x" %
                                         (label, w, dev[w], target[w], plat[w]))

    def row(self):
        return dict(ran=self.ran, states_diff=self.states_diff, words=self.words,
                    **{f: n for f, n in zip(LIBM_FUNCS, self.by_func)}, combo=self.combo, other=self.other)


def attribute_step(mixed, acc, label, before, world, dev_after, target_after, step, exclude=()):
    """Runs `step(mixed, state, world, mask)` (one phase or one tick) from `before` under the correctly rounded, platform
    and seven single-function masks, and charges the words where `dev_after` differs from `target_after` (outside the
    word indices in `exclude`) to acc."""
    acc.ran += 1
    diff = [int(w) for w in np.flatnonzero(np.asarray(dev_after) != np.asarray(target_after)) if int(w) not in exclude]
    if not diff:
        return
    runs = {}
    for mask in [0, PLATFORM_ALL] + [1 << f for f in range(len(LIBM_FUNCS))]:
        s = np.array(before, np.uint32, copy=True)
        step(mixed, s, world, mask)
        runs[mask] = s
    acc.add(label, diff, dev_after, target_after, runs[0], runs[PLATFORM_ALL],
            [runs[1 << f] for f in range(len(LIBM_FUNCS))])
