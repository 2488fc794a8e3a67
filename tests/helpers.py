"""Procedures shared by the tests and the golden-vector generator: how a decision sequence is replayed."""
import hashlib

import numpy as np

TICK_WORDS = 2567
ARENA_WORDS = 2784


def digest(words):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(words).tobytes()).digest()[:8], np.uint64)[0]


def digests(states, words=TICK_WORDS):
    return np.array([digest(s[:words]) for s in states], np.uint64)
