"""Shared helpers for the parity tests (seeds, comparisons)."""
import numpy as np

SEED_BASE = 0x4158494F4D000000


def seed(config, world=0):
    """Scene seed of a configuration (world > 0: the seed of batched world world-1).
    Seeds step by 2 because DeterministicRNG starts from seed | 1."""
    return SEED_BASE + 1000 * config + 2 * world


def pair_set(a, b):
    return set(zip(np.asarray(a).tolist(), np.asarray(b).tolist()))


def rel_err(x, y, floor=1.0):
    """max |x - y| / max(|y|, floor) elementwise."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.size == 0:
        return 0.0
    return float(np.max(np.abs(x - y) / np.maximum(np.abs(y), floor)))


def match_manifolds(g, o):
    """Index map from oracle manifolds to GPU manifolds by (a, b)."""
    gi = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(g["a"], g["b"]))}
    return [gi.get((int(a), int(b)), -1) for a, b in zip(o["a"], o["b"])]
