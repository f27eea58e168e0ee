"""Seeded synthetic targets, generated exactly as the reference experiment does.

`make_targets` restates `src/symplectic_projection/experiment.py:43-51`:
T_j = randn(2n, 2k, float64, Generator('cpu').manual_seed(s_j)) / sqrt(2n),
with s_j = BASE_SEED + 1000 * case_index + j (`reproduce.py:21`, `experiment.py:44-45`).
"""
from __future__ import annotations

import hashlib
import math
import random

import torch

BASE_SEED = 2026102100                # reproduce.py:21
GRADIENT_SEED_OFFSET = 10 ** 6        # upstream gradients G use BASE_SEED + 1e6 + j

# Case indices 0-5 are the reference's cases n in (2048, 4096) x k in (16, 32, 48)
# (reproduce.py:14-15); 6-8 extend the numbering for the extra workloads.
CASES = {
    "n2048_k16": (0, 2048, 16), "n2048_k32": (1, 2048, 32), "n2048_k48": (2, 2048, 48),
    "n4096_k16": (3, 4096, 16), "n4096_k32": (4, 4096, 32), "n4096_k48": (5, 4096, 48),
    "n256_k8": (6, 256, 8), "n8192_k64": (7, 8192, 64),
}
UNEVEN_CASE_INDEX = 8


def seeds_for(case_index, count):
    return [BASE_SEED + 1000 * case_index + j for j in range(count)]


def make_targets(n, k, count, case_index=None, seeds=None):
    """CPU float64 [count, 2n, 2k] targets with entries N(0, 1/(2n))."""
    if seeds is None:
        seeds = seeds_for(case_index, count)
    targets = torch.empty(len(seeds), 2 * n, 2 * k, dtype=torch.float64)
    for item, seed in enumerate(seeds):
        generator = torch.Generator(device="cpu").manual_seed(seed)
        targets[item] = torch.randn(2 * n, 2 * k, dtype=torch.float64,
                                    generator=generator) / math.sqrt(2 * n)
    return targets


def make_case(name, count=64):
    case_index, n, k = CASES[name]
    return make_targets(n, k, count, case_index=case_index)


def make_gradients(n, k, count):
    """Upstream gradients G ~ N(0, 1) of shape [count, 2n, 2k]."""
    out = torch.empty(count, 2 * n, 2 * k, dtype=torch.float64)
    for j in range(count):
        generator = torch.Generator(device="cpu").manual_seed(BASE_SEED + GRADIENT_SEED_OFFSET + j)
        out[j] = torch.randn(2 * n, 2 * k, dtype=torch.float64, generator=generator)
    return out


def uneven_shapes(count=128):
    """(n, k) pairs of the uneven workload, drawn with random.Random(BASE_SEED + 8)."""
    rng = random.Random(BASE_SEED + UNEVEN_CASE_INDEX)
    return [(rng.choice((1024, 2048, 4096, 8192)), rng.choice((8, 16, 32, 48, 64)))
            for _ in range(count)]


def make_uneven(count=128):
    """List of CPU float64 targets with mixed (n, k), seeded per target."""
    shapes = uneven_shapes(count)
    seeds = seeds_for(UNEVEN_CASE_INDEX, count)
    return [make_targets(n, k, 1, seeds=[s])[0] for (n, k), s in zip(shapes, seeds)]


def checksum(tensor):
    """SHA-256 of the raw float64 bytes (guards against RNG drift across torch versions)."""
    return hashlib.sha256(tensor.contiguous().numpy().tobytes()).hexdigest()
