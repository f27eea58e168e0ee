"""Writes the golden fixtures of this directory from the CPU FP64 oracle.

The reference repository holds no golden vectors, known-answer tests or fixtures
(it has no tests); its only pinned data are the seeds (reproduce.py:21-22,
experiment.py:43-51), the thresholds (reproduce.py:28) and the case list
(reproduce.py:14-15). These fixtures therefore freeze:
  - SHA-256 checksums of the seeded targets and upstream gradients (torch RNG drift),
  - the oracle's solve of the n=256, k=8 target (X, omega, Newton / CG / backtrack counts),
  - the oracle's solve of two n=2048, k=16 targets (omega, counts, the first 64 rows of X),
  - the oracle's implicit-differentiation gradient for two small targets.
Regenerate with `python tests/golden/make_golden.py` from the repository root.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import adjoint as oadj  # noqa: E402
from oracle import dual as odual  # noqa: E402
from symplectic_stiefel_nearness_amd import inputs  # noqa: E402

CHECKSUM_COUNT = 2
UNEVEN_CHECKSUM_COUNT = 4
ADJOINT_CASE = dict(n=20, k=3, seeds=[9001, 9002], grad_seed=9003)


def checksums():
    out = {"base_seed": inputs.BASE_SEED, "cases": {}, "uneven": {}, "gradients": {}}
    for name, (idx, n, k) in inputs.CASES.items():
        T = inputs.make_case(name, CHECKSUM_COUNT)
        out["cases"][name] = {"case_index": idx, "n": n, "k": k,
                              "sha256": [inputs.checksum(T[j]) for j in range(CHECKSUM_COUNT)]}
    shapes = inputs.uneven_shapes(128)
    targets = inputs.make_uneven(UNEVEN_CHECKSUM_COUNT)
    out["uneven"] = {"shapes": shapes, "sha256": [inputs.checksum(t) for t in targets]}
    G = inputs.make_gradients(4096, 32, CHECKSUM_COUNT)
    out["gradients"]["n4096_k32"] = [inputs.checksum(G[j]) for j in range(CHECKSUM_COUNT)]
    return out


def main():
    (HERE / "inputs_sha256.json").write_text(json.dumps(checksums(), indent=1) + "\n")

    T = inputs.make_case("n256_k8", 1).numpy()
    X, omega, iters, st = odual.solve(T, return_stats=True)
    np.savez(HERE / "n256_k8_solve.npz", T=T, X=X, omega=omega, iterations=iters,
             cg=np.asarray(st.cg), backtracks=np.asarray(st.backtracks))

    T = inputs.make_case("n2048_k16", 2).numpy()
    X, omega, iters, st = odual.solve(T, return_stats=True)
    np.savez(HERE / "n2048_k16_solve.npz", X_rows=X[:, :64, :], omega=omega, iterations=iters,
             cg=np.asarray(st.cg), backtracks=np.asarray(st.backtracks))

    c = ADJOINT_CASE
    T = inputs.make_targets(c["n"], c["k"], 2, seeds=c["seeds"]).numpy()
    G = np.random.default_rng(c["grad_seed"]).standard_normal(T.shape)
    dT, _, _ = oadj.projection_grad_batch(T, G)
    np.savez(HERE / "adjoint_n20_k3.npz", T=T, G=G, dT=dT)
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))


if __name__ == "__main__":
    main()
