"""Write the golden fixtures of this directory from published vectors and library routines.

No reference code runs here: the reference holds no tests, golden vectors or
fixtures for its RBRP path.  The fixtures pin the parts of the path whose
arithmetic is fixed by a published generator or by a library routine:

* ``philox4x64_10_kat.json``: the Random123 ``philox4x64_10`` known-answer
  vectors (Salmon, Moraes, Dror, Shaw, SC'11; ``kat_vectors`` of the Random123
  distribution), typed in from the publication, not computed here.
* ``uniforms.npz``: NumPy's ``Generator(Philox(key=seed, counter=offset)).random``
  stream, which is the uniform stream of the path (it replaces
  gpu/random.cu:26-54 and the ``default_rng`` of RBRP.ipynb c11:89).
* ``sampler_choice.npz``: ``Generator.choice(N, s, replace=False, p=p)``, the
  sampler of RBRP.ipynb c11:108, with the number of uniforms each call consumed.
* ``dgeqp3_pivots.npz``: LAPACK ``dgeqp3`` (through ``scipy.linalg.qr(pivoting=True)``)
  pivot order and ``R`` of small Gram-friendly panels, the pivot rule of
  ``magma_dgeqp3_gpu`` at gpu/rbrp.cu:404-417.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""

from __future__ import annotations

import json
import os

import numpy as np
import scipy
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1

# Random123 kat_vectors, philox4x64_10: (counter words, key words, expected output words)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0),
     (0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b)),
    ((M64, M64, M64, M64), (M64, M64),
     (0x87b092c3013fe90b, 0x438c3c67be8d0224, 0x9cc7d7c69cd777b6, 0xa09caebf594f0ba0)),
    ((0x243f6a8885a308d3, 0x13198a2e03707344, 0xa4093822299f31d0, 0x082efa98ec4e6c89),
     (0x452821e638d01377, 0xbe5466cf34e90c6c),
     (0xa528f45403e61d95, 0x38c72dbd566e9788, 0xa5a1610e72fd18b5, 0x57bd43b5e52b7fe6)),
]

# (seed, offset, first draw, count) of the uniform-stream fixtures
UNIFORM_CASES = [(0, 0, 0, 256), (42, 0, 3, 129), (1986, 7, 5, 64),
                 (2 ** 63 + 5, 2 ** 64 - 2, 1, 64), (M64, 2 ** 200, 0, 33)]

# (N, s, weight law, seed, offset, draws already consumed) of the sampler fixtures
SAMPLER_CASES = [(10, 4, "uniform", 0, 0, 0), (1, 1, "uniform", 1, 0, 0),
                 (64, 64, "uniform", 2, 3, 5), (300, 40, "skew6", 3, 0, 0),
                 (4000, 64, "dominant", 4, 9, 7), (4000, 64, "zeros", 5, 1, 2),
                 (5000, 128, "skew12", 6, 0, 11), (2000, 16, "ties", 7, 2 ** 64 - 1, 0),
                 (129, 120, "dominant", 8, 0, 3), (700, 64, "expdecay", 9, 0, 0)]

# (rows, columns, dependent pairs, seed) of the pivoted-QR panels
PANEL_CASES = [(160, 8, 0, 0), (160, 16, 1, 1), (160, 40, 0, 2), (160, 64, 2, 3),
               (160, 64, 0, 4), (200, 33, 3, 5)]


def _weights(N: int, law: str, rng: np.random.Generator) -> np.ndarray:
    if law == "uniform":
        w = rng.random(N) + 0.5
    elif law.startswith("skew"):
        w = rng.random(N) ** int(law[4:])
    elif law == "dominant":
        w = rng.random(N)
        w[rng.integers(0, N)] = w.sum() * 50.0     # forces redraw rounds
    elif law == "zeros":
        w = rng.random(N)
        w[rng.random(N) < 0.5] = 0.0
    elif law == "ties":
        w = np.repeat(rng.random(N // 8) + 0.1, 8)
    elif law == "expdecay":
        w = np.exp(-np.arange(N) / 50.0)
    else:
        raise ValueError(law)
    return w


def _draws_consumed(seed: int, offset: int, state_after: dict, bound: int) -> int:
    """Smallest t such that a fresh stream after t uniforms has the state ``state_after``."""
    for t in range(bound + 1):
        g = np.random.Generator(np.random.Philox(key=seed, counter=offset))
        g.random(t)
        st = g.bit_generator.state
        if (np.array_equal(st["state"]["counter"], state_after["state"]["counter"])
                and st["buffer_pos"] == state_after["buffer_pos"]):
            return t
    raise RuntimeError("draw count not found")


def make_philox() -> None:
    vecs = [{"counter": [f"{w:#018x}" for w in c], "key": [f"{w:#018x}" for w in k],
             "expected": [f"{w:#018x}" for w in o]} for c, k, o in PHILOX_KAT]
    with open(os.path.join(HERE, "philox4x64_10_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors, philox4x64_10 (Salmon et al., SC'11)",
                   "vectors": vecs}, f, indent=1)
        f.write("\n")


def make_uniforms() -> None:
    out = {}
    for q, (seed, off, d0, cnt) in enumerate(UNIFORM_CASES):
        g = np.random.Generator(np.random.Philox(key=seed, counter=off))
        u = g.random(d0 + cnt)[d0:]
        # seed and counter as decimal strings: the counter is a 256-bit integer
        out[f"case{q}_seed"] = np.array([str(seed)])
        out[f"case{q}_offset"] = np.array([str(off)])
        out[f"case{q}_draw0"] = np.array([d0], dtype=np.int64)
        out[f"case{q}_u"] = u
    out["n_cases"] = np.array([len(UNIFORM_CASES)])
    np.savez_compressed(os.path.join(HERE, "uniforms.npz"), **out)


def make_sampler() -> None:
    out = {}
    for q, (N, s, law, seed, off, pre) in enumerate(SAMPLER_CASES):
        w = _weights(N, law, np.random.default_rng(1000 + q))
        p = w / w.sum()
        g = np.random.Generator(np.random.Philox(key=seed, counter=off))
        g.random(pre)
        J = g.choice(N, size=s, replace=False, p=p)
        draws = _draws_consumed(seed, off, g.bit_generator.state, pre + 64 * s + 64)
        out[f"case{q}_p"] = p
        out[f"case{q}_meta"] = np.array([N, s, pre], dtype=np.int64)
        out[f"case{q}_seed"] = np.array([str(seed)])
        out[f"case{q}_offset"] = np.array([str(off)])
        out[f"case{q}_J"] = J.astype(np.int64)
        out[f"case{q}_draws"] = np.array([draws - pre], dtype=np.int64)
    out["n_cases"] = np.array([len(SAMPLER_CASES)])
    np.savez_compressed(os.path.join(HERE, "sampler_choice.npz"), **out)


def make_panels() -> None:
    out = {}
    for q, (m, s, deps, seed) in enumerate(PANEL_CASES):
        rng = np.random.default_rng(seed)
        V = rng.standard_normal((m, s)) * np.exp(-rng.random(s) * 8)
        for _ in range(deps):
            a, c = rng.integers(0, s, size=2)
            V[:, a] = V[:, c] * 2.0 + 1e-9 * rng.standard_normal(m)
        _, R, piv = sla.qr(V, mode="economic", pivoting=True)
        out[f"case{q}_V"] = V
        out[f"case{q}_piv"] = piv.astype(np.int64)
        out[f"case{q}_R"] = R      # row signs are LAPACK's convention; |R| and its row sums pin
    out["n_cases"] = np.array([len(PANEL_CASES)])
    out["scipy_version"] = np.array([scipy.__version__])
    np.savez_compressed(os.path.join(HERE, "dgeqp3_pivots.npz"), **out)


if __name__ == "__main__":
    make_philox()
    make_uniforms()
    make_sampler()
    make_panels()
    print("fixtures written to", HERE)
