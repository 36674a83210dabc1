"""Skeleton parity against the oracle with its near-tie margins.

A block whose decisions came within rounding of flipping (``first_near_tie``:
a uniform next to a CDF step, two nearly equal pivot candidates, a filter sum
at its threshold, an error at tol^2) may legitimately go the other way on the
device, which forms those quantities in another order or through another
factorization.  The skeleton is therefore required to be identical up to the
first such block (the whole of it when there is none, which is the case for
every matrix in these tests so far); the margins are printed either way.
"""

import numpy as np

from oracle import first_near_tie, margin_summary

# lines for the end-of-run summary (tests/conftest.py prints them)
REPORT = []


def report(line: str) -> None:
    print(line)
    REPORT.append(line)


def assert_skeleton(perm, k, ranks, draws, r, name="", quiet=False):
    """perm (device result, numpy), k, ranks (None: not reported, as for the
    batched form), draws against the oracle result r.  Returns True when the
    whole skeleton was required to match.  quiet: the margins line only when
    there is a near tie (batches of many problems)."""
    tie = first_near_tie(r)
    if not quiet or tie is not None:
        report(f"[{name}] margins: {margin_summary(r)}")
    if tie is None:
        assert k == r.k, name
        assert np.array_equal(perm[:k], r.perm[: r.k]), name
        if ranks is not None:
            assert np.array_equal(np.asarray(ranks), r.ranks), name
        assert draws == r.draws, name
        return True
    t = tie[0]
    # blocks before the near tie are identical
    if ranks is not None:
        assert np.array_equal(np.asarray(ranks)[:t], r.ranks[:t]), (name, tie)
    kt = int(r.ranks[t - 1]) if t > 0 else 0
    assert np.array_equal(perm[:kt], r.perm[:kt]), (name, tie)
    same = k == r.k and np.array_equal(perm[:k], r.perm[: r.k])
    report(f"[{name}] near tie at block {t}: skeleton identical through column {kt}, "
           f"whole skeleton {'identical' if same else 'differs after it'}")
    return same
