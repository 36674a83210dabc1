"""The committed fixtures of ``tests/golden`` against the oracle (CPU) and the HIP library (GPU).

The fixtures hold published Philox known-answer vectors and the outputs of
NumPy's ``Generator.choice`` and LAPACK ``dgeqp3`` on fixed inputs, written by
``tests/golden/make_golden.py``; they pin the uniform stream, the sampler and
the panel's pivot rule without running any library at test time.
"""

import ctypes
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import robust_filter, sample_without_replacement, uniform_stream

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _philox_kat():
    with open(os.path.join(GOLDEN, "philox4x64_10_kat.json")) as f:
        vecs = json.load(f)["vectors"]
    return [tuple(tuple(int(w, 16) for w in v[key]) for key in ("counter", "key", "expected"))
            for v in vecs]


def _cases(d):
    return range(int(d["n_cases"][0]))


# ---------------------------------------------------------------- CPU: the oracle

@pytest.mark.parametrize("ctr,key,want", _philox_kat())
def test_host_philox_matches_kat_fixture(librbrp, ctr, key, want):
    c = (ctypes.c_uint64 * 4)(*ctr)
    k = (ctypes.c_uint64 * 2)(*key)
    out = (ctypes.c_uint64 * 4)()
    librbrp.rbrp_philox4x64_10_host(c, k, out)
    assert tuple(out) == want


def test_oracle_uniform_stream_matches_fixture():
    d = _load("uniforms.npz")
    for q in _cases(d):
        seed, off = int(d[f"case{q}_seed"][0]), int(d[f"case{q}_offset"][0])
        d0, u = int(d[f"case{q}_draw0"][0]), d[f"case{q}_u"]
        got = uniform_stream(seed, off).random(d0 + len(u))[d0:]
        assert np.array_equal(got, u), q


def test_oracle_sampler_matches_choice_fixture():
    d = _load("sampler_choice.npz")
    for q in _cases(d):
        N, s, pre = (int(x) for x in d[f"case{q}_meta"])
        g = uniform_stream(int(d[f"case{q}_seed"][0]), int(d[f"case{q}_offset"][0]))
        g.random(pre)
        J, draws = sample_without_replacement(d[f"case{q}_p"], s, g)
        assert np.array_equal(J, d[f"case{q}_J"]), q
        assert draws == int(d[f"case{q}_draws"][0]), q


def test_oracle_panel_routine_matches_dgeqp3_fixture():
    """The oracle's panel is LAPACK dgeqp3 through SciPy: same pivots and |R| here."""
    d = _load("dgeqp3_pivots.npz")
    for q in _cases(d):
        V, piv, R = d[f"case{q}_V"], d[f"case{q}_piv"], d[f"case{q}_R"]
        _, R2, piv2 = sla.qr(V, mode="economic", pivoting=True)
        assert np.array_equal(piv2, piv), q
        assert np.abs(np.abs(R2) - np.abs(R)).max() <= 1e-12 * np.abs(R).max(), q


# ---------------------------------------------------------------- GPU: the HIP library

def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.gpu
def test_device_uniforms_match_fixture(librbrp):
    from robust_blockwise_random_pivoting_amd import uniforms
    d = _load("uniforms.npz")
    for q in _cases(d):
        seed, off = int(d[f"case{q}_seed"][0]), int(d[f"case{q}_offset"][0])
        if off >= 1 << 64:
            continue   # the C ABI takes a 64-bit stream offset
        d0, u = int(d[f"case{q}_draw0"][0]), d[f"case{q}_u"]
        got = uniforms(seed, off, d0, len(u)).cpu().numpy()
        assert np.array_equal(got, u), q


@pytest.mark.gpu
def test_device_sampler_matches_choice_fixture(librbrp, hooks):
    import torch
    d = _load("sampler_choice.npz")
    for q in _cases(d):
        N, s, pre = (int(x) for x in d[f"case{q}_meta"])
        seed, off = int(d[f"case{q}_seed"][0]), int(d[f"case{q}_offset"][0])
        dp = torch.as_tensor(d[f"case{q}_p"]).cuda()
        dJ = torch.zeros(s, dtype=torch.int32, device="cuda")
        s_out, draws = ctypes.c_int32(), ctypes.c_uint64()
        print(f"sampler fixture case {q}: N={N} s={s}", flush=True)
        code = hooks.rbrp_test_sample(_stream(), _ptr(dp), 0, N, N, s, seed, off, pre, _ptr(dJ),
                                        ctypes.byref(s_out), ctypes.byref(draws))
        assert code == 0
        assert s_out.value == s, q
        assert np.array_equal(dJ.cpu().numpy().astype(np.int64), d[f"case{q}_J"]), q
        assert draws.value == pre + int(d[f"case{q}_draws"][0]), q


@pytest.mark.gpu
def test_device_panel_matches_dgeqp3_fixture(librbrp, hooks):
    """Gram + pivoted Cholesky on the device: dgeqp3's pivots through the filter's kept count."""
    import torch
    d = _load("dgeqp3_pivots.npz")
    for q in _cases(d):
        V, piv, R = d[f"case{q}_V"], d[f"case{q}_piv"], d[f"case{q}_R"]
        s = V.shape[1]
        b = s
        i = 500
        J = (i + np.arange(s)[::-1]).astype(np.int32)      # any distinct labels
        G = V.T @ V
        dG = torch.as_tensor(np.ascontiguousarray(G.T)).cuda()
        dJ = torch.as_tensor(J).cuda()
        dJk = torch.zeros(b, dtype=torch.int32, device="cuda")
        dJin = torch.zeros(b, dtype=torch.int32, device="cuda")
        dJout = torch.zeros(b, dtype=torch.int32, device="cuda")
        dRb = torch.zeros(b * b, dtype=torch.float64, device="cuda")
        bt, ns, status = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        print(f"panel fixture case {q}: s={s}", flush=True)
        code = hooks.rbrp_test_panel(_stream(), _ptr(dG), s, b, 0, i, _ptr(dJ), _ptr(dJk),
                                       _ptr(dRb), _ptr(dJin), _ptr(dJout), ctypes.byref(bt),
                                       ctypes.byref(ns), ctypes.byref(status))
        assert code == 0 and status.value == 0, q
        bt_ref = robust_filter(R, s, b, "sampled")
        assert bt.value == bt_ref, q
        assert np.array_equal(dJk.cpu().numpy()[:bt_ref], J[piv[:bt_ref]]), q
        Rb = dRb.cpu().numpy().reshape(b, b).T[:bt_ref, :bt_ref]
        Rref = np.abs(R[:bt_ref, :bt_ref])
        assert np.abs(np.abs(Rb) - Rref).max() <= 1e-9 * Rref.max(), q
