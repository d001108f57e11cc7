"""The MPNN oracle (oracle/mpnn_ref.py) pinned by finite differences, plus
host-side checks of the network's C entry points that need no GPU."""
import ctypes as C

import numpy as np
import pytest

from oracle import mpnn_ref
from spherocyls_amd import native


def random_graph(n, rng, degree=3):
    """Symmetric random graph in CSR over receivers, senders ascending."""
    pairs = set()
    for i in range(n):
        for j in rng.choice(n, degree, replace=False):
            if i != j:
                pairs.add((i, int(j)))
                pairs.add((int(j), i))
    edges = sorted(pairs)  # (dst, src) sorted by dst then src
    dst = np.array([d for d, _ in edges])
    src = np.array([s for _, s in edges], dtype=np.int32)
    off = np.zeros(n + 1, dtype=np.int32)
    np.add.at(off, dst + 1, 1)
    off = np.cumsum(off).astype(np.int32)
    nf = np.column_stack([np.cos(rng.uniform(0, 6, n)), np.sin(rng.uniform(0, 6, n)),
                          rng.uniform(0.5, 0.7, n), rng.uniform(0.2, 1.5, n)])
    ef = rng.normal(0, 0.5, (len(edges), 5))
    return nf, off, src, ef


def random_params(rng, scale=0.08):
    p = rng.normal(0, scale, mpnn_ref.num_params())
    lay, _ = mpnn_ref.layout()
    o, K, N = lay["wo"]
    w = p[o:o + K * N].reshape(K, N)
    w[:, 3:] = 0.0
    return p


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nf, off, src, ef = random_graph(12, rng)
    params = random_params(rng)
    target = rng.normal(0, 1, (12, 3))
    loss, grad = mpnn_ref.loss_and_grad(params, nf, off, src, ef, target)
    lay, _ = mpnn_ref.layout()
    h = 1e-6
    checked = 0
    for name, (o, K, N) in lay.items():
        # rows that carry data: real inputs and the bias row
        for _ in range(6):
            idx = o + int(rng.integers(0, K * N))
            if abs(grad[idx]) < 1e-12:
                continue
            pp, pm = params.copy(), params.copy()
            pp[idx] += h
            pm[idx] -= h
            lp, _ = mpnn_ref.loss_and_grad(pp, nf, off, src, ef, target)
            lm, _ = mpnn_ref.loss_and_grad(pm, nf, off, src, ef, target)
            fd = (lp - lm) / (2 * h)
            assert fd == pytest.approx(grad[idx], rel=1e-5, abs=1e-10), (name, idx)
            checked += 1
    assert checked >= 20


def test_oracle_edge_order_and_bias_rows():
    """Permuting nodes permutes outputs (the network sees the graph, not the
    order), and the bias rows act as biases."""
    rng = np.random.default_rng(1)
    nf, off, src, ef = random_graph(10, rng)
    params = random_params(rng)
    y, _ = mpnn_ref.forward(params, nf, off, src, ef)
    perm = rng.permutation(10)
    inv = np.argsort(perm)
    # relabel node i -> inv[i]; rebuild CSR
    dst = np.repeat(np.arange(10), np.diff(off))
    edges = sorted((int(inv[d]), int(inv[s]), k) for k, (d, s) in enumerate(zip(dst, src)))
    src2 = np.array([s for _, s, _ in edges], dtype=np.int32)
    ef2 = ef[[k for _, _, k in edges]]
    off2 = np.zeros(11, dtype=np.int32)
    np.add.at(off2, np.array([d for d, _, _ in edges]) + 1, 1)
    off2 = np.cumsum(off2).astype(np.int32)
    nf2 = nf[perm]
    y2, _ = mpnn_ref.forward(params, nf2, off2, src2, ef2)
    np.testing.assert_allclose(y2, y[perm], rtol=1e-12, atol=1e-14)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -3.14159, 0.0])
    r = mpnn_ref.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -6        # tie rounds up to even
    assert abs(r[3] - (-3.14159)) < 2 ** -7 * 4
    assert r[4] == 0.0


def test_mpnn_entry_points_exported_and_need_gpu():
    lib = native.load()
    for name in ("sc_mpnn_create", "sc_mpnn_destroy", "sc_mpnn_forward", "sc_mpnn_backward",
                 "sc_mpnn_adam", "sc_mpnn_load_graph", "sc_mpnn_graph_from_system",
                 "sc_mpnn_export_graph", "sc_mpnn_warm_start", "sc_mpnn_num_params"):
        assert hasattr(lib, name)
    if lib.sc_device_count() == 0:
        h = C.c_void_p()
        assert lib.sc_mpnn_create(C.byref(h), 0, 1) == native.SC_ERR_NO_DEVICE
