"""Message-passing warm-start network on the GPU: contact graph, forward pass,
gradients and training against the CPU oracle (oracle/mpnn_ref.py, pinned by
finite differences in tests/test_mpnn_cpu.py)."""
import numpy as np
import pytest

import oracle
from oracle import mpnn_ref
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def packing():
    R, A = configs.bidisperse(300)
    return configs.make_packing(R, A, 0.84, 71, "mpnn graph")


@pytest.fixture(scope="module")
def graph(packing):
    s = native.System(packing, pads=0.1)
    m = native.Mpnn(seed=3)
    try:
        g = m.graph_of(s)
    finally:
        s.close()
        m.close()
    return g


def random_params(seed, scale=0.08):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, mpnn_ref.num_params())
    lay, _ = mpnn_ref.layout()
    o, K, N = lay["wo"]
    p[o:o + K * N].reshape(K, N)[:, 3:] = 0.0
    return p.astype(np.float32)


def test_graph_is_the_contact_graph(packing, graph):
    g = graph
    ob = oracle.OracleBox(packing.L, packing.x, packing.y, packing.phi, packing.A, packing.R, 0.1)
    cp, cq = ob.brute_contacts()
    want = set(zip(np.minimum(cp, cq).tolist(), np.maximum(cp, cq).tolist()))
    dst = np.repeat(np.arange(g.n), np.diff(g.offsets))
    got = set(zip(np.minimum(dst, g.src).tolist(), np.maximum(dst, g.src).tolist()))
    assert got == want and g.e == 2 * len(want)
    # senders ascend within each receiver; every edge has its reverse
    for i in range(g.n):
        s = g.src[g.offsets[i]:g.offsets[i + 1]]
        assert np.all(np.diff(s) > 0)
    ef = g.edge_feat
    assert np.all((ef[:, 0] > 0) & (ef[:, 0] < 1))
    np.testing.assert_allclose(np.hypot(ef[:, 1], ef[:, 2]), 1.0, rtol=1e-6)
    # branch vector: sender minus receiver, minimum image
    L = packing.L
    bx = packing.x[g.src] - packing.x[dst]
    by = packing.y[g.src] - packing.y[dst]
    bx -= L * np.round(bx / L)
    by -= L * np.round(by / L)
    np.testing.assert_allclose(ef[:, 3], bx, atol=1e-5)
    np.testing.assert_allclose(ef[:, 4], by, atol=1e-5)
    # node features
    np.testing.assert_allclose(g.node_feat[:, 0], np.cos(2 * packing.phi), atol=1e-6)
    np.testing.assert_allclose(g.node_feat[:, 2], packing.R, rtol=1e-7)


def test_forward_and_gradients_match_oracle(graph):
    g = graph
    p = random_params(5)
    rng = np.random.default_rng(6)
    target = rng.normal(0, 1, (g.n, 3)).astype(np.float32)
    m = native.Mpnn(seed=1)
    try:
        m.set_params(p)
        m.load_graph(g)
        y = m.forward()
        loss = m.backward(target)
        G = m.grads()
        y2 = m.forward()
        loss2 = m.backward(target)
        G2 = m.grads()
    finally:
        m.close()
    # deterministic: fixed-order sums everywhere
    assert np.array_equal(y, y2) and loss == loss2 and np.array_equal(G, G2)
    args = (p.astype(np.float64), g.node_feat, g.offsets, g.src, g.edge_feat)
    yr, _ = mpnn_ref.forward(*args, bf16=True)
    lr, Gr = mpnn_ref.loss_and_grad(*args, target, bf16=True)
    err_y = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    print("forward rel err", err_y, "loss", loss, lr, flush=True)
    assert err_y < 2e-3
    assert loss == pytest.approx(lr, rel=1e-3)
    lay, _ = mpnn_ref.layout()
    for name, (o, K, N) in lay.items():
        a, b = G[o:o + K * N], Gr[o:o + K * N]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        print(f"grad {name}: rel {rel:.2e} cos {cos:.6f}", flush=True)
        assert rel < 2e-2 and cos > 0.9995, name
    # against the unquantised FP64 network the BF16 operands cost ~1e-2
    _, G64 = mpnn_ref.loss_and_grad(*args, target, bf16=False)
    assert np.linalg.norm(G - G64) / np.linalg.norm(G64) < 5e-2


def test_training_reduces_loss(graph):
    """Adam on a smooth target (overlap-weighted sum of contact normals, the
    direction a relaxation step pushes a particle) lowers the loss 10x."""
    g = graph
    dst = np.repeat(np.arange(g.n), np.diff(g.offsets))
    t = np.zeros((g.n, 3), dtype=np.float32)
    np.add.at(t[:, 0], dst, g.edge_feat[:, 0] * g.edge_feat[:, 1] * 10)
    np.add.at(t[:, 1], dst, g.edge_feat[:, 0] * g.edge_feat[:, 2] * 10)
    m = native.Mpnn(seed=11)
    try:
        m.load_graph(g)
        losses = []
        for step in range(1, 301):
            m.forward()
            losses.append(m.backward(t))
            m.adam(3e-3, step)
    finally:
        m.close()
    print("training loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.1 * losses[0]


def test_batched_graph_is_the_concatenation(packing):
    """A batched system's graph, in caller order, is the concatenation of its
    packings' own graphs (indices offset by the packing's first particle)."""
    R, A = configs.bidisperse(180, aspect=3.0)
    other = configs.make_packing(R, A, 0.80, 73, "second")
    m = native.Mpnn(seed=3)
    graphs = []
    try:
        for pks in ([packing], [other], [packing, other]):
            s = native.System(pks, pads=0.1)
            try:
                graphs.append(m.graph_of(s))
            finally:
                s.close()
    finally:
        m.close()
    g1, g2, g12 = graphs
    ref = native.Graph.concat([g1, g2])
    np.testing.assert_array_equal(g12.offsets, ref.offsets)
    np.testing.assert_array_equal(g12.src, ref.src)
    np.testing.assert_array_equal(g12.node_feat, ref.node_feat)
    np.testing.assert_array_equal(g12.edge_feat, ref.edge_feat)
