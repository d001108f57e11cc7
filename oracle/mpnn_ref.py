"""ORACLE -- test infrastructure only (tests/, never the product path).

CPU restatement (numpy, FP64) of the message-passing warm-start network of
spherocyls_amd/csrc/sc_mpnn.hip: forward pass, mean-squared-error loss and the
gradient with respect to every parameter, written out by hand (the chain rule
through the residual layers, the edge sums and their reverse edges).  With
``bf16=True`` every GEMM operand is rounded to bfloat16 first, as the device
does, so device and oracle see the same quantised operands and differ only
by FP32 accumulation order and FP32 elementwise rounding on the device.

Pinned by central finite differences of the FP64 loss (tests/test_mpnn_cpu.py).
"""
from __future__ import annotations

import numpy as np

H, LAYERS = 64, 3
KX, KE, KU, KO, NO = 16, 144, 144, 80, 16
NODE_FEAT, EDGE_FEAT = 4, 5


def layout():
    o = 0
    out = {"we": (o, KX, H)}
    o += KX * H
    for l in range(LAYERS):
        out[f"wm{l}"] = (o, KE, H)
        o += KE * H
        out[f"wu{l}"] = (o, KU, H)
        o += KU * H
    out["wo"] = (o, KO, NO)
    o += KO * NO
    return out, o


def num_params():
    return layout()[1]


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _weights(params):
    lay, _ = layout()
    return {k: np.asarray(params[o:o + K * N], dtype=np.float64).reshape(K, N)
            for k, (o, K, N) in lay.items()}


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _dst(off):
    n = len(off) - 1
    return np.repeat(np.arange(n), np.diff(off))


def _rev(off, src):
    dst = _dst(off)
    key = {(int(d), int(s)): k for k, (d, s) in enumerate(zip(dst, src))}
    return np.array([key[(int(s), int(d))] for d, s in zip(dst, src)], dtype=np.int64)


def forward(params, nf, off, src, ef, bf16=False):
    """Returns (y [n,3], cache)."""
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W = {k: q(v) for k, v in _weights(params).items()}
    nf = np.asarray(nf, dtype=np.float64)
    ef = np.asarray(ef, dtype=np.float64).reshape(-1, EDGE_FEAT)
    off = np.asarray(off)
    src = np.asarray(src, dtype=np.int64)
    n, e = nf.shape[0], src.shape[0]
    dst = _dst(off)
    X = np.zeros((n, KX))
    X[:, :NODE_FEAT] = nf
    X[:, NODE_FEAT] = 1.0
    X = q(X)
    Ze = X @ W["we"]
    Hs = [silu(Ze)]
    cache = {"X": X, "Ze": Ze, "Ein": [], "Z": [], "Uin": [], "U": [], "dst": dst, "src": src,
             "off": off, "W": W, "n": n, "e": e}
    for l in range(LAYERS):
        h = Hs[-1]
        Ein = np.zeros((e, KE))
        Ein[:, :H] = h[dst]
        Ein[:, H:2 * H] = h[src]
        Ein[:, 2 * H:2 * H + EDGE_FEAT] = ef
        Ein[:, 2 * H + EDGE_FEAT] = 1.0
        Ein = q(Ein)
        Z = Ein @ W[f"wm{l}"]
        M = np.zeros((n, H))
        np.add.at(M, dst, silu(Z))
        Uin = np.zeros((n, KU))
        Uin[:, :H] = h
        Uin[:, H:2 * H] = M
        Uin[:, 2 * H] = 1.0
        Uin = q(Uin)
        U = Uin @ W[f"wu{l}"]
        Hs.append(h + silu(U))
        cache["Ein"].append(Ein)
        cache["Z"].append(Z)
        cache["Uin"].append(Uin)
        cache["U"].append(U)
    Hin = np.zeros((n, KO))
    Hin[:, :H] = Hs[-1]
    Hin[:, H] = 1.0
    Hin = q(Hin)
    Y = Hin @ W["wo"]
    cache.update(Hs=Hs, Hin=Hin, Y=Y)
    return Y[:, :3], cache


def loss_and_grad(params, nf, off, src, ef, target, bf16=False):
    """Mean over nodes and 3 outputs of (y - t)^2, and d loss / d params."""
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(params, nf, off, src, ef, bf16=bf16)
    n, e, W = c["n"], c["e"], c["W"]
    t = np.asarray(target, dtype=np.float64).reshape(n, 3)
    d = y - t
    loss = float(np.sum(d * d) / (3 * n))
    dY = np.zeros((n, NO))
    dY[:, :3] = 2.0 * d / (3 * n)
    dY = q(dY)
    lay, total = layout()
    G = np.zeros(total)

    def put(name, g):
        o, K, N = lay[name]
        G[o:o + K * N] = g.reshape(-1)

    put("wo", c["Hin"].T @ dY)
    dH = (dY @ W["wo"].T)[:, :H]
    rev = _rev(c["off"], c["src"])
    for l in reversed(range(LAYERS)):
        dU = q(dH * silu_grad(c["U"][l]))
        put(f"wu{l}", c["Uin"][l].T @ dU)
        dUin = dU @ W[f"wu{l}"].T
        dM = dUin[:, H:2 * H]
        dZ = q(dM[c["dst"]] * silu_grad(c["Z"][l]))
        put(f"wm{l}", c["Ein"][l].T @ dZ)
        dEin = dZ @ W[f"wm{l}"].T
        dprev = dH + dUin[:, :H]
        np.add.at(dprev, c["dst"], dEin[:, :H])
        np.add.at(dprev, c["dst"], dEin[rev, H:2 * H])
        dH = dprev
    dZe = q(dH * silu_grad(c["Ze"]))
    put("we", c["X"].T @ dZe)
    return loss, G
