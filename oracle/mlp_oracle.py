"""CPU restatement of the learned dispatch-block proposal's network (test infrastructure: only
``tests/`` load it; the product path runs ``k_mlp.hip``).

The network and its gradient in FP64 with the device's rounding points: inputs, weights and the
hidden activations rounded to BF16 (round to nearest even), FP32 outputs; the backward pass
rounds dL/dout and the masked hidden gradients to BF16 as the device's GEMM operands are. The
mixture head (log q(y|c), its gradient) and the instance features follow the definitions of
``include/esm.h`` (esm_mlp_*). No counterpart in the reference (the extension is new); pinned by
finite differences of the FP64 loss and by the exactness test of the learned move (the chain's
marginals against quadrature of exp(−β ΔC), with ΔC from the oracle's gas atom).
"""
from __future__ import annotations

import numpy as np

LOG_SQRT_2PI = 0.91893853320467274178


def bf16(a):
    """Round to BF16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where((u & 0x7F800000) == 0x7F800000, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def exact(a):
    return np.asarray(a, dtype=np.float64)


def forward(layers, X, rnd=bf16, out_rnd=f32):
    """Outputs and the inputs of every layer [X, H1, H2, H3]; rnd: the rounding of the inputs,
    weights and hidden activations (BF16 on the device, ``exact`` for plain FP64)."""
    acts = [rnd(X)]
    h = acts[0]
    for li, (W, b) in enumerate(layers):
        pre = h @ rnd(W).T + (f32(b) if rnd is bf16 else exact(b))
        if li < 3:
            h = rnd(np.maximum(pre, 0.0))
            acts.append(h)
        else:
            return (out_rnd(pre) if rnd is bf16 else pre), acts
    raise AssertionError


def mixture(out, mu0, K, B):
    """Per hour b: log weights [M, K, B] (log-softmax over k), means and log sds [M, K, B] of the
    rows; layout [logits K·B][means K·B][log sd K·B], index k·B + b (FP64; the FP32 arithmetic of
    the device is not emulated)."""
    M = out.shape[0]
    lg = out[:, :K * B].reshape(M, K, B)
    lw = lg - lg.max(axis=1, keepdims=True)
    lw = lw - np.log(np.exp(lw).sum(axis=1, keepdims=True))
    mu = mu0[:, None, :] + out[:, K * B:2 * K * B].reshape(M, K, B)
    raw = out[:, 2 * K * B:3 * K * B].reshape(M, K, B)
    ls = np.clip(raw, -7.0, 2.0)
    return lw, mu, ls, raw


def logq(out, mu0, y, K, B):
    """log q(y|c) = Σ_b log Σ_k w_kb N(y_b; μ_kb, σ_kb) [M] and the responsibilities [M, K, B]."""
    lw, mu, ls, _ = mixture(out, mu0, K, B)
    z = (y[:, None, :] - mu) * np.exp(-ls)
    lk = lw - 0.5 * z * z - ls - LOG_SQRT_2PI
    mx = lk.max(axis=1, keepdims=True)
    lqb = mx[:, 0, :] + np.log(np.exp(lk - mx).sum(axis=1))
    return lqb.sum(axis=1), np.exp(lk - lqb[:, None, :])


def head_grad(out, mu0, y, K, B, dout):
    """Loss Σ −log q and d(mean loss)/d out [M, dout] (zeros beyond 3KB)."""
    M = out.shape[0]
    lq, resp = logq(out, mu0, y, K, B)
    lw, mu, ls, raw = mixture(out, mu0, K, B)
    sd = np.exp(ls)
    z = (y[:, None, :] - mu) / sd
    g = np.zeros((M, dout))
    g[:, :K * B] = (np.exp(lw) - resp).reshape(M, -1)
    g[:, K * B:2 * K * B] = (-resp * z / sd).reshape(M, -1)
    free = (raw > -7.0) & (raw < 2.0)
    g[:, 2 * K * B:3 * K * B] = np.where(free, resp * (1.0 - z * z), 0.0).reshape(M, -1)
    return -lq.sum(), g / M


def backward(layers, acts, dout_grad, rnd=bf16):
    """Gradients [(dWt, db)] of the layers from d(mean loss)/d out (with the device's BF16 points,
    or none with rnd = exact)."""
    g = rnd(dout_grad)
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        a = acts[li]
        grads[li] = (g.T @ a, g.sum(axis=0))
        if li > 0:
            g = rnd((g @ rnd(W)) * (acts[li] > 0))
    return grads


def flatten(grads):
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def loss_fp64(layers, X, y, mu0, K, B):
    """The mean loss of the network in plain FP64 (no rounding): the finite-difference reference."""
    h = np.asarray(X, dtype=np.float64)
    for li, (W, b) in enumerate(layers):
        h = h @ W.T + b
        if li < 3:
            h = np.maximum(h, 0.0)
    return -logq(h, mu0, y, K, B)[0].mean()


def features(sys, v, k, x, r, F, B, R, din=160):
    """X [n_inst, din] (BF16 values), Y and MU0 [n_inst, B] of every instance
    ((node·nb + block)·R + replica) of the state (v [4][N][T][R], raw capacities k [6][N], raw
    capacity factors x [3][N][T][R], residual r [N][T][R], flows F [L][T][R], positive from bus0
    to bus1). Features: [Q/κ per hour][capacity factors per carrier and hour][capacities over the
    system means][1][net line inflow over κ per hour][largest loading |F|/s of the node's lines
    per hour][0 …]."""
    N, T = sys.N, sys.T
    nb = T // B
    kc = np.clip(k, sys.klo, sys.khi)
    kap = kc[3]
    kmean = k.mean(axis=1)
    rows, Ys, M0 = [], [], []
    for n in range(N):
        out_l = np.flatnonzero(sys.line_bus0 == n)
        in_l = np.flatnonzero(sys.line_bus1 == n)
        inc = np.concatenate([out_l, in_l])
        for d in range(nb):
            tau = np.arange(d * B, d * B + B)
            for j in range(R):
                y = v[1, n, tau, j]
                Q = r[n, tau, j] + kap[n] * np.clip(y, 0, 1)
                q = f32(Q / kap[n]) if kap[n] > 0 else np.zeros(B)
                feat = np.zeros(din)
                feat[:B] = q
                for c in range(3):
                    feat[B + c * B:B + (c + 1) * B] = x[c, n, tau, j]
                feat[4 * B:4 * B + 6] = np.where(kmean > 0, k[:, n] / np.where(kmean > 0, kmean, 1.0), 0.0)
                feat[4 * B + 6] = 1.0
                inflow = F[in_l][:, tau, j].sum(axis=0) - F[out_l][:, tau, j].sum(axis=0)
                feat[4 * B + 7:5 * B + 7] = inflow / kap[n] if kap[n] > 0 else 0.0
                load = np.abs(F[inc][:, tau, j]) / sys.line_s[inc][:, None]
                feat[5 * B + 7:6 * B + 7] = load.max(axis=0) if len(inc) else 0.0
                rows.append(bf16(feat))
                Ys.append(y)
                M0.append(np.clip(q, 0.0, 1.0))
    return np.array(rows), np.array(Ys), np.array(M0)
