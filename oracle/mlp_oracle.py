"""CPU oracle of the per-column MLP surrogate -- TEST INFRASTRUCTURE ONLY.

Float64 numpy restatement of the network the GPU trains (no counterpart in
the reference): layers l = 0..L-1 with weights W_l [out][in], biases b_l,
SiLU on every hidden layer, MSE loss.  `bf16=True` rounds to bfloat16 at
exactly the points where the GPU kernels do (inputs, weights, hidden
activations, output and hidden-layer gradients), so the GPU gradients can be
compared to a float64 computation whose only difference is the fp32
accumulation order.
"""
import numpy as np


def to_bf16(x):
    """Round float values to bfloat16 (round to nearest even), as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def forward(Ws, bs, X, bf16=True):
    """Returns (Y, pres, acts) with acts[0] = X."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [r(X)]
    pres = []
    h = acts[0]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        pre = h @ r(W).T + np.asarray(b, dtype=np.float64)
        pres.append(pre)
        if l < len(Ws) - 1:
            h = r(silu(pre))
            acts.append(h)
    return pres[-1], pres, acts


def loss_and_grads(Ws, bs, X, T, bf16=True):
    """MSE = mean((Y - T)^2) and its gradients with respect to W_l, b_l."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Y, pres, acts = forward(Ws, bs, X, bf16)
    n = Y.size
    loss = float(np.mean((Y - T) ** 2))
    d = r(2.0 * (Y - T) / n)
    gWs, gbs = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gWs[l] = d.T @ acts[l]
        gbs[l] = d.sum(axis=0)
        if l > 0:
            dact = d @ r(Ws[l])
            d = r(dact * dsilu(pres[l - 1]))
    return loss, gWs, gbs


def adam_step(params, grads, state, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (bias-corrected, weight decay added to the gradient) on
    a list of arrays, in place; `state` is a dict holding t, m, v."""
    if not state:
        state.update(t=0, m=[np.zeros_like(p) for p in params],
                     v=[np.zeros_like(p) for p in params])
    state["t"] += 1
    c1 = 1.0 - b1 ** state["t"]
    c2 = 1.0 - b2 ** state["t"]
    for p, g, m, v in zip(params, grads, state["m"], state["v"]):
        g = g + wd * p
        m *= b1
        m += (1.0 - b1) * g
        v *= b2
        v += (1.0 - b2) * g * g
        p -= lr * (m / c1) / (np.sqrt(v / c2) + eps)
