"""MLP surrogate: the float64 CPU oracle's gradients against finite
differences (CPU), and the GPU's bf16-MFMA GEMM, forward, gradients and
Adam training against the oracle (GPU)."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _params(sizes, seed):
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for i, o in zip(sizes[:-1], sizes[1:]):
        Ws.append(rng.standard_normal((o, i)) * np.sqrt(2.0 / i))
        bs.append(0.1 * rng.standard_normal(o))
    return Ws, bs


def test_oracle_gradients_match_finite_differences():
    sizes = [12, 8, 8, 6]
    Ws, bs = _params(sizes, 0)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((5, 12))
    T = rng.standard_normal((5, 6))
    loss, gW, gb = MO.loss_and_grads(Ws, bs, X, T, bf16=False)
    eps = 1e-6
    for l in range(len(Ws)):
        for (o, i) in [(0, 0), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1), (1, 2)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][o, i] += eps
            Wm[l][o, i] -= eps
            fd = (MO.loss_and_grads(Wp, bs, X, T, False)[0] -
                  MO.loss_and_grads(Wm, bs, X, T, False)[0]) / (2 * eps)
            assert abs(fd - gW[l][o, i]) <= 1e-7 * max(1.0, abs(fd)), (l, o, i, fd, gW[l][o, i])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += eps
        bm[l][0] -= eps
        fd = (MO.loss_and_grads(Ws, bp, X, T, False)[0] -
              MO.loss_and_grads(Ws, bm, X, T, False)[0]) / (2 * eps)
        assert abs(fd - gb[l][0]) <= 1e-7 * max(1.0, abs(fd))


def test_bf16_rounding_is_round_to_nearest_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even
    # mantissa (1.0); 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 and
    # goes up to the even one; above-half rounds up, below-half down; values
    # past the largest bf16 round to infinity.
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -12,
                  1.0 + 2 ** -9, -2.5, 0.0, 3.0e38, 3.4e38])
    r = MO.to_bf16(x)
    assert list(r[:7]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, 1.0, -2.5, 0.0]
    assert r[7] == 2.0 ** 127 * 226 / 128   # 3e38 = 2^127 * 1.7632.. -> 226/128
    assert np.isinf(r[8])


# ---------------------------------------------------------------------- GPU
class GpuMLP:
    def __init__(self, ctx, sizes, max_batch):
        from miso_amd import hip
        self.hip, self.ctx = hip, ctx
        h = C.c_void_p()
        hip.call("misohip_mlp_create", ctx.handle, sizes[0], sizes[1], len(sizes) - 2, sizes[-1],
                 max_batch, C.byref(h))
        self.h = h
        self.sizes = sizes

    def set(self, Ws, bs):
        for l, (W, b) in enumerate(zip(Ws, bs)):
            w32 = np.ascontiguousarray(W, dtype=np.float32)
            b32 = np.ascontiguousarray(b, dtype=np.float32)
            self.hip.call("misohip_mlp_set_layer", self.ctx.handle, self.h, l,
                          w32.ctypes.data_as(C.c_void_p), b32.ctypes.data_as(C.c_void_p))

    def get(self, l, which):
        o, i = self.sizes[l + 1], self.sizes[l]
        out = np.zeros((o, i) if which in (0, 2) else (o,), dtype=np.float32)
        self.hip.call("misohip_mlp_get_layer", self.ctx.handle, self.h, l, which,
                      out.ctypes.data_as(C.c_void_p))
        return out.astype(np.float64)

    def close(self):
        self.hip.lib().misohip_mlp_destroy(self.h)


def _bf16_dev(x):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to("cuda:0")
    return t.to(torch.bfloat16)


@pytest.mark.gpu
def test_gpu_forward_and_gradients_match_oracle(ctx):
    import torch
    sizes = [96, 64, 64, 64, 96]   # same structure as the surrogate, small
    Ws, bs = _params(sizes, 3)
    rng = np.random.default_rng(4)
    batch = 256
    X = MO.to_bf16(rng.standard_normal((batch, sizes[0])))
    T = rng.standard_normal((batch, sizes[-1]))
    net = GpuMLP(ctx, sizes, batch)
    net.set(Ws, bs)
    Wr = [MO.to_bf16(w.astype(np.float32)) for w in Ws]  # the GPU's bf16 copies
    br = [b.astype(np.float32).astype(np.float64) for b in bs]
    Xd = _bf16_dev(X)
    Y = torch.zeros((batch, sizes[-1]), dtype=torch.float32, device="cuda:0")
    net.hip.call("misohip_mlp_forward", ctx.handle, net.h, C.c_void_p(Xd.data_ptr()), batch,
                 C.c_void_p(Y.data_ptr()))
    torch.cuda.synchronize()
    Yref, _, _ = MO.forward(Wr, br, X, bf16=True)
    err = np.abs(Y.cpu().numpy() - Yref).max() / np.abs(Yref).max()
    assert err < 2e-3, err
    Td = torch.from_numpy(T.astype(np.float32)).to("cuda:0")
    loss = C.c_double()
    net.hip.call("misohip_mlp_grad", ctx.handle, net.h, C.c_void_p(Xd.data_ptr()),
                 C.c_void_p(Td.data_ptr()), batch, C.byref(loss))
    lref, gW, gb = MO.loss_and_grads(Wr, br, X, T.astype(np.float32).astype(np.float64), bf16=True)
    assert abs(loss.value - lref) <= 1e-3 * lref
    for l in range(len(Ws)):
        g = net.get(l, 2)
        rel = np.linalg.norm(g - gW[l]) / np.linalg.norm(gW[l])
        assert rel < 2e-2, (l, rel)
        gbl = net.get(l, 3)
        relb = np.linalg.norm(gbl - gb[l]) / np.linalg.norm(gb[l])
        assert relb < 2e-2, (l, relb)
    net.close()


@pytest.mark.gpu
def test_gpu_gemm_odd_shapes(ctx):
    """Forward GEMM with a batch that is not a tile multiple and the split-K
    weight gradient with a batch that is."""
    import torch
    sizes = [64, 32, 64]
    Ws, bs = _params(sizes, 5)
    batch = 160  # 1.25 tiles of 128 rows
    rng = np.random.default_rng(6)
    X = MO.to_bf16(rng.standard_normal((batch, sizes[0])))
    net = GpuMLP(ctx, sizes, 512)
    net.set(Ws, bs)
    Wr = [MO.to_bf16(w.astype(np.float32)) for w in Ws]
    br = [b.astype(np.float32).astype(np.float64) for b in bs]
    Y = torch.zeros((batch, sizes[-1]), dtype=torch.float32, device="cuda:0")
    net.hip.call("misohip_mlp_forward", ctx.handle, net.h, C.c_void_p(_bf16_dev(X).data_ptr()),
                 batch, C.c_void_p(Y.data_ptr()))
    torch.cuda.synchronize()
    Yref, _, _ = MO.forward(Wr, br, X)
    assert np.abs(Y.cpu().numpy() - Yref).max() <= 2e-3 * np.abs(Yref).max()
    net.close()


def _oracle_training(Ws, bs, Xs, Ts, epochs, batch, lr):
    Ws = [w.astype(np.float32).astype(np.float64) for w in Ws]
    bs = [b.astype(np.float32).astype(np.float64) for b in bs]
    st, losses = {}, []
    for _ in range(epochs):
        for b0 in range(0, len(Xs), batch):
            loss, gW, gb = MO.loss_and_grads(Ws, bs, Xs[b0:b0 + batch], Ts[b0:b0 + batch])
            MO.adam_step(Ws + bs, gW + gb, st, lr)
            losses.append(loss)
    return losses


def test_oracle_adam_single_step_closed_form():
    # first Adam step moves every parameter by lr * sign(g) (m/c1 = g,
    # sqrt(v/c2) = |g|), up to eps
    p = [np.array([1.0, -2.0, 3.0])]
    g = [np.array([0.5, -1e-3, 7.0])]
    MO.adam_step(p, g, {}, 0.01, eps=0.0)
    assert np.allclose(p[0], [0.99, -1.99, 2.99], rtol=0, atol=1e-15)


@pytest.mark.gpu
def test_gpu_training_tracks_oracle(ctx):
    """Adam on the GPU (bf16 GEMMs, fp32 master weights) follows the float64
    oracle's training curve on a smooth synthetic map."""
    import torch
    sizes = [96, 128, 128, 128, 64]
    Ws, bs = _params(sizes, 7)
    for w in Ws:
        w *= 0.5
    rng = np.random.default_rng(8)
    n, batch, epochs, lr = 4096, 1024, 30, 2e-3
    Xs = MO.to_bf16(rng.standard_normal((n, sizes[0])))
    A = rng.standard_normal((sizes[0], sizes[-1])) / np.sqrt(sizes[0])
    Ts = np.tanh(Xs @ A).astype(np.float32).astype(np.float64)
    net = GpuMLP(ctx, sizes, batch)
    net.set(Ws, bs)
    Xd = _bf16_dev(Xs)
    Td = torch.from_numpy(Ts.astype(np.float32)).to("cuda:0")
    losses = []
    for ep in range(epochs):
        for b0 in range(0, n, batch):
            loss = C.c_double()
            net.hip.call("misohip_mlp_grad", ctx.handle, net.h,
                         C.c_void_p(Xd[b0:b0 + batch].data_ptr()),
                         C.c_void_p(Td[b0:b0 + batch].data_ptr()), batch, C.byref(loss))
            net.hip.call("misohip_mlp_adam", ctx.handle, net.h, C.c_float(lr), C.c_float(0.9),
                         C.c_float(0.999), C.c_float(1e-8), C.c_float(0.0))
            losses.append(loss.value)
    net.close()
    ref = _oracle_training(Ws, bs, Xs, Ts, epochs, batch, lr)
    assert losses[0] == pytest.approx(ref[0], rel=1e-3)
    assert losses[-1] < 0.2 * losses[0]
    # bf16 rounding makes the curves drift apart slowly; they stay close
    for k in (3, 15, 60, len(ref) - 1):
        assert abs(losses[k] - ref[k]) <= 0.1 * ref[k] + 1e-4, (k, losses[k], ref[k])
