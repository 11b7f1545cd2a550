"""FP64 CPU reference of the chemistry-surrogate MLP: forward pass, Gaussian NLL and the
full parameter gradient by back-propagation, in numpy float64.

TEST INFRASTRUCTURE: imported only by tests/ as the checker of the GPU forward/backward
kernels.  The surrogate has no counterpart in the reference, so nothing here follows a
reference file; the back-propagation is pinned by central finite differences in
tests/test_surrogate_cpu.py.

Parameter layout (one flat vector): for each layer l the weights [din_l][dout_l]
row-major (din_0 = n_in rounded up to a multiple of 32, padded rows zero), then the
biases of every layer.  Hidden layers use ReLU; the last layer emits [mu | s] with s the
log-variance.  Loss = mean over (batch, output) of 0.5 (s' + (t - mu)^2 exp(-s')),
s' = clamp(s, -12, 12) (zero gradient outside the clamp).

bf16=True rounds to bfloat16 (round to nearest even) where the GPU stores BF16: the
packed inputs, the weights, every hidden activation, the output gradient and every
back-propagated hidden gradient.  Products are then exact and sums are done in float64,
so the only remaining difference to the GPU is its FP32 accumulation order.
"""
import numpy as np

S_CLAMP = 12.0


def bf16_round(x):
    """float -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_dims(n_in, n_hidden, n_layers, n_out):
    pad = (n_in + 31) // 32 * 32
    dims = []
    for l in range(n_layers + 1):
        di = pad if l == 0 else n_hidden
        do = 2 * n_out if l == n_layers else n_hidden
        dims.append((di, do))
    return dims


def unflatten(flat, dims):
    flat = np.asarray(flat, np.float64)
    Ws, bs, off = [], [], 0
    for di, do in dims:
        Ws.append(flat[off:off + di * do].reshape(di, do))
        off += di * do
    for di, do in dims:
        bs.append(flat[off:off + do])
        off += do
    assert off == flat.size
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([w.ravel() for w in Ws] + [b.ravel() for b in bs])


def forward(flat, dims, X, bf16=False):
    """Returns (Y [B, 2 n_out], list of layer inputs)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unflatten(flat, dims)
    pad = dims[0][0]
    B, n_in = X.shape
    x = np.zeros((B, pad))
    x[:, :n_in] = X
    h = r(x)
    ins = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        ins.append(h)
        z = h @ r(W) + b
        if l < len(Ws) - 1:
            h = r(np.maximum(z, 0.0))
        else:
            h = z
    return h, ins


def nll(Y, T):
    """Mean NLL and dL/dY."""
    no = T.shape[1]
    mu, s = Y[:, :no], Y[:, no:]
    sc = np.clip(s, -S_CLAMP, S_CLAMP)
    d = mu - T
    e = np.exp(-sc)
    cnt = T.size
    loss = 0.5 * (sc + d * d * e)
    g_mu = d * e / cnt
    g_s = np.where((s > -S_CLAMP) & (s < S_CLAMP), 0.5 * (1.0 - d * d * e), 0.0) / cnt
    return loss.sum() / cnt, np.concatenate([g_mu, g_s], axis=1)


def loss_and_grad(flat, dims, X, T, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unflatten(flat, dims)
    Y, ins = forward(flat, dims, X, bf16)
    loss, dY = nll(Y, T)
    d = r(dY)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW[l] = ins[l].T @ d
        gb[l] = d.sum(axis=0)
        if l == 0:
            break
        d = r((d @ r(Ws[l]).T) * (ins[l] > 0))
    return loss, flatten(gW, gb)
