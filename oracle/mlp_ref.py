"""Float64 numpy restatement of the radiance-cache MLP -- TEST INFRASTRUCTURE ONLY.

Checker for crystal2_amd's MFMA kernels (used by tests/ only): the same input
encoding, 64 -> 4 x 64 ReLU -> 3 network and relative-L2 loss
(|y - t|^2 / (y^2 + 0.01), denominator held constant), with the GPU's fp16
rounding of weights, activations and backpropagated gradients emulated so the
comparison isolates summation-order differences. The analytic backward pass is
itself checked against central finite differences in float64 (tests/test_nrc.py).
"""
import numpy as np

W_LAYERS, WIDTH, OUT = 5, 64, 3
NPARAMS = W_LAYERS * WIDTH * WIDTH + W_LAYERS * WIDTH


def unpack(params):
    p = np.asarray(params, np.float64)
    W = p[: W_LAYERS * WIDTH * WIDTH].reshape(W_LAYERS, WIDTH, WIDTH)  # [layer][in][out]
    b = p[W_LAYERS * WIDTH * WIDTH:].reshape(W_LAYERS, WIDTH)
    return W, b


def pack(W, b):
    return np.concatenate([W.reshape(-1), b.reshape(-1)])


def encode(raw, bmin, bmax):
    raw = np.asarray(raw, np.float64)
    n = len(raw)
    bmin = np.asarray(bmin, np.float64)
    ext = np.asarray(bmax, np.float64) - bmin
    binv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1), 1.0)
    out = np.zeros((n, 64))
    for c in range(3):
        u = (raw[:, c] - bmin[c]) * binv[c]
        for band in range(8):
            x = u * (1 << band) * np.pi
            out[:, 16 * c + 2 * band] = np.sin(x)
            out[:, 16 * c + 2 * band + 1] = np.cos(x)
    out[:, 48] = np.arctan2(raw[:, 5], raw[:, 3]) / (2 * np.pi) + 0.5
    out[:, 49] = raw[:, 4] * 0.5 + 0.5
    out[:, 50] = np.arctan2(raw[:, 8], raw[:, 6]) / (2 * np.pi) + 0.5
    out[:, 51] = raw[:, 7] * 0.5 + 0.5
    for k in range(4):
        c = (k + 0.5) * 0.25
        out[:, 52 + k] = np.exp(-((raw[:, 9] - c) * 4) ** 2)
    out[:, 56:59] = raw[:, 10:13]
    out[:, 59] = 1.0
    return out


def _h(x, fp16):
    return x.astype(np.float16).astype(np.float64) if fp16 else x


def forward(params, X, fp16=True):
    W, b = unpack(params)
    Wq = _h(W, fp16)
    H = [_h(X, fp16)]
    for L in range(4):
        H.append(_h(np.maximum(H[-1] @ Wq[L] + b[L], 0.0), fp16))
    y = H[-1] @ Wq[4][:, :OUT] + b[4][:OUT]
    return H, y


def loss_and_grad(params, X, targets, fp16=True):
    """Mean relative-L2 loss over n x 3 and its gradient (same layout as params)."""
    W, b = unpack(params)
    Wq = _h(W, fp16)
    H, y = forward(params, X, fp16)
    t = np.asarray(targets, np.float64)
    n = len(X)
    den = y * y + 0.01
    loss = np.sum((y - t) ** 2 / den) / (3 * n)
    D = _h(2 * (y - t) / den / 3.0, fp16)  # per-row gradient; the 1/n is applied to the sums
    gW = np.zeros_like(W)
    gb = np.zeros_like(b)
    gW[4][:, :OUT] = H[4].T @ D / n
    gb[4][:OUT] = D.sum(0) / n
    Dl = D
    Wl_out = Wq[4][:, :OUT]
    for L in range(3, -1, -1):
        Dl = _h((Dl @ Wl_out.T) * (H[L + 1] > 0), fp16)
        gW[L] = H[L].T @ Dl / n
        gb[L] = Dl.sum(0) / n
        Wl_out = Wq[L]
    return loss, pack(gW, gb)
