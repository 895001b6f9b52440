"""CPU checker for the return-map network (test infrastructure only).

A numpy restatement of the network the GPU kernels evaluate: 14 -> 64 -> 64 -> 64 -> 1,
SiLU, parameters in the padded flat layout of gy_mlp (W1 16x64, W2 64x64, W3 64x64,
W4 64x16, row-major [k][n], then b1, b2, b3, b4). With `bf16=True` it rounds to
bfloat16 (round to nearest even) exactly where the GPU kernels feed the matrix cores:
weights, inputs, activations, and the back-propagated errors dY and dZ; everything else
(accumulation, biases, pre-activations, losses, bias gradients of the output layer) is
carried in float64 here, float32 on the GPU.
"""
import numpy as np

NIN, NOUT, KIN, HID, KOUT = 14, 1, 16, 64, 16
SHAPES = [(KIN, HID), (HID, HID), (HID, HID), (HID, KOUT)]
NW = sum(a * b for a, b in SHAPES)
NPARAM = NW + 3 * HID + KOUT


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    Ws, o = [], 0
    for a, b in SHAPES:
        Ws.append(p[o:o + a * b].reshape(a, b))
        o += a * b
    bs = [p[o:o + HID], p[o + HID:o + 2 * HID], p[o + 2 * HID:o + 3 * HID], p[o + 3 * HID:o + 3 * HID + KOUT]]
    return Ws, bs


def pack(Ws, bs):
    return np.concatenate([w.ravel() for w in Ws] + [b.ravel() for b in bs])


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def silu_d(z):
    s = sig(z)
    return s * (1.0 + z * (1.0 - s))


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(p, x, bf16_ops=True, keep=False):
    Ws, bs = unpack(p)
    Wq = [_r(w, bf16_ops) for w in Ws]
    X = np.zeros((x.shape[0], KIN))
    X[:, :NIN] = x
    H = [_r(X, bf16_ops)]
    Z = []
    for l in range(3):
        z = H[-1] @ Wq[l] + bs[l]
        Z.append(z)
        H.append(_r(silu(z), bf16_ops))
    Y = H[-1] @ Wq[3] + bs[3]
    if keep:
        return Y[:, :NOUT], (Wq, H, Z, Y)
    return Y[:, :NOUT]


def gradient(p, x, t, bf16_ops=True):
    """Gradient of mean((y - t)^2) over samples and the 7 outputs, and the loss."""
    n = x.shape[0]
    y, (Wq, H, Z, Y) = forward(p, x, bf16_ops, keep=True)
    dY = np.zeros_like(Y)
    dY[:, :NOUT] = 2.0 * (y - t) / (n * NOUT)
    loss = float(np.mean((y - t) ** 2))
    dYq = _r(dY, bf16_ops)
    gW = [None] * 4
    gb = [None] * 4
    gW[3] = H[3].T @ dYq
    gb[3] = dY.sum(axis=0)
    dz = dYq
    for l in (2, 1, 0):
        dh = dz @ Wq[l + 1].T
        dz = _r(dh * silu_d(Z[l]), bf16_ops)
        gW[l] = H[l].T @ dz
        gb[l] = dz.sum(axis=0)
    return pack(gW, gb), loss
