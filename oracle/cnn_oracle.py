"""CPU oracle of the proposal network (test infrastructure only).

A plain numpy restatement, in float64, of what the gfx950 kernels of
vsg_cnn.hip compute: dilated 3x3 convolutions over NHWC maps of 32 channels
(padding = dilation), bias, ReLU, the mean pooling of channels 0 and 1 over
field windows, the weighted squared-error loss, and the analytic gradients of
all of them.  The learned form has no counterpart in the reference: nothing
here follows a reference file.  The gradients are pinned by central finite
differences in tests/test_cnn_oracle_cpu.py, the forward pass by a direct
per-pixel sum on tiny inputs.

Weights are [9][32][32] (tap, ci, co) with tap = (dy+1)*3 + (dx+1).
"""
import numpy as np

C = 32


def bf16_round(a):
    """round float32 values to the nearest bfloat16 (ties to even), as float64"""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shifted(x, dy, dx):
    """x[n, y+dy, x+dx, :] with zeros outside the image"""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, ye = max(0, -dy), min(h, h - dy)
    xs, xe = max(0, -dx), min(w, w - dx)
    if ys < ye and xs < xe:
        out[:, ys:ye, xs:xe, :] = x[:, ys + dy:ye + dy, xs + dx:xe + dx, :]
    return out


def conv(x, w, b, dil):
    """z[n,y,x,co] = b[co] + sum_{tap,ci} x[n, y+dy*dil, x+dx*dil, ci] * w[tap,ci,co]"""
    z = np.zeros(x.shape[:3] + (C,), dtype=np.float64)
    for tap in range(9):
        dy, dx = (tap // 3 - 1) * dil, (tap % 3 - 1) * dil
        z += np.einsum("nhwi,io->nhwo", _shifted(x, dy, dx), w[tap])
    if b is not None:
        z += b
    return z


def conv_data_grad(dz, w, dil):
    """d x from d z: the adjoint of conv (flipped taps, transposed channels)"""
    dx = np.zeros(dz.shape[:3] + (C,), dtype=np.float64)
    for tap in range(9):
        dy, dxx = (tap // 3 - 1) * dil, (tap % 3 - 1) * dil
        dx += np.einsum("nhwo,io->nhwi", _shifted(dz, -dy, -dxx), w[tap])
    return dx


def conv_weight_grad(x, dz, dil):
    """(dW [9,32,32], dB [32]) of sum(dz * conv(x, W, B))"""
    dW = np.zeros((9, C, C), dtype=np.float64)
    for tap in range(9):
        dy, dx = (tap // 3 - 1) * dil, (tap % 3 - 1) * dil
        dW[tap] = np.einsum("nhwi,nhwo->io", _shifted(x, dy, dx), dz)
    return dW, dz.sum(axis=(0, 1, 2))


def pool(m, windows):
    """windows: rows (sample, x0, y0, x1, y1); mean of channels 0, 1"""
    out = np.zeros((len(windows), 2), dtype=np.float64)
    for f, (n, x0, y0, x1, y1) in enumerate(windows):
        reg = m[n, y0:y1, x0:x1, :2]
        out[f] = reg.mean(axis=(0, 1)) if reg.size else 0.0
    return out


def pool_grad(dpred, shape, windows):
    dm = np.zeros(shape, dtype=np.float64)
    for f, (n, x0, y0, x1, y1) in enumerate(windows):
        cnt = (x1 - x0) * (y1 - y0)
        if cnt:
            dm[n, y0:y1, x0:x1, 0] += dpred[f, 0] / cnt
            dm[n, y0:y1, x0:x1, 1] += dpred[f, 1] / cnt
    return dm


def loss(pred, label, weight, inv):
    d = pred - label
    return float((weight[:, None] * d * d).sum() * inv), 2.0 * weight[:, None] * d * inv


class Net:
    """layers: list of (W [9,32,32], B [32], dil); ReLU after every layer but the last"""

    def __init__(self, layers):
        self.layers = layers

    def forward(self, x):
        acts = [x]
        for k, (w, b, dil) in enumerate(self.layers):
            z = conv(acts[-1], w, b, dil)
            acts.append(np.maximum(z, 0.0) if k + 1 < len(self.layers) else z)
        return acts

    def loss_and_grads(self, x, windows, label, weight, inv):
        acts = self.forward(x)
        pred = pool(acts[-1], windows)
        l, dpred = loss(pred, label, weight, inv)
        dz = pool_grad(dpred, acts[-1].shape, windows)
        grads = [None] * len(self.layers)
        for k in range(len(self.layers) - 1, -1, -1):
            w, b, dil = self.layers[k]
            grads[k] = conv_weight_grad(acts[k], dz, dil)
            if k > 0:
                dz = conv_data_grad(dz, w, dil) * (acts[k] > 0.0)
        return l, pred, grads
