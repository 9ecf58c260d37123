"""The proposal network of the learned detect form: a driver over the gfx950
kernels declared in vsg_cnn.h (forward, data / weight gradients, pooling,
loss, Adam).  torch provides device buffers and the stream only; every
arithmetic step runs in the project's own HIP kernels.

Network: 6 dilated 3x3 convolutions, 32 channels, BF16 weights and
activations, FP32 accumulation; input = the 4x4-downsampled luma of frames
t-1 (channel 0) and t (channel 1) as (mean - 128) / 64; output channels 0, 1
= displacement (dx, dy) in quarter-resolution pixels, averaged over each
field's window to give that field's proposal.

Weights file (little-endian): magic b"VSGP", int32 version, int32 layers,
int32 dilations[layers], float32 residual threshold (the window search's
acceptance bound in match units: twice the exact search's p95 match on the
training clips), then per layer float32 W[9][32][32] and B[32].
"""
import ctypes
import struct

import numpy as np

from ._paths import lib_path

C = 32
DILATIONS = (1, 2, 4, 8, 4, 1)
MAGIC = b"VSGP"

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(lib_path("libvidstab_amd.so"))
        vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
        L.vsgc_pack_weights.argtypes = [vp, vp, i32, vp]
        L.vsgc_conv.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, vp, vp, vp]
        L.vsgc_conv_k.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, i32, vp, vp, vp]
        L.vsgc_pack_weights16.argtypes = [vp, vp, vp]
        L.vsgc_conv16.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, i32, vp, vp, vp]
        L.vsgc_conv16_direct.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, i32, vp, vp, vp]
        L.vsgc_conv16_pair.argtypes = [vp, ctypes.c_size_t, vp, vp, i32, i32, i32, i32, vp, vp, i32, i32, vp, vp, vp]
        L.vsgc_conv16_pair2.argtypes = [vp, ctypes.c_size_t, vp, vp, i32, i32, i32, i32, vp, vp, i32, vp, vp, i32,
                                        i32, vp, vp, vp]
        L.vsgc_field_pool2.argtypes = [vp, i32, i32, vp, i32, vp, vp]
        L.vsgc_conv_wgrad.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, vp]
        L.vsgc_relu_mask.argtypes = [vp, vp, i64, vp]
        L.vsgc_f32_to_bf16.argtypes = [vp, vp, i64, vp]
        L.vsgc_downsample.argtypes = [vp, i32, i32, i32, vp, i32, vp]
        L.vsgc_field_pool.argtypes = [vp, i32, i32, vp, i32, vp, vp]
        L.vsgc_field_pool_bwd.argtypes = [vp, i32, i32, vp, i32, vp, vp]
        L.vsgc_field_loss.argtypes = [vp, vp, vp, i32, f32, vp, vp, vp]
        L.vsgc_adam.argtypes = [vp, vp, vp, vp, i32, f32, f32, f32, f32, i32, vp]
        _lib = L
    return _lib


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: HIP error {rc}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bf16_bits(a):
    """float array -> int16 tensor-ready bit patterns (round to nearest even)"""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16).view(np.int16)


def bits_to_float(b):
    b = np.ascontiguousarray(b).view(np.uint16).astype(np.uint32) << 16
    return b.view(np.float32)


def save_weights(path, layers, threshold):
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<ii", 1, len(layers)))
        f.write(struct.pack(f"<{len(layers)}i", *[d for (_, _, d) in layers]))
        f.write(struct.pack("<f", float(threshold)))
        for w, b, _ in layers:
            f.write(np.ascontiguousarray(w, dtype="<f4").tobytes())
            f.write(np.ascontiguousarray(b, dtype="<f4").tobytes())


def load_weights(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError("not a proposal weights file")
    ver, nl = struct.unpack_from("<ii", data, 4)
    dils = struct.unpack_from(f"<{nl}i", data, 12)
    (thr,) = struct.unpack_from("<f", data, 12 + 4 * nl)
    off = 16 + 4 * nl
    layers = []
    for k in range(nl):
        w = np.frombuffer(data, dtype="<f4", count=9 * C * C, offset=off).reshape(9, C, C).copy()
        off += 9 * C * C * 4
        b = np.frombuffer(data, dtype="<f4", count=C, offset=off).copy()
        off += C * 4
        layers.append((w, b, dils[k]))
    return layers, thr


class ProposalNet:
    """Device-resident network state: FP32 master weights, BF16 packings,
    Adam moments; forward / backward over NHWC-32 BF16 maps (int16 tensors)."""

    def __init__(self, layers=None, seed=0, dilations=DILATIONS):
        import torch
        self.torch = torch
        if layers is None:
            rng = np.random.default_rng(seed)
            layers = []
            for k, d in enumerate(dilations):
                fan_in = 9 * (2 if k == 0 else C)
                w = rng.standard_normal((9, C, C)).astype(np.float32) * np.sqrt(2.0 / fan_in)
                if k == 0:
                    w[:, 2:, :] = 0.0          # only channels 0, 1 carry input
                if k == len(dilations) - 1:
                    w[:, :, 2:] = 0.0          # only channels 0, 1 are read out
                    w *= 0.1
                layers.append((w, np.zeros(C, np.float32), d))
        dev = "cuda"
        self.dils = [d for (_, _, d) in layers]
        self.w = [torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(dev) for (w, _, _) in layers]
        self.b = [torch.from_numpy(np.ascontiguousarray(b, np.float32)).to(dev) for (_, b, _) in layers]
        self.pk = [torch.empty(9 * 4 * 64 * 4, dtype=torch.int16, device=dev) for _ in layers]
        self.pkT = [torch.empty(9 * 4 * 64 * 4, dtype=torch.int16, device=dev) for _ in layers]
        self.gw = [torch.zeros_like(w) for w in self.w]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.mw = [torch.zeros_like(w) for w in self.w]
        self.vw = [torch.zeros_like(w) for w in self.w]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step = 0
        self.pack()

    @property
    def nlayers(self):
        return len(self.w)

    def pack(self):
        L, st = lib(), _stream()
        for k in range(self.nlayers):
            _chk(L.vsgc_pack_weights(_p(self.w[k]), _p(self.pk[k]), 0, st), "pack")
            _chk(L.vsgc_pack_weights(_p(self.w[k]), _p(self.pkT[k]), 1, st), "pack")

    def layers_numpy(self):
        return [(self.w[k].cpu().numpy(), self.b[k].cpu().numpy(), self.dils[k]) for k in range(self.nlayers)]

    def forward(self, x):
        """x: int16 [N,H,W,32] (BF16). Returns the BF16 activations of layers
        0..L-1 (input first) and the FP32 output map."""
        torch = self.torch
        L, st = lib(), _stream()
        n, h, w = x.shape[0], x.shape[1], x.shape[2]
        acts = [x]
        out = None
        for k in range(self.nlayers):
            last = k == self.nlayers - 1
            cin = 2 if k == 0 else C      # the input map carries frames t-1, t in channels 0, 1
            if last:
                out = torch.empty((n, h, w, C), dtype=torch.float32, device=x.device)
                _chk(L.vsgc_conv_k(_p(acts[-1]), n, h, w, _p(self.pk[k]), _p(self.b[k]), self.dils[k], 2, cin, None,
                                   _p(out), st), "conv")
            else:
                y = torch.empty((n, h, w, C), dtype=torch.int16, device=x.device)
                _chk(L.vsgc_conv_k(_p(acts[-1]), n, h, w, _p(self.pk[k]), _p(self.b[k]), self.dils[k], 0, cin, _p(y),
                                   None, st), "conv")
                acts.append(y)
        return acts, out

    def backward(self, acts, dout):
        """dout: FP32 gradient at the output map; accumulates into gw / gb"""
        torch = self.torch
        L, st = lib(), _stream()
        n, h, w = acts[0].shape[0], acts[0].shape[1], acts[0].shape[2]
        cnt = n * h * w * C
        dz = torch.empty((n, h, w, C), dtype=torch.int16, device=dout.device)
        _chk(L.vsgc_f32_to_bf16(_p(dout), _p(dz), cnt, st), "f2bf")
        for k in range(self.nlayers - 1, -1, -1):
            _chk(L.vsgc_conv_wgrad(_p(acts[k]), _p(dz), n, h, w, self.dils[k], _p(self.gw[k]), _p(self.gb[k]), st),
                 "wgrad")
            if k == 0:
                break
            dx = torch.empty_like(dz)
            _chk(L.vsgc_conv(_p(dz), n, h, w, _p(self.pkT[k]), None, self.dils[k], 1, _p(dx), None, st), "dgrad")
            _chk(L.vsgc_relu_mask(_p(dx), _p(acts[k]), cnt, st), "relu mask")
            dz = dx

    def zero_grad(self):
        for g in self.gw + self.gb:
            g.zero_()

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        L, st = lib(), _stream()
        self.step += 1
        for k in range(self.nlayers):
            _chk(L.vsgc_adam(_p(self.w[k]), _p(self.gw[k]), _p(self.mw[k]), _p(self.vw[k]), self.w[k].numel(), lr,
                             b1, b2, eps, self.step, st), "adam")
            _chk(L.vsgc_adam(_p(self.b[k]), _p(self.gb[k]), _p(self.mb[k]), _p(self.vb[k]), self.b[k].numel(), lr,
                             b1, b2, eps, self.step, st), "adam")
        self.pack()

    def pool(self, out, windows):
        """windows: int32 tensor [nf, 5] -> FP32 [nf, 2]"""
        torch = self.torch
        pred = torch.empty((windows.shape[0], 2), dtype=torch.float32, device=out.device)
        _chk(lib().vsgc_field_pool(_p(out), out.shape[1], out.shape[2], _p(windows), windows.shape[0], _p(pred),
                                   _stream()), "pool")
        return pred

    def loss_grad(self, out, windows, label, weight):
        """weighted MSE of the pooled proposals; returns (loss tensor, dout)"""
        torch = self.torch
        L, st = lib(), _stream()
        nf = windows.shape[0]
        pred = self.pool(out, windows)
        dpred = torch.empty_like(pred)
        loss = torch.zeros(1, dtype=torch.float32, device=out.device)
        inv = 1.0 / max(1.0, float(nf))
        _chk(L.vsgc_field_loss(_p(pred), _p(label), _p(weight), nf, inv, _p(dpred), _p(loss), st), "loss")
        dout = torch.zeros_like(out)
        _chk(L.vsgc_field_pool_bwd(_p(dpred), out.shape[1], out.shape[2], _p(windows), nf, _p(dout), st),
             "pool bwd")
        return loss, pred, dout
