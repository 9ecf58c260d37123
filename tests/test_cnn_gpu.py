"""The proposal network's gfx950 kernels against the CPU oracle
(oracle/cnn_oracle.py): MFMA convolution forward in all three output modes,
the data gradient through the transposed packing, the weight and bias
gradients, and one full forward + loss + backward of the 6-layer network.

Tolerances: inputs and weights are BF16 on both sides, so every product is
exact; the device sums in FP32 (relative error well below 1e-4 for these
sizes) and rounds stored activations to BF16 (relative 2^-8).  Outputs
stored in BF16 are compared within 2^-7 relative plus a small absolute term;
FP32 results within 1e-3 of the largest magnitude.  The network gradients
pass through five BF16 roundings of activations and of dz on the device;
the oracle applies the same roundings, and the check is 3e-2 of the largest
gradient of each layer."""
import numpy as np
import pytest

import cnn_oracle as co

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev_bf16(a):
    from vidstab_amd import learned
    torch = _torch()
    return torch.from_numpy(learned.bf16_bits(a)).to("cuda")


def _host_f(t):
    from vidstab_amd import learned
    return learned.bits_to_float(t.cpu().numpy()).astype(np.float64)


def _rand_map(rng, n, h, w, scale=1.0):
    return co.bf16_round(rng.standard_normal((n, h, w, 32)).astype(np.float32) * scale)


@pytest.mark.parametrize("dil", [1, 2, 5])
def test_conv_forward_modes(dil):
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    rng = np.random.default_rng(10 + dil)
    n, h, w = 2, 13, 45
    x = _rand_map(rng, n, h, w)
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.1)
    b = rng.standard_normal(32).astype(np.float32) * 0.1
    ref = co.conv(x, wt, b.astype(np.float64), dil)
    dx = _dev_bf16(x)
    dw = torch.from_numpy(wt.astype(np.float32)).cuda()
    db = torch.from_numpy(b).cuda()
    pk = torch.empty(9 * 4 * 64 * 4, dtype=torch.int16, device="cuda")
    st = learned._stream()
    assert L.vsgc_pack_weights(learned._p(dw), learned._p(pk), 0, st) == 0
    outf = torch.empty((n, h, w, 32), dtype=torch.float32, device="cuda")
    assert L.vsgc_conv(learned._p(dx), n, h, w, learned._p(pk), learned._p(db), dil, 2, None,
                       learned._p(outf), st) == 0
    got = outf.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - ref)) <= 1e-3 * np.max(np.abs(ref))
    # the K = 16 instruction path (learned detect), fp32 and bf16 outputs
    pk16 = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights16(learned._p(dw), learned._p(pk16), st) == 0
    out16 = torch.empty((n, h, w, 32), dtype=torch.float32, device="cuda")
    assert L.vsgc_conv16(learned._p(dx), n, h, w, learned._p(pk16), learned._p(db), dil, 2, 32, None,
                         learned._p(out16), st) == 0
    assert np.max(np.abs(out16.cpu().numpy().astype(np.float64) - ref)) <= 1e-3 * np.max(np.abs(ref))
    for mode, fn in ((0, lambda z: np.maximum(z, 0.0)), (1, lambda z: z)):
        outb = torch.empty((n, h, w, 32), dtype=torch.int16, device="cuda")
        assert L.vsgc_conv(learned._p(dx), n, h, w, learned._p(pk), learned._p(db), dil, mode,
                           learned._p(outb), None, st) == 0
        got = _host_f(outb)
        want = fn(ref)
        assert np.all(np.abs(got - want) <= 2.0 ** -7 * np.abs(want) + 1e-3), f"mode {mode}"


@pytest.mark.parametrize("dil", [1, 3])
def test_conv_gradients(dil):
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    rng = np.random.default_rng(20 + dil)
    n, h, w = 2, 11, 37
    x = _rand_map(rng, n, h, w)
    dz = _rand_map(rng, n, h, w, 0.5)
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.1)
    st = learned._stream()
    # data gradient: convolution of dz with the flipped, transposed packing
    dw = torch.from_numpy(wt.astype(np.float32)).cuda()
    pkT = torch.empty(9 * 4 * 64 * 4, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights(learned._p(dw), learned._p(pkT), 1, st) == 0
    ddz = _dev_bf16(dz)
    dxo = torch.empty((n, h, w, 32), dtype=torch.int16, device="cuda")
    assert L.vsgc_conv(learned._p(ddz), n, h, w, learned._p(pkT), None, dil, 1, learned._p(dxo), None, st) == 0
    want = co.conv_data_grad(dz, wt, dil)
    got = _host_f(dxo)
    assert np.all(np.abs(got - want) <= 2.0 ** -7 * np.abs(want) + 1e-3)
    # weight and bias gradients
    gW = torch.zeros((9, 32, 32), dtype=torch.float32, device="cuda")
    gB = torch.zeros(32, dtype=torch.float32, device="cuda")
    assert L.vsgc_conv_wgrad(learned._p(_dev_bf16(x)), learned._p(ddz), n, h, w, dil, learned._p(gW),
                             learned._p(gB), st) == 0
    wW, wB = co.conv_weight_grad(x, dz, dil)
    assert np.max(np.abs(gW.cpu().numpy() - wW)) <= 1e-3 * np.max(np.abs(wW))
    assert np.max(np.abs(gB.cpu().numpy() - wB)) <= 1e-3 * np.max(np.abs(wB))


def test_narrow_input_paths_match_full_width():
    """the first layer's 2-channel input: the 8- and 16-channel K paths equal
    the full 32-channel computation"""
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    rng = np.random.default_rng(31)
    n, h, w, dil = 2, 9, 70, 1
    x = np.zeros((n, h, w, 32))
    x[..., :2] = co.bf16_round(rng.standard_normal((n, h, w, 2)).astype(np.float32))
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.2)
    ref = co.conv(x, wt, None, dil)
    st = learned._stream()
    dx = _dev_bf16(x)
    dw = torch.from_numpy(wt.astype(np.float32)).cuda()
    pk = torch.empty(9 * 4 * 64 * 4, dtype=torch.int16, device="cuda")
    pk16 = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights(learned._p(dw), learned._p(pk), 0, st) == 0
    assert L.vsgc_pack_weights16(learned._p(dw), learned._p(pk16), st) == 0
    for fn, packed, args in ((L.vsgc_conv_k, pk, 2), (L.vsgc_conv16, pk16, 2)):
        out = torch.empty((n, h, w, 32), dtype=torch.float32, device="cuda")
        assert fn(learned._p(dx), n, h, w, learned._p(packed), None, dil, 2, args, None, learned._p(out), st) == 0
        assert np.max(np.abs(out.cpu().numpy().astype(np.float64) - ref)) <= 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("dil,cin,mode", [(1, 32, 0), (2, 32, 0), (4, 32, 1), (8, 32, 2), (1, 2, 0), (3, 32, 0),
                                          (5, 16, 1), (7, 32, 0)])
def test_tiled_conv_equals_direct_kernel(dil, cin, mode):
    """the LDS-tiled kernel (4 rows spaced by the dilation per wave, several
    column bands and row tiles, image edges inside tiles) accumulates the taps
    in the direct kernel's order: outputs are bit-identical"""
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    rng = np.random.default_rng(40 + dil + cin)
    n, h, w = 3, 37, 300
    x = np.zeros((n, h, w, 32))
    x[..., :cin] = co.bf16_round(rng.standard_normal((n, h, w, cin)).astype(np.float32))
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.2)
    b = rng.standard_normal(32).astype(np.float32) * 0.1
    st = learned._stream()
    dx = _dev_bf16(x)
    pk16 = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights16(learned._p(torch.from_numpy(wt.astype(np.float32)).cuda()), learned._p(pk16), st) == 0
    db = torch.from_numpy(b).cuda()
    outs = []
    for fn in (L.vsgc_conv16, L.vsgc_conv16_direct):
        if mode == 2:
            o = torch.full((n, h, w, 32), 7.0, dtype=torch.float32, device="cuda")
            assert fn(learned._p(dx), n, h, w, learned._p(pk16), learned._p(db), dil, 2, cin, None, learned._p(o), st) == 0
        else:
            o = torch.full((n, h, w, 32), 7, dtype=torch.int16, device="cuda")
            assert fn(learned._p(dx), n, h, w, learned._p(pk16), learned._p(db), dil, mode, cin, learned._p(o), None,
                      st) == 0
        outs.append(o.cpu().numpy())
    assert outs[0].view(np.uint8).tobytes() == outs[1].view(np.uint8).tobytes()
    # and against the oracle
    ref = co.conv(x, wt, b.astype(np.float64), dil)
    if mode == 2:
        assert np.max(np.abs(outs[0].astype(np.float64) - ref)) <= 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("dil,mode", [(1, 0), (1, 3), (2, 1), (8, 2)])
def test_pair_input_layer_equals_conv_on_the_built_map(dil, mode):
    """vsgc_conv16_pair builds channel 0 (a previous plane, or the reference
    plane for a negative index) and channel 1 (the sample's own plane) in the
    staging step; its outputs equal vsgc_conv16 with inChannels = 2 on the
    NHWC-32 map holding those planes, bit for bit, in every output mode"""
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    P = learned._p
    rng = np.random.default_rng(70 + dil + mode)
    n, h, w, nplanes = 4, 29, 261, 7
    planes = co.bf16_round(rng.standard_normal((nplanes, h, w)).astype(np.float32))
    ref_plane = co.bf16_round(rng.standard_normal((h, w)).astype(np.float32))
    prev = np.array([-1, 2, 6, 0], dtype=np.int32)
    cur0 = 3
    x = np.zeros((n, h, w, 32), np.float32)
    for z in range(n):
        x[z, :, :, 0] = ref_plane if prev[z] < 0 else planes[prev[z]]
        x[z, :, :, 1] = planes[cur0 + z]
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.2)
    b = rng.standard_normal(32).astype(np.float32) * 0.1
    st = learned._stream()
    pk16 = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights16(P(torch.from_numpy(wt.astype(np.float32)).cuda()), P(pk16), st) == 0
    db = torch.from_numpy(b).cuda()
    dplanes = torch.from_numpy(learned.bf16_bits(planes)).cuda()
    dref = torch.from_numpy(learned.bf16_bits(ref_plane)).cuda()
    dprev = torch.from_numpy(prev).cuda()
    dx = _dev_bf16(x)
    oc = {0: 32, 1: 32, 2: 32, 3: 2}[mode]
    outs = []
    for pair in (True, False):
        if mode >= 2:
            o = torch.full((n, h, w, oc), 7.0, dtype=torch.float32, device="cuda")
            ob, of = None, P(o)
        else:
            o = torch.full((n, h, w, oc), 7, dtype=torch.int16, device="cuda")
            ob, of = P(o), None
        if pair:
            rc = L.vsgc_conv16_pair(P(dplanes), h * w, P(dref), P(dprev), cur0, n, h, w, P(pk16), P(db), dil, mode,
                                    ob, of, st)
        else:
            rc = L.vsgc_conv16(P(dx), n, h, w, P(pk16), P(db), dil, mode, 2, ob, of, st)
        assert rc == 0
        outs.append(o.cpu().numpy())
    assert outs[0].view(np.uint8).tobytes() == outs[1].view(np.uint8).tobytes()
    ref = co.conv(x.astype(np.float64), wt, b.astype(np.float64), dil)
    if mode >= 2:
        assert np.max(np.abs(outs[0].astype(np.float64) - ref[..., :oc])) <= 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("dil0,dil1,mode", [(1, 2, 0), (1, 1, 0), (2, 8, 1), (8, 4, 0), (3, 5, 1)])
def test_fused_first_two_layers_equal_two_launches(dil0, dil1, mode):
    """vsgc_conv16_pair2 runs layer 0 on the plane pair inside layer 1's band
    walker; its output equals vsgc_conv16_pair (BF16 + ReLU) followed by
    vsgc_conv16 on that map, bit for bit, on a frame whose width is not a
    multiple of the 64-px band (right-edge padding) and with a reference
    plane for the first sample"""
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    P = learned._p
    rng = np.random.default_rng(90 + 10 * dil0 + dil1 + mode)
    n, h, w, nplanes = 3, 37, 203, 6
    planes = co.bf16_round(rng.standard_normal((nplanes, h, w)).astype(np.float32))
    ref_plane = co.bf16_round(rng.standard_normal((h, w)).astype(np.float32))
    prev = np.array([-1, 1, 4], dtype=np.int32)
    cur0 = 2
    st = learned._stream()
    pks, dbs = [], []
    for k in range(2):
        wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.2)
        b = rng.standard_normal(32).astype(np.float32) * 0.1
        pk = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
        assert L.vsgc_pack_weights16(P(torch.from_numpy(wt).cuda()), P(pk), st) == 0
        pks.append(pk)
        dbs.append(torch.from_numpy(b).cuda())
    dplanes = torch.from_numpy(learned.bf16_bits(planes)).cuda()
    dref = torch.from_numpy(learned.bf16_bits(ref_plane)).cuda()
    dprev = torch.from_numpy(prev).cuda()

    def out_buf():
        o = torch.full((n, h, w, 32), 7, dtype=torch.int16, device="cuda")
        return o, P(o), None

    o0, ob, of = out_buf()
    assert L.vsgc_conv16_pair2(P(dplanes), h * w, P(dref), P(dprev), cur0, n, h, w, P(pks[0]), P(dbs[0]), dil0,
                               P(pks[1]), P(dbs[1]), dil1, mode, ob, of, st) == 0
    mid = torch.full((n, h, w, 32), 7, dtype=torch.int16, device="cuda")
    assert L.vsgc_conv16_pair(P(dplanes), h * w, P(dref), P(dprev), cur0, n, h, w, P(pks[0]), P(dbs[0]), dil0, 0,
                              P(mid), None, st) == 0
    o1, ob, of = out_buf()
    assert L.vsgc_conv16(P(mid), n, h, w, P(pks[1]), P(dbs[1]), dil1, mode, 32, ob, of, st) == 0
    a0, a1 = o0.cpu().numpy(), o1.cpu().numpy()
    assert a0.view(np.uint8).tobytes() == a1.view(np.uint8).tobytes()
    # FP32 outputs are not offered (layer 1 is never the last layer there)
    assert L.vsgc_conv16_pair2(P(dplanes), h * w, P(dref), P(dprev), cur0, n, h, w, P(pks[0]), P(dbs[0]), dil0,
                               P(pks[1]), P(dbs[1]), dil1, 2, None, P(o0), st) != 0


def test_two_channel_output_and_pooling():
    """mode 3 writes channels 0 and 1 of mode 2's FP32 output, and
    vsgc_field_pool2 over it equals vsgc_field_pool over the 32-channel map"""
    from vidstab_amd import learned
    torch = _torch()
    L = learned.lib()
    P = learned._p
    rng = np.random.default_rng(91)
    n, h, w = 2, 40, 150
    x = _rand_map(rng, n, h, w)
    wt = co.bf16_round(rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.1)
    b = rng.standard_normal(32).astype(np.float32) * 0.1
    st = learned._stream()
    pk16 = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
    assert L.vsgc_pack_weights16(P(torch.from_numpy(wt.astype(np.float32)).cuda()), P(pk16), st) == 0
    db = torch.from_numpy(b).cuda()
    dx = _dev_bf16(x)
    o32 = torch.empty((n, h, w, 32), dtype=torch.float32, device="cuda")
    o2 = torch.empty((n, h, w, 2), dtype=torch.float32, device="cuda")
    assert L.vsgc_conv16(P(dx), n, h, w, P(pk16), P(db), 4, 2, 32, None, P(o32), st) == 0
    assert L.vsgc_conv16(P(dx), n, h, w, P(pk16), P(db), 4, 3, 32, None, P(o2), st) == 0
    assert (o32[..., :2].cpu().numpy().view(np.uint32) == o2.cpu().numpy().view(np.uint32)).all()
    win = np.array([[0, 3, 2, 40, 30], [1, 100, 20, 150, 40], [1, -5, -5, 7, 9], [0, 60, 10, 61, 11]], np.int32)
    dwin = torch.from_numpy(win).cuda()
    p32 = torch.empty((4, 2), dtype=torch.float32, device="cuda")
    p2 = torch.empty((4, 2), dtype=torch.float32, device="cuda")
    assert L.vsgc_field_pool(P(o32), h, w, P(dwin), 4, P(p32), st) == 0
    assert L.vsgc_field_pool2(P(o2), h, w, P(dwin), 4, P(p2), st) == 0
    assert (p32.cpu().numpy().view(np.uint32) == p2.cpu().numpy().view(np.uint32)).all()


def test_network_loss_and_gradients_match_oracle():
    """one training step's gradients of the 6-layer network vs the oracle run
    with the device's BF16 rounding points"""
    from vidstab_amd import learned
    torch = _torch()
    rng = np.random.default_rng(7)
    n, h, w = 2, 24, 40
    net = learned.ProposalNet(seed=5)
    layers = net.layers_numpy()
    layers = [(co.bf16_round(wk), bk.astype(np.float64), d) for (wk, bk, d) in layers]
    x = np.zeros((n, h, w, 32))
    x[..., :2] = co.bf16_round(rng.standard_normal((n, h, w, 2)).astype(np.float32))
    windows = np.array([[0, 2, 3, 14, 15], [1, 20, 1, 39, 20], [0, 10, 10, 30, 23]], dtype=np.int32)
    label = rng.standard_normal((3, 2)) * 0.5
    weight = np.array([1.0, 1.0, 0.5])

    # oracle with the device's rounding points: activations stored in BF16
    acts = [x]
    for k, (wk, bk, d) in enumerate(layers):
        z = co.conv(acts[-1], wk, bk, d)
        acts.append(co.bf16_round(np.maximum(z, 0.0)) if k + 1 < len(layers) else z)
    pred = co.pool(acts[-1], windows)
    loss, dpred = co.loss(pred, label, weight, 1.0 / 3.0)
    dz = co.bf16_round(co.pool_grad(dpred, acts[-1].shape, windows))
    want = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        want[k] = co.conv_weight_grad(acts[k], dz, layers[k][2])
        if k:
            dz = co.bf16_round(co.conv_data_grad(dz, layers[k][0], layers[k][2])) * (acts[k] > 0)

    xt = _dev_bf16(x)
    acts_d, out = net.forward(xt)
    wins = torch.from_numpy(windows).cuda()
    lab = torch.from_numpy(label.astype(np.float32)).cuda()
    wgt = torch.from_numpy(weight.astype(np.float32)).cuda()
    net.zero_grad()
    l_d, pred_d, dout = net.loss_grad(out, wins, lab, wgt)
    net.backward(acts_d, dout)
    torch.cuda.synchronize()
    assert np.allclose(pred_d.cpu().numpy(), pred, rtol=2e-2, atol=2e-3)
    assert abs(float(l_d.item()) - loss) <= 2e-2 * abs(loss) + 1e-5
    for k in range(len(layers)):
        gw = net.gw[k].cpu().numpy()
        gb = net.gb[k].cpu().numpy()
        scale = max(np.max(np.abs(want[k][0])), 1e-12)
        assert np.max(np.abs(gw - want[k][0])) <= 3e-2 * scale, f"layer {k} dW"
        bscale = max(np.max(np.abs(want[k][1])), 1e-12)
        assert np.max(np.abs(gb - want[k][1])) <= 3e-2 * bscale, f"layer {k} dB"
