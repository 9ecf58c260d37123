"""Trains the proposal network of the learned detect form on the GPU.

Data: seeded synthetic clips of the 4K motion-blur workload (config 5 of the
generator; clips 100 .. 100+N-1, so the evaluation clip 0 is never seen),
one consecutive frame pair per clip at a seeded random time.  Labels: the
coarse LocalMotions that the exact full search (the plain GPU detect, which
is bit-identical to the reference) returns for the second frame of the pair:
per selected coarse field its centre and shift.  The target of a field is
its shift / 4 (quarter-resolution pixels), compared with the mean of the
network's two output channels over the field's window (field size / 4).

Every arithmetic step of training -- downsampling, forward convolutions,
pooling, loss, data and weight gradients, Adam -- runs in the project's own
HIP kernels (vsg_cnn.h); torch only allocates device memory.

    python -m vidstab_amd.train_proposal --clips 2000 --epochs 10 --out proposal.bin
"""
import argparse
import ctypes
import sys
import time

import numpy as np


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_dataset(nclips, first_clip, seed):
    import torch
    from . import api
    from .synth import Clip
    spec = Clip(5, first_clip).spec
    W, H = spec.width, spec.height
    kw = dict(shakiness=spec.shakiness, accuracy=spec.accuracy, stepsize=spec.stepsize,
              mincontrast=spec.mincontrast)
    det = api.Detector(W, H, **kw)
    S = det.md.fieldscoarse.fieldSize
    frames = torch.empty((nclips, 2, H, W), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(seed)
    labels = []
    host = np.empty((2, H, W), dtype=np.uint8)
    t0 = time.time()
    for i in range(nclips):
        clip = Clip(5, first_clip + i)
        t = int(rng.integers(0, clip.spec.nframes - 1))
        clip.render(t, out=host[0])
        clip.render(t + 1, out=host[1])
        frames[i].copy_(torch.from_numpy(host))
        base = frames[i].data_ptr()
        mots = det.detect_device([base, base + W * H], [W, W], keep=True)[1]
        rows = [(m[2], m[3], m[0], m[1], float(np.array([m[6]], dtype=np.uint64).view(np.float64)[0]))
                for m in mots if m[4] == S]
        labels.append(np.array(rows, dtype=np.float64).reshape(-1, 5))
        if (i + 1) % 100 == 0:
            log(f"data: {i + 1}/{nclips} pairs, {time.time() - t0:.0f} s")
    det.close()
    torch.cuda.synchronize()
    return frames, labels, S, W, H


def batch_tensors(idx, labels, S, Wq, Hq):
    import torch
    wins, lab = [], []
    h = S // 8
    for k, s in enumerate(idx):
        for fx, fy, vx, vy, _ in labels[s]:
            cx, cy = int(fx) // 4, int(fy) // 4
            wins.append((k, max(cx - h, 0), max(cy - h, 0), min(cx + h, Wq), min(cy + h, Hq)))
            lab.append((vx / 4.0, vy / 4.0))
    wins = torch.tensor(np.array(wins, dtype=np.int32).reshape(-1, 5), device="cuda")
    lab_np = np.array(lab, dtype=np.float32).reshape(-1, 2)
    return wins, torch.tensor(lab_np, device="cuda"), lab_np


def fill_inputs(L, xin, frames, idx, W, H, st):
    from .learned import _p, _chk
    Wq, Hq = W // 4, H // 4
    stride = Hq * Wq * 32 * 2
    for k, s in enumerate(idx):
        out = ctypes.c_void_p(xin.data_ptr() + k * stride)
        for ch in range(2):
            _chk(L.vsgc_downsample(_p(frames[s, ch]), W, W, H, out, ch, st), "downsample")


def evaluate(net, frames, labels, idx_all, S, W, H, batch):
    import torch
    from . import learned
    L, st = learned.lib(), learned._stream()
    Wq, Hq = W // 4, H // 4
    xin = torch.zeros((batch, Hq, Wq, 32), dtype=torch.int16, device="cuda")
    err = []
    for b0 in range(0, len(idx_all), batch):
        idx = idx_all[b0:b0 + batch]
        fill_inputs(L, xin, frames, idx, W, H, st)
        wins, _, lab_np = batch_tensors(idx, labels, S, Wq, Hq)
        _, out = net.forward(xin[:len(idx)])
        pred = net.pool(out, wins).cpu().numpy()
        err.append(np.abs(np.rint(pred * 4.0) - lab_np * 4.0).max(axis=1))
    e = np.concatenate(err) if err else np.zeros(0)
    return {"fields": int(e.size), "within1px": float(np.mean(e <= 1.0)) if e.size else 0.0,
            "exact": float(np.mean(e == 0.0)) if e.size else 0.0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--clips", type=int, default=2000)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--first-clip", type=int, default=100)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    import torch
    from . import learned
    frames, labels, S, W, H = build_dataset(args.clips, args.first_clip, args.seed)
    matches = np.concatenate([l[:, 4] for l in labels if len(l)])
    # the window search's acceptance bound: twice the exact search's p95
    # match.  The bound only has to catch a window whose interior minimum is
    # not the field's true match (those sit far above the typical residual);
    # at the p95 itself, one field in twenty went to the exact search by
    # construction, although its window minimum was right.
    p95 = float(np.percentile(matches, 95))
    threshold = 2.0 * p95
    log(f"data: {sum(len(l) for l in labels)} labelled coarse fields, field size {S}, "
        f"match p95 {p95:.3f}, acceptance bound {threshold:.3f}")
    n = len(labels)
    val_idx = list(range(n - args.val, n))
    tr_idx = np.arange(n - args.val)
    net = learned.ProposalNet(seed=args.seed)
    L, st = learned.lib(), learned._stream()
    Wq, Hq = W // 4, H // 4
    xin = torch.zeros((args.batch, Hq, Wq, 32), dtype=torch.int16, device="cuda")
    rng = np.random.default_rng(args.seed + 7)
    steps_per_epoch = len(tr_idx) // args.batch
    total = steps_per_epoch * args.epochs
    step = 0
    t0 = time.time()
    for ep in range(args.epochs):
        rng.shuffle(tr_idx)
        run = []
        for b in range(steps_per_epoch):
            idx = [int(s) for s in tr_idx[b * args.batch:(b + 1) * args.batch]]
            fill_inputs(L, xin, frames, idx, W, H, st)
            wins, lab, _ = batch_tensors(idx, labels, S, Wq, Hq)
            if wins.shape[0] == 0:
                continue
            wgt = torch.ones(wins.shape[0], dtype=torch.float32, device="cuda")
            acts, out = net.forward(xin)
            net.zero_grad()
            loss, _, dout = net.loss_grad(out, wins, lab, wgt)
            net.backward(acts, dout)
            lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * step / max(1, total)))
            net.adam(float(lr))
            step += 1
            run.append(float(loss.item()))
            if step % 25 == 0:
                log(f"epoch {ep} step {step}/{total} loss {np.mean(run[-25:]):.4f} lr {lr:.2e} "
                    f"{time.time() - t0:.0f} s")
        ev = evaluate(net, frames, labels, val_idx, S, W, H, args.batch)
        log(f"epoch {ep}: train loss {np.mean(run):.4f}; held-out {ev['fields']} fields: "
            f"{100 * ev['within1px']:.2f}% within 1 px, {100 * ev['exact']:.2f}% exact")
    learned.save_weights(args.out, net.layers_numpy(), threshold)
    log(f"saved {args.out}")
    print(f'{{"fields": {ev["fields"]}, "within1px": {ev["within1px"]:.4f}, "exact": {ev["exact"]:.4f}, '
          f'"threshold": {threshold:.4f}, "train_seconds": {time.time() - t0:.1f}}}', flush=True)


if __name__ == "__main__":
    main()
