"""Detect-pass benchmark on one MI355X.

A step is one pass of vsMotionDetection over a whole seeded synthetic clip
whose luma planes are already resident in HBM (vsMotionDetectionDevice, the
batched form of the same C API): 1920x1080 YUV420p, 300 frames, random-walk
motion, shakiness 8, accuracy 15, stepsize 6.  value = frames/s.

Also reported: the roofline of the dominant kernel (the coarse SAD grid
search; integer-VALU bound: v_sad_u8 does 4 abs-diffs per lane per cycle, so
the MI355X peak is 256 CU x 64 lanes x 2.4 GHz x 4 = 157.3 T abs-diff/s),
timed with HIP events on the library's own stream over the timed region, and
the reference's CPU path on a bounded sample of the same clip on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(REPO, "vid.stab_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "detect-pass frames/s at 1080p & 4K (plain vs learned form, same shift accuracy)"
SAD_PEAK_T = 256 * 64 * 2.4e9 * 4 / 1e12      # T abs-diff/s
HBM_PEAK = 8000.0                              # GB/s
COARSE_TRAFFIC_1080P = 1709.2e6 + 0.13e6       # bytes read + written per 100-frame coarse launch


class VSGProfile(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_longlong * 7), ("ms", ctypes.c_double * 7),
                ("work", ctypes.c_double * 7)]


KNAMES = ["blur", "fields", "select", "grid_coarse", "refine_coarse", "grid_fine", "refine_fine"]
K_GRID_COARSE = 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(spec, frames, nsample, threads):
    """reference vsMotionDetection (oracle/_ref) on a bounded sample; falls back
    to the project's C restatement when the reference build is absent"""
    import numpy as np
    import pyoracle
    kw = dict(shakiness=spec.shakiness, accuracy=spec.accuracy, stepsize=spec.stepsize,
              mincontrast=spec.mincontrast)
    kind = "reference"
    best = None
    for _ in range(3):
        try:
            det = pyoracle.RefDetector(spec.width, spec.height, threads=threads, **kw)
            run = lambda x: det.detect(x, keep=False)
        except RuntimeError:
            kind = "port"
            det = pyoracle.Detector(spec.width, spec.height, **kw)
            run = det.detect
        run(np.ascontiguousarray(frames[0]))            # warm-up frame fills prev
        t0 = time.perf_counter()
        for t in range(1, nsample + 1):
            run(np.ascontiguousarray(frames[t]))
        dt = time.perf_counter() - t0
        det.close()
        best = dt if best is None else min(best, dt)
    return {"value": round(nsample / best, 3), "unit": "frames/s", "cores": threads, "kind": kind,
            "sample": f"{nsample} frames of the same {spec.width}x{spec.height} clip after one warm-up "
                      f"frame, best of 3, {threads} OpenMP threads"}


def measure_network_layers(weights, n=32, h=540, w=960, reps=5):
    """the proposal network's layers one at a time (HIP events on torch's
    stream): the first layer reading the quarter-resolution planes, the
    32-channel layers, the two-channel FP32 last layer, over n frames"""
    import numpy as np
    import torch
    from vidstab_amd import learned
    L, st, P = learned.lib(), learned._stream(), learned._p
    layers, _ = learned.load_weights(weights)
    planes = torch.randint(-16000, 16000, (n + 1, h, w), dtype=torch.int16, device="cuda") & 0x3FFF
    prev = torch.arange(-1, n - 1, dtype=torch.int32, device="cuda")
    acts = [torch.empty((n, h, w, 32), dtype=torch.int16, device="cuda") for _ in range(2)]
    out2 = torch.empty((n, h, w, 2), dtype=torch.float32, device="cuda")
    packed = []
    for wt, b, d in layers:
        pk = torch.empty(9 * 2 * 64 * 8, dtype=torch.int16, device="cuda")
        if L.vsgc_pack_weights16(P(torch.from_numpy(np.ascontiguousarray(wt, np.float32)).cuda()), P(pk), st) != 0:
            raise RuntimeError("weight packing failed")
        packed.append((pk, torch.from_numpy(np.ascontiguousarray(b, np.float32)).cuda(), d))
    res = []
    for k, (pk, b, d) in enumerate(packed):
        last = k == len(packed) - 1
        mode = 3 if last else 0
        if k == 0:
            run = lambda: L.vsgc_conv16_pair(P(planes), h * w, P(planes), P(prev), 1, n, h, w, P(pk), P(b), d, mode,
                                             None if last else P(acts[0]), P(out2) if last else None, st)
        else:
            src, dst = acts[(k - 1) & 1], acts[k & 1]
            run = lambda: L.vsgc_conv16(P(src), n, h, w, P(pk), P(b), d, mode, 32, None if last else P(dst),
                                        P(out2) if last else None, st)
        if run() != 0:
            raise RuntimeError(f"layer {k} failed")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flop = 2.0 * 9 * 32 * (16 if k == 0 else 32) * n * h * w   # MFMA work issued (K = 16 or 32 per tap)
        res.append({"layer": k, "dilation": d, "ms": round(ms, 4), "tflops": round(flop / ms / 1e9, 1)})
        log(f"network layer {k} (dilation {d}): {ms:.3f} ms per {n} frames of {w}x{h}, {flop / ms / 1e9:.1f} TFLOP/s")
    log(f"network: {sum(r['ms'] for r in res):.3f} ms per {n} frames")
    if len(packed) >= 3:
        # layers 0 and 1 in one pass (layer 0's map stays on chip), as the
        # learned detect path runs them
        (pk0, b0, d0), (pk1, b1, d1) = packed[0], packed[1]
        run = lambda: L.vsgc_conv16_pair2(P(planes), h * w, P(planes), P(prev), 1, n, h, w, P(pk0), P(b0), d0,
                                          P(pk1), P(b1), d1, 0, P(acts[1]), None, st)
        if run() != 0:
            raise RuntimeError("fused layers 0+1 failed")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res.append({"layer": "0+1 fused", "dilation": [d0, d1], "ms": round(ms, 4)})
        net = ms + sum(r["ms"] for r in res[2:-1])
        log(f"network layers 0+1 fused: {ms:.3f} ms per {n} frames; network with the fused pair: {net:.3f} ms")
    return res


def network_roofline(layers, n=32, h=540, w=960):
    """the learned form's dominant kernel: the band walker on the 32-channel
    layers, each reading and writing n*h*w*64 bytes of BF16 activations (the
    algorithmic bytes; rocprofv3 FETCH_SIZE reads 1.06-1.09 GB for 32 frames),
    against the HBM peak"""
    single = [r for r in layers if isinstance(r["layer"], int)]
    mid = single[1:-1]   # neither the first (2-channel input) nor the last (2-channel output) layer
    byts = 2.0 * n * h * w * 64
    ach = [byts / (r["ms"] * 1e-3) / 1e9 for r in mid]
    avg_ms = sum(r["ms"] for r in mid) / len(mid)
    return {"bound": "hbm", "kernel": "k_conv16w (32-channel layers)", "achieved": round(byts / (avg_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK, "unit": "GB/s", "frac": round(byts / (avg_ms * 1e-3) / 1e9 / HBM_PEAK, 4),
            "per_layer_gbs": [round(a, 1) for a in ach], "bytes_per_launch": byts, "avg_launch_ms": round(avg_ms, 4),
            "layers": layers}


def measure_per_frame(nframes):
    """ffmpeg's pattern: one vsMotionDetection call per pageable host frame,
    each returning its frame's LocalMotions (the whole pass per call), on the
    1080p clip after 5 untimed calls"""
    from vidstab_amd import api
    from vidstab_amd.synth import Clip
    clip = Clip(2, 0)
    s = clip.spec
    frames = [clip.render(t) for t in range(nframes)]
    det = api.Detector(s.width, s.height, shakiness=s.shakiness, accuracy=s.accuracy, stepsize=s.stepsize,
                       mincontrast=s.mincontrast)
    for t in range(5):
        det.detect(frames[t])
    t0 = time.perf_counter()
    for t in range(5, nframes):
        det.detect(frames[t])
    dt = time.perf_counter() - t0
    det.close()
    fps = (nframes - 5) / dt
    log(f"per-frame vsMotionDetection, {s.width}x{s.height} host frames: {fps:.1f} frames/s, "
        f"{dt / (nframes - 5) * 1000:.3f} ms per call")
    return {"frames": nframes - 5, "fps": round(fps, 2), "ms_per_call": round(dt / (nframes - 5) * 1000, 3)}


def measure_per_frame_c(spec, host, nframes=120, warmup=10):
    """ffmpeg's pattern through C, no Python in the loop: the ffmpeg-shaped
    driver (tests/c/vidstabdetect_like.c: vsMotionDetection, vsWriteToFile and
    vs_vector_del per pageable host frame) on the first frames of the clip;
    the frames after `warmup` are timed by the driver itself"""
    import subprocess
    import tempfile
    import numpy as np
    exe = os.path.join(REPO, "vid.stab_amd", "lib", "vidstabdetect_like")
    n = min(nframes, host.shape[0])
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "frames.raw")
        host[:n].tofile(raw)
        bits = "%016x" % int(np.array([spec.mincontrast], dtype=np.float64).view(np.uint64)[0])
        cmd = [exe, raw, str(spec.width), str(spec.height), str(n), str(spec.width), str(spec.shakiness),
               str(spec.accuracy), str(spec.stepsize), bits, "0", "2", os.path.join(tmp, "out.trf"), str(warmup)]
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if p.returncode != 0:
        raise RuntimeError(f"vidstabdetect_like failed: {p.stderr[-500:]}")
    line = [l for l in p.stdout.splitlines() if l.startswith("per-frame ")][-1].split()
    res = {"frames": int(line[1]), "ms_per_call": float(line[2]), "fps": float(line[3]),
           "driver": "tests/c/vidstabdetect_like.c: one vsMotionDetection + vsWriteToFile + vs_vector_del "
                     "per pageable host frame (ffmpeg vidstabdetect's sequence), timed in C"}
    log(f"per-frame C API, {spec.width}x{spec.height} host frames: {res['fps']:.1f} frames/s, "
        f"{res['ms_per_call']:.3f} ms per call")
    return res


def measure_plain_4k_step4(steps=2, warmup=1):
    """the 3840x2160 step-4 clip with moving rectangles (config 3, 240
    frames), plain form, device-resident frames: frames/s and the coarse
    search's abs-diff rate from HIP events on its stream"""
    import numpy as np
    import torch
    from vidstab_amd import api
    from vidstab_amd.synth import Clip
    clip = Clip(3, 0)
    s = clip.spec
    n = s.nframes
    host = np.empty((n, s.height, s.width), dtype=np.uint8)
    for t in range(n):
        clip.render(t, out=host[t])
    dev = torch.from_numpy(host).to("cuda:0")
    del host
    torch.cuda.synchronize()
    fb = s.width * s.height
    ptrs = [dev.data_ptr() + i * fb for i in range(n)]
    L = api.lib()
    det = api.Detector(s.width, s.height, shakiness=s.shakiness, accuracy=s.accuracy, stepsize=s.stepsize,
                       mincontrast=s.mincontrast)
    for _ in range(warmup):
        det.detect_device(ptrs, [s.width] * n)
    L.vsg_profile_reset()
    L.vsg_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        det.detect_device(ptrs, [s.width] * n)
    api.release_wait()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    L.vsg_profile_enable(0)
    prof = VSGProfile()
    L.vsg_profile_get(ctypes.byref(prof))
    det.close()
    del dev
    k = K_GRID_COARSE
    avg_s = prof.ms[k] / prof.launches[k] / 1000.0
    achieved = prof.work[k] / prof.launches[k] / avg_s / 1e12
    res = {"workload": "3840x2160 YUV420p, 240-frame seeded clip with 3-6 moving textured rectangles (config 3), "
                       "shakiness 10, accuracy 15, step 4, device-resident frames",
           "frames": n, "fps": round(n / dt, 2), "ms_per_pass": round(dt * 1000.0, 2),
           "coarse_roofline": {"bound": "valu", "kernel": "k_sad_search (coarse)", "achieved": round(achieved, 3),
                               "peak": SAD_PEAK_T, "unit": "T abs-diff/s", "frac": round(achieved / SAD_PEAK_T, 4),
                               "avg_launch_ms": round(avg_s * 1000.0, 3)}}
    log(f"config 3 (4K, step 4) plain: {res['fps']:.1f} frames/s, coarse search {achieved:.1f} T abs-diff/s")
    return res


def measure_second_pass(spec, ptrs, ls):
    """the second pass's per-frame transforms (vsMotionsToTransformBatch, one
    launch for the clip, host LocalMotions in and transforms out) beside the
    oracle's restatement of vsMotionsToTransform on one host core"""
    import numpy as np
    import pyoracle
    from vidstab_amd import api
    det = api.Detector(spec.width, spec.height, shakiness=spec.shakiness, accuracy=spec.accuracy,
                       stepsize=spec.stepsize, mincontrast=spec.mincontrast)
    frames = det.detect_device(ptrs, ls, keep=True)
    det.close()
    # the C structures are built before either clock starts
    clip = api.ClipMotions(frames)
    arrays = [pyoracle.local_motions(m) for m in frames]
    api.motions_to_transforms(spec.width, spec.height, api.ClipMotions(frames[:8]))   # warm-up: module load
    t0 = time.perf_counter()
    got = api.motions_to_transforms(spec.width, spec.height, clip)
    gpu_ms = (time.perf_counter() - t0) * 1000.0
    OL = pyoracle.lib()
    res, rounds = ctypes.c_double(0.0), ctypes.c_int(0)
    t0 = time.perf_counter()
    want = [OL.or_motions_to_transform(spec.width, spec.height, a, len(m), 0, ctypes.byref(res), ctypes.byref(rounds))
            for a, m in zip(arrays, frames)]
    cpu_ms = (time.perf_counter() - t0) * 1000.0
    same = sum(all(np.float64(g[i]).view(np.uint64) == np.float64(v).view(np.uint64)
                   for i, v in enumerate((w.x, w.y, w.alpha, w.zoom))) and g[4] == w.extra
               for g, w in zip(got, want))
    out = {"frames": len(frames), "motions": sum(len(m) for m in frames), "gpu_ms": round(gpu_ms, 3),
           "cpu_port_ms_1core": round(cpu_ms, 3), "frames_bit_identical": same}
    log(f"second pass: {len(frames)} frames on the GPU in {gpu_ms:.2f} ms (host-inclusive), oracle on one core "
        f"{cpu_ms:.1f} ms; {same}/{len(frames)} transforms bit-identical")
    out["default_config"] = measure_second_pass_default(spec, frames, clip)
    return out


def measure_second_pass_default(spec, frames, clip):
    """vsLocalmotions2Transforms with the reference's default transform
    configuration (lens estimate on, fov 0) on the clip's LocalMotions: the
    GPU library against the reference's own build on the host, and the lens
    estimate alone (vsEstimateLensDistortion: Brent's search, one k_lens_frame
    launch over every frame per objective evaluation)"""
    import pyoracle
    from vidstab_amd import api
    api.localmotions_to_transforms(spec.width, spec.height, api.ClipMotions(frames[:8]))   # warm-up
    t0 = time.perf_counter()
    got, k = api.localmotions_to_transforms(spec.width, spec.height, clip)
    gpu_ms = (time.perf_counter() - t0) * 1000.0
    t0 = time.perf_counter()
    est = api.estimate_lens(spec.width, spec.height, clip)
    est_ms = (time.perf_counter() - t0) * 1000.0
    out = {"gpu_ms": round(gpu_ms, 3), "lens_k": k, "estimate_gpu_ms": round(est_ms, 3),
           "estimate_evaluations": est["iterations"], "estimate_determined": est["determined"],
           "estimate_ms_per_evaluation": round(est_ms / max(1, est["iterations"]), 4)}
    if os.path.exists(pyoracle.REF_PATH):
        t0 = time.perf_counter()
        want, rk = pyoracle.ref_localmotions_to_transforms(spec.width, spec.height, clip)
        ref_ms = (time.perf_counter() - t0) * 1000.0
        worst = max((abs(g[i] - w[i]) for g, w in zip(got, want) for i in range(4)), default=0.0)
        same = sum(g == w for g, w in zip(got, want))
        out.update({"reference_ms_1core": round(ref_ms, 3), "reference_lens_k": rk,
                    "frames_identical": same, "largest_difference": worst})
        log(f"second pass, default configuration (lens estimate on): GPU {gpu_ms:.2f} ms, reference on one core "
            f"{ref_ms:.1f} ms; lens k {k:.6g} vs {rk:.6g}; {same}/{len(got)} transforms identical, largest "
            f"difference {worst:.3g}; estimate alone {est_ms:.2f} ms over {est['iterations']} evaluations")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2,
                    help="workload: 2 = 1080p, 3 = 4K, 4 = 64 clips of uneven size (batched multi-clip "
                         "form), 5 = 4K motion blur, plain vs learned-proposal form")
    ap.add_argument("--weights", default=os.path.join(REPO, "vid.stab_amd", "learned", "proposal_c5.bin"))
    ap.add_argument("--frames", type=int, default=0, help="clip length (0: the workload's)")
    ap.add_argument("--cpu-frames", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-4k", action="store_true", help="skip the 4K plain vs learned-proposal measurement")
    ap.add_argument("--no-host", action="store_true", help="skip the host-frame (PCIe-inclusive) rate")
    ap.add_argument("--no-second-pass", action="store_true", help="skip the second-pass transform measurement")
    ap.add_argument("--clip", type=int, default=0, help="--config 5: the held-out clip (0..99)")
    ap.add_argument("--net-layers", action="store_true",
                    help="only time the proposal network's layers one at a time (--frames of 960x540, default 32)")
    ap.add_argument("--per-frame", action="store_true",
                    help="only time one vsMotionDetection call per host frame (ffmpeg's pattern) at 1080p")
    ap.add_argument("--per-frame-c", action="store_true",
                    help="only run the ffmpeg-shaped C driver (one vsMotionDetection call per host frame, "
                         "no Python in the loop) on the 1080p clip")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one 256 MiB device copy (known bytes) to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the detect pass shards by frame range or clip with no "
              "exchange between GPUs, and only the one-GPU figure is reported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from vidstab_amd import api
    from vidstab_amd.synth import Clip

    if args.per_frame:
        print(json.dumps({"per_frame_api": measure_per_frame(args.frames or 120)}), flush=True)
        return
    if args.per_frame_c:
        clip = Clip(2, 0)
        n = args.frames or 130
        host = np.empty((n, clip.spec.height, clip.spec.width), dtype=np.uint8)
        for t in range(n):
            clip.render(t, out=host[t])
        print(json.dumps({"per_frame_c_api": measure_per_frame_c(clip.spec, host, nframes=n)}), flush=True)
        return
    if args.net_layers:
        n = args.frames or 32
        print(json.dumps({"network_layers": measure_network_layers(args.weights, n=n)}), flush=True)
        return
    if args.config == 4:
        return run_multi(args)
    if args.config == 5:
        return run_learned(args)
    clip = Clip(args.config, 0)
    spec = clip.spec
    nfr = args.frames or spec.nframes
    log(f"generating {nfr} frames {spec.width}x{spec.height}")
    host = np.empty((nfr, spec.height, spec.width), dtype=np.uint8)
    for t in range(nfr):
        clip.render(t, out=host[t])
    dev = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    fb = spec.width * spec.height
    ptrs = [dev.data_ptr() + i * fb for i in range(nfr)]
    ls = [spec.width] * nfr

    L = api.lib()
    L.vsg_profile_get.argtypes = [ctypes.POINTER(VSGProfile)]
    det = api.Detector(spec.width, spec.height, shakiness=spec.shakiness, accuracy=spec.accuracy,
                       stepsize=spec.stepsize, mincontrast=spec.mincontrast)
    for w in range(args.warmup):
        n = det.detect_device(ptrs, ls)
        log(f"warmup {w}: {sum(n)} LocalMotions")
    L.vsg_profile_reset()
    L.vsg_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        det.detect_device(ptrs, ls)
    api.release_wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.vsg_profile_enable(0)
    prof = VSGProfile()
    L.vsg_profile_get(ctypes.byref(prof))
    det.close()
    if args.calib_copy:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0").fill_(7)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")
    ms_step = (t1 - t0) * 1000.0 / args.steps
    value = nfr * args.steps / (t1 - t0)
    for k, name in enumerate(KNAMES):
        if prof.launches[k]:
            log(f"kernel {name}: {prof.launches[k]} launches, avg {prof.ms[k] / prof.launches[k]:.4f} ms, "
                f"total {prof.ms[k] / args.steps:.3f} ms/step")
    L.vsg_sad_peak.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    peak_meas = ctypes.c_double(0.0)
    if L.vsg_sad_peak(0, ctypes.byref(peak_meas), None) != 0:
        peak_meas.value = 0.0
    log(f"measured v_sad_u8 ceiling: {peak_meas.value / 1e12:.1f} T abs-diff/s")
    k = K_GRID_COARSE
    avg_s = prof.ms[k] / prof.launches[k] / 1000.0
    work = prof.work[k] / prof.launches[k]
    achieved = work / avg_s / 1e12
    roofline = {"bound": "valu", "kernel": "k_sad_search (coarse)", "achieved": round(achieved, 3),
                "peak": round(SAD_PEAK_T, 1), "unit": "T abs-diff/s", "frac": round(achieved / SAD_PEAK_T, 4),
                # HBM bytes per 64-frame launch of this workload: rocprofv3 --pmc
                # FETCH_SIZE (x 2.00) and WRITE_SIZE (x 1.00), each scaled by what the
                # counter reads for a 256 MiB device copy (profiles/k_sad_search_1080p.md)
                "traffic": COARSE_TRAFFIC_1080P if spec.width == 1920 and spec.height == 1080 else None,
                "avg_launch_ms": round(avg_s * 1000.0, 4),
                "peak_measured": round(peak_meas.value / 1e12, 1),
                "frac_of_measured": round(achieved * 1e12 / peak_meas.value, 4) if peak_meas.value else None,
                "work_per_launch": work}
    bk = 0
    # in the pipeline the blur shares the CUs with other batches' searches;
    # the same 64-plane launch alone on the idle device beside it
    blur_gbs = prof.work[bk] / (prof.ms[bk] / 1000.0) / 1e9 if prof.ms[bk] else 0.0
    L.vsg_blur_time.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int),
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
    nb = min(64, nfr)
    blur_ms = ctypes.c_double(0.0)
    if L.vsg_blur_time(0, (ctypes.c_void_p * nb)(*ptrs[:nb]), (ctypes.c_int * nb)(*ls[:nb]), nb, spec.width,
                       spec.height, spec.stepsize, 20, ctypes.byref(blur_ms)) != 0:
        blur_ms.value = 0.0
    blur_alone_gbs = 2.0 * fb * nb / (blur_ms.value / 1000.0) / 1e9 if blur_ms.value else 0.0
    log(f"blur alone: {nb} planes in {blur_ms.value * 1000.0:.1f} us, {blur_alone_gbs:.1f} GB/s; "
        f"in the pipeline {blur_gbs:.1f} GB/s")

    host_fps = None
    if not args.no_host:
        # the same clip handed over as host frames (vsMotionDetectionBatch):
        # every plane crosses PCIe inside the timed pass
        det = api.Detector(spec.width, spec.height, shakiness=spec.shakiness, accuracy=spec.accuracy,
                           stepsize=spec.stepsize, mincontrast=spec.mincontrast)
        frames = [host[t] for t in range(nfr)]
        # one untimed pass: every slot set gets its pinned staging area
        det.detect_batch(frames, keep=False)
        t0 = time.perf_counter()
        det.detect_batch(frames, keep=False)
        api.release_wait()
        host_fps = nfr / (time.perf_counter() - t0)
        det.close()
        log(f"host frames (PCIe-inclusive): {host_fps:.1f} frames/s")
    per_frame = None if args.no_host else measure_per_frame_c(spec, host)
    second = None if args.no_second_pass else measure_second_pass(spec, ptrs, ls)
    del dev
    forms = plain4k = None
    if not args.no_4k:
        plain4k = measure_plain_4k_step4()
        forms = measure_forms_4k(0, 3, 1, args.weights)   # 3 timed passes of each form
        forms["network_roofline"] = network_roofline(measure_network_layers(args.weights))
    cpu = cpu1 = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        ns = min(args.cpu_frames, nfr - 1)
        log(f"cpu baseline: {ns} frames, {threads} threads")
        cpu = cpu_baseline(spec, host, ns, threads)
        # the reference on one host thread beside it (a shorter sample)
        cpu1 = cpu_baseline(spec, host, min(ns, 12), 1)
        log(f"cpu baseline: {cpu['value']} frames/s on {threads} threads, {cpu1['value']} on 1")
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"plain GPU detect, {spec.width}x{spec.height} YUV420p 8-bit, {nfr}-frame "
                               f"seeded value-noise clip (config {args.config}), device-resident frames",
                   "width": spec.width, "height": spec.height, "frames": nfr,
                   "shakiness": spec.shakiness, "accuracy": spec.accuracy, "stepsize": spec.stepsize,
                   "mincontrast": spec.mincontrast, "parallelism": "1 GPU"},
        "roofline": roofline,
        # plane bytes read + written (2 W H a plane) per second of the 64-plane
        # launch alone; in_pipeline: the same over its event interval inside
        # the timed passes, stretched by the concurrent searches
        "blur_roofline": {"bound": "hbm", "kernel": "k_blur_walk", "achieved": round(blur_alone_gbs, 1),
                          "peak": HBM_PEAK, "unit": "GB/s", "frac": round(blur_alone_gbs / HBM_PEAK, 4),
                          "avg_launch_ms": round(blur_ms.value, 4), "planes_per_launch": nb,
                          "in_pipeline_gbs": round(blur_gbs, 1)},
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "host_frames_fps": round(host_fps, 2) if host_fps else None,
        # the luma bytes those frames carry over PCIe per second (W*H per frame)
        "host_frames_h2d_gbs": round(host_fps * spec.width * spec.height / 1e9, 2) if host_fps else None,
        # the drop-in path: one vsMotionDetection call per host frame, from C
        "per_frame_fps": per_frame["fps"] if per_frame else None,
        "per_frame_c_api": per_frame,
        "second_pass": second,
        "forms_4k": forms,
        "plain_4k_step4": plain4k,
    }
    print(json.dumps(res), flush=True)


def run_multi(args):
    """64 clips of uneven resolution and length through vsMotionDetectMulti
    (persistent queue launches over the units of every clip, HIP graphs per
    slot set), against the same clips run one after another through the
    single-clip entry point."""
    import numpy as np
    import torch
    from vidstab_amd import api
    from vidstab_amd.synth import Clip

    nclips = 64
    clips, keep, frames = [], [], 0
    for c in range(nclips):
        clip = Clip(4, c)
        s = clip.spec
        n = args.frames or s.nframes
        host = np.empty((n, s.height, s.width), dtype=np.uint8)
        for t in range(n):
            clip.render(t, out=host[t])
        dev = torch.from_numpy(host).to("cuda:0")
        keep.append(dev)
        fb = s.width * s.height
        clips.append(dict(width=s.width, height=s.height, shakiness=s.shakiness, accuracy=s.accuracy,
                          stepsize=s.stepsize, mincontrast=s.mincontrast, linesize=s.width,
                          ptrs=[dev.data_ptr() + t * fb for t in range(n)]))
        frames += n
        del host
    torch.cuda.synchronize()
    log(f"config 4: {nclips} clips, {frames} frames")
    for w in range(args.warmup):
        api.detect_multi(clips, keep=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        api.detect_multi(clips, keep=False)
    api.release_wait()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    multi_fps = frames * args.steps / (t1 - t0)
    log(f"multi-clip: {multi_fps:.1f} frames/s")
    # the same clips one after another through the single-clip batched entry
    # point; each detector is set up and warmed up (its first call) outside
    # the clock, as the multi-clip form's state is after its warm-up call
    seq_s, t2 = 0.0, time.perf_counter()
    for cl in clips:
        det = api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                           stepsize=cl["stepsize"], mincontrast=cl["mincontrast"])
        det.detect_device(cl["ptrs"][:2], [cl["linesize"]] * 2)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        det.detect_device(cl["ptrs"], [cl["linesize"]] * len(cl["ptrs"]))
        api.release_wait()
        torch.cuda.synchronize()
        seq_s += time.perf_counter() - tc
        det.close()
    setup_fps = frames / (time.perf_counter() - t2)
    seq_fps = frames / seq_s
    log(f"per-clip sequential: {seq_fps:.1f} frames/s (set-up excluded; {setup_fps:.1f} frames/s with it)")
    del keep, clips
    small = measure_small_clips(args.steps, args.warmup)
    res = {
        "metric": METRIC, "value": round(multi_fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round((t1 - t0) * 1000.0 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"batched multi-clip detect: {nclips} seeded clips (config 4), resolutions "
                               "360p-2160p, 30-120 frames, device-resident frames", "clips": nclips,
                   "frames": frames, "parallelism": "1 GPU"},
        "per_clip_sequential_fps": round(seq_fps, 2),
        "per_clip_sequential_with_setup_fps": round(setup_fps, 2),
        "small_clips": small,
    }
    print(json.dumps(res), flush=True)


def measure_small_clips(steps, warmup, nclips=64, width=640, height=360, nframes=30):
    """where the multi-clip form pays off: many short low-resolution clips,
    each too small to fill the GPU alone.  The same clips through
    vsMotionDetectMulti and through detectors created beforehand, one clip
    after another (set-up outside both clocks)"""
    import numpy as np
    import torch
    from vidstab_amd import api
    from vidstab_amd.synth import Clip

    clips, keep = [], []
    for c in range(nclips):
        clip = Clip(4, c, width=width, height=height, nframes=nframes)
        s = clip.spec
        host = np.empty((nframes, height, width), dtype=np.uint8)
        for t in range(nframes):
            clip.render(t, out=host[t])
        dev = torch.from_numpy(host).to("cuda:0")
        keep.append(dev)
        clips.append(dict(width=width, height=height, shakiness=s.shakiness, accuracy=s.accuracy,
                          stepsize=s.stepsize, mincontrast=s.mincontrast, linesize=width,
                          ptrs=[dev.data_ptr() + t * width * height for t in range(nframes)]))
    torch.cuda.synchronize()
    frames = nclips * nframes
    for _ in range(warmup):
        api.detect_multi(clips, keep=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        api.detect_multi(clips, keep=False)
    api.release_wait()
    torch.cuda.synchronize()
    multi_fps = frames * steps / (time.perf_counter() - t0)
    dets = [api.Detector(cl["width"], cl["height"], shakiness=cl["shakiness"], accuracy=cl["accuracy"],
                         stepsize=cl["stepsize"], mincontrast=cl["mincontrast"]) for cl in clips]
    for det, cl in zip(dets, clips):   # warm-up: every detector's first call
        det.detect_device(cl["ptrs"][:2], [cl["linesize"]] * 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for det, cl in zip(dets, clips):
        det.detect_device(cl["ptrs"], [cl["linesize"]] * nframes)
    api.release_wait()
    torch.cuda.synchronize()
    seq_fps = frames / (time.perf_counter() - t0)
    for det in dets:
        det.close()
    log(f"{nclips} clips of {width}x{height} x {nframes} frames: multi-clip {multi_fps:.1f} frames/s, "
        f"one clip after another {seq_fps:.1f} frames/s ({multi_fps / seq_fps:.2f}x)")
    return {"workload": f"{nclips} clips of {width}x{height}, {nframes} frames each (config 4 generator)",
            "frames": frames, "multi_fps": round(multi_fps, 2), "sequential_fps": round(seq_fps, 2),
            "speedup": round(multi_fps / seq_fps, 3)}


def measure_forms_4k(nframes, steps, warmup, weights, clip_index=0):
    """4K motion-blur clip (config 5, clip 0 by default; the trainer uses
    clips 100 and up): the plain form and the learned-proposal form on the
    same device-resident frames; agreement with the full search per field
    and per frame transform"""
    import numpy as np
    import torch
    from vidstab_amd import api
    from vidstab_amd.synth import Clip

    clip = Clip(5, clip_index)
    s = clip.spec
    nfr = nframes or s.nframes
    log(f"generating {nfr} frames {s.width}x{s.height}")
    host = np.empty((nfr, s.height, s.width), dtype=np.uint8)
    for t in range(nfr):
        clip.render(t, out=host[t])
    dev = torch.from_numpy(host).to("cuda:0")
    del host
    torch.cuda.synchronize()
    fb = s.width * s.height
    ptrs = [dev.data_ptr() + t * fb for t in range(nfr)]
    ls = [s.width] * nfr
    kw = dict(shakiness=s.shakiness, accuracy=s.accuracy, stepsize=s.stepsize, mincontrast=s.mincontrast)

    def timed(det):
        for _ in range(warmup):
            det.detect_device(ptrs, ls)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            det.detect_device(ptrs, ls)
        api.release_wait()
        torch.cuda.synchronize()
        return nfr * steps / (time.perf_counter() - t0)

    plain = api.Detector(s.width, s.height, **kw)
    plain_fps = timed(plain)
    ref = plain.detect_device(ptrs, ls, keep=True)
    S = plain.md.fieldscoarse.fieldSize
    plain.close()
    def learned(walk):
        """the learned form with windows that may walk `walk` times before
        the full search (1: the first edge minimum falls back)"""
        old = os.environ.get("VIDSTAB_AMD_PROPOSAL_WALK")
        os.environ["VIDSTAB_AMD_PROPOSAL_WALK"] = str(walk)
        try:
            lrn = api.Detector(s.width, s.height, **kw)
            lrn.set_proposal(weights)
        finally:
            if old is None:
                del os.environ["VIDSTAB_AMD_PROPOSAL_WALK"]
            else:
                os.environ["VIDSTAB_AMD_PROPOSAL_WALK"] = old
        fps = timed(lrn)
        f0, b0 = lrn.proposal_stats(0)
        g0, c0 = lrn.proposal_stats(1)
        got = lrn.detect_device(ptrs, ls, keep=True)
        f1, b1 = lrn.proposal_stats(0)
        g1, c1 = lrn.proposal_stats(1)
        lrn.close()
        return fps, got, (f1 - f0, b1 - b0, g1 - g0, c1 - c0)

    learned_fps, got, (nf0, nb0, nf1, nb1) = learned(4)
    edge_fps, got1, (ef0, eb0, ef1, eb1) = learned(1)
    del dev
    cn, ce, an, ae = api.compare_forms(ref, got, S)
    errs = [api.transform_distance(s.width, s.height, api.simple_transform(s.width, s.height, p),
                                   api.simple_transform(s.width, s.height, q)) for p, q in zip(ref[1:], got[1:])]
    cn1, ce1, an1, ae1 = api.compare_forms(ref, got1, S)
    errs1 = [api.transform_distance(s.width, s.height, api.simple_transform(s.width, s.height, p),
                                    api.simple_transform(s.width, s.height, q)) for p, q in zip(ref[1:], got1[1:])]
    out = {"workload": f"{s.width}x{s.height}, {nfr} frames of config 5 clip {clip_index} (motion blur, not in "
                       "the training set), device-resident", "frames": nfr,
           "plain_fps": round(plain_fps, 2), "learned_fps": round(learned_fps, 2),
           "speedup_vs_plain": round(learned_fps / plain_fps, 3),
           "fields_equal_pct": round(100.0 * ae / max(an, 1), 3),
           "coarse_fields_equal_pct": round(100.0 * ce / max(cn, 1), 3),
           "fine_fields_equal_pct": round(100.0 * (ae - ce) / max(an - cn, 1), 3),
           "fallback_pct": round(100.0 * nb0 / max(nf0, 1), 3),
           "fine_fallback_pct": round(100.0 * nb1 / max(nf1, 1), 3),
           "global_transform_err_px": round(float(np.mean(errs)), 5),
           "window_walk": 4,
           # the rule as first specified: the full search at the first edge minimum
           "first_edge_fallback": {"learned_fps": round(edge_fps, 2), "speedup_vs_plain": round(edge_fps / plain_fps, 3),
                                   "fields_equal_pct": round(100.0 * ae1 / max(an1, 1), 3),
                                   "fallback_pct": round(100.0 * eb0 / max(ef0, 1), 3),
                                   "fine_fallback_pct": round(100.0 * eb1 / max(ef1, 1), 3),
                                   "global_transform_err_px": round(float(np.mean(errs1)), 5)}}
    log(f"4K: plain {plain_fps:.1f} fps, learned {learned_fps:.1f} fps ({learned_fps / plain_fps:.2f}x); fields "
        f"equal {out['fields_equal_pct']:.2f}% (fine {out['fine_fields_equal_pct']:.2f}%), fallbacks coarse "
        f"{out['fallback_pct']:.2f}% / fine {out['fine_fallback_pct']:.2f}%, mean transform error "
        f"{out['global_transform_err_px']:.4f} px")
    e = out["first_edge_fallback"]
    log(f"4K learned, full search at the first edge minimum: {e['learned_fps']:.1f} fps ({e['speedup_vs_plain']:.2f}x), "
        f"fields equal {e['fields_equal_pct']:.2f}%, fallbacks coarse {e['fallback_pct']:.2f}% / fine "
        f"{e['fine_fallback_pct']:.2f}%")
    return out


def run_learned(args):
    f = measure_forms_4k(args.frames, args.steps, args.warmup, args.weights, args.clip)
    res = {
        "metric": METRIC, "value": f["learned_fps"], "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(f["frames"] * 1000.0 / f["learned_fps"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8+bf16",
        "data": "synthetic",
        "config": {"workload": "learned-proposal detect, " + f["workload"], "parallelism": "1 GPU"},
    }
    res.update({k: v for k, v in f.items() if k not in ("workload", "frames")})
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
