#!/usr/bin/env python3
"""ST reconstruction throughput on one MI355X.

Workload (default): space-time reconstruction of 8 frames of a 256^3 fp32
volume, 45 golden-angle views per frame on a 512x512 detector (sid 600 mm,
sdd 1000 mm, 1 mm pitch), primal-dual sigma = tau = 0.2, SART alpha = 0.3,
lambda = 0.01, k = 0.1, e = 0.05, huber = 0.01, STV spatio-temporal prior,
voxel-driven BP, clamped cubic warp.  Synthetic data: seeded ellipsoid phantom deformed by a smooth
sinusoidal field; projections are the forward projections of the phantom.

One step = one ST iteration (a primal-dual pass over all frames: per frame
dual with backward warp, primal with forward warp, a full SART sweep over the
frame's views, extrapolation), run in the batched form captured in a hipGraph.
All inputs are resident in HBM before the timed region.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ST recon iters/sec at 256^3 x 8 frames; projector % HBM peak; vs CPU"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# angles: 0 even over 360, 1 golden angle, 2 van der Corput; P may be a list
# of per-frame view counts (uneven batches)
CONFIGS = {
    "st256": dict(vol=(256, 256, 256), det=(512, 512), F=8, P=45, amp=4.0, angles=1),
    "small": dict(vol=(64, 64, 64), det=(128, 128), F=2, P=8, amp=2.0, angles=1),
    # 32 frames with seeded view counts U{8..64} (seed 7), 384x384 detector
    "uneven128": dict(vol=(128, 128, 128), det=(384, 384), F=32,
                      P=[int(v) for v in np.random.default_rng(7).integers(8, 65, 32)], amp=3.0, angles=2),
    # full scale: 512^3 x 16 frames, 1024x1024 detector, 60 low-discrepancy views per frame
    "st512": dict(vol=(512, 512, 512), det=(1024, 1024), F=16, P=60, amp=4.0, angles=2),
}


def projector_kernel_name(tile):
    """Kernel of a projector shape (stt_set_projector_tile): 0, -32, -8 are
    the slab projector with 16x16, 32x8 and 8x32 rays per block; positive
    values the level-synchronous projector with that many rays per wave row."""
    if tile <= 0:
        tw = 16 if tile == 0 else -tile
        return f"k_fp_slab<SART,{tw},{256 // tw}>"
    return f"k_fp_level<SART,{tile}>"


def views_of(cfg, f):
    P = cfg["P"]
    return P[f] if isinstance(P, list) else P


def setup(stt, cfg, bp="voxel"):
    vol, det, F = cfg["vol"], cfg["det"], cfg["F"]
    g = stt.Geometry.make(vol, det, (1.0, 1.0, 1.0), 1.0, 600.0, 1000.0)
    p = stt.params_default()
    p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.1, 0.05, 0.01, 0.3
    p.bp = stt.STT_BP_RAY if bp == "ray" else stt.STT_BP_VOXEL
    p.prior = stt.STT_PRIOR_STV
    p.interp = stt.STT_INTERP_CUBIC_CLAMPED
    Ps = [views_of(cfg, f) for f in range(F)]
    ctx = stt.Context(g, p, Ps)
    angles = [stt.synth_angles(cfg["angles"], Ps[f], offset=sum(Ps[:f])) for f in range(F)]
    projections = []
    for f in range(F):
        t0 = time.time()
        P = Ps[f]
        truth = stt.synth_phantom_frame(vol, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=cfg["amp"])
        ctx.set_views(f, angles[f])
        ctx.set_volume(f, truth)
        b = np.empty((P, det[0] * det[1]), np.float32)
        for k in range(P):
            b[k] = ctx.forward_project(f, k)[0]
        ctx.set_projections(f, b)
        projections.append(b)
        if f < F - 1:
            ctx.set_flow(f, stt.synth_flow(vol, f, deform_seed=2, amp=cfg["amp"]))
        log(f"setup: frame {f} phantom + {P} projections ({time.time() - t0:.1f}s)")
    zero = np.zeros(int(np.prod(vol)), np.float32)
    for f in range(F + 1):
        ctx.set_volume(f, zero)
    ctx.pd_init()
    return g, p, ctx, angles, projections


def group_bench(args):
    """Configuration "uneven batched solves": 32 frames with seeded view
    counts U{8..64} (seed 7) and volume sizes drawn from {128^3, 192^3}
    (seed 8), 384x384 detector, van der Corput views; SART (one sweep) + PD
    ATV, k = e = 0, so the frames are independent solves.  Frames of one size
    share a context; both contexts run as one group whose iteration is one
    hipGraph.  Reports group iterations per second (all 32 frames)."""
    import spacetimetomography_amd as stt

    F, det = 32, (384, 384)
    P = [int(v) for v in np.random.default_rng(7).integers(8, 65, F)]
    sizes = [int(v) for v in np.random.default_rng(8).choice([128, 192], F)]
    ctxs, frames = [], {}
    for n in sorted(set(sizes)):
        idx = [f for f in range(F) if sizes[f] == n]
        frames[n] = idx
        vol = (n, n, n)
        g = stt.Geometry.make(vol, det, (1.0, 1.0, 1.0), 1.0, 600.0, 1000.0)
        p = stt.params_default()
        p.sigma, p.tau, p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.2, 0.2, 0.01, 0.0, 0.0, 0.01, 0.3
        p.prior = stt.STT_PRIOR_ATV
        ctx = stt.Context(g, p, [P[f] for f in idx])
        zero = np.zeros(n ** 3, np.float32)
        for j, f in enumerate(idx):
            ctx.set_views(j, stt.synth_angles(2, P[f], offset=sum(P[:f])))
            ctx.set_volume(j, stt.synth_phantom_frame(vol, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=3.0))
            ctx.set_projections(j, np.stack([ctx.forward_project(j, k)[0] for k in range(P[f])]))
        for j in range(len(idx) + 1):
            ctx.set_volume(j, zero)
        ctx.pd_init()
        ctxs.append(ctx)
        log(f"group: {len(idx)} frames of {n}^3, {sum(P[f] for f in idx)} views")
    grp = stt.Group(ctxs)
    grp.iterate(max(args.warmup, 1), use_graph=True)
    ms = grp.iterate(args.steps, use_graph=True)
    per = ms / args.steps
    log(f"group: {args.steps} iterations {ms:.1f} ms -> {per:.2f} ms per iteration")
    out = {
        "metric": "uneven batched solves: group iterations/s (32 frames, mixed 128^3/192^3)",
        "value": 1e3 / per,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantoms; projections are their forward projections)",
        "config": {"workload": f"32 frames, views U{{8..64}} (total {sum(P)}), sizes 128^3 x "
                               f"{len(frames.get(128, []))} and 192^3 x {len(frames.get(192, []))}, 384x384 "
                               "detector, SART(1) + PD ATV, one hipGraph per group iteration"},
    }
    grp.close()
    for c in ctxs:
        c.close()
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime (a plain
    copy kernel of known size: reads and writes exactly mib MiB), so that a
    counter pass can scale FETCH_SIZE / WRITE_SIZE against known bytes."""
    hip = C.CDLL("libamdhip64.so.7")
    n = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, n)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, n, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: {mib} MiB device-to-device")


def load_traffic(config, shape, box_kib):
    """HBM traffic per projector launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated against a copy kernel),
    or None when no summary matches this configuration (projector shape and,
    for the slab shapes, the staged box size)."""
    path = os.path.join(ROOT, "profiles", f"traffic_projector_{config}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("shape") != shape or (shape <= 0 and t.get("box_kib", 24) != box_kib):
        return None
    return t


def cpu_baseline(g, p, cfg, angles, projections, nviews, faithful):
    """Times the CPU oracle on one frame's PD step + nviews SART views and
    scales to one ST iteration (F frames x P views)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    vol, F = cfg["vol"], cfg["F"]
    Ps = [views_of(cfg, f) for f in range(F)]
    P = Ps[1]
    nviews = min(nviews, P)
    oc = O.OracleCtx(g, p, [Ps[0], P])
    for f in range(2):
        oc.set_views(f, angles[f])
        oc.set_projections(f, projections[f].reshape(-1))
    import spacetimetomography_amd as stt
    oc.set_flow(0, stt.synth_flow(vol, 0, deform_seed=2, amp=cfg["amp"]))
    zero = np.zeros(int(np.prod(vol)), np.float32)
    for f in range(3):
        oc.set_volume(f, zero)
    oc.pd_init()
    t_pd, t_v = C.c_double(), C.c_double()
    rc = O.orc.o_ctx_time_sample(oc.h, 1, nviews, int(faithful), C.byref(t_pd), C.byref(t_v))
    assert rc == 0
    per_view = t_v.value / nviews
    t_iter = F * t_pd.value + sum(Ps) * per_view
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / t_iter,
        "unit": "iters/s",
        "cores": threads,
        "kind": "port",
        "variant": "reference-faithful" if faithful else "lean",
        "sample": (f"frame 1 of {F}: PD dual+primal with both warps ({t_pd.value:.2f}s) and {nviews} of {P} "
                   f"SART views ({per_view:.2f}s/view), scaled to {F} frames and {sum(Ps)} views"),
    }


def flow_pair(stt, n_side, amp=2.0):
    """Two synthetic frames of the deformed ellipsoid phantom (frames 0 and 1)."""
    dims = (n_side,) * 3
    I1 = stt.synth_phantom_frame(dims, 0, n_ellipsoids=10, seed=1, deform_seed=2, amp=amp)
    I2 = stt.synth_phantom_frame(dims, 1, n_ellipsoids=10, seed=1, deform_seed=2, amp=amp)
    return dims, I1, I2


def flow_bench(args):
    """Optical flow (L1Huber_OF) line: PD iterations/s of one 256^3 level from
    two fixed-count runs (their difference cancels copies, warps and set-up),
    the roofline of an iteration's primal + dual launches, the reference-
    parameter pair solve time and the oracle's iteration rate."""
    import spacetimetomography_amd as stt

    n_side = 256
    dims, I1, I2 = flow_pair(stt, n_side)
    V = int(np.prod(dims))
    p = stt.of_params_reference()
    p.tol, p.warps = 0.0, 1  # no early stop: every run does exactly max_iter + 1 iterations
    zeros3, zeros10 = np.zeros(3 * V, np.float32), np.zeros(10 * V, np.float32)
    ext = np.asarray(dims, np.float32)

    def timed(iters):
        p.max_iter = iters
        t = time.perf_counter()
        stt.of_level(dims, ext, I1, I2, zeros3, zeros10, p, tuple(p.cut))
        return time.perf_counter() - t

    timed(49)  # warm-up (module load, graph instantiation)
    lo, hi = 49 + 50 * args.warmup, 49 + 50 * (args.warmup + args.steps)
    t_lo, t_hi = timed(lo), timed(hi)
    per_iter = (t_hi - t_lo) / (hi - lo)
    log(f"flow: {lo + 1} iterations {t_lo:.3f}s, {hi + 1} iterations {t_hi:.3f}s -> {per_iter * 1e3:.3f} ms/iteration")
    # primal, on the ROI: reads y (10 planes), ux uy uz (3), V (3), writes V' (3) = 19 planes;
    # dual, on every voxel: reads V' (3), V (3), ux uy uz ut (4), y (10), writes y (10) = 30 planes
    # (the step sizes tau and sigma2 are recomputed from ux uy uz)
    c = p.cut
    roi_vox = (n_side - c[0] - c[1]) * (n_side - c[2] - c[3]) * (n_side - c[4] - c[5])
    it_bytes = 4 * (19 * roi_vox + 30 * V)
    achieved = it_bytes / per_iter
    # measured HBM bytes of one primal + dual pair at this size (committed counter summary)
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_flow256.json")) as f:
            t = json.load(f)
        if t.get("voxels") == V:
            traffic = t["bytes_per_iteration"]
    except (OSError, ValueError, KeyError):
        pass
    pr = stt.of_params_reference()
    t = time.perf_counter()
    _, its = stt.of_multiscale(dims, I1, I2, zeros3, pr)
    t_pair = time.perf_counter() - t
    log(f"flow: reference-parameter pair solve {t_pair:.2f}s, iterations per level/warp {its.tolist()}")
    out = {
        "metric": "optical flow PD iterations/s (L1/Huber-TV, one 256^3 level)",
        "value": 1.0 / per_iter,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": hi - lo,
        "warmup": lo,
        "ms_per_step": per_iter * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (frames 0 and 1 of the seeded deformed ellipsoid phantom)",
        "config": {"workload": "L1Huber_OF primal+dual iteration, 256x256x256 fp32, ROI cut 8,8,8,45,8,8"},
        "roofline": {"bound": "hbm", "kernel": "k_of_primal + k_of_dual (one PD iteration)",
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "bytes_per_iteration": it_bytes},
        "pair_solve_s": t_pair,
        "pair_iterations": its.tolist(),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O

        def otimed(iters):
            t = time.perf_counter()
            O.of_level(dims, ext, I1, I2, zeros3, zeros10, tol=0.0, lam=p.lambda_, max_iter=iters, warps=1,
                       roi=tuple(p.cut))
            return time.perf_counter() - t

        a, b = otimed(4), otimed(14)
        cpu_per_iter = (b - a) / 10
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": 1.0 / cpu_per_iter, "unit": "iters/s", "cores": threads, "kind": "port",
                               "sample": "oracle L1Huber_OF at 256^3: 15 minus 5 iterations (difference)"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        log(f"flow cpu baseline {out['cpu_baseline']['value']:.3f} it/s")
    print(json.dumps(out), flush=True)


def io_bench(args):
    """On-disk formats line: TIFF read/write throughput of the projection
    stack of the 256^3 x 8 workload (512x512 detector, 8 x 45 pages, fp32,
    one file as the reference's -i input) and Analyze write/read of one
    3 x 256^3 flow field, through the library's codec; the same stack
    through libtiff (Pillow with its libtiff codec, the TIFF library CImg's
    load_tiff/save_tiff call) is the CPU baseline.  Files go to $TMPDIR (page
    cache), so the figures are codec throughput, not disk throughput."""
    import shutil
    import tempfile

    import spacetimetomography_amd as stt

    W, H, Z = 512, 512, 8 * 45
    rng = np.random.default_rng(5)
    stack = rng.standard_normal((Z, H, W), dtype=np.float32)
    flow = rng.standard_normal((3, 256, 256, 256), dtype=np.float32)
    reps = max(args.steps, 1)
    d = tempfile.mkdtemp(prefix="stt_io_", dir=os.environ.get("TMPDIR", "/tmp"))

    def best(fn):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return min(ts)

    try:
        tif, hdr = os.path.join(d, "proj.tif"), os.path.join(d, "flow.field.hdr")
        t_tw = best(lambda: stt.save_tiff(tif, stack))
        got = stt.load_image(tif)
        assert np.array_equal(got[0], stack), "TIFF round trip differs"
        del got
        t_tr = best(lambda: stt.load_image(tif))
        t_aw = best(lambda: stt.save_analyze(hdr, flow))
        t_ar = best(lambda: stt.load_image(hdr))
        nbytes, fbytes = stack.nbytes, flow.nbytes
        log(f"io: TIFF write {nbytes / t_tw / 1e9:.2f} GB/s, read {nbytes / t_tr / 1e9:.2f} GB/s; "
            f"Analyze write {fbytes / t_aw / 1e9:.2f} GB/s, read {fbytes / t_ar / 1e9:.2f} GB/s")
        out = {
            "metric": "projection-stack TIFF read throughput (codec, page cache)",
            "value": nbytes / t_tr / 1e9,
            "unit": "GB/s",
            "n_gpus": 0,
            "steps": reps,
            "warmup": 1,
            "ms_per_step": t_tr * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic (seeded normal values)",
            "config": {"workload": f"TIFF stack {W}x{H}x{Z} fp32 ({nbytes / 2**20:.0f} MiB), "
                                   "Analyze flow 3x256^3 fp32"},
            "tiff_write_gbs": nbytes / t_tw / 1e9,
            "analyze_write_gbs": fbytes / t_aw / 1e9,
            "analyze_read_gbs": fbytes / t_ar / 1e9,
            "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
        }
        if not args.no_cpu_baseline:
            from PIL import Image, TiffImagePlugin

            TiffImagePlugin.WRITE_LIBTIFF = True
            TiffImagePlugin.READ_LIBTIFF = True
            ptif = os.path.join(d, "pil.tif")

            def pil_write():
                ims = [Image.fromarray(stack[z]) for z in range(Z)]
                ims[0].save(ptif, save_all=True, append_images=ims[1:], compression="raw")

            def pil_read():
                with Image.open(ptif) as im:
                    for z in range(im.n_frames):
                        im.seek(z)
                        np.asarray(im)

            t_pw, t_pr = best(pil_write), best(pil_read)
            with Image.open(tif) as im:  # libtiff reads the file written here
                im.seek(Z - 1)
                assert np.array_equal(np.asarray(im), stack[Z - 1])
            out["cpu_baseline"] = {"value": nbytes / t_pr / 1e9, "unit": "GB/s", "cores": 1, "kind": "library",
                                   "sample": "the same stack read page by page by libtiff through Pillow "
                                             f"(write {nbytes / t_pw / 1e9:.2f} GB/s)"}
            log(f"io baseline: libtiff read {nbytes / t_pr / 1e9:.2f} GB/s, write {nbytes / t_pw / 1e9:.2f} GB/s")
    finally:
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps(out), flush=True)


def cli_bench(args):
    """Command-line driver line: the reference's Render loop end to end
    (stt-tomography, default profile) on 8 frames of 256^3 with 4 rounds of 10
    views per frame on a 512x512 detector: inputs written as the reference
    reads them (one TIFF projection stack, one TIFF volume per frame, an
    lds.txt), then `steps` outer iterations (flow of the 7 frame pairs with
    their dumps; from the second on, 2 PD passes with per-frame dumps).
    Reports the wall time per outer iteration after the first (the drain of
    the last iteration's files included) and the driver's own split into
    flow, PD and the warp + queueing of the dumps."""
    import re
    import shutil
    import subprocess
    import tempfile

    import spacetimetomography_amd as stt
    from spacetimetomography_amd.build import CLI

    dims, det, F, R = (256, 256, 256), (512, 512), 8, 4
    outer = max(args.steps, 2)
    d = tempfile.mkdtemp(prefix="stt_cli_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        start = stt.synth_angles(1, F * R).reshape(F, R)  # golden-angle start of every round
        g = stt.Geometry.make(dims, det, sid=600.0, sdd=1000.0)
        ctx = stt.Context(g, stt.params_default(), [10 * R] * F)
        stack = []
        for f in range(F):
            truth = stt.synth_phantom_frame(dims, f, n_ellipsoids=10, seed=1, deform_seed=2, amp=4.0)
            ctx.set_view_quaternions(f, stt.view_rounds(start[f]))
            ctx.set_volume(f, truth)
            for v in range(10 * R):
                stack.append(ctx.forward_project(f, v)[0].reshape(det[1], det[0]))
            stt.save_tiff(os.path.join(d, f"init-{f:03d}.tif"), (0.8 * truth).reshape(dims[::-1]))
        ctx.close()
        stt.save_tiff(os.path.join(d, "proj.tif"), np.stack(stack))
        del stack
        with open(os.path.join(d, "lds.txt"), "w") as fh:
            fh.write("\n".join(" ".join(repr(float(a)) for a in row) for row in start))
        cmd = [CLI, "-i", "proj.tif", "-f", "256", "-f", "256", "-f", "256", "-u", "512", "-u", "512",
               "--oxyz", "0", "--sdg", "0", "-v", str(F), "-r", str(R), "-g", "600", "-j", "1000",
               "-s", "0.2", "-t", "0.2", "-b", "2", "-p", "1", "--joint-iters", str(outer),
               "--init", "init-%03d.tif", "-o", "out/rec", "--flow-dir", "flow"]
        log("cli:", " ".join(cmd))
        t = time.perf_counter()
        proc = subprocess.run(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        wall = time.perf_counter() - t
        if proc.returncode != 0:
            log(proc.stdout[-4000:])
            sys.exit(proc.returncode)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    rows = [tuple(float(x) for x in m.groups()) for m in
            re.finditer(r"outer iteration (\d+): ([\d.]+) s wall; flow ([\d.]+) s, PD ([\d.]+) s, "
                        r"warp \+ dump queueing ([\d.]+) s", proc.stdout)]
    load = float(re.search(r"inputs read and uploaded in ([\d.]+) s", proc.stdout).group(1))
    drain = float(re.search(r"remaining dumps written in ([\d.]+) s", proc.stdout).group(1))
    for line in proc.stdout.splitlines():
        if line.startswith(("outer iteration", "inputs read", "remaining dumps", "done in")):
            log(line)
    # files are written by a background thread behind the next flow solve; the
    # last iteration's files have nothing to hide behind, so their drain time is
    # charged to the measured iterations
    sec = float((sum(r[1] for r in rows) + drain) / len(rows))
    out = {
        "metric": "CLI Render seconds per outer iteration (flow of every pair + 2 PD passes + dumps)",
        "value": sec,
        "unit": "s",
        "n_gpus": 1,
        "steps": len(rows),
        "warmup": 1,
        "ms_per_step": sec * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded deformed ellipsoid phantom; projections are its forward projections)",
        "config": {"workload": "stt-tomography, 256x256x256 x 8 frames, 40 views per frame (4 rounds of 10), "
                               "512x512 detector, reference flow parameters"},
        "flow_s": float(np.mean([r[2] for r in rows])),
        "pd_s": float(np.mean([r[3] for r in rows])),
        "warp_and_dump_queueing_s": float(np.mean([r[4] for r in rows])),
        "final_dump_drain_s": drain,
        "inputs_load_s": load,
        "process_wall_s": wall,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="st256", choices=sorted(CONFIGS))
    ap.add_argument("--form", default="batched-graph", choices=["batched-graph", "batched", "plain"])
    ap.add_argument("--cpu-views", type=int, default=0,
                    help="SART views in the CPU-baseline sample (default: all views of one frame)")
    ap.add_argument("--cpu-lean", action="store_true", help="also time the lean CPU variant")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--projector-tile", type=int, default=None,
                    help="projector shape (0, -32, -8 slab; 64/32/16/8 level-synchronous); default: autotune")
    ap.add_argument("--bp", default="voxel", choices=["voxel", "ray"],
                    help="backprojector of the SART update (the reference's --bp flag); the headline is voxel")
    ap.add_argument("--projector-box", type=int, default=None, choices=[16, 20, 24, 28],
                    help="slab projector's staged box in KiB; default: autotune with a slab shape, else 24")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first (counter calibration)")
    ap.add_argument("--flow", action="store_true",
                    help="measure the optical flow solver instead (steps/warmup count 50-iteration chunks)")
    ap.add_argument("--io", action="store_true",
                    help="measure the TIFF / Analyze codec instead (steps = repetitions, best kept)")
    ap.add_argument("--cli", action="store_true",
                    help="measure the command-line driver end to end instead (steps = outer iterations, >= 2)")
    ap.add_argument("--mixed", action="store_true",
                    help="measure uneven batched solves of mixed 128^3/192^3 frames as one group instead")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(frame-sharded multi-GPU is a design, not built)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
    if args.flow:
        flow_bench(args)
        return
    if args.io:
        io_bench(args)
        return
    if args.cli:
        cli_bench(args)
        return
    if args.mixed:
        group_bench(args)
        return

    import spacetimetomography_amd as stt

    cfg = CONFIGS[args.config]
    t0 = time.time()
    g, p, ctx, angles, projections = setup(stt, cfg, args.bp)
    log(f"setup done in {time.time() - t0:.1f}s")
    if args.projector_tile is not None:
        tile = args.projector_tile
        ctx.set_projector_tile(tile)
        log(f"projector tile {tile} (fixed)")
    else:
        tile, tile_ms = ctx.autotune()
        log(f"autotune: projector tile {tile} (ms per batched SART launch, shapes 0/-32/-8/64/32/16/8: "
            f"{[round(v, 4) for v in tile_ms]})")
    box = 24
    if args.projector_box is not None:
        box = args.projector_box
        ctx.set_projector_box(box)
        log(f"projector box {box} KiB (fixed)")
    elif tile <= 0:
        box, box_ms = ctx.autotune_box()
        log(f"autotune: projector box {box} KiB (ms per batched SART launch, boxes 16/20/24/28 KiB: "
            f"{[round(v, 4) for v in box_ms]})")
    batched = args.form != "plain"
    use_graph = args.form == "batched-graph"

    ctx.iterate(args.warmup, batched=batched, use_graph=use_graph)
    ctx.synchronize()
    log(f"warmup: {args.warmup} iterations")
    t_start = time.perf_counter()
    ctx.iterate(args.steps, batched=batched, use_graph=use_graph)
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    log(f"timed: {args.steps} iterations in {elapsed:.3f}s")

    # Per-kernel durations from HIP events on the context stream (eager launches).
    ksteps = max(1, min(args.steps, 3))
    fp_ms = bp_ms = dual_ms = primal_ms = aux_ms = 0.0
    fp_n = bp_n = 0
    for _ in range(ksteps):
        t = ctx.iterate(1, batched=batched, use_graph=False, timing=True)
        fp_ms += t.fp_ms
        fp_n += t.fp_launches
        bp_ms += t.bp_ms
        bp_n += t.bp_launches
        dual_ms += t.dual_ms
        primal_ms += t.primal_ms
        aux_ms += t.aux_ms
    ctx.synchronize()
    log(f"kernel timing: {ksteps} eager iterations")

    # The projector launches of one iteration in the launch shape the timed
    # hipGraph uses (single-frame launches on the frame-group streams,
    # replayed as a graph), and in the eager batched shape, HIP events around
    # each (stt_time_projector; outputs to scratch).
    proj_graph_ms = ctx.time_projector(1, reps=3) if use_graph else None
    proj_eager_ms = ctx.time_projector(0, reps=3)
    # Re-trace of the flows after a flow change (the command-line loop's
    # cadence: once per outer iteration, i.e. every 2 PD passes in Render):
    # rewrite every flow with its own values, then time the preparation.
    for f in range(cfg["F"] - 1):
        ctx.set_flow(f, ctx.get_buffer(stt.STT_BUF_FLOW, f))
    tr = ctx.iterate(0, batched=batched, use_graph=False, timing=True)
    log(f"projector phase: graph form {proj_graph_ms} ms, eager {proj_eager_ms:.3f} ms; "
        f"re-trace of {tr.trace_launches} flows {tr.trace_ms:.3f} ms")

    vol, det, F = cfg["vol"], cfg["det"], cfg["F"]
    Ps = [views_of(cfg, f) for f in range(F)]
    V = int(np.prod(vol))
    D = det[0] * det[1]
    # frame-views per projector launch: all frames' view k (batched), one (plain)
    frames_per_fp = sum(Ps) / max(Ps) if batched else 1
    fp_avg_s = fp_ms / max(fp_n, 1) / 1e3
    fp_bytes = frames_per_fp * (4 * V + 8 * D)  # volume read once + b read + correction write
    achieved = fp_bytes / fp_avg_s
    iter_bytes = sum(Ps) * (12 * V + 12 * D) + F * 124 * V
    ms_per_step = elapsed / args.steps * 1e3
    proj_iter_bytes = sum(Ps) * (4 * V + 8 * D)  # every frame-view's projector bytes

    views_desc = str(Ps[0]) if len(set(Ps)) == 1 else f"{min(Ps)}-{max(Ps)} (seeded, total {sum(Ps)})"
    traffic = load_traffic(args.config, tile, box) if batched else None
    out = {
        "metric": (METRIC if args.config == "st256" else
                   "ST recon iters/sec at %dx%dx%d x %d frames; projector %% HBM peak" % (*cfg["vol"], cfg["F"]))
        + ("" if args.bp == "voxel" else "; ray-driven BP"),
        "value": args.steps / elapsed,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded deformed ellipsoid phantom, projections = its forward projections)",
        "config": {
            "workload": (f"space-time PD+SART reconstruction {vol[0]}x{vol[1]}x{vol[2]} fp32 x {F} frames, "
                         f"{views_desc} {['even', 'golden-angle', 'van der Corput'][cfg['angles']]} views/frame, "
                         f"{det[0]}x{det[1]} detector, STV prior, {args.bp} BP, clamped-cubic warp"),
            "form": args.form,
            "volume": list(vol), "frames": F, "views_per_frame": cfg["P"], "detector": list(det),
            "projector_tile": tile,
            "projector_box_kib": box if tile <= 0 else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": (f"{projector_kernel_name(tile)} "
                       "(forward projection + fused SART correction, all frames' view k)"),
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None if traffic is None else traffic["bytes_per_launch"],
            "launch_avg_ms": fp_avg_s * 1e3,
            "bytes_per_launch": fp_bytes,
        },
        "traffic_source": None if traffic is None else traffic["source"],
        # the projector phase of one iteration in the timed form's own launch
        # shape (single-frame launches on 8 frame-group streams, one hipGraph),
        # and in the eager batched shape of `roofline`, both timed live
        "roofline_graph_form": None if proj_graph_ms is None else {
            "bound": "hbm",
            "kernel": "projector launches of one iteration, hipGraph of single-frame launches on frame-group streams",
            "achieved": proj_iter_bytes / (proj_graph_ms / 1e3) / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": proj_iter_bytes / (proj_graph_ms / 1e3) / HBM_PEAK,
            "ms_per_iteration": proj_graph_ms,
            "bytes_per_iteration": proj_iter_bytes,
        },
        "projector_eager_ms_per_iteration": proj_eager_ms,
        # The flows stay fixed over the timed iterations, as over the PD passes
        # of one outer iteration; after a flow change the PD warps' trace end
        # points are recomputed once (k_trace).  At the reference's cadence of
        # 2 PD passes per flow (ST_Tomography.cpp:477) that adds half a re-trace
        # to every iteration:
        "retrace_ms_per_flow_change": tr.trace_ms,
        "value_retrace_every_2_passes": 1e3 / (ms_per_step + tr.trace_ms / 2),
        "kernel_ms_per_step": {
            "fp": fp_ms / ksteps, "bp": bp_ms / ksteps, "pd_dual": dual_ms / ksteps, "pd_primal": primal_ms / ksteps,
            # the x<->z layout copy of a level projector shape; a slab shape
            # launches none, and the interval holds only the events' own gaps
            ("layout_copy" if tile > 0 else "event_gaps"): aux_ms / ksteps,
        },
        "iteration_hbm_frac": iter_bytes / (ms_per_step / 1e3) / HBM_PEAK,
    }
    if not args.no_cpu_baseline:
        t1 = time.time()
        nv = args.cpu_views if args.cpu_views > 0 else views_of(cfg, 1)
        out["cpu_baseline"] = cpu_baseline(g, p, cfg, angles, projections, nv, faithful=True)
        log(f"cpu baseline (faithful) {out['cpu_baseline']['value']:.5f} it/s in {time.time() - t1:.1f}s")
        if args.cpu_lean:
            out["cpu_baseline_lean"] = cpu_baseline(g, p, cfg, angles, projections, nv, faithful=False)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
