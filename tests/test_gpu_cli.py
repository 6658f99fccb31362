"""The command-line driver end to end on the GPU.

- Its files equal the same Render sequence (flow of every frame pair, warped
  frame + flow dumps, PD passes with per-frame volume dumps;
  ST_Tomography.cpp:397-614) issued in-process through the C-ABI, bit for bit.
- Its files equal the oracle's Render loop -- of_multiscale per frame pair,
  the clamped-cubic warp of the second frame by the new flow, then PD passes
  of o_ctx_iterate -- in the default profile and with --compat=reference.
"""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from helpers import rel_l2
from spacetimetomography_amd.build import CLI

pytestmark = pytest.mark.gpu

DIMS = (24, 22, 20)
DET = (32, 28)
START = [0.0, 7.0]  # one round of 10 views per frame


def _params(sigma, tau):
    p = stt.params_default()
    p.sigma, p.tau = sigma, tau
    return p


def _of_params():
    o = stt.of_params_reference()
    for k in range(6):
        o.cut[k] = 0
    return o


def _scene(tmp_path):
    g = stt.Geometry.make(DIMS, DET, sid=600.0, sdd=1000.0)
    truth = [stt.synth_phantom_frame(DIMS, t, n_ellipsoids=6, seed=3, deform_seed=4, amp=1.5) for t in range(2)]
    ctx = stt.Context(g, _params(0.2, 0.5), [10, 10])
    stack = []
    for f in range(2):
        ctx.set_view_quaternions(f, stt.view_rounds([START[f]]))
        ctx.set_volume(f, truth[f])
        for v in range(10):
            stack.append(ctx.forward_project(f, v)[0].reshape(DET[1], DET[0]))
    ctx.close()
    stt.save_tiff(str(tmp_path / "proj.tif"), np.stack(stack))
    os.makedirs(tmp_path / "Input")
    init = [(0.5 * t).astype(np.float32) for t in truth]
    for f in range(2):
        stt.save_tiff(str(tmp_path / "Input" / f"rose.-{f:03d}-frame-9iter.tiff"),
                      init[f].reshape(DIMS[2], DIMS[1], DIMS[0]))
    (tmp_path / "lds.txt").write_text(" ".join(str(s) for s in START))
    return g, np.stack(stack), init


def _expected_context(g, stack, init, sigma, tau):
    ctx = stt.Context(g, _params(sigma, tau), [10, 10])
    for f in range(2):
        ctx.set_projections(f, stack[10 * f:10 * f + 10].reshape(-1))
        ctx.set_view_quaternions(f, stt.view_rounds([START[f]]))
        ctx.set_volume(f, init[f])
    ctx.set_volume(2, init[1])
    return ctx


def _cli(tmp_path, *extra):
    cmd = [CLI, "-i", "proj.tif", "-f", str(DIMS[0]), "-f", str(DIMS[1]), "-f", str(DIMS[2]), "-u", str(DET[0]),
           "-u", str(DET[1]), "--oxyz", "0", "--sdg", "0", "-v", "2", "-r", "1", "-g", "600", "-j", "1000",
           "-o", "out/rec", "--joint-iters", "2", *extra]
    print("cli:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert p.returncode == 0, p.stdout
    return p.stdout


def _vol(path):
    return stt.load_image(str(path)).reshape(-1)


def test_cli_render_equals_in_process_sequence(tmp_path):
    g, stack, init = _scene(tmp_path)
    out = _cli(tmp_path, "-s", "0.2", "-t", "0.5", "-b", "2", "-p", "1")
    assert "done in" in out and "remaining dumps written in" in out
    ctx = _expected_context(g, stack, init, 0.2, 0.5)
    ofp = _of_params()
    base = "{:03d}.{:03d}.{:03d}.frame00.iter{:02d}".format(*DIMS, 0)
    for outer in range(2):
        ctx.optical_flow(0, ofp)
        uvw = ctx.get_buffer(stt.STT_BUF_FLOW, 0)
        warped = ctx.warp(uvw, ctx.get_volume(1), -1.0)
        base = tmp_path / "FlowOutput" / "{:03d}.{:03d}.{:03d}.frame00.iter{:02d}".format(*DIMS, outer)
        assert np.array_equal(_vol(str(base) + ".tiff"), warped)
        got = stt.load_image(str(base) + ".field.hdr")
        assert got.shape == (3, DIMS[2], DIMS[1], DIMS[0])
        assert np.array_equal(got.reshape(-1), uvw)
        if outer == 0:
            continue
        for n in (1, 2):
            ctx.iterate(1, batched=True, use_graph=True)
            for f in range(2):
                assert np.array_equal(_vol(tmp_path / "out" / f"rec-{f:03d}-frame-{n}iter.tiff"), ctx.get_volume(f))
    # the reconstruction moved towards the phantom's projections
    assert not np.array_equal(ctx.get_volume(0), init[0])
    ctx.close()


def test_cli_nested_sart_sweeps_equal_step_sequence(tmp_path):
    g, stack, init = _scene(tmp_path)
    _cli(tmp_path, "-s", "0.2", "-t", "0.5", "-b", "1", "-p", "2", "--no-flow")
    ctx = _expected_context(g, stack, init, 0.2, 0.5)
    ctx.pd_init()
    for f in range(2):
        ctx.pd_dual(f)
        ctx.pd_primal(f)
        ctx.sart_sweep(f)
        ctx.sart_sweep(f)
        ctx.pd_extrapolate(f)
    for f in range(2):
        assert np.array_equal(_vol(tmp_path / "out" / f"rec-{f:03d}-frame-1iter.tiff"), ctx.get_volume(f))
    assert not os.path.exists(tmp_path / "FlowOutput")
    ctx.close()


def test_cli_reports_failed_dump(tmp_path):
    """A file the background writer cannot create fails the run with its path."""
    _scene(tmp_path)
    os.makedirs(tmp_path / "out" / "rec-001-frame-1iter.tiff")  # a directory where a dump goes
    cmd = [CLI, "-i", "proj.tif", "-f", str(DIMS[0]), "-f", str(DIMS[1]), "-f", str(DIMS[2]), "-u", str(DET[0]),
           "-u", str(DET[1]), "--oxyz", "0", "--sdg", "0", "-v", "2", "-r", "1", "-o", "out/rec",
           "--joint-iters", "2", "-b", "2", "--no-flow"]
    print("cli:", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert p.returncode == 1, p.stdout
    assert "rec-001-frame-1iter.tiff failed" in p.stdout
    assert "done in" not in p.stdout


def _scene_at(tmp_path, dims, det):
    g = stt.Geometry.make(dims, det, sid=600.0, sdd=1000.0)
    truth = [stt.synth_phantom_frame(dims, t, n_ellipsoids=6, seed=3, deform_seed=4, amp=1.5) for t in range(2)]
    ctx = stt.Context(g, _params(0.2, 0.5), [10, 10])
    q = [stt.view_rounds([START[f]]) for f in range(2)]
    stack = []
    for f in range(2):
        ctx.set_view_quaternions(f, q[f])
        ctx.set_volume(f, truth[f])
        for v in range(10):
            stack.append(ctx.forward_project(f, v)[0].reshape(det[1], det[0]))
    ctx.close()
    stt.save_tiff(str(tmp_path / "proj.tif"), np.stack(stack))
    os.makedirs(tmp_path / "Input")
    init = [(0.5 * t).astype(np.float32) for t in truth]
    for f in range(2):
        stt.save_tiff(str(tmp_path / "Input" / f"rose.-{f:03d}-frame-9iter.tiff"),
                      init[f].reshape(dims[2], dims[1], dims[0]))
    (tmp_path / "lds.txt").write_text(" ".join(str(s) for s in START))
    return g, np.stack(stack), init, q


def _oracle_render(g, p, ofp, stack, init, q, outer, passes):
    """ST_Tomography::Render's loop on the oracle, in the driver's order:
    per outer iteration the flow of the frame pair from the previous flow
    (L1Huber_OF_MultiScale3D, :421-456) and the dump of frame 1 warped back by
    it (:462); from the second outer iteration on, `passes` PD passes (:477)
    with each frame's volume dumped after its sweep (:786-805)."""
    dims = tuple(g.vol)
    oc = O.OracleCtx(g, p, [10, 10])
    for f in range(2):
        assert O.orc.o_ctx_set_view_quaternions(oc.h, f, O.f32(np.ascontiguousarray(q[f].reshape(-1))), 10) == 0
        oc.set_projections(f, stack[10 * f:10 * f + 10].reshape(-1))
        oc.set_volume(f, init[f])
    oc.set_volume(2, init[1])
    oc.per_frame_raylen(True)
    flow = np.zeros(3 * int(np.prod(dims)), np.float32)
    flows, warped, vols = [], [], {}
    for o in range(outer):
        flow, _, _ = O.of_multiscale(dims, oc.get_volume(0), oc.get_volume(1), flow, tol=ofp.tol, lam=ofp.lambda_,
                                     max_iter=ofp.max_iter, norm=ofp.norm, warps=ofp.warps, eta=ofp.eta,
                                     sigma=ofp.sigma, levels=ofp.levels, huber=ofp.huber, cut=tuple(ofp.cut))
        oc.set_flow(0, flow)
        flows.append(flow)
        warped.append(O.warp(dims, flow, oc.get_volume(1), -1.0, interp=stt.STT_INTERP_CUBIC_CLAMPED))
        if o == 0:
            continue
        if o == 1:
            oc.pd_init()
        for _ in range(passes):
            oc.iterate(1)
            for f in range(2):
                vols.setdefault(f, []).append(oc.get_volume(f))
    return flows, warped, vols


@pytest.mark.parametrize("profile", ["default", "reference"])
def test_cli_render_matches_oracle(tmp_path, profile):
    dims, det = (DIMS, DET) if profile == "default" else ((64, 64, 64), (96, 96))
    print(f"[cli-oracle] profile={profile} {dims} det {det}, 2 frames x 10 views, 2 outer iterations", flush=True)
    g, stack, init, q = _scene_at(tmp_path, dims, det)
    cmd = [CLI, "-i", "proj.tif", "-f", str(dims[0]), "-f", str(dims[1]), "-f", str(dims[2]), "-u", str(det[0]),
           "-u", str(det[1]), "--oxyz", "0", "--sdg", "0", "-v", "2", "-r", "1", "-g", "600", "-j", "1000",
           "-o", "out/rec", "--joint-iters", "2", "-k", "0.1", "-e", "0.05"]
    if profile == "default":
        cmd += ["-s", "0.2", "-t", "0.5", "-b", "2", "-p", "1"]
        p = _params(0.2, 0.5)
        ofp = _of_params()
        passes = 2
    else:
        cmd += ["--compat=reference"]
        p = stt.params_reference()
        ofp = stt.of_params_reference()
        passes = 2
    p.k_tp, p.e_bc = 0.1, 0.05
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    flows, warped, vols = _oracle_render(g, p, ofp, stack, init, q, 2, passes)
    for o in range(2):
        base = tmp_path / "FlowOutput" / "{:03d}.{:03d}.{:03d}.frame00.iter{:02d}".format(*dims, o)
        ef = rel_l2(stt.load_image(str(base) + ".field.hdr").reshape(-1), flows[o])
        ew = rel_l2(_vol(str(base) + ".tiff"), warped[o])
        print(f"  outer {o}: flow relL2 vs oracle {ef:.2e}, warped frame {ew:.2e}", flush=True)
        assert ef <= 1e-4 and ew <= 1e-4, (o, ef, ew)
    for f in range(2):
        for n in range(passes):
            got = _vol(tmp_path / "out" / f"rec-{f:03d}-frame-{n + 1}iter.tiff")
            e = rel_l2(got, vols[f][n])
            print(f"  frame {f} pass {n + 1}: relL2 vs oracle {e:.2e}", flush=True)
            assert e <= 1e-4, (f, n, e)
        assert not np.array_equal(vols[f][-1], init[f])
