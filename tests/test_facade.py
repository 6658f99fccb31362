"""The C++ facade (include/stt.hpp): the reference's class names --
TypicalRayProducer, RayCaster, VolumeData, ST_Tomography -- over the C-ABI.

tests/facade_check.cpp is compiled against the header and libstt.so:
- its failure paths run on the CPU (no device needed: bad sizes, calls out
  of order, a missing lds.txt, a foreign sample distance, a projection stack
  that does not split into rounds of 10 views);
- on the GPU, TypicalRayProducer's forward projection and ray-BP scatter of
  views set through SetViewMatrix / SetProjectionMatrix (the Scene's
  matrices of each rotation) equal the oracle's, and ST_Tomography::Render
  equals the oracle's Render loop.
"""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt
from conftest import HAVE_GPU
from helpers import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "spacetimetomography_amd", "lib")


@pytest.fixture(scope="module")
def facade(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", exe, os.path.join(HERE, "facade_check.cpp"), "-L", LIBDIR, "-lstt",
                    "-Wl,-rpath," + LIBDIR], check=True)
    return exe


def _run(exe, args, cwd):
    r = subprocess.run([exe, *map(str, args)], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    print(r.stdout)
    return r


def test_facade_failure_paths(facade, tmp_path):
    r = _run(facade, ["errors", "device" if HAVE_GPU else "nodevice"], tmp_path)
    assert r.returncode == 0 and "FAIL" not in r.stdout
    assert r.stdout.count("OK") >= (9 if HAVE_GPU else 10)


@pytest.mark.gpu
def test_facade_producer_matches_oracle(facade, tmp_path):
    dims, det, ds, sid, sdd = (24, 22, 26), (40, 36), 1.0, 60.0, 100.0
    print(f"[facade-producer] {dims} det {det}: TypicalRayProducer vs oracle, 3 views", flush=True)
    g = stt.Geometry.make(dims, det, (1, 1, 1), ds, sid, sdd)
    V, D = int(np.prod(dims)), det[0] * det[1]
    vol = stt.synth_phantom(dims, 6, seed=9)
    q = stt.view_rounds([13.0])[:3]
    corr = np.random.default_rng(10).standard_normal(D).astype(np.float32)
    vol.tofile(tmp_path / "vol.raw")
    np.ascontiguousarray(q, np.float32).tofile(tmp_path / "q.raw")
    corr.tofile(tmp_path / "corr.raw")
    r = _run(facade, ["producer", tmp_path, *dims, *det, ds, sid, sdd, len(q)], tmp_path)
    assert r.returncode == 0, r.stdout
    raylen = np.zeros(D, np.float32)  # the producer's ray lengths persist across views until Initial()
    for k in range(len(q)):
        oax, raylen, ow = O.fp(g, q[k], vol, raylen=raylen)
        ax = np.fromfile(tmp_path / f"ax{k}.raw", np.float32)
        assert np.array_equal(np.fromfile(tmp_path / f"w{k}.raw", np.float32), ow), k
        assert np.array_equal(np.fromfile(tmp_path / f"rl{k}.raw", np.float32), raylen), k
        assert rel_l2(ax, oax) <= 1e-5, (k, rel_l2(ax, oax))
        oc, ow2 = O.bp_ray_scatter(g, q[k], corr)
        assert rel_l2(np.fromfile(tmp_path / f"s{k}.raw", np.float32), ow2) <= 1e-6, k
        assert rel_l2(np.fromfile(tmp_path / f"c{k}.raw", np.float32), oc) <= 1e-5, k


@pytest.mark.gpu
def test_facade_render_matches_oracle(facade, tmp_path):
    from test_gpu_cli import _oracle_render, _scene_at

    dims, det, k, e = (64, 64, 64), (96, 96), 0.1, 0.05
    print(f"[facade-render] {dims} det {det}: ST_Tomography::Render (2 outer iterations) vs oracle", flush=True)
    g, stack, init, q = _scene_at(tmp_path, dims, det)
    for f in range(2):
        init[f].tofile(tmp_path / f"init{f}.raw")
    r = _run(facade, ["render", tmp_path, *dims, *det, 600.0, 1000.0, k, e, 2, 2], tmp_path)
    assert r.returncode == 0, r.stdout
    p = stt.params_reference()
    p.lambda_tv, p.k_tp, p.e_bc, p.huber, p.alpha = 0.1, k, e, 0.1, 0.1
    flows, warped, vols = _oracle_render(g, p, stt.of_params_reference(), stack, init, q, 2, 2)
    for o in range(2):
        base = tmp_path / "FlowOutput" / "{:03d}.{:03d}.{:03d}.frame00.iter{:02d}".format(*dims, o)
        ef = rel_l2(stt.load_image(str(base) + ".field.hdr").reshape(-1), flows[o])
        ew = rel_l2(stt.load_image(str(base) + ".tiff").reshape(-1), warped[o])
        print(f"  outer {o}: flow relL2 {ef:.2e}, warped frame {ew:.2e}", flush=True)
        assert ef <= 1e-4 and ew <= 1e-4
    for f in range(2):
        for n in range(2):
            got = stt.load_image(str(tmp_path / "out" / f"rec-{f:03d}-frame-{n + 1}iter.tiff")).reshape(-1)
            err = rel_l2(got, vols[f][n])
            print(f"  frame {f} pass {n + 1}: relL2 vs oracle {err:.2e}", flush=True)
            assert err <= 1e-4
