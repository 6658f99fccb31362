"""The command-line driver (spacetimetomography_amd/lib/stt-tomography):
argtable2 flag semantics of the reference's main.cpp:13-95 and the two
profiles, checked through --dry-run (no device)."""
import json
import os
import subprocess

import numpy as np
import pytest

import spacetimetomography_amd as stt
from spacetimetomography_amd.build import CLI


@pytest.fixture()
def scene(tmp_path):
    """A 2-frame, 2-round scene on disk: 32x24 detector, 18x20x16 volume."""
    os.makedirs(tmp_path / "Input")
    stt.save_tiff(str(tmp_path / "proj.tif"), np.zeros((40, 24, 32), np.float32))
    for f in range(2):
        stt.save_tiff(str(tmp_path / "Input" / f"rose.-{f:03d}-frame-9iter.tiff"), np.zeros((16, 20, 18), np.float32))
    (tmp_path / "lds.txt").write_text("0 18\n5 23\n")
    return tmp_path


BASE = ["-i", "proj.tif", "-f", "18", "-f", "20", "-f", "16", "-u", "32", "-u", "24", "--oxyz", "0", "--sdg", "0",
        "-v", "2", "-r", "2"]


def run(cwd, *args):
    return subprocess.run([CLI, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=60)


def plan(cwd, *args):
    p = run(cwd, "--dry-run", *args)
    assert p.returncode == 0, p.stdout
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_required_options_as_argtable_mincount(scene):
    p = run(scene)
    assert p.returncode == 1
    for opt in ("-i|--projsvolume", "-f|--XYZ", "-u|--imgUV", "--oxyz", "--sdg"):
        assert f"missing option {opt}" in p.stdout
    assert "Try 'stt-tomography --help'" in p.stdout
    h = run(scene, "--help")
    assert h.returncode == 0 and "--AlgorithmIter=<int>" in h.stdout and "--bp={Voxelbased,Raybased}" in h.stdout


def test_code_defaults_of_main(scene):
    # main.cpp:67-95: huber 0.1, alpha 0.1, lambda 0.1, k = e = 0, sid 400, sdd 200, spacing 1, ds 1
    j = plan(scene, *BASE, "--sdg", "1", "--sdg", "2", "--sdg", "3")
    assert j["profile"] == "default"
    assert (j["sid"], j["sdd"], j["ds"], j["spacing"]) == (400, 200, 1, [1, 1, 1])
    assert (j["lambda"], j["k"], j["e"], j["huber"], j["alpha"]) == (0.1, 0, 0, 0.1, 0.1)
    assert (j["sigma"], j["tau"]) == (0.3, 0.3)
    assert (j["pd_passes"], j["sart_sweeps"], j["outer"]) == (2, 2, 20)
    assert j["output"] == "Test/test" and j["flow_dir"] == "FlowOutput"
    assert j["views"] == [20, 20] and j["stack"] == [32, 24, 40]
    assert j["start_deg"] == [0, 1, 2, 3]
    assert j["roi_cut"] == [0] * 8 and j["of_cut"] == [0] * 6


def test_reference_profile_overrides_ignored_flags(scene):
    j = plan(scene, *BASE, "--compat=reference", "-s", "9", "-t", "9", "--prior=STV", "-b", "7", "-p", "5",
             "-l", "0.05", "-k", "0.2", "-e", "0.3", "--huber", "0.5", "-a", "0.25", "--bp=Raybased")
    assert (j["sigma"], j["tau"], j["prior"]) == (0.1125, 1, stt.STT_PRIOR_ATV)
    assert (j["outer"], j["pd_passes"], j["sart_sweeps"]) == (20, 2, 1)
    assert j["start_deg"] == [0, 18, 5, 23]  # lds.txt, frame-major; --sdg ignored
    assert j["roi_cut"] == [8, 8, 8, 45, 8, 8, 30, 20] and j["roi_jtop"] == 7
    assert j["of_cut"] == [8, 8, 8, 45, 8, 8]
    # flags the reference does honour
    assert (j["lambda"], j["k"], j["e"], j["huber"], j["alpha"], j["bp"]) == (0.05, 0.2, 0.3, 0.5, 0.25, 1)


def test_argtable_syntax(scene):
    args = ["-i", "proj.tif", "-f18", "-f", "0x14", "-f", "16", "-u", "32", "-u", "24", "--oxyz=1.5", "--oxyz",
            "-2", "--sdg", "0", "--nframes", "2", "-r2", "--Sig", "0.2", "--Tau=0.25", "--prior", "TV",
            "--Bright constancy prior=0.125", "--detector spacing", "0.5", "-g", "600", "-j1000", "-c", "0.75"]
    j = plan(scene, *args, "--lds", "lds.txt")
    assert j["vol"] == [18, 20, 16]             # 0x14 = 20
    assert j["nframes"] == 2                    # --nframes is -v (first exact match)
    assert j["oxyz"] == [1.5, -2, 0]            # missing values keep the code default
    assert (j["sigma"], j["tau"], j["prior"]) == (0.2, 0.25, stt.STT_PRIOR_TV)
    assert (j["e"], j["ds"], j["sid"], j["sdd"], j["spacing"]) == (0.125, 0.5, 600, 1000, [0.75] * 3)
    # a partial -f keeps main.cpp's 63 for the rest
    p = run(scene, "--dry-run", "-i", "proj.tif", "-f", "18", "-u", "32", "-u", "24", "--oxyz", "0", "--sdg", "0",
            "-v", "2", "-r", "2", "--init", "zero")
    j = json.loads(p.stdout.strip().splitlines()[-1])
    assert j["vol"] == [18, 63, 63]


def test_argument_errors(scene):
    p = run(scene, *BASE, "-f", "4")
    assert p.returncode == 1 and "excess option -f|--XYZ" in p.stdout
    p = run(scene, *BASE, "-s", "abc")
    assert p.returncode == 1 and 'invalid argument "abc"' in p.stdout
    p = run(scene, *BASE, "--frobnicate")
    assert p.returncode == 1 and "invalid option" in p.stdout
    p = run(scene, *BASE, "stray")
    assert p.returncode == 1 and 'unexpected argument "stray"' in p.stdout


def test_input_checks(scene):
    p = run(scene, "--dry-run", *BASE[:-4], "-v", "2", "-r", "1", "--sdg", "0")  # 20 views per frame, 1 round
    assert p.returncode == 1 and "1 rounds of 10 views need 10" in p.stdout
    p = run(scene, "--dry-run", *BASE, "--lds", "none.txt", "--compat=reference")
    assert p.returncode == 1 and "cannot open 'none.txt'" in p.stdout
    p = run(scene, "--dry-run", *[a if a != "24" else "25" for a in BASE])
    assert p.returncode == 1 and "projection images are 32x24, -u says 32x25" in p.stdout
    p = run(scene, "--dry-run", *BASE, "--sdg", "1", "--sdg", "2", "--sdg", "3", "--init", "Input/v-%s.tif")
    assert p.returncode == 1 and "--init needs one integer conversion" in p.stdout
    p = run(scene, "--dry-run", *BASE, "--sdg", "1", "--sdg", "2", "--sdg", "3", "--init", "Input/v-%03d.tif")
    assert p.returncode == 1 and "cannot open 'Input/v-000.tif'" in p.stdout
    stt.save_tiff(str(scene / "Input" / "rose.-001-frame-9iter.tiff"), np.zeros((16, 20, 17), np.float32))
    p = run(scene, "--dry-run", *BASE, "--compat=reference")
    assert p.returncode == 1 and "is 17x20x16x1, the volume is 18x20x16" in p.stdout
