"""torch.autograd wrappers of the forward/VJP pairs (projector and its
transpose, trilinear warp and its VJP) with a per-frame checkpointed
brightness-constancy term, run in a child process (tests/autograd_harness.py):
the checkpointed graph keeps no warped volume, its gradient equals the plain
graph's, and a directional derivative matches FP64 central differences of
the loss."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_autograd_wrappers_and_checkpointing():
    print("[autograd] child process tests/autograd_harness.py", flush=True)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "autograd_harness.py")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-3000:]
    r = json.loads(p.stdout.strip().splitlines()[-1])
    print(" ", r, flush=True)
    # same loss and the same gradient through both graphs
    assert abs(r["loss_checkpointed"] - r["loss_plain"]) <= 1e-9 * abs(r["loss_plain"])
    assert r["grad_rel_diff"] <= 1e-5
    # the checkpointed pairs keep no residual volume; the plain graph keeps one per pair
    assert r["saved_intermediate_bytes_plain"] - r["saved_intermediate_bytes_checkpointed"] == \
        r["pairs"] * r["residual_volume_bytes"]
    # the gradient is the loss's derivative (fp32 operators, so a loose bar)
    assert r["directional_rel_err"] <= 1e-2
