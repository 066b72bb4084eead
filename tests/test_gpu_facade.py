"""The drop-in C++ API (include/mpm/material_point_method.hpp) driving the
reference demo's loop (realtime-deformations/main.cpp:48-54,119-218): the
prebuilt realtime-deformations_amd/lib/facade_demo against the oracle run on
the same particles and collider poses."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "realtime-deformations_amd", "lib", "facade_demo")


def test_facade_demo_loop_matches_oracle(tmp_path):
    frames = 60
    prefix = str(tmp_path / "demo")
    out = subprocess.run([DEMO, str(frames), prefix], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    init = np.fromfile(prefix + ".init.f32", np.float32).reshape(-1, 7)
    cols = np.fromfile(prefix + ".colliders.f32", np.float32).reshape(-1, 13)
    fin = np.fromfile(prefix + ".final.f32", np.float32).reshape(-1, 15)
    n = init.shape[0]
    assert n == 2147 and cols.shape[0] == 3
    p = O.make_params((20, 20, 20), 0.05, bc_mode=O.BC_COLLIDERS, theta_s=5e-3)
    s = O.Sim(p, n)
    s.set_particles(init[:, 0:3], init[:, 3:6], init[:, 6], np.full(n, 1.4e5, np.float32))
    s.set_colliders(cols)
    s.stage(O.STAGE_P2G)
    s.stage(O.STAGE_VOLUMES)
    s.substeps(1e-5, frames)
    ref = s.particles()
    pos = np.fromfile(prefix + ".positions.f32", np.float32).reshape(-1, 3)
    np.testing.assert_array_equal(pos, fin[:, 0:3])  # getPositions() == getParticles() positions
    assert rel_err(fin[:, 0:3], ref["x"]) < 1e-5
    assert rel_err(fin[:, 3:6], ref["v"]) < 1e-4
    assert rel_err(fin[:, 6:15], ref["FE"]) < 1e-4
    # the viewer loop calls getParticles() once per particle every frame
    # (main.cpp:257-271): one device read per frame, not one per call
    assert f"particle reads {frames} in {frames} frames of {n} getParticles() calls each" in out.stdout, out.stdout
    drawn = np.fromfile(prefix + ".drawn.f32", np.float32).reshape(-1, 4)
    np.testing.assert_array_equal(drawn[:, 0:3], fin[:, 0:3])
    assert np.all(drawn[:, 3] == np.float32(0.02))
