"""Calibrate the oracle's CPU speed against the reference's own CPU path.

    python oracle/calibrate.py            # writes oracle/calibration.json

Times the oracle (FP32, one thread, the reference's single-threaded stage
order) on the workload of a measurement of the compiled reference:
8,192 particles of the reference-sized snowball on a 64^3 grid (h = 1/64) with
the reference's constants (E = 1.4e5, nu = 0.2, theta_c = 2.5e-2,
theta_s = 5e-3, xi = 10, dt = 1e-5, no colliders).  The reference measured
83.3 ms per substep there, 9.8e4 particle-substeps/s on one core of the same
Intel Xeon @ 2.70 GHz; the ratio of the two rates converts the oracle's CPU
baseline on any host into reference-equivalent throughput.  Run it in the
container where the probe ran (test infrastructure only).
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "realtime-deformations_amd", "python")]

REFERENCE_PER_CORE = 9.8e4  # particle-substeps/s, compiled reference, 1 core, Xeon @2.70 GHz, same workload


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    import oracle as O
    from mpm_amd import scenes

    sc = scenes.snowball()
    n = sc["x"].shape[0]
    p = O.make_params(sc["dims"], sc["h"], compose=O.LITERAL, bc_mode=0, E=1.4e5, nu=0.2, xi=10.0,
                      theta_c=2.5e-2, theta_s=5e-3, threads=1)
    s = O.Sim(p, n)
    s.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    s.stage(O.STAGE_P2G)
    s.stage(O.STAGE_VOLUMES)
    steps = 100
    t0 = time.perf_counter()
    s.substeps(1e-5, steps)
    secs = time.perf_counter() - t0
    rate = n * steps / secs
    out = dict(workload="snowball, 8192 particles, 64^3, reference constants, dt 1e-5, 100 substeps",
               oracle_per_core=rate, reference_per_core=REFERENCE_PER_CORE,
               oracle_over_reference=rate / REFERENCE_PER_CORE, ms_per_substep=secs / steps * 1e3,
               cpu=cpu_model(), command="python oracle/calibrate.py")
    with open(os.path.join(HERE, "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
