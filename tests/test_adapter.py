"""The C++ ReactorBase adapter (include/pelelm_reactor.hpp) under a PeleLM-shaped
driver (tests/adapter/driver.cpp): the driver compiles against the header unchanged,
runs the reference's call sequence, and on the GPU reacts a box given as Array4 views
into an STemp-like multi-component array (rhoY, rhoH, T planes, Source/PeleLM.cpp:5728-5730)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pelelm_amd import cells

HERE = os.path.dirname(os.path.abspath(__file__))
ADAPTER = os.path.join(HERE, "adapter")
DRIVER = os.path.join(ADAPTER, "driver")


def test_driver_builds_and_runs_lifecycle_on_host():
    subprocess.run(["make", "-s", "-C", ADAPTER], check=True)
    out = subprocess.run([DRIVER, "host"], check=True, capture_output=True, text=True).stdout
    assert "host ok nspec 9" in out


@pytest.mark.gpu
def test_driver_reacts_box_on_gpu(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("tests/adapter/driver not built")
    n = 8
    d = cells.generate_workload("lidryer_16", ncell=n ** 3)
    N = d["rY"].shape[0]
    inp = np.concatenate([d["rY"].reshape(-1), d["T"], d["rE"]])
    fi, fo = tmp_path / "in.bin", tmp_path / "out.bin"
    inp.tofile(fi)
    print(f"adapter driver: {n}^3 LiDryer cells", flush=True)
    r = subprocess.run([DRIVER, "gpu", str(n), str(fi), str(fo)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fo)
    ncell = n ** 3
    rY = out[: N * ncell].reshape(N, ncell)
    T = out[N * ncell:(N + 1) * ncell]
    rE = out[(N + 1) * ncell:(N + 2) * ncell]
    FC = out[(N + 2) * ncell:]
    ref = O.Mech("LiDryer").react(d["rY"], d["T"], d["rE"], d["dt"], rtol=1e-6, atol=1e-10)
    tol = 10 * (1e-6 * np.abs(ref["rY"]) + 1e-10)
    assert np.max(np.abs(rY - ref["rY"]) / tol) < 1.0
    assert np.max(np.abs(T - ref["T"])) < 1e-2
    np.testing.assert_array_equal(rE, d["rE"])
    assert np.all(FC > 0)
