"""A Fortran host (ISO_C_BINDING, the reference's language) drives the GPU
solve through the C ABI: uniform slab, basal stress = driving stress."""
import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
DRIVER = os.path.join(ROOT, "imau-ice_amd", "fortran", "build", "slab_driver")


@pytest.mark.gpu
def test_fortran_driver_slab():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(line.split("=") for line in r.stdout.splitlines() if "=" in line)
    taub, taud = float(vals["taub  "]), float(vals["tau_d "])
    assert abs(taub - taud) <= 1e-6 * taud


def test_fortran_driver_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available() or not os.path.exists(DRIVER):
        pytest.skip("needs the built driver and no GPU")
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "no HIP device" in r.stdout
