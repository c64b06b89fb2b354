"""CPU check that the wall-distance bisection of the particle kernels
(mglc_amd/csrc/sphere_geom.hpp: halvings on the quadratic alone far from the
surface, the reference's own test near it; shared by host and device builds)
equals the reference's plain calQ loop (particle_bounceback.F90:101-141) bit
for bit on every boundary link of 300 random spheres (~8e5 links; integer
radii, node-centred spheres and node coordinates up to ~30000 included)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_band_bisection_equals_plain_loop(tmp_path):
    exe = tmp_path / "calq_check"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-I", ROOT,
                    os.path.join(ROOT, "tests", "calq_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe), "300"], check=False, capture_output=True, text=True, timeout=120)
    print(out.stdout)
    fields = dict(zip(out.stdout.split()[0::2], map(int, out.stdout.split()[1::2])))
    assert out.returncode == 0
    assert fields["links"] > 500000 and fields["mismatches"] == 0
