"""The C++ drop-in (include/sphe_fluid_system.hpp) driven the way the
reference's front end drives FluidSystemSPH, against the reference step."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_front_end_matches_reference(tmp_path):
    exe = tmp_path / "front"
    lib_dir = os.path.join(ROOT, "sph-erosion_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "front_end.cpp"), "-o", str(exe),
                        os.path.join(lib_dir, "libsphe.so"), "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    steps = 5
    out = subprocess.run([str(exe), str(steps)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    R = O.ref32_step(O.lattice(1000), O.default_params(dt=0.01), steps)
    for line in lines[:2]:
        f = line.split()
        pid = int(f[0])
        pos = np.array([float(x) for x in f[1:4]])
        vel = np.array([float(x) for x in f[4:7]])
        rho, nb = float(f[7]), int(f[8])
        assert np.abs(pos - R["Position"][pid]).max() <= 1e-5
        assert np.abs(vel - R["Velocity"][pid]).max() <= 1e-3 * max(1.0, np.abs(R["Velocity"][pid]).max())
        assert abs(rho - R["Density"][pid]) <= 1e-5 * R["Density"][pid]
        assert nb == R["NeighbId"][pid]
    assert lines[2] == "n 1125"   # AddParticles(125) appends a 5^3 lattice
    assert lines[3] == "n 1000"   # Reset re-creates Initialize's count
    assert lines[4] == "range_error 1"
