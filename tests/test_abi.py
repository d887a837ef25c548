"""The C ABI library loads and exports every symbol include/*.h declares.
No GPU work: only host-side entry points are called."""
import ctypes as C
import os
import re
import subprocess

import pytest

import sphe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in os.listdir(INCLUDE):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**(sphe_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_built_and_loads():
    assert os.path.exists(sphe.LIB_PATH), "run make -C sph-erosion_amd"
    L = sphe.lib()
    assert L.sphe_version().decode().startswith("sphe")


def test_every_declared_symbol_is_exported():
    decl = declared_functions()
    assert len(decl) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", sphe.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = decl - exported
    assert not missing, missing
    # every declared function has a ctypes signature in the Python mirror
    assert decl == set(sphe.SIGNATURES), decl ^ set(sphe.SIGNATURES)


def test_struct_layouts():
    assert C.sizeof(sphe.Particle) == 112  # FluidParticle, fluid_system.h:49-64
    assert C.sizeof(sphe.Params) == 4 * 19  # 19 four-byte fields of sphe_params


def test_default_params_are_the_reference_defaults():
    p = sphe.default_params()
    f32 = lambda x: float(C.c_float(x).value)  # noqa: E731
    # Erosion/fluid_system.h:460-478
    assert list(p.gravity) == [0.0, f32(-9.82), 0.0]
    assert p.dt == 0.0
    assert p.rest_density == f32(998.29)
    assert p.mass == f32(0.02)
    assert p.visc == f32(3.5)
    assert p.surf_tens == f32(0.0728)
    assert p.stiffness == 3.0
    assert p.h == f32(0.0457)
    assert p.box_half == f32(0.2)
    assert list(p.box_center) == [0.0, 0.0, 0.0]
    assert p.erosion == sphe.EROSION_OFF


def test_null_arguments_rejected_without_gpu():
    L = sphe.lib()
    assert L.sphe_create(7, None) == -1
    out = C.c_void_p()
    assert L.sphe_create(9, C.byref(out)) == -1
    assert L.sphe_destroy(None) == -1
    # per-scene entry points on a null context: no scene, nothing touched
    dptr, n = C.c_void_p(), C.c_size_t()
    assert L.sphe_positions_device(None, 0, C.byref(dptr), C.byref(n)) == -5
    assert L.sphe_terrain_mesh_device(None, 0, C.byref(dptr), C.byref(n)) == -5
    buf = (C.c_float * 3)()
    assert L.sphe_read_buffer(None, 0, 0, buf, C.sizeof(buf)) == -5


@pytest.mark.skipif(subprocess.run(["which", "g++"], capture_output=True).returncode != 0, reason="no g++")
def test_cpp_dropin_header_compiles(tmp_path):
    """A translation unit written against the reference's FluidSystemSPH
    method names compiles against include/sphe_fluid_system.hpp and links
    with libsphe.so."""
    src = tmp_path / "front.cpp"
    src.write_text(open(os.path.join(ROOT, "tests", "cpp", "front_end.cpp")).read())
    exe = tmp_path / "front"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", INCLUDE, str(src), "-o", str(exe),
                        os.path.join(ROOT, "sph-erosion_amd", "libsphe.so"),
                        "-Wl,-rpath," + os.path.join(ROOT, "sph-erosion_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
