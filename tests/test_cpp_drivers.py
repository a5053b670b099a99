"""The C++ host side: ASL-style drivers over the asl:: compatibility layer.

GPU: the C++ parity programs (tests/cpp/*.cc, built by build() into
asl_amd/bin) run the lid-driven cavity and the sphere channel through
asl::LBGK / asl::generateBC* exactly as ASL's examples/flow drivers do, and
compare with the CPU oracle; they exit with EXIT_SUCCESS like ASL's tests.

CPU: ASL's own flow drivers compile and link unchanged against the
compatibility headers (skipped where the ASL source tree is not mounted).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "asl_amd", "bin")
REF_FLOW = "/root/reference/examples/flow"


def _run(name, *args, timeout=600):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "%s not built (run __graft_entry__.build())" % exe
    p = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(p.stdout, p.stderr)
    return p


@pytest.mark.gpu
def test_cpp_cavity_64_1000():
    p = _run("cavity_parity", 64, 1000)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_sphere_channel():
    p = _run("sphere_parity", 64, 32, 300)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout and "cell types match masks: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_sphere_channel_full_size():
    # the benchmark lattice (256x128x128, sphere r = 16 at (64,64,64)), 20 steps
    p = _run("sphere_parity", 256, 128, 20, timeout=900)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout and "cell types match masks: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_stl_body_turbulence(tmp_path):
    # locomotive-style: STL body via readSurface(file, Block), LBGKTurbulence D3Q15,
    # tunnel velocity map, visualisation BCs, surface force, .vti output
    from test_gpu_surface_map import icosphere
    from test_oracle_surface_map import write_stl_binary
    path = tmp_path / "body.stl"
    write_stl_binary(path, icosphere(0.8, (2.6, 1.5, 3.0)))  # file (x, y, z) = ASL (z, y, x)
    p = _run("stl_body_parity", path, 200, tmp_path / "out")
    assert p.returncode == 0
    assert "map bit-identical: yes" in p.stdout and ", bit-identical: yes" in p.stdout
    assert (tmp_path / "out_0.vti").stat().st_size > 10000


@pytest.mark.skipif(not os.path.isdir(REF_FLOW), reason="ASL source tree not mounted")
@pytest.mark.parametrize("driver", ["flow", "flow2", "flow3", "flowRotatingCylinders", "locomotive",
                                    "locomotive_laminar", "locomotive_stability", "bus_wind", "compressor"])
def test_asl_flow_drivers_compile_unchanged(driver, tmp_path):
    cmd = ["g++", "-std=c++17", "-O0", "-Wno-ignored-qualifiers", "-I", os.path.join(ROOT, "asl_amd", "compat", "include"),
           "-I", os.path.join(ROOT, "include"), os.path.join(REF_FLOW, driver + ".cc"), "-o", str(tmp_path / driver),
           "-L", os.path.join(ROOT, "asl_amd"), "-laslcompat", "-laslhip"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
