"""ABD (ASL Binary Dump) files: the Python codec, the compat layer's C++ stream
(src/writers/aslABDFormat.h:36-135, src/writers/aslABDFormat.cxx:35-81) reading
and writing the same bytes, and ASL's own ABD test built unchanged against the
compat headers."""
import os
import subprocess

import numpy as np
import pytest

from asl_amd.abd import read_abd, write_abd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TEST = "/root/reference/test/testABD/testABDFormat.cc"


def _build(src, exe):
    cmd = ["g++", "-std=c++17", "-O0", "-Wno-ignored-qualifiers", "-I", os.path.join(ROOT, "asl_amd", "compat", "include"),
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", os.path.join(ROOT, "asl_amd"),
           "-laslcompat", "-laslhip", "-Wl,-rpath," + os.path.join(ROOT, "asl_amd")]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]


def test_python_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    comp = rng.normal(size=(19, 5, 6, 7))
    write_abd(tmp_path / "a.abd", "f", comp, dx=0.25, position=(1.0, 2.0, 3.0))
    name, dx, size, pos, back = read_abd(tmp_path / "a.abd")
    assert (name, dx, size, pos) == ("f", 0.25, (5, 6, 7), (1.0, 2.0, 3.0))
    assert np.array_equal(back, comp)
    # layout: 4+1 (name) + 8 (dx) + 4+12 (size) + 4+24 (position) + 4 (count) + data
    assert os.path.getsize(tmp_path / "a.abd") == 5 + 8 + 16 + 28 + 4 + 8 * comp.size


CPP = r'''
#include <cstdio>
#include <writers/aslABDFormat.h>
int main(int argc, char **argv) {
  // read the file Python wrote with the ABD stream operators
  asl::ABDFileIn in(argv[1]);
  std::string name; asl::Block b; unsigned int nc = 0;
  in >> name >> b >> nc;
  const int len = b.getSize()[0] * b.getSize()[1] * b.getSize()[2];
  std::vector<double> v(len);
  double s = 0.;
  for (unsigned int c = 0; c < nc; ++c) {
    in >> std::make_pair(v.data(), (unsigned int)len);
    for (int i = 0; i < len; ++i) s += v[i] * (c + 1);
  }
  std::printf("%s %g %d %d %d %g %g %g %u %.17g\n", name.c_str(), b.dx, b.getSize()[0], b.getSize()[1],
              b.getSize()[2], b.position[0], b.position[1], b.position[2], nc, s);
  // write a container with writeABD
  asl::Block ib(asl::makeAVec(3, 4, 2), 0.5, asl::makeAVec(0., 0., 0.));
  asl::DataWithGhostNodesACLData d(ib, 2, acl::TYPE_DOUBLE, 1);
  for (unsigned int c = 0; c < 2; ++c)
    for (size_t i = 0; i < d.values_[c].size(); ++i) d.values_[c][i] = 100. * c + i;
  asl::writeABD(argv[2], d, "pair");
  return 0;
}
'''


def test_compat_stream_reads_and_writes_the_same_bytes(tmp_path):
    src = tmp_path / "abd_io.cc"
    src.write_text(CPP)
    _build(src, tmp_path / "abd_io")
    rng = np.random.default_rng(3)
    comp = rng.normal(size=(3, 4, 5, 6))
    write_abd(tmp_path / "in.abd", "f", comp, dx=0.5, position=(-0.5, 1.5, 2.5))
    p = subprocess.run([str(tmp_path / "abd_io"), str(tmp_path / "in.abd"), str(tmp_path / "out.abd")],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    f = p.stdout.split()
    assert f[:9] == ["f", "0.5", "4", "5", "6", "-0.5", "1.5", "2.5", "3"]
    ref = sum((c + 1) * comp[c].sum() for c in range(3))
    assert float(f[9]) == pytest.approx(ref, rel=1e-12)
    name, dx, size, pos, back = read_abd(tmp_path / "out.abd")
    # the container's ghost-inclusive block: the internal block grown by one
    assert (name, dx, size, pos) == ("pair", 0.5, (5, 6, 4), (-0.5, -0.5, -0.5))
    n = 5 * 6 * 4
    assert np.array_equal(back.reshape(2, n), np.array([np.arange(n), 100 + np.arange(n)], dtype=np.float64))


@pytest.mark.skipif(not os.path.exists(REF_TEST), reason="ASL source tree not mounted")
def test_asl_abd_test_passes_unchanged(tmp_path):
    _build(REF_TEST, tmp_path / "testABDFormat")
    p = subprocess.run([str(tmp_path / "testABDFormat")], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("Test passed") == 4
