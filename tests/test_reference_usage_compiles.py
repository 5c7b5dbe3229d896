"""The drop-in headers accept code written the way the reference's programs
use them (src/benchmark_depth_super_res_cli.cpp:25-122,
src/depth_super_res_dataset.h:38-85): two-argument EdgeAwareSolver<S, O>
subclasses with overrides, ProblemImageData returned by value, and
EdgeAwareOptimize on its members.  CPU only: compiles and links
tests/cxx/reference_usage.cpp against include/ and libdts_amd.so."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "reference_usage.cpp")
PKG = os.path.join(ROOT, "domain_transform_solver_amd")


def test_reference_style_program_compiles_and_links(tmp_path):
    from domain_transform_solver_amd import _lib
    assert os.path.exists(_lib.LIB_PATH)
    exe = tmp_path / "reference_usage"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-parameter",
           "-I", os.path.join(ROOT, "include"), SRC, "-o", str(exe), "-L", PKG, "-ldts_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr


def test_plane_image_copies_own_their_pixels(tmp_path):
    """PlaneImage is the default image type without OpenCV: a copy or a move
    keeps `data` pointing at the object's own pixels."""
    src = tmp_path / "pi.cpp"
    src.write_text(r'''
#include "edge_aware_solver.h"
using domain_transform::PlaneImage;
static PlaneImage make() { PlaneImage p(2, 3, 4); p.ptr<float>()[5] = 7.f; return p; }
int main() {
  PlaneImage a = make();
  PlaneImage b = a;
  if (b.data == a.data || b.ptr<float>()[5] != 7.f) return 1;
  PlaneImage c(std::move(b));
  if (c.data != c.storage.data() || b.data != nullptr) return 2;
  PlaneImage d; d = c;
  if (d.data == c.data || d.ptr<float>()[5] != 7.f || d.rows != 2 || d.cols != 3) return 3;
  PlaneImage e; if (e.data != nullptr || !e.empty()) return 4;
  domain_transform::ProblemImageData pd; pd.target = a;
  if (pd.target.data == a.data) return 5;
  return 0;
}
''')
    exe = tmp_path / "pi"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-fsanitize=address,undefined",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", PKG, "-ldts_amd", "-Wl,-rpath," + PKG],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr)
