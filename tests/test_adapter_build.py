"""The boundary from compiled code without a GPU: include/vct.h compiles as
C99 and as C++17, and the compiled pass adapter driver
(tests/adapter/clipmap_frame.cpp over include/vct_passes.hpp) builds with
-Werror, links against the in-tree libvct.so and the HIP runtime, and starts
(a missing input file is reported before any HIP call)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_header_compiles_as_c99_and_cpp17(tmp_path):
    c = tmp_path / "t.c"
    c.write_text('#include "vct.h"\nint main(void) { vct_config c = {0}; (void)c; return (int)sizeof(vct_mesh) - 120; }\n')
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(c),
                    "-o", str(tmp_path / "c")], check=True)
    assert subprocess.run([str(tmp_path / "c")]).returncode == 0
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "vct_passes.hpp"\nint main() { vct::FrameLayout f; return (int)f.frame_index; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", f"-I{ROOT / 'include'}", "-fsyntax-only", str(cpp)], check=True)


def test_adapter_driver_builds_links_and_starts(tmp_path):
    subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "tests" / "adapter")], check=True)
    exe = ROOT / "tests" / "adapter" / "_build" / "clipmap_frame"
    r = subprocess.run([str(exe), str(tmp_path / "missing.rec"), str(tmp_path / "out.rec")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr, r.stderr
