"""CPU builds under AddressSanitizer + UndefinedBehaviorSanitizer.

* The FP64 oracle (oracle/cpu_ref.c, oracle/mlp_ref.c) with a driver that
  calls every entry point on every lattice / collision / forcing / force /
  equilibrium, degenerate extents, every entropic branch and both network
  shapes (oracle/sanitize_main.c, `make -C oracle sanitize`).
* The host-only rules of the C ABI -- configuration validation, kernel-shape
  choice, workgroup size (metalbm_amd/csrc/config_rules.hpp) -- and the HDF5
  field layout (metalbm_amd/csrc/hdf5_layout.hpp), driven with hostile
  configuration values (tests/native/host_rules.cpp).
Each build aborts at the first finding (-fno-sanitize-recover=all); a
planted out-of-bounds read shows the sanitizer is live."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="gcc/g++ not present")


def run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    return subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True,
                   timeout=300)
    r = run(os.path.join(ROOT, "oracle", "_native", "sanitize_oracle"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "640 step cases, 0 failures" in r.stdout


def build_host(tmp_path, extra=()):
    exe = str(tmp_path / "host_rules")
    cmd = (["g++", "-std=c++17"] + SAN + list(extra) +
           ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "metalbm_amd", "csrc"),
            os.path.join(ROOT, "tests", "native", "host_rules.cpp"), "-o", exe])
    subprocess.run(cmd, check=True, timeout=300)
    return exe


def test_host_rules_under_sanitizers(tmp_path):
    r = run(build_host(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert ", 0 failures" in r.stdout


def test_sanitizer_is_live(tmp_path):
    src = tmp_path / "oob.cpp"
    src.write_text("#include <vector>\n#include <cstdio>\n"
                   "int main(int argc, char**) { std::vector<int> v(4); "
                   "std::printf(\"%d\\n\", v.data()[4 + argc]); return 0; }\n")
    exe = str(tmp_path / "oob")
    subprocess.run(["g++", "-std=c++17"] + SAN + [str(src), "-o", exe], check=True, timeout=120)
    r = run(exe)
    assert r.returncode != 0 and "AddressSanitizer" in r.stderr
