"""Host-side checks of bench.py's contract: the workload table is consistent
with the library's configuration, more than one GPU is refused with a
message, and the bytes-per-cell model matches 2 Q sizeof(storage) plus the
entropic alpha state and the spectral force field."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("name", sorted(bench.WORKLOADS))
def test_workload_configs_build(name):
    wl = bench.WORKLOADS[name]
    cfg = bench.make_config(wl["cfg"])  # enum names resolve
    assert cfg.cells == cfg.nx * cfg.ny * cfg.nz
    assert wl["init"][0] in ("noise", "tgv", "spectral")
    assert wl["steps"] > 0


def test_bytes_per_cell_model():
    wl = bench.WORKLOADS
    b = lambda n: bench.bytes_per_cell(bench.make_config(wl[n]["cfg"]), wl[n]["cfg"])
    assert b("d3q19-bgk-256-f32") == 2 * 19 * 4
    assert b("d3q19-bgk-256-f64") == 2 * 19 * 8
    assert b("d3q27-bgk-384-f16") == 2 * 27 * 2
    assert b("d2q9-elbm-2048-newton") == 2 * 9 * 4 + 2 * 4       # + alpha state
    assert b("d2q9-bgk-turbulent2d-2048-f32") == 2 * 9 * 4 + 2 * 4  # + force field
    assert bench.DEFAULT == "d3q19-bgk-256-f32"
