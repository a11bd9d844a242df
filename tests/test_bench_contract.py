"""bench.py's CPU-side contract: the one-GPU refusal and the roofline byte figures (no GPU needed)."""
from __future__ import annotations

import importlib.util
import subprocess
import sys

from conftest import ROOT

# the 6.3M-tet bump channel: 256 x 64 x 64 hexes, 6 tets each
BUMP_CELLS, BUMP_FACES = 6_291_456, 12_656_640


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_cell_stage():
    b = load_bench()
    # FP64: 3 x 40 B of states + 32 B of cell metrics, 32 B per face (3 x f64 + 2 x int32)
    assert b.algorithmic_bytes_per_cell_stage(BUMP_CELLS, BUMP_FACES) == 216.375
    # a closed hexahedral cell has 6 faces, each shared by 2 cells: 3 faces per cell
    assert b.algorithmic_bytes_per_cell_stage(1000, 3000) == 248.0
    # mixed: 20 + 40 + 40 + 20 B of states, 16 B of FP32 metrics, 20 B per face
    assert b.mixed_bytes_per_cell_stage(BUMP_CELLS, BUMP_FACES) == 176.234375


def test_recorded_traffic_is_at_least_the_algorithmic_bytes():
    b = load_bench()
    for kernel, per_cell in (("kStageFused", b.algorithmic_bytes_per_cell_stage(BUMP_CELLS, BUMP_FACES)),
                             ("kMixStage", b.mixed_bytes_per_cell_stage(BUMP_CELLS, BUMP_FACES))):
        t = b.measured_traffic(kernel, BUMP_CELLS, 0)
        assert t is not None, kernel
        ratio = t / (per_cell * BUMP_CELLS)
        assert 1.0 <= ratio < 1.5, (kernel, ratio)
    assert b.measured_traffic("kStageFused", 12345, 0) is None
