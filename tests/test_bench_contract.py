"""bench.py host logic: the one-GPU contract and the algorithmic byte counts
the roofline figures are computed from."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_of_the_timed_mode():
    u = 227673985 / 442368000          # unique / element-local nodes, 192x192x96 p=4
    mode = bench.DEFAULT_MODE
    # element kernel's share: stress 24 r+w, BF16 compensation 12 r+w, memory
    # variables 72 r+w, moduli and tau 12 r, velocity 12u r (the force it
    # writes for the node kernel is outside the algorithmic bytes)
    assert bench.elem_kernel_bytes(mode, u, True) == pytest.approx(228 + 12 * u)
    assert bench.elem_kernel_bytes(mode, u, True) == pytest.approx(234.18, abs=0.01)
    # whole step: adds the node kernel's v write, its compensation r+w and
    # dt/M per unique node (40u)
    assert bench.step_bytes(mode, u, True) == pytest.approx(254.76, abs=0.01)
    assert bench.step_bytes(mode, u, True) == pytest.approx(
        bench.elem_kernel_bytes(mode, u, True) + 40 * u)
    # 16 sources sharing the statics
    assert bench.step_bytes(mode, u, True, S=16) == pytest.approx(
        bench.step_bytes(mode, u, True) - (12 + 4 * u) * 15 / 16)


def test_five_memory_words_and_the_scope_table():
    """The five-word memory form drops 12 B read and 12 B written per node;
    with six words the step bytes equal the scope table's figures for the
    192x192x96 workload (FP64 436.8, FP32 + Kahan(v) 230.8, + Kahan(sigma)
    278.8, + BF16 statics 223.7 B per element-local node)."""
    u = 227673985 / 442368000
    mode = bench.DEFAULT_MODE
    assert bench.elem_kernel_bytes(mode, u, True, 5) == pytest.approx(204 + 12 * u)
    assert bench.elem_kernel_bytes(mode, u, True, 5) == pytest.approx(210.18, abs=0.01)
    assert bench.step_bytes(mode, u, True, nm=5) == pytest.approx(bench.step_bytes(mode, u, True) - 24)
    for m, b in (("fp64", 436.8), ("fp32+kahan_v", 230.8), ("fp32+kahan_v+kahan_sigma", 278.8),
                 ("fp32+kahan_v+bf16_static", 223.7)):
        assert bench.step_bytes(m, u, True) == pytest.approx(b, abs=0.05), m


def test_fp64_and_fp32_modes_sizes():
    assert bench.mode_sizes("fp64") == (8, 0, 0, 8)
    assert bench.mode_sizes("fp32+kahan_v+kahan_sigma_bf16") == (4, 4, 2, 4)
    assert bench.mode_sizes("fp32+kahan_v+bf16_static")[3] == 2
    assert bench.mode_sizes("fp32+kahan_v_bf16+kahan_sigma_bf16") == (4, 2, 2, 4)
