"""The benchmark's bookkeeping, on the CPU: the algorithmic byte counts it
prices the roofline with, the traffic lookup by launch shape, and the
window phase that gives a timed run its share of output steps."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_unit_and_stored_bytes():
    # the 600 B FP32 unit: 73 float words read and written plus one 8-byte gating word read and written
    assert bench.UNIT_BYTES[1] == 2 * 73 * 4 + 16
    assert bench.UNIT_BYTES[0] == 2 * 73 * 8 + 16
    # stored: the 73 words and 3 packed 32-bit words, each read once and written once
    assert bench.stored_bytes_per_cru_step(1) == 608
    assert bench.stored_bytes_per_cru_step(0) == 1192


def test_kernel_bytes_add_up_to_the_step_plus_the_shared_reads():
    # the two kernels both read ci/cnsr of sub-voxels 0-1, cs, cp and cjsr: FP32 8 + 8 + 4 + 4 + 4 = 28 B
    kb = bench.kernel_bytes(1)
    assert kb["k_cyto"] + kb["k_cru_step"] == bench.stored_bytes_per_cru_step(1) + 28
    kb = bench.kernel_bytes(0)
    assert kb["k_cyto"] + kb["k_cru_step"] == bench.stored_bytes_per_cru_step(0) + 56
    # the gating in the sub-voxel launch moves the RyR and flag words (+16) off the CRU kernel's writes (-4)
    assert bench.kernel_bytes(1, 3) == {"k_cyto": 524 + 16, "k_cru_step": 112 - 4}
    assert bench.kernel_bytes(1, 4) == {"k_one": 524 + 112}


def test_traffic_lookup_by_launch_shape():
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        t = json.load(f)
    exact = t["batch@128,64,2"]["hbm_bytes_per_step"]
    assert bench.traffic_from_profiles("batch", [128, 64, 2]) == exact
    # the bytes do not depend on the workgroup sizes: another size of the same sweep form and schedule
    assert bench.traffic_from_profiles("batch", [64, 128, 2]) == exact
    assert bench.traffic_from_profiles("batch", [1024, 64, 2]) == t["batch@1024,128,2"]["hbm_bytes_per_step"]
    assert bench.traffic_from_profiles("cell", [256, 256, 4]) == t["cell@256,512,4"]["hbm_bytes_per_step"]
    # no measurement for the schedule or the workload: none reported
    assert bench.traffic_from_profiles("batch", [128, 64, 0]) is None
    assert bench.traffic_from_profiles("cell_f64", [256, 256, 4]) is None
    for k, e in t.items():
        if not k.startswith("_"):
            assert e["hbm_bytes_per_step"] == pytest.approx(sum(v["hbm_bytes_per_launch"] for v in e["kernels"].values()))


@pytest.mark.parametrize("steps", [1, 20, 99, 100, 101, 150, 151, 249, 250, 8000])
def test_window_phase_gives_the_production_share_of_output_steps(steps):
    spo = 100
    s0 = bench.window_start(steps, spo)
    outputs = sum(1 for n in range(s0, s0 + steps) if n % spo == 0)
    assert outputs == round(steps / spo)
