"""The benchmark's byte models and its committed traffic records (CPU only).

The dominant-kernel roofline of the bench line divides algorithmic bytes per
launch by the launch time; the CN family moves 4 units per launch with one
component per launch and 10 (p read once) with the three components per
launch.  `roofline.traffic` comes from the committed counter record only when
that record was taken on the same workload, family and kernel form."""
import importlib.util
import json
import os

import pytest

from distributednavierstokes_amd.workloads import channel_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_cn_bytes_per_launch(bench):
    prob = channel_problem((512, 384, 512), "implicit", "fp64")
    N = 512 * 384 * 512
    assert bench.family_bytes_per_launch(prob, 5, 8) == 4 * N * 8
    assert bench.family_bytes_per_launch(prob, 5, 8, cn3=True) == 10 * N * 8
    # three components per launch: fewer bytes than three launches of one
    assert bench.family_bytes_per_launch(prob, 5, 8, cn3=True) < 3 * bench.family_bytes_per_launch(prob, 5, 8)


def test_traffic_record_matches_its_form(bench):
    prob = channel_problem((512, 384, 512), "implicit", "fp64")
    path = os.path.join(ROOT, "profiles", "cn_ydiffusion_traffic_channel_512x384x512_fp64.json")
    with open(path) as f:
        rec = json.load(f)
    got = bench.measured_traffic(prob, "cn_ydiffusion", rec["form"])
    assert got == rec["bytes_per_launch"]
    # counter traffic of the recorded kernel: the sum of its read and write MB
    k = next(iter(rec["kernels"].values()))
    assert abs(got / 1e6 - (k["read_mb"] + k["write_mb"])) < 0.2
    # at least the algorithmic bytes of that form
    cn3 = rec["form"] == "seg3"
    assert got >= bench.family_bytes_per_launch(prob, 5, 8, cn3)
    # another kernel form, precision or family: no record
    other = "seg" if cn3 else "seg3"
    assert bench.measured_traffic(prob, "cn_ydiffusion", other) is None
    assert bench.measured_traffic(channel_problem((512, 384, 512), "implicit", "mixed"), "cn_ydiffusion", "seg") is None
    assert bench.measured_traffic(prob, "momentum_rk") is None
