"""The benchmark's host-side contract, checked without a GPU: it refuses
more than one GPU, and it reports the committed HBM traffic of the contact
kernel only while the kernel sources are the ones it was measured on."""

import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_traffic_is_reported_only_for_the_measured_sources(tmp_path, monkeypatch):
    with open(os.path.join(ROOT, "profiles", "traffic_contact.json")) as fh:
        committed = json.load(fh)
    # A copy of the repository layout the lookup reads: the profile and the
    # kernel sources.
    (tmp_path / "profiles").mkdir()
    csrc = tmp_path / "sphere_amd" / "csrc"
    csrc.mkdir(parents=True)
    for f in ("dem_kernels.hip", "dem_math.hpp", "dem_types.hpp", "dem_arrays.hpp"):
        with open(os.path.join(ROOT, "sphere_amd", "csrc", f), "rb") as src:
            (csrc / f).write_bytes(src.read())
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    t = dict(committed)
    t["kernel_source_sha256"] = bench.kernel_source_hash()
    (tmp_path / "profiles" / "traffic_contact.json").write_text(json.dumps(t))
    got = bench.measured_traffic(3.0, 1)
    assert got is not None and got[0] == committed["bytes_per_launch"]
    # Any change to the sources makes the figure stale.
    with open(csrc / "dem_types.hpp", "a") as fh:
        fh.write("\n// changed\n")
    assert bench.measured_traffic(3.0, 1) is None
