"""bench.py's command line and its algorithmic byte counts (no GPU)."""
import subprocess
import sys

from conftest import ROOT
import bench
from imex_sflow2d_dev_amd import workloads


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_per_active_cell():
    # 8 (3 n_v + 2) + 16 n_s S_settle + 48 + 9 per active cell and step
    c1, _ = workloads.config1(nx=16, ny=16)
    assert bench.kernel_bytes_per_cell(c1)["step"] == 193
    c2, _ = workloads.config2(nx=16, ny=16)
    per = bench.kernel_bytes_per_cell(c2)
    assert per["step"] == 249
    assert per["k_flux"] == 64 and per["k_mask"] == 9
    assert per["k_update"] == 176
    assert set(per) == set(bench.KERNELS) | {"step"}


def test_every_named_workload_validates():
    for name in workloads.WORKLOADS:
        fn = workloads.WORKLOADS[name]
        cfg, B = fn(nx=32, ny=32) if name != "c2-small" else fn(nx=32, ny=32)
        assert B.shape == (cfg.ny, cfg.nx)
    cfg, B = workloads.config4_member(3, min_side=40, max_side=60)
    assert B.shape == (cfg.ny, cfg.nx)


def test_traffic_is_read_only_for_the_profiled_window(tmp_path, monkeypatch):
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "c2_step_kernels_traffic.md").write_text(
        "window: steps 2493..2512\nbytes_per_cell k_update: 180.5\n")
    monkeypatch.setattr(bench, "HERE", str(tmp_path))
    assert bench.read_traffic("k_update", "c2", (2493, 2512)) == 180.5
    assert bench.read_traffic("k_update", "c2", (1, 5000)) is None
    assert bench.read_traffic("k_flux", "c2", (2493, 2512)) is None


def test_host_cores_counts_physical_cores():
    n, topo = bench.host_cores()
    assert 1 <= n <= 16
    assert "socket" in topo
