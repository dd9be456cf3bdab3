"""bench.py's command line on the CPU: the multi-GPU refusal, and the HBM
traffic summary's arithmetic on synthetic counter files (the timed legs
need the GPU)."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run_bench(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_multi_gpu_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def write_counters(directory, counter, rows):
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})


def test_pmc_summary_scaling(tmp_path):
    """Each kernel's median per-dispatch KiB x 1024 x (calibration bytes /
    what the counter read for the calibration copy); the FP32 call sums
    k_bin, k_sub and the FP32 halo pass, the FP64 call k_bin, k_tile4<1, 0>
    and the FP64 halo pass."""
    cal = 268435456
    copy = "__amd_rocclr_copyBuffer"
    fetch_rows = [(copy, cal / 1024 / 2.0),  # the copy reads half what it moved: factor 2
                  ("void hpgx::k_bin<1, true>(args)", 1000.0),
                  ("void hpgx::k_bin<1, true>(args)", 3000.0),
                  ("void hpgx::k_bin<1, true>(args)", 2000.0),
                  ("void hpgx::k_sub<1, 8, true, true, false>(args)", 500.0),
                  ("void hpgx::k_halo<HIP_vector_type<float, 2u> >(args)", 100.0),
                  ("void hpgx::k_tile4<1, 0, 0>(args)", 700.0),
                  ("void hpgx::k_halo<HIP_vector_type<double, 2u> >(args)", 200.0),
                  ("fft_rtc_back(args)", 9999.0)]
    write_rows = [(copy, cal / 1024),  # factor 1
                  ("void hpgx::k_bin<1, true>(args)", 10.0),
                  ("void hpgx::k_sub<1, 8, true, true, false>(args)", 20.0),
                  ("void hpgx::k_halo<HIP_vector_type<float, 2u> >(args)", 30.0),
                  ("void hpgx::k_tile4<1, 0, 0>(args)", 40.0),
                  ("void hpgx::k_halo<HIP_vector_type<double, 2u> >(args)", 50.0)]
    fd, wd = str(tmp_path / "fetch"), str(tmp_path / "write")
    write_counters(fd, "FETCH_SIZE", fetch_rows)
    write_counters(wd, "WRITE_SIZE", write_rows)
    r = run_bench("--pmc-summary", fd, wd)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout[r.stdout.index("{"):])
    assert abs(res["FETCH_SIZE_factor"] - 2.0) < 1e-12
    assert abs(res["WRITE_SIZE_factor"] - 1.0) < 1e-12
    k = res["kernels"]
    assert "fft_rtc_back" not in k
    assert k["k_bin<1, true>"]["FETCH_SIZE_bytes"] == 2000.0 * 1024 * 2.0  # the median
    fp32c = (2000 + 500 + 100) * 1024 * 2.0 + (10 + 20 + 30) * 1024
    fp64 = (2000 + 700 + 200) * 1024 * 2.0 + (10 + 40 + 50) * 1024
    assert abs(res["per_call_bytes"]["fp32c"] - fp32c) < 1e-6
    assert abs(res["per_call_bytes"]["fp64"] - fp64) < 1e-6
