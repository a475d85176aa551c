"""bench.py's measurement helpers on the CPU: the step-kernel variant of a
dispatch name as rocprofv3 writes it (demangled, or mangled for the FP16
instances), and the traffic record built from counter CSVs (per-dispatch
FETCH_SIZE / WRITE_SIZE in KiB, scaled by what the counters read for a
copy of exactly 4 GiB), checked on synthetic counter files whose values
are chosen so the expected bytes per node are exact."""
import csv
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
import bench  # noqa: E402

HDR = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size", "Kernel_Id",
       "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
       "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def test_step_variant_names():
    assert bench._step_variant("void lbm::step_kernel<float, false, true, true, 1>(lbm::StepArgs)") == \
        ("fp32", False, True, 1, False)
    assert bench._step_variant("void lbm::step_kernel<double, true, true, false, 1>(lbm::StepArgs)") == \
        ("fp64", True, True, 1, False)
    assert bench._step_variant("_ZN3lbmL11step_kernelIDF16_Lb0ELb0ELb1ELi2EEEvNS_8StepArgsE") == \
        ("fp16", False, False, 2, False)
    # with the z-piece flag (the kernel's sixth template argument)
    assert bench._step_variant("void lbm::step_kernel<float, false, true, true, 1, false>(lbm::StepArgs)") == \
        ("fp32", False, True, 1, False)
    assert bench._step_variant("void lbm::step_kernel<double, false, true, true, 1, true>(lbm::StepArgs)") == \
        ("fp64", False, True, 1, True)
    assert bench._step_variant("_ZN3lbmL11step_kernelIDF16_Lb0ELb1ELb1ELi1ELb1EEEvNS_8StepArgsE") == \
        ("fp16", False, True, 1, True)
    assert bench._step_variant("_ZN3lbmL11step_kernelIDF16_Lb0ELb1ELb1ELi2ELb0EEEvNS_8StepArgsE") == \
        ("fp16", False, True, 2, False)
    assert bench._step_variant("lbm::finalize_kernel(lbm::FinalizeArgs)") is None


def write_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HDR)
        for i, (name, grid, wg, counter, value) in enumerate(rows):
            w.writerow([i + 1, i + 1, "Agent 1", 1, 1, 1, grid, 7, name, wg, 0, 0, 64, 0, 96, counter, value, 0, 1])


def test_traffic_summary_from_counter_files(tmp_path):
    nodes = 1536 * 768 * 384
    k32 = "void lbm::step_kernel<float, false, true, true, 1>(lbm::StepArgs)"
    # FETCH_SIZE reads half of the bytes (scale 2), WRITE_SIZE all of them;
    # FP32 reads 80 and writes 76 bytes per node at the chosen shape
    fetch_kib = 80 * nodes / 1024 / 2
    write_kib = 76 * nodes / 1024
    # a trial shape first, the chosen shape (512 blocks x 256 threads) last
    write_csv(tmp_path / "f.csv", [(k32, 1024 * 128, 128, "FETCH_SIZE", 2 * fetch_kib),
                                   (k32, 512 * 256, 256, "FETCH_SIZE", fetch_kib),
                                   (k32, 512 * 256, 256, "FETCH_SIZE", fetch_kib)])
    write_csv(tmp_path / "w.csv", [(k32, 1024 * 128, 128, "WRITE_SIZE", write_kib),
                                   (k32, 512 * 256, 256, "WRITE_SIZE", write_kib),
                                   (k32, 512 * 256, 256, "WRITE_SIZE", write_kib)])
    copy = "void copy_f4(float4 const*, float4*, unsigned long)"
    write_csv(tmp_path / "cf.csv", [(copy, 262144, 256, "FETCH_SIZE", 2 * 1024 * 1024)] * 3)
    write_csv(tmp_path / "cw.csv", [(copy, 262144, 256, "WRITE_SIZE", 4 * 1024 * 1024)] * 3)
    arg = ",".join(str(tmp_path / n) for n in ("f.csv", "w.csv", "cf.csv", "cw.csv"))
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-summary", arg],
                         capture_output=True, text=True, check=True, timeout=120).stdout
    rec = json.loads(out)
    assert rec["fetch_scale"] == 2.0 and rec["write_scale"] == 1.0
    ent = rec["fp32_stats"]
    assert ent["chosen_shape"] == "512x256"
    assert abs(ent["read_B_per_node"] - 80.0) < 1e-3 and abs(ent["write_B_per_node"] - 76.0) < 1e-3
    assert abs(ent["traffic_bytes_per_launch"] - 156 * nodes) < 1e-6 * 156 * nodes
    trial = ent["by_shape"]["1024x128"]
    assert abs(trial["read_B_per_node"] - 160.0) < 1e-3 and trial["dispatches"] == 1


def test_roofline_fields():
    r = {"achieved_gbs": 5000.0, "kernel_ms": 10.0, "bytes_per_node": 152}
    rf = bench.roofline_of(r, 60e9, "step_kernel<float,false,true,true>")
    assert rf["frac"] == 0.625 and rf["frac_of_copy_ceiling"] == round(5000.0 / bench.CEILING_GBS, 4)
    assert rf["traffic_gbs"] == 6000.0


def test_ensemble_traffic_summary_from_counter_files(tmp_path):
    """The multi-case step kernel's record (bench.py --workload
    texture_sweep_64 --traffic-summary): the grid of the last step dispatch,
    bytes per node of the 64-solve pool, and the record reduction's reads."""
    nodes = 231473152  # the seeded 64-request pool (dnsserve_generate, seed 5)
    k = "void lbm::ens_step_kernel<float, true>(lbm::EnsMember const*, lbm::EnsTile const*, int, int, int)"
    rk = "lbm::ens_reduce_kernel(lbm::EnsMember const*, int)"
    fk = 86 * nodes / 1024 / 2
    wk = 78 * nodes / 1024
    write_csv(tmp_path / "f.csv", [(k, 512 * 256, 256, "FETCH_SIZE", 2 * fk), (k, 768 * 256, 256, "FETCH_SIZE", fk),
                                   (rk, 444 * 1024, 1024, "FETCH_SIZE", 2.5 * nodes / 1024 / 2),
                                   (k, 768 * 256, 256, "FETCH_SIZE", fk)])
    write_csv(tmp_path / "w.csv", [(k, 512 * 256, 256, "WRITE_SIZE", wk), (k, 768 * 256, 256, "WRITE_SIZE", wk),
                                   (k, 768 * 256, 256, "WRITE_SIZE", wk)])
    copy = "void copy_f4(float4 const*, float4*, unsigned long)"
    write_csv(tmp_path / "cf.csv", [(copy, 262144, 256, "FETCH_SIZE", 2 * 1024 * 1024)] * 3)
    write_csv(tmp_path / "cw.csv", [(copy, 262144, 256, "WRITE_SIZE", 4 * 1024 * 1024)] * 3)
    arg = ",".join(str(tmp_path / n) for n in ("f.csv", "w.csv", "cf.csv", "cw.csv"))
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "texture_sweep_64",
                          "--traffic-summary", arg], capture_output=True, text=True, check=True, timeout=120).stdout
    rec = json.loads(out)
    assert rec["nodes"] == nodes and rec["fetch_scale"] == 2.0 and rec["write_scale"] == 1.0
    ent = rec["fp32"]
    assert ent["grid_blocks"] == 768 and ent["dispatches"] == 2
    assert abs(ent["read_B_per_node"] - 86.0) < 1e-3 and abs(ent["write_B_per_node"] - 78.0) < 1e-3
    assert abs(ent["reduce_read_B_per_node"] - 2.5) < 1e-3
    assert abs(ent["traffic_bytes_per_launch"] - 164 * nodes) < 1e-6 * 164 * nodes
