"""Counter-run post-processing: calibration and the per-launch byte table."""
import os

from xtb_amd import profiling

HEADER = ('"Correlation_Id","Dispatch_Id","Agent_Id","Queue_Id","Process_Id","Thread_Id","Grid_Size",'
          '"Kernel_Id","Kernel_Name","Workgroup_Size","LDS_Block_Size","Scratch_Size","VGPR_Count",'
          '"Accum_VGPR_Count","SGPR_Count","Counter_Name","Counter_Value","Start_Timestamp","End_Timestamp"\n')


def _write(path, counter, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w") as f:
        f.write(HEADER)
        for i, (name, value, t0, t1) in enumerate(rows):
            f.write(f'{i},{i},"Agent 1",1,1,1,64,1,"{name}",64,0,0,4,0,16,"{counter}",{value},{t0},{t1}\n')


def test_short_names():
    assert profiling.kernel_short_name(
        "xtbamd::(anonymous namespace)::k_copy(double const*, double*, long long)") == "k_copy"
    assert profiling.kernel_short_name("void xtbamd::k_jac_pairsP<32, 64>(JacobiProb const*, int)") == "k_jac_pairsP"


def test_calibrated_table_roundtrip(tmp_path):
    copy = "xtbamd::(anonymous namespace)::k_copy(double const*, double*, long long)"
    kern = "xtbamd::k_work(double*)"
    # the fetch counter reads half the true bytes, the write counter reads them as is
    _write(tmp_path / "F", "FETCH_SIZE", [(copy, 512.0, 0, 10), (kern, 100.0, 20, 30), (kern, 300.0, 40, 50)])
    _write(tmp_path / "W", "WRITE_SIZE", [(copy, 1024.0, 0, 10), (kern, 50.0, 20, 30), (kern, 50.0, 40, 50)])
    rows, fs, ws = profiling.traffic(str(tmp_path / "F"), str(tmp_path / "W"), 1.0)
    assert abs(fs - 2.0) < 1e-12 and abs(ws - 1.0) < 1e-12
    work = {r["kernel"]: r for r in rows}["k_work"]
    assert work["launches"] == 2
    assert abs(work["bytes"] - (200.0 * 2.0 + 50.0) * 1024.0) < 1e-6
    md = tmp_path / "t.md"
    md.write_text(profiling.markdown(rows, fs, ws, 1.0))
    table = profiling.traffic_table(str(md))
    assert abs(table["k_work"] - work["bytes"]) <= 1e-5 * work["bytes"]
    assert abs(table["k_copy"] - 2 * 1024.0 * 1024.0) <= 1e-5 * 2 * 1024.0 * 1024.0


def test_committed_table_has_dominant_kernels():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    table = profiling.traffic_table(os.path.join(root, "profiles", "batched512_hbm_traffic.md"))
    assert table["k_jac_pairsR"] > 0 and table["k_jac_apply_R"] > 0 and table["k_gemm_mfma"] > 0
