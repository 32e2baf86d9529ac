"""bench.py on a machine without a GPU: its command-line contract (it
measures one GPU only, so any other GPU count exits with status 2 and a
message that says so, before touching a device) and the parsing of the
rocprofv3 counter files its counter passes write."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_contract_flags():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_counter_files_give_the_stage_means_and_the_copy_scale(tmp_path):
    """the counter passes' CSV parsing: rows of one dispatch are summed, the
    stage kernel's dispatches averaged, other kernels ignored, the largest
    copy taken as the calibration copy; traffic = FETCH x (copy bytes / copy
    FETCH) + WRITE x (copy bytes / copy WRITE)"""
    b = _bench_module()
    stage = "e3_stage_direct<5, 1, 1, double, false, 1, double>"
    head = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
    rows = [
        (1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 10.0),
        (2, "__amd_rocclr_copyBuffer", "FETCH_SIZE", b.CALIBRATION_COPY_BYTES / 2048.0),
        (3, f"void atmo::{stage}(atmo::E3Params)", "FETCH_SIZE", 100.0),
        (3, f"void atmo::{stage}(atmo::E3Params)", "FETCH_SIZE", 20.0),  # second row, same dispatch
        (4, f"void atmo::{stage}(atmo::E3Params)", "FETCH_SIZE", 140.0),
        (5, "void atmo::e3_stage_direct<5, 1, 1, double, false, 2, double>(x)", "FETCH_SIZE", 1e9),
        (6, "__amd_rocclr_copyBuffer", "WRITE_SIZE", b.CALIBRATION_COPY_BYTES / 1024.0),
        (7, f"void atmo::{stage}(atmo::E3Params)", "WRITE_SIZE", 50.0),
        (8, f"void atmo::{stage}(atmo::E3Params)", "SQ_INSTS_VALU_FMA_F64", 10.0),
        (8, f"void atmo::{stage}(atmo::E3Params)", "SQ_INSTS_VALU_MUL_F64", 1.0),
        (8, f"void atmo::{stage}(atmo::E3Params)", "SQ_INSTS_VALU_ADD_F64", 2.0),
        (8, f"void atmo::{stage}(atmo::E3Params)", "SQ_INSTS_VALU_TRANS_F64", 3.0),
    ]
    f = tmp_path / "run_counter_collection.csv"
    f.write_text(head + "".join(f'{d},"{k}",{c},{v}\n' for d, k, c, v in rows))
    values = {}
    for counters in b.COUNTER_PASSES:
        got, why = b.counter_means([str(f)], stage, counters)
        assert got is not None, why
        values.update(got)
    assert values["FETCH_SIZE"] == 130.0 and values["FETCH_SIZE_dispatches"] == 2
    assert values["copy_FETCH_SIZE"] == b.CALIBRATION_COPY_BYTES / 2048.0
    s = b.counter_summary(values, stage, nodes=4)
    assert s["read_factor"] == 2.0 and s["write_factor"] == 1.0
    assert s["traffic"] == (130.0 * 2.0 + 50.0) * 1024.0
    assert s["exec_fp64_flops_per_node"] == (64.0 * 6.0 + 128.0 * 10.0) / 4
    missing, why = b.counter_means([str(f)], stage.replace("5,", "6,"), ("FETCH_SIZE",))
    assert missing is None and "no FETCH_SIZE" in why


def test_counter_files_without_the_expected_columns_are_an_error_not_a_crash(tmp_path):
    """a counter file of another layout makes counter_means raise, which
    bench.py's main catches (the line then falls back to the committed
    profile and says so)"""
    import pytest
    b = _bench_module()
    f = tmp_path / "run_counter_collection.csv"
    f.write_text("Kernel,Counter,Value\nx,FETCH_SIZE,1\n")
    with pytest.raises(KeyError):
        b.counter_means([str(f)], "e3_stage_direct<5, 1, 1, double, false, 1, double>",
                        ("FETCH_SIZE",))
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert "except Exception as exc" in src and "counter passes failed" in src
