"""bench.py's rocprofv3 summaries (the arithmetic behind profiles/ and the
roofline's traffic and MFMA figures) on small hand-made csv files (CPU).

The figures are fixed by hand: a copy kernel of 512 MiB that reads half its
bytes in FETCH_SIZE gives a fetch scale of 2; a kernel's working launches are
the ones above the thresholds; the SQ notes are the counter ratios the design
notes quote (matrix-core busy = MFMA x cycles per MFMA / (cycles per XCD x
1024 SIMDs), and so on)."""
import csv
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import bench  # noqa: E402

TRACE_HEAD = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
              "Correlation_Id", "Start_Timestamp", "End_Timestamp"]
PMC_HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
            "Kernel_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]

APPLY_F = "void gmk::k_apply<2, float>(gmk::ApplyArgs)"
APPLY_D = "void gmk::k_apply<2, double>(gmk::ApplyArgs)"
UPDATE = "gmk::k_update(gmk::UpdateArgs)"
COPY = "__amd_rocclr_copyBuffer"


def write(path, head, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(head)
        w.writerows(rows)
    return str(path)


def trace(tmp_path):
    rows, t = [], 1000
    # (kernel, duration in ns): two working FP32 applies of 200 us and one
    # early return of 5 us; an update of 80 us
    for i, (k, d) in enumerate([(APPLY_F, 200_000), (APPLY_F, 200_000), (APPLY_F, 5_000), (UPDATE, 80_000),
                                (APPLY_D, 300_000)]):
        rows.append(["KERNEL_DISPATCH", "Agent 1", 1, 0, 1, i + 1, 7, k, i + 1, t, t + d])
        t += d + 1000
    return write(tmp_path / "trace.csv", TRACE_HEAD, rows)


def pmc(tmp_path, name, counter, values):
    rows = [[i + 1, i + 1, "Agent 1", 1, 1, 1, 512, 7, k, counter, v, 0, 1] for i, (k, v) in enumerate(values)]
    return write(tmp_path / name, PMC_HEAD, rows)


def test_trace_summary_counts_working_launches(tmp_path):
    t = bench.trace_launches(trace(tmp_path))
    assert sorted(t["k_apply<2, float>"]) == [5.0, 200.0, 200.0]
    text = bench.summary_trace(trace(tmp_path))
    line = next(l for l in text.splitlines() if l.startswith("k_apply<2, float>"))
    f = line.split()
    # calls, total ms, average us, percent | working, average working us
    assert f[2] == "3" and f[3] == "0.405" and f[4] == "135.0"
    assert f[-2] == "2" and f[-1] == "200.0"


def test_traffic_summary_scales_by_the_copy_calibration(tmp_path):
    moved = (1 << 29) / 1024.0  # KiB each way
    # KiB per dispatch; a working launch fetches more than 1 MiB
    f = pmc(tmp_path, "f.csv", "FETCH_SIZE", [(APPLY_F, 100000.0), (APPLY_F, 100000.0), (APPLY_F, 10.0),
                                              (APPLY_D, 300000.0), (UPDATE, 500000.0)])
    w = pmc(tmp_path, "w.csv", "WRITE_SIZE", [(APPLY_F, 40000.0), (APPLY_F, 40000.0), (APPLY_F, 1.0),
                                              (APPLY_D, 80000.0), (UPDATE, 200000.0)])
    cf = pmc(tmp_path, "cf.csv", "FETCH_SIZE", [(COPY, moved / 2)] * 3)
    cw = pmc(tmp_path, "cw.csv", "WRITE_SIZE", [(COPY, moved)] * 3)
    rows, fs, ws, _, _ = bench._traffic(f, w, cf, cw)
    assert fs == pytest.approx(2.0) and ws == pytest.approx(1.0)
    n, fk, wk, b = rows["k_apply<2, float>"]
    # the 10 KiB early return is not a working launch
    assert n == 2 and fk == 100000.0 and wk == 40000.0
    assert b == pytest.approx((100000.0 * 2 + 40000.0) * 1024)
    text = bench.summary_traffic(f, w, cf, cw)
    assert "traffic_bytes_per_launch k_apply_fp32 %.6g" % ((100000.0 * 2 + 40000.0) * 1024) in text
    assert "traffic_bytes_per_launch k_apply %.6g" % ((300000.0 * 2 + 80000.0) * 1024) in text
    # rates: bytes over the mean working duration of the trace
    rates = bench.summary_rates(trace(tmp_path), f, w, cf, cw)
    line = next(l for l in rates.splitlines() if l.startswith("k_apply<2, float>"))
    gbs = float(line.split()[-2])
    assert gbs == pytest.approx((100000.0 * 2 + 40000.0) * 1024 / 200.0 / 1e3, rel=1e-3)


def test_sq_summary_and_the_roofline_reader(tmp_path, monkeypatch):
    k = APPLY_F
    p1 = [[1, 1, "Agent 1", 1, 1, 1, 512, 7, k, c, v, 0, 1] for c, v in
          (("SQ_WAVES", 8192), ("SQ_INSTS_MFMA", 4.0e6), ("SQ_INSTS_VALU", 3.6e7), ("SQ_INSTS_SALU", 2e7),
           ("SQ_INSTS_VMEM_RD", 5e6), ("SQ_WAVE_CYCLES", 4.8e8), ("SQ_WAIT_INST_ANY", 2.4e8))]
    p2 = [[1, 1, "Agent 1", 1, 1, 1, 512, 7, k, c, v, 0, 1] for c, v in
          (("GRBM_GUI_ACTIVE", 8 * 500000), ("SQ_VALU_MFMA_BUSY_CYCLES", 1.28e8), ("SQ_BUSY_CU_CYCLES", 9e7),
           ("SQ_INSTS_LDS", 2.7e6), ("SQ_INSTS_SMEM", 3.4e6))]
    s1 = write(tmp_path / "s1.csv", PMC_HEAD, p1)
    s2 = write(tmp_path / "s2.csv", PMC_HEAD, p2)
    text = bench.summary_sq(s1, s2, trace(tmp_path))
    note = next(l for l in text.splitlines() if l.startswith("k_apply<2, float>:"))
    # 500000 cycles per XCD; MFMA 4e6 x 32 / (5e5 x 1024) = 25 %; VALU 3.6e7 x 4 / 5.12e8 = 28 %;
    # waves 4.8e8 x 4 / 5.12e8 = 3.8; waits 2.4e8 / 4.8e8 = 50 %
    assert "500000 cycles per XCD" in note and "busy 25%" in note and "instruction 28%" in note
    assert "per SIMD 3.8" in note and "s_waitcnt 50%" in note
    # bench.mfma_macs reads INSTS_MFMA x 1024 from the first table of the committed file
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "k_apply_sq_counters_l20_n16384.txt").write_text(text)
    monkeypatch.setattr(bench, "REPO", str(tmp_path))
    assert bench.mfma_macs("k_apply<2, float>") == pytest.approx(4.0e6 * 1024)


def test_committed_profiles_feed_the_roofline():
    """The committed summaries hold the lines bench.py reads for the roofline."""
    assert bench.measured_traffic("k_tapply") > 0
    assert bench.mfma_macs("k_tapply<20>") > 0
