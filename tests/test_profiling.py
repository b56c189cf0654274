"""The rocprofv3 summaries of fluids_amd.profiling on small hand-made CSVs
whose expected figures are worked out by hand: kernel names, the labelling
of launch shapes, the copy calibration of the traffic summary and the busy
fractions of the pipe summary (one XCD's cycles = GRBM_GUI_ACTIVE / 8)."""
import csv
import os

import pytest

from fluids_amd import profiling as P

HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
        "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
        "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]
WALK = "void sph::k_lists_density<false>(int, float const*)"
TAIL = "void sph::k_update_pressure<false>(int, int)"


def _write(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for k, (name, grid, wg, counter, value, ns) in enumerate(rows):
            w.writerow([k, k, "Agent 1", 1, 1, 1, grid, 1, name, wg, 0, 0, 50, 0, 48, counter, value, 1000, 1000 + ns])


def test_short_name_and_shape_labels():
    assert P.short_name(WALK) == "lists_density"
    assert P.short_name("void sph::k_gather_fluid(int)") == "gather_ranges"
    assert P.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    rows = ([{"Kernel_Name": TAIL, "Grid_Size": "1000192", "Workgroup_Size": "256", "Start_Timestamp": "0",
              "End_Timestamp": "50000"}] * 20
            + [{"Kernel_Name": TAIL, "Grid_Size": "65536", "Workgroup_Size": "256", "Start_Timestamp": "0",
                "End_Timestamp": "2000"}] * 30
            + [{"Kernel_Name": TAIL, "Grid_Size": "4096", "Workgroup_Size": "64", "Start_Timestamp": "0",
                "End_Timestamp": "9000"}] * 3)
    labels = [k for k, _ in P._full_grid(rows, lambda r: P.short_name(r["Kernel_Name"]))]
    # the shape with the most time keeps the name; rare shapes are dropped
    assert labels.count("update_pressure") == 20
    assert labels.count("update_pressure@65536/256") == 30
    assert len(labels) == 50


def test_traffic_summary_calibration(tmp_path):
    # a 1 MiB copy reads FETCH 512 KiB and WRITE 1024 KiB: factors 2 and 1
    _write(tmp_path / "cf", [("__amd_rocclr_copyBuffer", 1024, 256, "FETCH_SIZE", 512.0, 100)] * 3)
    _write(tmp_path / "cw", [("__amd_rocclr_copyBuffer", 1024, 256, "WRITE_SIZE", 1024.0, 100)] * 3)
    _write(tmp_path / "f", [(WALK, 1000000, 64, "FETCH_SIZE", 100.0, 1000)] * 16)
    _write(tmp_path / "w", [(WALK, 1000000, 64, "WRITE_SIZE", 50.0, 1000)] * 16)
    out = P.pmc_summary(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                        1 << 20)
    assert "fetch KiB x 2.000 and write KiB x 1.000" in out
    want = 1024 * (100.0 * 2 + 50.0)
    assert f"traffic lists_density {want:.0f}" in out
    md = tmp_path / "t.md"
    md.write_text(out)
    assert P.read_traffic(str(md), "lists_density") == pytest.approx(want)
    assert P.read_traffic(str(md), "normals") is None


def test_units_summary_fractions(tmp_path):
    ns, cus, xcds = 1000, 256, 8
    cyc = 2000  # one XCD's cycles over a 1000 ns dispatch: 2.00 GHz
    ta = [("GRBM_GUI_ACTIVE", cyc * xcds), ("TA_BUSY_avr", 0.5 * cyc), ("TA_BUSY_max", 0.75 * cyc),
          ("TA_ADDR_STALLED_BY_TC_CYCLES_sum", 0.1 * cyc * cus), ("TD_TD_BUSY_sum", 0.8 * cyc * cus),
          ("TD_TC_STALL_sum", 0.2 * cyc * cus)]
    sq = [("GRBM_GUI_ACTIVE", cyc * xcds), ("SQ_WAVES", 1000), ("SQ_INSTS_VALU", 0.25 * cus * cyc),
          ("SQ_INSTS_VMEM_RD", 30000), ("SQ_INSTS_VMEM_WR", 2000), ("SQ_ACTIVE_INST_VALU", 0.3 * cus * cyc),
          ("SQ_INSTS_SALU", 0), ("SQ_INSTS_LDS", 0), ("SQ_BUSY_CU_CYCLES", 0)]
    tcp = [("GRBM_GUI_ACTIVE", cyc * xcds), ("TCP_TOTAL_CACHE_ACCESSES_sum", 600000), ("TCP_TCC_READ_REQ_sum", 12345),
           ("TCP_PENDING_STALL_CYCLES_sum", 0.4 * cyc * cus), ("TCP_TCP_TA_DATA_STALL_CYCLES_sum", 0)]
    for name, counters in (("ta", ta), ("sq", sq), ("tcp", tcp)):
        assert sorted(c for c, _ in counters) == sorted(P.UNIT_PASSES["ta_td" if name == "ta" else name])
        _write(tmp_path / name, [(WALK, 1000000, 64, c, v, ns) for c, v in counters] * 16)
    out = P.units_summary(str(tmp_path / "ta"), str(tmp_path / "sq"), str(tmp_path / "tcp"), cus, xcds)
    row = [l for l in out.splitlines() if l.startswith("| lists_density ")][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    assert cells[:5] == ["lists_density", "1000000/64", "16", "1.0", "2.00"]
    # TA busy, TA max, TA stalled, TD busy, TD stalled, VALU issue, VALUBusy
    assert cells[5:12] == ["50 %", "75 %", "10 %", "80 %", "20 %", "25 %", "30 %"]
    # VALU / wave, VMEM rd / wave, VMEM wr / wave, tags / read, L1 -> L2 reads, pending stall
    assert cells[12:] == [f"{0.25 * cus * cyc / 1000:.0f}", "30.0", "2.0", "20.0", "12345", "40 %"]
