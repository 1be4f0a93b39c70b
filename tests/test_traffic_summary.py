"""The HBM traffic summary (gsbn_amd.traffic) on synthetic rocprofv3 output:
request counts by size become bytes, each kernel is averaged over its last
`launches` dispatches, and the copy kernel checks the formula."""

import csv
import json

from gsbn_amd import traffic

FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write_run(path, rows):
    path.mkdir(parents=True)
    with open(path / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (kernel, counters) in enumerate(rows):
            for name, value in counters.items():
                w.writerow({"Correlation_Id": i + 1, "Dispatch_Id": i + 1, "Kernel_Name": kernel,
                            "Counter_Name": name, "Counter_Value": value})


COPY = "(anonymous namespace)::copy_kernel(HIP_vector_type<unsigned int, 4u>*, long)"
POP = "(anonymous namespace)::pop_kernel(gsbn_pop_args)"
RC = "void (anonymous namespace)::rowcol_kernel<0>(gsbn_rowcol_args)"


def test_traffic_summary(tmp_path):
    mib = 1
    rd = [(COPY, {"TCC_EA0_RDREQ_32B_sum": 0, "TCC_EA0_RDREQ_64B_sum": 0,
                  "TCC_EA0_RDREQ_128B_sum": (mib << 20) // 128})]
    wr = [(COPY, {"TCC_EA0_WRREQ_sum": (mib << 20) // 64, "TCC_EA0_WRREQ_64B_sum": (mib << 20) // 64})]
    # three pop launches; the window is the last two
    for k, (r32, r64, r128, w, w64) in enumerate([(9, 9, 9, 9, 9), (1, 2, 3, 10, 4), (3, 2, 1, 20, 6)]):
        rd.append((POP, {"TCC_EA0_RDREQ_32B_sum": r32, "TCC_EA0_RDREQ_64B_sum": r64,
                         "TCC_EA0_RDREQ_128B_sum": r128}))
        wr.append((POP, {"TCC_EA0_WRREQ_sum": w, "TCC_EA0_WRREQ_64B_sum": w64}))
        rd.append((RC, {"TCC_EA0_RDREQ_32B_sum": 0, "TCC_EA0_RDREQ_64B_sum": 0,
                        "TCC_EA0_RDREQ_128B_sum": 10}))
        wr.append((RC, {"TCC_EA0_WRREQ_sum": 5, "TCC_EA0_WRREQ_64B_sum": 5}))
    _write_run(tmp_path / "rd", rd)
    _write_run(tmp_path / "wr", wr)
    bench = {"roofline": {"window": "w", "step_plan": "split:2", "deferred_columns": 8,
                          "kernels": {"pop_kernel": {"launches": 2, "bytes_per_launch": 100.0},
                                      "rowcol_kernel": {"launches": 3, "bytes_per_launch": 640.0}}}}
    (tmp_path / "b.json").write_text("noise\n" + json.dumps(bench) + "\n")
    s = traffic.summarize(str(tmp_path / "rd"), str(tmp_path / "wr"),
                          traffic.last_line(str(tmp_path / "b.json")), "100x100", mib)
    assert s["copy_check"]["read_ratio"] == 1.0 and s["copy_check"]["write_ratio"] == 1.0
    pop = s["kernels"]["pop_kernel"]
    # launches 2 and 3: reads (32+128+384 + 96+128+128)/2, writes (4*64+6*32 + 6*64+14*32)/2
    assert pop["read_bytes_per_launch"] == (544 + 352) // 2
    assert pop["write_bytes_per_launch"] == (448 + 832) // 2
    assert pop["plan"] == "any" and pop["launches_averaged"] == 2
    assert abs(pop["ratio"] - (448 + 640) / 100.0) < 1e-12
    rc = s["kernels"]["rowcol_kernel"]
    assert rc["hbm_bytes_per_launch"] == 1280 + 320 and rc["plan"] == "split:2"
    assert rc["ratio"] == 2.5
