"""The bench's profile summaries, on the CPU: a kernel trace's last steps kernel by kernel (`bench.py
--summarize-steps`), from a small trace written here in rocprofv3's CSV columns, plain and gzipped."""

import csv
import gzip
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
          "Correlation_Id", "Start_Timestamp", "End_Timestamp"]


def write_trace(path, steps, opener=open):
    """`steps` steps, each: two round launches of 1,000 and 3,000 ns, a walls launch of 500 ns, then k_show; written
    out of time order, as a trace may be."""
    rows, t = [], 0
    for _ in range(steps):
        for name, ns in (("void vr::(anonymous namespace)::k_round_wide<true, false>(vr::SphereDev)", 1000),
                         ("void vr::(anonymous namespace)::k_round_tail<true, false, 512>(vr::SphereDev)", 3000),
                         ("vr::(anonymous namespace)::k_walls(vr::SphereDev)", 500),
                         ("void vr::(anonymous namespace)::k_show<true>(vr::SphereDev)", 100)):
            rows.append({"Kind": "KERNEL_DISPATCH", "Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + ns})
            t += ns + 50
    with opener(path, "wt") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS, restval="0")
        w.writeheader()
        for r in reversed(rows):
            w.writerow(r)


@pytest.mark.parametrize("gz", [False, True])
def test_the_last_steps_are_summarized_kernel_by_kernel(tmp_path, gz):
    trace = str(tmp_path / ("t.csv.gz" if gz else "t.csv"))
    write_trace(trace, 7, gzip.open if gz else open)
    out = str(tmp_path / "w.md")
    bench.summarize_steps(trace, 3, out)
    text = open(out).read()
    assert "The last 3 steps: 12 dispatches, 0.00 ms of kernel time a step." in text
    assert "| `k_round_tail<true, false, 512>` | 3 | 3.00 | 0.01 | 65.22 % |" in text
    assert "| `k_walls` | 3 | 0.50 | 0.00 | 10.87 % |" in text  # 500 of the 4,600 ns a step
    assert "Round kernels (`k_round_wide`, `k_round_tail`): 6 dispatches (2 a step), mean 2.000 us" in text


def test_a_trace_shorter_than_the_window_is_refused(tmp_path):
    trace = str(tmp_path / "t.csv")
    write_trace(trace, 3)
    with pytest.raises(SystemExit):
        bench.summarize_steps(trace, 3, str(tmp_path / "w.md"))
