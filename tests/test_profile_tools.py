"""The kernel-trace summaries on the CPU: synthetic traces whose per-call
figures follow from their timestamps."""
import csv

from skinny_matmul_amd import profile_summary as ps


def _rows(seq):
    """[(kernel name, duration ns)] -> trace rows back to back."""
    out, t = [], 1000
    for name, ns in seq:
        out.append({"Kernel_Name": f"void skinny::{name}(args)", "Start_Timestamp": str(t),
                    "End_Timestamp": str(t + ns)})
        t += ns + 10
    return out


def test_svqb_window_starts_after_the_marker(capsys):
    # input construction: Gram and TSMM calls that are not replays
    seq = [("tsmttsm_kernel<double, 1, 1, 1, 16, true>", 5000),
           ("tsmm_mfma_kernel<double, 4, 1, true, true, false>", 7000)] * 3
    seq.append(("fill_kernel<2>", 10))  # the marker
    replay = [("tsmttsm_kernel<double, 1, 1, 1, 16, true>", 100),
              ("tsmttsm_reduce_kernel<double>", 20),
              ("svqb_eig_kernel", 300),
              ("tsmm_mfma_kernel<double, 4, 1, true, true, false>", 200)] * 2
    seq += replay * 3
    assert ps.svqb_summary(_rows(seq)) == 0
    out = capsys.readouterr().out
    assert "3 replays" in out
    # per replay: two launches of each kernel, e.g. 2 x 200 ns of TSMM
    assert "| `tsmm_mfma_kernel<double, 4, 1, true, true, false>` | 2 | 0.000 |" in out
    assert "Kernel time per replay: 0.001 ms" in out  # 2 x 620 ns


def test_headline_table_from_trace(tmp_path, capsys, monkeypatch):
    ops = [(op, M) for M in ps.WIDTHS for op in ("tsmttsm", "tsmm")]
    seq = [("some_tuning_kernel", 50)] * 5
    for sweep in range(2):  # one warm-up sweep, one timed
        for op, M in ops:
            if op == "tsmttsm":
                seq.append((f"tsmttsm_kernel<double, {M}>", 1000 * M + sweep))
                if M != 8:  # M = 8: partials summed in the same launch
                    seq.append(("tsmttsm_reduce_kernel<double>", 7))
            else:
                seq.append((f"tsmm_mfma_kernel<double, {M}>", 2000 * M))
    rows = _rows(seq)
    p = tmp_path / "t.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerows(rows)
    monkeypatch.setattr("sys.argv", ["x", str(p), "--steps", "1", "--warmup", "1",
                                     "--logK", "10"])
    assert ps.main() == 0
    out = capsys.readouterr().out
    # timed sweep only: TSMTTSM M=N=64 product 64001 ns plus 7 ns partial sums
    line = next(l for l in out.splitlines() if l.startswith("| tsmttsm M=N=64 "))
    assert "| 0.064 | 0.0000 | 0.064 |" in line
    assert line.split("|")[7].strip() == str(2 * 64 * 1024 * 8)
    # a one-launch call: no partial-sum kernel follows it
    line = next(l for l in out.splitlines() if l.startswith("| tsmttsm M=N=8 "))
    assert "| 0.008 | 0.0000 | 0.008 |" in line
    line = next(l for l in out.splitlines() if l.startswith("| tsmm M=N=4 "))
    assert line.split("|")[7].strip() == str((2 * 4 * 1024 + 16) * 8)
