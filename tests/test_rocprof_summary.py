"""pecs_amd.rocprof_summary on a synthetic counter CSV: the solve windows
between the marker kernels, the per-kernel averages, the copy calibration
and the MFMA figures derived from the raw counters."""
import csv

from pecs_amd import rocprof_summary as rs


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def _rows():
    seq = [("void pecs::(anonymous namespace)::k_rhs_carrier<2, true>(...)", 10.0),
           ("void pecs::(anonymous namespace)::k_schur_rhs<2>(...)", 1.0),
           ("k_sweep_mfma<32, 1, true>(...)", 100.0), ("k_fwd_gather_il<32>(...)", 5.0),
           ("k_sweep_mfma<32, 1, false>(...)", 80.0),
           ("void pecs::(anonymous namespace)::k_schur_recover<2>(...)", 1.0),
           ("void pecs::(anonymous namespace)::k_rhs_carrier<2, true>(...)", 30.0),
           ("void pecs::(anonymous namespace)::k_schur_rhs<2>(...)", 1.0),
           ("k_sweep_mfma<32, 1, true>(...)", 120.0),
           ("void pecs::(anonymous namespace)::k_schur_recover<2>(...)", 1.0),
           ("__amd_rocclr_copyBuffer", 512.0)]
    rows = []
    for d, (n, v) in enumerate(seq):
        rows.append(dict(Dispatch_Id=d, Kernel_Name=n, Counter_Name="SQ_INSTS_VALU_MFMA_MOPS_F64", Counter_Value=v))
        rows.append(dict(Dispatch_Id=d, Kernel_Name=n, Counter_Name="SQ_VALU_MFMA_BUSY_CYCLES", Counter_Value=v * 64))
        rows.append(dict(Dispatch_Id=d, Kernel_Name=n, Counter_Name="GRBM_GUI_ACTIVE", Counter_Value=100.0))
    return rows


def test_windows_kernels_and_derived(tmp_path, capsys):
    p = tmp_path / "c.csv"
    _write(p, _rows())
    rs.pmc(str(p), simds=8)
    out = capsys.readouterr().out
    # windows: (100 + 5 + 80) and (120): averages 152.5 MOPS; GRBM sums 300 and 100
    assert "solve window carrier: SQ_INSTS_VALU_MFMA_MOPS_F64 per solve avg 152.5 over 2 solves" in out
    assert "solve window carrier per solve: FP64 MFMA flops 78080" in out
    # busy (185 + 120) / 2 x 64 = 9760 cycles over 200 x 8 SIMD-cycles
    assert "MFMA busy 610.00% of 8 SIMDs x 200 active cycles" in out
    rs.pmc(str(p), kernel="k_rhs_carrier", simds=8)
    out = capsys.readouterr().out
    assert "kernel k_rhs_carrier: 2 dispatches" in out
    assert "kernel k_rhs_carrier: SQ_INSTS_VALU_MFMA_MOPS_F64 per dispatch avg 20" in out


def test_copy_calibration(tmp_path, capsys):
    p = tmp_path / "c.csv"
    _write(p, [dict(Dispatch_Id=0, Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name="FETCH_SIZE",
                    Counter_Value=524288.0)])
    rs.pmc(str(p), copy_bytes=1073741824)
    out = capsys.readouterr().out
    assert "factor 2.0000 B per counted B" in out
