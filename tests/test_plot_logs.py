"""The CSV-log performance view (apps/plot_logs.py) on a log in the format
dg25DBenchmark writes: runs averaged per configuration, Tflops = GFlops/1024."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "distributed-gemm-thesis_amd", "apps", "plot_logs.py")
HEADER = ("Algo,M,N,K,TotalNodes,TotalGPUs,dRow,dCol,StackSize,DecompositionTime,ExecutionTime,GatherTime,GFlops,"
          "DataLocation,CommunicationTime,GemmCallTime")


def load():
    spec = importlib.util.spec_from_file_location("plot_logs", SCRIPT)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def write_log(path):
    rows = [
        "2.5DCANNON,8192,8192,8192,1,4,2,2,1,0.1,0.02,0.1,54000,devices,0,0.02",
        "2.5DCANNON,8192,8192,8192,1,4,2,2,1,0.1,0.03,0.1,56000,devices,0,0.03",
        "2.5DCANNON,16384,16384,16384,1,4,2,2,1,0.4,0.13,0.4,70000,devices,0,0.13",
        "2.5DCANNON,16384,8192,32768,1,16,4,4,2,0.4,0.13,0.4,69000,devices,0,0.13",
    ]
    with open(path, "w") as f:
        f.write(HEADER + "\n" + "\n".join(rows) + "\n")


def test_average_and_tflops(tmp_path):
    pytest.importorskip("pandas")
    mod = load()
    log = tmp_path / "log.csv"
    write_log(log)
    df = mod.with_performance(mod.group_by_average(mod.read_logs([str(log)])))
    assert len(df) == 3
    r = df[(df["M"] == 8192) & (df["TotalGPUs"] == 4)].iloc[0]
    assert r["GFlops"] == 55000.0 and abs(r["ExecutionTime"] - 0.025) < 1e-12
    assert r["Performance (Tflops)"] == 55000.0 / 1024
    sq = mod.square_only(df)
    assert sorted(sq["M"].tolist()) == [8192, 16384]


def test_cli_prints_and_draws(tmp_path, capsys):
    pytest.importorskip("pandas")
    pytest.importorskip("matplotlib")
    mod = load()
    log = tmp_path / "log.csv"
    write_log(log)
    png = tmp_path / "out.png"
    assert mod.main([str(log), "--png", str(png)]) == 0
    out = capsys.readouterr().out
    assert "Performance (Tflops)" in out and "53.710938" in out
    assert png.exists() and png.stat().st_size > 1000
