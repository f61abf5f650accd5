"""The drop-in C++ host API (Summa25Decomposer, drivers, pblasDecomposer,
BlockSequentialDecomposer, GridLayout, validation) through its self-test
program, and the dg25DBenchmark command line with the reference's
cublas25DBenchmark arguments and CSV log."""
import glob
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "distributed-gemm-thesis_amd", "bin")


def run(args, cwd):
    p = subprocess.run(args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(p.stdout)
    return p


def test_host_api_selftest(dev, tmp_path):
    p = run([os.path.join(BIN, "hostapi_selftest")], str(tmp_path))
    assert p.returncode == 0
    assert "PASSED: 0 failure(s)" in p.stdout


@pytest.mark.parametrize("extra", [[], ["fullOffload"]])
def test_cli_csv_log(dev, tmp_path, extra):
    p = run([os.path.join(BIN, "dg25DBenchmark"), "NN", "1024", "1024", "1024", "1.0", "0.0", "2", "2", "1", "3",
             "devices"] + extra, str(tmp_path))
    assert p.returncode == 0
    logs = glob.glob(str(tmp_path / "DGEMM_execution_logs-*-2.5DCANNON-cuBLAS.csv"))
    assert len(logs) == 1
    lines = open(logs[0]).read().splitlines()
    assert lines[0] == ("Algo,M,N,K,TotalNodes,TotalGPUs,dRow,dCol,StackSize,DecompositionTime,ExecutionTime,"
                        "GatherTime,GFlops,DataLocation,CommunicationTime,GemmCallTime")
    assert len(lines) == 4
    for l in lines[1:]:
        f = l.split(",")
        assert f[0] == "2.5D CANNON cuBLAS" and f[1:4] == ["1024"] * 3 and f[5] == "4" and float(f[12]) > 0


def test_cli_usage_error(dev, tmp_path):
    p = run([os.path.join(BIN, "dg25DBenchmark"), "NN", "64"], str(tmp_path))
    assert p.returncode == 1 and "Usage" in p.stdout
