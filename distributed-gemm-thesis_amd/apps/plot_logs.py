"""Performance view of the benchmark CSV logs written by dg25DBenchmark.

The runs of each (Algo, M, N, K, TotalGPUs, DataLocation) are averaged. Each
group's rate is shown as "Performance (Tflops)" = GFlops / 1024, the column
and unit of the reference's plots. That follows plotters/executionAnalysis.ipynb
and plotters/thesisPlotters.ipynb: readLogFile, groupByAverage,
squareDatasetLimits, and dataset['GFlops'] / 1024.

The script prints the averaged table. With --png it also draws Tflops against M,
one panel per device count, with that count's FP64 matrix peak as a dashed
line (78.6 TFLOP/s per MI355X, where the reference drew 19.5 per A100).

    python plot_logs.py LOG.csv [LOG2.csv ...] [--square] [--png out.png]
"""
import argparse
import sys

GROUP_KEYS = ["Algo", "M", "N", "K", "TotalGPUs", "DataLocation"]
PEAK_PER_GPU = 78.6  # TFLOP/s, MI355X dense FP64 matrix


def read_logs(paths):
    """All logs in one frame, sorted by M, N, K (readLogFile)."""
    import pandas as pd
    frames = [pd.read_csv(p, header=0, delimiter=",") for p in paths]
    df = pd.concat(frames, ignore_index=True)
    return df.sort_values(by=["M", "N", "K"]).reset_index(drop=True)


def group_by_average(df):
    """Mean of every numeric column over the runs of one configuration
    (groupByAverage)."""
    return df.groupby(GROUP_KEYS).mean(numeric_only=True).reset_index()


def square_only(df, lower=0, upper=None):
    """M = N = K with lower^3 < MNK < upper^3 (squareDatasetLimits)."""
    mnk = df["M"] * df["N"] * df["K"]
    keep = (df["M"] == df["N"]) & (df["K"] == df["M"]) & (mnk > lower ** 3)
    if upper is not None:
        keep &= mnk < upper ** 3
    return df[keep]


def with_performance(df):
    """Adds "Performance (Tflops)" = GFlops / 1024."""
    out = df.copy()
    out["Performance (Tflops)"] = out["GFlops"] / 1024
    return out


def plot(df, path):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    counts = sorted(df["TotalGPUs"].unique())
    fig, axes = plt.subplots(nrows=len(counts), ncols=1, figsize=(8, 2.2 * len(counts)), squeeze=False)
    for ax, g in zip(axes[:, 0], counts):
        sub = df[df["TotalGPUs"] == g]
        for algo, rows in sub.groupby("Algo"):
            rows = rows.sort_values(by="M")
            ax.plot(rows["M"], rows["Performance (Tflops)"], marker="o", linewidth=2, markersize=5, label=algo)
        ax.axhline(y=PEAK_PER_GPU * g, color="k", linestyle="--", linewidth=1)
        ax.set_ylim([0, PEAK_PER_GPU * g * 1.1])
        ax.set_ylabel("Tflops (%d GPU%s)" % (g, "" if g == 1 else "s"), fontsize=9)
    axes[-1, 0].set_xlabel("Matrix dimension M", fontsize=10)
    axes[0, 0].legend(fontsize=8)
    fig.tight_layout()
    fig.savefig(path, dpi=120)
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("logs", nargs="+", help="CSV logs written by dg25DBenchmark")
    ap.add_argument("--square", action="store_true", help="keep only M = N = K")
    ap.add_argument("--png", help="also draw Tflops against M into this file")
    args = ap.parse_args(argv)
    df = read_logs(args.logs)
    if args.square:
        df = square_only(df)
    df = with_performance(group_by_average(df))
    cols = GROUP_KEYS + ["StackSize", "ExecutionTime", "GFlops", "Performance (Tflops)"]
    print(df[[c for c in cols if c in df.columns]].to_string(index=False))
    if args.png:
        plot(df, args.png)
    return 0


if __name__ == "__main__":
    sys.exit(main())
