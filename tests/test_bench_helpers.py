"""CPU checks of bench.py's helpers (no GPU): the saved autotune table
parser that `bench.py --tuning` uses."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_bench_reads_saved_tuning_table(tmp_path):
    """bench.py --tuning re-applies the "best" line of each mode of a saved
    gi_autotune table."""
    p = tmp_path / "t.txt"
    p.write_text("# gi_autotune tables\n\n## mode fp64_batched (tuned in 1 s)\n"
                 "halfls_tile 3 128x128 4.0 ms\nbest halfls_tile=3 corr_tile=1 pipe_rounds=4\n"
                 "\n## mode split (tuned in 1 s)\nsplit_halfls_tile 0 64x64 2.0 ms\n"
                 "best halfls_tile=3 corr_tile=0 pipe_rounds=3 split_tile=1 pipe_share=60\n")
    assert bench.read_tuning(str(p)) == {"fp64_batched": (3, 1, 4, 0, 50),
                                         "split": (3, 0, 3, 1, 60)}
    committed = bench.read_tuning(os.path.join(ROOT, "profiles", "autotune_c2.txt"))
    assert set(committed) == {"fp64_batched", "split"}
