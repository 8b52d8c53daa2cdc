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


def _fake_record():
    mode = lambda v: {"value": v, "ms_per_step": 1e3 / v * 1e6, "transitions": 10**6}
    lam = {str(l): {"p50_ms": 1.0, "p99_ms": 4.0} for l in (10, 50, 100, 200, 500)}
    return {
        "metric": bench.METRIC, "value": 2.5e8, "unit": "transitions/s", "n_gpus": 1,
        "steps": 20, "warmup": 5, "ms_per_step": 6.8, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C2v", "mode": "fp64_batched", "transitions": 1742477,
                   "lower_levels": 2968, "parallelism": "1 GPU", "nvib": 2000},
        "split_dtype": "i8 digits x5/x5 radix 254 (split products), f64 elsewhere",
        "roofline": {"bound": "mfma", "achieved": 45.0, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": 0.57, "traffic": 9.2e8, "kernel": "half_ls_mfma",
                     "avg_launch_ms": 1.5, "count": "x" * 120, "measured_copy_GBps": 5000.0,
                     "work": "w"},
        "cpu_baseline": {"value": 2.9e3, "unit": "transitions/s", "cores": 16, "kind": "port",
                         "sample": "s" * 200, "one_thread": {"value": 234}},
        "modes": {"fp64": mode(4e7), "fp64_batched": mode(2.5e8), "split": mode(4.2e8)},
        "kernel_rooflines": {"split": {"half_ls_split": {"frac": 0.3},
                                       "corr_split": {"frac": 0.2}}},
        "parity": {"fp64": {"same_transition_set": True, "amp_within_1e-10": True,
                            "max_rel_S": 4e-11},
                   "split": {"same_transition_set": True, "amp_within_1e-10": False,
                             "within_1e-6": True, "max_rel_S": 1e-7,
                             "max_rel_S_strong": 1.4e-7},
                   "split_1500K": {"same_transition_set": True, "within_1e-6": True,
                                   "max_rel_S_strong": 1.8e-7}},
        "parity_ok": True,
        "multi_j": {"modes": {"fp64": mode(3e7), "fp64_batched": mode(2.7e8),
                              "split": mode(4.2e8)},
                    "parity": {"sample": "x", "fp64": {"same_transition_set": True,
                                                       "amp_within_1e-10": True}}},
        "c3v_multi_j": {"modes": {"fp64_batched": mode(1.8e7), "split": mode(3e7)}},
        "serving": {"workload": "w" * 300, "fp64_batched": lam, "split": lam},
        "full_record": "gpurun_out/bench_full.json",
    }


def test_compact_line_is_small_and_carries_the_contract():
    """The last stdout line stays well under the few KB of stdout the driver
    keeps, and carries the contract fields, roofline and CPU baseline."""
    import json
    line = bench.compact_line(_fake_record())
    s = json.dumps(line)
    assert len(s) < 4000, len(s)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in line
    assert line["dtype"] == "f64"
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in line["roofline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in line["cpu_baseline"]
    assert line["split"]["vs_fp64_plain"] == 4.2e8 / 4e7
    assert line["split"]["max_rel_S_1500K"] == 1.8e-7


def test_parity_gate():
    """A failed sampled check (transition set, FP64 amplitude tolerance, or
    the split bar) makes parity_ok false; the split mode's FP64 amplitude flag
    is not part of its bar."""
    rec = _fake_record()
    assert bench.parity_ok(rec["parity"])
    assert bench.parity_ok(None)
    bad = dict(rec["parity"], fp64=dict(rec["parity"]["fp64"], **{"amp_within_1e-10": False}))
    assert not bench.parity_ok(bad)
    bad = dict(rec["parity"], split=dict(rec["parity"]["split"], **{"within_1e-6": False}))
    assert not bench.parity_ok(bad)
    bad = dict(rec["parity"],
               split_1500K=dict(rec["parity"]["split_1500K"], same_transition_set=False))
    assert not bench.parity_ok(bad)
