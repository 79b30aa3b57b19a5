"""bench.py's bookkeeping on the host: which launches the roofline averages
cover, and how profiler rows map to iterations (no GPU)."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def test_algorithmic_bytes_match_the_gather_model():
    B = bench.algorithmic_bytes(np.array([10]), np.array([30]), np.array([20]), S=2, P=4)
    assert B["update"][0] == (38 + 2) * 10      # coord 4, 7 taps 28, source 2, phi 4, flag 1, mask 1
    assert B["mark"][0] == 33 * 10               # coord 4, flag 1, 7 taps 28
    assert B["keep"][0] + B["compact"][0] == 10 * 30 + 5 * 20


def test_launch_parts_cover_non_empty_iterations_only():
    # 5 launched iterations, the last two over an empty list
    n = np.array([100, 200, 300, 0, 0])
    c = np.array([150, 250, 350, 0, 0])
    m = np.array([200, 300, 0, 0, 0])
    L = 5
    per = np.zeros(3 * L + 1)
    per[0:3 * L:3] = [1.0, 2.0, 3.0, 0.1, 0.1]      # update / fused
    per[1:3 * L:3] = [0.5, 0.6, 0.7, 0.05, 0.05]     # mark
    per[2:3 * L:3] = [0.4, 0.4, 0.4, 0.04, 0.04]     # keep
    parts, got_L = bench.launch_parts(per, n, c, m, S=1, P=4)
    assert got_L == L
    uf = parts["update_fused"]
    assert uf["launches"] == 2                      # iterations 1 and 2
    want_bytes = ((39 * 200 + 5 * 150 + 5 * 200) + (39 * 300 + 5 * 250 + 5 * 300)) / 2
    assert abs(uf["bytes_per_launch"] - want_bytes) < 1e-6
    assert abs(uf["avg_launch_us"] - 2500.0) < 1e-9
    assert parts["mark"]["launches"] == 3 and abs(parts["mark"]["avg_launch_us"] - 600.0) < 1e-9


def test_profiler_rows_map_to_iterations():
    names = ["lsseg::k_set_phi<float>", "void lsseg::k_compact<3>(lsseg::KArgs, int)",
             "void lsseg::k_update<unsigned short, false, float>(lsseg::KArgs, int)",
             "void lsseg::k_mark<float>(lsseg::KArgs, int)", "void lsseg::k_keep<1>(lsseg::KArgs, int)"]
    for _ in range(3):
        names += ["void lsseg::k_update_fused<unsigned short, false, float>(lsseg::KArgs, int)",
                  "void lsseg::k_mark<float>(lsseg::KArgs, int)", "void lsseg::k_keep<1>(lsseg::KArgs, int)"]
    names.append("void lsseg::k_compact<2>(lsseg::KArgs, int)")
    rows = [{"Kernel_Name": k, "Dispatch_Id": str(i)} for i, k in enumerate(names)]
    rows = bench._dispatch_order(rows[::-1])
    got = bench.iteration_dispatches(rows, np.array([5, 6, 0, 0]))
    assert [i for i, _ in got["k_update_fused"]] == [1]
    assert [i for i, _ in got["k_mark"]] == [0, 1]
    assert [i for i, _ in got["k_keep"]] == [0, 1]
    # an earlier run in the same process (the autotuner's) is skipped: the
    # profiled run starts at the last k_update
    earlier = names[2:5] + names[5:8] * 2 + [names[-1]]
    rows2 = [{"Kernel_Name": k, "Dispatch_Id": str(i)} for i, k in enumerate(names[:2] + earlier + names[2:])]
    got2 = bench.iteration_dispatches(bench._dispatch_order(rows2), np.array([5, 6, 0, 0]))
    for k in got:
        assert [(i, r["Kernel_Name"]) for i, r in got2[k]] == [(i, r["Kernel_Name"]) for i, r in got[k]]
        assert all(int(r["Dispatch_Id"]) > len(earlier) + 1 for _, r in got2[k])


def test_profiler_rows_of_the_deferred_keep_step_map_to_iterations():
    """The deferred keep step's kernels (k_update_fused_defer, k_keep_defer)
    count as the fused update and keep launches; the traversal rebuild at the
    end of a run (k_settle_traversal) is no iteration launch."""
    names = ["void lsseg::k_update<float, true, float>(lsseg::KArgs, int)",
             "void lsseg::k_mark<float>(lsseg::KArgs, int)", "void lsseg::k_keep_defer<1>(lsseg::KArgs, int)"]
    for _ in range(2):
        names += ["void lsseg::k_update_fused_defer<float, true, float>(lsseg::KArgs, int)",
                  "void lsseg::k_mark<float>(lsseg::KArgs, int)", "void lsseg::k_keep_defer<1>(lsseg::KArgs, int)"]
    names += ["void lsseg::k_compact<2>(lsseg::KArgs, int)", "lsseg::k_settle_traversal(lsseg::KArgs, int)"]
    rows = [{"Kernel_Name": k, "Dispatch_Id": str(i)} for i, k in enumerate(names)]
    got = bench.iteration_dispatches(bench._dispatch_order(rows), np.array([4, 3, 2]))
    assert [i for i, _ in got["k_update_fused"]] == [1, 2]
    assert [i for i, _ in got["k_mark"]] == [0, 1, 2]
    assert [i for i, _ in got["k_keep"]] == [0, 1, 2]
    assert bench.SUMMARY_KEY_OF_PART["update_fused"] == "k_update_fused"
