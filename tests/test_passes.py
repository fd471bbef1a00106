"""The step's bytes model and the counter summaries built on it (CPU only)."""
import csv

import pytest

from flip3d_amd import passes as P


def _u(n=128, np_=4048638, cf=None):
    u = P.units(n, np_, cf if cf is not None else n ** 3)
    return u


def test_units_follow_grid_size():
    u = P.units(128, 10, 7)
    assert u["C"] == 128 ** 3
    assert u["F"] == 3 * 128 * 128 * 129  # all three face components
    assert (u["Np"], u["Cf"]) == (10, 7)


def test_totals_match_per_unit_table():
    """With every cell fluid the passes of one step add up to 328 Np + 150 C
    + 72 F, plus 117 C per solver iteration (the sort runs twice a step)."""
    for n, np_ in ((32, 59498), (128, 4048638)):
        u = _u(n, np_)
        per_step = 0
        for p in P.PASSES:
            b = p["bytes"](u)
            c = P.COEFFS[p["name"]]
            assert b == c[0] * u["Np"] + c[1] * u["C"] + c[2] * u["F"] + c[3] * u["Cf"], p["name"]
            if p["name"] == "solver_iteration":
                assert b == 117 * u["C"]
                continue
            per_step += b * (2 if p["name"] == "sort" else 1)
        assert per_step == 328 * u["Np"] + 150 * u["C"] + 72 * u["F"]


def test_p2g_credit_at_benchmark_size():
    """P2G with save_grid and the first boundary pass fused in, at N=128."""
    u = _u()
    assert P.pass_bytes("p2g", u) == 24 * 4048638 + 5 * 128 ** 3 + 20 * 3 * 128 * 128 * 129


@pytest.mark.parametrize("sym,want", [
    ("k_p2g_lds<0>(int, HIP_vector_type<float, 4u> const*, int const*, signed char const*, float*, float*, float)",
     "p2g"),
    ("k_p2g_prep_unit", "p2g"), ("k_p2g_lds3", "p2g"), ("k_sort_gather", "sort"), ("k_sort_order_big", "sort"), ("k_fill_i32", "sort"),
    ("k_density<true, true>", "density"), ("k_density_lds<true>", "density"), ("k_flags", "flags"),
    ("k_divergence", "divergence"),
    ("k_pcg_a", "solver_iteration"), ("k_pcg_b<true>", "solver_iteration"), ("k_fcg_b_pre0", "solver_iteration"),
    ("k_gradient<2>", "gradient"), ("k_extrapolate<1>", "extrapolate"), ("k_gradient3", "gradient"),
    ("k_extrapolate3", "extrapolate"), ("k_g2p_flip2", "g2p_flip"),
    ("k_advect", "advect"), ("k_stuck", "advect"), ("k_spring", "corrector"), ("k_spring_fix_wave", "corrector"),
    ("k_jitter_draw", "corrector"), ("k_resample", "corrector"), ("k_correct_update", "corrector"),
    ("k_resample_levels", None), ("k_ml_pre0", None), ("k_conv3_fwd_sk<16, 16, 1, true, 4>", None),
    ("k_pcg_check", None), ("k_scan_apply", None), ("k_pcg_begin", None), ("k_pcg_init<true>", None),
    ("void (anonymous namespace)::k_conv3_fwd_sk<16, 16, 1, true, 4>(unsigned short const*)", None),
    ("(anonymous namespace)::k_sort_gather(long, int const*)", "sort"),
])
def test_kernel_symbols_map_to_passes(sym, want):
    assert P.kernel_pass(sym) == want


def test_step_passes_counts_invocations():
    agg = {"sort_count": (0.1, 2), "sort_scan": (0.05, 2), "sort_scatter": (0.15, 2), "sort_gather": (0.2, 2),
           "p2g_prep": (0.05, 1), "p2g": (0.6, 1), "pcg_a": (0.2, 10), "pcg_b": (0.1, 10),
           "fcg_a": (0.05, 3), "fcg_b": (0.05, 3), "spring": (0.4, 1), "resample": (0.4, 1), "unet": (0.6, 3)}
    u = _u(cf=500000)
    ps = P.step_passes(agg, u)
    assert ps["sort"]["invocations"] == 2 and ps["sort"]["ms"] == pytest.approx(0.5)
    assert ps["solver_iteration"]["invocations"] == 13
    assert ps["p2g"]["avg_ms"] == pytest.approx(0.65)
    want = P.pass_bytes("corrector", u) / 0.8e-3 / 1e9
    assert ps["corrector"]["achieved_gbs"] == pytest.approx(want, abs=0.1)
    assert "unet" not in ps
    assert P.dominant(ps)[0] == "corrector"


def _write_counter_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Kernel_Name", "Counter_Name",
                    "Counter_Value"])
        for d, k, c, v in rows:
            w.writerow([d, d, 1, 1, k, c, v])


def test_traffic_summary_scales_by_copy(tmp_path):
    """Counters are summed over the instances of a dispatch, scaled by the
    copy calibration, and a pass's traffic per invocation is the traffic of
    all its dispatches over its counting kernel's dispatches."""
    copy_kib = 1024 * 1024
    fetch, write = [], []
    for d in range(3):  # copies: FETCH reads half of what was copied, WRITE all of it
        fetch += [(d, "__amd_rocclr_copyBuffer", "FETCH_SIZE", copy_kib / 4)] * 2
        write += [(d, "__amd_rocclr_copyBuffer", "WRITE_SIZE", copy_kib)]
    d = 10
    for _ in range(2):  # two sorts: gather + count kernels
        for k, fk, wk in (("k_sort_count<false>", 100.0, 10.0), ("k_sort_gather", 300.0, 200.0)):
            fetch.append((d, k, "FETCH_SIZE", fk))
            write.append((d, k, "WRITE_SIZE", wk))
            d += 1
    for k in ("k_p2g_prep_unit", "k_p2g_lds<0>", "k_p2g_lds<1>", "k_p2g_lds<2>"):
        fetch.append((d, k, "FETCH_SIZE", 50.0))
        write.append((d, k, "WRITE_SIZE", 25.0))
        d += 1
    fetch.append((d, "k_ml_pre0", "FETCH_SIZE", 1.0))
    write.append((d, "k_ml_pre0", "WRITE_SIZE", 1.0))
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _write_counter_csv(fp, fetch)
    _write_counter_csv(wp, write)
    u = _u(cf=1000)
    s = P.traffic_summary(str(fp), str(wp), copy_kib, u)
    assert s["scale_fetch"] == pytest.approx(2.0) and s["scale_write"] == pytest.approx(1.0)
    assert s["kernels"]["k_sort_gather"]["bytes"] == pytest.approx(1024 * (300 * 2 + 200))
    sort = s["passes"]["sort"]
    assert sort["invocations"] == 2
    assert sort["traffic_per_invocation"] == pytest.approx(1024 * ((100 * 2 + 10) + (300 * 2 + 200)))
    p2g = s["passes"]["p2g"]
    assert p2g["invocations"] == 1 and p2g["traffic_per_invocation"] == pytest.approx(4 * 1024 * (100 + 25))
    assert "k_ml_pre0" in s["kernels"] and s["kernels"]["k_ml_pre0"]["pass"] is None
    text = P.traffic_text(s, "test")
    assert "traffic_per_invocation sort " in text
