"""The CPU-oracle checks that bench.py attaches to its lines (runs without a GPU).

The checker itself is checked here: fed the oracle's own float32 record it reports zero error and
passes; fed a record with one planted mistake (a probe sample off by 1e-3 relative) it fails; its
float64 figure is the float32 / float64 distance of the oracle itself, small but non-zero; a
record or field component at rounding level (below the float64 floor) is listed and kept out of
the relative maxima.  The state comparison finds one planted field error.  The least-work sample
takes cases in ascending work up to its budget, at least one.
"""
import dataclasses

import numpy as np

import bench


def _case(T=60):
    from maxwell_solver_amd import scenarios as S
    sc = S.small_variant(S.config2(T=T), 40, T)
    return dataclasses.replace(sc, snap_every=0)


def _oracle_record(sc, steps, dtype=np.float32):
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    o = OracleSim(dtype)
    S.setup(o, sc)
    o.run(0, steps, sc.mode)
    rec = o.probes().copy()
    o.close()
    return rec


def test_sample_by_work_takes_least_work_first_within_budget():
    assert bench.sample_by_work([5, 1, 3, 2], 6) == [1, 3, 2]
    assert bench.sample_by_work([5, 1, 3, 2], 0) == [1]  # at least one
    assert bench.sample_by_work([7, 9], 100) == [0, 1]
    # one case per class first (whatever its work), then the least work within the budget
    works, groups = [5, 1, 3, 2, 40, 30], ["a", "a", "a", "a", "b", "b"]
    assert bench.sample_by_work(works, 0, groups) == [1, 5]
    assert bench.sample_by_work(works, 34, groups) == [1, 3, 5]
    assert bench.sample_by_work(works, 100, groups) == [1, 3, 2, 0, 5, 4]


def test_size_class_of_the_largest_edge():
    from types import SimpleNamespace as G
    b = (0, 64, 96, 128, 192)
    assert bench.size_class(G(I=32, J=63, K=40), b) == 0
    assert bench.size_class(G(I=64, J=32, K=32), b) == 1
    assert bench.size_class(G(I=32, J=32, K=191), b) == 3
    assert bench.size_class(G(I=268, J=140, K=134), b) == 4


def test_oracle_check_passes_the_oracle_and_fails_a_planted_mistake():
    sc, steps = _case(), 50
    rec = _oracle_record(sc, steps)
    assert np.abs(rec).max() > 0  # the plane wave has reached the probes
    good = bench.oracle_check([("good", sc, steps, rec)])
    assert good["cases"] == 1 and good["within_tol"]
    assert good["probe_rel_l2_f32_max"] == 0.0
    assert 0.0 < good["probe_rel_l2_f64_max"] < 1e-4
    bad = rec.copy()
    q = np.unravel_index(np.argmax(np.abs(bad)), bad.shape)
    bad[q] *= 1.0 + 1e-3 * np.sqrt(bad.size)  # relative L2 of about 1e-3 if all samples were alike
    out = bench.oracle_check([("good", sc, steps, rec), ("planted", sc, steps, bad)])
    assert not out["within_tol"]
    assert out["per_case"][0]["probe_rel_l2_f32"] == 0.0
    assert out["per_case"][1]["probe_rel_l2_f32"] > bench.REL_L2_TOL


def test_records_without_a_field_are_listed_below_the_float64_floor():
    """One step: no field has reached the probes, the float64 record is zero, and its relative L2
    measures nothing -- the case is listed and left out of the float64 maximum."""
    sc = _case()
    empty = _oracle_record(sc, 1)
    assert not np.any(empty)
    alone = bench.oracle_check([("empty", sc, 1, empty)])
    assert alone["within_tol"] and alone["cases_below_f64_floor"] == ["empty"]
    assert alone["probe_rel_l2_f64_max"] is None
    rec = _oracle_record(sc, 50)
    both = bench.oracle_check([("empty", sc, 1, empty), ("field", sc, 50, rec)])
    assert both["cases_below_f64_floor"] == ["empty"]
    assert both["probe_rel_l2_f64_max"] == both["per_case"][1]["probe_rel_l2_f64"] > 0.0
    assert both["per_case"][1]["probe_rms_f64"] >= bench.F64_SIGNAL_FLOOR


def test_state_diff_of_the_oracle_with_itself_and_a_planted_field_error():
    from maxwell_solver_amd import scenarios as S
    from oracle.oracle import OracleSim
    sc = _case()
    o = OracleSim(np.float32)
    S.setup(o, sc)
    o.run(0, 50, sc.mode)
    state = (o.probes().copy(), {c: o.field(c).copy() for c in bench.FIELDS})
    same = bench.state_diff(state, o)
    assert same["probe_rel_l2"] == same["field_rel_l2_max"] == 0.0
    assert same["probe_max_abs"] == same["field_max_abs"] == 0.0
    assert same["field_rms"] > bench.F64_SIGNAL_FLOOR and same["probe_rms"] > 0.0
    hz = state[1]["Hz"]
    q = np.unravel_index(np.argmax(np.abs(hz)), hz.shape)
    hz[q] += 0.5 * abs(hz[q])  # one cell of one component off by half its value
    off = bench.state_diff(state, o)
    assert off["probe_rel_l2"] == 0.0 and off["field_rel_l2_max"] > bench.REL_L2_TOL
    assert off["field_max_abs"] == np.float32(0.5 * abs(o.field("Hz")[q]))
    # a component that is zero in the reference (as one the polarisation never excites) against
    # rounding-sized residue in every cell: listed below the floor, out of the per-component
    # maximum, and counted in the whole-state figure
    o.set_field("Ez", np.zeros_like(o.field("Ez")))
    state[1]["Hz"][q] = o.field("Hz")[q]
    state[1]["Ez"] = np.full_like(state[1]["Ez"], 1e-7)
    quiet = bench.state_diff(state, o)
    assert quiet["field_components_below_floor"] == ["Ez"]
    assert quiet["field_rel_l2_max"] == 0.0
    assert 0.0 < quiet["field_rel_l2"] < 1e-3 and quiet["field_max_abs"] == np.float32(1e-7)
    o.close()


def test_diff_is_relative_l2_and_max_abs():
    a = np.array([1.0, 2.0, 2.0], dtype=np.float32)
    b = np.array([1.0, 2.0, 1.0], dtype=np.float32)
    r, m = bench._diff(a, b)
    assert m == 1.0 and abs(r - 1.0 / np.sqrt(6.0)) < 1e-15
    assert bench._diff(np.zeros(3), np.zeros(3)) == (0.0, 0.0)


def test_run_roofline_is_bytes_over_time_against_the_peak():
    r = bench.run_roofline(4e12, 2.0)
    assert r["achieved"] == 2000.0 and r["frac"] == 0.25 and r["peak"] == bench.HBM_PEAK_GBS
    assert r["bound"] == "hbm" and r["algorithmic_bytes"] == 4_000_000_000_000
