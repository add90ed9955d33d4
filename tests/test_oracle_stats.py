"""The timing and stats contract's host side: the library's FINE_SECTIONS and
CG_CATS names against the reference's (tests/golden/stats_contract_names.json,
read from src/gpu/timing.rs and src/profiler/mod.rs by
tests/golden/make_stats_names.py), and the restatement in
oracle/stats_oracle.py pinned by formatting cases that decimal arithmetic
fixes, the roll-up and bucketing identities and the reference's key orders.
No GPU."""
import json
import os

import numpy as np
import pytest

import fluidstep as fsm
import frame_oracle as fo
import stats_oracle as so

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "stats_contract_names.json")))


def test_library_names_are_the_references():
    assert fsm.fine_sections() == GOLD["fine_sections"]
    assert fsm.cg_categories() == GOLD["cg_cats"]
    L = fsm.lib()
    assert L.fs_fine_section_name(-1) is None and L.fs_fine_section_name(26) is None
    assert L.fs_cg_category_name(4) is None


def test_rollup_boundaries():
    names = GOLD["fine_sections"]
    assert names.index(so.PREP_END_NAME) == 17 and names.index(so.FINISH_START_NAME) == 19
    prep, pres, fin = so.rollup([1.0] * 26, names, 5.0)
    assert (prep, pres, fin) == (17.0, 7.0, 7.0)
    secs = [float(2 ** i) for i in range(26)]
    prep, pres, fin = so.rollup(secs, names, 0.0)
    assert prep + pres + fin == sum(secs) and prep == 2 ** 17 - 1 and fin == sum(secs[19:])


def test_cg_bucket_follows_the_solver_order():
    # spmv . dot(d.q) . alpha . update . dot(r.r) . beta+dir
    assert GOLD["cg_bucket"] == [0, 1, 3, 2, 1, 3]
    cats = so.bucket([[1, 10, 100, 1000, 10000, 100000], [2, 20, 200, 2000, 20000, 200000]], GOLD["cg_bucket"])
    assert cats == [3, 30 + 30000, 3000, 300 + 300000]


@pytest.mark.parametrize("v,want", [
    (0.0, "0"), (2.0, "2"), (10.0, "10"), (100.0, "100"), (1.5, "1.5"), (16.666666, "16.667"),
    (0.0004, "0"), (0.0006, "0.001"), (0.125, "0.125"), (1234.5678, "1234.568"), (0.1 + 0.2, "0.3"),
    (3.14159, "3.142"), (0.05, "0.05"),
])
def test_fmt_ms(v, want):
    assert so.fmt_ms(v) == want


@pytest.mark.parametrize("v,want", [
    (0.0, "0"), (1.0, "1"), (0.833333, "0.8333"), (0.00004, "0"), (2.5, "2.5"), (0.12346, "0.1235"),
])
def test_fmt_ratio(v, want):
    assert so.fmt_ratio(v) == want


def facts(**kw):
    f = dict(substeps=2, nx=64, ny=32, nz=48, h=0.02, liquid_cells=1234, particles=99999,
             pressure_iterations=40, buffer_bytes=123456789, dispatches_per_substep=77, engine="reference")
    f.update(kw)
    return f


def readout(detailed=True, fused=()):
    return dict(detailed=detailed, substeps=2, prep_ms=1.25, pressure_ms=3.5, finish_ms=0.75,
                sections=[0.01 * i for i in range(26)], cg_cats=[1.0, 0.5, 0.25, 0.125], cg_total_ms=1.875,
                cg_iters=40, fused=list(fused))


def ordered_subsequence(keys, ref):
    it = iter(ref)
    return all(any(k == r for r in it) for k in keys)


def test_stats_json_keys_follow_the_references_order():
    ts = fo.Timestep(1.0 / 120.0, 2)
    ts.steps_for_frame(1.0 / 60.0, 1.0)
    stats = {k: v for k, v in ts.last.items()}
    gpu = so.gpu_object(readout(), GOLD["fine_sections"], GOLD["cg_cats"], 1234)
    text = so.stats_json(facts(), stats, float(ts.dropped), gpu)
    obj = json.loads(text)
    keys = [k for k in obj if k != "engine"]
    assert ordered_subsequence(keys, GOLD["stats_json_keys"]) and keys[-1] == "gpu"
    g = obj["gpu"]
    assert list(g)[:8] == GOLD["gpu_keys"] and list(g)[8:] == ["sections", "cg"]
    assert list(g["sections"]) == GOLD["fine_sections"]
    assert list(g["cg"]) == GOLD["cg_keys"]
    assert obj["timing"] == "gpu-timestamp" and obj["timestep_policy"] == "fixed-drop-cap"
    assert obj["substeps_this_frame"] == 2 and obj["fixed_dt_ms"] == 8.333
    assert obj["dispatches_this_frame"] == 154 and obj["grid_res"] == "64x32x48"
    assert obj["sim_buffers_mb"] == 123.5 and obj["total_cells"] == 64 * 32 * 48
    assert g["sim_ms"] == 5.5 and g["render_ms"] is None and g["cg"]["avg_ms_per_iter"] == 0.047


def test_stats_json_without_timing_or_frame():
    obj = json.loads(so.stats_json(facts(), None, 0.0, None))
    assert obj["gpu"] is None and obj["timing"] == "off" and obj["timestep_policy"] == ""
    assert obj["substeps_this_frame"] == 0 and obj["dispatches_this_frame"] == 0
    coarse = json.loads(so.gpu_object(readout(detailed=False), GOLD["fine_sections"], GOLD["cg_cats"], 7))
    assert list(coarse) == GOLD["gpu_keys"] and coarse["detailed"] is False


def test_fused_readout_marks_categories_null():
    r = readout(fused=["normalize_v", "cg_init"])
    r["cg_cats"] = [None] * 4
    g = json.loads(so.gpu_object(r, GOLD["fine_sections"], GOLD["cg_cats"], 1))
    assert g["cg"]["spmv_ms"] is None and g["cg"]["total_ms"] == 1.875
    assert g["fused_sections"] == ["normalize_v", "cg_init"]


def test_timestep_fields_convert_in_f32():
    F = np.float32
    ts = fo.Timestep(F(1.0) / F(120.0), 1)
    n = ts.steps_for_frame(F(0.05), F(1.0))  # clamped to 1/30 s, capped at 1: drops the rest
    assert n == 1
    got = dict(so.timestep_fields(ts.last, ts.dropped))
    assert got["substep_cap_hit"] == "true"
    assert got["wall_raf_ms"] == so.fmt_ms(F(F(0.05) * F(1000.0))) == "50"
    assert float(got["dropped_sim_time_ms"]) > 0 and got["total_dropped_sim_time_ms"] == got["dropped_sim_time_ms"]
