"""Restatement of fluid-lab's timing and stats contract: the detailed
readout's roll-ups and CG bucketing (src/gpu/timing.rs:105-116, 440-481) and
the simulation's part of Profiler::stats_json with its number formatting
(src/profiler/mod.rs:172-193, 416-612).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
product's fs_get_detailed_timing / fs_stats_json; the product library never
imports it.

Pins (tests/test_oracle_stats.py): the names and key orders in
tests/golden/stats_contract_names.json (read from the reference's sources by
tests/golden/make_stats_names.py), formatting cases that decimal arithmetic
fixes, and the roll-up identities.
"""
from __future__ import annotations

import json

import numpy as np

F = np.float32
PREP_END_NAME, FINISH_START_NAME = "divergence", "gradient_u"  # src/gpu/timing.rs:69-70
POLICY_LABELS = {0: "fixed-drop-cap", 1: "fixed-explicit-export"}  # src/timestep.rs:18, 193


def _fixed(v, decimals):
    if v == 0.0:
        return "0"
    s = "%.*f" % (decimals, v)
    return s.rstrip("0").rstrip(".")


def fmt_ms(v):
    """fmt_ms (src/profiler/mod.rs:594-602): 3 decimals, trailing zeros and
    point trimmed, 0 as "0"."""
    return _fixed(float(v), 3)


def fmt_ratio(v):
    """fmt_ratio (src/profiler/mod.rs:604-612): the same at 4 decimals."""
    return _fixed(float(v), 4)


def bucket(pass_ms, cg_bucket):
    """CG_BUCKET (src/gpu/timing.rs:116): per-iteration lists of six timed
    pass times summed into the four categories."""
    cats = [0.0] * 4
    for it in pass_ms:
        for k, ms in enumerate(it):
            cats[cg_bucket[k]] += ms
    return cats


def rollup(sections, names, cg_total):
    """prep / pressure / finish from the fine sections (src/gpu/timing.rs:472-481)."""
    pe, fs = names.index(PREP_END_NAME), names.index(FINISH_START_NAME)
    return sum(sections[:pe]), sum(sections[pe:fs]) + cg_total, sum(sections[fs:])


def gpu_object(readout, names, cg_names, liquid_cells):
    """The "gpu" value (src/profiler/mod.rs:433-475) for a readout dict:
    detailed, prep_ms, pressure_ms, finish_ms, substeps and, when detailed,
    sections (list by FINE_SECTIONS), cg_cats (4, None when not split),
    cg_total_ms, cg_iters, fused (names, fused engine only)."""
    prep, pres, fin = readout["prep_ms"], readout["pressure_ms"], readout["finish_ms"]
    head = ('{"sim_ms":%s,"prep_ms":%s,"pressure_ms":%s,"finish_ms":%s,"render_ms":null,'
            '"liquid_cells":%d,"substeps":%d,"detailed":%s'
            % (fmt_ms(prep + pres + fin), fmt_ms(prep), fmt_ms(pres), fmt_ms(fin), liquid_cells,
               readout["substeps"], "true" if readout["detailed"] else "false"))
    if not readout["detailed"]:
        return head + "}"
    secs = ",".join('"%s":%s' % (n, fmt_ms(v)) for n, v in zip(names, readout["sections"]))
    total = F(readout["cg_total_ms"])
    it = readout["cg_iters"]
    avg = fmt_ms(F(total / F(it))) if it > 0 else "null"
    cats = [("null" if c is None else fmt_ms(c)) for c in readout["cg_cats"]]
    cg = ('"cg":{"total_ms":%s,"avg_ms_per_iter":%s,"spmv_ms":%s,"reductions_ms":%s,"updates_ms":%s,'
          '"scalars_ms":%s,"iters":%d}' % (fmt_ms(total), avg, cats[0], cats[1], cats[2], cats[3], it))
    out = head + ',"sections":{' + secs + "}," + cg
    if readout.get("fused"):
        out += ',"fused_sections":[' + ",".join('"%s"' % n for n in readout["fused"]) + "]"
    return out + "}"


def timestep_fields(ts, total_dropped_s):
    """set_timestep_stats (src/profiler/mod.rs:172-193) then the format
    (:572-585); ts a TimestepStats dict or None (the profiler's defaults)."""
    if ts is None:
        z = F(0.0)
        ts = dict(substeps=0, fixed_dt=z, max_substeps=0, natural_substeps=0, substep_cap_hit=0,
                  sim_advanced=z, wall_dt=z, sim_time_scale=z, real_time_factor=z, accumulated_before=z,
                  accumulated_after=z, dropped_this_frame=z, policy=None)
        total_dropped_s = z
    ms = lambda v: fmt_ms(F(F(v) * F(1000.0)))  # noqa: E731
    return [
        ("substeps_this_frame", str(ts["substeps"])),
        ("fixed_dt_ms", ms(ts["fixed_dt"])),
        ("max_substeps", str(ts["max_substeps"])),
        ("natural_substeps", str(ts["natural_substeps"])),
        ("substep_cap_hit", "true" if ts["substep_cap_hit"] else "false"),
        ("sim_advanced_ms", ms(ts["sim_advanced"])),
        ("wall_raf_ms", ms(ts["wall_dt"])),
        ("sim_time_scale", fmt_ratio(F(ts["sim_time_scale"]))),
        ("real_time_factor", fmt_ratio(F(ts["real_time_factor"]))),
        ("timestep_policy", json.dumps(POLICY_LABELS.get(ts["policy"], ""))),
        ("accumulated_before_ms", ms(ts["accumulated_before"])),
        ("accumulated_after_ms", ms(ts["accumulated_after"])),
        ("dropped_sim_time_ms", ms(ts["dropped_this_frame"])),
        ("total_dropped_sim_time_ms", ms(total_dropped_s)),
    ]


def stats_json(facts, ts, total_dropped_s, gpu):
    """The object text (src/profiler/mod.rs:539-591) restricted to the
    simulation's fields, in the reference's key order, with "engine" before
    "gpu".  facts: timing_on, substeps, nx, ny, nz, h, liquid_cells,
    particles, pressure_iterations, buffer_bytes, dispatches_per_substep,
    engine; gpu: the "gpu" value text or None."""
    f = facts
    cells = f["nx"] * f["ny"] * f["nz"]
    mb = "%.1f" % (f["buffer_bytes"] / 1.0e6)
    parts = [
        ("timing", json.dumps("gpu-timestamp" if gpu is not None else "off")),
        ("substeps", str(f["substeps"])),
        ("grid_n", str(f["nx"])),
        ("grid_res", json.dumps("%dx%dx%d" % (f["nx"], f["ny"], f["nz"]))),
        ("total_cells", str(cells)),
        ("filled_volume", "%.6f" % (f["liquid_cells"] * (float(f["h"]) * float(f["h"]) * float(f["h"])))),
        ("liquid_fraction", "%.6f" % (f["liquid_cells"] / cells if cells else 0.0)),
        ("particles", str(f["particles"])),
        ("pressure_iterations", str(f["pressure_iterations"])),
        ("gpu_buffer_mb", mb), ("sim_buffers_mb", mb), ("total_tracked_mb", mb),
    ]
    parts += timestep_fields(ts, total_dropped_s)
    dps = f["dispatches_per_substep"]
    subs = ts["substeps"] if ts is not None else 0
    parts += [("dispatches_per_substep", str(dps)), ("dispatches_this_frame", str(dps * subs)),
              ("engine", json.dumps(f["engine"])), ("gpu", gpu if gpu is not None else "null")]
    return "{" + ",".join('"%s":%s' % kv for kv in parts) + "}"
