"""Writes stats_contract_names.json: the names of the reference's timing and
stats contract, read as text from its sources —
  FINE_SECTIONS, CG_CATS, CG_BUCKET        src/gpu/timing.rs:32-116
  stats_json's top-level keys, in order     src/profiler/mod.rs:540
  the "gpu" object's keys, in order         src/profiler/mod.rs:466
  the "cg" object's keys, in order          src/profiler/mod.rs:453
Run from the repository root: python tests/golden/make_stats_names.py
"""
import json
import os
import re

REF = "/root/reference/app/crates/fluid-lab/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def str_array(text, name):
    m = re.search(r"const " + name + r": \[&str; \d+\] = \[(.*?)\];", text, re.S)
    return re.findall(r'"([^"]+)"', m.group(1))


def int_array(text, name):
    m = re.search(r"const " + name + r": \[usize; [A-Z_]+\] = \[(.*?)\];", text, re.S)
    return [int(x) for x in re.findall(r"\d+", m.group(1))]


def keys_of(fmt):
    return re.findall(r'\\?"([a-z_0-9]+)\\?":', fmt)


def main():
    timing = open(os.path.join(REF, "gpu/timing.rs")).read()
    prof = open(os.path.join(REF, "profiler/mod.rs")).read()
    top = re.search(r'r#"\{\{"timing".*?"#', prof).group(0)
    gpu = re.search(r'r#"\{\{"sim_ms".*?"#', prof).group(0)
    cg = re.search(r'\\"cg\\":\{\{.*?\}\}"', prof).group(0)
    out = {
        "fine_sections": str_array(timing, "FINE_SECTIONS"),
        "cg_cats": str_array(timing, "CG_CATS"),
        "cg_bucket": int_array(timing, "CG_BUCKET"),
        "stats_json_keys": keys_of(top),
        "gpu_keys": keys_of(gpu),
        "cg_keys": keys_of(cg)[1:],
    }
    with open(os.path.join(HERE, "stats_contract_names.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
