// fs_stats.cpp — the timing and stats contract: the reference's detailed
// section and CG category names (src/gpu/timing.rs:32-116), the totals the
// step's boundary events collect, and the simulation's part of
// Profiler::stats_json (src/profiler/mod.rs:416-602).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "fluidstep.h"
#include "fs_host_common.h"
#include "fs_tank.h"

using namespace fs;

namespace {

// FINE_SECTIONS (src/gpu/timing.rs:32-59).
const char* const kSectionNames[FS_FINE_SECTIONS] = {
    "clear",       "mark",        "sort",        "classify",     "scatter",      "normalize_u",  "normalize_v",
    "normalize_w", "savevel_u",   "savevel_v",   "savevel_w",    "forces_u",     "forces_v",     "forces_w",
    "bound_pre_u", "bound_pre_v", "bound_pre_w", "divergence",   "cg_init",      "gradient_u",   "gradient_v",
    "gradient_w",  "bound_post_u", "bound_post_v", "bound_post_w", "g2p"};
// CG_CATS (src/gpu/timing.rs:105).
const char* const kCgNames[FS_CG_CATEGORIES] = {"cg_spmv", "cg_reduce", "cg_update", "cg_scalars"};
// PREP_END and FINISH_START (src/gpu/timing.rs:69-70).
constexpr int kPrepEnd = 17, kFinishStart = 19;

// Sections whose work the fused engine does inside an earlier section's
// kernel: the face epilogue (normalize_u) normalizes, saves, forces and
// enforces all three axes; div_cginit (divergence) also initialises the CG;
// gradient_enforce (gradient_u) does every axis and both enforcements.
bool fused_section(int i) { return (i >= 6 && i <= 16) || i == 18 || (i >= 20 && i <= 24); }

// fmt_ms / fmt_ratio (src/profiler/mod.rs:594-612): fixed decimals, trailing
// zeros and a trailing point trimmed, exactly 0 as "0".  printf rounds the
// exact binary value to nearest (ties to even), as Rust's formatter does.
std::string fmt_fixed(double v, int decimals) {
    if (!std::isfinite(v)) return "null";
    if (v == 0.0) return "0";
    char b[64];
    std::snprintf(b, sizeof b, "%.*f", decimals, v);
    std::string s = b;
    while (!s.empty() && s.back() == '0') s.pop_back();
    if (!s.empty() && s.back() == '.') s.pop_back();
    return s;
}
std::string fmt_ms(double v) { return fmt_fixed(v, 3); }
std::string fmt_ratio(double v) { return fmt_fixed(v, 4); }
std::string fmt_mb(uint64_t bytes) {
    char b[64];
    std::snprintf(b, sizeof b, "%.1f", (double)bytes / 1.0e6);
    return b;
}

void fill_detailed(const fs_tank* t, fs_detailed_timing& d) {
    std::memset(&d, 0, sizeof d);
    d.detailed = (uint32_t)t->detailed;
    d.substeps = t->substeps_timed;
    d.cg_iters = t->cg_iters_timed;
    const bool split = t->prm.engine == FS_ENGINE_REFERENCE;
    d.cg_split = split ? 1u : 0u;
    for (int i = 0; i < FS_FINE_SECTIONS; ++i) {
        d.section_ms[i] = t->sec_ms[i];
        d.fused[i] = (!split && fused_section(i)) ? 1u : 0u;
    }
    for (int c = 0; c < FS_CG_CATEGORIES; ++c) d.cg_ms[c] = split ? t->cg_cat_ms[c] : NAN;
    d.cg_total_ms = t->cg_total_ms;
    for (int i = 0; i < kPrepEnd; ++i) d.prep_ms += d.section_ms[i];
    for (int i = kPrepEnd; i < kFinishStart; ++i) d.pressure_ms += d.section_ms[i];
    d.pressure_ms += d.cg_total_ms;
    for (int i = kFinishStart; i < FS_FINE_SECTIONS; ++i) d.finish_ms += d.section_ms[i];
}

const char* policy_label(uint32_t p) {
    return p == FS_TIMESTEP_FIXED_EXPLICIT_EXPORT ? "fixed-explicit-export" : "fixed-drop-cap";
}

}  // namespace

extern "C" {

FS_API int fs_set_detailed_timing(fs_tank* t, int enabled) {
    if (!t) return fail(FS_ERR_INVALID_ARGUMENT, "null tank handle");
    t->detailed = enabled ? 1 : 0;
    return FS_OK;
}

FS_API int fs_get_detailed_timing(fs_tank* t, fs_detailed_timing* out) {
    if (!t || !out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    FS_HIP(hipStreamSynchronize(t->stream));
    fill_detailed(t, *out);
    return FS_OK;
}

FS_API const char* fs_fine_section_name(int i) {
    return i >= 0 && i < FS_FINE_SECTIONS ? kSectionNames[i] : nullptr;
}
FS_API const char* fs_cg_category_name(int i) { return i >= 0 && i < FS_CG_CATEGORIES ? kCgNames[i] : nullptr; }

FS_API uint32_t fs_dispatches_per_substep(const fs_tank* t) { return t ? t->dispatches_last : 0u; }

FS_API int fs_stats_json(fs_tank* t, const fs_frame* f, char* buf, size_t cap, size_t* len_out) {
    if (!t || !len_out) return fail(FS_ERR_INVALID_ARGUMENT, "null argument");
    fs_stats st;
    if (int rc = fs_get_stats(t, &st)) return rc;
    fs_detailed_timing d;
    fill_detailed(t, d);
    const fs_params& q = t->prm;
    const uint64_t total_cells = (uint64_t)q.nx * q.ny * q.nz;

    // GpuSample -> "gpu" (src/profiler/mod.rs:433-475); every time a total
    // over the substeps of the last fs_step call.
    std::string gpu = "null";
    if (t->detailed || t->timing) {
        double prep, pres, fin;
        uint32_t subs;
        std::string detail;
        if (t->detailed) {
            prep = d.prep_ms, pres = d.pressure_ms, fin = d.finish_ms, subs = d.substeps;
            detail = ",\"sections\":{";
            for (int i = 0; i < FS_FINE_SECTIONS; ++i) {
                if (i) detail += ',';
                detail += std::string("\"") + kSectionNames[i] + "\":" + fmt_ms(d.section_ms[i]);
            }
            detail += '}';
            const float cg_total = (float)d.cg_total_ms;
            const std::string avg = d.cg_iters > 0 ? fmt_ms(cg_total / (float)d.cg_iters) : "null";
            detail += ",\"cg\":{\"total_ms\":" + fmt_ms(cg_total) + ",\"avg_ms_per_iter\":" + avg +
                      ",\"spmv_ms\":" + fmt_ms(d.cg_ms[0]) + ",\"reductions_ms\":" + fmt_ms(d.cg_ms[1]) +
                      ",\"updates_ms\":" + fmt_ms(d.cg_ms[2]) + ",\"scalars_ms\":" + fmt_ms(d.cg_ms[3]) +
                      ",\"iters\":" + std::to_string(d.cg_iters) + "}";
            if (!d.cg_split) {
                // Not in the reference's object: the sections the fused
                // kernels absorb.
                detail += ",\"fused_sections\":[";
                bool first = true;
                for (int i = 0; i < FS_FINE_SECTIONS; ++i)
                    if (d.fused[i]) {
                        detail += std::string(first ? "" : ",") + "\"" + kSectionNames[i] + "\"";
                        first = false;
                    }
                detail += ']';
            }
        } else {
            prep = st.ms_prep, pres = st.ms_pressure, fin = st.ms_finish, subs = t->call_substeps;
        }
        gpu = "{\"sim_ms\":" + fmt_ms(prep + pres + fin) + ",\"prep_ms\":" + fmt_ms(prep) +
              ",\"pressure_ms\":" + fmt_ms(pres) + ",\"finish_ms\":" + fmt_ms(fin) +
              ",\"render_ms\":null,\"liquid_cells\":" + std::to_string(st.liquid_cells) +
              ",\"substeps\":" + std::to_string(subs) + ",\"detailed\":" + (t->detailed ? "true" : "false") +
              detail + "}";
    }

    // set_timestep_stats (src/profiler/mod.rs:172-193): seconds to ms in f32.
    fs_timestep_stats ts;
    std::memset(&ts, 0, sizeof ts);
    float dropped_total = 0.0f;
    std::string policy;
    if (f) {
        if (int rc = fs_frame_timestep_stats(f, &ts, &dropped_total)) return rc;
        policy = policy_label(ts.policy);
    }
    const uint32_t dps = t->dispatches_last;

    char res[96];
    std::snprintf(res, sizeof res, "%ux%ux%u", q.nx, q.ny, q.nz);
    const double h = q.h;
    std::string j = "{";
    j += std::string("\"timing\":\"") + (gpu != "null" ? "gpu-timestamp" : "off") + "\"";
    j += ",\"substeps\":" + std::to_string(t->call_substeps);
    j += ",\"grid_n\":" + std::to_string(q.nx);
    j += std::string(",\"grid_res\":\"") + res + "\"";
    j += ",\"total_cells\":" + std::to_string(total_cells);
    char fv[64], lf[64];
    const double cell_volume = h * h * h;  // H^3 (src/profiler/mod.rs:525)
    std::snprintf(fv, sizeof fv, "%.6f", (double)st.liquid_cells * cell_volume);
    std::snprintf(lf, sizeof lf, "%.6f", total_cells ? (double)st.liquid_cells / (double)total_cells : 0.0);
    j += std::string(",\"filled_volume\":") + fv + ",\"liquid_fraction\":" + lf;
    j += ",\"particles\":" + std::to_string(t->N);
    j += ",\"pressure_iterations\":" + std::to_string(q.pressure_iterations);
    const uint64_t bytes = fs_buffer_bytes(t);
    j += ",\"gpu_buffer_mb\":" + fmt_mb(bytes) + ",\"sim_buffers_mb\":" + fmt_mb(bytes) +
         ",\"total_tracked_mb\":" + fmt_mb(bytes);
    j += ",\"substeps_this_frame\":" + std::to_string(ts.substeps);
    j += ",\"fixed_dt_ms\":" + fmt_ms(ts.fixed_dt * 1000.0f);
    j += ",\"max_substeps\":" + std::to_string(ts.max_substeps);
    j += ",\"natural_substeps\":" + std::to_string(ts.natural_substeps);
    j += std::string(",\"substep_cap_hit\":") + (ts.substep_cap_hit ? "true" : "false");
    j += ",\"sim_advanced_ms\":" + fmt_ms(ts.sim_advanced * 1000.0f);
    j += ",\"wall_raf_ms\":" + fmt_ms(ts.wall_dt * 1000.0f);
    j += ",\"sim_time_scale\":" + fmt_ratio(ts.sim_time_scale);
    j += ",\"real_time_factor\":" + fmt_ratio(ts.real_time_factor);
    j += ",\"timestep_policy\":\"" + policy + "\"";
    j += ",\"accumulated_before_ms\":" + fmt_ms(ts.accumulated_before * 1000.0f);
    j += ",\"accumulated_after_ms\":" + fmt_ms(ts.accumulated_after * 1000.0f);
    j += ",\"dropped_sim_time_ms\":" + fmt_ms(ts.dropped_this_frame * 1000.0f);
    j += ",\"total_dropped_sim_time_ms\":" + fmt_ms(dropped_total * 1000.0f);
    j += ",\"dispatches_per_substep\":" + std::to_string(dps);
    j += ",\"dispatches_this_frame\":" + std::to_string((uint64_t)dps * ts.substeps);
    j += std::string(",\"engine\":\"") + (q.engine == FS_ENGINE_FUSED ? "fused" : "reference") + "\"";
    j += ",\"gpu\":" + gpu + "}";

    *len_out = j.size();
    if (!buf) return FS_OK;
    if (cap <= j.size()) return fail(FS_ERR_INVALID_ARGUMENT, "stats_json: buffer too small");
    std::memcpy(buf, j.c_str(), j.size() + 1);
    return FS_OK;
}

}  // extern "C"
