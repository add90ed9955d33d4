/*
 * flip_oracle_body.h — precision-generic body of the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see flip_oracle.h).  Included by flip_oracle_f32.c
 * (R = float, ACC = uint32_t wrapping, the reference's i32 atomics) and by
 * flip_oracle_f64.c (R = double, ACC = int64_t, the FP64 oracle).
 *
 * Every arithmetic expression keeps the operand order of the WGSL it restates,
 * and the file is compiled with -ffp-contract=off so no multiply-add is fused:
 * that is what lets the HIP step match this code bit for bit in FP32.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "flip_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(CAT(orc_, name), CAT(_, SFX))
#define OUTS CAT(orc_outputs_, SFX)

typedef R real;

/* WGSL clamp(x,lo,hi) = min(max(x,lo),hi); fmax/fmin return the non-NaN operand. */
static inline real clampr(real x, real lo, real hi) { return RMIN(RMAX(x, lo), hi); }

/* WGSL i32(f) saturates; NaN maps to 0 here (and in the HIP kernels). */
static inline int64_t sat_i32(real x) {
    if (x != x) return 0;
    if (x >= (real)2147483647.0) return 2147483647;
    if (x <= (real)-2147483648.0) return -2147483647 - 1;
    return (int64_t)x;
}

/* Cell key, sh/mark.wgsl:35-43 and src/lib.rs:1161-1172:
 * g = (pos - origin) / h (a DIVISION), (i,j,k) = clamp(floor(g), 0, n-1). */
static inline uint32_t cell_key(const orc_params *P, real x, real y, real z) {
    const real h = (real)P->h;
    real gx = (x - (real)P->origin[0]) / h;
    real gy = (y - (real)P->origin[1]) / h;
    real gz = (z - (real)P->origin[2]) / h;
    uint32_t i = (uint32_t)clampr(RFLOOR(gx), (real)0, (real)(P->nx - 1u));
    uint32_t j = (uint32_t)clampr(RFLOOR(gy), (real)0, (real)(P->ny - 1u));
    uint32_t k = (uint32_t)clampr(RFLOOR(gz), (real)0, (real)(P->nz - 1u));
    return i + P->nx * (j + P->ny * k);
}

uint32_t FN(cell_key)(const orc_params *P, real x, real y, real z) { return cell_key(P, x, y, z); }

/* Occupancy histogram, sh/mark.wgsl:27-45 (atomicAdd(&occ[c], 1)). */
void FN(histogram)(const orc_params *P, uint32_t n, const real *px, const real *py,
                   const real *pz, uint32_t *occ) {
    const uint32_t C = P->nx * P->ny * P->nz;
    memset(occ, 0, sizeof(uint32_t) * C);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
        uint32_t c = cell_key(P, px[p], py[p], pz[p]);
#pragma omp atomic
        occ[c] += 1u;
    }
}

/* Cell typing, sh/classify.wgsl:25-61: boundary -> Solid 0; occ >= max(thresh,1)
 * -> Liquid 1, or (dilation >= 1 and a 6-neighbour is filled) -> Liquid;
 * else Air 2.  Returns stats[0] (liquid count). */
static uint32_t classify(const orc_params *P, const uint32_t *occ, uint32_t *ct) {
    const uint32_t nx = P->nx, ny = P->ny, nz = P->nz, C = nx * ny * nz;
    const uint32_t thresh = (uint32_t)P->liquid_threshold > 1u ? (uint32_t)P->liquid_threshold : 1u;
    const int dil = (uint32_t)P->surface_dilation >= 1u;
    uint32_t liquid = 0;
#pragma omp parallel for schedule(static) reduction(+ : liquid)
    for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
        uint32_t c = (uint32_t)cc;
        uint32_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
            ct[c] = 0u;
            continue;
        }
        int liq = occ[c] >= thresh;
        if (!liq && dil) {
            liq = occ[c - 1] >= thresh || occ[c + 1] >= thresh || occ[c - nx] >= thresh ||
                  occ[c + nx] >= thresh || occ[c - nx * ny] >= thresh || occ[c + nx * ny] >= thresh;
        }
        if (liq) {
            ct[c] = 1u;
            liquid += 1;
        } else {
            ct[c] = 2u;
        }
    }
    return liquid;
}

/* Face grid dims for axis a: n+1 on the component's own axis. */
static inline void face_dims(const orc_params *P, int a, uint32_t d[3]) {
    d[0] = P->nx; d[1] = P->ny; d[2] = P->nz;
    d[a] += 1u;
}

static const real k_off[3][3] = {
    /* u */ {(real)0.0, (real)-0.5, (real)-0.5},
    /* v */ {(real)-0.5, (real)0.0, (real)-0.5},
    /* w */ {(real)-0.5, (real)-0.5, (real)0.0},
};

/* P2G for one particle and one axis, sh/scatter.wgsl:44-76 (scatter_component). */
static inline void p2g_component(const real g[3], const uint32_t dim[3], real vcomp, real scale,
                                 ACC *num, ACC *den) {
    int64_t bx = sat_i32(RFLOOR(g[0])), by = sat_i32(RFLOOR(g[1])), bz = sat_i32(RFLOOR(g[2]));
    real tx = g[0] - (real)bx, ty = g[1] - (real)by, tz = g[2] - (real)bz;
    for (int dk = 0; dk < 2; ++dk) {
        int64_t kk = bz + dk;
        if (kk < 0 || kk >= (int64_t)dim[2]) continue;
        real wz = dk == 1 ? tz : (real)1.0 - tz;
        for (int dj = 0; dj < 2; ++dj) {
            int64_t jj = by + dj;
            if (jj < 0 || jj >= (int64_t)dim[1]) continue;
            real wy = dj == 1 ? ty : (real)1.0 - ty;
            for (int di = 0; di < 2; ++di) {
                int64_t ii = bx + di;
                if (ii < 0 || ii >= (int64_t)dim[0]) continue;
                real wx = di == 1 ? tx : (real)1.0 - tx;
                real w = wx * wy * wz;
                if (w <= (real)0.0) continue;
                uint64_t idx = (uint64_t)ii + dim[0] * ((uint64_t)jj + dim[1] * (uint64_t)kk);
                ACC a = (ACC)sat_i32(RRINT(w * vcomp * scale));
                ACC b = (ACC)sat_i32(RRINT(w * scale));
#pragma omp atomic
                num[idx] += a;
#pragma omp atomic
                den[idx] += b;
            }
        }
    }
}

/* Fused P2G, sh/scatter.wgsl:78-118.  num/den are cleared first
 * (sh/clear.wgsl via src/gpu/fluid.rs:1465-1476). */
static void p2g(const orc_params *P, uint32_t n, const real *px, const real *py, const real *pz,
                const real *vx, const real *vy, const real *vz, ACC *num[3], ACC *den[3]) {
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        size_t cnt = (size_t)d[0] * d[1] * d[2];
        memset(num[a], 0, sizeof(ACC) * cnt);
        memset(den[a], 0, sizeof(ACC) * cnt);
    }
    const real inv_h = RINVH(P);
    const real scale = (real)P->fixed_scale;
    uint32_t dims[3][3];
    for (int a = 0; a < 3; ++a) face_dims(P, a, dims[a]);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
        real pos[3] = {(px[p] - (real)P->origin[0]) * inv_h, (py[p] - (real)P->origin[1]) * inv_h,
                       (pz[p] - (real)P->origin[2]) * inv_h};
        real pv[3] = {vx[p], vy[p], vz[p]};
        for (int a = 0; a < 3; ++a) {
            real g[3] = {pos[0] + k_off[a][0], pos[1] + k_off[a][1], pos[2] + k_off[a][2]};
            p2g_component(g, dims[a], pv[a], scale, num[a], den[a]);
        }
    }
}

/* normalize (sh/normalize.wgsl:15-27), save_vel (sh/save_vel.wgsl:14-19),
 * forces (sh/forces.wgsl:30-84) and enforce (sh/boundaries.wgsl:20-61) for one
 * axis, in that order (src/gpu/fluid.rs:1444-1460). */
static void enforce_axis(const orc_params *P, int a, const uint32_t *ct, real *vel);

/* forces and the pre-pressure enforce, after normalize and save. */
static void forces_enforce(const orc_params *P, const uint32_t *ct, real *vel[3]);

static void p2g_epilogue(const orc_params *P, const uint32_t *ct, ACC *num[3], ACC *den[3],
                         real *vel[3], real *saved[3]) {
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        const int64_t cnt = (int64_t)d[0] * d[1] * d[2];
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < cnt; ++f) {
            NUMT dd = (NUMT)den[a][f];
            vel[a][f] = dd > 0 ? (real)(NUMT)num[a][f] / (real)dd : (real)0.0;
        }
        memcpy(saved[a], vel[a], sizeof(real) * (size_t)cnt);
    }
    forces_enforce(P, ct, vel);
}

static void forces_enforce(const orc_params *P, const uint32_t *ct, real *vel[3]) {
    const uint32_t nx = P->nx, ny = P->ny;
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        const uint32_t nax = a == 0 ? P->nx : (a == 1 ? P->ny : P->nz);
        const int64_t cnt = (int64_t)d[0] * d[1] * d[2];
        const real g_axis = (real)P->grav[a];
        const real dt = (real)P->dt;
#pragma omp parallel for schedule(static)
        for (int64_t ff = 0; ff < cnt; ++ff) {
            uint32_t f = (uint32_t)ff;
            uint32_t i = f % d[0], j = (f / d[0]) % d[1], k = f / (d[0] * d[1]);
            uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
            int liquid = 0;
            if (comp > 0) {
                uint32_t ci = i, cj = j, ck = k;
                if (a == 0) ci -= 1; else if (a == 1) cj -= 1; else ck -= 1;
                if (ct[ci + nx * (cj + ny * ck)] == 1u) liquid = 1;
            }
            if (comp < nax) {
                if (ct[i + nx * (j + ny * k)] == 1u) liquid = 1;
            }
            if (liquid) vel[a][f] = vel[a][f] + g_axis * dt;
        }
    }
    for (int a = 0; a < 3; ++a) enforce_axis(P, a, ct, vel[a]);
}

/* sh/boundaries.wgsl:20-61. */
static void enforce_axis(const orc_params *P, int a, const uint32_t *ct, real *vel) {
    const uint32_t nx = P->nx, ny = P->ny, nz = P->nz;
    uint32_t d[3];
    face_dims(P, a, d);
    const uint32_t nax = a == 0 ? nx : (a == 1 ? ny : nz);
    const int64_t cnt = (int64_t)d[0] * d[1] * d[2];
#pragma omp parallel for schedule(static)
    for (int64_t ff = 0; ff < cnt; ++ff) {
        uint32_t f = (uint32_t)ff;
        uint32_t i = f % d[0], j = (f / d[0]) % d[1], k = f / (d[0] * d[1]);
        uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
        int lo_solid = comp == 0;
        int hi_solid = comp == nax;
        if (!lo_solid) {
            uint32_t ci = i, cj = j, ck = k;
            if (a == 0) ci -= 1; else if (a == 1) cj -= 1; else ck -= 1;
            if (ct[ci + nx * (cj + ny * ck)] == 0u) lo_solid = 1;
        }
        if (!hi_solid) {
            uint32_t ci = i < nx - 1 ? i : nx - 1;
            uint32_t cj = j < ny - 1 ? j : ny - 1;
            uint32_t ck = k < nz - 1 ? k : nz - 1;
            if (ct[ci + nx * (cj + ny * ck)] == 0u) hi_solid = 1;
        }
        if (lo_solid || hi_solid) vel[f] = (real)0.0;
    }
}

/* Divergence with the volume-drift source, sh/divergence.wgsl:22-59. */
static void divergence(const orc_params *P, real *const vel[3], const uint32_t *ct,
                       const uint32_t *occ, real *div) {
    const uint32_t nx = P->nx, ny = P->ny, nz = P->nz, C = nx * ny * nz;
    const real inv_h = RINVH(P);
    const real rest = (real)P->rest, stiff = (real)P->stiff, clamp_d = (real)P->drift_clamp;
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
        uint32_t c = (uint32_t)cc;
        if (ct[c] != 1u) {
            div[c] = (real)0.0;
            continue;
        }
        uint32_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        real u_hi = vel[0][(i + 1u) + (nx + 1u) * (j + ny * k)];
        real u_lo = vel[0][i + (nx + 1u) * (j + ny * k)];
        real v_hi = vel[1][i + nx * ((j + 1u) + (ny + 1u) * k)];
        real v_lo = vel[1][i + nx * (j + (ny + 1u) * k)];
        real w_hi = vel[2][i + nx * (j + ny * (k + 1u))];
        real w_lo = vel[2][i + nx * (j + ny * k)];
        real dv = ((u_hi - u_lo) + (v_hi - v_lo) + (w_hi - w_lo)) * inv_h;
        if (stiff > (real)0.0 && rest > (real)0.0) {
            real over = RMAX((real)0.0, (real)occ[c] - rest) / rest;
            real drift = RMIN(stiff * over, clamp_d);
            dv = dv - drift;
        }
        div[c] = dv;
    }
    (void)nz;
}

/* (A x)_c = n_c x_c - sum_{liquid nb} x_nb over liquid cells; neighbour order
 * -x,+x,-y,+y,-z,+z and float counter, sh/cg_spmv.wgsl:37-68 (host form:
 * src/sim/pressure.rs:145-179). */
static inline real apply_cell(const uint32_t *ct, const real *x, uint32_t c, uint32_t nx,
                              uint32_t nxny, real xc) {
    real cnt = (real)0.0, sum = (real)0.0;
    const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
    for (int s = 0; s < 6; ++s) {
        uint32_t t = ct[nb[s]];
        if (t != 0u) {
            cnt += (real)1.0;
            if (t == 1u) sum += x[nb[s]];
        }
    }
    return cnt * xc - sum;
}

static void apply_poisson(const orc_params *P, const uint32_t *ct, const real *x, real *out) {
    const uint32_t nx = P->nx, nxny = P->nx * P->ny, C = nxny * P->nz;
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
        uint32_t c = (uint32_t)cc;
        out[c] = ct[c] == 1u ? apply_cell(ct, x, c, nx, nxny, x[c]) : (real)0.0;
    }
}

/* Two-level dot product.  mode 0 restates sh/cg_reduce.wgsl:25-54 then
 * sh/cg_reduce_final.wgsl:23-54 exactly (256-element trees in R); mode 1 is
 * the f64 dot of src/sim/pressure.rs:182-188 evaluated per 256-cell block
 * then over the blocks in index order (thread-count independent). */
static double dot2(uint32_t C, const real *a, const real *b, int mode, int active, real *partials) {
    const uint32_t nb = (C + 255u) / 256u;
    if (mode == 0) {
#pragma omp parallel for schedule(static)
        for (int64_t blk = 0; blk < (int64_t)nb; ++blk) {
            real s[256];
            for (uint32_t li = 0; li < 256; ++li) {
                uint64_t idx = (uint64_t)blk * 256u + li;
                s[li] = (active && idx < C) ? a[idx] * b[idx] : (real)0.0;
            }
            for (uint32_t d = 128; d >= 1; d >>= 1)
                for (uint32_t li = 0; li < d; ++li) s[li] += s[li + d];
            partials[blk] = s[0];
        }
        real s[256];
        for (uint32_t li = 0; li < 256; ++li) {
            real acc = (real)0.0;
            if (active)
                for (uint32_t idx = li; idx < nb; idx += 256u) acc += partials[idx];
            s[li] = acc;
        }
        for (uint32_t d = 128; d >= 1; d >>= 1)
            for (uint32_t li = 0; li < d; ++li) s[li] += s[li + d];
        return (double)s[0];
    }
    double *pd = (double *)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < (int64_t)nb; ++blk) {
        double acc = 0.0;
        uint64_t e = (uint64_t)blk * 256u + 256u;
        if (e > C) e = C;
        for (uint64_t idx = (uint64_t)blk * 256u; idx < e; ++idx) acc += (double)a[idx] * (double)b[idx];
        pd[blk] = acc;
    }
    double acc = 0.0;
    for (uint32_t blk = 0; blk < nb; ++blk) acc += pd[blk];
    free(pd);
    return active ? acc : 0.0;
}

/* CG pressure solve, the dispatch sequence of src/gpu/fluid.rs:1622-1701:
 * cg_init (sh/cg_init.wgsl:50-87) -> reduce r.r -> set_rsold
 * (sh/cg_set_rsold.wgsl) -> pressure_iters x [spmv, reduce d.q, alpha
 * (sh/cg_alpha.wgsl), update (sh/cg_update.wgsl), reduce r.r, beta
 * (sh/cg_beta.wgsl), dir (sh/cg_dir.wgsl)].
 *
 * reduce_mode 0 keeps every scalar in R with the WGSL f32 trees.
 * reduce_mode 1 (mixed) keeps rs_old/rs_initial/dq/rs_new in f64 and casts
 * alpha and beta to R, as src/sim/pressure.rs:282,296 do; the stop test is
 * rs_new <= tol^2 * rs_initial in f64 (src/sim/pressure.rs:262-268,286-292). */
uint32_t FN(pressure_solve)(const orc_params *P, const uint32_t *ct, const real *div, real *p,
                            orc_stats *st) {
    const uint32_t nx = P->nx, nxny = P->nx * P->ny, C = nxny * P->nz;
    const int mode = (int)P->reduce_mode;
    real *r = (real *)malloc(sizeof(real) * C);
    real *d = (real *)malloc(sizeof(real) * C);
    real *q = (real *)malloc(sizeof(real) * C);
    real *b = (real *)malloc(sizeof(real) * C);
    real *partials = (real *)malloc(sizeof(real) * ((C + 255u) / 256u));
    const real h = (real)P->h;
    const real scale = (real)P->rho * h * h / (real)P->dt;
    const int warm = P->warm_start != 0u;
    /* cg_init: r = rhs - A p_initial (warm) or rhs (cold); d = r. */
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
        uint32_t c = (uint32_t)cc;
        if (ct[c] == 1u) {
            real rhs = -scale * div[c];
            real ap = (real)0.0;
            if (warm) ap = apply_cell(ct, p, c, nx, nxny, p[c]);
            real rc = rhs - ap;
            r[c] = rc;
            d[c] = rc;
            b[c] = rhs;
        } else {
            r[c] = (real)0.0;
            d[c] = (real)0.0;
            b[c] = (real)0.0;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
        uint32_t c = (uint32_t)cc;
        if (ct[c] != 1u || !warm) p[c] = (real)0.0;
    }
    /* scalars: rs_old, rs_initial as R (mode 0) or f64 (mode 1). */
    double rs0 = dot2(C, r, r, mode, 1, partials);
    real rs_old_r = (real)rs0, rs_init_r = (real)rs0;
    double rs_old = rs0, rs_init = rs0;
    int active = 1;
    uint32_t iters = 0;
    const real tol_sq_r = (real)P->tol_sq;
    const double tol64 = (double)P->tol;
    const double tol_sq64 = tol64 * tol64;
    for (uint32_t it = 0; it < P->pressure_iters; ++it) {
        if (!active) break; /* every later dispatch is a no-op (active == 0) */
        iters++;
        apply_poisson(P, ct, d, q);
        double dq = dot2(C, d, q, mode, 1, partials);
        real alpha;
        if (mode == 0) {
            real dqr = (real)dq;
            alpha = RFABS(dqr) > (real)1e-30f ? rs_old_r / dqr : (real)0.0;
        } else {
            alpha = fabs(dq) > 1e-30 ? (real)(rs_old / dq) : (real)0.0;
        }
#pragma omp parallel for schedule(static)
        for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
            p[cc] += alpha * d[cc];
            r[cc] -= alpha * q[cc];
        }
        double rr = dot2(C, r, r, mode, 1, partials);
        real beta;
        if (mode == 0) {
            real rs_new = (real)rr;
            beta = rs_old_r > (real)1e-30f ? rs_new / rs_old_r : (real)0.0;
            rs_old_r = rs_new;
            if (tol_sq_r > (real)0.0 && rs_new <= tol_sq_r * rs_init_r) {
                active = 0;
                beta = (real)0.0;
            }
        } else {
            beta = rs_old > 1e-30 ? (real)(rr / rs_old) : (real)0.0;
            rs_old = rr;
            if (tol64 > 0.0 && rr <= tol_sq64 * rs_init) {
                active = 0;
                beta = (real)0.0;
            }
        }
        if (active) {
#pragma omp parallel for schedule(static)
            for (int64_t cc = 0; cc < (int64_t)C; ++cc) d[cc] = r[cc] + beta * d[cc];
        }
    }
    if (st) {
        st->cg_iterations = iters;
        st->rs_initial = mode == 0 ? (double)rs_init_r : rs_init;
        st->rs_final = mode == 0 ? (double)rs_old_r : rs_old;
        /* True residual |b - A p|^2 in f64 over liquid cells. */
        double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
        for (int64_t cc = 0; cc < (int64_t)C; ++cc) {
            uint32_t c = (uint32_t)cc;
            if (ct[c] != 1u) continue;
            double cnt = 0.0, sum = 0.0;
            const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
            for (int s = 0; s < 6; ++s) {
                uint32_t t = ct[nb[s]];
                if (t != 0u) {
                    cnt += 1.0;
                    if (t == 1u) sum += (double)p[nb[s]];
                }
            }
            double e = (double)b[c] - (cnt * (double)p[c] - sum);
            acc += e * e;
        }
        st->true_residual_sq = acc;
    }
    free(r); free(d); free(q); free(b); free(partials);
    return iters;
}

/* Gradient subtraction, sh/gradient.wgsl:23-67, then enforce x3. */
static void gradient_enforce(const orc_params *P, const real *pr, const uint32_t *ct, real *vel[3]) {
    const uint32_t nx = P->nx, ny = P->ny, nz = P->nz;
    const real coeff = ((real)P->dt / (real)P->rho) / (real)P->h;
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        const uint32_t nax = a == 0 ? nx : (a == 1 ? ny : nz);
        const int64_t cnt = (int64_t)d[0] * d[1] * d[2];
#pragma omp parallel for schedule(static)
        for (int64_t ff = 0; ff < cnt; ++ff) {
            uint32_t f = (uint32_t)ff;
            uint32_t i = f % d[0], j = (f / d[0]) % d[1], k = f / (d[0] * d[1]);
            uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
            if (comp == 0 || comp == nax) continue;
            uint32_t hi = i + nx * (j + ny * k);
            uint32_t lo = a == 0 ? (i - 1u) + nx * (j + ny * k)
                                 : (a == 1 ? i + nx * ((j - 1u) + ny * k) : i + nx * (j + ny * (k - 1u)));
            uint32_t lt = ct[lo], ht = ct[hi];
            if (lt == 0u || ht == 0u) continue;
            if (lt != 1u && ht != 1u) continue;
            vel[a][f] = vel[a][f] - coeff * (pr[hi] - pr[lo]);
        }
    }
    for (int a = 0; a < 3; ++a) enforce_axis(P, a, ct, vel[a]);
}

void FN(gradient_enforce)(const orc_params *P, const real *p, const uint32_t *ct, real *vel[3]) {
    gradient_enforce(P, p, ct, vel);
}

/* Static-wall test of sh/g2p.wgsl:35-54 (u: i<=1 ... etc). */
static inline int touches_static_solid(const orc_params *P, int a, int64_t ii, int64_t jj, int64_t kk) {
    const int64_t nx = P->nx, ny = P->ny, nz = P->nz;
    const int64_t lx = a == 0 ? 1 : 0, ly = a == 1 ? 1 : 0, lz = a == 2 ? 1 : 0;
    return ii <= lx || ii >= nx - 1 || jj <= ly || jj >= ny - 1 || kk <= lz || kk >= nz - 1;
}

/* sample_u/v/w of sh/g2p.wgsl:56-129 (wall_aware) or the unaware variant of
 * src/sim/mod.rs:462-525.  Returns (final, saved) / wsum, or 0 when wsum == 0. */
static inline void sample_axis(const orc_params *P, int a, const real Pp[3], const real *vel,
                               const real *saved, int wall_aware, real *ofin, real *osav) {
    uint32_t dim[3];
    face_dims(P, a, dim);
    const real inv_h = RINVH(P);
    real g[3];
    for (int c = 0; c < 3; ++c) g[c] = (Pp[c] - (real)P->origin[c]) * inv_h + k_off[a][c];
    int64_t bx = sat_i32(RFLOOR(g[0])), by = sat_i32(RFLOOR(g[1])), bz = sat_i32(RFLOOR(g[2]));
    real tx = g[0] - (real)bx, ty = g[1] - (real)by, tz = g[2] - (real)bz;
    real fin = (real)0.0, sav = (real)0.0, wsum = (real)0.0;
    for (int dk = 0; dk < 2; ++dk) {
        int64_t kk = bz + dk;
        if (kk < 0 || kk >= (int64_t)dim[2]) continue;
        real wz = dk == 1 ? tz : (real)1.0 - tz;
        for (int dj = 0; dj < 2; ++dj) {
            int64_t jj = by + dj;
            if (jj < 0 || jj >= (int64_t)dim[1]) continue;
            real wy = dj == 1 ? ty : (real)1.0 - ty;
            for (int di = 0; di < 2; ++di) {
                int64_t ii = bx + di;
                if (ii < 0 || ii >= (int64_t)dim[0]) continue;
                if (wall_aware && touches_static_solid(P, a, ii, jj, kk)) continue;
                real wx = di == 1 ? tx : (real)1.0 - tx;
                real w = wx * wy * wz;
                uint64_t idx = (uint64_t)ii + dim[0] * ((uint64_t)jj + dim[1] * (uint64_t)kk);
                fin += w * vel[idx];
                sav += w * saved[idx];
                wsum += w;
            }
        }
    }
    if (wsum > (real)0.0) {
        *ofin = fin / wsum;
        *osav = sav / wsum;
    } else {
        *ofin = (real)0.0;
        *osav = (real)0.0;
    }
}

void FN(sample_axis)(const orc_params *P, int axis, real x, real y, real z, const real *vel,
                     const real *saved, int wall_aware, real *out_fin, real *out_sav) {
    real Pp[3] = {x, y, z};
    sample_axis(P, axis, Pp, vel, saved, wall_aware, out_fin, out_sav);
}

/* G2P + FLIP/PIC blend + CFL clamp + wall friction + advection + wall clamp,
 * sh/g2p.wgsl:131-198. */
static void g2p(const orc_params *P, uint32_t n, real *px, real *py, real *pz, real *vx, real *vy,
                real *vz, real *const vel[3], real *const saved[3]) {
    const real h = (real)P->h, dt = (real)P->dt, alpha = (real)P->flip_blend;
    const real cfl = RMAX((real)P->cfl, (real)1.0);
    const real maxs = cfl * h / dt;
    const real ext[3] = {(real)(float)P->nx * h, (real)(float)P->ny * h, (real)(float)P->nz * h};
    const real hband = h * (real)1.05f;
    real lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
        lo[c] = (real)P->origin[c] + hband;
        hi[c] = (real)P->origin[c] + ext[c] - hband;
    }
    const real friction = (real)P->wall_friction;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
        real Pp[3] = {px[p], py[p], pz[p]};
        real pv[3] = {vx[p], vy[p], vz[p]};
        real pic[3], sv[3];
        for (int a = 0; a < 3; ++a) sample_axis(P, a, Pp, vel[a], saved[a], 1, &pic[a], &sv[a]);
        real nv[3];
        for (int a = 0; a < 3; ++a) {
            real flip = pv[a] + (pic[a] - sv[a]);
            nv[a] = pic[a] * ((real)1.0 - alpha) + flip * alpha; /* WGSL mix */
        }
        real sp = RSQRT(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        if (sp > maxs && sp > (real)0.0) {
            real s = maxs / sp;
            for (int a = 0; a < 3; ++a) nv[a] = nv[a] * s;
        }
        if (friction > (real)0.0) {
            const real band = h * (real)2.5;
            real near[3];
            for (int a = 0; a < 3; ++a) {
                real dlo = (Pp[a] - lo[a]) / band;
                real dhi = (hi[a] - Pp[a]) / band;
                near[a] = clampr((real)1.0 - RMIN(dlo, dhi), (real)0.0, (real)1.0);
            }
            nv[0] = nv[0] * ((real)1.0 - friction * RMAX(near[1], near[2]));
            nv[1] = nv[1] * ((real)1.0 - friction * RMAX(near[0], near[2]));
            nv[2] = nv[2] * ((real)1.0 - friction * RMAX(near[0], near[1]));
        }
        real np[3];
        for (int a = 0; a < 3; ++a) np[a] = Pp[a] + nv[a] * dt;
        for (int a = 0; a < 3; ++a) {
            if (np[a] < lo[a]) { np[a] = lo[a]; nv[a] = (real)0.0; }
            if (np[a] > hi[a]) { np[a] = hi[a]; nv[a] = (real)0.0; }
        }
        px[p] = np[0]; py[p] = np[1]; pz[p] = np[2];
        vx[p] = nv[0]; vy[p] = nv[1]; vz[p] = nv[2];
    }
}

void FN(g2p)(const orc_params *P, uint32_t n, real *px, real *py, real *pz, real *vx, real *vy,
             real *vz, real *const vel[3], real *const saved[3]) {
    g2p(P, n, px, py, pz, vx, vy, vz, vel, saved);
}

void FN(divergence)(const orc_params *P, real *const vel[3], const uint32_t *ct, const uint32_t *occ,
                    real *div) {
    divergence(P, vel, ct, occ, div);
}

void FN(apply_poisson)(const orc_params *P, const uint32_t *ct, const real *x, real *out) {
    apply_poisson(P, ct, x, out);
}

void FN(p2g)(const orc_params *P, uint32_t n, const real *px, const real *py, const real *pz,
             const real *vx, const real *vy, const real *vz, ACC_PUB *num[3], ACC_PUB *den[3]) {
    p2g(P, n, px, py, pz, vx, vy, vz, (ACC **)num, (ACC **)den);
}

void FN(p2g_epilogue)(const orc_params *P, const uint32_t *ct, ACC_PUB *num[3], ACC_PUB *den[3], real *vel[3],
                      real *saved[3]) {
    p2g_epilogue(P, ct, (ACC **)num, (ACC **)den, vel, saved);
}

static void permute(uint32_t n, const uint32_t *perm, real *a, real *tmp) {
#pragma omp parallel for schedule(static)
    for (int64_t d = 0; d < (int64_t)n; ++d) tmp[d] = a[perm[d]];
    memcpy(a, tmp, sizeof(real) * n);
}

/* Wall time of each phase of FN(run), summed over its substeps since the
 * last reset (the CPU baseline's per-phase figures): mark, sort, classify,
 * scatter (P2G), face epilogue (normalize .. enforce), divergence, CG,
 * gradient + enforce, G2P. */
#define ORC_PHASES 9
static double CAT(g_phase_ms_, SFX)[ORC_PHASES];
static inline double orc_now_ms(void) {
#ifdef _OPENMP
    return omp_get_wtime() * 1e3;
#else
    return 0.0;
#endif
}
void FN(phase_ms)(double *out) { memcpy(out, CAT(g_phase_ms_, SFX), sizeof(double) * ORC_PHASES); }
void FN(phase_reset)(void) { memset(CAT(g_phase_ms_, SFX), 0, sizeof(double) * ORC_PHASES); }

/* One or more substeps in the order of src/gpu/mod.rs:993-1033. */
int FN(run)(const orc_params *P, uint32_t n, real *px, real *py, real *pz, real *vx, real *vy,
            real *vz, real *pressure, uint32_t steps, uint32_t sort_period, uint32_t tick0,
            OUTS *out) {
    const uint32_t C = P->nx * P->ny * P->nz;
    uint32_t fc[3];
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        fc[a] = d[0] * d[1] * d[2];
    }
    uint32_t *occ = (uint32_t *)malloc(sizeof(uint32_t) * C);
    uint32_t *ct = (uint32_t *)malloc(sizeof(uint32_t) * C);
    uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    real *tmp = (real *)malloc(sizeof(real) * (n ? n : 1));
    real *div = (real *)malloc(sizeof(real) * C);
    ACC *num[3], *den[3];
    real *vel[3], *saved[3];
    for (int a = 0; a < 3; ++a) {
        num[a] = (ACC *)malloc(sizeof(ACC) * fc[a]);
        den[a] = (ACC *)malloc(sizeof(ACC) * fc[a]);
        vel[a] = (real *)malloc(sizeof(real) * fc[a]);
        saved[a] = (real *)malloc(sizeof(real) * fc[a]);
    }
    orc_stats st;
    memset(&st, 0, sizeof st);
    int sorted_last = 0;
    double *ph = CAT(g_phase_ms_, SFX);
    for (uint32_t s = 0; s < steps; ++s) {
        double t0 = orc_now_ms(), t1;
        /* clear (pressure_a only when warm start is off) + mark */
        if (!P->warm_start) memset(pressure, 0, sizeof(real) * C);
        FN(histogram)(P, n, px, py, pz, occ);
        t1 = orc_now_ms(); ph[0] += t1 - t0; t0 = t1;
        const uint32_t tick = tick0 + s;
        sorted_last = sort_period > 0 && (tick % sort_period) == 0u;
        if (sorted_last) {
            for (uint32_t p = 0; p < n; ++p) keys[p] = cell_key(P, px[p], py[p], pz[p]);
            orc_stable_permutation(n, C, keys, perm);
            permute(n, perm, px, tmp);
            permute(n, perm, py, tmp);
            permute(n, perm, pz, tmp);
            permute(n, perm, vx, tmp);
            permute(n, perm, vy, tmp);
            permute(n, perm, vz, tmp);
            if (out && out->cell_offset) orc_exclusive_scan_u32(C, occ, out->cell_offset);
        }
        t1 = orc_now_ms(); ph[1] += t1 - t0; t0 = t1;
        st.liquid_cells = classify(P, occ, ct);
        t1 = orc_now_ms(); ph[2] += t1 - t0; t0 = t1;
        p2g(P, n, px, py, pz, vx, vy, vz, num, den);
        t1 = orc_now_ms(); ph[3] += t1 - t0; t0 = t1;
        p2g_epilogue(P, ct, num, den, vel, saved);
        t1 = orc_now_ms(); ph[4] += t1 - t0; t0 = t1;
        if (out) {
            for (int a = 0; a < 3; ++a) {
                if (out->num[a]) memcpy(out->num[a], num[a], sizeof(ACC) * fc[a]);
                if (out->den[a]) memcpy(out->den[a], den[a], sizeof(ACC) * fc[a]);
                if (out->vel[a]) memcpy(out->vel[a], vel[a], sizeof(real) * fc[a]);
                if (out->saved[a]) memcpy(out->saved[a], saved[a], sizeof(real) * fc[a]);
            }
        }
        t0 = orc_now_ms();
        divergence(P, vel, ct, occ, div);
        t1 = orc_now_ms(); ph[5] += t1 - t0; t0 = t1;
        FN(pressure_solve)(P, ct, div, pressure, &st);
        t1 = orc_now_ms(); ph[6] += t1 - t0; t0 = t1;
        gradient_enforce(P, pressure, ct, vel);
        t1 = orc_now_ms(); ph[7] += t1 - t0; t0 = t1;
        g2p(P, n, px, py, pz, vx, vy, vz, vel, saved);
        t1 = orc_now_ms(); ph[8] += t1 - t0;
    }
    if (out) {
        if (out->occupancy) memcpy(out->occupancy, occ, sizeof(uint32_t) * C);
        if (out->cell_type) memcpy(out->cell_type, ct, sizeof(uint32_t) * C);
        if (out->perm && sorted_last) memcpy(out->perm, perm, sizeof(uint32_t) * n);
        if (out->divergence) memcpy(out->divergence, div, sizeof(real) * C);
        for (int a = 0; a < 3; ++a)
            if (out->vel_final[a]) memcpy(out->vel_final[a], vel[a], sizeof(real) * fc[a]);
        if (out->stats) *out->stats = st;
    }
    free(occ); free(ct); free(keys); free(perm); free(tmp); free(div);
    for (int a = 0; a < 3; ++a) {
        free(num[a]); free(den[a]); free(vel[a]); free(saved[a]);
    }
    return sorted_last;
}

/* Host Jacobi, src/sim/pressure.rs:71-137 (KAT use only). */
void FN(jacobi_solve)(const orc_params *P, const uint32_t *ct, const real *div, uint32_t iters,
                      real *pout) {
    const uint32_t nx = P->nx, nxny = P->nx * P->ny, C = nxny * P->nz;
    const real h = (real)P->h;
    const real scale = (real)P->rho * h * h / (real)P->dt;
    real *a = (real *)calloc(C, sizeof(real));
    real *b = (real *)calloc(C, sizeof(real));
    for (uint32_t it = 0; it < iters; ++it) {
        for (uint32_t c = 0; c < C; ++c) {
            if (ct[c] != 1u) {
                b[c] = (real)0.0;
                continue;
            }
            const uint32_t nb[6] = {c - 1u, c + 1u, c - nx, c + nx, c - nxny, c + nxny};
            real sum = (real)0.0, cnt = (real)0.0;
            for (int s = 0; s < 6; ++s) {
                uint32_t t = ct[nb[s]];
                if (t == 2u) cnt += (real)1.0;
                else if (t == 1u) { sum += a[nb[s]]; cnt += (real)1.0; }
            }
            b[c] = cnt > (real)0.0 ? (sum - scale * div[c]) / cnt : (real)0.0;
        }
        real *t = a; a = b; b = t;
    }
    memcpy(pout, a, sizeof(real) * C);
    free(a); free(b);
}

#include "flip_oracle_adjoint.h"
