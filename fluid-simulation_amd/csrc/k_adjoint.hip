// k_adjoint.hip — reverse-mode kernels of one fused substep (the adjoint of
// the step of sh/scatter.wgsl .. sh/g2p.wgsl as restated by the fused engine).
//
// Conventions (the forward's discrete decisions are held fixed):
//   * the sort order, cell types, stencil masks, the static-wall filter of the
//     G2P gather, which faces are wet/solid/gradient-applicable, the CFL
//     branch and the wall-clamp branch are constants;
//   * fixed-point rounding of the P2G terms is the identity (num = sum w v,
//     den = sum w);
//   * the pressure solve is exact, p = A^-1 b, so its adjoint is one more
//     solve with the same SPD operator, b_bar = A^-1 p_bar (the warm start
//     does not enter);
//   * the trilinear weights are differentiated inside their cell (d w/d x is
//     piecewise constant in the fractional coordinate).
// Reverse order of one substep:
//   G2P + blend + CFL + advection + clamp   -> k_adj_g2p (particles; scatters
//                                              into the face adjoints)
//   gradient + enforce                      -> k_adj_pbar (cells)
//   pressure solve                          -> adjoint CG (k_adj_cg_init + the
//                                              persistent CG kernel)
//   divergence + forces/enforce + normalize -> k_adj_faces (faces)
//   P2G                                     -> k_adj_p2g (particles)
//   sort                                    -> k_adj_unsort
#include "fs_cell.h"
#include "fs_particle.h"

namespace fs {

namespace {
inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
constexpr int kAdjBlock = 256;
}  // namespace

// Adjoint of sample_axis (fs_particle.h): given S_fin, S_sav (the forward
// samples) and their adjoints bfin, bsav, scatter bfin*w/W and bsav*w/W into
// the face adjoints and add dS/dx to xbar.
__device__ __forceinline__ void sample_axis_adjoint(const StepParams& P, int a, const float Pp[3],
                                                    const float* __restrict__ vel, const float* __restrict__ sav,
                                                    float s_fin, float s_sav, float bfin, float bsav,
                                                    float* __restrict__ vbar, float* __restrict__ sbar,
                                                    float xbar[3]) {
    uint32_t dim[3];
    face_dims(P, a, dim);
    float g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[c] = (Pp[c] - P.origin[c]) * P.inv_h + stagger(a, c);
    const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
    const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;
    const int32_t lx = a == 0 ? 1 : 0, ly = a == 1 ? 1 : 0, lz = a == 2 ? 1 : 0;
    auto valid = [&](int32_t ii, int32_t jj, int32_t kk) {
        if (ii < 0 || ii >= (int32_t)dim[0] || jj < 0 || jj >= (int32_t)dim[1] || kk < 0 || kk >= (int32_t)dim[2])
            return false;
        return !(ii <= lx || ii >= nx - 1 || jj <= ly || jj >= ny - 1 || kk <= lz || kk >= nz - 1);
    };
    float W = 0.0f;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                if (!valid(bx + di, by + dj, bz + dk)) continue;
                const float wx = di ? tx : 1.0f - tx, wy = dj ? ty : 1.0f - ty, wz = dk ? tz : 1.0f - tz;
                W += wx * wy * wz;
            }
    if (!(W > 0.0f)) return;  // the sample was the constant 0
    const float invW = 1.0f / W;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int32_t ii = bx + di, jj = by + dj, kk = bz + dk;
                if (!valid(ii, jj, kk)) continue;
                const float wx = di ? tx : 1.0f - tx, wy = dj ? ty : 1.0f - ty, wz = dk ? tz : 1.0f - tz;
                const float sx = di ? 1.0f : -1.0f, sy = dj ? 1.0f : -1.0f, sz = dk ? 1.0f : -1.0f;
                const float w = wx * wy * wz;
                const uint32_t idx = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
                if (bfin != 0.0f) atomicAdd(&vbar[idx], bfin * w * invW);
                if (bsav != 0.0f) atomicAdd(&sbar[idx], bsav * w * invW);
                const float c = bfin * (vel[idx] - s_fin) + bsav * (sav[idx] - s_sav);
                dx += sx * wy * wz * c;
                dy += wx * sy * wz * c;
                dz += wx * wy * sz * c;
            }
    xbar[0] += dx * invW * P.inv_h;
    xbar[1] += dy * invW * P.inv_h;
    xbar[2] += dz * invW * P.inv_h;
}

// Adjoint of g2p_particle (fs_particle.h) for friction 0.  In: the sorted
// particles before G2P, the final and saved faces, the adjoint of the
// particle outputs (gxo, gvo).  Out: gxs, gvs (adjoint of the particle inputs
// through this part of the step) and the face adjoints vbar, sbar (added).
__global__ __launch_bounds__(kAdjBlock) void k_adj_g2p(StepParams P, ParticleSoA sp, AdjFaces faces, AdjParticles out,
                                                       AdjParticlesIn in) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.N) return;
    const float h = P.h, dt = P.dt, alpha = P.flip;
    const float Pp[3] = {sp.pos[0][p], sp.pos[1][p], sp.pos[2][p]};
    const float pv[3] = {((const float*)sp.vel[0])[p], ((const float*)sp.vel[1])[p], ((const float*)sp.vel[2])[p]};
    float pic[3], sav[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) sample_axis(P, a, Pp, faces.vel[a], faces.saved[a], pic[a], sav[a]);
    float nv0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float flip = pv[a] + (pic[a] - sav[a]);
        nv0[a] = pic[a] * (1.0f - alpha) + flip * alpha;
    }
    const float cfl = fmaxf(P.cfl, 1.0f);
    const float maxs = cfl * h / dt;
    const float spd = sqrtf(nv0[0] * nv0[0] + nv0[1] * nv0[1] + nv0[2] * nv0[2]);
    const bool capped = spd > maxs && spd > 0.0f;
    const float s = capped ? maxs / spd : 1.0f;
    float nv1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) nv1[a] = capped ? nv0[a] * s : nv0[a];
    const float hband = h * 1.05f;
    const float ext[3] = {(float)P.nx * h, (float)P.ny * h, (float)P.nz * h};
    float nv1b[3], xb[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = P.origin[a] + hband, hi = P.origin[a] + ext[a] - hband;
        const float np = Pp[a] + nv1[a] * dt;
        const bool clamped = np < lo || np > hi;
        const float gx = in.gx[a][p], gv = in.gv[a][p];
        nv1b[a] = clamped ? 0.0f : gv + gx * dt;
        xb[a] = clamped ? 0.0f : gx;
    }
    float nv0b[3];
    if (capped) {
        const float n[3] = {nv0[0] / spd, nv0[1] / spd, nv0[2] / spd};
        const float dn = n[0] * nv1b[0] + n[1] * nv1b[1] + n[2] * nv1b[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) nv0b[a] = s * (nv1b[a] - n[a] * dn);
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) nv0b[a] = nv1b[a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        sample_axis_adjoint(P, a, Pp, faces.vel[a], faces.saved[a], pic[a], sav[a], nv0b[a], -alpha * nv0b[a],
                            faces.vbar[a], faces.sbar[a], xb);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        out.gx[a][p] = xb[a];
        out.gv[a][p] = alpha * nv0b[a];
    }
}

// The faces whose velocity the pressure gradient changes (face_gradient).
__device__ __forceinline__ bool grad_applies(const StepParams& P, int a, const uint8_t* __restrict__ ct, uint32_t i,
                                             uint32_t j, uint32_t k) {
    const uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
    if (comp == 0u || comp == axis_n(P, a)) return false;
    const uint32_t hi = i + P.nx * (j + P.ny * k);
    const uint32_t lo = a == 0 ? hi - 1u : (a == 1 ? hi - P.nx : hi - P.nx * P.ny);
    const uint8_t lt = ct[lo], ht = ct[hi];
    if (lt == kSolid || ht == kSolid) return false;
    return lt == kLiquid || ht == kLiquid;
}

// p_bar of liquid cells: vel_f -= coeff (p[hi] - p[lo]) on applicable faces.
__global__ __launch_bounds__(kAdjBlock) void k_adj_pbar(StepParams P, const uint8_t* __restrict__ ct, AdjFaces faces,
                                                        float* __restrict__ pbar) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.C) return;
    if (ct[c] != kLiquid) {
        pbar[c] = 0.0f;
        return;
    }
    const uint32_t i = c % P.nx, j = (c / P.nx) % P.ny, k = c / (P.nx * P.ny);
    const float coeff = (P.dt / P.rho) / P.h;
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t d[3];
        face_dims(P, a, d);
        const uint32_t own = i + d[0] * (j + d[1] * k);
        const uint32_t ni = i + (a == 0), nj = j + (a == 1), nk = k + (a == 2);
        const uint32_t next = ni + d[0] * (nj + d[1] * nk);
        if (grad_applies(P, a, ct, i, j, k)) acc -= coeff * faces.vbar[a][own];
        if (grad_applies(P, a, ct, ni, nj, nk)) acc += coeff * faces.vbar[a][next];
    }
    pbar[c] = acc;
}

// Right-hand side of the adjoint solve (liquid cells), zero start.
__global__ __launch_bounds__(kAdjBlock) void k_adj_cg_init(uint32_t C, const uint16_t* __restrict__ nmask,
                                                           const float* __restrict__ pbar, float* __restrict__ rhs,
                                                           float* __restrict__ z, float* __restrict__ r,
                                                           float* __restrict__ d) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const float v = (nmask[c] & kMaskLiquid) ? pbar[c] : 0.0f;
    rhs[c] = v;
    r[c] = v;
    d[c] = v;
    z[c] = 0.0f;
}

// Per face: adjoint of gradient/enforce (pass-through), divergence, the
// epilogue (forces, enforce, save) and normalize.  z = A^-1 p_bar.
__global__ __launch_bounds__(kAdjBlock) void k_adj_faces(StepParams P, const uint8_t* __restrict__ ct,
                                                         const float* __restrict__ z, AdjFaces faces,
                                                         double* __restrict__ grav) {
    __shared__ double s_g[3][kAdjBlock / 64];
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    int a = 0;
    uint32_t f = 0;
    double gcontrib[3] = {0.0, 0.0, 0.0};
    if (split_face(P, gi, a, f)) {
        uint32_t d[3], i, j, k;
        face_dims(P, a, d);
        face_ijk(f, d, i, j, k);
        const bool solid = face_solid(P, a, ct, i, j, k);
        const bool wet = face_wet(P, a, ct, i, j, k);
        float ve = solid ? 0.0f : faces.vbar[a][f];
        const float bscale = P.rho * P.h * P.h / P.dt;
        const uint32_t comp = a == 0 ? i : (a == 1 ? j : k);
        if (comp > 0u) {
            const uint32_t lo = a == 0 ? (i - 1u) + P.nx * (j + P.ny * k)
                                       : (a == 1 ? i + P.nx * ((j - 1u) + P.ny * k) : i + P.nx * (j + P.ny * (k - 1u)));
            if (ct[lo] == kLiquid) ve += P.inv_h * (-bscale * z[lo]);
        }
        if (comp < axis_n(P, a)) {
            const uint32_t hi = i + P.nx * (j + P.ny * k);
            if (ct[hi] == kLiquid) ve -= P.inv_h * (-bscale * z[hi]);
        }
        if (!solid && wet) gcontrib[a] = (double)ve * (double)P.dt;
        const float vn = (solid ? 0.0f : ve) + faces.sbar[a][f];
        const int32_t den = faces.den[a][f];
        float nb = 0.0f, db = 0.0f;
        if (den > 0) {
            const float den_r = (float)den / P.scale;
            nb = vn / den_r;
            db = -vn * faces.saved[a][f] / den_r;
        }
        faces.numbar[a][f] = nb;
        faces.denbar[a][f] = db;
    }
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v = wave_sum(gcontrib[c]);
        if (lane == 0) s_g[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < kAdjBlock / 64; ++w) t += s_g[threadIdx.x][w];
        if (t != 0.0) atomicAdd(&grav[threadIdx.x], t);
    }
}

// Adjoint of P2G (num = sum w v, den = sum w per face, the taps of
// sh/scatter.wgsl:52-75): adds to gxs, gvs.
__global__ __launch_bounds__(kAdjBlock) void k_adj_p2g(StepParams P, ParticleSoA sp, AdjFaces faces,
                                                       AdjParticles out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.N) return;
    const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                          (sp.pos[2][p] - P.origin[2]) * P.inv_h};
    float xb[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float va = ((const float*)sp.vel[a])[p];
        uint32_t dim[3];
        face_dims(P, a, dim);
        const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
        const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
        const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
        float vb = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
#pragma unroll
        for (int dk = 0; dk < 2; ++dk) {
            const int32_t kk = bz + dk;
            if (kk < 0 || kk >= (int32_t)dim[2]) continue;
            const float wz = dk ? tz : 1.0f - tz, sz = dk ? 1.0f : -1.0f;
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const int32_t jj = by + dj;
                if (jj < 0 || jj >= (int32_t)dim[1]) continue;
                const float wy = dj ? ty : 1.0f - ty, sy = dj ? 1.0f : -1.0f;
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int32_t ii = bx + di;
                    if (ii < 0 || ii >= (int32_t)dim[0]) continue;
                    const float wx = di ? tx : 1.0f - tx, sx = di ? 1.0f : -1.0f;
                    const float w = wx * wy * wz;
                    if (w <= 0.0f) continue;
                    const uint32_t idx = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
                    const float nb = faces.numbar[a][idx], db = faces.denbar[a][idx];
                    vb += w * nb;
                    const float c = va * nb + db;
                    dx += sx * wy * wz * c;
                    dy += wx * sy * wz * c;
                    dz += wx * wy * sz * c;
                }
            }
        }
        out.gv[a][p] += vb;
        xb[0] += dx * P.inv_h;
        xb[1] += dy * P.inv_h;
        xb[2] += dz * P.inv_h;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) out.gx[c][p] += xb[c];
}

// Back to the order before the sort: g_in[perm[d]] = g_sorted[d].
__global__ __launch_bounds__(kAdjBlock) void k_adj_unsort(uint32_t n, const uint32_t* __restrict__ perm,
                                                          AdjParticles sorted, AdjParticles in_order) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const uint32_t s = perm[d];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        in_order.gx[a][s] = sorted.gx[a][d];
        in_order.gv[a][s] = sorted.gv[a][d];
    }
}

hipError_t launch_adj_g2p(const StepParams& P, const ParticleSoA& sorted, const AdjFaces& faces,
                          const AdjParticles& out, const AdjParticlesIn& in, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_g2p, dim3(grid_for(P.N, kAdjBlock)), dim3(kAdjBlock), 0, s, P, sorted, faces, out, in);
    return hipGetLastError();
}

hipError_t launch_adj_pbar(const StepParams& P, const uint8_t* ct, const AdjFaces& faces, float* pbar,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_adj_pbar, dim3(grid_for(P.C, kAdjBlock)), dim3(kAdjBlock), 0, s, P, ct, faces, pbar);
    return hipGetLastError();
}

hipError_t launch_adj_cg_init(const StepParams& P, const uint16_t* nmask, const float* pbar, float* rhs, float* z,
                              float* r, float* d, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_cg_init, dim3(grid_for(P.C, kAdjBlock)), dim3(kAdjBlock), 0, s, P.C, nmask, pbar, rhs,
                       z, r, d);
    return hipGetLastError();
}

hipError_t launch_adj_faces(const StepParams& P, const uint8_t* ct, const float* z, const AdjFaces& faces,
                            double* grav, hipStream_t s) {
    const size_t tot = (size_t)P.F[0] + P.F[1] + P.F[2];
    hipLaunchKernelGGL(k_adj_faces, dim3(grid_for(tot, kAdjBlock)), dim3(kAdjBlock), 0, s, P, ct, z, faces, grav);
    return hipGetLastError();
}

hipError_t launch_adj_p2g(const StepParams& P, const ParticleSoA& sorted, const AdjFaces& faces,
                          const AdjParticles& out, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_p2g, dim3(grid_for(P.N, kAdjBlock)), dim3(kAdjBlock), 0, s, P, sorted, faces, out);
    return hipGetLastError();
}

hipError_t launch_adj_unsort(uint32_t n, const uint32_t* perm, const AdjParticles& sorted,
                             const AdjParticles& in_order, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_unsort, dim3(grid_for(n, kAdjBlock)), dim3(kAdjBlock), 0, s, n, perm, sorted, in_order);
    return hipGetLastError();
}

}  // namespace fs
