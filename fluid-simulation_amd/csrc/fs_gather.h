// fs_gather.h — the sorted-segment P2G gather of one brick of cells, shared by
// the single-tank and the batched kernels.
//
// One workgroup owns a brick of cells (Brick<V>) and every MAC face
// whose cell coordinate (own-axis coordinate clamped to n-1) lies in the
// brick.  A face's taps come from particles whose cell is in the brick grown
// by one cell on every side; after the sort those particles are, per (y,z)
// row of that halo region, one contiguous segment of the sorted arrays.  The
// workgroup walks the segments, evaluates each particle's 24 taps exactly as
// sh/scatter.wgsl:78-118 does, and adds the ones it owns into LDS with
// integer atomics (order-free, so the sums equal the reference's i32 atomics
// bit for bit).  The epilogue normalizes, saves, applies gravity and enforces
// walls per owned face (fs_face.h), and stores vel/saved once.
//
// Particles whose P2G cell floor((pos-origin)*inv_h) differs from their sort
// cell (clamped floor((pos-origin)/h); the two differ by one ulp-level
// rounding when h is not a power of two) are listed by the key pass.  A
// listed particle whose sort cell lies outside the halo may still reach the
// brick, so every workgroup also walks the list and adds those particles'
// owned taps; particles inside the halo were already counted by the segment
// walk.  Each (particle, face) term is therefore added exactly once.
#pragma once
#include <type_traits>
#include "fs_particle.h"

namespace fs {

// Workgroup sizes of the gather (measured on MI355X).  One tank: 1024 threads
// when the grid has at most kGSmallGrid bricks, so a brick's particles are
// walked by more lanes at once (32^3, 64^3), else 256 (128^3 equal, 256^3
// faster than 512).  The batch, with its many small tanks, uses 512.
constexpr int kGThreads = 512;
constexpr int kGThreadsSmall = 1024;
constexpr int kGThreadsLarge = 256;
constexpr uint32_t kGSmallGrid = 1024;
constexpr int kGBatch = 4;  // particles whose loads a thread issues together

// Brick shapes of the gather menu: 0 = 8x8x8, 1 = 16x8x8, 2 = 32x4x4 cells
// (0 measured fastest at 128^3, 256^3 and on the 64-tank batch).
template <int V> struct Brick;
template <> struct Brick<0> { static constexpr int X = 8, Y = 8, Z = 8; };
template <> struct Brick<1> { static constexpr int X = 16, Y = 8, Z = 8; };
template <> struct Brick<2> { static constexpr int X = 32, Y = 4, Z = 4; };

struct GatherOut {
    float* vel[3];
    float* saved[3];
};

// Brick at cell (x0, y0, z0).  sp: the sorted side; src: the side the keys
// were computed on (listed indices refer to it); [pbeg, pend): the particle
// range of this tank in both sides.  With perm, sp is not sorted: sorted slot
// d holds particle perm[d] of sp (the particles were not moved).  weight_out
// (nullable) receives the number of particles the brick's halo holds.
template <int BF16, int V, int THREADS = kGThreads>
__device__ __forceinline__ void p2g_gather_brick(const StepParams& P, const ParticleSoA& sp, const ParticleSoA& src,
                                                 const uint32_t* __restrict__ occ, const uint32_t* __restrict__ off,
                                                 const uint8_t* __restrict__ ct, const GatherOut& out, int32_t x0,
                                                 int32_t y0, int32_t z0, const uint32_t* __restrict__ mis_list,
                                                 uint32_t mis_n, uint32_t pbeg, uint32_t pend,
                                                 const uint32_t* __restrict__ perm = nullptr,
                                                 uint32_t* __restrict__ weight_out = nullptr) {
    constexpr int kGBX = Brick<V>::X, kGBY = Brick<V>::Y, kGBZ = Brick<V>::Z;
    constexpr int kGRowsMax = (kGBY + 2) * (kGBZ + 2);
    constexpr int kGU = (kGBX + 1) * kGBY * kGBZ;
    constexpr int kGV = kGBX * (kGBY + 1) * kGBZ;
    constexpr int kGW = kGBX * kGBY * (kGBZ + 1);
    __shared__ int32_t s_num[kGU + kGV + kGW];
    __shared__ int32_t s_den[kGU + kGV + kGW];
    __shared__ uint32_t s_row_start[kGRowsMax];
    __shared__ uint32_t s_row_pref[kGRowsMax + 1];
    // Cell types of the brick grown by one cell on every side: every cell the
    // epilogue's wet and solid tests read for an owned face (cells outside
    // the grid are never read).  Loaded here, where their latency overlaps the
    // row-segment loads, so the epilogue waits on LDS only.
    constexpr int kCX = kGBX + 2, kCY = kGBY + 2, kCZ = kGBZ + 2;
    __shared__ uint8_t s_ct[kCX * kCY * kCZ];
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;

    for (int t = threadIdx.x; t < kGU + kGV + kGW; t += THREADS) {
        s_num[t] = 0;
        s_den[t] = 0;
    }
    for (int c = threadIdx.x; c < kCX * kCY * kCZ; c += THREADS) {
        const int32_t ci = x0 - 1 + c % kCX, cj = y0 - 1 + (c / kCX) % kCY, ck = z0 - 1 + c / (kCX * kCY);
        uint8_t v = 0xFFu;
        if (ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz)
            v = ct[(uint32_t)ci + P.nx * ((uint32_t)cj + P.ny * (uint32_t)ck)];
        s_ct[c] = v;
    }
    // Halo cell range (inclusive), clamped to the grid.
    const int32_t xs = max(x0 - 1, 0), xe = min(x0 + kGBX, nx - 1);
    const int32_t ys = max(y0 - 1, 0), ye = min(y0 + kGBY, ny - 1);
    const int32_t zs = max(z0 - 1, 0), ze = min(z0 + kGBZ, nz - 1);
    const int32_t rows_y = ye - ys + 1;
    const int32_t rows = rows_y * (ze - zs + 1);
    if (threadIdx.x < 64) {
        // Row segments and their prefix sums (one wave, rows <= kGRowsMax).
        uint32_t carry = 0;
        for (int r0 = 0; r0 < rows; r0 += 64) {
            const int r = r0 + (int)threadIdx.x;
            uint32_t len = 0, start = 0;
            if (r < rows) {
                const int32_t cy = ys + r % rows_y, cz = zs + r / rows_y;
                const uint32_t cf = (uint32_t)xs + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
                const uint32_t cl = (uint32_t)xe + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
                start = off[cf];
                len = off[cl] + occ[cl] - start;
                s_row_start[r] = start;
            }
            uint32_t inc = len;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(inc, o, 64);
                if ((int)threadIdx.x >= o) inc += t;
            }
            if (r < rows) s_row_pref[r + 1] = carry + inc;
            carry += __shfl(inc, 63, 64);
        }
        if (threadIdx.x == 0) s_row_pref[0] = 0;
    }
    __syncthreads();
    const uint32_t total = s_row_pref[rows];
    if (weight_out && threadIdx.x == 0) *weight_out = total;
    const int32_t ownx_hi = (x0 + kGBX >= nx) ? kGBX : kGBX - 1;  // own-axis local max index per axis
    const int32_t owny_hi = (y0 + kGBY >= ny) ? kGBY : kGBY - 1;
    const int32_t ownz_hi = (z0 + kGBZ >= nz) ? kGBZ : kGBZ - 1;

    auto add_particle = [&](const float px, const float py, const float pz, const float pv[3]) {
        const float pos[3] = {(px - P.origin[0]) * P.inv_h, (py - P.origin[1]) * P.inv_h,
                              (pz - P.origin[2]) * P.inv_h};
        // Below 2^24 grid units every g - floor(g) is exact and in [0, 1], so
        // every weight is in [0, 1].
        const bool moderate = fabsf(pos[0]) < 16777216.0f && fabsf(pos[1]) < 16777216.0f &&
                              fabsf(pos[2]) < 16777216.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t dim[3];
            face_dims(P, a, dim);
            // pos + stagger; the reference's + 0.0 on the own axis is left
            // out: it only turns -0 into +0, which gives the same integers.
            const float g[3] = {a == 0 ? pos[0] : pos[0] - 0.5f, a == 1 ? pos[1] : pos[1] - 0.5f,
                                a == 2 ? pos[2] : pos[2] - 0.5f};
            const int32_t lx_dim = a == 0 ? kGBX + 1 : kGBX;
            const int32_t ly_dim = a == 1 ? kGBY + 1 : kGBY;
            const uint32_t base = a == 0 ? 0u : (a == 1 ? (uint32_t)kGU : (uint32_t)(kGU + kGV));
            // A face index c is owned when 0 <= c - corner <= span: inside the
            // brick's owned range (own-axis high face only at the grid's end)
            // and inside the grid (c >= 0 follows from corner >= 0).  In
            // unsigned arithmetic that is one compare, and a saturated base
            // index wraps to a value above every span.
            const uint32_t spx = (uint32_t)min(a == 0 ? ownx_hi : kGBX - 1, (int32_t)dim[0] - 1 - x0);
            const uint32_t spy = (uint32_t)min(a == 1 ? owny_hi : kGBY - 1, (int32_t)dim[1] - 1 - y0);
            const uint32_t spz = (uint32_t)min(a == 2 ? ownz_hi : kGBZ - 1, (int32_t)dim[2] - 1 - z0);
            // The taps of p2g_taps (sh/scatter.wgsl:78-118) as straight-line
            // code: per dimension, whether each of the two face indices is
            // owned by this brick; a tap is added when all three are.  Same
            // expressions, same sums.
            const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
            const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
            const uint32_t ux = (uint32_t)bx - (uint32_t)x0, uy = (uint32_t)by - (uint32_t)y0,
                           uz = (uint32_t)bz - (uint32_t)z0;
            const bool ox[2] = {ux <= spx, ux + 1u <= spx}, oy[2] = {uy <= spy, uy + 1u <= spy},
                       oz[2] = {uz <= spz, uz + 1u <= spz};
            if (!((ox[0] || ox[1]) && (oy[0] || oy[1]) && (oz[0] || oz[1]))) continue;
            const float vcomp = pv[a];
            const float wx[2] = {1.0f - tx, tx}, wy[2] = {1.0f - ty, ty}, wz[2] = {1.0f - tz, tz};
            float wxy[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) wxy[q] = wx[q & 1] * wy[q >> 1];
            // LDS slots of the taps: constant offsets from s_num + li0 (and
            // s_den + li0).  The empty asm keeps the byte offset computed once
            // per axis; otherwise the compiler sinks the shift into each tap's
            // branch.
            const uint32_t li0 = base + ux + (uint32_t)lx_dim * (uy + (uint32_t)ly_dim * uz);
            uint32_t lb = li0 * 4u;
            asm volatile("" : "+v"(lb));
            int32_t* pn = (int32_t*)((char*)s_num + lb);
            int32_t* pd = (int32_t*)((char*)s_den + lb);
            // Fixed point, rint then i32 (sh/scatter.wgsl:78-118).  The
            // reference skips taps with w <= 0; here w is never negative on an
            // owned tap, and w == 0 (or NaN, from a NaN position) converts to
            // 0 in both terms, so adding it changes no sum.  P.scale is a
            // power of two of at most 2^22 (2^16; 2^22 where the adjoint
            // replays a step).  The short form (moderate position,
            // |v| < 1e30): v * scale is exact, so rint(w * (v * scale))
            // equals rint((w * v) * scale) (both are 0 where w * v is
            // subnormal); w * scale is exact and at most 2^22, so one fused
            // add of 2^23 rounds it to the nearest even integer, which then
            // sits in the low bits.  Other lanes take the reference's
            // expressions with saturating conversions.
            const float vs = vcomp * P.scale;
            auto add_taps = [&](auto short_form) {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int di = t & 1, dj = (t >> 1) & 1, dk = t >> 2;
                    const float w = wxy[t & 3] * wz[dk];  // wx * wy * wz
                    if (ox[di] && oy[dj] && oz[dk]) {
                        const int32_t lo = di + lx_dim * (dj + ly_dim * dk);
                        int32_t tn, td;
                        if constexpr (decltype(short_form)::value) {
                            tn = sat_i32(rintf(w * vs));
                            td = __float_as_int(__builtin_fmaf(w, P.scale, 8388608.0f)) - 0x4B000000;
                        } else {
                            tn = sat_i32(rintf(w * vcomp * P.scale));
                            td = sat_i32(rintf(w * P.scale));
                        }
                        atomicAdd(pn + lo, tn);
                        atomicAdd(pd + lo, td);
                    }
                }
            };
            if (moderate && fabsf(vcomp) < 1.0e30f)
                add_taps(std::true_type{});
            else
                add_taps(std::false_type{});
        }
    };

    // Items are taken kGBatch at a time per thread: the permutation entries,
    // then the positions and velocities of all of them are loaded before any
    // is evaluated, so a thread keeps kGBatch particles' loads in flight
    // instead of waiting out one particle's load chain at a time.
    int row = 0;
    for (uint32_t base = 0; base < total; base += THREADS * kGBatch) {
        uint32_t pk[kGBatch];
        bool live[kGBatch];
#pragma unroll
        for (int k = 0; k < kGBatch; ++k) {
            // Each wave takes every (THREADS / 64)-th item of its chunk, so its lanes fall in
            // different cells (sorted particles of one cell share faces, and
            // same-address LDS atomics within a wave serialize).
            const uint32_t t = base + (uint32_t)k * THREADS + (threadIdx.x & 63u) * (THREADS / 64) + (threadIdx.x >> 6);
            live[k] = t < total;
            pk[k] = 0;
            if (live[k]) {
                // The row holding flattened item t (t grows with k; linear advance).
                while (s_row_pref[row + 1] <= t) ++row;
                pk[k] = s_row_start[row] + (t - s_row_pref[row]);
            }
        }
        if (perm) {
#pragma unroll
            for (int k = 0; k < kGBatch; ++k)
                if (live[k]) pk[k] = ld_at(perm, pk[k]);
        }
        float px[kGBatch], py[kGBatch], pz[kGBatch], vx[kGBatch], vy[kGBatch], vz[kGBatch];
#pragma unroll
        for (int k = 0; k < kGBatch; ++k) {
            if (!live[k]) continue;
            const uint32_t p = pk[k];
            px[k] = ld_at(sp.pos[0], p);
            py[k] = ld_at(sp.pos[1], p);
            pz[k] = ld_at(sp.pos[2], p);
            vx[k] = load_vel<BF16>(sp.vel[0], p);
            vy[k] = load_vel<BF16>(sp.vel[1], p);
            vz[k] = load_vel<BF16>(sp.vel[2], p);
        }
#pragma unroll
        for (int k = 0; k < kGBatch; ++k) {
            if (!live[k]) continue;
            const float pv[3] = {vx[k], vy[k], vz[k]};
            add_particle(px[k], py[k], pz[k], pv);
        }
    }
    for (uint32_t e = threadIdx.x; e < mis_n; e += THREADS) {
        const uint32_t p = mis_list[e];
        if (p < pbeg || p >= pend) continue;
        const float x = src.pos[0][p], y = src.pos[1][p], z = src.pos[2][p];
        const uint32_t key = cell_key(P, x, y, z);
        const int32_t ci = (int32_t)(key % P.nx), cj = (int32_t)((key / P.nx) % P.ny), ck = (int32_t)(key / (P.nx * P.ny));
        if (ci >= xs && ci <= xe && cj >= ys && cj <= ye && ck >= zs && ck <= ze) continue;
        const float pv[3] = {load_vel<BF16>(src.vel[0], p), load_vel<BF16>(src.vel[1], p),
                             load_vel<BF16>(src.vel[2], p)};
        add_particle(x, y, z, pv);
    }
    __syncthreads();
    // Epilogue over owned faces of the three axes.
    for (int a = 0; a < 3; ++a) {
        const int32_t lxd = a == 0 ? kGBX + 1 : kGBX;
        const int32_t lyd = a == 1 ? kGBY + 1 : kGBY;
        const int32_t lzd = a == 2 ? kGBZ + 1 : kGBZ;
        const int32_t hx = a == 0 ? ownx_hi : kGBX - 1;
        const int32_t hy = a == 1 ? owny_hi : kGBY - 1;
        const int32_t hz = a == 2 ? ownz_hi : kGBZ - 1;
        const int32_t base = a == 0 ? 0 : (a == 1 ? kGU : kGU + kGV);
        float* vel = out.vel[a];
        float* sav = out.saved[a];
        uint32_t dim[3];
        face_dims(P, a, dim);
        for (int t = threadIdx.x; t < lxd * lyd * lzd; t += THREADS) {
            const int32_t lx = t % lxd, ly = (t / lxd) % lyd, lz = t / (lxd * lyd);
            if (lx > hx || ly > hy || lz > hz) continue;
            const int32_t i = x0 + lx, j = y0 + ly, k = z0 + lz;
            if (i >= (int32_t)dim[0] || j >= (int32_t)dim[1] || k >= (int32_t)dim[2]) continue;
            float vv, ss;
            face_epilogue_by(P, a, (uint32_t)i, (uint32_t)j, (uint32_t)k, s_num[base + t], s_den[base + t], vv, ss,
                             [&](uint32_t ci, uint32_t cj, uint32_t ck) {
                                 return s_ct[(ci - (uint32_t)(x0 - 1)) +
                                             (uint32_t)kCX * ((cj - (uint32_t)(y0 - 1)) +
                                                              (uint32_t)kCY * (ck - (uint32_t)(z0 - 1)))];
                             });
            const uint32_t f = (uint32_t)i + dim[0] * ((uint32_t)j + dim[1] * (uint32_t)k);
            vel[f] = vv;
            sav[f] = ss;
        }
    }
}

// ---- particle-owned form: every particle's taps evaluated once ----
// One workgroup per 8x8x8 brick takes the particles whose sort cell lies in
// the brick (one contiguous segment per (y,z) row of the brick after the
// sort) and adds all 24 taps of each into LDS with integer atomics.  For axis
// a the LDS region holds the faces such a particle can reach: face indices
// [x0, x0 + 8] along a and [x0 - 1, x0 + 8] across it (9 x 10 x 10 faces), so
// no ownership test and no halo walk is needed, and no tap is evaluated by
// two workgroups.  A tap outside the grid adds 0 to a region slot that is
// never flushed; a particle whose P2G cell differs from its sort cell (the
// misfit of the gather above) or that lies outside the tank may reach past
// the region, and only its taps take the per-tap path, adding out-of-region
// taps to num/den with global atomics.  Afterwards the region's non-zero
// sums are added to num/den with global i32 atomics: a face's total is the
// sum of its terms over every brick that reached it, in any order (wrapping
// integer adds), so num/den equal the reference's per-tap atomics bit for
// bit.  num/den start cleared; the face epilogue runs as its own pass.
constexpr int kSB = 8;                      // brick edge (cells)
constexpr int kSRA = kSB + 1, kSRO = kSB + 2;  // region extent along / across the axis
constexpr int kSRegion = kSRA * kSRO * kSRO;   // faces per axis
constexpr uint32_t kSPerThread = 4;         // consecutive particle slots a thread takes per round
constexpr int kSThreads = 256;              // workgroup size; small grids (<= kGSmallGrid bricks) use
constexpr int kSThreadsSmall = 512;         // more lanes per brick
constexpr uint32_t kSDefer = 256;           // particles per brick whose P2G cell is not in it, listed

// The taps of one particle and axis whose stencil leaves the brick's region
// (a misfit or a particle outside the tank), as p2g_taps: in-region taps into
// LDS, the others straight into num/den with global atomics.  Rare.
__device__ __forceinline__ void scatter_taps_outside(const StepParams& P, int a, const float g[3], float vcomp,
                                                  int32_t x0, int32_t y0, int32_t z0, int32_t* s_num,
                                                  int32_t* s_den, int32_t* num, int32_t* den) {
    uint32_t dim[3];
    face_dims(P, a, dim);
    const int32_t rdx = a == 0 ? kSRA : kSRO, rdy = a == 1 ? kSRA : kSRO, rdz = a == 2 ? kSRA : kSRO;
    const int32_t rx = x0 - (a != 0), ry = y0 - (a != 1), rz = z0 - (a != 2);
    p2g_taps(g, dim, vcomp, P.scale, [&](int32_t ii, int32_t jj, int32_t kk, int32_t tn, int32_t td) {
        const int32_t mx = ii - rx, my = jj - ry, mz = kk - rz;
        if (mx >= 0 && mx < rdx && my >= 0 && my < rdy && mz >= 0 && mz < rdz) {
            const int32_t li = a * kSRegion + mx + rdx * (my + rdy * mz);
            atomicAdd(&s_num[li], tn);
            atomicAdd(&s_den[li], td);
        } else {
            const uint32_t f = (uint32_t)ii + dim[0] * ((uint32_t)jj + dim[1] * (uint32_t)kk);
            atomicAdd(&num[f], tn);
            atomicAdd(&den[f], td);
        }
    });
}

template <int BF16, int THREADS>
__device__ __forceinline__ void p2g_scatter_brick(const StepParams& P, const ParticleSoA& sp,
                                                  const uint32_t* __restrict__ occ, const uint32_t* __restrict__ off,
                                                  int32_t* const num[3], int32_t* const den[3], int32_t x0, int32_t y0,
                                                  int32_t z0, const uint32_t* __restrict__ perm) {
    __shared__ int32_t s_num[3 * kSRegion];
    __shared__ int32_t s_den[3 * kSRegion];
    __shared__ uint32_t s_row_start[kSB * kSB];
    __shared__ uint32_t s_row_pref[kSB * kSB + 1];
    __shared__ uint32_t s_defer[kSDefer];
    __shared__ uint32_t s_ndefer;
    const int32_t nx = (int32_t)P.nx, ny = (int32_t)P.ny, nz = (int32_t)P.nz;
    const int32_t xe = min(x0 + kSB - 1, nx - 1);
    if (threadIdx.x < 64) {
        // Row r = (y0 + r % 8, z0 + r / 8): its particles are those of cells
        // [x0, xe] of that row, one segment of the sorted order.
        const int r = (int)threadIdx.x;
        const int32_t cy = y0 + r % kSB, cz = z0 + r / kSB;
        uint32_t len = 0, start = 0;
        if (cy < ny && cz < nz) {
            const uint32_t cf = (uint32_t)x0 + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
            const uint32_t cl = (uint32_t)xe + P.nx * ((uint32_t)cy + P.ny * (uint32_t)cz);
            start = off[cf];
            len = off[cl] + occ[cl] - start;
        }
        s_row_start[r] = start;
        uint32_t inc = len;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o, 64);
            if (r >= o) inc += t;
        }
        s_row_pref[r + 1] = inc;
        if (r == 0) {
            s_row_pref[0] = 0;
            s_ndefer = 0;
        }
    }
    for (int t = threadIdx.x; t < 3 * kSRegion; t += THREADS) {
        s_num[t] = 0;
        s_den[t] = 0;
    }
    __syncthreads();
    const uint32_t total = s_row_pref[kSB * kSB];
    if (total == 0) return;  // no particle: nothing to add (num/den start cleared)

    // The particle's P2G cell floor((pos - origin) * inv_h) lies in the brick
    // for a particle whose sort cell does, except a misfit or a particle
    // outside the tank; then every tap of every axis has a region slot.
    auto in_brick = [&](float px, float py, float pz) -> bool {
        const float q[3] = {(px - P.origin[0]) * P.inv_h, (py - P.origin[1]) * P.inv_h,
                            (pz - P.origin[2]) * P.inv_h};
        const int32_t o[3] = {x0, y0, z0};
        bool in = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int32_t cc = sat_i32(floorf(q[c]));
            in = in && cc >= o[c] && cc < o[c] + kSB;
        }
        return in;
    };
    // All 24 taps of a particle whose P2G cell lies in the brick, into LDS.
    auto add_particle = [&](const float px, const float py, const float pz, const float pv[3]) {
        const float pos[3] = {(px - P.origin[0]) * P.inv_h, (py - P.origin[1]) * P.inv_h,
                              (pz - P.origin[2]) * P.inv_h};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t dim[3];
            face_dims(P, a, dim);
            const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
            const int32_t bx = sat_i32(floorf(g[0])), by = sat_i32(floorf(g[1])), bz = sat_i32(floorf(g[2]));
            const float tx = g[0] - (float)bx, ty = g[1] - (float)by, tz = g[2] - (float)bz;
            // Region of axis a: origin (x0, y0, z0) less one cell across the
            // axis; extents kSRA along it, kSRO across.
            const int32_t rdx = a == 0 ? kSRA : kSRO, rdy = a == 1 ? kSRA : kSRO;
            const int32_t lx = bx - (x0 - (a != 0)), ly = by - (y0 - (a != 1)), lz = bz - (z0 - (a != 2));
            bool gx[2], gy[2], gz[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                gx[d] = bx + d >= 0 && bx + d < (int32_t)dim[0];
                gy[d] = by + d >= 0 && by + d < (int32_t)dim[1];
                gz[d] = bz + d >= 0 && bz + d < (int32_t)dim[2];
            }
            const float vcomp = pv[a];
            const float wx[2] = {1.0f - tx, tx}, wy[2] = {1.0f - ty, ty}, wz[2] = {1.0f - tz, tz};
            float wxy[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) wxy[q] = wx[q & 1] * wy[q >> 1];
            const int32_t base = a * kSRegion + lx + rdx * (ly + rdy * lz);
            // Fixed point as in p2g_taps (sh/scatter.wgsl:78-118): w in (0, 1]
            // at scale 2^16 never saturates; the velocity term cannot when
            // |v| * scale < 2e9 (false for NaN), else the saturating form.
            // A tap that is skipped (outside the grid, or w <= 0) adds 0 to
            // its slot instead of branching; slots outside the grid are never
            // flushed.
            auto add_taps = [&](auto plain_v) {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int di = t & 1, dj = (t >> 1) & 1, dk = t >> 2;
                    const float w = wxy[t & 3] * wz[dk];  // wx * wy * wz
                    const bool take = gx[di] && gy[dj] && gz[dk] && w > 0.0f;
                    const float tv = rintf(w * vcomp * P.scale);
                    const int32_t tn = take ? (decltype(plain_v)::value ? (int32_t)tv : sat_i32(tv)) : 0;
                    const int32_t td = take ? (int32_t)rintf(w * P.scale) : 0;
                    const int32_t li = base + di + rdx * (dj + rdy * dk);
                    atomicAdd(&s_num[li], tn);
                    atomicAdd(&s_den[li], td);
                }
            };
            if (fabsf(vcomp) * P.scale < 2.0e9f)
                add_taps(std::true_type{});
            else
                add_taps(std::false_type{});
        }
    };
    // A particle whose P2G cell is not in the brick, tap by tap as p2g_taps:
    // taps in the region into LDS, the others into num/den with global atomics.
    auto add_outside = [&](uint32_t p) {
        const float pos[3] = {(sp.pos[0][p] - P.origin[0]) * P.inv_h, (sp.pos[1][p] - P.origin[1]) * P.inv_h,
                              (sp.pos[2][p] - P.origin[2]) * P.inv_h};
        const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p),
                             load_vel<BF16>(sp.vel[2], p)};
        for (int a = 0; a < 3; ++a) {
            const float g[3] = {pos[0] + stagger(a, 0), pos[1] + stagger(a, 1), pos[2] + stagger(a, 2)};
            scatter_taps_outside(P, a, g, pv[a], x0, y0, z0, s_num, s_den, num[a], den[a]);
        }
    };
    // The brick's particle slots: each round gives a thread kSPerThread
    // consecutive items, taken one after another; the lanes of a wave are
    // kSPerThread * THREADS / 64 items apart, so they fall in different
    // cells, whose taps mostly land on different LDS slots (same-address
    // atomics within a wave serialize).
    constexpr uint32_t kLaneStride = kSPerThread * (THREADS / 64);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    auto walk = [&](auto second) {
        int row = 0;
        for (uint32_t base = 0; base < total; base += THREADS * kSPerThread) {
#pragma unroll 1
            for (uint32_t m = 0; m < kSPerThread; ++m) {
                const uint32_t t = base + lane * kLaneStride + wave * kSPerThread + m;
                if (t >= total) break;
                while (s_row_pref[row + 1] <= t) ++row;
                uint32_t p = s_row_start[row] + (t - s_row_pref[row]);
                if (perm) p = perm[p];
                const float px = sp.pos[0][p], py = sp.pos[1][p], pz = sp.pos[2][p];
                if (decltype(second)::value) {
                    // Only when the deferred list overflowed: the particles
                    // not in the list, tap by tap.
                    if (in_brick(px, py, pz)) continue;
                    bool listed = false;
                    for (uint32_t e = 0; e < kSDefer; ++e) listed = listed || s_defer[e] == p;
                    if (!listed) add_outside(p);
                    continue;
                }
                if (!in_brick(px, py, pz)) {
                    const uint32_t slot = atomicAdd(&s_ndefer, 1u);
                    if (slot < kSDefer) s_defer[slot] = p;
                    continue;
                }
                const float pv[3] = {load_vel<BF16>(sp.vel[0], p), load_vel<BF16>(sp.vel[1], p),
                                     load_vel<BF16>(sp.vel[2], p)};
                add_particle(px, py, pz, pv);
            }
        }
    };
    walk(std::false_type{});
    __syncthreads();
    // Particles whose P2G cell is not in the brick (rare): the list, and if it
    // overflowed, a second walk for the rest.
    const uint32_t nd = s_ndefer;
    for (uint32_t e = threadIdx.x; e < (nd < kSDefer ? nd : kSDefer); e += THREADS) add_outside(s_defer[e]);
    if (nd > kSDefer) walk(std::true_type{});
    __syncthreads();
    // Flush: the region's in-grid, non-zero sums into num/den (the axis loop
    // is unrolled, so every array is picked at compile time).
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t rdx = a == 0 ? kSRA : kSRO, rdy = a == 1 ? kSRA : kSRO;
        uint32_t dim[3];
        face_dims(P, a, dim);
        int32_t* __restrict__ na = num[a];
        int32_t* __restrict__ da = den[a];
        for (int li = threadIdx.x; li < kSRegion; li += THREADS) {
            const int32_t n_ = s_num[a * kSRegion + li], d_ = s_den[a * kSRegion + li];
            if ((n_ | d_) == 0) continue;
            const int32_t i = x0 - (a != 0) + li % rdx, j = y0 - (a != 1) + (li / rdx) % rdy,
                          k = z0 - (a != 2) + li / (rdx * rdy);
            if (i < 0 || j < 0 || k < 0 || i >= (int32_t)dim[0] || j >= (int32_t)dim[1] || k >= (int32_t)dim[2])
                continue;
            const uint32_t f = (uint32_t)i + dim[0] * ((uint32_t)j + dim[1] * (uint32_t)k);
            atomicAdd(&na[f], n_);
            atomicAdd(&da[f], d_);
        }
    }
}

}  // namespace fs
