// fs_cg.h — building blocks of the mixed-precision CG shared by the
// single-tank and batched solvers: deterministic f64 block/partial sums and
// the per-tile work of one iteration (kU tiles of 64 cells per wave step).
#pragma once
#include "fs_kernels.h"
#include "fs_face.h"

namespace fs {

constexpr int kTile = 64;  // cells per CG tile (one wave)
constexpr int kU = 4;      // default tiles a wave loads before computing (memory-level parallelism)
constexpr long long kSpinTimeoutTicks = 400000000LL;  // wall_clock64 runs at 100 MHz
// Stencil mask of a liquid cell: bits 0-2 non-solid neighbour count, bits 3-8
// liquid neighbour in the order -x,+x,-y,+y,-z,+z, bit 9 the cell is liquid.
constexpr uint16_t kMaskLiquid = 1u << 9;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double* s_w) {
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) s_w[wv] = v;
    __syncthreads();
    double bs = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < THREADS / 64; ++i) bs += s_w[i];
    return bs;
}

template <int THREADS>
__device__ __forceinline__ double sum_partials(const double* partials, uint32_t n, double* s_tree) {
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += THREADS)
        acc += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_tree[threadIdx.x] = acc;
    __syncthreads();
    for (int d = THREADS / 2; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) s_tree[threadIdx.x] += s_tree[threadIdx.x + d];
        __syncthreads();
    }
    return s_tree[0];
}

// Block partials of NV values: wave sums (wave_sum, one fixed order), then
// thread 0 adds the waves in order.  Valid in thread 0 only.
template <int THREADS, int NV>
__device__ __forceinline__ void block_sum_nv(const double (&v)[NV], double (&bs)[NV]) {
    __shared__ double s_w[NV][THREADS / 64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const double t = wave_sum(v[j]);
        if (lane == 0) s_w[j][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double t = 0.0;
            for (int i = 0; i < THREADS / 64; ++i) t += s_w[j][i];
            bs[j] = t;
        }
    }
}

// Total of G published partials in one fixed order: thread i has added
// partials i, i + THREADS, ... into `loaded`; waves reduce with wave_sum;
// thread 0 adds the wave results in order.  Every block
// therefore forms the bit-identical total, returned to all threads.
template <int THREADS, int NV>
__device__ __forceinline__ void reduce_loaded(const double (&loaded)[NV], double (&out)[NV]) {
    __shared__ double s_w[NV][THREADS / 64];
    __shared__ double s_out[NV];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const double t = wave_sum(loaded[j]);
        if (lane == 0) s_w[j][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double t = 0.0;
            for (int i = 0; i < THREADS / 64; ++i) t += s_w[j][i];
            s_out[j] = t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) out[j] = s_out[j];
}

// Epoch-tagged slots of the counter-free exchanges.  A slot is two 64-bit
// words: the high half of an f64 under tag `epoch`, and the low half under
// tag (epoch << 1 | flag).  Each word is written and read whole with an
// agent-scope relaxed atomic (memory-side, around the per-XCD L2), so a
// reader that sees both tags equal to the epoch it waits for holds that
// epoch's value.  Epochs are below 2^31 and start at 1 over cleared slots.
// A tagged slot is two u64 words (hi: epoch | high half of the value; lo:
// epoch << 1 | flag | low half), `stride` words apart from the next slot.
// The single-tank solve spaces its slots kSlotStride words (256 bytes)
// apart: every workgroup polls every slot, and slots packed into a few
// cache lines make those lines' memory channels the bottleneck of the
// exchange (measured at 208 workgroups on MI355X: 10.9 us per iteration
// packed, 9.6 us spaced).
constexpr uint32_t kSlotStride = 32;
__device__ __forceinline__ void publish_tagged(uint64_t* half, uint32_t slot, uint32_t epoch, double v,
                                               uint32_t flag, uint32_t stride = 2u) {
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    __hip_atomic_store(half + stride * slot, ((uint64_t)epoch << 32) | (bits >> 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(half + stride * slot + 1u, ((uint64_t)((epoch << 1) | (flag & 1u)) << 32) | (bits & 0xffffffffull),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits for slot `slot` to carry `epoch`; false (and the abort flag raised)
// on abort or after `limit` wall_clock64 ticks.
__device__ __forceinline__ bool poll_tagged(const uint64_t* half, uint32_t slot, uint32_t epoch, uint32_t* abort_flag,
                                            double& v, uint32_t& flag, long long limit = kSpinTimeoutTicks,
                                            uint32_t stride = 2u) {
    const long long t0 = wall_clock64();
    for (uint32_t spin = 0;; ++spin) {
        const uint64_t hi = __hip_atomic_load(half + stride * slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t lo = __hip_atomic_load(half + stride * slot + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)(hi >> 32) == epoch && (uint32_t)(lo >> 33) == epoch) {
            v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
            flag = (uint32_t)(lo >> 32) & 1u;
            return true;
        }
        if ((spin & 31u) == 31u && (ld_relaxed(abort_flag) || wall_clock64() - t0 > limit)) {
            __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
}

// Per-cell work, U tiles at a time: all loads of the batch are issued before
// any result is computed, so a wave keeps several memory round trips in flight.
struct TileWalk {
    uint32_t begin, end, stride;
};

__device__ __forceinline__ TileWalk tile_walk(uint32_t nt, uint32_t logical_block, uint32_t G, uint32_t waves) {
    TileWalk t;
    const uint32_t b0 = (uint32_t)(((uint64_t)nt * logical_block) / G);
    t.begin = b0 + (threadIdx.x >> 6);
    t.end = (uint32_t)(((uint64_t)nt * (logical_block + 1u)) / G);
    t.stride = waves;
    return t;
}

// Store of a value other blocks read after the next grid barrier: an
// agent-scope relaxed atomic store goes through the XCD's L2 to memory at
// once, so the barrier's release finds the line clean instead of writing
// it back then.
__device__ __forceinline__ void store_through(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Element i read with an agent-scope relaxed atomic load (an sc1 load): it
// bypasses the CU's L1 and returns the value at the agent's coherence point,
// so it sees every store_through whose acknowledgement its writer waited for
// before the grid barrier it crossed (coherent_sum), with no cache
// invalidation between.
template <typename T>
__device__ __forceinline__ T ld_coherent_at(const T* base, uint32_t i) {
    return __hip_atomic_load(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T)),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Phase A for tiles ti, ti+stride, ... (U of them): dn = first ? d : r + beta*d_prev,
// q = A dn (liquid cells; 0 elsewhere); returns this lane's dn.q.  COH (the
// fence-free barrier) stores d with store_through and reads the neighbours
// with ld_coherent_at; a lane's own cells are read back by the wave that
// wrote them.
template <int U, bool COH = false, typename TileId>
__device__ __forceinline__ double phase_a_batch(const StepParams& P, const uint16_t* __restrict__ nmask,
                                                const float* __restrict__ r, const float* __restrict__ dprev,
                                                float* __restrict__ dnext, float* __restrict__ q, float beta,
                                                bool first, const TileWalk& tw, uint32_t ti, TileId&& tile_id) {
    const uint32_t lane = lane_id(), C = P.C;
    const int32_t off[6] = {-1, 1, -(int32_t)P.nx, (int32_t)P.nx, -(int32_t)(P.nx * P.ny), (int32_t)(P.nx * P.ny)};
    uint32_t c[U];
    uint16_t m[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        c[u] = t < tw.end ? tile_id(t) * kTile + lane : C;
    }
    // Every load is issued by every lane, with no branch and without waiting
    // for the mask: indices are clamped into the grid (a lane past the tile
    // list reads cell C - 1), and a value is used only for a cell in the grid
    // whose stencil mask has that term (r only after the first iteration).
    // A wave thus pays one memory round trip per batch of tiles.
    uint16_t mv[U];
    float rc[U], dc[U], rn[U][6], dnb[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t cc = c[u] < C ? c[u] : C - 1u;
        mv[u] = ld_at(nmask, cc);
        rc[u] = ld_at(r, cc);
        dc[u] = ld_at(dprev, cc);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            int32_t nb = (int32_t)c[u] + off[s];
            nb = nb < 0 ? 0 : (nb >= (int32_t)C ? (int32_t)C - 1 : nb);
            rn[u][s] = COH ? ld_coherent_at(r, (uint32_t)nb) : ld_at(r, (uint32_t)nb);
            dnb[u][s] = COH ? ld_coherent_at(dprev, (uint32_t)nb) : ld_at(dprev, (uint32_t)nb);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) m[u] = c[u] < C ? mv[u] : (uint16_t)0;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c[u] >= C) continue;
        const float dn = first ? dc[u] : rc[u] + beta * dc[u];
        float qc = 0.0f;
        if (m[u] & kMaskLiquid) {
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < 6; ++s)
                if ((m[u] >> (3 + s)) & 1u) sum += first ? dnb[u][s] : rn[u][s] + beta * dnb[u][s];
            qc = (float)(m[u] & 7u) * dn - sum;
        }
        if (!first) {
            if (COH)
                store_through(&dnext[c[u]], dn);
            else
                st_at(dnext, c[u], dn);
        }
        st_at(q, c[u], qc);
        acc += (double)dn * (double)qc;
    }
    return acc;
}

// Phase B: r -= alpha q and r.r, and the p update selected by PMODE:
// kPOne p += alpha dn; kPNone p left alone; kPTwo p = (p + alpha_prev dprev)
// + alpha dn, the deferred update of iteration k-1 and the one of iteration
// k in their own order (the same f32 operations as one update per
// iteration, so deferring every other update changes no bit).
// COH stores r with store_through (the fence-free barrier).
constexpr int kPNone = 0, kPOne = 1, kPTwo = 2;
template <int U, int PMODE, bool COH = false, typename TileId>
__device__ __forceinline__ double phase_b_batch(const StepParams& P, const float* __restrict__ dn,
                                                const float* __restrict__ q, float* __restrict__ p,
                                                float* __restrict__ r, float alpha, const float* __restrict__ dprev,
                                                float alpha_prev, const TileWalk& tw, uint32_t ti, TileId&& tile_id) {
    const uint32_t lane = lane_id(), C = P.C;
    uint32_t c[U];
    float pv[U], dv[U], dpv[U], rv[U], qv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        c[u] = t < tw.end ? tile_id(t) * kTile + lane : C;
    }
    // Unconditional loads at a clamped index, as in phase_a_batch.
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t cc = c[u] < C ? c[u] : C - 1u;
        if (PMODE != kPNone) {
            pv[u] = ld_at(p, cc);
            dv[u] = ld_at(dn, cc);
        }
        if (PMODE == kPTwo) dpv[u] = ld_at(dprev, cc);
        rv[u] = ld_at(r, cc);
        qv[u] = ld_at(q, cc);
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c[u] >= C) continue;
        if (PMODE == kPOne) st_at(p, c[u], pv[u] + alpha * dv[u]);
        if (PMODE == kPTwo) st_at(p, c[u], (pv[u] + alpha_prev * dpv[u]) + alpha * dv[u]);
        const float rn = rv[u] - alpha * qv[u];
        if (COH)
            store_through(&r[c[u]], rn);
        else
            st_at(r, c[u], rn);
        acc += (double)rn * (double)rn;
    }
    return acc;
}

// The last pending p update (p += alpha dn) after an iteration whose update
// was deferred.
template <int U, typename TileId>
__device__ __forceinline__ void p_update_batch(const StepParams& P, const float* __restrict__ dn,
                                               float* __restrict__ p, float alpha, const TileWalk& tw, uint32_t ti,
                                               TileId&& tile_id) {
    const uint32_t lane = lane_id(), C = P.C;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        const uint32_t c = t < tw.end ? tile_id(t) * kTile + lane : C;
        if (c < C) st_at(p, c, ld_at(p, c) + alpha * ld_at(dn, c));
    }
}

// ---- register-resident form of the two phases ----
// When a wave's tiles fit in one batch of U, its cells' mask, r, d, p and q
// stay in registers for the whole solve: phase A loads only the neighbours'
// r and d and stores d (read by neighbours next iteration); phase B loads
// nothing and stores r; p is stored once after the loop.  The f32
// expressions are those of phase_a_batch / phase_b_batch, so both forms give
// the same bits.
template <int U>
struct ResidentCells {
    uint32_t c[U];
    uint16_t m[U];
    float r[U], d[U], p[U], q[U];
};

template <int U, typename TileId>
__device__ __forceinline__ double resident_load(const StepParams& P, const uint16_t* __restrict__ nmask,
                                                const float* __restrict__ r, const float* __restrict__ d,
                                                const float* __restrict__ p, const TileWalk& tw, TileId&& tile_id,
                                                ResidentCells<U>& rc) {
    const uint32_t lane = lane_id(), C = P.C;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = tw.begin + (uint32_t)u * tw.stride;
        rc.c[u] = t < tw.end ? tile_id(t) * kTile + lane : C;
        const bool ok = rc.c[u] < C;
        rc.m[u] = ok ? ld_at(nmask, rc.c[u]) : (uint16_t)0;
        rc.r[u] = ok ? ld_at(r, rc.c[u]) : 0.0f;
        rc.d[u] = ok ? ld_at(d, rc.c[u]) : 0.0f;
        rc.p[u] = ok ? ld_at(p, rc.c[u]) : 0.0f;
        rc.q[u] = 0.0f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (rc.c[u] < C) acc += (double)rc.r[u] * (double)rc.r[u];
    return acc;
}

// COH reads the neighbours with ld_coherent_at (the fence-free barrier).
template <int U, bool COH>
__device__ __forceinline__ double resident_phase_a(const StepParams& P, const float* __restrict__ r,
                                                   const float* __restrict__ dprev, float* __restrict__ dnext,
                                                   float beta, bool first, ResidentCells<U>& rc) {
    auto ld = [](const float* base, uint32_t i) -> float { return COH ? ld_coherent_at(base, i) : ld_at(base, i); };
    const uint32_t C = P.C;
    const int32_t off[6] = {-1, 1, -(int32_t)P.nx, (int32_t)P.nx, -(int32_t)(P.nx * P.ny), (int32_t)(P.nx * P.ny)};
    float rn[U][6], dnb[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        // The cell index is made opaque here (an empty asm, no instruction) so
        // the 6 x U neighbour addresses are formed per iteration instead of
        // being hoisted out of the solve loop and held in registers throughout.
        uint32_t cu = rc.c[u];
        asm volatile("" : "+v"(cu));
        // Every lane loads (the index is clamped into the grid), so no load
        // sits behind a branch; a value is used only for a cell in the grid
        // whose stencil mask has that neighbour, and r only after the first
        // iteration.
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            int32_t nb = (int32_t)cu + off[s];
            nb = nb < 0 ? 0 : (nb >= (int32_t)C ? (int32_t)C - 1 : nb);
            rn[u][s] = ld(r, (uint32_t)nb);
            dnb[u][s] = ld(dprev, (uint32_t)nb);
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (rc.c[u] >= C) continue;
        const uint16_t m = rc.m[u];
        const float dn = first ? rc.d[u] : rc.r[u] + beta * rc.d[u];
        float qc = 0.0f;
        if (m & kMaskLiquid) {
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < 6; ++s)
                if ((m >> (3 + s)) & 1u) sum += first ? dnb[u][s] : rn[u][s] + beta * dnb[u][s];
            qc = (float)(m & 7u) * dn - sum;
        }
        if (!first) store_through(&dnext[rc.c[u]], dn);
        rc.d[u] = dn;
        rc.q[u] = qc;
        acc += (double)dn * (double)qc;
    }
    return acc;
}

template <int U>
__device__ __forceinline__ double resident_phase_b(const StepParams& P, float* __restrict__ r, float alpha,
                                                   ResidentCells<U>& rc) {
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (rc.c[u] >= P.C) continue;
        rc.p[u] = rc.p[u] + alpha * rc.d[u];
        const float rn = rc.r[u] - alpha * rc.q[u];
        store_through(&r[rc.c[u]], rn);
        rc.r[u] = rn;
        acc += (double)rn * (double)rn;
    }
    return acc;
}

// ---- single-reduction form (Chronopoulos & Gear 1989) ----
// With s = A pd and w = A r kept as vectors, one sweep per iteration does
//     pd = r + beta pd;  s = w + beta s;  x += alpha pd;  r' = r - alpha s;
//     w' = A r';  partial sums of r'.r' and w'.r'
// where alpha and beta come from the previous sweep's two sums.  A r' needs r'
// at the six neighbours: the sweep recomputes it there from the neighbours'
// r, w and s with the owner's exact f32 expression, so one grid barrier per
// iteration suffices.  r, w and s are read at neighbours while their next
// values are written, so they alternate between two buffers.
struct CgSweepIo {
    const float* r_in;   // r_k
    const float* w_in;   // w_k = A r_k
    const float* s_in;   // s_{k-1} (unread when first)
    float* r_out;        // r_{k+1}
    float* w_out;        // w_{k+1}
    float* s_out;        // s_k
    float* pd;           // search direction, own cells only
    float* x;            // solution, own cells only
};

// Initial sweep: w_0 = A r_0 over the tiles; returns partial r.r and w.r.
template <int U, typename TileId>
__device__ __forceinline__ void cgs_init_batch(const StepParams& P, const uint16_t* __restrict__ nmask,
                                               const float* __restrict__ r, float* __restrict__ w_out,
                                               const TileWalk& tw, uint32_t ti, TileId&& tile_id, double& gam,
                                               double& del) {
    const uint32_t lane = lane_id(), C = P.C;
    const int32_t off[6] = {-1, 1, -(int32_t)P.nx, (int32_t)P.nx, -(int32_t)(P.nx * P.ny), (int32_t)(P.nx * P.ny)};
    uint32_t c[U];
    uint16_t m[U];
    float rc[U], rn[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        c[u] = t < tw.end ? tile_id(t) * kTile + lane : C;
        m[u] = c[u] < C ? ld_at(nmask, c[u]) : (uint16_t)0;
        const bool ok = c[u] < C;
        rc[u] = ok ? r[c[u]] : 0.0f;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            int32_t nb = (int32_t)c[u] + off[s];
            nb = nb < 0 ? 0 : (nb >= (int32_t)C ? (int32_t)C - 1 : nb);
            rn[u][s] = ok ? r[nb] : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c[u] >= C) continue;
        float wc = 0.0f;
        if (m[u] & kMaskLiquid) {
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < 6; ++s)
                if ((m[u] >> (3 + s)) & 1u) sum += rn[u][s];
            wc = (float)(m[u] & 7u) * rc[u] - sum;
        }
        w_out[c[u]] = wc;
        gam += (double)rc[u] * (double)rc[u];
        del += (double)wc * (double)rc[u];
    }
}

// One iteration's sweep over U tiles; accumulates r'.r' and w'.r'.
template <int U, typename TileId>
__device__ __forceinline__ void cgs_sweep_batch(const StepParams& P, const uint16_t* __restrict__ nmask,
                                                const CgSweepIo& io, float alpha, float beta, bool first,
                                                const TileWalk& tw, uint32_t ti, TileId&& tile_id, double& gam,
                                                double& del) {
    const uint32_t lane = lane_id(), C = P.C;
    const int32_t off[6] = {-1, 1, -(int32_t)P.nx, (int32_t)P.nx, -(int32_t)(P.nx * P.ny), (int32_t)(P.nx * P.ny)};
    const float* __restrict__ r = io.r_in;
    const float* __restrict__ w = io.w_in;
    const float* __restrict__ sp = io.s_in;
    uint32_t c[U];
    uint16_t m[U];
    float rc[U], wc[U], sc[U], pc[U], xc[U], rn[U][6], wn[U][6], sn[U][6];
    // Every load is issued before any result is used (clamped neighbour
    // indices; the mask selects the terms afterwards).
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        c[u] = t < tw.end ? tile_id(t) * kTile + lane : C;
        m[u] = c[u] < C ? ld_at(nmask, c[u]) : (uint16_t)0;
        const bool ok = c[u] < C;
        rc[u] = ok ? r[c[u]] : 0.0f;
        wc[u] = ok ? w[c[u]] : 0.0f;
        sc[u] = (ok && !first) ? sp[c[u]] : 0.0f;
        pc[u] = (ok && !first) ? io.pd[c[u]] : 0.0f;
        xc[u] = ok ? io.x[c[u]] : 0.0f;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            int32_t nb = (int32_t)c[u] + off[s];
            nb = nb < 0 ? 0 : (nb >= (int32_t)C ? (int32_t)C - 1 : nb);
            rn[u][s] = ok ? r[nb] : 0.0f;
            wn[u][s] = ok ? w[nb] : 0.0f;
            sn[u][s] = (ok && !first) ? sp[nb] : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (c[u] >= C) continue;
        const float s_k = first ? wc[u] : wc[u] + beta * sc[u];
        const float pd_k = first ? rc[u] : rc[u] + beta * pc[u];
        const float r1 = rc[u] - alpha * s_k;
        float w1 = 0.0f;
        if (m[u] & kMaskLiquid) {
            float sum = 0.0f;
#pragma unroll
            for (int s = 0; s < 6; ++s)
                if ((m[u] >> (3 + s)) & 1u) {
                    const float s_nb = first ? wn[u][s] : wn[u][s] + beta * sn[u][s];
                    sum += rn[u][s] - alpha * s_nb;
                }
            w1 = (float)(m[u] & 7u) * r1 - sum;
        }
        io.pd[c[u]] = pd_k;
        io.x[c[u]] = xc[u] + alpha * pd_k;
        io.s_out[c[u]] = s_k;
        io.r_out[c[u]] = r1;
        io.w_out[c[u]] = w1;
        gam += (double)r1 * (double)r1;
        del += (double)w1 * (double)r1;
    }
}

template <int U, typename TileId>
__device__ __forceinline__ double rr_batch(const StepParams& P, const float* __restrict__ r, const TileWalk& tw,
                                           uint32_t ti, TileId&& tile_id) {
    const uint32_t lane = lane_id();
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        const uint32_t c = t < tw.end ? tile_id(t) * kTile + lane : P.C;
        if (c < P.C) {
            const double rv = (double)ld_at(r, c);
            acc += rv * rv;
        }
    }
    return acc;
}

template <int U, typename TileId>
__device__ __forceinline__ double true_res_batch(const StepParams& P, const uint16_t* __restrict__ nmask,
                                                 const float* __restrict__ b, const float* __restrict__ p,
                                                 const TileWalk& tw, uint32_t ti, TileId&& tile_id) {
    const uint32_t lane = lane_id(), C = P.C;
    const int32_t off[6] = {-1, 1, -(int32_t)P.nx, (int32_t)P.nx, -(int32_t)(P.nx * P.ny), (int32_t)(P.nx * P.ny)};
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t t = ti + (uint32_t)u * tw.stride;
        const uint32_t c = t < tw.end ? tile_id(t) * kTile + lane : C;
        if (c >= C) continue;
        const uint16_t m = ld_at(nmask, c);
        if (!(m & kMaskLiquid)) continue;
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < 6; ++s)
            if ((m >> (3 + s)) & 1u) sum += (double)p[(uint32_t)((int32_t)c + off[s])];
        const double e = (double)b[c] - ((double)(m & 7u) * (double)p[c] - sum);
        acc += e * e;
    }
    return acc;
}

}  // namespace fs
