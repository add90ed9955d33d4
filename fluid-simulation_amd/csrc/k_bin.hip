// k_bin.hip — particle binning: clears, cell keys + occupancy histogram,
// exclusive scans and the stable counting-sort permutation.
//
// Reference: sh/mark.wgsl:27-45 (histogram), sh/sort_scan_block|spine|add.wgsl
// (exclusive scan of the histogram = bin starts), sh/sort_scatter.wgsl:35-54
// (permutation).  The reference's permutation claims slots with atomicAdd and
// so orders particles inside a bin arbitrarily; this file produces the stable
// order of the host mirror src/lib.rs:1188-1211 with an LSD radix sort on the
// cell key (each 8-bit digit pass is stable), so the permutation is bit-exact
// and deterministic.
#include "fs_kernels.h"
#include "fs_particle.h"

namespace fs {

namespace {
constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;     // 2048 elements per scan block
constexpr int kRadixWaves = kBlock / 64;
constexpr int kRadixBins = 256;
constexpr int kRadixMinTile = kBlock * 8;          // smallest menu entry (workspace sizing)
// Keys per lane of the radix workgroup, by sort tile variant (0: 4096 keys,
// 1: 2048, 2: 8192 per workgroup).
constexpr int kRadixItemsMenu[3] = {16, 8, 32};

inline unsigned grid_for(size_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }
}  // namespace

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f32(float* __restrict__ p, float v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Fill that runs only when *guard > above (used for the P2G fallback path).
__global__ void k_fill_u32_guarded(uint32_t* __restrict__ p, uint32_t v, size_t n, const uint32_t* guard,
                                   uint32_t above) {
    if (*guard <= above) return;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32_guarded(uint32_t* p, uint32_t value, size_t count, const uint32_t* guard,
                                   uint32_t above, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned g = grid_for(count, kBlock);
    if (g > 8192) g = 8192;
    FS_LAUNCH(k_fill_u32_guarded, dim3(g), dim3(kBlock), 0, s, p, value, count, guard, above);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned g = grid_for(count, kBlock);
    if (g > 8192) g = 8192;
    FS_LAUNCH(k_fill_u32, dim3(g), dim3(kBlock), 0, s, p, value, count);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float value, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned g = grid_for(count, kBlock);
    if (g > 8192) g = 8192;
    FS_LAUNCH(k_fill_f32, dim3(g), dim3(kBlock), 0, s, p, value, count);
    return hipGetLastError();
}

// sh/mark.wgsl:27-45: one atomicAdd per particle.
__global__ void k_mark(StepParams P, const float* __restrict__ px, const float* __restrict__ py,
                       const float* __restrict__ pz, uint32_t* __restrict__ occ) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.N) return;
    atomicAdd(&occ[cell_key(P, px[p], py[p], pz[p])], 1u);
}

hipError_t launch_mark(const StepParams& P, const float* const pos[3], uint32_t* occ, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    FS_LAUNCH(k_mark, dim3(grid_for(P.N, 64)), dim3(64), 0, s, P, pos[0], pos[1], pos[2], occ);
    return hipGetLastError();
}

// Keys + histogram.  Lanes holding equal consecutive keys (the common case once
// particles are sorted) are merged into one atomic per run; the counts are
// integers, so the result equals one atomic per particle exactly.
__global__ void k_key_hist(StepParams P, const float* __restrict__ px, const float* __restrict__ py,
                           const float* __restrict__ pz, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ occ, uint32_t* __restrict__ mis_count,
                           uint32_t* __restrict__ mis_list, uint32_t mis_cap) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P.N;
    uint32_t key = 0xffffffffu;
    bool mis = false;
    if (valid) {
        // The P2G cell (sh/scatter.wgsl:94: multiply by inv_h, no clamp) must be
        // the sort cell for the brick gather to see every tap from the segments;
        // particles where the two differ are counted and listed.
        key = particle_key(P, px[p], py[p], pz[p], mis);
        keys[p] = key;
    }
    list_misfits(mis, p, mis_count, mis_list, mis_cap, nullptr);
    hist_key_runs(key, valid, occ);
}

hipError_t launch_key_hist(const StepParams& P, const float* const pos[3], uint32_t* keys, uint32_t* occ,
                           uint32_t* mis_count, uint32_t* mis_list, uint32_t mis_cap, hipStream_t s) {
    if (P.N == 0) return hipSuccess;
    FS_LAUNCH(k_key_hist, dim3(grid_for(P.N, kBlock)), dim3(kBlock), 0, s, P, pos[0], pos[1],
                       pos[2], keys, occ, mis_count, mis_list, mis_cap);
    return hipGetLastError();
}

// Cell order without stability: particle p takes slot off[key] + (a count of
// the particles of its cell placed before it), one atomic per run of equal
// keys among a wave's lanes.  The slots of cell k are exactly
// [off[k], off[k] + occ[k]) when occ counted the same keys, so `order` is a
// permutation grouping the particles by cell in key order; the order inside
// a cell is whatever the atomics give.
// A thread takes kOrderItems particles kBlock apart (a wave's lanes still hold
// consecutive particles), so a wave has that many independent load, atomic
// and store chains in flight.
namespace {
constexpr int kOrderItems = 4;
}

__global__ __launch_bounds__(kBlock) void k_cell_order(const uint32_t* __restrict__ keys, uint32_t n,
                                                       const uint32_t* __restrict__ off,
                                                       uint32_t* __restrict__ cursor, uint32_t* __restrict__ order) {
    const uint32_t p0 = blockIdx.x * (kBlock * kOrderItems) + threadIdx.x;
    const uint32_t lane = lane_id();
    uint32_t key[kOrderItems], base[kOrderItems], hl[kOrderItems], ok[kOrderItems];
#pragma unroll
    for (int i = 0; i < kOrderItems; ++i) {
        const uint32_t p = p0 + (uint32_t)i * kBlock;
        key[i] = p < n ? keys[p] : 0xffffffffu;
    }
#pragma unroll
    for (int i = 0; i < kOrderItems; ++i) {
        const uint32_t p = p0 + (uint32_t)i * kBlock;
        const bool valid = p < n;
        const uint32_t prev = __shfl_up(key[i], 1, 64);
        const bool head = valid && (lane == 0 || prev != key[i]);
        const uint64_t heads = __ballot(head);
        const uint64_t act = __ballot(valid);
        ok[i] = valid ? off[key[i]] : 0u;
        base[i] = 0;
        if (head) {
            const uint64_t later = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
            const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : (uint32_t)(64 - __builtin_clzll(act));
            base[i] = atomicAdd(&cursor[key[i]], end - lane);
        }
        // The head of this lane's run: the highest head lane at or below it.
        const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        hl[i] = upto ? 63u - (uint32_t)__builtin_clzll(upto) : 0u;
    }
#pragma unroll
    for (int i = 0; i < kOrderItems; ++i) {
        const uint32_t p = p0 + (uint32_t)i * kBlock;
        const uint32_t b = (uint32_t)__shfl((int)base[i], (int)hl[i], 64);
        if (p < n) order[ok[i] + b + (lane - hl[i])] = p;
    }
}

hipError_t launch_cell_order(const uint32_t* keys, uint32_t n, const uint32_t* cell_offset, uint32_t* cursor,
                             uint32_t ncells, uint32_t* order, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipError_t e = launch_fill_u32(cursor, 0u, ncells, s);
    if (e != hipSuccess) return e;
    FS_LAUNCH(k_cell_order, dim3(grid_for(n, kBlock * kOrderItems)), dim3(kBlock), 0, s, keys, n, cell_offset,
              cursor, order);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exclusive scan (u32, wrapping adds): tiles of 2048, tile sums scanned
// recursively, then added back.  Integer prefix sums have one answer, so the
// result equals the three WGSL scan passes exactly.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_wave[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t i = 0; i < (blockDim.x >> 6); ++i) {
        uint32_t x = s_wave[i];
        if (i < w) base += x;
        tot += x;
    }
    total = tot;
    return base + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ tile_sums, size_t len) {
    __shared__ uint32_t s_data[kScanTile];
    __shared__ uint32_t s_wave[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kBlock) {
        size_t g = base + i;
        s_data[i] = g < len ? in[g] : 0u;
    }
    __syncthreads();
    uint32_t loc[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        loc[i] = sum;
        sum += s_data[threadIdx.x * kScanItems + i];
    }
    uint32_t total;
    uint32_t pre = block_excl_scan(sum, s_wave, total);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s_data[threadIdx.x * kScanItems + i] = pre + loc[i];
    __syncthreads();
    for (int i = threadIdx.x; i < kScanTile; i += kBlock) {
        size_t g = base + i;
        if (g < len) out[g] = s_data[i];
    }
    if (tile_sums && threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_off, size_t len) {
    size_t g = (size_t)blockIdx.x * kScanTile + threadIdx.x;
    const uint32_t add = tile_off[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i, g += kBlock)
        if (g < len) out[g] += add;
}

size_t scan_workspace(size_t len) {
    size_t total = 0;
    while (len > (size_t)kScanTile) {
        len = (len + kScanTile - 1) / kScanTile;
        total += 2 * len;
    }
    return total + 16;
}

hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t len, uint32_t* ws,
                                 hipStream_t s) {
    if (len == 0) return hipSuccess;
    size_t tiles = (len + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        FS_LAUNCH(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, in, out, (uint32_t*)nullptr, len);
        return hipGetLastError();
    }
    uint32_t* sums = ws;
    uint32_t* sums_scanned = ws + tiles;
    FS_LAUNCH(k_scan_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, out, sums, len);
    hipError_t e = launch_exclusive_scan(sums, sums_scanned, tiles, ws + 2 * tiles, s);
    if (e != hipSuccess) return e;
    FS_LAUNCH(k_scan_add, dim3((unsigned)tiles), dim3(kBlock), 0, s, out, sums_scanned, len);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort over 8-bit digits.
//
// Per-tile digit counts.  A wave's 64 lanes read consecutive keys, and the
// keys come nearly sorted (the particles were sorted on an earlier substep and
// moved little), so neighbouring lanes mostly share a digit: one LDS atomic per
// run of equal digits (hist_key_runs) instead of one per key, which would
// serialize on the shared bin.  The counts are integers, so the table is the
// same either way.
template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                        uint32_t shift, uint32_t ntiles,
                                                        uint32_t* __restrict__ hist) {
    // XCD-contiguous tiles, as in k_radix_scatter: neighbouring tiles' table
    // entries share lines.
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    __shared__ uint32_t s_h[kRadixBins];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)tile * (ITEMS * kBlock);
#pragma unroll 4
    for (int i = 0; i < ITEMS; ++i) {
        const size_t idx = base + threadIdx.x + (size_t)i * kBlock;
        const bool valid = idx < n;
        hist_key_runs(valid ? (keys[idx] >> shift) & 255u : 0u, valid, s_h);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + tile] = s_h[threadIdx.x];
}

// Each wave ranks 64*ITEMS consecutive keys in ITEMS rounds of 64; equal
// digits inside a round are found with eight ballots.  Order of items =
// (wave, round, lane) = input order, so equal digits keep their input order.
// Output position of an item = start of its digit's run for this tile (from
// the scanned histogram) + the counts of its digit in earlier waves + its
// rank.  Tiles of up to 4096 keys are staged in LDS in that order, digit by
// digit, and written out as contiguous runs per digit (coalesced stores);
// larger tiles write each item directly.  The positions are the same either
// way.
template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ offsets, uint32_t n, uint32_t shift, uint32_t ntiles,
    int write_keys) {
    constexpr bool kStage = ITEMS <= 16;
    constexpr int kTileKeys = ITEMS * kBlock;
    static_assert(kBlock == kRadixBins, "one thread per digit");
    __shared__ uint32_t s_cnt[kRadixWaves][kRadixBins];
    __shared__ uint32_t s_lbase[kRadixBins];  // tile-local start of each digit's run
    __shared__ uint32_t s_gbase[kRadixBins];  // global start of each digit's run of this tile
    __shared__ uint32_t s_wtot[kRadixWaves];
    __shared__ uint32_t s_k[kStage ? kTileKeys : 1], s_v[kStage ? kTileKeys : 1];
    // Each XCD takes a contiguous share of the tiles, so the runs that
    // neighbouring tiles write side by side are completed in one L2.
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    for (int i = threadIdx.x; i < kRadixWaves * kRadixBins; i += kBlock) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    const size_t tile_base = (size_t)tile * kTileKeys;
    const size_t base = tile_base + (size_t)w * (ITEMS * 64);
    uint32_t kk[ITEMS], vv[ITEMS], rk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        size_t idx = base + (size_t)r * 64 + lane;
        bool valid = idx < n;
        uint32_t key = valid ? keys_in[idx] : 0u;
        uint32_t val = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
        uint32_t digit = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (digit >> b) & 1u;
            uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        uint32_t old = s_cnt[w][digit];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) s_cnt[w][digit] = old + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        kk[r] = key;
        vv[r] = val;
        rk[r] = valid ? old + rank : 0xffffffffu;
    }
    __syncthreads();
    {
        // Thread t owns digit t: prefix of its counts over the waves, the
        // global start of its run, and (block scan) its tile-local start.
        const uint32_t t = threadIdx.x;
        uint32_t tot = 0;
        for (int ww = 0; ww < kRadixWaves; ++ww) {
            const uint32_t c = s_cnt[ww][t];
            s_cnt[ww][t] = tot;
            tot += c;
        }
        s_gbase[t] = offsets[(size_t)t * ntiles + tile];
        uint32_t inc = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t up = __shfl_up(inc, o, 64);
            if ((int)lane >= o) inc += up;
        }
        if (lane == 63) s_wtot[w] = inc;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t ww = 0; ww < w; ++ww) before += s_wtot[ww];
        s_lbase[t] = before + inc - tot;
    }
    __syncthreads();
    if constexpr (kStage) {
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            if (rk[r] == 0xffffffffu) continue;
            const uint32_t d = (kk[r] >> shift) & 255u;
            const uint32_t lp = s_lbase[d] + s_cnt[w][d] + rk[r];
            s_k[lp] = kk[r];
            s_v[lp] = vv[r];
        }
        __syncthreads();
        const uint32_t count = (uint32_t)min((size_t)kTileKeys, (size_t)n - tile_base);
        for (uint32_t j = threadIdx.x; j < count; j += kBlock) {
            const uint32_t k = s_k[j];
            const uint32_t d = (k >> shift) & 255u;
            const uint32_t pos = s_gbase[d] + (j - s_lbase[d]);
            if (write_keys) keys_out[pos] = k;
            vals_out[pos] = s_v[j];
        }
    } else {
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            if (rk[r] == 0xffffffffu) continue;
            const uint32_t d = (kk[r] >> shift) & 255u;
            const uint32_t pos = s_gbase[d] + s_cnt[w][d] + rk[r];
            if (write_keys) keys_out[pos] = kk[r];
            vals_out[pos] = vv[r];
        }
    }
}

size_t radix_workspace(size_t n) {
    size_t ntiles = (n + kRadixMinTile - 1) / kRadixMinTile;
    size_t table = ntiles * kRadixBins;
    return 2 * table + scan_workspace(table) + 16;
}

hipError_t launch_stable_permutation(const uint32_t* keys, uint32_t n, uint32_t key_limit,
                                     uint32_t* keys_tmp0, uint32_t* keys_tmp1, uint32_t* vals_tmp,
                                     uint32_t* perm, uint32_t* ws, hipStream_t s, int tile_variant) {
    if (n == 0) return hipSuccess;
    if (tile_variant < 0 || tile_variant > 2) return hipErrorInvalidValue;
    const int items = kRadixItemsMenu[tile_variant];
    const uint32_t tile = (uint32_t)(items * kBlock);
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)key_limit) ++bits;
    const uint32_t passes = (bits + 7) / 8;
    const uint32_t ntiles = (n + tile - 1) / tile;
    const uint32_t grid8 = (ntiles + 7u) / 8u * 8u;  // whole XCD groups for xcd_remap
    const size_t table = (size_t)ntiles * kRadixBins;
    uint32_t* hist = ws;
    uint32_t* offs = ws + table;
    uint32_t* scan_ws = ws + 2 * table;
    // Value buffers: the last pass must land in `perm`.
    const uint32_t* kin = keys;
    const uint32_t* vin = nullptr;
    uint32_t* kbuf[2] = {keys_tmp0, keys_tmp1};
    uint32_t* vbuf[2];
    if (passes % 2 == 1) { vbuf[0] = perm; vbuf[1] = vals_tmp; }
    else { vbuf[0] = vals_tmp; vbuf[1] = perm; }
    for (uint32_t ps = 0; ps < passes; ++ps) {
        const uint32_t shift = 8 * ps;
        if (items == 16)
            FS_LAUNCH(k_radix_hist<16>, dim3(grid8), dim3(kBlock), 0, s, kin, n, shift, ntiles, hist);
        else if (items == 8)
            FS_LAUNCH(k_radix_hist<8>, dim3(grid8), dim3(kBlock), 0, s, kin, n, shift, ntiles, hist);
        else
            FS_LAUNCH(k_radix_hist<32>, dim3(grid8), dim3(kBlock), 0, s, kin, n, shift, ntiles, hist);
        hipError_t e = launch_exclusive_scan(hist, offs, table, scan_ws, s);
        if (e != hipSuccess) return e;
        uint32_t* kout = kbuf[ps & 1];
        uint32_t* vout = vbuf[ps & 1];
        const int write_keys = ps + 1 < passes;
        if (items == 16)
            FS_LAUNCH(k_radix_scatter<16>, dim3(grid8), dim3(kBlock), 0, s, kin, vin, kout, vout, offs,
                               n, shift, ntiles, write_keys);
        else if (items == 8)
            FS_LAUNCH(k_radix_scatter<8>, dim3(grid8), dim3(kBlock), 0, s, kin, vin, kout, vout, offs,
                               n, shift, ntiles, write_keys);
        else
            FS_LAUNCH(k_radix_scatter<32>, dim3(grid8), dim3(kBlock), 0, s, kin, vin, kout, vout, offs,
                               n, shift, ntiles, write_keys);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        kin = kout;
        vin = vout;
    }
    return hipSuccess;
}

template <int BF16>
__global__ void k_permute(uint32_t n, const uint32_t* __restrict__ perm, ParticleSoA src, ParticleSoA dst) {
    uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const uint32_t sidx = perm[d];
#pragma unroll
    for (int a = 0; a < 3; ++a) dst.pos[a][d] = src.pos[a][sidx];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (BF16) ((uint16_t*)dst.vel[a])[d] = ((const uint16_t*)src.vel[a])[sidx];
        else ((float*)dst.vel[a])[d] = ((const float*)src.vel[a])[sidx];
    }
}

hipError_t launch_permute(uint32_t n, const uint32_t* perm, const ParticleSoA& src, const ParticleSoA& dst,
                          int bf16, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (bf16) FS_LAUNCH(k_permute<1>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, perm, src, dst);
    else FS_LAUNCH(k_permute<0>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, perm, src, dst);
    return hipGetLastError();
}

}  // namespace fs
