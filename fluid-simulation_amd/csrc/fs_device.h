// fs_device.h — step constants and exact-arithmetic device helpers shared by
// every kernel of the FLIP/PIC step (gfx950, wave64).
//
// All kernels are compiled with -ffp-contract=off: each multiply and add rounds
// separately, in the operand order of the WGSL pass it implements, so the FP32
// results match the CPU restatement bit for bit.  f32 division and sqrt are
// the correctly rounded HIP defaults.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs {

// Step constants: the reference's 128-byte Params uniform
// (src/gpu/fluid.rs:25-39) unpacked, plus derived sizes.  Passed by value as
// a kernel argument.
struct StepParams {
    uint32_t nx, ny, nz, C;      // cell grid
    uint32_t F[3];               // face counts (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1)
    uint32_t N;                  // particles (dims.y)
    float h, inv_h, dt, scale;   // geom
    float rho, flip, friction, cfl;
    float origin[3];
    float grav[3];
    float rest, stiff, drift;    // spc
    uint32_t thresh;             // max(u32(cls.x), 1)
    uint32_t dilation;           // u32(cls.y) >= 1
    uint32_t warm;               // dims.w
    uint32_t iters;              // dims.z (>= 1)
    float tol_sq;                // cg_scalars[6] = clamp(tol,0,0.1)^2 in f32
    float tol;                   // clamp(tol,0,0.1) (f64 stop test of the mixed solve)
};

enum CellType : uint8_t { kSolid = 0, kLiquid = 1, kAir = 2 };

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
    return fminf(fmaxf(x, lo), hi);  // WGSL clamp; fmin/fmax drop a NaN operand
}

// WGSL i32(f32): saturating, NaN -> 0 (v_cvt_i32_f32 behaves the same).
__device__ __forceinline__ int32_t sat_i32(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

// Cell key of sh/mark.wgsl:35-43: g = (pos - origin) / h (division), clamped
// floor.  The same expression the sort keys use (sh/sort_scatter.wgsl:43-51).
__device__ __forceinline__ uint32_t cell_key(const StepParams& P, float x, float y, float z) {
    float gx = (x - P.origin[0]) / P.h;
    float gy = (y - P.origin[1]) / P.h;
    float gz = (z - P.origin[2]) / P.h;
    uint32_t i = (uint32_t)clampf(floorf(gx), 0.0f, (float)(P.nx - 1u));
    uint32_t j = (uint32_t)clampf(floorf(gy), 0.0f, (float)(P.ny - 1u));
    uint32_t k = (uint32_t)clampf(floorf(gz), 0.0f, (float)(P.nz - 1u));
    return i + P.nx * (j + P.ny * k);
}

__device__ __forceinline__ void face_dims(const StepParams& P, int a, uint32_t d[3]) {
    d[0] = P.nx; d[1] = P.ny; d[2] = P.nz;
    d[a] += 1u;
}

// Per-axis staggering of the MAC gather/scatter (sh/scatter.wgsl:97-116).
__device__ __forceinline__ float stagger(int a, int c) { return a == c ? 0.0f : -0.5f; }

// Wave64 helpers.
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// bf16 storage of particle velocity: round-to-nearest-even on store.
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t r = 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

}  // namespace fs
