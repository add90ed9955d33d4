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

// Element i of an array as a 32-bit byte offset from a uniform base, so the
// access is one VGPR offset on an SGPR base instead of a 64-bit address per
// lane (arrays are below 4 GB: grids are limited to 2^30 faces, tanks to
// 2^30 particles).
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_at(T* base, uint32_t i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T)) = v;
}

// WGSL i32(f32): truncating, saturating, NaN -> 0.  v_cvt_i32_f32 does
// exactly that (out-of-range values clamp to INT_MIN/INT_MAX, NaN gives 0),
// so the conversion is that one instruction; the C cast is undefined out of
// range and written with its checks compiled to three nested branches.
__device__ __forceinline__ int32_t sat_i32(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Cell key of sh/mark.wgsl:35-43: g = (pos - origin) / h (division), clamped
// floor.  The same expression the sort keys use (sh/sort_scatter.wgsl:43-51).
// The clamped cell coordinates of a position and their row-major key.
__device__ __forceinline__ uint32_t cell_ijk(const StepParams& P, float x, float y, float z, uint32_t& i,
                                             uint32_t& j, uint32_t& k) {
    float gx = (x - P.origin[0]) / P.h;
    float gy = (y - P.origin[1]) / P.h;
    float gz = (z - P.origin[2]) / P.h;
    i = (uint32_t)clampf(floorf(gx), 0.0f, (float)(P.nx - 1u));
    j = (uint32_t)clampf(floorf(gy), 0.0f, (float)(P.ny - 1u));
    k = (uint32_t)clampf(floorf(gz), 0.0f, (float)(P.nz - 1u));
    return i + P.nx * (j + P.ny * k);
}

__device__ __forceinline__ uint32_t cell_key(const StepParams& P, float x, float y, float z) {
    uint32_t i, j, k;
    return cell_ijk(P, x, y, z, i, j, k);
}

// Per-axis parameters by selects: indexing a parameter array (or a local
// one) with a run-time axis makes the compiler copy it to per-lane scratch
// memory, which costs a whole pass of scratch traffic in a per-face kernel.
__device__ __forceinline__ void face_dims(const StepParams& P, int a, uint32_t d[3]) {
    d[0] = P.nx + (a == 0 ? 1u : 0u);
    d[1] = P.ny + (a == 1 ? 1u : 0u);
    d[2] = P.nz + (a == 2 ? 1u : 0u);
}
// One of three uniform values by a run-time index.  The empty asm makes the
// three opaque, so the compiler cannot fold the selects back into an indexed
// load of the array they came from.
template <typename T>
__device__ __forceinline__ T pick3(int a, T x0, T x1, T x2) {
    asm volatile("" : "+s"(x0), "+s"(x1), "+s"(x2));
    return a == 0 ? x0 : (a == 1 ? x1 : x2);
}
__device__ __forceinline__ float grav_axis(const StepParams& P, int a) { return pick3(a, P.grav[0], P.grav[1], P.grav[2]); }
__device__ __forceinline__ uint32_t faces_axis(const StepParams& P, int a) { return pick3(a, P.F[0], P.F[1], P.F[2]); }

// Per-axis staggering of the MAC gather/scatter (sh/scatter.wgsl:97-116).
__device__ __forceinline__ float stagger(int a, int c) { return a == c ? 0.0f : -0.5f; }

// Wave64 helpers.
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

// v of the partner lane given by DPP control CTRL (a permutation within rows
// of 16 lanes), both 32-bit halves.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum of v over the wave's 64 lanes, returned to every lane; every lane must
// be active.  Four DPP steps add the partner lane^1, lane^2, the mirror
// within 8 lanes and the mirror within 16: each partner adds the same two
// values in the other order, so after every step all lanes of a row of 16
// hold the identical bits, and after the fourth, the row's sum.  The row
// sums (lanes 0, 16, 32, 48) are then added as (r0 + r1) + (r2 + r3).  The
// order is fixed, so equal inputs give equal totals in every wave.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
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
