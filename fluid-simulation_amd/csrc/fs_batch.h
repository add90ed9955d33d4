// fs_batch.h — device tables and launchers of the batched-tank step.
//
// A batch keeps every tank's arrays concatenated: particles at part_off,
// cells at cell_off (a multiple of 64, so a 64-cell CG tile never spans two
// tanks), faces of axis a at face_off[a].  Each kernel phase is one launch
// whose workgroup w reads a WorkItem (tank, first local index) from the
// phase's work table, so a workgroup never spans two tanks and the tank's
// constants are uniform.  Work tables are ordered so that consecutive items
// of a tank land on the same XCD.
#pragma once
#include "fs_kernels.h"

namespace fs {

struct BatchTank {
    StepParams P;          // the tank's step constants (P.N = its particle count)
    uint32_t part_off;     // first particle
    uint32_t cell_off;     // first cell (multiple of 64)
    uint32_t tile_off;     // cell_off / 64
    uint32_t ntiles;       // ceil(P.C / 64)
    uint32_t face_off[3];  // first face of each axis
    uint32_t cg_base;      // first partial slot of the tank's CG group
    uint32_t cg_nb;        // workgroups in the tank's CG group
    uint32_t pad;
};

struct WorkItem {
    uint32_t tank;   // kNoTank: padding, the workgroup exits
    uint32_t start;  // first local particle / cell / face / brick index; CG: rank in the group
};
constexpr uint32_t kNoTank = 0xffffffffu;
constexpr int kBatchBlock = 256;  // particles, cells or faces per workgroup

struct BatchCgState {
    double rs_init, rs_final, true_res_sq;
    uint32_t iterations;
    uint32_t error;   // nonzero: the group's barrier timed out in the last solve
    uint32_t failed;  // sticky: a solve of this tank timed out since the host
                      // last cleared it (kernels only ever set it)
    uint32_t pad;
};

struct BatchCgWork {
    uint16_t* nmask;       // [C_tot] stencil masks
    uint32_t* tile_flags;  // [tiles_tot + 1], last entry stays 0
    uint32_t* tile_scan;   // [tiles_tot + 1]
    uint32_t* tiles;       // [tiles_tot] active tile list (global tile ids)
    uint32_t* n_tiles;     // [1]
    double* partials;      // [2 * grid] one slot per workgroup, two halves
    uint32_t* flags;       // [2 * grid] the slot's tank still iterates
    uint32_t* arrive;      // [9 * 32]: eight group counters and the grid counter, one 128-byte line each
    uint32_t* gen;         // [32]: grid barrier generation
    uint32_t* abort_flag;  // [1]
    uint64_t* slots;       // [2 x (2 * grid * 2)] epoch-tagged partials: the exchange after phase A, then
                           // the fence-free barrier after phase B; cleared per solve
    uint32_t slot_words;   // u32 words of `slots`
    BatchCgState* state;   // [T]
};

struct BatchGrid {
    float* vel[3];
    float* saved[3];
    int32_t* num[3];
    int32_t* den[3];
};

// Counters per tank: [4 t + 0] liquid cells, [4 t + 1] misfit particles.
hipError_t launch_batch_key_hist(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const ParticleSoA& sp,
                                 uint32_t* keys, uint32_t* occ, uint32_t* counters, uint32_t* mis_count,
                                 uint32_t* mis_list, hipStream_t s);
hipError_t launch_batch_classify(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const uint32_t* occ,
                                 uint8_t* ct, uint32_t* counters, hipStream_t s);
hipError_t launch_batch_gather(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int brick_variant, int bf16,
                               const ParticleSoA& sorted, const ParticleSoA& src, const uint32_t* occ,
                               const uint32_t* cell_offset, const uint8_t* ct, const BatchGrid& g,
                               const uint32_t* mis_count, const uint32_t* mis_list, hipStream_t s,
                               const uint32_t* perm = nullptr);  // with perm: slot d of `sorted` is perm[d]
// Atomic P2G + face epilogue, run only when *mis_count > kMisfitCap (num/den
// cleared beforehand by guarded fills).
hipError_t launch_batch_p2g_fallback(const BatchTank* T, const WorkItem* wl_part, uint32_t npart,
                                     const WorkItem* wl_face, uint32_t nface, int bf16, const ParticleSoA& sorted,
                                     const uint8_t* ct, const BatchGrid& g, const uint32_t* mis_count,
                                     hipStream_t s);
hipError_t launch_batch_div_cginit(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const BatchGrid& g,
                                   const uint8_t* ct, const uint32_t* occ, float* div, float* b, float* p, float* r,
                                   float* d, const BatchCgWork& w, hipStream_t s);
// One cooperative launch; each tank's group of workgroups runs that tank's
// solve.  cg_unroll: 0 = 4, 1 = 2, 2 = 1 tiles in flight per wave.
hipError_t launch_batch_cg(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int cg_unroll, const float* b,
                           float* p, float* r, float* d0, float* d1, float* q, const BatchCgWork& w, hipStream_t s);
int batch_cg_max_blocks();
hipError_t launch_batch_gradient_enforce(const BatchTank* T, const WorkItem* wl, uint32_t nwl, const float* p,
                                         const uint8_t* ct, const BatchGrid& g, hipStream_t s);
// With from and perm, particle d of sp is read from slot perm[d] of *from
// (the sort's permutation applied as G2P writes).
hipError_t launch_batch_g2p(const BatchTank* T, const WorkItem* wl, uint32_t nwl, int bf16, const ParticleSoA& sp,
                            const BatchGrid& g, hipStream_t s, const ParticleSoA* from = nullptr,
                            const uint32_t* perm = nullptr);

}  // namespace fs
