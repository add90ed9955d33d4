/*
 * flip_oracle_f32.c — plain-FP32 instantiation of the CPU restatement, plus
 * the precision-independent integer parts (scan, stable permutation).
 * TEST INFRASTRUCTURE ONLY (see flip_oracle.h).
 */
#define R float
#define SFX f32
#define ACC uint32_t       /* i32 atomics of the reference wrap: add as u32 */
#define ACC_PUB int32_t
#define NUMT int32_t
#define RMIN fminf
#define RMAX fmaxf
#define RFLOOR floorf
#define RRINT rintf        /* WGSL round() is round-half-to-even */
#define RFABS fabsf
#define RSQRT sqrtf
#define RINVH(P) ((P)->inv_h)
#include "flip_oracle_body.h"

/* Exclusive prefix sum, the net effect of sh/sort_scan_block.wgsl ->
 * sh/sort_scan_spine.wgsl -> sh/sort_scan_add.wgsl; host mirror
 * src/lib.rs:1175-1183 (wrapping u32 adds, like the shaders). */
void orc_exclusive_scan_u32(uint32_t len, const uint32_t *in, uint32_t *out) {
    uint32_t acc = 0;
    for (uint32_t c = 0; c < len; ++c) {
        uint32_t v = in[c];
        out[c] = acc;
        acc += v;
    }
}

/* Stable counting sort, src/lib.rs:1188-1211: histogram -> exclusive scan ->
 * cursor scatter in input order, so perm[d] = source of sorted slot d and
 * particles sharing a key keep their input order. */
void orc_stable_permutation(uint32_t n, uint32_t cells, const uint32_t *keys, uint32_t *perm) {
    uint32_t *cur = (uint32_t *)calloc(cells ? cells : 1, sizeof(uint32_t));
    for (uint32_t p = 0; p < n; ++p) cur[keys[p]] += 1u;
    uint32_t acc = 0;
    for (uint32_t c = 0; c < cells; ++c) {
        uint32_t v = cur[c];
        cur[c] = acc;
        acc += v;
    }
    for (uint32_t p = 0; p < n; ++p) perm[cur[keys[p]]++] = p;
    free(cur);
}

uint32_t orc_classify_u32(const orc_params *P, const uint32_t *occ, uint32_t *cell_type) {
    return classify(P, occ, cell_type);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
