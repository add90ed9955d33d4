/*
 * flip_oracle_f64.c — FP64 instantiation of the CPU restatement: the same
 * algorithm with every real in double and the P2G fixed-point sums in int64.
 * This is the higher-precision oracle the precision reports compare against.
 * TEST INFRASTRUCTURE ONLY (see flip_oracle.h).
 */
#define R double
#define SFX f64
#define ACC int64_t
#define ACC_PUB int64_t
#define NUMT int64_t
#define RMIN fmin
#define RMAX fmax
#define RFLOOR floor
#define RRINT rint
#define RFABS fabs
#define RSQRT sqrt
#define RINVH(P) (1.0 / (double)(P)->h)
#include "flip_oracle_body.h"
