/*
 * ct2_oracle.c -- CPU restatement of ct-2's reconstruction hot path
 * (emil_code/view_proj3D.f, view_proj3D_tomo.f, ray_proj3D.f, totalvar.f,
 * tomo3DfortranAux.f).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels and the CPU
 * baseline leg of bench.py.  Nothing in the product imports or links it.
 *
 * Built twice from ora_body.h: *_f32 symbols are Fortran-faithful FP32
 * (same loop order and operation order as the reference, compiled with
 * -ffp-contract=off so no FMA contraction), *_f64 symbols are the same
 * algorithm in FP64 (the precision oracle).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))
#define ORA_LINEAR 0
#define ORA_NEAREST 1

#define REAL float
#define FN(name) name##_f32
#define SQRT(x) sqrtf(x)
#define FLOOR(x) floorf(x)
#define ABS(x) fabsf(x)
#define POW(x, y) powf(x, y)
#define COS(x) cosf(x)
#define SIN(x) sinf(x)
#include "ora_body.h"
#undef REAL
#undef FN
#undef SQRT
#undef FLOOR
#undef ABS
#undef POW
#undef COS
#undef SIN
#undef R

#define REAL double
#define FN(name) name##_f64
#define SQRT(x) sqrt(x)
#define FLOOR(x) floor(x)
#define ABS(x) fabs(x)
#define POW(x, y) pow(x, y)
#define COS(x) cos(x)
#define SIN(x) sin(x)
#include "ora_body.h"
