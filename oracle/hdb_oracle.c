/*
 * CPU oracle for the batched HDBSCAN hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product path (hdbscan_amd) never links, loads or calls it.
 *
 * The algorithm restates the reference's per-window pipeline
 * (reference hdbscan_gpu/_kernels.py:45-760, hdbscan_gpu/_cuda.py:114-118)
 * serially in fp32 and fp64.  Which checks pin each function is listed in
 * tests/test_oracle.py; every exported function is pinned there.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* NumPy's scalar quicksort switches to insertion sort at this span (hi-lo). */
#define ORC_SMALL_QUICKSORT 15

#define REAL float
#define SFX f32
#define SQRTF sqrtf
#define FMAXF fmaxf
#define REAL_MAXV 3.4028234663852886e+38F
#include "hdb_oracle_body.h"
#undef REAL
#undef SFX
#undef SQRTF
#undef FMAXF
#undef REAL_MAXV

#define REAL double
#define SFX f64
#define SQRTF sqrt
#define FMAXF fmax
#define REAL_MAXV 1.7976931348623157e+308
#include "hdb_oracle_body.h"
#undef REAL
#undef SFX
#undef SQRTF
#undef FMAXF
#undef REAL_MAXV

int orc_version(void) { return 1; }
