/*
 * oracle.c — builds the CPU oracle (TEST INFRASTRUCTURE) in two precisions:
 * the *_f64 functions are the FP64 oracle, the *_f32 functions restate the
 * reference's Float32 pipeline with its Float64 islands. Only tests/, the
 * smoke check and the CPU-baseline leg of bench.py load this library.
 */
#include "orc_common.h"

#define REAL double
#define SFX _f64
#include "orc_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f32
#include "orc_impl.h"
#undef REAL
#undef SFX

/* Offsets of the Wieneke covariance sums in the order the stats use. */
int orc_uq_offset_table(int32_t *out /* 2 x 40 */, int32_t *ring_start /* 5 */)
{
    int off[ORC_UQ_NOFF][2], rings[ORC_UQ_MAXOFF + 1];
    int n = orc_uq_offsets(off, rings);
    for (int k = 0; k < n; k++) {
        out[2 * k] = off[k][0];
        out[2 * k + 1] = off[k][1];
    }
    for (int r = 0; r <= ORC_UQ_MAXOFF; r++) ring_start[r] = rings[r];
    return n;
}
