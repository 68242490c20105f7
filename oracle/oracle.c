/*
 * oracle.c -- CPU oracle for the dense free-tensor hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of RoughPy's reference algorithms
 * used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product (roughpy_amd) never links,
 * loads or calls it.
 *
 * Every function is instantiated for double (suffix _f64) and float
 * (suffix _f32) from oracle_ops.h.  Each cites the reference file:line it
 * follows.  Parity pins for each function are listed in oracle/README.md.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

#define ORC_MAX_DEPTH 32

/* db[0] = 0, db[k+1] = 1 + w*db[k]  (roughpy/compute/_src/tensor_basis.c:
 * 42-58; roughpy_compute/common/basis.hpp:8-18). */
void orc_degree_begin(int w, int D, int64_t *db)
{
    db[0] = 0;
    for (int k = 0; k <= D; ++k) db[k + 1] = 1 + (int64_t)w * db[k];
}

int orc_max_depth(void) { return ORC_MAX_DEPTH; }

#define REAL double
#define SFX(name) name##_f64
#include "oracle_ops.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "oracle_ops.h"
#undef REAL
#undef SFX
