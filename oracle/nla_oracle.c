/*
 * CPU oracle for the tile-QR / recursive TRSM-TRMM hot path of NextLA.jl.
 * TEST INFRASTRUCTURE ONLY: linked by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the shipped HIP library never uses it.
 * The algorithms are restated in nla_oracle_t.h, instantiated here for
 * double (nlao_d*) and float (nlao_s*).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define R double
#define P(x) nlao_d##x
#include "nla_oracle_t.h"
#undef R
#undef P

#define R float
#define P(x) nlao_s##x
#include "nla_oracle_t.h"
#undef R
#undef P

int nlao_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
