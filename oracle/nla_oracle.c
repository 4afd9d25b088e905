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

/* The scalar-generic QR-family restatement (real and complex), nlaq_{d,s,z,c}*. */
#include <complex.h>

#define R double
#define RR double
#define P(x) nlaq_d##x
#define CONJ(x) (x)
#define RE(x) (x)
#define IM(x) ((RR)0)
#define MKC(re, im) ((R)(re))
#define RSQRT(x) sqrt(x)
#include "nla_oracle_qr_t.h"
#undef R
#undef RR
#undef P
#undef MKC
#undef RSQRT

#define R float
#define RR float
#define P(x) nlaq_s##x
#define MKC(re, im) ((R)(re))
#define RSQRT(x) sqrtf(x)
#include "nla_oracle_qr_t.h"
#undef R
#undef RR
#undef P
#undef CONJ
#undef RE
#undef IM
#undef MKC
#undef RSQRT

#define R double _Complex
#define RR double
#define P(x) nlaq_z##x
#define CONJ(x) conj(x)
#define RE(x) creal(x)
#define IM(x) cimag(x)
#define MKC(re, im) ((R)(re) + (R)(im) * I)
#define RSQRT(x) sqrt(x)
#include "nla_oracle_qr_t.h"
#undef R
#undef RR
#undef P
#undef CONJ
#undef RE
#undef IM
#undef MKC
#undef RSQRT

#define R float _Complex
#define RR float
#define P(x) nlaq_c##x
#define CONJ(x) conjf(x)
#define RE(x) crealf(x)
#define IM(x) cimagf(x)
#define MKC(re, im) ((R)(re) + (R)(im) * I)
#define RSQRT(x) sqrtf(x)
#include "nla_oracle_qr_t.h"
#undef R
#undef RR
#undef P
#undef CONJ
#undef RE
#undef IM
#undef MKC
#undef RSQRT

int nlao_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
