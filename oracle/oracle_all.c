/* One translation unit holding both precisions of the oracle (test
 * infrastructure only; see egg_oracle.c). The float build's FP64 reductions
 * call the double metric, so both live in the same library. */
#define REAL double
#define SFX _f64
#include "egg_oracle.c"
#undef REAL
#undef SFX
#define REAL float
#define SFX _f32
#include "egg_oracle.c"

#ifdef _OPENMP
#include <omp.h>
#endif
/* Thread count of the oracle's OpenMP loops (CPU baseline timing). */
void o_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
