/*
 * lbm_oracle.c -- CPU oracle for the D3Q19 AA-pattern hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library (as what the tests compare
 * against, or as the timed CPU baseline).  The product library never links or calls it.
 *
 * It restates, as plain scalar C:
 *   init_kernel       src/cl_programs/lbm_init.cl:32-237
 *   lbm_kernel_alpha  src/cl_programs/lbm_alpha.cl:12-498
 *   lbm_kernel_beta   src/cl_programs/lbm_beta.cl:13-802 (LDS summation order)
 *   copy_buffer_rect  src/cl_programs/copy_buffer_rect.cl:12-52
 *   region get/set    src/CLbmSolver.hpp:695-757
 * plus the new Smagorinsky LES rate, the body force and the seeded initial
 * perturbation.  Each function is pinned by tests/test_oracle.py against
 * closed forms, conservation laws and a brute-force textbook stream/collide.
 *
 * Build: gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off (oracle/build_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* flag values (src/common.h:19-22, src/CLbmSolver.hpp:30-33) */
#define FLAG_OBSTACLE 1
#define FLAG_FLUID 2
#define FLAG_INJECTION 4
#define FLAG_GHOST 8

/* D3Q19 directions in the reference's order (src/main.cpp:38-66,
 * src/cl_programs/lbm_header.h:12-19); opp(i) = i^1 for i < 18. */
static const int ORACLE_E[19][3] = {
    {1, 0, 0},  {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},   {0, -1, -1},
    {0, 1, -1}, {0, -1, 1},  {0, 0, 1},  {0, 0, -1},  {0, 0, 0}};

/* Counter-based generator for the seeded initial perturbation: splitmix64
 * finaliser (Steele, Lea & Flood 2014) over (seed, case, key, component). */
static inline uint64_t oracle_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double oracle_unit(uint64_t seed, uint64_t case_id, uint64_t key, int comp) {
    uint64_t h = oracle_mix64(seed ^ oracle_mix64(case_id ^ oracle_mix64(key * 4ull + (uint64_t)comp)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

int oracle_direction(int i, int c) { return ORACLE_E[i][c]; }

#define REAL float
#define SFX _f
#define ORACLE_IS_DOUBLE 0
#define SQRT sqrtf
#include "lbm_oracle_impl.h"
#undef REAL
#undef SFX
#undef ORACLE_IS_DOUBLE
#undef SQRT

#define REAL double
#define SFX _d
#define ORACLE_IS_DOUBLE 1
#define SQRT sqrt
#include "lbm_oracle_impl.h"
#undef REAL
#undef SFX
#undef ORACLE_IS_DOUBLE
#undef SQRT

int oracle_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
