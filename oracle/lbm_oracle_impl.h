/*
 * lbm_oracle_impl.h -- body of the CPU oracle, included twice by lbm_oracle.c
 * (REAL=float with SFX _f, REAL=double with SFX _d).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar CPU restatement of the
 * reference's D3Q19 AA-pattern kernels.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the reference the
 * tests compare against or the timed CPU baseline.  The product path never calls it.
 *
 * Every function names the reference file:line it follows.  Arithmetic is
 * written as separate IEEE operations in the reference's evaluation order and
 * the file is compiled with -ffp-contract=off, so for REAL=float the results
 * are the reference's un-contracted FP32 results.
 *
 * Precision note: the reference's double instantiation keeps the float-rounded
 * weights (T)(1.0f/18.0f) (src/cl_programs/lbm_header.h:68-94).  The FP64
 * oracle uses exact double constants instead, because it is the precision
 * yardstick.  For REAL=float both forms give the same constants.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* lattice constants (src/cl_programs/lbm_header.h:68-94) */
#define W1 (ORACLE_IS_DOUBLE ? (REAL)(1.0 / 18.0) : (REAL)(1.0f / 18.0f))
#define W2 (ORACLE_IS_DOUBLE ? (REAL)(1.0 / 36.0) : (REAL)(1.0f / 36.0f))
#define W0 (ORACLE_IS_DOUBLE ? (REAL)(1.0 / 3.0) : (REAL)(1.0f / 3.0f))
#define C3 ((REAL)3.0)
#define C45 ((REAL)4.5)
#define C15 ((REAL)1.5)

/* eq_dd_a0 / eq_dd_a1 / eq_dd4 / eq_dd5 / eq_dd18 (src/cl_programs/lbm_header.h:68-94).
 * Evaluation order of the macro body: W * ((r + 3*v) + 4.5*v2). */
static inline REAL FN(eq_p1)(REAL v, REAL v2, REAL r) {
    REAL a = C3 * v;
    REAL b = r + a;
    REAL c = C45 * v2;
    REAL d = b + c;
    return W1 * d;
}
static inline REAL FN(eq_m1)(REAL v, REAL v2, REAL r) {
    REAL a = ((REAL)-3.0) * v;
    REAL b = r + a;
    REAL c = C45 * v2;
    REAL d = b + c;
    return W1 * d;
}
static inline REAL FN(eq_p2)(REAL v, REAL v2, REAL r) {
    REAL a = C3 * v;
    REAL b = r + a;
    REAL c = C45 * v2;
    REAL d = b + c;
    return W2 * d;
}
static inline REAL FN(eq_m2)(REAL v, REAL v2, REAL r) {
    REAL a = ((REAL)-3.0) * v;
    REAL b = r + a;
    REAL c = C45 * v2;
    REAL d = b + c;
    return W2 * d;
}

/* Equilibrium for all 19 directions, given dd_param = rho - 1.5|u|^2 and u.
 * Order of the intermediate products follows lbm_alpha.cl:352-478
 * (the velocity-injection branch, which is the plain equilibrium). */
static void FN(oracle_feq)(REAL dd_param, REAL vx, REAL vy, REAL vz, REAL *out) {
    REAL v2, ab, ab2;
    v2 = vx * vx;
    out[0] = FN(eq_p1)(vx, v2, dd_param);
    out[1] = FN(eq_m1)(vx, v2, dd_param);
    v2 = vy * vy;
    out[2] = FN(eq_p1)(vy, v2, dd_param);
    out[3] = FN(eq_m1)(vy, v2, dd_param);
    ab = vx + vy; ab2 = ab * ab;
    out[4] = FN(eq_p2)(ab, ab2, dd_param);
    out[5] = FN(eq_m2)(ab, ab2, dd_param);
    ab = vx - vy; ab2 = ab * ab;
    out[6] = FN(eq_p2)(ab, ab2, dd_param);
    out[7] = FN(eq_m2)(ab, ab2, dd_param);
    ab = vx + vz; ab2 = ab * ab;
    out[8] = FN(eq_p2)(ab, ab2, dd_param);
    out[9] = FN(eq_m2)(ab, ab2, dd_param);
    ab = vx - vz; ab2 = ab * ab;
    out[10] = FN(eq_p2)(ab, ab2, dd_param);
    out[11] = FN(eq_m2)(ab, ab2, dd_param);
    ab = vy + vz; ab2 = ab * ab;
    out[12] = FN(eq_p2)(ab, ab2, dd_param);
    out[13] = FN(eq_m2)(ab, ab2, dd_param);
    ab = vy - vz; ab2 = ab * ab;
    out[14] = FN(eq_p2)(ab, ab2, dd_param);
    out[15] = FN(eq_m2)(ab, ab2, dd_param);
    v2 = vz * vz;
    out[16] = FN(eq_p1)(vz, v2, dd_param);
    out[17] = FN(eq_m1)(vz, v2, dd_param);
    out[18] = W0 * dd_param;
}

/* f_eq(rho, u) with dd_param = rho - 1.5|u|^2 (public entry for tests). */
static inline REAL FN(vel2_of)(REAL vx, REAL vy, REAL vz);
void FN(oracle_equilibrium)(REAL rho, REAL vx, REAL vy, REAL vz, REAL *out) {
    REAL dd_param = rho - ((REAL)1.5) * FN(vel2_of)(vx, vy, vz);
    FN(oracle_feq)(dd_param, vx, vy, vz, out);
}

/* |u|^2 as vel2 = vx*vx + vy*vy + vz*vz (lbm_alpha.cl:180, lbm_beta.cl:498). */
static inline REAL FN(vel2_of)(REAL vx, REAL vy, REAL vz) {
    REAL a = vx * vx;
    REAL b = vy * vy;
    REAL c = a + b;
    REAL d = vz * vz;
    return c + d;
}

/* Smagorinsky LES relaxation rate (new; no reference code exists).
 * Hou, Sterling, Chen & Doolen (1996): tau_eff = (tau0 + sqrt(tau0^2 +
 * 18*sqrt(2)*Cs^2*|Pi_neq|/rho)) / 2, with Pi_neq_ab = sum_i e_ia e_ib f_i
 * - rho/3 delta_ab - j_a j_b for the incompressible equilibrium.  les_coeff
 * is 18*sqrt(2)*Cs^2 rounded to REAL once on the host. */
static REAL FN(oracle_les_inv_tau)(const REAL *f, REAL rho, REAL jx, REAL jy, REAL jz,
                                   REAL tau0, REAL les_coeff) {
    REAL pxx = f[0] + f[1];
    pxx = pxx + f[4]; pxx = pxx + f[5]; pxx = pxx + f[6]; pxx = pxx + f[7];
    pxx = pxx + f[8]; pxx = pxx + f[9]; pxx = pxx + f[10]; pxx = pxx + f[11];
    REAL pyy = f[2] + f[3];
    pyy = pyy + f[4]; pyy = pyy + f[5]; pyy = pyy + f[6]; pyy = pyy + f[7];
    pyy = pyy + f[12]; pyy = pyy + f[13]; pyy = pyy + f[14]; pyy = pyy + f[15];
    REAL pzz = f[8] + f[9];
    pzz = pzz + f[10]; pzz = pzz + f[11]; pzz = pzz + f[12]; pzz = pzz + f[13];
    pzz = pzz + f[14]; pzz = pzz + f[15]; pzz = pzz + f[16]; pzz = pzz + f[17];
    REAL pxy = (f[4] + f[5]) - (f[6] + f[7]);
    REAL pxz = (f[8] + f[9]) - (f[10] + f[11]);
    REAL pyz = (f[12] + f[13]) - (f[14] + f[15]);
    REAL r3 = rho * W0;
    REAL qxx = (pxx - r3) - jx * jx;
    REAL qyy = (pyy - r3) - jy * jy;
    REAL qzz = (pzz - r3) - jz * jz;
    REAL qxy = pxy - jx * jy;
    REAL qxz = pxz - jx * jz;
    REAL qyz = pyz - jy * jz;
    REAL diag = (qxx * qxx + qyy * qyy) + qzz * qzz;
    REAL off = (qxy * qxy + qxz * qxz) + qyz * qyz;
    REAL q2 = diag + ((REAL)2.0) * off;
    REAL q = SQRT(q2);
    REAL t = (les_coeff * q) / rho;
    REAL s = SQRT(tau0 * tau0 + t);
    REAL tau_eff = ((REAL)0.5) * (tau0 + s);
    return ((REAL)1.0) / tau_eff;
}

/* Exported for the tests: the Smagorinsky rate of one cell's populations,
 * with rho and j summed here in index order. */
REAL FN(oracle_les_rate)(const REAL *f, REAL tau0, REAL les_coeff) {
    REAL rho = (REAL)0, jx = (REAL)0, jy = (REAL)0, jz = (REAL)0;
    for (int i = 0; i < 19; ++i) {
        rho = rho + f[i];
        jx = jx + (REAL)ORACLE_E[i][0] * f[i];
        jy = jy + (REAL)ORACLE_E[i][1] * f[i];
        jz = jz + (REAL)ORACLE_E[i][2] * f[i];
    }
    return FN(oracle_les_inv_tau)(f, rho, jx, jy, jz, tau0, les_coeff);
}

/* Body force (new, forcing mode 1): f_i += 3 w_i (e_i . F), F = rho*g.
 * Adds momentum rho*g and no mass per step. */
static void FN(oracle_body_force)(REAL *f, REAL rho, REAL gx, REAL gy, REAL gz) {
    REAL c1 = ORACLE_IS_DOUBLE ? (REAL)(1.0 / 6.0) : (REAL)(1.0f / 6.0f);
    REAL c2 = ORACLE_IS_DOUBLE ? (REAL)(1.0 / 12.0) : (REAL)(1.0f / 12.0f);
    REAL fx = rho * gx, fy = rho * gy, fz = rho * gz;
    REAL t;
    t = c1 * fx;          f[0] = f[0] + t;   f[1] = f[1] - t;
    t = c1 * fy;          f[2] = f[2] + t;   f[3] = f[3] - t;
    t = c2 * (fx + fy);   f[4] = f[4] + t;   f[5] = f[5] - t;
    t = c2 * (fx - fy);   f[6] = f[6] + t;   f[7] = f[7] - t;
    t = c2 * (fx + fz);   f[8] = f[8] + t;   f[9] = f[9] - t;
    t = c2 * (fx - fz);   f[10] = f[10] + t; f[11] = f[11] - t;
    t = c2 * (fy + fz);   f[12] = f[12] + t; f[13] = f[13] - t;
    t = c2 * (fy - fz);   f[14] = f[14] + t; f[15] = f[15] - t;
    t = c1 * fz;          f[16] = f[16] + t; f[17] = f[17] - t;
}

/* BGK relaxation f_i += inv_tau*(feq_i - f_i) (lbm_alpha.cl:185-300,
 * lbm_beta.cl:502-558). */
static inline void FN(oracle_relax)(REAL *f, const REAL *feq, REAL w) {
    for (int i = 0; i < 19; ++i) {
        REAL d = feq[i] - f[i];
        REAL e = w * d;
        f[i] = f[i] + e;
    }
}

/* The forcing chain of the alpha kernel, mirrored with its alias quirk:
 * `#define tmp rho` makes every term reuse the previous term as "rho"
 * (lbm_alpha.cl:173,183-298 for FLUID; :359-476 for VELOCITY_INJECTION).
 * Returns the final value of the aliased variable (what the kernel stores as
 * density). */
static REAL FN(oracle_alpha_ref_forcing)(REAL *f, REAL rho, REAL gx, REAL gy, REAL gz) {
    REAL tmp;
    tmp = (gx * W1) * rho;                  rho = tmp;  f[1] = f[1] - tmp; f[0] = f[0] + tmp;
    tmp = (gy * (-W1)) * rho;               rho = tmp;  f[3] = f[3] - tmp; f[2] = f[2] + tmp;
    tmp = ((gx - gy) * W2) * rho;           rho = tmp;  f[5] = f[5] - tmp; f[4] = f[4] + tmp;
    tmp = ((gx + gy) * W2) * rho;           rho = tmp;  f[7] = f[7] - tmp; f[6] = f[6] + tmp;
    tmp = ((gx + gz) * W2) * rho;           rho = tmp;  f[9] = f[9] - tmp; f[8] = f[8] + tmp;
    tmp = ((gx - gz) * W2) * rho;           rho = tmp;  f[11] = f[11] - tmp; f[10] = f[10] + tmp;
    tmp = ((gz - gy) * W2) * rho;           rho = tmp;  f[13] = f[13] - tmp; f[12] = f[12] + tmp;
    tmp = ((gz + gy) * (-W2)) * rho;        rho = tmp;  f[15] = f[15] - tmp; f[14] = f[14] + tmp;
    tmp = (gz * W1) * rho;                  rho = tmp;  f[17] = f[17] - tmp; f[16] = f[16] + tmp;
    return rho;
}

typedef struct {
    REAL inv_tau, tau0, les_coeff;
    REAL gx, gy, gz;
    REAL lid;
    int forcing;        /* 0: reference alpha chain, 1: body force in both kernels */
    int store_velocity, store_density;
} FN(oracle_params);

/* FLUID collision shared by both kernels: LES rate (if les_coeff > 0), BGK. */
static void FN(oracle_collide_fluid)(REAL *f, REAL rho, REAL vx, REAL vy, REAL vz,
                                     const FN(oracle_params) *p, REAL dd_param) {
    REAL w = p->inv_tau;
    if (p->les_coeff > (REAL)0)
        w = FN(oracle_les_inv_tau)(f, rho, vx, vy, vz, p->tau0, p->les_coeff);
    REAL feq[19];
    FN(oracle_feq)(dd_param, vx, vy, vz, feq);
    FN(oracle_relax)(f, feq, w);
}

static inline long long FN(wrap)(long long a, long long C) {
    if (a < 0) a += C;
    if (a >= C) a -= C;
    return a;
}

/* lbm_kernel_alpha (src/cl_programs/lbm_alpha.cl:12-498), all cells. */
void FN(oracle_alpha)(REAL *dd, const int32_t *flags, REAL *vel, REAL *dens,
                      int nx, int ny, int nz, const FN(oracle_params) *p) {
    const long long C = (long long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long long g = 0; g < C; ++g) {
        int flag = flags[g];
        if (flag == FLAG_GHOST) continue;                      /* :30-32 */
        REAL f[19];
        for (int i = 0; i < 19; ++i) f[i] = dd[i * C + g];    /* :61-156 */
        /* moments in the kernel's summation order (:61-156) */
        REAL rho = f[0], vx = f[0], vy, vz;
        rho = rho + f[1];  vx = vx - f[1];
        rho = rho + f[2];  vy = f[2];
        rho = rho + f[3];  vy = vy - f[3];
        rho = rho + f[4];  vx = vx + f[4];  vy = vy + f[4];
        rho = rho + f[5];  vx = vx - f[5];  vy = vy - f[5];
        rho = rho + f[6];  vx = vx + f[6];  vy = vy - f[6];
        rho = rho + f[7];  vx = vx - f[7];  vy = vy + f[7];
        rho = rho + f[8];  vx = vx + f[8];  vz = f[8];
        rho = rho + f[9];  vx = vx - f[9];  vz = vz - f[9];
        rho = rho + f[10]; vx = vx + f[10]; vz = vz - f[10];
        rho = rho + f[11]; vx = vx - f[11]; vz = vz + f[11];
        rho = rho + f[12]; vy = vy + f[12]; vz = vz + f[12];
        rho = rho + f[13]; vy = vy - f[13]; vz = vz - f[13];
        rho = rho + f[14]; vy = vy + f[14]; vz = vz - f[14];
        rho = rho + f[15]; vy = vy - f[15]; vz = vz + f[15];
        rho = rho + f[16]; vz = vz + f[16];
        rho = rho + f[17]; vz = vz - f[17];
        rho = rho + f[18];
        REAL dens_out = rho;
        int store_dd = 0;
        if (flag == FLAG_FLUID) {                              /* :179-303 */
            REAL dd_param = rho - C15 * FN(vel2_of)(vx, vy, vz);
            FN(oracle_collide_fluid)(f, rho, vx, vy, vz, p, dd_param);
            if (p->forcing == 0) {
                dens_out = FN(oracle_alpha_ref_forcing)(f, rho, p->gx, p->gy, p->gz);
            } else {
                FN(oracle_body_force)(f, rho, p->gx, p->gy, p->gz);
            }
            store_dd = 1;
        } else if (flag == FLAG_OBSTACLE) {                    /* :305-343 */
            vx = 0; vy = 0; vz = 0;
        } else if (flag == FLAG_INJECTION) {                   /* :345-480 */
            vx = p->lid; vy = 0; vz = 0;
            REAL r1 = (REAL)1.0;
            REAL dd_param = r1 - C15 * FN(vel2_of)(vx, vy, vz);
            FN(oracle_feq)(dd_param, vx, vy, vz, f);
            if (p->forcing == 0) dens_out = FN(oracle_alpha_ref_forcing)(f, r1, p->gx, p->gy, p->gz);
            else dens_out = r1;
            store_dd = 1;
        }
        if (store_dd)
            for (int i = 0; i < 19; ++i) dd[(i ^ (i < 18)) * C + g] = f[i];
        if (p->store_velocity && vel) {                        /* :486-492 */
            vel[g] = vx; vel[C + g] = vy; vel[2 * C + g] = vz;
        }
        if (p->store_density && dens) dens[g] = dens_out;      /* :494-497 */
    }
}

/* lbm_kernel_beta (src/cl_programs/lbm_beta.cl:13-802), all cells.
 * Gather f_i = S[opp(i)][W(g - d_i)], scatter S[i][W(g + d_i)] = f_i*.
 * Every slot is read and written by exactly one cell, so the sequential and
 * the parallel loops give the same result. */
void FN(oracle_beta)(REAL *dd, const int32_t *flags, REAL *vel, REAL *dens,
                     int nx, int ny, int nz, const FN(oracle_params) *p) {
    const long long C = (long long)nx * ny * nz;
    const long long X = nx, XY = (long long)nx * ny;
    long long d[19];
    for (int i = 0; i < 19; ++i) d[i] = ORACLE_E[i][0] + X * ORACLE_E[i][1] + XY * ORACLE_E[i][2];
#pragma omp parallel for schedule(static)
    for (long long g = 0; g < C; ++g) {
        int flag = flags[g];
        REAL f[19];
        for (int i = 0; i < 18; ++i) f[i] = dd[(i ^ 1) * C + FN(wrap)(g - d[i], C)];
        f[18] = dd[18 * C + g];
        /* moments in the LDS kernel's summation order (lbm_beta.cl:268-483) */
        REAL rho = f[3], vx, vy = -f[3], vz;
        rho = rho + f[2];  vy = vy + f[2];
        rho = rho + f[0];  vx = f[0];
        rho = rho + f[1];  vx = vx - f[1];
        rho = rho + f[4];  vx = vx + f[4];  vy = vy + f[4];
        rho = rho + f[5];  vx = vx - f[5];  vy = vy - f[5];
        rho = rho + f[6];  vx = vx + f[6];  vy = vy - f[6];
        rho = rho + f[7];  vx = vx - f[7];  vy = vy + f[7];
        rho = rho + f[8];  vx = vx + f[8];  vz = f[8];
        rho = rho + f[9];  vx = vx - f[9];  vz = vz - f[9];
        rho = rho + f[10]; vx = vx + f[10]; vz = vz - f[10];
        rho = rho + f[11]; vx = vx - f[11]; vz = vz + f[11];
        rho = rho + f[13]; vy = vy - f[13]; vz = vz - f[13];
        rho = rho + f[12]; vy = vy + f[12]; vz = vz + f[12];
        rho = rho + f[15]; vy = vy - f[15]; vz = vz + f[15];
        rho = rho + f[14]; vy = vy + f[14]; vz = vz - f[14];
        rho = rho + f[17]; vz = vz - f[17];
        rho = rho + f[16]; vz = vz + f[16];
        rho = rho + f[18];
        REAL dens_out = rho;
        if (flag == FLAG_FLUID) {                              /* :497-559 */
            REAL dd_param = rho - C15 * FN(vel2_of)(vx, vy, vz);
            FN(oracle_collide_fluid)(f, rho, vx, vy, vz, p, dd_param);
            if (p->forcing == 1) {
                FN(oracle_body_force)(f, rho, p->gx, p->gy, p->gz);
                dens_out = rho;
            } else {
                dens_out = dd_param;                            /* #define dd_param rho (:493) */
            }
        } else if (flag == FLAG_OBSTACLE) {                    /* :561-580 */
            vx = 0; vy = 0; vz = 0;
            for (int i = 0; i < 18; i += 2) { REAL t = f[i + 1]; f[i + 1] = f[i]; f[i] = t; }
        } else if (flag == FLAG_INJECTION) {                   /* :582-653 */
            vx = p->lid; vy = 0; vz = 0;
            REAL dd_param = ((REAL)1.0) - C15 * FN(vel2_of)(vx, vy, vz);
            FN(oracle_feq)(dd_param, vx, vy, vz, f);
            dens_out = dd_param;
        }
        for (int i = 0; i < 18; ++i) dd[i * C + FN(wrap)(g + d[i], C)] = f[i];  /* :756-784 */
        dd[18 * C + g] = f[18];
        if (flag == FLAG_GHOST) continue;                      /* :787-788 */
        if (p->store_velocity && vel) { vel[g] = vx; vel[C + g] = vy; vel[2 * C + g] = vz; }
        if (p->store_density && dens) dens[g] = dens_out;
    }
}

/* init_kernel (src/cl_programs/lbm_init.cl:32-237) plus the seeded
 * perturbation (new).  Flags: x faces first, then y, then z (:55-68).
 * Cells on an OBSTACLE face keep f = feq(1, 0); every other cell gets
 * f = feq(1, du) with du_c = amp*(2*U[0,1) - 1) drawn from
 * oracle_unit(seed, case_id, key, c), key = the cell's linear index in the
 * key grid (key_origin + local position, key_dims).  The state is stored in
 * the swapped slots S[opp(i)] because the first step is a beta step. */
void FN(oracle_init)(REAL *dd, int32_t *flags, REAL *vel, REAL *dens,
                     int nx, int ny, int nz, const int *bc,
                     double amp, uint64_t seed, uint64_t case_id,
                     const int *key_origin, const int *key_dims) {
    const long long C = (long long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long long g = 0; g < C; ++g) {
        int x = (int)(g % nx), y = (int)((g / nx) % ny), z = (int)(g / ((long long)nx * ny));
        int flag = FLAG_FLUID;
        if (x == 0) flag = bc[0];
        else if (x == nx - 1) flag = bc[1];
        else if (y == 0) flag = bc[2];
        else if (y == ny - 1) flag = bc[3];
        else if (z == 0) flag = bc[4];
        else if (z == nz - 1) flag = bc[5];
        flags[g] = flag;
        int outer = (x == 0 && bc[0] == FLAG_OBSTACLE) || (x == nx - 1 && bc[1] == FLAG_OBSTACLE) ||
                    (y == 0 && bc[2] == FLAG_OBSTACLE) || (y == ny - 1 && bc[3] == FLAG_OBSTACLE) ||
                    (z == 0 && bc[4] == FLAG_OBSTACLE) || (z == nz - 1 && bc[5] == FLAG_OBSTACLE);
        REAL u[3] = {0, 0, 0};
        if (amp != 0.0 && !outer) {
            long long kx = key_origin[0] + x, ky = key_origin[1] + y, kz = key_origin[2] + z;
            uint64_t key = (uint64_t)(kx + (long long)key_dims[0] * (ky + (long long)key_dims[1] * kz));
            for (int c = 0; c < 3; ++c)
                u[c] = (REAL)(amp * (2.0 * oracle_unit(seed, case_id, key, c) - 1.0));
        }
        REAL r1 = (REAL)1.0;
        REAL dd_param = r1 - C15 * FN(vel2_of)(u[0], u[1], u[2]);
        REAL f[19];
        FN(oracle_feq)(dd_param, u[0], u[1], u[2], f);
        for (int i = 0; i < 19; ++i) dd[(i ^ (i < 18)) * C + g] = f[i];
        if (vel) { vel[g] = u[0]; vel[C + g] = u[1]; vel[2 * C + g] = u[2]; }
        if (dens) dens[g] = r1;
    }
}

/* copy_buffer_rect (src/cl_programs/copy_buffer_rect.cl:12-52): copy a
 * block from an x-fastest array into another, with element offsets. */
void FN(oracle_copy_rect)(const REAL *src, long long src_offset, const int *src_origin, const int *src_size,
                          REAL *dst, long long dst_offset, const int *dst_origin, const int *dst_size,
                          const int *block) {
    for (int k = 0; k < block[2]; ++k)
        for (int j = 0; j < block[1]; ++j) {
            long long s = src_offset + src_origin[0] + (long long)(src_origin[1] + j) * src_size[0] +
                          (long long)(src_origin[2] + k) * src_size[0] * src_size[1];
            long long d = dst_offset + dst_origin[0] + (long long)(dst_origin[1] + j) * dst_size[0] +
                          (long long)(dst_origin[2] + k) * dst_size[0] * dst_size[1];
            for (int x = 0; x < block[0]; ++x) dst[d + x] = src[s + x];
        }
}

/* storeDensityDistribution(T*, origin, size) (src/CLbmSolver.hpp:695-717):
 * dense [f][size] buffer from the [f][C] distribution array. */
void FN(oracle_store_dd_region)(const REAL *dd, const int *dims, const int *origin, const int *size, REAL *out) {
    long long C = (long long)dims[0] * dims[1] * dims[2];
    long long E = (long long)size[0] * size[1] * size[2];
    int zero[3] = {0, 0, 0};
    for (int f = 0; f < 19; ++f)
        FN(oracle_copy_rect)(dd, f * C, origin, dims, out, f * E, zero, size, size);
}

/* setDensityDistribution(T*, origin, size[, norm]) (src/CLbmSolver.hpp:719-757):
 * norm == NULL copies all 19 populations, otherwise only those with
 * norm . e_f > 0. */
void FN(oracle_set_dd_region)(REAL *dd, const int *dims, const int *origin, const int *size,
                              const REAL *src, const int *norm) {
    long long C = (long long)dims[0] * dims[1] * dims[2];
    long long E = (long long)size[0] * size[1] * size[2];
    int zero[3] = {0, 0, 0};
    for (int f = 0; f < 19; ++f) {
        if (norm) {
            int dot = norm[0] * ORACLE_E[f][0] + norm[1] * ORACLE_E[f][1] + norm[2] * ORACLE_E[f][2];
            if (dot <= 0) continue;
        }
        FN(oracle_copy_rect)(src, f * E, zero, size, dd, f * C, origin, dims, size);
    }
}

/* Direct device-side analogue of one CComm transfer: copy a box of the
 * distributions of one subdomain into another (all 19, or filtered by norm). */
void FN(oracle_copy_dd_box)(const REAL *src, const int *src_dims, const int *src_origin,
                            REAL *dst, const int *dst_dims, const int *dst_origin,
                            const int *size, const int *norm) {
    long long Cs = (long long)src_dims[0] * src_dims[1] * src_dims[2];
    long long Cd = (long long)dst_dims[0] * dst_dims[1] * dst_dims[2];
    for (int f = 0; f < 19; ++f) {
        if (norm) {
            int dot = norm[0] * ORACLE_E[f][0] + norm[1] * ORACLE_E[f][1] + norm[2] * ORACLE_E[f][2];
            if (dot <= 0) continue;
        }
        FN(oracle_copy_rect)(src, f * Cs, src_origin, src_dims, dst, f * Cd, dst_origin, dst_dims, size);
    }
}

/* Density and momentum from the distributions, read in the step-parity
 * dependent layout: after an odd number of steps (last kernel beta) slot i
 * holds f_i; after an even number (last kernel alpha) slot opp(i) holds f_i.
 * Sums are taken in double, in direction order. */
void FN(oracle_moments)(const REAL *dd, int nx, int ny, int nz, long long steps,
                        double *rho, double *jx, double *jy, double *jz) {
    const long long C = (long long)nx * ny * nz;
    int swapped = (steps % 2) == 0;
#pragma omp parallel for schedule(static)
    for (long long g = 0; g < C; ++g) {
        double r = 0, a = 0, b = 0, c = 0;
        for (int i = 0; i < 19; ++i) {
            int slot = swapped ? (i ^ (i < 18)) : i;
            double v = (double)dd[slot * C + g];
            r += v;
            a += v * ORACLE_E[i][0];
            b += v * ORACLE_E[i][1];
            c += v * ORACLE_E[i][2];
        }
        rho[g] = r; jx[g] = a; jy[g] = b; jz[g] = c;
    }
}

#undef W1
#undef W2
#undef W0
#undef C3
#undef C45
#undef C15
#undef FN
#undef CAT
#undef CAT2
