// lbm_device.hpp -- gfx950 device code of the D3Q19 AA-pattern step.
//
// One work-item updates one cell.  The alpha kernel reads the 19 slots of its
// own cell and writes them back swapped (src/cl_programs/lbm_alpha.cl:12-498);
// the beta kernel gathers from the 18 upwind neighbours and scatters to the
// 18 downwind neighbours (src/cl_programs/lbm_beta.cl:13-802).  Each slot is
// read and written by exactly one work-item per kernel, so both kernels run
// in place without barriers.
//
// Arithmetic is written as separate IEEE operations in the reference's
// evaluation order and the library is built with -ffp-contract=off, so the
// FP32 results are bit-identical to the un-contracted reference arithmetic.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int kFlagObstacle = 1;
constexpr int kFlagFluid = 2;
constexpr int kFlagInjection = 4;
constexpr int kFlagGhost = 8;

// D3Q19 directions, reference order (src/cl_programs/lbm_header.h:12-19).
__host__ __device__ constexpr int ex(int i) {
    return (i == 0 || i == 4 || i == 6 || i == 8 || i == 10) ? 1
           : (i == 1 || i == 5 || i == 7 || i == 9 || i == 11) ? -1
                                                               : 0;
}
__host__ __device__ constexpr int ey(int i) {
    return (i == 2 || i == 4 || i == 7 || i == 12 || i == 14) ? 1
           : (i == 3 || i == 5 || i == 6 || i == 13 || i == 15) ? -1
                                                                : 0;
}
__host__ __device__ constexpr int ez(int i) {
    return (i == 8 || i == 11 || i == 12 || i == 15 || i == 16) ? 1
           : (i == 9 || i == 10 || i == 13 || i == 14 || i == 17) ? -1
                                                                  : 0;
}
__host__ __device__ constexpr int opp(int i) { return i < 18 ? (i ^ 1) : 18; }

// Lattice constants of the compute type.  FP32 uses (float)(1.0f/18.0f) etc.
// exactly as the reference (src/cl_programs/lbm_header.h:68-94); FP64 uses
// exact double constants (the reference's double build keeps float-rounded
// weights; the FP64 path here is the precision yardstick).
template <typename T>
struct LC;
template <>
struct LC<float> {
    static constexpr float w1 = 1.0f / 18.0f, w2 = 1.0f / 36.0f, w0 = 1.0f / 3.0f;
    static constexpr float c1 = 1.0f / 6.0f, c2 = 1.0f / 12.0f;
};
template <>
struct LC<double> {
    static constexpr double w1 = 1.0 / 18.0, w2 = 1.0 / 36.0, w0 = 1.0 / 3.0;
    static constexpr double c1 = 1.0 / 6.0, c2 = 1.0 / 12.0;
};
template <typename T>
__host__ __device__ constexpr T weight(int i) {
    return i == 18 ? LC<T>::w0 : (i < 4 || i >= 16) ? LC<T>::w1 : LC<T>::w2;
}

// ---- storage of the distributions -------------------------------------------
// 0: f32, 1: f64, 2: f16 holding f - w, 3: bf16 holding f - w.
// Accesses go through address-space-1 (global) pointers: a uniform 64-bit
// plane base in SGPRs plus a 32-bit per-lane byte offset, which the compiler
// turns into `global_load/store ... v_off, s[base]` (no 64-bit VALU address
// arithmetic, no flat instructions).
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) unsigned short gushort;
typedef __attribute__((address_space(1))) unsigned char guchar;

__device__ __forceinline__ unsigned short bf16_bits_rne(float v) {
    unsigned u = __float_as_uint(v);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Raw element access; NT selects the non-temporal (streaming) cache policy.
template <bool NT, typename U>
__device__ __forceinline__ U raw_ld(const gchar *base, unsigned off) {
    const __attribute__((address_space(1))) U *p = (const __attribute__((address_space(1))) U *)(base + off);
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT, typename U>
__device__ __forceinline__ void raw_st(gchar *base, unsigned off, U v) {
    __attribute__((address_space(1))) U *p = (__attribute__((address_space(1))) U *)(base + off);
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int ST>
struct Storage;
template <>
struct Storage<0> {
    using E = float;
    using T = float;
    static constexpr unsigned kBytes = 4;
    template <bool NT = false>
    __device__ static T ld(const gchar *base, unsigned off, int) { return raw_ld<NT, float>(base, off); }
    template <bool NT = false>
    __device__ static void st(gchar *base, unsigned off, int, T v) { raw_st<NT, float>(base, off, v); }
};
template <>
struct Storage<1> {
    using E = double;
    using T = double;
    static constexpr unsigned kBytes = 8;
    template <bool NT = false>
    __device__ static T ld(const gchar *base, unsigned off, int) { return raw_ld<NT, double>(base, off); }
    template <bool NT = false>
    __device__ static void st(gchar *base, unsigned off, int, T v) { raw_st<NT, double>(base, off, v); }
};
template <>
struct Storage<2> {
    using E = __half;
    using T = float;
    static constexpr unsigned kBytes = 2;
    template <bool NT = false>
    __device__ static T ld(const gchar *base, unsigned off, int i) {
        return __half2float(__ushort_as_half(raw_ld<NT, unsigned short>(base, off))) + weight<float>(i);
    }
    template <bool NT = false>
    __device__ static void st(gchar *base, unsigned off, int i, T v) {
        raw_st<NT, unsigned short>(base, off, __half_as_ushort(__float2half(v - weight<float>(i))));
    }
};
template <>
struct Storage<3> {
    using E = __hip_bfloat16;
    using T = float;
    static constexpr unsigned kBytes = 2;
    template <bool NT = false>
    __device__ static T ld(const gchar *base, unsigned off, int i) {
        return __uint_as_float((unsigned)raw_ld<NT, unsigned short>(base, off) << 16) + weight<float>(i);
    }
    template <bool NT = false>
    __device__ static void st(gchar *base, unsigned off, int i, T v) {
        raw_st<NT, unsigned short>(base, off, bf16_bits_rne(v - weight<float>(i)));
    }
};

// Step-kernel variants (tuning knobs): bit 0 non-temporal loads, bit 1
// non-temporal stores.
constexpr int kVarNtLoad = 1;
constexpr int kVarNtStore = 2;

// ---- per-(sub)domain launch descriptor ---------------------------------------
struct DomDev {
    void *dd;          // 19 planes of `plane` elements
    uint8_t *flags;    // one byte per cell
    void *vel;         // [3][cells] compute type, or null
    void *dens;        // [cells] compute type, or null
    long long plane;   // plane stride in elements
    long long cells;
    int nx, ny, nz;
    int tile_begin;    // first tile of this domain in the launch
    int ntiles;
    int forcing;
    int store_velocity;
    int store_density;
    double inv_tau, tau0, les_coeff, gx, gy, gz, lid;
};

// ---- collision arithmetic (same operation order as the oracle) -------------
// eq_dd_a0/a1/eq_dd4/eq_dd5 (src/cl_programs/lbm_header.h:68-88):
// W * ((r + (+-3)*v) + 4.5*v2)
template <typename T>
__device__ __forceinline__ T eqv(T w, T c3, T v, T v2, T r) {
    T a = c3 * v;
    T b = r + a;
    T c = T(4.5) * v2;
    T d = b + c;
    return w * d;
}

// Equilibrium of all 19 directions from dd_param = rho - 1.5|u|^2 and u
// (product order of src/cl_programs/lbm_alpha.cl:352-478).
template <typename T>
__device__ __forceinline__ void feq19(T dd_param, T vx, T vy, T vz, T *o) {
    const T w1 = LC<T>::w1, w2 = LC<T>::w2;
    T v2, ab, ab2;
    v2 = vx * vx;
    o[0] = eqv(w1, T(3), vx, v2, dd_param);
    o[1] = eqv(w1, T(-3), vx, v2, dd_param);
    v2 = vy * vy;
    o[2] = eqv(w1, T(3), vy, v2, dd_param);
    o[3] = eqv(w1, T(-3), vy, v2, dd_param);
    ab = vx + vy; ab2 = ab * ab;
    o[4] = eqv(w2, T(3), ab, ab2, dd_param);
    o[5] = eqv(w2, T(-3), ab, ab2, dd_param);
    ab = vx - vy; ab2 = ab * ab;
    o[6] = eqv(w2, T(3), ab, ab2, dd_param);
    o[7] = eqv(w2, T(-3), ab, ab2, dd_param);
    ab = vx + vz; ab2 = ab * ab;
    o[8] = eqv(w2, T(3), ab, ab2, dd_param);
    o[9] = eqv(w2, T(-3), ab, ab2, dd_param);
    ab = vx - vz; ab2 = ab * ab;
    o[10] = eqv(w2, T(3), ab, ab2, dd_param);
    o[11] = eqv(w2, T(-3), ab, ab2, dd_param);
    ab = vy + vz; ab2 = ab * ab;
    o[12] = eqv(w2, T(3), ab, ab2, dd_param);
    o[13] = eqv(w2, T(-3), ab, ab2, dd_param);
    ab = vy - vz; ab2 = ab * ab;
    o[14] = eqv(w2, T(3), ab, ab2, dd_param);
    o[15] = eqv(w2, T(-3), ab, ab2, dd_param);
    v2 = vz * vz;
    o[16] = eqv(w1, T(3), vz, v2, dd_param);
    o[17] = eqv(w1, T(-3), vz, v2, dd_param);
    o[18] = LC<T>::w0 * dd_param;
}

template <typename T>
__device__ __forceinline__ T vel2_of(T vx, T vy, T vz) {
    T a = vx * vx;
    T b = vy * vy;
    T c = a + b;
    T d = vz * vz;
    return c + d;
}

// Smagorinsky rate (Hou, Sterling, Chen & Doolen 1996):
// tau_eff = (tau0 + sqrt(tau0^2 + les_coeff*|Pi_neq|/rho)) / 2,
// les_coeff = 18 sqrt(2) Cs^2, Pi_neq_ab = sum e_a e_b f - rho/3 d_ab - j_a j_b.
// The 19 populations are in this work-item's registers: no cross-lane traffic.
template <typename T>
__device__ __forceinline__ T les_inv_tau(const T *f, T rho, T jx, T jy, T jz, T tau0, T les_coeff) {
    T pxx = f[0] + f[1];
    pxx = pxx + f[4]; pxx = pxx + f[5]; pxx = pxx + f[6]; pxx = pxx + f[7];
    pxx = pxx + f[8]; pxx = pxx + f[9]; pxx = pxx + f[10]; pxx = pxx + f[11];
    T pyy = f[2] + f[3];
    pyy = pyy + f[4]; pyy = pyy + f[5]; pyy = pyy + f[6]; pyy = pyy + f[7];
    pyy = pyy + f[12]; pyy = pyy + f[13]; pyy = pyy + f[14]; pyy = pyy + f[15];
    T pzz = f[8] + f[9];
    pzz = pzz + f[10]; pzz = pzz + f[11]; pzz = pzz + f[12]; pzz = pzz + f[13];
    pzz = pzz + f[14]; pzz = pzz + f[15]; pzz = pzz + f[16]; pzz = pzz + f[17];
    T pxy = (f[4] + f[5]) - (f[6] + f[7]);
    T pxz = (f[8] + f[9]) - (f[10] + f[11]);
    T pyz = (f[12] + f[13]) - (f[14] + f[15]);
    T r3 = rho * LC<T>::w0;
    T qxx = (pxx - r3) - jx * jx;
    T qyy = (pyy - r3) - jy * jy;
    T qzz = (pzz - r3) - jz * jz;
    T qxy = pxy - jx * jy;
    T qxz = pxz - jx * jz;
    T qyz = pyz - jy * jz;
    T diag = (qxx * qxx + qyy * qyy) + qzz * qzz;
    T off = (qxy * qxy + qxz * qxz) + qyz * qyz;
    T q2 = diag + T(2) * off;
    T q = sqrt(q2);
    T t = (les_coeff * q) / rho;
    T s = sqrt(tau0 * tau0 + t);
    T tau_eff = T(0.5) * (tau0 + s);
    return T(1) / tau_eff;
}

template <typename T>
__device__ __forceinline__ void body_force(T *f, T rho, T gx, T gy, T gz) {
    const T c1 = LC<T>::c1, c2 = LC<T>::c2;
    T fx = rho * gx, fy = rho * gy, fz = rho * gz;
    T t;
    t = c1 * fx;        f[0] = f[0] + t;   f[1] = f[1] - t;
    t = c1 * fy;        f[2] = f[2] + t;   f[3] = f[3] - t;
    t = c2 * (fx + fy); f[4] = f[4] + t;   f[5] = f[5] - t;
    t = c2 * (fx - fy); f[6] = f[6] + t;   f[7] = f[7] - t;
    t = c2 * (fx + fz); f[8] = f[8] + t;   f[9] = f[9] - t;
    t = c2 * (fx - fz); f[10] = f[10] + t; f[11] = f[11] - t;
    t = c2 * (fy + fz); f[12] = f[12] + t; f[13] = f[13] - t;
    t = c2 * (fy - fz); f[14] = f[14] + t; f[15] = f[15] - t;
    t = c1 * fz;        f[16] = f[16] + t; f[17] = f[17] - t;
}

// Alpha-kernel forcing with the reference's `#define tmp rho` alias: each term
// is scaled by the previous term (src/cl_programs/lbm_alpha.cl:173,183-298).
// Returns the aliased variable, which the reference stores as density.
template <typename T>
__device__ __forceinline__ T alpha_ref_forcing(T *f, T rho, T gx, T gy, T gz) {
    const T w1 = LC<T>::w1, w2 = LC<T>::w2;
    T tmp;
    tmp = (gx * w1) * rho;           rho = tmp; f[1] = f[1] - tmp;   f[0] = f[0] + tmp;
    tmp = (gy * (-w1)) * rho;        rho = tmp; f[3] = f[3] - tmp;   f[2] = f[2] + tmp;
    tmp = ((gx - gy) * w2) * rho;    rho = tmp; f[5] = f[5] - tmp;   f[4] = f[4] + tmp;
    tmp = ((gx + gy) * w2) * rho;    rho = tmp; f[7] = f[7] - tmp;   f[6] = f[6] + tmp;
    tmp = ((gx + gz) * w2) * rho;    rho = tmp; f[9] = f[9] - tmp;   f[8] = f[8] + tmp;
    tmp = ((gx - gz) * w2) * rho;    rho = tmp; f[11] = f[11] - tmp; f[10] = f[10] + tmp;
    tmp = ((gz - gy) * w2) * rho;    rho = tmp; f[13] = f[13] - tmp; f[12] = f[12] + tmp;
    tmp = ((gz + gy) * (-w2)) * rho; rho = tmp; f[15] = f[15] - tmp; f[14] = f[14] + tmp;
    tmp = (gz * w1) * rho;           rho = tmp; f[17] = f[17] - tmp; f[16] = f[16] + tmp;
    return rho;
}

// FLUID collision shared by both kernels: optional LES rate, then BGK
// f_i += w (feq_i - f_i) (src/cl_programs/lbm_alpha.cl:185-300).
template <typename T>
__device__ __forceinline__ void collide_fluid(T *f, T rho, T vx, T vy, T vz, T dd_param, T inv_tau, T tau0,
                                              T les_coeff) {
    T w = inv_tau;
    if (les_coeff > T(0)) w = les_inv_tau(f, rho, vx, vy, vz, tau0, les_coeff);
    T q[19];
    feq19(dd_param, vx, vy, vz, q);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        T d = q[i] - f[i];
        T e = w * d;
        f[i] = f[i] + e;
    }
}

// Block -> tile map that gives each of the 8 XCDs one contiguous run of tiles
// (work-groups are dispatched to XCDs round-robin by block index).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    int q = nb >> 3, r = nb & 7;
    int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ int wrap_idx(int a, int cells) {
    a = a < 0 ? a + cells : a;
    return a >= cells ? a - cells : a;
}

// ---- the step kernel -------------------------------------------------------
// Gather of the beta kernel: f_i = S[opp(i)][W(g - d_i)].  WRAP is false for
// blocks whose cells are all at least max|d| away from both ends of the
// domain, where the linear wrap W (src/cl_programs/wrap.h:112-127) is the
// identity.
template <class SP, bool WRAP, bool NT>
__device__ __forceinline__ void beta_gather(gchar *const *pl, unsigned g, int cells, const int *d,
                                            typename SP::T *f) {
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        unsigned idx = WRAP ? (unsigned)wrap_idx((int)g - d[i], cells) : g - (unsigned)d[i];
        f[i] = SP::template ld<NT>(pl[opp(i)], idx * SP::kBytes, opp(i));
    }
    f[18] = SP::template ld<NT>(pl[18], g * SP::kBytes, 18);
}

template <class SP, bool WRAP, bool NT>
__device__ __forceinline__ void beta_scatter(gchar *const *pl, unsigned g, int cells, const int *d,
                                             const typename SP::T *f) {
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        unsigned idx = WRAP ? (unsigned)wrap_idx((int)g + d[i], cells) : g + (unsigned)d[i];
        SP::template st<NT>(pl[i], idx * SP::kBytes, i, f[i]);
    }
    SP::template st<NT>(pl[18], g * SP::kBytes, 18, f[18]);
}

template <int ST, bool BETA, int VAR = 0>
__global__ void __launch_bounds__(1024) step_kernel(const DomDev *__restrict__ doms, const int *__restrict__ tile_dom,
                                                    int xcd_map) {
    using SP = Storage<ST>;
    using T = typename SP::T;
    constexpr bool NTL = (VAR & kVarNtLoad) != 0;
    constexpr bool NTS = (VAR & kVarNtStore) != 0;
    const int nb = (int)gridDim.x;
    const int t = xcd_map ? xcd_tile((int)blockIdx.x, nb) : (int)blockIdx.x;
    const DomDev *D = doms + tile_dom[t];
    const int cells = (int)D->cells;
    const int block_first = (t - D->tile_begin) * (int)blockDim.x;
    const int gi = block_first + (int)threadIdx.x;
    if (gi >= cells) return;
    const unsigned g = (unsigned)gi;
    gchar *pl[19];
    {
        gchar *base = (gchar *)D->dd;
        const size_t pb = (size_t)D->plane * SP::kBytes;
#pragma unroll
        for (int i = 0; i < 19; ++i) pl[i] = base + (size_t)i * pb;
    }
    const int flag = ((const guchar *)D->flags)[g];

    T f[19];
    int d[19];
    bool wrap = false;
    if (BETA) {
        const int X = D->nx, XY = D->nx * D->ny;
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = ex(i) + X * ey(i) + XY * ez(i);
        const int dmax = 1 + X + XY;
        wrap = block_first < dmax || block_first + (int)blockDim.x + dmax > cells;
        if (wrap)
            beta_gather<SP, true, NTL>(pl, g, cells, d, f);
        else
            beta_gather<SP, false, NTL>(pl, g, cells, d, f);
    } else {
        // the 19 loads do not wait for the flag; a GHOST cell (lbm_alpha.cl:30-32)
        // matches no branch below and stores nothing
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = SP::template ld<NTL>(pl[i], g * SP::kBytes, i);
    }

    T rho, vx, vy, vz;
    if (BETA) {  // summation order of lbm_beta.cl:268-483
        rho = f[3]; vy = -f[3];
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
    } else {  // summation order of lbm_alpha.cl:61-156
        rho = f[0]; vx = f[0];
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
    }

    T dens_out = rho;
    bool write_dd = BETA;
    if (flag == kFlagFluid) {
        const T dd_param = rho - T(1.5) * vel2_of(vx, vy, vz);
        collide_fluid(f, rho, vx, vy, vz, dd_param, (T)D->inv_tau, (T)D->tau0, (T)D->les_coeff);
        if (D->forcing == 1) {
            body_force(f, rho, (T)D->gx, (T)D->gy, (T)D->gz);
        } else if (!BETA) {
            dens_out = alpha_ref_forcing(f, rho, (T)D->gx, (T)D->gy, (T)D->gz);
        } else {
            dens_out = dd_param;  // `#define dd_param rho` (lbm_beta.cl:493)
        }
        write_dd = true;
    } else if (flag == kFlagObstacle) {
        vx = T(0); vy = T(0); vz = T(0);
        if (BETA) {  // lbm_beta.cl:570-578
#pragma unroll
            for (int i = 0; i < 18; i += 2) {
                T tt = f[i + 1];
                f[i + 1] = f[i];
                f[i] = tt;
            }
        }
    } else if (flag == kFlagInjection) {  // lbm_alpha.cl:345-480, lbm_beta.cl:582-653
        vx = (T)D->lid; vy = T(0); vz = T(0);
        const T one = T(1);
        const T dd_param = one - T(1.5) * vel2_of(vx, vy, vz);
        feq19(dd_param, vx, vy, vz, f);
        if (BETA) dens_out = dd_param;
        else if (D->forcing == 0) dens_out = alpha_ref_forcing(f, one, (T)D->gx, (T)D->gy, (T)D->gz);
        else dens_out = one;
        write_dd = true;
    }

    if (write_dd) {
        if (BETA) {
            if (wrap)
                beta_scatter<SP, true, NTS>(pl, g, cells, d, f);
            else
                beta_scatter<SP, false, NTS>(pl, g, cells, d, f);
        } else {
#pragma unroll
            for (int i = 0; i < 19; ++i) SP::template st<NTS>(pl[opp(i)], g * SP::kBytes, opp(i), f[i]);
        }
    }
    if (flag == kFlagGhost) return;  // lbm_alpha.cl:30-32, lbm_beta.cl:787-788
    if (D->store_velocity) {
        gchar *v = (gchar *)D->vel;
        const unsigned cs = (unsigned)sizeof(T);
        *(__attribute__((address_space(1))) T *)(v + g * cs) = vx;
        *(__attribute__((address_space(1))) T *)(v + (size_t)cells * cs + g * cs) = vy;
        *(__attribute__((address_space(1))) T *)(v + 2 * (size_t)cells * cs + g * cs) = vz;
    }
    if (D->store_density) *(__attribute__((address_space(1))) T *)((gchar *)D->dens + g * (unsigned)sizeof(T)) = dens_out;
}

}  // namespace lbm
