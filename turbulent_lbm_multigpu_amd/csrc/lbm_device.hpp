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

#include "exchange_plan.hpp"  // kMaxChain, kChainRows
#include "launch_shape.hpp"   // variant bits of the XCD tile map

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
//
// Layout in HBM: blocks of kBlk = 4096 consecutive cells (cell order
// g = x + X*(y + Y*z)) hold their 19 populations one after the other, so
// population i of cell g is element
//     (g / kBlk) * 19 * kBlk + i * kBlk + g % kBlk.
// Inside a block each population is a contiguous run of cells, so a wave's
// accesses are as coalesced as in a [19][cells] layout, but the 38 streams of
// a work-group stay inside one 19 * kBlk element window instead of 19 planes
// a whole domain apart (on MI355X the in-place 19-load/19-store pattern over
// eight 128^3 domains moves ~20% more bytes per second this way).
//
// Accesses go through address-space-1 (global) pointers: a uniform 64-bit
// population base (dd + i * kBlk elements) in SGPRs plus a 32-bit per-lane
// byte offset (cell_base(g) elements), which the compiler turns into
// `global_load/store ... v_off, s[base]` (no 64-bit VALU address arithmetic,
// no flat instructions).  The whole array of one domain therefore stays below
// 4 GiB.
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) unsigned short gushort;
typedef __attribute__((address_space(1))) unsigned char guchar;

constexpr unsigned kBlkShift = 12;
constexpr unsigned kBlk = 1u << kBlkShift;
// element offset of cell g from its population base
__host__ __device__ __forceinline__ unsigned cell_base(unsigned g) {
    return (g >> kBlkShift) * (19u * kBlk) + (g & (kBlk - 1u));
}

// float -> bf16 bits, round to nearest even: gfx950's v_cvt_pk_bf16_f32. On
// every finite float (all 2^32 patterns checked on the device) it gives the
// bits of the integer rounding u + 0x7fff + ((u >> 16) & 1) that the CPU
// oracle uses; NaNs stay NaNs (their payload bits may differ).
__device__ __forceinline__ unsigned short bf16_bits_rne(float v) {
    return __builtin_bit_cast(unsigned short, (__bf16)v);
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

// Storage traits: U is the stored bit pattern, T the compute type; from_bits /
// to_bits convert one element of population slot i.
template <int ST>
struct Storage;
// kShifted storages hold f - w_i: from_bits/to_bits convert to and from f,
// from_raw/to_raw keep the shifted value (the step kernel collides in it).
template <>
struct Storage<0> {
    using U = unsigned;
    using T = float;
    static constexpr unsigned kBytes = 4;
    static constexpr bool kShifted = false;
    __device__ static T from_bits(U u, int) { return __uint_as_float(u); }
    __device__ static U to_bits(T v, int) { return __float_as_uint(v); }
    __device__ static T from_raw(U u) { return __uint_as_float(u); }
    __device__ static U to_raw(T v) { return __float_as_uint(v); }
};
template <>
struct Storage<1> {
    using U = unsigned long long;
    using T = double;
    static constexpr unsigned kBytes = 8;
    static constexpr bool kShifted = false;
    __device__ static T from_bits(U u, int) { return __longlong_as_double((long long)u); }
    __device__ static U to_bits(T v, int) { return (U)__double_as_longlong(v); }
    __device__ static T from_raw(U u) { return __longlong_as_double((long long)u); }
    __device__ static U to_raw(T v) { return (U)__double_as_longlong(v); }
};
template <>
struct Storage<2> {
    using U = unsigned short;
    using T = float;
    static constexpr unsigned kBytes = 2;
    static constexpr bool kShifted = true;
    __device__ static T from_raw(U u) { return __half2float(__ushort_as_half(u)); }
    __device__ static U to_raw(T v) { return __half_as_ushort(__float2half(v)); }
    __device__ static T from_bits(U u, int i) { return from_raw(u) + weight<float>(i); }
    __device__ static U to_bits(T v, int i) { return to_raw(v - weight<float>(i)); }
};
template <>
struct Storage<3> {
    using U = unsigned short;
    using T = float;
    static constexpr unsigned kBytes = 2;
    static constexpr bool kShifted = true;
    __device__ static T from_raw(U u) { return __uint_as_float((unsigned)u << 16); }
    __device__ static U to_raw(T v) { return bf16_bits_rne(v); }
    __device__ static T from_bits(U u, int i) { return from_raw(u) + weight<float>(i); }
    __device__ static U to_bits(T v, int i) { return to_raw(v - weight<float>(i)); }
};

// Conversions of the step kernel: shifted storages stay shifted in registers.
template <class SP>
__device__ __forceinline__ typename SP::T kin(typename SP::U u, int i) {
    if (SP::kShifted) return SP::from_raw(u);
    return SP::from_bits(u, i);
}
template <class SP>
__device__ __forceinline__ typename SP::U kout(typename SP::T v, int i) {
    if (SP::kShifted) return SP::to_raw(v);
    return SP::to_bits(v, i);
}

// one element of population slot i at byte offset `off` from its population base
template <class SP, bool NT = false>
__device__ __forceinline__ typename SP::T ld(const gchar *base, unsigned off, int i) {
    return SP::from_bits(raw_ld<NT, typename SP::U>(base, off), i);
}
template <class SP, bool NT = false>
__device__ __forceinline__ void st(gchar *base, unsigned off, int i, typename SP::T v) {
    raw_st<NT, typename SP::U>(base, off, SP::to_bits(v, i));
}

// the step kernel's element access (kin/kout conversions)
template <class SP, bool NT = false>
__device__ __forceinline__ typename SP::T kld(const gchar *base, unsigned off, int i) {
    return kin<SP>(raw_ld<NT, typename SP::U>(base, off), i);
}
template <class SP, bool NT = false>
__device__ __forceinline__ void kst(gchar *base, unsigned off, int i, typename SP::T v) {
    raw_st<NT, typename SP::U>(base, off, kout<SP>(v, i));
}

// N consecutive elements as one vector access (alignment: one element),
// with the step kernel's conversions.
template <class SP, int N>
__device__ __forceinline__ void vld(const gchar *base, unsigned off, int i, typename SP::T *out) {
    typedef typename SP::U V __attribute__((ext_vector_type(N)));
    V v;
    __builtin_memcpy(&v, base + off, sizeof(V));
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = kin<SP>(v[j], i);
}
template <class SP, int N>
__device__ __forceinline__ void vst(gchar *base, unsigned off, int i, const typename SP::T *in) {
    typedef typename SP::U V __attribute__((ext_vector_type(N)));
    V v;
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = kout<SP>(in[j], i);
    __builtin_memcpy(base + off, &v, sizeof(V));
}

// Step-kernel variant bits (tuning knobs): bit 0 non-temporal loads and bit 1
// non-temporal stores of the distributions on the one-cell path (with the
// blocked layout the streaming cache policy moves the in-place pattern ~13 %
// faster), bit 2 no XCD tile map (host side), bits 3-4 log2 of the number of
// x-consecutive cells one work-item updates (1, 2 or 4).
constexpr int kVarNtLoad = 1;
constexpr int kVarNtStore = 2;
__host__ __device__ constexpr int cells_per_lane(int var) { return 1 << ((var >> 3) & 3); }

// ---- per-(sub)domain launch descriptor ---------------------------------------
struct DomDev {
    void *dd;          // blocks of kBlk cells x 19 populations
    uint8_t *flags;    // one byte per cell
    void *vel;         // [3][cells] compute type, or null
    void *dens;        // [cells] compute type, or null
    long long cells;
    int nx, ny, nz;
    int tile_begin;    // first tile of this domain in the launch
    int ntiles;
    int forcing;
    int store_velocity;
    int store_density;
    double inv_tau, tau0, les_coeff, gx, gy, gz, lid;
    long long stop_at;  // launch half-step index from which this domain stays frozen
    unsigned xmask_a[2];  // slots this domain's x- / x+ ghost layer takes after alpha in a chain tile
    int seam_wpr;         // > 0: periodic x seam inside the pair kernel, wavefronts per row (seam_pairs)
};

// ---- fused x chains ----------------------------------------------------------
// Same-sized subdomains joined across x faces form a chain (a line, or a ring
// when x is periodic).  A chain tile holds the same `rows` (y, z) rows of every
// member, so both columns next to each shared face are in one work-group: the
// face values move through LDS and each population slot is stored only by the
// work-item that owns it in the AA pattern (whole cache lines, no stores into
// another subdomain), with the result of the copy-based exchange.  The host
// side (joins, chains, tile shape) is in exchange_plan.hpp.
struct XChain {
    int nmem, rows, nx, tile_begin;
    int lanes;             // whole-row tiles: work-items per member, rows * nx rounded up to whole wavefronts
    int nedge;             // run tiles: 64-cell runs holding columns 0, 1, X-2, X-1 (else 0)
    int edge[3];           //   their run indices
    int ninner, ipasses;   //   interior runs 1 .. ninner per member and row, in ipasses tiles after each handoff tile
    int mem[kMaxChain];    // descriptor index of each position
    int left[kMaxChain];   // position whose x = X-2 column feeds my x = 0 ghost layer, or -1
    int right[kMaxChain];  // position whose x = 1 column feeds my x = X-1 ghost layer, or -1
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

// Block -> tile map in chunks: the k-th block dispatched to XCD x takes
// offset k % K of chunk (k / K) * 8 + x, so each XCD walks K consecutive tiles
// while the 8 XCDs together sweep 8 K adjacent tiles.  The blocks of the
// last incomplete round of 8 K keep their own index.  A bijection of
// [0, nb) for any nb and K > 0.
__device__ __forceinline__ int xcd_chunk_tile(int b, int nb, int K) {
    const int full = nb / (8 * K) * (8 * K);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / K) * 8 + x) * K + (k - (k / K) * K);
}

__device__ __forceinline__ int wrap_idx(int a, int cells) {
    a = a < 0 ? a + cells : a;
    return a >= cells ? a - cells : a;
}

// ---- the step kernel -------------------------------------------------------
// The two-byte storages (f[] holds f_i - w_i in FP32): the update of one cell
// from the sums s_k = f_a + f_b and differences d_k = f_a - f_b of the nine
// opposite pairs (a, b) = (2k, 2k + 1), k over the directions
//   x, y, x+y, x-y, x+z, x-z, y+z, y-z, z.
// rho, j and the Smagorinsky stress come from them with 48 additions instead
// of 81.  The equilibrium, the relaxation and the body force of a pair along
// v = e_a . j take one product, one fused product-sum and two sums each:
//   P = omega w (dd + 4.5 v^2),  Q = 3 omega w v + c F_v,
//   f_a <- (1 - omega) f_a + (P + Q),  f_b <- (1 - omega) f_b + (P - Q),
// which is f += omega (f_eq - f), then f += the force term, regrouped.  In
// shifted values rho sums to rho - 1, j is unchanged (sum w_i e_i = 0), the
// equilibrium from dd = (rho - 1) - 1.5 |j|^2 is f_eq - w, and the stress loses
// the w terms exactly (sum e_a e_b w_i = delta_ab / 3 cancels rho / 3); the
// full density enters only the LES divisor, the forcing and the density
// output.  The LES rate uses the hardware square root and reciprocal (1 ulp in
// FP32), far below the storage's rounding.  Only these storages take this
// form: the FP32 and FP64 kernels keep the reference's expression order, which
// their bit parity with the CPU oracle rests on.
template <bool BETA>
__device__ __forceinline__ bool cell_update_pairs(float *f, int flag, const DomDev *D, float &vx, float &vy,
                                                  float &vz, float &dens_out) {
    float s[9], d[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        s[k] = f[2 * k] + f[2 * k + 1];
        d[k] = f[2 * k] - f[2 * k + 1];
    }
    const float s23 = s[2] + s[3], s45 = s[4] + s[5], s67 = s[6] + s[7];
    const float rho = ((s[0] + s[1]) + (s[8] + f[18])) + ((s23 + s45) + s67);  // rho - 1
    vx = ((d[0] + d[2]) + (d[3] + d[4])) + d[5];
    vy = ((d[1] + d[2]) - (d[3] - d[6])) + d[7];
    vz = ((d[8] + d[4]) - (d[5] - d[6])) - d[7];
    const float rho_full = 1.0f + rho;
    dens_out = rho_full;
    if (flag == kFlagFluid) {
        const float ddp = rho - 1.5f * vel2_of(vx, vy, vz);
        float om = (float)D->inv_tau;
        const float les = (float)D->les_coeff;
        if (les > 0.0f) {
            const float r3 = rho * LC<float>::w0;
            const float qxx = ((s[0] + s23) + (s45 - r3)) - vx * vx;
            const float qyy = ((s[1] + s23) + (s67 - r3)) - vy * vy;
            const float qzz = ((s[8] + s45) + (s67 - r3)) - vz * vz;
            const float qxy = (s[2] - s[3]) - vx * vy;
            const float qxz = (s[4] - s[5]) - vx * vz;
            const float qyz = (s[6] - s[7]) - vy * vz;
            const float diag = (qxx * qxx + qyy * qyy) + qzz * qzz;
            const float off = (qxy * qxy + qxz * qxz) + qyz * qyz;
            const float q = __builtin_amdgcn_sqrtf(diag + 2.0f * off);
            const float tau0 = (float)D->tau0;
            const float t = (les * q) * __builtin_amdgcn_rcpf(rho_full);
            const float root = __builtin_amdgcn_sqrtf(tau0 * tau0 + t);
            om = 2.0f * __builtin_amdgcn_rcpf(tau0 + root);
        }
        const float om1 = 1.0f - om;
        const float o1 = om * LC<float>::w1, o2 = om * LC<float>::w2;
        const float a1 = o1 * ddp, a2 = o2 * ddp, b1 = 4.5f * o1, b2 = 4.5f * o2, c1 = 3.0f * o1, c2 = 3.0f * o2;
        const float v[9] = {vx, vy, vx + vy, vx - vy, vx + vz, vx - vz, vy + vz, vy - vz, vz};
        float qv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) qv[k] = (k < 2 || k == 8 ? c1 : c2) * v[k];
        const bool body = D->forcing == 1;
        if (body) {
            const float fx = rho_full * (float)D->gx, fy = rho_full * (float)D->gy, fz = rho_full * (float)D->gz;
            const float fv[9] = {fx, fy, fx + fy, fx - fy, fx + fz, fx - fz, fy + fz, fy - fz, fz};
#pragma unroll
            for (int k = 0; k < 9; ++k) qv[k] += (k < 2 || k == 8 ? LC<float>::c1 : LC<float>::c2) * fv[k];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const bool axis = k < 2 || k == 8;
            const float pv = (axis ? b1 : b2) * (v[k] * v[k]) + (axis ? a1 : a2);
            f[2 * k] = om1 * f[2 * k] + (pv + qv[k]);
            f[2 * k + 1] = om1 * f[2 * k + 1] + (pv - qv[k]);
        }
        f[18] = om1 * f[18] + (om * LC<float>::w0) * ddp;
        if (!body) {
            if (!BETA) dens_out = alpha_ref_forcing(f, rho_full, (float)D->gx, (float)D->gy, (float)D->gz);
            else dens_out = 1.0f + ddp;  // `#define dd_param rho` (lbm_beta.cl:493)
        }
        return true;
    }
    if (flag == kFlagObstacle) {
        vx = 0.0f; vy = 0.0f; vz = 0.0f;
        if (BETA) {  // lbm_beta.cl:570-578
#pragma unroll
            for (int i = 0; i < 18; i += 2) {
                const float tt = f[i + 1];
                f[i + 1] = f[i];
                f[i] = tt;
            }
        }
        return BETA;
    }
    if (flag == kFlagInjection) {  // lbm_alpha.cl:345-480, lbm_beta.cl:582-653
        vx = (float)D->lid; vy = 0.0f; vz = 0.0f;
        const float dd_param = -1.5f * vel2_of(vx, vy, vz);
        feq19(dd_param, vx, vy, vz, f);
        if (BETA) dens_out = 1.0f + dd_param;
        else if (D->forcing == 0) dens_out = alpha_ref_forcing(f, 1.0f, (float)D->gx, (float)D->gy, (float)D->gz);
        else dens_out = 1.0f;
        return true;
    }
    return BETA;
}

// Moments in the kernel's summation order, the flag branch and the collision
// of one cell.  f[] holds the gathered (beta) or local (alpha) populations on
// entry and the values to store on exit; returns whether they are stored
// (alpha stores FLUID and VELOCITY_INJECTION cells only, beta every cell).
// SHIFT (the two-byte storages, f[] holding f_i - w_i) takes the pair form
// above.
template <typename T, bool BETA, bool SHIFT = false>
__device__ __forceinline__ bool cell_update(T *f, int flag, const DomDev *D, T &vx, T &vy, T &vz, T &dens_out) {
    if constexpr (SHIFT) return cell_update_pairs<BETA>(f, flag, D, vx, vy, vz, dens_out);
    T rho;
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
    const T rho_full = rho;
    dens_out = rho_full;
    bool write_dd = BETA;
    if (flag == kFlagFluid) {
        const T dd_param = rho - T(1.5) * vel2_of(vx, vy, vz);
        collide_fluid(f, rho, vx, vy, vz, dd_param, (T)D->inv_tau, (T)D->tau0, (T)D->les_coeff);
        if (D->forcing == 1) {
            body_force(f, rho_full, (T)D->gx, (T)D->gy, (T)D->gz);
        } else if (!BETA) {
            dens_out = alpha_ref_forcing(f, rho_full, (T)D->gx, (T)D->gy, (T)D->gz);
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
    return write_dd;
}

template <typename T>
__device__ __forceinline__ void store_macros(const DomDev *D, unsigned g, int cells, T vx, T vy, T vz, T dens) {
    typedef __attribute__((address_space(1))) T gT;
    const unsigned cs = (unsigned)sizeof(T);
    if (D->store_velocity) {
        gchar *v = (gchar *)D->vel;
        *(gT *)(v + g * cs) = vx;
        *(gT *)(v + (size_t)cells * cs + g * cs) = vy;
        *(gT *)(v + 2 * (size_t)cells * cs + g * cs) = vz;
    }
    if (D->store_density) *(gT *)((gchar *)D->dens + g * cs) = dens;
}

// One cell, element accesses.  With WRAP the neighbour index goes through the
// linear wrap W (src/cl_programs/wrap.h:112-127); without it the block is far
// enough from both ends of the domain for W to be the identity.
// A wavefront whose 64 lanes hold the consecutive cells g0 .. g0+63, none
// wrapping: each shifted beta access is then one run s .. s+63, and where the
// run lies in one layout block its address is a wavefront-uniform base plus
// the lane's byte offset, with no per-lane index arithmetic.  Used for 4- and
// 8-byte storages (kRunAddressing): it cuts beta's vector instructions by a
// third; the two-byte kernels, which need every wave of occupancy, lose more
// to its registers than they gain.
struct Run {
    bool ld, st;  // every gather (g0 - d_i) / scatter (g0 + d_i) run lies in one layout block
    unsigned g0;  // first cell of the wavefront (uniform)
    unsigned lo;  // this lane's byte offset from it (lane * element bytes)
};

__device__ __forceinline__ bool run_in_block(unsigned s) { return (s & (kBlk - 1u)) <= kBlk - 64u; }

template <class SP>
constexpr bool kRunAddressing = SP::kBytes >= 4;

// The run descriptor of a wavefront of consecutive cells g0 .. g0+63 (uniform
// checks; a block-crossing run sends the whole wavefront to per-lane addresses
// for that direction of access).
__device__ __forceinline__ Run make_run(unsigned g0, unsigned lo, const int *d) {
    bool ld = true, st = true;
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        ld = ld && run_in_block(g0 - (unsigned)d[i]);
        st = st && run_in_block(g0 + (unsigned)d[i]);
    }
    return Run{ld, st, g0, lo};
}

template <class SP, bool BETA, bool NTL>
__device__ __forceinline__ void cell_load(gchar *const *pl, unsigned g, int cells, const int *d, bool wrap,
                                          typename SP::T *f, const Run &run = Run{false, false, 0u, 0u}) {
    const unsigned own = cell_base(g) * SP::kBytes;
    if (BETA) {
        if (run.ld) {
#pragma unroll
            for (int i = 0; i < 18; ++i)
                f[i] = kld<SP>(pl[opp(i)] + (size_t)cell_base(run.g0 - (unsigned)d[i]) * SP::kBytes, run.lo, opp(i));
        } else {
#pragma unroll
            for (int i = 0; i < 18; ++i) {
                const unsigned idx = wrap ? (unsigned)wrap_idx((int)g - d[i], cells) : g - (unsigned)d[i];
                f[i] = kld<SP>(pl[opp(i)], cell_base(idx) * SP::kBytes, opp(i));
            }
        }
        f[18] = kld<SP>(pl[18], own, 18);
    } else {
        // the 19 loads do not wait for the flag
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = kld<SP, NTL>(pl[i], own, i);
    }
}

// One beta scatter store.  A wavefront's run for a population with e_x != 0
// is shifted by one element, so its first and last cache lines are shared
// with the neighbouring runs: those stores keep the default policy (the line
// stays in L2 until the other part arrives); the aligned ones may stream.
template <class SP, bool NTS>
__device__ __forceinline__ void beta_st(gchar *base, unsigned off, int i, typename SP::T v) {
    if (NTS && ex(i) == 0) kst<SP, true>(base, off, i, v);
    else kst<SP, false>(base, off, i, v);
}

// Beta scatters f[i] to slot i of W(g + d_i) (lbm_beta.cl:708-784).  Alpha
// stores every slot of its cell: f[] already holds slot k's value at k.
// Beta scatter of population i for the lane's cell g: a uniform base plus the
// lane offset when every scatter run of the wavefront lies in one block.
template <class SP, bool NTS>
__device__ __forceinline__ void beta_scatter(gchar *p, unsigned g, int cells, int di, bool wrap, const Run &run, int i,
                                             typename SP::T v) {
    if (run.st) {
        beta_st<SP, NTS>(p + (size_t)cell_base(run.g0 + (unsigned)di) * SP::kBytes, run.lo, i, v);
    } else {
        const unsigned idx = wrap ? (unsigned)wrap_idx((int)g + di, cells) : g + (unsigned)di;
        beta_st<SP, NTS>(p, cell_base(idx) * SP::kBytes, i, v);
    }
}

template <class SP, bool BETA, bool NTS>
__device__ __forceinline__ void cell_store(gchar *const *pl, unsigned g, int cells, const int *d, bool wrap,
                                           const typename SP::T *f, const Run &run = Run{false, false, 0u, 0u}) {
    const unsigned own = cell_base(g) * SP::kBytes;
    if (BETA) {
#pragma unroll
        for (int i = 0; i < 18; ++i) beta_scatter<SP, NTS>(pl[i], g, cells, d[i], wrap, run, i, f[i]);
        kst<SP, NTS>(pl[18], own, 18, f[18]);
    } else {
#pragma unroll
        for (int k = 0; k < 19; ++k) kst<SP, NTS>(pl[k], own, k, f[k]);
    }
}

// cell_store restricted to the slots (beta: populations) k whose bit in
// `mask` equals `bit`; the mask is wavefront-uniform, so whole-wavefront
// stores are kept.
template <class SP, bool BETA, bool NTS>
__device__ __forceinline__ void cell_store_part(gchar *const *pl, unsigned g, int cells, const int *d, bool wrap,
                                                const typename SP::T *f, unsigned mask, unsigned bit,
                                                const Run &run = Run{false, false, 0u, 0u}) {
    const unsigned own = cell_base(g) * SP::kBytes;
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        if (((mask >> k) & 1u) != bit) continue;
        if (BETA && k < 18) {
            beta_scatter<SP, NTS>(pl[k], g, cells, d[k], wrap, run, k, f[k]);
        } else {
            kst<SP, NTS>(pl[k], own, k, f[k]);
        }
    }
}

// Alpha's slot values: slot k takes f[opp(k)] where the cell collided, else
// its loaded value f[k] (the reference stores nothing there,
// lbm_alpha.cl:305-343; rewriting the same value keeps memory identical and
// every cache line fully written).
template <typename T>
__device__ __forceinline__ void alpha_slots(T *f, bool write_dd) {
#pragma unroll
    for (int k = 0; k < 18; k += 2) {
        const T a = f[k], b = f[k + 1];
        f[k] = write_dd ? b : a;
        f[k + 1] = write_dd ? a : b;
    }
}

template <class SP, bool BETA, bool NTS, bool NTL = false>
__device__ __forceinline__ void cell_scalar(gchar *const *pl, unsigned g, int cells, const int *d, bool wrap,
                                            const DomDev *D, const Run &run = Run{false, false, 0u, 0u}) {
    using T = typename SP::T;
    const int flag = ((const guchar *)D->flags)[g];
    T f[19];
    cell_load<SP, BETA, NTL>(pl, g, cells, d, wrap, f, run);
    T vx, vy, vz, dens;
    const bool write_dd = cell_update<T, BETA, SP::kShifted>(f, flag, D, vx, vy, vz, dens);
    if (!BETA) alpha_slots(f, write_dd);
    cell_store<SP, BETA, NTS>(pl, g, cells, d, wrap, f, run);  // beta stores every cell
    if (flag == kFlagGhost) return;  // lbm_alpha.cl:30-32, lbm_beta.cl:787-788
    store_macros<T>(D, g, cells, vx, vy, vz, dens);
}

// N cells s .. s+N-1 of population base p: one N-element vector access when
// they lie in one layout block, element accesses when they straddle two.
template <class SP, int N>
__device__ __forceinline__ void vld_any(const gchar *p, unsigned s, int i, typename SP::T *out) {
    if ((s & (kBlk - 1u)) <= kBlk - (unsigned)N) {
        vld<SP, N>(p, cell_base(s) * SP::kBytes, i, out);
    } else {
#pragma unroll
        for (int j = 0; j < N; ++j) out[j] = kld<SP>(p, cell_base(s + j) * SP::kBytes, i);
    }
}
template <class SP, int N>
__device__ __forceinline__ void vst_any(gchar *p, unsigned s, int i, const typename SP::T *in) {
    if ((s & (kBlk - 1u)) <= kBlk - (unsigned)N) {
        vst<SP, N>(p, cell_base(s) * SP::kBytes, i, in);
    } else {
#pragma unroll
        for (int j = 0; j < N; ++j) kst<SP>(p, cell_base(s + j) * SP::kBytes, i, in[j]);
    }
}

// N x-consecutive cells g0 .. g0+N-1 away from the domain ends (g0 a multiple
// of N, so they share a layout block): each of the 19 (beta: 18 shifted + 1)
// population accesses is one N-element vector access, or N element accesses
// where a shifted run straddles two blocks.
template <class SP, bool BETA, int N>
__device__ __forceinline__ void cells_vector(gchar *const *pl, unsigned g0, int cells, const int *d,
                                             const DomDev *D) {
    using T = typename SP::T;
    const unsigned es = SP::kBytes;
    int flag[N];
    {
        typedef unsigned char VF __attribute__((ext_vector_type(N)));
        VF fv;
        __builtin_memcpy(&fv, (const gchar *)D->flags + g0, sizeof(VF));
#pragma unroll
        for (int j = 0; j < N; ++j) flag[j] = fv[j];
    }
    const unsigned own = cell_base(g0) * es;
    T F[19][N];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int src = BETA ? opp(i) : i;
        if (BETA && i < 18) vld_any<SP, N>(pl[src], g0 - (unsigned)d[i], src, F[i]);
        else vld<SP, N>(pl[src], own, src, F[i]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        T f[19], vx, vy, vz, dn;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = F[i][j];
        const bool wr = cell_update<T, BETA, SP::kShifted>(f, flag[j], D, vx, vy, vz, dn);
        // slot values in place: beta scatters f[i] to slot i, alpha stores
        // f[opp(k)] in slot k (see cell_scalar); the loaded values where the
        // cell did not collide
#pragma unroll
        for (int k = 0; k < 19; ++k) F[k][j] = wr ? (BETA ? f[k] : f[opp(k)]) : F[k][j];
        // the macroscopic outputs go out at once instead of staying live
        // through the other cells' collisions
        if (flag[j] != kFlagGhost) store_macros<T>(D, g0 + j, cells, vx, vy, vz, dn);
    }
    // Store addresses from a fresh copy of the cell index: reusing the loads'
    // address arithmetic, the compiler keeps 64-bit per-lane addresses live
    // through the collisions (FP32 two cells per lane: 108 -> 79 VGPRs in beta).
    unsigned gs = g0;
    asm volatile("" : "+v"(gs));
    const unsigned own_st = cell_base(gs) * es;
    if (BETA) {
#pragma unroll
        for (int i = 0; i < 18; ++i) vst_any<SP, N>(pl[i], gs + (unsigned)d[i], i, F[i]);
        vst<SP, N>(pl[18], own_st, 18, F[18]);
    } else {
#pragma unroll
        for (int k = 0; k < 19; ++k) vst<SP, N>(pl[k], own_st, k, F[k]);
    }
}

// Cells s and s+1 of population base p as one 32-bit word (cell s in the low
// half), for the two-byte storages: one access when both lie in one layout
// block (at any 2-byte alignment: gfx950 serves unaligned dword accesses),
// two element accesses at a block's last cell.
typedef __attribute__((address_space(1))) unsigned __attribute__((aligned(2))) gpair;
__device__ __forceinline__ unsigned pair_ld(const gchar *p, unsigned s) {
    if ((s & (kBlk - 1u)) != kBlk - 1u) return *(const gpair *)(p + cell_base(s) * 2u);
    return (unsigned)raw_ld<false, unsigned short>(p, cell_base(s) * 2u) |
           ((unsigned)raw_ld<false, unsigned short>(p, cell_base(s + 1u) * 2u) << 16);
}
template <bool NT>
__device__ __forceinline__ void pair_st(gchar *p, unsigned s, unsigned v) {
    if ((s & (kBlk - 1u)) != kBlk - 1u) {
        if (NT && (s & 1u) == 0u) {
            raw_st<true, unsigned>(p, cell_base(s) * 2u, v);
        } else {
            *(gpair *)(p + cell_base(s) * 2u) = v;
        }
        return;
    }
    raw_st<NT, unsigned short>(p, cell_base(s) * 2u, (unsigned short)v);
    raw_st<NT, unsigned short>(p, cell_base(s + 1u) * 2u, (unsigned short)(v >> 16));
}

// pair_ld / pair_st through the linear wrap W where the run nears a domain end
__device__ __forceinline__ unsigned pair_ld_w(const gchar *p, int s, int cells, bool wrap) {
    if (!wrap) return pair_ld(p, (unsigned)s);
    return (unsigned)raw_ld<false, unsigned short>(p, cell_base((unsigned)wrap_idx(s, cells)) * 2u) |
           ((unsigned)raw_ld<false, unsigned short>(p, cell_base((unsigned)wrap_idx(s + 1, cells)) * 2u) << 16);
}
__device__ __forceinline__ void pair_st_w(gchar *p, int s, int cells, bool wrap, bool nt, unsigned v) {
    if (!wrap) {
        if (nt) pair_st<true>(p, (unsigned)s, v);
        else pair_st<false>(p, (unsigned)s, v);
        return;
    }
    raw_st<false, unsigned short>(p, cell_base((unsigned)wrap_idx(s, cells)) * 2u, (unsigned short)v);
    raw_st<false, unsigned short>(p, cell_base((unsigned)wrap_idx(s + 1, cells)) * 2u, (unsigned short)(v >> 16));
}

// Pair runs: a wavefront whose 64 lanes hold the packed pairs of the
// consecutive cells c0 .. c0+127 (c0 uniform).  Where a shifted run
// c0 - d_i (gather) or c0 + d_i (scatter) lies in one layout block, lane l's
// pair of it is the run's wavefront-uniform base plus 4 l bytes: the address
// costs scalar instructions only, where the per-lane form spends about six
// vector instructions per access (index, block split, straddle check), for
// each of beta's 36 shifted accesses.  The two-byte kernels issue vector
// instructions near the SIMDs' limit, so these count.
__device__ __forceinline__ bool pair_run_in_block(unsigned c0) { return (c0 & (kBlk - 1u)) <= kBlk - 128u; }

// every gather run (ld) / every scatter run (st) of the wavefront in one block
__device__ __forceinline__ void pair_runs(unsigned c0, const int *d, bool &ld, bool &st) {
    ld = st = true;
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        ld = ld && pair_run_in_block(c0 - (unsigned)d[i]);
        st = st && pair_run_in_block(c0 + (unsigned)d[i]);
    }
}

// beta's 18 gathered pairs (slot opp(i) at g - d_i) and the own slot 18, per
// lane (pair_ld) or, with `run`, from the runs' uniform bases (lo = 4 l)
__device__ __forceinline__ void pair_gather(gchar *const *pl, unsigned g, unsigned c0, unsigned lo, bool run,
                                            const int *d, int cells, bool wrap, unsigned *P) {
    if (run) {
#pragma unroll
        for (int i = 0; i < 18; ++i)
            P[i] = *(const gpair *)(pl[opp(i)] + (size_t)cell_base(c0 - (unsigned)d[i]) * 2u + lo);
    } else {
#pragma unroll
        for (int i = 0; i < 18; ++i) P[i] = pair_ld_w(pl[opp(i)], (int)g - d[i], cells, wrap);
    }
    P[18] = raw_ld<false, unsigned>(pl[18], cell_base(g) * 2u);
}

// beta's 18 scattered pairs (slot q at g + d_q) and the own slot 18; the
// non-temporal policy only for the e.x = 0 slots at 4-byte alignment
template <bool NTS>
__device__ __forceinline__ void pair_scatter(gchar *const *pl, unsigned g, unsigned c0, unsigned lo, bool run,
                                             const int *d, int cells, bool wrap, const unsigned *P) {
    if (run) {
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            const unsigned s0 = c0 + (unsigned)d[q];
            gchar *b = pl[q] + (size_t)cell_base(s0) * 2u;
            if (NTS && ex(q) == 0 && (s0 & 1u) == 0u) raw_st<true, unsigned>(b, lo, P[q]);  // uniform
            else *(gpair *)(b + lo) = P[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < 18; ++q) pair_st_w(pl[q], (int)g + d[q], cells, wrap, NTS && ex(q) == 0, P[q]);
    }
    raw_st<NTS, unsigned>(pl[18], cell_base(g) * 2u, P[18]);
}

// Two x-consecutive cells g0, g0+1 per work-item in a two-byte storage (g0
// even, away from the domain ends): every population access moves one 32-bit
// word holding both cells, and the words stay packed in registers.  The two
// cells collide one after the other, each unpacking its 19 halves and packing
// its results back in place, so only one cell's values are live as floats
// (cells_vector keeps all 2 x 19).  Same conversions, arithmetic and stored
// values as two calls of cell_scalar.
template <class SP, bool BETA, bool NTS, bool NTL>
__device__ __forceinline__ void cells_pair2(gchar *const *pl, unsigned g0, int cells, const int *d, const DomDev *D) {
    static_assert(SP::kBytes == 2, "packed pairs hold two-byte storages");
    using T = typename SP::T;
    const unsigned fl = raw_ld<false, unsigned short>((const gchar *)D->flags, g0);
    unsigned P[19];
    // beta: pair runs when all 64 lanes are here (consecutive pairs from the
    // lane-0 cell c0; lanes at the domain end take another path)
    unsigned c0 = g0, lo = 0u;
    bool rl = false, rs = false;
    if (BETA) {
        c0 = __builtin_amdgcn_readfirstlane(g0);
        lo = (g0 - c0) * 2u;
        if (__ballot(1) == ~0ull) pair_runs(c0, d, rl, rs);
        pair_gather(pl, g0, c0, lo, rl, d, cells, false, P);
    } else {
        const unsigned own = cell_base(g0) * 2u;
#pragma unroll
        for (int i = 0; i < 19; ++i) P[i] = raw_ld<NTL, unsigned>(pl[i], own);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int flag = (int)((fl >> (8 * j)) & 0xffu);
        T f[19], vx, vy, vz, dn;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = kin<SP>((typename SP::U)(P[i] >> (16 * j)), i);
        const bool wr = cell_update<T, BETA, SP::kShifted>(f, flag, D, vx, vy, vz, dn);
        if (wr) {
            // beta: slot i of the scatter gets f[i]; alpha: slot k of the cell gets f[opp(k)]
#pragma unroll
            for (int k = 0; k < 19; ++k) {
                const unsigned h = kout<SP>(BETA ? f[k] : f[opp(k)], k);
                P[k] = j ? ((P[k] & 0xffffu) | (h << 16)) : ((P[k] & 0xffff0000u) | h);
            }
        }
        // the macroscopic outputs go out at once: held for the end, they stay
        // live through the other cell's collision
        if (flag != kFlagGhost) store_macros<T>(D, g0 + j, cells, vx, vy, vz, dn);
    }
    // The store addresses start from a fresh copy of the cell index: reusing
    // the loads' address arithmetic, the compiler keeps 19 64-bit per-lane
    // addresses live through both collisions (~40 VGPRs).
    unsigned gs = g0;
    asm volatile("" : "+v"(gs));
    if (BETA) {
        pair_scatter<NTS>(pl, gs, c0, lo, rs, d, cells, false, P);
    } else {
        const unsigned own_st = cell_base(gs) * 2u;
#pragma unroll
        for (int k = 0; k < 19; ++k) raw_st<NTS, unsigned>(pl[k], own_st, P[k]);
    }
}

// An interior tile of a run-tiled chain: the `pass`-th share of the interior
// runs of the rows of row tile `rt` (columns 64 .. X-3 of every member, one
// 64-cell run per wavefront).  No column of a joined face: no LDS, no
// barrier.
template <class SP, bool BETA, bool NTS, bool NTL>
__device__ __forceinline__ void chain_interior(const DomDev *__restrict__ doms, const XChain *__restrict__ C, int rt,
                                               int pass, long long h) {
    const int X = C->nx, inner = C->ninner, per_row = C->nmem * inner;
    const int u = pass * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (u >= per_row * C->rows) return;  // uniform per wavefront
    const int r = u / per_row;
    const int m = __builtin_amdgcn_readfirstlane((u - r * per_row) / inner);
    const DomDev *D = doms + C->mem[m];
    if (h >= D->stop_at) return;
    const int cells = (int)D->cells;
    const int gs = (rt * C->rows + r) * X + 64 * (1 + u - r * per_row - m * inner);
    if (gs >= cells) return;  // rows past the last of the domain
    gchar *pl[19];
    {
        gchar *base = (gchar *)D->dd;
#pragma unroll
        for (int i = 0; i < 19; ++i) pl[i] = base + (size_t)i * kBlk * SP::kBytes;
    }
    int d[19];
    bool wrap = false;
    if (BETA) {
        const int XY = X * D->ny;
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = ex(i) + X * ey(i) + XY * ez(i);
        const int dmax = 1 + X + XY;
        wrap = gs < dmax || gs + 64 + dmax > cells;
    } else {
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = 0;
    }
    Run run{false, false, 0u, 0u};
    if (BETA && kRunAddressing<SP> && !wrap) run = make_run((unsigned)gs, (threadIdx.x & 63u) * SP::kBytes, d);
    cell_scalar<SP, BETA, NTS, NTL>(pl, (unsigned)(gs + (int)(threadIdx.x & 63)), cells, d, wrap, D, run);
}

// One chain tile.  Senders put their face values in LDS; after the barrier a
// ghost cell of a joined face takes them in place of its own:
//   alpha: the slots its receiver mask names (the sender's slot values after
//          alpha, src/CController.hpp:265-320);
//   beta:  the populations it scatters into the inner layer, n.e > 0 (the
//          sender's values scattered into its own ghost layer,
//          src/CController.hpp:322-383).
// A frozen member stores nothing; as a sender it feeds its slots as they
// stand, which is what the copy exchange reads from it.  Two shapes:
//   whole rows (nedge == 0): R rows of every member, `lanes` work-items per
//     member (position m, row r, column x per work-item);
//   R rows in 64-cell runs (nedge > 0): the runs holding columns 0, 1, X-2
//     and X-1 of every member and row, with the handoff; the interior runs of
//     those rows follow in `ipasses` tiles of their own (chain_interior).
// Each wavefront holds one member, so the member and every address base
// derived from it stay uniform per wavefront.
template <class SP, bool BETA, bool NTS, bool NTL>
__device__ __forceinline__ void chain_tile(const DomDev *__restrict__ doms, const XChain *__restrict__ C, int rt,
                                           long long h) {
    using T = typename SP::T;
    __shared__ T xv[kChainRows * 2 * 19];
    const int X = C->nx, R = C->rows, span = R * X;
    const int q = (int)threadIdx.x, w = q >> 6;
    int mm, r, x;
    if (C->nedge == 0) {
        mm = q / C->lanes;
        r = (q - mm * C->lanes) / X;
        x = q - mm * C->lanes - r * X;
    } else {
        const int per_row = C->nmem * C->nedge;  // wavefronts of one row's face runs
        r = w / per_row;
        mm = (w - r * per_row) / C->nedge;
        x = 64 * C->edge[w - r * per_row - mm * C->nedge] + (q & 63);
    }
    const int m = __builtin_amdgcn_readfirstlane(mm);
    const bool lane = m < C->nmem && r < R && x < X;
    const DomDev *D = doms + C->mem[m < C->nmem ? m : 0];
    const int cells = (int)D->cells;
    const int gfirst = rt * span;  // first cell of the tile's rows (equal dims along the chain)
    const int g = gfirst + r * X + x;
    const bool in = lane && g < cells;
    const bool live = in && h < D->stop_at;
    T f[19];
    int flag = kFlagGhost;
    T vx = T(0), vy = T(0), vz = T(0), dens = T(0);
    gchar *pl[19];
    {
        gchar *base = (gchar *)D->dd;
#pragma unroll
        for (int i = 0; i < 19; ++i) pl[i] = base + (size_t)i * kBlk * SP::kBytes;
    }
    int d[19];
    bool wrap = false;
    const int dmax = 1 + X + X * D->ny;
    if (BETA) {
        const int XY = X * D->ny;
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = ex(i) + X * ey(i) + XY * ez(i);
        wrap = gfirst < dmax || gfirst + span + dmax > cells;
    } else {
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = 0;
    }
    // senders: x = X-2 feeds the x = 0 ghost layer of the right partner,
    // x = 1 the x = X-1 ghost layer of the left one
    const int to = !in ? -1 : x == X - 2 ? C->right[m] : (x == 1 ? C->left[m] : -1);
    const int sside = x == X - 2 ? 0 : 1;
    T *so = xv + (((to < 0 ? 0 : to) * R + r) * 2 + sside) * 19;
    // alpha: the slots a ghost cell of this wavefront takes from LDS wait for
    // the barrier, the wavefront stores the rest before it (uniform per
    // wavefront).  Beta stores everything after it: its scatter addresses
    // would stay live across the barrier.
    const int from = !live ? -1 : x == 0 ? C->left[m] : (x == X - 1 ? C->right[m] : -1);
    const bool side0 = __ballot(from >= 0 && x == 0) != 0, side1 = __ballot(from >= 0 && x == X - 1) != 0;
    const unsigned late = BETA ? (1u << 19) - 1u : (side0 ? D->xmask_a[0] : 0u) | (side1 ? D->xmask_a[1] : 0u);
    // every lane of the wavefront live: the cells are consecutive (one member's
    // rows, or one 64-cell run)
    Run run{false, false, 0u, 0u};
    if (BETA && kRunAddressing<SP> && !wrap && __ballot(live) == ~0ull) {
        const unsigned w0 = __builtin_amdgcn_readfirstlane((unsigned)g);
        run = make_run(w0, ((unsigned)g - w0) * SP::kBytes, d);
    }
    if (live) {
        flag = ((const guchar *)D->flags)[g];
        cell_load<SP, BETA, NTL>(pl, (unsigned)g, cells, d, wrap, f, run);
        const bool write_dd = cell_update<T, BETA, SP::kShifted>(f, flag, D, vx, vy, vz, dens);
        if (!BETA) alpha_slots(f, write_dd);
        if (to >= 0) {
#pragma unroll
            for (int k = 0; k < 19; ++k)
                if (!BETA || ex(k) == (sside ? -1 : 1)) so[k] = f[k];
        }
        cell_store_part<SP, BETA, NTS>(pl, (unsigned)g, cells, d, wrap, f, late, 0u, run);
    } else if (to >= 0) {
        // frozen sender: alpha's slot values are its own slots, beta's are the
        // slots its scatter would write
#pragma unroll
        for (int k = 0; k < 19; ++k) {
            if (!BETA) {
                so[k] = kld<SP>(pl[k], cell_base((unsigned)g) * SP::kBytes, k);
            } else if (ex(k) == (sside ? -1 : 1)) {
                const unsigned idx = wrap ? (unsigned)wrap_idx(g + d[k], cells) : (unsigned)(g + d[k]);
                so[k] = kld<SP>(pl[k], cell_base(idx) * SP::kBytes, k);
            }
        }
    }
    __syncthreads();
    if (live) {
        if (from >= 0) {
            const int side = x == 0 ? 0 : 1;
            const T *o = xv + ((m * R + r) * 2 + side) * 19;
            const unsigned mask = D->xmask_a[side];
#pragma unroll
            for (int k = 0; k < 19; ++k) {
                const bool take = BETA ? ex(k) == (side ? -1 : 1) : ((mask >> k) & 1u) != 0;
                if (take) f[k] = o[k];
            }
        }
        cell_store_part<SP, BETA, NTS>(pl, (unsigned)g, cells, d, wrap, f, late, 1u, run);
        if (flag != kFlagGhost) store_macros<T>(D, (unsigned)g, cells, vx, vy, vz, dens);
    }
}

// ---- periodic x seam of one domain, packed pairs ---------------------------
// A two-byte domain whose x faces are each other's neighbours (a periodic
// ring of one), stepped with two cells per work-item: every row is split into
// `seam_wpr` wavefronts, the first holding pairs 0..31 and the row's last 32
// pairs (cells 0..63 and X-64..X-1), the others the pairs in between.  The
// x = 0 ghost cell (lane 0, low half) and its sender x = X-2 (lane 63, low
// half), and the x = X-1 ghost (lane 63, high half) and its sender x = 1
// (lane 0, high half), are then in one wavefront: the face values move by
// lane reads, with no LDS, no barrier and no copy launch.  The ghost cells
// take what the copy exchange would give them: after alpha the slots of
// their receiver masks, in beta the populations they scatter into the inner
// layer (e.x = +1 at x = 0, e.x = -1 at x = X-1).
template <class SP, bool BETA, bool NTS, bool NTL>
__device__ __forceinline__ void seam_pairs(const DomDev *__restrict__ D, int tile) {
    using T = typename SP::T;
    const int X = D->nx, NP = X >> 1, wpr = D->seam_wpr;
    const int cells = (int)D->cells;
    const int gw = tile * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);  // wavefront of the domain
    const int row = gw / wpr, k = gw - row * wpr;
    if (row >= D->ny * D->nz) return;  // uniform per wavefront
    const int l = (int)(threadIdx.x & 63);
    const int p = k == 0 ? (l < 32 ? l : NP - 64 + l) : 32 + 64 * (k - 1) + l;
    if (k != 0 && p >= NP - 32) return;  // past the row's interior pairs (never in the first wavefront)
    const int g = row * X + 2 * p;
    gchar *pl[19];
    {
        gchar *base = (gchar *)D->dd;
#pragma unroll
        for (int i = 0; i < 19; ++i) pl[i] = base + (size_t)i * kBlk * SP::kBytes;
    }
    int d[19];
    bool wrap = false;
    if (BETA) {
        const int XY = X * D->ny;
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = ex(i) + X * ey(i) + XY * ez(i);
        const int dmax = 1 + X + XY;
        wrap = row * X < dmax || row * X + X + dmax > cells;
    } else {
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = 0;
    }
    // the pair's loads and collisions (as cells_pair2, with the linear wrap per half)
    const unsigned fl = raw_ld<false, unsigned short>((const gchar *)D->flags, (unsigned)g);
    unsigned P[19];
    // beta: pair runs for the wavefronts of consecutive pairs (not the split
    // first one), all 64 lanes present, away from the domain ends
    unsigned c0 = (unsigned)g, lo = 0u;
    bool rl = false, rs = false;
    if (BETA) {
        c0 = __builtin_amdgcn_readfirstlane((unsigned)g);
        lo = ((unsigned)g - c0) * 2u;
        if (k != 0 && !wrap && __ballot(1) == ~0ull) pair_runs(c0, d, rl, rs);
        pair_gather(pl, (unsigned)g, c0, lo, rl, d, cells, wrap, P);
    } else {
        const unsigned own = cell_base((unsigned)g) * 2u;
#pragma unroll
        for (int i = 0; i < 19; ++i) P[i] = raw_ld<NTL, unsigned>(pl[i], own);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int flag = (int)((fl >> (8 * j)) & 0xffu);
        T f[19], vx, vy, vz, dn;
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = kin<SP>((typename SP::U)(P[i] >> (16 * j)), i);
        const bool wr = cell_update<T, BETA, SP::kShifted>(f, flag, D, vx, vy, vz, dn);
        if (wr) {
#pragma unroll
            for (int q = 0; q < 19; ++q) {
                const unsigned hv = kout<SP>(BETA ? f[q] : f[opp(q)], q);
                P[q] = j ? ((P[q] & 0xffffu) | (hv << 16)) : ((P[q] & 0xffff0000u) | hv);
            }
        }
        if (flag != kFlagGhost) store_macros<T>(D, (unsigned)(g + j), cells, vx, vy, vz, dn);
    }
    if (k == 0) {  // the seam: lane 0 holds cells 0, 1 and lane 63 cells X-2, X-1
        const unsigned m0 = D->xmask_a[0], m1 = D->xmask_a[1];
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const unsigned last = __builtin_amdgcn_readlane(P[q], 63), first = __builtin_amdgcn_readlane(P[q], 0);
            const bool take0 = BETA ? ex(q) == 1 : ((m0 >> q) & 1u) != 0;
            const bool take1 = BETA ? ex(q) == -1 : ((m1 >> q) & 1u) != 0;
            if (l == 0 && take0) P[q] = (P[q] & 0xffff0000u) | (last & 0xffffu);
            if (l == 63 && take1) P[q] = (P[q] & 0xffffu) | (first & 0xffff0000u);
        }
    }
    // store addresses from a fresh copy of the cell index (see cells_pair2)
    int gs = g;
    asm volatile("" : "+v"(gs));
    if (BETA) {
        pair_scatter<NTS>(pl, (unsigned)gs, c0, lo, rs, d, cells, wrap, P);
    } else {
        const unsigned own_st = cell_base((unsigned)gs) * 2u;
#pragma unroll
        for (int q = 0; q < 19; ++q) raw_st<NTS, unsigned>(pl[q], own_st, P[q]);
    }
}

template <int ST, bool BETA, int VAR = 0>
__global__ void __launch_bounds__(512) step_kernel(const DomDev *__restrict__ doms, const int *__restrict__ tile_dom,
                                                    const XChain *__restrict__ chains, int xcd_map, long long *ctr) {
    using SP = Storage<ST>;
    constexpr bool NTS = (VAR & kVarNtStore) != 0;
    constexpr bool NTL = (VAR & kVarNtLoad) != 0;
    constexpr int N = cells_per_lane(VAR);
    // Half-step counter of a group launch: beta reads slot 0 and publishes
    // h + 1 in slot 1 for the alpha launch after it, which does the reverse.
    // Readers and the single writer of one launch use different slots, and
    // the next launch on the stream sees the write.
    long long h = 0;
    if (ctr) {
        h = ctr[BETA ? 0 : 1];
        if (blockIdx.x == 0 && threadIdx.x == 0) ctr[BETA ? 1 : 0] = h + 1;
    }
    const int nb = (int)gridDim.x;
    // xcd_map: < 0 one contiguous eighth per XCD, 0 none, K > 0 chunks of K
    const int t = xcd_map < 0   ? xcd_tile((int)blockIdx.x, nb)
                  : xcd_map > 0 ? xcd_chunk_tile((int)blockIdx.x, nb, xcd_map)
                                : (int)blockIdx.x;
    const int td = tile_dom[t];
    if (td < 0) {  // a tile of fused x chain -td-1 (uniform per work-group: the barrier is reached by all)
        // per row tile: the handoff tile, then the interior tiles of a one-row chain
        const XChain *C = chains + (-td - 1);
        const int per = 1 + C->ipasses, k = t - C->tile_begin;
        const int rt = k / per, part = k - rt * per;
        if (part == 0) chain_tile<SP, BETA, NTS, NTL>(doms, C, rt, h);
        else chain_interior<SP, BETA, NTS, NTL>(doms, C, rt, part - 1, h);
        return;
    }
    const DomDev *D = doms + td;
    if (h >= D->stop_at) return;  // a finished member of a continuous batch keeps its state
    if constexpr (N == 2 && SP::kBytes == 2) {
        if (D->seam_wpr > 0) {  // uniform per work-group
            seam_pairs<SP, BETA, NTS, NTL>(D, t - D->tile_begin);
            return;
        }
    }
    const int cells = (int)D->cells;
    const int block_cells = (int)blockDim.x * N;
    const int block_first = (t - D->tile_begin) * block_cells;
    const int g0 = block_first + (int)threadIdx.x * N;
    if (g0 >= cells) return;
    gchar *pl[19];
    {
        gchar *base = (gchar *)D->dd;
#pragma unroll
        for (int i = 0; i < 19; ++i) pl[i] = base + (size_t)i * kBlk * SP::kBytes;
    }
    int d[19];
    bool wrap = false;
    if (BETA) {
        const int X = D->nx, XY = D->nx * D->ny;
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = ex(i) + X * ey(i) + XY * ez(i);
        const int dmax = 1 + X + XY;
        wrap = block_first < dmax || block_first + block_cells + dmax > cells;
    } else {
#pragma unroll
        for (int i = 0; i < 19; ++i) d[i] = 0;
    }
    if (N > 1 && !wrap && g0 + N <= cells) {
        if constexpr (N == 2 && SP::kBytes == 2) cells_pair2<SP, BETA, NTS, NTL>(pl, (unsigned)g0, cells, d, D);
        else cells_vector<SP, BETA, N>(pl, (unsigned)g0, cells, d, D);
        return;
    }
    Run run{false, false, 0u, 0u};
    if (BETA && kRunAddressing<SP> && N == 1 && !wrap) {
        // lane 0 holds the wavefront's first cell (lanes past the end have returned)
        const unsigned w0 = __builtin_amdgcn_readfirstlane((unsigned)g0);
        if (w0 + 63u < (unsigned)cells) run = make_run(w0, ((unsigned)g0 - w0) * SP::kBytes, d);
    }
    // one cell at a time: interleaving N copies of the scalar update would set
    // the kernel's register count for a path only the domain ends take
#pragma unroll 1
    for (int j = 0; j < N; ++j)
        if (g0 + j < cells) cell_scalar<SP, BETA, NTS, NTL>(pl, (unsigned)(g0 + j), cells, d, wrap, D, run);
}

// Launch of step_kernel for storage ST: the half-step and the variant bits
// become template arguments; bit 2 of the variant turns the XCD tile map off
// and bit 5 deals the tiles to the XCDs in chunks of `xcd_chunk` tiles
// instead of eighths; `lds` reserves LDS per work-group to cap residency.  The 4-
// and 8-byte storages are instantiated in lbm_runtime.hip, the two-byte ones
// in lbm_step_2byte.hip (a translation unit of their own: see there).
template <int ST>
hipError_t launch_step_storage(bool beta, int variant, int lds, int tiles, int wg, const DomDev *descs, const int *tile_table,
                       hipStream_t stream, long long *ctr, const XChain *chains, int xcd_chunk = kXcdChunkTiles) {
    const int xcd = (variant & kVarNoXcdMap) ? 0 : (variant & kVarXcdChunks) ? xcd_chunk : -1;
    auto go = [&](auto b, auto vv) {
        constexpr bool B = decltype(b)::value;
        constexpr int V = decltype(vv)::value;
        step_kernel<ST, B, V><<<tiles, wg, lds, stream>>>(descs, tile_table, chains, xcd, ctr);
    };
    auto pick = [&](auto b) {
        switch (variant & ~(kVarNoXcdMap | kVarXcdChunks)) {
            case 0: go(b, std::integral_constant<int, 0>()); break;
            case 1: go(b, std::integral_constant<int, 1>()); break;
            case 2: go(b, std::integral_constant<int, 2>()); break;
            case 3: go(b, std::integral_constant<int, 3>()); break;
            case 8: go(b, std::integral_constant<int, 8>()); break;
            case 11: go(b, std::integral_constant<int, 11>()); break;
            case 16: go(b, std::integral_constant<int, 16>()); break;
            default: go(b, std::integral_constant<int, 0>()); break;
        }
    };
    if (beta) pick(std::true_type());
    else pick(std::false_type());
    return hipGetLastError();
}

// launch_step_storage for the FP16 / BF16 storages (2, 3), defined in lbm_step_2byte.hip
hipError_t launch_step_2byte(int storage, bool beta, int variant, int lds, int tiles, int wg, const DomDev *descs,
                             const int *tile_table, hipStream_t stream, long long *ctr, const XChain *chains,
                             int xcd_chunk);

}  // namespace lbm
