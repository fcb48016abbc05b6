// lbm_runtime.hip -- host runtime and auxiliary kernels behind include/lbm_capi.h.
//
// Objects:
//   lbm_solver  one (sub)domain: CLbmSolver<T> (src/CLbmSolver.hpp:67-1124)
//   lbm_group   N solvers stepped by one fused kernel per half-step plus the
//               ghost-layer exchange of CController::syncAlpha/syncBeta
//               (src/CController.hpp:265-391) as on-device copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_capi.h"
#include "exchange_plan.hpp"
#include "launch_shape.hpp"
#include "lbm_device.hpp"

using namespace lbm;

namespace {

thread_local std::string g_last_error;

int record(std::string *slot, int code, const std::string &msg) {
    if (slot) *slot = msg;
    g_last_error = msg;
    return code;
}

struct Box {
    int o[3];
    int s[3];
};

// ---- auxiliary kernels ------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unit01(uint64_t seed, uint64_t case_id, uint64_t key, int comp) {
    uint64_t h = mix64(seed ^ mix64(case_id ^ mix64(key * 4ull + (uint64_t)comp)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// base of population i (global address space); cell g sits cell_base(g)
// elements further
template <class SP>
__device__ __forceinline__ gchar *pop_ptr(const DomDev &D, int i) {
    return (gchar *)D.dd + (size_t)i * kBlk * SP::kBytes;
}
template <class SP>
__device__ __forceinline__ unsigned cell_off(long long g) {
    return cell_base((unsigned)g) * SP::kBytes;
}

struct InitArgs {
    int bc[6];
    double amp;
    uint64_t seed, case_id;
    int key_origin[3];
    int key_dims[3];
};

// init_kernel (src/cl_programs/lbm_init.cl:32-237) + seeded perturbation.
template <int ST>
__global__ void init_kernel(DomDev D, InitArgs a) {
    using SP = Storage<ST>;
    using T = typename SP::T;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= D.cells) return;
    const int x = (int)(g % D.nx), y = (int)((g / D.nx) % D.ny), z = (int)(g / ((long long)D.nx * D.ny));
    int flag = kFlagFluid;
    if (x == 0) flag = a.bc[0];
    else if (x == D.nx - 1) flag = a.bc[1];
    else if (y == 0) flag = a.bc[2];
    else if (y == D.ny - 1) flag = a.bc[3];
    else if (z == 0) flag = a.bc[4];
    else if (z == D.nz - 1) flag = a.bc[5];
    D.flags[g] = (uint8_t)flag;
    const bool outer = (x == 0 && a.bc[0] == kFlagObstacle) || (x == D.nx - 1 && a.bc[1] == kFlagObstacle) ||
                       (y == 0 && a.bc[2] == kFlagObstacle) || (y == D.ny - 1 && a.bc[3] == kFlagObstacle) ||
                       (z == 0 && a.bc[4] == kFlagObstacle) || (z == D.nz - 1 && a.bc[5] == kFlagObstacle);
    T u[3] = {T(0), T(0), T(0)};
    if (a.amp != 0.0 && !outer) {
        const long long kx = a.key_origin[0] + x, ky = a.key_origin[1] + y, kz = a.key_origin[2] + z;
        const uint64_t key = (uint64_t)(kx + (long long)a.key_dims[0] * (ky + (long long)a.key_dims[1] * kz));
        for (int c = 0; c < 3; ++c) u[c] = (T)(a.amp * (2.0 * unit01(a.seed, a.case_id, key, c) - 1.0));
    }
    const T one = T(1);
    const T dd_param = one - T(1.5) * vel2_of(u[0], u[1], u[2]);
    T f[19];
    feq19(dd_param, u[0], u[1], u[2], f);
#pragma unroll
    for (int i = 0; i < 19; ++i) st<SP>(pop_ptr<SP>(D, opp(i)), cell_off<SP>(g), opp(i), f[i]);
    if (D.vel) {
        T *v = (T *)D.vel;
        v[g] = u[0];
        v[D.cells + g] = u[1];
        v[2 * D.cells + g] = u[2];
    }
    if (D.dens) ((T *)D.dens)[g] = one;
}

__device__ __forceinline__ long long box_cell(const DomDev &D, const Box &b, long long e) {
    const int bx = (int)(e % b.s[0]);
    const int by = (int)((e / b.s[0]) % b.s[1]);
    const int bz = (int)(e / ((long long)b.s[0] * b.s[1]));
    return (long long)(b.o[0] + bx) + (long long)D.nx * ((b.o[1] + by) + (long long)D.ny * (b.o[2] + bz));
}

// Region -> dense [component][box] buffer (copy_buffer_rect,
// src/cl_programs/copy_buffer_rect.cl:12-52, all components in one launch).
template <int ST>
__global__ void pack_kernel(DomDev D, int field, Box b, void *out) {
    using SP = Storage<ST>;
    using T = typename SP::T;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const long long g = box_cell(D, b, e);
    if (field == LBM_FIELD_DD_RAW) {  // stored bit patterns, no conversion
        using U = typename SP::U;
        U *o = (U *)out;
#pragma unroll
        for (int i = 0; i < 19; ++i) o[i * n + e] = raw_ld<false, U>(pop_ptr<SP>(D, i), cell_off<SP>(g));
    } else if (field == LBM_FIELD_DD) {
        T *o = (T *)out;
#pragma unroll
        for (int i = 0; i < 19; ++i) o[i * n + e] = ld<SP>(pop_ptr<SP>(D, i), cell_off<SP>(g), i);
    } else if (field == LBM_FIELD_VELOCITY) {
        const T *v = (const T *)D.vel;
        T *o = (T *)out;
        for (int c = 0; c < 3; ++c) o[c * n + e] = v[c * D.cells + g];
    } else if (field == LBM_FIELD_DENSITY) {
        ((T *)out)[e] = ((const T *)D.dens)[g];
    } else {
        ((int32_t *)out)[e] = (int32_t)D.flags[g];
    }
}

// Dense buffer -> region; for distributions only populations in `mask`.
template <int ST>
__global__ void unpack_kernel(DomDev D, int field, Box b, const void *in, uint32_t mask) {
    using SP = Storage<ST>;
    using T = typename SP::T;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const long long g = box_cell(D, b, e);
    if (field == LBM_FIELD_DD_RAW) {
        using U = typename SP::U;
        const U *s = (const U *)in;
#pragma unroll
        for (int i = 0; i < 19; ++i)
            if (mask & (1u << i)) raw_st<false, U>(pop_ptr<SP>(D, i), cell_off<SP>(g), s[i * n + e]);
    } else if (field == LBM_FIELD_DD) {
        const T *s = (const T *)in;
#pragma unroll
        for (int i = 0; i < 19; ++i)
            if (mask & (1u << i)) st<SP>(pop_ptr<SP>(D, i), cell_off<SP>(g), i, s[i * n + e]);
    } else if (field == LBM_FIELD_VELOCITY) {
        T *v = (T *)D.vel;
        const T *s = (const T *)in;
        for (int c = 0; c < 3; ++c) v[c * D.cells + g] = s[c * n + e];
    } else if (field == LBM_FIELD_DENSITY) {
        ((T *)D.dens)[g] = ((const T *)in)[e];
    } else {
        D.flags[g] = (uint8_t)((const int32_t *)in)[e];
    }
}

// One ghost-layer transfer: a box of raw distribution elements from one
// subdomain into another (populations listed in `pops`).  One work-item moves
// one element: consecutive work-items walk x, then y, then z of the box for
// one population, so y- and z-face transfers are contiguous runs.
struct CopyOp {
    const void *src;
    void *dst;
    int src_nx, src_ny, dst_nx, dst_ny;
    int so[3], dof[3], sz[3];
    uint32_t mask;
    int npop;
    unsigned char pops[20];
    int block_begin;
    int cells;
    int vec;            // x-consecutive elements per work-item: 4 where both x runs are 4-aligned, else 1
    long long stop_at;  // half-step index from which the receiving domain is frozen
};

// ctr: the group's half-step counter slot that the kernel just run has
// written (h + 1), or null outside a counted group step.
// ops: nops copies, followed (when nops > 1) by the op index of every block.
template <unsigned B>
__global__ void __launch_bounds__(256) exchange_kernel(const CopyOp *__restrict__ ops, int nops,
                                                       const long long *ctr) {
    const int k = nops == 1 ? 0 : ((const int *)(ops + nops))[blockIdx.x];
    const CopyOp &op = ops[k];
    if (ctr && *ctr - 1 >= op.stop_at) return;
    const int t = ((int)blockIdx.x - op.block_begin) * (int)blockDim.x + (int)threadIdx.x;
    const int per = op.cells / op.vec;  // work-items per population
    if (t >= per * op.npop) return;
    const int p = t / per;
    const int e = (t - p * per) * op.vec;
    const int bx = e % op.sz[0];
    const int r = e / op.sz[0];
    const int by = r % op.sz[1];
    const int bz = r / op.sz[1];
    const unsigned gs = (unsigned)((op.so[0] + bx) + op.src_nx * ((op.so[1] + by) + op.src_ny * (op.so[2] + bz)));
    const unsigned gd = (unsigned)((op.dof[0] + bx) + op.dst_nx * ((op.dof[1] + by) + op.dst_ny * (op.dof[2] + bz)));
    const int i = op.pops[p];
    const gchar *sp = (const gchar *)op.src + (size_t)i * kBlk * B + cell_base(gs) * B;
    gchar *dp = (gchar *)op.dst + (size_t)i * kBlk * B + cell_base(gd) * B;
    if (op.vec == 4) {  // gs, gd multiples of 4: the four cells are contiguous in one layout block
        typedef unsigned V __attribute__((ext_vector_type(B)));  // 4 elements of B bytes
        *(__attribute__((address_space(1))) V *)dp = *(const __attribute__((address_space(1))) V *)sp;
        return;
    }
    if (B == 8)
        *(__attribute__((address_space(1))) unsigned long long *)dp =
            *(const __attribute__((address_space(1))) unsigned long long *)sp;
    else if (B == 4)
        *(__attribute__((address_space(1))) unsigned *)dp = *(const __attribute__((address_space(1))) unsigned *)sp;
    else
        *(gushort *)dp = *(const gushort *)sp;
}

// rho and j = sum f_i e_i read in the current step parity: after an even
// number of steps slot opp(i) holds f_i, after an odd number slot i does.
template <int ST>
__device__ __forceinline__ void cell_moments(const DomDev &D, long long g, bool swapped, double &r, double &a,
                                             double &bb, double &c) {
    using SP = Storage<ST>;
    r = a = bb = c = 0.0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int slot = swapped ? opp(i) : i;
        const double v = (double)ld<SP>(pop_ptr<SP>(D, slot), cell_off<SP>(g), slot);
        r += v;
        a += v * ex(i);
        bb += v * ey(i);
        c += v * ez(i);
    }
}

template <int ST>
__global__ void moments_kernel(DomDev D, Box b, int swapped, void *rho_out, void *j_out) {
    using T = typename Storage<ST>::T;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const long long g = box_cell(D, b, e);
    double r, a, bb, c;
    cell_moments<ST>(D, g, swapped != 0, r, a, bb, c);
    ((T *)rho_out)[e] = (T)r;
    T *j = (T *)j_out;
    j[e] = (T)a;
    j[n + e] = (T)bb;
    j[2 * n + e] = (T)c;
}

// FP64 mass and kinetic energy over FLUID cells, first pass: one partial
// pair per block (wavefront shuffles, then LDS across the block's waves).
template <int ST>
__global__ void __launch_bounds__(256) mass_energy_kernel(DomDev D, int swapped, double *partials) {
    __shared__ double sm[2][4];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0, k = 0.0;
    if (g < D.cells && D.flags[g] == kFlagFluid) {
        double r, a, bb, c;
        cell_moments<ST>(D, g, swapped != 0, r, a, bb, c);
        m = r;
        k = 0.5 * (a * a + bb * bb + c * c) / r;
    }
    for (int off = 32; off > 0; off >>= 1) {
        m += __shfl_down(m, off, 64);
        k += __shfl_down(k, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sm[0][wave] = m;
        sm[1][wave] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0, kk = 0.0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) {
            mm += sm[0][w];
            kk += sm[1][w];
        }
        partials[2 * blockIdx.x] = mm;
        partials[2 * blockIdx.x + 1] = kk;
    }
}

__global__ void fill_int_kernel(int *p, int n, int v) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

// Second pass: one block sums the partials in a fixed order (deterministic).
__global__ void __launch_bounds__(256) sum_pairs_kernel(const double *partials, int n, double *out) {
    __shared__ double sm[2][256];
    double m = 0.0, k = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        m += partials[2 * i];
        k += partials[2 * i + 1];
    }
    sm[0][threadIdx.x] = m;
    sm[1][threadIdx.x] = k;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0][0];
        out[1] = sm[1][0];
    }
}

size_t elem_size(int storage) {
    switch (storage) {
        case LBM_STORAGE_F32: return 4;
        case LBM_STORAGE_F64: return 8;
        default: return 2;
    }
}
size_t compute_size(int storage) { return storage == LBM_STORAGE_F64 ? 8 : 4; }

uint32_t norm_mask(const int *norm) {
    if (!norm) return (1u << 19) - 1;
    uint32_t m = 0;
    for (int i = 0; i < 19; ++i)
        if (norm[0] * ex(i) + norm[1] * ey(i) + norm[2] * ez(i) > 0) m |= 1u << i;
    return m;
}

}  // namespace

// ---- objects ----------------------------------------------------------------

// One device allocation holding the buffers of several solvers
// (lbm_solver_colocate); the last of them to be destroyed frees it.
struct SharedBlock {
    void *base = nullptr;
    int refs = 0;
};

struct lbm_solver {
    lbm_solver_config cfg{};
    long long cells = 0, padded = 0;  // padded: cells rounded up to whole layout blocks
    int wg = 256;
    void *dd = nullptr;
    uint8_t *flags = nullptr;
    void *vel = nullptr;
    void *dens = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    long long step_counter = 0;
    std::string err;
    DomDev host_desc{};
    DomDev *d_desc = nullptr;  // single-domain launch descriptor
    int *d_tile = nullptr;
    int ntiles = 0;
    void *staging = nullptr;
    size_t staging_bytes = 0;
    double *d_partials = nullptr;
    size_t partial_count = 0;
    unsigned version = 0;      // bumps when kernel parameters change
    lbm_group *group = nullptr;
    bool external = false;     // buffers live in caller-provided device memory
    SharedBlock *block = nullptr;  // buffers live in a block shared with other solvers (lbm_solver_colocate)
    long long stop_at = std::numeric_limits<long long>::max();  // group half-step index of its freeze
};

struct GroupComm {
    int member, dst;
    int send_origin[3], recv_origin[3], size[3], normal[3];
};

// The launch tables of one work-group size: member descriptors (with their
// tile ranges), the tile -> member or chain map, and the chain descriptors.
struct LaunchTables {
    DomDev *d_descs = nullptr;
    size_t descs_cap = 0;
    int *d_tiles = nullptr;
    size_t tiles_cap = 0;
    int ntiles = 0;
    XChain *d_chains = nullptr;
    size_t chains_cap = 0;
    int xcd_chunk = kXcdChunkTiles;  // tiles per XCD chunk of the chunked tile map
};

// Tiles per XCD chunk for a launch whose members all have `plane` tiles per z
// plane (0: unknown or not uniform).  Beta gathers from the z neighbours,
// `plane` tiles away; when 8 chunks of K tiles divide the plane, the tile a
// plane ahead falls to the same XCD a round or more later, its rows still in
// that XCD's L2.  Keeps kXcdChunkTiles when that already holds or nothing in
// [12, 24] does.
static int xcd_chunk_for_plane(long long plane) {
    if (plane <= 0 || plane % (8 * kXcdChunkTiles) == 0) return kXcdChunkTiles;
    for (int d = 1; d <= 8; ++d)
        for (int K : {kXcdChunkTiles + d, kXcdChunkTiles - d})
            if (K >= 12 && K <= 24 && plane % (8 * K) == 0) return K;
    return kXcdChunkTiles;
}

struct lbm_group {
    std::vector<lbm_solver *> members;
    std::vector<hipStream_t> saved_streams;
    hipStream_t stream = nullptr;
    int storage = 0, wg = 256, device = 0;
    int variant = 0, lds = 0;  // launch shape of the fused step kernel (tuner-selected)
    int wg_alpha = 0;          // alpha's work-group size when it differs from wg (0: wg)
    std::string err;
    LaunchTables lt[2];        // [0]: at wg (beta, and alpha unless wg_alpha is set), [1]: alpha at wg_alpha
    long long *d_ctr = nullptr;  // device half-step counter (two slots, see step_kernel)
    std::vector<unsigned> seen_versions;
    std::vector<GroupComm> comms;
    bool plan_built = false;
    bool descs_dirty = false;
    int exchange_mode = 0;
    bool fuse_x = false;  // x faces exchanged inside the step kernel where the geometry allows
    bool fuse_auto = false;  // LBM_FUSED_X_AUTO: two-byte storages fuse only seam rings of packed pairs
    // [0]=after alpha, [1]=after beta; per phase x/y/z, and [3] the x-face copies
    // that the step kernel performs when fused (launched only by lbm_group_sync)
    CopyOp *d_ops[2][4] = {{nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};
    size_t ops_cap[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    int nops[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    int nblocks[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    std::vector<XJoin> xfuse;                 // per member, from the plan
    std::vector<std::vector<int>> chains;     // joined members in x order, from the plan
    long long step_counter = 0;
    bool use_graph = true;
    int graph_pairs = 64;  // most beta+alpha pairs one graph launch replays
    struct Graph {
        int pairs;
        hipGraph_t graph;
        hipGraphExec_t exec;
        unsigned long long used;
    };
    std::vector<Graph> graphs;           // captured step sequences by pair count, all of graph_shape
    unsigned long long graph_clock = 0;  // LRU order of graphs
    std::vector<long long> graph_shape;  // launch sizes and table addresses baked into the captured graphs
    long long graph_launches = 0;        // hipGraphLaunch calls of lbm_group_step
    long long eager_half_steps = 0;      // half-steps lbm_group_step launched without a graph
    long long graphs_captured = 0;       // graphs captured and instantiated
    bool timing = false;
    struct Ev {
        hipEvent_t a, b;
        int kind;
    };
    std::vector<Ev> events;
    double acc_ms[3] = {0, 0, 0};   // beta kernel, exchange, alpha kernel
    long long acc_n[3] = {0, 0, 0};
};

namespace {

int fail(lbm_solver *s, int code, const std::string &msg) { return record(s ? &s->err : nullptr, code, msg); }
int gfail(lbm_group *g, int code, const std::string &msg) { return record(g ? &g->err : nullptr, code, msg); }

#define SCHK(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(s, LBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define GCHK(expr)                                                                                          \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return gfail(g, LBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename F>
int dispatch_storage(int storage, F &&f) {
    switch (storage) {
        case LBM_STORAGE_F32: return f(std::integral_constant<int, 0>());
        case LBM_STORAGE_F64: return f(std::integral_constant<int, 1>());
        case LBM_STORAGE_F16_SHIFTED: return f(std::integral_constant<int, 2>());
        default: return f(std::integral_constant<int, 3>());
    }
}

void fill_desc(lbm_solver *s) {
    DomDev &d = s->host_desc;
    d.dd = s->dd;
    d.flags = s->flags;
    d.vel = s->vel;
    d.dens = s->dens;
    d.cells = s->cells;
    d.nx = s->cfg.dims[0];
    d.ny = s->cfg.dims[1];
    d.nz = s->cfg.dims[2];
    d.tile_begin = 0;
    d.ntiles = s->ntiles;
    d.forcing = s->cfg.forcing;
    d.store_velocity = s->cfg.store_velocity && s->vel;
    d.store_density = s->cfg.store_density && s->dens;
    d.inv_tau = s->cfg.inv_tau;
    d.tau0 = s->cfg.tau;
    const double cs = s->cfg.smagorinsky_cs;
    d.les_coeff = cs > 0.0 ? 18.0 * std::sqrt(2.0) * cs * cs : 0.0;
    d.gx = s->cfg.gravitation[0];
    d.gy = s->cfg.gravitation[1];
    d.gz = s->cfg.gravitation[2];
    d.lid = s->cfg.lid_velocity;
    d.stop_at = s->group ? s->stop_at : std::numeric_limits<long long>::max();
    d.seam_wpr = 0;
}

int upload_desc(lbm_solver *s) {
    fill_desc(s);
    SCHK(hipMemcpyAsync(s->d_desc, &s->host_desc, sizeof(DomDev), hipMemcpyHostToDevice, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    s->version++;
    return LBM_OK;
}

// Launch of the step kernel (launch_step_storage in lbm_device.hpp): the FP32 and
// FP64 kernels are instantiated here, the two-byte storages' in
// lbm_step_2byte.hip.
hipError_t launch_step_kernel(int storage, bool beta, int variant, int lds, int tiles, int wg, const DomDev *descs,
                              const int *tile_table, hipStream_t stream, long long *ctr = nullptr,
                              const XChain *chains = nullptr, int xcd_chunk = kXcdChunkTiles) {
    switch (storage) {
        case LBM_STORAGE_F32:
            return launch_step_storage<0>(beta, variant, lds, tiles, wg, descs, tile_table, stream, ctr, chains,
                                          xcd_chunk);
        case LBM_STORAGE_F64:
            return launch_step_storage<1>(beta, variant, lds, tiles, wg, descs, tile_table, stream, ctr, chains,
                                          xcd_chunk);
        default:
            return launch_step_2byte(storage, beta, variant, lds, tiles, wg, descs, tile_table, stream, ctr, chains,
                                     xcd_chunk);
    }
}

int launch_step(lbm_solver *s, bool beta) {
    SCHK(launch_step_kernel(s->cfg.storage, beta, s->cfg.kernel_variant, s->cfg.lds_per_block, s->ntiles, s->wg,
                            s->d_desc, s->d_tile, s->stream));
    return LBM_OK;
}

int ensure_staging(lbm_solver *s, size_t bytes) {
    if (s->staging_bytes >= bytes) return LBM_OK;
    if (s->external)
        return fail(s, LBM_ERR_CAPACITY,
                    "a solver in caller-provided memory has " + std::to_string(s->staging_bytes) +
                        " staging bytes; this transfer needs " + std::to_string(bytes) + ": transfer it in regions");
    if (s->staging) SCHK(hipFree(s->staging));
    s->staging = nullptr;
    s->staging_bytes = 0;
    if (hipMalloc(&s->staging, bytes) != hipSuccess)
        return fail(s, LBM_ERR_OUT_OF_MEMORY, "staging buffer allocation of " + std::to_string(bytes) + " bytes failed");
    s->staging_bytes = bytes;
    return LBM_OK;
}

int check_box(lbm_solver *s, const int *origin, const int *size, Box &b) {
    for (int a = 0; a < 3; ++a) {
        b.o[a] = origin ? origin[a] : 0;
        b.s[a] = origin ? size[a] : s->cfg.dims[a];
        if (b.s[a] <= 0 || b.o[a] < 0 || b.o[a] + b.s[a] > s->cfg.dims[a])
            return fail(s, LBM_ERR_OUT_OF_RANGE,
                        "region origin " + std::to_string(b.o[a]) + " size " + std::to_string(b.s[a]) +
                            " outside axis " + std::to_string(a) + " of extent " + std::to_string(s->cfg.dims[a]));
    }
    return LBM_OK;
}

int field_components(int field) {
    return (field == LBM_FIELD_DD || field == LBM_FIELD_DD_RAW) ? 19 : field == LBM_FIELD_VELOCITY ? 3 : 1;
}

// bytes of one component value of a region transfer
size_t field_elem_bytes(const lbm_solver *s, int field) {
    if (field == LBM_FIELD_FLAGS) return 4;
    if (field == LBM_FIELD_DD_RAW) return elem_size(s->cfg.storage);
    return compute_size(s->cfg.storage);
}

}  // namespace

// ---- library -----------------------------------------------------------------

extern "C" const char *lbm_version(void) { return "turbulent_lbm_multigpu_amd 0.1 (gfx950)"; }
extern "C" const char *lbm_last_error(void) { return g_last_error.c_str(); }

extern "C" int lbm_device_synchronize(int device) {
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return record(nullptr, LBM_ERR_DEVICE, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    return LBM_OK;
}

extern "C" void lbm_solver_config_default(lbm_solver_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->dims[0] = c->dims[1] = c->dims[2] = 32;
    for (int i = 0; i < 6; ++i) c->bc[i] = LBM_FLAG_OBSTACLE;
    c->storage = LBM_STORAGE_F32;
    c->forcing = LBM_FORCING_REFERENCE;
    c->tau = 0.953575;  // src/CLbmSkeleton.hpp:142
    c->inv_tau = 1.0 / 0.953575;
    c->work_group_size = 256;
    c->device = 0;
}

// ---- solver ------------------------------------------------------------------

// Device-memory layout of one solver: every buffer 256-byte aligned, in the
// order distributions, flags, velocity, density, launch descriptor, tile
// table, staging.  The tile table is sized for the most tiles any launch
// shape needs (work-groups of 64, one cell per lane), so the size does not
// depend on work_group_size or kernel_variant.  Staging holds one byte per
// cell (obstacle masks) or the 19 populations of the largest face, whichever
// is more; larger transfers go region by region.
struct SolverLayout {
    size_t dd = 0, flags = 0, vel = 0, dens = 0, desc = 0, tile = 0, staging = 0, staging_bytes = 0, total = 0;
};

static SolverLayout solver_layout(const lbm_solver_config &c, long long cells, long long padded) {
    const long long max_tiles = (cells + 63) / 64;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t es = elem_size(c.storage), cs = compute_size(c.storage);
    SolverLayout L;
    size_t off = 0;
    L.dd = off;
    off += al((size_t)19 * padded * es);
    L.flags = off;
    off += al((size_t)cells);
    if (c.store_velocity) {
        L.vel = off;
        off += al((size_t)3 * cells * cs);
    }
    if (c.store_density) {
        L.dens = off;
        off += al((size_t)cells * cs);
    }
    L.desc = off;
    off += al(sizeof(DomDev));
    L.tile = off;
    off += al(sizeof(int) * (size_t)max_tiles);
    L.staging = off;
    {
        const size_t X = c.dims[0], Y = c.dims[1], Z = c.dims[2];
        const size_t face = std::max(X * Y, std::max(Y * Z, X * Z));
        L.staging_bytes = std::max((size_t)cells, (size_t)19 * cs * face);
    }
    off += al(L.staging_bytes);
    L.total = off;
    return L;
}

static int validate_config(const lbm_solver_config &c) {
    for (int a = 0; a < 3; ++a)
        if (c.dims[a] < 3)
            return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "dims must be >= 3 on every axis");
    const long long cells = (long long)c.dims[0] * c.dims[1] * c.dims[2];
    if (cells > (1ll << 31) - (1ll << 20))
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "a (sub)domain is limited to 2^31 - 2^20 cells");
    for (int i = 0; i < 6; ++i)
        if (c.bc[i] < 0 || c.bc[i] > 255)
            return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "boundary flags must be in [0, 255]");
    if (c.storage < 0 || c.storage > 3) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "unknown storage");
    // kernels address the distributions of a domain with 32-bit byte offsets
    const unsigned long long padded = (unsigned long long)(cells + kBlk - 1) / kBlk * kBlk;
    if (19ull * padded * elem_size(c.storage) >= (1ull << 32))
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT,
                      "the distributions of one (sub)domain must stay below 4 GiB: split the domain into subdomains");
    if (c.forcing != 0 && c.forcing != 1) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "unknown forcing");
    const int wg = c.work_group_size == 0 ? 256 : c.work_group_size;
    if (!lbm_work_group_supported(wg))
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "work_group_size must be 64, 128, 192 or 256");
    if (!(c.inv_tau > 0.0) || !std::isfinite(c.inv_tau))
        return record(nullptr, LBM_ERR_PARAMETRIZATION, "inv_tau must be positive and finite");
    if (c.smagorinsky_cs < 0.0 || (c.smagorinsky_cs > 0.0 && !(c.tau > 0.5)))
        return record(nullptr, LBM_ERR_PARAMETRIZATION, "LES needs smagorinsky_cs >= 0 and tau > 0.5");
    if (!lbm_variant_supported(c.kernel_variant) || c.lds_per_block < 0 || c.lds_per_block > 65536)
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT,
                      "kernel_variant must be one of 0, 1, 2, 3, 8, 11, 16 (optionally | 4 or | 32); lds_per_block in [0, 65536]");
    return LBM_OK;
}

static void solver_geometry(const lbm_solver_config &c, long long &cells, long long &padded, int &wg, int &ntiles) {
    cells = (long long)c.dims[0] * c.dims[1] * c.dims[2];
    padded = (cells + kBlk - 1) / kBlk * kBlk;
    wg = c.work_group_size == 0 ? 256 : c.work_group_size;
    const long long per_tile = (long long)wg * cells_per_lane(c.kernel_variant);
    ntiles = (int)((cells + per_tile - 1) / per_tile);
}

static int create_solver(const lbm_solver_config *cfg, void *memory, size_t memory_bytes, hipStream_t stream,
                         lbm_solver **out) {
    if (!cfg || !out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "lbm_solver_create: null argument");
    *out = nullptr;
    const lbm_solver_config &c = *cfg;
    int rc0 = validate_config(c);
    if (rc0) return rc0;
    long long cells, padded;
    int wg, ntiles;
    solver_geometry(c, cells, padded, wg, ntiles);
    const SolverLayout L = solver_layout(c, cells, padded);
    if (memory && (memory_bytes < L.total || ((uintptr_t)memory & 255) != 0))
        return record(nullptr, LBM_ERR_CAPACITY,
                      "device memory block of " + std::to_string(memory_bytes) + " bytes (256-byte aligned) is smaller than the " +
                          std::to_string(L.total) + " bytes this solver needs");

    lbm_solver *s = new lbm_solver();
    s->cfg = c;
    s->cfg.work_group_size = wg;
    s->wg = wg;
    s->cells = cells;
    s->padded = padded;
    s->ntiles = ntiles;
    s->external = memory != nullptr;
    auto bail = [&](int code, const std::string &m) {
        record(nullptr, code, m);
        lbm_solver_destroy(s);
        return code;
    };
    if (hipSetDevice(c.device) != hipSuccess) return bail(LBM_ERR_DEVICE, "hipSetDevice failed");
    if (stream) {
        s->stream = stream;
    } else {
        if (hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess)
            return bail(LBM_ERR_DEVICE, "hipStreamCreate failed");
        s->stream = s->own_stream;
    }
    if (memory) {
        char *base = (char *)memory;
        s->dd = base + L.dd;
        s->flags = (uint8_t *)(base + L.flags);
        s->vel = c.store_velocity ? base + L.vel : nullptr;
        s->dens = c.store_density ? base + L.dens : nullptr;
        s->d_desc = (DomDev *)(base + L.desc);
        s->d_tile = (int *)(base + L.tile);
        s->staging = base + L.staging;
        s->staging_bytes = L.staging_bytes;
    } else {
        const size_t es = elem_size(c.storage), cs = compute_size(c.storage);
        if (hipMalloc(&s->dd, (size_t)19 * s->padded * es) != hipSuccess)
            return bail(LBM_ERR_OUT_OF_MEMORY, "distribution buffer allocation failed");
        if (hipMalloc((void **)&s->flags, (size_t)cells) != hipSuccess)
            return bail(LBM_ERR_OUT_OF_MEMORY, "flag buffer allocation failed");
        if (c.store_velocity && hipMalloc(&s->vel, (size_t)3 * cells * cs) != hipSuccess)
            return bail(LBM_ERR_OUT_OF_MEMORY, "velocity buffer allocation failed");
        if (c.store_density && hipMalloc(&s->dens, (size_t)cells * cs) != hipSuccess)
            return bail(LBM_ERR_OUT_OF_MEMORY, "density buffer allocation failed");
        if (hipMalloc((void **)&s->d_desc, sizeof(DomDev)) != hipSuccess ||
            hipMalloc((void **)&s->d_tile, sizeof(int) * (size_t)s->ntiles) != hipSuccess)
            return bail(LBM_ERR_OUT_OF_MEMORY, "descriptor allocation failed");
    }
    // all tiles of the single-domain launch belong to domain 0 (a fill
    // kernel, not a memset: callers may capture later launches in graphs)
    fill_int_kernel<<<(unsigned)((s->ntiles + 255) / 256), 256, 0, s->stream>>>(s->d_tile, s->ntiles, 0);
    if (hipGetLastError() != hipSuccess) return bail(LBM_ERR_DEVICE, "tile table initialisation failed");
    if (upload_desc(s) != LBM_OK) return bail(LBM_ERR_DEVICE, s->err);
    int rc = lbm_solver_reset(s);
    if (rc != LBM_OK) return bail(rc, s->err);
    *out = s;
    return LBM_OK;
}

extern "C" int lbm_solver_create(const lbm_solver_config *cfg, lbm_solver **out) {
    return create_solver(cfg, nullptr, 0, nullptr, out);
}

extern "C" int lbm_solver_required_bytes(const lbm_solver_config *cfg, size_t *bytes) {
    if (!cfg || !bytes) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    int rc = validate_config(*cfg);
    if (rc) return rc;
    long long cells, padded;
    int wg, ntiles;
    solver_geometry(*cfg, cells, padded, wg, ntiles);
    *bytes = solver_layout(*cfg, cells, padded).total;
    return LBM_OK;
}

extern "C" int lbm_solver_create_in(const lbm_solver_config *cfg, void *device_memory, size_t bytes, void *stream,
                                    lbm_solver **out) {
    if (!device_memory) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "lbm_solver_create_in: null device memory");
    return create_solver(cfg, device_memory, bytes, (hipStream_t)stream, out);
}

extern "C" int lbm_solver_destroy(lbm_solver *s) {
    if (!s) return LBM_OK;
    if (s->group) return fail(s, LBM_ERR_STATE, "destroy the group before its member solvers");
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->block) {
        (void)hipFree(s->staging);  // allocated on its own when a transfer needed it
        if (--s->block->refs == 0) {
            (void)hipFree(s->block->base);
            delete s->block;
        }
    } else if (!s->external) {
        (void)hipFree(s->dd);
        (void)hipFree(s->flags);
        (void)hipFree(s->vel);
        (void)hipFree(s->dens);
        (void)hipFree(s->d_desc);
        (void)hipFree(s->d_tile);
        (void)hipFree(s->staging);
    }
    if (s->d_partials) (void)hipFree(s->d_partials);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    delete s;
    return LBM_OK;
}

extern "C" int lbm_solver_colocate(lbm_solver *const *solvers, int n) {
    if (!solvers || n <= 0) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "lbm_solver_colocate: no solvers");
    std::vector<size_t> off((size_t)n);
    size_t total = 0;
    const int device = solvers[0] ? solvers[0]->cfg.device : -1;
    for (int i = 0; i < n; ++i) {
        lbm_solver *s = solvers[i];
        const std::string who = "lbm_solver_colocate: solver " + std::to_string(i);
        if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, who + " is null");
        for (int k = 0; k < i; ++k)
            if (solvers[k] == s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, who + " is listed twice");
        if (s->group) return record(nullptr, LBM_ERR_STATE, who + " is a group member: colocate before grouping");
        if (s->external || s->block)
            return record(nullptr, LBM_ERR_STATE, who + " already lives in a shared or caller-owned block");
        if (s->cfg.device != device) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, who + " is on another device");
        // the persistent buffers of the solver's layout; staging stays separate
        off[(size_t)i] = total;
        total += solver_layout(s->cfg, s->cells, s->padded).staging;
    }
    if (hipSetDevice(device) != hipSuccess) return record(nullptr, LBM_ERR_DEVICE, "hipSetDevice failed");
    void *base = nullptr;
    if (hipMalloc(&base, total) != hipSuccess)
        return record(nullptr, LBM_ERR_OUT_OF_MEMORY,
                      "lbm_solver_colocate: allocation of " + std::to_string(total) + " bytes failed");
    SharedBlock *blk = new SharedBlock{base, 0};
    for (int i = 0; i < n; ++i) {
        lbm_solver *s = solvers[i];
        const SolverLayout L = solver_layout(s->cfg, s->cells, s->padded);
        char *b = (char *)base + off[(size_t)i];
        const size_t es = elem_size(s->cfg.storage), cs = compute_size(s->cfg.storage);
        void *dd = b + L.dd, *vel = s->vel ? b + L.vel : nullptr, *dens = s->dens ? b + L.dens : nullptr;
        uint8_t *flags = (uint8_t *)(b + L.flags);
        int *tile = (int *)(b + L.tile);
        hipError_t e = hipMemcpyAsync(dd, s->dd, (size_t)19 * s->padded * es, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(flags, s->flags, (size_t)s->cells, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess && vel)
            e = hipMemcpyAsync(vel, s->vel, (size_t)3 * s->cells * cs, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess && dens)
            e = hipMemcpyAsync(dens, s->dens, (size_t)s->cells * cs, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(tile, s->d_tile, sizeof(int) * (size_t)s->ntiles, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            // solvers before this one have moved and hold the block; this one is unchanged
            const std::string m = std::string("lbm_solver_colocate: copy failed: ") + hipGetErrorString(e);
            if (blk->refs == 0) {
                (void)hipFree(base);
                delete blk;
            }
            return record(nullptr, LBM_ERR_DEVICE, m);
        }
        (void)hipFree(s->dd);
        (void)hipFree(s->flags);
        (void)hipFree(s->vel);
        (void)hipFree(s->dens);
        (void)hipFree(s->d_desc);
        (void)hipFree(s->d_tile);
        s->dd = dd;
        s->flags = flags;
        s->vel = vel;
        s->dens = dens;
        s->d_desc = (DomDev *)(b + L.desc);
        s->d_tile = tile;
        s->block = blk;
        blk->refs++;
        const int rc = upload_desc(s);
        if (rc) return rc;
    }
    return LBM_OK;
}

extern "C" const char *lbm_solver_error(const lbm_solver *s) { return s ? s->err.c_str() : g_last_error.c_str(); }

static int init_common(lbm_solver *s, const InitArgs &a) {
    s->step_counter = 0;
    return dispatch_storage(s->cfg.storage, [&](auto st) -> int {
        constexpr int ST = decltype(st)::value;
        const int blocks = (int)((s->cells + 255) / 256);
        init_kernel<ST><<<blocks, 256, 0, s->stream>>>(s->host_desc, a);
        SCHK(hipGetLastError());
        SCHK(hipStreamSynchronize(s->stream));
        return LBM_OK;
    });
}

extern "C" int lbm_solver_reset(lbm_solver *s) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    InitArgs a{};
    for (int i = 0; i < 6; ++i) a.bc[i] = s->cfg.bc[i];
    a.amp = 0.0;
    for (int i = 0; i < 3; ++i) a.key_dims[i] = s->cfg.dims[i];
    return init_common(s, a);
}

extern "C" int lbm_solver_init_perturbed(lbm_solver *s, double amplitude, uint64_t seed, uint64_t case_id,
                                         const int key_origin[3], const int key_dims[3]) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    if (!std::isfinite(amplitude)) return fail(s, LBM_ERR_INVALID_ARGUMENT, "amplitude must be finite");
    InitArgs a{};
    for (int i = 0; i < 6; ++i) a.bc[i] = s->cfg.bc[i];
    a.amp = amplitude;
    a.seed = seed;
    a.case_id = case_id;
    for (int i = 0; i < 3; ++i) {
        a.key_origin[i] = key_origin ? key_origin[i] : 0;
        a.key_dims[i] = key_dims ? key_dims[i] : s->cfg.dims[i];
        if (a.key_dims[i] <= 0 || a.key_origin[i] < 0 || a.key_origin[i] + s->cfg.dims[i] > a.key_dims[i])
            return fail(s, LBM_ERR_OUT_OF_RANGE, "key grid must contain the domain");
    }
    return init_common(s, a);
}

extern "C" int lbm_solver_simulation_step_alpha(lbm_solver *s) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    return launch_step(s, false);
}
extern "C" int lbm_solver_simulation_step_beta(lbm_solver *s) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    return launch_step(s, true);
}
extern "C" int lbm_solver_simulation_step(lbm_solver *s) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    int rc = launch_step(s, (s->step_counter & 1) == 0);
    if (rc == LBM_OK) s->step_counter++;
    return rc;
}
extern "C" int lbm_solver_run(lbm_solver *s, int steps) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    if (steps < 0) return fail(s, LBM_ERR_INVALID_ARGUMENT, "steps must be >= 0");
    for (int i = 0; i < steps; ++i) {
        int rc = lbm_solver_simulation_step(s);
        if (rc) return rc;
    }
    return LBM_OK;
}
extern "C" int lbm_solver_wait(lbm_solver *s) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    SCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}
extern "C" int lbm_solver_get_step_counter(const lbm_solver *s, long long *out) {
    if (!s || !out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *out = s->step_counter;
    return LBM_OK;
}
extern "C" int lbm_solver_set_step_counter(lbm_solver *s, long long counter) {
    if (!s || counter < 0) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad argument");
    s->step_counter = counter;
    return LBM_OK;
}
extern "C" int lbm_solver_get_config(const lbm_solver *s, lbm_solver_config *out) {
    if (!s || !out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *out = s->cfg;
    return LBM_OK;
}
extern "C" int lbm_solver_set_lid_velocity(lbm_solver *s, double lid) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    s->cfg.lid_velocity = lid;
    return upload_desc(s);
}
extern "C" int lbm_solver_set_relaxation(lbm_solver *s, double inv_tau, double tau, double cs) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    if (!(inv_tau > 0.0) || cs < 0.0 || (cs > 0.0 && !(tau > 0.5)))
        return fail(s, LBM_ERR_PARAMETRIZATION, "bad relaxation parameters");
    s->cfg.inv_tau = inv_tau;
    s->cfg.tau = tau;
    s->cfg.smagorinsky_cs = cs;
    return upload_desc(s);
}

extern "C" int lbm_solver_store(lbm_solver *s, int field, const int origin[3], const int size[3], void *host_dst) {
    if (!s || !host_dst || (origin && !size)) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    if (field < 0 || field > 4) return fail(s, LBM_ERR_INVALID_ARGUMENT, "unknown field");
    if (field == LBM_FIELD_VELOCITY && !s->vel)
        return fail(s, LBM_ERR_STATE, "velocity is not stored by this solver (store_velocity = 0)");
    if (field == LBM_FIELD_DENSITY && !s->dens)
        return fail(s, LBM_ERR_STATE, "density is not stored by this solver (store_density = 0)");
    Box b;
    int rc = check_box(s, origin, size, b);
    if (rc) return rc;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const size_t bytes = (size_t)n * field_components(field) * field_elem_bytes(s, field);
    rc = ensure_staging(s, bytes);
    if (rc) return rc;
    rc = dispatch_storage(s->cfg.storage, [&](auto st) -> int {
        constexpr int ST = decltype(st)::value;
        pack_kernel<ST><<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(s->host_desc, field, b, s->staging);
        SCHK(hipGetLastError());
        return LBM_OK;
    });
    if (rc) return rc;
    SCHK(hipMemcpyAsync(host_dst, s->staging, bytes, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

extern "C" int lbm_solver_set(lbm_solver *s, int field, const int origin[3], const int size[3], const void *host_src,
                              const int norm[3]) {
    if (!s || !host_src || (origin && !size)) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    if (field < 0 || field > 4) return fail(s, LBM_ERR_INVALID_ARGUMENT, "unknown field");
    if (norm && field != LBM_FIELD_DD && field != LBM_FIELD_DD_RAW)
        return fail(s, LBM_ERR_INVALID_ARGUMENT, "norm filter applies to distributions only");
    if (field == LBM_FIELD_VELOCITY && !s->vel)
        return fail(s, LBM_ERR_STATE, "velocity is not stored by this solver (store_velocity = 0)");
    if (field == LBM_FIELD_DENSITY && !s->dens)
        return fail(s, LBM_ERR_STATE, "density is not stored by this solver (store_density = 0)");
    Box b;
    int rc = check_box(s, origin, size, b);
    if (rc) return rc;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    if (field == LBM_FIELD_FLAGS) {
        const int32_t *fl = (const int32_t *)host_src;
        for (long long i = 0; i < n; ++i)
            if (fl[i] < 0 || fl[i] > 255)
                return fail(s, LBM_ERR_INVALID_ARGUMENT, "flag values must be in [0, 255]");
    }
    const size_t bytes = (size_t)n * field_components(field) * field_elem_bytes(s, field);
    rc = ensure_staging(s, bytes);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s->staging, host_src, bytes, hipMemcpyHostToDevice, s->stream));
    const uint32_t mask = norm_mask(norm);
    rc = dispatch_storage(s->cfg.storage, [&](auto st) -> int {
        constexpr int ST = decltype(st)::value;
        unpack_kernel<ST><<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(s->host_desc, field, b, s->staging, mask);
        SCHK(hipGetLastError());
        return LBM_OK;
    });
    if (rc) return rc;
    SCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

extern "C" int lbm_solver_velocity_checksum(lbm_solver *s, float *out) {
    if (!s || !out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    if (!s->vel) return fail(s, LBM_ERR_STATE, "velocity is not stored by this solver (store_velocity = 0)");
    // src/CLbmSolver.hpp:1103-1123: float sum over FLUID cells, cell order
    std::vector<int32_t> fl((size_t)s->cells);
    int rc = lbm_solver_store(s, LBM_FIELD_FLAGS, nullptr, nullptr, fl.data());
    if (rc) return rc;
    float checksum = 0.0f;
    if (s->cfg.storage == LBM_STORAGE_F64) {
        std::vector<double> v((size_t)s->cells * 3);
        rc = lbm_solver_store(s, LBM_FIELD_VELOCITY, nullptr, nullptr, v.data());
        if (rc) return rc;
        for (long long a = 0; a < s->cells; ++a)
            if (fl[a] == LBM_FLAG_FLUID) checksum += (float)(v[a] + v[s->cells + a] + v[2 * s->cells + a]);
    } else {
        std::vector<float> v((size_t)s->cells * 3);
        rc = lbm_solver_store(s, LBM_FIELD_VELOCITY, nullptr, nullptr, v.data());
        if (rc) return rc;
        for (long long a = 0; a < s->cells; ++a)
            if (fl[a] == LBM_FLAG_FLUID) checksum += v[a] + v[s->cells + a] + v[2 * s->cells + a];
    }
    *out = checksum;
    return LBM_OK;
}

namespace {

// The two reduction passes of the FP64 mass and kinetic energy into
// `partials` (2 * (blocks + 1) doubles: the result pair, then one pair per
// block; every element is written before it is read, so nothing needs
// clearing), then the result pair to the host.
int mass_energy_passes(lbm_solver *s, double *partials, int blocks, double *mass, double *ke) {
    const int swapped = (s->step_counter % 2) == 0;
    int rc = dispatch_storage(s->cfg.storage, [&](auto st) -> int {
        constexpr int ST = decltype(st)::value;
        mass_energy_kernel<ST><<<blocks, 256, 0, s->stream>>>(s->host_desc, swapped, partials + 2);
        SCHK(hipGetLastError());
        return LBM_OK;
    });
    if (rc) return rc;
    sum_pairs_kernel<<<1, 256, 0, s->stream>>>(partials + 2, blocks, partials);
    SCHK(hipGetLastError());
    double res[2];
    SCHK(hipMemcpyAsync(res, partials, sizeof(res), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    *mass = res[0];
    *ke = res[1];
    return LBM_OK;
}

}  // namespace

extern "C" int lbm_solver_mass_energy(lbm_solver *s, double *mass, double *ke) {
    if (!s || !mass || !ke) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    const int blocks = (int)((s->cells + 255) / 256);
    if (s->partial_count < (size_t)blocks + 1) {
        if (s->d_partials) SCHK(hipFree(s->d_partials));
        s->d_partials = nullptr;
        if (hipMalloc((void **)&s->d_partials, sizeof(double) * 2 * ((size_t)blocks + 1)) != hipSuccess)
            return fail(s, LBM_ERR_OUT_OF_MEMORY, "reduction buffer allocation failed");
        s->partial_count = (size_t)blocks + 1;
    }
    return mass_energy_passes(s, s->d_partials, blocks, mass, ke);
}

// ---- allocation-free reduction (no reference counterpart) -------------------
// Scratch of lbm_solver_mass_energy_noalloc: the result pair and one partial
// pair per 256-cell block, for solvers of at most kScratchBlocks * 256 cells.
// Each device holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress gives on that device.
constexpr int kScratchBlocks = 1 << 17;
static __device__ __align__(256) double scratch[2 * (kScratchBlocks + 1)];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on `device`, resolved once per device with
// that device current.
int scratch_address(lbm_solver *s, int device, double **out) {
    if (device < 0 || device >= kScratchDevices)
        return fail(s, LBM_ERR_DEVICE, "device ordinal " + std::to_string(device) + " is beyond the " +
                                           std::to_string(kScratchDevices) + " devices the scratch table covers");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        SCHK(hipGetDevice(&prev));
        SCHK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        SCHK(hipSetDevice(prev));
        if (e != hipSuccess || !p)
            return fail(s, LBM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return LBM_OK;
}

}  // namespace

extern "C" int lbm_solver_mass_energy_noalloc(lbm_solver *s, double *mass, double *ke) {
    if (!s || !mass || !ke) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    const long long blocks = (s->cells + 255) / 256;
    if (blocks > kScratchBlocks)
        return fail(s, LBM_ERR_CAPACITY,
                    "lbm_solver_mass_energy_noalloc: " + std::to_string(s->cells) + " cells exceed the " +
                        std::to_string((long long)kScratchBlocks * 256) +
                        " cells its fixed scratch covers (use lbm_solver_mass_energy)");
    double *ptr = nullptr;
    const int rc = scratch_address(s, s->cfg.device, &ptr);
    if (rc) return rc;
    return mass_energy_passes(s, ptr, (int)blocks, mass, ke);
}

extern "C" int lbm_solver_moments(lbm_solver *s, const int origin[3], const int size[3], void *host_rho, void *host_j) {
    if (!s || !host_rho || !host_j || (origin && !size)) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    Box b;
    int rc = check_box(s, origin, size, b);
    if (rc) return rc;
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const size_t cs = compute_size(s->cfg.storage);
    rc = ensure_staging(s, (size_t)n * 4 * cs);
    if (rc) return rc;
    const int swapped = (s->step_counter % 2) == 0;
    char *base = (char *)s->staging;
    rc = dispatch_storage(s->cfg.storage, [&](auto st) -> int {
        constexpr int ST = decltype(st)::value;
        moments_kernel<ST><<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(s->host_desc, b, swapped, base,
                                                                                 base + (size_t)n * cs);
        SCHK(hipGetLastError());
        return LBM_OK;
    });
    if (rc) return rc;
    SCHK(hipMemcpyAsync(host_rho, base, (size_t)n * cs, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipMemcpyAsync(host_j, base + (size_t)n * cs, (size_t)n * 3 * cs, hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

extern "C" int lbm_solver_device_pointers(lbm_solver *s, void **dd, void **flags, long long *block_cells) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    if (dd) *dd = s->dd;
    if (flags) *flags = s->flags;
    if (block_cells) *block_cells = kBlk;
    return LBM_OK;
}

static int fill_op(lbm_solver *src, const int *so, lbm_solver *dst, const int *dof, const int *size, uint32_t mask,
                   CopyOp &op, std::string &why) {
    for (int a = 0; a < 3; ++a) {
        if (size[a] <= 0 || so[a] < 0 || dof[a] < 0 || so[a] + size[a] > src->cfg.dims[a] ||
            dof[a] + size[a] > dst->cfg.dims[a]) {
            why = "ghost-layer box outside a subdomain on axis " + std::to_string(a);
            return LBM_ERR_OUT_OF_RANGE;
        }
    }
    op.src = src->dd;
    op.dst = dst->dd;
    op.src_nx = src->cfg.dims[0];
    op.src_ny = src->cfg.dims[1];
    op.dst_nx = dst->cfg.dims[0];
    op.dst_ny = dst->cfg.dims[1];
    for (int a = 0; a < 3; ++a) {
        op.so[a] = so[a];
        op.dof[a] = dof[a];
        op.sz[a] = size[a];
    }
    op.mask = mask;
    op.npop = 0;
    for (int i = 0; i < 19; ++i)
        if (mask & (1u << i)) op.pops[op.npop++] = (unsigned char)i;
    const long long cells = (long long)size[0] * size[1] * size[2];
    if (cells * op.npop > (1ll << 30)) {
        why = "ghost-layer box too large for one transfer";
        return LBM_ERR_OUT_OF_RANGE;
    }
    op.cells = (int)cells;
    op.vec = (so[0] % 4 == 0 && dof[0] % 4 == 0 && size[0] % 4 == 0 && op.src_nx % 4 == 0 && op.dst_nx % 4 == 0)
                 ? 4
                 : 1;
    op.block_begin = 0;
    op.stop_at = dst->stop_at;
    return LBM_OK;
}

static hipError_t exchange_launch(int storage, const CopyOp *d_ops, int nops, int nblocks, hipStream_t stream,
                                  const long long *ctr = nullptr) {
    switch (elem_size(storage)) {
        case 8: exchange_kernel<8><<<nblocks, 256, 0, stream>>>(d_ops, nops, ctr); break;
        case 4: exchange_kernel<4><<<nblocks, 256, 0, stream>>>(d_ops, nops, ctr); break;
        default: exchange_kernel<2><<<nblocks, 256, 0, stream>>>(d_ops, nops, ctr); break;
    }
    return hipGetLastError();
}

extern "C" int lbm_copy_box(lbm_solver *src, const int src_origin[3], lbm_solver *dst, const int dst_origin[3],
                            const int size[3], const int norm[3]) {
    if (!src || !dst || !src_origin || !dst_origin || !size)
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    lbm_solver *s = dst;
    if (src->cfg.storage != dst->cfg.storage) return fail(s, LBM_ERR_INVALID_ARGUMENT, "storage types differ");
    CopyOp op;
    std::string why;
    int rc = fill_op(src, src_origin, dst, dst_origin, size, norm_mask(norm), op, why);
    if (rc) return fail(s, rc, why);
    if (op.npop == 0) return LBM_OK;
    SCHK(hipStreamSynchronize(src->stream));
    CopyOp *d_op = nullptr;
    SCHK(hipMalloc((void **)&d_op, sizeof(CopyOp)));
    SCHK(hipMemcpyAsync(d_op, &op, sizeof(CopyOp), hipMemcpyHostToDevice, s->stream));
    SCHK(exchange_launch(s->cfg.storage, d_op, 1, (op.cells / op.vec * op.npop + 255) / 256, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    SCHK(hipFree(d_op));
    return LBM_OK;
}

// ---- seeded spherical obstacles ------------------------------------------------

namespace {
uint64_t host_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double host_unit(uint64_t seed, uint64_t key) {
    return (double)(host_mix64(seed ^ host_mix64(key)) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void mark_obstacles_kernel(DomDev D, const uint8_t *solid, Box b) {
    const long long n = (long long)b.s[0] * b.s[1] * b.s[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || !solid[e]) return;
    const long long g = box_cell(D, b, e);
    const uint8_t f = D.flags[g];
    if (f == kFlagFluid) D.flags[g] = (uint8_t)kFlagObstacle;
}
}  // namespace

// Spheres with centres uniform in the key grid's interior and radii uniform in
// [rmin, rmax], drawn from splitmix64(seed, k) in order k = 0, 1, ... until the
// union covers `fill` of the key grid's interior cells; every FLUID cell of the
// solver inside a sphere (at its key-grid position) becomes OBSTACLE.  The
// same (seed, key grid) gives the same obstacles in every subdomain.
extern "C" int lbm_solver_place_spheres(lbm_solver *s, uint64_t seed, double fill, double rmin, double rmax,
                                        const int key_origin[3], const int key_dims[3], double *achieved) {
    if (!s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null solver");
    if (!(fill >= 0.0 && fill < 0.9) || !(rmin > 0.0) || !(rmax >= rmin))
        return fail(s, LBM_ERR_INVALID_ARGUMENT, "need 0 <= fill < 0.9 and 0 < rmin <= rmax");
    int ko[3], kd[3];
    for (int a = 0; a < 3; ++a) {
        ko[a] = key_origin ? key_origin[a] : 0;
        kd[a] = key_dims ? key_dims[a] : s->cfg.dims[a];
        if (kd[a] < 3 || ko[a] < 0 || ko[a] + s->cfg.dims[a] > kd[a])
            return fail(s, LBM_ERR_OUT_OF_RANGE, "key grid must contain the domain");
    }
    const long long KX = kd[0], KY = kd[1], KZ = kd[2];
    const long long interior = (KX - 2) * (KY - 2) * (KZ - 2);
    if (KX * KY * KZ > (1ll << 33)) return fail(s, LBM_ERR_INVALID_ARGUMENT, "key grid too large");
    std::vector<uint8_t> solid((size_t)(KX * KY * KZ), 0);
    long long marked = 0, k = 0;
    const long long target = (long long)std::ceil(fill * (double)interior);
    while (marked < target && k < 4000000) {
        const double cx = 1.0 + host_unit(seed, 4 * k + 0) * (double)(KX - 2);
        const double cy = 1.0 + host_unit(seed, 4 * k + 1) * (double)(KY - 2);
        const double cz = 1.0 + host_unit(seed, 4 * k + 2) * (double)(KZ - 2);
        const double r = rmin + host_unit(seed, 4 * k + 3) * (rmax - rmin);
        ++k;
        const long long x0 = std::max(1ll, (long long)std::floor(cx - r)), x1 = std::min(KX - 2, (long long)std::ceil(cx + r));
        const long long y0 = std::max(1ll, (long long)std::floor(cy - r)), y1 = std::min(KY - 2, (long long)std::ceil(cy + r));
        const long long z0 = std::max(1ll, (long long)std::floor(cz - r)), z1 = std::min(KZ - 2, (long long)std::ceil(cz + r));
        for (long long z = z0; z <= z1; ++z)
            for (long long y = y0; y <= y1; ++y)
                for (long long x = x0; x <= x1; ++x) {
                    const double dx = x - cx, dy = y - cy, dz = z - cz;
                    if (dx * dx + dy * dy + dz * dz > r * r) continue;
                    uint8_t &c = solid[(size_t)(x + KX * (y + KY * z))];
                    if (!c) {
                        c = 1;
                        ++marked;
                    }
                }
    }
    if (achieved) *achieved = interior ? (double)marked / (double)interior : 0.0;
    // the solver's box inside the key grid
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.o[a] = 0;
        b.s[a] = s->cfg.dims[a];
    }
    const long long n = s->cells;
    std::vector<uint8_t> local((size_t)n);
    for (int z = 0; z < s->cfg.dims[2]; ++z)
        for (int y = 0; y < s->cfg.dims[1]; ++y)
            for (int x = 0; x < s->cfg.dims[0]; ++x)
                local[(size_t)(x + (long long)s->cfg.dims[0] * (y + (long long)s->cfg.dims[1] * z))] =
                    solid[(size_t)((ko[0] + x) + KX * ((ko[1] + y) + KY * (long long)(ko[2] + z)))];
    int rc = ensure_staging(s, (size_t)n);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s->staging, local.data(), (size_t)n, hipMemcpyHostToDevice, s->stream));
    mark_obstacles_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(s->host_desc, (const uint8_t *)s->staging, b);
    SCHK(hipGetLastError());
    SCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

// ---- device copy bandwidth (calibration of the roofline) --------------------

__global__ void __launch_bounds__(256) stream_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Each work-item moves 4 float4 spaced one block apart: four loads in flight,
// then four stores; NT selects the non-temporal cache policy.
typedef float v4f __attribute__((ext_vector_type(4)));
template <bool NT>
__global__ void __launch_bounds__(256) stream_copy4_kernel(const v4f *__restrict__ src, v4f *__restrict__ dst,
                                                           long long n) {
    const long long base = (long long)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    v4f v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long i = base + (long long)k * blockDim.x;
        if (i < n) v[k] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long i = base + (long long)k * blockDim.x;
        if (i < n) {
            if (NT) __builtin_nontemporal_store(v[k], dst + i);
            else dst[i] = v[k];
        }
    }
}

extern "C" int lbm_device_copy_bandwidth(int device, size_t bytes, int reps, int variant, double *gbs,
                                         double *kernel_ms) {
    if (bytes < 4096 || reps <= 0 || !gbs || variant < 0 || variant > 2)
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad probe arguments");
    bytes &= ~(size_t)15;
    hipError_t e = hipSetDevice(device);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    const long long n = (long long)(bytes / 16);
    auto launch = [&]() {
        if (variant == 0) stream_copy_kernel<<<256 * 32, 256>>>((const float4 *)a, (float4 *)b, n);
        else if (variant == 1) stream_copy4_kernel<false><<<(unsigned)((n + 1023) / 1024), 256>>>((const v4f *)a, (v4f *)b, n);
        else stream_copy4_kernel<true><<<(unsigned)((n + 1023) / 1024), 256>>>((const v4f *)a, (v4f *)b, n);
    };
    if (e == hipSuccess) e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        launch();
        e = hipEventRecord(e0, 0);
    }
    if (e == hipSuccess) {
        for (int r = 0; r < reps; ++r) launch();
        e = hipEventRecord(e1, 0);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return record(nullptr, LBM_ERR_DEVICE, std::string("copy probe: ") + hipGetErrorString(e));
    const double per = (double)ms / reps;
    *gbs = 2.0 * (double)(n * 16) / (per * 1e-3) / 1e9;
    if (kernel_ms) *kernel_ms = per;
    return LBM_OK;
}

// ---- group -----------------------------------------------------------------

static void group_drop_graph(lbm_group *g) {
    for (auto &e : g->graphs) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    g->graphs.clear();
    g->graph_shape.clear();
}

static void group_free_plan(lbm_group *g) {
    for (int k = 0; k < 2; ++k)
        for (int p = 0; p < 4; ++p) {
            if (g->d_ops[k][p]) (void)hipFree(g->d_ops[k][p]);
            g->d_ops[k][p] = nullptr;
            g->ops_cap[k][p] = 0;
            g->nops[k][p] = 0;
            g->nblocks[k][p] = 0;
        }
    group_drop_graph(g);
    g->plan_built = false;
}

// Device tables grow by doubling and are never shrunk, so membership changes
// of a continuous batch rewrite them in place (no free, no device-wide wait).
template <typename T>
static hipError_t grow_table(T *&ptr, size_t &cap, size_t n) {
    if (ptr && n <= cap) return hipSuccess;
    size_t c = std::max<size_t>(std::max<size_t>(n, 2 * cap), 16);
    if (ptr) {
        hipError_t e = hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc((void **)&ptr, sizeof(T) * c);
    if (e == hipSuccess) cap = c;
    return e;
}

// The group's step kernels hold packed pairs: a two-byte storage launched with
// two cells per work-item (kernel variants 8 and 11).
static bool group_pairs(const lbm_group *g) {
    return elem_size(g->storage) == 2 && cells_per_lane(g->variant & ~4) == 2;
}

// Fills the launch tables of the step kernel at work-group size wg: every
// member's descriptor with its tile range, and the tile map (member index,
// or -chain-1 for the tiles of a chain).
static int build_tables(lbm_group *g, int wg, LaunchTables &T) {
    const size_t nm = g->members.size();
    std::vector<DomDev> descs(nm);
    std::vector<int> tiles;
    std::vector<XChain> xch;
    const long long per_tile = (long long)wg * cells_per_lane(g->variant);
    // members of a chain (plan): chain index and position
    std::vector<int> chain_of(nm, -1), pos(nm, -1);
    for (size_t c = 0; c < g->chains.size(); ++c)
        for (size_t k = 0; k < g->chains[c].size(); ++k) {
            chain_of[g->chains[c][k]] = (int)c;
            pos[g->chains[c][k]] = (int)k;
        }
    for (size_t i = 0; i < nm; ++i) {
        lbm_solver *m = g->members[i];
        fill_desc(m);
        descs[i] = m->host_desc;
        if (i < g->xfuse.size())
            for (int side = 0; side < 2; ++side) descs[i].xmask_a[side] = g->xfuse[i].mask[side];
    }
    long long plane = -1;  // tiles per z plane, shared by every member so far (0: not shared or not whole)
    auto note_plane = [&plane](long long p) { plane = (plane < 0 || plane == p) ? p : 0; };
    for (size_t i = 0; i < nm; ++i) {
        const int c = chain_of[i];
        if (c < 0) {
            const long long pc = (long long)g->members[i]->cfg.dims[0] * g->members[i]->cfg.dims[1];
            note_plane(pc % per_tile == 0 ? pc / per_tile : 0);
            const int nt = (int)((g->members[i]->cells + per_tile - 1) / per_tile);
            descs[i].tile_begin = (int)tiles.size();
            descs[i].ntiles = nt;
            tiles.insert(tiles.end(), (size_t)nt, (int)i);
            continue;
        }
        if (pos[i] != 0) continue;  // a chain's tiles go with its first member
        if (g->chains[c].size() == 1 && g->xfuse[i].nb[0] == (int)i &&
            seam_ring_ok(group_pairs(g), g->members[i]->cfg.dims[0])) {
            // a periodic ring of one in packed pairs: its seam stays in one
            // wavefront per row (seam_pairs), tiles of whole wavefronts
            const int X = g->members[i]->cfg.dims[0];
            const long long pw = (long long)seam_waves_per_row(X) * g->members[i]->cfg.dims[1];
            note_plane(pw % (wg / 64) == 0 ? pw / (wg / 64) : 0);
            const long long waves = pw * g->members[i]->cfg.dims[2];
            const int nt = (int)((waves + wg / 64 - 1) / (wg / 64));
            descs[i].tile_begin = (int)tiles.size();
            descs[i].ntiles = nt;
            descs[i].seam_wpr = seam_waves_per_row(X);
            tiles.insert(tiles.end(), (size_t)nt, (int)i);
            continue;
        }
        // a chain tile: the same R rows of all M members, each member on
        // ceil(R X / 64) wavefronts, M of those within the work-group and
        // R M <= kChainRows; the most rows that fit
        const std::vector<int> &ch = g->chains[c];
        const int M = (int)ch.size();
        const lbm_solver *m0 = g->members[ch[0]];
        const int X = m0->cfg.dims[0];
        const long long rows = (long long)m0->cfg.dims[1] * m0->cfg.dims[2];
        const ChainShape sh = chain_shape(M, X, wg);  // fits: the plan kept only chains that fit every wg
        XChain xc{};
        xc.nmem = M;
        xc.nx = X;
        xc.rows = sh.rows;
        xc.lanes = sh.lanes;
        xc.nedge = sh.nedge;
        for (int k = 0; k < 3; ++k) xc.edge[k] = sh.edge[k];
        xc.ninner = sh.ninner;
        xc.ipasses = sh.ipasses;
        const int R = xc.rows;
        xc.tile_begin = (int)tiles.size();
        for (int k = 0; k < M; ++k) {
            const int mem = ch[k];
            xc.mem[k] = mem;
            xc.left[k] = g->xfuse[mem].nb[0] >= 0 ? pos[g->xfuse[mem].nb[0]] : -1;
            xc.right[k] = g->xfuse[mem].nb[1] >= 0 ? pos[g->xfuse[mem].nb[1]] : -1;
        }
        const int nt = (int)((rows + R - 1) / R) * (1 + xc.ipasses);
        const long long Y = m0->cfg.dims[1];
        note_plane(Y % R == 0 ? Y / R * (1 + xc.ipasses) : 0);
        for (int k = 0; k < M; ++k) {
            descs[ch[k]].tile_begin = xc.tile_begin;
            descs[ch[k]].ntiles = nt;
        }
        tiles.insert(tiles.end(), (size_t)nt, -(int)xch.size() - 1);
        xch.push_back(xc);
    }
    GCHK(grow_table(T.d_tiles, T.tiles_cap, tiles.size()));
    GCHK(grow_table(T.d_descs, T.descs_cap, descs.size()));
    if (!xch.empty()) GCHK(grow_table(T.d_chains, T.chains_cap, xch.size()));
    T.ntiles = (int)tiles.size();
    T.xcd_chunk = xcd_chunk_for_plane(plane > 0 ? plane : 0);
    if (!tiles.empty())
        GCHK(hipMemcpyAsync(T.d_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, g->stream));
    if (!xch.empty())
        GCHK(hipMemcpyAsync(T.d_chains, xch.data(), sizeof(XChain) * xch.size(), hipMemcpyHostToDevice, g->stream));
    if (!descs.empty())
        GCHK(hipMemcpyAsync(T.d_descs, descs.data(), sizeof(DomDev) * descs.size(), hipMemcpyHostToDevice, g->stream));
    // the host vectors are released on return: the copies have to finish first
    GCHK(hipStreamSynchronize(g->stream));
    return LBM_OK;
}

static bool alpha_split(const lbm_group *g) { return g->wg_alpha != 0 && g->wg_alpha != g->wg; }

// Descriptor and tile tables of the fused launches (beta's, and alpha's at
// its own work-group size) for the group's current members and launch shape;
// also re-publishes the host half-step count to the device counter.
static int group_upload_descs(lbm_group *g) {
    int rc = build_tables(g, g->wg, g->lt[0]);
    if (!rc && alpha_split(g)) rc = build_tables(g, g->wg_alpha, g->lt[1]);
    if (rc) return rc;
    const long long ctr[2] = {g->step_counter, g->step_counter};
    GCHK(hipMemcpyAsync(g->d_ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice, g->stream));
    GCHK(hipStreamSynchronize(g->stream));
    for (size_t i = 0; i < g->members.size(); ++i) g->seen_versions[i] = g->members[i]->version;
    g->descs_dirty = false;
    return LBM_OK;
}

static int group_init(lbm_group *g) {
    GCHK(hipSetDevice(g->device));
    GCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    GCHK(hipMalloc((void **)&g->d_ctr, 2 * sizeof(long long)));
    return group_upload_descs(g);
}

extern "C" int lbm_group_create(lbm_solver *const *solvers, int n, lbm_group **out) {
    if (!solvers || n <= 0 || !out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad group arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!solvers[i]) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null member solver");
        if (solvers[i]->group) return record(nullptr, LBM_ERR_STATE, "a solver can belong to one group only");
        if (solvers[i]->cfg.storage != solvers[0]->cfg.storage || solvers[i]->wg != solvers[0]->wg ||
            solvers[i]->cfg.device != solvers[0]->cfg.device ||
            solvers[i]->cfg.kernel_variant != solvers[0]->cfg.kernel_variant ||
            solvers[i]->cfg.lds_per_block != solvers[0]->cfg.lds_per_block)
            return record(nullptr, LBM_ERR_INVALID_ARGUMENT,
                          "group members need the same storage, work-group size, kernel variant and device");
        for (int j = 0; j < i; ++j)
            if (solvers[j] == solvers[i]) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "duplicate member solver");
        // members advance together, so their step counters must share a parity
        // (the AA pattern's layout); each keeps its own count
        if ((solvers[i]->step_counter & 1) != (solvers[0]->step_counter & 1))
            return record(nullptr, LBM_ERR_STATE, "member step counters differ in parity");
    }
    lbm_group *g = new lbm_group();
    g->storage = solvers[0]->cfg.storage;
    g->wg = solvers[0]->wg;
    g->device = solvers[0]->cfg.device;
    g->variant = solvers[0]->cfg.kernel_variant;
    g->lds = solvers[0]->cfg.lds_per_block;
    g->step_counter = solvers[0]->step_counter;
    if (hipSetDevice(g->device) != hipSuccess) {
        delete g;
        return record(nullptr, LBM_ERR_DEVICE, "hipSetDevice failed");
    }
    for (int i = 0; i < n; ++i) {
        lbm_solver *m = solvers[i];
        if (hipStreamSynchronize(m->stream) != hipSuccess) {
            lbm_group_destroy(g);
            return record(nullptr, LBM_ERR_DEVICE, "member stream sync failed");
        }
        g->members.push_back(m);
        g->saved_streams.push_back(m->stream);
        g->seen_versions.push_back(~0u);
        m->group = g;
        m->stop_at = std::numeric_limits<long long>::max();
    }
    int rc = group_init(g);
    if (rc) {
        std::string msg = g->err;
        lbm_group_destroy(g);
        return record(nullptr, rc, msg);
    }
    for (auto *m : g->members) m->stream = g->stream;
    *out = g;
    return LBM_OK;
}

extern "C" int lbm_group_create_empty(int device, int storage, int work_group_size, int kernel_variant,
                                      lbm_group **out) {
    if (!out) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (storage < 0 || storage > 3 || !lbm_work_group_supported(work_group_size) ||
        !lbm_variant_supported(kernel_variant))
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad storage or launch shape");
    lbm_group *g = new lbm_group();
    g->storage = storage;
    g->wg = work_group_size;
    g->device = device;
    g->variant = kernel_variant;
    int rc = group_init(g);
    if (rc) {
        std::string msg = g->err;
        lbm_group_destroy(g);
        return record(nullptr, rc, msg);
    }
    *out = g;
    return LBM_OK;
}

extern "C" int lbm_group_destroy(lbm_group *g) {
    if (!g) return LBM_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto &e : g->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    group_free_plan(g);
    for (size_t i = 0; i < g->members.size(); ++i) {
        lbm_solver *m = g->members[i];
        if (i < g->saved_streams.size()) m->stream = g->saved_streams[i];
        m->group = nullptr;
        m->stop_at = std::numeric_limits<long long>::max();
    }
    for (LaunchTables &T : g->lt) {
        (void)hipFree(T.d_descs);
        (void)hipFree(T.d_tiles);
        (void)hipFree(T.d_chains);
    }
    (void)hipFree(g->d_ctr);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return LBM_OK;
}

extern "C" const char *lbm_group_error(const lbm_group *g) { return g ? g->err.c_str() : g_last_error.c_str(); }

extern "C" int lbm_group_add_member(lbm_group *g, lbm_solver *s, int *index) {
    if (!g || !s) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    if (s->group) return gfail(g, LBM_ERR_STATE, "a solver can belong to one group only");
    if (s->cfg.storage != g->storage || s->cfg.device != g->device)
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "a member needs the group's storage and device");
    if ((s->step_counter & 1) != (g->step_counter & 1))
        return gfail(g, LBM_ERR_STATE, "the solver's step counter differs in parity from the group's");
    GCHK(hipSetDevice(g->device));
    GCHK(hipStreamSynchronize(s->stream));
    GCHK(hipStreamSynchronize(g->stream));
    g->members.push_back(s);
    g->saved_streams.push_back(s->stream);
    g->seen_versions.push_back(~0u);
    s->group = g;
    s->stop_at = std::numeric_limits<long long>::max();
    s->stream = g->stream;
    g->descs_dirty = true;
    g->plan_built = false;
    if (index) *index = (int)g->members.size() - 1;
    return LBM_OK;
}

extern "C" int lbm_group_remove_member(lbm_group *g, int index) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (index < 0 || index >= (int)g->members.size()) return gfail(g, LBM_ERR_INVALID_ARGUMENT, "member index out of range");
    GCHK(hipStreamSynchronize(g->stream));
    lbm_solver *m = g->members[index];
    m->stream = g->saved_streams[index];
    m->group = nullptr;
    m->stop_at = std::numeric_limits<long long>::max();
    g->members.erase(g->members.begin() + index);
    g->saved_streams.erase(g->saved_streams.begin() + index);
    g->seen_versions.erase(g->seen_versions.begin() + index);
    std::vector<GroupComm> kept;
    for (GroupComm c : g->comms) {
        if (c.member == index || c.dst == index) continue;
        if (c.member > index) --c.member;
        if (c.dst > index) --c.dst;
        kept.push_back(c);
    }
    g->comms.swap(kept);
    g->descs_dirty = true;
    g->plan_built = false;
    if (upload_desc(m) != LBM_OK) return gfail(g, LBM_ERR_DEVICE, m->err);
    return LBM_OK;
}

extern "C" int lbm_group_get_stream(const lbm_group *g, void **stream) {
    if (!g || !stream) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *stream = (void *)g->stream;
    return LBM_OK;
}

extern "C" int lbm_group_member_count(const lbm_group *g, int *count) {
    if (!g || !count) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *count = (int)g->members.size();
    return LBM_OK;
}

extern "C" int lbm_group_set_member_stop(lbm_group *g, int index, long long steps) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (index < 0 || index >= (int)g->members.size()) return gfail(g, LBM_ERR_INVALID_ARGUMENT, "member index out of range");
    lbm_solver *m = g->members[index];
    m->stop_at = steps < 0 ? std::numeric_limits<long long>::max() : g->step_counter + steps;
    m->version++;
    g->plan_built = false;  // the exchange copies carry the receiver's stop
    return LBM_OK;
}

extern "C" int lbm_group_add_comm(lbm_group *g, int member, int dst_member, const int send_origin[3],
                                  const int recv_origin[3], const int size[3], const int normal[3]) {
    if (!g || !send_origin || !recv_origin || !size || !normal)
        return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    const int n = (int)g->members.size();
    if (member < 0 || member >= n || dst_member < 0 || dst_member >= n)
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "member index out of range");
    int nz = 0;
    for (int a = 0; a < 3; ++a) nz += normal[a] != 0;
    if (nz != 1 || std::abs(normal[0] + normal[1] + normal[2]) != 1)
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "comm normal must be a unit axis vector");
    GroupComm c;
    c.member = member;
    c.dst = dst_member;
    for (int a = 0; a < 3; ++a) {
        c.send_origin[a] = send_origin[a];
        c.recv_origin[a] = recv_origin[a];
        c.size[a] = size[a];
        c.normal[a] = normal[a];
    }
    g->comms.push_back(c);
    g->plan_built = false;
    return LBM_OK;
}

extern "C" int lbm_group_set_exchange_mode(lbm_group *g, int mode) {
    if (!g || (mode != 0 && mode != 1)) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "exchange mode must be 0 or 1");
    g->exchange_mode = mode;
    g->plan_built = false;
    return LBM_OK;
}

extern "C" int lbm_group_set_fused_x(lbm_group *g, int mode) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (mode != 0 && mode != 1 && mode != LBM_FUSED_X_AUTO)
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "fused-x mode must be 0, 1 or LBM_FUSED_X_AUTO");
    g->fuse_auto = mode == LBM_FUSED_X_AUTO;
    g->fuse_x = mode != 0;
    g->plan_built = false;
    return LBM_OK;
}

extern "C" int lbm_group_set_graph(lbm_group *g, int enable) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    g->use_graph = enable != 0;
    return LBM_OK;
}

// The copy plan of the exchange.  With fuse_x, an x-face comm between
// equally sized members with the reference's face geometry is done by the
// step kernel instead: members joined that way form chains along x whose
// tiles hold the same rows of every member (XChain), and the face values move
// through LDS; the copies of a joined face are kept for lbm_group_sync only.
static bool fusable_x(const lbm_group *g, const GroupComm &c, const GroupComm &p) {
    if (!g->fuse_x || c.normal[0] == 0) return false;
    const lbm_solver *me = g->members[c.member], *nb = g->members[c.dst];
    const int X = me->cfg.dims[0], Y = me->cfg.dims[1], Z = me->cfg.dims[2];
    if (X < 4 || nb->cfg.dims[0] != X || nb->cfg.dims[1] != Y || nb->cfg.dims[2] != Z) return false;
    if (c.size[0] != 1 || c.size[1] != Y || c.size[2] != Z || p.size[0] != 1 || p.size[1] != Y || p.size[2] != Z)
        return false;
    for (int a = 1; a < 3; ++a)
        if (c.send_origin[a] || c.recv_origin[a] || p.send_origin[a] || p.recv_origin[a]) return false;
    if (c.normal[0] == 1)  // my x- face: ghost 0, inner 1; the neighbour's x+ face
        return c.recv_origin[0] == 0 && c.send_origin[0] == 1 && p.send_origin[0] == X - 2 &&
               p.recv_origin[0] == X - 1;
    return c.recv_origin[0] == X - 1 && c.send_origin[0] == X - 2 && p.send_origin[0] == 1 && p.recv_origin[0] == 0;
}

// after alpha: the slots of my ghost layer the neighbour's inner layer
// replaces (src/CController.hpp:265-320; mode 1: those an inner cell reads)
static uint32_t alpha_mask(const lbm_group *g, const GroupComm &c) {
    if (g->exchange_mode != 1) return (1u << 19) - 1;
    uint32_t m = 0;
    for (int i = 0; i < 19; ++i)
        if (c.normal[0] * ex(i) + c.normal[1] * ey(i) + c.normal[2] * ez(i) < 0) m |= 1u << i;
    return m;
}

// Joins the fusable x faces into chains (lines, or rings where x is
// periodic) and keeps those a chain tile can hold (chain_shape).
static void plan_chains(lbm_group *g, const std::vector<char> &fuse) {
    std::vector<JoinCandidate> cand;
    for (size_t ci = 0; ci < g->comms.size(); ++ci) {
        if (!fuse[ci]) continue;
        const GroupComm &c = g->comms[ci];
        cand.push_back({c.member, c.normal[0] == 1 ? 0 : 1, c.dst, alpha_mask(g, c)});  // inward +1: my x- face
    }
    g->xfuse = resolve_joins((int)g->members.size(), cand);
    std::vector<int> X;
    for (const lbm_solver *m : g->members) X.push_back(m->cfg.dims[0]);
    // chains that a work-group of either launch holds (beta at wg, alpha at wg_alpha)
    g->chains = find_chains(g->xfuse, X, alpha_split(g) ? std::min(g->wg, g->wg_alpha) : g->wg, group_pairs(g));
    if (g->fuse_auto && elem_size(g->storage) == 2) {
        // two-byte storages: the chain tiles' handoff costs their kernels more
        // than the copy launch does; the seam wavefront of a ring of one with
        // packed pairs does not
        std::vector<std::vector<int>> kept;
        for (auto &ch : g->chains) {
            if (ch.size() == 1 && g->xfuse[ch[0]].nb[0] == ch[0] && seam_ring_ok(group_pairs(g), X[ch[0]])) {
                kept.push_back(ch);
                continue;
            }
            for (int m : ch) g->xfuse[m] = XJoin{{-1, -1}, {0u, 0u}};
        }
        g->chains.swap(kept);
    }
}

static int build_plan(lbm_group *g) {
    std::vector<CopyOp> ops[2][4];
    std::vector<HOp> hops[2][4];
    std::vector<const GroupComm *> partner(g->comms.size(), nullptr);
    std::vector<char> fuse(g->comms.size(), 0);
    for (size_t ci = 0; ci < g->comms.size(); ++ci) {
        const GroupComm &c = g->comms[ci];
        for (const GroupComm &q : g->comms)
            if (q.member == c.dst && q.dst == c.member && q.normal[0] == -c.normal[0] && q.normal[1] == -c.normal[1] &&
                q.normal[2] == -c.normal[2]) {
                partner[ci] = &q;
                break;
            }
        if (!partner[ci])
            return gfail(g, LBM_ERR_STATE,
                         "comm of member " + std::to_string(c.member) + " toward " + std::to_string(c.dst) +
                             " has no counterpart");
        fuse[ci] = fusable_x(g, c, *partner[ci]);
    }
    plan_chains(g, fuse);
    for (size_t ci = 0; ci < g->comms.size(); ++ci) {
        const GroupComm &c = g->comms[ci];
        const GroupComm *p = partner[ci];
        const int axis = c.normal[0] ? 0 : c.normal[1] ? 1 : 2;
        lbm_solver *me = g->members[c.member], *nb = g->members[c.dst];
        std::string why;
        const uint32_t amask = alpha_mask(g, c);
        const bool fused = fuse[ci] && g->xfuse[c.member].nb[c.normal[0] == 1 ? 0 : 1] == c.dst;
        const int slot = fused ? 3 : axis;
        HOp a{c.dst, c.member, {}, {}, {}, amask};
        // after beta: neighbour's ghost layer -> my inner layer, n.e > 0 only (src/CController.hpp:322-383)
        HOp b{c.dst, c.member, {}, {}, {}, norm_mask(c.normal)};
        for (int ax = 0; ax < 3; ++ax) {
            a.so[ax] = p->send_origin[ax];
            a.dof[ax] = c.recv_origin[ax];
            b.so[ax] = p->recv_origin[ax];
            b.dof[ax] = c.send_origin[ax];
            a.sz[ax] = b.sz[ax] = c.size[ax];
        }
        // each box is checked against both subdomains before any merging
        CopyOp chk;
        int rc = fill_op(nb, a.so, me, a.dof, a.sz, a.mask, chk, why);
        if (!rc) rc = fill_op(nb, b.so, me, b.dof, b.sz, b.mask, chk, why);
        if (rc) return gfail(g, rc, why);
        hops[0][slot].push_back(a);
        hops[1][slot].push_back(b);
    }
    // One copy launch per half-step instead of one per axis where the order
    // resolves on the host.  Not with frozen members: a frozen receiver skips
    // its copies, which the resolved plan would read through.
    bool frozen = false;
    for (const lbm_solver *m : g->members) frozen |= m->stop_at != std::numeric_limits<long long>::max();
    for (int k = 0; k < 2 && !frozen; ++k) {
        std::vector<HOp> merged;
        if (!merge_phases(hops[k], merged)) continue;
        hops[k][0].swap(merged);
        hops[k][1].clear();
        hops[k][2].clear();
    }
    for (int k = 0; k < 2; ++k)
        for (int q = 0; q < 4; ++q)
            for (const HOp &h : hops[k][q]) {
                if (!h.mask) continue;
                CopyOp op;
                std::string why;
                const int rc = fill_op(g->members[h.sm], h.so, g->members[h.dm], h.dof, h.sz, h.mask, op, why);
                if (rc) return gfail(g, rc, why);
                ops[k][q].push_back(op);
            }
    // each table: the ops, then the op index of every block (exchange_kernel)
    std::vector<int> block_op[2][4];
    for (int k = 0; k < 2; ++k)
        for (int p = 0; p < 4; ++p) {
            int blocks = 0;
            for (size_t i = 0; i < ops[k][p].size(); ++i) {
                CopyOp &op = ops[k][p][i];
                op.block_begin = blocks;
                const int nb = (op.cells / op.vec * op.npop + 255) / 256;
                block_op[k][p].insert(block_op[k][p].end(), (size_t)nb, (int)i);
                blocks += nb;
            }
            g->nops[k][p] = (int)ops[k][p].size();
            g->nblocks[k][p] = blocks;
            if (ops[k][p].empty()) continue;
            const size_t n = ops[k][p].size();
            const size_t extra = (sizeof(int) * (size_t)blocks + sizeof(CopyOp) - 1) / sizeof(CopyOp);
            GCHK(grow_table(g->d_ops[k][p], g->ops_cap[k][p], n + extra));
            GCHK(hipMemcpyAsync(g->d_ops[k][p], ops[k][p].data(), sizeof(CopyOp) * n, hipMemcpyHostToDevice,
                                g->stream));
            GCHK(hipMemcpyAsync(g->d_ops[k][p] + n, block_op[k][p].data(), sizeof(int) * (size_t)blocks,
                                hipMemcpyHostToDevice, g->stream));
        }
    GCHK(hipStreamSynchronize(g->stream));
    g->plan_built = true;
    g->descs_dirty = true;  // the descriptors carry the fused faces
    return LBM_OK;
}

// Brings descriptors and the exchange plan up to date before launches.
static int group_prepare(lbm_group *g) {
    if (!g->plan_built) {
        int rc = build_plan(g);
        if (rc) return rc;
    }
    bool stale = g->descs_dirty;
    for (size_t i = 0; i < g->members.size(); ++i) stale |= g->seen_versions[i] != g->members[i]->version;
    if (stale) {
        int rc = group_upload_descs(g);
        if (rc) return rc;
    }
    return LBM_OK;
}

// Everything a captured graph bakes into its launches: sizes and the device
// tables' addresses (a table that grew was reallocated; its contents are read
// at replay time and may change in place).
static std::vector<long long> group_shape(const lbm_group *g) {
    std::vector<long long> v = {g->wg, g->wg_alpha, g->variant, g->lds, (long long)(uintptr_t)g->d_ctr};
    for (const LaunchTables &T : g->lt) {
        v.push_back(T.ntiles);
        v.push_back((long long)(uintptr_t)T.d_descs);
        v.push_back((long long)(uintptr_t)T.d_tiles);
        v.push_back((long long)(uintptr_t)T.d_chains);
    }
    for (int k = 0; k < 2; ++k)
        for (int p = 0; p < 4; ++p) {
            v.push_back(g->nops[k][p]);
            v.push_back(g->nblocks[k][p]);
            v.push_back((long long)(uintptr_t)g->d_ops[k][p]);
        }
    return v;
}

// A start/stop event pair for the timing of one launch, owned by g->events.
static hipError_t group_event_pair(lbm_group *g, int kind, hipEvent_t *a, hipEvent_t *b) {
    hipError_t e = hipEventCreate(a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(*a);
        return e;
    }
    g->events.push_back({*a, *b, kind});
    return hipSuccess;
}

static int group_launch_half(lbm_group *g, bool beta, bool timed) {
    if (g->lt[0].ntiles == 0) return LBM_OK;
    const bool split = !beta && alpha_split(g);
    const LaunchTables &T = g->lt[split ? 1 : 0];
    // timing events belong to g->events from creation on, so every exit
    // path leaves them to drain_events / lbm_group_destroy
    hipEvent_t ea = nullptr, eb = nullptr;
    if (timed) {
        GCHK(group_event_pair(g, beta ? 0 : 2, &ea, &eb));
        GCHK(hipEventRecord(ea, g->stream));
    }
    GCHK(launch_step_kernel(g->storage, beta, g->variant, g->lds, T.ntiles, split ? g->wg_alpha : g->wg, T.d_descs,
                            T.d_tiles, g->stream, g->d_ctr, T.d_chains, T.xcd_chunk));
    if (timed) GCHK(hipEventRecord(eb, g->stream));
    // the kernel just run was beta -> syncBeta, alpha -> syncAlpha (src/CController.hpp:385-391)
    const int k = beta ? 1 : 0;
    for (int p = 0; p < 3; ++p) {
        if (!g->nops[k][p]) continue;
        hipEvent_t xa = nullptr, xb = nullptr;
        if (timed) {
            GCHK(group_event_pair(g, 1, &xa, &xb));
            GCHK(hipEventRecord(xa, g->stream));
        }
        GCHK(exchange_launch(g->storage, g->d_ops[k][p], g->nops[k][p], g->nblocks[k][p], g->stream,
                             g->d_ctr + (beta ? 1 : 0)));
        if (timed) GCHK(hipEventRecord(xb, g->stream));
    }
    return LBM_OK;
}

// Advances every member's own step counter by the half-steps it took part in
// (a member frozen at its stop keeps its count).
static void group_advance_counters(lbm_group *g, long long before, long long done) {
    for (auto *m : g->members) {
        const long long room = m->stop_at - before;
        m->step_counter += room <= 0 ? 0 : std::min(done, room);
    }
    g->step_counter = before + done;
}

// The graph replaying `pairs` beta+alpha pairs (with their exchanges):
// captured once per launch shape and pair count, at most kMaxGraphs kept
// (least recently used dropped first).  Capturing records launches without
// running them.
static constexpr size_t kMaxGraphs = 4;
static int group_graph(lbm_group *g, int pairs, hipGraphExec_t *out) {
    if (!g->graphs.empty() && g->graph_shape != group_shape(g)) group_drop_graph(g);
    for (auto &e : g->graphs)
        if (e.pairs == pairs) {
            e.used = ++g->graph_clock;
            *out = e.exec;
            return LBM_OK;
        }
    if (g->graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(g->graphs.begin(), g->graphs.end(),
                                    [](const lbm_group::Graph &a, const lbm_group::Graph &b) { return a.used < b.used; });
        (void)hipGraphExecDestroy(lru->exec);
        (void)hipGraphDestroy(lru->graph);
        g->graphs.erase(lru);
    }
    GCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = LBM_OK;
    for (int p = 0; p < pairs && !rc; ++p) {
        rc = group_launch_half(g, true, false);
        if (!rc) rc = group_launch_half(g, false, false);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
    if (rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        GCHK(ec);
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        GCHK(ei);
    }
    if (g->graphs.empty()) g->graph_shape = group_shape(g);
    g->graphs.push_back({pairs, graph, exec, ++g->graph_clock});
    g->graphs_captured += 1;
    *out = exec;
    return LBM_OK;
}

// Walks `steps` half-steps from the group's counter: graph replays of up to
// graph_pairs pairs where a pair starts on a beta step, single eager launches
// otherwise.  With run == false it only captures the graphs the walk needs.
static int group_walk(lbm_group *g, int steps, bool run, int *done_out) {
    long long h = g->step_counter;
    int done = 0;
    while (done < steps) {
        const bool beta = (h & 1) == 0;
        const int pairs = (steps - done) / 2;
        if (g->use_graph && !g->timing && beta && pairs >= 1) {
            const int P = std::min(pairs, g->graph_pairs);
            hipGraphExec_t exec = nullptr;
            int rc = group_graph(g, P, &exec);
            if (rc) return rc;
            if (run) {
                GCHK(hipGraphLaunch(exec, g->stream));
                g->graph_launches += 1;
            }
            h += 2 * P;
            done += 2 * P;
        } else {
            if (run) {
                int rc = group_launch_half(g, beta, g->timing);
                if (rc) return rc;
                g->eager_half_steps += 1;
            }
            h += 1;
            done += 1;
        }
    }
    *done_out = done;
    return LBM_OK;
}

extern "C" int lbm_group_step(lbm_group *g, int steps) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (steps < 0) return gfail(g, LBM_ERR_INVALID_ARGUMENT, "steps must be >= 0");
    GCHK(hipSetDevice(g->device));
    int rc = group_prepare(g);
    if (rc) return rc;
    const long long before = g->step_counter;
    if (g->lt[0].ntiles == 0) {  // no members: counters only (re-published with the next upload)
        g->descs_dirty = true;
        group_advance_counters(g, before, steps);
        return LBM_OK;
    }
    int done = 0;
    rc = group_walk(g, steps, true, &done);
    if (rc) return rc;
    group_advance_counters(g, before, done);
    return LBM_OK;
}

extern "C" int lbm_group_prepare_steps(lbm_group *g, int steps) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (steps < 0) return gfail(g, LBM_ERR_INVALID_ARGUMENT, "steps must be >= 0");
    GCHK(hipSetDevice(g->device));
    int rc = group_prepare(g);
    if (rc || g->lt[0].ntiles == 0) return rc;
    int done = 0;
    return group_walk(g, steps, false, &done);
}

extern "C" int lbm_group_get_launch_counts(const lbm_group *g, long long *graph_launches, long long *eager_half_steps,
                                           long long *graphs_captured) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (graph_launches) *graph_launches = g->graph_launches;
    if (eager_half_steps) *eager_half_steps = g->eager_half_steps;
    if (graphs_captured) *graphs_captured = g->graphs_captured;
    return LBM_OK;
}

extern "C" int lbm_group_set_graph_pairs(lbm_group *g, int pairs) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (pairs < 1 || pairs > 4096) return gfail(g, LBM_ERR_INVALID_ARGUMENT, "graph pairs must be in [1, 4096]");
    g->graph_pairs = pairs;
    return LBM_OK;
}

extern "C" int lbm_group_set_launch(lbm_group *g, int work_group_size, int kernel_variant) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (!lbm_work_group_supported(work_group_size) || !lbm_variant_supported(kernel_variant))
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "unsupported launch shape");
    GCHK(hipStreamSynchronize(g->stream));
    g->wg = work_group_size;
    g->variant = kernel_variant;
    group_drop_graph(g);
    if (g->fuse_x) g->plan_built = false;  // which chains a work-group holds depends on its size
    g->descs_dirty = true;
    return g->plan_built ? group_upload_descs(g) : LBM_OK;
}

extern "C" int lbm_group_get_launch(const lbm_group *g, int *work_group_size, int *kernel_variant) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (work_group_size) *work_group_size = g->wg;
    if (kernel_variant) *kernel_variant = g->variant;
    return LBM_OK;
}

// Alpha's launches at their own work-group size (0: the group's).  Chain
// tiles then hold as many rows per member as alpha's work-group takes, while
// beta keeps the group's; the plan keeps the chains that both sizes hold.
extern "C" int lbm_group_set_alpha_work_group(lbm_group *g, int work_group_size) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    if (work_group_size != 0 && !lbm_work_group_supported(work_group_size))
        return gfail(g, LBM_ERR_INVALID_ARGUMENT, "unsupported alpha work-group size");
    GCHK(hipStreamSynchronize(g->stream));
    g->wg_alpha = work_group_size;
    group_drop_graph(g);
    if (g->fuse_x) g->plan_built = false;  // the chains kept depend on both sizes
    g->descs_dirty = true;
    return g->plan_built ? group_upload_descs(g) : LBM_OK;
}

extern "C" int lbm_group_get_alpha_work_group(const lbm_group *g, int *work_group_size) {
    if (!g || !work_group_size) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null argument");
    *work_group_size = g->wg_alpha ? g->wg_alpha : g->wg;
    return LBM_OK;
}

// Autotuner: times `pairs` beta+alpha step pairs (with their exchanges) for
// every work-group size in {64, 128, 256, 512} and every variant in {0, 3, 8, 11}
// (1 or 2 cells per lane, each with the default or the non-temporal cache
// policy) and keeps the fastest; then, with that shape for beta, times
// alpha at each other work-group size (which sets how many rows of each
// member an alpha chain tile holds) and keeps alpha's own size if it is
// faster; last, it times the chosen shape with the two other XCD tile maps
// (one contiguous eighth of the tiles per XCD, chunks of consecutive tiles
// dealt in turn, or none: variant bits 32 and 4) and keeps the fastest.  The steps it runs are
// ordinary simulation steps: every launch shape computes the same
// arithmetic, so the state is the one the same number of steps gives with
// any shape.
extern "C" int lbm_group_autotune(lbm_group *g, int pairs, int *best_wg, int *best_variant, double *best_ms_per_step) {
    if (!g || pairs <= 0) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad autotune arguments");
    GCHK(hipSetDevice(g->device));
    if ((g->step_counter & 1) != 0) {
        int rc = lbm_group_step(g, 1);  // start the pairs on a beta step
        if (rc) return rc;
    }
    {
        int rc = group_prepare(g);
        if (rc) return rc;
    }
    if (g->lt[0].ntiles == 0) return gfail(g, LBM_ERR_STATE, "autotune needs at least one member");
    const int wgs[4] = {64, 128, 256, 512};
    const int vars[4] = {0, 3, 8, 11};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    struct Events {
        hipEvent_t *a, *b;
        ~Events() {
            if (*a) (void)hipEventDestroy(*a);
            if (*b) (void)hipEventDestroy(*b);
        }
    } events_guard{&e0, &e1};
    GCHK(hipEventCreate(&e0));
    GCHK(hipEventCreate(&e1));
    // ms per step of the group's current launch shape: a warm-up pair, then
    // `pairs` timed pairs
    auto time_shape = [&](double *per) -> int {
        int rc = group_prepare(g);
        if (!rc) rc = group_launch_half(g, true, false);
        if (!rc) rc = group_launch_half(g, false, false);
        if (rc) return rc;
        GCHK(hipEventRecord(e0, g->stream));
        for (int p = 0; p < pairs && !rc; ++p) {
            rc = group_launch_half(g, true, false);
            if (!rc) rc = group_launch_half(g, false, false);
        }
        if (rc) return rc;
        GCHK(hipEventRecord(e1, g->stream));
        GCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        GCHK(hipEventElapsedTime(&ms, e0, e1));
        group_advance_counters(g, g->step_counter, 2 * (pairs + 1));
        *per = (double)ms / (2.0 * pairs);
        return LBM_OK;
    };
    double best = 1e300;
    const int map = lbm_variant_map_bits(g->variant);
    int bw = g->wg, bv = g->variant, ba = 0;
    int rc = lbm_group_set_alpha_work_group(g, 0);  // stage 1: one size for both kernels
    if (rc) return rc;
    for (int wi = 0; wi < 4; ++wi)
        for (int vi = 0; vi < 4; ++vi) {
            double per = 0.0;
            rc = lbm_group_set_launch(g, wgs[wi], vars[vi] | map);
            if (!rc) rc = time_shape(&per);
            if (rc) return rc;
            if (per < best) {
                best = per;
                bw = wgs[wi];
                bv = vars[vi] | map;
            }
        }
    rc = lbm_group_set_launch(g, bw, bv);
    if (rc) return rc;
    for (int wi = 0; wi < 4; ++wi) {  // stage 2: alpha at a size of its own
        if (wgs[wi] == bw) continue;
        double per = 0.0;
        rc = lbm_group_set_alpha_work_group(g, wgs[wi]);
        if (!rc) rc = time_shape(&per);
        if (rc) return rc;
        if (per < best) {
            best = per;
            ba = wgs[wi];
        }
    }
    rc = lbm_group_set_alpha_work_group(g, ba);
    if (rc) return rc;
    {  // stage 3: the two other XCD tile maps
        const int maps[3] = {0, kVarXcdChunks, kVarNoXcdMap};
        const int base = bv & ~(kVarNoXcdMap | kVarXcdChunks);
        int keep = bv;
        for (int mi = 0; mi < 3; ++mi) {
            const int v = base | maps[mi];
            if (v == bv) continue;
            double per = 0.0;
            rc = lbm_group_set_launch(g, bw, v);
            if (!rc) rc = time_shape(&per);
            if (rc) return rc;
            if (per < best) {
                best = per;
                keep = v;
            }
        }
        bv = keep;
    }
    rc = lbm_group_set_launch(g, bw, bv);
    if (!rc) rc = group_prepare(g);
    if (rc) return rc;
    if (best_wg) *best_wg = bw;
    if (best_variant) *best_variant = bv;
    if (best_ms_per_step) *best_ms_per_step = best;
    return LBM_OK;
}

extern "C" int lbm_group_sync(lbm_group *g, int after_beta) {
    if (!g || (after_beta != 0 && after_beta != 1)) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "bad argument");
    GCHK(hipSetDevice(g->device));
    {
        int rc = group_prepare(g);
        if (rc) return rc;
    }
    // fused x copies first: they belong to the x phase, disjoint from the rest of it
    const int order[4] = {3, 0, 1, 2};
    for (int q = 0; q < 4; ++q) {
        const int p = order[q];
        if (!g->nops[after_beta][p]) continue;
        GCHK(exchange_launch(g->storage, g->d_ops[after_beta][p], g->nops[after_beta][p], g->nblocks[after_beta][p],
                             g->stream));
    }
    return LBM_OK;
}

extern "C" int lbm_group_wait(lbm_group *g) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    GCHK(hipStreamSynchronize(g->stream));
    return LBM_OK;
}

extern "C" int lbm_group_set_timing(lbm_group *g, int enable) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    GCHK(hipStreamSynchronize(g->stream));
    for (auto &e : g->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    g->events.clear();
    for (int k = 0; k < 3; ++k) {
        g->acc_ms[k] = 0;
        g->acc_n[k] = 0;
    }
    g->timing = enable != 0;
    return LBM_OK;
}

// Folds the recorded launch times into the averages and destroys every event
// pair, also when one of them cannot be read (a launch that failed between
// its two records); the first such error is reported.
static int drain_events(lbm_group *g) {
    const hipError_t sync = hipStreamSynchronize(g->stream);
    hipError_t first = sync;
    for (auto &e : g->events) {
        float ms = 0.f;
        const hipError_t te = sync == hipSuccess ? hipEventElapsedTime(&ms, e.a, e.b) : sync;
        if (te == hipSuccess) {
            g->acc_ms[e.kind] += ms;
            g->acc_n[e.kind] += 1;
        } else if (first == hipSuccess) {
            first = te;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    g->events.clear();
    if (first != hipSuccess) return gfail(g, LBM_ERR_DEVICE, std::string("launch timing: ") + hipGetErrorString(first));
    return LBM_OK;
}

extern "C" int lbm_group_get_timing_detail(lbm_group *g, double *beta_ms, double *alpha_ms, double *exchange_ms) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    int rc = drain_events(g);
    if (rc) return rc;
    if (beta_ms) *beta_ms = g->acc_n[0] ? g->acc_ms[0] / g->acc_n[0] : 0.0;
    if (alpha_ms) *alpha_ms = g->acc_n[2] ? g->acc_ms[2] / g->acc_n[2] : 0.0;
    if (exchange_ms) *exchange_ms = g->acc_n[1] ? g->acc_ms[1] / g->acc_n[1] : 0.0;
    return LBM_OK;
}

extern "C" int lbm_group_get_timing(lbm_group *g, double *step_ms, long long *step_n, double *xch_ms, long long *xch_n) {
    if (!g) return record(nullptr, LBM_ERR_INVALID_ARGUMENT, "null group");
    int rc = drain_events(g);
    if (rc) return rc;
    const long long ns = g->acc_n[0] + g->acc_n[2];
    if (step_ms) *step_ms = ns ? (g->acc_ms[0] + g->acc_ms[2]) / ns : 0.0;
    if (step_n) *step_n = ns;
    if (xch_ms) *xch_ms = g->acc_n[1] ? g->acc_ms[1] / g->acc_n[1] : 0.0;
    if (xch_n) *xch_n = g->acc_n[1];
    return LBM_OK;
}
