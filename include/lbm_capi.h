/*
 * lbm_capi.h -- C-ABI of the MI355X-native D3Q19 lattice-Boltzmann hot path.
 *
 * The entry points replace the device-facing surface of the reference's
 * CLbmSolver<T> (src/CLbmSolver.hpp:67-1124) and the halo exchange of
 * CController (src/CController.hpp:265-391).  Plain pointers and sizes only;
 * every call returns an LBM_* status, and lbm_solver_error()/lbm_last_error()
 * give the message (the reference accumulates the same text in a CError,
 * src/lib/CError.hpp:42-80).  Host buffers are caller-owned; set* copies at
 * call time and store* blocks until the data is on the host, as in the
 * reference.  A solver, group or batch is not thread-safe (one host thread
 * per object, as the reference's one command queue per solver).
 *
 * Layout: distributions are 19 planes, plane i holding population i for all
 * cells in x-fastest order (x + X*(y + Y*z)); populations are numbered as in
 * src/cl_programs/lbm_header.h:12-19.  Region buffers are dense
 * [component][z][y][x] of the region size (src/CLbmSolver.hpp:695-957).
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_CAPI_VERSION 1

/* status codes */
#define LBM_OK 0
#define LBM_ERR_INVALID_ARGUMENT 1
#define LBM_ERR_OUT_OF_RANGE 2
#define LBM_ERR_DEVICE 3
#define LBM_ERR_OUT_OF_MEMORY 4
#define LBM_ERR_PARAMETRIZATION 5
#define LBM_ERR_DECOMPOSITION 6
#define LBM_ERR_STATE 7
#define LBM_ERR_CAPACITY 8

/* cell flags: src/common.h:19-22, src/CLbmSolver.hpp:30-33 */
#define LBM_FLAG_OBSTACLE 1
#define LBM_FLAG_FLUID 2
#define LBM_FLAG_VELOCITY_INJECTION 4
#define LBM_FLAG_GHOST_LAYER 8

/* storage of the distributions (compute type: f32 except LBM_STORAGE_F64).
 * FP32 and FP64 evaluate the reference's expressions rounding by rounding
 * (no fused multiply-adds); the two-byte storages' step kernels fuse them
 * (FP contraction), which changes their last bits, not their error bounds. */
#define LBM_STORAGE_F32 0
#define LBM_STORAGE_F64 1
#define LBM_STORAGE_F16_SHIFTED 2  /* stores f_i - w_i as IEEE half */
#define LBM_STORAGE_BF16_SHIFTED 3 /* stores f_i - w_i as bfloat16 */

/* forcing model */
#define LBM_FORCING_REFERENCE 0 /* the alpha kernel's term chain, src/cl_programs/lbm_alpha.cl:173-298 */
#define LBM_FORCING_BODY 1      /* f_i += 3 w_i e_i.(rho g) after collision, both kernels */

/* fields of the region get/set calls */
#define LBM_FIELD_DD 0       /* 19 components, compute type */
#define LBM_FIELD_VELOCITY 1 /* 3 components, compute type */
#define LBM_FIELD_DENSITY 2  /* 1 component, compute type */
#define LBM_FIELD_FLAGS 3    /* 1 component, int32 */
#define LBM_FIELD_DD_RAW 4   /* 19 components, the stored bit patterns (f32/f64, or f - w as f16/bf16):
                                an exact checkpoint of any storage (new) */

typedef struct lbm_solver lbm_solver;
typedef struct lbm_group lbm_group;

/* Construction parameters of one (sub)domain; lattice units throughout.
 * Replaces the argument list of CLbmSolver::CLbmSolver
 * (src/CLbmSolver.hpp:221-260) after CLbmSkeleton's parametrization
 * (src/CLbmSkeleton.hpp:81-190), which the C++/Python host layers perform. */
typedef struct lbm_solver_config {
    int dims[3];            /* cells per axis, boundary and ghost layers included */
    int bc[6];              /* flag of the x-, x+, y-, y+, z-, z+ faces (BC[3][2]) */
    int storage;            /* LBM_STORAGE_* */
    int forcing;            /* LBM_FORCING_* */
    double inv_tau;         /* 1/tau of the BGK collision */
    double tau;             /* tau0 of the Smagorinsky rate */
    double smagorinsky_cs;  /* 0 gives plain BGK, bit for bit */
    double gravitation[3];  /* lattice body force / gravitation */
    double lid_velocity;    /* lattice lid speed (x) of VELOCITY_INJECTION cells */
    int store_velocity;     /* kernels write the velocity field (STORE_VELOCITY) */
    int store_density;      /* kernels write the density field (STORE_DENSITY) */
    int work_group_size;    /* threads per work-group of the step kernels: 64..512, whole wavefronts; 0 = 256 */
    int device;             /* HIP device ordinal */
    int kernel_variant;     /* tuning bits: 1 non-temporal loads, 2 non-temporal stores, 4 no XCD tile map,
                               8 / 16 two / four x-consecutive cells per work-item, 32 the XCD tile map in
                               chunks of about 16 consecutive tiles per XCD in turn, matched to the tiles
                               of a z plane (instead of one contiguous eighth of the tiles per XCD); supported: 0-3, 8, 11 (two cells with the
                               non-temporal policy; two-byte storages keep the pair packed in one 32-bit
                               word) and 16, each optionally | 4 or | 32 (not both) */
    int lds_per_block;      /* bytes of LDS reserved per work-group (caps residency); 0 = none */
} lbm_solver_config;

/* ---- library ---------------------------------------------------------- */
const char *lbm_version(void);
/* message of the last failed call on this thread (creation failures included) */
const char *lbm_last_error(void);
int lbm_device_synchronize(int device);
void lbm_solver_config_default(lbm_solver_config *cfg);
/* read+write bandwidth of a plain device copy of `bytes`, averaged over `reps`
 * launches (variant 0: grid-stride float4, 1: four float4 per work-item,
 * 2: as 1 with non-temporal loads/stores): the achievable-HBM reference of the
 * roofline and the known-size kernel that calibrates the FETCH_SIZE and
 * WRITE_SIZE counters (new; no reference counterpart) */
int lbm_device_copy_bandwidth(int device, size_t bytes, int reps, int variant, double *gbs, double *kernel_ms);

/* ---- solver: CLbmSolver<T> (src/CLbmSolver.hpp:67-1124) ----------------- */
/* constructor + reload() + reset() (src/CLbmSolver.hpp:221-260, 272-635) */
int lbm_solver_create(const lbm_solver_config *cfg, lbm_solver **out);
int lbm_solver_destroy(lbm_solver *s);
/* device bytes of a solver built by lbm_solver_create_in: distributions,
 * flags, optional velocity/density, launch tables and a staging area of
 * max(1 byte per cell, 19 compute values per cell of the largest face)
 * (new; the reference allocates per buffer, src/CLbmSolver.hpp:381-403) */
int lbm_solver_required_bytes(const lbm_solver_config *cfg, size_t *bytes);
/* constructor over caller-owned device memory (256-byte aligned, at least
 * lbm_solver_required_bytes): no device allocation and no free on destroy,
 * so a resident case set can live in one arena.  `stream` is the hipStream_t
 * the solver works on, NULL for a stream of its own.  Region transfers larger
 * than the staging area fail with LBM_ERR_CAPACITY (new) */
int lbm_solver_create_in(const lbm_solver_config *cfg, void *device_memory, size_t bytes, void *stream,
                         lbm_solver **out);
/* moves the device buffers of n solvers (created by lbm_solver_create, not yet
 * in a group) into one device allocation, side by side, contents unchanged;
 * each solver is still destroyed on its own and the allocation goes with the
 * last of them.  Where buffers land in HBM changes the step time by a few
 * percent either way depending on the machine, so this is a placement
 * choice, not a default (new; the reference allocates per buffer and per
 * rank, src/CLbmSolver.hpp:381-403).  Needs the new block while the old
 * buffers are still held. */
int lbm_solver_colocate(lbm_solver *const *solvers, int n);
/* CError text of this solver (src/lib/CError.hpp:42-80) */
const char *lbm_solver_error(const lbm_solver *s);
/* reset(): init kernel, step counter 0 (src/CLbmSolver.hpp:619-635) */
int lbm_solver_reset(lbm_solver *s);
/* reset() with f = f_eq(1, du) on every cell off OBSTACLE faces, du drawn per
 * component from a counter-based generator keyed by (seed, case_id, cell of
 * the key grid); state written in the swapped layout (new; no reference) */
int lbm_solver_init_perturbed(lbm_solver *s, double amplitude, uint64_t seed, uint64_t case_id,
                              const int key_origin[3], const int key_dims[3]);
/* simulationStep(): beta when the counter is even, alpha when odd
 * (src/CLbmSolver.hpp:664-676); Alpha/Beta (src/CLbmSolver.hpp:637-659) */
int lbm_solver_simulation_step(lbm_solver *s);
int lbm_solver_simulation_step_alpha(lbm_solver *s);
int lbm_solver_simulation_step_beta(lbm_solver *s);
int lbm_solver_run(lbm_solver *s, int steps);
/* wait(): clFinish (src/CLbmSolver.hpp:681-683) */
int lbm_solver_wait(lbm_solver *s);
int lbm_solver_get_step_counter(const lbm_solver *s, long long *out);
int lbm_solver_set_step_counter(lbm_solver *s, long long counter);
int lbm_solver_get_config(const lbm_solver *s, lbm_solver_config *out);
/* kernel argument update of addDrivenCavityValue (src/CLbmSolver.hpp:262-270) */
int lbm_solver_set_lid_velocity(lbm_solver *s, double lattice_lid);
int lbm_solver_set_relaxation(lbm_solver *s, double inv_tau, double tau, double smagorinsky_cs);
/* storeDensityDistribution / storeVelocity / storeDensity / storeFlags, full
 * (origin == NULL) or region (src/CLbmSolver.hpp:688-717, 764-795, 827-870, 907-945) */
int lbm_solver_store(lbm_solver *s, int field, const int origin[3], const int size[3], void *host_dst);
/* setDensityDistribution (with optional norm filter), setVelocity,
 * setDensity, setFlags (src/CLbmSolver.hpp:719-757, 804-821, 879-905, 947-978);
 * norm != NULL copies only populations with norm . e_f > 0 (LBM_FIELD_DD) */
int lbm_solver_set(lbm_solver *s, int field, const int origin[3], const int size[3], const void *host_src,
                   const int norm[3]);
/* getVelocityChecksum (src/CLbmSolver.hpp:1103-1123) */
int lbm_solver_velocity_checksum(lbm_solver *s, float *out);
/* FP64 reductions over FLUID cells of rho and kinetic energy 1/2 |j|^2/rho,
 * moments taken from the distributions in the current step parity (new) */
int lbm_solver_mass_energy(lbm_solver *s, double *mass, double *kinetic_energy);
/* lbm_solver_mass_energy without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and summation order,
 * hence the same bits.  Its partial sums live in a fixed per-device scratch
 * array of 2 MiB, resolved once per device and held until the process ends:
 * solvers of more than 33,554,432 cells (2^25) are rejected with
 * LBM_ERR_CAPACITY.  Every call on one device shares that array, so calls
 * that overlap on one device (from several host threads) race; calls from
 * one thread do not, since each returns after its result is on the host. */
int lbm_solver_mass_energy_noalloc(lbm_solver *s, double *mass, double *kinetic_energy);
/* rho and j = sum f_i e_i of a region, from the distributions in the current
 * step parity, in the compute type: rho[size], j[3][size] (new) */
int lbm_solver_moments(lbm_solver *s, const int origin[3], const int size[3], void *host_rho, void *host_j);
/* device-to-device box copy of distributions between two solvers of the same
 * storage (one CComm transfer without the host round trip of
 * src/CController.hpp:265-383); norm == NULL copies all 19 populations */
int lbm_copy_box(lbm_solver *src, const int src_origin[3], lbm_solver *dst, const int dst_origin[3],
                 const int size[3], const int norm[3]);
/* seeded spherical obstacles (new): spheres with centres uniform in the key
 * grid's interior and radii in [rmin, rmax] are drawn until they cover `fill`
 * of it; FLUID cells inside become OBSTACLE.  achieved = covered fraction. */
int lbm_solver_place_spheres(lbm_solver *s, uint64_t seed, double fill, double rmin, double rmax,
                             const int key_origin[3], const int key_dims[3], double *achieved);
/* raw device pointers (for callers that bring their own kernels).  The
 * distributions are stored in blocks of `block_cells` consecutive cells (cell
 * order x + X*(y + Y*z)), each block holding its 19 populations one after the
 * other: population i of cell g is element
 *   (g / block_cells) * 19 * block_cells + i * block_cells + g % block_cells
 * (the reference's [19][cells] planes, src/CLbmSolver.hpp:381-382, blocked);
 * flags are one byte per cell. */
int lbm_solver_device_pointers(lbm_solver *s, void **dd, void **flags, long long *block_cells);

/* ---- group: N subdomains stepped together on one device ----------------- */
/* One fused step kernel over all member solvers per half-step, then the
 * ghost-layer exchange of CController::syncAlpha/syncBeta
 * (src/CController.hpp:265-391) as on-device copies, x faces, then y, then z. */
int lbm_group_create(lbm_solver *const *solvers, int n, lbm_group **out);
int lbm_group_destroy(lbm_group *g);
const char *lbm_group_error(const lbm_group *g);
/* CController::addCommunication(new CComm(...)) (src/CController.hpp:524-526,
 * src/CComm.hpp:8-79) for member `member`: dst_member is the neighbour */
int lbm_group_add_comm(lbm_group *g, int member, int dst_member, const int send_origin[3],
                       const int recv_origin[3], const int size[3], const int normal[3]);
/* exchange mode: 0 = all 19 populations after alpha (reference),
 * 1 = only the 5 populations an inner cell reads from a ghost */
int lbm_group_set_exchange_mode(lbm_group *g, int mode);
/* x faces exchanged inside the step kernel (new): equally sized members
 * joined across x with the reference's face geometry form chains (at most 8,
 * lines or periodic rings); one work-group holds the same rows of every
 * member of a chain and hands the face values over in LDS, so the x-face copy
 * launch disappears (the x faces are the strided ones: one element per cache
 * line).  A periodic ring of one member in a two-byte storage launched with
 * packed pairs (kernel variants 8, 11) and rows of an even length >= 128
 * instead keeps each row's seam in one wavefront (cells 0..63 and X-64..X-1)
 * and moves the face values by lane reads.  Every slot of every member is
 * identical to the copy-based exchange.  Chains whose rows (or, for wide
 * rows, whose 64-cell face runs) do not fit the work-group keep their
 * copies.  lbm_group_sync still copies.  `mode`: 0 = copy launch, 1 = inside
 * the step kernel, LBM_FUSED_X_AUTO = inside the step kernel for FP32/FP64
 * storage; for FP16/BF16 only the seam wavefronts, other x faces by the copy
 * launch (measured: the two-byte kernels lose more to the handoff tiles than
 * the copy launch costs). */
#define LBM_FUSED_X_AUTO 2
int lbm_group_set_fused_x(lbm_group *g, int mode);
/* replay the step pairs as hipGraphs (1) or launch eagerly (0) */
int lbm_group_set_graph(lbm_group *g, int enable);
/* most beta+alpha step pairs one graph launch replays (1..4096, default 64):
 * lbm_group_step(g, n) replays its pairs in ceil(pairs / this) graph launches.
 * A graph is captured once per launch shape and pair count; the last few are
 * kept (new; the reference enqueues every kernel, src/CLbmSolver.hpp:664-676) */
int lbm_group_set_graph_pairs(lbm_group *g, int pairs);
/* captures the graphs a following lbm_group_step(g, steps) from the current
 * step counter replays, without running anything, so that call only launches
 * (new) */
int lbm_group_prepare_steps(lbm_group *g, int steps);
/* how lbm_group_step has launched work so far: graph launches, half-steps
 * launched without a graph, and graphs captured (new) */
int lbm_group_get_launch_counts(const lbm_group *g, long long *graph_launches, long long *eager_half_steps,
                                long long *graphs_captured);
/* computeNextStep() for all members, `steps` times (src/CController.hpp:385-391) */
int lbm_group_step(lbm_group *g, int steps);
/* one exchange on its own: syncAlpha (after_beta = 0) or syncBeta
 * (after_beta = 1) of every member (src/CController.hpp:265-383) */
int lbm_group_sync(lbm_group *g, int after_beta);
int lbm_group_wait(lbm_group *g);
/* average device time per launch of the step kernels and the exchange kernels
 * since the last reset, measured with events on the group's stream when
 * timing is on */
int lbm_group_set_timing(lbm_group *g, int enable);
int lbm_group_get_timing(lbm_group *g, double *step_kernel_ms, long long *step_kernel_launches,
                         double *exchange_ms, long long *exchange_launches);
/* launch shape of the fused step kernel: work-group size (64..512) and
 * kernel variant (see lbm_solver_config.kernel_variant); results do not
 * depend on it */
int lbm_group_set_launch(lbm_group *g, int work_group_size, int kernel_variant);
int lbm_group_get_launch(const lbm_group *g, int *work_group_size, int *kernel_variant);
/* alpha's launches at a work-group size of their own (64..512, 0 = the
 * group's; new): chain tiles of the x-face exchange then hold as many rows
 * per member as alpha's work-group takes (two 128-cell rows per member of a
 * two-member chain at 512) while beta keeps the group's size.  The chains
 * kept are those both sizes hold.  Results do not depend on it.  The
 * autotuner picks it too.  get returns the size alpha launches with. */
int lbm_group_set_alpha_work_group(lbm_group *g, int work_group_size);
int lbm_group_get_alpha_work_group(const lbm_group *g, int *work_group_size);
/* autotuner (new): times `pairs` step pairs for each work-group size in
 * {64,128,256,512} x kernel variant in {0,3,8,11} (1 or 2 cells per lane,
 * each with the default or the non-temporal cache policy) and keeps the
 * fastest; then, with that shape, times alpha at each other work-group size
 * (lbm_group_set_alpha_work_group: the rows per member of its chain tiles)
 * and keeps alpha's own size if it is faster (read it back with
 * lbm_group_get_alpha_work_group; a size set before is replaced); last, the
 * chosen shape with the two other XCD tile maps (eighths, chunks, none:
 * variant bits 32 and 4).  The timed steps are real steps of the simulation
 * (42 * (pairs + 1) of them, plus one if the counter was odd) */
int lbm_group_autotune(lbm_group *g, int pairs, int *best_wg, int *best_variant, double *best_ms_per_step);
/* the same averages split into beta and alpha step kernels */
int lbm_group_get_timing_detail(lbm_group *g, double *beta_ms, double *alpha_ms, double *exchange_ms);
/* Membership of a continuous batch (new; no reference counterpart): a group
 * with no members yet; members join between steps when their step counter has
 * the group's parity and leave with their state in place.  A member given a
 * stop freezes after `steps` more half-steps of the group: the step and
 * exchange kernels skip it from then on (a device-side half-step counter
 * decides, so chunks of many steps run as one graph replay) and its own step
 * counter stops there; steps < 0 removes the stop.  Removing a member
 * renumbers the later ones and drops the comms that name it. */
int lbm_group_create_empty(int device, int storage, int work_group_size, int kernel_variant, lbm_group **out);
int lbm_group_add_member(lbm_group *g, lbm_solver *s, int *index);
int lbm_group_remove_member(lbm_group *g, int index);
int lbm_group_member_count(const lbm_group *g, int *count);
int lbm_group_set_member_stop(lbm_group *g, int index, long long steps);
/* the hipStream_t all members work on while in the group (for events) */
int lbm_group_get_stream(const lbm_group *g, void **stream);

/* ---- server: continuous batching of independent cases (new) ------------- */
/* Cases (cavities, periodic channels) are admitted against a device-memory
 * budget into one arena, stepped together per storage type as members of a
 * group (one fused launch per half-step, graph-replayed chunks), frozen at
 * their own step count and read back while the next chunk runs.  Requests
 * that arrive while others run join at the next chunk boundary.  A case's
 * result does not depend on what it was batched with: the per-cell
 * arithmetic is the same in every launch shape and membership. */
typedef struct lbm_server lbm_server;

#define LBM_CASE_CAVITY 0  /* walls on all faces, lid layer at y = ny-2 moving in x */
#define LBM_CASE_CHANNEL 1 /* periodic in x, walls in y and z, body force g_x, seeded spheres */

#define LBM_CASE_QUEUED 0
#define LBM_CASE_RUNNING 1
#define LBM_CASE_DONE 2
#define LBM_CASE_FAILED 3 /* setup failed: lbm_server_fetch reports why and releases it */

typedef struct lbm_server_config {
    int device;
    size_t memory_budget;  /* device bytes for case state (one arena) */
    int max_active;        /* cases resident at once, 0 = limited by memory only */
    double chunk_ms;       /* target device time of one scheduling chunk */
    int max_chunk_steps;   /* steps per chunk at most (rounded to even) */
    int work_group_size;   /* launch shape of the fused kernels */
    int kernel_variant;
    int autotune;          /* 1: tune the launch shape per storage on its first chunk */
    int backfill;          /* 1: admit later requests that fit when the head of the queue does not */
    long long max_active_cells; /* cells resident at once, 0 = no limit: the queue is then admitted in
                                   order (no backfill past a case that would exceed it) unless nothing
                                   is resident, so large cases run beside few others */
} lbm_server_config;

typedef struct lbm_case {
    int kind;                /* LBM_CASE_* */
    int dims[3];             /* cells per axis including boundary/ghost layers */
    int steps;
    int storage;             /* LBM_STORAGE_*; results are float for every storage */
    double tau;              /* BGK relaxation time (tau0 of the LES rate) */
    double smagorinsky_cs;
    double lid_velocity;     /* cavity lid speed */
    double body_force;       /* channel g_x */
    double obstacle_fill;    /* channel: sphere volume fraction */
    double r_min, r_max;     /* channel: sphere radii */
    double perturbation;     /* amplitude of the seeded initial velocity perturbation (0: rest) */
    uint64_t seed;
} lbm_case;

typedef struct lbm_case_status {
    long long id;
    int state;             /* LBM_CASE_* state */
    long long steps_done;
    long long cells;
    double submit_s, admit_s, done_s;  /* server clock (lbm_server_time) */
} lbm_case_status;

typedef struct lbm_server_stats {
    long long submitted, completed, failed, chunks, steps_cells;  /* steps_cells: sum of cells x steps run */
    size_t bytes_in_use, bytes_peak;
    int queued, active, ready;
    double device_ms;  /* sum over chunks of the device time of the fused launches */
    long long cells_peak;  /* most cells resident at once (in the pools or admitted for the next chunk) */
} lbm_server_stats;

void lbm_server_config_default(lbm_server_config *cfg);
void lbm_case_default(lbm_case *c);
int lbm_server_create(const lbm_server_config *cfg, lbm_server **out);
int lbm_server_destroy(lbm_server *s);
const char *lbm_server_error(const lbm_server *s);
/* device bytes a case occupies while resident */
int lbm_server_case_bytes(const lbm_case *c, size_t *bytes);
/* queue a case; fails if it can never fit the budget */
int lbm_server_submit(lbm_server *s, const lbm_case *c, long long *id);
/* one scheduling round: wait for the chunk in flight, retire frozen cases
 * and add admitted ones, launch the next chunk, then (while it runs) read
 * back the retired cases and admit queued ones.  *completed = cases whose
 * results became available in this call. */
int lbm_server_progress(lbm_server *s, int *completed);
/* oldest completed case not yet returned by this call (LBM_ERR_STATE: none) */
int lbm_server_next_done(lbm_server *s, lbm_case_status *st);
int lbm_server_status(const lbm_server *s, long long id, lbm_case_status *st);
/* copy a completed case's rho[cells] and j[3][cells] (float, cell order
 * x + X*(y + Y*z), from the distributions after its last step) and release
 * them; null pointers only release */
int lbm_server_fetch(lbm_server *s, long long id, float *rho, float *j);
int lbm_server_stats_get(const lbm_server *s, lbm_server_stats *st);
double lbm_server_time(const lbm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* LBM_CAPI_H */
