// ref_runner.hip -- launcher of the reference's own kernels (test
// infrastructure only).
//
// Loads a code object built by oracle/build_ref.py from the unmodified
// reference OpenCL sources and replays the reference host's sequence for one
// solver: reload() allocates dd[19][C], flags[C], velocity[3][C], density[C],
// bc[6] (src/CLbmSolver.hpp:381-403); reset() runs init_kernel
// (src/CLbmSolver.hpp:619-635); the caller's flags and (optionally) initial
// distributions are then written, as setFlags / setDensityDistribution would;
// simulationStep() runs lbm_kernel_beta on even and lbm_kernel_alpha on odd
// counters over ceil(C/WG)*WG work-items (src/CLbmSolver.hpp:637-676).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

static std::string g_err;

extern "C" const char *ref_last_error(void) { return g_err.c_str(); }

#define RCHK(x)                                                               \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);           \
            rc = -1;                                                          \
            goto done;                                                        \
        }                                                                     \
    } while (0)

// dims/wg must match the code object's compile-time prefix; cells % wg == 0.
extern "C" int ref_run(const char *co_path, int nx, int ny, int nz, int wg, const int *bc, float inv_tau, float gx,
                       float gy, float gz, float lid, const int *flags_in, const float *dd_in, int steps,
                       float *dd_out, float *vel_out, float *dens_out, int *flags_out) {
    int rc = 0;
    const long long C = (long long)nx * ny * nz;
    hipModule_t mod = nullptr;
    hipFunction_t k_init = nullptr, k_alpha = nullptr, k_beta = nullptr;
    void *dd = nullptr, *fl = nullptr, *vel = nullptr, *dens = nullptr, *dbc = nullptr;
    if (wg <= 0 || wg > 1024 || C % wg != 0 || C > (1ll << 28)) {
        g_err = "ref_run: cells must be a positive multiple of the work-group size";
        return -2;
    }
    const unsigned blocks = (unsigned)(C / wg);
    RCHK(hipModuleLoad(&mod, co_path));
    RCHK(hipModuleGetFunction(&k_init, mod, "init_kernel"));
    RCHK(hipModuleGetFunction(&k_alpha, mod, "lbm_kernel_alpha"));
    RCHK(hipModuleGetFunction(&k_beta, mod, "lbm_kernel_beta"));
    RCHK(hipMalloc(&dd, sizeof(float) * 19 * C));
    RCHK(hipMalloc(&fl, sizeof(int) * C));
    RCHK(hipMalloc(&vel, sizeof(float) * 3 * C));
    RCHK(hipMalloc(&dens, sizeof(float) * C));
    RCHK(hipMalloc(&dbc, sizeof(int) * 6));
    RCHK(hipMemcpy(dbc, bc, sizeof(int) * 6, hipMemcpyHostToDevice));
    {
        void *args[] = {&dd, &fl, &vel, &dens, &dbc, &lid};
        RCHK(hipModuleLaunchKernel(k_init, blocks, 1, 1, wg, 1, 1, 0, nullptr, args, nullptr));
        RCHK(hipDeviceSynchronize());
    }
    if (flags_in) RCHK(hipMemcpy(fl, flags_in, sizeof(int) * C, hipMemcpyHostToDevice));
    if (dd_in) RCHK(hipMemcpy(dd, dd_in, sizeof(float) * 19 * C, hipMemcpyHostToDevice));
    for (int s = 0; s < steps; ++s) {
        void *args[] = {&dd, &fl, &vel, &dens, &inv_tau, &gx, &gy, &gz, &lid};
        RCHK(hipModuleLaunchKernel((s & 1) ? k_alpha : k_beta, blocks, 1, 1, wg, 1, 1, 0, nullptr, args, nullptr));
    }
    RCHK(hipDeviceSynchronize());
    if (dd_out) RCHK(hipMemcpy(dd_out, dd, sizeof(float) * 19 * C, hipMemcpyDeviceToHost));
    if (vel_out) RCHK(hipMemcpy(vel_out, vel, sizeof(float) * 3 * C, hipMemcpyDeviceToHost));
    if (dens_out) RCHK(hipMemcpy(dens_out, dens, sizeof(float) * C, hipMemcpyDeviceToHost));
    if (flags_out) RCHK(hipMemcpy(flags_out, fl, sizeof(int) * C, hipMemcpyDeviceToHost));
done:
    if (dd) (void)hipFree(dd);
    if (fl) (void)hipFree(fl);
    if (vel) (void)hipFree(vel);
    if (dens) (void)hipFree(dens);
    if (dbc) (void)hipFree(dbc);
    if (mod) (void)hipModuleUnload(mod);
    return rc;
}
