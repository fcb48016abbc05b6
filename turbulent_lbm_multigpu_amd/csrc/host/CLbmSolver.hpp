// CLbmSolver.hpp -- the reference's solver class (src/CLbmSolver.hpp:67-1124)
// kept as the drop-in C++ surface; every method forwards to the gfx950 HIP
// library through include/lbm_capi.h.  The three OpenCL handles of the
// reference constructor (queue, context, device) become one lbm::CDeviceContext.
#pragma once

#include <cstdio>
#include <iomanip>
#include <iostream>
#include <list>
#include <string>
#include <typeinfo>
#include <vector>

#include "../../../include/lbm_capi.h"
#include "CError.hpp"
#include "CLbmSkeleton.hpp"
#include "CVector.hpp"

#define LBM_FLAG_OBSTACLE_HOST (1 << 0)
#define LBM_FLAG_FLUID_HOST (1 << 1)
#define LBM_FLAG_VELOCITY_INJECTION_HOST (1 << 2)
#define LBM_FLAG_GHOST_LAYER_HOST (1 << 3)

namespace lbm {

// Device selection replacing CCL::CCommandQueue/CContext/CDevice.
struct CDeviceContext {
    int device = 0;
};

// D3Q19 directions in the reference's order (src/main.cpp:38-66).
inline const CVector<3, int> &lattice_unit(int f) {
    static const CVector<3, int> e[19] = {
        CVector<3, int>(1, 0, 0),  CVector<3, int>(-1, 0, 0),  CVector<3, int>(0, 1, 0),   CVector<3, int>(0, -1, 0),
        CVector<3, int>(1, 1, 0),  CVector<3, int>(-1, -1, 0), CVector<3, int>(1, -1, 0),  CVector<3, int>(-1, 1, 0),
        CVector<3, int>(1, 0, 1),  CVector<3, int>(-1, 0, -1), CVector<3, int>(1, 0, -1),  CVector<3, int>(-1, 0, 1),
        CVector<3, int>(0, 1, 1),  CVector<3, int>(0, -1, -1), CVector<3, int>(0, 1, -1),  CVector<3, int>(0, -1, 1),
        CVector<3, int>(0, 0, 1),  CVector<3, int>(0, 0, -1),  CVector<3, int>(0, 0, 0)};
    return e[f];
}

// Lattice-unit parameters that bypass the physical parametrization (the
// `<lattice>` block of the configuration): tau 0.6 with no gravity, LES,
// body force, storage and kernel tuning cannot be expressed through
// CLbmSkeleton alone.
template <typename T>
struct CLatticeOverrides {
    bool enabled = false;
    double tau = 0.6;
    double lid_velocity = 0.0;
    double body_force[3] = {0, 0, 0};
    double smagorinsky_cs = 0.0;
    int forcing = LBM_FORCING_REFERENCE;
    int storage = -1;         // -1: from T (f32 / f64)
    int kernel_variant = 0;
};

}  // namespace lbm

template <typename T>
class CLbmSolver : public CLbmSkeleton<T> {
public:
    CVector<4, T> drivenCavityVelocity;
    static const size_t SIZE_DD_HOST = 19;
    size_t simulation_step_counter = 0;
    CError error;
    std::list<int> lbm_opencl_number_of_work_items_list;
    std::list<int> lbm_opencl_number_of_registers_list;

private:
    int _UID;
    int _BC[3][2];
    lbm::CDeviceContext _ctx;
    size_t computation_kernel_count;
    bool store_velocity, store_density;
    lbm::CLatticeOverrides<T> _lattice;
    lbm_solver *_h = nullptr;

    bool fail(int rc) {
        if (rc == LBM_OK) return false;
        error << "lbm: " << (_h ? lbm_solver_error(_h) : lbm_last_error()) << CError::endl;
        std::cerr << "lbm: " << (_h ? lbm_solver_error(_h) : lbm_last_error()) << std::endl;
        return true;
    }
    static int storage_of_T() { return typeid(T) == typeid(double) ? LBM_STORAGE_F64 : LBM_STORAGE_F32; }
    int box(CVector<3, int> &v, int out[3]) const {
        for (int i = 0; i < 3; ++i) out[i] = v[i];
        return 0;
    }

public:
    // src/CLbmSolver.hpp:221-260
    CLbmSolver(int UID, lbm::CDeviceContext &ctx, int BC[3][2], CDomain<T> &domain, CVector<3, T> &p_d_gravitation,
               T p_d_viscosity, size_t p_computation_kernel_count, bool p_store_velocity, bool p_store_density,
               T /*p_d_timestep: unused, as in the reference*/, CVector<4, T> &_drivenCavityVelocity,
               std::list<int> &p_work_items_list, std::list<int> &p_registers_list,
               const lbm::CLatticeOverrides<T> *lattice = nullptr)
        : CLbmSkeleton<T>(CDomain<T>(domain), _drivenCavityVelocity),
          drivenCavityVelocity(_drivenCavityVelocity),
          lbm_opencl_number_of_work_items_list(p_work_items_list),
          lbm_opencl_number_of_registers_list(p_registers_list),
          _UID(UID),
          _ctx(ctx),
          computation_kernel_count(p_computation_kernel_count),
          store_velocity(p_store_velocity),
          store_density(p_store_density) {
        if (lattice) _lattice = *lattice;
        CLbmSkeleton<T>::init(p_d_gravitation, p_d_viscosity, (T)1.0, (T)0.0001, (T)0.953575, _lattice.enabled);
        if (_lattice.enabled) {
            // lattice values given directly: no parametrization, no tau guard
            CLbmSkeleton<T>::error.clear();
            this->tau = (T)_lattice.tau;
            this->inv_tau = (T)1.0 / this->tau;
            CLbmSkeleton<T>::drivenCavityVelocity[0] = (T)_lattice.lid_velocity;
            for (int a = 0; a < 3; ++a) this->gravitation[a] = (T)_lattice.body_force[a];
        }
        if (CLbmSkeleton<T>::error()) error << CLbmSkeleton<T>::error.getString();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) _BC[i][j] = BC[i][j];
        reload();
    }

    ~CLbmSolver() {
        if (_h) lbm_solver_destroy(_h);
    }

    lbm_solver *handle() { return _h; }
    int getUid() const { return _UID; }

    // src/CLbmSolver.hpp:262-270: new lid speed = (U + value) * dt
    void addDrivenCavityValue(T value) {
        drivenCavityVelocity[0] += value;
        CVector<4, T> param = drivenCavityVelocity;
        param *= CLbmSkeleton<T>::d_timestep;
        if (_h) fail(lbm_solver_set_lid_velocity(_h, (double)param[0]));
    }

    // src/CLbmSolver.hpp:272-617: work-group size from the list (the alpha
    // entry, which drives the step kernels) or computation_kernel_count;
    // buffers, kernel arguments, then reset().
    void reload() {
        if (_h) {
            lbm_solver_destroy(_h);
            _h = nullptr;
        }
        if (error()) return;
        lbm_solver_config c;
        solverConfig(c);
        if (fail(lbm_solver_create(&c, &_h))) return;
        simulation_step_counter = 0;
    }

    // the configuration reload() creates the device solver with: the
    // skeleton's lattice values, the boundary flags and the launch shape
    void solverConfig(lbm_solver_config &c) const {
        lbm_solver_config_default(&c);
        CVector<3, int> cells = this->domain_cells;
        for (int a = 0; a < 3; ++a) c.dims[a] = cells[a];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) c.bc[i * 2 + j] = _BC[i][j];
        c.storage = _lattice.storage >= 0 ? _lattice.storage : storage_of_T();
        c.forcing = _lattice.forcing;
        c.inv_tau = (double)this->inv_tau;
        c.tau = (double)this->tau;
        c.smagorinsky_cs = _lattice.enabled ? _lattice.smagorinsky_cs : 0.0;
        for (int a = 0; a < 3; ++a) c.gravitation[a] = (double)this->gravitation[a];
        // the lattice lid U dt of the parametrization, not the physical U of
        // this class's own member (src/CLbmSolver.hpp:567)
        c.lid_velocity = (double)CLbmSkeleton<T>::drivenCavityVelocity[0];
        c.store_velocity = store_velocity;
        c.store_density = store_density;
        int wg = (int)computation_kernel_count;
        auto it = lbm_opencl_number_of_work_items_list.begin();
        if (it != lbm_opencl_number_of_work_items_list.end()) ++it;  // init kernel entry
        if (it != lbm_opencl_number_of_work_items_list.end() && *it != 0) wg = *it;
        c.work_group_size = wg;
        c.kernel_variant = _lattice.kernel_variant;
        c.device = _ctx.device;
    }

    // src/CLbmSolver.hpp:619-635
    void reset() {
        simulation_step_counter = 0;
        if (_h) fail(lbm_solver_reset(_h));
    }
    void simulationStepAlpha() {
        if (_h) fail(lbm_solver_simulation_step_alpha(_h));
    }
    void simulationStepBeta() {
        if (_h) fail(lbm_solver_simulation_step_beta(_h));
    }
    // src/CLbmSolver.hpp:664-676
    void simulationStep() {
        if (!_h) return;
        if (fail(lbm_solver_simulation_step(_h))) return;
        simulation_step_counter++;
    }
    // keeps the public counter in step with group stepping
    void syncStepCounter() {
        long long v = 0;
        if (_h && lbm_solver_get_step_counter(_h, &v) == LBM_OK) simulation_step_counter = (size_t)v;
    }
    void wait() {
        if (_h) fail(lbm_solver_wait(_h));
    }

    // ---- distributions (src/CLbmSolver.hpp:688-757)
    void storeDensityDistribution(T *dst) {
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_DD, nullptr, nullptr, dst));
    }
    void storeDensityDistribution(T *dst, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_DD, o, s, dst));
    }
    void setDensityDistribution(T *src, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_set(_h, LBM_FIELD_DD, o, s, src, nullptr));
    }
    void setDensityDistribution(T *src, CVector<3, int> &origin, CVector<3, int> &size, CVector<3, int> norm) {
        int o[3], s[3], n[3];
        box(origin, o);
        box(size, s);
        box(norm, n);
        if (_h) fail(lbm_solver_set(_h, LBM_FIELD_DD, o, s, src, n));
    }

    // ---- velocity / density / flags (src/CLbmSolver.hpp:764-978)
    void storeVelocity(T *dst) {
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_VELOCITY, nullptr, nullptr, dst));
    }
    void storeVelocity(T *dst, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_VELOCITY, o, s, dst));
    }
    void setVelocity(T *src, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_set(_h, LBM_FIELD_VELOCITY, o, s, src, nullptr));
    }
    void storeDensity(T *dst) {
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_DENSITY, nullptr, nullptr, dst));
    }
    void storeDensity(T *dst, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_DENSITY, o, s, dst));
    }
    void setDensity(T *src, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_set(_h, LBM_FIELD_DENSITY, o, s, src, nullptr));
    }
    void storeFlags(int *dst) {
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_FLAGS, nullptr, nullptr, dst));
    }
    void storeFlags(int *dst, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_store(_h, LBM_FIELD_FLAGS, o, s, dst));
    }
    void setFlags(int *src, CVector<3, int> &origin, CVector<3, int> &size) {
        int o[3], s[3];
        box(origin, o);
        box(size, s);
        if (_h) fail(lbm_solver_set(_h, LBM_FIELD_FLAGS, o, s, src, nullptr));
    }

    // ---- debugging (src/CLbmSolver.hpp:1032-1123)
    void debug_print() {
        const size_t n = (size_t)this->domain_cells.elements();
        std::vector<T> dd(n * SIZE_DD_HOST);
        storeDensityDistribution(dd.data());
        std::cout << "DENSITY DISTRIBUTIONS:";
        print_rows(dd.data(), dd.size(), 16);
        std::vector<int> fl(n);
        storeFlags(fl.data());
        std::cout << std::endl << "FLAGS:";
        for (size_t i = 0; i < n; ++i) {
            if (i % 16 == 0) std::cout << std::endl << (i / 16) << ": ";
            std::cout << fl[i] << " ";
        }
        std::cout << std::endl;
    }
    void debugDD(size_t dd_id = 0, size_t wrap_size = 16, size_t empty_line = 16) {
        const size_t n = (size_t)this->domain_cells.elements();
        std::vector<T> dd(n * SIZE_DD_HOST);
        storeDensityDistribution(dd.data());
        std::streamsize ss = std::cout.precision();
        std::cout.precision(4);
        std::cout.setf(std::ios::fixed, std::ios::floatfield);
        for (size_t i = n * dd_id; i < n * (dd_id + 1); ++i) {
            if (empty_line != wrap_size && i % empty_line == 0 && i != n * dd_id) std::cout << std::endl;
            if (i % wrap_size == 0) {
                if (i != n * dd_id) std::cout << std::endl;
                std::cout << (i / wrap_size) << ": ";
            }
            std::cout << dd[i] << " ";
        }
        std::cout.precision(ss);
        std::cout << std::resetiosflags(std::ios::fixed) << std::endl;
    }
    // float sum of vx+vy+vz over FLUID cells (src/CLbmSolver.hpp:1103-1123)
    float getVelocityChecksum() {
        float v = 0.f;
        if (_h) fail(lbm_solver_velocity_checksum(_h, &v));
        return v;
    }

private:
    void print_rows(const T *buf, size_t n, size_t wrap) {
        std::streamsize ss = std::cout.precision();
        std::cout.precision(4);
        std::cout.setf(std::ios::fixed, std::ios::floatfield);
        for (size_t i = 0; i < n; ++i) {
            if (i % wrap == 0) std::cout << std::endl << (i / wrap) << ": ";
            std::cout << buf[i] << " ";
        }
        std::cout.precision(ss);
        std::cout << std::resetiosflags(std::ios::fixed);
    }
};
