// CController.hpp -- one subdomain: its solver, boundary conditions, CComm
// list and geometry (src/CController.hpp:54-563).  The reference moves ghost
// layers through host buffers and MPI after every kernel
// (src/CController.hpp:265-383); here every subdomain lives on the same GPU and
// the manager's lbm_group performs syncAlpha/syncBeta as device copies.
#pragma once

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "CComm.hpp"
#include "CConfiguration.hpp"
#include "CDomain.hpp"
#include "CLbmSolver.hpp"
#include "CLbmVisualizationVTK.hpp"

#define FLAG_OBSTACLE (1 << 0)
#define FLAG_FLUID (1 << 1)
#define FLAG_VELOCITY_INJECTION (1 << 2)
#define FLAG_GHOST_LAYER (1 << 3)
#define FLAG_GHOST_LAYER_BETA (FLAG_GHOST_LAYER | (1 << 4))

#define BENCHMARK_OUTPUT_DIR "output/benchmark"
#define PROFILE_OUTPUT_DIR "output/profile"
#define VTK_OUTPUT_DIR "output/vtk"

template <typename T>
class CController {
    int _UID;
    CDomain<T> _domain;
    CLbmVisualizationVTK<T> *cLbmVisualization = nullptr;
    CLbmSolver<T> *cLbmPtr = nullptr;
    int _BC[3][2];
    std::vector<CComm<T> *> _comm_container;
    lbm::CDeviceContext _ctx;
    lbm_group *_group = nullptr;  // owned by the manager when several subdomains share the device
    bool _drives_group = false;   // this controller's computeNextStep steps _group

    // src/CController.hpp:83-226 (device selection and solver construction)
    int initLBMSolver() {
        auto *cfg = Singleton<CConfiguration<T>>::Instance();
        _ctx.device = cfg->device_nr < 0 ? 0 : cfg->device_nr;
        const bool store = cfg->do_visualization || cfg->debug_mode || cfg->do_validate;
        cLbmPtr = new CLbmSolver<T>(_UID, _ctx, _BC, _domain, cfg->gravitation, cfg->viscosity,
                                    cfg->computation_kernel_count, store, store, cfg->timestep,
                                    cfg->drivenCavityVelocity, cfg->lbm_opencl_number_of_threads_list,
                                    cfg->lbm_opencl_number_of_threads_list,
                                    cfg->lattice.enabled ? &cfg->lattice : nullptr);
        if (cLbmPtr->error()) {
            std::cout << cLbmPtr->error.getString();
            return -1;
        }
        cLbmPtr->wait();
        return 0;
    }

public:
    CController(int UID, CDomain<T> domain, int BC[3][2]) : _UID(UID), _domain(domain) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) _BC[i][j] = BC[i][j];
        if (-1 == initLBMSolver()) throw "Initialization of LBM Solver failed!";
    }
    ~CController() {
        delete cLbmVisualization;
        delete cLbmPtr;
        for (auto *c : _comm_container) delete c;
    }

    // `drives`: this controller steps the group (one per group)
    void attachGroup(lbm_group *g, bool drives = true) {
        _group = g;
        _drives_group = drives;
    }
    lbm_group *group() const { return _group; }

    // after alpha: neighbours' inner layers -> this group's ghost layers
    void syncAlpha() {
        if (_group && !_comm_container.empty()) lbm_group_sync(_group, 0);
    }
    // after beta: ghost layers -> neighbours' inner layers, n.e > 0 only
    void syncBeta() {
        if (_group && !_comm_container.empty()) lbm_group_sync(_group, 1);
    }

    // src/CController.hpp:385-391.  A controller attached to a group steps the
    // whole group: every subdomain advances together in one fused launch.  The
    // reference runs one controller per rank; here one controller per group
    // drives the loop (CManager::run), so a caller must not loop over the
    // controllers of one group, or the group would advance once per
    // controller.  Only the group's driving controller steps it; any other
    // controller of the group throws.
    void computeNextStep() {
        if (_group) {
            if (!_drives_group) throw "computeNextStep: this controller's group is stepped by its driving controller";
            if (lbm_group_step(_group, 1) != LBM_OK) throw "lbm_group_step failed";
            cLbmPtr->syncStepCounter();
            return;
        }
        cLbmPtr->simulationStep();
    }

    // src/CController.hpp:396-522 for a standalone subdomain
    int run() {
        auto *cfg = Singleton<CConfiguration<T>>::Instance();
        int loops = cfg->loops < 0 ? 100 : cfg->loops;
        double floats_per_cell = 19.0 * 2.0 + 1.0;
        if (cfg->do_visualization || cfg->debug_mode) floats_per_cell += 3;
        if (cfg->do_visualization) {
            cLbmVisualization = new CLbmVisualizationVTK<T>(_UID, std::string("./") + VTK_OUTPUT_DIR + "/OUTPUT");
            cLbmVisualization->setup(cLbmPtr);
        }
        auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < loops; ++i) {
            computeNextStep();
            if (cfg->do_visualization) cLbmVisualization->render(i);
        }
        cLbmPtr->wait();
        if (_group) lbm_group_wait(_group);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        report(std::cout, _domain.getSize(), secs, loops, (double)cLbmPtr->domain_cells.elements(), floats_per_cell);
        return EXIT_SUCCESS;
    }

    // the reference's end-of-run lines (src/CController.hpp:478-489)
    static void report(std::ostream &os, CVector<3, int> cube, double secs, int loops, double cells,
                       double floats_per_cell) {
        const double fps = (double)loops / secs;
        const double mlups = fps * cells * 0.000001;
        os << std::endl << "Cube: " << cube << std::endl;
        os << "Seconds: " << secs << std::endl;
        os << "FPS: " << fps << std::endl;
        os << "MLUPS: " << mlups << std::endl;
        os << "Bandwidth: " << (mlups * floats_per_cell * (double)sizeof(T)) << " MB/s (RW, bidirectional)" << std::endl;
    }

    void addCommunication(CComm<T> *comm) { _comm_container.push_back(comm); }
    const std::vector<CComm<T> *> &getCommunications() const { return _comm_container; }

    // src/CController.hpp:531-546: VELOCITY_INJECTION layer at y = sy-2
    void setGeometry() {
        CVector<3, int> origin(1, _domain.getSize()[1] - 2, 1);
        CVector<3, int> size(_domain.getSize()[0] - 2, 1, _domain.getSize()[2] - 2);
        std::vector<int> src((size_t)size.elements(), FLAG_VELOCITY_INJECTION);
        cLbmPtr->setFlags(src.data(), origin, size);
    }

    CLbmSolver<T> *getSolver() const { return cLbmPtr; }
    void setSolver(CLbmSolver<T> *s) { cLbmPtr = s; }
    CDomain<T> getDomain() const { return _domain; }
    int getUid() const { return _UID; }
    const int (*getBC() const)[2] { return _BC; }
};
