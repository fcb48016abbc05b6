// CManager.hpp -- grid decomposition and the simulation loop
// (src/CManager.hpp:15-224).  Subdomain `id` sits at (id % Nx, id / Nx % Ny,
// id / (Nx*Ny)); outer faces are OBSTACLE, internal faces GHOST with a CComm
// to the neighbour; neighbours overlap by two cells; the top row of
// subdomains gets the lid (setGeometry).  The reference builds one subdomain
// per MPI rank; here initSimulation builds every subdomain on the one device,
// joins them in an lbm_group (one fused step kernel per half-step plus
// on-device ghost exchange, x, then y, then z faces) and getController()
// returns the controller of `my_rank`.  A periodic axis (new) makes the outer
// faces of that axis GHOST faces whose neighbour wraps around.
#pragma once

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "CComm.hpp"
#include "CConfiguration.hpp"
#include "CController.hpp"
#include "CDomain.hpp"
#include "CVector.hpp"

template <typename T>
class CManager {
    CDomain<T> _domain;
    CVector<3, int> _subdomain_size;
    CVector<3, T> _subdomain_length;
    CVector<3, int> _subdomain_nums;
    std::vector<CController<T> *> _controllers;
    int _my_rank = 0;
    lbm_group *_group = nullptr;
    bool _periodic[3] = {false, false, false};

public:
    CManager(CDomain<T> domain, CVector<3, int> subdomainNums, const bool *periodic = nullptr) : _domain(domain) {
        if (periodic)
            for (int a = 0; a < 3; ++a) _periodic[a] = periodic[a];
        setSubdomainNums(subdomainNums);
    }
    ~CManager() {
        if (_group) lbm_group_destroy(_group);
        for (auto *c : _controllers) delete c;
    }

    CDomain<T> getDomain() const { return _domain; }
    void setDomain(CDomain<T> grid) { _domain = grid; }
    CVector<3, int> getSubdomainNums() const { return _subdomain_nums; }
    CVector<3, int> getSubdomainSize() const { return _subdomain_size; }
    lbm_group *getGroup() const { return _group; }

    // src/CManager.hpp:49-76
    void setSubdomainNums(CVector<3, int> nums) {
        CVector<3, int> ds = _domain.getSize();
        for (int a = 0; a < 3; ++a)
            if (nums[a] <= 0 || ds[a] % nums[a] != 0) throw "Number of subdomains does not match with the grid size!";
        for (int a = 0; a < 3; ++a) {
            _subdomain_size[a] = ds[a] / nums[a];
            _subdomain_length[a] = _domain.getLength()[a] / nums[a];
        }
        _subdomain_nums = nums;
    }

    // src/CManager.hpp:78-204, for every subdomain of the grid
    void initSimulation(int my_rank) {
        if (!_controllers.empty()) throw "CManager: simulation already initialised";
        auto *cfg = Singleton<CConfiguration<T>>::Instance();
        _my_rank = my_rank < 0 ? 0 : my_rank;
        const int n = _subdomain_nums.elements();
        if (_my_rank >= n) throw "CManager: rank outside the decomposition";
        const CVector<3, int> s = _subdomain_size;
        for (int id = 0; id < n; ++id) {
            int t = id;
            const int c[3] = {t % _subdomain_nums[0], (t / _subdomain_nums[0]) % _subdomain_nums[1],
                              t / (_subdomain_nums[0] * _subdomain_nums[1])};
            CVector<3, int> origin(c[0] * s[0], c[1] * s[1], c[2] * s[2]);
            CDomain<T> sub(id, s, origin, _subdomain_length);
            int BC[3][2];
            for (int a = 0; a < 3; ++a) {
                BC[a][0] = BC[a][1] = FLAG_GHOST_LAYER;
                if (!_periodic[a]) {
                    if (c[a] == 0) BC[a][0] = FLAG_OBSTACLE;
                    if (c[a] == _subdomain_nums[a] - 1) BC[a][1] = FLAG_OBSTACLE;
                }
            }
            auto *ctl = new CController<T>(id, sub, BC);
            _controllers.push_back(ctl);
            const int stride[3] = {1, _subdomain_nums[0], _subdomain_nums[0] * _subdomain_nums[1]};
            for (int a = 0; a < 3; ++a)
                for (int side = 0; side < 2; ++side) {
                    if (BC[a][side] != FLAG_GHOST_LAYER) continue;
                    int nc[3] = {c[0], c[1], c[2]};
                    nc[a] = (c[a] + (side ? 1 : -1) + _subdomain_nums[a]) % _subdomain_nums[a];
                    const int dst = nc[0] * stride[0] + nc[1] * stride[1] + nc[2] * stride[2];
                    CVector<3, int> size = s;
                    size[a] = 1;
                    CVector<3, int> send_o(0, 0, 0), recv_o(0, 0, 0), normal(0, 0, 0);
                    if (side == 0) {
                        send_o[a] = 1;
                        recv_o[a] = 0;
                        normal[a] = 1;
                    } else {
                        send_o[a] = s[a] - 2;
                        recv_o[a] = s[a] - 1;
                        normal[a] = -1;
                    }
                    ctl->addCommunication(new CComm<T>(dst, size, size, send_o, recv_o, normal));
                }
        }
        // seeded initial field keyed by the position in the equivalent single domain
        if (cfg->perturbation_amplitude != 0.0) {
            int key_dims[3];
            for (int a = 0; a < 3; ++a) key_dims[a] = _domain.getSize()[a] - 2 * (_subdomain_nums[a] - 1);
            for (int id = 0; id < n; ++id) {
                int key_o[3] = {(id % _subdomain_nums[0]) * (s[0] - 2),
                                ((id / _subdomain_nums[0]) % _subdomain_nums[1]) * (s[1] - 2),
                                (id / (_subdomain_nums[0] * _subdomain_nums[1])) * (s[2] - 2)};
                lbm_solver *h = _controllers[id]->getSolver()->handle();
                if (h && lbm_solver_init_perturbed(h, cfg->perturbation_amplitude, cfg->perturbation_seed, 0, key_o,
                                                   key_dims) != LBM_OK)
                    throw "perturbed initialisation failed";
            }
        }
        for (int id = 0; id < n; ++id) {
            const int ny = (id / _subdomain_nums[0]) % _subdomain_nums[1];
            if (ny == _subdomain_nums[1] - 1) _controllers[id]->setGeometry();  // src/CManager.hpp:200-202
        }
        std::vector<lbm_solver *> hs;
        for (auto *ctl : _controllers) hs.push_back(ctl->getSolver()->handle());
        if (lbm_group_create(hs.data(), (int)hs.size(), &_group) != LBM_OK) throw "lbm_group_create failed";
        for (auto *ctl : _controllers)
            for (auto *cm : ctl->getCommunications()) {
                int so[3], ro[3], sz[3], nm[3];
                for (int a = 0; a < 3; ++a) {
                    so[a] = cm->getSendOrigin()[a];
                    ro[a] = cm->getRecvOrigin()[a];
                    sz[a] = cm->getSendSize()[a];
                    nm[a] = cm->getCommDirection()[a];
                }
                if (lbm_group_add_comm(_group, ctl->getUid(), cm->getDstId(), so, ro, sz, nm) != LBM_OK)
                    throw "lbm_group_add_comm failed";
            }
        // ghost layers start as copies of their owners (all 19 populations)
        if (lbm_group_set_exchange_mode(_group, 0) != LBM_OK || lbm_group_sync(_group, 0) != LBM_OK)
            throw "CManager: initial ghost-layer exchange failed";
        if (lbm_group_set_exchange_mode(_group, cfg->exchange_mode) != LBM_OK)
            throw "CManager: <exchange-mode> must be 0 or 1";
        if (lbm_group_set_fused_x(_group, cfg->fused_x) != LBM_OK)
            throw "CManager: <fused-x> must be 0, 1 or 2";
        if (lbm_group_set_alpha_work_group(_group, cfg->alpha_work_group) != LBM_OK)
            throw "CManager: <alpha-work-group> must be 0 or 64..512, a multiple of 64";
        // the first controller drives the group; stepping through any other throws
        for (size_t i = 0; i < _controllers.size(); ++i) _controllers[i]->attachGroup(_group, i == 0);
    }

    // src/CManager.hpp:206-211 + the loop of src/CController.hpp:396-522 for all subdomains
    void startSimulation() {
        if (_controllers.empty()) throw "CManager: Initialize the simulation before starting it!";
        auto *cfg = Singleton<CConfiguration<T>>::Instance();
        const int loops = cfg->loops < 0 ? 100 : cfg->loops;
        double floats_per_cell = 19.0 * 2.0 + 1.0;
        if (cfg->do_visualization || cfg->debug_mode) floats_per_cell += 3;
        std::vector<CLbmVisualizationVTK<T> *> vis;
        if (cfg->do_visualization) {
            mkdir("output", 0755);
            mkdir(VTK_OUTPUT_DIR, 0755);
            for (auto *ctl : _controllers) {
                vis.push_back(new CLbmVisualizationVTK<T>(ctl->getUid(), std::string("./") + VTK_OUTPUT_DIR + "/OUTPUT"));
                vis.back()->setup(ctl->getSolver());
            }
        }
        lbm_group_wait(_group);
        auto t0 = std::chrono::steady_clock::now();
        if (vis.empty()) {
            if (lbm_group_step(_group, loops) != LBM_OK) throw "lbm_group_step failed";
        } else {
            for (int i = 0; i < loops; ++i) {
                if (lbm_group_step(_group, 1) != LBM_OK) throw "lbm_group_step failed";
                for (auto *v : vis) v->render(i);
            }
        }
        lbm_group_wait(_group);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto *ctl : _controllers) ctl->getSolver()->syncStepCounter();
        for (auto *v : vis) delete v;
        const double cells = (double)_domain.getSize().elements();
        CController<T>::report(std::cout, _domain.getSize(), secs, loops, cells, floats_per_cell);
        if (cfg->benchmark_output) write_benchmark(secs, loops, cells, floats_per_cell);
        std::cout << "done." << std::endl;
    }

    // benchmark_<n>.ini as in src/CController.hpp:445-477
    void write_benchmark(double secs, int loops, double cells, double floats_per_cell) const {
        mkdir("output", 0755);
        mkdir(BENCHMARK_OUTPUT_DIR, 0755);
        std::stringstream name;
        name << "./" << BENCHMARK_OUTPUT_DIR << "/benchmark_" << _subdomain_nums.elements() << ".ini";
        std::ofstream f(name.str(), std::ios::out | std::ios::app);
        if (!f.is_open()) {
            std::cout << "Unable to open file";
            return;
        }
        const double fps = (double)loops / secs;
        const double mlups = fps * cells * 0.000001;
        f << "CUBE_X : " << _domain.getSize()[0] << std::endl;
        f << "CUBE_Y : " << _domain.getSize()[1] << std::endl;
        f << "CUBE_Z : " << _domain.getSize()[2] << std::endl;
        f << "SECONDS : " << secs << std::endl;
        f << "FPS : " << fps << std::endl;
        f << "MLUPS : " << mlups << std::endl;
        f << "BANDWIDTH : " << mlups * floats_per_cell * (double)sizeof(T) << std::endl << std::endl;
    }

    CController<T> *getController() const { return _controllers.empty() ? nullptr : _controllers[_my_rank]; }
    CController<T> *getController(int id) const { return _controllers.at(id); }
    const std::vector<CController<T> *> &getControllers() const { return _controllers; }
    void setController(CController<T> *c) {
        if (!_controllers.empty()) _controllers[_my_rank] = c;
    }
};
