// lbm_main.cpp -- command-line driver with the reference's options and modes
// (src/main.cpp:96-434): getopt flags or `-c conf.xml`, normal mode and
// validation mode.  Validation runs the decomposed grid and the equivalent
// single domain of N - 2(k - 1) cells in one process and compares the interior
// velocity of every subdomain (the reference compares rank 0 over MPI,
// src/main.cpp:309-408).
#include <getopt.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <list>
#include <string>
#include <vector>

#include "host/CConfiguration.hpp"
#include "host/CController.hpp"
#include "host/CDomain.hpp"
#include "host/CManager.hpp"

typedef float T;
typedef Singleton<CConfiguration<T>> ConfigSingleton;

static void extract_comma_separated_integers(std::list<int> &out, const std::string &s) {
    size_t start = 0, comma;
    while ((comma = s.find(',', start)) != std::string::npos) {
        std::string sub = s.substr(start, comma - start);
        out.push_back(sub.empty() ? 0 : atoi(sub.c_str()));
        start = comma + 1;
    }
    out.push_back(atoi(s.substr(start).c_str()));
}

static int usage(const char *argv0) {
    std::cout << "usage: " << argv0 << std::endl
              << "\t\t[-x resolution_x, default: 32]" << std::endl
              << "\t\t[-y resolution_y, default: 32]" << std::endl
              << "\t\t[-z resolution_z, default: 32]" << std::endl
              << "\t\t[-X subdomain resolution_x, default: 1]" << std::endl
              << "\t\t[-Y subdomain resolution_y, default: 1]" << std::endl
              << "\t\t[-Z subdomain resolution_z, default: 1]" << std::endl
              << "\t\t[-S cube resolution]" << std::endl
              << "\t\t[-G gravitation in down direction, default: -9.81]" << std::endl
              << "\t\t[-r viscosity, default: 0.001308]" << std::endl
              << "\t\t[-n -m -p domain length x, y, z, default: 0.1]" << std::endl
              << "\t\t[-v]\t(debug mode on, be verbose)" << std::endl
              << "\t\t[-d device_num]" << std::endl
              << "\t\t[-k work-group size of the step kernels]" << std::endl
              << "\t\t[-g]\t(VTK output every step)" << std::endl
              << "\t\t[-l loops]" << std::endl
              << "\t\t[-R registers list]\t(accepted, not used on gfx950)" << std::endl
              << "\t\t[-T work group threads]\t(comma separated list: init,alpha,beta,copy)" << std::endl
              << "\t\t[-t timestep]\t(default: -1, not used: derived by the parametrization)" << std::endl
              << "\t\t[-u suite]\t(unit tests live in tests/ and run with pytest)" << std::endl
              << "\t\t[-c config.xml]" << std::endl;
    return -1;
}

int main(int argc, char **argv) {
    bool debug = false;
    CVector<3, int> domain_size(32, 32, 32);
    CVector<3, int> subdomain_nums(1, 1, 1);
    CVector<3, T> domain_length(0.1f, 0.1f, 0.1f);
    CVector<3, T> gravitation(0, -9.81f, 0);
    CVector<4, T> drivenCavityVelocity(100, 0, 0, 1);
    T viscosity = 0.001308f;
    T timestep = -1.0f;
    int loops = -1;
    bool do_visualisation = false;
    size_t computation_kernel_count = 128;
    std::string registers, threads, conf_file, test_suite;
    bool use_config_file = false, unit_test = false;
    int device_nr = 0;
    int opt;
    while ((opt = getopt(argc, argv, "x:y:z:d:vr:k:gG:t:sl:R:T:X:Y:Z:S:u:c:n:m:p:")) > 0) {
        switch (opt) {
            case 'd': device_nr = atoi(optarg); break;
            case 'v': debug = true; break;
            case 'R': registers = optarg; break;
            case 'T': threads = optarg; break;
            case 'x': domain_size[0] = atoi(optarg); break;
            case 'y': domain_size[1] = atoi(optarg); break;
            case 'z': domain_size[2] = atoi(optarg); break;
            case 'X': subdomain_nums[0] = atoi(optarg); break;
            case 'Y': subdomain_nums[1] = atoi(optarg); break;
            case 'Z': subdomain_nums[2] = atoi(optarg); break;
            case 'S': domain_size[0] = domain_size[1] = domain_size[2] = atoi(optarg); break;
            case 'l': loops = atoi(optarg); break;
            case 'k': computation_kernel_count = (size_t)atoi(optarg); break;
            case 'r': viscosity = (T)atof(optarg); break;
            case 'n': domain_length[0] = (T)atof(optarg); break;
            case 'm': domain_length[1] = (T)atof(optarg); break;
            case 'p': domain_length[2] = (T)atof(optarg); break;
            case 'g': do_visualisation = true; break;
            case 'G': gravitation[1] = (T)atof(optarg); break;
            case 't': timestep = (T)atof(optarg); break;
            case 's': break;
            case 'u': unit_test = true; test_suite = optarg; break;
            case 'c': use_config_file = true; conf_file = optarg; break;
            default: return usage(argv[0]);
        }
    }
    auto *cfg = ConfigSingleton::Instance();
    try {
        if (use_config_file) {
            cfg->loadFile(conf_file);
        } else {
            cfg->domain_size = domain_size;
            cfg->subdomain_num = subdomain_nums;
            cfg->domain_length = domain_length;
            cfg->gravitation = gravitation;
            cfg->viscosity = viscosity;
            cfg->computation_kernel_count = computation_kernel_count;
            cfg->device_nr = device_nr;
            cfg->do_visualization = do_visualisation;
            cfg->timestep = timestep;
            cfg->loops = loops;
            cfg->drivenCavityVelocity = drivenCavityVelocity;
            cfg->do_validate = false;
        }
        if (!threads.empty()) extract_comma_separated_integers(cfg->lbm_opencl_number_of_threads_list, threads);
        if (!registers.empty()) extract_comma_separated_integers(cfg->lbm_opencl_number_of_registers_list, registers);
        cfg->debug_mode = debug;
        if (debug) cfg->printMe();
        if (unit_test) {
            std::cout << "unit tests run with pytest (tests/); suite '" << test_suite << "' ignored"
                      << std::endl;
            return 0;
        }
        CVector<3, int> origin(0, 0, 0);
        if (cfg->do_validate) {
            const CVector<3, int> nums = cfg->subdomain_num;
            CDomain<T> domain(-1, cfg->domain_size, origin, cfg->domain_length);
            CManager<T> manager(domain, nums, cfg->periodic);
            manager.initSimulation(0);
            manager.startSimulation();
            CVector<3, int> vsize(cfg->domain_size[0] - 2 * (nums[0] - 1), cfg->domain_size[1] - 2 * (nums[1] - 1),
                                  cfg->domain_size[2] - 2 * (nums[2] - 1));
            CVector<3, T> vlen;
            for (int a = 0; a < 3; ++a) vlen[a] = vsize[a] * (cfg->domain_length[a] / (T)cfg->domain_size[a]);
            CDomain<T> vdomain(-2, vsize, origin, vlen);
            CManager<T> vmanager(vdomain, CVector<3, int>(1, 1, 1), cfg->periodic);
            std::cout << nums.elements() << "--> Compute the results for one domain." << std::endl;
            vmanager.initSimulation(-1);
            vmanager.startSimulation();
            const double tolerance = 1.0e-15;
            int failed_total = 0;
            for (auto *ctl : manager.getControllers()) {
                const int id = ctl->getUid();
                CVector<3, int> s = ctl->getDomain().getSize();
                CVector<3, int> inner(s[0] - 2, s[1] - 2, s[2] - 2);
                std::vector<T> local((size_t)inner.elements() * 3), global((size_t)inner.elements() * 3);
                CVector<3, int> lo(1, 1, 1);
                ctl->getSolver()->storeVelocity(local.data(), lo, inner);
                const int c[3] = {id % nums[0], (id / nums[0]) % nums[1], id / (nums[0] * nums[1])};
                CVector<3, int> go(1 + c[0] * inner[0], 1 + c[1] * inner[1], 1 + c[2] * inner[2]);
                vmanager.getController()->getSolver()->storeVelocity(global.data(), go, inner);
                int failed = 0;
                for (size_t i = 0; i < local.size(); ++i)
                    if (std::fabs(global[i] - local[i]) > tolerance) ++failed;
                failed_total += failed;
                std::cout << "--> PROC. RANK: " << id << " TOLERANCE: " << tolerance
                          << " NUMBER OF FAILED CELLS/TOTAL NUMBER OF CELLS: " << failed << "/" << inner.elements()
                          << std::endl;
            }
            return failed_total == 0 ? 0 : 2;
        }
        CDomain<T> domain(-1, cfg->domain_size, origin, cfg->domain_length);
        CManager<T> manager(domain, cfg->subdomain_num, cfg->periodic);
        manager.initSimulation(0);
        manager.startSimulation();
        if (debug) {
            // the debug checksum line of src/CController.hpp:490-497 (fixed, 8 decimals)
            auto *s = manager.getController()->getSolver();
            const std::streamsize ss = std::cout.precision();
            std::cout.precision(8);
            std::cout.setf(std::ios::fixed, std::ios::floatfield);
            std::cout << "Checksum: " << (s->getVelocityChecksum() * 1000.0f) << std::endl;
            std::cout.precision(ss);
            std::cout.unsetf(std::ios::floatfield);
        }
    } catch (const char *msg) {
        std::cerr << "error: " << msg << std::endl;
        return 1;
    }
    return 0;
}
