// The lbm_solver_config that CLbmSolver hands the library, from the
// reference's physical constructor arguments (no device needed: the device
// solver's creation may fail, the configuration is printed either way).
// Usage: host_solver_selftest <nx> <length> <gy> <nu> <U>
#include <cstdio>
#include <cstdlib>
#include <list>

#include "host/CLbmSolver.hpp"

int main(int argc, char **argv) {
    if (argc < 6) return 2;
    const int n = atoi(argv[1]);
    const float L = (float)atof(argv[2]);
    CVector<3, float> g(0.f, (float)atof(argv[3]), 0.f);
    const float nu = (float)atof(argv[4]);
    CVector<4, float> U((float)atof(argv[5]), 0.f, 0.f, 1.f);
    CDomain<float> d(0, CVector<3, int>(n, n, n), CVector<3, int>(0, 0, 0), CVector<3, float>(L, L, L));
    int bc[3][2] = {{1, 1}, {1, 1}, {1, 1}};
    lbm::CDeviceContext ctx;
    std::list<int> wi, regs;
    CLbmSolver<float> s(0, ctx, bc, d, g, nu, 256, true, true, -1.0f, U, wi, regs);
    lbm_solver_config c;
    s.solverConfig(c);
    printf("tau %.9g inv_tau %.9g lid %.9g g %.9g %.9g %.9g wg %d\n", c.tau, c.inv_tau, c.lid_velocity,
           c.gravitation[0], c.gravitation[1], c.gravitation[2], c.work_group_size);
    return 0;
}
