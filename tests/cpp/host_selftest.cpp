// Host-only checks of the C++ surface (no device calls): prints the parsed
// configuration and the parametrization of CLbmSkeleton for given inputs.
// Usage: host_selftest config <file.xml> | skeleton <nx> <length> <gy> <nu> <U>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "host/CConfiguration.hpp"
#include "host/CLbmSkeleton.hpp"

int main(int argc, char **argv) {
    if (argc >= 3 && !strcmp(argv[1], "config")) {
        CConfiguration<float> c;
        try {
            c.loadFile(argv[2]);
        } catch (const char *m) {
            printf("error %s\n", m);
            return 3;
        }
        printf("viscosity %.9g\n", c.viscosity);
        printf("gravitation %.9g %.9g %.9g\n", c.gravitation[0], c.gravitation[1], c.gravitation[2]);
        printf("cavity %.9g %.9g %.9g %.9g\n", c.drivenCavityVelocity[0], c.drivenCavityVelocity[1],
               c.drivenCavityVelocity[2], c.drivenCavityVelocity[3]);
        printf("domain %d %d %d\n", c.domain_size[0], c.domain_size[1], c.domain_size[2]);
        printf("subdomains %d %d %d\n", c.subdomain_num[0], c.subdomain_num[1], c.subdomain_num[2]);
        printf("length %.9g %.9g %.9g\n", c.domain_length[0], c.domain_length[1], c.domain_length[2]);
        printf("loops %d timestep %.9g vtk %d validate %d\n", c.loops, c.timestep, (int)c.do_visualization,
               (int)c.do_validate);
        printf("kernel_count %zu device %d\n", c.computation_kernel_count, c.device_nr);
        printf("lattice %d tau %.9g lid %.9g cs %.9g storage %d periodic %d %d %d exchange %d\n",
               (int)c.lattice.enabled, c.lattice.tau, c.lattice.lid_velocity, c.lattice.smagorinsky_cs,
               c.lattice.storage, (int)c.periodic[0], (int)c.periodic[1], (int)c.periodic[2], c.exchange_mode);
        return 0;
    }
    if (argc >= 7 && !strcmp(argv[1], "skeleton")) {
        const int n = atoi(argv[2]);
        const float L = (float)atof(argv[3]);
        CVector<3, float> g(0.f, (float)atof(argv[4]), 0.f);
        const float nu = (float)atof(argv[5]);
        CVector<4, float> U((float)atof(argv[6]), 0.f, 0.f, 1.f);
        CDomain<float> d(0, CVector<3, int>(n, n, n), CVector<3, int>(0, 0, 0), CVector<3, float>(L, L, L));
        CLbmSkeleton<float> s(d, U);
        s.init(g, nu, 1.0f, 0.0001f, 0.953575f, true);
        printf("dx %.9g dt %.9g tau %.9g inv_tau %.9g g %.9g %.9g %.9g lid %.9g error %d\n", s.d_cell_length,
               s.d_timestep, s.tau, s.inv_tau, s.gravitation[0], s.gravitation[1], s.gravitation[2],
               s.drivenCavityVelocity[0], (int)s.error());
        return 0;
    }
    return 2;
}
