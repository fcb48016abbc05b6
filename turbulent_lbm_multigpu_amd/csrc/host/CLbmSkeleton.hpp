// CLbmSkeleton.hpp -- physical -> lattice parametrization of the solver
// (src/CLbmSkeleton.hpp:38-199), same arithmetic in T:
//   dt  = dx^2 (2 tau - 1) / (6 nu sqrt(m))
//   g   = g_phys dt^2 / dx; if |g| >= max_g: dt = sqrt(max_g dx / |g_phys|),
//         g = g_phys dt^2 / dx, tau = 0.5 (dt nu sqrt(m) 6) / dx^2 + 0.5
//   tau outside [0.51, 2.5] is an error; inv_tau = 1 / tau; lid = U dt.
#pragma once

#include <cmath>
#include <iostream>

#include "CDomain.hpp"
#include "CError.hpp"
#include "CVector.hpp"

template <typename T>
class CLbmSkeleton {
public:
    CError error;
    bool debug = false;

    CDomain<T> domain;
    CVector<3, int> domain_cells;
    int domain_cells_count = 0;
    CVector<4, T> d_drivenCavityVelocity;

    T d_domain_x_length = 0;
    CVector<3, T> d_gravitation;
    T d_viscosity = 0;
    T mass_exchange_factor = 1;

    T d_cell_length = 0;
    T d_timestep = 0;
    T d_reynolds = 0;
    T tau = 0;
    T inv_tau = 0;
    T inv_trt_tau = 0;

    CVector<4, T> drivenCavityVelocity;
    CVector<3, T> gravitation;
    T max_sim_gravitation_length = 0;

    void updateValues(bool info_output = false) {
        d_timestep = (d_cell_length * d_cell_length) * ((T)2.0 * tau - (T)1.0) /
                     ((T)6.0 * d_viscosity * (T)std::sqrt(mass_exchange_factor));
        gravitation = d_gravitation * ((d_timestep * d_timestep) / d_cell_length);
        if (gravitation.length() >= max_sim_gravitation_length) {
            if (info_output) std::cout << "limiting timestep (gravitation: " << gravitation << ")" << std::endl;
            d_timestep = (T)std::sqrt((max_sim_gravitation_length * d_cell_length) / d_gravitation.length());
            gravitation = d_gravitation * ((d_timestep * d_timestep) / d_cell_length);
            tau = (T)0.5 * (d_timestep * d_viscosity * (T)std::sqrt(mass_exchange_factor) * (T)6.0) /
                      (d_cell_length * d_cell_length) +
                  (T)0.5;
        }
        if (tau < 0.51 || tau > 2.5) {
            error << "relaxation time tau = " << tau << " is outside the stable range 0.51 <= tau <= 2.5" << std::endl;
        }
        inv_tau = (T)1.0 / tau;
        inv_trt_tau = (T)1.0 / ((T)0.5 + (T)3.0 / ((T)16.0 * tau - (T)8.0));
    }

    void setGravitation(CVector<3, T> g) {
        d_gravitation = g;
        updateValues();
    }

    void init(CVector<3, T> &p_d_gravitation, T p_d_viscosity, T p_mass_exchange_factor,
              T p_max_sim_gravitation_length = (T)0.0001, T p_tau = (T)0.953575, bool quiet = false) {
        domain_cells = domain.getSize();
        d_domain_x_length = domain.getLength()[0];
        d_gravitation = p_d_gravitation;
        domain_cells_count = domain_cells.elements();
        d_viscosity = p_d_viscosity;
        mass_exchange_factor = p_mass_exchange_factor;
        d_cell_length = d_domain_x_length / (T)domain_cells[0];
        max_sim_gravitation_length = p_max_sim_gravitation_length;
        tau = p_tau;
        updateValues(!quiet);
        drivenCavityVelocity = d_drivenCavityVelocity * d_timestep;
        d_reynolds = d_domain_x_length * d_drivenCavityVelocity[0] / d_viscosity;
        if (!quiet) std::cout << "dim reynolds number: " << d_reynolds << std::endl;
    }

    CLbmSkeleton(CDomain<T> d, CVector<4, T> v) : domain(d), d_drivenCavityVelocity(v) {}
};
