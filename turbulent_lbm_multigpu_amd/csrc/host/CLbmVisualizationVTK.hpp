// CLbmVisualizationVTK.hpp -- legacy-ASCII VTK output of a (sub)domain
// (src/libvis/CLbmVisualizationVTK.hpp:8-100, src/libvis/VTK_Common.hpp):
// STRUCTURED_GRID with (X+1)(Y+1)(Z+1) points at the subdomain's physical
// position and CELL_DATA density, flag and velocity, one file
// `<name>.<uid>.<step>.vtk` per call of render().
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "CLbmSolver.hpp"

template <typename T>
class CLbmVisualizationVTK {
    int _UID;
    std::string _file_name;
    CLbmSolver<T> *_solver = nullptr;
    std::vector<T> velocity, density;
    std::vector<int> flags;

public:
    CLbmVisualizationVTK(int UID, std::string file_name) : _UID(UID), _file_name(std::move(file_name)) {}

    void setup(CLbmSolver<T> *solver) {
        _solver = solver;
        const size_t n = (size_t)solver->domain_cells.elements();
        velocity.assign(3 * n, T(0));
        density.assign(n, T(0));
        flags.assign(n, 0);
    }

    // returns false when the file cannot be written
    bool render(int increment = -1) {
        if (!_solver) return false;
        const int imax = _solver->domain_cells[0], jmax = _solver->domain_cells[1], kmax = _solver->domain_cells[2];
        const size_t total = (size_t)imax * jmax * kmax;
        _solver->storeVelocity(velocity.data());
        _solver->storeDensity(density.data());
        _solver->storeFlags(flags.data());
        const T dx = _solver->d_cell_length;
        const CVector<3, int> origin = _solver->domain.getOrigin();
        char name[512];
        snprintf(name, sizeof(name), "%s.%i.%i.vtk", _file_name.c_str(), _UID, increment);
        FILE *fp = fopen(name, "w");
        if (!fp) {
            fprintf(stderr, "Failed to open %s\n", name);
            return false;
        }
        fprintf(fp, "# vtk DataFile Version 3.1\nD3Q19 lattice Boltzmann, MI355X\nASCII\n\n");
        fprintf(fp, "DATASET STRUCTURED_GRID\n");
        fprintf(fp, "DIMENSIONS  %i %i %i \n", imax + 1, jmax + 1, kmax + 1);
        fprintf(fp, "POINTS %i float\n\n", (imax + 1) * (jmax + 1) * (kmax + 1));
        const T ox = origin[0] * dx, oy = origin[1] * dx, oz = origin[2] * dx;
        for (int k = 0; k <= kmax; ++k)
            for (int j = 0; j <= jmax; ++j)
                for (int i = 0; i <= imax; ++i)
                    fprintf(fp, "%f %f %f\n", (double)(ox + i * dx), (double)(oy + j * dx), (double)(oz + k * dx));
        fprintf(fp, "\nCELL_DATA %i \n", (int)total);
        fprintf(fp, "SCALARS density float 1 \nLOOKUP_TABLE default \n");
        for (size_t a = 0; a < total; ++a) fprintf(fp, "%f\n", (double)density[a]);
        fprintf(fp, "\nSCALARS flag INT 1 \nLOOKUP_TABLE default \n");
        for (size_t a = 0; a < total; ++a) fprintf(fp, "%i\n", flags[a]);
        fprintf(fp, "\nVECTORS velocity float\n");
        for (size_t a = 0; a < total; ++a)
            fprintf(fp, "%f %f %f\n", (double)velocity[a], (double)velocity[total + a], (double)velocity[2 * total + a]);
        return fclose(fp) == 0;
    }
};
