// Device copy of one particle source (start_generate) and its launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "start_gpu.h"

namespace sg {

struct SourceK {
    int32_t kind, energy, angle, flags;  // flags: species << kSpeciesShift (run 0)
    int64_t first, count;                // particles [first, first + count) of the population
    double energy_eV, p0, p1, mass;
    double pos[3];
    double theta, phi, expansion, displace;
    int64_t cell0, n_cells;              // surface source: its cells in the shared cell arrays
};

void launch_generate(hipStream_t s, const SourceK& S, uint32_t s_index, uint64_t seed, const double* centres,
                     const double* normals, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                     int32_t* flags);

}  // namespace sg
