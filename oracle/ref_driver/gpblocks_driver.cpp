// Driver for the reference's scalar CPU solver, used only to produce golden numbers.
//
// The reference's GrossPitaevskii/GPBlocks.cpp is compiled here from its own source
// file (included below with its `main` renamed, since that `main` returns right
// after emitting mesh code).  This driver then calls the reference's own
// VortexState relaxation and its integrateInTime() staggered-leapfrog loop
// (GPBlocks.cpp:451-609) for one small configuration and prints, at 17 digits, the
// per-iteration "normsq=... error=..." lines that loop writes, plus the vortex
// profile and the BCC face table.  GPBlocks.cpp:251 calls an unqualified `max`,
// so <algorithm> and `using std::max` are made visible first.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
using std::max;
#define main gpblocks_reference_main
#include "GPBlocks.cpp"
#undef main

static ddouble ratio_potential(const ddouble r, const ddouble z) { return potentialRZ(r, z); }

int main(int argc, char **argv)
{
    // argv: kappa g rmax zmax rstep zstep maxiters block_scale period iterations
    if (argc < 11) {
        std::fprintf(stderr, "usage: kappa g rmax zmax rstep zstep maxiters block_scale period iterations\n");
        return 2;
    }
    const uint kappa = (uint)std::atoi(argv[1]);
    const ddouble g = std::atof(argv[2]);
    const ddouble rmax = std::atof(argv[3]), zmax = std::atof(argv[4]);
    const ddouble rstep = std::atof(argv[5]), zstep = std::atof(argv[6]);
    const uint maxiters = (uint)std::atoi(argv[7]);
    const ddouble block_scale = std::atof(argv[8]);
    const ddouble period = std::atof(argv[9]);
    const uint iterations = (uint)std::atoi(argv[10]);
    std::cout.precision(17);

    VortexState state;
    state.setKappa(kappa);
    state.setG(g);
    state.setRange(0.0, rmax, zmax, rstep, zstep);
    state.iterateSolution(ratio_potential, maxiters, 1e-29);
    std::cout << "VORTEX_MU " << state.getMu() << std::endl;
    for (int i = 0; i <= 40; ++i) {
        const ddouble r = 0.15 * i, z = 0.11 * i;
        std::cout << "VORTEX_F " << r << " " << z << " " << state.getFunction(r, z) << " "
                  << state.getFunction(r, 0.0) << std::endl;
    }

    Buffer<uint> ind;
    const uint nx = 12, ny = 1000, nz = 1000000;
    getLaplacian(ind, nx, ny, nz);
    for (uint i = 0; i < ind.size(); ++i) std::cout << "BCC_IND " << i << " " << ind[i] << std::endl;
    Buffer<Vector3> pos;
    getPositions(pos);
    for (uint i = 0; i < pos.size(); ++i)
        std::cout << "BCC_POS " << i << " " << pos[i].x << " " << pos[i].y << " " << pos[i].z << std::endl;

    integrateInTime(state, block_scale, Vector3(-rmax, -rmax, -zmax), Vector3(rmax, rmax, zmax), period, iterations);
    return 0;
}
