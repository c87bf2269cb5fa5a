// Driver for the reference's scalar CPU solver, used only to produce golden numbers.
//
// The reference's GrossPitaevskii/GPBlocks.cpp is compiled here from its own source
// file (included below with its `main` renamed, since that `main` returns right
// after emitting mesh code).  This driver then calls the reference's own
// VortexState relaxation and its integrateInTime() staggered-leapfrog loop
// (GPBlocks.cpp:451-609) for one small configuration and prints, at 17 digits, the
// per-iteration "normsq=... error=..." lines that loop writes, plus the vortex
// profile, the BCC face table, and the grid and steps per iteration of the run (so that
// the time between two of those lines can be turned into node-updates per second).  GPBlocks.cpp:251 calls an unqualified `max`,
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

    // Grid and steps per iteration of the call below, restated from integrateInTime's own
    // set-up (GPBlocks.cpp:455-501), so that a caller timing the per-iteration output lines
    // knows how many unknowns and sweeps each iteration holds.
    {
        const Vector3 minp(-rmax, -rmax, -zmax), maxp(rmax, rmax, zmax);
        const Vector3 dom = maxp - minp;
        const uint xs = uint(dom.x / (block_scale * BLOCK_WIDTH.x)) + 1;
        const uint ys = uint(dom.y / (block_scale * BLOCK_WIDTH.y)) + 1;
        const uint zs = uint(dom.z / (block_scale * BLOCK_WIDTH.z)) + 1;
        const Vector3 q0 = 0.5 * (minp + maxp - block_scale * Vector3(BLOCK_WIDTH.x * xs, BLOCK_WIDTH.y * ys,
                                                                       BLOCK_WIDTH.z * zs));
        ddouble maxpot = 0.0;
        for (uint k = 0; k < zs; k++)
            for (uint j = 0; j < ys; j++)
                for (uint i = 0; i < xs; i++)
                    for (uint l = 0; l < pos.size(); l++) {
                        const Vector3 p(q0.x + block_scale * (i * BLOCK_WIDTH.x + pos[l].x),
                                        q0.y + block_scale * (j * BLOCK_WIDTH.y + pos[l].y),
                                        q0.z + block_scale * (k * BLOCK_WIDTH.z + pos[l].z));
                        const ddouble v = potentialV3(p) + g * state.getPsi(p).normsq();
                        if (v > maxpot) maxpot = v;
                    }
        Buffer<uint> lind;
        const uint bs = pos.size();
        const ddouble lapfac = -0.5 * getLaplacian(lind, bs, (xs + 1) * bs, (ys + 1) * (xs + 1) * bs) /
                               (block_scale * block_scale);
        const ddouble lapfac0 = (lind.size() / bs) * (-lapfac);
        const uint steps = uint(period * (maxpot + lapfac0)) + 1;
        std::cout << "GRID " << xs << " " << ys << " " << zs << " " << bs << std::endl;
        std::cout << "STEPS_PER_ITERATION " << steps << std::endl;
    }

    integrateInTime(state, block_scale, Vector3(-rmax, -rmax, -zmax), Vector3(rmax, rmax, zmax), period, iterations);
    return 0;
}
