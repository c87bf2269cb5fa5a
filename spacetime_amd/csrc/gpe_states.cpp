// Host-side initial states of libgpe: the axisymmetric vortex relaxation of the reference
// (GrossPitaevskii/VortexState.cpp) and its sampling on the DEC unknowns.
#include "gpe.h"
#include "lattice_tables.h"

#include <cmath>
#include <vector>

namespace {

struct Profile {
    int rsize, zsize;
    double rstep, zstep, rplug;
};

bool profile_of(const gpe_vortex_params *vp, Profile &p)
{
    if (!vp || !(vp->rstep > 0) || !(vp->zstep > 0) || !(vp->rmax > vp->rmin) || !(vp->zmax > 0)) return false;
    // setRange (VortexState.cpp:42-47) then setDimensions' clamps (:12-18)
    const unsigned rs = (unsigned)((vp->rmax - vp->rmin) / vp->rstep + 0.5);
    const unsigned zs = (unsigned)(vp->zmax / vp->zstep + 0.5);
    if (rs == 0 || zs == 0) return false;
    p.rsize = (int)rs;
    p.zsize = (int)zs;
    p.rstep = (vp->rmax - vp->rmin) / (double)rs;
    p.zstep = vp->zmax / (double)zs;
    p.rplug = vp->rmin > 0.0 ? vp->rmin : 0.0;
    return true;
}

// VortexState::getFunction (VortexState.cpp:220-238), bilinear in (r, |z|); the corner
// sample f[ii + zsize + 1] is taken as the reference does, including in its last z-row.
double profile_value(const Profile &p, const double *f, double r, double z)
{
    const double rr = (r - p.rplug) / p.rstep;
    if (rr < 0.0) return 0.0;
    const unsigned ir = (unsigned)rr;
    if (ir >= (unsigned)p.rsize) return 0.0;
    const double dr = rr - (double)ir;
    const double zz = std::fabs(z) / p.zstep;
    const unsigned iz = (unsigned)zz;
    if (iz >= (unsigned)p.zsize) return 0.0;
    const double dz = zz - (double)iz;
    const unsigned ii = ir * p.zsize + iz, n = (unsigned)(p.rsize * p.zsize);
    const double fr = ir + 1 < (unsigned)p.rsize ? f[ii + p.zsize] : 0.0;
    const double fz = iz + 1 < (unsigned)p.zsize ? f[ii + 1] : 0.0;
    const double frz = ii + p.zsize + 1 < n ? f[ii + p.zsize + 1] : 0.0;
    return (1.0 - dz) * ((1.0 - dr) * f[ii] + dr * fr) + dz * ((1.0 - dr) * fz + dr * frz);
}

}  // namespace

extern "C" {

int gpe_vortex_size(const gpe_vortex_params *vp, int *rsize, int *zsize)
{
    Profile p;
    if (!profile_of(vp, p)) return GPE_EINVAL;
    if (rsize) *rsize = p.rsize;
    if (zsize) *zsize = p.zsize;
    return p.rsize * p.zsize;
}

int gpe_vortex_relax(const gpe_vortex_params *vp, double *f, double *mu_out)
{
    Profile p;
    if (!profile_of(vp, p) || !f || !mu_out || vp->maxiters < 0) return GPE_EINVAL;
    const int R = p.rsize, Z = p.zsize;
    const double two_pi = 6.283185307179586476925286766559;
    // initial guess (setDimensions, VortexState.cpp:20-39)
    double prob = 0.0;
    for (int i = 0, ii = 0; i < R; ++i) {
        const double fr = 1.0 - std::cos(i * two_pi / R);
        const double r = p.rplug + i * p.rstep;
        int times = 1;
        for (int j = 0; j < Z; ++j, ++ii) {
            f[ii] = (1.0 + std::cos(j * (0.5 * two_pi) / Z)) * fr;
            prob += times * f[ii] * f[ii] * r;
            times = 2;
        }
    }
    prob *= two_pi * p.rstep * p.zstep;
    for (int ii = 0; ii < R * Z; ++ii) f[ii] *= 1.0 / std::sqrt(prob);
    double mu = 1.0;
    // relaxation (iterateSolution, VortexState.cpp:109-183): Gauss-Seidel sweeps over
    // (r, z) solving the radial+axial discretisation for f, then mu = <H>/<f|f> and renorm.
    const double w2 = two_pi * p.rstep * p.zstep;
    const double kk = 0.5 * vp->kappa * vp->kappa;
    const double lr = 0.5 / (p.rstep * p.rstep);
    const double lz = Z > 1 ? 0.5 / (p.zstep * p.zstep) : 0.0;
    const double ratio2 = vp->ratio * vp->ratio;
    for (int j = 0; j < Z; ++j) f[j] = 0.0;
    for (int it = 0; it < vp->maxiters; ++it) {
        double diff = 0.0;
        for (int i = 1, ii = Z; i < R; ++i) {
            const double r = p.rplug + i * p.rstep;
            const double lap = kk / (r * r) + 2.0 * (lr + lz);
            const double gr = 0.25 / (r * p.rstep);
            double dv = w2 * r;
            for (int j = 0; j < Z; ++j, ++ii) {
                const double zc = j * p.zstep;
                const double pot = 0.5 * (r * r + ratio2 * zc * zc) + vp->g * f[ii] * f[ii];
                const double fo = i + 1 < R ? f[ii + Z] : 0.0;  // outer r neighbour
                const double fi = f[ii - Z];                    // inner r neighbour
                const double fu = j + 1 < Z ? f[ii + 1] : 0.0;  // upper z neighbour
                const double fd = j > 0 ? f[ii - 1] : fu;       // lower (mirror at z = 0)
                const double old = f[ii];
                f[ii] = (gr * (fi - fo) - lr * (fi + fo) - lz * (fd + fu)) / (mu - lap - pot);
                const double d = old - f[ii];
                diff += d * d * dv;
                if (j == 0) dv *= 2.0;
            }
        }
        double nrm = 0.0, hf = 0.0;
        for (int i = 1, ii = Z; i < R; ++i) {
            const double r = p.rplug + i * p.rstep;
            const double lap = kk / (r * r) + 2.0 * (lr + lz);
            const double gr = 0.25 / (r * p.rstep);
            double dv = w2 * r;
            for (int j = 0; j < Z; ++j, ++ii) {
                const double zc = j * p.zstep;
                const double pot = 0.5 * (r * r + ratio2 * zc * zc) + vp->g * f[ii] * f[ii];
                const double fo = i + 1 < R ? f[ii + Z] : 0.0;
                const double fi = f[ii - Z];
                const double fu = j + 1 < Z ? f[ii + 1] : 0.0;
                const double fd = j > 0 ? f[ii - 1] : fu;
                nrm += f[ii] * f[ii] * dv;
                hf += ((pot + lap) * f[ii] + gr * (fi - fo) - lr * (fi + fo) - lz * (fd + fu)) * f[ii] * dv;
                if (j == 0) dv *= 2.0;
            }
        }
        mu = hf / nrm;
        const double sc = 1.0 / std::sqrt(nrm);
        for (int ii = Z; ii < R * Z; ++ii) f[ii] *= sc;
        if (diff < vp->eps) break;
    }
    double sum = 0.0;
    for (int ii = Z; ii < R * Z; ++ii) sum += f[ii];
    if (sum < 0.0)
        for (int ii = Z; ii < R * Z; ++ii) f[ii] = -f[ii];
    *mu_out = mu;
    return GPE_OK;
}

int gpe_vortex_fill(const gpe_config *cfg, const gpe_vortex_params *vp, const double *f, double *psi)
{
    Profile p;
    if (!cfg || !f || !psi || !profile_of(vp, p)) return GPE_EINVAL;
    if (cfg->lattice < 0 || cfg->lattice > 5 || cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return GPE_EINVAL;
    if (cfg->ncomp != 1 && cfg->ncomp != 3 && cfg->ncomp != 5) return GPE_EINVAL;
    const gpe::LatticeTable &lat = gpe::kLattices[cfg->lattice];
    const int V = lat.V;
    const double s = cfg->block_scale, *w = lat.block_width;
    const size_t n = (size_t)cfg->nx * cfg->ny * cfg->nz * V;
    for (size_t i = 0; i < 2 * n * (size_t)cfg->ncomp; ++i) psi[i] = 0.0;
    for (int z = 0; z < cfg->nz; ++z)
        for (int y = 0; y < cfg->ny; ++y)
            for (int x = 0; x < cfg->nx; ++x)
                for (int v = 0; v < V; ++v) {
                    const double px = cfg->p0[0] + s * (x * w[0] + lat.pos[3 * v + 0]);
                    const double py = cfg->p0[1] + s * (y * w[1] + lat.pos[3 * v + 1]);
                    const double pz = cfg->p0[2] + s * (z * w[2] + lat.pos[3 * v + 2]);
                    const double r = std::sqrt(px * px + py * py);
                    double re = 0.0, im = 0.0;
                    if (!(r < 1e-30)) {
                        re = profile_value(p, f, r, pz);
                        const double nr = px / r, ni = py / r;
                        for (int k = 0; k < vp->kappa; ++k) {
                            const double t = re * nr - im * ni;
                            im = re * ni + im * nr;
                            re = t;
                        }
                    }
                    const size_t i = (((size_t)z * cfg->ny + y) * cfg->nx + x) * V + v;
                    psi[2 * i] = re;
                    psi[2 * i + 1] = im;
                }
    return GPE_OK;
}

}  // extern "C"
