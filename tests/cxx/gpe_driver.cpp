// A C++ host program over the C ABI alone (no Python), in the shape of the reference's
// integrateInTime (Spin1GpeBCC.cu:1105-1743): set-up, a device-side start (Thomas-Fermi fill,
// imaginary time, polar projection, spinor noise, renormalisation), a knot ramp, forward-Euler
// start and leapfrog steps, a checkpoint pair written mid-run and resumed in a second context,
// observables and a density picture.  Prints one JSON line: the observables of both runs, whether the resumed run
// equals the uninterrupted one bit for bit, and the Hamiltonian bound.
//
// usage: gpe_driver <directory for the checkpoint files>
#include "gpe.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define CHECK(ctx, call)                                                                         \
    do {                                                                                         \
        const int rc_ = (call);                                                                  \
        if (rc_ != GPE_OK) {                                                                     \
            std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, gpe_last_error(ctx));       \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

static gpe_config make_config()
{
    gpe_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    cfg.lattice = GPE_LATTICE_BCC;
    cfg.ncomp = 3;
    cfg.precision = GPE_FP64;
    cfg.scheme = GPE_SCHEME_LEAPFROG3;
    cfg.nx = 12;
    cfg.ny = 10;
    cfg.nz = 8;
    const double cell = 0.8;  // block edge; the BCC table's block width is 2 sqrt 2
    cfg.block_scale = cell / (2.0 * std::sqrt(2.0));
    cfg.p0[0] = -0.5 * cell * cfg.nx;
    cfg.p0[1] = -0.5 * cell * cfg.ny;
    cfg.p0[2] = -0.5 * cell * cfg.nz;
    cfg.dt = 1e-4;
    cfg.trap[0] = 1.0;
    cfg.trap[1] = 1.0;
    cfg.trap[2] = 1.3;
    cfg.trap_off_time = 1e300;
    cfg.c0 = 1.0e3;
    cfg.c1 = 50.0;
    cfg.quad_sign = 1.0;
    cfg.q = 0.1;
    return cfg;
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    const gpe_config cfg = make_config();
    const int V = gpe_lattice_values(cfg.lattice);
    const size_t n = (size_t)cfg.ncomp * cfg.nx * cfg.ny * cfg.nz * V;
    const double knots[8] = {0.05, -0.02, 0.03, 0.0, -0.04, 0.02, 0.01, -0.03};
    const int steps = 100;

    gpe_ctx *a = nullptr;
    if (gpe_create(&cfg, &a) != GPE_OK) {
        std::fprintf(stderr, "gpe_create failed\n");
        return 1;
    }
    const double mu = 0.5 * std::pow(15.0 * cfg.c0 * 1.3 / (4.0 * M_PI), 0.4);
    const double w[6] = {0.0, 0.0, 1.0, 0.0, 0.0, 0.0};  // m = 0
    CHECK(a, gpe_fill_thomas_fermi(a, cfg.c0, mu, w, 0.0, 0, 0.0));
    CHECK(a, gpe_itp(a, 50, 5e-5));
    CHECK(a, gpe_project_spinor(a, 0, GPE_PHASE_POLAR, 0.0));
    CHECK(a, gpe_add_noise(a, 0, 0.01, 54363));
    gpe_observables_t o0;
    CHECK(a, gpe_observables(a, 0, &o0));
    CHECK(a, gpe_scale(a, 0, 1.0 / std::sqrt(o0.norm)));
    std::vector<double> psi0(2 * n);
    CHECK(a, gpe_get_state(a, 0, psi0.data()));

    CHECK(a, gpe_set_knot_ramp(a, 8, knots, 2 * steps * cfg.dt, 4e-3));
    CHECK(a, gpe_set_state(a, 0, psi0.data(), 0.0));
    CHECK(a, gpe_start(a, GPE_START_EULER, 0.0));
    CHECK(a, gpe_step(a, steps));
    char tag[512];
    CHECK(a, gpe_save_checkpoint(a, dir, "drv_", tag, (int)sizeof tag));
    CHECK(a, gpe_step(a, steps));
    gpe_observables_t oa;
    CHECK(a, gpe_observables(a, 0, &oa));
    double hmax = 0.0;
    CHECK(a, gpe_max_hamiltonian(a, 0, &hmax));
    std::vector<double> fa(2 * n);
    CHECK(a, gpe_get_state(a, 0, fa.data()));
    // the density picture the reference draws at each save (drawDensity, Spin1GpeBCC.cu:1557)
    const std::string picture = std::string(dir) + "/drv_" + tag + "ms.bmp";
    CHECK(a, gpe_write_bmp(a, 0, picture.c_str(), GPE_PICTURE_DENSITY, GPE_BASIS_Z));

    gpe_ctx *b = nullptr;
    if (gpe_create(&cfg, &b) != GPE_OK) {
        std::fprintf(stderr, "gpe_create failed\n");
        return 1;
    }
    CHECK(b, gpe_set_knot_ramp(b, 8, knots, 2 * steps * cfg.dt, 4e-3));
    int levels = 0;
    CHECK(b, gpe_load_checkpoint(b, dir, "drv_", tag, &levels));
    CHECK(b, gpe_step(b, steps));
    gpe_observables_t ob;
    CHECK(b, gpe_observables(b, 0, &ob));
    std::vector<double> fb(2 * n);
    CHECK(b, gpe_get_state(b, 0, fb.data()));
    const bool same = std::memcmp(fa.data(), fb.data(), fa.size() * sizeof(double)) == 0;

    std::printf("{\"unknowns\": %zu, \"tag\": \"%s\", \"levels\": %d, \"bitwise_resume\": %s, "
                "\"norm\": %.17g, \"energy\": %.17g, \"mz\": %.17g, \"time\": %.17g, "
                "\"norm_resumed\": %.17g, \"energy_resumed\": %.17g, \"max_hamiltonian\": %.17g}\n",
                n / cfg.ncomp, tag, levels, same ? "true" : "false", oa.norm, oa.energy, oa.magnetization[2],
                oa.time, ob.norm, ob.energy, hmax);
    gpe_destroy(b);
    gpe_destroy(a);
    return same ? 0 : 2;
}
