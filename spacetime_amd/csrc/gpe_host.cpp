// Host runtime of libgpe: contexts, device memory, lattice tables, field ramps and the
// sweep loop behind the C ABI declared in include/gpe.h.
#include "gpe.h"
#include "gpe_internal.h"
#include "lattice_tables.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

using namespace gpe;

namespace {

struct Ramp {
    std::vector<double> q_val, q_dur, b_val, b_dur;
    std::vector<int> q_type, b_type;
    double tau = 1.0, time_scale = 1.0;
    double scale_q = 1.0, scale_b = 1.0, qscale_q = 0.0, qscale_b = 0.0;
    bool active = false;
};

// One channel of the piecewise ramp (getSignal, Spin1GpeBCC/AliceRingRamps.h:80-138).
void ramp_channel(const std::vector<double> &val, const std::vector<double> &dur, const std::vector<int> &type,
                  int width, double tau, double t, double *out)
{
    const int n = (int)dur.size();
    if (n == 0) {
        for (int k = 0; k < width; ++k) out[k] = 0.0;
        return;
    }
    int idx = 0;
    for (; idx < n; ++idx) {
        const double tin = t - dur[idx];
        if (tin < 0) break;
        t = tin;
    }
    if (idx >= n) {
        for (int k = 0; k < width; ++k) out[k] = val[(n - 1) * width + k];
        return;
    }
    for (int k = 0; k < width; ++k) {
        const double prev = idx > 0 ? val[(idx - 1) * width + k] : 0.0;
        const double cur = val[idx * width + k];
        switch (type[idx]) {
        case 1: out[k] = prev + t * (cur - prev) / dur[idx]; break;
        case 2: out[k] = prev + (cur - prev) * std::sqrt(t / dur[idx]); break;
        case 3: out[k] = prev + (cur - prev) * (1.0 - std::exp(-t / tau)); break;
        default: out[k] = cur; break;
        }
    }
}

}  // namespace

struct gpe_ctx {
    gpe_config cfg{};
    const LatticeTable *lat = nullptr;
    Geometry g{};    // one x per lane: observables, packing, adjoint
    Geometry gs{};   // the sweep's geometry (1 or 2 x per lane)
    Physics ph{};
    int fp64 = 0;
    size_t elem_bytes = 0;   // bytes of one complex element
    long long level_elems = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    void *level[2] = {nullptr, nullptr};
    // tables
    int *d_face = nullptr, *d_nface = nullptr, *d_facex = nullptr;
    float *d_w32 = nullptr, *d_wsum32 = nullptr, *d_pos32 = nullptr;
    double *d_w64 = nullptr, *d_wsum64 = nullptr, *d_pos64 = nullptr, *d_dv = nullptr;
    // fields per sweep
    FieldRec *d_fields = nullptr;
    int fields_cap = 0;
    FieldRec *h_fields = nullptr;  // pinned staging
    // reductions
    double *d_partials = nullptr;
    double *d_scalars = nullptr;   // [0..15] results / scales
    double *d_red = nullptr;       // FIN_SCRATCH doubles: first-stage results of launch_finalize
    // static field and ramp
    double bq = 0, bb[3] = {0, 0, 0}, bqq = 0, bbq[3] = {0, 0, 0};
    Ramp ramp;
    // time bookkeeping
    double t0 = 0.0;
    long long sweeps = 0;
    bool level1_valid = false;
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms_per_sweep = 0.0;
    std::string err;
};

namespace {

int fail(gpe_ctx *c, int code, const char *what)
{
    if (c) c->err = what;
    return code;
}

int hip_fail(gpe_ctx *c, hipError_t e, const char *where)
{
    if (c) {
        c->err = std::string(where) + ": " + hipGetErrorString(e);
    }
    return GPE_EDEVICE;
}

#define HIPCALL(ctx, call)                                     \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

double sweep_dt(const gpe_ctx *c) { return c->cfg.scheme == GPE_SCHEME_STAGGERED ? 0.5 * c->cfg.dt : c->cfg.dt; }

FieldRec field_at(const gpe_ctx *c, double t)
{
    FieldRec f{};
    f.t = t;
    f.sigma = 1.0;
    if (c->cfg.expansion_rate != 0.0 && t > c->cfg.expansion_start)
        f.sigma = 1.0 + c->cfg.expansion_rate * (t - c->cfg.expansion_start);
    if (c->ramp.active) {
        const Ramp &r = c->ramp;
        double sq, sb[3];
        const double tr = t * r.time_scale;
        ramp_channel(r.q_val, r.q_dur, r.q_type, 1, r.tau, tr, &sq);
        ramp_channel(r.b_val, r.b_dur, r.b_type, 3, r.tau, tr, sb);
        f.bq = r.scale_q * sq;
        f.bqq = r.qscale_q * sq;
        for (int k = 0; k < 3; ++k) {
            f.bb[k] = r.scale_b * sb[k];
            f.bbq[k] = r.qscale_b * sb[k];
        }
    } else {
        f.bq = c->bq;
        f.bqq = c->bqq;
        for (int k = 0; k < 3; ++k) {
            f.bb[k] = c->bb[k];
            f.bbq[k] = c->bbq[k];
        }
    }
    f.fbq = (float)f.bq;
    f.fbqq = (float)f.bqq;
    for (int k = 0; k < 3; ++k) {
        f.fbb[k] = (float)f.bb[k];
        f.fbbq[k] = (float)f.bbq[k];
    }
    f.fsigma = (float)f.sigma;
    f.fksc = (float)(1.0 / (f.sigma * f.sigma));
    return f;
}

// FP32 mirror of the physics constants (after any change to c->ph).
void physics_to_float(Physics &ph)
{
    PhysicsF &f = ph.f;
    for (int k = 0; k < 3; ++k) {
        f.cell[k] = (float)ph.cell[k];
        f.p0[k] = (float)ph.p0[k];
        f.trap2[k] = (float)ph.trap2[k];
        f.center[k] = (float)ph.center[k];
    }
    f.trap_offset = (float)ph.trap_offset;
    f.c0 = (float)ph.c0;
    f.c1 = (float)ph.c1;
    f.c2 = (float)ph.c2;
    f.alpha = (float)ph.alpha;
    f.quad_sign = (float)ph.quad_sign;
    f.q = (float)ph.q;
}

bool quad_active(const gpe_ctx *c)
{
    if (c->cfg.ncomp == 1) return false;
    if (c->ramp.active) return c->ramp.qscale_q != 0.0 || c->ramp.qscale_b != 0.0;
    return c->bqq != 0.0 || c->bbq[0] != 0.0 || c->bbq[1] != 0.0 || c->bbq[2] != 0.0;
}

Tables tables(const gpe_ctx *c)
{
    Tables t;
    t.face = c->d_face;
    t.w32 = c->d_w32;
    t.w64 = c->d_w64;
    t.wsum32 = c->d_wsum32;
    t.wsum64 = c->d_wsum64;
    t.pos32 = c->d_pos32;
    t.pos64 = c->d_pos64;
    t.dv = c->d_dv;
    t.nface = c->d_nface;
    t.facex = c->d_facex;
    return t;
}

int ensure_fields(gpe_ctx *c, int n)
{
    if (n <= c->fields_cap) return GPE_OK;
    int cap = c->fields_cap ? c->fields_cap : 1024;
    while (cap < n) cap *= 2;
    if (c->d_fields) {
        HIPCALL(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->d_fields);
        (void)hipHostFree(c->h_fields);
        c->d_fields = nullptr;
        c->h_fields = nullptr;
    }
    HIPCALL(c, hipMalloc((void **)&c->d_fields, sizeof(FieldRec) * cap));
    HIPCALL(c, hipHostMalloc((void **)&c->h_fields, sizeof(FieldRec) * cap, hipHostMallocDefault));
    c->fields_cap = cap;
    return GPE_OK;
}

// Upload the fields of sweeps [first, first + n) to d_fields[0..n).
int upload_fields(gpe_ctx *c, long long first, int n)
{
    int rc = ensure_fields(c, n);
    if (rc) return rc;
    HIPCALL(c, hipStreamSynchronize(c->stream));  // staging buffer may still feed a copy
    const double sdt = sweep_dt(c);
    for (int k = 0; k < n; ++k) c->h_fields[k] = field_at(c, c->t0 + (double)(first + k) * sdt);
    HIPCALL(c, hipMemcpyAsync(c->d_fields, c->h_fields, sizeof(FieldRec) * n, hipMemcpyHostToDevice, c->stream));
    return GPE_OK;
}

FastDiv make_fast_div(unsigned d)
{
    FastDiv f;
    f.d = d;
    unsigned l = 0;
    while ((1ull << l) < d) ++l;
    f.l = l;
    f.m = d <= 1 ? 1u : (unsigned)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    return f;
}

// Segments, tiles, waves and workgroups of a traversal from nx, ny, nz, V, ty and xpl.
// Wave indices must stay below 2^31 (the decode's divisions).
// Lanes of a traversal's waves that carry values: rows * pairs pair slots in whole waves.
static double lane_use(int rows, int pairs)
{
    const int slots = rows * pairs;
    return (double)slots / (64.0 * ((slots + 63) / 64));
}

bool finish_geometry(Geometry &g)
{
    g.ty = g.ty < g.ny ? g.ty : g.ny;
    g.pairs = (g.nx + 1) / 2;
    // x pairs: one row per wave leaves lanes idle unless a line holds a multiple of 128 values.
    // Lines of <= 64 (<= 32) values: the wave's pair slots spread over a group of 2 (4) rows of
    // the tile (a group size that divides the tile's rows).  Unit cells of <= 12 unknowns: the
    // group of 1-8 rows (dividing the tile's rows) with the best lane use when it fills at least
    // 5% more lanes (BCC 70^3 +9%, FCC 45^3 +7%, scalar FCC 160^3 +17%, spin-2 Cube 293^3 +8%,
    // spin-1 BCC 192^3 +3%; scalar Cube 293^3 -2%).  Larger cells (A15, C15, Z) take a group
    // only where it packs >= 90% of the lanes: groups of 4 rows (85%) cost A15 82^3 and Z 68^3
    // spin-2 5-7%, while 3 rows of A15's 41 pairs (96%, tiles of 6 or 12 rows) gain 6-8% and 7
    // rows of Z's 34 pairs (93%, tiles of 14) 5%; the tile heights with such divisors are
    // autotuner candidates.
    int rows = 1;
    if (g.xpl == 2 && g.slabs && g.ty > 1 && ((g.ty * g.lx) % 16 == 0 || g.ty >= g.ny)) {
        // slab layout, tiles of rows holding a multiple of 16 values (or one tile): the tile is
        // the group, one contiguous run of pairs from a 128-byte boundary
        rows = g.ty;
    } else if (g.xpl == 2) {
        const int want = g.nx <= 32 ? 4 : (g.nx <= 64 ? 2 : 1);
        while (rows < want && g.ty % (2 * rows) == 0) rows *= 2;
        const double floor_use = g.V <= 12 ? 0.0 : 0.9;
        double best = lane_use(rows, g.pairs) + 0.05;
        for (int r = 2; r <= 8; ++r)
            if (g.ty % r == 0 && lane_use(r, g.pairs) >= floor_use && lane_use(r, g.pairs) > best + 1e-9) {
                best = lane_use(r, g.pairs);
                rows = r;
            }
    }
    g.rows = rows;
    g.tyr = g.ty / rows;
    g.nseg = rows > 1 ? (rows * g.pairs + 63) / 64 : (g.nx + 64 * g.xpl - 1) / (64 * g.xpl);
    g.ntile = (g.ny + g.ty - 1) / g.ty;
    g.waves = (long long)g.ntile * g.tyr * g.nz * g.V * g.nseg;
    const long long blocks = (g.waves + 3) / 4;
    if (blocks >= (1ll << 29)) return false;  // wave indices below 2^31 (the decode's divisions)
    g.blocks = (int)blocks;
    g.f_nseg = make_fast_div((unsigned)g.nseg);
    g.f_V = make_fast_div((unsigned)g.V);
    g.f_ty = make_fast_div((unsigned)g.tyr);
    g.f_nz = make_fast_div((unsigned)g.nz);
    g.f_pairs = make_fast_div((unsigned)g.pairs);
    return true;
}

// Layout of a configuration's levels and its default traversal (one x per lane, tiles of 8 rows,
// XCD remap), before finish_geometry.
Geometry base_geometry(const gpe_config &cfg)
{
    Geometry g{};
    g.nx = cfg.nx;
    g.ny = cfg.ny;
    g.nz = cfg.nz;
    g.V = kLattices[cfg.lattice].V;
    g.deg = kLattices[cfg.lattice].deg;
    g.C = cfg.ncomp;
    g.xpl = 1;
    g.ty = cfg.ny < 8 ? cfg.ny : 8;
    // x-lines start on 128-byte boundaries unless that pads a line by more than 1/20 (the padding
    // is fetched and written in part with every line), then on 16 / 32 bytes (an even number of
    // values, which keeps the x-pair accesses 16-byte aligned).  Spin-2 FP32, even against
    // 128-byte lines: BCC 115^3 (11% padding) +8.5%, BCC 118^3 (8.5%) +5%, A15 76^3 (5.3%) +2.4%;
    // lines padded by less ran 1-3% faster with the 128-byte start.
    const long long align = cfg.precision == GPE_FP64 ? 8 : 16;
    g.lx = ((long long)cfg.nx + align - 1) / align * align;
    if (20 * g.lx > 21 * (long long)cfg.nx) g.lx = ((long long)cfg.nx + 1) / 2 * 2;
    // FP32 lines whose even length is not a whole number of 128-byte lines: component slabs of
    // contiguous rows of that even length, whose row groups start on 128-byte boundaries
    // (finish_geometry) with no padding fetched (Cube 293^3: +2-3% against 304-value x-lines)
    const long long even = ((long long)cfg.nx + 1) / 2 * 2;
    g.slabs = cfg.precision != GPE_FP64 && even % align != 0;
    if (g.slabs) g.lx = even;
    if (g.slabs) {
        const long long ls = ((long long)cfg.ny * g.lx + align - 1) / align * align;
        g.ys = g.lx;
        g.cs = ls;
        g.vs = (long long)cfg.ncomp * ls;
        g.zs = g.vs * g.V;
    } else {
        g.cs = g.lx;
        g.vs = (long long)cfg.ncomp * g.lx;
        g.ys = g.vs * g.V;
        g.zs = g.ys * cfg.ny;
    }
    g.xcd_remap = 1;
    return g;
}

SweepArgs sweep_args(gpe_ctx *c, int mode_coef_sweep, const void *cur, void *next, double coef)
{
    SweepArgs a;
    a.g = c->gs;
    a.tb = tables(c);
    a.ph = c->ph;
    a.cur = cur;
    a.next = next;
    a.prev = nullptr;
    a.fields = c->d_fields;
    a.sweep = mode_coef_sweep;
    a.order = 0;
    a.quad = quad_active(c) ? 1 : 0;
    a.loss = c->cfg.alpha != 0.0 ? 1 : 0;
    a.coef = coef;
    a.scale_in = nullptr;
    a.partials = c->d_partials;
    return a;
}

int check_level(gpe_ctx *c, int level)
{
    if (level != 0 && level != 1) return fail(c, GPE_EINVAL, "level must be 0 or 1");
    return GPE_OK;
}

}  // namespace

namespace gpe {

int ctx_view(gpe_ctx *c, CtxView *out)
{
    if (!c || !out) return GPE_EINVAL;
    out->g = c->g;
    out->fp64 = c->fp64;
    out->ncomp = c->cfg.ncomp;
    out->device = c->device;
    out->stream = c->stream;
    out->level[0] = c->level[0];
    out->level[1] = c->level[1];
    return GPE_OK;
}

int ctx_fail(gpe_ctx *c, int code, const char *what) { return fail(c, code, what); }

}  // namespace gpe

extern "C" {

int gpe_lattice_values(int lattice)
{
    if (lattice < 0 || lattice > 5) return GPE_EINVAL;
    return kLattices[lattice].V;
}

const char *gpe_last_error(const gpe_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int gpe_create(const gpe_config *cfg, gpe_ctx **out)
{
    if (!cfg || !out) return GPE_EINVAL;
    *out = nullptr;
    if (cfg->lattice < 0 || cfg->lattice > 5) return GPE_EINVAL;
    if (cfg->ncomp != 1 && cfg->ncomp != 3 && cfg->ncomp != 5) return GPE_EINVAL;
    if (cfg->precision != GPE_FP32 && cfg->precision != GPE_FP64) return GPE_EINVAL;
    if (cfg->scheme != GPE_SCHEME_LEAPFROG3 && cfg->scheme != GPE_SCHEME_STAGGERED) return GPE_EINVAL;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return GPE_EINVAL;
    if (!(cfg->block_scale > 0.0) || !(cfg->dt > 0.0)) return GPE_EINVAL;
    if (!(cfg->expansion_rate >= 0.0)) return GPE_EINVAL;
    gpe_ctx *c = new (std::nothrow) gpe_ctx();
    if (!c) return GPE_ENOMEM;
    c->cfg = *cfg;
    c->lat = &kLattices[cfg->lattice];
    c->fp64 = cfg->precision == GPE_FP64;
    c->elem_bytes = c->fp64 ? 16 : 8;

    Geometry &g = c->g;
    g = base_geometry(*cfg);
    c->gs = g;
    c->gs.xpl = default_xpl(c->fp64, g.deg, cfg->ncomp, cfg->nx);
    if (!finish_geometry(g) || !finish_geometry(c->gs)) {
        delete c;
        return GPE_EINVAL;
    }
    // 32-bit index arithmetic in the kernels: offsets within a z plane and to its neighbours
    // (line_index, the per-face line offsets, row strides: below 4 zs) and line numbers
    // ((z ny + y) V + v)
    if (4 * g.zs >= (1ll << 31) || (long long)g.nz * g.ny * g.V >= (1ll << 31)) {
        delete c;
        return GPE_EINVAL;
    }
    c->level_elems = g.nz * g.zs;

    Physics &ph = c->ph;
    for (int k = 0; k < 3; ++k) ph.cell[k] = cfg->block_scale * c->lat->block_width[k];
    for (int k = 0; k < 3; ++k) {
        ph.p0[k] = cfg->p0[k];
        ph.trap[k] = cfg->trap[k];
        ph.trap2[k] = cfg->trap[k] * cfg->trap[k];
        ph.center[k] = cfg->center[k];
    }
    ph.trap_offset = cfg->trap_offset;
    ph.trap_off_time = cfg->trap_off_time;
    ph.c0 = cfg->c0;
    ph.c1 = cfg->c1;
    ph.c2 = cfg->c2;
    ph.alpha = cfg->alpha;
    ph.quad_sign = cfg->quad_sign;
    ph.q = cfg->q;
    physics_to_float(ph);

    int rc = GPE_OK;
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) rc = hip_fail(c, e, "hipGetDevice");
    if (!rc && (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        rc = hip_fail(c, e, "hipStreamCreate");
    const size_t lbytes = (size_t)c->level_elems * c->elem_bytes;
    for (int l = 0; l < 2 && !rc; ++l) {
        if ((e = hipMalloc(&c->level[l], lbytes)) != hipSuccess)
            rc = GPE_ENOMEM;
        else if ((e = hipMemsetAsync(c->level[l], 0, lbytes, c->stream)) != hipSuccess)
            rc = hip_fail(c, e, "hipMemset");
    }
    // tables: weights w_f = -1/2 hodge_f / s^2 (Spin1GpeBCC.cu:859,1169)
    const int V = g.V, deg = g.deg;
    const double s = cfg->block_scale;
    std::vector<double> w64(V * deg), wsum64(V, 0.0), pos64(3 * V), dv(V);
    std::vector<float> w32(V * deg), wsum32(V), pos32(3 * V);
    std::vector<int> nface(V), facex(2 * V * deg);
    for (int v = 0; v < V; ++v) {
        for (int f = 0; f < deg; ++f) {
            const int pk = c->lat->face[v * deg + f];
            const long long dy = ((pk >> 2) & 3) - 1, dz = ((pk >> 4) & 3) - 1, t = pk >> 8;
            facex[2 * (v * deg + f)] = pk;
            // |offset| < 3 zs, checked against the int range at creation
            facex[2 * (v * deg + f) + 1] = (int)(dz * g.zs + dy * g.ys + (t - v) * g.vs);
        }
        int nf = deg;  // pad faces (zero Hodge weight) trail the real ones
        while (nf > 0 && c->lat->hodge[v * deg + nf - 1] == 0.0) --nf;
        nface[v] = nf;
        for (int f = 0; f < deg; ++f) {
            const double w = -0.5 * c->lat->hodge[v * deg + f] / (s * s);
            w64[v * deg + f] = w;
            w32[v * deg + f] = (float)w;
            wsum64[v] += w;
        }
        wsum32[v] = (float)wsum64[v];
        for (int k = 0; k < 3; ++k) {
            pos64[3 * v + k] = s * c->lat->pos[3 * v + k];
            pos32[3 * v + k] = (float)pos64[3 * v + k];
        }
        dv[v] = s * s * s * c->lat->vol[v];
    }
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (rc) return;
        hipError_t ee = hipMalloc(dst, bytes);
        if (ee != hipSuccess) {
            rc = GPE_ENOMEM;
            return;
        }
        ee = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (ee != hipSuccess) rc = hip_fail(c, ee, "hipMemcpy tables");
    };
    up((void **)&c->d_face, c->lat->face, sizeof(int) * V * deg);
    up((void **)&c->d_nface, nface.data(), sizeof(int) * V);
    up((void **)&c->d_facex, facex.data(), sizeof(int) * 2 * V * deg);
    up((void **)&c->d_w32, w32.data(), sizeof(float) * V * deg);
    up((void **)&c->d_w64, w64.data(), sizeof(double) * V * deg);
    up((void **)&c->d_wsum32, wsum32.data(), sizeof(float) * V);
    up((void **)&c->d_wsum64, wsum64.data(), sizeof(double) * V);
    up((void **)&c->d_pos32, pos32.data(), sizeof(float) * 3 * V);
    up((void **)&c->d_pos64, pos64.data(), sizeof(double) * 3 * V);
    up((void **)&c->d_dv, dv.data(), sizeof(double) * V);
    if (!rc && hipMalloc((void **)&c->d_partials, sizeof(double) * OBS_COUNT * (size_t)g.blocks) != hipSuccess)
        rc = GPE_ENOMEM;
    if (!rc && hipMalloc((void **)&c->d_scalars, sizeof(double) * 64) != hipSuccess) rc = GPE_ENOMEM;
    if (!rc && hipMalloc((void **)&c->d_red, sizeof(double) * FIN_SCRATCH) != hipSuccess) rc = GPE_ENOMEM;
    if (!rc) rc = ensure_fields(c, 1024);
    if (!rc && (e = hipEventCreate(&c->ev0)) != hipSuccess) rc = hip_fail(c, e, "hipEventCreate");
    if (!rc && (e = hipEventCreate(&c->ev1)) != hipSuccess) rc = hip_fail(c, e, "hipEventCreate");
    if (!rc && (e = hipStreamSynchronize(c->stream)) != hipSuccess) rc = hip_fail(c, e, "create sync");
    if (rc) {
        gpe_destroy(c);
        return rc;
    }
    *out = c;
    return GPE_OK;
}

int gpe_destroy(gpe_ctx *c)
{
    if (!c) return GPE_EINVAL;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int l = 0; l < 2; ++l)
        if (c->level[l]) (void)hipFree(c->level[l]);
    void *ptrs[] = {c->d_face, c->d_nface, c->d_facex, c->d_w32, c->d_w64, c->d_wsum32, c->d_wsum64, c->d_pos32,
                    c->d_pos64, c->d_dv, c->d_fields, c->d_partials, c->d_scalars, c->d_red};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_fields) (void)hipHostFree(c->h_fields);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GPE_OK;
}

static int set_state_impl(gpe_ctx *c, int level, const double *psi, double t, hipMemcpyKind kind)
{
    if (!c) return GPE_EINVAL;
    if (!psi) return fail(c, GPE_EINVAL, "null state pointer");
    int rc = check_level(c, level);
    if (rc) return rc;
    const Geometry &g = c->g;
    const size_t nbytes = (size_t)c->cfg.ncomp * g.nz * g.ny * g.nx * g.V * 2 * sizeof(double);
    const double *src = psi;
    double *staging = nullptr;
    if (kind == hipMemcpyHostToDevice) {
        HIPCALL(c, hipMallocAsync((void **)&staging, nbytes, c->stream));
        HIPCALL(c, hipMemcpyAsync(staging, psi, nbytes, hipMemcpyHostToDevice, c->stream));
        src = staging;
    }
    hipError_t e = launch_pack(c->fp64, g, c->cfg.ncomp, src, c->level[level], c->stream);
    if (staging) (void)hipFreeAsync(staging, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "pack");
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (level == 0) {
        c->t0 = t;
        c->sweeps = 0;
        c->level1_valid = false;
    } else {
        c->level1_valid = true;
    }
    return GPE_OK;
}

int gpe_set_state(gpe_ctx *c, int level, const double *psi, double t)
{
    return set_state_impl(c, level, psi, t, hipMemcpyHostToDevice);
}

int gpe_set_state_device(gpe_ctx *c, int level, const double *dpsi, double t)
{
    return set_state_impl(c, level, dpsi, t, hipMemcpyDeviceToDevice);
}

static int get_state_impl(gpe_ctx *c, int level, double *psi, bool host)
{
    if (!c) return GPE_EINVAL;
    if (!psi) return fail(c, GPE_EINVAL, "null state pointer");
    int rc = check_level(c, level);
    if (rc) return rc;
    const Geometry &g = c->g;
    const size_t nbytes = (size_t)c->cfg.ncomp * g.nz * g.ny * g.nx * g.V * 2 * sizeof(double);
    double *dst = psi;
    double *staging = nullptr;
    if (host) {
        HIPCALL(c, hipMallocAsync((void **)&staging, nbytes, c->stream));
        dst = staging;
    }
    hipError_t e = launch_unpack(c->fp64, g, c->cfg.ncomp, c->level[level], dst, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "unpack");
    if (host) {
        HIPCALL(c, hipMemcpyAsync(psi, staging, nbytes, hipMemcpyDeviceToHost, c->stream));
        (void)hipFreeAsync(staging, c->stream);
    }
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_get_state(gpe_ctx *c, int level, double *psi) { return get_state_impl(c, level, psi, true); }
int gpe_get_state_device(gpe_ctx *c, int level, double *dpsi) { return get_state_impl(c, level, dpsi, false); }

int gpe_set_field(gpe_ctx *c, double bq, const double bb[3], double bq_quad, const double bb_quad[3])
{
    if (!c) return GPE_EINVAL;
    c->ramp.active = false;
    c->bq = bq;
    c->bqq = bq_quad;
    for (int k = 0; k < 3; ++k) {
        c->bb[k] = bb ? bb[k] : 0.0;
        c->bbq[k] = bb_quad ? bb_quad[k] : 0.0;
    }
    return GPE_OK;
}

int gpe_set_ramp(gpe_ctx *c, const gpe_ramp *r)
{
    if (!c) return GPE_EINVAL;
    if (!r) {
        c->ramp.active = false;
        return GPE_OK;
    }
    if (r->nq < 0 || r->nb < 0) return fail(c, GPE_EINVAL, "negative segment count");
    if ((r->nq && (!r->q_val || !r->q_dur || !r->q_type)) || (r->nb && (!r->b_val || !r->b_dur || !r->b_type)))
        return fail(c, GPE_EINVAL, "null ramp array");
    Ramp &d = c->ramp;
    d.q_val.assign(r->q_val, r->q_val + r->nq);
    d.q_dur.assign(r->q_dur, r->q_dur + r->nq);
    d.q_type.assign(r->q_type, r->q_type + r->nq);
    d.b_val.assign(r->b_val, r->b_val + 3 * r->nb);
    d.b_dur.assign(r->b_dur, r->b_dur + r->nb);
    d.b_type.assign(r->b_type, r->b_type + r->nb);
    for (int t : d.q_type)
        if (t < 0 || t > 3) return fail(c, GPE_EINVAL, "bad ramp segment type");
    for (int t : d.b_type)
        if (t < 0 || t > 3) return fail(c, GPE_EINVAL, "bad ramp segment type");
    d.tau = r->tau;
    d.time_scale = r->time_scale;
    d.scale_q = r->scale_q;
    d.scale_b = r->scale_b;
    d.qscale_q = r->qscale_q;
    d.qscale_b = r->qscale_b;
    d.active = true;
    return GPE_OK;
}

int gpe_start(gpe_ctx *c, int mode, double mu)
{
    if (!c) return GPE_EINVAL;
    const double sdt = sweep_dt(c);
    if (mode == GPE_START_EULER) {
        int rc = upload_fields(c, c->sweeps, 1);
        if (rc) return rc;
        SweepArgs a = sweep_args(c, 0, c->level[0], c->level[1], sdt);
        hipError_t e = launch_sweep(c->fp64, c->cfg.ncomp, MODE_EULER, a, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "forward Euler");
        std::swap(c->level[0], c->level[1]);
        c->sweeps += 1;
    } else if (mode == GPE_START_PHASE) {
        // psi(t - sdt) = exp(+i mu sdt) psi(t)  (GPBlocks.cpp:516-523)
        hipError_t e = launch_phase_copy(c->fp64, c->level[0], c->level[1], c->level_elems, std::cos(mu * sdt),
                                         std::sin(mu * sdt), c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "phase start");
    } else if (mode == GPE_START_COPY) {
        HIPCALL(c, hipMemcpyAsync(c->level[1], c->level[0], (size_t)c->level_elems * c->elem_bytes,
                                  hipMemcpyDeviceToDevice, c->stream));
    } else {
        return fail(c, GPE_EINVAL, "unknown start mode");
    }
    c->level1_valid = true;
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_step(gpe_ctx *c, int nsteps)
{
    if (!c) return GPE_EINVAL;
    if (nsteps < 0) return fail(c, GPE_EINVAL, "negative step count");
    if (!c->level1_valid) return fail(c, GPE_ESTATE, "second time level not set: call gpe_start");
    if (nsteps == 0) return GPE_OK;
    const int per = c->cfg.scheme == GPE_SCHEME_STAGGERED ? 2 : 1;
    const long long total = (long long)nsteps * per;
    const double coef = 2.0 * sweep_dt(c);
    const int chunk = 1 << 16;
    double ms_total = 0.0;
    for (long long done = 0; done < total;) {
        const int n = (int)std::min<long long>(chunk, total - done);
        int rc = upload_fields(c, c->sweeps, n);
        if (rc) return rc;
        HIPCALL(c, hipEventRecord(c->ev0, c->stream));
        for (int k = 0; k < n; ++k) {
            SweepArgs a = sweep_args(c, k, c->level[0], c->level[1], coef);
            a.order = (int)(c->sweeps & 1);
            hipError_t e = launch_sweep(c->fp64, c->cfg.ncomp, MODE_UPDATE, a, c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "sweep launch");
            std::swap(c->level[0], c->level[1]);
            c->sweeps += 1;
        }
        HIPCALL(c, hipEventRecord(c->ev1, c->stream));
        HIPCALL(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCALL(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        ms_total += ms;
        done += n;
    }
    c->last_ms_per_sweep = ms_total / (double)total;
    return GPE_OK;
}

int gpe_itp(gpe_ctx *c, int nsteps, double dt_itp)
{
    if (!c) return GPE_EINVAL;
    if (nsteps < 0 || !(dt_itp > 0.0)) return fail(c, GPE_EINVAL, "bad ITP arguments");
    int rc = upload_fields(c, c->sweeps, 1);  // field and time of level 0
    if (rc) return rc;
    double *norm = c->d_scalars + 0;
    double *scale[2] = {c->d_scalars + 1, c->d_scalars + 2};
    // initial scale from the norm of level 0
    ObsArgs oa;
    oa.g = c->g;
    oa.tb = tables(c);
    oa.ph = c->ph;
    oa.psi = c->level[0];
    oa.field = field_at(c, c->t0 + c->sweeps * sweep_dt(c));
    oa.quad = quad_active(c);
    oa.partials = c->d_partials;
    hipError_t e = launch_observables(c->fp64, c->cfg.ncomp, oa, c->stream);
    if (e == hipSuccess) e = launch_finalize(c->d_partials, c->g.blocks, OBS_COUNT, OBS_MAXDENS, c->d_scalars + 8, c->d_red, c->stream);
    if (e == hipSuccess) e = launch_norm_to_scale(c->d_scalars + 8 + OBS_NORM, 1.0, scale[0], c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "ITP initial norm");
    for (int i = 0; i < nsteps; ++i) {
        SweepArgs a = sweep_args(c, 0, c->level[0], c->level[1], dt_itp);
        a.scale_in = scale[i & 1];
        a.order = i & 1;
        e = launch_sweep(c->fp64, c->cfg.ncomp, MODE_ITP, a, c->stream);
        if (e == hipSuccess) e = launch_finalize(c->d_partials, c->gs.blocks, 1, -1, norm, c->d_red, c->stream);
        if (e == hipSuccess) e = launch_norm_to_scale(norm, 1.0, scale[(i + 1) & 1], c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "ITP step");
        std::swap(c->level[0], c->level[1]);
    }
    e = launch_scale(c->fp64, c->level[0], c->level_elems, 1.0, scale[nsteps & 1], c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "ITP final scale");
    c->level1_valid = false;
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_observables(gpe_ctx *c, int level, gpe_observables_t *out)
{
    if (!c) return GPE_EINVAL;
    if (!out) return fail(c, GPE_EINVAL, "null output");
    int rc = check_level(c, level);
    if (rc) return rc;
    const double t = c->t0 + (double)(c->sweeps - level) * sweep_dt(c);
    ObsArgs oa;
    oa.g = c->g;
    oa.tb = tables(c);
    oa.ph = c->ph;
    oa.psi = c->level[level];
    oa.field = field_at(c, t);
    oa.quad = quad_active(c);
    oa.partials = c->d_partials;
    hipError_t e = launch_observables(c->fp64, c->cfg.ncomp, oa, c->stream);
    if (e == hipSuccess) e = launch_finalize(c->d_partials, c->g.blocks, OBS_COUNT, OBS_MAXDENS, c->d_scalars + 8, c->d_red, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "observables");
    double h[OBS_COUNT];
    HIPCALL(c, hipMemcpyAsync(h, c->d_scalars + 8, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    out->norm = h[OBS_NORM];
    out->energy = h[OBS_ENERGY];
    out->expect_h = h[OBS_EXPECT_H];
    out->magnetization[0] = h[OBS_MX];
    out->magnetization[1] = h[OBS_MY];
    out->magnetization[2] = h[OBS_MZ];
    out->max_density = h[OBS_MAXDENS];
    out->time = t;
    out->finite = (h[OBS_NONFINITE] == 0.0 && std::isfinite(h[OBS_NORM]) && std::isfinite(h[OBS_ENERGY])) ? 1 : 0;
    out->spin = h[OBS_SPIN];
    out->moment[0] = h[OBS_RX];
    out->moment[1] = h[OBS_RY];
    out->moment[2] = h[OBS_RZ];
    out->ref_spin = h[OBS_REF_SPIN];
    out->ref_magnetization_z = h[OBS_REF_MZ];
    return out->finite ? GPE_OK : fail(c, GPE_ERANGE, "non-finite value in the state");
}

// ------------------------------------------------------------------ observable series --

// The reference's time-series file (Spin1GpeBCC.cu:1497-1504, 1659-1712): eight MATLAB
// vectors that each saved file extends ("times = [times, ...];"), one column per append.
struct gpe_series {
    std::vector<double> col[8];  // times, Bq, Bz, Spin, mag_x, mag_y, mag_z, norm
};

int gpe_series_create(gpe_series **out)
{
    if (!out) return GPE_EINVAL;
    *out = new (std::nothrow) gpe_series();
    return *out ? GPE_OK : GPE_ENOMEM;
}

int gpe_series_destroy(gpe_series *s)
{
    delete s;
    return GPE_OK;
}

int gpe_series_append(gpe_series *s, gpe_ctx *c, int level)
{
    if (!s || !c) return GPE_EINVAL;
    gpe_observables_t o;
    const int rc = gpe_observables(c, level, &o);
    if (rc && rc != GPE_ERANGE) return rc;
    const FieldRec f = field_at(c, o.time);
    const double v[8] = {o.time * (c->ramp.active ? c->ramp.time_scale : 1.0),
                         f.bq,
                         f.bb[0] + f.bb[1] + f.bb[2],
                         o.ref_spin,
                         o.magnetization[0],
                         o.magnetization[1],
                         o.ref_magnetization_z,
                         o.norm};
    for (int k = 0; k < 8; ++k) s->col[k].push_back(v[k]);
    return rc;
}

int gpe_series_save(gpe_series *s, const char *path)
{
    if (!s || !path) return GPE_EINVAL;
    static const char *const names[8] = {"times", "Bq", "Bz", "Spin", "mag_x", "mag_y", "mag_z", "norm"};
    FILE *f = std::fopen(path, "w");
    if (!f) return GPE_EINVAL;
    bool ok = true;
    for (int k = 0; k < 8 && ok; ++k) {
        ok = std::fprintf(f, "%s = [%s", names[k], names[k]) > 0;
        for (double x : s->col[k]) ok = ok && std::fprintf(f, ", %.18f", x) > 0;
        ok = ok && std::fprintf(f, "];\n") > 0;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return GPE_EDEVICE;
    for (auto &v : s->col) v.clear();  // the next file starts a new series, as the reference's does
    return GPE_OK;
}

// k_aux over a level with an optional device reference state; h[0] = maximum, h[1] = sum.
static int aux_reduce(gpe_ctx *c, int level, const double *dref, double pre, double pim, double h[AUX_COUNT])
{
    const double t = c->t0 + (double)(c->sweeps - level) * sweep_dt(c);
    AuxArgs aa;
    aa.g = c->g;
    aa.tb = tables(c);
    aa.ph = c->ph;
    aa.psi = c->level[level];
    aa.ref = (const double2 *)dref;
    aa.phase_re = pre;
    aa.phase_im = pim;
    aa.field = field_at(c, t);
    aa.partials = c->d_partials;
    hipError_t e = launch_aux(c->fp64, c->cfg.ncomp, aa, c->stream);
    if (e == hipSuccess) e = launch_finalize(c->d_partials, c->g.blocks, AUX_COUNT, AUX_MAXHAM, c->d_scalars + 8, c->d_red, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "local reduction");
    HIPCALL(c, hipMemcpyAsync(h, c->d_scalars + 8, sizeof(double) * AUX_COUNT, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_max_hamiltonian(gpe_ctx *c, int level, double *out)
{
    if (!c) return GPE_EINVAL;
    if (!out) return fail(c, GPE_EINVAL, "null output");
    if (c->cfg.ncomp == 5) return fail(c, GPE_EINVAL, "the Hamiltonian bound is defined for scalar and spin-1 states");
    int rc = check_level(c, level);
    if (rc) return rc;
    double h[AUX_COUNT];
    if ((rc = aux_reduce(c, level, nullptr, 1.0, 0.0, h))) return rc;
    *out = h[AUX_MAXHAM];
    return std::isfinite(*out) ? GPE_OK : fail(c, GPE_ERANGE, "non-finite value in the state");
}

int gpe_weighted_diff_device(gpe_ctx *c, int level, const double *dref, double phase_re, double phase_im, double *out)
{
    if (!c) return GPE_EINVAL;
    if (!dref || !out) return fail(c, GPE_EINVAL, "null reference state or output");
    int rc = check_level(c, level);
    if (rc) return rc;
    double h[AUX_COUNT];
    if ((rc = aux_reduce(c, level, dref, phase_re, phase_im, h))) return rc;
    *out = h[AUX_WDIFF];
    return GPE_OK;
}

int gpe_weighted_diff(gpe_ctx *c, int level, const double *ref, double phase_re, double phase_im, double *out)
{
    if (!c) return GPE_EINVAL;
    if (!ref || !out) return fail(c, GPE_EINVAL, "null reference state or output");
    int rc = check_level(c, level);
    if (rc) return rc;
    const size_t bytes = 2 * sizeof(double) * (size_t)c->cfg.ncomp * c->cfg.nx * c->cfg.ny * c->cfg.nz * c->g.V;
    double *d = nullptr;
    if (hipMalloc((void **)&d, bytes) != hipSuccess) return fail(c, GPE_ENOMEM, "reference state does not fit");
    hipError_t e = hipMemcpyAsync(d, ref, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) rc = gpe_weighted_diff_device(c, level, d, phase_re, phase_im, out);
    else rc = hip_fail(c, e, "reference upload");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return rc;
}

int gpe_field_width(int kind)
{
    switch (kind) {
    case GPE_FIELD_DENSITY: return 1;
    case GPE_FIELD_SPIN: return 3;
    case GPE_FIELD_SPIN_NORM: return 1;
    case GPE_FIELD_UV_THETA: return 7;
    default: return GPE_EINVAL;
    }
}

int gpe_local_fields(gpe_ctx *c, int level, int kind, double *out)
{
    if (!c) return GPE_EINVAL;
    const int width = gpe_field_width(kind);
    if (width < 0) return fail(c, GPE_EINVAL, "unknown field kind");
    if (!out) return fail(c, GPE_EINVAL, "null output");
    if (kind == GPE_FIELD_UV_THETA && c->cfg.ncomp != 3) return fail(c, GPE_EINVAL, "u, v, theta need spin-1");
    int rc = check_level(c, level);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)width * (size_t)c->cfg.nx * c->cfg.ny * c->cfg.nz * c->g.V;
    double *d = nullptr;
    if (hipMalloc((void **)&d, bytes) != hipSuccess) return fail(c, GPE_ENOMEM, "field map does not fit");
    hipError_t e = launch_local_fields(c->fp64, c->cfg.ncomp, c->g, c->level[level], kind, d, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return e == hipSuccess ? GPE_OK : hip_fail(c, e, "local fields");
}

int gpe_scale(gpe_ctx *c, int level, double factor)
{
    if (!c) return GPE_EINVAL;
    int rc = check_level(c, level);
    if (rc) return rc;
    hipError_t e = launch_scale(c->fp64, c->level[level], c->level_elems, factor, nullptr, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "scale");
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_sync(gpe_ctx *c)
{
    if (!c) return GPE_EINVAL;
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return GPE_OK;
}

int gpe_time(const gpe_ctx *c, double *t, long long *sweeps)
{
    if (!c) return GPE_EINVAL;
    if (t) *t = c->t0 + (double)c->sweeps * sweep_dt(c);
    if (sweeps) *sweeps = c->sweeps;
    return GPE_OK;
}

int gpe_fill_thomas_fermi(gpe_ctx *c, double g, double mu, const double *weights, double noise, unsigned long long seed,
                          double t)
{
    if (!c) return GPE_EINVAL;
    if (!(g > 0.0) || !weights || noise < 0.0) return fail(c, GPE_EINVAL, "bad Thomas-Fermi fill arguments");
    FillArgs f;
    f.mu = mu;
    f.g = g;
    f.amp = noise;
    f.seed = seed;
    for (int k = 0; k < 10; ++k) f.w[k] = k < 2 * c->cfg.ncomp ? weights[k] : 0.0;
    hipError_t e = launch_fill_tf(c->fp64, c->g, tables(c), c->ph, c->cfg.ncomp, f, c->level[0], c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "fill");
    HIPCALL(c, hipStreamSynchronize(c->stream));
    c->t0 = t;
    c->sweeps = 0;
    c->level1_valid = false;
    return GPE_OK;
}

int gpe_set_trap(gpe_ctx *c, const double trap[3])
{
    if (!c || !trap) return GPE_EINVAL;
    for (int k = 0; k < 3; ++k) {
        c->cfg.trap[k] = trap[k];
        c->ph.trap[k] = trap[k];
        c->ph.trap2[k] = trap[k] * trap[k];
    }
    physics_to_float(c->ph);
    return GPE_OK;
}

int gpe_get_config(const gpe_ctx *c, gpe_config *out)
{
    if (!c || !out) return GPE_EINVAL;
    *out = c->cfg;
    return GPE_OK;
}

int gpe_set_time(gpe_ctx *c, double t0, long long sweeps)
{
    if (!c) return GPE_EINVAL;
    if (sweeps < 0 || !std::isfinite(t0)) return fail(c, GPE_EINVAL, "bad clock");
    c->t0 = t0;
    c->sweeps = sweeps;
    return GPE_OK;
}

int gpe_clock(const gpe_ctx *c, double *t0, long long *sweeps)
{
    if (!c) return GPE_EINVAL;
    if (t0) *t0 = c->t0;
    if (sweeps) *sweeps = c->sweeps;
    return GPE_OK;
}

int gpe_project_spinor(gpe_ctx *c, int level, int phase, double param)
{
    if (!c) return GPE_EINVAL;
    int rc = check_level(c, level);
    if (rc) return rc;
    const int C = c->cfg.ncomp;
    const bool spin1 = C == 3 && (phase == GPE_PHASE_POLAR || phase == GPE_PHASE_FERRO || phase == GPE_PHASE_MIXED);
    const bool spin2 = C == 5 && (phase == GPE_PHASE_FERRO || (phase >= GPE_PHASE_UN && phase <= GPE_PHASE_CYCLIC));
    if (!spin1 && !spin2) return fail(c, GPE_EINVAL, "phase does not exist for this component count");
    if (phase == GPE_PHASE_MIXED && !(param >= 0.0 && param <= 1.0)) return fail(c, GPE_EINVAL, "mix outside [0, 1]");
    hipError_t e = launch_project(c->fp64, C, c->g, c->level[level], phase, param, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? GPE_OK : hip_fail(c, e, "spinor projection");
}

int gpe_add_noise(gpe_ctx *c, int level, double amplitude, unsigned long long seed)
{
    if (!c) return GPE_EINVAL;
    int rc = check_level(c, level);
    if (rc) return rc;
    if (!(amplitude >= 0.0)) return fail(c, GPE_EINVAL, "negative noise amplitude");
    hipError_t e = launch_noise(c->fp64, c->cfg.ncomp, c->g, c->level[level], amplitude, seed, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? GPE_OK : hip_fail(c, e, "noise");
}

int gpe_set_knot_ramp(gpe_ctx *c, int n, const double *knots, double duration, double bq)
{
    if (!c) return GPE_EINVAL;
    if (n < 2 || !knots || !(duration > 0.0)) return fail(c, GPE_EINVAL, "need >= 2 knots and a positive duration");
    std::vector<double> bval(3 * (size_t)n, 0.0), bdur(n), qval{bq}, qdur{1e300};
    std::vector<int> btype(n), qtype{0};
    for (int j = 0; j < n; ++j) {
        bval[3 * j + 2] = knots[j];
        bdur[j] = j == 0 ? 0.0 : duration / (n - 1);
        btype[j] = j == 0 ? 0 : 1;
    }
    gpe_ramp r{};
    r.nq = 1;
    r.q_val = qval.data();
    r.q_dur = qdur.data();
    r.q_type = qtype.data();
    r.nb = n;
    r.b_val = bval.data();
    r.b_dur = bdur.data();
    r.b_type = btype.data();
    r.tau = 1.0;
    r.time_scale = 1.0;
    r.scale_q = r.scale_b = 1.0;
    r.qscale_q = r.qscale_b = 0.0;
    return gpe_set_ramp(c, &r);
}

// Ramp sensitivities: the ramp value is linear in the segment values, so
// d value(t) / d val[j] = value(t) evaluated with val = e_j.
static double ramp_basis(const std::vector<double> &dur, const std::vector<int> &type, double tau, double t, int j)
{
    std::vector<double> e(dur.size(), 0.0);
    e[j] = 1.0;
    double out;
    ramp_channel(e, dur, type, 1, tau, t, &out);
    return out;
}

// The adjoint of an FP32 context with FP64 psi and mu: the run (configuration, fields, ramp,
// clock and level 0) is copied into a temporary FP64 context, the adjoint taken there, and the
// final two levels copied back.
static int adjoint_fp64_copy(gpe_ctx *c, const gpe_adjoint_spec *spec, gpe_adjoint_result *res)
{
    gpe_config cfg = c->cfg;
    cfg.precision = GPE_FP64;
    gpe_ctx *d = nullptr;
    int rc = gpe_create(&cfg, &d);
    if (rc) return fail(c, rc, "the FP64 copy of the run does not fit");
    d->ramp = c->ramp;
    d->bq = c->bq;
    d->bqq = c->bqq;
    for (int k = 0; k < 3; ++k) {
        d->bb[k] = c->bb[k];
        d->bbq[k] = c->bbq[k];
    }
    const size_t abi = (size_t)cfg.ncomp * c->g.nz * c->g.ny * c->g.nx * c->g.V * 2 * sizeof(double);
    double *tmp = nullptr;
    if (hipMalloc((void **)&tmp, abi) != hipSuccess) {
        gpe_destroy(d);
        return fail(c, GPE_ENOMEM, "staging for the FP64 copy does not fit");
    }
    const double t0 = c->t0;
    rc = gpe_get_state_device(c, 0, tmp);
    if (!rc) rc = gpe_set_state_device(d, 0, tmp, t0);
    gpe_adjoint_spec s64 = *spec;
    s64.fp64_trajectory = 0;
    if (!rc) {
        rc = gpe_adjoint(d, &s64, res);
        if (rc) c->err = "FP64 copy: " + d->err;
    }
    if (!rc) rc = gpe_get_state_device(d, 0, tmp);
    if (!rc) rc = gpe_set_state_device(c, 0, tmp, t0);
    if (!rc) rc = gpe_get_state_device(d, 1, tmp);
    if (!rc) rc = gpe_set_state_device(c, 1, tmp, t0);
    if (!rc) {
        c->sweeps = d->sweeps;
        c->level1_valid = true;
    } else {
        c->level1_valid = false;
    }
    (void)hipFree(tmp);
    gpe_destroy(d);
    return rc;
}

int gpe_adjoint(gpe_ctx *c, const gpe_adjoint_spec *spec, gpe_adjoint_result *res)
{
    if (!c) return GPE_EINVAL;
    if (!spec || !res) return fail(c, GPE_EINVAL, "null adjoint spec or result");
    if (c->cfg.scheme != GPE_SCHEME_LEAPFROG3) return fail(c, GPE_EINVAL, "adjoint needs the 3-level scheme");
    if (c->cfg.alpha != 0.0) return fail(c, GPE_EINVAL, "adjoint needs alpha = 0");
    if (c->cfg.expansion_rate != 0.0) return fail(c, GPE_EINVAL, "adjoint needs a static grid (expansion_rate = 0)");
    if (c->ramp.active && (c->ramp.qscale_q != 0.0 || c->ramp.qscale_b != 0.0))
        return fail(c, GPE_EINVAL, "adjoint does not differentiate a ramped quadratic field");
    const int N = spec->nsweeps, K = spec->checkpoint_every;
    if (N < 1 || K < 2) return fail(c, GPE_EINVAL, "need nsweeps >= 1 and checkpoint_every >= 2");
    if (spec->objective < 0 || spec->objective > 2) return fail(c, GPE_EINVAL, "unknown objective");
    if (spec->objective == GPE_OBJ_OVERLAP && !spec->target) return fail(c, GPE_EINVAL, "overlap needs a target");
    if (c->sweeps != 0) return fail(c, GPE_ESTATE, "adjoint must start right after gpe_set_state");
    if (spec->fp64_trajectory && !c->fp64) return adjoint_fp64_copy(c, spec, res);

    const size_t lbytes = (size_t)c->level_elems * c->elem_bytes;
    const int nck = (N + K - 1) / K;
    std::vector<void *> bufs;
    auto alloc = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) return GPE_ENOMEM;
        bufs.push_back(*p);
        return GPE_OK;
    };
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (void *p : bufs) (void)hipFree(p);
        bufs.clear();
    };
    std::vector<void *> ck(2 * nck, nullptr), seg(K, nullptr);
    void *mu[2] = {nullptr, nullptr};
    void *target = nullptr;
    double *d_steps = nullptr, *d_part = nullptr;
    int rc = GPE_OK;
    for (int i = 0; i < 2 * nck && !rc; ++i) rc = alloc(&ck[i], lbytes);
    for (int i = 2; i < K && !rc; ++i) rc = alloc(&seg[i], lbytes);  // seg[0], seg[1] alias checkpoints
    for (int i = 0; i < 2 && !rc; ++i) rc = alloc(&mu[i], lbytes);
    if (!rc && spec->objective == GPE_OBJ_OVERLAP) rc = alloc(&target, lbytes);
    if (!rc) rc = alloc((void **)&d_steps, sizeof(double) * ADJ_COUNT * (size_t)N);
    if (!rc) rc = alloc((void **)&d_part, sizeof(double) * ADJ_COUNT * (size_t)c->g.blocks);
    // per-sweep sensitivity partials of up to Q consecutive adjoint sweeps, summed in one launch
    // (one two-stage finalize per sweep cost ~15% of the backward pass in launches)
    // the adjoint sweeps run on the forward sweep's traversal (x pairs in FP32)
    const int adj_blocks = c->gs.blocks;
    const size_t sweep_part = (size_t)ADJ_COUNT * adj_blocks;
    const int Q = (int)std::max<size_t>(1, std::min<size_t>((size_t)K, ((size_t)256 << 20) / (sweep_part * 8)));
    double *d_qpart = nullptr;
    if (!rc) rc = alloc((void **)&d_qpart, sizeof(double) * sweep_part * Q);
    if (rc) {
        release();
        return fail(c, rc, "adjoint buffers do not fit");
    }
    res->checkpoint_bytes = (double)lbytes * 2 * nck;
    res->segment_bytes = (double)lbytes * (K - 2 + 2);
    hipError_t e = hipSuccess;
    // on a failure the buffers are freed and the levels hold an intermediate state: the
    // context needs gpe_set_state (and gpe_start) before it steps again
#define ADJ_CHECK(call, what)                 \
    do {                                      \
        e = (call);                           \
        if (e != hipSuccess) {                \
            release();                        \
            c->level1_valid = false;          \
            return hip_fail(c, e, what);      \
        }                                     \
    } while (0)

    const double sdt = sweep_dt(c);
    rc = upload_fields(c, 0, N);
    if (rc) {
        release();
        return rc;
    }
    // ---- forward pass, keeping (psi_jK, psi_jK+1)
    ADJ_CHECK(hipEventRecord(c->ev0, c->stream), "adjoint");
    ADJ_CHECK(hipMemcpyAsync(ck[0], c->level[0], lbytes, hipMemcpyDeviceToDevice, c->stream), "checkpoint");
    {
        SweepArgs a = sweep_args(c, 0, c->level[0], c->level[1], sdt);
        ADJ_CHECK(launch_sweep(c->fp64, c->cfg.ncomp, MODE_EULER, a, c->stream), "forward Euler");
        std::swap(c->level[0], c->level[1]);
    }
    ADJ_CHECK(hipMemcpyAsync(ck[1], c->level[0], lbytes, hipMemcpyDeviceToDevice, c->stream), "checkpoint");
    for (int k = 1; k < N; ++k) {  // produces psi_{k+1}
        SweepArgs a = sweep_args(c, k, c->level[0], c->level[1], 2.0 * sdt);
        a.order = k & 1;
        ADJ_CHECK(launch_sweep(c->fp64, c->cfg.ncomp, MODE_UPDATE, a, c->stream), "forward sweep");
        std::swap(c->level[0], c->level[1]);
        const int m = k + 1;
        if (m < N && m % K == 0)
            ADJ_CHECK(hipMemcpyAsync(ck[2 * (m / K)], c->level[0], lbytes, hipMemcpyDeviceToDevice, c->stream), "ck");
        if (m <= N && m % K == 1 && m / K < nck)
            ADJ_CHECK(hipMemcpyAsync(ck[2 * (m / K) + 1], c->level[0], lbytes, hipMemcpyDeviceToDevice, c->stream),
                      "ck");
    }
    c->sweeps = N;
    c->level1_valid = true;
    ADJ_CHECK(hipEventRecord(c->ev1, c->stream), "adjoint");
    ADJ_CHECK(hipEventSynchronize(c->ev1), "adjoint");
    float fms = 0.f;
    ADJ_CHECK(hipEventElapsedTime(&fms, c->ev0, c->ev1), "adjoint");
    res->forward_ms = fms;

    // ---- objective and mu_N
    ADJ_CHECK(hipEventRecord(c->ev0, c->stream), "adjoint");
    ObjArgs oa;
    oa.g = c->g;
    oa.tb = tables(c);
    oa.ph = c->ph;
    oa.field = field_at(c, c->t0 + N * sdt);
    oa.psi = c->level[0];
    oa.target = target;
    oa.grad = mu[1];
    oa.scalars = c->d_scalars + 32;
    oa.partials = d_part;
    ADJ_CHECK(hipMemsetAsync(mu[0], 0, lbytes, c->stream), "clear mu");
    double J = 0.0;
    double explicit_mom[4] = {0, 0, 0, 0};  // energy: sum dv r_a^2 n, s sum dv (dx Fx + dy Fy - 2 dz Fz)
    double explicit_mag[3] = {0, 0, 0};
    if (spec->objective == GPE_OBJ_OVERLAP) {
        const Geometry &g = c->g;
        const size_t abi = (size_t)c->cfg.ncomp * g.nz * g.ny * g.nx * g.V * 2 * sizeof(double);
        double *staging = nullptr;
        ADJ_CHECK(hipMalloc((void **)&staging, abi), "staging");
        bufs.push_back(staging);
        ADJ_CHECK(hipMemcpyAsync(staging, spec->target, abi, hipMemcpyHostToDevice, c->stream), "target");
        ADJ_CHECK(launch_pack(c->fp64, g, c->cfg.ncomp, staging, target, c->stream), "pack target");
        oa.mode = OBJ_PARTIALS;
        ADJ_CHECK(launch_objective(c->fp64, c->cfg.ncomp, oa, c->stream), "overlap");
        ADJ_CHECK(launch_finalize(d_part, g.blocks, 2, -1, c->d_scalars + 32, c->d_red, c->stream), "overlap sum");
        oa.mode = OBJ_OVERLAP;
        ADJ_CHECK(launch_objective(c->fp64, c->cfg.ncomp, oa, c->stream), "overlap grad");
        double ov[2];
        ADJ_CHECK(hipMemcpyAsync(ov, c->d_scalars + 32, sizeof(ov), hipMemcpyDeviceToHost, c->stream), "adjoint");
        ADJ_CHECK(hipStreamSynchronize(c->stream), "adjoint");
        J = ov[0] * ov[0] + ov[1] * ov[1];
    } else {
        gpe_observables_t o;
        int orc = gpe_observables(c, 0, &o);
        if (orc) {
            release();
            return orc;
        }
        if (spec->objective == GPE_OBJ_ENERGY) {
            J = o.energy;
            for (int k = 0; k < 3; ++k) explicit_mag[k] = o.magnetization[k];
            oa.mode = OBJ_EXPLICIT;
            ADJ_CHECK(launch_objective(c->fp64, c->cfg.ncomp, oa, c->stream), "explicit terms");
            ADJ_CHECK(launch_finalize(d_part, c->g.blocks, 4, -1, c->d_scalars + 40, c->d_red, c->stream), "explicit sums");
            ADJ_CHECK(hipMemcpyAsync(explicit_mom, c->d_scalars + 40, sizeof(explicit_mom), hipMemcpyDeviceToHost,
                                      c->stream), "adjoint");
            ADJ_CHECK(hipStreamSynchronize(c->stream), "adjoint");
            // H psi_N at the time of psi_N into mu[1], then scaled by 2 dv
            int frc = upload_fields(c, N, 1);
            if (frc) {
                release();
                return frc;
            }
            SweepArgs a = sweep_args(c, 0, c->level[0], mu[1], 0.0);
            ADJ_CHECK(launch_sweep(c->fp64, c->cfg.ncomp, MODE_HPSI, a, c->stream), "H psi");
            oa.mode = OBJ_ENERGY;
        } else {
            J = o.magnetization[2];
            oa.mode = OBJ_MAGZ;
        }
        ADJ_CHECK(launch_objective(c->fp64, c->cfg.ncomp, oa, c->stream), "objective grad");
        int frc = upload_fields(c, 0, N);
        if (frc) {
            release();
            return frc;
        }
    }
    res->objective = J;

    // ---- backward pass over the segments, recomputing psi inside each
    // A = mu_{k+2}, B = mu_{k+1}
    void *A = mu[0], *B = mu[1];
    for (int j = nck - 1; j >= 0; --j) {
        const int klo = j * K, khi = std::min((j + 1) * K - 1, N - 1);
        seg[0] = ck[2 * j];
        seg[1] = ck[2 * j + 1];
        for (int m = klo + 1; m < khi; ++m) {  // psi_{m+1} = psi_{m-1} - 2i dt H psi_m
            SweepArgs a = sweep_args(c, m, seg[m - klo], seg[m + 1 - klo], 2.0 * sdt);
            a.prev = seg[m - 1 - klo];
            a.order = m & 1;
            ADJ_CHECK(launch_sweep(c->fp64, c->cfg.ncomp, MODE_UPDATE, a, c->stream), "recompute");
        }
        int qhi = khi, nq = 0;  // pending sweeps qhi, qhi - 1, ... in slots 0, 1, ...
        for (int k = khi; k >= klo; --k) {
            AdjArgs aa;
            aa.g = c->gs;
            aa.tb = tables(c);
            aa.ph = c->ph;
            aa.psi = seg[k - klo];
            aa.mu1 = B;
            aa.mu2 = A;
            aa.fields = c->d_fields;
            aa.sweep = k;
            aa.order = k & 1;
            aa.quad = quad_active(c);
            aa.coef = 2.0 * sdt;
            aa.partials = d_qpart + sweep_part * nq;
            ADJ_CHECK(launch_adjoint(c->fp64, c->cfg.ncomp, aa, c->stream), "adjoint sweep");
            if (++nq == Q || k == klo) {
                ADJ_CHECK(launch_sum_tables(d_qpart, nq, adj_blocks, ADJ_COUNT, d_steps + (size_t)ADJ_COUNT * qhi,
                                            -(long long)ADJ_COUNT, c->stream),
                          "adjoint sums");
                qhi = k - 1;
                nq = 0;
            }
            std::swap(A, B);
        }
    }
    std::vector<double> steps((size_t)ADJ_COUNT * N);
    ADJ_CHECK(hipMemcpyAsync(steps.data(), d_steps, sizeof(double) * steps.size(), hipMemcpyDeviceToHost, c->stream), "adjoint");
    ADJ_CHECK(hipEventRecord(c->ev1, c->stream), "adjoint");
    ADJ_CHECK(hipEventSynchronize(c->ev1), "adjoint");
    float bms = 0.f;
    ADJ_CHECK(hipEventElapsedTime(&bms, c->ev0, c->ev1), "adjoint");
    res->backward_ms = bms;
    release();

    // ---- chain rule to the parameters (FP64 on the host)
    for (int a = 0; a < 3; ++a) {
        res->grad_trap[a] = 0.0;
        res->grad_static_bb[a] = 0.0;
    }
    res->grad_static_bq = 0.0;
    const Ramp &r = c->ramp;
    if (res->grad_ramp_q)
        for (size_t j = 0; j < r.q_dur.size(); ++j) res->grad_ramp_q[j] = 0.0;
    if (res->grad_ramp_b)
        for (size_t j = 0; j < 3 * r.b_dur.size(); ++j) res->grad_ramp_b[j] = 0.0;
    for (int k = 0; k < N; ++k) {
        const double *p = &steps[(size_t)ADJ_COUNT * k];
        const double w = (k == 0) ? sdt : 2.0 * sdt;
        for (int a = 0; a < 3; ++a) res->grad_trap[a] += w * p[ADJ_WX + a];
        if (!r.active) {
            for (int a = 0; a < 3; ++a) res->grad_static_bb[a] += w * p[ADJ_BX + a];
            res->grad_static_bq += w * p[ADJ_BQ];
        } else {
            const double tr = (c->t0 + k * sdt) * r.time_scale;
            if (res->grad_ramp_q)
                for (size_t j = 0; j < r.q_dur.size(); ++j)
                    res->grad_ramp_q[j] += w * r.scale_q * ramp_basis(r.q_dur, r.q_type, r.tau, tr, (int)j) * p[ADJ_BQ];
            if (res->grad_ramp_b)
                for (size_t j = 0; j < r.b_dur.size(); ++j) {
                    const double phi = ramp_basis(r.b_dur, r.b_type, r.tau, tr, (int)j);
                    if (phi == 0.0) continue;
                    for (int a = 0; a < 3; ++a) res->grad_ramp_b[3 * j + a] += w * r.scale_b * phi * p[ADJ_BX + a];
                }
        }
    }
    if (spec->objective == GPE_OBJ_ENERGY) {
        // explicit dependence of E[psi_N] on the parameters at t_N
        const double tN = c->t0 + N * sdt;
        if (tN < c->ph.trap_off_time)
            for (int a = 0; a < 3; ++a) res->grad_trap[a] += c->ph.trap[a] * explicit_mom[a];
        if (c->cfg.ncomp > 1) {
            const double s = c->cfg.ncomp == 3 ? 1.0 : -1.0;
            if (!r.active) {
                for (int a = 0; a < 3; ++a) res->grad_static_bb[a] += s * explicit_mag[a];
                res->grad_static_bq += explicit_mom[3];
            } else {
                const double tr = tN * r.time_scale;
                if (res->grad_ramp_q)
                    for (size_t j = 0; j < r.q_dur.size(); ++j)
                        res->grad_ramp_q[j] += r.scale_q * ramp_basis(r.q_dur, r.q_type, r.tau, tr, (int)j) * explicit_mom[3];
                if (res->grad_ramp_b)
                    for (size_t j = 0; j < r.b_dur.size(); ++j) {
                        const double phi = ramp_basis(r.b_dur, r.b_type, r.tau, tr, (int)j);
                        for (int a = 0; a < 3; ++a) res->grad_ramp_b[3 * j + a] += r.scale_b * phi * s * explicit_mag[a];
                    }
            }
        }
    }
#undef ADJ_CHECK
    return GPE_OK;
}

// ------------------------------------------------------------------------- ensembles ---

// Members sharing (face degree, x per lane) form a group launched by one specialised kernel.
struct EnsembleGroup {
    int deg, form;      // face degree and lane_form of every member
    int first, count;   // members [first, first + count) of the sorted member list
    int blocks;         // workgroups of the group's launch
};

struct gpe_ensemble {
    std::vector<gpe_ctx *> m;          // sorted by group
    std::vector<Geometry> trav;        // each member's sweep traversal when the ensemble was made
    std::vector<EnsembleMember> h;
    std::vector<EnsembleGroup> groups;
    EnsembleMember *d = nullptr;
    int *d_base = nullptr;
    int ncomp = 0, fp64 = 0, per = 1;
    hipStream_t stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_sweeps = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, unknowns = 0.0;
    std::string err;
};

static int ens_fail(gpe_ensemble *e, int code, const std::string &what)
{
    if (e) e->err = what;
    return code;
}

int gpe_ensemble_create(gpe_ctx *const *members, int n, gpe_ensemble **out)
{
    if (!members || n < 1 || !out) return GPE_EINVAL;
    *out = nullptr;
    gpe_ctx *c0 = members[0];
    for (int i = 0; i < n; ++i) {
        gpe_ctx *c = members[i];
        if (!c) return GPE_EINVAL;
        if (c->cfg.ncomp != c0->cfg.ncomp || c->fp64 != c0->fp64 || c->cfg.scheme != c0->cfg.scheme ||
            c->device != c0->device)
            return GPE_EINVAL;
        if (!c->level1_valid) return GPE_ESTATE;
    }
    gpe_ensemble *e = new (std::nothrow) gpe_ensemble();
    if (!e) return GPE_ENOMEM;
    e->m.assign(members, members + n);
    std::stable_sort(e->m.begin(), e->m.end(), [](const gpe_ctx *a, const gpe_ctx *b) {
        return a->gs.deg != b->gs.deg ? a->gs.deg < b->gs.deg
                                      : lane_form(a->gs.xpl, a->gs.rows) < lane_form(b->gs.xpl, b->gs.rows);
    });
    e->ncomp = c0->cfg.ncomp;
    e->fp64 = c0->fp64;
    e->per = c0->cfg.scheme == GPE_SCHEME_STAGGERED ? 2 : 1;
    e->h.resize(n);
    int begin = 0;
    bool prev_solo = false;
    for (int i = 0; i < n; ++i) {
        const gpe_ctx *c = e->m[i];
        const int form = lane_form(c->gs.xpl, c->gs.rows);
        const bool solo = c->gs.blocks >= ENS_SOLO_BLOCKS;
        if (i == 0 || solo || prev_solo || c->gs.deg != e->groups.back().deg || form != e->groups.back().form ||
            e->groups.back().count == ENS_GROUP_MAX) {
            e->groups.push_back(EnsembleGroup{c->gs.deg, form, i, 0, 0});
            begin = 0;
        }
        prev_solo = solo;
        e->groups.back().count += 1;
        e->h[i].begin = begin;
        e->h[i].blocks = c->gs.blocks;
        begin += (c->gs.blocks + 7) / 8 * 8;
        e->groups.back().blocks = begin;
        e->trav.push_back(c->gs);
        e->unknowns += (double)c->cfg.nx * c->cfg.ny * c->cfg.nz * c->g.V;
    }
    int rc = GPE_OK;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = GPE_EDEVICE;
    if (!rc && hipMalloc((void **)&e->d, sizeof(EnsembleMember) * n) != hipSuccess) rc = GPE_ENOMEM;
    if (!rc && hipMalloc((void **)&e->d_base, sizeof(int) * 4) != hipSuccess) rc = GPE_ENOMEM;
    if (!rc && (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)) rc = GPE_EDEVICE;
    if (rc) {
        gpe_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return GPE_OK;
}

int gpe_ensemble_destroy(gpe_ensemble *e)
{
    if (!e) return GPE_EINVAL;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->d) (void)hipFree(e->d);
    if (e->d_base) (void)hipFree(e->d_base);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return GPE_OK;
}

const char *gpe_ensemble_last_error(const gpe_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

#define ENSCALL(e, call)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return ens_fail(e, GPE_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int gpe_ensemble_step(gpe_ensemble *e, int nsteps, int graph_sweeps)
{
    if (!e) return GPE_EINVAL;
    if (nsteps < 0) return ens_fail(e, GPE_EINVAL, "negative step count");
    const int nsw = nsteps * e->per;
    if (nsw == 0) return GPE_OK;
    if (graph_sweeps < 0 || (graph_sweeps > 0 && (graph_sweeps % 2 != 0 || nsw % graph_sweeps != 0)))
        return ens_fail(e, GPE_EINVAL, "graph_sweeps must be even and divide the sweep count");
    for (size_t i = 0; i < e->m.size(); ++i) {
        gpe_ctx *c = e->m[i];
        if (!c->level1_valid) return ens_fail(e, GPE_ESTATE, "member not started");
        const Geometry &t = e->trav[i];
        if (c->gs.blocks != t.blocks || c->gs.xpl != t.xpl || c->gs.rows != t.rows || c->gs.ty != t.ty ||
            c->gs.xcd_remap != t.xcd_remap || c->gs.v_inner != t.v_inner)
            return ens_fail(e, GPE_ESTATE, "member traversal changed after the ensemble was created");
        int rc = upload_fields(c, c->sweeps, nsw);
        if (rc) return ens_fail(e, rc, c->err);
        ENSCALL(e, hipStreamSynchronize(c->stream));
        const double coef = 2.0 * sweep_dt(c);
        e->h[i].args[0] = sweep_args(c, 0, c->level[0], c->level[1], coef);
        e->h[i].args[1] = sweep_args(c, 0, c->level[1], c->level[0], coef);
    }
    ENSCALL(e, hipMemcpyAsync(e->d, e->h.data(), sizeof(EnsembleMember) * e->h.size(), hipMemcpyHostToDevice,
                              e->stream));
    // one launch per group per sweep; groups are independent of each other
    auto launch_group = [&](const EnsembleGroup &gr, int k, const int *base, hipStream_t st) -> hipError_t {
        EnsembleArgs a;
        a.members = e->d + gr.first;
        a.nmember = gr.count;
        a.parity = k & 1;
        a.sweep = k;
        a.sweep_base = base;
        for (int i = 0; i < gr.count; ++i) a.begin[i] = e->h[gr.first + i].begin;
        return launch_ensemble_sweep(e->fp64, e->ncomp, gr.deg, gr.form, a, gr.blocks, st);
    };
    auto launch_all = [&](int k, const int *base) -> hipError_t {
        for (const EnsembleGroup &gr : e->groups) {
            hipError_t err = launch_group(gr, k, base, e->stream);
            if (err != hipSuccess) return err;
        }
        return hipSuccess;
    };
    // K sweeps of every group, one group after another: chains of independent groups forked
    // onto concurrent streams ran 3% slower (their kernels share each XCD's L2)
    auto capture_chains = [&](int K) -> hipError_t {
        hipError_t err = hipSuccess;
        for (size_t g = 0; g < e->groups.size() && err == hipSuccess; ++g)
            for (int k = 0; k < K && err == hipSuccess; ++k) err = launch_group(e->groups[g], k, e->d_base, e->stream);
        return err;
    };
    ENSCALL(e, hipEventRecord(e->ev0, e->stream));
    if (graph_sweeps > 0) {
        if (!e->gexec || e->graph_sweeps != graph_sweeps) {
            if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
            if (e->graph) (void)hipGraphDestroy(e->graph);
            e->gexec = nullptr;
            e->graph = nullptr;
            ENSCALL(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            hipError_t cap = capture_chains(graph_sweeps);
            if (cap == hipSuccess) cap = launch_advance_base(e->d_base, graph_sweeps, e->stream);
            hipGraph_t captured = nullptr;
            const hipError_t end = hipStreamEndCapture(e->stream, &captured);
            if (cap != hipSuccess || end != hipSuccess) {
                if (captured) (void)hipGraphDestroy(captured);
                return ens_fail(e, GPE_EDEVICE, std::string("graph capture: ") +
                                                    hipGetErrorString(cap != hipSuccess ? cap : end));
            }
            e->graph = captured;
            ENSCALL(e, hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
            e->graph_sweeps = graph_sweeps;
        }
        ENSCALL(e, launch_set_base(e->d_base, 0, e->stream));
        for (int r = 0; r < nsw / graph_sweeps; ++r) ENSCALL(e, hipGraphLaunch(e->gexec, e->stream));
    } else {
        for (int k = 0; k < nsw; ++k) ENSCALL(e, launch_all(k, nullptr));
    }
    ENSCALL(e, hipEventRecord(e->ev1, e->stream));
    ENSCALL(e, hipEventSynchronize(e->ev1));
    float ms = 0.f;
    ENSCALL(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms = ms / nsw;
    for (gpe_ctx *c : e->m) {
        if (nsw & 1) std::swap(c->level[0], c->level[1]);
        c->sweeps += nsw;
    }
    return GPE_OK;
}

int gpe_ensemble_timing(const gpe_ensemble *e, double *ms_per_sweep, double *unknowns)
{
    if (!e) return GPE_EINVAL;
    if (ms_per_sweep) *ms_per_sweep = e->last_ms;
    if (unknowns) *unknowns = e->unknowns;
    return GPE_OK;
}

// ------------------------------------------------------------------------ autotuning ---

// Apply a traversal to both geometries; on failure (a workgroup count past the 2^29 limit of
// the decode) both are left as they were and false is returned.  flags: bit 0 the XCD remap,
// bit 1 the unknown v inside the x segment in the traversal order.
static bool set_traversal(gpe_ctx *c, int ty, int flags, int xpl)
{
    Geometry g = c->g, gs = c->gs;
    gs.xpl = xpl;
    for (Geometry *p : {&g, &gs}) {
        p->ty = ty;
        p->xcd_remap = flags & 1;
        p->v_inner = (flags >> 1) & 1;
        if (!finish_geometry(*p)) return false;
    }
    c->g = g;
    c->gs = gs;
    return true;
}

// Reduction partials are sized by the workgroups of the one-x-per-lane traversal.
static int grow_partials(gpe_ctx *c, int blocks_before)
{
    if (c->g.blocks <= blocks_before) return GPE_OK;
    HIPCALL(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->d_partials);
    c->d_partials = nullptr;
    if (hipMalloc((void **)&c->d_partials, sizeof(double) * OBS_COUNT * (size_t)c->g.blocks) != hipSuccess)
        return fail(c, GPE_ENOMEM, "partials");
    return GPE_OK;
}

int gpe_set_traversal(gpe_ctx *c, int tile_rows, int xcd_remap, int x_per_lane)
{
    if (!c) return GPE_EINVAL;
    if (tile_rows < 1 || tile_rows > 1024 || xcd_remap < 0 || xcd_remap > 3 || (x_per_lane != 1 && x_per_lane != 2))
        return fail(c, GPE_EINVAL, "bad traversal");
    if (x_per_lane == 2 && !pair_lanes_ok(c->fp64, c->g.deg, c->cfg.ncomp))
        return fail(c, GPE_EINVAL, "two x per lane needs FP32");
    const int before = c->g.blocks;
    if (!set_traversal(c, tile_rows, xcd_remap, x_per_lane))
        return fail(c, GPE_EINVAL, "traversal needs more workgroups than the decode supports");
    return grow_partials(c, before);
}

int gpe_traversal(const gpe_ctx *c, int *tile_rows, int *xcd_remap, int *x_per_lane)
{
    if (!c) return GPE_EINVAL;
    if (tile_rows) *tile_rows = c->gs.ty;
    if (xcd_remap) *xcd_remap = c->gs.xcd_remap | (c->gs.v_inner << 1);
    if (x_per_lane) *x_per_lane = c->gs.xpl;
    return GPE_OK;
}

int gpe_plan_traversal(const gpe_config *cfg, int tile_rows, int xcd_remap, int x_per_lane, gpe_traversal_plan *out)
{
    if (!cfg || !out) return GPE_EINVAL;
    if (cfg->lattice < 0 || cfg->lattice > 5 || (cfg->ncomp != 1 && cfg->ncomp != 3 && cfg->ncomp != 5)) return GPE_EINVAL;
    if (cfg->precision != GPE_FP32 && cfg->precision != GPE_FP64) return GPE_EINVAL;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return GPE_EINVAL;
    if (tile_rows < 1 || tile_rows > 1024 || xcd_remap < 0 || xcd_remap > 3 || (x_per_lane != 1 && x_per_lane != 2))
        return GPE_EINVAL;
    Geometry g = base_geometry(*cfg);
    if (x_per_lane == 2 && !pair_lanes_ok(cfg->precision == GPE_FP64, g.deg, cfg->ncomp)) return GPE_EINVAL;
    g.ty = tile_rows;
    g.xcd_remap = xcd_remap & 1;
    g.v_inner = (xcd_remap >> 1) & 1;
    g.xpl = x_per_lane;
    if (!finish_geometry(g)) return GPE_EINVAL;
    out->nx = g.nx;
    out->ny = g.ny;
    out->nz = g.nz;
    out->V = g.V;
    out->C = g.C;
    out->x_per_lane = g.xpl;
    out->rows = g.rows;
    out->pairs = g.pairs;
    out->nseg = g.nseg;
    out->tile_rows = g.ty;
    out->row_groups = g.tyr;
    out->ntile = g.ntile;
    out->xcd_remap = g.xcd_remap | (g.v_inner << 1);
    out->lx = g.lx;
    out->slabs = g.slabs;
    out->ys = g.ys;
    out->vs = g.vs;
    out->cs = g.cs;
    out->zs = g.zs;
    out->waves = g.waves;
    out->blocks = g.blocks;
    const FastDiv *d[5] = {&g.f_nseg, &g.f_V, &g.f_ty, &g.f_nz, &g.f_pairs};
    unsigned *o[5] = {out->div_nseg, out->div_V, out->div_rows, out->div_nz, out->div_pairs};
    for (int k = 0; k < 5; ++k) {
        o[k][0] = d[k]->d;
        o[k][1] = d[k]->m;
        o[k][2] = d[k]->l;
    }
    return GPE_OK;
}

typedef std::tuple<int, int, int, int, int, int> TuneKey;
static std::map<TuneKey, std::tuple<int, int, int>> g_tune_cache;  // rows, remap, x per lane
static std::mutex g_tune_mutex;

int gpe_autotune(gpe_ctx *c, int trial_sweeps, int *tile_rows, int *xcd_remap, int *x_per_lane, double *ms_out)
{
    if (!c) return GPE_EINVAL;
    if (trial_sweeps < 0) return fail(c, GPE_EINVAL, "negative trial sweeps");
    const TuneKey key(c->cfg.lattice, c->cfg.ncomp, c->fp64, c->cfg.nx, c->cfg.ny, c->cfg.nz);
    const int max_blocks_before = c->g.blocks;
    int best_ty = c->g.ty, best_remap = c->gs.xcd_remap | (c->gs.v_inner << 1), best_xpl = c->gs.xpl;
    double best_ms = -1.0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            std::tie(best_ty, best_remap, best_xpl) = it->second;
            trial_sweeps = 0;
        }
    }
    if (trial_sweeps > 0) {
        const size_t lbytes = (size_t)c->level_elems * c->elem_bytes;
        // scratch copies of the two levels, freed when the trials end (also on failure), when
        // the context's traversal on entry is restored too (the winner is applied below)
        struct Trial {
            gpe_ctx *c;
            Geometry g, gs;
            void *s0 = nullptr, *s1 = nullptr;
            ~Trial()
            {
                if (s0 || s1) (void)hipStreamSynchronize(c->stream);
                if (s0) (void)hipFree(s0);
                if (s1) (void)hipFree(s1);
                c->g = g;
                c->gs = gs;
            }
        } trial{c, c->g, c->gs};
        if (hipMalloc(&trial.s0, lbytes) != hipSuccess || hipMalloc(&trial.s1, lbytes) != hipSuccess)
            return fail(c, GPE_ENOMEM, "autotune scratch does not fit");
        HIPCALL(c, hipMemcpyAsync(trial.s0, c->level[0], lbytes, hipMemcpyDeviceToDevice, c->stream));
        HIPCALL(c, hipMemcpyAsync(trial.s1, c->level[1], lbytes, hipMemcpyDeviceToDevice, c->stream));
        int rc = upload_fields(c, c->sweeps, 1);
        if (rc) return rc;
        // tile heights: powers of two, and multiples of 3, 5, 6 and 7 rows, which are
        // candidates only where their row groups are not powers of two; and the whole plane
        // (one tile: each XCD a range of planes that share their y neighbours in its L2;
        // FP32 A15 82^3 one component +17%, spin-2 +8%, Cube 293^3 one component +11%)
        std::vector<int> tys = {1, 2, 4, 8, 16, 32, 3, 6, 12, 24, 5, 10, 20, 7, 14, 28};
        if (c->g.ny > 32 && c->g.ny <= 1024) tys.push_back(c->g.ny);
        const int nxpl = pair_lanes_ok(c->fp64, c->g.deg, c->cfg.ncomp) ? 2 : 1;
        for (int xpl = 1; xpl <= nxpl; ++xpl)
        for (int ty : tys) {
            if (ty > c->g.ny && ty != 1) continue;
            const bool pow2 = (ty & (ty - 1)) == 0;
            // flags: XCD remap, and the unknown inside the x segment (only where lines span several waves)
            for (int remap = 0; remap < 4; ++remap) {
                if (!set_traversal(c, ty, remap, xpl)) continue;  // too many workgroups: not a candidate
                if (!pow2 && (c->gs.rows & (c->gs.rows - 1)) == 0) continue;
                if ((remap & 2) && c->gs.nseg == 1) continue;
                void *cur = trial.s0, *nxt = trial.s1;
                for (int k = 0; k < trial_sweeps + 2; ++k) {
                    if (k == 2) HIPCALL(c, hipEventRecord(c->ev0, c->stream));
                    SweepArgs a = sweep_args(c, 0, cur, nxt, 2.0 * sweep_dt(c));
                    a.order = k & 1;
                    hipError_t e = launch_sweep(c->fp64, c->cfg.ncomp, MODE_UPDATE, a, c->stream);
                    if (e != hipSuccess) return hip_fail(c, e, "autotune sweep");
                    std::swap(cur, nxt);
                }
                HIPCALL(c, hipEventRecord(c->ev1, c->stream));
                HIPCALL(c, hipEventSynchronize(c->ev1));
                float ms = 0.f;
                HIPCALL(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
                const double per = ms / trial_sweeps;
                if (best_ms < 0 || per < best_ms) {
                    best_ms = per;
                    best_ty = c->g.ty;
                    best_remap = remap;
                    best_xpl = xpl;
                }
            }
        }
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune_cache[key] = std::make_tuple(best_ty, best_remap, best_xpl);
    }
    if (!set_traversal(c, best_ty, best_remap, best_xpl))
        return fail(c, GPE_EINVAL, "traversal needs more workgroups than the decode supports");
    int rc = grow_partials(c, max_blocks_before);
    if (rc) return rc;
    if (tile_rows) *tile_rows = best_ty;
    if (xcd_remap) *xcd_remap = best_remap;
    if (x_per_lane) *x_per_lane = best_xpl;
    if (ms_out) *ms_out = best_ms;
    return GPE_OK;
}

int gpe_stream_bandwidth(long long bytes, int reps, double *copy_gbs, double *update_gbs)
{
    if (bytes < (1 << 20) || reps < 1) return GPE_EINVAL;
    bytes = bytes / 16 * 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GPE_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = GPE_ENOMEM;
    if (!rc && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = GPE_EDEVICE;
    if (!rc && (hipMemsetAsync(a, 0, bytes, s) != hipSuccess || hipMemsetAsync(b, 0, bytes, s) != hipSuccess))
        rc = GPE_EDEVICE;
    double *out[2] = {copy_gbs, update_gbs};
    const double moved[2] = {2.0 * bytes, 3.0 * bytes};
    for (int kind = 0; kind < 2 && !rc; ++kind) {
        bool ok = launch_stream(kind, a, b, bytes / 16, s) == hipSuccess && hipEventRecord(e0, s) == hipSuccess;
        for (int r = 0; r < reps && ok; ++r) ok = launch_stream(kind, a, b, bytes / 16, s) == hipSuccess;
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float ms = 0.f;
        ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (!ok) rc = GPE_EDEVICE;
        else if (out[kind]) *out[kind] = moved[kind] * reps / (ms * 1e-3) / 1e9;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int gpe_last_step_timing(const gpe_ctx *c, double *ms_per_sweep, double *bytes_per_sweep)
{
    if (!c) return GPE_EINVAL;
    if (ms_per_sweep) *ms_per_sweep = c->last_ms_per_sweep;
    if (bytes_per_sweep) {
        const double n = (double)c->cfg.nx * c->cfg.ny * c->cfg.nz * c->g.V;
        *bytes_per_sweep = n * 3.0 * c->cfg.ncomp * 2.0 * (c->fp64 ? 8.0 : 4.0);
    }
    return GPE_OK;
}

}  // extern "C"
