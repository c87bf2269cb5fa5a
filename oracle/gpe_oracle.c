/*
 * FP64 CPU oracle for the DEC Gross-Pitaevskii step -- TEST INFRASTRUCTURE ONLY.
 * See gpe_oracle.h for what it restates and for the state layout.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 */
#include "gpe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cx(double re, double im) { cplx c = {re, im}; return c; }
static inline cplx cadd(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cx(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cscale(double s, cplx a) { return cx(s * a.re, s * a.im); }
static inline cplx cconj(cplx a) { return cx(a.re, -a.im); }
static inline double cnorm2(cplx a) { return a.re * a.re + a.im * a.im; }

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ---------------------------------------------------------------- ramps ---- */

/* One channel of getSignal (AliceRingRamps.h:80-107 for Bq, :111-138 for Bb). */
static void ramp_channel(int n, const double *val, const double *dur, const int *type, int width,
                         double tau, double t, double *out)
{
    int k;
    if (n <= 0) {
        for (k = 0; k < width; ++k) out[k] = 0.0;
        return;
    }
    int idx = 0;
    for (; idx < n; ++idx) {
        double tin = t - dur[idx];
        if (tin < 0) break;
        t = tin;
    }
    if (idx >= n) { /* past the end: hold the last value */
        for (k = 0; k < width; ++k) out[k] = val[(n - 1) * width + k];
        return;
    }
    for (k = 0; k < width; ++k) {
        double prev = idx > 0 ? val[(idx - 1) * width + k] : 0.0;
        double cur = val[idx * width + k];
        switch (type[idx]) {
        case 1: out[k] = prev + t * (cur - prev) / dur[idx]; break;
        case 2: out[k] = prev + (cur - prev) * sqrt(t / dur[idx]); break;
        case 3: out[k] = prev + (cur - prev) * (1.0 - exp(-t / tau)); break;
        default: out[k] = cur; break;
        }
    }
}

void orc_ramp_signal(const orc_ramp *r, double t, double *bq, double bb[3])
{
    double tr = t * r->time_scale;
    ramp_channel(r->nq, r->q_val, r->q_dur, r->q_type, 1, r->tau, tr, bq);
    ramp_channel(r->nb, r->b_val, r->b_dur, r->b_type, 3, r->tau, tr, bb);
}

void orc_ramp_apply(const orc_ramp *r, double t, orc_params *p)
{
    double bq, bb[3];
    int k;
    orc_ramp_signal(r, t, &bq, bb);
    p->bq = r->scale_q * bq;
    p->bq_quad = r->qscale_q * bq;
    for (k = 0; k < 3; ++k) {
        p->bb[k] = r->scale_b * bb[k];
        p->bb_quad[k] = r->qscale_b * bb[k];
    }
}

/* ---------------------------------------------------------- local terms ---- */

/* (W . f) psi for spin F in the m = F..-F ordering, accumulated into out.
 *   (f+ psi)_m = sqrt(F(F+1) - (m-1) m) psi_{m-1},  (f- psi)_m = sqrt(F(F+1) - (m+1) m) psi_{m+1}
 *   W . f = Wz fz + (W- f+ + W+ f-) / 2,  W+- = Wx +- i Wy.
 * Spin-1 expanded form: Spin1GpeBCC.cu:890-895; spin-2: Spin2GpeBCC.cu:939-970. */
static void spin_dot_apply(int F, const double W[3], const cplx *psi, cplx *out)
{
    const double ff = F * (F + 1.0);
    const cplx wm = cx(0.5 * W[0], -0.5 * W[1]); /* W-/2 */
    const cplx wp = cx(0.5 * W[0], 0.5 * W[1]);  /* W+/2 */
    int c;
    for (c = 0; c <= 2 * F; ++c) {
        int m = F - c;
        cplx acc = cscale(W[2] * m, psi[c]);
        if (c + 1 <= 2 * F) /* psi_{m-1} */
            acc = cadd(acc, cmul(wm, cscale(sqrt(ff - (m - 1.0) * m), psi[c + 1])));
        if (c - 1 >= 0) /* psi_{m+1} */
            acc = cadd(acc, cmul(wp, cscale(sqrt(ff - (m + 1.0) * m), psi[c - 1])));
        out[c] = cadd(out[c], acc);
    }
}

/* Spin density F = psi^dagger f psi: Spin1GpeBCC.cu:886-888, Spin2GpeBCC.cu:939,954-957 */
static void spin_density(int F, const cplx *psi, double Fv[3])
{
    const double ff = F * (F + 1.0);
    cplx fp = cx(0, 0);
    double fz = 0;
    int c;
    for (c = 0; c <= 2 * F; ++c) {
        int m = F - c;
        fz += m * cnorm2(psi[c]);
        if (c + 1 <= 2 * F)
            fp = cadd(fp, cscale(sqrt(ff - (m - 1.0) * m), cmul(cconj(psi[c]), psi[c + 1])));
    }
    Fv[0] = fp.re;
    Fv[1] = fp.im;
    Fv[2] = fz;
}

static double trap_potential(const orc_params *p, const double r[3], double t)
{
    /* Spin1GpeBCC.cu:112-122 (the offset is an explicit parameter here) */
    if (t >= p->trap_off_time) return 0.0;
    double x = p->trap[0] * r[0], y = p->trap[1] * r[1], z = p->trap[2] * r[2];
    return 0.5 * (x * x + y * y + z * z) + p->trap_offset;
}

static void field_at(double bq, const double bb[3], const double ctr[3], const double r[3], double B[3])
{
    /* Spin1GpeBCC.cu:128-142 */
    B[0] = bq * (r[0] - ctr[0]) + bb[0];
    B[1] = bq * (r[1] - ctr[1]) + bb[1];
    B[2] = -2.0 * bq * (r[2] - ctr[2]) + bb[2];
}

static void local_h(const orc_params *p, const double r[3], double t, const cplx *psi, cplx *h)
{
    const int C = p->ncomp;
    const int F = (C - 1) / 2;
    double n = 0;
    int c;
    for (c = 0; c < C; ++c) n += cnorm2(psi[c]);
    /* density terms: V + c0 n - i alpha n^2 (Spin1GpeBCC.cu:876-884) */
    const cplx diag = cx(trap_potential(p, r, t) + p->c0 * n, -p->alpha * n * n);
    for (c = 0; c < C; ++c) h[c] = cmul(diag, psi[c]);
    if (F == 0) return;

    /* linear Zeeman + spin exchange: W = c1 F + s B, s = +1 (spin-1), -1 (spin-2) */
    double Fv[3], B[3], W[3];
    spin_density(F, psi, Fv);
    field_at(p->bq, p->bb, p->center, r, B);
    const double s = (F == 1) ? 1.0 : -1.0;
    for (c = 0; c < 3; ++c) W[c] = p->c1 * Fv[c] + s * B[c];
    spin_dot_apply(F, W, psi, h);

    if (F == 2 && p->c2 != 0.0) {
        /* singlet pairing: H_m += (c2/5) Theta (-1)^m psi*_{-m},
         * Theta = sum_m (-1)^m psi_m psi_{-m}  (Spin2GpeBCC.cu:941-970) */
        cplx theta = cx(0, 0);
        for (c = 0; c < C; ++c) {
            cplx term = cmul(psi[c], psi[C - 1 - c]);
            theta = (c & 1) ? csub(theta, term) : cadd(theta, term);
        }
        for (c = 0; c < C; ++c) {
            cplx term = cscale(p->c2 / 5.0, cmul(theta, cconj(psi[C - 1 - c])));
            h[c] = (c & 1) ? csub(h[c], term) : cadd(h[c], term);
        }
    }

    /* quadratic Zeeman quad_sign (Bquad . f)^2 (Spin1GpeBCC.cu:897-912) */
    if (p->bq_quad != 0.0 || p->bb_quad[0] != 0.0 || p->bb_quad[1] != 0.0 || p->bb_quad[2] != 0.0) {
        double Bqd[3];
        cplx tmp[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
        cplx tmp2[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
        field_at(p->bq_quad, p->bb_quad, p->center, r, Bqd);
        spin_dot_apply(F, Bqd, psi, tmp);
        spin_dot_apply(F, Bqd, tmp, tmp2);
        for (c = 0; c < C; ++c) h[c] = cadd(h[c], cscale(p->quad_sign, tmp2[c]));
    }
    /* constant quadratic Zeeman q f_z^2 */
    if (p->q != 0.0)
        for (c = 0; c < C; ++c) {
            int m = F - c;
            h[c] = cadd(h[c], cscale(p->q * m * m, psi[c]));
        }
}

void orc_local_h(const orc_params *p, const double r[3], double t, const double *psi_node, double *h_node)
{
    local_h(p, r, t, (const cplx *)psi_node, (cplx *)h_node);
}

/* ------------------------------------------------------------- geometry ---- */

static inline size_t ncells(const orc_params *p) { return (size_t)p->nx * p->ny * p->nz; }

void orc_position(const orc_lattice *L, const orc_params *p, int x, int y, int z, int v, double r[3])
{
    /* GPBlocks.cpp:487 */
    const double s = p->block_scale, *w = L->block_width;
    r[0] = p->p0[0] + s * (x * w[0] + L->pos[3 * v + 0]);
    r[1] = p->p0[1] + s * (y * w[1] + L->pos[3 * v + 1]);
    r[2] = p->p0[2] + s * (z * w[2] + L->pos[3 * v + 2]);
}

/* H psi at block (x,y,z) unknown v for all components.
 * Kinetic: sum_f w_f (psi_t(f) - psi_v), w_f = -0.5 hodge_f / s^2 with zero outside the box
 * (Spin1GpeBCC.cu:827-865, GPBlocks.cpp:569-579, Spin1GpeBCC.cu:1169). */
static void h_at(const orc_lattice *L, const orc_params *p, double t, const cplx *psi, int x, int y, int z,
                 int v, cplx *h)
{
    const int C = p->ncomp, V = L->V;
    const size_t nc = ncells(p) * V;
    const size_t me = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
    const double inv_s2 = 1.0 / (p->block_scale * p->block_scale);
    cplx own[5], kin[5];
    int c, f;
    for (c = 0; c < C; ++c) {
        own[c] = psi[c * nc + me];
        kin[c] = cx(0, 0);
    }
    for (f = L->face_start[v]; f < L->face_start[v + 1]; ++f) {
        const int nx_ = x + L->face_d[3 * f], ny_ = y + L->face_d[3 * f + 1], nz_ = z + L->face_d[3 * f + 2];
        const double w = -0.5 * L->face_hodge[f] * inv_s2;
        int inside = nx_ >= 0 && nx_ < p->nx && ny_ >= 0 && ny_ < p->ny && nz_ >= 0 && nz_ < p->nz;
        size_t nb = inside ? ((size_t)(nz_ * p->ny + ny_) * p->nx + nx_) * V + L->face_tgt[f] : 0;
        for (c = 0; c < C; ++c) {
            cplx o = inside ? psi[c * nc + nb] : cx(0, 0);
            kin[c] = cadd(kin[c], cscale(w, csub(o, own[c])));
        }
    }
    double r[3];
    orc_position(L, p, x, y, z, v, r);
    local_h(p, r, t, own, h);
    for (c = 0; c < C; ++c) h[c] = cadd(h[c], kin[c]);
}

void orc_apply_h(const orc_lattice *L, const orc_params *p, double t, const double *psi_, double *hpsi_)
{
    const cplx *psi = (const cplx *)psi_;
    cplx *hpsi = (cplx *)hpsi_;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V;
    int z;
#pragma omp parallel for schedule(static)
    for (z = 0; z < p->nz; ++z) {
        int x, y, v, c;
        cplx h[5];
        for (y = 0; y < p->ny; ++y)
            for (x = 0; x < p->nx; ++x)
                for (v = 0; v < V; ++v) {
                    size_t me = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
                    h_at(L, p, t, psi, x, y, z, v, h);
                    for (c = 0; c < C; ++c) hpsi[c * nc + me] = h[c];
                }
    }
}

void orc_update(const orc_lattice *L, const orc_params *p, double t, double coef, double *next_,
                const double *cur_)
{
    /* next += coef * (-i) H(cur): GPBlocks.cpp:578-579, Spin1GpeBCC.cu:914-916 */
    const cplx *cur = (const cplx *)cur_;
    cplx *next = (cplx *)next_;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V;
    int z;
#pragma omp parallel for schedule(static)
    for (z = 0; z < p->nz; ++z) {
        int x, y, v, c;
        cplx h[5];
        for (y = 0; y < p->ny; ++y)
            for (x = 0; x < p->nx; ++x)
                for (v = 0; v < V; ++v) {
                    size_t me = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
                    h_at(L, p, t, cur, x, y, z, v, h);
                    for (c = 0; c < C; ++c) {
                        cplx *o = &next[c * nc + me];
                        o->re += coef * h[c].im;
                        o->im -= coef * h[c].re;
                    }
                }
    }
}

void orc_forward_euler(const orc_lattice *L, const orc_params *p, double t, double dt, double *next,
                       const double *cur)
{
    /* Spin1GpeBCC.cu:788-790 */
    memcpy(next, cur, sizeof(double) * 2 * p->ncomp * ncells(p) * L->V);
    orc_update(L, p, t, dt, next, cur);
}

void orc_run_leapfrog3(const orc_lattice *L, const orc_params *p_, const orc_ramp *ramp, double t0, double dt,
                       int n, double *a, double *b)
{
    /* host loop Spin1GpeBCC.cu:1571-1603: sweep j computes psi_{k+j+1} from psi_{k+j} at
     * time t0 + j dt, in place over psi_{k+j-1}. */
    orc_params p = *p_;
    const size_t sz = sizeof(double) * 2 * p.ncomp * ncells(&p) * L->V;
    int j;
    for (j = 0; j < n; ++j) {
        double t = t0 + j * dt;
        if (ramp) orc_ramp_apply(ramp, t, &p);
        orc_update(L, &p, t, 2.0 * dt, a, b);
        double *tmp = a;
        a = b;
        b = tmp;
    }
    if (n & 1) { /* results live in swapped buffers: move them back */
        double *buf = (double *)malloc(sz);
        memcpy(buf, a, sz);
        memcpy(a, b, sz);
        memcpy(b, buf, sz);
        free(buf);
    }
}

void orc_run_staggered(const orc_lattice *L, const orc_params *p, double dt, int n, double *even, double *odd)
{
    /* GPBlocks.cpp:560-606 */
    int j;
    for (j = 0; j < n; ++j) {
        orc_update(L, p, 0.0, dt, odd, even);
        orc_update(L, p, 0.0, dt, even, odd);
    }
}

/* ----------------------------------------------------------- reductions ---- */

double orc_norm(const orc_lattice *L, const orc_params *p, const double *psi_)
{
    /* density + integrate (Spin1GpeBCC.cu:191-211, 328-349) with per-unknown dual volume */
    const cplx *psi = (const cplx *)psi_;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V, ncell = ncells(p);
    const double s3 = p->block_scale * p->block_scale * p->block_scale;
    double total = 0;
    size_t i;
    int v, c;
    for (v = 0; v < V; ++v) {
        double acc = 0;
        for (i = 0; i < ncell; ++i)
            for (c = 0; c < C; ++c) acc += cnorm2(psi[c * nc + i * V + v]);
        total += acc * L->vol[v] * s3;
    }
    return total;
}

void orc_normalize(const orc_lattice *L, const orc_params *p, double *psi, double target)
{
    /* normalize (Spin1GpeBCC.cu:412-434) */
    const double nrm = orc_norm(L, p, psi);
    const double sc = sqrt(target / nrm);
    const size_t n = 2 * (size_t)p->ncomp * ncells(p) * L->V;
    size_t i;
    for (i = 0; i < n; ++i) psi[i] *= sc;
}

double orc_itp_step(const orc_lattice *L, const orc_params *p, double dt, double *psi_, double *scratch)
{
    /* H then next -= dt H, then normalise (GPUGrossPitaevskii_BCC.cu:845-849, Spin1GpeBCC.cu:659-661) */
    const size_t n = 2 * (size_t)p->ncomp * ncells(p) * L->V;
    size_t i;
    orc_apply_h(L, p, 0.0, psi_, scratch);
    for (i = 0; i < n; ++i) psi_[i] -= dt * scratch[i];
    double nrm = orc_norm(L, p, psi_);
    const double sc = 1.0 / sqrt(nrm);
    for (i = 0; i < n; ++i) psi_[i] *= sc;
    return nrm;
}

double orc_expect_h(const orc_lattice *L, const orc_params *p, double t, const double *psi_, double *scratch)
{
    /* Re <psi|H psi> dv: the reference's "energy" (GPUGrossPitaevskii_BCC.cu:204-226) */
    const cplx *psi = (const cplx *)psi_;
    const cplx *hp = (const cplx *)scratch;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V, ncell = ncells(p);
    const double s3 = p->block_scale * p->block_scale * p->block_scale;
    double total = 0;
    size_t i;
    int v, c;
    orc_apply_h(L, p, t, psi_, scratch);
    for (v = 0; v < V; ++v) {
        double acc = 0;
        for (i = 0; i < ncell; ++i)
            for (c = 0; c < C; ++c) {
                size_t k = c * nc + i * V + v;
                acc += psi[k].re * hp[k].re + psi[k].im * hp[k].im;
            }
        total += acc * L->vol[v] * s3;
    }
    return total;
}

double orc_energy(const orc_lattice *L, const orc_params *p, double t, const double *psi_)
{
    /* GP energy functional (not computed by the reference, whose "energy" is <psi|H psi>):
     * E = sum dv [ psi^dag K psi + V n + c0/2 n^2 + c1/2 |F|^2 + s B.F + c2/10 |Theta|^2
     *              + quad_sign psi^dag (Bquad.f)^2 psi + q psi^dag fz^2 psi ]
     * with K psi = -1/2 L psi.  dE/dpsi*_i = dv_i (H psi)_i for the Hermitian part of H. */
    const cplx *psi = (const cplx *)psi_;
    const int V = L->V, C = p->ncomp, F = (C - 1) / 2;
    const size_t nc = ncells(p) * V;
    const double s3 = p->block_scale * p->block_scale * p->block_scale;
    const double inv_s2 = 1.0 / (p->block_scale * p->block_scale);
    double total = 0;
    int x, y, z, v, c, f;
    for (z = 0; z < p->nz; ++z)
        for (y = 0; y < p->ny; ++y)
            for (x = 0; x < p->nx; ++x)
                for (v = 0; v < V; ++v) {
                    const size_t me = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
                    cplx own[5];
                    double e = 0, n = 0, r[3];
                    for (c = 0; c < C; ++c) {
                        own[c] = psi[c * nc + me];
                        n += cnorm2(own[c]);
                    }
                    /* kinetic: psi*_v sum_f w_f (psi_t - psi_v) */
                    for (f = L->face_start[v]; f < L->face_start[v + 1]; ++f) {
                        const int ax = x + L->face_d[3 * f], ay = y + L->face_d[3 * f + 1],
                                  az = z + L->face_d[3 * f + 2];
                        const double w = -0.5 * L->face_hodge[f] * inv_s2;
                        int inside = ax >= 0 && ax < p->nx && ay >= 0 && ay < p->ny && az >= 0 && az < p->nz;
                        size_t nb = inside ? ((size_t)(az * p->ny + ay) * p->nx + ax) * V + L->face_tgt[f] : 0;
                        for (c = 0; c < C; ++c) {
                            cplx o = inside ? psi[c * nc + nb] : cx(0, 0);
                            cplx d = csub(o, own[c]);
                            e += w * (own[c].re * d.re + own[c].im * d.im);
                        }
                    }
                    orc_position(L, p, x, y, z, v, r);
                    e += trap_potential(p, r, t) * n + 0.5 * p->c0 * n * n;
                    if (F > 0) {
                        double Fv[3], B[3];
                        spin_density(F, own, Fv);
                        field_at(p->bq, p->bb, p->center, r, B);
                        const double s = (F == 1) ? 1.0 : -1.0;
                        e += 0.5 * p->c1 * (Fv[0] * Fv[0] + Fv[1] * Fv[1] + Fv[2] * Fv[2]);
                        e += s * (B[0] * Fv[0] + B[1] * Fv[1] + B[2] * Fv[2]);
                        if (F == 2) {
                            cplx theta = cx(0, 0);
                            for (c = 0; c < C; ++c) {
                                cplx term = cmul(own[c], own[C - 1 - c]);
                                theta = (c & 1) ? csub(theta, term) : cadd(theta, term);
                            }
                            e += p->c2 / 10.0 * cnorm2(theta);
                        }
                        if (p->bq_quad != 0.0 || p->bb_quad[0] != 0.0 || p->bb_quad[1] != 0.0 ||
                            p->bb_quad[2] != 0.0) {
                            double Bqd[3];
                            cplx tmp[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
                            field_at(p->bq_quad, p->bb_quad, p->center, r, Bqd);
                            spin_dot_apply(F, Bqd, own, tmp);
                            for (c = 0; c < C; ++c) e += p->quad_sign * cnorm2(tmp[c]);
                        }
                        for (c = 0; c < C; ++c) {
                            int m = F - c;
                            e += p->q * m * m * cnorm2(own[c]);
                        }
                    }
                    total += e * L->vol[v] * s3;
                }
    return total;
}

void orc_magnetization(const orc_lattice *L, const orc_params *p, const double *psi_, double out[3])
{
    /* sum dv F (integrateSpinAndDensity weights n * F/n, Spin1GpeBCC.cu:235-271,992-1020) */
    const cplx *psi = (const cplx *)psi_;
    const int V = L->V, C = p->ncomp, F = (C - 1) / 2;
    const size_t nc = ncells(p) * V, ncell = ncells(p);
    const double s3 = p->block_scale * p->block_scale * p->block_scale;
    size_t i;
    int v, c, k;
    out[0] = out[1] = out[2] = 0;
    if (F == 0) return;
    for (v = 0; v < V; ++v) {
        double acc[3] = {0, 0, 0};
        for (i = 0; i < ncell; ++i) {
            cplx own[5];
            double Fv[3];
            for (c = 0; c < C; ++c) own[c] = psi[c * nc + i * V + v];
            spin_density(F, own, Fv);
            for (k = 0; k < 3; ++k) acc[k] += Fv[k];
        }
        for (k = 0; k < 3; ++k) out[k] += acc[k] * L->vol[v] * s3;
    }
}

double orc_max_hamiltonian(const orc_lattice *L, const orc_params *p, double t, const double *psi_)
{
    /* maxHamilton + reduceMax (Spin1GpeBCC.cu:144-189,1022-1037): per unknown p = V + c0 n,
     * B = B(r) + c1 psi^dag F psi (not divided by n), max(p + |Bx|/sqrt2, p + |By|/sqrt2,
     * p + |Bz|); the scalar form is the max of V + g n of GPBlocks.cpp:488-495.  Spin-1 or
     * scalar only (returns NaN for spin-2, which has no such bound in the reference). */
    const cplx *psi = (const cplx *)psi_;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V;
    double best = -INFINITY;
    int x, y, z, v, c;
    if (C == 5) return NAN;
    for (z = 0; z < p->nz; ++z)
        for (y = 0; y < p->ny; ++y)
            for (x = 0; x < p->nx; ++x)
                for (v = 0; v < V; ++v) {
                    const size_t i = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
                    cplx own[3];
                    double r[3], n = 0, h;
                    orc_position(L, p, x, y, z, v, r);
                    for (c = 0; c < C; ++c) {
                        own[c] = psi[c * nc + i];
                        n += cnorm2(own[c]);
                    }
                    const double pot = trap_potential(p, r, t) + p->c0 * n;
                    h = pot;
                    if (C == 3) {
                        double Fv[3], B[3];
                        spin_density(1, own, Fv);
                        field_at(p->bq, p->bb, p->center, r, B);
                        for (c = 0; c < 3; ++c) B[c] += p->c1 * Fv[c];
                        h = fmax(pot + fabs(B[0]) / sqrt(2.0), fmax(pot + fabs(B[1]) / sqrt(2.0), pot + fabs(B[2])));
                    }
                    if (h > best) best = h;
                }
    return best;
}

double orc_weighted_diff(const orc_lattice *L, const orc_params *p, const double *psi_, const double *ref_,
                         double phase_re, double phase_im)
{
    /* weightedDiff + integrate (Spin1GpeBCC.cu:436-461,1620-1633): sum dv |a|^2 |psi - a|^2
     * over unknowns, a = phase * ref. */
    const cplx *psi = (const cplx *)psi_, *ref = (const cplx *)ref_;
    const int V = L->V, C = p->ncomp;
    const size_t nc = ncells(p) * V, ncell = ncells(p);
    const double s3 = p->block_scale * p->block_scale * p->block_scale;
    const cplx ph = cx(phase_re, phase_im);
    double total = 0;
    size_t i;
    int v, c;
    for (v = 0; v < V; ++v) {
        double acc = 0;
        for (i = 0; i < ncell; ++i) {
            double an = 0, dn = 0;
            for (c = 0; c < C; ++c) {
                const cplx a = cmul(ph, ref[c * nc + i * V + v]);
                an += cnorm2(a);
                dn += cnorm2(csub(psi[c * nc + i * V + v], a));
            }
            acc += an * dn;
        }
        total += acc * L->vol[v] * s3;
    }
    return total;
}

/* --------------------------------------------------------- vortex state ---- */

void orc_vortex_init(int rsize, double rstep, double rplug, int zsize, double zstep, double *f, double *mu)
{
    /* VortexState::setDimensions (VortexState.cpp:12-40) */
    const double PIx2 = 6.283185307179586476925286766559;
    const double PI_ = 3.1415926535897932384626433832795;
    double prob = 0.0;
    int i, j, ii = 0;
    for (i = 0; i < rsize; i++) {
        const double fr = 1.0 - cos(i * PIx2 / (double)rsize);
        const double r = rplug + i * rstep;
        int times = 1;
        for (j = 0; j < zsize; j++, ii++) {
            f[ii] = (1.0 + cos(j * PI_ / (double)zsize)) * fr;
            prob += times * f[ii] * f[ii] * r;
            times = 2;
        }
    }
    prob *= PIx2 * rstep * zstep;
    const double scale = 1.0 / sqrt(prob);
    for (ii = 0; ii < rsize * zsize; ii++) f[ii] *= scale;
    *mu = 1.0;
}

double orc_vortex_iterate(int kappa, double g, int rsize, double rstep, double rplug, int zsize, double zstep,
                          double ratio, double *f, double *mu, int maxiters, double eps)
{
    /* VortexState::iterateSolution (VortexState.cpp:109-196), potential
     * 0.5 (r^2 + ratio^2 z^2) as potentialRZ (GPBlocks.cpp:28-31). Returns the last diff. */
    const double PIx2 = 6.283185307179586476925286766559;
    const double PIx2drdz = PIx2 * rstep * zstep;
    const double kk = 0.5 * (double)(kappa * kappa);
    const double laplr = 0.5 / (rstep * rstep);
    const double laplz = (zsize > 1 ? 0.5 / (zstep * zstep) : 0.0);
    const int fsize = rsize * zsize;
    double diff = 0;
    int i, j, k, ii;
    for (j = 0; j < zsize; j++) f[j] = 0.0;
    for (k = 0; k < maxiters; k++) {
        diff = 0.0;
        for (i = 1, ii = zsize; i < rsize; i++) {
            const double r = rplug + i * rstep;
            const double lapl = kk / (r * r) + 2.0 * (laplr + laplz);
            const double gradr = 0.25 / (r * rstep);
            double dv = PIx2drdz * r;
            for (j = 0; j < zsize; j++, ii++) {
                const double zz = j * zstep;
                const double pot = 0.5 * (r * r + ratio * ratio * zz * zz) + g * f[ii] * f[ii];
                const double fiir = (i + 1 < rsize ? f[ii + zsize] : 0.0);
                const double frii = f[ii - zsize];
                const double fiiz = (j + 1 < zsize ? f[ii + 1] : 0.0);
                const double fzii = (j > 0 ? f[ii - 1] : fiiz);
                double dfii = f[ii];
                f[ii] = (gradr * (frii - fiir) - laplr * (frii + fiir) - laplz * (fzii + fiiz)) / (*mu - lapl - pot);
                dfii -= f[ii];
                diff += dfii * dfii * dv;
                if (j == 0) dv *= 2.0;
            }
        }
        double prob = 0.0, prob2 = 0.0;
        for (i = 1, ii = zsize; i < rsize; i++) {
            const double r = rplug + i * rstep;
            const double lapl = kk / (r * r) + 2.0 * (laplr + laplz);
            const double gradr = 0.25 / (r * rstep);
            double dv = PIx2drdz * r;
            for (j = 0; j < zsize; j++, ii++) {
                const double zz = j * zstep;
                const double pot = 0.5 * (r * r + ratio * ratio * zz * zz) + g * f[ii] * f[ii];
                const double fiir = (i + 1 < rsize ? f[ii + zsize] : 0.0);
                const double frii = f[ii - zsize];
                const double fiiz = (j + 1 < zsize ? f[ii + 1] : 0.0);
                const double fzii = (j > 0 ? f[ii - 1] : fiiz);
                prob += f[ii] * f[ii] * dv;
                prob2 += ((pot + lapl) * f[ii] + gradr * (frii - fiir) - laplr * (frii + fiir) -
                          laplz * (fzii + fiiz)) * f[ii] * dv;
                if (j == 0) dv *= 2.0;
            }
        }
        *mu = prob2 / prob;
        const double scale = 1.0 / sqrt(prob);
        for (ii = zsize; ii < fsize; ii++) f[ii] *= scale;
        if (diff < eps) break;
    }
    double sum = 0.0;
    for (ii = zsize; ii < fsize; ii++) sum += f[ii];
    if (sum < 0.0)
        for (ii = zsize; ii < fsize; ii++) f[ii] = -f[ii];
    return diff;
}

double orc_vortex_function(const double *f, int rsize, double rstep, double rplug, int zsize, double zstep,
                           double r, double z)
{
    /* VortexState::getFunction (VortexState.cpp:220-238) */
    const double rr = (r - rplug) / rstep;
    if (rr < 0.0) return 0.0;
    const unsigned ir = (unsigned)rr;
    if (ir >= (unsigned)rsize) return 0.0;
    const double dr = rr - (double)ir;
    const double zz = fabs(z) / zstep;
    const unsigned iz = (unsigned)zz;
    if (iz >= (unsigned)zsize) return 0.0;
    const double dz = zz - (double)iz;
    const unsigned ii = ir * zsize + iz;
    const unsigned fsize = rsize * zsize;
    const double fiir = (ir + 1 < (unsigned)rsize ? f[ii + zsize] : 0.0);
    const double fiiz = (iz + 1 < (unsigned)zsize ? f[ii + 1] : 0.0);
    const double fiirz = (ii + zsize + 1 < fsize ? f[ii + zsize + 1] : 0.0);
    return (1.0 - dz) * ((1.0 - dr) * f[ii] + dr * fiir) + dz * ((1.0 - dr) * fiiz + dr * fiirz);
}

/* --------------------------------------------------------------- MT19937 --- */
/* Textbook MT19937 (Matsumoto & Nishimura 1998), as Types/Random.cpp:12-118. */

void orc_mt_seed(unsigned int *st, unsigned int seed)
{
    int i;
    st[0] = seed;
    for (i = 1; i < 624; i++) st[i] = (1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (unsigned)i);
    st[624] = 624;
}

static unsigned int mt_next(unsigned int *st)
{
    if (st[624] >= 624) {
        int i;
        for (i = 0; i < 624; i++) {
            unsigned int y = (st[i] & 0x80000000u) | (st[(i + 1) % 624] & 0x7fffffffu);
            st[i] = st[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        st[624] = 0;
    }
    unsigned int y = st[st[624]++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orc_mt_uniform_circle(unsigned int *st, double out[2])
{
    /* Random::getUniformCircle (Random.cpp:77-86): int(u) / 2^31 in [-1, 1), rejection to the disk */
    const double INT2DOUBLE = 1.0 / 2147483648.0;
    do {
        out[0] = INT2DOUBLE * (int)mt_next(st);
        out[1] = INT2DOUBLE * (int)mt_next(st);
    } while (out[0] * out[0] + out[1] * out[1] > 1.0);
}

/* ------------------------------------------- GPBlocks integrateInTime ---- */

int orc_gpblocks_integrate(const orc_lattice *L, int kappa, double g, const double *f, int rsize, double rstep,
                           double rplug, int zsize, double zstep, double mu, double block_scale, const double minp[3],
                           const double maxp[3], unsigned period, unsigned iterations, double *out_normsq,
                           double *out_error, double *out_dt, int *out_steps)
{
    /* Restates GrossPitaevskii/GPBlocks.cpp:451-609 for the BCC table: box sizing (:456-460),
     * stationary state and potential (:473-495), dt rule (:497-511), odd = e^{i mu dt/2} even
     * (:516-523), per-iteration norm and overlap error (:530-538), staggered loop (:560-606).
     * Returns 0, or -1 on allocation failure. */
    orc_params p;
    memset(&p, 0, sizeof(p));
    const double *w = L->block_width;
    p.ncomp = 1;
    p.nx = (int)((unsigned)((maxp[0] - minp[0]) / (block_scale * w[0])) + 1);
    p.ny = (int)((unsigned)((maxp[1] - minp[1]) / (block_scale * w[1])) + 1);
    p.nz = (int)((unsigned)((maxp[2] - minp[2]) / (block_scale * w[2])) + 1);
    p.block_scale = block_scale;
    int a;
    for (a = 0; a < 3; ++a) {
        int n = a == 0 ? p.nx : (a == 1 ? p.ny : p.nz);
        p.p0[a] = 0.5 * (minp[a] + maxp[a] - block_scale * w[a] * n);
    }
    p.trap[0] = p.trap[1] = p.trap[2] = 1.0;
    p.trap_off_time = 1e300;
    p.c0 = g;
    const size_t N = (size_t)p.nx * p.ny * p.nz * L->V;
    cplx *psi0 = (cplx *)calloc(N, sizeof(cplx));
    cplx *even = (cplx *)calloc(N, sizeof(cplx));
    cplx *odd = (cplx *)calloc(N, sizeof(cplx));
    if (!psi0 || !even || !odd) {
        free(psi0); free(even); free(odd);
        return -1;
    }
    double maxpot = 0.0;
    int x, y, z, v, k;
    for (z = 0; z < p.nz; ++z)
        for (y = 0; y < p.ny; ++y)
            for (x = 0; x < p.nx; ++x)
                for (v = 0; v < L->V; ++v) {
                    double r[3];
                    size_t me = ((size_t)(z * p.ny + y) * p.nx + x) * L->V + v;
                    orc_position(L, &p, x, y, z, v, r);
                    /* VortexState::getPsi (VortexState.cpp:204-212) */
                    const double rr = sqrt(r[0] * r[0] + r[1] * r[1]);
                    cplx c = cx(0, 0);
                    if (!(rr < 1e-30)) {
                        c = cx(orc_vortex_function(f, rsize, rstep, rplug, zsize, zstep, rr, r[2]), 0.0);
                        const cplx nn = cx(r[0] / rr, r[1] / rr);
                        for (k = 0; k < kappa; ++k) c = cmul(c, nn);
                    }
                    psi0[me] = c;
                    const double pot = 0.5 * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                    const double poti = pot + g * cnorm2(c);
                    if (poti > maxpot) maxpot = poti;
                }
    /* dt rule: lapfac0 = degree * 0.5 * hodge / s^2 with the table's (uniform) hodge */
    const int deg = L->face_start[1] - L->face_start[0];
    const double lapfac0 = deg * 0.5 * L->face_hodge[0] / (block_scale * block_scale);
    const unsigned steps = (unsigned)(period * (maxpot + lapfac0)) + 1;
    const double dt = period / (double)steps;
    *out_dt = dt;
    *out_steps = (int)steps;
    const cplx odd_phase = cx(cos(-0.5 * dt * mu), -sin(-0.5 * dt * mu));
    size_t i;
    for (i = 0; i < N; ++i) {
        even[i] = psi0[i];
        odd[i] = cmul(odd_phase, even[i]);
    }
    const double vol = block_scale * block_scale * block_scale * L->vol[0];
    unsigned iter = 0;
    while (1) {
        double normsq = 0.0;
        cplx err = cx(0, 0);
        for (i = 0; i < N; ++i) {
            normsq += cnorm2(even[i]) * vol;
            err = cadd(err, cscale(vol, cmul(cconj(psi0[i]), even[i])));
        }
        out_normsq[iter] = normsq;
        out_error[iter] = normsq - sqrt(cnorm2(err));
        if (++iter > iterations) break;
        orc_run_staggered(L, &p, dt, (int)steps, (double *)even, (double *)odd);
    }
    free(psi0); free(even); free(odd);
    return 0;
}

/* ------------------------------------------------------------ discrete adjoint ---- */

/* G'(psi)^T mu at one unknown for the local part, in the general spin-F form:
 *   i [ (V + c0 n) mu + (c1 F + s B).f mu + quad_sign (Q.f)^2 mu + q fz^2 mu ]
 *   - 2 c0 Im(psi^dag mu) psi - 2 c1 sum_a Im(psi^dag f_a mu) f_a psi
 *   + i (2 c2/5) u (u^dag mu) - i (c2/5) Theta S conj(mu_rev)    (spin-2, u = S conj(psi_rev))
 * plus i * kin, supplied by the caller.  Result accumulated into out. */
static void adj_local(const orc_params *p, const double r[3], double t, const cplx *psi, const cplx *mu,
                      const cplx *kin, cplx *out)
{
    const int C = p->ncomp, F = (C - 1) / 2;
    cplx h[5];
    double n = 0;
    int c, a;
    for (c = 0; c < C; ++c) n += cnorm2(psi[c]);
    const double d = trap_potential(p, r, t) + p->c0 * n;
    for (c = 0; c < C; ++c) h[c] = cadd(kin[c], cscale(d, mu[c]));
    if (F > 0) {
        double Fv[3], B[3], W[3];
        spin_density(F, psi, Fv);
        field_at(p->bq, p->bb, p->center, r, B);
        const double s = (F == 1) ? 1.0 : -1.0;
        for (a = 0; a < 3; ++a) W[a] = p->c1 * Fv[a] + s * B[a];
        spin_dot_apply(F, W, mu, h);
        if (p->bq_quad != 0.0 || p->bb_quad[0] != 0.0 || p->bb_quad[1] != 0.0 || p->bb_quad[2] != 0.0) {
            double Q[3];
            cplx t1[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}, t2[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
            field_at(p->bq_quad, p->bb_quad, p->center, r, Q);
            spin_dot_apply(F, Q, mu, t1);
            spin_dot_apply(F, Q, t1, t2);
            for (c = 0; c < C; ++c) h[c] = cadd(h[c], cscale(p->quad_sign, t2[c]));
        }
        for (c = 0; c < C; ++c) {
            int m = F - c;
            h[c] = cadd(h[c], cscale(p->q * m * m, mu[c]));
        }
    }
    /* i h */
    for (c = 0; c < C; ++c) out[c] = cadd(out[c], cx(-h[c].im, h[c].re));
    /* -2 c0 Im(psi^dag mu) psi */
    double ima = 0;
    for (c = 0; c < C; ++c) ima += cmul(cconj(psi[c]), mu[c]).im;
    for (c = 0; c < C; ++c) out[c] = csub(out[c], cscale(2 * p->c0 * ima, psi[c]));
    if (F > 0) {
        double Y[3];
        for (a = 0; a < 3; ++a) {
            double e[3] = {0, 0, 0};
            cplx fm[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
            e[a] = 1.0;
            spin_dot_apply(F, e, mu, fm);
            double im = 0;
            for (c = 0; c < C; ++c) im += cmul(cconj(psi[c]), fm[c]).im;
            Y[a] = im;
        }
        cplx yp[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
        spin_dot_apply(F, Y, psi, yp);
        for (c = 0; c < C; ++c) out[c] = csub(out[c], cscale(2 * p->c1, yp[c]));
        if (F == 2 && p->c2 != 0.0) {
            cplx u[5], udmu = cx(0, 0), theta = cx(0, 0);
            for (c = 0; c < C; ++c) {
                const double sg = (c & 1) ? -1.0 : 1.0;
                u[c] = cscale(sg, cconj(psi[C - 1 - c]));
                udmu = cadd(udmu, cmul(cconj(u[c]), mu[c]));
                theta = cadd(theta, cscale(sg, cmul(psi[c], psi[C - 1 - c])));
            }
            for (c = 0; c < C; ++c) {
                const double sg = (c & 1) ? -1.0 : 1.0;
                cplx tp = cscale(0.4 * p->c2, cmul(u[c], udmu));                          /* (2c2/5) u (u^dag mu) */
                cplx tq = cscale(0.2 * p->c2 * sg, cmul(theta, cconj(mu[C - 1 - c]))); /* (c2/5) Theta S mu*_rev */
                out[c] = cadd(out[c], cx(-tp.im, tp.re));
                out[c] = cadd(out[c], cx(tq.im, -tq.re));
            }
        }
    }
}

/* One adjoint sweep over the grid: mu2 += coef G'(psi)^T mu1; sens[7] += plain-inner-product
 * sensitivities sum Im(conj(mu1) X psi) for X = s f_x, s f_y, s f_z, s (dx fx + dy fy - 2 dz fz),
 * w_x x^2, w_y y^2, w_z z^2. */
void orc_adjoint_sweep(const orc_lattice *L, const orc_params *p, double t, double coef, const double *psi_,
                       const double *mu1_, double *mu2_, double sens[7])
{
    const cplx *psi = (const cplx *)psi_, *mu1 = (const cplx *)mu1_;
    cplx *mu2 = (cplx *)mu2_;
    const int V = L->V, C = p->ncomp, F = (C - 1) / 2;
    const size_t nc = ncells(p) * V;
    const double inv_s2 = 1.0 / (p->block_scale * p->block_scale);
    /* planes in parallel; each plane's sensitivities are summed in its own slot and the slots
     * in plane order afterwards, so the result does not depend on the thread count */
    double *plane = (double *)calloc((size_t)7 * p->nz, sizeof(double));
    int z, k;
#pragma omp parallel for schedule(static)
    for (z = 0; z < p->nz; ++z) {
        int x, y, v, c, f, k2;
        double *sp = plane + 7 * (size_t)z;
        for (y = 0; y < p->ny; ++y)
            for (x = 0; x < p->nx; ++x)
                for (v = 0; v < V; ++v) {
                    const size_t me = ((size_t)(z * p->ny + y) * p->nx + x) * V + v;
                    cplx ps[5], mu[5], kin[5], out[5];
                    double r[3];
                    for (c = 0; c < C; ++c) {
                        ps[c] = psi[c * nc + me];
                        mu[c] = mu1[c * nc + me];
                        kin[c] = cx(0, 0);
                        out[c] = cx(0, 0);
                    }
                    for (f = L->face_start[v]; f < L->face_start[v + 1]; ++f) {
                        const int ax = x + L->face_d[3 * f], ay = y + L->face_d[3 * f + 1], az = z + L->face_d[3 * f + 2];
                        const double w = -0.5 * L->face_hodge[f] * inv_s2;
                        int inside = ax >= 0 && ax < p->nx && ay >= 0 && ay < p->ny && az >= 0 && az < p->nz;
                        size_t nb = inside ? ((size_t)(az * p->ny + ay) * p->nx + ax) * V + L->face_tgt[f] : 0;
                        for (c = 0; c < C; ++c) {
                            cplx o = inside ? mu1[c * nc + nb] : cx(0, 0);
                            kin[c] = cadd(kin[c], cscale(w, csub(o, mu[c])));
                        }
                    }
                    orc_position(L, p, x, y, z, v, r);
                    adj_local(p, r, t, ps, mu, kin, out);
                    for (c = 0; c < C; ++c) mu2[c * nc + me] = cadd(mu2[c * nc + me], cscale(coef, out[c]));
                    /* sensitivities */
                    double im_mp = 0;
                    for (c = 0; c < C; ++c) im_mp += cmul(cconj(mu[c]), ps[c]).im;
                    if (t < p->trap_off_time)
                        for (k2 = 0; k2 < 3; ++k2) sp[4 + k2] += p->trap[k2] * r[k2] * r[k2] * im_mp;
                    if (F > 0) {
                        const double s = (F == 1) ? 1.0 : -1.0;
                        double pb[3];
                        for (k2 = 0; k2 < 3; ++k2) {
                            double e[3] = {0, 0, 0};
                            cplx fp[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
                            e[k2] = 1.0;
                            spin_dot_apply(F, e, ps, fp);
                            double im = 0;
                            for (c = 0; c < C; ++c) im += cmul(cconj(mu[c]), fp[c]).im;
                            pb[k2] = s * im;
                            sp[k2] += pb[k2];
                        }
                        sp[3] += (r[0] - p->center[0]) * pb[0] + (r[1] - p->center[1]) * pb[1] -
                                 2.0 * (r[2] - p->center[2]) * pb[2];
                    }
                }
    }
    for (k = 0; k < 7; ++k) sens[k] = 0;
    for (z = 0; z < p->nz; ++z)
        for (k = 0; k < 7; ++k) sens[k] += plane[7 * (size_t)z + k];
    free(plane);
}
