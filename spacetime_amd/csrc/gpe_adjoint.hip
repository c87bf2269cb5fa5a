// Discrete adjoint of the 3-level leapfrog (no counterpart in the reference).
//
// Forward (real-pair view of psi, inner product <a, b> = Re sum conj(a) b):
//     psi_1     = psi_0 + dt G(psi_0, t_0)                  G(psi) = -i H(psi)
//     psi_{k+1} = psi_{k-1} + 2 dt G(psi_k, t_k),  k >= 1
// Reverse, for an objective J(psi_N):
//     mu_N = dJ/dpsi_N,  mu_{N+1} = 0
//     mu_k = mu_{k+2} + 2 dt G'(psi_k)^T mu_{k+1}
//     dJ/dtheta = dt <mu_1, dG/dtheta(psi_0)> + sum_{k>=1} 2 dt <mu_{k+1}, dG/dtheta(psi_k)>
// with the transposed Jacobian of the local + kinetic Hamiltonian
//     G'(psi)^T mu = i H_lin(psi) mu - 2 c0 Im(psi^dag mu) psi - 2 c1 sum_a Im(psi^dag f_a mu) f_a psi
//                    [+ i (2 c2/5) u (u^dag mu) - i (c2/5) Theta S conj(mu_rev),  u = S conj(psi_rev)]
// where H_lin(psi) is H with the density and spin density frozen at psi (Hermitian, so its
// transpose in the real view is i H_lin).  Parameter derivatives of H: bias field
// s f_a psi, quadrupole gradient s (dx f_x + dy f_y - 2 dz f_z) psi, trap frequency
// w_a r_a^2 psi (s = +1 spin-1, -1 spin-2).
#include "gpe_device.h"

namespace gpe {

// Physics and field constants of one adjoint sweep in its precision (the FP32 copies made on
// the host for FP32), as the forward sweep takes them.
template <typename R>
struct AdjConsts {
    R cell[3], p0[3], trap2[3], trap_offset, center[3], c0, c1, c2, quad_sign, q;
    R bq, bb[3], bqq, bbq[3];
};

template <typename R>
__device__ __forceinline__ AdjConsts<R> adj_consts(const Physics &ph, const FieldRec &fr)
{
    AdjConsts<R> k;
    if constexpr (std::is_same<R, float>::value) {
        const PhysicsF &f = ph.f;
        for (int i = 0; i < 3; ++i) {
            k.cell[i] = f.cell[i];
            k.p0[i] = f.p0[i];
            k.trap2[i] = f.trap2[i];
            k.center[i] = f.center[i];
            k.bb[i] = fr.fbb[i];
            k.bbq[i] = fr.fbbq[i];
        }
        k.trap_offset = f.trap_offset;
        k.c0 = f.c0;
        k.c1 = f.c1;
        k.c2 = f.c2;
        k.quad_sign = f.quad_sign;
        k.q = f.q;
        k.bq = fr.fbq;
        k.bqq = fr.fbqq;
    } else {
        for (int i = 0; i < 3; ++i) {
            k.cell[i] = ph.cell[i];
            k.p0[i] = ph.p0[i];
            k.trap2[i] = ph.trap2[i];
            k.center[i] = ph.center[i];
            k.bb[i] = fr.bb[i];
            k.bbq[i] = fr.bbq[i];
        }
        k.trap_offset = ph.trap_offset;
        k.c0 = ph.c0;
        k.c1 = ph.c1;
        k.c2 = ph.c2;
        k.quad_sign = ph.quad_sign;
        k.q = ph.q;
        k.bq = fr.bq;
        k.bqq = fr.bqq;
    }
    return k;
}

// G'(psi)^T mu at one unknown, given the kinetic part of H_lin mu in (hr, hi) (the stencil on mu,
// already weighted): the local part of H_lin added, then
//   out = i H_lin mu - 2 c0 Im(psi^dag mu) psi - 2 c1 sum_a Im(psi^dag f_a mu) f_a psi
//         [+ i (2 c2 / 5) u (u^dag mu) - i (c2 / 5) Theta S conj(mu_rev)]
template <typename R, int C>
__device__ __forceinline__ void adj_point(const R *pr, const R *pi, const R *mr, const R *mi, R *hr, R *hi, R rx,
                                          R ry, R rz, bool trap_on, int quad, const AdjConsts<R> &k, R *orr, R *oi)
{
    R vtrap = 0;
    if (trap_on)
        vtrap = (R)0.5 * (k.trap2[0] * rx * rx + k.trap2[1] * ry * ry + k.trap2[2] * rz * rz) + k.trap_offset;
    R n = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) n += pr[c] * pr[c] + pi[c] * pi[c];
    const R d = vtrap + k.c0 * n;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        hr[c] += d * mr[c];
        hi[c] += d * mi[c];
    }
    if constexpr (C > 1) {
        const R dxr = rx - k.center[0], dyr = ry - k.center[1], dzr = rz - k.center[2];
        const R Bx = k.bq * dxr + k.bb[0], By = k.bq * dyr + k.bb[1], Bz = (R)-2 * k.bq * dzr + k.bb[2];
        const R s = (C == 3) ? (R)1 : (R)-1;
        R tr[C], ti[C];
        R Fx = 0, Fy = 0, Fz = 0;  // spin density of psi
        spin_dot<R, C>(1, 0, 0, pr, pi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) Fx += pr[c] * tr[c] + pi[c] * ti[c];
        spin_dot<R, C>(0, 1, 0, pr, pi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) Fy += pr[c] * tr[c] + pi[c] * ti[c];
        spin_dot<R, C>(0, 0, 1, pr, pi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) Fz += pr[c] * tr[c] + pi[c] * ti[c];
        spin_dot<R, C>(k.c1 * Fx + s * Bx, k.c1 * Fy + s * By, k.c1 * Fz + s * Bz, mr, mi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] += tr[c];
            hi[c] += ti[c];
        }
        if (quad) {
            const R Qx = k.bqq * dxr + k.bbq[0], Qy = k.bqq * dyr + k.bbq[1], Qz = (R)-2 * k.bqq * dzr + k.bbq[2];
            R ur[C], ui[C];
            spin_dot<R, C>(Qx, Qy, Qz, mr, mi, tr, ti);
            spin_dot<R, C>(Qx, Qy, Qz, tr, ti, ur, ui);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                hr[c] += k.quad_sign * ur[c];
                hi[c] += k.quad_sign * ui[c];
            }
        }
        constexpr int F = (C - 1) / 2;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const R mm = k.q * (R)((F - c) * (F - c));
            hr[c] += mm * mr[c];
            hi[c] += mm * mi[c];
        }
    }
    R ima = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        orr[c] = -hi[c];
        oi[c] = hr[c];
        ima += pr[c] * mi[c] - pi[c] * mr[c];
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        orr[c] -= 2 * k.c0 * ima * pr[c];
        oi[c] -= 2 * k.c0 * ima * pi[c];
    }
    if constexpr (C > 1) {
        R tr[C], ti[C], Y[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            spin_dot<R, C>(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0, mr, mi, tr, ti);
            R im = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) im += pr[c] * ti[c] - pi[c] * tr[c];
            Y[ax] = im;
        }
        spin_dot<R, C>(Y[0], Y[1], Y[2], pr, pi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            orr[c] -= 2 * k.c1 * tr[c];
            oi[c] -= 2 * k.c1 * ti[c];
        }
        if constexpr (C == 5) {
            // u_c = S_c conj(psi_{4-c}); u^dag mu = sum S_c psi_{4-c} mu_c; Theta = sum S_c psi_c psi_{4-c}
            R udr = 0, udi = 0, thr = 0, thi = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const R sg = (c & 1) ? (R)-1 : (R)1;
                udr += sg * (pr[4 - c] * mr[c] - pi[4 - c] * mi[c]);
                udi += sg * (pr[4 - c] * mi[c] + pi[4 - c] * mr[c]);
                thr += sg * (pr[c] * pr[4 - c] - pi[c] * pi[4 - c]);
                thi += sg * (pr[c] * pi[4 - c] + pi[c] * pr[4 - c]);
            }
            const R kP = (R)0.4 * k.c2, kQ = (R)0.2 * k.c2;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const R sg = (c & 1) ? (R)-1 : (R)1;
                // u_c (u^dag mu), u_c = sg (pr[4-c], -pi[4-c]); i kP (p + i q) = (-kP q, kP p)
                const R ur = sg * pr[4 - c], ui = -sg * pi[4 - c];
                const R p_ = ur * udr - ui * udi, q_ = ur * udi + ui * udr;
                orr[c] += -kP * q_;
                oi[c] += kP * p_;
                // -i kQ Theta sg conj(mu_{4-c})
                const R cr = mr[4 - c], ci = -mi[4 - c];
                const R tr2 = sg * (thr * cr - thi * ci), ti2 = sg * (thr * ci + thi * cr);
                orr[c] += kQ * ti2;
                oi[c] += -kQ * tr2;
            }
        }
    }
}

// Parameter sensitivities of one unknown, in FP64 from the stored values and FP64 positions:
//   sum Im(conj(mu) X psi) for X = s f_x, s f_y, s f_z (bias field), s (dx f_x + dy f_y - 2 dz f_z)
//   (quadrupole gradient), w_a r_a^2 (trap frequencies); s = +1 spin-1, -1 spin-2.
template <typename R, int C>
__device__ __forceinline__ void adj_sens(const R *pr, const R *pi, const R *mr, const R *mi, const Physics &ph,
                                         const double *pos64, int x, int y, int z, bool trap_on, double *part)
{
    const double rx = ph.p0[0] + ph.cell[0] * x + pos64[0];
    const double ry = ph.p0[1] + ph.cell[1] * y + pos64[1];
    const double rz = ph.p0[2] + ph.cell[2] * z + pos64[2];
    double im_mp = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) im_mp += (double)mr[c] * pi[c] - (double)mi[c] * pr[c];
    if (trap_on) {
        part[ADJ_WX] += ph.trap[0] * rx * rx * im_mp;
        part[ADJ_WY] += ph.trap[1] * ry * ry * im_mp;
        part[ADJ_WZ] += ph.trap[2] * rz * rz * im_mp;
    }
    if constexpr (C > 1) {
        const double s = (C == 3) ? 1.0 : -1.0;
        R tr[C], ti[C];
        double pb[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            spin_dot<R, C>(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0, pr, pi, tr, ti);
            double im = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) im += (double)mr[c] * ti[c] - (double)mi[c] * tr[c];
            pb[ax] = s * im;
        }
        part[ADJ_BX] += pb[0];
        part[ADJ_BY] += pb[1];
        part[ADJ_BZ] += pb[2];
        part[ADJ_BQ] += (rx - ph.center[0]) * pb[0] + (ry - ph.center[1]) * pb[1] - 2.0 * (rz - ph.center[2]) * pb[2];
    }
}

// The same sensitivities for the FP32 pair body, formed in FP32 per unknown and summed over the
// lane's two x before they join the FP64 partials (one conversion per parameter and lane instead
// of one per product); the trap terms leave out the frequency factor omega_a, applied in FP64 when
// the lane's sums are added (they are linear in it).
template <int C>
__device__ __forceinline__ void adj_sens32(const float *pr, const float *pi, const float *mr, const float *mi,
                                           float rx, float ry, float rz, const AdjConsts<float> &k, bool trap_on,
                                           float *s)
{
    float im_mp = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) im_mp += mr[c] * pi[c] - mi[c] * pr[c];
    if (trap_on) {
        s[ADJ_WX] += rx * rx * im_mp;
        s[ADJ_WY] += ry * ry * im_mp;
        s[ADJ_WZ] += rz * rz * im_mp;
    }
    if constexpr (C > 1) {
        const float sg = (C == 3) ? 1.f : -1.f;
        float tr[C], ti[C], pb[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            spin_dot<float, C>(ax == 0 ? 1.f : 0.f, ax == 1 ? 1.f : 0.f, ax == 2 ? 1.f : 0.f, pr, pi, tr, ti);
            float im = 0.f;
#pragma unroll
            for (int c = 0; c < C; ++c) im += mr[c] * ti[c] - mi[c] * tr[c];
            pb[ax] = sg * im;
        }
        s[ADJ_BX] += pb[0];
        s[ADJ_BY] += pb[1];
        s[ADJ_BZ] += pb[2];
        s[ADJ_BQ] += (rx - k.center[0]) * pb[0] + (ry - k.center[1]) * pb[1] - 2.f * (rz - k.center[2]) * pb[2];
    }
}

// One x value per lane (FP64, and FP32 where the pair form is not chosen): the forward sweep's
// single body on mu_{k+1} -- interior fast path through the per-face line offsets, wave-uniform
// weights on same-x faces, zero-weight pad faces skipped.
template <typename R, int C, int DEG>
__device__ __forceinline__ void adjoint_single(const AdjArgs &a, int b, int nb, double *part)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const WaveCoord wc = decode(g, b, nb);
    if (!wc.valid) return;
    const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
    const bool act = x < g.nx;
    const int xc = act ? x : g.nx - 1;
    const V2 *__restrict__ psi = (const V2 *)a.psi;
    const V2 *__restrict__ mu1 = (const V2 *)a.mu1;
    V2 *__restrict__ mu2 = (V2 *)a.mu2;
    const long long cstride = comp_stride(g);
    const long long own_line = line_index(g, 0, z, y, v);

    R pr[C], pi[C], mr[C], mi[C], qr[C], qi[C], ar[C], ai[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const V2 p = psi[own_line + c * cstride + xc];
        const V2 m = mu1[own_line + c * cstride + xc];
        const V2 q = nt_load<R>(&mu2[own_line + c * cstride + xc]);
        pr[c] = p.x;
        pi[c] = p.y;
        mr[c] = m.x;
        mi[c] = m.y;
        qr[c] = q.x;
        qi[c] = q.y;
        ar[c] = 0;
        ai[c] = 0;
    }
    const int2 *fx = (const int2 *)a.tb.facex + v * DEG;
    int face[DEG], fd[DEG];
#pragma unroll
    for (int f = 0; f < DEG; ++f) {
        const int2 e = fx[f];
        face[f] = e.x;
        fd[f] = e.y;
    }
    const R *wt = table_w<R>(a.tb) + v * DEG;
    R wu[DEG];
    bool xok[DEG];
    int dxf[DEG], noff[DEG];
    const V2 *nbase[DEG];
#pragma unroll
    for (int f = 0; f < DEG; ++f) {
        const int dx = (face[f] & 3) - 1;
        const int xn = xc + dx;
        xok[f] = (unsigned)xn < (unsigned)g.nx;
        dxf[f] = dx;
        noff[f] = xn < 0 ? 0 : (xn >= g.nx ? g.nx - 1 : xn);
    }
    if (z >= 1 && z <= g.nz - 2 && y >= 1 && y <= g.ny - 2) {
        const V2 *own = mu1 + own_line;
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            wu[f] = wt[f];
            nbase[f] = own + fd[f];
        }
    } else {
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int pk = face[f];
            const int dy = ((pk >> 2) & 3) - 1, dz = ((pk >> 4) & 3) - 1, t = pk >> 8;
            int zn = z + dz, yn = y + dy;
            const bool inside = (unsigned)zn < (unsigned)g.nz && (unsigned)yn < (unsigned)g.ny;
            if (!inside) {
                zn = z;
                yn = y;
            }
            wu[f] = inside ? wt[f] : (R)0;
            nbase[f] = mu1 + line_index(g, 0, zn, yn, t);
        }
    }
    const int nf = DEG == 8 ? a.tb.nface[v] : DEG;
    V2 nbv[DEG][C];
#pragma unroll
    for (int f = 0; f < DEG; ++f)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (f < nf) {
                nbv[f][c] = nbase[f][c * cstride + noff[f]];
            } else {
                nbv[f][c].x = 0;
                nbv[f][c].y = 0;
            }
        }
#pragma unroll
    for (int f = 0; f < DEG; ++f) {
        const R w = (dxf[f] == 0 || xok[f]) ? wu[f] : (R)0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            ar[c] += w * nbv[f][c].x;
            ai[c] += w * nbv[f][c].y;
        }
    }
    const R wsum = table_wsum<R>(a.tb)[v];
    R hr[C], hi[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        hr[c] = ar[c] - wsum * mr[c];
        hi[c] = ai[c] - wsum * mi[c];
    }
    const FieldRec &fr = a.fields[a.sweep];
    const AdjConsts<R> k = adj_consts<R>(a.ph, fr);
    const bool trap_on = fr.t < a.ph.trap_off_time;
    const R *pos = table_pos<R>(a.tb) + 3 * v;
    const R rx = k.p0[0] + (k.cell[0] * (R)x + pos[0]);
    const R ry = k.p0[1] + (k.cell[1] * (R)y + pos[1]);
    const R rz = k.p0[2] + (k.cell[2] * (R)z + pos[2]);
    R orr[C], oi[C];
    adj_point<R, C>(pr, pi, mr, mi, hr, hi, rx, ry, rz, trap_on, a.quad, k, orr, oi);
    const R coef = (R)a.coef;
    if (act) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            V2 o;
            o.x = qr[c] + coef * orr[c];
            o.y = qi[c] + coef * oi[c];
            nt_store<R>(o, &mu2[own_line + c * cstride + xc]);
        }
        adj_sens<R, C>(pr, pi, mr, mi, a.ph, a.tb.pos64 + 3 * v, x, y, z, trap_on, part);
    }
}

typedef float f4a __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(8)));

// Two consecutive x values per lane (FP32), optionally spread over a group of y rows: the
// forward sweep's pair body (gpe_kernels.hip, sweep_body_pair) on mu_{k+1}, with psi_k and
// mu_{k+2} read as own pairs, mu_{k+2} only after G'^T mu is formed (136 -> 110 VGPRs for spin-1:
// 4 waves per SIMD instead of 3; the forward sweep, whose psi_{k-1} could be read as late, ran
// 2-20% slower that way and keeps its early load).  A neighbour pair starts at x0 + dx, clamped
// into the line and shifted back; every neighbour value outside the box is replaced by zero with
// a select, so line padding is never used as a value.  The lone last x of an odd-length line
// writes its partner slot, which is line padding.
template <int C, int DEG, bool ROWS>
__device__ __forceinline__ void adjoint_pair(const AdjArgs &a, int b, int nb, double *part)
{
    const Geometry &g = a.g;
    int x0, v, y0, z, r = 0;
    bool valid;
    if constexpr (ROWS) {
        const RowCoord rc = decode_rows(g, b, nb);
        x0 = rc.x, v = rc.v, y0 = rc.y0, z = rc.z, r = rc.r, valid = rc.valid;
    } else {
        const WaveCoord wc = decode(g, b, nb);
        x0 = wc.x, v = wc.v, y0 = wc.y, z = wc.z, valid = wc.valid;
    }
    if (!valid) return;
    bool rowok = true;
    int y = y0, roff = 0;
    const int rowstride = (int)g.ys;
    int ylo = y0, yhi = y0;  // the rows of the wave's lanes (a lane's row grows with its lane number)
    if constexpr (ROWS) {
        const int nrow = g.ny - y0 < g.rows ? g.ny - y0 : g.rows;
        rowok = r < nrow;
        const int rr = rowok ? r : nrow - 1;
        y = y0 + rr;
        roff = rr * rowstride;
        ylo = y0 + __builtin_amdgcn_readfirstlane(rr);
        yhi = y0 + __builtin_amdgcn_readlane(rr, 63);
    }
    const bool act0 = rowok && x0 < g.nx, act1 = rowok && x0 + 1 < g.nx;
    const int xc = x0 < g.nx ? x0 : ((g.nx - 1) & ~1);
    const int lo = roff + xc;
    const float2 *__restrict__ psi = (const float2 *)a.psi;
    const float2 *__restrict__ mu1 = (const float2 *)a.mu1;
    float2 *__restrict__ mu2 = (float2 *)a.mu2;
    const long long cstride = comp_stride(g);
    const long long own_line = line_index(g, 0, z, y0, v);
    const int lxm2 = (int)g.lx - 2;

    float pr[2][C], pi[2][C], mr[2][C], mi[2][C], ar[2][C], ai[2][C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const f4a p = *(const f4a *)(psi + own_line + c * cstride + lo);
        const f4a m = *(const f4a *)(mu1 + own_line + c * cstride + lo);
        pr[0][c] = p.x;
        pi[0][c] = p.y;
        pr[1][c] = p.z;
        pi[1][c] = p.w;
        mr[0][c] = m.x;
        mi[0][c] = m.y;
        mr[1][c] = m.z;
        mi[1][c] = m.w;
        ar[0][c] = ar[1][c] = 0.f;
        ai[0][c] = ai[1][c] = 0.f;
    }
    const int2 *fx = (const int2 *)a.tb.facex + v * DEG;
    int face[DEG], fd[DEG];
#pragma unroll
    for (int f = 0; f < DEG; ++f) {
        const int2 e = fx[f];
        face[f] = e.x;
        fd[f] = e.y;
    }
    const float *wt = a.tb.w32 + v * DEG;
    const float2 *nbase[DEG];
    int sh[DEG], dxf[DEG], scf[DEG];
    float wf[DEG];
    bool ok0[DEG], ok1[DEG];
#pragma unroll
    for (int f = 0; f < DEG; ++f) {
        const int dx = (face[f] & 3) - 1;
        dxf[f] = dx;
        const int s0 = xc + dx;
        const int sc = s0 < 0 ? 0 : (s0 > lxm2 ? lxm2 : s0);
        scf[f] = sc;
        sh[f] = s0 - sc;
        ok0[f] = (unsigned)(x0 + dx) < (unsigned)g.nx;
        ok1[f] = (unsigned)(x0 + 1 + dx) < (unsigned)g.nx;
    }
    if (z >= 1 && z <= g.nz - 2 && ylo >= 1 && yhi <= g.ny - 2) {
        const float2 *own = mu1 + own_line;
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            wf[f] = wt[f];
            nbase[f] = own + (fd[f] + roff + scf[f]);  // one 32-bit lane offset
        }
    } else {
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int pk = face[f];
            const int dy = ((pk >> 2) & 3) - 1, dz = ((pk >> 4) & 3) - 1, t = pk >> 8;
            int zn = z + dz, yn = y + dy;
            const bool inside = (unsigned)zn < (unsigned)g.nz && (unsigned)yn < (unsigned)g.ny;
            wf[f] = inside ? wt[f] : 0.f;
            if constexpr (ROWS) {
                if ((unsigned)zn >= (unsigned)g.nz) zn = z;
                nbase[f] = mu1 + line_index(g, 0, zn, y0, t) + (roff + (inside ? dy * rowstride : 0) + scf[f]);
            } else {
                if (!inside) {
                    zn = z;
                    yn = y;
                }
                nbase[f] = mu1 + line_index(g, 0, zn, yn, t) + scf[f];
            }
        }
    }
    const int nf = DEG == 8 ? a.tb.nface[v] : DEG;
    constexpr int FB = (DEG == 8 && C == 5 && !ROWS) ? 4 : DEG;
#pragma unroll
    for (int f0 = 0; f0 < DEG; f0 += FB) {
        if (f0 > 0 && f0 >= nf) break;
        f4u nbv[FB][C];
#pragma unroll
        for (int kk = 0; kk < FB; ++kk)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (f0 + kk < nf)
                    nbv[kk][c] = *(const f4u *)(nbase[f0 + kk] + c * cstride);
                else
                    nbv[kk][c] = (f4u)(0.f);
            }
#pragma unroll
        for (int kk = 0; kk < FB; ++kk) {
            const int f = f0 + kk;
            const float w = wf[f];
            if (dxf[f] == 0) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const f4u e = nbv[kk][c];
                    ar[0][c] += w * e.x;
                    ai[0][c] += w * e.y;
                    ar[1][c] += w * e.z;
                    ai[1][c] += w * e.w;
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const f4u e = nbv[kk][c];
                    float n0r = sh[f] == 0 ? e.x : e.z, n0i = sh[f] == 0 ? e.y : e.w;
                    float n1r = sh[f] == 0 ? e.z : e.x, n1i = sh[f] == 0 ? e.w : e.y;
                    n0r = (ok0[f] && sh[f] >= 0) ? n0r : 0.f;
                    n0i = (ok0[f] && sh[f] >= 0) ? n0i : 0.f;
                    n1r = (ok1[f] && sh[f] <= 0) ? n1r : 0.f;
                    n1i = (ok1[f] && sh[f] <= 0) ? n1i : 0.f;
                    ar[0][c] += w * n0r;
                    ai[0][c] += w * n0i;
                    ar[1][c] += w * n1r;
                    ai[1][c] += w * n1i;
                }
            }
        }
    }
    const float wsum = a.tb.wsum32[v];
    const FieldRec &fr = a.fields[a.sweep];
    const AdjConsts<float> k = adj_consts<float>(a.ph, fr);
    const bool trap_on = fr.t < a.ph.trap_off_time;
    const float *pos = a.tb.pos32 + 3 * v;
    const float ry = k.p0[1] + (k.cell[1] * (float)y + pos[1]);
    const float rz = k.p0[2] + (k.cell[2] * (float)z + pos[2]);
    const float coef = (float)a.coef;
    float2 out[2][C];
    float s32[ADJ_COUNT];
#pragma unroll
    for (int i = 0; i < ADJ_COUNT; ++i) s32[i] = 0.f;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        float hr[C], hi[C], orr[C], oi[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] = ar[e][c] - wsum * mr[e][c];
            hi[c] = ai[e][c] - wsum * mi[e][c];
        }
        const float rx = k.p0[0] + (k.cell[0] * (float)(x0 + e) + pos[0]);
        adj_point<float, C>(pr[e], pi[e], mr[e], mi[e], hr, hi, rx, ry, rz, trap_on, a.quad, k, orr, oi);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            out[e][c].x = coef * orr[c];
            out[e][c].y = coef * oi[c];
        }
        if (e == 0 ? act0 : act1)
            adj_sens32<C>(pr[e], pi[e], mr[e], mi[e], rx, ry, rz, k, trap_on, s32);
    }
    part[ADJ_BX] += (double)s32[ADJ_BX];
    part[ADJ_BY] += (double)s32[ADJ_BY];
    part[ADJ_BZ] += (double)s32[ADJ_BZ];
    part[ADJ_BQ] += (double)s32[ADJ_BQ];
    part[ADJ_WX] += a.ph.trap[0] * (double)s32[ADJ_WX];
    part[ADJ_WY] += a.ph.trap[1] * (double)s32[ADJ_WY];
    part[ADJ_WZ] += a.ph.trap[2] * (double)s32[ADJ_WZ];
    // mu_{k+2} is read only now: its registers are free while G'^T mu is formed
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const f4a q = __builtin_nontemporal_load((const f4a *)(mu2 + own_line + c * cstride + lo));
        out[0][c].x += q.x;
        out[0][c].y += q.y;
        out[1][c].x += q.z;
        out[1][c].y += q.w;
    }
    if (act0) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            f4a o;
            o.x = out[0][c].x;
            o.y = out[0][c].y;
            o.z = out[1][c].x;
            o.w = out[1][c].y;
            __builtin_nontemporal_store(o, (f4a *)(mu2 + own_line + c * cstride + lo));
        }
    }
}

// Adjoint sweep mu_{k+2} -> mu_k = mu_{k+2} + coef G'(psi_k)^T mu_{k+1} over the sweep traversal
// (X: the lane form of gpe_internal.h), with the FP64 parameter sensitivities of the workgroup's
// unknowns summed into partials[b].
template <typename R, int C, int DEG, int X>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs a)
{
    const int b = sweep_order(a.g, blockIdx.x, gridDim.x, a.order);
    double part[ADJ_COUNT];
#pragma unroll
    for (int i = 0; i < ADJ_COUNT; ++i) part[i] = 0.0;
    if constexpr (X == 1)
        adjoint_single<R, C, DEG>(a, b, gridDim.x, part);
    else
        adjoint_pair<C, DEG, X == 3>(a, b, gridDim.x, part);
    __shared__ double red[4][ADJ_COUNT];
    wave_sum_n<ADJ_COUNT>(part);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int i = 0; i < ADJ_COUNT; ++i) red[threadIdx.x >> 6][i] = part[i];
    }
    __syncthreads();
    if (threadIdx.x < ADJ_COUNT) {
        const int i = threadIdx.x;
        a.partials[(long long)b * ADJ_COUNT + i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
    }
}

// Objective support: overlap partials (Re, Im of sum dv conj(T) psi) and the gradient
// dJ/dpsi_N in the real-pair sense, written into g.
//   OBJ_OVERLAP: g = 2 c dv T with c = <T|psi> (device scalar pair)
//   OBJ_ENERGY:  g = 2 dv g (g holds H psi on entry)
//   OBJ_MAGZ:    g = 2 dv f_z psi
//   OBJ_EXPLICIT: partials of sum dv r_a^2 n (a = x, y, z) and s sum dv (dx Fx + dy Fy - 2 dz Fz),
//                 the explicit parameter dependence of the energy functional
template <typename R, int C>
__global__ __launch_bounds__(256) void k_objective(ObjArgs a)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const WaveCoord wc = decode(g);
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    if (a.mode == OBJ_EXPLICIT && wc.valid && wc.x < g.nx) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const long long cstride = comp_stride(g);
        const long long own = line_index(g, 0, z, y, v) + x;
        const double dv = a.tb.dv[v];
        const V2 *psi = (const V2 *)a.psi;
        double pr[C], pi[C], n = 0;
        for (int c = 0; c < C; ++c) {
            const V2 p = psi[own + c * cstride];
            pr[c] = p.x;
            pi[c] = p.y;
            n += pr[c] * pr[c] + pi[c] * pi[c];
        }
        const double *pos = a.tb.pos64 + 3 * v;
        const double rx = a.ph.p0[0] + a.ph.cell[0] * x + pos[0];
        const double ry = a.ph.p0[1] + a.ph.cell[1] * y + pos[1];
        const double rz = a.ph.p0[2] + a.ph.cell[2] * z + pos[2];
        p0 = dv * rx * rx * n;
        p1 = dv * ry * ry * n;
        p2 = dv * rz * rz * n;
        if constexpr (C > 1) {
            double tr[C], ti[C], F[3] = {0, 0, 0};
            for (int ax = 0; ax < 3; ++ax) {
                spin_dot<double, C>(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0, pr, pi, tr, ti);
                for (int c = 0; c < C; ++c) F[ax] += pr[c] * tr[c] + pi[c] * ti[c];
            }
            const double s = (C == 3) ? 1.0 : -1.0;
            p3 = s * dv *
                 ((rx - a.ph.center[0]) * F[0] + (ry - a.ph.center[1]) * F[1] - 2.0 * (rz - a.ph.center[2]) * F[2]);
        }
    }
    if (a.mode != OBJ_EXPLICIT && wc.valid && wc.x < g.nx) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const long long cstride = comp_stride(g);
        const long long own = line_index(g, 0, z, y, v) + x;
        const double dv = a.tb.dv[v];
        const V2 *psi = (const V2 *)a.psi;
        const V2 *tgt = (const V2 *)a.target;
        V2 *gr = (V2 *)a.grad;
        constexpr int F = (C - 1) / 2;
        for (int c = 0; c < C; ++c) {
            const long long i = own + c * cstride;
            if (a.mode == OBJ_PARTIALS) {
                const V2 t = tgt[i], p = psi[i];
                p0 += dv * ((double)t.x * p.x + (double)t.y * p.y);
                p1 += dv * ((double)t.x * p.y - (double)t.y * p.x);
            } else if (a.mode == OBJ_OVERLAP) {
                const V2 t = tgt[i];
                const double cr = a.scalars[0], ci = a.scalars[1];
                V2 o;
                o.x = (R)(2 * dv * (cr * t.x - ci * t.y));
                o.y = (R)(2 * dv * (cr * t.y + ci * t.x));
                gr[i] = o;
            } else if (a.mode == OBJ_ENERGY) {
                V2 o = gr[i];
                o.x = (R)(2 * dv * o.x);
                o.y = (R)(2 * dv * o.y);
                gr[i] = o;
            } else {
                const V2 p = psi[i];
                const double m = (double)(F - c);
                V2 o;
                o.x = (R)(2 * dv * m * p.x);
                o.y = (R)(2 * dv * m * p.y);
                gr[i] = o;
            }
        }
    }
    if (a.mode == OBJ_PARTIALS || a.mode == OBJ_EXPLICIT) {
        const int width = a.mode == OBJ_PARTIALS ? 2 : 4;
        __shared__ double red[4][4];
        p0 = wave_sum(p0);
        p1 = wave_sum(p1);
        p2 = wave_sum(p2);
        p3 = wave_sum(p3);
        if ((threadIdx.x & 63) == 0) {
            red[threadIdx.x >> 6][0] = p0;
            red[threadIdx.x >> 6][1] = p1;
            red[threadIdx.x >> 6][2] = p2;
            red[threadIdx.x >> 6][3] = p3;
        }
        __syncthreads();
        if ((int)threadIdx.x < width)
            a.partials[(long long)blockIdx.x * width + threadIdx.x] =
                (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    }
}

template <typename R, int C, int DEG>
static hipError_t adj_form(const AdjArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    const int form = lane_form(a.g.xpl, a.g.rows);
    if (form == 1) {
        hipLaunchKernelGGL((k_adjoint<R, C, DEG, 1>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    if constexpr (pair_lanes_ok(std::is_same<R, double>::value, DEG, C)) {
        if (form == 2) hipLaunchKernelGGL((k_adjoint<R, C, DEG, 2>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_adjoint<R, C, DEG, 3>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

template <typename R, int C>
static hipError_t adj_deg(const AdjArgs &a, hipStream_t s)
{
    switch (a.g.deg) {
    case 4: return adj_form<R, C, 4>(a, s);
    case 6: return adj_form<R, C, 6>(a, s);
    case 8: return adj_form<R, C, 8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R>
static hipError_t adj_comp(int ncomp, const AdjArgs &a, hipStream_t s)
{
    switch (ncomp) {
    case 1: return adj_deg<R, 1>(a, s);
    case 3: return adj_deg<R, 3>(a, s);
    case 5: return adj_deg<R, 5>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint(int fp64, int ncomp, const AdjArgs &a, hipStream_t s)
{
    return fp64 ? adj_comp<double>(ncomp, a, s) : adj_comp<float>(ncomp, a, s);
}

template <typename R>
static hipError_t obj_comp(int ncomp, const ObjArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    switch (ncomp) {
    case 1: hipLaunchKernelGGL((k_objective<R, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_objective<R, 3>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_objective<R, 5>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_objective(int fp64, int ncomp, const ObjArgs &a, hipStream_t s)
{
    return fp64 ? obj_comp<double>(ncomp, a, s) : obj_comp<float>(ncomp, a, s);
}

}  // namespace gpe
