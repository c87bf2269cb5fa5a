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

template <typename R, int C, int DEG>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs a)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const int b = sweep_order(g, blockIdx.x, gridDim.x, a.order);
    const WaveCoord wc = decode(g, b, gridDim.x);
    double part[ADJ_COUNT];
#pragma unroll
    for (int i = 0; i < ADJ_COUNT; ++i) part[i] = 0.0;

    if (wc.valid) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const bool act = x < g.nx;
        const int xc = act ? x : g.nx - 1;
        const V2 *__restrict__ psi = (const V2 *)a.psi;
        const V2 *__restrict__ mu1 = (const V2 *)a.mu1;
        V2 *__restrict__ mu2 = (V2 *)a.mu2;
        const long long cstride = comp_stride(g);
        const long long own = line_index(g, 0, z, y, v);

        R pr[C], pi[C], mr[C], mi[C], qr[C], qi[C], ar[C], ai[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const V2 p = psi[own + c * cstride + xc];
            const V2 m = mu1[own + c * cstride + xc];
            const V2 q = mu2[own + c * cstride + xc];
            pr[c] = p.x;
            pi[c] = p.y;
            mr[c] = m.x;
            mi[c] = m.y;
            qr[c] = q.x;
            qi[c] = q.y;
            ar[c] = 0;
            ai[c] = 0;
        }
        const int *face = a.tb.face + v * DEG;
        const R *wt = table_w<R>(a.tb) + v * DEG;
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int pk = face[f];
            const int dx = (pk & 3) - 1, dy = ((pk >> 2) & 3) - 1, dz = ((pk >> 4) & 3) - 1, t = pk >> 8;
            int zn = z + dz, yn = y + dy;
            const bool inside = (unsigned)zn < (unsigned)g.nz && (unsigned)yn < (unsigned)g.ny;
            if (!inside) {
                zn = z;
                yn = y;
            }
            const int xn = x + dx;
            const bool xok = (unsigned)xn < (unsigned)g.nx;
            const int xnc = xn < 0 ? 0 : (xn >= g.nx ? g.nx - 1 : xn);
            const R wf = (inside && xok) ? wt[f] : (R)0;
            const long long nl = line_index(g, 0, zn, yn, t);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const V2 nb = mu1[nl + c * cstride + xnc];
                ar[c] += wf * nb.x;
                ai[c] += wf * nb.y;
            }
        }
        const R wsum = table_wsum<R>(a.tb)[v];
        R hr[C], hi[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] = ar[c] - wsum * mr[c];
            hi[c] = ai[c] - wsum * mi[c];
        }
        const Physics &ph = a.ph;
        const FieldRec &fr = a.fields[a.sweep];
        const R *pos = table_pos<R>(a.tb) + 3 * v;
        const R rx = (R)ph.p0[0] + (R)ph.cell[0] * (R)x + pos[0];
        const R ry = (R)ph.p0[1] + (R)ph.cell[1] * (R)y + pos[1];
        const R rz = (R)ph.p0[2] + (R)ph.cell[2] * (R)z + pos[2];
        const bool trap_on = fr.t < ph.trap_off_time;
        R vtrap = 0;
        if (trap_on)
            vtrap = (R)0.5 * ((R)ph.trap2[0] * rx * rx + (R)ph.trap2[1] * ry * ry + (R)ph.trap2[2] * rz * rz) +
                    (R)ph.trap_offset;
        R n = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) n += pr[c] * pr[c] + pi[c] * pi[c];
        const R c0 = (R)ph.c0, c1 = (R)ph.c1;
        // H_lin mu: diagonal density part
        const R d = vtrap + c0 * n;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] += d * mr[c];
            hi[c] += d * mi[c];
        }
        R dxr = 0, dyr = 0, dzr = 0;
        if constexpr (C > 1) {
            dxr = rx - (R)ph.center[0];
            dyr = ry - (R)ph.center[1];
            dzr = rz - (R)ph.center[2];
            const R bq = (R)fr.bq;
            const R Bx = bq * dxr + (R)fr.bb[0], By = bq * dyr + (R)fr.bb[1], Bz = (R)-2 * bq * dzr + (R)fr.bb[2];
            const R s = (C == 3) ? (R)1 : (R)-1;
            R tr[C], ti[C];
            // spin density of psi
            R Fx = 0, Fy = 0, Fz = 0;
            spin_dot<R, C>(1, 0, 0, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fx += pr[c] * tr[c] + pi[c] * ti[c];
            spin_dot<R, C>(0, 1, 0, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fy += pr[c] * tr[c] + pi[c] * ti[c];
            spin_dot<R, C>(0, 0, 1, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fz += pr[c] * tr[c] + pi[c] * ti[c];
            spin_dot<R, C>(c1 * Fx + s * Bx, c1 * Fy + s * By, c1 * Fz + s * Bz, mr, mi, tr, ti);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                hr[c] += tr[c];
                hi[c] += ti[c];
            }
            if (a.quad) {
                const R bqq = (R)fr.bqq;
                const R Qx = bqq * dxr + (R)fr.bbq[0], Qy = bqq * dyr + (R)fr.bbq[1],
                        Qz = (R)-2 * bqq * dzr + (R)fr.bbq[2];
                R ur[C], ui[C];
                spin_dot<R, C>(Qx, Qy, Qz, mr, mi, tr, ti);
                spin_dot<R, C>(Qx, Qy, Qz, tr, ti, ur, ui);
                const R qs = (R)ph.quad_sign;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    hr[c] += qs * ur[c];
                    hi[c] += qs * ui[c];
                }
            }
            constexpr int F = (C - 1) / 2;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const R mm = (R)ph.q * (R)((F - c) * (F - c));
                hr[c] += mm * mr[c];
                hi[c] += mm * mi[c];
            }
        }
        // out = i H_lin mu - 2 c0 Im(psi^dag mu) psi - ...
        R orr[C], oi[C];
        R ima = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            orr[c] = -hi[c];
            oi[c] = hr[c];
            ima += pr[c] * mi[c] - pi[c] * mr[c];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            orr[c] -= 2 * c0 * ima * pr[c];
            oi[c] -= 2 * c0 * ima * pi[c];
        }
        if constexpr (C > 1) {
            R tr[C], ti[C], Y[3];
            for (int ax = 0; ax < 3; ++ax) {
                spin_dot<R, C>(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0, mr, mi, tr, ti);
                R im = 0;
                for (int c = 0; c < C; ++c) im += pr[c] * ti[c] - pi[c] * tr[c];
                Y[ax] = im;
            }
            spin_dot<R, C>(Y[0], Y[1], Y[2], pr, pi, tr, ti);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                orr[c] -= 2 * c1 * tr[c];
                oi[c] -= 2 * c1 * ti[c];
            }
            if constexpr (C == 5) {
                const R c2 = (R)ph.c2;
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
                const R kP = (R)0.4 * c2, kQ = (R)0.2 * c2;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const R sg = (c & 1) ? (R)-1 : (R)1;
                    // u_c (u^dag mu), u_c = sg (pr[4-c], -pi[4-c])
                    const R ur = sg * pr[4 - c], ui = -sg * pi[4 - c];
                    const R pr_ = ur * udr - ui * udi, pi_ = ur * udi + ui * udr;
                    // i kP (pr_ + i pi_) = (-kP pi_, kP pr_)
                    orr[c] += -kP * pi_;
                    oi[c] += kP * pr_;
                    // -i kQ Theta sg conj(mu_{4-c})
                    const R cr = mr[4 - c], ci = -mi[4 - c];
                    const R tr2 = sg * (thr * cr - thi * ci), ti2 = sg * (thr * ci + thi * cr);
                    orr[c] += kQ * ti2;
                    oi[c] += -kQ * tr2;
                }
            }
        }
        const R coef = (R)a.coef;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            V2 o;
            o.x = qr[c] + coef * orr[c];
            o.y = qi[c] + coef * oi[c];
            if (act) mu2[own + c * cstride + xc] = o;
        }
        // parameter sensitivities: sum Im(conj(mu) * X psi)
        if (act) {
            double im_mp = 0;
            for (int c = 0; c < C; ++c) im_mp += (double)mr[c] * pi[c] - (double)mi[c] * pr[c];
            if (trap_on) {
                part[ADJ_WX] = ph.trap[0] * (double)rx * rx * im_mp;
                part[ADJ_WY] = ph.trap[1] * (double)ry * ry * im_mp;
                part[ADJ_WZ] = ph.trap[2] * (double)rz * rz * im_mp;
            }
            if constexpr (C > 1) {
                const double s = (C == 3) ? 1.0 : -1.0;
                R tr[C], ti[C];
                double pb[3];
                for (int ax = 0; ax < 3; ++ax) {
                    spin_dot<R, C>(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0, pr, pi, tr, ti);
                    double im = 0;
                    for (int c = 0; c < C; ++c) im += (double)mr[c] * ti[c] - (double)mi[c] * tr[c];
                    pb[ax] = s * im;
                }
                part[ADJ_BX] = pb[0];
                part[ADJ_BY] = pb[1];
                part[ADJ_BZ] = pb[2];
                part[ADJ_BQ] = (double)dxr * pb[0] + (double)dyr * pb[1] - 2.0 * (double)dzr * pb[2];
            }
        }
    }
    __shared__ double red[4][ADJ_COUNT];
#pragma unroll
    for (int i = 0; i < ADJ_COUNT; ++i) {
        const double r = wave_sum(part[i]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = r;
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

template <typename R, int C>
static hipError_t adj_deg(const AdjArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    switch (a.g.deg) {
    case 4: hipLaunchKernelGGL((k_adjoint<R, C, 4>), grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_adjoint<R, C, 6>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_adjoint<R, C, 8>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
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
