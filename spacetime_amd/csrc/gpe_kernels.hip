// HIP kernels of the DEC Gross-Pitaevskii step for CDNA4 (gfx950).
//
// One sweep = one pass over every unknown of a time level:
//     next += coef * (-i) H(cur)           (3-level leapfrog, Spin1GpeBCC.cu:793-917;
//                                           staggered half-step, GPBlocks.cpp:560-606)
// with H = -1/2 L + V(r) + c0 n [-i alpha n^2] + (c1 F +- B).f + singlet + quadratic Zeeman.
//
// Work mapping: a wave owns 64 consecutive x of one (z, y, unknown v) line of the device
// layout (gpe_internal.h), so the face table row, the neighbour line (z+dz, y+dy, target)
// and the field are wave-uniform (scalar loads), and each neighbour access is one
// contiguous 64-lane load shifted by dx.  Out-of-box neighbours are folded into the face
// weight (weight 0 and a clamped in-bounds address), so the sweep has no branches.
// Workgroups are remapped so that each XCD walks a contiguous slab of the traversal
// order (y-tiles, then z, then rows of the tile): the z +- 1 and y +- 1 rows a wave needs
// were fetched into the same XCD's L2 by its neighbours shortly before.
#include "gpe_device.h"

namespace gpe {

// ------------------------------------------------------------------------- the sweep --

// Physics and field constants of one sweep in the sweep's precision: the FP32 copies made on
// the host for FP32 (no per-wave conversions), the FP64 values for FP64.
template <typename R>
struct LocalConsts {
    R cell[3], p0[3], trap2[3], trap_offset, center[3], c0, c1, c2, alpha, quad_sign, q;
    R bq, bb[3], bqq, bbq[3], sigma, ksc;
};

template <typename R>
__device__ __forceinline__ LocalConsts<R> local_consts(const Physics &ph, const FieldRec &fr)
{
    LocalConsts<R> k;
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
        k.alpha = f.alpha;
        k.quad_sign = f.quad_sign;
        k.q = f.q;
        k.bq = fr.fbq;
        k.bqq = fr.fbqq;
        k.sigma = fr.fsigma;
        k.ksc = fr.fksc;
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
        k.alpha = ph.alpha;
        k.quad_sign = ph.quad_sign;
        k.q = ph.q;
        k.bq = fr.bq;
        k.bqq = fr.bqq;
        k.sigma = fr.sigma;
        k.ksc = 1.0 / (fr.sigma * fr.sigma);
    }
    return k;
}

// One x value per lane (FP64, and FP32 where the neighbour pairs would not fit).
template <typename R, int C, int DEG, int MODE>
__device__ __forceinline__ void sweep_body_single(const SweepArgs &a, int sweep, int b, int nb)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const WaveCoord wc = decode(g, b, nb);
    double part = 0.0;

    if (wc.valid) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const bool act = x < g.nx;
        const int xc = act ? x : g.nx - 1;
        const V2 *__restrict__ cur = (const V2 *)a.cur;
        V2 *__restrict__ nxt = (V2 *)a.next;
        const V2 *__restrict__ prv = (const V2 *)(a.prev ? a.prev : a.next);
        const long long cstride = comp_stride(g);  // elements between components of a line
        const long long own_line = line_index(g, 0, z, y, v);

        // issue every load of the wave before any use: own values, psi_{k-1}, and the
        // DEG neighbour lines (wave-uniform line bases, per-lane 32-bit offsets)
        R pr[C], pi[C], qr[C], qi[C], ar[C], ai[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const V2 o = cur[own_line + c * cstride + xc];
            pr[c] = o.x;
            pi[c] = o.y;
            ar[c] = 0;
            ai[c] = 0;
        }
        if constexpr (MODE == MODE_UPDATE) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const V2 o = nt_load<R>(&prv[own_line + c * cstride + xc]);
                qr[c] = o.x;
                qi[c] = o.y;
            }
        }
        const int2 *fx = (const int2 *)a.tb.facex + v * DEG;  // (packed face, line offset) pairs
        int face[DEG], fd[DEG];
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int2 e = fx[f];
            face[f] = e.x;
            fd[f] = e.y;
        }
        const R *wt = table_w<R>(a.tb) + v * DEG;
        R wu[DEG];    // wave-uniform weight: 0 when the neighbour's y or z is outside the box
        bool xok[DEG];
        int dxf[DEG];
        const V2 *nbase[DEG];
        int noff[DEG];
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int dx = (face[f] & 3) - 1;
            const int xn = xc + dx;
            xok[f] = (unsigned)xn < (unsigned)g.nx;
            dxf[f] = dx;
            noff[f] = xn < 0 ? 0 : (xn >= g.nx ? g.nx - 1 : xn);
        }
        if (z >= 1 && z <= g.nz - 2 && y >= 1 && y <= g.ny - 2) {
            // every y / z neighbour inside the box: own line plus the table's per-face offsets
            const V2 *own = cur + own_line;
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
                nbase[f] = cur + line_index(g, 0, zn, yn, t);
            }
        }
        // degree-8 tables pad some unknowns with zero-weight faces: skip their loads
        const int nf = DEG == 8 ? a.tb.nface[v] : DEG;
        V2 nb[DEG][C];
#pragma unroll
        for (int f = 0; f < DEG; ++f)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (f < nf) {
                    nb[f][c] = nbase[f][c * cstride + noff[f]];
                } else {
                    nb[f][c].x = 0;
                    nb[f][c].y = 0;
                }
            }
        // stencil: sum_f w_f psi_t(f); out-of-box neighbours carry weight 0.  Same-x faces need
        // no per-lane test (a live lane's own x is inside the box): wave-uniform weight.
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const R w = (dxf[f] == 0 || xok[f]) ? wu[f] : (R)0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                ar[c] += w * nb[f][c].x;
                ai[c] += w * nb[f][c].y;
            }
        }
        const R wsum = table_wsum<R>(a.tb)[v];
        const Physics &ph = a.ph;
        const FieldRec &fr = a.fields[sweep];
        const LocalConsts<R> k = local_consts<R>(ph, fr);
        R hr[C], hi[C];
        R sc = (R)1;
        if constexpr (MODE == MODE_ITP) sc = (R)(*a.scale_in);
        const R sig = k.sigma;        // expanding grid: lengths x sigma
        const R ksc = sc * k.ksc;     // Laplacian / sigma^2
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] = ksc * (ar[c] - wsum * pr[c]);
            hi[c] = ksc * (ai[c] - wsum * pi[c]);
            pr[c] *= sc;
            pi[c] *= sc;
        }
        // position, trap and field
        const R *pos = table_pos<R>(a.tb) + 3 * v;
        const R rx = k.p0[0] + sig * (k.cell[0] * (R)x + pos[0]);
        const R ry = k.p0[1] + sig * (k.cell[1] * (R)y + pos[1]);
        const R rz = k.p0[2] + sig * (k.cell[2] * (R)z + pos[2]);
        R vtrap = 0;
        if (fr.t < ph.trap_off_time)
            vtrap = (R)0.5 * (k.trap2[0] * rx * rx + k.trap2[1] * ry * ry + k.trap2[2] * rz * rz) + k.trap_offset;
        R Bx = 0, By = 0, Bz = 0, Qx = 0, Qy = 0, Qz = 0;
        if constexpr (C > 1) {
            const R dxr = rx - k.center[0], dyr = ry - k.center[1], dzr = rz - k.center[2];
            Bx = k.bq * dxr + k.bb[0];
            By = k.bq * dyr + k.bb[1];
            Bz = (R)-2 * k.bq * dzr + k.bb[2];
            Qx = k.bqq * dxr + k.bbq[0];
            Qy = k.bqq * dyr + k.bbq[1];
            Qz = (R)-2 * k.bqq * dzr + k.bbq[2];
        }
        add_local_h<R, C>(pr, pi, vtrap, Bx, By, Bz, Qx, Qy, Qz, k.c0, k.c1, k.c2, k.alpha, k.quad_sign, k.q,
                          a.quad, a.loss, hr, hi);

        const R coef = (R)a.coef;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const long long idx = own_line + c * cstride + xc;
            V2 o;
            if constexpr (MODE == MODE_UPDATE) {
                o.x = qr[c] + coef * hi[c];
                o.y = qi[c] - coef * hr[c];
            } else if constexpr (MODE == MODE_EULER) {
                o.x = pr[c] + coef * hi[c];
                o.y = pi[c] - coef * hr[c];
            } else if constexpr (MODE == MODE_ITP) {
                o.x = pr[c] - coef * hr[c];
                o.y = pi[c] - coef * hi[c];
                if (act) part += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
            } else {
                o.x = hr[c];
                o.y = hi[c];
            }
            if (act) nt_store<R>(o, &nxt[idx]);
        }
        if constexpr (MODE == MODE_ITP) part *= a.tb.dv[v];
    }

    if constexpr (MODE == MODE_ITP) {
        __shared__ double red[4];
        part = wave_sum(part);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) a.partials[b] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

// Two consecutive x values per lane (FP32): every access is one 16-byte load or store of a
// pair of complex values.  Lane l of segment s owns x0 = 128 s + 2 l and x0 + 1.  A
// neighbour pair starts at x0 + dx, which is 8-byte aligned only; the start is clamped to
// [0, lx - 2] (inside the line, so inside the allocation) and the two values are shifted
// back into place, and any value outside the box (x, y or z) is replaced by zero by a
// select, so the line padding x >= nx is never used whatever it holds.
// ROWS: the wave's pairs lie in a group of g.rows consecutive y rows of one (z, v); a lane's
// row is then a per-lane offset from the group's first row y0, a lane past the group or the box
// is clamped to the group's last row in the box and masked, and a neighbour row outside the
// box gets weight 0 per lane.
typedef float f4a __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(8)));
typedef float f2n __attribute__((ext_vector_type(2)));

template <int C, int DEG, int MODE, bool ROWS>
__device__ __forceinline__ void sweep_body_pair(const SweepArgs &a, int sweep, int b, int nb)
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
    double part = 0.0;

    if (valid) {
        // the lane's row: y0 (+ r under ROWS, clamped into the box) at element offset roff
        bool rowok = true;
        int y = y0, roff = 0;
        const int rowstride = (int)g.ys;  // elements between rows y and y + 1
        int ylo = y0, yhi = y0;           // the rows of the wave's lanes
        if constexpr (ROWS) {
            const int nrow = g.ny - y0 < g.rows ? g.ny - y0 : g.rows;  // the group's rows in the box
            rowok = r < nrow;
            const int rr = rowok ? r : nrow - 1;
            y = y0 + rr;
            roff = rr * rowstride;
            // a lane's row grows with its lane number: lanes 0 and 63 bound the wave's rows
            const int r0 = __builtin_amdgcn_readfirstlane(rr), r63 = __builtin_amdgcn_readlane(rr, 63);
            ylo = y0 + r0;
            yhi = y0 + r63;
        }
        const bool act0 = rowok && x0 < g.nx, act1 = rowok && x0 + 1 < g.nx;
        const int xc = x0 < g.nx ? x0 : ((g.nx - 1) & ~1);  // even, so xc + 1 <= lx - 1
        const int lo = roff + xc;                            // lane offset from the row-y0 line
        const float2 *__restrict__ cur = (const float2 *)a.cur;
        float2 *__restrict__ nxt = (float2 *)a.next;
        const float2 *__restrict__ prv = (const float2 *)(a.prev ? a.prev : a.next);
        const long long cstride = comp_stride(g);
        const long long own_line = line_index(g, 0, z, y0, v);
        const int lxm2 = (int)g.lx - 2;

        float pr[2][C], pi[2][C], qr[2][C], qi[2][C], ar[2][C], ai[2][C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const f4a o = *(const f4a *)(cur + own_line + c * cstride + lo);
            pr[0][c] = o.x;
            pi[0][c] = o.y;
            pr[1][c] = o.z;
            pi[1][c] = o.w;
            ar[0][c] = ar[1][c] = 0.f;
            ai[0][c] = ai[1][c] = 0.f;
        }
        if constexpr (MODE == MODE_UPDATE) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const f4a o = __builtin_nontemporal_load((const f4a *)(prv + own_line + c * cstride + lo));
                qr[0][c] = o.x;
                qi[0][c] = o.y;
                qr[1][c] = o.z;
                qi[1][c] = o.w;
            }
        }
        // packed face and own-line-to-neighbour-line offset side by side: one scalar load
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
            sh[f] = s0 - sc;  // -1, 0 or +1
            ok0[f] = (unsigned)(x0 + dx) < (unsigned)g.nx;
            ok1[f] = (unsigned)(x0 + 1 + dx) < (unsigned)g.nx;
        }
        // a wave whose rows and plane are all inside [1, n - 2] has every y / z neighbour inside
        // the box: its neighbour lines are the own line plus the per-face offsets of the table
        if (z >= 1 && z <= g.nz - 2 && ylo >= 1 && yhi <= g.ny - 2) {
            const float2 *own = cur + own_line;
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
                wf[f] = inside ? wt[f] : 0.f;  // wave-uniform unless ROWS: y / z outside the box
                if constexpr (ROWS) {
                    // row-y0 line of (zn, t), the lane's own row when the neighbour is outside
                    if ((unsigned)zn >= (unsigned)g.nz) zn = z;
                    nbase[f] = cur + line_index(g, 0, zn, y0, t) + (roff + (inside ? dy * rowstride : 0) + scf[f]);
                } else {
                    if (!inside) {
                        zn = z;
                        yn = y;
                    }
                    nbase[f] = cur + line_index(g, 0, zn, yn, t) + scf[f];
                }
            }
        }
        // degree-8 tables (FCC) pad the 4-face unknowns with 4 zero-weight faces: skip their
        // loads.  With 5 components the degree-8 faces go in two batches
        // of 4, so a wave holds one batch of neighbour pairs at a time (the second batch, loaded
        // after the first is summed and only by 8-face waves, cannot be hoisted past the
        // wave-uniform test): 159 instead of 228 VGPRs, 3 waves per SIMD instead of 2, +6-12% on
        // FCC boxes of 96-128 blocks; pairs over row groups of slab rows +1-1.5% on FCC 45-100^3.
        // Fewer components gain nothing from it.
        const int nf = DEG == 8 ? a.tb.nface[v] : DEG;
        constexpr int FB = (DEG == 8 && C == 5) ? 4 : DEG;
#pragma unroll
        for (int f0 = 0; f0 < DEG; f0 += FB) {
            if (f0 > 0 && f0 >= nf) break;
            f4u nbv[FB][C];
#pragma unroll
            for (int k = 0; k < FB; ++k)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    if (f0 + k < nf)
                        nbv[k][c] = *(const f4u *)(nbase[f0 + k] + c * cstride);
                    else
                        nbv[k][c] = (f4u)(0.f);
                }
#pragma unroll
            for (int k = 0; k < FB; ++k) {
                const int f = f0 + k;
                const float w = wf[f];
                if (dxf[f] == 0) {
                    // same x: the pair is (x0, x0 + 1) unshifted; a live value's x is inside the
                    // box (the second value of a lone last x is never stored): no per-lane masks
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const f4u e = nbv[k][c];
                        ar[0][c] += w * e.x;
                        ai[0][c] += w * e.y;
                        ar[1][c] += w * e.z;
                        ai[1][c] += w * e.w;
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const f4u e = nbv[k][c];
                        // values at x0 + dx and x0 + 1 + dx from the clamped pair
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
        float sc = 1.f;
        if constexpr (MODE == MODE_ITP) sc = (float)(*a.scale_in);
        const Physics &ph = a.ph;
        const PhysicsF &pf = ph.f;
        const FieldRec &fr = a.fields[sweep];
        const float sig = fr.fsigma;        // expanding grid: lengths x sigma
        const float ksc = sc * fr.fksc;     // Laplacian / sigma^2
        const float *pos = a.tb.pos32 + 3 * v;
        const float ry = pf.p0[1] + sig * (pf.cell[1] * (float)y + pos[1]);
        const float rz = pf.p0[2] + sig * (pf.cell[2] * (float)z + pos[2]);
        const bool trap_on = fr.t < ph.trap_off_time;
        const float coef = (float)a.coef;
        float2 out[2][C];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            float hr[C], hi[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                hr[c] = ksc * (ar[e][c] - wsum * pr[e][c]);
                hi[c] = ksc * (ai[e][c] - wsum * pi[e][c]);
                pr[e][c] *= sc;
                pi[e][c] *= sc;
            }
            const float rx = pf.p0[0] + sig * (pf.cell[0] * (float)(x0 + e) + pos[0]);
            float vtrap = 0.f;
            if (trap_on)
                vtrap = 0.5f * (pf.trap2[0] * rx * rx + pf.trap2[1] * ry * ry + pf.trap2[2] * rz * rz) +
                        pf.trap_offset;
            float Bx = 0, By = 0, Bz = 0, Qx = 0, Qy = 0, Qz = 0;
            if constexpr (C > 1) {
                const float dxr = rx - pf.center[0], dyr = ry - pf.center[1], dzr = rz - pf.center[2];
                Bx = fr.fbq * dxr + fr.fbb[0];
                By = fr.fbq * dyr + fr.fbb[1];
                Bz = -2.f * fr.fbq * dzr + fr.fbb[2];
                Qx = fr.fbqq * dxr + fr.fbbq[0];
                Qy = fr.fbqq * dyr + fr.fbbq[1];
                Qz = -2.f * fr.fbqq * dzr + fr.fbbq[2];
            }
            add_local_h<float, C>(pr[e], pi[e], vtrap, Bx, By, Bz, Qx, Qy, Qz, pf.c0, pf.c1, pf.c2, pf.alpha,
                                  pf.quad_sign, pf.q, a.quad, a.loss, hr, hi);
            const bool live = e == 0 ? act0 : act1;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float2 o;
                if constexpr (MODE == MODE_UPDATE) {
                    o.x = qr[e][c] + coef * hi[c];
                    o.y = qi[e][c] - coef * hr[c];
                } else if constexpr (MODE == MODE_EULER) {
                    o.x = pr[e][c] + coef * hi[c];
                    o.y = pi[e][c] - coef * hr[c];
                } else if constexpr (MODE == MODE_ITP) {
                    o.x = pr[e][c] - coef * hr[c];
                    o.y = pi[e][c] - coef * hi[c];
                    if (live) part += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
                } else {
                    o.x = hr[c];
                    o.y = hi[c];
                }
                out[e][c] = o;
            }
        }
        // one 16-byte store per live lane: the lone last x of an odd-length line also writes its
        // partner slot x = nx, which is line padding (lines have an even length), never read as a
        // value -- every load of it is masked by a select and every reduction and copy to the ABI
        // stops at nx.  (A separate 8-byte store for that lane made every wave run both store
        // paths: odd-length lines ran 8-10% slower.)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float2 *dst = nxt + own_line + c * cstride + lo;
            if (act0) {
                f4a o;
                o.x = out[0][c].x;
                o.y = out[0][c].y;
                o.z = out[1][c].x;
                o.w = out[1][c].y;
                __builtin_nontemporal_store(o, (f4a *)dst);
            }
        }
        if constexpr (MODE == MODE_ITP) part *= a.tb.dv[v];
    }

    if constexpr (MODE == MODE_ITP) {
        __shared__ double red[4];
        part = wave_sum(part);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) a.partials[b] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

// X: the lane form (lane_form): 1 one x per lane, 2 x pairs, 3 x pairs over a group of 2-8 y rows.
template <typename R, int C, int DEG, int MODE, int X>
__device__ __forceinline__ void sweep_body(const SweepArgs &a, int sweep, int b, int nb)
{
    if constexpr (X == 3)
        sweep_body_pair<C, DEG, MODE, true>(a, sweep, b, nb);
    else if constexpr (X == 2)
        sweep_body_pair<C, DEG, MODE, false>(a, sweep, b, nb);
    else
        sweep_body_single<R, C, DEG, MODE>(a, sweep, b, nb);
}

template <typename R, int C, int DEG, int MODE, int X>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a)
{
    sweep_body<R, C, DEG, MODE, X>(a, a.sweep, sweep_order(a.g, blockIdx.x, gridDim.x, a.order), gridDim.x);
}

// One leapfrog sweep of every member of an ensemble group in a single launch (a group: the
// members sharing component count, face degree and lanes per x, so the kernel is as
// specialised as a single run's).  Member m owns workgroups [begin_m, begin_m + blocks_m)
// (begin_m a multiple of 8, so the XCD remap of the member's own range stays valid); its
// arguments for parity p are m.args[p].
template <typename R, int C, int DEG, int X>
__global__ __launch_bounds__(256) void k_ensemble_sweep(EnsembleArgs e)
{
    // parity-1 sweeps take the whole launch in reverse (members included), keeping each
    // workgroup's XCD, so a sweep starts on the member the previous one ended with
    const int b = e.parity ? reverse_in_xcd(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    int lo = 0, hi = e.nmember - 1;
    while (lo < hi) {  // last member with begin <= b
        const int mid = (lo + hi + 1) >> 1;
        if (e.begin[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const EnsembleMember &m = e.members[lo];
    const int local = b - e.begin[lo];
    if (local >= m.blocks) return;  // padding workgroup (uniform)
    const SweepArgs &a = m.args[e.parity];
    const int sweep = e.sweep + (e.sweep_base ? *e.sweep_base : 0);
    sweep_body<R, C, DEG, MODE_UPDATE, X>(a, sweep, local, m.blocks);
}

__global__ void k_advance_base(int *base, int by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

__global__ void k_set_base(int *base, int value)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base = value;
}

// ------------------------------------------------------------------------ observables --

template <typename R, int C, int DEG>
__global__ __launch_bounds__(256) void k_observables(ObsArgs a)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const WaveCoord wc = decode(g);
    double acc[OBS_COUNT];
#pragma unroll
    for (int i = 0; i < OBS_COUNT; ++i) acc[i] = 0.0;

    if (wc.valid && wc.x < g.nx) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const V2 *__restrict__ psi = (const V2 *)a.psi;
        double pr[C], pi[C], kr[C], ki[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const V2 o = psi[line_index(g, c, z, y, v) + x];
            pr[c] = o.x;
            pi[c] = o.y;
            kr[c] = 0;
            ki[c] = 0;
        }
        const int *face = a.tb.face + v * DEG;
        const double *wt = a.tb.w64 + v * DEG;
#pragma unroll
        for (int f = 0; f < DEG; ++f) {
            const int pk = face[f];
            const int dx = (pk & 3) - 1, dy = ((pk >> 2) & 3) - 1, dz = ((pk >> 4) & 3) - 1, t = pk >> 8;
            const int zn = z + dz, yn = y + dy, xn = x + dx;
            if ((unsigned)zn < (unsigned)g.nz && (unsigned)yn < (unsigned)g.ny && (unsigned)xn < (unsigned)g.nx) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const V2 nb = psi[line_index(g, c, zn, yn, t) + xn];
                    kr[c] += wt[f] * ((double)nb.x - pr[c]);
                    ki[c] += wt[f] * ((double)nb.y - pi[c]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    kr[c] -= wt[f] * pr[c];
                    ki[c] -= wt[f] * pi[c];
                }
            }
        }
        const Physics &ph = a.ph;
        const double *pos = a.tb.pos64 + 3 * v;
        const double sig = a.field.sigma, ik = 1.0 / (sig * sig);
        const double rx = ph.p0[0] + sig * (ph.cell[0] * x + pos[0]);
        const double ry = ph.p0[1] + sig * (ph.cell[1] * y + pos[1]);
        const double rz = ph.p0[2] + sig * (ph.cell[2] * z + pos[2]);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            kr[c] *= ik;
            ki[c] *= ik;
        }
        double vtrap = 0;
        if (a.field.t < ph.trap_off_time)
            vtrap = 0.5 * (ph.trap2[0] * rx * rx + ph.trap2[1] * ry * ry + ph.trap2[2] * rz * rz) + ph.trap_offset;
        double n = 0, ekin = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            n += pr[c] * pr[c] + pi[c] * pi[c];
            ekin += pr[c] * kr[c] + pi[c] * ki[c];
        }
        double e = ekin + vtrap * n + 0.5 * ph.c0 * n * n;
        double eh = ekin + vtrap * n + ph.c0 * n * n;
        double Fx = 0, Fy = 0, Fz = 0;
        if constexpr (C > 1) {
            const double dxr = rx - ph.center[0], dyr = ry - ph.center[1], dzr = rz - ph.center[2];
            const double Bx = a.field.bq * dxr + a.field.bb[0];
            const double By = a.field.bq * dyr + a.field.bb[1];
            const double Bz = -2 * a.field.bq * dzr + a.field.bb[2];
            // F = psi^dag f psi through (e_x . f) psi etc. with the same ladder code
            double tr[C], ti[C];
            spin_dot<double, C>(1, 0, 0, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fx += pr[c] * tr[c] + pi[c] * ti[c];
            spin_dot<double, C>(0, 1, 0, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fy += pr[c] * tr[c] + pi[c] * ti[c];
            spin_dot<double, C>(0, 0, 1, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) Fz += pr[c] * tr[c] + pi[c] * ti[c];
            const double sgn = (C == 3) ? 1.0 : -1.0;
            const double f2 = Fx * Fx + Fy * Fy + Fz * Fz;
            const double bf = sgn * (Bx * Fx + By * Fy + Bz * Fz);
            e += 0.5 * ph.c1 * f2 + bf;
            eh += ph.c1 * f2 + bf;
            if constexpr (C == 5) {
                const double thr = 2 * (pr[0] * pr[4] - pi[0] * pi[4]) - 2 * (pr[1] * pr[3] - pi[1] * pi[3]) +
                                   (pr[2] * pr[2] - pi[2] * pi[2]);
                const double thi = 2 * (pr[0] * pi[4] + pi[0] * pr[4]) - 2 * (pr[1] * pi[3] + pi[1] * pr[3]) +
                                   2 * pr[2] * pi[2];
                const double th2 = thr * thr + thi * thi;
                e += ph.c2 / 10.0 * th2;
                eh += ph.c2 / 5.0 * th2;
            }
            if (a.quad) {
                const double Qx = a.field.bqq * dxr + a.field.bbq[0];
                const double Qy = a.field.bqq * dyr + a.field.bbq[1];
                const double Qz = -2 * a.field.bqq * dzr + a.field.bbq[2];
                spin_dot<double, C>(Qx, Qy, Qz, pr, pi, tr, ti);
                double m2 = 0;
                for (int c = 0; c < C; ++c) m2 += tr[c] * tr[c] + ti[c] * ti[c];
                e += ph.quad_sign * m2;
                eh += ph.quad_sign * m2;
            }
            constexpr int F = (C - 1) / 2;
            double qz = 0;
            for (int c = 0; c < C; ++c) qz += (double)((F - c) * (F - c)) * (pr[c] * pr[c] + pi[c] * pi[c]);
            e += ph.q * qz;
            eh += ph.q * qz;
        }
        const double dv = a.tb.dv[v];
        acc[OBS_NORM] = n * dv;
        acc[OBS_ENERGY] = e * dv;
        acc[OBS_EXPECT_H] = eh * dv;
        acc[OBS_MX] = Fx * dv;
        acc[OBS_MY] = Fy * dv;
        acc[OBS_MZ] = Fz * dv;
        acc[OBS_MAXDENS] = n;
        acc[OBS_NONFINITE] = (isfinite(n) && isfinite(e)) ? 0.0 : 1.0;
        acc[OBS_SPIN] = sqrt(Fx * Fx + Fy * Fy + Fz * Fz) * dv;
        acc[OBS_RX] = n * rx * dv;
        acc[OBS_RY] = n * ry * dv;
        acc[OBS_RZ] = n * rz * dv;
        if constexpr (C > 1) {
            // the reference's local spin (localAvgSpinAndDensity, Spin1GpeBCC.cu:235-271; Spin2GpeBCC.cu
            // 214-262 applies the same spin-1 expressions to m = 1, 0, -1): x, y of the spinor divided
            // by sqrt(n), z = |psi_1|^2 - |psi_-1|^2 undivided; its integrals weight it by n
            // (integrateSpinAndDensity, Spin1GpeBCC.cu:992-1020).  Unknowns with n = 0 add nothing.
            constexpr int o = C == 5 ? 1 : 0;
            const double zr = (pr[o] * pr[o] + pi[o] * pi[o]) - (pr[o + 2] * pr[o + 2] + pi[o + 2] * pi[o + 2]);
            double sx = 0, sy = 0;
            if (n > 0) {
                const double sq = sqrt(n);
                double ur[3], ui[3];
                for (int k = 0; k < 3; ++k) {
                    ur[k] = pr[o + k] / sq;
                    ui[k] = pi[o + k] / sq;
                }
                // sqrt2 (conj(u1) u0 + conj(u0) u-1)
                sx = 1.4142135623730951 * ((ur[0] * ur[1] + ui[0] * ui[1]) + (ur[1] * ur[2] + ui[1] * ui[2]));
                sy = 1.4142135623730951 * ((ur[0] * ui[1] - ui[0] * ur[1]) + (ur[1] * ui[2] - ui[1] * ur[2]));
            }
            acc[OBS_REF_SPIN] = n * sqrt(sx * sx + sy * sy + zr * zr) * dv;
            acc[OBS_REF_MZ] = n * zr * dv;
        }
    }
    __shared__ double red[4][OBS_COUNT];
#pragma unroll
    for (int i = 0; i < OBS_COUNT; ++i) {
        const double r = (i == OBS_MAXDENS) ? wave_max(acc[i]) : wave_sum(acc[i]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][i] = r;
    }
    __syncthreads();
    if (threadIdx.x < OBS_COUNT) {
        const int i = threadIdx.x;
        double r;
        if (i == OBS_MAXDENS)
            r = fmax(fmax(red[0][i], red[1][i]), fmax(red[2][i], red[3][i]));
        else
            r = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
        a.partials[(long long)blockIdx.x * OBS_COUNT + i] = r;
    }
}

// Two local reductions over a level, one x per lane:
//   max over unknowns of the local Hamiltonian bound (Spin1GpeBCC.cu:144-189): with
//   p = V(r) + c0 n and B = B(r, t) + c1 <F> (<F> = psi^dag f psi, not divided by n),
//   max(p + |Bx| / sqrt2, p + |By| / sqrt2, p + |Bz|); a scalar state gives V + c0 n;
//   sum over unknowns of dv |a|^2 |psi - a|^2 with a = phase * ref (Spin1GpeBCC.cu:436-461
//   with the volume factor of its integrate reduction, :1622-1631; spin-2: the same sum,
//   RelativisticSpin2Gpe.cu:227).  Spin-2 has no Hamiltonian bound: its column holds V + c0 n.
template <typename R, int C>
__global__ __launch_bounds__(256) void k_aux(AuxArgs a)
{
    using V2 = vec2_t<R>;
    const Geometry &g = a.g;
    const WaveCoord wc = decode(g);
    double hmax = -INFINITY, wsum = 0.0;
    if (wc.valid && wc.x < g.nx) {
        const int x = wc.x, v = wc.v, y = wc.y, z = wc.z;
        const V2 *__restrict__ psi = (const V2 *)a.psi;
        double pr[C], pi[C];
        double n = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const V2 o = psi[line_index(g, c, z, y, v) + x];
            pr[c] = o.x;
            pi[c] = o.y;
            n += pr[c] * pr[c] + pi[c] * pi[c];
        }
        const Physics &ph = a.ph;
        const double *pos = a.tb.pos64 + 3 * v;
        const double sig = a.field.sigma;
        const double rx = ph.p0[0] + sig * (ph.cell[0] * x + pos[0]);
        const double ry = ph.p0[1] + sig * (ph.cell[1] * y + pos[1]);
        const double rz = ph.p0[2] + sig * (ph.cell[2] * z + pos[2]);
        double vtrap = 0;
        if (a.field.t < ph.trap_off_time)
            vtrap = 0.5 * (ph.trap2[0] * rx * rx + ph.trap2[1] * ry * ry + ph.trap2[2] * rz * rz) + ph.trap_offset;
        const double p = vtrap + ph.c0 * n;
        hmax = p;
        if constexpr (C == 3) {
            const double tr = (pr[0] * pr[1] + pi[0] * pi[1]) + (pr[1] * pr[2] + pi[1] * pi[2]);
            const double ti = (pr[0] * pi[1] - pi[0] * pr[1]) + (pr[1] * pi[2] - pi[1] * pr[2]);
            const double s2 = 1.4142135623730951;
            const double Fx = s2 * tr, Fy = s2 * ti;
            const double Fz = (pr[0] * pr[0] + pi[0] * pi[0]) - (pr[2] * pr[2] + pi[2] * pi[2]);
            const double Bx = a.field.bq * (rx - ph.center[0]) + a.field.bb[0] + ph.c1 * Fx;
            const double By = a.field.bq * (ry - ph.center[1]) + a.field.bb[1] + ph.c1 * Fy;
            const double Bz = -2 * a.field.bq * (rz - ph.center[2]) + a.field.bb[2] + ph.c1 * Fz;
            hmax = fmax(p + fabs(Bx) / s2, fmax(p + fabs(By) / s2, p + fabs(Bz)));
        }
        if (a.ref) {
            const long long line = (long long)g.nx * g.V;
            const long long plane = line * g.ny, comp = plane * g.nz;
            const long long at = (long long)z * plane + (long long)y * line + (long long)x * g.V + v;
            double an = 0, dn = 0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const double2 r = a.ref[(long long)c * comp + at];
                const double ar = a.phase_re * r.x - a.phase_im * r.y;
                const double ai = a.phase_re * r.y + a.phase_im * r.x;
                an += ar * ar + ai * ai;
                dn += (pr[c] - ar) * (pr[c] - ar) + (pi[c] - ai) * (pi[c] - ai);
            }
            wsum = an * dn * a.tb.dv[v];
        }
    }
    __shared__ double red[4][AUX_COUNT];
    const double m = wave_max(hmax), s = wave_sum(wsum);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][AUX_MAXHAM] = m;
        red[threadIdx.x >> 6][AUX_WDIFF] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.partials[(long long)blockIdx.x * AUX_COUNT + AUX_MAXHAM] =
            fmax(fmax(red[0][AUX_MAXHAM], red[1][AUX_MAXHAM]), fmax(red[2][AUX_MAXHAM], red[3][AUX_MAXHAM]));
        a.partials[(long long)blockIdx.x * AUX_COUNT + AUX_WDIFF] =
            (red[0][AUX_WDIFF] + red[1][AUX_WDIFF]) + (red[2][AUX_WDIFF] + red[3][AUX_WDIFF]);
    }
}

// --------------------------------------------------------------------- small kernels --

// Deterministic two-stage reduction of per-workgroup partials (n rows x width columns; column
// maxcol takes the maximum, the others the sum): stage 1 gives each of up to FIN_BLOCKS
// workgroups a fixed contiguous chunk of rows, stage 2 reduces the chunk results in one
// workgroup.  Fixed chunks, strides and trees: the same inputs give the same bits.
constexpr int FIN_BLOCKS = 256;

__device__ __forceinline__ double fin_op(bool is_max, double a, double b) { return is_max ? fmax(a, b) : a + b; }

__global__ __launch_bounds__(256) void k_finalize_stage(const double *partials, int n, int width, int maxcol,
                                                        int chunk, double *out)
{
    __shared__ double red[256];
    const int lo = blockIdx.x * chunk;
    const int hi = min(n, lo + chunk);
    for (int col = 0; col < width; ++col) {
        const bool is_max = col == maxcol;
        double acc = is_max ? -INFINITY : 0.0;
        for (int i = lo + (int)threadIdx.x; i < hi; i += 256) acc = fin_op(is_max, acc, partials[(long long)i * width + col]);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] = fin_op(is_max, red[threadIdx.x], red[threadIdx.x + st]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(long long)blockIdx.x * width + col] = red[0];
        __syncthreads();
    }
}

// Column sums of `nslot` independent partial tables, one workgroup per table: table s holds n
// records of `width` doubles and its sums go to out + s * out_stride.  A fixed stride order
// and tree, so identical inputs give identical bits.
__global__ __launch_bounds__(256) void k_sum_tables(const double *partials, int n, int width, double *out,
                                                    long long out_stride)
{
    __shared__ double red[256];
    const double *tab = partials + (long long)blockIdx.x * n * width;
    double *dst = out + (long long)blockIdx.x * out_stride;
    for (int col = 0; col < width; ++col) {
        double acc = 0.0;
        for (int i = (int)threadIdx.x; i < n; i += 256) acc += tab[(long long)i * width + col];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) dst[col] = red[0];
        __syncthreads();
    }
}

__global__ void k_norm_to_scale(const double *in, double target, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = sqrt(target / in[0]);
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale(vec2_t<R> *psi, long long n, double factor, const double *dfactor)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const R f = (R)(dfactor ? *dfactor : factor);
    vec2_t<R> o = psi[i];
    o.x *= f;
    o.y *= f;
    psi[i] = o;
}

template <typename R>
__global__ __launch_bounds__(256) void k_phase_copy(const vec2_t<R> *src, vec2_t<R> *dst, long long n, double pre,
                                                     double pim)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const vec2_t<R> s = src[i];
    vec2_t<R> o;
    o.x = (R)(pre * (double)s.x - pim * (double)s.y);
    o.y = (R)(pre * (double)s.y + pim * (double)s.x);
    dst[i] = o;
}

// Counter-based uniform variate in [0, 1) from (seed, index, stream): a 64-bit mix
// (splitmix64 finaliser) of the counter, so every unknown draws independently of the
// traversal and of the launch shape.
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long idx, unsigned s)
{
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + idx * 0xD1B54A32D192ED03ull + (unsigned long long)s * 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// psi_c = sqrt(max(mu - V, 0) / g) (w_c + amp * eps_c), eps_c uniform in the unit disk: multiplicative
// noise for one component with w = 1, spinor noise around the spinor w otherwise.
template <typename R>
__global__ __launch_bounds__(256) void k_fill_tf(Geometry g, Tables tb, Physics ph, int ncomp, FillArgs f, void *psi)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long total = (long long)g.nz * g.ny * g.V * g.nx;
    if (i >= total) return;
    long long r = i;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int v = (int)(r % g.V);
    r /= g.V;
    const int y = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    const double *pos = tb.pos64 + 3 * v;
    const double rx = ph.p0[0] + ph.cell[0] * x + pos[0];
    const double ry = ph.p0[1] + ph.cell[1] * y + pos[1];
    const double rz = ph.p0[2] + ph.cell[2] * z + pos[2];
    const double V = 0.5 * (ph.trap2[0] * rx * rx + ph.trap2[1] * ry * ry + ph.trap2[2] * rz * rz) + ph.trap_offset;
    const double a = sqrt(fmax(f.mu - V, 0.0) / f.g);
    const unsigned long long node = (((unsigned long long)z * g.ny + y) * g.nx + x) * g.V + v;
    vec2_t<R> *out = (vec2_t<R> *)psi;
    for (int c = 0; c < ncomp; ++c) {
        double er = 0, ei = 0;
        if (f.amp != 0.0) {
            const double rad = sqrt(hash_uniform(f.seed, node, 2 * c));
            const double ang = 6.283185307179586 * hash_uniform(f.seed, node, 2 * c + 1);
            er = rad * cos(ang);
            ei = rad * sin(ang);
        }
        vec2_t<R> o;  // a (w_c + amp eps_c)
        o.x = (R)(a * (f.w[2 * c] + f.amp * er));
        o.y = (R)(a * (f.w[2 * c + 1] + f.amp * ei));
        out[line_index(g, c, z, y, v) + x] = o;
    }
}

// Unknown (x, v, y, z) of flat index i, x fastest: consecutive lanes take consecutive x of one
// line of the device layout.
struct Unknown {
    int x, v, y, z;
};

__device__ __forceinline__ Unknown unknown_of(const Geometry &g, long long i)
{
    Unknown u;
    u.x = (int)(i % g.nx);
    i /= g.nx;
    u.v = (int)(i % g.V);
    i /= g.V;
    u.y = (int)(i % g.ny);
    u.z = (int)(i / g.ny);
    return u;
}

// Spinor phase projection keeping the density n of every unknown (phases of gpe.h
// GPE_PHASE_*): spin-1 polar (0, sqrt n, 0), ferromagnetic (sqrt n, 0, 0), mixed
// (sqrt(n mix), sqrt(n (1 - mix)), 0) (Spin1GpeBCC.cu:463-551); spin-2 UN, horizontal and vertical
// BN, cyclic with the relative phase `param` (Spin2GpeBCC.cu:417-558), ferromagnetic.  The
// density is summed over m = F..-F and the amplitudes formed as in the reference.
template <typename R, int C>
__global__ __launch_bounds__(256) void k_project(Geometry g, vec2_t<R> *psi, int phase, double param)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)g.nz * g.ny * g.V * g.nx) return;
    const Unknown u = unknown_of(g, i);
    double n = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const vec2_t<R> o = psi[line_index(g, c, u.z, u.y, u.v) + u.x];
        const double re = o.x, im = o.y;
        n += re * re + im * im;
    }
    double outr[C], outi[C];
#pragma unroll
    for (int c = 0; c < C; ++c) outr[c] = outi[c] = 0.0;
    const double cp = cos(param), sp = sin(param);
    if constexpr (C == 3) {
        if (phase == 0) {
            outr[1] = sqrt(n);
        } else if (phase == 1) {
            outr[0] = sqrt(n);
        } else {
            outr[0] = sqrt(n * param);
            outr[1] = sqrt(n * (1.0 - param));
        }
    } else {
        if (phase == 1) {
            outr[0] = sqrt(n);
        } else if (phase == 3) {
            outr[2] = sqrt(n);
        } else if (phase == 4 || phase == 5) {
            const double a = sqrt(n / 2);
            const int lo = phase == 4 ? 0 : 1;
            outr[lo] = a;
            outr[4 - lo] = cp * a;
            outi[4 - lo] = sp * a;
        } else {
            const double a2 = sqrt(n * 1 / 3), a1 = sqrt(n * 2 / 3);
            outr[0] = a2;
            outr[3] = cp * a1;
            outi[3] = sp * a1;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        vec2_t<R> o;
        o.x = (R)outr[c];
        o.y = (R)outi[c];
        psi[line_index(g, c, u.z, u.y, u.v) + u.x] = o;
    }
}

// psi_m += amp sqrt(n) eps_m, eps_m uniform in the unit disk from the counter hash of (seed,
// unknown in ABI order, 2m / 2m + 1) -- the variates of k_fill_tf -- n the density before.
template <typename R, int C>
__global__ __launch_bounds__(256) void k_noise(Geometry g, vec2_t<R> *psi, double amp, unsigned long long seed)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)g.nz * g.ny * g.V * g.nx) return;
    const Unknown u = unknown_of(g, i);
    const unsigned long long node = (((unsigned long long)u.z * g.ny + u.y) * g.nx + u.x) * g.V + u.v;
    double pr[C], pi[C], n = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const vec2_t<R> o = psi[line_index(g, c, u.z, u.y, u.v) + u.x];
        pr[c] = o.x;
        pi[c] = o.y;
        n += pr[c] * pr[c] + pi[c] * pi[c];
    }
    const double a = amp * sqrt(n);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double rad = sqrt(hash_uniform(seed, node, 2 * c));
        const double ang = 6.283185307179586 * hash_uniform(seed, node, 2 * c + 1);
        vec2_t<R> o;
        o.x = (R)(pr[c] + a * rad * cos(ang));
        o.y = (R)(pi[c] + a * rad * sin(ang));
        psi[line_index(g, c, u.z, u.y, u.v) + u.x] = o;
    }
}

hipError_t launch_project(int fp64, int ncomp, const Geometry &g, void *psi, int phase, double param, hipStream_t s)
{
    const unsigned grid = (unsigned)(((long long)g.nz * g.ny * g.V * g.nx + 255) / 256);
    if (ncomp == 3) {
        if (fp64) hipLaunchKernelGGL((k_project<double, 3>), dim3(grid), dim3(256), 0, s, g, (double2 *)psi, phase, param);
        else hipLaunchKernelGGL((k_project<float, 3>), dim3(grid), dim3(256), 0, s, g, (float2 *)psi, phase, param);
    } else if (ncomp == 5) {
        if (fp64) hipLaunchKernelGGL((k_project<double, 5>), dim3(grid), dim3(256), 0, s, g, (double2 *)psi, phase, param);
        else hipLaunchKernelGGL((k_project<float, 5>), dim3(grid), dim3(256), 0, s, g, (float2 *)psi, phase, param);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_noise(int fp64, int ncomp, const Geometry &g, void *psi, double amp, unsigned long long seed,
                        hipStream_t s)
{
    const unsigned grid = (unsigned)(((long long)g.nz * g.ny * g.V * g.nx + 255) / 256);
#define NZ(R, C) hipLaunchKernelGGL((k_noise<R, C>), dim3(grid), dim3(256), 0, s, g, (vec2_t<R> *)psi, amp, seed)
    if (fp64) {
        if (ncomp == 1) NZ(double, 1); else if (ncomp == 3) NZ(double, 3); else NZ(double, 5);
    } else {
        if (ncomp == 1) NZ(float, 1); else if (ncomp == 3) NZ(float, 3); else NZ(float, 5);
    }
#undef NZ
    return hipGetLastError();
}

hipError_t launch_fill_tf(int fp64, const Geometry &g, const Tables &tb, const Physics &ph, int ncomp, const FillArgs &f,
                          void *psi, hipStream_t s)
{
    const long long total = (long long)g.nz * g.ny * g.V * g.nx;
    const unsigned grid = (unsigned)((total + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_fill_tf<double>, dim3(grid), dim3(256), 0, s, g, tb, ph, ncomp, f, psi);
    else
        hipLaunchKernelGGL(k_fill_tf<float>, dim3(grid), dim3(256), 0, s, g, tb, ph, ncomp, f, psi);
    return hipGetLastError();
}

// Per-unknown field maps in ABI order (z, y, x, v), `width` doubles each.
//   FIELD_DENSITY    n = sum_m |psi_m|^2
//   FIELD_SPIN       <F> = psi^dag F psi / n (0 where n = 0)
//   FIELD_SPIN_NORM  |psi^dag F psi|
//   FIELD_UV_THETA   spin-1 Cartesian vector a = (psi_-1 - psi_1, -i (psi_-1 + psi_1), sqrt2 psi_0) / sqrt2
//                    = m + i n; theta = atan2(-2 m.n, |m|^2 - |n|^2) / 2 in [0, pi) makes u = Re(a e^{i theta}),
//                    v = Im(a e^{i theta}) orthogonal; (u, v) ordered so that |u| >= |v|; out = u, v, theta.
template <typename R, int C>
__global__ __launch_bounds__(256) void k_local_fields(Geometry g, const vec2_t<R> *psi, int kind, double *out)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long total = (long long)g.nz * g.ny * g.nx * g.V;
    if (i >= total) return;
    long long r = i;
    const int v = (int)(r % g.V);
    r /= g.V;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int y = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    double pr[C], pi[C];
    double n = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const vec2_t<R> o = psi[line_index(g, c, z, y, v) + x];
        pr[c] = o.x;
        pi[c] = o.y;
        n += pr[c] * pr[c] + pi[c] * pi[c];
    }
    if (kind == FIELD_DENSITY) {
        out[i] = n;
        return;
    }
    if (kind == FIELD_UV_THETA) {
        if constexpr (C == 3) {
            const double s = 0.7071067811865476;
            const double m[3] = {s * (pr[2] - pr[0]), s * (pi[2] + pi[0]), pr[1]};
            const double q[3] = {s * (pi[2] - pi[0]), -s * (pr[2] + pr[0]), pi[1]};
            const double mq = m[0] * q[0] + m[1] * q[1] + m[2] * q[2];
            const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
            const double qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
            double th = 0.5 * atan2(-2.0 * mq, mm - qq);
            if (th < 0) th += 3.141592653589793;
            const double ct = cos(th), st = sin(th);
            double u[3], w[3], uu = 0, ww = 0;
            for (int k = 0; k < 3; ++k) {
                u[k] = m[k] * ct - q[k] * st;
                w[k] = m[k] * st + q[k] * ct;
                uu += u[k] * u[k];
                ww += w[k] * w[k];
            }
            double *o = out + 7 * i;
            const bool keep = uu >= ww;
            for (int k = 0; k < 3; ++k) {
                o[k] = keep ? u[k] : w[k];
                o[3 + k] = keep ? w[k] : u[k];
            }
            o[6] = th;
        }
        return;
    }
    double F[3] = {0, 0, 0};
    if constexpr (C > 1) {
        double tr[C], ti[C];
        for (int a = 0; a < 3; ++a) {
            spin_dot<double, C>(a == 0, a == 1, a == 2, pr, pi, tr, ti);
            for (int c = 0; c < C; ++c) F[a] += pr[c] * tr[c] + pi[c] * ti[c];
        }
    }
    if (kind == FIELD_SPIN_NORM) {
        out[i] = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    } else {
        const double inv = n > 0 ? 1.0 / n : 0.0;
        for (int a = 0; a < 3; ++a) out[3 * i + a] = F[a] * inv;
    }
}

hipError_t launch_local_fields(int fp64, int ncomp, const Geometry &g, const void *psi, int kind, double *out,
                               hipStream_t s)
{
    const long long total = (long long)g.nz * g.ny * g.nx * g.V;
    const unsigned grid = (unsigned)((total + 255) / 256);
#define LF(R, C) hipLaunchKernelGGL((k_local_fields<R, C>), dim3(grid), dim3(256), 0, s, g, (const vec2_t<R> *)psi, kind, out)
    if (fp64) {
        if (ncomp == 1) LF(double, 1); else if (ncomp == 3) LF(double, 3); else LF(double, 5);
    } else {
        if (ncomp == 1) LF(float, 1); else if (ncomp == 3) LF(float, 3); else LF(float, 5);
    }
#undef LF
    return hipGetLastError();
}

// Streaming reference kernels for the achievable HBM rate: a copy and an in-place
// read-2-write-1 update (the sweep's own HBM pattern), 16-byte non-temporal accesses.
typedef float f4s __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const f4s *__restrict__ a, f4s *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

__global__ __launch_bounds__(256) void k_stream_update(const f4s *__restrict__ a, f4s *b, long long n, float s)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&b[i]) + s * a[i], &b[i]);
}

hipError_t launch_stream(int kind, const void *a, void *b, long long n16, hipStream_t s)
{
    const unsigned grid = (unsigned)((n16 + 255) / 256);
    if (kind == 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const f4s *)a, (f4s *)b, n16);
    else
        hipLaunchKernelGGL(k_stream_update, dim3(grid), dim3(256), 0, s, (const f4s *)a, (f4s *)b, n16, 0.5f);
    return hipGetLastError();
}

// ABI index (c, z, y, x, v) <-> device index (c, z, y, v, x)
template <typename R, bool PACK>
__global__ __launch_bounds__(256) void k_pack(Geometry g, int ncomp, const double2 *abi_in, double2 *abi_out,
                                              const vec2_t<R> *dev_in, vec2_t<R> *dev_out)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long total = (long long)ncomp * g.nz * g.ny * g.nx * g.V;
    if (i >= total) return;
    long long r = i;
    const int v = (int)(r % g.V);
    r /= g.V;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int y = (int)(r % g.ny);
    r /= g.ny;
    const int z = (int)(r % g.nz);
    const int c = (int)(r / g.nz);
    const long long d = line_index(g, c, z, y, v) + x;
    if constexpr (PACK) {
        const double2 s = abi_in[i];
        vec2_t<R> o;
        o.x = (R)s.x;
        o.y = (R)s.y;
        dev_out[d] = o;
    } else {
        const vec2_t<R> s = dev_in[d];
        abi_out[i] = make_double2((double)s.x, (double)s.y);
    }
}

// ------------------------------------------------------------------------- launchers --

static inline unsigned grid1(long long n) { return (unsigned)((n + 255) / 256); }

template <typename R, int C, int DEG, int X>
static hipError_t sweep_mode_x(int mode, const SweepArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    switch (mode) {
    case MODE_UPDATE: hipLaunchKernelGGL((k_sweep<R, C, DEG, MODE_UPDATE, X>), grid, block, 0, s, a); break;
    case MODE_EULER: hipLaunchKernelGGL((k_sweep<R, C, DEG, MODE_EULER, X>), grid, block, 0, s, a); break;
    case MODE_ITP: hipLaunchKernelGGL((k_sweep<R, C, DEG, MODE_ITP, X>), grid, block, 0, s, a); break;
    case MODE_HPSI: hipLaunchKernelGGL((k_sweep<R, C, DEG, MODE_HPSI, X>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R, int C, int DEG>
static hipError_t sweep_mode(int mode, const SweepArgs &a, hipStream_t s)
{
    const int form = lane_form(a.g.xpl, a.g.rows);
    if (form == 1) return sweep_mode_x<R, C, DEG, 1>(mode, a, s);
    if constexpr (pair_lanes_ok(std::is_same<R, double>::value, DEG, C)) {
        if (form == 2) return sweep_mode_x<R, C, DEG, 2>(mode, a, s);
        if (form == 3) return sweep_mode_x<R, C, DEG, 3>(mode, a, s);
    }
    return hipErrorInvalidValue;
}

template <typename R, int C>
static hipError_t sweep_deg(int mode, const SweepArgs &a, hipStream_t s)
{
    switch (a.g.deg) {
    case 4: return sweep_mode<R, C, 4>(mode, a, s);
    case 6: return sweep_mode<R, C, 6>(mode, a, s);
    case 8: return sweep_mode<R, C, 8>(mode, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R>
static hipError_t sweep_comp(int ncomp, int mode, const SweepArgs &a, hipStream_t s)
{
    switch (ncomp) {
    case 1: return sweep_deg<R, 1>(mode, a, s);
    case 3: return sweep_deg<R, 3>(mode, a, s);
    case 5: return sweep_deg<R, 5>(mode, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep(int fp64, int ncomp, int mode, const SweepArgs &a, hipStream_t s)
{
    return fp64 ? sweep_comp<double>(ncomp, mode, a, s) : sweep_comp<float>(ncomp, mode, a, s);
}

template <typename R, int C>
static hipError_t obs_deg(const ObsArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    switch (a.g.deg) {
    case 4: hipLaunchKernelGGL((k_observables<R, C, 4>), grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_observables<R, C, 6>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_observables<R, C, 8>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename R>
static hipError_t obs_comp(int ncomp, const ObsArgs &a, hipStream_t s)
{
    switch (ncomp) {
    case 1: return obs_deg<R, 1>(a, s);
    case 3: return obs_deg<R, 3>(a, s);
    case 5: return obs_deg<R, 5>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_observables(int fp64, int ncomp, const ObsArgs &a, hipStream_t s)
{
    return fp64 ? obs_comp<double>(ncomp, a, s) : obs_comp<float>(ncomp, a, s);
}

template <typename R>
static hipError_t aux_comp(int ncomp, const AuxArgs &a, hipStream_t s)
{
    dim3 grid(a.g.blocks), block(256);
    switch (ncomp) {
    case 1: hipLaunchKernelGGL((k_aux<R, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_aux<R, 3>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_aux<R, 5>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_aux(int fp64, int ncomp, const AuxArgs &a, hipStream_t s)
{
    return fp64 ? aux_comp<double>(ncomp, a, s) : aux_comp<float>(ncomp, a, s);
}

template <typename R, int C, int DEG>
static hipError_t ens_x(int form, const EnsembleArgs &e, int blocks, hipStream_t s)
{
    dim3 grid(blocks), block(256);
    if (form == 1) {
        hipLaunchKernelGGL((k_ensemble_sweep<R, C, DEG, 1>), grid, block, 0, s, e);
        return hipGetLastError();
    }
    if constexpr (pair_lanes_ok(std::is_same<R, double>::value, DEG, C)) {
        if (form == 2) {
            hipLaunchKernelGGL((k_ensemble_sweep<R, C, DEG, 2>), grid, block, 0, s, e);
            return hipGetLastError();
        }
        if (form == 3) {
            hipLaunchKernelGGL((k_ensemble_sweep<R, C, DEG, 3>), grid, block, 0, s, e);
            return hipGetLastError();
        }
    }
    return hipErrorInvalidValue;
}

template <typename R, int C>
static hipError_t ens_deg(int deg, int form, const EnsembleArgs &e, int blocks, hipStream_t s)
{
    switch (deg) {
    case 4: return ens_x<R, C, 4>(form, e, blocks, s);
    case 6: return ens_x<R, C, 6>(form, e, blocks, s);
    case 8: return ens_x<R, C, 8>(form, e, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R>
static hipError_t ens_comp(int ncomp, int deg, int form, const EnsembleArgs &e, int blocks, hipStream_t s)
{
    switch (ncomp) {
    case 1: return ens_deg<R, 1>(deg, form, e, blocks, s);
    case 3: return ens_deg<R, 3>(deg, form, e, blocks, s);
    case 5: return ens_deg<R, 5>(deg, form, e, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ensemble_sweep(int fp64, int ncomp, int deg, int form, const EnsembleArgs &e, int blocks,
                                 hipStream_t s)
{
    return fp64 ? ens_comp<double>(ncomp, deg, form, e, blocks, s)
                : ens_comp<float>(ncomp, deg, form, e, blocks, s);
}

hipError_t launch_advance_base(int *base, int by, hipStream_t s)
{
    hipLaunchKernelGGL(k_advance_base, dim3(1), dim3(64), 0, s, base, by);
    return hipGetLastError();
}

hipError_t launch_set_base(int *base, int value, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, s, base, value);
    return hipGetLastError();
}

hipError_t launch_finalize(const double *partials, int n, int width, int maxcol, double *out, double *scratch,
                           hipStream_t s)
{
    if (n <= 0 || width <= 0 || width > FIN_WIDTH_MAX) return hipErrorInvalidValue;
    int nb = (n + 1023) / 1024;
    if (nb > FIN_BLOCKS) nb = FIN_BLOCKS;
    if (nb == 1) {
        hipLaunchKernelGGL(k_finalize_stage, dim3(1), dim3(256), 0, s, partials, n, width, maxcol, n, out);
        return hipGetLastError();
    }
    const int chunk = (n + nb - 1) / nb;
    nb = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_finalize_stage, dim3(nb), dim3(256), 0, s, partials, n, width, maxcol, chunk, scratch);
    hipLaunchKernelGGL(k_finalize_stage, dim3(1), dim3(256), 0, s, (const double *)scratch, nb, width, maxcol, nb,
                       out);
    return hipGetLastError();
}

hipError_t launch_sum_tables(const double *partials, int nslot, int n, int width, double *out, long long out_stride,
                             hipStream_t s)
{
    if (nslot <= 0 || n <= 0 || width <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sum_tables, dim3(nslot), dim3(256), 0, s, partials, n, width, out, out_stride);
    return hipGetLastError();
}

hipError_t launch_norm_to_scale(const double *in, double target, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_norm_to_scale, dim3(1), dim3(64), 0, s, in, target, out);
    return hipGetLastError();
}

hipError_t launch_scale(int fp64, void *psi, long long nelem, double factor, const double *dfactor, hipStream_t s)
{
    if (nelem <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL(k_scale<double>, dim3(grid1(nelem)), dim3(256), 0, s, (double2 *)psi, nelem, factor, dfactor);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(grid1(nelem)), dim3(256), 0, s, (float2 *)psi, nelem, factor, dfactor);
    return hipGetLastError();
}

hipError_t launch_phase_copy(int fp64, const void *src, void *dst, long long nelem, double pre, double pim,
                             hipStream_t s)
{
    if (nelem <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL(k_phase_copy<double>, dim3(grid1(nelem)), dim3(256), 0, s, (const double2 *)src,
                           (double2 *)dst, nelem, pre, pim);
    else
        hipLaunchKernelGGL(k_phase_copy<float>, dim3(grid1(nelem)), dim3(256), 0, s, (const float2 *)src,
                           (float2 *)dst, nelem, pre, pim);
    return hipGetLastError();
}

hipError_t launch_pack(int fp64, const Geometry &g, int ncomp, const double *abi, void *dev, hipStream_t s)
{
    const long long total = (long long)ncomp * g.nz * g.ny * g.nx * g.V;
    if (fp64)
        hipLaunchKernelGGL((k_pack<double, true>), dim3(grid1(total)), dim3(256), 0, s, g, ncomp,
                           (const double2 *)abi, nullptr, nullptr, (double2 *)dev);
    else
        hipLaunchKernelGGL((k_pack<float, true>), dim3(grid1(total)), dim3(256), 0, s, g, ncomp,
                           (const double2 *)abi, nullptr, nullptr, (float2 *)dev);
    return hipGetLastError();
}

hipError_t launch_unpack(int fp64, const Geometry &g, int ncomp, const void *dev, double *abi, hipStream_t s)
{
    const long long total = (long long)ncomp * g.nz * g.ny * g.nx * g.V;
    if (fp64)
        hipLaunchKernelGGL((k_pack<double, false>), dim3(grid1(total)), dim3(256), 0, s, g, ncomp, nullptr,
                           (double2 *)abi, (const double2 *)dev, nullptr);
    else
        hipLaunchKernelGGL((k_pack<float, false>), dim3(grid1(total)), dim3(256), 0, s, g, ncomp, nullptr,
                           (double2 *)abi, (const float2 *)dev, nullptr);
    return hipGetLastError();
}

}  // namespace gpe
