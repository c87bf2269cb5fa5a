// Device-side helpers shared by the HIP kernel files: work decode, layout indexing, the
// local spinor Hamiltonian and wave reductions.
#pragma once

#include "gpe_internal.h"

#include <type_traits>

namespace gpe {

template <typename R>
using vec2_t = typename std::conditional<std::is_same<R, float>::value, float2, double2>::type;

// Non-temporal (streaming) access to complex values through the native 2-vector types.
template <typename R>
struct native2;
template <>
struct native2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct native2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename R>
__device__ __forceinline__ vec2_t<R> nt_load(const vec2_t<R> *p)
{
    const typename native2<R>::type v = __builtin_nontemporal_load((const typename native2<R>::type *)p);
    vec2_t<R> o;
    o.x = v.x;
    o.y = v.y;
    return o;
}

template <typename R>
__device__ __forceinline__ void nt_store(const vec2_t<R> &o, vec2_t<R> *p)
{
    typename native2<R>::type v;
    v.x = o.x;
    v.y = o.y;
    __builtin_nontemporal_store(v, (typename native2<R>::type *)p);
}

__device__ __forceinline__ int logical_block(int b, int nb)
{
    // hardware dispatch is round-robin over the 8 XCDs: give XCD k a contiguous range
    const int q = nb >> 3, r = nb & 7, k = b & 7;
    return k * q + (k < r ? k : r) + (b >> 3);
}

// The workgroup whose tiles workgroup b of nb takes: b itself for order 0; for order 1 the
// reverse order, within each XCD's range when the XCD remap is on (b and the result share
// b mod 8, so logical_block of the result runs backwards through the same XCD's range).
// Alternating the order between consecutive sweeps starts each sweep on the lines the one
// before it touched last, which the memory-side cache still holds.
__device__ __forceinline__ int reverse_in_xcd(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, k = b & 7;
    const int len = q + (k < r ? 1 : 0);  // workgroups b' = k + 8 i, i < len
    return k + 8 * (len - 1 - (b >> 3));
}

__device__ __forceinline__ int sweep_order(const Geometry &g, int b, int nb, int order)
{
    if (!order) return b;
    return g.xcd_remap ? reverse_in_xcd(b, nb) : nb - 1 - b;
}

struct WaveCoord {
    int x, v, y, z;
    bool valid;
};

// n / d for n < 2^31 (wave and lane indices; the geometry keeps them there): with
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(n, m) <= n, so t + n does not wrap and
// (t + n) >> l is the quotient; d = 1 has m = 1, l = 0 and gives n.  Three scalar
// instructions and no branch.
__device__ __forceinline__ unsigned fast_div(unsigned n, const FastDiv &f)
{
    return (__umulhi(n, f.m) + n) >> f.l;
}

// The two fastest traversal indices of wave w: the unknown v and the x segment, the segment
// fastest, or (g.v_inner) v fastest so that a workgroup's waves take neighbouring unknowns of
// one segment.  Returns w / (V nseg).
__device__ __forceinline__ unsigned split_v_seg(const Geometry &g, unsigned w, int *v, int *xseg)
{
    if (g.v_inner) {
        const unsigned q0 = fast_div(w, g.f_V);
        *v = (int)(w - q0 * (unsigned)g.V);
        const unsigned q1 = fast_div(q0, g.f_nseg);
        *xseg = (int)(q0 - q1 * (unsigned)g.nseg);
        return q1;
    }
    const unsigned q0 = fast_div(w, g.f_nseg);
    *xseg = (int)(w - q0 * (unsigned)g.nseg);
    const unsigned q1 = fast_div(q0, g.f_V);
    *v = (int)(q0 - q1 * (unsigned)g.V);
    return q1;
}

// Wave coordinates of workgroup b of nb (b, nb: the workgroup's place in its own run) for a
// traversal with one y row per wave (g.rows == 1).
// Traversal order, slowest to fastest: y-tile, z, row in tile, unknown v, x-segment (or x-segment,
// unknown v: split_v_seg).
__device__ __forceinline__ WaveCoord decode(const Geometry &g, int b, int nb)
{
    WaveCoord wc;
    const unsigned wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned w = (unsigned)(g.xcd_remap ? logical_block(b, nb) : b) * 4u + wid;
    int xseg;
    const unsigned q1 = split_v_seg(g, w, &wc.v, &xseg);
    const unsigned q2 = fast_div(q1, g.f_ty);
    const int yin = (int)(q1 - q2 * (unsigned)g.ty);
    const unsigned q3 = fast_div(q2, g.f_nz);
    wc.z = (int)(q2 - q3 * (unsigned)g.nz);
    wc.y = (int)q3 * g.ty + yin;
    wc.valid = ((long long)w < g.waves) && (wc.y < g.ny);
    wc.x = (xseg * 64 + (int)(threadIdx.x & 63)) * g.xpl;
    return wc;
}

__device__ __forceinline__ WaveCoord decode(const Geometry &g) { return decode(g, blockIdx.x, gridDim.x); }

// Lane coordinates of a traversal with row groups of g.rows y rows (x pairs on lines that one
// row per wave would leave partly idle): the group's rows * pairs pair slots, row-major, fill
// g.nseg waves; lane l of segment s takes slot p = 64 s + l, row r = p / pairs of the group
// and the pair at x = 2 (p - r pairs).  Order: tile, z, row group, v, segment, as in decode().
// valid is wave-uniform; lanes with r >= rows or a row past ny are the caller's to mask.
struct RowCoord {
    int x, v, y0, r, z;
    bool valid;
};

__device__ __forceinline__ RowCoord decode_rows(const Geometry &g, int b, int nb)
{
    RowCoord rc;
    const unsigned wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned w = (unsigned)(g.xcd_remap ? logical_block(b, nb) : b) * 4u + wid;
    int xseg;
    const unsigned q1 = split_v_seg(g, w, &rc.v, &xseg);
    const unsigned q2 = fast_div(q1, g.f_ty);
    const int yg = (int)(q1 - q2 * (unsigned)g.tyr);
    const unsigned q3 = fast_div(q2, g.f_nz);
    rc.z = (int)(q2 - q3 * (unsigned)g.nz);
    rc.y0 = (int)q3 * g.ty + yg * g.rows;
    rc.valid = ((long long)w < g.waves) && (rc.y0 < g.ny);
    const unsigned p = (unsigned)(xseg * 64) + (threadIdx.x & 63);
    const unsigned r = fast_div(p, g.f_pairs);
    rc.r = (int)r;
    rc.x = 2 * (int)(p - r * (unsigned)g.pairs);
    return rc;
}

// First element of the x-line (z, y, v, c) in either layout; offsets within a z plane are int.
__device__ __forceinline__ long long line_index(const Geometry &g, int c, int z, int y, int v)
{
    return (long long)z * g.zs + (y * (int)g.ys + v * (int)g.vs + c * (int)g.cs);
}

// Elements between the same x of consecutive components of one line.
__device__ __forceinline__ long long comp_stride(const Geometry &g) { return g.cs; }

// ------------------------------------------------------------------ local Hamiltonian --

// (W . f) psi for spin-1 (m = +1, 0, -1): Spin1GpeBCC.cu:890-895 expanded form.
template <typename R>
__device__ __forceinline__ void spin1_dot(R wx, R wy, R wz, const R *pr, const R *pi, R *orr, R *oi)
{
    const R s = (R)0.70710678118654752440;
    const R pre = s * wx, pim = s * wy;  // W+ / sqrt2 = (pre, pim); W- / sqrt2 = (pre, -pim)
    // m = +1: wz psi1 + W-/sqrt2 psi0
    orr[0] = wz * pr[0] + pre * pr[1] + pim * pi[1];
    oi[0] = wz * pi[0] + pre * pi[1] - pim * pr[1];
    // m = 0: W+/sqrt2 psi1 + W-/sqrt2 psi-1
    orr[1] = pre * pr[0] - pim * pi[0] + pre * pr[2] + pim * pi[2];
    oi[1] = pre * pi[0] + pim * pr[0] + pre * pi[2] - pim * pr[2];
    // m = -1: W+/sqrt2 psi0 - wz psi-1
    orr[2] = pre * pr[1] - pim * pi[1] - wz * pr[2];
    oi[2] = pre * pi[1] + pim * pr[1] - wz * pi[2];
}

// (W . f) psi for spin-2 (m = 2..-2): ladder factors 2, sqrt6, sqrt6, 2 (Spin2GpeBCC.cu:954-970).
template <typename R>
__device__ __forceinline__ void spin2_dot(R wx, R wy, R wz, const R *pr, const R *pi, R *orr, R *oi)
{
    const R s6 = (R)2.44948974278317809820;
    const R hr = (R)0.5 * wx, hi = (R)0.5 * wy;  // W+/2 = (hr, hi), W-/2 = (hr, -hi)
    // lower(m) = W-/2 * a_m * psi_{m-1}; upper(m) = W+/2 * b_m * psi_{m+1}
    R lr, li, ur, ui;
    // m = 2
    {
        R a0 = 2 * (hr * pr[1] + hi * pi[1]), a1 = 2 * (hr * pi[1] - hi * pr[1]);
        orr[0] = 2 * wz * pr[0] + a0;
        oi[0] = 2 * wz * pi[0] + a1;
    }
    // m = 1
    lr = s6 * (hr * pr[2] + hi * pi[2]); li = s6 * (hr * pi[2] - hi * pr[2]);
    ur = 2 * (hr * pr[0] - hi * pi[0]); ui = 2 * (hr * pi[0] + hi * pr[0]);
    orr[1] = wz * pr[1] + lr + ur;
    oi[1] = wz * pi[1] + li + ui;
    // m = 0
    lr = s6 * (hr * pr[3] + hi * pi[3]); li = s6 * (hr * pi[3] - hi * pr[3]);
    ur = s6 * (hr * pr[1] - hi * pi[1]); ui = s6 * (hr * pi[1] + hi * pr[1]);
    orr[2] = lr + ur;
    oi[2] = li + ui;
    // m = -1
    lr = 2 * (hr * pr[4] + hi * pi[4]); li = 2 * (hr * pi[4] - hi * pr[4]);
    ur = s6 * (hr * pr[2] - hi * pi[2]); ui = s6 * (hr * pi[2] + hi * pr[2]);
    orr[3] = -wz * pr[3] + lr + ur;
    oi[3] = -wz * pi[3] + li + ui;
    // m = -2
    ur = 2 * (hr * pr[3] - hi * pi[3]); ui = 2 * (hr * pi[3] + hi * pr[3]);
    orr[4] = -2 * wz * pr[4] + ur;
    oi[4] = -2 * wz * pi[4] + ui;
}

template <typename R, int C>
__device__ __forceinline__ void spin_dot(R wx, R wy, R wz, const R *pr, const R *pi, R *orr, R *oi)
{
    if constexpr (C == 3)
        spin1_dot<R>(wx, wy, wz, pr, pi, orr, oi);
    else
        spin2_dot<R>(wx, wy, wz, pr, pi, orr, oi);
}

struct LocalCoef {
    double vtrap;
    double Bx, By, Bz;
    double Qx, Qy, Qz;
};

// hr/hi += local part of H psi at one unknown.
template <typename R, int C>
__device__ __forceinline__ void add_local_h(const R *pr, const R *pi, R vtrap, R Bx, R By, R Bz, R Qx, R Qy, R Qz,
                                            R c0, R c1, R c2, R alpha, R quad_sign, R q, int quad, int loss,
                                            R *hr, R *hi)
{
    R n = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) n += pr[c] * pr[c] + pi[c] * pi[c];
    const R d = vtrap + c0 * n;
    const R l = loss ? alpha * n * n : (R)0;
    // (d - i l) psi
#pragma unroll
    for (int c = 0; c < C; ++c) {
        hr[c] += d * pr[c] + l * pi[c];
        hi[c] += d * pi[c] - l * pr[c];
    }
    if constexpr (C > 1) {
        R Fx, Fy, Fz, wx, wy, wz;
        if constexpr (C == 3) {
            // F+ = sqrt2 (conj(psi1) psi0 + conj(psi0) psi-1), Fz = |psi1|^2 - |psi-1|^2
            const R s2 = (R)1.41421356237309504880;
            Fx = s2 * (pr[0] * pr[1] + pi[0] * pi[1] + pr[1] * pr[2] + pi[1] * pi[2]);
            Fy = s2 * (pr[0] * pi[1] - pi[0] * pr[1] + pr[1] * pi[2] - pi[1] * pr[2]);
            Fz = pr[0] * pr[0] + pi[0] * pi[0] - pr[2] * pr[2] - pi[2] * pi[2];
            wx = Bx + c1 * Fx;
            wy = By + c1 * Fy;
            wz = Bz + c1 * Fz;
        } else {
            const R s6 = (R)2.44948974278317809820;
            // F+ = 2 c(psi2) psi1 + s6 c(psi1) psi0 + s6 c(psi0) psi-1 + 2 c(psi-1) psi-2
            Fx = 2 * (pr[0] * pr[1] + pi[0] * pi[1]) + s6 * (pr[1] * pr[2] + pi[1] * pi[2]) +
                 s6 * (pr[2] * pr[3] + pi[2] * pi[3]) + 2 * (pr[3] * pr[4] + pi[3] * pi[4]);
            Fy = 2 * (pr[0] * pi[1] - pi[0] * pr[1]) + s6 * (pr[1] * pi[2] - pi[1] * pr[2]) +
                 s6 * (pr[2] * pi[3] - pi[2] * pr[3]) + 2 * (pr[3] * pi[4] - pi[3] * pr[4]);
            Fz = 2 * (pr[0] * pr[0] + pi[0] * pi[0]) + pr[1] * pr[1] + pi[1] * pi[1] - pr[3] * pr[3] -
                 pi[3] * pi[3] - 2 * (pr[4] * pr[4] + pi[4] * pi[4]);
            // spin-2 linear Zeeman enters with -B (Spin2GpeBCC.cu:942-957)
            wx = c1 * Fx - Bx;
            wy = c1 * Fy - By;
            wz = c1 * Fz - Bz;
        }
        R tr[C], ti[C];
        spin_dot<R, C>(wx, wy, wz, pr, pi, tr, ti);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            hr[c] += tr[c];
            hi[c] += ti[c];
        }
        if constexpr (C == 5) {
            // singlet: H_m += (c2/5) Theta (-1)^m conj(psi_-m), Theta = 2 p2 p-2 - 2 p1 p-1 + p0^2
            const R thr = 2 * (pr[0] * pr[4] - pi[0] * pi[4]) - 2 * (pr[1] * pr[3] - pi[1] * pi[3]) +
                          (pr[2] * pr[2] - pi[2] * pi[2]);
            const R thi = 2 * (pr[0] * pi[4] + pi[0] * pr[4]) - 2 * (pr[1] * pi[3] + pi[1] * pr[3]) +
                          2 * pr[2] * pi[2];
            const R k = c2 * (R)0.2;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const R sg = (c & 1) ? -k : k;
                const R cr = pr[C - 1 - c], ci = -pi[C - 1 - c];
                hr[c] += sg * (thr * cr - thi * ci);
                hi[c] += sg * (thr * ci + thi * cr);
            }
        }
        if (quad) {
            R ur[C], ui[C];
            spin_dot<R, C>(Qx, Qy, Qz, pr, pi, tr, ti);
            spin_dot<R, C>(Qx, Qy, Qz, tr, ti, ur, ui);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                hr[c] += quad_sign * ur[c];
                hi[c] += quad_sign * ui[c];
            }
        }
        constexpr int F = (C - 1) / 2;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const R mm = (R)((F - c) * (F - c));
            hr[c] += q * mm * pr[c];
            hi[c] += q * mm * pi[c];
        }
    }
}

template <typename R>
__device__ __forceinline__ const R *table_w(const Tables &t);
template <>
__device__ __forceinline__ const float *table_w<float>(const Tables &t) { return t.w32; }
template <>
__device__ __forceinline__ const double *table_w<double>(const Tables &t) { return t.w64; }
template <typename R>
__device__ __forceinline__ const R *table_wsum(const Tables &t);
template <>
__device__ __forceinline__ const float *table_wsum<float>(const Tables &t) { return t.wsum32; }
template <>
__device__ __forceinline__ const double *table_wsum<double>(const Tables &t) { return t.wsum64; }
template <typename R>
__device__ __forceinline__ const R *table_pos(const Tables &t);
template <>
__device__ __forceinline__ const float *table_pos<float>(const Tables &t) { return t.pos32; }
template <>
__device__ __forceinline__ const double *table_pos<double>(const Tables &t) { return t.pos64; }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// N wave sums at once: each butterfly step exchanges all N values before any is added, so a
// wave waits on N lane exchanges once per step instead of once per step and value.
template <int N>
__device__ __forceinline__ void wave_sum_n(double *v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o[N];
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] = __shfl_xor(v[i], off, 64);
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += o[i];
    }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace gpe
