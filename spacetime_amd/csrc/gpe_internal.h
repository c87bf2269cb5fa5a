// Internal interfaces between the host runtime (gpe_host.cpp) and the HIP kernels
// (gpe_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gpe.h"

namespace gpe {

// Division of a 32-bit unsigned n by an invariant d >= 1 as a multiply-high and shifts
// (Granlund & Montgomery 1994; Warren, Hacker's Delight, section 10-8): with l = ceil(log2 d)
// and m = floor(2^32 (2^l - d) / d) + 1, n / d = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(n, m),
// exact for every n < 2^32.  d = 1 is passed through.
struct FastDiv {
    unsigned d, m, l;
};

// Device layout of one time level, in one of two orders (Geometry::slabs):
//   lines (0):  element(z, y, v, c, x) = ((((z * ny + y) * V + v) * C + c) * lx + x
//   slabs (1):  element(z, v, c, y, x) = ((z * V + v) * C + c) * ls + y * lx + x
// x-lines (FP64, and FP32 lines of a multiple of 16 values once rounded up to even): lx >= nx
// rounded up so every x-line starts on a 128-byte boundary (or, where that would pad a line by
// more than 1/20, lx = nx rounded up to even).  Slabs (the other FP32 lines): lx = nx rounded
// up to even, ls = ny lx rounded up to 128 bytes.  Both are addressed through the strides ys,
// vs, cs, zs (line_index).
// Lines: the C components of a line are adjacent (C consecutive x-lines), so a wave's own,
// previous and next values are each one contiguous run of C line segments, which streams
// ~12% faster from HBM than C separate component planes (measured with the sweep's
// read-2-write-1 pattern).  Slabs:
// the rows of one (z, v, c) follow each other, so x pairs spread over a tile of rows are one
// contiguous run that starts on a 128-byte boundary when the tile's rows hold a multiple of
// 16 values -- lines of 82 or 58 values would otherwise start inside a cache line and every
// quarter-wave access would touch three lines instead of two.  One wave handles 64
// consecutive x (or x pairs) of one (z, y, v) line or of a group of rows, so the unknown v
// and the face table row are uniform across the wave.  Out-of-box neighbours read as zero
// (Dirichlet, as the reference's never-written halo).
struct Geometry {
    int nx, ny, nz, V, deg;
    int C;             // complex components (1, 3, 5)
    FastDiv f_nseg, f_V, f_ty, f_nz;  // the decode's divisors (f_ty divides by tyr)
    FastDiv f_pairs;   // divides a lane's pair slot by `pairs` (rows > 1)
    int xpl;           // consecutive x values per lane (1, or 2 for the FP32 sweep)
    int rows;          // y rows of a row group (1; 2-8, or the tile's rows in the slab layout
                       // when they start on 128-byte boundaries: x pairs spread over the rows)
    int pairs;         // x pairs per row: ceil(nx / 2)
    int nseg;          // waves per x-line (rows == 1: ceil(nx / (64 xpl))) or per row group
                       // (rows > 1: ceil(rows * pairs / 64))
    int ty;            // y rows per tile of the traversal order
    int tyr;           // row groups per tile: ty / rows
    int ntile;         // ceil(ny / ty)
    long long lx;      // elements per x-line
    int slabs;         // layout: 0 x-lines in (z, y, v, c) order, 1 component slabs of rows
    long long ys, vs, cs, zs;  // elements between rows, unknowns, components and z planes
                               // (4 zs < 2^31: offsets within a plane and its neighbours are int)
    long long waves;   // total waves of a sweep
    int blocks;        // workgroups of a sweep (4 waves each)
    int xcd_remap;     // give each XCD a contiguous range of the traversal order
    int v_inner;       // 1: the unknown v varies faster than the x segment (lines of several waves)
};

// x values per lane of the sweep kernel.  Two (one 16-byte load per component and neighbour,
// the access width of the FP64 form) exist for FP32; which one runs is a traversal choice
// (default_xpl, then the autotuner).
__host__ __device__ constexpr bool pair_lanes_ok(int fp64, int deg, int ncomp)
{
    return !fp64 && deg >= 1 && ncomp >= 1;
}

// Sweep body of a traversal: 1 one x per lane, 2 x pairs, 3 x pairs over a group of 2-8 y rows.
__host__ __device__ constexpr int lane_form(int xpl, int rows) { return xpl == 1 ? 1 : (rows > 1 ? 3 : 2); }

// Default: pairs in FP32.  Measured on spin-1 / spin-2 BCC, FCC and Cube boxes of 96-192 blocks,
// pairs run 30-40% faster than single values, also at degree 8 with 5 components where they cut
// occupancy to 2 waves per SIMD; lines of <= 64 values share a wave between 2 or 4 y rows, so
// pairs idle no more lanes than single values do.
__host__ __device__ constexpr int default_xpl(int fp64, int deg, int ncomp, int nx)
{
    return (pair_lanes_ok(fp64, deg, ncomp) && nx >= 1) ? 2 : 1;
}

// Per-context constant tables on the device (weights already scaled by -1/2 / s^2).
struct Tables {
    const int *face;       // V*deg packed faces
    const float *w32;      // V*deg
    const double *w64;     // V*deg
    const float *wsum32;   // V: sum of the weights of unknown v
    const double *wsum64;  // V
    const float *pos32;    // V*3: s * position of unknown v inside its block
    const double *pos64;   // V*3
    const double *dv;      // V: s^3 * dual volume of unknown v
    const int *nface;      // V: faces of unknown v before its zero-weight pad faces
    const int *facex;      // V*deg pairs (packed face, elements from unknown v's own line to the
                           // face's neighbour line: dz zs + dy ys + (t - v) vs, y / z inside)
};

// FP32 copies of the physics constants, converted once on the host for the FP32 sweep.
struct PhysicsF {
    float cell[3], p0[3], trap2[3], trap_offset, center[3];
    float c0, c1, c2, alpha, quad_sign, q;
};

struct Physics {
    double cell[3];        // s * block edges (x, y, z)
    double p0[3];
    double trap[3];        // trap frequencies
    double trap2[3];       // squared trap frequencies
    double trap_offset, trap_off_time;
    double c0, c1, c2, alpha, quad_sign, q;
    double center[3];
    PhysicsF f;            // the same in FP32
};

// Field and time of one sweep (computed on the host from the ramp).
struct FieldRec {
    double bq, bb[3];
    double bqq, bbq[3];
    double t;
    double sigma;         // block-scale factor of the expanding grid (1: static)
    // FP32 copies for the FP32 sweep: bq, bb[3], bqq, bbq[3], sigma, 1 / sigma^2
    float fbq, fbb[3], fbqq, fbbq[3], fsigma, fksc, pad[2];
};

enum Mode { MODE_UPDATE = 0, MODE_EULER = 1, MODE_ITP = 2, MODE_HPSI = 3 };

struct SweepArgs {
    Geometry g;
    Tables tb;
    Physics ph;
    const void *cur;          // level read by the stencil (psi_k)
    void *next;               // level written (psi_{k-1} -> psi_{k+1})
    const void *prev;         // update: psi_{k-1} read from here instead of next (may be null)
    const FieldRec *fields;   // field of sweep k at fields[k]
    int sweep;                // index into fields
    int order;                // 1: the workgroups take their tiles in reverse order (sweep_order)
    int quad, loss;           // quadratic Zeeman / three-body loss active
    double coef;              // update: next += coef * (-i) H cur ; ITP: dt
    const double *scale_in;   // ITP: scale applied to cur before H (device scalar)
    double *partials;         // ITP: one FP64 partial norm per workgroup
};

// Observables partial sums per workgroup.
enum {
    OBS_NORM = 0, OBS_ENERGY, OBS_EXPECT_H, OBS_MX, OBS_MY, OBS_MZ, OBS_MAXDENS, OBS_NONFINITE,
    OBS_SPIN, OBS_RX, OBS_RY, OBS_RZ, OBS_REF_SPIN, OBS_REF_MZ, OBS_COUNT
};
enum { FIELD_DENSITY = 0, FIELD_SPIN = 1, FIELD_SPIN_NORM = 2, FIELD_UV_THETA = 3 };
hipError_t launch_local_fields(int fp64, int ncomp, const Geometry &g, const void *psi, int kind, double *out,
                               hipStream_t s);

struct ObsArgs {
    Geometry g;
    Tables tb;
    Physics ph;
    const void *psi;
    FieldRec field;
    int quad;
    double *partials;         // OBS_COUNT per workgroup
};

// Per-unknown bound of the local Hamiltonian (maxHamilton) and the weighted distance to a
// reference state (weightedDiff), reduced per workgroup: partials[2 b] = max, [2 b + 1] = sum.
enum { AUX_MAXHAM = 0, AUX_WDIFF, AUX_COUNT };

struct AuxArgs {
    Geometry g;
    Tables tb;
    Physics ph;
    const void *psi;
    const double2 *ref;       // reference state, complex128 ABI layout (null: no distance)
    double phase_re, phase_im; // the reference is multiplied by this phase
    FieldRec field;
    double *partials;         // AUX_COUNT per workgroup
};

// Adjoint sweep: mu2 <- mu2 + coef G'(psi)^T mu1, plus per-workgroup parameter sensitivities.
enum { ADJ_BX = 0, ADJ_BY, ADJ_BZ, ADJ_BQ, ADJ_WX, ADJ_WY, ADJ_WZ, ADJ_COUNT };

struct AdjArgs {
    Geometry g;
    Tables tb;
    Physics ph;
    const void *psi;          // psi_k
    const void *mu1;          // mu_{k+1} (stencil operand)
    void *mu2;                // mu_{k+2} in, mu_k out
    const FieldRec *fields;
    int sweep;
    int order;                // 1: reverse workgroup order (sweep_order)
    int quad;
    double coef;
    double *partials;         // ADJ_COUNT per workgroup
};

enum { OBJ_PARTIALS = 0, OBJ_OVERLAP = 1, OBJ_ENERGY = 2, OBJ_MAGZ = 3, OBJ_EXPLICIT = 4 };

struct ObjArgs {
    Geometry g;
    Tables tb;
    Physics ph;
    FieldRec field;           // field at the time of psi (OBJ_EXPLICIT)
    int mode;
    const void *psi;
    const void *target;
    void *grad;
    const double *scalars;    // OBJ_OVERLAP: (Re, Im) of <target|psi>
    double *partials;         // OBJ_PARTIALS: 2 per workgroup; OBJ_EXPLICIT: 4 per workgroup
};

hipError_t launch_adjoint(int fp64, int ncomp, const AdjArgs &a, hipStream_t s);
hipError_t launch_objective(int fp64, int ncomp, const ObjArgs &a, hipStream_t s);

// Ensemble: one launch advances every member by one sweep.
struct EnsembleMember {
    SweepArgs args[2];        // parity 0: cur = level 0, next = level 1; parity 1: swapped
    int begin;                // first workgroup (multiple of 8)
    int blocks;               // workgroups of this member
};

// Members per launch: their first workgroups travel in the kernel arguments, so the
// workgroup -> member search reads the kernel-argument segment, not the member records.
constexpr int ENS_GROUP_MAX = 128;
// A member of at least this many workgroups fills the GPU by itself (16 per CU) and gets a
// launch of its own: consecutive sweeps of one launch then meet their own lines in the
// memory-side cache, which a batch shares among all its members.
constexpr int ENS_SOLO_BLOCKS = 4096;

struct EnsembleArgs {
    const EnsembleMember *members;
    int nmember;
    int parity;
    int sweep;                // sweep index within the uploaded field tables
    const int *sweep_base;    // device counter added to sweep (graph replays), may be null
    int begin[ENS_GROUP_MAX]; // members[i].begin
};

// All members of one launch share the component count, precision, face degree and lanes.
// form: lane_form of the group's traversal
hipError_t launch_ensemble_sweep(int fp64, int ncomp, int deg, int form, const EnsembleArgs &e, int blocks,
                                 hipStream_t s);
hipError_t launch_advance_base(int *base, int by, hipStream_t s);
hipError_t launch_set_base(int *base, int value, hipStream_t s);

// kind 0: b = a; kind 1: b = b + a / 2; n16 16-byte elements.
hipError_t launch_stream(int kind, const void *a, void *b, long long n16, hipStream_t s);

// Device-side Thomas-Fermi fill with spinor weights and counter-based noise.
struct FillArgs {
    double mu, g, amp;
    unsigned long long seed;
    double w[10];             // complex weight of each component
};
hipError_t launch_fill_tf(int fp64, const Geometry &g, const Tables &tb, const Physics &ph, int ncomp, const FillArgs &f,
                          void *psi, hipStream_t s);

// launchers (gpe_kernels.hip)
hipError_t launch_sweep(int fp64, int ncomp, int mode, const SweepArgs &a, hipStream_t s);
hipError_t launch_aux(int fp64, int ncomp, const AuxArgs &a, hipStream_t s);
hipError_t launch_observables(int fp64, int ncomp, const ObsArgs &a, hipStream_t s);
// sum `n` partial records of `width` doubles into out[width] (max for column maxcol, if >= 0)
// scratch: FIN_SCRATCH doubles of device memory owned by the caller's context.
constexpr int FIN_WIDTH_MAX = 16;
constexpr int FIN_SCRATCH = 256 * FIN_WIDTH_MAX;
// column sums of nslot tables of n records of width doubles (table s at partials + s n width),
// table s's sums to out + s out_stride; one workgroup per table
hipError_t launch_sum_tables(const double *partials, int nslot, int n, int width, double *out, long long out_stride,
                             hipStream_t s);
hipError_t launch_finalize(const double *partials, int n, int width, int maxcol, double *out, double *scratch,
                           hipStream_t s);
// out = sqrt(target / in[0])
hipError_t launch_norm_to_scale(const double *in, double target, double *out, hipStream_t s);
// level *= factor (device scalar if dfactor != nullptr)
hipError_t launch_scale(int fp64, void *psi, long long nelem, double factor, const double *dfactor,
                        hipStream_t s);
// ABI complex128 [c][z][y][x][v] <-> device layout
hipError_t launch_pack(int fp64, const Geometry &g, int ncomp, const double *abi, void *dev, hipStream_t s);
hipError_t launch_unpack(int fp64, const Geometry &g, int ncomp, const void *dev, double *abi, hipStream_t s);
// dst = phase * src (complex), for the staggered start
hipError_t launch_phase_copy(int fp64, const void *src, void *dst, long long nelem, double pre, double pim,
                             hipStream_t s);
// What other source files of the library see of a context (gpe_host.cpp owns the struct).
struct CtxView {
    Geometry g;          // one x per lane: the layout of a level
    int fp64, ncomp, device;
    hipStream_t stream;
    void *level[2];
};
int ctx_view(gpe_ctx *c, CtxView *out);
int ctx_fail(gpe_ctx *c, int code, const char *what);

// spinor phase projection (GPE_PHASE_*) and spinor noise of a level, in place
hipError_t launch_project(int fp64, int ncomp, const Geometry &g, void *psi, int phase, double param, hipStream_t s);
hipError_t launch_noise(int fp64, int ncomp, const Geometry &g, void *psi, double amp, unsigned long long seed,
                        hipStream_t s);

}  // namespace gpe
