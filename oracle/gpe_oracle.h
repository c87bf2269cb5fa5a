/*
 * FP64 CPU oracle for the DEC Gross-Pitaevskii step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is compared
 * against; it is linked or called only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in spacetime_amd/ may use it.
 *
 * It restates, in plain single-threaded (optionally OpenMP) C, the discrete update of
 * the reference solvers:
 *   - scalar staggered leapfrog      GrossPitaevskii/GPBlocks.cpp:560-606
 *   - spin-1 3-level leapfrog        Spin1GpeBCC/Spin1GpeBCC.cu:793-917 (start :667-791)
 *   - spin-2 3-level leapfrog        Spin2GpeBCC/Spin2GpeBCC.cu:845-990 (start :698-843)
 *   - imaginary-time step + renorm   GrossPitaevskiiGpuBcc/GPUGrossPitaevskii_BCC.cu:287-325,431-443
 *                                    Spin1GpeBCC/Spin1GpeBCC.cu:554-662,955-967
 *   - reductions                     Spin1GpeBCC/Spin1GpeBCC.cu:191-461
 *   - axisymmetric vortex relaxation GrossPitaevskii/VortexState.cpp:12-238
 * The spinor terms are written in the general spin-F form (F = 0, 1, 2) with the
 * ladder matrices sqrt(F(F+1) - m(m+-1)); that equals the reference's hand-expanded
 * matrices (checked in tests against dense textbook spin matrices).
 *
 * State layout (the C-ABI layout): for component c (m = F, F-1, ..., -F), cell
 * (x, y, z) and unknown v of the block, the complex value sits at
 *     psi[2 * (c * ncell * V + ((z * ny + y) * nx + x) * V + v) + {0 re, 1 im}]
 * Values outside the box are zero (Dirichlet), as the reference's never-written halo.
 *
 * The tests state, per function, what pins it (tests/test_oracle.py, tests/test_tables.py).
 */
#ifndef GPE_ORACLE_H
#define GPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int V;                     /* unknowns per cubic block                        */
    int nfaces;                /* total faces in the block                        */
    const int *face_start;     /* [V+1] first face of each unknown                */
    const int *face_d;         /* [nfaces*3] neighbour block offset dx, dy, dz    */
    const int *face_tgt;       /* [nfaces] target unknown in that block           */
    const double *face_hodge;  /* [nfaces] Hodge ratio, table units               */
    const double *pos;         /* [V*3] unknown positions in the block, table units */
    const double *vol;         /* [V] dual volume of each unknown, table units    */
    double block_width[3];     /* block edges (x, y, z), table units: cubic but Z  */
} orc_lattice;

typedef struct {
    int ncomp;                 /* 1, 3 or 5                                        */
    int nx, ny, nz;            /* blocks per axis                                  */
    double block_scale;        /* physical length of one table unit               */
    double p0[3];              /* physical position of block (0,0,0) origin        */
    double trap[3];            /* trap frequencies: V = 0.5 * sum (w_a r_a)^2      */
    double trap_offset;        /* constant added to V (reference uses 100)         */
    double trap_off_time;      /* V = 0 for t >= trap_off_time                     */
    double c0;                 /* density interaction (g for the scalar)          */
    double c1;                 /* spin exchange: spin-1 "c2", spin-2 "c2" (ref)    */
    double c2;                 /* spin-2 singlet: reference "c4"                   */
    double alpha;              /* three-body loss rate: H += -i alpha n^2          */
    double bq;                 /* quadrupole gradient of the linear-Zeeman field  */
    double bb[3];              /* uniform bias of the linear-Zeeman field          */
    double bq_quad;            /* quadrupole gradient of the quadratic field       */
    double bb_quad[3];         /* bias of the quadratic field                      */
    double quad_sign;          /* H += quad_sign (Bquad . f)^2                     */
    double q;                  /* constant quadratic Zeeman: H += q f_z^2          */
    double center[3];          /* quadrupole centre                                */
} orc_params;

/* Piecewise field ramp, restating getSignal (Spin1GpeBCC/AliceRingRamps.h:74-140,
 * Spin2GpeBCC/KnotRamps.h:49-116).  Two channels: quadrupole gradient Bq (scalar) and
 * bias Bb (3-vector), each a list of segments (value, duration, type).  Segment types:
 * 0 constant, 1 linear from the previous value, 2 sqrt(t/T) "fast extraction" (spin-1),
 * 3 1-exp(-t/tau) "fast extraction" (spin-2).  Past the last segment the last value
 * is held (the reference reads past its arrays there).  The field used by H is
 *   B(r) = scale_q * Bq * (x-cx, y-cy, -2(z-cz)) + scale_b * Bb
 *   Bquad(r) = qscale_q * Bq * (...) + qscale_b * Bb
 * with ramp time = time_scale * simulation time (Spin1GpeBCC.cu:1577-1587). */
typedef struct {
    int nq;  const double *q_val;  const double *q_dur;  const int *q_type;
    int nb;  const double *b_val;  const double *b_dur;  const int *b_type;   /* b_val: nb*3 */
    double tau;                /* time constant of type 3 */
    double time_scale;
    double scale_q, scale_b, qscale_q, qscale_b;
} orc_ramp;

/* Evaluate the ramp signal at simulation time t. */
void orc_ramp_signal(const orc_ramp *r, double t, double *bq, double bb[3]);
/* Set p->bq, p->bb, p->bq_quad, p->bb_quad from the ramp at simulation time t. */
void orc_ramp_apply(const orc_ramp *r, double t, orc_params *p);

/* Local (non-kinetic) Hamiltonian at one unknown at physical position r, time t. */
void orc_local_h(const orc_params *p, const double r[3], double t, const double *psi_node,
                 double *h_node);

/* H psi for the whole grid (kinetic + local). psi/hpsi in the ABI layout. */
void orc_apply_h(const orc_lattice *L, const orc_params *p, double t, const double *psi, double *hpsi);

/* next += -i * coef * H(cur)    (coef = dt: staggered half-sweep, 2 dt: 3-level)   */
void orc_update(const orc_lattice *L, const orc_params *p, double t, double coef,
                double *next, const double *cur);

/* next = cur - i * dt * H(cur)   (forward-Euler start of the 3-level scheme)        */
void orc_forward_euler(const orc_lattice *L, const orc_params *p, double t, double dt,
                       double *next, const double *cur);

/* psi <- (psi - dt H psi) / ||psi - dt H psi||   ; returns the norm before scaling  */
double orc_itp_step(const orc_lattice *L, const orc_params *p, double dt, double *psi,
                    double *scratch);

/* Reductions (FP64, per-unknown dual volume dv_v = s^3 vol[v]) */
double orc_norm(const orc_lattice *L, const orc_params *p, const double *psi);
void orc_normalize(const orc_lattice *L, const orc_params *p, double *psi, double target);
double orc_energy(const orc_lattice *L, const orc_params *p, double t, const double *psi);
double orc_expect_h(const orc_lattice *L, const orc_params *p, double t, const double *psi,
                    double *scratch);
void orc_magnetization(const orc_lattice *L, const orc_params *p, const double *psi, double out[3]);
/* max over unknowns of the local Hamiltonian bound (maxHamilton); NaN for spin-2. */
double orc_max_hamiltonian(const orc_lattice *L, const orc_params *p, double t, const double *psi);
/* sum dv |a|^2 |psi - a|^2, a = (phase_re + i phase_im) ref (weightedDiff + integrate). */
double orc_weighted_diff(const orc_lattice *L, const orc_params *p, const double *psi, const double *ref,
                         double phase_re, double phase_im);

/* Position of unknown v of block (x, y, z). */
void orc_position(const orc_lattice *L, const orc_params *p, int x, int y, int z, int v, double r[3]);

/* Run n sweeps of the 3-level scheme: a = psi_{k-1}, b = psi_k on entry; on exit a and b
 * hold psi_{k+n-1}, psi_{k+n} (swapped in place as needed).  Time of sweep j is t0 + j dt. */
void orc_run_leapfrog3(const orc_lattice *L, const orc_params *p, const orc_ramp *ramp,
                       double t0, double dt, int n, double *a, double *b);

/* Staggered scheme (GPBlocks.cpp:560-606): n steps, each = odd += -i dt H(even),
 * even += -i dt H(odd). */
void orc_run_staggered(const orc_lattice *L, const orc_params *p, double dt, int n,
                       double *even, double *odd);

/* Axisymmetric vortex relaxation, restating VortexState.cpp:12-238.
 * f has rsize*zsize entries (r-major).  Returns mu. */
void orc_vortex_init(int rsize, double rstep, double rplug, int zsize, double zstep,
                     double *f, double *mu);
double orc_vortex_iterate(int kappa, double g, int rsize, double rstep, double rplug, int zsize,
                          double zstep, double ratio, double *f, double *mu, int maxiters, double eps);
double orc_vortex_function(const double *f, int rsize, double rstep, double rplug, int zsize,
                           double zstep, double r, double z);

/* MT19937 uniform-disk sample as Types/Random.cpp:77-86 (state = 625 uint32). */
void orc_mt_seed(unsigned int *state, unsigned int seed);
void orc_mt_uniform_circle(unsigned int *state, double out[2]);

/* Restatement of GPBlocks.cpp integrateInTime (BCC, staggered scheme) for a vortex state f.
 * Writes iterations+1 values of normsq and overlap error, the dt and steps per iteration. */
int orc_gpblocks_integrate(const orc_lattice *L, int kappa, double g, const double *f, int rsize, double rstep,
                           double rplug, int zsize, double zstep, double mu, double block_scale, const double minp[3],
                           const double maxp[3], unsigned period, unsigned iterations, double *out_normsq,
                           double *out_error, double *out_dt, int *out_steps);

/* Discrete adjoint of the 3-level leapfrog, one reverse sweep (see gpe_oracle.c):
 * mu2 += coef G'(psi)^T mu1 and the seven parameter sensitivities at this sweep. */
void orc_adjoint_sweep(const orc_lattice *L, const orc_params *p, double t, double coef, const double *psi,
                       const double *mu1, double *mu2, double sens[7]);

int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
