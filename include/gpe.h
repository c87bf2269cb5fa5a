/*
 * gpe.h -- C ABI of the MI355X DEC Gross-Pitaevskii stepper (libgpe.so).
 *
 * The reference has no library boundary: each solver is an executable whose host loop
 * (integrateInTime) allocates pitched buffers and launches one kernel per sweep.  These
 * entry points replace that loop piece by piece; each declaration names the reference
 * code it stands in for.  All calls return 0 on success or a negative GPE_E* code, never
 * exit or abort; gpe_last_error() describes the last failure of a context.
 *
 * State layout across the ABI (host or device arrays, complex interleaved re/im):
 *   psi[2 * (((c * nz + z) * ny + y) * nx + x) * V + v) + {0, 1}]
 * component c (m = F, F-1, ..., -F), block (x, y, z), unknown v of the block.  The device
 * keeps its own layout (x-fastest lines per (component, z, y, unknown), or for FP32 short
 * lines component slabs of contiguous rows; gpe_internal.h, gpe_plan_traversal); the ABI
 * layout is independent of it.
 *
 * Threading: a context owns one HIP stream; calls on one context are stream-ordered and
 * not thread-safe; distinct contexts may be driven from distinct threads.  Calls that
 * return host data (get_state, observables) synchronise the stream.
 */
#ifndef GPE_H
#define GPE_H

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define GPE_OK 0
#define GPE_EINVAL (-22)  /* bad argument (lattice id, sizes, pointers)          */
#define GPE_ENOMEM (-12)  /* device or host allocation failed                     */
#define GPE_EDEVICE (-5)  /* HIP runtime error                                    */
#define GPE_ESTATE (-71)  /* call not valid in the current state (e.g. no start)  */
#define GPE_ERANGE (-34)  /* non-finite value detected                            */

/* lattices: the unit-cell tables of the reference's mesh.h files */
#define GPE_LATTICE_CUBE 0 /* Spin1GpeCube/mesh.h            1 unknown / block  */
#define GPE_LATTICE_BCC 1  /* Spin1GpeBCC/mesh.h            12 unknowns / block */
#define GPE_LATTICE_FCC 2  /* GrossPitaevskiiGpuFcc/mesh.h  12 unknowns / block */
#define GPE_LATTICE_A15 3  /* GrossPitaevskiiGpuA15/mesh.h  46 unknowns / block */
#define GPE_LATTICE_C15 4  /* GrossPitaevskiiGpuC15/mesh.h 136 unknowns / block */
#define GPE_LATTICE_Z 5    /* GrossPitaevskiiGpuZ/mesh.h    80 unknowns / block; block edges (w, sqrt(3) w, w) */

#define GPE_FP32 0 /* FP32 state and stencil arithmetic, FP64 reductions */
#define GPE_FP64 1 /* FP64 throughout                                    */

/* time integrators */
#define GPE_SCHEME_LEAPFROG3 0 /* psi_{k+1} = psi_{k-1} - 2i dt H psi_k   (Spin1GpeBCC.cu:914-916)  */
#define GPE_SCHEME_STAGGERED 1 /* odd += -i dt H(even); even += -i dt H(odd) (GPBlocks.cpp:560-606) */

/* how the second time level is made before stepping */
#define GPE_START_EULER 0 /* psi_1 = psi_0 - i dt H psi_0 (Spin1GpeBCC.cu:788-790)            */
#define GPE_START_PHASE 1 /* psi_{-1/2} = exp(+i mu dt/2) psi_0 (GPBlocks.cpp:516-523)       */
#define GPE_START_COPY 2  /* both levels = psi_0 (GPUGrossPitaevskii_BCC.cu:589-590)        */

typedef struct gpe_ctx gpe_ctx;

/* Physical and numerical set-up of one run.  Replaces the compile-time constants and the
 * integrateInTime arguments of the reference solvers (e.g. Spin1GpeBCC.cu:34-110,1105-1112). */
typedef struct {
    int lattice;          /* GPE_LATTICE_*                                             */
    int ncomp;            /* 1 (scalar), 3 (spin-1), 5 (spin-2)                        */
    int precision;        /* GPE_FP32 / GPE_FP64                                       */
    int scheme;           /* GPE_SCHEME_*                                              */
    int nx, ny, nz;       /* blocks per axis                                           */
    double block_scale;   /* physical length of one table unit                        */
    double p0[3];         /* physical position of the origin of block (0,0,0)          */
    double dt;            /* time step (3-level: between levels; staggered: full step) */
    double trap[3];       /* V = 0.5 sum_a (trap_a r_a)^2 + trap_offset               */
    double trap_offset;   /* the reference adds 100 (Spin1GpeBCC.cu:121)              */
    double trap_off_time; /* V = 0 for t >= trap_off_time (Spin1GpeBCC.cu:114-116)     */
    double c0;            /* density coupling (g for the scalar)                       */
    double c1;            /* spin exchange (reference "c2")                            */
    double c2;            /* spin-2 singlet (reference "c4")                           */
    double alpha;         /* three-body loss: H += -i alpha n^2 (Spin1GpeBCC.cu:876-880) */
    double quad_sign;     /* H += quad_sign (Bquad . f)^2; reference +1 spin-1, -1 spin-2 */
    double q;             /* constant quadratic Zeeman H += q f_z^2                    */
    double center[3];     /* quadrupole field centre (Spin1GpeBCC.cu:124-137)          */
    /* Expanding grid after the trap is switched off (Spin1GpeBCC.cu:1579-1582, Spin2GpeBCC.cu:
     * 1523-1540, where expansionBlockScale grows by k * block_scale per ms): the block scale used
     * by the Laplacian and the positions is block_scale * sigma(t), sigma = 1 + expansion_rate *
     * max(0, t - expansion_start) (rate = k * ms per time unit); the dual volumes of the
     * integrals stay those of block_scale, as in the reference.  expansion_rate = 0: static. */
    double expansion_start;
    double expansion_rate;
} gpe_config;

/* Piecewise field ramp (getSignal, Spin1GpeBCC/AliceRingRamps.h:74-140 and
 * Spin2GpeBCC/KnotRamps.h:49-116).  Segment type: 0 constant, 1 linear from the previous
 * value, 2 sqrt(t/T), 3 1-exp(-t/tau).  Past the last segment the last value is held.
 *   B(r)     = scale_q  * Bq * (x-cx, y-cy, -2(z-cz)) + scale_b  * Bb
 *   Bquad(r) = qscale_q * Bq * (x-cx, y-cy, -2(z-cz)) + qscale_b * Bb
 * evaluated at ramp time = time_scale * t (Spin1GpeBCC.cu:1577-1587). */
typedef struct {
    int nq;
    const double *q_val, *q_dur;
    const int *q_type;
    int nb;
    const double *b_val; /* nb * 3 */
    const double *b_dur;
    const int *b_type;
    double tau, time_scale;
    double scale_q, scale_b, qscale_q, qscale_b;
} gpe_ramp;

/* FP64 observables of one time level (Spin1GpeBCC.cu:191-271,955-1020). */
typedef struct {
    double norm;           /* sum dv |psi|^2                                   */
    double energy;         /* GP energy functional E[psi]                       */
    double expect_h;       /* Re <psi|H psi>: the reference's "energy" (F5)     */
    double magnetization[3]; /* sum dv F                                       */
    double max_density;    /* max |psi|^2                                       */
    double time;           /* simulation time of this level                    */
    int finite;            /* 0 if any non-finite value was seen                */
    double spin;           /* sum dv |psi^dag F psi| (Spin1GpeBCC.cu:235-271,992-1020; the
                              reference's local spin z part is not divided by n, :258-261) */
    double moment[3];      /* sum dv |psi|^2 r: the centre of mass of a unit-norm state
                              (GPUGrossPitaevskii_BCC.cu:349-413, which copies back only x) */
    double ref_spin;       /* the reference's "Spin" (integrateSpinAndDensity, Spin1GpeBCC.cu:
                              992-1020): sum dv n |s| with s = (x, y of the spinor divided by
                              sqrt n, z = |psi_1|^2 - |psi_-1|^2 undivided) (:235-271); spin-2
                              takes the same expressions on m = 1, 0, -1 (Spin2GpeBCC.cu:214-262);
                              0 for one component; unknowns with n = 0 add 0 (the reference NaN) */
    double ref_magnetization_z; /* the reference's mag_z: sum dv n (|psi_1|^2 - |psi_-1|^2); its
                              mag_x, mag_y are magnetization[0], [1] */
} gpe_observables_t;

/* Create a context: allocates both time levels on the current HIP device and uploads
 * the lattice tables.  Replaces integrateInTime's set-up (Spin1GpeBCC.cu:1105-1172). */
int gpe_create(const gpe_config *cfg, gpe_ctx **out);
int gpe_destroy(gpe_ctx *ctx);
const char *gpe_last_error(const gpe_ctx *ctx);

/* Number of unknowns per block for a lattice id (VALUES_IN_BLOCK), or GPE_EINVAL. */
int gpe_lattice_values(int lattice);

/* Upload / download one time level: level 0 = current psi_k, level 1 = previous psi_{k-1}.
 * Host pointers, complex128, ABI layout.  Replaces cudaMemcpy3D (Spin1GpeBCC.cu:1344-1345,
 * 1617,1646).  set_state(level 0) also sets the time of the run to t. */
int gpe_set_state(gpe_ctx *ctx, int level, const double *psi, double t);
int gpe_get_state(gpe_ctx *ctx, int level, double *psi);
/* Same with device pointers (complex128, ABI layout) on the context's device. */
int gpe_set_state_device(gpe_ctx *ctx, int level, const double *dpsi, double t);
int gpe_get_state_device(gpe_ctx *ctx, int level, double *dpsi);
/* gpe_get_state that allocates nothing (no hipMalloc / hipFree, no allocating library call):
 * for callers that cannot allocate device memory, such as tight serving loops and harnesses
 * that forbid it.  No counterpart in the reference.  The level streams to the host through a
 * file-scope device array of 32 MiB (held by every GPU the library runs on until the process
 * ends) in chunks of at most that size, so any level size is accepted; each chunk is written
 * in full before it is copied.  Calls that overlap on one device share its array and race.
 * Synchronises the context's stream. */
int gpe_get_state_noalloc(gpe_ctx *ctx, int level, double *psi);

/* Static linear and quadratic Zeeman fields (MagFields Bs, Spin1GpeBCC.cu:1429-1432). */
int gpe_set_field(gpe_ctx *ctx, double bq, const double bb[3], double bq_quad, const double bb_quad[3]);
/* Time-dependent fields; replaces the per-launch getSignal (Spin1GpeBCC.cu:1583-1587). */
int gpe_set_ramp(gpe_ctx *ctx, const gpe_ramp *ramp);

/* Make the second time level from level 0 (GPE_START_*; mu used by GPE_START_PHASE).
 * GPE_START_EULER replaces forwardEuler (Spin1GpeBCC.cu:667-791, launch :1424-1434). */
int gpe_start(gpe_ctx *ctx, int mode, double mu);

/* Advance n time steps (3-level: n sweeps; staggered: 2n sweeps).  Replaces the leapfrog
 * launches of the host loop (Spin1GpeBCC.cu:1571-1603, GPUGrossPitaevskii_BCC.cu:883-889). */
int gpe_step(gpe_ctx *ctx, int nsteps);

/* n imaginary-time steps psi <- N(psi - dt_itp H psi), renormalised to norm 1 each step.
 * Works on level 0; level 1 becomes invalid (call gpe_start again).  Replaces itp +
 * normalize_h (Spin1GpeBCC.cu:554-662,955-967; GPUGrossPitaevskii_BCC.cu:287-325,431-443,845-849). */
int gpe_itp(gpe_ctx *ctx, int nsteps, double dt_itp);

/* FP64 observables of a level.  Replaces density/innerProduct/localAvgSpinAndDensity and
 * the integrate trees (Spin1GpeBCC.cu:191-271,328-392,955-1020). */
int gpe_observables(gpe_ctx *ctx, int level, gpe_observables_t *out);

/* The reference's time-series file of a run (Spin1GpeBCC.cu:1497-1504, 1659-1712): each append
 * adds one column -- the level's time times the ramp's time scale (the reference's ms clock; 1
 * without a ramp), the field's Bq and Bb x + y + z at that time, ref_spin, magnetization[0],
 * magnetization[1], ref_magnetization_z and norm -- and each save writes
 *   times = [times, t1, t2, ...];  Bq = [Bq, ...];  Bz = [Bz, ...];  Spin = [Spin, ...];
 *   mag_x = [mag_x, ...];  mag_y = [...];  mag_z = [...];  norm = [norm, ...];
 * one per line, values as "%.18f" (the reference's toString), then starts a new series, so
 * running the saved files in order concatenates them as the reference's do.  append returns
 * GPE_ERANGE (after appending) for a non-finite state. */
typedef struct gpe_series gpe_series;
int gpe_series_create(gpe_series **out);
int gpe_series_destroy(gpe_series *series);
int gpe_series_append(gpe_series *series, gpe_ctx *ctx, int level);
int gpe_series_save(gpe_series *series, const char *path);

/* Maximum over the unknowns of a level of the local Hamiltonian bound: with p = V(r) + c0 n
 * and B = B(r, t) + c1 psi^dag F psi, max(p + |Bx|/sqrt2, p + |By|/sqrt2, p + |Bz|) for spin-1,
 * p for a scalar state (the GPBlocks dt rule's max potential).  Replaces maxHamilton +
 * reduceMax (Spin1GpeBCC.cu:144-189,1022-1037; GPBlocks.cpp:488-495).  GPE_EINVAL for spin-2. */
int gpe_max_hamiltonian(gpe_ctx *ctx, int level, double *out);

/* sum over unknowns of dv |a|^2 |psi - a|^2 with a = (phase_re + i phase_im) ref: the
 * reference's error against an analytic state, weightedDiff + integrate
 * (Spin1GpeBCC.cu:436-461,1620-1633; RelativisticSpin2Gpe.cu:227).  ref: complex128, ABI
 * layout, a host array (copied to the device per call) or, for _device, device memory. */
int gpe_weighted_diff(gpe_ctx *ctx, int level, const double *ref, double phase_re, double phase_im, double *out);
int gpe_weighted_diff_device(gpe_ctx *ctx, int level, const double *dref, double phase_re, double phase_im,
                             double *out);

/* Per-unknown field maps of a level, copied to `out` in ABI order (z, y, x, v) with
 * gpe_field_width(kind) doubles per unknown.  Replaces localAvgSpinAndDensity and uvTheta
 * (Spin1GpeBCC.cu:235-326), which feed the reference's VTK/state writers.
 *   GPE_FIELD_DENSITY    |psi|^2
 *   GPE_FIELD_SPIN       psi^dag F psi / |psi|^2 (0 where the density is 0)
 *   GPE_FIELD_SPIN_NORM  |psi^dag F psi|
 *   GPE_FIELD_UV_THETA   spin-1 only: u (3), v (3), theta with a e^{i theta} = u + i v, u.v = 0,
 *                        |u| >= |v| (else swapped), theta in [0, pi)                        */
#define GPE_FIELD_DENSITY 0
#define GPE_FIELD_SPIN 1
#define GPE_FIELD_SPIN_NORM 2
#define GPE_FIELD_UV_THETA 3
int gpe_field_width(int kind);
int gpe_local_fields(gpe_ctx *ctx, int level, int kind, double *out);

/* Scale a level by a real factor (normalize, Spin1GpeBCC.cu:412-434). */
int gpe_scale(gpe_ctx *ctx, int level, double factor);

/* Wait for all work of the context (the reference's blocking copies synchronise instead). */
int gpe_sync(gpe_ctx *ctx);

/* Simulation time of level 0 and the number of sweeps taken since set_state. */
int gpe_time(const gpe_ctx *ctx, double *t, long long *sweeps);

/* Average duration in ms of the step kernel over the last gpe_step call, measured with
 * HIP events on the context's stream; bytes_per_sweep gets the algorithmic bytes of one
 * sweep (3 touches x ncomp x 2 x sizeof(real) per unknown). */
int gpe_last_step_timing(const gpe_ctx *ctx, double *ms_per_sweep, double *bytes_per_sweep);

/* ---------------------------------------------------------------------- ensembles ---- */
/* No counterpart in the reference (one executable per run there): many independent runs,
 * each an ordinary started context (own lattice, size, dt, couplings, ramp), advanced
 * together: per sweep, one kernel launch for each group of members sharing face degree and
 * lane form (a member that fills the GPU alone gets its own launch).  Members must share
 * ncomp, precision and scheme and live on the current device. */
typedef struct gpe_ensemble gpe_ensemble;

int gpe_ensemble_create(gpe_ctx *const *members, int n, gpe_ensemble **out);
int gpe_ensemble_destroy(gpe_ensemble *ens);
const char *gpe_ensemble_last_error(const gpe_ensemble *ens);
/* Advance every member by nsteps.  graph_sweeps > 0 (even, dividing the sweep count)
 * replays the launches from a captured hipGraph of graph_sweeps sweeps; 0 launches eagerly. */
int gpe_ensemble_step(gpe_ensemble *ens, int nsteps, int graph_sweeps);
/* Average ms per batched sweep of the last gpe_ensemble_step, and total unknowns. */
int gpe_ensemble_timing(const gpe_ensemble *ens, double *ms_per_sweep, double *unknowns);

/* -------------------------------------------------------------------- autotuning ---- */
/* Time the traversal candidates (y-tile rows 1..32, and 3..28 where they give row groups that
 * are not powers of two, and the whole plane of ny <= 1024 rows, x the order flags of gpe_set_traversal -- the second only where lines
 * span several waves -- x one or two x values per lane, the latter for FP32) on scratch copies
 * of the context's two levels for trial_sweeps sweeps each, and keep the fastest.  Results are cached per
 * (lattice, ncomp, precision, nx, ny, nz) for the process; pass trial_sweeps = 0 to apply a
 * cached choice only.  Outputs may be NULL.  No counterpart in the reference, which fixes its
 * 16 x 2 x 1 thread blocks at compile time (Spin1GpeBCC.cu:39-41). */
int gpe_autotune(gpe_ctx *ctx, int trial_sweeps, int *tile_rows, int *xcd_remap, int *x_per_lane,
                 double *ms_per_sweep);

/* Set / read the sweep traversal directly: y-tile rows (1..1024; ny or more: one tile of the
 * whole plane), order flags (0..3: bit 0 gives each XCD a contiguous range of the traversal,
 * bit 1 puts the unknown v inside the x segment
 * so that a workgroup's waves take neighbouring unknowns of one stretch of x; the second
 * matters only where lines span several waves), x values per lane (1, or 2 for FP32).  No
 * counterpart in the reference (fixed thread blocks, Spin1GpeBCC.cu:39-41).  Results do not
 * depend on it beyond FP32 instruction selection and the order of FP64 partial sums. */
int gpe_set_traversal(gpe_ctx *ctx, int tile_rows, int xcd_remap, int x_per_lane);
int gpe_traversal(const gpe_ctx *ctx, int *tile_rows, int *xcd_remap, int *x_per_lane);

/* The launch a sweep of a configuration would make for a traversal, computed on the host
 * without a device: workgroups of four waves, each wave a 64-lane segment of one (z, y, v)
 * x-line (rows = 1) or of a group of `rows` y rows (x pairs spread over the rows; in the slab
 * layout a tile whose rows hold a multiple of 16 values is one group), in the
 * order y-tile, z, row (group) in tile, unknown v, segment (or segment, v: flag bit 1; the
 * plan's xcd_remap holds the flags); and the constants of the wave
 * decode's divisions by invariants (d, m, l: n / d = (t + n) >> l for n < 2^31, t the high
 * 32 bits of n m).  No counterpart in the reference, whose
 * launch shape is fixed (Spin1GpeBCC.cu:1376-1380). */
typedef struct {
    int nx, ny, nz, V, C;
    int x_per_lane, rows, pairs, nseg, tile_rows, row_groups, ntile, xcd_remap;
    long long lx, waves;
    int blocks;
    unsigned div_nseg[3], div_V[3], div_rows[3], div_nz[3], div_pairs[3]; /* d, m, l */
    /* level layout: element (z, y, v, c, x) at z zs + y ys + v vs + c cs + x; slabs = 1 when
     * the rows of one (z, v, c) are contiguous (FP32 lines whose even length is not a whole
     * number of 128-byte lines; lx = that even length), 0 when the C components of one
     * (z, y, v) are adjacent x-lines */
    int slabs;
    long long ys, vs, cs, zs;
} gpe_traversal_plan;
int gpe_plan_traversal(const gpe_config *cfg, int tile_rows, int xcd_remap, int x_per_lane, gpe_traversal_plan *out);

/* Achievable HBM rate on the current device, from streaming kernels over two buffers of
 * `bytes` each (16-byte non-temporal accesses), averaged over `reps` launches: a copy
 * (2 x bytes moved) and an in-place b += a / 2 (3 x bytes: the sweep's read-2-write-1 pattern).
 * GB/s out; either output may be NULL.  No counterpart in the reference. */
int gpe_stream_bandwidth(long long bytes, int reps, double *copy_gbs, double *update_gbs);

/* ---------------------------------------------------------------- initial states ---- */
/* Stationary vortex state f(r, z) e^{i kappa phi} of the scalar GPE in the trap
 * 0.5 (r^2 + ratio^2 z^2), by the reference's axisymmetric relaxation
 * (VortexState::setRange + iterateSolution, GrossPitaevskii/VortexState.cpp:42-196).
 * The profile lives on rsize x zsize points (r-major) from rmin with steps rstep, zstep. */
typedef struct {
    int kappa;
    double g;
    double rmin, rmax, zmax, rstep, zstep;
    double ratio;
    int maxiters;
    double eps;
} gpe_vortex_params;

/* Number of profile points (rsize * zsize) the parameters give, or GPE_EINVAL. */
int gpe_vortex_size(const gpe_vortex_params *vp, int *rsize, int *zsize);
/* Relax the profile: f has rsize * zsize entries; mu receives the chemical potential. */
int gpe_vortex_relax(const gpe_vortex_params *vp, double *f, double *mu);
/* Sample psi0 = f(r, z) ((x + i y) / r)^kappa (VortexState::getPsi, VortexState.cpp:204-238) at
 * every unknown of cfg's grid into a host ABI array (component 0 of ncomp; the others 0). */
int gpe_vortex_fill(const gpe_config *cfg, const gpe_vortex_params *vp, const double *f, double *psi);

/* Fill level 0 on the device: psi_c = sqrt(max(mu - V, 0) / g) (w_c + noise * eps_c), eps_c
 * uniform in the unit disk from a counter-based hash of (seed, unknown, component), so the
 * result does not depend on the launch (w = 1: multiplicative noise; a unit spinor w: spinor
 * noise of relative size `noise`).  weights: ncomp complex values (re, im).  Sets the run time
 * to t like gpe_set_state.  For large grids where host-side states do not pay.  No counterpart
 * in the reference, whose states come from files or host loops (Spin1GpeBCC.cu:1228-1292). */
int gpe_fill_thomas_fermi(gpe_ctx *ctx, double g, double mu, const double *weights, double noise,
                          unsigned long long seed, double t);

/* ---------------------------------------------------------------- discrete adjoint ---- */
/* No counterpart in the reference: gradients of a final-state objective with respect to
 * the field-ramp segment values and the trap frequencies, by the discrete adjoint of the
 * 3-level leapfrog with checkpoints every K sweeps and recomputation (FP64 accumulation).
 * Starts from level 0 (call right after gpe_set_state); requires GPE_SCHEME_LEAPFROG3,
 * alpha = 0 and no ramp on the quadratic field (qscale_q = qscale_b = 0).  On return the
 * context holds psi_N (level 0) and psi_{N-1} (level 1). */
#define GPE_OBJ_OVERLAP 0         /* J = |<target|psi_N>|^2, <a|b> = sum dv conj(a) b */
#define GPE_OBJ_ENERGY 1          /* J = E[psi_N] (GP energy functional)              */
#define GPE_OBJ_MAGNETIZATION_Z 2 /* J = sum dv psi^dag f_z psi                         */

typedef struct {
    int nsweeps;          /* N >= 1: forward Euler + (N-1) leapfrog sweeps                */
    int checkpoint_every; /* K >= 2: a (psi_jK, psi_jK+1) pair is kept every K sweeps     */
    int objective;        /* GPE_OBJ_*                                                    */
    const double *target; /* GPE_OBJ_OVERLAP: host state, ABI layout; else NULL           */
    int fp64_trajectory;  /* FP32 context: 1 keeps psi and mu in FP64 (the run is copied to a
                             temporary FP64 context and the final levels copied back); 0 FP32 */
} gpe_adjoint_spec;

/* Precision limit.  With FP32 storage of psi and mu each sweep rounds both to 6e-8; gradients
 * that are small residues of cancelling per-sweep terms do not survive it.  Measured (spin-1
 * BCC, c0 1e4, 4000 sweeps, overlap near its maximum): the trap-frequency gradients (~6e-7,
 * per-sweep terms ~1e-2) come out 14-17% off from an FP32 context against central
 * differences, and on the FP64 oracle rounding only the stored trajectory to FP32 moves them
 * by 19-47%, rounding only mu by up to 2.7%; the ramp-knot gradients stay within 3e-5 either
 * way.  fp64_trajectory = 1 gives the FP64 result (within 1e-4 there) on an FP32 context. */

typedef struct {
    double objective;          /* J                                                      */
    double grad_trap[3];       /* dJ / d trap[a]                                         */
    double grad_static_bq;     /* dJ / d bq       (static field, gpe_set_field)          */
    double grad_static_bb[3];  /* dJ / d bb[a]    (static field)                         */
    double *grad_ramp_q;       /* caller array [nq] or NULL: dJ / d q_val[j] of the ramp  */
    double *grad_ramp_b;       /* caller array [nb*3] or NULL: dJ / d b_val[j*3+a]        */
    double forward_ms, backward_ms;
    double checkpoint_bytes, segment_bytes;
} gpe_adjoint_result;

int gpe_adjoint(gpe_ctx *ctx, const gpe_adjoint_spec *spec, gpe_adjoint_result *res);

/* Replace the trap frequencies of a context (V = 0.5 sum (trap_a r_a)^2 + offset): the
 * lambda_x,y,z of the reference's trap() (Spin1GpeBCC.cu:112-122), a parameter here so the
 * adjoint's trap gradients can be checked by finite differences. */
int gpe_set_trap(gpe_ctx *ctx, const double trap[3]);

/* ------------------------------------------------------------- context queries ---- */
/* The configuration of a context (the trap as last set by gpe_set_trap). */
int gpe_get_config(const gpe_ctx *ctx, gpe_config *out);
/* Set the run clock without touching the levels: time t0 of level 0 when `sweeps` sweeps have
 * been taken since then (gpe_time reports t0 + sweeps * sweep dt).  Restores a saved clock
 * exactly on resume; the reference keeps its clock t in the state file names instead
 * (Spin1GpeBCC.cu:1229,1690). */
int gpe_set_time(gpe_ctx *ctx, double t0, long long sweeps);
/* The clock gpe_set_time sets: t0 and the sweeps taken since (gpe_time gives t itself). */
int gpe_clock(const gpe_ctx *ctx, double *t0, long long *sweeps);

/* ------------------------------------------------------- spinor states on the device ---- */
/* Project every unknown of a level onto a spinor phase, keeping its density n = |psi|^2:
 *   spin-1 (m = 1, 0, -1)                       Spin1GpeBCC.cu:463-551
 *     GPE_PHASE_POLAR     (0, sqrt n, 0)
 *     GPE_PHASE_FERRO     (sqrt n, 0, 0)
 *     GPE_PHASE_MIXED     (sqrt(n param), sqrt(n (1 - param)), 0)     param = ferro fraction
 *   spin-2 (m = 2..-2)                          Spin2GpeBCC.cu:417-558
 *     GPE_PHASE_UN            (0, 0, sqrt n, 0, 0)
 *     GPE_PHASE_BN_HORIZONTAL (sqrt(n/2), 0, 0, 0, e^{i param} sqrt(n/2))
 *     GPE_PHASE_BN_VERTICAL   (0, sqrt(n/2), 0, e^{i param} sqrt(n/2), 0)
 *     GPE_PHASE_CYCLIC        (sqrt(n/3), 0, 0, e^{i param} sqrt(2n/3), 0)
 *     GPE_PHASE_FERRO         (sqrt n, 0, 0, 0, 0)
 * Replaces the reference's projection kernels, launched on its state before stepping. */
#define GPE_PHASE_POLAR 0
#define GPE_PHASE_FERRO 1
#define GPE_PHASE_MIXED 2
#define GPE_PHASE_UN 3
#define GPE_PHASE_BN_HORIZONTAL 4
#define GPE_PHASE_BN_VERTICAL 5
#define GPE_PHASE_CYCLIC 6
int gpe_project_spinor(gpe_ctx *ctx, int level, int phase, double param);
/* psi_m += amplitude sqrt(n) eps_m on a level, eps_m uniform in the unit disk from a
 * counter-based hash of (seed, unknown, component), n the density before the noise.  The
 * reference draws normal variates on the host instead (Spin1GpeBCC.cu:1257-1292). */
int gpe_add_noise(gpe_ctx *ctx, int level, double amplitude, unsigned long long seed);
/* Piecewise-linear bias ramp Bz(t) through n >= 2 equally spaced knots over `duration` (the
 * first knot held at t = 0, then linear segments; LINEAR segments of getSignal,
 * AliceRingRamps.h:74-135) with a constant quadrupole gradient bq: gpe_set_ramp with those
 * segments, time_scale 1 and unit scales. */
int gpe_set_knot_ramp(gpe_ctx *ctx, int n, const double *bz_knots, double duration, double bq);

/* ------------------------------------------------------- state files and checkpoints ---- */
/* The reference solvers' state file: the raw padded block array they dump with
 * ofstream::write (Spin1GpeBCC.cu:1690-1699, read back :1228-1237): (nz+2)(ny+2)(nx+2) blocks (a
 * one-block zero halo), block (i, j, k) at index (k+1)(ny+2)(nx+2) + (j+1)(nx+2) + (i+1), each
 * holding V unknowns of ncomp complex doubles, m = F..-F.  gpe_read_dat sets the level like
 * gpe_set_state (level 0: run time t). */
int gpe_write_dat(gpe_ctx *ctx, int level, const char *path);
int gpe_read_dat(gpe_ctx *ctx, int level, const char *path, double t);
/* Checkpoint pair of the reference (written :1690-1699, resumed :1228-1307): files
 *   <dir>/<prefix><T>.dat        level 1, psi_{k-1} (the reference's odd array)
 *   <dir>/<prefix>even_<T>.dat   level 0, psi_k
 *   <dir>/<prefix><T>.clock      text: t0 and sweeps of the run clock (exact resume)
 * T: the run time printed as the reference's toString (fixed, 18 decimals, utils.h:102-108),
 * copied to tag (tag_cap bytes, may be NULL).  Needs a started context. */
int gpe_save_checkpoint(gpe_ctx *ctx, const char *dir, const char *prefix, char *tag, int tag_cap);
/* Resume from <dir>/<prefix><tag>.dat.  With the even file present both levels are loaded and
 * the context steps on directly (*levels = 2, the reference skips its forward-Euler step);
 * without it level 0 is loaded and the caller starts (gpe_start) as the reference does
 * (*levels = 1).  The clock comes from the .clock file when present, else t = the tag's value. */
int gpe_load_checkpoint(gpe_ctx *ctx, const char *dir, const char *prefix, const char *tag, int *levels);

/* -------------------------------------------------------------------- VTK files ---- */
/* Legacy binary VTK POLYDATA files laid out byte for byte as the reference writes them, so its
 * viewers (ferro_skyrmion.py, bn_skyrmion.py, integrate_dens.py: arrays read by index) read
 * them unchanged.  Points: every unknown of the one-block-padded grid in the order z, x, y, v,
 * at p0 + s (i_padded * block edge + pos_v) as big-endian float32 (the padded index, as in the
 * reference, which places the points one block edge above the unknowns' positions; halo
 * points carry zeros).  Values 0 where the density is at most 1e-4 (utils.h:789).
 *   spinor:    SCALARS r_m=F, i_m=F, ..., r_m=-F, i_m=-F: psi_m / sqrt(n)      (saveSpinor,
 *              Spin1GpeBCC/utils.h:1079-1314, Spin2GpeBCC/utils.h:1071-1744); spin-2 basis
 *              GPE_BASIS_Z, _Y, _X quantisation axis (Spin2GpeBCC/AliceRingRamps.h:6-10)
 *   preimage:  spin-1: the spinor arrays on the interior points whose spin direction
 *              psi^dag F psi / |psi^dag F psi| has |component axis| > threshold (reference: y,
 *              0.9996; savePreImageSpinor, Spin1GpeBCC/utils.h:1351-1715)
 *   volume:    spin-1: density, spinNorm, theta, VECTORS localAvgSpin, VECTORS u
 *              (saveVolume, Spin1GpeBCC/utils.h:792-1077; localAvgSpin as the reference's
 *              localAvgSpinAndDensity, Spin1GpeBCC.cu:235-271, u and theta as uvTheta,
 *              :273-326); spin-2: SCALARS m=2..m=-2 densities (Spin2GpeBCC/utils.h:809-1069) */
#define GPE_BASIS_Z 0
#define GPE_BASIS_Y 1
#define GPE_BASIS_X 2
int gpe_write_vtk_spinor(gpe_ctx *ctx, int level, const char *path, int basis);
int gpe_write_vtk_preimage(gpe_ctx *ctx, int level, const char *path, int axis, double threshold);
int gpe_write_vtk_volume(gpe_ctx *ctx, int level, const char *path, int basis);

/* The same formats for a state held on the host (ABI layout, complex128) on the grid of cfg,
 * without a context or a device: post-processing of saved states, and the host side of the
 * checkpoint pair.  gpe_host_read_dat fills psi (ncomp * nz * ny * nx * V complex values). */
int gpe_host_write_dat(const gpe_config *cfg, const double *psi, const char *path);
int gpe_host_read_dat(const gpe_config *cfg, const char *path, double *psi);
int gpe_host_write_vtk_spinor(const gpe_config *cfg, const double *psi, const char *path, int basis);
int gpe_host_write_vtk_preimage(const gpe_config *cfg, const double *psi, const char *path, int axis,
                                double threshold);
int gpe_host_write_vtk_volume(const gpe_config *cfg, const double *psi, const char *path, int basis);

/* ------------------------------------------------------------------ BMP pictures ---- */
/* The reference's density pictures, byte for byte (Picture::save, Output/Picture.cpp:83-138:
 * 24-bit BMP, rows from y = 0 up, channel value v stored as 256 v truncated and clamped to
 * 0..255; the header's file and image sizes count 4 bytes per pixel as the reference's do).
 * With W = 2 (nx + 2), H = 2 (ny + 2), D = 2 (nz + 2) pixels along the padded grid (two pixels
 * per block edge, the zero halo included):
 *   GPE_PICTURE_DENSITY  spinor states: per component m, the density |psi_m|^2 in the basis
 *       `basis` summed along one axis, grey, scaled by 1 / the plane's largest sum.
 *       spin-1 (drawDensity, Spin1GpeBCC/utils.h:151-366): 3 W x 2 H, XZ (y < D) and XY
 *       (from y = H), both mirrored in x, XY also in y.  spin-2 (drawDensity,
 *       Spin2GpeBCC/utils.h:287-796): 5 W x 3 H, XZ, YZ (scaled by the XZ maximum, as there),
 *       XY.  Grey 0.5 separator lines between the panels.
 *   GPE_PICTURE_RI  spinor states: red = sum of Re^2, green = sum of Im^2.  spin-1
 *       (drawDensityRI, Spin1GpeBCC/utils.h:368-580): 3 W x 2 H, XZ and XY, times 2 / the
 *       plane's largest density sum; spin-2 (drawIandR, Spin2GpeBCC/utils.h:131-285): 5 W x 2 H,
 *       unscaled, z basis only.
 *   GPE_PICTURE_SLICE  scalar states (GPUGrossPitaevskii_BCC.cu:700-723): W x H, the padded
 *       block layer nz / 2 + 1, unknown 0: red = 20 Re^2, green = 20 Im^2.
 * The reference names its pictures dir + "/" + name + t + "ms.bmp" with t in fixed notation
 * and 18 decimals (the checkpoint tag's format); path is the whole file name here.
 * GPE_EINVAL for a style the state's component count has no picture for, or a basis other
 * than z for a spin-2 GPE_PICTURE_RI picture. */
#define GPE_PICTURE_DENSITY 0
#define GPE_PICTURE_RI 1
#define GPE_PICTURE_SLICE 2
int gpe_write_bmp(gpe_ctx *ctx, int level, const char *path, int style, int basis);
int gpe_host_write_bmp(const gpe_config *cfg, const double *psi, const char *path, int style, int basis);

#ifdef __cplusplus
}
#endif
#endif
