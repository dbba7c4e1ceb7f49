/*
 * pentrack_gpu.h -- C ABI of the MI355X tracking engine (libpentrack_gpu.so).
 *
 * The reference (PENTrack) has no plugin loader or FFI: its hot path is the C++
 * call TTracker::IntegrateParticle(p, tmax, particleconf, mc, geom, field)
 * (include/tracking.h:46-47, called from src/main.cpp:159,166), with the world
 * model behind TGeometry (include/geometry.h:84-155) and TFieldManager
 * (include/fields.h:19-62).  This header is the drop-in boundary for that path:
 * plain pointers and sizes, no exceptions, status codes instead of throws.
 * All pointers are host pointers unless the name says otherwise; the library
 * owns every device allocation it makes.
 */
#ifndef PENTRACK_GPU_H_
#define PENTRACK_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PT_OK 0
#define PT_EINVAL (-1)   /* invalid argument / inconsistent input */
#define PT_ENOMEM (-2)   /* host or device allocation failed */
#define PT_EDEVICE (-3)  /* HIP runtime error or no usable GPU */

/* ---- per-particle stop codes (include/globals.h:13-23) ----------------- */
#define PT_ID_UNKNOWN 0
#define PT_ID_NOT_FINISH (-1)
#define PT_ID_HIT_BOUNDARIES (-2)
#define PT_ID_ODEINT_ERROR (-3)
#define PT_ID_DECAYED (-4)
#define PT_ID_CGAL_ERROR (-6)
#define PT_ID_GEOMETRY_ERROR (-7)
#define PT_ID_ABSORBED_IN_MATERIAL 1
#define PT_ID_ABSORBED_ON_SURFACE 2
/* extensions: a reference runtime_error (src/tracking.cpp:38,286,333,350) becomes a
   per-particle invariant stop; a configuration outside the implemented path
   (micro-roughness model, more than PT_MAX_HITS intersections of one segment, more
   than PT_MAX_SOLIDS nested solids) stops the particle as unsupported */
#define PT_ID_INVARIANT (-8)
#define PT_ID_UNSUPPORTED (-9)

#define PT_STATE 9 /* x y z vx vy vz proper_time polarisation path_length */
#define PT_SPIN 5
#define PT_MAX_HITS 16
#define PT_MAX_SOLIDS 8

/* ---- geometry ---------------------------------------------------------- */
typedef struct pt_geometry pt_geometry;

/*
 * Replaces TGeometry::TGeometry + TTriangleMesh::ReadFile + the CGAL AABB tree
 * (src/geometry.cpp:93-137, src/trianglemesh.cpp:17-130).
 * tri_vertices   [n_tri][3][3] float, as stored in the STL files, meshes concatenated
 *                in solid load order; coordinates with |v| < 1e-8 are snapped to 0
 *                (src/trianglemesh.cpp:42)
 * tri_solid      [n_tri] index into the solid table
 * solid_id       [n_solid] solid IDs (> 1), solid_mat [n_solid] material index
 * ign_off/ign    ignore intervals of solid s: ign[2k], ign[2k+1] for k in [ign_off[s], ign_off[s+1])
 * default_mat    material of the default solid (ID 1)
 * mats           [n_mat][9] FermiReal FermiImag DiffProb SpinflipProb RMSRoughness CorrelLength
 *                InternalBField ModifiedLambertProb LossPerBounce (include/geometry.h:20-31)
 * leaf_size      BVH leaf size (triangles per leaf), 1..32
 * Besides the BVH it builds, for the tracking kernels' wall test, a clearance grid (128
 * cells on the bounding box's longest axis) and per cell a candidate list (copies of the
 * triangles whose boxes meet the cell widened by two cells, at most 32) and a BVH subtree
 * root (of the triangles meeting the cell widened by four): for the 19,600-triangle
 * benchmark cylinder 1.8 s of host time (16 threads) and about 310 MB of device memory.
 */
int pt_geometry_create(int n_tri, const float *tri_vertices, const int32_t *tri_solid,
                       int n_solid, const uint32_t *solid_id, const int32_t *solid_mat,
                       const int32_t *ign_off, const double *ign, int default_mat,
                       int n_mat, const double *mats, int leaf_size, pt_geometry **out);
void pt_geometry_destroy(pt_geometry *g);
/*
 * Host-only diagnostic (no device memory): the clearance grid pt_geometry_create
 * builds over the mesh bounding box, a lower bound per cell of the distance to the
 * nearest triangle.  No counterpart in the reference (its AABB tree is queried for
 * every segment); with out == NULL only gn[3] and lo[3] are filled.
 */
int pt_clearance_grid(int n_tri, const float *tri_vertices, int cells_long, float *out, int32_t *gn,
                      double *lo, double *inv_h);
/* number of BVH nodes and depth of the built tree (diagnostics) */
int pt_geometry_info(const pt_geometry *g, int *n_nodes, int *depth);

/* ---- fields ------------------------------------------------------------ */
#define PT_FIELD_TRICUBIC 1      /* TabField3 (src/field_3d.cpp) */
#define PT_FIELD_UNIFORM_E 2     /* TEDMStaticEField (src/edmfields.cpp:175-187): e */
#define PT_FIELD_CONDUCTOR 3     /* TConductorField (src/conductor.cpp:9-172): p = I x1 y1 z1 x2 y2 z2 */
#define PT_FIELD_EXPONENTIAL_X 4 /* TExponentialFieldX (src/analyticFields.cpp:11-33): p = a1 a2 a3 c1 c2 */
#define PT_FIELD_LINEAR_Z 5      /* TLinearFieldZ (src/analyticFields.cpp:37-57): p = a1 a2 */
#define PT_FIELD_B0_GRAD_Z 6     /* TB0GradZ (src/analyticFields.cpp:61-81): p = a1 a2 z0 */
#define PT_FIELD_B0_GRAD_X2 7    /* TB0GradX2 (src/analyticFields.cpp:85-105): p = a1 a2 a3 z0 */
#define PT_FIELD_B0_GRAD_XY 8    /* TB0GradXY (src/analyticFields.cpp:109-129): p = a1 a2 z0 */
#define PT_FIELD_B0_XY 9         /* TB0_XY (src/analyticFields.cpp:133-153): p = a1 z0 */
#define PT_FIELD_TAB2D 10        /* TabField (src/field_2d.cpp): axisymmetric r-z map, bicubic */

typedef struct {
    int type;
    double bscale, escale;      /* constant TFieldScaler factors (src/field.cpp:7-45) */
    double box[6];              /* xmax xmin ymax ymin zmax zmin; unbounded unless min<max (src/field.cpp:57-69) */
    double bw;                  /* boundary width (src/field.cpp:71-125) */
    int nx, ny, nz;             /* tricubic grid (sorted unique coordinates) */
    const double *xg, *yg, *zg;
    int has_b, has_v;           /* coefficient blocks per cell: Bx By Bz if has_b, then V if has_v */
    const double *coef;         /* [(nx-1)(ny-1)(nz-1)][ncomp][64], cell (ix*(ny-1)+iy)*(nz-1)+iz */
    /* PT_FIELD_TAB2D uses nx radial points in xg and nz axial points in zg (both strictly
       increasing), has_b = mask of the Br (1), Bphi (2), Bz (4) columns, has_v = mask of
       the V (1), Er (2), Ephi (4), Ez (8) columns (V not together with E: the reference
       ignores the potential when E is tabulated), and coef from pt_bicubic_precompute
       with the present columns in the order Br Bphi Bz V Er Ephi Ez. */
    double e[3];                /* uniform E */
    double p[8];                /* analytic field and conductor parameters (see the type codes) */
    /* TFieldScaler formulas of the time t (src/field.cpp:7-45) in exprtk syntax, replacing
       bscale / escale when not NULL; formula names already resolved (src/fields.cpp ResolveFormula).
       pt_fields_create rejects, with PT_EINVAL and a message, the forms whose exprtk value
       the compiler does not reproduce (chains of + - or * / with constants or divisions,
       literal 0 operands of + * / on sub-expressions, '%', '&&', '||', other functions). */
    const char *bscale_formula, *escale_formula;
    /* PT_FIELD_TRICUBIC with coef == NULL: the samples [ncomp][nx][ny][nz] (host memory),
       pre-interpolated on the GPU by pt_fields_create (bit-identical to pt_tricubic_precompute) */
    const double *values;
    /* PT_FIELD_TRICUBIC, mixed-precision tables: the largest gradient error [field unit / m]
       (value error: times the smallest cell size) that a cell component's FP32 remainder may
       carry by its worst-case bound; components above it keep their FP64 coefficients
       (0: PT_MIXED_TOL_DEFAULT) */
    double mixed_tol;
} pt_field_desc;

#define PT_MIXED_TOL_DEFAULT 1e-10

typedef struct pt_fields pt_fields;

/*
 * Replaces TFieldManager::TFieldManager (src/fields.cpp:19-136): descriptors are
 * kept in the given order (the reference's lexicographic FIELDS key order) and
 * summed in that order (at most 16 containers).  with_mixed also builds, on the GPU, the mixed-precision
 * table of every tricubic map for PT_VARIANT_MIXED: per cell and component the
 * constant and linear coefficients in FP64 and the other 60 in FP32 (288 B instead of
 * 512 B), or the FP64 block where the FP32 remainder's error bound exceeds the map's
 * mixed_tol.  The FP64 tables are kept as well.
 */
int pt_fields_create(int n, const pt_field_desc *descs, int with_mixed, pt_fields **out);
void pt_fields_destroy(pt_fields *f);
/* No counterpart in the reference: blocks (cells x components) of the mixed-precision
   tables and how many of them kept their FP64 coefficients */
int pt_fields_mixed_info(const pt_fields *f, int64_t *blocks, int64_t *fp64_blocks);

/*
 * TabField3::PreInterpol (src/field_3d.cpp:367-414): spline derivatives along each
 * axis (ALGLIB spline1dgriddiffcubic, parabolic end conditions) and the 64
 * Lekien-Marsden coefficients of every cell (libtricubic tricubic_get_coeff).
 * values   [ncomp][nx][ny][nz] samples on the grid
 * coef     [(nx-1)(ny-1)(nz-1)][ncomp][64] output
 */
int pt_tricubic_precompute(int nx, int ny, int nz, const double *xg, const double *yg, const double *zg,
                           int ncomp, const double *values, double *coef, int threads);
/* pt_tricubic_precompute on the GPU (seven batched spline-derivative passes, one thread per
   grid line, then the 64 x 64 products per cell); host arrays in and out, same layout and
   bit-identical results; synchronous */
int pt_tricubic_precompute_gpu(int nx, int ny, int nz, const double *xg, const double *yg, const double *zg,
                               int ncomp, const double *values, double *coef);
/*
 * TabField's bicubic splines (src/field_2d.cpp:235-285: ALGLIB spline2dbuildbicubicv, the
 * derivative tables of spline2d_bicubiccalcderivatives).
 * m, rg     radial grid, n, zg axial grid (strictly increasing)
 * values    [ncomp][n][m] samples, values[c][iz][ir]
 * coef      [(n-1)(m-1)][ncomp][16] output, cell iz*(m-1)+ir: f, df/dr, df/dz, d2f/drdz
 *           at the corners (iz,ir), (iz,ir+1), (iz+1,ir), (iz+1,ir+1)
 */
int pt_bicubic_precompute(int m, int n, const double *rg, const double *zg, int ncomp, const double *values,
                          double *coef);
/* the 64x64 integer matrix of libtricubic's tricubic_get_coeff (coefficients = A x),
   row-major into out[64*64]; for tests */
int pt_tricubic_matrix(int32_t *out);
/* the kernels' segment/triangle intersection (CGAL intersection(Triangle_3, Segment_3) as
   TTriangleMesh::Collision uses it, src/trianglemesh.cpp:141-160), run on the host: 1 with
   the point in pt[3], 0 none, -1 coplanar overlap; for tests */
int pt_seg_tri(const double *a, const double *b, const double *c, const double *p, const double *q, double *pt);
/* spline derivatives dy/dx at the n grid points of one line, as ALGLIB's
   spline1dgriddiffcubic with parabolic end conditions (src/field_3d.cpp:356); for tests */
int pt_spline_grid_diff(int n, const double *x, const double *y, double *d);
/* TFieldScaler::scalingFactor (src/field.cpp:7-10): the exprtk formula f(t) compiled
   as for a field container and evaluated at n times on the host (for tests) */
int pt_formula_eval(const char *formula, int n, const double *t, double *out);

/* ---- particles --------------------------------------------------------- */
typedef struct {
    int kind;          /* 0 neutron, 1 proton, 2 electron */
    double q, m, mu, gamma; /* charge [C], mass [eV/c^2], magnetic moment [J/T], gyromagnetic ratio [1/(T s)] */
    double tau;        /* particleconf "tau": mean lifetime (<=0: none) */
    double tmax;       /* particleconf "tmax": proper-time limit when tau<=0 */
    double lmax;       /* particleconf "lmax" */
    int flipspin;      /* particleconf "flipspin" */
    double Bmax;       /* particleconf "Bmax" */
    int n_spintimes;   /* particleconf "spintimes": BMT spin tracking in [spintimes[2k], spintimes[2k+1]) */
    double spintimes[8];
    int interpolatefields; /* particleconf "interpolatefields": precession axis from splines over each step */
} pt_species;

typedef struct {
    double t, y[PT_STATE], spin[PT_SPIN];
    int32_t stopID, Nhit, Nstep, Nspinflip;
    int64_t n_acc, n_rej;        /* accepted / rejected dopri5 trial steps */
    double Hmax, noflipprob, tau;
    int32_t solidend;            /* ID of the current solid at the end */
    int32_t n_hits;              /* hits logged (may exceed the log capacity) */
} pt_particle_out;

typedef struct {
    double t;                    /* time of the bracketing point before the wall */
    int32_t tri;                 /* triangle index (input order) whose normal was used */
    int32_t leaving, entering;   /* solid IDs */
    double normal[3], v1[3], v2[3], pol1, pol2;
    double pos[3];               /* position of the bracketing point before the wall */
} pt_hit;

#define PT_VARIANT_FP64 0        /* plain FP64, one thread per particle */
#define PT_VARIANT_MIXED 1       /* mixed-precision field tables (FP64 linear part, FP32 remainder
                                    within mixed_tol), FP64 state, control, bisection, triangle test;
                                    tables that keep every block in FP64 run the FP64 kernels
                                    (the same arithmetic) */

typedef struct {
    int variant;
    int block;                   /* threads per block (64..1024) */
    int steps_per_launch;        /* accepted steps per particle per pt_batch_run call (0 = unlimited) */
    int persistent;              /* 1: persistent kernel refilling slots from a device work queue */
    int grid;                    /* persistent grid size in blocks (0 = automatic) */
    int rounds;                  /* step-kernel / collision-kernel rounds per pt_batch_run call
                                    (0 = 4); each round resolves at most one wall collision per particle */
    int graph;                   /* 1: capture the call's launch sequence into a hipGraph on first use
                                    and replay it (needs a non-default stream; no per-kernel timing) */
    int spatial_order;           /* persistent kernel: 1 hands out particles in Morton order of their
                                    position (32^3 buckets over the mesh box) so concurrent waves
                                    gather nearby field cells; 2 also splits that list into one
                                    contiguous chunk per XCD, taken first by the blocks on that XCD
                                    (neighbouring particles share its L2); 0 in index order.
                                    Results are equal. */
    int collision_order;         /* 1: the collision kernel takes the parked particles ordered by the
                                    number of candidate triangles of their chord piece (counted and
                                    counting-sorted first), so a wave's lanes loop over lists of
                                    similar length; 0 in queue order.  Results are equal. */
} pt_launch_cfg;

typedef struct pt_batch pt_batch;

/*
 * Uploads n particles (the state TTracker::IntegrateParticle starts from) and
 * allocates their resumable tracker state in device memory (at most 8 species).
 * species_idx [n], t0 [n], y0 [n][9], spin0 [n][5], pid [n] (random-stream ids).
 * hit_cap: hits logged per particle (0 = none).
 */
int pt_batch_create(const pt_geometry *g, const pt_fields *f, int n_species, const pt_species *species,
                    double tmax_global, int n, const int32_t *species_idx, const double *t0,
                    const double *y0, const double *spin0, uint64_t seed, const uint64_t *pid,
                    int hit_cap, pt_batch **out);
void pt_batch_destroy(pt_batch *b);
/* Advance every unfinished particle (stream-ordered, asynchronous; stream may be NULL). */
int pt_batch_run(pt_batch *b, void *stream, const pt_launch_cfg *cfg);
/* Number of particles not yet stopped (synchronises the stream). */
int pt_batch_active(pt_batch *b, void *stream, int *active);
/* Copy results to the host (synchronises the stream).  hits may be NULL. */
int pt_batch_download(pt_batch *b, void *stream, pt_particle_out *out, pt_hit *hits);

/*
 * No counterpart in the reference (which has LossPerBounce per material only,
 * include/geometry.h:30): a loss probability per bounce for each of the n particles, replacing
 * the material's LossPerBounce in TNeutron::OnHit's absorption probability on total reflection
 * (src/neutron.cpp:241) at every wall that particle meets; values in [0, 1].  loss == NULL
 * switches it off.  Call before pt_batch_run.
 */
int pt_batch_set_loss(pt_batch *b, const double *loss);

/* a particle's state at a snapshot time, with its spin and counters at that point */
typedef struct {
    double t, y[9], spin[5], Hmax, noflipprob;
    int32_t stopID, Nhit, Nstep, Nspinflip;
} pt_snapshot;
/*
 * TLogger::PrintSnapshot (src/logger.cpp:128-140, called after each step's collision
 * checks, src/tracking.cpp:126): after every step from x1 (the stepper's step start) to x
 * (the step's end after collisions), the first snapshot time in [x1, x) is recorded from
 * the stepper's dense output.  Times are sorted and de-duplicated; cap records are kept
 * per particle (more are counted).  Call before the first pt_batch_run; n_times = 0
 * switches snapshots off.
 */
int pt_batch_set_snapshots(pt_batch *b, int n_times, const double *times, int cap);
/* out [n][cap] records, count [n] records taken (may exceed cap; synchronises the stream) */
int pt_batch_snapshots(pt_batch *b, void *stream, pt_snapshot *out, int32_t *count);

/* a step's end state for the track log and the ID of the particle's current solid */
typedef struct {
    double t, y[9];
    int32_t solid, pad;
} pt_track_point;
/*
 * TLogger::PrintTrack (src/logger.cpp:142-196) as IntegrateParticle calls it after each step
 * (src/tracking.cpp:140): the step's end state (after the spin step) is recorded when
 * int(l1 / interval) != int(l / interval) for the trajectory lengths l1 at the stepper's
 * step start and l at the end.  The start point (src/tracking.cpp:43) is not recorded here.
 * cap records are kept per particle (more are counted); interval = 0 switches it off.
 */
int pt_batch_set_track(pt_batch *b, double interval, int cap);
/* out [n][cap] records, count [n] records taken (may exceed cap; synchronises the stream) */
int pt_batch_track(pt_batch *b, void *stream, pt_track_point *out, int32_t *count);

/* a spin-log row: time, position, spin and the BMT precession axis W there */
typedef struct {
    double t, pos[3], spin[3], W[3];
} pt_spin_point;
/*
 * TLogger::PrintSpin (src/logger.cpp:237-284) as IntegrateSpin calls it (src/tracking.cpp:431,
 * :448): inside the BMT spin-tracking windows a row is recorded at the start of each spin
 * integration and after each spin step whose end x satisfies int(x1 / interval) !=
 * int(x / interval) for the spin stepper's previous time x1 (the end is clamped to the
 * trajectory step's end, where the row holds the dense-output spin).  At the first call of
 * an integration the fresh stepper's previous time is 0.0, so its start row is skipped when
 * the integration starts in (0, interval).  cap records are kept per particle (more are
 * counted); interval = 0 switches it off.
 */
int pt_batch_set_spinlog(pt_batch *b, double interval, int cap);
/* out [n][cap] records, count [n] records taken (may exceed cap; synchronises the stream) */
int pt_batch_spinlog(pt_batch *b, void *stream, pt_spin_point *out, int32_t *count);
/* Sums over the batch (synchronises the stream): accepted / rejected dopri5 steps and
   field-map cell evaluations of the step kernel (each gathers one coefficient block per
   component); any pointer may be NULL. */
int pt_batch_step_counts(pt_batch *b, void *stream, int64_t *accepted, int64_t *rejected,
                         int64_t *field_evals);
/* Kernel timing (no counterpart in the reference): with profiling on, pt_batch_run records
   HIP events around every step-kernel and collision-kernel launch on its stream;
   pt_batch_kernel_times synchronises and returns the accumulated milliseconds and
   launch counts, ms[0]/launches[0] for the step kernel, ms[1]/launches[1] for the
   collision kernel (with its queue reset). */
int pt_batch_set_profiling(pt_batch *b, int on);

/* a decay product: TNeutron::Decay (src/neutron.cpp:285-362) creates a proton and an
   electron at the decay point with these arguments of the TProton / TElectron
   constructors (src/particle.cpp:36-90) */
typedef struct {
    int32_t parent;              /* index of the decayed neutron in the batch */
    int32_t kind;                /* 1 proton, 2 electron */
    double t, pos[3];            /* decay time and point */
    double E, phi, theta;        /* kinetic energy [eV], azimuth and polar angle of the velocity */
    int32_t pol, pad;            /* polarisation (= spin projection), +-1 */
} pt_secondary;

/*
 * TParticle::DoDecay / TNeutron::Decay (src/particle.cpp:219-223, src/neutron.cpp:285-362)
 * for the neutrons of the batch that stopped with ID_DECAYED (the secondaries the
 * reference appends, src/tracking.cpp:152-155): *count = 2 x decayed neutrons, the
 * first min(cap, *count) records written to out, ordered by parent (proton, then
 * electron).  The draws continue each neutron's random stream; the products are the
 * same on every call.  Synchronises the stream.
 */
int pt_batch_secondaries(pt_batch *b, void *stream, int cap, pt_secondary *out, int *count);
/* for tests: the proton recoil spectrum as std::piecewise_linear_distribution holds it
   (n_int + 1 abscissas, spectrum values and normalised densities, n_int cumulative
   probabilities and slopes); with x == NULL only *n_int is returned */
int pt_proton_beta_table(int cap, double *x, double *raw, double *den, double *cp, double *m, int *n_int);
/*
 * parse_distribution (include/mc.h:310-335): the piecewise linear distribution
 * (std::piecewise_linear_distribution parameter) of a formula of x on [lo, hi], as
 * the reference builds the source's energy spectrum and angular distributions;
 * the formula may call ProtonBetaSpectrum(x).  With x == NULL only *n_int.
 */
int pt_distribution_table(const char *formula, double lo, double hi, int cap, double *x, double *raw,
                          double *den, double *cp, double *m, int *n_int);
/* TFieldManager::BField / EField (src/fields.cpp:139-182) at n points pts[n][3] and times t[n]:
   B [n][3], dB [n][3][3], V [n], E [n][3] (host arrays; FP64 tables; synchronises the stream).
   dB may be NULL: then B is the value-only evaluation BField(x, y, z, t, B) (which for 2D
   maps differs from the one with derivatives in the last bits, as in the reference). */
int pt_fields_eval(const pt_fields *f, void *stream, int n, const double *pts, const double *t, double *B,
                   double *dB, double *V, double *E);
/* The same field sum evaluated as the tracking kernels evaluate it: variant PT_VARIANT_FP64
   or PT_VARIANT_MIXED (FP64 linear part plus FP32 remainder; the fields must have been
   created with their mixed-precision tables), with per-lane coefficient loads
   (cooperative 0) or the step kernels' wave-cooperative gathers (cooperative 1; the mixed
   tables are read per lane either way, as the step kernels read them).  The two gathers
   give bit-identical results; variant FP64 with cooperative 0 is pt_fields_eval.
   No counterpart in the reference: it exposes the kernels' variants of
   TFieldManager::BField / EField (src/fields.cpp:139-182) for parity checks. */
int pt_fields_eval_tables(const pt_fields *f, void *stream, int n, const double *pts, const double *t, int variant,
                          int cooperative, double *B, double *dB, double *V, double *E);
/* TGeometry::GetSolid (src/geometry.cpp:151-175) at n points and times: the ID of the
   highest-priority solid not ignored at t (1: default solid; -1: too deeply nested;
   -2: reserved for an incomplete BVH traversal, which the stackless traversal excludes) */
int pt_geometry_solid_at(const pt_geometry *g, void *stream, int n, const double *pts, const double *t,
                         int32_t *solid_id);
/* The intersection count TTracker::CheckHit takes from TGeometry::GetCollisions
   (src/tracking.cpp:193-195, src/trianglemesh.cpp:127-160) for n segments p[i] -> q[i]
   (p, q [n][3]): the number of triangles each segment meets, -1 for a segment coplanar
   with a triangle it overlaps (ID_CGAL_ERROR), -2 for more than 16 intersections.
   mode 0 counts as the tracking kernels do (clearance grid, then the candidate list of the
   segment's cell, then the BVH), mode 1 from the BVH alone (both give the same counts);
   mode 2 returns which way mode 0 takes (0 cleared by the grid, 1 cell list, 2 BVH). */
int pt_geometry_segment_hits(const pt_geometry *g, void *stream, int n, const double *p, const double *q, int mode,
                             int32_t *out);
/* for tests: the proton energies sampled for the variates p[0..n-1] in [0, 1) */
int pt_beta_sample(int n, const double *p, double *out);
/* for tests: the decay kinematics on the host for stream (key, ctr) and neutron velocity v[3]:
   out6 = Ep phip thetap Ee phie thetae, pol2 = polarisations, *ctr_out = stream position after */
int pt_neutron_decay_host(uint64_t key, uint64_t ctr, const double *v, double *out6, int32_t *pol2,
                          uint64_t *ctr_out);
int pt_batch_kernel_times(pt_batch *b, void *stream, double *ms, int64_t *launches);

/*
 * One-shot equivalent of calling TTracker::IntegrateParticle for every particle:
 * upload, run to completion, download.
 */
int pt_track_batch(void *stream, const pt_geometry *g, const pt_fields *f, int n_species,
                   const pt_species *species, double tmax_global, int n, const int32_t *species_idx,
                   const double *t0, const double *y0, const double *spin0, uint64_t seed,
                   const uint64_t *pid, pt_particle_out *out, pt_hit *hits, int hit_cap,
                   const pt_launch_cfg *cfg);

/* ---- differentiable path: fixed-step RK4 and its adjoint ----------------- */
/*
 * No counterpart in the reference (its odeint stepper has no derivatives): smooth
 * neutron motion (TParticle::EquationOfMotion, src/particle.cpp:105-135, in the
 * non-relativistic limit) through B = sum_k s_k B_k, gravity, and specular box walls
 * applied as fixed maps after each step.  All array pointers are DEVICE pointers;
 * states are structure-of-arrays [6][n] (x y z vx vy vz).
 */
#define PT_ADJ_MAX_MAPS 8

typedef struct {
    int nmaps;                  /* number of maps k (<= PT_ADJ_MAX_MAPS), all on one grid */
    int nx, ny, nz;             /* grid (strictly increasing coordinates) */
    double x0[3], inv_h[3];     /* first coordinate and (n-1)/extent per axis (cell-search guess) */
    const double *g[3];         /* grid coordinates (device) */
    const double *coef[PT_ADJ_MAX_MAPS]; /* map k: [(nx-1)(ny-1)(nz-1)][3][64] blocks, device layout
                                           a[16i+4j+k] = coefficient of x^i y^j z^k */
    double s[PT_ADJ_MAX_MAPS];  /* scale factors s_k */
} pt_adj_maps;

typedef struct {
    pt_adj_maps maps;
    double lo[3], hi[3];        /* box walls */
    double dt;                  /* RK4 step [s] */
    double coupling;            /* pol * mu / m  [m^2 / (s^2 T)] */
    double grav;                /* g [m/s^2] along -z */
} pt_adj_params;

/* Forward: yT = state after nsteps steps from y0; ckpt (may be NULL) receives the state
   before every K-th step, [ceil(nsteps/K)][6][n]. */
int pt_rk4_forward_ckpt(void *stream, const pt_adj_params *params, int n, const double *y0, double *yT,
                        double *ckpt, int nsteps, int K);
/* Adjoint: lam holds dJ/dyT on entry and dJ/dy0 on exit ([6][n]); sbar receives dJ/ds_k per
   particle ([nmaps][n]); seg ([K][24][n] doubles) and flags ([K][n] ints) are workspace. */
int pt_rk4_adjoint(void *stream, const pt_adj_params *params, int n, const double *ckpt, double *lam,
                   double *sbar, double *seg, int *flags, int nsteps, int K);
/*
 * No counterpart in the reference.  pt_rk4_adjoint for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same kernel, with its workspace (seg and flags above) in a
 * device array of PT_ADJ_NOALLOC_MAX_BYTES that the library holds on every GPU it runs
 * on until the process ends.  The call allocates nothing; the caller still passes ckpt,
 * lam and sbar.  Returns PT_EINVAL, with a message in pt_last_error(), when
 * min(K, nsteps) * n * 196 bytes (plus alignment) exceed PT_ADJ_NOALLOC_MAX_BYTES.
 * Calls on one GPU share that array: calls that overlap in time on one GPU (different
 * streams or host threads) race and must be serialised by the caller.
 */
#define PT_ADJ_NOALLOC_MAX_BYTES (64u << 20)
int pt_rk4_adjoint_noalloc(void *stream, const pt_adj_params *params, int n, const double *ckpt, double *lam,
                           double *sbar, int nsteps, int K);

/* ---- measurement aid ------------------------------------------------------ */
/*
 * No counterpart in the reference: the bandwidth ceilings the tracking kernels are compared
 * with.  gather_gbs: `threads` lanes each read `gathers` blocks of block_bytes (512: one
 * component's coefficients, 1536: a cell's three; contiguous per lane, 16-byte loads, a
 * 128-byte line's eight issued together) from uniformly random blocks of a table of
 * table_bytes, the access pattern of the tricubic field evaluation; copy_gbs: a streaming copy of the table (bytes
 * read + written).  Both averaged over `reps` launches after a warm-up; allocates and frees
 * 2 x table_bytes of device memory; synchronous on the default stream.
 */
int pt_gather_bandwidth(uint64_t table_bytes, int block_bytes, int threads, int gathers, int reps,
                        double *gather_gbs, double *copy_gbs);
/* No counterpart in the reference: the same gather of 1536-byte blocks in the step kernels'
 * cooperative form (the 64 lanes of a wave load each other's blocks half a 128-byte line
 * per lane at a time, four lanes per 64 contiguous bytes, the chunks passed to their
 * owners through LDS; two 256-thread blocks of 80 KB of LDS per CU, as the step kernels
 * run).  Allocates and frees table_bytes of device memory; synchronous on the default
 * stream. */
int pt_gather_bandwidth_coop(uint64_t table_bytes, int threads, int gathers, int reps, double *gather_gbs);

/* Cancellation flag mirroring the reference's global `quit` (include/globals.h:33):
   pt_batch_run copies it to the device when it is called, and its step kernels
   then advance no particle; state stays resumable (clear the flag and run again). */
void pt_request_quit(int on);

/* Last HIP error string of this thread (diagnostics). */
const char *pt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PENTRACK_GPU_H_ */
