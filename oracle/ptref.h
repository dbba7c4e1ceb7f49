/*
 * ptref.h -- data layout of the CPU oracle for PENTrack's tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker the GPU path is compared
 * against; only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The product never links or calls it.
 *
 * All arrays are plain host memory, laid out exactly like the inputs of the
 * product's C ABI (include/pentrack_gpu.h) so that both can be fed from the
 * same numpy buffers.
 */
#ifndef PTREF_H_
#define PTREF_H_

#include <stdint.h>

#define PTREF_STATE 9
#define PTREF_SPIN 5

typedef struct {
    int n_tri;
    const double *tri;        /* [n_tri][9]: v0, v1, v2 (x,y,z each), after the |v|<1e-8 snap */
    const int32_t *tri_solid; /* [n_tri]: index into the solid table */
    int n_solid;              /* meshes, in load order (lexicographic GEOMETRY keys) */
    const uint32_t *solid_id; /* [n_solid] solid IDs (>1) */
    const int32_t *solid_mat; /* [n_solid] material index */
    const int32_t *ign_off;   /* [n_solid+1] offsets into ign */
    const double *ign;        /* [n_ign][2]: ignore intervals [start, end) */
    int default_mat;          /* material of the default solid (ID 1) */
    int n_mat;
    const double *mats;       /* [n_mat][9]: FermiReal FermiImag DiffProb SpinflipProb RMSRoughness
                                 CorrelLength InternalBField ModifiedLambertProb LossPerBounce */
} ptref_geom;

enum {
    PTREF_FIELD_TRICUBIC = 1, PTREF_FIELD_UNIFORM_E = 2, PTREF_FIELD_CONDUCTOR = 3,
    PTREF_FIELD_EXPONENTIAL_X = 4, PTREF_FIELD_LINEAR_Z = 5, PTREF_FIELD_B0_GRAD_Z = 6,
    PTREF_FIELD_B0_GRAD_X2 = 7, PTREF_FIELD_B0_GRAD_XY = 8, PTREF_FIELD_B0_XY = 9, PTREF_FIELD_TAB2D = 10
};

/* one node of a field-scaling formula tree (oracle/formula.py nodes()): children
   precede parents, the root is the last node */
typedef struct {
    int op;                   /* oracle/formula.py codes: 0 const, 1 t, 2 neg, ... */
    int a, b, c;              /* child node indices */
    double v;                 /* constant value; exponent count of an integer power */
} ptref_fnode;

typedef struct {
    int type;
    double bscale, escale;    /* constant scale factors of the container (used when no formula) */
    int n_bform, n_eform;     /* formula trees of t for the scale factors (0: constant) */
    const ptref_fnode *bform, *eform;
    double p[8];              /* analytic field parameters (include/pentrack_gpu.h PT_FIELD_*) */
    double box[6];            /* xmax xmin ymax ymin zmax zmin; no bounds unless min<max on all axes */
    double bw;                /* boundary width of the smooth taper */
    /* tricubic map */
    int nx, ny, nz;
    const double *xg, *yg, *zg; /* sorted unique grid coordinates */
    int has_b, has_v;           /* coefficient blocks present per cell: Bx By Bz (if has_b), V (if has_v) */
    const double *coef;         /* [(nx-1)(ny-1)(nz-1)][ncomp][64], cell index (ix*(ny-1)+iy)*(nz-1)+iz */
    /* uniform E */
    double e[3];
    /* axisymmetric map: nx radial points in xg, nz axial points in zg; has_b = mask of Br (1),
       Bphi (2), Bz (4), has_v = mask of V (1), Er (2), Ephi (4), Ez (8); bic = ALGLIB's four
       tables per present column (ptref_bicubic_tables) */
    const double *bic;          /* [ncomp][4][nz][nx]: f, df/dr, df/dz, d2f/drdz */
} ptref_field;

typedef struct {
    int kind;                 /* 0 neutron, 1 proton, 2 electron */
    double tau;               /* mean lifetime (<=0: none; then tmax is the proper-time limit) */
    double tmax;              /* species tmax (used as lifetime when tau<=0) */
    double lmax;              /* max trajectory length */
    int flipspin;
    double Bmax;
    int n_spintimes;          /* spin-tracking windows [spintimes[2k], spintimes[2k+1]) (BMT) */
    double spintimes[8];
    int interpolatefields;    /* precession axis from cubic splines over 11 points of each step */
} ptref_species;

typedef struct {
    double t, y[PTREF_STATE], spin[PTREF_SPIN];
    int32_t stopID, Nhit, Nstep, Nspinflip;
    int64_t n_acc, n_rej;
    double Hmax, noflipprob, tau;
    int32_t solidend;         /* ID of the solid the particle ended in */
    int32_t n_hits_logged;
    int64_t n_rhs, n_field, n_tri_tests, n_bisect, n_pieces;  /* op counters */
    uint64_t ctr_end;         /* position of the particle's random stream at the end */
} ptref_out;

typedef struct {
    double t;
    int32_t tri, leaving, entering;
    double normal[3], v1[3], v2[3], pol1, pol2;
    double pos[3];
} ptref_hit;

/* a step's end state for the track log (TLogger::PrintTrack) and the current solid's ID */
typedef struct {
    double t, y[9];
    int32_t solid, pad;
} ptref_trackpt;

/* a spin-log row (TLogger::PrintSpin): time, position, spin, BMT precession axis */
typedef struct {
    double t, pos[3], spin[3], W[3];
} ptref_spinpt;

/* a particle's state at a snapshot time (TLogger::PrintSnapshot) and its counters then */
typedef struct {
    double t, y[9], spin[5], Hmax, noflipprob;
    int32_t stopID, Nhit, Nstep, Nspinflip;
} ptref_snap;

#endif /* PTREF_H_ */
