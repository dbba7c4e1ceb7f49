/*
 * pt_device.h -- device-side physics of the MI355X tracking engine.
 *
 * One GPU thread advances one particle through TTracker::IntegrateParticle's
 * loop (src/tracking.cpp:71-148).  The reference's recursion (CheckHit,
 * iterate_collision) becomes loops with bounded state; its CGAL AABB tree
 * becomes a flattened BVH walked with a short per-thread stack; its
 * Boost.Odeint dense-output stepper keeps its stage vectors in registers.
 * Floating-point expressions keep the reference's evaluation order and the
 * build disables FP contraction, so the FP64 path reproduces the reference
 * arithmetic (with long double replaced by double) bit for bit.
 */
#ifndef PT_DEVICE_H_
#define PT_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pt_detmath.h"

#define PTD __device__ __forceinline__
#define PTD_NOINLINE __device__ __noinline__
#define PTHD __host__ __device__ __forceinline__
#define PTHD_NOINLINE inline __host__ __device__ __noinline__

/* Loads and stores through a pointer to the library's device allocations, issued as
   global (not flat) memory instructions: flat accesses also count against the wait
   counter of LDS and scalar loads, so every LDS read would wait for them too */
template <class T> PTD T gload(const T *p) { return *(const __attribute__((address_space(1))) T *)(p); }
template <class T> PTD void gstore(T *p, T v) { *(__attribute__((address_space(1))) T *)(p) = v; }
/* relaxed device-scope atomic add on global memory: with the address space known the
   compiler combines the lanes of a wave that add to one address into a single atomic
   (a flat atomic would be issued lane by lane) */
template <class T> PTD T gatomic_add(T *p, T v) {
    return __hip_atomic_fetch_add((__attribute__((address_space(1))) T *)(p), v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}
/* HIP's vector structs copy through a generic reference: load their native vector type */
typedef double pt_d2v __attribute__((ext_vector_type(2)));
typedef float pt_f4v __attribute__((ext_vector_type(4)));
PTD double2 gload(const double2 *p) {
    const pt_d2v v = *(const __attribute__((address_space(1))) pt_d2v *)(p);
    return make_double2(v.x, v.y);
}
PTD float4 gload(const float4 *p) {
    const pt_f4v v = *(const __attribute__((address_space(1))) pt_f4v *)(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

namespace ptd {

/* ---- constants (src/globals.cpp:13-32, double-rounded) ------------------- */
constexpr double ELE_E = 1.602176487E-19;
constexpr double GRAV = 9.80665;
constexpr double C_0 = 299792458.;
constexpr double PI = 3.1415926535897932384626;
constexpr double HBAR = 1.05457266e-34;
constexpr double MAX_TRACK_DEVIATION = 0.001;   /* include/particle.h:20 */
constexpr double REFLECT_TOLERANCE = 1e-8;      /* include/trianglemesh.h:39 */
constexpr double DBL_EPS = 2.220446049250313e-16;
constexpr double DBL_MAX_V = 1.7976931348623157e308;   /* largest finite double */

enum {
    ID_UNKNOWN = 0, ID_NOT_FINISH = -1, ID_HIT_BOUNDARIES = -2, ID_ODEINT_ERROR = -3,
    ID_DECAYED = -4, ID_CGAL_ERROR = -6, ID_GEOMETRY_ERROR = -7,
    ID_ABSORBED_IN_MATERIAL = 1, ID_ABSORBED_ON_SURFACE = 2,
    ID_INVARIANT = -8, ID_UNSUPPORTED = -9
};

constexpr int MAX_HITS = 16;     /* intersections kept per segment query */
constexpr int MAX_SOLIDS = 8;    /* entries of the per-particle solid stack */
constexpr int BVH_STACK = 48;

/* ---- device-side data ---------------------------------------------------- */

/* BVH nodes in preorder (an internal node's left child is the next node).  On the device
   an internal node holds its escape index, the first node after its subtree, so the
   traversals need no stack: a box that is hit is entered at the next node, one that is
   missed is left for its escape; leaves and their triangles are visited in the same
   depth-first order as with a stack.  The host's build keeps the left child in a. */
struct BvhNode {          /* 32 B: two 16-B loads */
    float lo[3]; int a;   /* internal: a = escape index (host build: left child); leaf: a = -(first + 1) */
    float hi[3]; int b;   /* internal: b = right child; leaf: b = count */
};

struct Tri {              /* 48 B: three 16-B loads, vertices exact in float */
    float v[9];
    int solid;            /* solid index */
    int orig;             /* triangle index in input order */
    int pad;
};

struct Geom {
    int n_tri, n_node, n_solid, default_mat;
    const Tri *tri;           /* BVH leaf order */
    const double *nrm;        /* [n_tri][3] face normals, BVH leaf order */
    const BvhNode *node;
    const uint32_t *solid_id;
    const int *solid_mat;
    const int *ign_off;
    const double *ign;
    const double *mats;       /* [n_mat][9] */
    const double *mesh_bbox;  /* [n_solid][6] xmin ymin zmin xmax ymax zmax */
    /* clearance grid: per cell a lower bound of the distance from its centre to the
       nearest triangle less half its diagonal (clear_of_walls) */
    const float *clear;       /* [gn0][gn1][gn2] */
    int gn[3];
    double glo[3], ginv_h[3];
    /* candidate lists of the clearance cells: the triangles (copies, BVH leaf order) whose
       padded boxes meet the cell widened by cl_pad on every side; cl_rng[cell] = (first,
       count), count < 0 where the list would be longer than CELL_LIST_MAX */
    const int2 *cl_rng;
    const Tri *cl_tri;
    double cl_pad;
    /* per cell the deepest BVH node whose subtree holds every triangle whose padded box
       meets the cell widened by rt_pad (-2: no such triangle) */
    const int *cl_root;
    double rt_pad;
};
constexpr int CELL_LIST_MAX = 32;

struct Field {
    int type;                 /* PT_FIELD_* of include/pentrack_gpu.h */
    int has_b, has_v, ncomp;
    int blen, elen;           /* lengths of the scale-formula programs (0: constant bscale / escale) */
    const uint32_t *bcode, *ecode;
    const double *bcst, *ecst;
    double p[8];              /* analytic parameters; conductor: I x1 y1 z1 x2 y2 z2, mu0 I / (4 pi) */
    int nx, ny, nz;
    double bscale, escale, bw;
    double box[6];
    double e[3];
    double x0[3], inv_h[3];   /* first grid coordinate and 1/spacing guess for the cell search */
    const double *g[3];       /* grid coordinates */
    const double *coef;       /* FP64 table */
    const unsigned char *mix; /* mixed-precision table: mix_cell_bytes(ncomp) per cell */
};

/* A cell record of the mixed-precision table: first every component's constant and linear
   coefficients in FP64 (32 bytes each, device order: slots 0 = 1, 1 = z, 4 = y, 16 = x),
   padded to a multiple of 128 bytes, then every component's 64 coefficients in FP32 with
   those four slots zeroed (256 bytes each).  Records are 128-byte aligned, so the FP64 part
   of a 3-component cell is one 128-byte line and each FP32 x-power line (64 bytes) lies
   within one.  A component whose FP32 remainder may err by more than the map's tolerance
   keeps its FP64 block instead, marked by a NaN constant term. */
constexpr int MIX_LIN_BYTES = 4 * 8, MIX_REM_BYTES = 64 * 4;
PTHD size_t mix_lin_bytes(int ncomp) { return ((size_t)ncomp * MIX_LIN_BYTES + 127) & ~(size_t)127; }
PTHD size_t mix_cell_bytes(int ncomp) { return mix_lin_bytes(ncomp) + (size_t)ncomp * MIX_REM_BYTES; }

struct Species {
    int kind;
    double q, m, mu, gamma, tau, tmax, lmax, Bmax;
    int flipspin;
    int n_spintimes;          /* BMT spin-tracking windows [spintimes[2k], spintimes[2k+1]) */
    double spintimes[8];
    int interpolatefields;    /* precession axis from cubic splines over 11 points of each step */
};

struct Hit {
    double t;
    int tri, leaving, entering;
    double normal[3], v1[3], v2[3], pol1, pol2;
    double pos[3];
};

/* a step's end state for the track log (TLogger::PrintTrack) and the current solid's ID */
struct TrackRec {
    double t, y[9];
    int solid, pad;
};

/* a spin-log row (TLogger::PrintSpin): time, position, spin and BMT precession axis */
struct SpinRec {
    double t, pos[3], spin[3], W[3];
};

/* a particle's state at a snapshot time (TLogger::PrintSnapshot) with its counters then */
struct Snap {
    double t, y[9], spin[5], Hmax, noflip;
    int stop, Nhit, Nstep, Nspinflip;
};

/* resumable per-particle state, structure of arrays */
struct State {
    int n;
    double *t, *y, *xc, *dxc, *ts, *dt, *spin, *Hmax, *noflip, *tau;  /* y,xc,dxc: [9][n]; spin: [5][n] */
    int *stop, *flags, *Nhit, *Nstep, *Nspinflip, *nsolid, *nhitlog, *species;
    long long *n_acc, *n_rej;
    unsigned long long *ctr, *key, *solids;
    Hit *hits;
    int hit_cap;
    const double *loss;       /* per-particle loss per bounce replacing the materials' (nullptr: off) */
    /* snapshots: sorted times, records [n][snap_cap], records per particle */
    const double *snap_t;
    int snap_n, snap_cap, snap_sorted;
    Snap *snaps;
    int *nsnap;
    /* track log: length interval (0: off), records [n][trk_cap], records per particle */
    double trk_interval;
    int trk_cap;
    TrackRec *trk;
    int *ntrk;
    /* spin log: time interval (0: off), records [n][slog_cap], records per particle */
    double slog_interval;
    int slog_cap;
    SpinRec *slog;
    int *nslog;
    double *Bc, *Vc;          /* field at the stepper's current state, [3][n] and [n] (valid if FL_BVALID) */
    double *pend;             /* pending wall-collision piece: x1, y1[9], x2, y2[9] -> [20][n] */
    double *dense;            /* dense-output state of the last step: [DENSE][n] */
    long long *n_field;       /* field-map cell evaluations made by the step kernel */
    int *budget;              /* accepted steps left in the current run call */
    int *queue;               /* indices of particles parked on a wall collision */
    int *qcount;              /* queue length */
    int *work;                /* persistent variant: indices of particles with steps to do */
    int *wcount;              /* [0] work list length, [1] next work item; [WC_CHUNK0 + WC_STRIDE c]:
                                 next item of chunk c of the list split by XCD */
    int *pkey;                /* spatial bucket of each particle (work-list ordering) */
    int *bhist;               /* [NBUCKET + 1] bucket counts, then running offsets */
    int *qkey;                /* candidate count of each queue entry's chord piece (collision order) */
    int *qsorted;             /* the collision queue ordered by qkey */
    int *qhist;               /* [QKEYS + 1] counts of the keys, then running offsets */
};

constexpr int BUCKET_BITS = 5;                     /* 32 buckets per axis */
constexpr int NBUCKET = 1 << (3 * BUCKET_BITS);     /* Morton-ordered spatial buckets */
/* work list split into NXCD contiguous chunks, chunk c taken first by the blocks that run
   on XCD c (workgroups are dispatched to the XCDs round robin, block b on XCD b % NXCD);
   each chunk's counter sits in its own 64-byte line */
constexpr int NXCD = 8, WC_CHUNK0 = 16, WC_STRIDE = 16, WCOUNT_LEN = WC_CHUNK0 + NXCD * WC_STRIDE;
constexpr int DENSE = 59;     /* xo[8] dxo[8] k3..k6[32] dxc[8] t t_old pol */
constexpr int QKEYS = 26;     /* collision-order keys: candidate counts 0..24, 25 = no list */

/* field containers and species held in the parameter block itself, so that the kernels
   read them through the parameter pointer (uniform scalar loads) */
constexpr int MAX_FIELDS = 16;
constexpr int MAX_SPECIES = 8;

struct Params {
    Geom geom;
    int n_fields;
    int b_reuse;              /* 0 if a 2D map tabulates B: its value-only B differs from the one with derivatives */
    Field fields[MAX_FIELDS];
    Species species[MAX_SPECIES];
    double tmax_global;
    State st;
    const int *quit;
    /* the device copy of these parameters, which the kernels read through one pointer
       argument: a kernel argument struct whose address escapes to a non-inlined function
       would be copied into per-lane scratch memory and read back through flat loads */
    const Params *self;
};

/* ---- small helpers ------------------------------------------------------- */

PTD double dabs(double x) { return __builtin_fabs(x); }

/* ---- CGAL Simple_cartesian<double> predicates ---------------------------- */

/* Points are passed by value (P3) through the predicates: a pointer to a per-lane array
   that reached a non-inlined function, or a pointer chosen between two arrays at run
   time, would force the arrays into scratch memory, stored and reloaded for every
   triangle test. */
struct P3 { double x, y, z; };
PTHD P3 p3(const double *v) { return P3{v[0], v[1], v[2]}; }

PTHD int orient3v(P3 p, P3 q, P3 r, P3 s) {
    double a00 = q.x - p.x, a01 = r.x - p.x, a02 = s.x - p.x;
    double a10 = q.y - p.y, a11 = r.y - p.y, a12 = s.y - p.y;
    double a20 = q.z - p.z, a21 = r.z - p.z, a22 = s.z - p.z;
    double m01 = a00 * a11 - a10 * a01;
    double m02 = a00 * a21 - a20 * a01;
    double m12 = a10 * a21 - a20 * a11;
    double det = m01 * a22 - m02 * a12 + m12 * a02;
    return (det > 0) - (det < 0);
}
PTHD int orient3(const double *p, const double *q, const double *r, const double *s) {
    return orient3v(p3(p), p3(q), p3(r), p3(s));
}

PTHD int orient2(double px, double py, double qx, double qy, double rx, double ry) {
    double d = (qx - px) * (ry - py) - (rx - px) * (qy - py);
    return (d > 0) - (d < 0);
}
PTHD int cop_orient(P3 p, P3 q, P3 r) {
    int o = orient2(p.x, p.y, q.x, q.y, r.x, r.y);
    if (o != 0) return o;
    o = orient2(p.y, p.z, q.y, q.z, r.y, r.z);
    if (o != 0) return o;
    return orient2(p.x, p.z, q.x, q.z, r.x, r.z);
}
/* CGAL do_intersect_coplanar(Triangle_3, Segment_3) */
PTHD_NOINLINE int coplanar_seg_tri(P3 A, P3 B, P3 C, P3 p, P3 q) {
    P3 a = A, b = B, c = C;
    if (cop_orient(A, B, C) <= 0) { b = C; c = B; }
    int pqa = cop_orient(p, q, a), pqb = cop_orient(p, q, b), pqc = cop_orient(p, q, c);
    if (pqa > 0) {
        if (pqb > 0) {
            if (pqc > 0) return 0;
            return cop_orient(b, c, q) >= 0 && cop_orient(c, a, p) >= 0;
        }
        if (pqc > 0) return cop_orient(a, b, q) >= 0 && cop_orient(b, c, p) >= 0;
        return cop_orient(a, b, q) >= 0 && cop_orient(c, a, p) >= 0;
    } else if (pqa < 0) {
        if (pqb > 0) {
            if (pqc > 0) return cop_orient(a, b, p) >= 0 && cop_orient(c, a, q) >= 0;
            return cop_orient(a, b, p) >= 0 && cop_orient(b, c, q) >= 0;
        }
        if (pqb < 0) {
            if (pqc < 0) return 0;
            return cop_orient(b, c, p) >= 0 && cop_orient(c, a, q) >= 0;
        }
        if (pqc < 0) return cop_orient(a, b, p) >= 0 && cop_orient(b, c, q) >= 0;
        return cop_orient(a, b, p) >= 0 && cop_orient(c, a, q) >= 0;
    } else {
        if (pqb > 0) {
            if (pqc > 0) return cop_orient(a, b, p) >= 0 && cop_orient(c, a, q) >= 0;
            return cop_orient(a, b, p) >= 0 && cop_orient(b, c, q) >= 0;
        }
        if (pqb < 0) {
            if (pqc < 0) return cop_orient(a, b, q) >= 0 && cop_orient(c, a, p) >= 0;
            return cop_orient(a, b, q) >= 0 && cop_orient(b, c, p) >= 0;
        }
        if (pqc > 0) return cop_orient(b, c, p) >= 0 && cop_orient(c, a, q) >= 0;
        return cop_orient(b, c, q) >= 0 && cop_orient(c, a, p) >= 0;
    }
}

/* CGAL intersection(Triangle_3, Segment_3): 1 hit (point in pt), 0 none, -1 coplanar overlap.
   CGAL's case analysis on the sides of p and q, without its branches: in every case with
   an end off the plane the segment meets the triangle iff the three edges are oriented
   non-positively about the segment directed from p to q (p above the plane, or p on it
   and q below) or from q to p (otherwise); the same orientation predicates are evaluated
   on the same points, so each case's decision is CGAL's, and a wave whose lanes fall in
   different cases runs one set of predicates instead of one per case. */
PTHD int seg_tri(const double *a_, const double *b_, const double *c_, const double *p_, const double *q_, double pt[3]) {
    const P3 a = p3(a_), b = p3(b_), c = p3(c_), p = p3(p_), q = p3(q_);
    const int abcp = orient3v(a, b, c, p);
    const int abcq = orient3v(a, b, c, q);
    if (abcp == 0 && abcq == 0) return coplanar_seg_tri(a, b, c, p, q) ? -1 : 0;
    const bool apart = (abcp > 0 && abcq > 0) || (abcp < 0 && abcq < 0);
    const bool fwd = abcp > 0 || (abcp == 0 && abcq < 0);
    const P3 u = fwd ? p : q, w = fwd ? q : p;
    const int o1 = orient3v(u, w, a, b), o2 = orient3v(u, w, b, c), o3 = orient3v(u, w, c, a);
    if (apart || o1 > 0 || o2 > 0 || o3 > 0) return 0;
    if (abcq == 0) { pt[0] = q.x; pt[1] = q.y; pt[2] = q.z; return 1; }
    if (abcp == 0) { pt[0] = p.x; pt[1] = p.y; pt[2] = p.z; return 1; }
    double rpx = a.x - c.x, rpy = a.y - c.y, rpz = a.z - c.z;
    double rqx = b.x - c.x, rqy = b.y - c.y, rqz = b.z - c.z;
    double pa = rpy * rqz - rqy * rpz;
    double pb = rpz * rqx - rqz * rpx;
    double pc = rpx * rqy - rqx * rpy;
    double pd = -pa * c.x - pb * c.y - pc * c.z;
    double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    double num = pa * p.x + pb * p.y + pc * p.z + pd;
    double den = pa * dx + pb * dy + pc * dz;
    if (den == 0) return 0;
    double hx = den * p.x - num * dx, hy = den * p.y - num * dy, hz = den * p.z - num * dz;
    if (den != 1.0) { pt[0] = hx / den; pt[1] = hy / den; pt[2] = hz / den; }
    else { pt[0] = hx; pt[1] = hy; pt[2] = hz; }
    return 1;
}

/* CGAL do_intersect(Triangle_3, Ray_3) for the +z ray of TTriangleMesh::GetSolids */
PTD int ray_tri(const double *a, const double *b, const double *c, const double *p) {
    double q[3] = {p[0], p[1], p[2] + 1.0};
    double ad[3] = {a[0], a[1], a[2] + 1.0};
    int dir = orient3(a, b, c, ad);
    if (dir == 0) return 0;
    int abcp = orient3(a, b, c, p);
    if (abcp > 0) {
        if (dir > 0) return 0;
        return !(orient3(p, q, a, b) > 0 || orient3(p, q, b, c) > 0 || orient3(p, q, c, a) > 0);
    } else if (abcp < 0) {
        if (dir < 0) return 0;
        return !(orient3(q, p, a, b) > 0 || orient3(q, p, b, c) > 0 || orient3(q, p, c, a) > 0);
    }
    if (dir > 0) return orient3(q, p, a, b) <= 0 && orient3(q, p, b, c) <= 0 && orient3(q, p, c, a) <= 0;
    return orient3(p, q, a, b) <= 0 && orient3(p, q, b, c) <= 0 && orient3(p, q, c, a) <= 0;
}

/* CGAL do_intersect(Segment_3, Bbox_3) (Bbox_3_Segment_3_do_intersect.h, bounded, unfiltered) */
PTD int seg_bbox(const double *P, const double *Q, const double *bb) {
    double px = P[0], py = P[1], pz = P[2], qx = Q[0], qy = Q[1], qz = Q[2];
    double bxmin = bb[0], bymin = bb[1], bzmin = bb[2], bxmax = bb[3], bymax = bb[4], bzmax = bb[5];
    if ((px >= bxmin && px <= bxmax && py >= bymin && py <= bymax && pz >= bzmin && pz <= bzmax) ||
        (qx >= bxmin && qx <= bxmax && qy >= bymin && qy <= bymax && qz >= bzmin && qz <= bzmax))
        return 1;
    double dmin, tmin, tmax, dmax;
    if (qx >= px) {
        if (px > bxmax || qx < bxmin) return 0;
        if (bxmax > qx) { tmax = 1; dmax = 1; } else { tmax = bxmax - px; dmax = qx - px; }
        tmin = bxmin - px; dmin = qx - px;
    } else {
        if (qx > bxmax || px < bxmin) return 0;
        if (bxmin < qx) { tmax = 1; dmax = 1; } else { tmax = px - bxmin; dmax = px - qx; }
        tmin = px - bxmax; dmin = px - qx;
    }
    if (tmin < 0) tmin = 0;
    double dymin, tymin, tymax, dymax;
    if (qy >= py) {
        if (py > bymax || qy < bymin) return 0;
        if (bymax > qy) { tymax = 1; dymax = 1; } else { tymax = bymax - py; dymax = qy - py; }
        tymin = bymin - py; dymin = qy - py;
    } else {
        if (qy > bymax || py < bymin) return 0;
        if (bymin < qy) { tymax = 1; dymax = 1; } else { tymax = py - bymin; dymax = py - qy; }
        tymin = py - bymax; dymin = py - qy;
    }
    if (tymin < 0) tymin = 0;
    if (py != qy && px != qx) {
        if (dymax * tmin > dmin * tymax) return 0;
        if (dmax * tymin > dymin * tmax) return 0;
    }
    if (px == qx || (py != qy && dmin * tymin > dymin * tmin)) { tmin = tymin; dmin = dymin; }
    if (px == qx || (py != qy && dymax * tmax > dmax * tymax)) { tmax = tymax; dmax = dymax; }
    double dzmin, tzmin, tzmax, dzmax;
    if (qz >= pz) {
        if (pz > bzmax || qz < bzmin) return 0;
        if (bzmax > qz) { tzmax = 1; dzmax = 1; } else { tzmax = bzmax - pz; dzmax = qz - pz; }
        tzmin = bzmin - pz; dzmin = qz - pz;
    } else {
        if (qz > bzmax || pz < bzmin) return 0;
        if (bzmin < qz) { tzmax = 1; dzmax = 1; } else { tzmax = pz - bzmin; dzmax = pz - qz; }
        tzmin = pz - bzmax; dzmin = pz - qz;
    }
    if (tzmin < 0) tzmin = 0;
    if ((px != qx || py != qy) && pz != qz) {
        if (dzmax * tmin > dmin * tzmax) return 0;
        if (dmax * tzmin > dzmin * tmax) return 0;
    }
    return 1;
}

/* conservative segment-vs-node test for BVH culling: the box is widened so that
   rounding never rejects a node holding a triangle the exact test would hit */
PTD bool node_hit(const BvhNode &nd, const double *p, const double *inv, const double *d) {
    double t0 = 0.0, t1 = 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double lo = (double)nd.lo[k], hi = (double)nd.hi[k];
        double pad = 1e-9 + 1e-7 * (dabs(lo) + dabs(hi));
        lo -= pad; hi += pad;
        if (d[k] == 0.0) {
            if (p[k] < lo || p[k] > hi) return false;
        } else {
            double ta = (lo - p[k]) * inv[k], tb = (hi - p[k]) * inv[k];
            if (ta > tb) { double s = ta; ta = tb; tb = s; }
            t0 = ta > t0 ? ta : t0;
            t1 = tb < t1 ? tb : t1;
            if (t0 > t1) return false;
        }
    }
    return true;
}

PTD void load_tri(const Tri *T, double a[3], double b[3], double c[3], int &solid, int &orig) {
    const float4 *q = reinterpret_cast<const float4 *>(T);
    float4 w0 = q[0], w1 = q[1], w2 = q[2];
    a[0] = w0.x; a[1] = w0.y; a[2] = w0.z;
    b[0] = w0.w; b[1] = w1.x; b[2] = w1.y;
    c[0] = w1.z; c[1] = w1.w; c[2] = w2.x;
    solid = __float_as_int(w2.y);
    orig = __float_as_int(w2.z);
}

PTD void load_node(const BvhNode *N, BvhNode &nd) {
    const float4 *q = reinterpret_cast<const float4 *>(N);
    float4 w0 = q[0], w1 = q[1];
    nd.lo[0] = w0.x; nd.lo[1] = w0.y; nd.lo[2] = w0.z; nd.a = __float_as_int(w0.w);
    nd.hi[0] = w1.x; nd.hi[1] = w1.y; nd.hi[2] = w1.z; nd.b = __float_as_int(w1.w);
}

/* sorted intersection list of one segment: (s asc, solid ID desc, input index asc) */
struct HitList {
    int n;
    double s[MAX_HITS];
    int tri[MAX_HITS];     /* leaf-order index */
    int orig[MAX_HITS];
    unsigned id[MAX_HITS];
};

/*
 * TGeometry::GetCollisions (src/geometry.cpp:139-148) over the BVH.
 * list == nullptr: count only.  Returns the number of intersections, or
 * -1 (coplanar overlap: the reference throws) / -2 (more than MAX_HITS).
 */
PTD_NOINLINE int2 seg_query_v(const Geom &G, double p0, double p1, double p2, double q0, double q1, double q2, HitList *list,
                              int start = 0) {
    const double p[3] = {p0, p1, p2}, q[3] = {q0, q1, q2};
    int ntests = 0;
    double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    double sl = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    /* the subtree of start: the nodes from start up to its escape index */
    int node = start, end = G.n_node;
    int count = 0;
    while (node < end) {
        BvhNode nd;
        load_node(G.node + node, nd);
        if (node == start && start > 0) end = nd.a >= 0 ? nd.a : start + 1;
        const bool hit = node_hit(nd, p, inv, d);
        if (nd.a >= 0) {
            node = hit ? node + 1 : nd.a;
            continue;
        }
        if (hit) {
            {
                int first = -nd.a - 1;
                for (int k = 0; k < nd.b; ++k) {
                    double a[3], b[3], c[3], pt[3];
                    int solid, orig;
                    load_tri(G.tri + first + k, a, b, c, solid, orig);
                    ++ntests;
                    int r = seg_tri(a, b, c, p, q, pt);
                    if (r == 0) continue;
                    if (r < 0) return make_int2(-1, ntests);
                    if (count >= MAX_HITS) return make_int2(-2, ntests);
                    ++count;
                    if (list) {
                        double v[3] = {pt[0] - p[0], pt[1] - p[1], pt[2] - p[2]};
                        double s = (v[0] * d[0] + v[1] * d[1] + v[2] * d[2]) / sl;
                        unsigned id = G.solid_id[solid];
                        int j = list->n;
                        while (j > 0) {
                            double so = list->s[j - 1];
                            unsigned io = list->id[j - 1];
                            bool before = (s == so) ? (id > io || (id == io && orig < list->orig[j - 1])) : (s < so);
                            if (!before) break;
                            list->s[j] = so; list->id[j] = io; list->tri[j] = list->tri[j - 1]; list->orig[j] = list->orig[j - 1];
                            --j;
                        }
                        list->s[j] = s; list->id[j] = id; list->tri[j] = first + k; list->orig[j] = orig;
                        list->n++;
                    }
                }
            }
        }
        ++node;
    }
    return make_int2(count, ntests);
}

/* Uniform (scalar-path) forms of the node and triangle loads: with a wave-uniform index
   the read-only BVH is read through the constant address space, i.e. the scalar data cache,
   which the field-map gathers of the vector path do not congest. */
typedef const __attribute__((address_space(4))) pt_f4v pt_cf4v;
PTD void load_node_u(const BvhNode *N, int idx, BvhNode &nd) {
    pt_cf4v *q = (pt_cf4v *)(reinterpret_cast<const float4 *>(N) + 2 * (size_t)idx);
    const pt_f4v w0 = q[0], w1 = q[1];
    nd.lo[0] = w0.x; nd.lo[1] = w0.y; nd.lo[2] = w0.z; nd.a = __float_as_int(w0.w);
    nd.hi[0] = w1.x; nd.hi[1] = w1.y; nd.hi[2] = w1.z; nd.b = __float_as_int(w1.w);
}
PTD void load_tri_u(const Tri *T, int idx, double a[3], double b[3], double c[3]) {
    pt_cf4v *q = (pt_cf4v *)(reinterpret_cast<const float4 *>(T) + 3 * (size_t)idx);
    const pt_f4v w0 = q[0], w1 = q[1], w2 = q[2];
    a[0] = w0.x; a[1] = w0.y; a[2] = w0.z;
    b[0] = w0.w; b[1] = w1.x; b[2] = w1.y;
    c[0] = w1.z; c[1] = w1.w; c[2] = w2.x;
}

template <class T> PTD const T *uniform_ptr(const T *p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return reinterpret_cast<const T *>(((unsigned long long)hi << 32) | lo);
}

/* seg_query_v's count for a segment and start node that are the same on every lane of the
   wave: the traversal's control flow is uniform and its loads take the scalar path */
PTD int2 seg_count_uniform(const BvhNode *nodes, const Tri *tris, int n_node, const double p[3], const double q[3],
                           int start) {
    int ntests = 0;
    const double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    const double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    /* every lane computes the same values; readfirstlane tells the compiler so, which keeps
       the control flow and the load addresses scalar */
    int node = __builtin_amdgcn_readfirstlane(start), end = n_node;
    int count = 0;
    while (node < end) {
        BvhNode nd;
        load_node_u(nodes, node, nd);
        if (node == start && start > 0) end = nd.a >= 0 ? nd.a : start + 1;
        const bool hit = __builtin_amdgcn_readfirstlane((int)node_hit(nd, p, inv, d)) != 0;
        if (nd.a >= 0) {
            node = __builtin_amdgcn_readfirstlane(hit ? node + 1 : nd.a);
            continue;
        }
        if (hit) {
            const int first = -nd.a - 1;
            for (int k = 0; k < nd.b; ++k) {
                double a[3], b[3], c[3], pt[3];
                load_tri_u(tris, first + k, a, b, c);
                ++ntests;
                const int r = __builtin_amdgcn_readfirstlane(seg_tri(a, b, c, p, q, pt));
                if (r == 0) continue;
                if (r < 0) return make_int2(-1, ntests);
                if (count >= MAX_HITS) return make_int2(-2, ntests);
                ++count;
            }
        }
        node = __builtin_amdgcn_readfirstlane(node + 1);
    }
    return make_int2(count, ntests);
}

PTD double readlane_d(double v, int l) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* The counts of the active lanes' segments p -> q from their start nodes: the lanes take
   turns, each one's traversal run uniformly by the wave (seg_count_uniform) */
PTD_NOINLINE int2 seg_count_lanes(const Geom &G, double p0, double p1, double p2, double q0, double q1, double q2,
                                  int start) {
    int2 mine = make_int2(0, 0);
    unsigned long long pend = __builtin_amdgcn_read_exec();
    const int me = (int)__lane_id();
    const BvhNode *nodes = uniform_ptr(G.node);
    const Tri *tris = uniform_ptr(G.tri);
    const int n_node = __builtin_amdgcn_readfirstlane(G.n_node);
    while (pend) {
        const int l = __builtin_ctzll(pend);
        pend &= pend - 1;
        const double p[3] = {readlane_d(p0, l), readlane_d(p1, l), readlane_d(p2, l)};
        const double q[3] = {readlane_d(q0, l), readlane_d(q1, l), readlane_d(q2, l)};
        const int st = __builtin_amdgcn_readlane(start, l);
        const int2 r = seg_count_uniform(nodes, tris, n_node, p, q, st);
        if (me == l) mine = r;
    }
    return mine;
}

/*
 * Candidate list of one chord piece: the triangles whose bounding boxes meet a box
 * that contains every sub-segment the collision search can test on that piece,
 * collected once from the BVH in its depth-first leaf order.  A query whose two
 * end points lie in the box is answered from the list: the box is convex, so the
 * segment lies in it and only candidates can meet it; visiting them in BVH leaf
 * order keeps the first error (coplanar overlap, too many hits) the same as the
 * full traversal's.  n < 0: no list (too many candidates).
 */
constexpr int MAX_CAND = 24;
static_assert(QKEYS == MAX_CAND + 2, "a key per candidate count and one for no list");
struct Cand {
    int n;
    double lo[3], hi[3];
    int idx[MAX_CAND];
};

/* The vertices of a lane's first CAND_LDS candidates, staged in LDS by cand_build for the
   bisection's counting queries (cand_lds[k][vertex coordinate][lane], 36 B per candidate
   and lane: 36 KB for a block of CAND_LANES threads, four blocks per CU): an LDS read
   instead of an index and a triangle fetched from memory for every candidate test.  Only
   kernels whose blocks have at most CAND_LANES threads may reach cand_build. */
constexpr int CAND_LDS = 4, CAND_LANES = 256;
__shared__ float cand_lds[CAND_LDS][9][CAND_LANES];

/* the candidates of a chord piece: triangles whose padded box meets [lo, hi], in BVH leaf
   order; stage = false counts them without staging their vertices in LDS */
PTD void cand_build(const Geom &G, const double *lo, const double *hi, Cand &C, bool stage = true) {
    C.n = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { C.lo[k] = lo[k]; C.hi[k] = hi[k]; }
    int node = 0;
    while (node < G.n_node) {
        BvhNode nd;
        load_node(G.node + node, nd);
        bool hit = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double pad = 1e-9 + 1e-7 * (dabs((double)nd.lo[k]) + dabs((double)nd.hi[k]));
            hit = hit && ((double)nd.lo[k] - pad <= hi[k]) && ((double)nd.hi[k] + pad >= lo[k]);
        }
        if (nd.a >= 0) {
            node = hit ? node + 1 : nd.a;
            continue;
        }
        if (hit) {
            {
                int first = -nd.a - 1;
                for (int k = 0; k < nd.b; ++k) {
                    double a[3], b[3], c[3];
                    int solid, orig;
                    load_tri(G.tri + first + k, a, b, c, solid, orig);
                    bool in = true;
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) {
                        double tlo = a[ax] < b[ax] ? a[ax] : b[ax];
                        tlo = tlo < c[ax] ? tlo : c[ax];
                        double thi = a[ax] > b[ax] ? a[ax] : b[ax];
                        thi = thi > c[ax] ? thi : c[ax];
                        double pad = 1e-9 + 1e-7 * (dabs(tlo) + dabs(thi));
                        in = in && (tlo - pad <= hi[ax]) && (thi + pad >= lo[ax]);
                    }
                    if (!in) continue;
                    if (C.n >= MAX_CAND) { C.n = -1; return; }
                    if (stage && C.n < CAND_LDS) {
                        const int l = threadIdx.x;
                        cand_lds[C.n][0][l] = (float)a[0]; cand_lds[C.n][1][l] = (float)a[1]; cand_lds[C.n][2][l] = (float)a[2];
                        cand_lds[C.n][3][l] = (float)b[0]; cand_lds[C.n][4][l] = (float)b[1]; cand_lds[C.n][5][l] = (float)b[2];
                        cand_lds[C.n][6][l] = (float)c[0]; cand_lds[C.n][7][l] = (float)c[1]; cand_lds[C.n][8][l] = (float)c[2];
                    }
                    C.idx[C.n++] = first + k;
                }
            }
        }
        ++node;
    }
}

/* the box CheckHit's candidates are collected in for the chord piece y1 -> y2: every sub-segment
   of the search has its ends on the dense-output path between y1 and y2; a path of length l with
   chord d strays at most sqrt(l^2 - d^2)/2 from the chord, so the box of the chord widened by
   twice that contains them (ends outside it are detected and fall back to the BVH) */
PTD void piece_cand_box(const double *y1, const double *y2, double lo[3], double hi[3]) {
    double l2 = (y2[8] - y1[8]) * (y2[8] - y1[8]);
    double d2 = (y2[0] - y1[0]) * (y2[0] - y1[0]) + (y2[1] - y1[1]) * (y2[1] - y1[1]) + (y2[2] - y1[2]) * (y2[2] - y1[2]);
    double dev = l2 > d2 ? __builtin_sqrt(l2 - d2) : 0.0;
    double w = dev + 1e-6 + 1e-9 * __builtin_sqrt(d2);
    for (int k = 0; k < 3; ++k) {
        lo[k] = (y1[k] < y2[k] ? y1[k] : y2[k]) - w;
        hi[k] = (y1[k] < y2[k] ? y2[k] : y1[k]) + w;
    }
}

PTD bool cand_covers(const Cand &C, const double *p, const double *q) {
    if (C.n < 0) return false;
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        in = in && p[k] >= C.lo[k] && p[k] <= C.hi[k] && q[k] >= C.lo[k] && q[k] <= C.hi[k];
    return in;
}

/* seg_query_v restricted to the candidates (same tests, same sorted list, same errors) */
PTD int cand_query(const Geom &G, const Cand &C, const double *p, const double *q, HitList *list, long long *tests) {
    double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
    double sl = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    int count = 0;
    /* the test counter is added once, not updated in memory per candidate */
    *tests += C.n > 0 ? C.n : 0;
    for (int k = 0; k < C.n; ++k) {
        double a[3], b[3], c[3], pt[3];
        int solid = 0, orig = 0, t = 0;
        if (!list && k < CAND_LDS) {
            /* a counting query: the vertices staged by cand_build (the STL's floats, exact) */
            const int l = threadIdx.x;
            a[0] = cand_lds[k][0][l]; a[1] = cand_lds[k][1][l]; a[2] = cand_lds[k][2][l];
            b[0] = cand_lds[k][3][l]; b[1] = cand_lds[k][4][l]; b[2] = cand_lds[k][5][l];
            c[0] = cand_lds[k][6][l]; c[1] = cand_lds[k][7][l]; c[2] = cand_lds[k][8][l];
        } else {
            t = C.idx[k];
            load_tri(G.tri + t, a, b, c, solid, orig);
        }
        int r = seg_tri(a, b, c, p, q, pt);
        if (r == 0) continue;
        if (r < 0) return -1;
        if (count >= MAX_HITS) return -2;
        ++count;
        if (list) {
            double v[3] = {pt[0] - p[0], pt[1] - p[1], pt[2] - p[2]};
            double s = (v[0] * d[0] + v[1] * d[1] + v[2] * d[2]) / sl;
            unsigned id = G.solid_id[solid];
            int j = list->n;
            while (j > 0) {
                double so = list->s[j - 1];
                unsigned io = list->id[j - 1];
                bool before = (s == so) ? (id > io || (id == io && orig < list->orig[j - 1])) : (s < so);
                if (!before) break;
                list->s[j] = so; list->id[j] = io; list->tri[j] = list->tri[j - 1]; list->orig[j] = list->orig[j - 1];
                --j;
            }
            list->s[j] = s; list->id[j] = id; list->tri[j] = t; list->orig[j] = orig;
            list->n++;
        }
    }
    return count;
}

PTD int seg_query(const Geom &G, const double *p, const double *q, HitList *list, long long *tests) {
    int2 r = seg_query_v(G, p[0], p[1], p[2], q[0], q[1], q[2], list);
    *tests += r.y;
    return r.x;
}

/* +z ray parity per solid (TTriangleMesh::GetSolids, include/trianglemesh.h:181-188):
   bit m of the result is set if the ray from p crosses solid m an odd number of times.
   *overflow reports a traversal that could not visit every node; the stackless traversal
   always can, so it is 0 (kept for the callers' checks). */
PTD_NOINLINE unsigned long long ray_parity(const Geom &G, double p0, double p1, double p2, int *overflow) {
    const double p[3] = {p0, p1, p2};
    unsigned long long par = 0;
    *overflow = 0;
    int node = 0;
    while (node < G.n_node) {
        BvhNode nd;
        load_node(G.node + node, nd);
        double pad0 = 1e-9 + 1e-7 * (dabs(nd.lo[0]) + dabs(nd.hi[0]));
        double pad1 = 1e-9 + 1e-7 * (dabs(nd.lo[1]) + dabs(nd.hi[1]));
        double pad2 = 1e-9 + 1e-7 * (dabs(nd.lo[2]) + dabs(nd.hi[2]));
        bool hit = p[0] >= (double)nd.lo[0] - pad0 && p[0] <= (double)nd.hi[0] + pad0 &&
                   p[1] >= (double)nd.lo[1] - pad1 && p[1] <= (double)nd.hi[1] + pad1 &&
                   p[2] <= (double)nd.hi[2] + pad2;
        if (nd.a >= 0) {
            node = hit ? node + 1 : nd.a;
            continue;
        }
        if (hit) {
            int first = -nd.a - 1;
            for (int k = 0; k < nd.b; ++k) {
                double a[3], b[3], c[3];
                int solid, orig;
                load_tri(G.tri + first + k, a, b, c, solid, orig);
                if (ray_tri(a, b, c, p)) par ^= (1ull << solid);
            }
        }
        ++node;
    }
    return par;
}

/* ---- solid stack: up to MAX_SOLIDS entries packed in one 64-bit word -------
   entry = (solid index + 1) | ignored << 7; index 0 stands for the default solid */
PTHD int sk_sidx(unsigned long long s, int i) { return (int)((s >> (8 * i)) & 0x7f) - 1; }
PTHD int sk_ign(unsigned long long s, int i) { return (int)((s >> (8 * i + 7)) & 1); }
PTHD unsigned long long sk_set(unsigned long long s, int i, int sidx, int ign) {
    unsigned long long e = (unsigned long long)((sidx + 1) | (ign << 7));
    return (s & ~(0xffull << (8 * i))) | (e << (8 * i));
}
PTHD unsigned long long sk_erase(unsigned long long s, int i) {
    unsigned long long low = (i == 0) ? 0ull : (s & ((1ull << (8 * i)) - 1));
    unsigned long long high = (i >= 7) ? 0ull : ((s >> (8 * (i + 1))) << (8 * i));
    return low | high;
}

PTD unsigned solid_ID(const Geom &G, int sidx) { return sidx < 0 ? 1u : G.solid_id[sidx]; }
PTD const double *solid_mat(const Geom &G, int sidx) {
    int m = sidx < 0 ? G.default_mat : G.solid_mat[sidx];
    return G.mats + 9 * m;
}
PTD int solid_ignored(const Geom &G, int sidx, double t) {
    if (sidx < 0) return 0;
    for (int k = G.ign_off[sidx]; k < G.ign_off[sidx + 1]; ++k)
        if (t >= G.ign[2 * k] && t < G.ign[2 * k + 1]) return 1;
    return 0;
}
/* TTracker::GetCurrentsolid (std::max_element with the reference comparator) */
PTD int current_solid(const Geom &G, unsigned long long s, int n) {
    if (n <= 1) return sk_sidx(s, 0);
    int best = 0;
    unsigned best_id = solid_ID(G, sk_sidx(s, 0));
#pragma unroll
    for (int i = 1; i < MAX_SOLIDS; ++i) {
        if (i < n) {
            const unsigned id = solid_ID(G, sk_sidx(s, i));
            if (sk_ign(s, best) || (!sk_ign(s, i) && best_id < id)) { best = i; best_id = id; }
        }
    }
    return sk_sidx(s, best);
}

/* true if segment p->q provably misses every triangle (clearance grid).  A cell's entry is
   a lower bound of d - r, d the distance from the cell's centre c to the nearest triangle
   and r half the cell's diagonal (0 when that is not positive).  Every point of the segment
   lies within |m - c| + L/2 of c (m its midpoint, L its length), so the segment misses the
   walls when that is below d; the entry plus r stands for d, with a relative and an
   absolute margin for the rounding of the entry, the centre and the distances.  cell[]
   returns the midpoint's cell (cell[0] = -1 outside the grid). */
PTD bool clear_of_walls(const Geom &G, const double *p, const double *q, int cell[3]) {
    cell[0] = -1;
    if (G.clear == nullptr) return false;
    int c[3];
    double m[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        m[a] = 0.5 * (p[a] + q[a]);
        double u = (m[a] - G.glo[a]) * G.ginv_h[a];
        if (!(u >= 0.0) || u >= (double)G.gn[a]) return false;
        c[a] = (int)u;
    }
    cell[0] = c[0]; cell[1] = c[1]; cell[2] = c[2];
    const double bound = (double)G.clear[((size_t)c[0] * G.gn[1] + c[1]) * G.gn[2] + c[2]];
    if (!(bound > 0.0)) return false;
    double r2 = 0.0, o2 = 0.0, L2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double h = 1.0 / G.ginv_h[a];
        const double o = m[a] - (G.glo[a] + ((double)c[a] + 0.5) * h);
        const double d = q[a] - p[a];
        r2 += h * h;
        o2 += o * o;
        L2 += d * d;
    }
    const double reach = __builtin_sqrt(o2) + 0.5 * __builtin_sqrt(L2);
    return reach < 0.999 * (bound + 0.5 * __builtin_sqrt(r2)) - 1e-9;
}

PTD bool clear_of_walls(const Geom &G, const double *p, const double *q) {
    int cell[3];
    return clear_of_walls(G, p, q, cell);
}

/* p and q lie in cell c[] widened by pad on every side */
PTD bool cell_covers(const Geom &G, const int c[3], double pad, const double *p, const double *q) {
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double h = 1.0 / G.ginv_h[a];
        const double lo = G.glo[a] + (double)c[a] * h - pad, hi = G.glo[a] + (double)(c[a] + 1) * h + pad;
        in = in && p[a] >= lo && p[a] <= hi && q[a] >= lo && q[a] <= hi;
    }
    return in;
}

/* the candidate list of cell c[] when it exists and its widened cell contains p and q
   (then it contains the segment and every triangle the segment can meet is on the list):
   returns (first, count), count < 0 when there is none */
PTD int2 cell_list(const Geom &G, const int c[3], const double *p, const double *q) {
    if (G.cl_rng == nullptr || c[0] < 0 || !cell_covers(G, c, G.cl_pad, p, q)) return make_int2(0, -1);
    return G.cl_rng[((size_t)c[0] * G.gn[1] + c[1]) * G.gn[2] + c[2]];
}

/* the BVH node to start the segment's traversal from: the root of cell c[]'s triangles when
   its widened cell contains p and q (a subtree holds a contiguous run of the leaf order, so
   its traversal tests those triangles in the same order; -2: none to test), else 0 */
PTD int cell_root(const Geom &G, const int c[3], const double *p, const double *q) {
    if (G.cl_root == nullptr || c[0] < 0 || !cell_covers(G, c, G.rt_pad, p, q)) return 0;
    return G.cl_root[((size_t)c[0] * G.gn[1] + c[1]) * G.gn[2] + c[2]];
}

/* seg_query_v's count (and its first error, the list being in BVH leaf order) from a cell's
   candidate list; the triangles are fetched two at a time (four or eight at a time keep more
   registers live across the call and cost the step kernel 1.5% and 20%) */
PTD_NOINLINE int2 cell_list_count(const Geom &G, int first, int n, double p0, double p1, double p2, double q0,
                                  double q1, double q2) {
    const double p[3] = {p0, p1, p2}, q[3] = {q0, q1, q2};
    int count = 0;
    for (int k0 = 0; k0 < n; k0 += 2) {
        float4 w[2][3];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (k0 + j < n) {
                const float4 *t = reinterpret_cast<const float4 *>(G.cl_tri + first + k0 + j);
                w[j][0] = t[0]; w[j][1] = t[1]; w[j][2] = t[2];
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (k0 + j < n) {
                const double a[3] = {w[j][0].x, w[j][0].y, w[j][0].z};
                const double b[3] = {w[j][0].w, w[j][1].x, w[j][1].y};
                const double c[3] = {w[j][1].z, w[j][1].w, w[j][2].x};
                double pt[3];
                const int r = seg_tri(a, b, c, p, q, pt);
                if (r == 0) continue;
                if (r < 0) return make_int2(-1, n);
                if (count >= MAX_HITS) return make_int2(-2, n);
                ++count;
            }
        }
    }
    return make_int2(count, n);
}

/* TGeometry::CheckSegment -> TTriangleMesh::InBoundingBox */
PTD bool check_segment(const Geom &G, const double *p1, const double *p2) {
    for (int m = 0; m < G.n_solid; ++m)
        if (seg_bbox(p1, p2, G.mesh_bbox + 6 * m)) return true;
    return false;
}

} // namespace ptd

#endif /* PT_DEVICE_H_ */
