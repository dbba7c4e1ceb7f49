/*
 * pt_track.hip -- the tracking kernels of the MI355X engine.
 *
 * One thread per particle runs TTracker::IntegrateParticle's loop
 * (src/tracking.cpp:17-173) on resumable state kept in HBM.  A run call is a
 * few rounds of two kernels:
 *  - the step kernel keeps the dopri5 stepper in registers (the polarisation,
 *    whose derivative is identically zero, is carried outside the stage
 *    vectors) and finishes every step whose chord pieces touch no wall; a
 *    particle whose chord piece meets the mesh parks its piece and dense-output
 *    state in HBM and enters the collision queue;
 *  - the collision kernel runs the queued particles, compacted into full waves,
 *    through CheckHit's bisection, DoHit and OnHit and the rest of their step.
 * Built with -ffp-contract=off: the FP64 path evaluates every expression in
 * the reference's order, with the reference's long double constants rounded
 * to double, so results match the FP64 CPU restatement bit for bit.
 */
#include "../../include/pentrack_gpu.h"
#include "pt_decay.h"
#include "pt_device.h"
#include "pt_formula.h"
#include "pt_kernels.h"

/* This file is compiled three times per kernel mode (-DPT_MODE_TU=0..3 with
   -DPT_KERNEL_SET=0: the mode's set-up and step kernels, =2: its persistent step kernel,
   =1: its collision kernel and field queries) and once
   without PT_MODE_TU (the mode-independent kernels and the launchers that dispatch on
   the mode), so the modes build in parallel. */
#ifndef PT_MODE_TU
#define PT_MODE_TU (-1)
#define PT_COMMON_TU 1
#else
#define PT_COMMON_TU 0
#endif
#ifndef PT_KERNEL_SET
#define PT_KERNEL_SET 0
#endif

namespace ptd {

/* ------------------------------------------------------------------------ */
/* fields                                                                    */
/* ------------------------------------------------------------------------ */

/* The cell k of grid g (n strictly increasing nodes) with g[k] <= p < g[k+1], or -1 when
   p lies outside [g[0], g[n-1]) or is NaN: a binary search, for the rare points where the
   uniform-spacing guess of tric_locate is off (rounding at a cell edge, a non-uniform grid) */
PTD_NOINLINE int grid_cell_search(const double *g, int n, double p) {
    if (!(g[0] <= p && p < g[n - 1])) return -1;
    int lo = 0, hi = n - 1;   /* g[lo] <= p < g[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g[mid] <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* TabField3::Interpolate cell search (std::upper_bound semantics, src/field_3d.cpp:490-500):
   the unique cell whose nodes bracket the point, from the uniform-spacing guess or, when
   the guess does not bracket it, from grid_cell_search */
PTD bool tric_locate(const Field &f, double x, double y, double z, long &cell, double r[3], double dist[3]) {
    double p[3] = {x, y, z};
    int n[3] = {f.nx, f.ny, f.nz};
    int idx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double *g = f.g[a];
        double gs = (p[a] - f.x0[a]) * f.inv_h[a];
        int k;
        if (!(gs >= 0.0)) k = 0;
        else if (gs >= (double)(n[a] - 2)) k = n[a] - 2;
        else k = (int)gs;
        double gk = gload(g + k), gk1 = gload(g + k + 1);
        if (__builtin_expect(!(gk <= p[a] && p[a] < gk1), 0)) {
            k = grid_cell_search(g, n[a], p[a]);
            if (k < 0) return false;
            gk = gload(g + k);
            gk1 = gload(g + k + 1);
        }
        idx[a] = k;
        dist[a] = gk1 - gk;
        r[a] = (p[a] - gk) / dist[a];
    }
    cell = ((long)idx[0] * (f.ny - 1) + idx[1]) * (f.nz - 1) + idx[2];
    return true;
}

/*
 * Value and gradient of one tricubic block: the four Horner evaluations of
 * tricubic_eval_fast (src/field_3d.cpp:36-64, 82-121) fused over one pass of the
 * coefficients.  Multiplications by fact[..][0] = 1 are exact, so sharing the
 * inner z- and y-Horner partial sums between value and derivatives leaves every
 * result bit-identical to the four separate evaluations.
 * Device layout of a block: a[16 i + 4 j + k] holds the coefficient of x^i y^j z^k,
 * so the 16 coefficients of one power of x are one contiguous 128-byte chunk.
 */
PTD void load_line64(const double *__restrict__ a, double c[16]) {
    const double2 *c2 = reinterpret_cast<const double2 *>(a);
#pragma unroll
    for (int q = 0; q < 8; ++q) { double2 v = gload(c2 + q); c[2 * q] = v.x; c[2 * q + 1] = v.y; }
}

/* one row j (the coefficients of x^i y^j z^0..3) of the z- and y-Horner sums of an x-power line */
PTD void tric_row(double a0, double a1, double a2, double a3, int j, double y, double z, bool grad, double &rY,
                  double &rYd, double &rYz) {
    double rz = 0.;
    rz += a3; rz *= z;
    rz += a2; rz *= z;
    rz += a1; rz *= z;
    rz += a0;
    rY += rz;
    if (grad) {
        double rzd = 0.;
        rzd += 3.0 * a3; rzd *= z;
        rzd += 2.0 * a2; rzd *= z;
        rzd += a1;
        rYz += rzd;
        if (j > 0) {
            rYd += (double)j * rz;
            if (j > 1) rYd *= y;
        }
        if (j > 0) rYz *= y;
    }
    if (j > 0) rY *= y;
}

/* the z- and y-Horner sums of one x-power line (the 16 coefficients of x^i), rows 3 to 0 */
PTD void tric_line(const double c[16], double y, double z, bool grad, double &rYo, double &rYdo, double &rYzo) {
    double rY = 0., rYd = 0., rYz = 0.;
#pragma unroll
    for (int j = 3; j >= 0; --j) tric_row(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3], j, y, z, grad, rY, rYd, rYz);
    rYo = rY; rYdo = rYd; rYzo = rYz;
}

/* the x-Horner sums over the four lines' results: value and the three derivatives */
PTD void tric_finish(const double ry[4], const double ryd[4], const double ryz[4], double x, bool grad, double &F,
                     double &Fx, double &Fy, double &Fz) {
    double rx = 0.;
#pragma unroll
    for (int i = 3; i >= 0; --i) { rx += ry[i]; if (i > 0) rx *= x; }
    F = rx;
    if (grad) {
        double dx = 0.;
        dx += 3.0 * ry[3]; dx *= x;
        dx += 2.0 * ry[2]; dx *= x;
        dx += ry[1];
        Fx = dx;
        double dy = 0.;
#pragma unroll
        for (int i = 3; i >= 0; --i) { dy += ryd[i]; if (i > 0) dy *= x; }
        Fy = dy;
        double dz = 0.;
#pragma unroll
        for (int i = 3; i >= 0; --i) { dz += ryz[i]; if (i > 0) dz *= x; }
        Fz = dz;
    }
}

PTD void tric_eval_grad64(const double *__restrict__ a, double x, double y, double z, bool grad,
                          double &F, double &Fx, double &Fy, double &Fz) {
    double ry[4], ryd[4], ryz[4];
    /* two coefficient lines in flight: the next power of x is requested before the
       current one is used (the lines are independent, the arithmetic unchanged) */
    double c[16], cn[16];
    load_line64(a + 48, cn);
#pragma unroll
    for (int i = 3; i >= 0; --i) {
#pragma unroll
        for (int q = 0; q < 16; ++q) c[q] = cn[q];
        if (i > 0) load_line64(a + 16 * (i - 1), cn);
        tric_line(c, y, z, grad, ry[i], ryd[i], ryz[i]);
    }
    tric_finish(ry, ryd, ryz, x, grad, F, Fx, Fy, Fz);
}

/*
 * Wave-cooperative gathers of the step kernels' trial steps.  A lane's 1536-byte cell
 * block lies at a random place in the table, so one 16-byte load per lane touches as many
 * 128-byte lines as the wave has lanes and the vector L1 serves it a lane at a time.  The
 * active lanes of a wave therefore load each other's coefficients: a unit is one half
 * (64 bytes, two rows of four coefficients) of one x-power line of every active lane's
 * block, four lanes of consecutive rank read one half line together (the lane of rank r
 * reads, for k = 0..3, chunk v & 3 of the half line of rank v >> 2, v = k n + r, n active
 * lanes), so a load instruction of a full wave touches 16 lines instead of 64, and the
 * chunks reach their owners through 4 KB of LDS per wave.  Each lane evaluates its own
 * coefficients in the order tric_eval_grad64 does (rows 3 and 2 of a line, then rows 1
 * and 0), so the results are bit-identical.
 * LDS layout of a wave's slice: chunk q (0..3) of rank p at byte 1024 q + 16 (p ^ 2q), so
 * the eight lanes of every 16-byte access of a full wave cover all banks and the reads
 * are one XOR from a per-lane base plus a constant offset.  Before the units, the slice's
 * first 256 bytes pass each rank's block index to the lanes that load for it.
 */
constexpr int COOP_WAVES = 4;   /* waves of a 256-thread block */
__shared__ double2 coop_lds[COOP_WAVES][4][64];

/* the map's table is addressable in 16-byte units with 32-bit indices */
PTD bool coop_fits(const Field &f) {
    return (uint64_t)(f.nx - 1) * (uint64_t)(f.ny - 1) * (uint64_t)(f.nz - 1) * (uint64_t)f.ncomp * 32ull < (1ull << 32);
}

struct Chunks4 { double2 a, b, c, d; };

/* this lane's four chunk loads of one unit: src[k] is the 16-byte index of the chunk it
   loads for the k-th of its four (rank, chunk) pairs */
PTD Chunks4 coop_load(const double2 *__restrict__ t, const uint32_t *src, uint32_t off) {
    Chunks4 u;
    u.a = gload(t + (src[0] + off));
    u.b = gload(t + (src[1] + off));
    u.c = gload(t + (src[2] + off));
    u.d = gload(t + (src[3] + off));
    return u;
}

/* rank of this lane among the wave's active lanes (its lane index when all are active) */
PTD uint32_t active_rank() {
    uint32_t r;
    asm volatile("v_mbcnt_lo_u32_b32 %0, exec_lo, 0\n\tv_mbcnt_hi_u32_b32 %0, exec_hi, %0" : "=v"(r));
    return r;
}

#define COOP_LDS_FENCE() asm volatile("" ::: "memory")
#ifndef COOP_DEPTH
#define COOP_DEPTH 2   /* units whose loads are in flight ahead of the one being staged */
#endif

/* Value and gradient of the three B components of every active lane's cell (FP64 table):
   dist-scaled into Bt / dBt as container_B does */
PTD void tric_eval_B_coop(const Field &f, long cell, const double r[3], const double dist[3], bool grad,
                          double Bt[3], double dBt[3][3]) {
    const uint32_t n = (uint32_t)__builtin_popcountll(__builtin_amdgcn_read_exec());
    const uint32_t rank = active_rank();
    char *S = reinterpret_cast<char *>(coop_lds[threadIdx.x >> 6]);
    const double2 *t2 = reinterpret_cast<const double2 *>(f.coef);
    const uint32_t cs = (uint32_t)f.ncomp * 32u;           /* 16-byte units per cell */
    /* block indices by rank, then this lane's four (rank, chunk) pairs */
    reinterpret_cast<uint32_t *>(S)[rank] = (uint32_t)cell * cs;
    COOP_LDS_FENCE();
    uint32_t src[4], wb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = (uint32_t)k * n + rank, o = v >> 2, c = v & 3;
        src[k] = reinterpret_cast<const uint32_t *>(S)[o] + c;
        wb[k] = (c << 10) | ((o ^ (2 * c)) << 4);
    }
    COOP_LDS_FENCE();
    const double y = r[1], z = r[2];
    /* unit u = 8 i + 2 (3 - L) + (1 - h) (component i, line L, half h) at 16-byte offset
       32 i + 8 L + 4 h of a cell: the upper half (rows 3 and 2) of a line comes first.
       COOP_DEPTH units are in flight ahead of the one being staged; the loop is unrolled,
       so the unit buffers stay in registers and no loads in flight cross a back edge. */
    constexpr int NU = 24;
    auto unit_off = [](int u) -> uint32_t {
        const int i = u / 8, L = 3 - (u % 8) / 2, h = 1 - (u & 1);
        return (uint32_t)(32 * i + 8 * L + 4 * h);
    };
    Chunks4 buf[COOP_DEPTH + 1];
#pragma unroll
    for (int u = 0; u < COOP_DEPTH; ++u) buf[u] = coop_load(t2, src, unit_off(u));
    double ry[4], ryd[4], ryz[4], rY = 0., rYd = 0., rYz = 0.;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int i = u / 8, L = 3 - (u % 8) / 2, h = 1 - (u & 1);
        /* the loads COOP_DEPTH units ahead go out before this unit is staged */
        if (u + COOP_DEPTH < NU) buf[(u + COOP_DEPTH) % (COOP_DEPTH + 1)] = coop_load(t2, src, unit_off(u + COOP_DEPTH));
        const Chunks4 &vc = buf[u % (COOP_DEPTH + 1)];
        *reinterpret_cast<double2 *>(S + wb[0]) = vc.a;
        *reinterpret_cast<double2 *>(S + wb[1]) = vc.b;
        *reinterpret_cast<double2 *>(S + wb[2]) = vc.c;
        *reinterpret_cast<double2 *>(S + wb[3]) = vc.d;
        COOP_LDS_FENCE();
        /* the read base is rebuilt here from the rank, not kept live (or spilled) across
           the evaluation */
        const uint32_t rb = active_rank() << 4;
        double2 c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            c[q] = *reinterpret_cast<const double2 *>(S + ((rb ^ (uint32_t)(32 * q)) + 1024u * q));
        COOP_LDS_FENCE();
        if (h == 1) { rY = 0.; rYd = 0.; rYz = 0.; }
        tric_row(c[2].x, c[2].y, c[3].x, c[3].y, 2 * h + 1, y, z, grad, rY, rYd, rYz);
        tric_row(c[0].x, c[0].y, c[1].x, c[1].y, 2 * h, y, z, grad, rY, rYd, rYz);
        if (h == 0) { ry[L] = rY; ryd[L] = rYd; ryz[L] = rYz; }
        if (h == 0 && L == 0) {
            double Fv, Fx = 0, Fy = 0, Fz = 0;
            tric_finish(ry, ryd, ryz, r[0], grad, Fv, Fx, Fy, Fz);
            double d0 = Fx / dist[0], d1 = Fy / dist[1], d2 = Fz / dist[2];
            Bt[i] = Fv;
            if (grad) { dBt[i][0] = d0; dBt[i][1] = d1; dBt[i][2] = d2; }
        }
    }
}

/* FP32 Horner (with FMA) of a block held in registers, for the remainder of the
   mixed-precision tables: one row j (a0..a3 the coefficients of x^i y^j z^0..3), the
   x-Horner over the four lines' sums, and the whole block */
PTD void tric_row32(float a0, float a1, float a2, float a3, int j, float y, float z, bool grad, float &rY, float &rYd,
                    float &rYz) {
    const float rz = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(a3, z, a2), z, a1), z, a0);
    rY = __builtin_fmaf(rY, y, rz);
    if (grad) {
        const float rzd = __builtin_fmaf(__builtin_fmaf(3.0f * a3, z, 2.0f * a2), z, a1);
        rYz = __builtin_fmaf(rYz, y, rzd);
        if (j > 0) rYd = __builtin_fmaf(rYd, y, (float)j * rz);
    }
}

PTD void tric_finish32(const float ry[4], const float ryd[4], const float ryz[4], float x, bool grad, float &F,
                       float &Fx, float &Fy, float &Fz) {
    F = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ry[3], x, ry[2]), x, ry[1]), x, ry[0]);
    if (grad) {
        Fx = __builtin_fmaf(__builtin_fmaf(3.0f * ry[3], x, 2.0f * ry[2]), x, ry[1]);
        Fy = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ryd[3], x, ryd[2]), x, ryd[1]), x, ryd[0]);
        Fz = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ryz[3], x, ryz[2]), x, ryz[1]), x, ryz[0]);
    }
}

PTD void tric_eval_grad32(const float a[64], float x, float y, float z, bool grad,
                          float &F, float &Fx, float &Fy, float &Fz) {
    float ry[4], ryd[4], ryz[4];
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        const float *c = a + 16 * i;
        float rY = 0.f, rYd = 0.f, rYz = 0.f;
#pragma unroll
        for (int j = 3; j >= 0; --j) tric_row32(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3], j, y, z, grad, rY, rYd, rYz);
        ry[i] = rY; ryd[i] = rYd; ryz[i] = rYz;
    }
    tric_finish32(ry, ryd, ryz, x, grad, F, Fx, Fy, Fz);
}

/* Mixed precision: value and gradient of component comp of a cell from the mixed table,
   the FP64 linear part (in FP64 arithmetic on the FP64 cell coordinates) plus the FP32
   Horner of the remainder; or the FP64 block where the component kept it. */
PTD void tric_eval_mixed(const Field &f, size_t cell, int comp, const double r[3], bool grad, double &F, double &Fx,
                         double &Fy, double &Fz) {
    const size_t blk = cell * (size_t)f.ncomp + (size_t)comp;
    const unsigned char *rec = f.mix + cell * mix_cell_bytes(f.ncomp);
    const double *lin = reinterpret_cast<const double *>(rec + (size_t)comp * MIX_LIN_BYTES);
    /* the whole block is loaded before the marker is tested, so the loads of the common
       (mixed) case do not wait for the marker's round trip to memory */
    const double2 l01 = reinterpret_cast<const double2 *>(lin)[0];
    const double2 l23 = reinterpret_cast<const double2 *>(lin)[1];
    float c[64];
    const float4 *c4 = reinterpret_cast<const float4 *>(rec + mix_lin_bytes(f.ncomp) + (size_t)comp * MIX_REM_BYTES);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float4 v = c4[q];
        c[4 * q] = v.x; c[4 * q + 1] = v.y; c[4 * q + 2] = v.z; c[4 * q + 3] = v.w;
    }
    if (l01.x != l01.x) {
        tric_eval_grad64(f.coef + blk * 64, r[0], r[1], r[2], grad, F, Fx, Fy, Fz);
        return;
    }
    float fv, fx = 0.f, fy = 0.f, fz = 0.f;
    tric_eval_grad32(c, (float)r[0], (float)r[1], (float)r[2], grad, fv, fx, fy, fz);
    F = l01.x + l01.y * r[2] + l23.x * r[1] + l23.y * r[0] + (double)fv;
    if (grad) {
        Fx = l23.y + (double)fx;
        Fy = l23.x + (double)fy;
        Fz = l01.y + (double)fz;
    }
}

#undef COOP_LDS_FENCE

PTD bool in_bounds(const Field &f, double x, double y, double z) {
    const double *b = f.box;
    bool has = b[1] < b[0] && b[3] < b[2] && b[5] < b[4];
    if (!has) return true;
    return x >= b[1] && x < b[0] && y >= b[3] && y < b[2] && z >= b[5] && z < b[4];
}
PTD double smth(double x) { return ((6 * x - 15) * x + 10) * x * x * x; }
PTD double smthd(double x) { return ((30 * x - 60) * x + 30) * x * x; }

/* TFieldBoundaryBox::scaleScalarFieldAtBounds (src/field.cpp:71-112); F and its gradient
   (if grad) in, scaled values out */
struct FG { double F, d0, d1, d2; };
PTD_NOINLINE FG scale_at_bounds(const Field &f, double x, double y, double z, double F, double g0, double g1, double g2, bool grad) {
    const double *b = f.box;
    double xmax = b[0], xmin = b[1], ymax = b[2], ymin = b[3], zmax = b[4], zmin = b[5], bw = f.bw;
    double dF[3] = {g0, g1, g2};
    FG out = {F, g0, g1, g2};
    bool has = xmin < xmax && ymin < ymax && zmin < zmax;
    if (!has || (F == 0 && !grad) || (F == 0 && dF[0] == 0 && dF[1] == 0 && dF[2] == 0)) return out;
    if (!in_bounds(f, x, y, z)) {
        out.F = 0.;
        if (grad) { out.d0 = 0.; out.d1 = 0.; out.d2 = 0.; }
    } else if (bw > 0) {
        double Fscale = 1.;
        double dFadd[3] = {0., 0., 0.};
        double lo[3] = {(x - xmin) / bw, (y - ymin) / bw, (z - zmin) / bw};
        double hi[3] = {(xmax - x) / bw, (ymax - y) / bw, (zmax - z) / bw};
        for (int i = 0; i < 3; ++i) {
            if (0 <= lo[i] && lo[i] <= 1) {
                Fscale *= smth(lo[i]);
                if (grad) dFadd[i] += F * smthd(lo[i]) / bw / smth(lo[i]);
            } else if (0 <= hi[i] && hi[i] <= 1) {
                Fscale *= smth(hi[i]);
                if (grad) dFadd[i] += -F * smthd(hi[i]) / bw / smth(hi[i]);
            }
        }
        if (Fscale != 1.) {
            out.F = F * Fscale;
            if (grad) {
                out.d0 = dF[0] * Fscale + dFadd[0] * Fscale;
                out.d1 = dF[1] * Fscale + dFadd[1] * Fscale;
                out.d2 = dF[2] * Fscale + dFadd[2] * Fscale;
            }
        }
    }
    return out;
}

PTD bool has_box(const Field &f) { return f.box[1] < f.box[0] && f.box[3] < f.box[2] && f.box[5] < f.box[4]; }

/* kernel modes: mixed-precision field tables; containers with scale formulas,
   analytic fields or conductors present (without it the field sum handles tricubic maps
   and uniform E with constant scales only, the code of the common case) */
constexpr int MODE_MIXED = 1, MODE_FIELDS = 2;

/* TFieldScaler::scalingFactor (src/field.cpp:7-10) from the compiled formula */
PTD double scale_formula(const uint32_t *code, int n, const double *cst, double t) {
    return ptf_run(code, n, cst, t);
}
PTD double scale_B(const Field &f, double t) { return f.blen ? scale_formula(f.bcode, f.blen, f.bcst, t) : f.bscale; }
PTD double scale_E(const Field &f, double t) { return f.elen ? scale_formula(f.ecode, f.elen, f.ecst, t) : f.escale; }

/* analytic fields (src/analyticFields.cpp:17-147, operations in the reference's order) and
   the finite straight wire (src/conductor.cpp:21-172) by the Biot-Savart law of a segment
   A->B: with e the unit vector along the wire, r1 = P - A, r2 = P - B, w = e x r1,
   d^2 = |w|^2, c_i = (e . r_i)/|r_i|:  B = k (c1 - c2)/d^2 w, k = mu0 I/(4 pi), and the
   Jacobian from d(d^2)/dP = 2 (r1 - (e.r1) e), dc_i/dP = e/|r_i| - (e.r_i) r_i/|r_i|^3,
   dw/dP = [e]x.  B and dB (row-major 3x3) are zeroed first. */
PTD void analytic_B(const Field &f, double x, double y, double z, double *B, double *dB, bool grad) {
    const double *p = f.p;
    for (int i = 0; i < 3; ++i) B[i] = 0;
    if (grad)
        for (int i = 0; i < 9; ++i) dB[i] = 0;
    switch (f.type) {
    case PT_FIELD_CONDUCTOR: {
        if (p[0] == 0) return;
        const double k = p[7];
        double l0 = p[4] - p[1], l1 = p[5] - p[2], l2 = p[6] - p[3];
        double L = __builtin_sqrt(l0 * l0 + l1 * l1 + l2 * l2);
        double e0 = l0 / L, e1 = l1 / L, e2 = l2 / L;
        double r10 = x - p[1], r11 = y - p[2], r12 = z - p[3];
        double r20 = x - p[4], r21 = y - p[5], r22 = z - p[6];
        double n1 = __builtin_sqrt(r10 * r10 + r11 * r11 + r12 * r12), n2 = __builtin_sqrt(r20 * r20 + r21 * r21 + r22 * r22);
        double s1 = e0 * r10 + e1 * r11 + e2 * r12, s2 = e0 * r20 + e1 * r21 + e2 * r22;
        double w0 = e1 * r12 - e2 * r11, w1 = e2 * r10 - e0 * r12, w2 = e0 * r11 - e1 * r10;
        double d2 = w0 * w0 + w1 * w1 + w2 * w2;
        double c = s1 / n1 - s2 / n2;
        double g = k * c / d2;
        B[0] = g * w0; B[1] = g * w1; B[2] = g * w2;
        if (grad) {
            double q1 = s1 / (n1 * n1 * n1), q2 = s2 / (n2 * n2 * n2);
            double rp0 = r10 - s1 * e0, rp1 = r11 - s1 * e1, rp2 = r12 - s1 * e2;
            double dc0 = (e0 / n1 - q1 * r10) - (e0 / n2 - q2 * r20);
            double dc1 = (e1 / n1 - q1 * r11) - (e1 / n2 - q2 * r21);
            double dc2 = (e2 / n1 - q1 * r12) - (e2 / n2 - q2 * r22);
            double dd0 = 2 * rp0, dd1 = 2 * rp1, dd2 = 2 * rp2;
            double dg0 = k * (dc0 / d2 - c * dd0 / (d2 * d2));
            double dg1 = k * (dc1 / d2 - c * dd1 / (d2 * d2));
            double dg2 = k * (dc2 / d2 - c * dd2 / (d2 * d2));
            const double E[9] = {0, -e2, e1, e2, 0, -e0, -e1, e0, 0};
            const double w[3] = {w0, w1, w2};
            const double dg[3] = {dg0, dg1, dg2};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) dB[3 * i + j] = g * E[3 * i + j] + w[i] * dg[j];
        }
        return;
    }
    case PT_FIELD_EXPONENTIAL_X: {
        double a1 = p[0], a2 = p[1], a3 = p[2], c1 = p[3], c2 = p[4];
        double e = ptm_exp(-a2 * x + a3);
        B[0] = a1 * e + c1;
        B[1] = y * a1 * a2 / 2 * e + c2;
        B[2] = z * a1 * a2 / 2 * e + c2;
        if (grad) {
            dB[0] = -a2 * a1 * e;
            dB[3] = -a2 * y * a1 * a2 / 2 * e;
            dB[4] = a1 * a2 / 2 * e;
            dB[6] = -a2 * z * a1 * a2 / 2 * e;
            dB[8] = a1 * a2 / 2 * e;
        }
        return;
    }
    case PT_FIELD_LINEAR_Z:
        B[2] = p[0] * x + p[1];
        if (grad) dB[6] = p[0];
        return;
    case PT_FIELD_B0_GRAD_Z: {
        double a1 = p[0], a2 = p[1], z0 = p[2];
        B[0] = -(a1 * z + a2) / 2 * x;
        B[1] = -(a1 * z + a2) / 2 * y;
        B[2] = a1 * z * z / 2 + a2 * z + z0;
        if (grad) {
            dB[0] = -(a1 * z + a2) / 2;
            dB[2] = -a1 * x / 2;
            dB[4] = -(a1 * z + a2) / 2;
            dB[5] = -a1 * y / 2;
            dB[8] = a1 * z + a2;
        }
        return;
    }
    case PT_FIELD_B0_GRAD_X2: {
        double a1 = p[0], a2 = p[1], a3 = p[2], z0 = p[3];
        B[0] = -a1 / 6 * x * x * x - a2 / 4 * x * x - a3 / 2 * x;
        B[1] = -(a1 * x * x + a2 * x + a3) / 2 * y;
        B[2] = (a1 * x * x + a2 * x + a3) * z + z0;
        if (grad) {
            dB[0] = -(a1 * x * x + a2 * x + a3) / 2;
            dB[3] = (-a1 * x - a2 / 2) * y;
            dB[4] = -(a1 * x * x + a2 * x + a3) / 2;
            dB[6] = (2 * a1 * x + a2) * z;
            dB[8] = a1 * x * x + a2 * x + a3;
        }
        return;
    }
    case PT_FIELD_B0_GRAD_XY: {
        double a1 = p[0], a2 = p[1], z0 = p[2];
        B[0] = -a1 * x * x * y / 4 - a2 * x / 2;
        B[1] = -a1 * x * y * y / 4 - a2 * y / 2;
        B[2] = a1 * x * y * z + a2 * z + z0;
        if (grad) {
            dB[0] = -(a1 * x * y + a2) / 2;
            dB[1] = -a1 * x * x / 4;
            dB[3] = -a1 * y * y / 4;
            dB[4] = -(a1 * x * y + a2) / 2;
            dB[6] = a1 * y * z;
            dB[7] = a1 * x * z;
            dB[8] = a1 * x * y + a2;
        }
        return;
    }
    case PT_FIELD_B0_XY: {
        double a1 = p[0], z0 = p[1];
        B[0] = a1 * y * z;
        B[1] = a1 * x * z;
        B[2] = a1 * x * y + z0;
        if (grad) {
            dB[1] = a1 * z; dB[2] = a1 * y;
            dB[3] = a1 * z; dB[5] = a1 * x;
            dB[6] = a1 * y; dB[7] = a1 * x;
        }
        return;
    }
    default:
        return;
    }
}

/* Axisymmetric maps (TabField, src/field_2d.cpp:289-400) with ALGLIB's bicubic splines.
   The cell is ALGLIB's bisection result: the last node below the coordinate, clamped to
   [0, n-2] (found from a uniform-grid guess).  Value-only evaluations follow spline2dcalc
   (t = (x - x_l) * (1/h)), evaluations with derivatives spline2ddiff (t = (x - x_l) / h),
   as the reference calls one or the other; sums in ALGLIB's order. */
PTD int bic_cell(const double *g, int n, double x0, double inv_h, double v) {
    const double gd = (v - x0) * inv_h;
    int l = gd > 0 ? (gd < (double)(n - 2) ? (int)gd : n - 2) : 0;
    while (l > 0 && !(g[l] < v)) --l;
    while (l < n - 2 && g[l + 1] < v) ++l;
    return l;
}

struct BicAt { double t, u, dt, du; size_t cell; };

PTD BicAt bic_locate(const Field &f, double r, double z, bool diff) {
    BicAt b;
    const int ir = bic_cell(f.g[0], f.nx, f.x0[0], f.inv_h[0], r);
    const int iz = bic_cell(f.g[2], f.nz, f.x0[2], f.inv_h[2], z);
    const double r0 = f.g[0][ir], hr = f.g[0][ir + 1] - r0;
    const double z0 = f.g[2][iz], hz = f.g[2][iz + 1] - z0;
    b.dt = 1.0 / hr;
    b.du = 1.0 / hz;
    if (diff) { b.t = (r - r0) / hr; b.u = (z - z0) / hz; }
    else { b.t = (r - r0) * b.dt; b.u = (z - z0) * b.du; }
    b.cell = (size_t)iz * (f.nx - 1) + ir;
    return b;
}

/* one component: c[16] = f, df/dr, df/dz, d2f/drdz at the corners s1..s4 */
PTD void bic_eval(const double *c, const BicAt &b, bool diff, double &F, double &Fr, double &Fz) {
    const double t = b.t, u = b.u;
    const double t2 = t * t, t3 = t * t2, u2 = u * u, u3 = u * u2;
    const double ht00 = 2 * t3 - 3 * t2 + 1, ht01 = -2 * t3 + 3 * t2;
    const double ht10 = (t3 - 2 * t2 + t) / b.dt, ht11 = (t3 - t2) / b.dt;
    const double hu00 = 2 * u3 - 3 * u2 + 1, hu01 = -2 * u3 + 3 * u2;
    const double hu10 = (u3 - 2 * u2 + u) / b.du, hu11 = (u3 - u2) / b.du;
    F = 0.0 + c[0] * ht00 * hu00 + c[1] * ht01 * hu00 + c[2] * ht00 * hu01 + c[3] * ht01 * hu01;
    F = F + c[4] * ht10 * hu00 + c[5] * ht11 * hu00 + c[6] * ht10 * hu01 + c[7] * ht11 * hu01;
    F = F + c[8] * ht00 * hu10 + c[9] * ht01 * hu10 + c[10] * ht00 * hu11 + c[11] * ht01 * hu11;
    F = F + c[12] * ht10 * hu10 + c[13] * ht11 * hu10 + c[14] * ht10 * hu11 + c[15] * ht11 * hu11;
    if (!diff) return;
    const double dht00 = (6 * t2 - 6 * t) * b.dt, dht01 = (-6 * t2 + 6 * t) * b.dt;
    const double dht10 = 3 * t2 - 4 * t + 1, dht11 = 3 * t2 - 2 * t;
    const double dhu00 = (6 * u2 - 6 * u) * b.du, dhu01 = (-6 * u2 + 6 * u) * b.du;
    const double dhu10 = 3 * u2 - 4 * u + 1, dhu11 = 3 * u2 - 2 * u;
    Fr = 0.0 + c[0] * dht00 * hu00 + c[1] * dht01 * hu00 + c[2] * dht00 * hu01 + c[3] * dht01 * hu01;
    Fr = Fr + c[4] * dht10 * hu00 + c[5] * dht11 * hu00 + c[6] * dht10 * hu01 + c[7] * dht11 * hu01;
    Fr = Fr + c[8] * dht00 * hu10 + c[9] * dht01 * hu10 + c[10] * dht00 * hu11 + c[11] * dht01 * hu11;
    Fr = Fr + c[12] * dht10 * hu10 + c[13] * dht11 * hu10 + c[14] * dht10 * hu11 + c[15] * dht11 * hu11;
    Fz = 0.0 + c[0] * ht00 * dhu00 + c[1] * ht01 * dhu00 + c[2] * ht00 * dhu01 + c[3] * ht01 * dhu01;
    Fz = Fz + c[4] * ht10 * dhu00 + c[5] * ht11 * dhu00 + c[6] * ht10 * dhu01 + c[7] * ht11 * dhu01;
    Fz = Fz + c[8] * ht00 * dhu10 + c[9] * ht01 * dhu10 + c[10] * ht00 * dhu11 + c[11] * ht01 * dhu11;
    Fz = Fz + c[12] * ht10 * dhu10 + c[13] * ht11 * dhu10 + c[14] * ht10 * dhu11 + c[15] * ht11 * dhu11;
}

PTD bool tab2d_in_range(const Field &f, double r, double z) {
    return r >= f.p[0] && r < f.p[1] && z >= f.p[2] && z < f.p[3];
}

/* TabField::BField (src/field_2d.cpp:289-337): B and dB (row-major) zeroed first */
PTD void tab2d_B(const Field &f, double x, double y, double z, double *B, double *dB, bool grad, long long &nfield) {
    for (int i = 0; i < 3; ++i) B[i] = 0;
    if (grad)
        for (int i = 0; i < 9; ++i) dB[i] = 0;
    const double r = __builtin_sqrt(x * x + y * y);
    if (!tab2d_in_range(f, r, z)) return;
    ++nfield;
    const double phi = ptm_atan2(y, x);
    const double cp = ptm_cos(phi), sp = ptm_sin(phi);
    const BicAt b = bic_locate(f, r, z, grad);
    const double *c = f.coef + b.cell * f.ncomp * 16;
    double Br = 0, Bphi = 0, Bz = 0, dBrdr = 0, dBrdz = 0, dBphidr = 0, dBphidz = 0, dBzdr = 0, dBzdz = 0;
    if (f.has_b & 1) { bic_eval(c, b, grad, Br, dBrdr, dBrdz); c += 16; }
    if (f.has_b & 2) { bic_eval(c, b, grad, Bphi, dBphidr, dBphidz); c += 16; }
    if (f.has_b & 4) bic_eval(c, b, grad, Bz, dBzdr, dBzdz);
    if (grad) {
        if (r > 0) {
            dB[0] = dBrdr * cp * cp - dBphidr * cp * sp + (Br * sp * sp + Bphi * cp * sp) / r;
            dB[1] = dBrdr * cp * sp - dBphidr * sp * sp - (Br * cp * sp + Bphi * cp * cp) / r;
            dB[3] = dBrdr * cp * sp + dBphidr * cp * cp - (Br * cp * sp - Bphi * sp * sp) / r;
            dB[4] = dBrdr * sp * sp + dBphidr * cp * sp + (Br * cp * cp - Bphi * cp * sp) / r;
        }
        dB[2] = dBrdz * cp - dBphidz * sp;           /* CylToCart (src/field_2d.cpp:32-35) */
        dB[5] = dBrdz * sp + dBphidz * cp;
        if (f.has_b & 4) { dB[6] = dBzdr * cp; dB[7] = dBzdr * sp; dB[8] = dBzdz; }
    }
    B[0] = Br * cp - Bphi * sp;
    B[1] = Br * sp + Bphi * cp;
    B[2] = Bz;
}

/* TabField::EField (src/field_2d.cpp:340-400): a tabulated E (value-only splines) takes
   precedence over the potential (splines with derivatives, E = -grad V) */
PTD void tab2d_E(const Field &f, double x, double y, double z, double &V, double E[3], long long &nfield) {
    V = 0; E[0] = E[1] = E[2] = 0;
    const double r = __builtin_sqrt(x * x + y * y);
    if (!tab2d_in_range(f, r, z)) return;
    ++nfield;
    const double *c0 = f.coef + __builtin_popcount((unsigned)f.has_b) * 16;
    if (f.has_v & 14) {
        const double phi = ptm_atan2(y, x);
        const BicAt b = bic_locate(f, r, z, false);
        const double *c = c0 + b.cell * f.ncomp * 16;
        double Er = 0, Ephi = 0, Ez = 0, dummy;
        if (f.has_v & 2) { bic_eval(c, b, false, Er, dummy, dummy); c += 16; }
        if (f.has_v & 4) { bic_eval(c, b, false, Ephi, dummy, dummy); c += 16; }
        if (f.has_v & 8) bic_eval(c, b, false, Ez, dummy, dummy);
        const double cp = ptm_cos(phi), sp = ptm_sin(phi);
        E[0] = Er * cp - Ephi * sp;
        E[1] = Er * sp + Ephi * cp;
        E[2] = Ez;
    } else if (f.has_v & 1) {
        const BicAt b = bic_locate(f, r, z, true);
        double dVdr = 0, dVdz = 0;
        bic_eval(c0 + b.cell * f.ncomp * 16, b, true, V, dVdr, dVdz);
        const double phi = ptm_atan2(y, x);
        E[0] = -dVdr * ptm_cos(phi);
        E[1] = -dVdr * ptm_sin(phi);
        E[2] = -dVdz;
    }
}

/* TFieldContainer::BField (src/field.cpp:139-153) with the scale factor bs already
   evaluated, accumulated into the manager sum (TFieldManager::BField, src/fields.cpp:139-164);
   tricubic maps and, through analytic_B, the analytic fields and conductors */
template <int MODE, bool COOP = false>
PTD void container_B(const Field &f, double x, double y, double z, double bs, double B[3], double dB[3][3], bool grad,
                     long long &nfield) {
    double Bt[3] = {0, 0, 0};
    double dBt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (in_bounds(f, x, y, z) && bs != 0.) {
        if ((MODE & MODE_FIELDS) && f.type >= PT_FIELD_CONDUCTOR) {
            double Ba[3], dBa[9];
            if (f.type == PT_FIELD_TAB2D) tab2d_B(f, x, y, z, Ba, dBa, grad, nfield);
            else analytic_B(f, x, y, z, Ba, dBa, grad);
            Bt[0] = Ba[0]; Bt[1] = Ba[1]; Bt[2] = Ba[2];
            if (grad) {
                dBt[0][0] = dBa[0]; dBt[0][1] = dBa[1]; dBt[0][2] = dBa[2];
                dBt[1][0] = dBa[3]; dBt[1][1] = dBa[4]; dBt[1][2] = dBa[5];
                dBt[2][0] = dBa[6]; dBt[2][1] = dBa[7]; dBt[2][2] = dBa[8];
            }
        } else if (f.type == 1 && f.has_b) {
            long cell; double r[3], dist[3];
            if (tric_locate(f, x, y, z, cell, r, dist)) {
                ++nfield;
                /* the mixed tables keep per-lane loads: gathered cooperatively (the FP32 lines
                   as units, the components kept in FP64 dealt to the wave's lanes in one
                   FP64 pass) they ran at 0.79x the per-lane form on a map with 26% FP64
                   components */
                if (COOP && !(MODE & MODE_MIXED) && f.ncomp >= 3 && coop_fits(f)) {
                    tric_eval_B_coop(f, cell, r, dist, grad, Bt, dBt);
                } else {
#pragma unroll 1
                for (int i = 0; i < 3; ++i) {
                    double Fv, Fx = 0, Fy = 0, Fz = 0;
                    if (MODE & MODE_MIXED) {
                        tric_eval_mixed(f, (size_t)cell, i, r, grad, Fv, Fx, Fy, Fz);
                    } else {
                        tric_eval_grad64(f.coef + ((size_t)cell * f.ncomp + i) * 64, r[0], r[1], r[2], grad, Fv, Fx, Fy, Fz);
                    }
                    double d0 = Fx / dist[0], d1 = Fy / dist[1], d2 = Fz / dist[2];
                    /* write the component without a dynamically indexed array */
                    if (i == 0) { Bt[0] = Fv; if (grad) { dBt[0][0] = d0; dBt[0][1] = d1; dBt[0][2] = d2; } }
                    else if (i == 1) { Bt[1] = Fv; if (grad) { dBt[1][0] = d0; dBt[1][1] = d1; dBt[1][2] = d2; } }
                    else { Bt[2] = Fv; if (grad) { dBt[2][0] = d0; dBt[2][1] = d1; dBt[2][2] = d2; } }
                }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Bt[i] *= bs;
            if (grad) { dBt[i][0] *= bs; dBt[i][1] *= bs; dBt[i][2] *= bs; }
        }
        if (has_box(f)) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                FG r = scale_at_bounds(f, x, y, z, Bt[i], dBt[i][0], dBt[i][1], dBt[i][2], grad);
                Bt[i] = r.F;
                if (grad) { dBt[i][0] = r.d0; dBt[i][1] = r.d1; dBt[i][2] = r.d2; }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        B[i] += Bt[i];
        if (grad) {
#pragma unroll
            for (int j = 0; j < 3; j++) dB[i][j] += dBt[i][j];
        }
    }
}

template <int MODE, bool COOP = false>
PTD void field_B(const Params &P, double x, double y, double z, double t, double B[3], double dB[3][3], bool grad, long long &nfield) {
#pragma unroll
    for (int i = 0; i < 3; i++) { B[i] = 0; if (grad) { dB[i][0] = 0; dB[i][1] = 0; dB[i][2] = 0; } }
    for (int k = 0; k < P.n_fields; ++k) {
        const Field &f = P.fields[k];
        /* TFieldScaler::scalingFactor (src/field.cpp:7-10), evaluated where the reference does */
        const double bs = (!(MODE & MODE_FIELDS) || f.blen == 0) ? f.bscale : (in_bounds(f, x, y, z) ? scale_B(f, t) : 0.);
        container_B<MODE, COOP>(f, x, y, z, bs, B, dB, grad, nfield);
    }
}

/* TFieldContainer::EField (src/field.cpp:155-169) with the scale factor es already evaluated */
template <int MODE>
PTD void container_E(const Field &f, double x, double y, double z, double es, double &Vt, double Et[3], long long &nfield) {
    Vt = 0; Et[0] = Et[1] = Et[2] = 0;
    if (in_bounds(f, x, y, z) && es != 0.) {
        if (f.type == 2) {
            Et[0] = f.e[0]; Et[1] = f.e[1]; Et[2] = f.e[2];
            Vt = Et[2] * z;
        } else if ((MODE & MODE_FIELDS) && f.type == PT_FIELD_TAB2D) {
            tab2d_E(f, x, y, z, Vt, Et, nfield);
        } else if (f.type == 1 && f.has_v) {
            long cell; double r[3], dist[3];
            double dV[3] = {0., 0., 0.};
            if (tric_locate(f, x, y, z, cell, r, dist)) {
                ++nfield;
                double Fv, Fx, Fy, Fz;
                int comp = f.has_b ? 3 : 0;
                if (MODE & MODE_MIXED) {
                    tric_eval_mixed(f, (size_t)cell, comp, r, true, Fv, Fx, Fy, Fz);
                } else {
                    tric_eval_grad64(f.coef + ((size_t)cell * f.ncomp + comp) * 64, r[0], r[1], r[2], true, Fv, Fx, Fy, Fz);
                }
                Vt = Fv;
                dV[0] = Fx / dist[0]; dV[1] = Fy / dist[1]; dV[2] = Fz / dist[2];
            }
            for (int i = 0; i < 3; i++) Et[i] = -dV[i];
        }
        Vt *= es;
        Et[0] *= es; Et[1] *= es; Et[2] *= es;
        if (has_box(f)) {
            FG r = scale_at_bounds(f, x, y, z, Vt, Et[0], Et[1], Et[2], true);
            Vt = r.F; Et[0] = r.d0; Et[1] = r.d1; Et[2] = r.d2;
        }
    }
}

/* TFieldManager::EField (src/fields.cpp:167-182) */
template <int MODE>
PTD void field_E(const Params &P, double x, double y, double z, double t, double *V, double E[3], long long &nfield) {
    *V = 0; E[0] = E[1] = E[2] = 0;
    for (int k = 0; k < P.n_fields; ++k) {
        const Field &f = P.fields[k];
        double Vt, Et[3];
        const double es = (!(MODE & MODE_FIELDS) || f.elen == 0) ? f.escale : (in_bounds(f, x, y, z) ? scale_E(f, t) : 0.);
        container_E<MODE>(f, x, y, z, es, Vt, Et, nfield);
        *V += Vt;
        for (int i = 0; i < 3; i++) E[i] += Et[i];
    }
}

/* ------------------------------------------------------------------------ */
/* particle                                                                  */
/* ------------------------------------------------------------------------ */

/* TParticle::EquationOfMotion, src/particle.cpp:105-135 */
PTD void eom(const Species &sp, const double *y, const double B[3], const double dB[3][3], const double E[3], double *dydx) {
    dydx[0] = y[3]; dydx[1] = y[4]; dydx[2] = y[5];
    double F[3] = {0, 0, 0};
    F[2] += -GRAV * sp.m * ELE_E;
    if (sp.q != 0) {
        F[0] += sp.q * (E[0] + y[4] * B[2] - y[5] * B[1]);
        F[1] += sp.q * (E[1] + y[5] * B[0] - y[3] * B[2]);
        F[2] += sp.q * (E[2] + y[3] * B[1] - y[4] * B[0]);
    }
    if (sp.mu != 0 && y[7] != 0 && (B[0] != 0 || B[1] != 0 || B[2] != 0)) {
        double Babs = __builtin_sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        double d0 = (B[0] * dB[0][0] + B[1] * dB[1][0] + B[2] * dB[2][0]) / Babs;
        double d1 = (B[0] * dB[0][1] + B[1] * dB[1][1] + B[2] * dB[2][1]) / Babs;
        double d2 = (B[0] * dB[0][2] + B[1] * dB[1][2] + B[2] * dB[2][2]) / Babs;
        F[0] += y[7] * sp.mu * d0;
        F[1] += y[7] * sp.mu * d1;
        F[2] += y[7] * sp.mu * d2;
    }
    double v2 = y[3] * y[3] + y[4] * y[4] + y[5] * y[5];
    double inversegamma = __builtin_sqrt(1 - v2 / (C_0 * C_0));
    double f = inversegamma / sp.m / ELE_E;
    dydx[3] = f * (F[0] - (y[3] * y[3] * F[0] + y[3] * y[4] * F[1] + y[3] * y[5] * F[2]) / C_0 / C_0);
    dydx[4] = f * (F[1] - (y[4] * y[3] * F[0] + y[4] * y[4] * F[1] + y[4] * y[5] * F[2]) / C_0 / C_0);
    dydx[5] = f * (F[2] - (y[5] * y[3] * F[0] + y[5] * y[4] * F[1] + y[5] * y[5] * F[2]) / C_0 / C_0);
    dydx[6] = inversegamma;
    dydx[7] = 0;
    dydx[8] = __builtin_sqrt(v2);
}

/* evaluation counters kept per thread */
struct Ctr {
    long long rhs, field, tri;
};

/* TParticle::GetKineticEnergy, src/particle.cpp:237-251 */
PTD double kinetic_energy(const Species &sp, const double *v) {
    double beta2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / C_0 / C_0;
    double gm1;
    if (ptm_pow(beta2, 5) < DBL_EPS)
        gm1 = (0.5 + (3.0 / 8.0 + (5.0 / 16.0 + 35.0 / 128.0 * beta2) * beta2) * beta2) * beta2;
    else
        gm1 = 1.0 / __builtin_sqrt(1.0 - beta2) - 1.0;
    return C_0 * C_0 * sp.m * gm1;
}

/* TParticle::GetPotentialEnergy, src/particle.cpp:254-269; Bc/Vc: field values already
   known at this point (from the equation-of-motion evaluation), or nullptr */
template <int MODE>
PTD double potential_energy(const Params &P, const Species &sp, const double *y, double t, const double *Bc, const double *Vc, Ctr &c) {
    double result = 0;
    if (sp.q != 0 || sp.mu != 0) {
        if (sp.mu != 0) {
            double B[3], dB[3][3];
            if (Bc) { B[0] = Bc[0]; B[1] = Bc[1]; B[2] = Bc[2]; }
            else field_B<MODE>(P, y[0], y[1], y[2], t, B, dB, false, c.field);
            result += -y[7] * sp.mu / ELE_E * __builtin_sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        }
        if (sp.q != 0) {
            double E[3], V;
            if (Vc) V = *Vc;
            else field_E<MODE>(P, y[0], y[1], y[2], t, &V, E, c.field);
            result += sp.q / ELE_E * V;
        }
    }
    result += sp.m * GRAV * y[2];
    return result;
}

/* ------------------------------------------------------------------------ */
/* Boost.Odeint dense-output dopri5 (runge_kutta_dopri5 + controlled + dense) */
/* ------------------------------------------------------------------------ */

/* dense-output state of the collision path (all nine state components) */
struct Dopri {
    double xc[9], dxc[9];
    double xo[9], dxo[9];
    double k3[9], k4[9], k5[9], k6[9];
    double t, t_old, dt;
    bool deriv_init;
};

/* runge_kutta_dopri5::calc_state */
PTD void dopri_calc_state(const Dopri &s, double t, double *x) {
    constexpr double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192, b5 = -2187.0 / 6784, b6 = 11.0 / 84;
    const double dt = (s.t - s.t_old);
    const double theta = (t - s.t_old) / dt;
    const double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
    const double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
    const double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
    const double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
    const double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
    const double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
    const double theta_m_1 = theta - 1.0;
    const double theta_sq = theta * theta;
    const double A = theta_sq * (3.0 - 2.0 * theta);
    const double B = theta_sq * theta_m_1;
    const double C = theta_sq * theta_m_1 * theta_m_1;
    const double D = theta * theta_m_1 * theta_m_1;
    const double h1 = dt * (A * b1 - C * X1 + D);
    const double h3 = dt * (A * b3 + C * X3);
    const double h4 = dt * (A * b4 - C * X4);
    const double h5 = dt * (A * b5 + C * X5);
    const double h6 = dt * (A * b6 - C * X6);
    const double h7 = dt * (B + C * X7);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        x[i] = s.xo[i] + h1 * s.dxo[i] + h3 * s.k3[i] + h4 * s.k4[i] + h5 * s.k5[i] + h6 * s.k6[i] + h7 * s.dxc[i];
}

PTD void to9(const double *s, double pol, double *y) {
    y[0] = s[0]; y[1] = s[1]; y[2] = s[2]; y[3] = s[3]; y[4] = s[4]; y[5] = s[5];
    y[6] = s[6]; y[7] = pol; y[8] = s[7];
}

/* TParticle::derivs (src/particle.cpp:96-103) on the eight stepped components
   (x, y, z, vx, vy, vz, proper time, path length); the polarisation, whose
   derivative is identically +0.0, is passed separately.  B and V return the
   field the equation of motion used (hasB: B was evaluated). */
template <int MODE, bool COOP = false>
PTD void derivs8(const Params &P, const Species &sp, const double *s, double t, double pol, double *d, double B[3],
                 double &V, bool &hasB, Ctr &c) {
    double y[9];
    to9(s, pol, y);
    double dB[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, E[3] = {0, 0, 0};
    B[0] = 0; B[1] = 0; B[2] = 0; V = 0;
    ++c.rhs;
    const bool needB = (sp.q != 0 || (sp.mu != 0 && pol != 0));
    if (needB) field_B<MODE, COOP>(P, y[0], y[1], y[2], t, B, dB, true, c.field);
    /* B may stand in for later value-only evaluations at this point unless a 2D map
       computes those differently (TabField: spline2dcalc vs spline2ddiff) */
    hasB = needB && (!(MODE & MODE_FIELDS) || P.b_reuse);
    if (sp.q != 0) field_E<MODE>(P, y[0], y[1], y[2], t, &V, E, c.field);
    double d9[9];
    eom(sp, y, B, dB, E, d9);
    d[0] = d9[0]; d[1] = d9[1]; d[2] = d9[2]; d[3] = d9[3]; d[4] = d9[4]; d[5] = d9[5]; d[6] = d9[6]; d[7] = d9[8];
}

/* runge_kutta_dopri5::calc_state from the dense-output data in HBM */
/* dense-output data of particle i in HBM: [59][n] = xo dxo k3 k4 k5 k6 k7 (8 each), t, t_old, pol */
struct DenseRef { const Params *P; int i; };
PTD void dense_calc(const DenseRef &r, double t, double *y) {
    const double *ds = r.P->st.dense;
    const size_t n = (size_t)r.P->st.n;
    const int i = r.i;
    constexpr double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192, b5 = -2187.0 / 6784, b6 = 11.0 / 84;
    const double st_t = gload(ds + 56 * n + i), st_told = gload(ds + 57 * n + i), polo = gload(ds + 58 * n + i);
    const double dt = (st_t - st_told);
    const double theta = (t - st_told) / dt;
    const double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
    const double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
    const double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
    const double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
    const double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
    const double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
    const double theta_m_1 = theta - 1.0;
    const double theta_sq = theta * theta;
    const double A = theta_sq * (3.0 - 2.0 * theta);
    const double B = theta_sq * theta_m_1;
    const double C = theta_sq * theta_m_1 * theta_m_1;
    const double D = theta * theta_m_1 * theta_m_1;
    const double h1 = dt * (A * b1 - C * X1 + D);
    const double h3 = dt * (A * b3 + C * X3);
    const double h4 = dt * (A * b4 - C * X4);
    const double h5 = dt * (A * b5 + C * X5);
    const double h6 = dt * (A * b6 - C * X6);
    const double h7 = dt * (B + C * X7);
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        x[k] = gload(ds + k * n + i) + h1 * gload(ds + (8 + k) * n + i) + h3 * gload(ds + (16 + k) * n + i) +
               h4 * gload(ds + (24 + k) * n + i) + h5 * gload(ds + (32 + k) * n + i) + h6 * gload(ds + (40 + k) * n + i) +
               h7 * gload(ds + (48 + k) * n + i);
    to9(x, polo + h1 * 0.0 + h3 * 0.0 + h4 * 0.0 + h5 * 0.0 + h6 * 0.0 + h7 * 0.0, y);
}

PTD void dense_calc(const Dopri &s, double t, double *y) { dopri_calc_state(s, t, y); }

/* TLogger::PrintSnapshot (src/logger.cpp:128-140): the index of the first snapshot time in
   the configured order (find_if) with x1 <= t < x2, -1 if none; a sorted list by bisection */
#define NOINLINE_SNAP __attribute__((noinline))
PTD int snapshot_in(const State &S, double x1, double x2) {
    if (!S.snap_sorted) {
        for (int j = 0; j < S.snap_n; ++j)
            if (x1 <= S.snap_t[j] && S.snap_t[j] < x2) return j;
        return -1;
    }
    int lo = 0, hi = S.snap_n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (S.snap_t[mid] < x1) lo = mid + 1;
        else hi = mid;
    }
    return (lo < S.snap_n && S.snap_t[lo] < x2) ? lo : -1;
}

/* TLogger::PrintTrack (src/logger.cpp:142-153, src/tracking.cpp:140): the step's end state
   when the trajectory length crossed a multiple of trackloginterval since the stepper's step
   start (l1) */
PTD void record_track(const State &S, int i, double l1, double x, const double *y, unsigned solid) {
    if (y[8] > 0 && (int)(l1 / S.trk_interval) == (int)(y[8] / S.trk_interval)) return;
    const int k = S.ntrk[i];
    S.ntrk[i] = k + 1;
    if (k >= S.trk_cap) return;
    TrackRec &r = S.trk[(size_t)i * S.trk_cap + k];
    r.t = x;
    for (int q = 0; q < 9; ++q) r.y[q] = y[q];
    r.solid = (int)solid;
    r.pad = 0;
}

/* record the stepper's dense-output state at the snapshot time in [x1, x2), if any, with
   the particle's spin and counters at that point of the step */
template <class DenseSrc>
PTD NOINLINE_SNAP void record_snapshot(const State &S, int i, const DenseSrc &d, double x1, double x2,
                                       const double *spin_or_null, int stop, int Nhit, int Nstep, int Nspinflip,
                                       double Hmax, double noflip) {
    const int j = snapshot_in(S, x1, x2);
    if (j < 0) return;
    const int k = S.nsnap[i];
    S.nsnap[i] = k + 1;
    if (k >= S.snap_cap) return;
    Snap &r = S.snaps[(size_t)i * S.snap_cap + k];
    r.t = S.snap_t[j];
    double y[9];
    dense_calc(d, r.t, y);
    for (int q = 0; q < 9; ++q) r.y[q] = y[q];
    for (int q = 0; q < 5; ++q) r.spin[q] = spin_or_null ? spin_or_null[q] : S.spin[(size_t)q * S.n + i];
    r.Hmax = Hmax; r.noflip = noflip;
    r.stop = stop; r.Nhit = Nhit; r.Nstep = Nstep; r.Nspinflip = Nspinflip;
}

struct Track {
    int i;                      /* particle index */
    Species sp;
    uint64_t key, ctr;
    unsigned long long solids;  /* packed solid stack (TTracker::currentsolids) */
    int nsolid;
    int stop, Nhit, Nstep, Nspinflip, nhitlog;
    long long nacc, nrej;
    double Hmax, noflip;
    Ctr c;
    Cand cand;                  /* candidate triangles of the chord piece under test */
};

PTD double rnd_uniform(Track &T, double a, double b) {
    return ptm_canonical(ptm_draw_u64(T.key, &T.ctr)) * (b - a) + a;
}
PTD double rnd_sincos(Track &T, double amin, double amax) {
    double x = ptm_canonical(ptm_draw_u64(T.key, &T.ctr));
    double cmax = ptm_cos(amax), cmin = ptm_cos(amin);
    return ptm_acos(__builtin_sqrt(x * (cmax * cmax - cmin * cmin) + cmin * cmin));
}

PTD bool state_eq(const double *a, const double *b) {
    bool eq = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) eq = eq && (a[i] == b[i]);
    return eq;
}

/* RotateVector, src/globals.cpp:40-64 */
PTD void rotate_vector(double v[3], const double n[3], const double *x) {
    double xv[3], zv[3] = {n[0], n[1], n[2]};
    double pa = zv[0], pb = zv[1], pc = zv[2];
    double pd = -zv[0] * 0.0 - zv[1] * 0.0 - zv[2] * 0.0;
    bool use_base = (x == nullptr);
    if (!use_base) {
        double num = pa * x[0] + pb * x[1] + pc * x[2] + pd;
        double den = pa * pa + pb * pb + pc * pc;
        double lambda = num / den;
        xv[0] = x[0] - lambda * pa;
        xv[1] = x[1] - lambda * pb;
        xv[2] = x[2] - lambda * pc;
        if (xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2] < 1e-30) use_base = true;
    }
    if (use_base) {
        if (pa == 0) { xv[0] = 1; xv[1] = 0; xv[2] = 0; }
        else if (pb == 0) { xv[0] = 0; xv[1] = 1; xv[2] = 0; }
        else if (pc == 0) { xv[0] = 0; xv[1] = 0; xv[2] = 1; }
        else {
            double a = dabs(pa), b = dabs(pb), c = dabs(pc);
            if (a <= b && a <= c) { xv[0] = 0; xv[1] = -pc; xv[2] = pb; }
            else if (b <= a && b <= c) { xv[0] = -pc; xv[1] = 0; xv[2] = pa; }
            else { xv[0] = -pb; xv[1] = pa; xv[2] = 0; }
        }
    }
    double lx = __builtin_sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    xv[0] = xv[0] / lx; xv[1] = xv[1] / lx; xv[2] = xv[2] / lx;
    double lz = __builtin_sqrt(zv[0] * zv[0] + zv[1] * zv[1] + zv[2] * zv[2]);
    zv[0] = zv[0] / lz; zv[1] = zv[1] / lz; zv[2] = zv[2] / lz;
    double yv[3] = {zv[1] * xv[2] - zv[2] * xv[1], zv[2] * xv[0] - zv[0] * xv[2], zv[0] * xv[1] - zv[1] * xv[0]};
    double v0 = v[0], v1 = v[1], v2 = v[2];
    v[0] = xv[0] * v0 + yv[0] * v1 + zv[0] * v2;
    v[1] = xv[1] * v0 + yv[1] * v1 + zv[1] * v2;
    v[2] = xv[2] * v0 + yv[2] * v1 + zv[2] * v2;
}

struct V3 { double x, y, z; };
/* RotateVector with values in and out (no address-taken arrays on the step path) */
PTD_NOINLINE V3 rotate_vector_v(V3 v, V3 n) {
    double a[3] = {v.x, v.y, v.z}, nn[3] = {n.x, n.y, n.z};
    rotate_vector(a, nn, nullptr);
    V3 r = {a[0], a[1], a[2]};
    return r;
}

/* neutron wall interaction, TNeutron::OnHit and helpers (src/neutron.cpp:25-264) */
PTD double material_potential(const Species &sp, const double *mat, const double *y) {
    return mat[0] * 1e-9 - mat[6] * sp.mu * y[7] / ELE_E;
}

PTD void neutron_reflect(const double *y1, double &x2, double x1, double *y2, const double *normal) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    x2 = x1;
    y2[0] = y1[0]; y2[1] = y1[1]; y2[2] = y1[2];
    y2[3] -= 2 * vnormal * normal[0];
    y2[4] -= 2 * vnormal * normal[1];
    y2[5] -= 2 * vnormal * normal[2];
    y2[6] = y1[6];
    y2[8] = y1[8];
}

PTD_NOINLINE void neutron_reflect_lambert(Track &T, const double *y1, double &x2, double x1, double *y2, const double *normal, const double *mat) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double n[3] = {normal[0], normal[1], normal[2]};
    if (vnormal > 0) { vnormal *= -1; n[0] *= -1; n[1] *= -1; n[2] *= -1; }
    x2 = x1;
    y2[0] = y1[0]; y2[1] = y1[1]; y2[2] = y1[2];
    double vabs = __builtin_sqrt(y1[3] * y1[3] + y1[4] * y1[4] + y1[5] * y1[5]);
    do {
        double phi_r = rnd_uniform(T, 0, 2. * PI);
        double theta_r = rnd_sincos(T, 0, PI / 2.);
        y2[3] = vabs * ptm_cos(phi_r) * ptm_sin(theta_r);
        y2[4] = vabs * ptm_sin(phi_r) * ptm_sin(theta_r);
        y2[5] = vabs * ptm_cos(theta_r);
        if (mat[2] > 0) {
            rotate_vector(y2 + 3, n, y1 + 3);
        } else {
            double sr[3] = {y1[3] - 2 * vnormal * n[0], y1[4] - 2 * vnormal * n[1], y1[5] - 2 * vnormal * n[2]};
            rotate_vector(y2 + 3, sr, y1 + 3);
        }
    } while (y2[3] * n[0] + y2[4] * n[1] + y2[5] * n[2] <= 0);
    y2[6] = y1[6];
    y2[8] = y1[8];
}

PTD void neutron_transmit(const Species &sp, const double *y1, double *y2, const double *normal, double Estep) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double Enormal = 0.5 * sp.m * vnormal * vnormal;
    double k1 = __builtin_sqrt(Enormal);
    double k2 = __builtin_sqrt(Enormal - Estep);
    for (int i = 0; i < 3; i++) y2[i + 3] += (k2 / k1 - 1) * (normal[i] * vnormal);
}

PTD_NOINLINE void neutron_transmit_lambert(Track &T, const double *y1, double *y2, const double *normal, double Estep, const double *mat) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double n[3] = {normal[0], normal[1], normal[2]};
    if (vnormal < 0) { vnormal *= -1; n[0] *= -1; n[1] *= -1; n[2] *= -1; }
    double Enormal = 0.5 * T.sp.m * vnormal * vnormal;
    double k1 = __builtin_sqrt(Enormal);
    double k2 = __builtin_sqrt(Enormal - Estep);
    double st[3] = {y2[3], y2[4], y2[5]};
    for (int i = 0; i < 3; ++i) st[i] += ((k2 / k1 - 1) * n[i]) * vnormal;
    double vabs = __builtin_sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
    do {
        double phi_t = rnd_uniform(T, 0, 2. * PI);
        double theta_t = rnd_sincos(T, 0, PI / 2.);
        y2[3] = vabs * ptm_cos(phi_t) * ptm_sin(theta_t);
        y2[4] = vabs * ptm_sin(phi_t) * ptm_sin(theta_t);
        y2[5] = vabs * ptm_cos(theta_t);
        if (mat[2] > 0) rotate_vector(y2 + 3, n, y1 + 3);
        else rotate_vector(y2 + 3, st, y1 + 3);
    } while (y2[3] * n[0] + y2[4] * n[1] + y2[5] * n[2] <= 0);
}

PTD double refl_prob(double E1r, double E1i, double E2r, double E2i) {
    double k1r, k1i, k2r, k2i, qr, qi;
    ptm_csqrt(E1r, E1i, &k1r, &k1i);
    ptm_csqrt(E2r, E2i, &k2r, &k2i);
    ptm_cdiv(k1r - k2r, k1i - k2i, k1r + k2r, k1i + k2i, &qr, &qi);
    double a = ptm_hypot(qr, qi);
    return a * a;
}

PTD_NOINLINE void neutron_onhit(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2,
                                const double *normal, int leaving, int entering) {
    const Geom &G = P.geom;
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double Enormal = 0.5 * T.sp.m * vnormal * vnormal;
    const double *lmat = solid_mat(G, leaving), *emat = solid_mat(G, entering);
    const double *mat = vnormal < 0 ? emat : lmat;
    if (rnd_uniform(T, 0, 1) < mat[3]) y2[7] *= -1;
    double Estep = material_potential(T.sp, emat, y2) - material_potential(T.sp, lmat, y2);
    if (mat[4] != 0 && mat[5] != 0) { T.stop = ID_UNSUPPORTED; return; }
    double prob = rnd_uniform(T, 0, 1);
    double reflprob = refl_prob(Enormal, -lmat[1] * 1e-9, Enormal - Estep, -emat[1] * 1e-9);
    if (Enormal > Estep) {
        if (prob < 0 + 0 + reflprob * (1 - 0 - 0)) {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_reflect_lambert(T, y1, x2, x1, y2, normal, mat);
            else neutron_reflect(y1, x2, x1, y2, normal);
        } else {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_transmit_lambert(T, y1, y2, normal, Estep, mat);
            else neutron_transmit(T.sp, y1, y2, normal, Estep);
        }
    } else {
        /* LossPerBounce of the material, or the particle's own where a per-particle loss is set */
        const double loss = P.st.loss ? P.st.loss[T.i] : mat[8];
        double absprob = 1 - reflprob + loss;
        if (prob < 0 + 0 + absprob * (1 - 0 - 0)) {
            T.stop = ID_ABSORBED_ON_SURFACE;
        } else {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_reflect_lambert(T, y1, x2, x1, y2, normal, mat);
            else neutron_reflect(y1, x2, x1, y2, normal);
        }
    }
}

/* OnStep hooks (src/neutron.cpp:267-282, src/proton.cpp:25-33, src/electron.cpp:25-33) */
template <class DS>
PTD void on_step(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2, const DS &st, int cursolid) {
    if (T.sp.kind == 0) {
        const double *mat = solid_mat(P.geom, cursolid);
        if (mat[1] > 0) {
            double mn = T.sp.m, e = ELE_E, hb = HBAR;
            double Er = 0.5 * mn * (y1[3] * y1[3] + y1[4] * y1[4] + y1[5] * y1[5]);
            double Ei = mat[1] * 1e-9;
            double kr, ki;
            ptm_csqrt((2 * mn) * Er, (2 * mn) * Ei, &kr, &ki);
            ki = ki * e;
            ki = ki / hb;
            double l = __builtin_sqrt((y2[0] - y1[0]) * (y2[0] - y1[0]) + (y2[1] - y1[1]) * (y2[1] - y1[1]) + (y2[2] - y1[2]) * (y2[2] - y1[2]));
            double abspath = -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(T.key, &T.ctr))) / (2 * ki);
            if (abspath < l) {
                x2 = x1 + abspath / l * (x2 - x1);
                dense_calc(st, x2, y2);
                T.stop = ID_ABSORBED_IN_MATERIAL;
            }
        }
    } else {
        if (solid_ID(P.geom, cursolid) > 1) {
            x2 = x1;
            for (int k = 0; k < 9; ++k) y2[k] = y1[k];
            T.stop = ID_ABSORBED_IN_MATERIAL;
        }
    }
}

/* TTracker::DoStep + TParticle::DoStep (src/tracking.cpp:268-278, src/particle.cpp:196-206) */
template <int MODE>
PTD_NOINLINE bool do_step_hook(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2, const Dopri &st, int cursolid) {
    double x2temp = x2, y2temp[9];
    for (int k = 0; k < 9; ++k) y2temp[k] = y2[k];
    on_step(P, T, x1, y1, x2, y2, st, cursolid);
    double H = kinetic_energy(T.sp, y2 + 3) + potential_energy<MODE>(P, T.sp, y2, x2, nullptr, nullptr, T.c);
    T.Hmax = (H < T.Hmax) ? T.Hmax : H;
    if (y2temp[7] != y2[7]) T.Nspinflip++;
    T.Nstep++;
    return !(x2temp == x2 && state_eq(y2temp, y2));
}

/* TTracker::DoHit (src/tracking.cpp:280-365) */
PTD_NOINLINE bool do_hit(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2) {
    const Geom &G = P.geom;
    HitList L;
    L.n = 0;
    int nc = cand_covers(T.cand, y1, y2) ? cand_query(G, T.cand, y1, y2, &L, &T.c.tri) : seg_query(G, y1, y2, &L, &T.c.tri);
    if (nc <= 0) { T.stop = nc == -1 ? ID_CGAL_ERROR : (nc == -2 ? ID_UNSUPPORTED : ID_INVARIANT); return true; }
    double sv[3] = {y2[0] - y1[0], y2[1] - y1[1], y2[2] - y1[2]};
    unsigned long long ns = T.solids;
    int nn = T.nsolid;
    bool trajectoryaltered = false, traversed = true;
    for (int k = 0; k < nc; ++k) {
        unsigned id = L.id[k];
        const double *n = G.nrm + 3 * L.tri[k];
        double distnormal = sv[0] * n[0] + sv[1] * n[1] + sv[2] * n[2];
        int solid = 0;
        for (int m = 0; m < G.n_solid; ++m) if (G.solid_id[m] == id) { solid = m; break; }
        int found = -1;
        for (int i = 0; i < nn; ++i) if (solid_ID(G, sk_sidx(ns, i)) == id) { found = i; break; }
        if (distnormal < 0) {
            if (found >= 0) {
                if (L.s[k] > 0) {
                    if (nn >= MAX_SOLIDS) { T.stop = ID_UNSUPPORTED; return true; }
                    ns = sk_set(ns, nn, solid, sk_ign(ns, found)); nn++;
                }
            } else {
                if (nn >= MAX_SOLIDS) { T.stop = ID_UNSUPPORTED; return true; }
                int ign = solid_ignored(G, solid, x1 + (x2 - x1) * L.s[k]);
                ns = sk_set(ns, nn, solid, ign); nn++;
            }
        } else if (distnormal > 0) {
            if (found < 0) {
                if (L.s[k] > 0) { T.stop = ID_GEOMETRY_ERROR; return true; }
            } else {
                ns = sk_erase(ns, found); nn--;
            }
        } else {
            T.stop = ID_GEOMETRY_ERROR;
            return true;
        }
    }
    int leaving = current_solid(G, T.solids, T.nsolid);
    int entering = -1;
    for (int i = 0; i < nn; ++i)
        if (!sk_ign(ns, i) && solid_ID(G, sk_sidx(ns, i)) > solid_ID(G, entering)) entering = sk_sidx(ns, i);
    unsigned lid = solid_ID(G, leaving), eid = solid_ID(G, entering);
    if (lid != eid) {
        int sel = -1;
        for (int k = 0; k < nc; ++k) {
            if (L.id[k] == lid) { sel = k; break; }
            if (L.id[k] == eid) {
                int solid = 0;
                for (int m = 0; m < G.n_solid; ++m) if (G.solid_id[m] == eid) { solid = m; break; }
                if (!solid_ignored(G, solid, x1 + (x2 - x1) * L.s[k])) { sel = k; break; }
            }
        }
        if (sel < 0) { T.stop = ID_INVARIANT; return true; }
        const double *np = G.nrm + 3 * L.tri[sel];
        double normal[3] = {np[0], np[1], np[2]};
        double x2temp = x2, y2temp[9];
        for (int k = 0; k < 9; ++k) y2temp[k] = y2[k];
        if (T.sp.kind == 0) neutron_onhit(P, T, x1, y1, x2, y2, normal, leaving, entering);
        if (y2temp[7] != y2[7]) T.Nspinflip++;
        T.Nhit++;
        if (x2temp == x2 && state_eq(y2temp, y2)) {
            trajectoryaltered = false; traversed = true;
        } else {
            trajectoryaltered = true;
            if (x2 == x2temp && y2[0] == y2temp[0] && y2[1] == y2temp[1] && y2[2] == y2temp[2]) traversed = true;
            else if (x2 == x1 && y2[0] == y1[0] && y2[1] == y1[1] && y2[2] == y1[2]) traversed = false;
            else { T.stop = ID_INVARIANT; return true; }
        }
        if (T.nhitlog < P.st.hit_cap) {
            Hit &h = P.st.hits[(size_t)T.i * P.st.hit_cap + T.nhitlog];
            h.t = x1; h.tri = L.orig[sel]; h.leaving = (int)lid; h.entering = (int)eid;
            for (int k = 0; k < 3; ++k) {
                h.normal[k] = normal[k]; h.v1[k] = y1[3 + k]; h.v2[k] = y2[3 + k]; h.pos[k] = y1[k];
            }
            h.pol1 = y1[7]; h.pol2 = y2[7];
        }
        T.nhitlog++;
    }
    if (traversed) { T.solids = ns; T.nsolid = nn; }
    return trajectoryaltered || T.stop != ID_UNKNOWN;
}

/* Number of intersections of segment p1->p2 with the mesh (-1 coplanar overlap, -2 more
   than MAX_HITS): nothing when the clearance grid clears the segment, else from the
   candidate list of its midpoint's cell when that covers it, else from the BVH subtree of
   the cell's triangles when the cell widened by rt_pad covers it, else from the whole BVH;
   all give the BVH query's result.  TURNS: the traversals of the lanes that need one run
   one after the other on the scalar path (seg_count_lanes) instead of side by side. */
template <bool TURNS>
PTD int count_hits_geom(const Geom &G, const double *p1, const double *p2, long long &tests) {
    int cell[3];
    if (clear_of_walls(G, p1, p2, cell)) return 0;
    const int2 cl = cell_list(G, cell, p1, p2);
    int2 r;
    if (cl.y >= 0) {
        r = cell_list_count(G, cl.x, cl.y, p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]);
    } else {
        const int root = cell_root(G, cell, p1, p2);
        if (root == -2) return 0;
        /* the step kernels' lanes rarely need a traversal, and then it runs faster on the
           scalar path, one lane's at a time; the collision kernel's lanes often need one */
        r = TURNS ? seg_count_lanes(G, p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], root)
                  : seg_query_v(G, p1[0], p1[1], p1[2], p2[0], p2[1], p2[2], nullptr, root);
    }
    tests += r.y;
    return r.x;
}

/* on the collision path: from the piece's candidates when they cover the segment (an exact
   count; the clearance grid never applies this close to a wall), else as on the step path */
PTD int count_hits(const Params &P, Track &T, const double *p1, const double *p2) {
    if (cand_covers(T.cand, p1, p2)) return cand_query(P.geom, T.cand, p1, p2, nullptr, &T.c.tri);
    return count_hits_geom<false>(P.geom, p1, p2, T.c.tri);
}

PTD int count_hits(const Params &P, Ctr &c, const double *p1, const double *p2) {
    return count_hits_geom<true>(P.geom, p1, p2, c.tri);
}

/*
 * TTracker::iterate_collision (src/tracking.cpp:228-265) without recursion.
 * The recursion is a depth-first search over dyadic sub-intervals; a frame is
 * identified by the bit string of first/second-half choices leading to it, so
 * the search keeps only that bit string (depth <= 100) and recomputes interval
 * ends by replaying the halvings and states by dense output.  Returns 1 with the
 * final bracket in (x1,y1,x2,y2), 0 if every branch failed, -1 on a geometry error.
 */
PTD_NOINLINE int iterate_collision(const Params &P, Track &T, const Dopri &st, double &X1, double *Y1, double &X2, double *Y2) {
    unsigned long long ph[2] = {0ull, 0ull};   /* phase bit per level: 0 first half, 1 second half */
    int d = 0;
    double lo = X1, hi = X2;
    double plo[3] = {Y1[0], Y1[1], Y1[2]}, phi[3] = {Y2[0], Y2[1], Y2[2]};
    double pmid[3] = {0., 0., 0.};
    double mid = 0.;
    /* one segment query per iteration: second = false tests the first half of frame d
       (after the frame's stop checks and midpoint), second = true its second half; the
       queries and the decisions come in the recursion's order */
    bool second = false;
    for (;;) {
        double a[3], b[3];
        if (!second) {
            double dx0 = phi[0] - plo[0], dx1 = phi[1] - plo[1], dx2 = phi[2] - plo[2];
            if (dx0 * dx0 + dx1 * dx1 + dx2 * dx2 < REFLECT_TOLERANCE * REFLECT_TOLERANCE) break;
            if (hi - lo < 4 * (lo + hi) * DBL_EPS) break;
            if (d >= 100) break;
            mid = lo + (hi - lo) * 0.5;
            double ymid[9];
            dopri_calc_state(st, mid, ymid);
            pmid[0] = ymid[0]; pmid[1] = ymid[1]; pmid[2] = ymid[2];
            a[0] = plo[0]; a[1] = plo[1]; a[2] = plo[2];
            b[0] = pmid[0]; b[1] = pmid[1]; b[2] = pmid[2];
        } else {
            a[0] = pmid[0]; a[1] = pmid[1]; a[2] = pmid[2];
            b[0] = phi[0]; b[1] = phi[1]; b[2] = phi[2];
        }
        const int nh = count_hits(P, T, a, b);
        if (nh < 0) return -1;
        if (nh > 0) {
            /* descend into the half that holds a collision */
            if (!second) {
                ph[d >> 6] &= ~(1ull << (d & 63));
                hi = mid; phi[0] = pmid[0]; phi[1] = pmid[1]; phi[2] = pmid[2];
            } else {
                ph[d >> 6] |= (1ull << (d & 63));
                lo = mid; plo[0] = pmid[0]; plo[1] = pmid[1]; plo[2] = pmid[2];
            }
            ++d;
            second = false;
            continue;
        }
        if (!second) {
            second = true;
            continue;
        }
        /* backtrack to the nearest ancestor whose first half was tried */
        for (;;) {
            if (d == 0) return 0;
            --d;
            if (((ph[d >> 6] >> (d & 63)) & 1ull) == 0ull) break;
        }
        /* rebuild frame d: interval by replaying the halvings, states by dense output */
        lo = X1; hi = X2;
        bool hiroot = true;
        for (int k = 0; k < d; ++k) {
            double m = lo + (hi - lo) * 0.5;
            if ((ph[k >> 6] >> (k & 63)) & 1ull) { lo = m; }
            else { hi = m; hiroot = false; }
        }
        mid = lo + (hi - lo) * 0.5;
        double ymid[9];
        dopri_calc_state(st, mid, ymid);
        pmid[0] = ymid[0]; pmid[1] = ymid[1]; pmid[2] = ymid[2];
        if (hiroot) { phi[0] = Y2[0]; phi[1] = Y2[1]; phi[2] = Y2[2]; }
        else { double yh[9]; dopri_calc_state(st, hi, yh); phi[0] = yh[0]; phi[1] = yh[1]; phi[2] = yh[2]; }
        second = true;
    }
    /* success: the final bracket; its ends are root states or dense-output states */
    bool loroot = true, hiroot = true;
    for (int k = 0; k < d; ++k) {
        if ((ph[k >> 6] >> (k & 63)) & 1ull) loroot = false;
        else hiroot = false;
    }
    if (!loroot) { X1 = lo; dopri_calc_state(st, lo, Y1); }
    if (!hiroot) { X2 = hi; dopri_calc_state(st, hi, Y2); }
    return 1;
}

/* TTracker::CheckHit (src/tracking.cpp:176-226); the tail recursion is a loop */
template <int MODE>
PTD_NOINLINE bool check_hit(const Params &P, Track &T, double x1, const double *y1in, double &x2, double *y2, const Dopri &st) {
    double y1[9];
    for (int k = 0; k < 9; ++k) y1[k] = y1in[k];
    {
        double lo[3], hi[3];
        piece_cand_box(y1, y2, lo, hi);
        cand_build(P.geom, lo, hi, T.cand);
    }
    for (;;) {
        if (!check_segment(P.geom, y1, y2)) { T.stop = ID_HIT_BOUNDARIES; return true; }
        if (x2 == x1) return false;
        int cursolid = current_solid(P.geom, T.solids, T.nsolid);
        int n = count_hits(P, T, y1, y2);
        if (n < 0) { T.stop = n == -1 ? ID_CGAL_ERROR : ID_UNSUPPORTED; return true; }
        if (n == 0) return do_step_hook<MODE>(P, T, x1, y1, x2, y2, st, cursolid);
        double xc1 = x1, xc2 = x2, yc1[9], yc2[9];
        for (int k = 0; k < 9; ++k) { yc1[k] = y1[k]; yc2[k] = y2[k]; }
        int r = iterate_collision(P, T, st, xc1, yc1, xc2, yc2);
        if (r < 0) { T.stop = ID_CGAL_ERROR; return true; }
        if (r == 0) return false;
        if (xc1 > x1 && do_step_hook<MODE>(P, T, x1, y1, xc1, yc1, st, cursolid)) {
            x2 = xc1;
            for (int k = 0; k < 9; ++k) y2[k] = yc1[k];
            return true;
        }
        if (do_hit(P, T, xc1, yc1, xc2, yc2)) {
            x2 = xc2;
            for (int k = 0; k < 9; ++k) y2[k] = yc2[k];
            return true;
        }
        x1 = xc2;
        for (int k = 0; k < 9; ++k) y1[k] = yc2[k];
    }
}

/* TParticle::DoPolarize (src/particle.cpp:225-234) */
PTD void do_polarize(Track &T, double *y, double polarization, int flipspin) {
    double flipprob = 0.5 * (1 - y[7] * polarization);
    T.noflip *= 1. - flipprob;
    if (flipspin) {
        double prevpol = y[7];
        y[7] = (ptm_canonical(ptm_draw_u64(T.key, &T.ctr)) < 0.5 * (polarization + 1)) ? 1 : -1;
        if (y[7] != prevpol) ++T.Nspinflip;
    }
}

/* spin-tracking windows of IntegrateSpin (src/tracking.cpp:397-400) */
PTD bool spin_window(const Species &sp, double t) {
    bool in = false;
    for (int k = 0; k + 1 < sp.n_spintimes; k += 2) in |= (t >= sp.spintimes[k] && t < sp.spintimes[k + 1]);
    return in;
}

/* TParticle::SpinDerivs without field interpolation (src/particle.cpp:141-190): the
   precession axis (BMT with Thomas precession) from the trajectory's dense output at t */
template <int MODE>
PTD void spin_axis(const Params &P, Track &T, const Dopri &traj, double t, double O[3], double *pos = nullptr) {
    double yt[9], d9[9], B[3], dB[3][3], E[3] = {0, 0, 0}, V = 0;
    dopri_calc_state(traj, t, yt);
    if (pos) { pos[0] = yt[0]; pos[1] = yt[1]; pos[2] = yt[2]; }
    field_B<MODE>(P, yt[0], yt[1], yt[2], t, B, dB, true, T.c.field);
    field_E<MODE>(P, yt[0], yt[1], yt[2], t, &V, E, T.c.field);
    eom(T.sp, yt, B, dB, E, d9);
    const double gamma = T.sp.gamma;
    const double v2 = d9[0] * d9[0] + d9[1] * d9[1] + d9[2] * d9[2];
    const double gamma_rel = 1. / __builtin_sqrt(1. - v2 / C_0 / C_0);
    const double Bdotv = B[0] * d9[0] + B[1] * d9[1] + B[2] * d9[2];
    double Bp[3];
    for (int j = 0; j < 3; j++) Bp[j] = Bdotv * d9[j] / v2;
    O[0] = -gamma / gamma_rel * ((1 - gamma_rel) * Bp[0] + gamma_rel * B[0] - gamma_rel * (d9[1] * E[2] - d9[2] * E[1]) / C_0 / C_0) +
           gamma_rel * gamma_rel / (gamma_rel + 1) * (d9[4] * d9[2] - d9[5] * d9[1]) / C_0 / C_0;
    O[1] = -gamma / gamma_rel * ((1 - gamma_rel) * Bp[1] + gamma_rel * B[1] - gamma_rel * (d9[2] * E[0] - d9[0] * E[2]) / C_0 / C_0) +
           gamma_rel * gamma_rel / (gamma_rel + 1) * (d9[5] * d9[0] - d9[3] * d9[2]) / C_0 / C_0;
    O[2] = -gamma / gamma_rel * ((1 - gamma_rel) * Bp[2] + gamma_rel * B[2] - gamma_rel * (d9[0] * E[1] - d9[1] * E[0]) / C_0 / C_0) +
           gamma_rel * gamma_rel / (gamma_rel + 1) * (d9[3] * d9[1] - d9[4] * d9[0]) / C_0 / C_0;
}

/* TLogger::PrintSpin (src/logger.cpp:237-284): a row at time x unless x lies after the spin
   stepper's previous time x1 in the same multiple of spinloginterval.  For the row right
   after the stepper's initialize (src/tracking.cpp:430-431) x1 is the fresh stepper's
   value-initialised previous time, 0.0.  The spin at x is the integrator's. */
template <int MODE>
PTD void record_spin(const Params &P, Track &T, const Dopri &traj, double x1, double x, const double *spin) {
    const State &S = P.st;
    if (x > x1 && (int)(x1 / S.slog_interval) == (int)(x / S.slog_interval)) return;
    const int k = S.nslog[T.i];
    S.nslog[T.i] = k + 1;
    if (k >= S.slog_cap) return;
    SpinRec &r = S.slog[(size_t)T.i * S.slog_cap + k];
    double O[3], pos[3];
    spin_axis<MODE>(P, T, traj, x, O, pos);
    r.t = x;
    for (int q = 0; q < 3; ++q) { r.pos[q] = pos[q]; r.spin[q] = spin[q]; r.W[q] = O[q]; }
}

/* interpolatefields: ALGLIB spline1dbuildcubic (parabolic ends) of each axis component over
   SPIN_PTS + 1 points of the step (src/tracking.cpp:404-420), spline1dcalc by bisection and
   the Hermite piece c0 + s (c1 + s (c2 + s c3)) */
constexpr int SPIN_PTS = 10;
struct SpinSplines { double ts[SPIN_PTS + 1]; double c[3][4 * SPIN_PTS]; };

template <int MODE>
PTD void spin_splines_build(const Params &P, Track &T, const Dopri &traj, double x1, double x2, SpinSplines &S) {
    constexpr int n = SPIN_PTS + 1;
    double om[3][n];
    for (int i = 0; i < n; i++) {
        const double t = x1 + i * (x2 - x1) / SPIN_PTS;
        double O[3];
        S.ts[i] = t;
        spin_axis<MODE>(P, T, traj, t, O);
        om[0][i] = O[0]; om[1][i] = O[1]; om[2][i] = O[2];
    }
    const double *x = S.ts;
    for (int k = 0; k < 3; ++k) {
        const double *y = om[k];
        /* spline1dgriddiffcubic: tridiagonal system with parabolic ends (Thomas algorithm) */
        double a2[n], r[n], d[n];
        a2[0] = 1; r[0] = 2 * (y[1] - y[0]) / (x[1] - x[0]);
        double a3p = 1;
        for (int i = 1; i <= n - 2; i++) {
            const double a1 = x[i + 1] - x[i], a2i = 2 * (x[i + 1] - x[i - 1]), a3 = x[i] - x[i - 1];
            const double b = 3 * (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (x[i + 1] - x[i]) +
                             3 * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) * (x[i] - x[i - 1]);
            const double tt = a1 / a2[i - 1];
            a2[i] = a2i - tt * a3p;
            r[i] = b - tt * r[i - 1];
            a3p = a3;
        }
        {
            const double b = 2 * (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
            const double tt = 1.0 / a2[n - 2];
            a2[n - 1] = 1.0 - tt * a3p;
            r[n - 1] = b - tt * r[n - 2];
        }
        d[n - 1] = r[n - 1] / a2[n - 1];
        for (int i = n - 2; i >= 0; --i) d[i] = (r[i] - (i == 0 ? 1.0 : x[i] - x[i - 1]) * d[i + 1]) / a2[i];
        for (int i = 0; i <= n - 2; i++) {
            const double delta = x[i + 1] - x[i], delta2 = delta * delta, delta3 = delta * delta2;
            S.c[k][4 * i + 0] = y[i];
            S.c[k][4 * i + 1] = d[i];
            S.c[k][4 * i + 2] = (3 * (y[i + 1] - y[i]) - 2 * d[i] * delta - d[i + 1] * delta) / delta2;
            S.c[k][4 * i + 3] = (2 * (y[i] - y[i + 1]) + d[i] * delta + d[i + 1] * delta) / delta3;
        }
    }
}

PTD double spin_spline_calc(const SpinSplines &S, int k, double x) {
    int l = 0, r = SPIN_PTS;
    while (l != r - 1) {
        const int m = (l + r) / 2;
        if (S.ts[m] >= x) r = m;
        else l = m;
    }
    x = x - S.ts[l];
    const double *c = S.c[k] + 4 * l;
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}

/* TParticle::SpinDerivs (src/particle.cpp:177-190) */
template <int MODE>
PTD void spin_derivs(const Params &P, Track &T, const Dopri &traj, const SpinSplines *spl, const double *y,
                     double *dydx, double t) {
    double O[3];
    if (spl) {
        O[0] = spin_spline_calc(*spl, 0, t);
        O[1] = spin_spline_calc(*spl, 1, t);
        O[2] = spin_spline_calc(*spl, 2, t);
    } else {
        spin_axis<MODE>(P, T, traj, t, O);
    }
    dydx[0] = O[1] * y[2] - O[2] * y[1];
    dydx[1] = O[2] * y[0] - O[0] * y[2];
    dydx[2] = O[0] * y[1] - O[1] * y[0];
    dydx[3] = 1.;
    dydx[4] = __builtin_sqrt(O[0] * O[0] + O[1] * O[1] + O[2] * O[2]);
}

/* the spin integration of IntegrateSpin (src/tracking.cpp:424-450): a dense-output dopri5
   (rtol = atol = 1e-12) on the five spin variables from x1, first step half a Larmor
   period, until it passes x2; the state at x2 from its dense output.  false if a step
   fails 500 times. */
template <int MODE>
PTD_NOINLINE bool bmt_integrate(const Params &P, Track &T, const Dopri &traj, double *spin, double x1, double x2,
                                double Babs1) {
    constexpr double a2 = 1.0 / 5, a3 = 3.0 / 10, a4 = 4.0 / 5, a5 = 8.0 / 9;
    constexpr double b21 = 1.0 / 5, b31 = 3.0 / 40, b32 = 9.0 / 40, b41 = 44.0 / 45, b42 = -56.0 / 15, b43 = 32.0 / 9;
    constexpr double b51 = 19372.0 / 6561, b52 = -25360.0 / 2187, b53 = 64448.0 / 6561, b54 = -212.0 / 729;
    constexpr double b61 = 9017.0 / 3168, b62 = -355.0 / 33, b63 = 46732.0 / 5247, b64 = 49.0 / 176, b65 = -5103.0 / 18656;
    constexpr double c1 = 35.0 / 384, c3 = 500.0 / 1113, c4 = 125.0 / 192, c5 = -2187.0 / 6784, c6 = 11.0 / 84;
    constexpr double dc1 = c1 - 5179.0 / 57600, dc3 = c3 - 7571.0 / 16695, dc4 = c4 - 393.0 / 640,
                     dc5 = c5 - (-92097.0 / 339200), dc6 = c6 - 187.0 / 2100, dc7 = -1.0 / 40;
    double x[5], k1[5], xo[5], k1o[5], k3[5], k4[5], k5[5], k6[5];
    for (int i = 0; i < 5; ++i) x[i] = spin[i];
    double t = x1, told = x1;
    double dt = __builtin_fabs(PI / T.sp.gamma / Babs1);
    SpinSplines splines;
    const SpinSplines *spl = nullptr;
    if (T.sp.interpolatefields) {
        spin_splines_build<MODE>(P, T, traj, x1, x2, splines);
        spl = &splines;
    }
    const bool slog = P.st.slog_interval > 0;
    if (slog) record_spin<MODE>(P, T, traj, 0.0, x1, x);
    spin_derivs<MODE>(P, T, traj, spl, x, k1, t);
    for (;;) {
        int tries = 0;
        for (;;) {
            double tmp[5], k2[5], xn[5], kn[5];
            for (int i = 0; i < 5; ++i) tmp[i] = 1.0 * x[i] + (dt * b21) * k1[i];
            spin_derivs<MODE>(P, T, traj, spl, tmp, k2, t + dt * a2);
            for (int i = 0; i < 5; ++i) tmp[i] = 1.0 * x[i] + (dt * b31) * k1[i] + (dt * b32) * k2[i];
            spin_derivs<MODE>(P, T, traj, spl, tmp, k3, t + dt * a3);
            for (int i = 0; i < 5; ++i) tmp[i] = 1.0 * x[i] + (dt * b41) * k1[i] + (dt * b42) * k2[i] + (dt * b43) * k3[i];
            spin_derivs<MODE>(P, T, traj, spl, tmp, k4, t + dt * a4);
            for (int i = 0; i < 5; ++i)
                tmp[i] = 1.0 * x[i] + (dt * b51) * k1[i] + (dt * b52) * k2[i] + (dt * b53) * k3[i] + (dt * b54) * k4[i];
            spin_derivs<MODE>(P, T, traj, spl, tmp, k5, t + dt * a5);
            for (int i = 0; i < 5; ++i)
                tmp[i] = 1.0 * x[i] + (dt * b61) * k1[i] + (dt * b62) * k2[i] + (dt * b63) * k3[i] + (dt * b64) * k4[i] + (dt * b65) * k5[i];
            spin_derivs<MODE>(P, T, traj, spl, tmp, k6, t + dt);
            for (int i = 0; i < 5; ++i)
                xn[i] = 1.0 * x[i] + (dt * c1) * k1[i] + (dt * c3) * k3[i] + (dt * c4) * k4[i] + (dt * c5) * k5[i] + (dt * c6) * k6[i];
            spin_derivs<MODE>(P, T, traj, spl, xn, kn, t + dt);
            double err = 0;
            for (int i = 0; i < 5; ++i) {
                double xe = (dt * dc1) * k1[i] + (dt * dc3) * k3[i] + (dt * dc4) * k4[i] + (dt * dc5) * k5[i] + (dt * dc6) * k6[i] + (dt * dc7) * kn[i];
                double e = __builtin_fabs(xe) / (1e-12 + 1e-12 * (1.0 * __builtin_fabs(x[i]) + dt * __builtin_fabs(k1[i])));
                err = (err < e) ? e : err;
            }
            if (err > 1.0) {
                double f = 9.0 / 10 * ptm_pow(err, -1.0 / 3);
                double m = 1.0 / 5;
                dt = dt * ((f < m) ? m : f);
                if (tries++ >= 500) return false;
                continue;
            }
            told = t;
            t = t + dt;
            if (err < 0.5) {
                double emin = ptm_pow(5.0, -5.0);
                err = (emin < err) ? err : emin;
                dt = dt * (9.0 / 10 * ptm_pow(err, -1.0 / 5));
            }
            for (int i = 0; i < 5; ++i) { xo[i] = x[i]; k1o[i] = k1[i]; x[i] = xn[i]; k1[i] = kn[i]; }
            break;
        }
        if (t > x2) {
            /* runge_kutta_dopri5::calc_state at x2 */
            constexpr double e1 = 35.0 / 384, e3 = 500.0 / 1113, e4 = 125.0 / 192, e5 = -2187.0 / 6784, e6 = 11.0 / 84;
            const double h = (t - told);
            const double theta = (x2 - told) / h;
            const double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
            const double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
            const double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
            const double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
            const double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
            const double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
            const double tm1 = theta - 1.0, tsq = theta * theta;
            const double A = tsq * (3.0 - 2.0 * theta), B = tsq * tm1, C = tsq * tm1 * tm1, D = theta * tm1 * tm1;
            const double b1t = A * e1 - C * X1 + D, b3t = A * e3 + C * X3, b4t = A * e4 - C * X4;
            const double b5t = A * e5 + C * X5, b6t = A * e6 - C * X6, b7t = B + C * X7;
            for (int i = 0; i < 5; ++i)
                spin[i] = 1.0 * xo[i] + (h * b1t) * k1o[i] + (h * b3t) * k3[i] + (h * b4t) * k4[i] + (h * b5t) * k5[i] +
                          (h * b6t) * k6[i] + (h * b7t) * k1[i];
            if (slog) record_spin<MODE>(P, T, traj, told, x2, spin);
            return true;
        }
        for (int i = 0; i < 5; ++i) spin[i] = x[i];
        if (slog) record_spin<MODE>(P, T, traj, told, t, spin);
        if (t >= x2) return true;
    }
}

/* TTracker::IntegrateSpin (src/tracking.cpp:373-483) on the collision path: the BMT
   integration inside the spin-tracking windows, parallel transport outside */
template <int MODE>
PTD_NOINLINE double integrate_spin(const Params &P, Track &T, double *spin, const Dopri &st, double x2, const double *y2) {
    double x1 = st.t_old;
    const double *y1 = st.xo;
    double B1[3], B2[3], dB[3][3];
    field_B<MODE>(P, y1[0], y1[1], y1[2], x1, B1, dB, false, T.c.field);
    field_B<MODE>(P, y2[0], y2[1], y2[2], x2, B2, dB, false, T.c.field);
    double Babs1 = __builtin_sqrt(B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]);
    double Babs2 = __builtin_sqrt(B2[0] * B2[0] + B2[1] * B2[1] + B2[2] * B2[2]);
    double Bmax = T.sp.Bmax;
    if (T.sp.gamma == 0 || Babs2 == 0) {
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return y1[7];
    } else if (Babs1 == 0 && Babs2 > 0) {
        int sp = ptm_uniform_int01(T.key, &T.ctr) * 2 - 1;
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return sp;
    } else if ((x2 > x1) && (Babs1 < Bmax || Babs2 < Bmax)) {
        if (T.sp.n_spintimes > 0 && (spin_window(T.sp, x1) || spin_window(T.sp, x2))) {
            if (spin[0] == 0 && spin[1] == 0 && spin[2] == 0) {
                spin[0] = B1[0] * y1[7] / Babs1;
                spin[1] = B1[1] * y1[7] / Babs1;
                spin[2] = B1[2] * y1[7] / Babs1;
            }
            if (!bmt_integrate<MODE>(P, T, st, spin, x1, x2, Babs1)) T.stop = ID_ODEINT_ERROR;
        } else if (spin[0] != 0 || spin[1] != 0 || spin[2] != 0) {
            double sp = (spin[0] * B1[0] + spin[1] * B1[1] + spin[2] * B1[2]) / Babs1 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
            if (sp * sp >= 1) { spin[0] = 0; spin[1] = 0; }
            else {
                double spinaz = rnd_uniform(T, 0, 2. * PI);
                spin[0] = __builtin_sqrt(1 - sp * sp) * ptm_sin(spinaz);
                spin[1] = __builtin_sqrt(1 - sp * sp) * ptm_cos(spinaz);
            }
            spin[2] = sp;
            rotate_vector(spin, B2, nullptr);
        }
    }
    if (Babs2 > Bmax && (spin[0] != 0 || spin[1] != 0 || spin[2] != 0)) {
        double sp = (spin[0] * B2[0] + spin[1] * B2[1] + spin[2] * B2[2]) / Babs2 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return sp;
    }
    return y1[7];
}

/* ------------------------------------------------------------------------ */
/* resumable state, step path and collision path                             */
/* ------------------------------------------------------------------------ */

/* flags bits of the resumable state */
/* FL_DETACHED: the stepper's state (S.xc, S.ts) is not the tracker's (x, y); this
   happens after a decay truncation that does not stop the particle, when the
   stepper goes on from its own end of step (src/tracking.cpp:86-89) */
constexpr int FL_STARTED = 1, FL_DERIV = 2, FL_RESET = 4, FL_BVALID = 8, FL_DETACHED = 16;

/*
 * Per-lane state of the step path, in registers.  y[0..6] and y[8] double as the
 * stepper's current state xc unless FL_RESET is set (then the next step restarts
 * from y, as stepper.initialize does); pol is the stepper's own polarisation.
 * Bc/Vc: the field of the FSAL evaluation at xc.  Statistics that change rarely
 * (spin, no-flip probability, random-stream counter) stay in HBM.
 */
struct Lane {
    int sidx, stop, flags, nsolid, Nstep;
    int tries;                  /* trials of the current step so far; -1 before its first */
    long long nacc, nrej;
    unsigned long long solids;
    double x, dt, pol, tau, Hmax;
    double y[9], dxc[8];
    double Bc[3], Vc;
};

PTD void load_lane(const Params &P, int i, Lane &L) {
    const State &S = P.st;
    const size_t n = (size_t)S.n;
    L.sidx = gload(S.species + i);
    L.stop = gload(S.stop + i); L.flags = gload(S.flags + i); L.nsolid = gload(S.nsolid + i);
    L.Nstep = gload(S.Nstep + i);
    L.nacc = gload(S.n_acc + i); L.nrej = gload(S.n_rej + i);
    L.solids = gload(S.solids + i);
    L.x = gload(S.t + i); L.dt = gload(S.dt + i); L.tau = gload(S.tau + i); L.Hmax = gload(S.Hmax + i);
    L.pol = gload(S.xc + 7 * n + i);
#pragma unroll
    for (int k = 0; k < 9; ++k) L.y[k] = gload(S.y + k * n + i);
#pragma unroll
    for (int k = 0; k < 6; ++k) L.dxc[k] = gload(S.dxc + k * n + i);
    L.dxc[6] = gload(S.dxc + 6 * n + i);
    L.dxc[7] = gload(S.dxc + 8 * n + i);
    L.Bc[0] = gload(S.Bc + i); L.Bc[1] = gload(S.Bc + n + i); L.Bc[2] = gload(S.Bc + 2 * n + i);
    L.Vc = gload(S.Vc + i);
    L.tries = -1;
}

/* write the lane back (the stepper's xc is y with the stepper's polarisation) */
PTD void store_lane(const Params &P, int i, const Lane &L) {
    const State &S = P.st;
    const size_t n = (size_t)S.n;
    gstore(S.stop + i, L.stop); gstore(S.flags + i, L.flags); gstore(S.nsolid + i, L.nsolid);
    gstore(S.Nstep + i, L.Nstep);
    gstore(S.n_acc + i, L.nacc); gstore(S.n_rej + i, L.nrej);
    gstore(S.solids + i, L.solids);
    gstore(S.t + i, L.x); gstore(S.dt + i, L.dt); gstore(S.tau + i, L.tau); gstore(S.Hmax + i, L.Hmax);
#pragma unroll
    for (int k = 0; k < 9; ++k) gstore(S.y + k * n + i, L.y[k]);
    if (!(L.flags & FL_DETACHED)) {
        gstore(S.ts + i, L.x);
#pragma unroll
        for (int k = 0; k < 9; ++k) gstore(S.xc + k * n + i, (k == 7) ? L.pol : L.y[k]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) gstore(S.dxc + k * n + i, L.dxc[k]);
    gstore(S.dxc + 6 * n + i, L.dxc[6]);
    gstore(S.dxc + 7 * n + i, 0.0);
    gstore(S.dxc + 8 * n + i, L.dxc[7]);
    gstore(S.Bc + i, L.Bc[0]); gstore(S.Bc + n + i, L.Bc[1]); gstore(S.Bc + 2 * n + i, L.Bc[2]);
    gstore(S.Vc + i, L.Vc);
}

/* OnStep of the step path (src/neutron.cpp:267-282, src/proton.cpp:25-33,
   src/electron.cpp:25-33); random draws come from the particle's stream in HBM */
PTD void on_step_A(const Params &P, const Species &sp, int i, int &stop, double x1, const double *y1, double &x2,
                   double *y2, const DenseRef &dref, int cursolid) {
    if (sp.kind == 0) {
        const double *mat = solid_mat(P.geom, cursolid);
        if (mat[1] > 0) {
            double mn = sp.m, e = ELE_E, hb = HBAR;
            double Er = 0.5 * mn * (y1[3] * y1[3] + y1[4] * y1[4] + y1[5] * y1[5]);
            double Ei = mat[1] * 1e-9;
            double kr, ki;
            ptm_csqrt((2 * mn) * Er, (2 * mn) * Ei, &kr, &ki);
            ki = ki * e;
            ki = ki / hb;
            double l = __builtin_sqrt((y2[0] - y1[0]) * (y2[0] - y1[0]) + (y2[1] - y1[1]) * (y2[1] - y1[1]) + (y2[2] - y1[2]) * (y2[2] - y1[2]));
            uint64_t key = gload(P.st.key + i), ctr = gload(P.st.ctr + i);
            double abspath = -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(key, &ctr))) / (2 * ki);
            P.st.ctr[i] = ctr;
            if (abspath < l) {
                x2 = x1 + abspath / l * (x2 - x1);
                dense_calc(dref, x2, y2);
                stop = ID_ABSORBED_IN_MATERIAL;
            }
        }
    } else {
        if (solid_ID(P.geom, cursolid) > 1) {
            x2 = x1;
            for (int k = 0; k < 9; ++k) y2[k] = y1[k];
            stop = ID_ABSORBED_IN_MATERIAL;
        }
    }
}

/* TParticle::DoStep on the step path (src/particle.cpp:196-206); at_end: (x2, y2) is the
   unmodified end of the step, where the last equation-of-motion evaluation left the field */
template <int MODE>
PTD bool do_step_hook_A(const Params &P, const Species &sp, Lane &L, int i, const DenseRef &dref, double x1,
                        const double *y1, double &x2, double *y2, int cursolid, bool at_end, Ctr &c) {
    double x2temp = x2, y2temp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) y2temp[k] = y2[k];
    on_step_A(P, sp, i, L.stop, x1, y1, x2, y2, dref, cursolid);
    bool changed = !(x2temp == x2 && state_eq(y2temp, y2));
    bool cached = at_end && !changed && (L.flags & FL_BVALID);
    double H = kinetic_energy(sp, y2 + 3) +
               potential_energy<MODE>(P, sp, y2, x2, cached ? L.Bc : nullptr, cached ? &L.Vc : nullptr, c);
    L.Hmax = (H < L.Hmax) ? L.Hmax : H;
    if (y2temp[7] != y2[7]) gstore(P.st.Nspinflip + i, gload(P.st.Nspinflip + i) + (1));
    L.Nstep++;
    return changed;
}

/* TTracker::IntegrateSpin without the BMT branch (src/tracking.cpp:373-394, 473-483) and
   TParticle::DoPolarize (src/particle.cpp:225-234) on the step path.  B1/B2: the field at the
   step start / end when the stepper evaluated it there (else nullptr).  Returns the new
   tracker polarisation in y[7], or false, with nothing changed, when the step needs the BMT
   integration (a spin-tracking window): the collision path then does the spin step. */
template <int MODE>
PTD bool spin_step(const Params &P, int i, const Species &sp, double x1, double ys0, double ys1, double ys2,
                            double polo, double x2, double *y, bool haveB1, double B10, double B11, double B12,
                            bool haveB2, double B20, double B21, double B22, Ctr &c) {
    const State &S = P.st;
    const size_t n = (size_t)S.n;
    double B1[3], B2[3], dB[3][3];
    if (haveB1) { B1[0] = B10; B1[1] = B11; B1[2] = B12; }
    else field_B<MODE>(P, ys0, ys1, ys2, x1, B1, dB, false, c.field);
    if (haveB2) { B2[0] = B20; B2[1] = B21; B2[2] = B22; }
    else field_B<MODE>(P, y[0], y[1], y[2], x2, B2, dB, false, c.field);
    double Babs1 = __builtin_sqrt(B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]);
    double Babs2 = __builtin_sqrt(B2[0] * B2[0] + B2[1] * B2[1] + B2[2] * B2[2]);
    double spin[5];
    for (int k = 0; k < 5; ++k) spin[k] = gload(S.spin + k * n + i);
    const double spin_in[3] = {spin[0], spin[1], spin[2]};
    uint64_t key = gload(S.key + i), ctr = gload(S.ctr + i);
    double spinProjection;
    bool done = false;
    if (sp.gamma == 0 || Babs2 == 0) {
        for (int k = 0; k < 5; ++k) spin[k] = 0;
        spinProjection = polo;
        done = true;
    } else if (Babs1 == 0 && Babs2 > 0) {
        int pr = ptm_uniform_int01(key, &ctr) * 2 - 1;
        for (int k = 0; k < 5; ++k) spin[k] = 0;
        spinProjection = pr;
        done = true;
    } else if ((x2 > x1) && (Babs1 < sp.Bmax || Babs2 < sp.Bmax)) {
        if (sp.n_spintimes > 0 && (spin_window(sp, x1) || spin_window(sp, x2))) return false;
        if (spin[0] != 0 || spin[1] != 0 || spin[2] != 0) {
            double pr = (spin[0] * B1[0] + spin[1] * B1[1] + spin[2] * B1[2]) / Babs1 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
            if (pr * pr >= 1) { spin[0] = 0; spin[1] = 0; }
            else {
                double spinaz = ptm_canonical(ptm_draw_u64(key, &ctr)) * (2. * PI - 0) + 0;
                spin[0] = __builtin_sqrt(1 - pr * pr) * ptm_sin(spinaz);
                spin[1] = __builtin_sqrt(1 - pr * pr) * ptm_cos(spinaz);
            }
            spin[2] = pr;
            rotate_vector(spin, B2, nullptr);
        }
    }
    if (!done) {
        if (Babs2 > sp.Bmax && (spin[0] != 0 || spin[1] != 0 || spin[2] != 0)) {
            double pr = (spin[0] * B2[0] + spin[1] * B2[1] + spin[2] * B2[2]) / Babs2 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
            for (int k = 0; k < 5; ++k) spin[k] = 0;
            spinProjection = pr;
        } else {
            spinProjection = polo;
        }
    }
    /* src/tracking.cpp:128-138 */
    bool spinHasFlipped = polo != y[7];
    if (spinProjection != polo) {
        double flipprob = 0.5 * (1 - y[7] * spinProjection);
        gstore(S.noflip + i, gload(S.noflip + i) * (1. - flipprob));
        if (sp.flipspin) {
            double prevpol = y[7];
            y[7] = (ptm_canonical(ptm_draw_u64(key, &ctr)) < 0.5 * (spinProjection + 1)) ? 1 : -1;
            if (y[7] != prevpol) gstore(S.Nspinflip + i, gload(S.Nspinflip + i) + (1));
            if (spinHasFlipped) y[7] *= -1;
        }
    } else if (spinHasFlipped) {
        for (int k = 0; k < 5; ++k) spin[k] = 0;
    }
    S.ctr[i] = ctr;
    for (int k = 0; k < 5; ++k) gstore(S.spin + k * n + i, spin[k]);
    (void)spin_in;
    return true;
}

/* The dopri5 stage vectors k2..k5 of the step kernels' trial steps live in LDS, one slice per
   lane: stage_lds[stage][component][threadIdx.x] (k2 = stage 0 ... k5 = stage 3; 64 KB per
   block of at most 256 threads, pt_batch_run's limit).  Keeping them out of the registers
   leaves the register file to the field evaluations between the stages; each lane touches
   only its own slice, so no barrier is needed. */
constexpr int STAGE_LDS = 4, STAGE_LANES = 256;
__shared__ double stage_lds[STAGE_LDS][8][STAGE_LANES];
static_assert(sizeof(stage_lds) + sizeof(coop_lds) == PT_STEP_LDS_BYTES,
              "the host checks PT_STEP_LDS_BYTES against the device");
static_assert(COOP_WAVES * 64 == STAGE_LANES, "one LDS staging slice per wave of a step-kernel block");

/* An accepted dopri5 trial step, handed from trialA to acceptA: the stepper's state at
   the step start (xc, ts), the new state and its derivative and field, and whether the
   rest of the step can need the dense output (then the stage vectors are in HBM). */
struct Trial {
    double xc[8], ts, xnew[8], k7[8], Bn[3], Vn, polnew, dt_used;
    bool bn, need_dense;
};

/*
 * One trial step of odeint's controlled dopri5 (dense_output_runge_kutta::do_step,
 * src/tracking.cpp:71-93) on the step path: the stage vectors, the error norm and the
 * step-size controller.  Returns 0 when the trial is rejected (L.dt shrunk), 1 when it is
 * accepted (T filled, L.dt grown), 2 when the stepper fails (L.stop set).  The step
 * kernels call it once per loop iteration for every lane and run acceptA only on the
 * lanes whose trial was accepted, so a wave does not repeat trials until its last lane
 * accepts: each lane's arithmetic is unchanged, only the lanes' schedule differs.
 * L.tries < 0 marks the first trial of a step.
 */
template <int MODE>
PTD int trialA(const Params &P, int i, Lane &L, Trial &T, Ctr &c) {
    const Geom &G = P.geom;
    const Species &sp = P.species[L.sidx];
    const double tmax = P.tmax_global;
    const size_t n = (size_t)P.st.n;
    double *ds = P.st.dense;
    if (L.tries < 0) {
        if (L.flags & FL_RESET) {
            /* stepper.initialize(y, x, current_time_step()) (src/tracking.cpp:72-74) */
            L.pol = L.y[7];
            L.flags &= ~(FL_DERIV | FL_DETACHED);
        }
        L.tries = 0;
    }
    /* the stepper starts from its own state: the tracker's unless detached */
    double *xc = T.xc;
    double ts;
    if (L.flags & FL_DETACHED) {
#pragma unroll
        for (int k = 0; k < 8; ++k) xc[k] = P.st.xc[(k < 7 ? k : 8) * n + i];
        ts = P.st.ts[i];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) xc[k] = L.y[k < 7 ? k : 8];
        ts = L.x;
    }
    T.ts = ts;
    if (!(L.flags & FL_DERIV)) {
        bool hb;
        derivs8<MODE>(P, sp, xc, ts, L.pol, L.dxc, L.Bc, L.Vc, hb, c);
        L.flags = (L.flags | FL_DERIV);
        L.flags = hb ? (L.flags | FL_BVALID) : (L.flags & ~FL_BVALID);
    }
    const int cursolid = current_solid(G, L.solids, L.nsolid);
    const bool absorb = sp.kind == 0 && solid_mat(G, cursolid)[1] > 0;

    /* ---- dense_output_runge_kutta::do_step (controlled dopri5, abs = rel = 1e-9) ---- */
    constexpr double b21 = 1.0 / 5;
    constexpr double b31 = 3.0 / 40, b32 = 9.0 / 40;
    constexpr double b41 = 44.0 / 45, b42 = -56.0 / 15, b43 = 32.0 / 9;
    constexpr double b51 = 19372.0 / 6561, b52 = -25360.0 / 2187, b53 = 64448.0 / 6561, b54 = -212.0 / 729;
    constexpr double b61 = 9017.0 / 3168, b62 = -355.0 / 33, b63 = 46732.0 / 5247, b64 = 49.0 / 176, b65 = -5103.0 / 18656;
    constexpr double c1 = 35.0 / 384, c3 = 500.0 / 1113, c4 = 125.0 / 192, c5 = -2187.0 / 6784, c6 = 11.0 / 84;
    constexpr double dc1 = c1 - 5179.0 / 57600, dc3 = c3 - 7571.0 / 16695, dc4 = c4 - 393.0 / 640,
                     dc5 = c5 - (-92097.0 / 339200), dc6 = c6 - 187.0 / 2100, dc7 = -1.0 / 40;
    double *xnew = T.xnew, *k7 = T.k7;
    /* this lane's slices of the stage vectors k2..k5 */
    const int lane = threadIdx.x;
#define K2(k) stage_lds[0][k][lane]
#define K3(k) stage_lds[1][k][lane]
#define K4(k) stage_lds[2][k][lane]
#define K5(k) stage_lds[3][k][lane]
    /* a compiler-only barrier after each stage's stores: later uses read LDS, so the
       compiler cannot keep the stage vectors in registers by forwarding its own stores */
#define STAGE_STORED() asm volatile("" ::: "memory")
    {
        const double dt = L.dt;
        double tmp[8], kst[8], k6[8], Bd[3], Vd;
        bool bd;
        const double h21 = dt * b21;
#pragma unroll
        for (int k = 0; k < 8; ++k) tmp[k] = xc[k] + h21 * L.dxc[k];
        derivs8<MODE, true>(P, sp, tmp, ts + dt * (1.0 / 5), L.pol, kst, Bd, Vd, bd, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) K2(k) = kst[k];
        STAGE_STORED();
        const double h31 = dt * b31, h32 = dt * b32;
#pragma unroll
        for (int k = 0; k < 8; ++k) tmp[k] = xc[k] + h31 * L.dxc[k] + h32 * K2(k);
        derivs8<MODE, true>(P, sp, tmp, ts + dt * (3.0 / 10), L.pol, kst, Bd, Vd, bd, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) K3(k) = kst[k];
        STAGE_STORED();
        const double h41 = dt * b41, h42 = dt * b42, h43 = dt * b43;
#pragma unroll
        for (int k = 0; k < 8; ++k) tmp[k] = xc[k] + h41 * L.dxc[k] + h42 * K2(k) + h43 * K3(k);
        derivs8<MODE, true>(P, sp, tmp, ts + dt * (4.0 / 5), L.pol, kst, Bd, Vd, bd, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) K4(k) = kst[k];
        STAGE_STORED();
        const double h51 = dt * b51, h52 = dt * b52, h53 = dt * b53, h54 = dt * b54;
#pragma unroll
        for (int k = 0; k < 8; ++k) tmp[k] = xc[k] + h51 * L.dxc[k] + h52 * K2(k) + h53 * K3(k) + h54 * K4(k);
        derivs8<MODE, true>(P, sp, tmp, ts + dt * (8.0 / 9), L.pol, kst, Bd, Vd, bd, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) K5(k) = kst[k];
        STAGE_STORED();
        const double h61 = dt * b61, h62 = dt * b62, h63 = dt * b63, h64 = dt * b64, h65 = dt * b65;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            tmp[k] = xc[k] + h61 * L.dxc[k] + h62 * K2(k) + h63 * K3(k) + h64 * K4(k) + h65 * K5(k);
        derivs8<MODE, true>(P, sp, tmp, ts + dt, L.pol, k6, Bd, Vd, bd, c);
        const double g1 = dt * c1, g3 = dt * c3, g4 = dt * c4, g5 = dt * c5, g6 = dt * c6;
#pragma unroll
        for (int k = 0; k < 8; ++k) xnew[k] = xc[k] + g1 * L.dxc[k] + g3 * K3(k) + g4 * K4(k) + g5 * K5(k) + g6 * k6[k];
        T.polnew = L.pol + g1 * 0.0 + g3 * 0.0 + g4 * 0.0 + g5 * 0.0 + g6 * 0.0;
        const double e1 = dt * dc1, e3 = dt * dc3, e4 = dt * dc4, e5 = dt * dc5, e6 = dt * dc6, e7 = dt * dc7;
        double errp[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) errp[k] = e1 * L.dxc[k] + e3 * K3(k) + e4 * K4(k) + e5 * K5(k) + e6 * k6[k];
        /* can the rest of this step need the dense output? */
        {
            double l2 = (xnew[7] - xc[7]) * (xnew[7] - xc[7]);
            double d2 = (xnew[0] - xc[0]) * (xnew[0] - xc[0]) + (xnew[1] - xc[1]) * (xnew[1] - xc[1]) + (xnew[2] - xc[2]) * (xnew[2] - xc[2]);
            bool split = 0.25 * (l2 - d2) > MAX_TRACK_DEVIATION * MAX_TRACK_DEVIATION;
            bool trunc = (L.tau > 0 && xnew[6] > L.tau) || (ts + dt > tmax);
            /* detached: the chord starts at the tracker's point, not at xc */
            T.need_dense = trunc || split || absorb || (L.flags & FL_DETACHED) || !clear_of_walls(G, xc, xnew) ||
                         (P.st.snap_n > 0 && snapshot_in(P.st, ts, ts + dt) >= 0) ||
                         (sp.n_spintimes > 0 && (spin_window(sp, ts) || spin_window(sp, ts + dt)));
        }
        if (T.need_dense) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                gstore(ds + k * n + i, xc[k]);
                gstore(ds + (8 + k) * n + i, L.dxc[k]);
                gstore(ds + (16 + k) * n + i, K3(k));
                gstore(ds + (24 + k) * n + i, K4(k));
                gstore(ds + (32 + k) * n + i, K5(k));
                gstore(ds + (40 + k) * n + i, k6[k]);
            }
        }
        derivs8<MODE, true>(P, sp, xnew, ts + dt, T.polnew, k7, T.Bn, T.Vn, T.bn, c);
        double err = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double xe = errp[k] + e7 * k7[k];
            double e = dabs(xe) / (1e-9 + 1e-9 * (dabs(xc[k]) + dt * dabs(L.dxc[k])));
            err = (err < e) ? e : err;
        }
        bool ok = !(err > 1.0);
        if (!ok) {
            double f = 9.0 / 10 * ptm_pow(err, -1.0 / 3);
            double m = 1.0 / 5;
            L.dt = dt * ((f < m) ? m : f);
            ++L.nrej;
        } else {
            T.dt_used = dt;
            if (err < 0.5) {
                double emin = ptm_pow(5.0, -5.0);
                err = (emin < err) ? err : emin;
                L.dt = dt * (9.0 / 10 * ptm_pow(err, -1.0 / 5));
            }
            ++L.nacc;
        }
        if (L.tries++ >= 500) {
            /* the failed-step checker throws; IntegrateParticle keeps x, y and stops (src/tracking.cpp:78-85) */
            L.stop = ID_ODEINT_ERROR;
            return 2;
        }
        if (!ok) return 0;
    }
#undef K2
#undef K3
#undef K4
#undef K5
#undef STAGE_STORED
    {
        /* an accepted step with a non-finite state (a stage beyond the speed of light makes
           1/gamma a NaN, which the maximum of the error norm passes over) stops as a stepper
           failure, x and y kept at the step start (src/tracking.cpp:78-85) */
        bool finite = T.polnew == T.polnew;
#pragma unroll
        for (int k = 0; k < 8; ++k) finite = finite && (dabs(xnew[k]) <= DBL_MAX_V);
        if (!finite) {
            L.stop = ID_ODEINT_ERROR;
            return 2;
        }
    }
    L.tries = -1;
    return 1;
}

/*
 * The rest of one step of TTracker::IntegrateParticle (src/tracking.cpp:94-148) on the
 * step path, after trialA accepted a trial: decay/tmax truncation through the dense
 * output (written to HBM only when the step can need it: truncation, chord splitting,
 * bulk absorption, or a chord that comes near a wall), chord splitting and the
 * collision test of each piece, spin step, logs and stop checks.  Returns pending =
 * true, with the chord piece in HBM, when a chord piece intersects the mesh; the
 * collision path then finishes the step.
 */
template <int MODE>
PTD void acceptA(const Params &P, int i, Lane &L, bool &pending, Ctr &c, const Trial &T) {
    const Geom &G = P.geom;
    const Species &sp = P.species[L.sidx];
    const double tmax = P.tmax_global;
    const size_t n = (size_t)P.st.n;
    double *ds = P.st.dense;
    const int cursolid = current_solid(G, L.solids, L.nsolid);
    const double x1 = L.x;
    const double *xc = T.xc, *xnew = T.xnew, *k7 = T.k7;
    const double ts = T.ts, polnew = T.polnew, dt_used = T.dt_used;
    const bool need_dense = T.need_dense;
    const double polo = L.pol;
    const double told = ts;
    const double tnew = ts + dt_used;
    if (need_dense) {
#pragma unroll
        for (int k = 0; k < 8; ++k) gstore(ds + (48 + k) * n + i, k7[k]);
        gstore(ds + 56 * n + i, tnew);
        gstore(ds + 57 * n + i, told);
        gstore(ds + 58 * n + i, polo);
    }
    /* accepted: the FSAL derivative and field move to the new state */
    const bool b1v = (L.flags & FL_BVALID) != 0;
    const double B1c[3] = {L.Bc[0], L.Bc[1], L.Bc[2]};
#pragma unroll
    for (int k = 0; k < 8; ++k) L.dxc[k] = k7[k];
    L.pol = polnew;
    L.Bc[0] = T.Bn[0]; L.Bc[1] = T.Bn[1]; L.Bc[2] = T.Bn[2]; L.Vc = T.Vn;
    L.flags = T.bn ? (L.flags | FL_BVALID) : (L.flags & ~FL_BVALID);

    double x = tnew, y[9], y1[9];
    to9(xnew, polnew, y);
#pragma unroll
    for (int k = 0; k < 9; ++k) y1[k] = L.y[k];
    const DenseRef dref = {&P, i};
    bool at_end = true;
    if (L.tau > 0 && y[6] > L.tau) {
        x = x1 + (x - x1) * (L.tau - y1[6]) / (y[6] - y1[6]);
        dense_calc(dref, x, y);
        at_end = false;
    }
    if (x > tmax) {
        x = tmax;
        dense_calc(dref, x, y);
        at_end = false;
    }
    if (!at_end) {
        /* the tracker's end of step is interpolated; the stepper keeps its own */
#pragma unroll
        for (int k = 0; k < 9; ++k) gstore(P.st.xc + k * n + i, (k == 7) ? polnew : xnew[k < 7 ? k : 7]);
        gstore(P.st.ts + i, tnew);
        L.flags |= FL_DETACHED;
    } else {
        L.flags &= ~FL_DETACHED;
    }
    bool reset = (L.flags & FL_RESET) != 0;
    bool altered = false;
    double xa = x1;
    while (xa < x) {
        double l2 = (y[8] - y1[8]) * (y[8] - y1[8]);
        double d2 = (y[0] - y1[0]) * (y[0] - y1[0]) + (y[1] - y1[1]) * (y[1] - y1[1]) + (y[2] - y1[2]) * (y[2] - y1[2]);
        double dev2 = 0.25 * (l2 - d2);
        double x2 = x, y2[9];
        bool piece_end = at_end;
#pragma unroll
        for (int k = 0; k < 9; ++k) y2[k] = y[k];
        if (dev2 > MAX_TRACK_DEVIATION * MAX_TRACK_DEVIATION) {
            x2 = xa + (x - xa) / __builtin_ceil(__builtin_sqrt(dev2) / MAX_TRACK_DEVIATION);
            dense_calc(dref, x2, y2);
            piece_end = false;
        }
        bool r;
        /* TTracker::CheckHit (src/tracking.cpp:176-226) up to the collision query */
        if (!check_segment(G, y1, y2)) {
            L.stop = ID_HIT_BOUNDARIES;
            r = true;
        } else if (x2 == xa) {
            r = false;
        } else {
            int nh = count_hits(P, c, y1, y2);
            if (nh < 0) {
                L.stop = nh == -1 ? ID_CGAL_ERROR : ID_UNSUPPORTED;
                r = true;
            } else if (nh > 0) {
                /* defer to the collision path: chord piece, step end and stepper state to HBM */
                double *pd = P.st.pend;
                gstore(pd + i, xa);
                gstore(pd + 10 * n + i, x2);
#pragma unroll
                for (int k = 0; k < 9; ++k) { gstore(pd + (1 + k) * n + i, y1[k]); gstore(pd + (11 + k) * n + i, y2[k]); }
                L.x = x;
#pragma unroll
                for (int k = 0; k < 9; ++k) L.y[k] = y[k];
                L.flags = reset ? (L.flags | FL_RESET) : (L.flags & ~FL_RESET);
                pending = true;
                return;
            } else {
                r = do_step_hook_A<MODE>(P, sp, L, i, dref, xa, y1, x2, y2, cursolid, piece_end, c);
            }
        }
        reset = r;
        if (r) {
            altered = true;
            x = x2;
#pragma unroll
            for (int k = 0; k < 9; ++k) y[k] = y2[k];
        }
        xa = x2;
#pragma unroll
        for (int k = 0; k < 9; ++k) y1[k] = y2[k];
    }
    L.flags = reset ? (L.flags | FL_RESET) : (L.flags & ~FL_RESET);
    if (P.st.snap_n > 0)
        record_snapshot(P.st, i, dref, told, x, nullptr, L.stop, P.st.Nhit[i], L.Nstep, P.st.Nspinflip[i], L.Hmax,
                        P.st.noflip[i]);
    const bool y_is_xc = at_end && !altered && (L.flags & FL_BVALID);
    if (!spin_step<MODE>(P, i, sp, told, xc[0], xc[1], xc[2], polo, x, y, b1v, B1c[0], B1c[1], B1c[2],
                         y_is_xc, L.Bc[0], L.Bc[1], L.Bc[2], c)) {
        /* a spin-tracking window: the collision path finishes the step from its spin step
           (a chord piece of zero length at the step's end) */
        double *pd = P.st.pend;
        gstore(pd + i, x);
        gstore(pd + 10 * n + i, x);
#pragma unroll
        for (int k = 0; k < 9; ++k) { gstore(pd + (1 + k) * n + i, y[k]); gstore(pd + (11 + k) * n + i, y[k]); }
        L.x = x;
#pragma unroll
        for (int k = 0; k < 9; ++k) L.y[k] = y[k];
        pending = true;
        return;
    }
    if (P.st.trk_interval > 0)
        record_track(P.st, i, xc[7], x, y, solid_ID(G, current_solid(G, L.solids, L.nsolid)));
    if (L.stop == ID_UNKNOWN && y[6] >= L.tau) L.stop = ID_DECAYED;
    else if (L.stop == ID_UNKNOWN && (x >= tmax || y[8] >= sp.lmax)) L.stop = ID_NOT_FINISH;
    L.x = x;
#pragma unroll
    for (int k = 0; k < 9; ++k) L.y[k] = y[k];
}

/* stop checks at the end of a step (src/tracking.cpp:144-147) */
PTD void finish_checks(Track &T, double x, const double *y, double tau, double tmax) {
    if (T.stop == ID_UNKNOWN && y[6] >= tau) T.stop = ID_DECAYED;
    else if (T.stop == ID_UNKNOWN && (x >= tmax || y[8] >= T.sp.lmax)) T.stop = ID_NOT_FINISH;
}

/*
 * Collision path: finishes the step of particle i from the state in HBM, starting
 * with the chord piece whose collision the step path deferred.  This is the rest of
 * the loop body of TTracker::IntegrateParticle with CheckHit's recursion
 * (src/tracking.cpp:96-147).
 */
template <int MODE>
PTD void collision_path(const Params &P, int i, Ctr &c) {
    const State &S = P.st;
    const int n = S.n;
    Track T;
    T.i = i;
    T.cand.n = -1;
    T.sp = P.species[S.species[i]];
    T.key = S.key[i]; T.ctr = S.ctr[i];
    T.stop = S.stop[i];
    T.Nhit = S.Nhit[i]; T.Nstep = S.Nstep[i]; T.Nspinflip = S.Nspinflip[i];
    T.nhitlog = S.nhitlog[i];
    T.nacc = S.n_acc[i]; T.nrej = S.n_rej[i];
    T.Hmax = S.Hmax[i]; T.noflip = S.noflip[i];
    T.solids = S.solids[i]; T.nsolid = S.nsolid[i];
    T.c = c;
    int flags = S.flags[i];
    double tau = S.tau[i];
    const double tmax = P.tmax_global;
    Dopri st;
    const double *ds = S.dense;
    double xo8[8], dxo8[8], k38[8], k48[8], k58[8], k68[8], dxc8[8];
    for (int k = 0; k < 8; ++k) {
        xo8[k] = ds[(size_t)k * n + i];
        dxo8[k] = ds[(size_t)(8 + k) * n + i];
        k38[k] = ds[(size_t)(16 + k) * n + i];
        k48[k] = ds[(size_t)(24 + k) * n + i];
        k58[k] = ds[(size_t)(32 + k) * n + i];
        k68[k] = ds[(size_t)(40 + k) * n + i];
        dxc8[k] = ds[(size_t)(48 + k) * n + i];
    }
    st.t = ds[(size_t)56 * n + i];
    st.t_old = ds[(size_t)57 * n + i];
    double polo = ds[(size_t)58 * n + i];
    to9(xo8, polo, st.xo);
    to9(dxo8, 0.0, st.dxo);
    to9(k38, 0.0, st.k3);
    to9(k48, 0.0, st.k4);
    to9(k58, 0.0, st.k5);
    to9(k68, 0.0, st.k6);
    to9(dxc8, 0.0, st.dxc);
    double x = S.t[i], y[9], spin[5], x1, y1[9], x2, y2[9];
    for (int k = 0; k < 9; ++k) y[k] = S.y[(size_t)k * n + i];
    for (int k = 0; k < 5; ++k) spin[k] = S.spin[(size_t)k * n + i];
    const double *pd = S.pend;
    x1 = pd[i];
    x2 = pd[(size_t)10 * n + i];
    for (int k = 0; k < 9; ++k) { y1[k] = pd[(size_t)(1 + k) * n + i]; y2[k] = pd[(size_t)(11 + k) * n + i]; }

    /* a zero-length piece: the step path handed over only the spin step */
    const bool spin_only = !(x2 > x1);
    bool reset = spin_only ? (flags & FL_RESET) != 0 : check_hit<MODE>(P, T, x1, y1, x2, y2, st);
    if (reset && !spin_only) {
        x = x2;
        for (int k = 0; k < 9; ++k) y[k] = y2[k];
    }
    x1 = x2;
    for (int k = 0; k < 9; ++k) y1[k] = y2[k];
    while (x1 < x) {
        double l2 = (y[8] - y1[8]) * (y[8] - y1[8]);
        double d2 = (y[0] - y1[0]) * (y[0] - y1[0]) + (y[1] - y1[1]) * (y[1] - y1[1]) + (y[2] - y1[2]) * (y[2] - y1[2]);
        double dev2 = 0.25 * (l2 - d2);
        x2 = x;
        for (int k = 0; k < 9; ++k) y2[k] = y[k];
        if (dev2 > MAX_TRACK_DEVIATION * MAX_TRACK_DEVIATION) {
            x2 = x1 + (x - x1) / __builtin_ceil(__builtin_sqrt(dev2) / MAX_TRACK_DEVIATION);
            dopri_calc_state(st, x2, y2);
        }
        reset = check_hit<MODE>(P, T, x1, y1, x2, y2, st);
        if (reset) {
            x = x2;
            for (int k = 0; k < 9; ++k) y[k] = y2[k];
        }
        x1 = x2;
        for (int k = 0; k < 9; ++k) y1[k] = y2[k];
    }
    flags = reset ? (flags | FL_RESET) : (flags & ~FL_RESET);
    if (S.snap_n > 0 && !spin_only)
        record_snapshot(S, i, st, st.t_old, x, spin, T.stop, T.Nhit, T.Nstep, T.Nspinflip, T.Hmax, T.noflip);
    double spinProjection = integrate_spin<MODE>(P, T, spin, st, x, y);
    bool spinHasFlipped = st.xo[7] != y[7];
    if (spinProjection != st.xo[7]) {
        do_polarize(T, y, spinProjection, T.sp.flipspin);
        if (T.sp.flipspin && spinHasFlipped) y[7] *= -1;
    } else if (spinHasFlipped) {
        for (int k = 0; k < 5; ++k) spin[k] = 0;
    }
    if (S.trk_interval > 0)
        record_track(S, i, st.xo[8], x, y, solid_ID(P.geom, current_solid(P.geom, T.solids, T.nsolid)));
    finish_checks(T, x, y, tau, tmax);

    S.t[i] = x;
    for (int k = 0; k < 9; ++k) S.y[(size_t)k * n + i] = y[k];
    for (int k = 0; k < 5; ++k) S.spin[(size_t)k * n + i] = spin[k];
    S.flags[i] = flags;
    S.ctr[i] = T.ctr;
    S.stop[i] = T.stop;
    S.Nhit[i] = T.Nhit; S.Nstep[i] = T.Nstep; S.Nspinflip[i] = T.Nspinflip;
    S.nhitlog[i] = T.nhitlog;
    S.n_acc[i] = T.nacc; S.n_rej[i] = T.nrej;
    S.Hmax[i] = T.Hmax; S.noflip[i] = T.noflip;
    S.solids[i] = T.solids; S.nsolid[i] = T.nsolid;
    c = T.c;
}

/*
 * Particle set-up: the TParticle constructor's Hmax (src/particle.cpp:89) and the start
 * of TTracker::IntegrateParticle (src/tracking.cpp:19-68): lifetime draw, initial
 * step size, and the solid stack by +z ray parity.
 */
template <int MODE>
__global__ void __launch_bounds__(256) init_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n || (S.flags[i] & FL_STARTED)) return;
    const int n = S.n;
    Species sp = P.species[S.species[i]];
    uint64_t key = S.key[i], ctr = S.ctr[i];
    double x = S.t[i], y[9];
    for (int k = 0; k < 9; ++k) y[k] = S.y[(size_t)k * n + i];
    Ctr c = {0, 0, 0};
    double Hmax = kinetic_energy(sp, y + 3) + potential_energy<MODE>(P, sp, y, x, nullptr, nullptr, c);
    double tau;
    if (sp.tau > 0) tau = -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(key, &ctr))) / (1. / sp.tau);
    else tau = sp.tmax;
    int stop = ID_UNKNOWN;
    unsigned long long sk = sk_set(0ull, 0, -1, 0);
    int ns = 1;
    if (x > P.tmax_global) {
        stop = ID_INVARIANT;
    } else {
        for (int k = 0; k < 9; ++k) S.xc[(size_t)k * n + i] = y[k];
        S.ts[i] = x;
        S.dt[i] = 10. * MAX_TRACK_DEVIATION / __builtin_sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        int ovf;
        unsigned long long par = ray_parity(P.geom, y[0], y[1], y[2], &ovf);
        if (ovf) stop = ID_INVARIANT;
        for (int m = 0; m < P.geom.n_solid && !ovf; ++m) {
            if ((par >> m) & 1ull) {
                if (ns >= MAX_SOLIDS) { stop = ID_UNSUPPORTED; break; }
                sk = sk_set(sk, ns, m, solid_ignored(P.geom, m, x));
                ++ns;
            }
        }
    }
    S.Hmax[i] = Hmax;
    S.tau[i] = tau;
    S.ctr[i] = ctr;
    S.stop[i] = stop;
    S.solids[i] = sk;
    S.nsolid[i] = ns;
    S.flags[i] = FL_STARTED;
}

/*
 * Step kernel: one thread per particle advances its particle by accepted steps
 * until its step budget is spent, it stops, or one of its chord pieces meets the
 * mesh.  In the last case the particle parks (piece, step end and dense-output
 * state in HBM) and its index goes to the collision queue.
 */
#ifndef PT_STEP_MIN_WAVES
#define PT_STEP_MIN_WAVES 2
#endif
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PT_STEP_MIN_WAVES))) step_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n) return;
    if (S.stop[i] != ID_UNKNOWN || *P.quit != 0) return;
    int budget = S.budget[i];
    if (budget <= 0) return;
    Lane L;
    load_lane(P, i, L);
    Ctr c = {0, 0, 0};
    bool pending = false;
    do {
        Trial T;
        const int r = trialA<MODE>(P, i, L, T, c);
        if (r == 1) acceptA<MODE>(P, i, L, pending, c, T);
        if (r != 0) --budget;
    } while (!pending && L.stop == ID_UNKNOWN && budget > 0);
    store_lane(P, i, L);
    S.budget[i] = budget;
    S.n_field[i] += c.field;
    if (pending) {
        int slot = gatomic_add(S.qcount, 1);
        S.queue[slot] = i;
    }
}

/*
 * Persistent step kernel: a grid sized to the machine, each lane a slot that takes the
 * next particle from the work list (the particles with steps to do, compacted by
 * work_kernel) whenever its particle parks, stops or spends its step budget, so waves
 * stay full while particles of very different lifetimes drop out.  Every lane leaves
 * when the list is exhausted.  With split set the list (in Morton order of the particle
 * positions) is taken as NXCD chunks: a block takes from the chunk of its XCD first, so
 * the particles an XCD steps at one time are neighbours and share field-map cells in its
 * L2; when its chunk is done it helps with the next ones.
 */
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PT_STEP_MIN_WAVES))) step_kernel_persistent(const Params *__restrict__ Pp, int split) {
    const Params &P = *Pp;
    const State &S = P.st;
    const int count = S.wcount[0];
    const bool quit = *P.quit != 0;
    Lane L;
    Ctr c = {0, 0, 0};
    int i = -1, budget = 0;
    bool have = false, pending = false;
    for (;;) {
        if (!have) {
            if (quit) break;
            int w = -1;
            if (split) {
                const int x0 = (int)(blockIdx.x % NXCD);
                for (int d = 0; d < NXCD && w < 0; ++d) {
                    const int ch = (x0 + d) % NXCD;
                    const int lo = (int)((long long)count * ch / NXCD), hi = (int)((long long)count * (ch + 1) / NXCD);
                    if (gload(S.wcount + WC_CHUNK0 + WC_STRIDE * ch) + lo >= hi) continue;
                    const int k = gatomic_add(S.wcount + WC_CHUNK0 + WC_STRIDE * ch, 1);
                    if (lo + k < hi) w = lo + k;
                }
                if (w < 0) break;
            } else {
                w = gatomic_add(S.wcount + 1, 1);
                if (w >= count) break;
            }
            i = S.work[w];
            budget = S.budget[i];
            load_lane(P, i, L);
            c = Ctr{0, 0, 0};
            pending = false;
            have = true;
        }
        /* one trial per iteration on every lane; the accepted lanes finish their step */
        Trial T;
        const int r = trialA<MODE>(P, i, L, T, c);
        if (r == 0) continue;
        if (r == 1) acceptA<MODE>(P, i, L, pending, c, T);
        --budget;
        if (pending || L.stop != ID_UNKNOWN || budget <= 0) {
            store_lane(P, i, L);
            S.budget[i] = budget;
            gstore(S.n_field + i, gload(S.n_field + i) + (c.field));
            if (pending) {
                int slot = gatomic_add(S.qcount, 1);
                S.queue[slot] = i;
            }
            have = false;
        }
    }
}

/* work list of the persistent kernel: particles that are running and have steps left */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) work_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n) return;
    if (S.stop[i] == ID_UNKNOWN && (S.flags[i] & FL_STARTED) && S.budget[i] > 0) {
        int w = gatomic_add(S.wcount, 1);
        S.work[w] = i;
    }
}
#endif

#if PT_COMMON_TU
__global__ void reset_work_kernel(int *wcount) { wcount[0] = 0; wcount[1] = 0; }
#endif

/* ---- decay products (TParticle::DoDecay, src/particle.cpp:219-223) ---------------- */
#if PT_COMMON_TU
__global__ void reset_count_kernel(int *count) { *count = 0; }
#endif

/* TNeutron::Decay (src/neutron.cpp:285-362) for every neutron stopped with ID_DECAYED, from
   its final state and a copy of its stream position */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) decay_kernel(const Params *__restrict__ Pp, pt_beta_table T, pt_secondary *out, int *count) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n) return;
    if (S.stop[i] != ID_DECAYED || P.species[S.species[i]].kind != 0) return;
    const size_t n = (size_t)S.n;
    uint64_t ctr = S.ctr[i];
    const double v[3] = {S.y[3 * n + i], S.y[4 * n + i], S.y[5 * n + i]};
    pt_decay_out o;
    ptk_neutron_decay(&T, S.key[i], &ctr, v, &o);
    const int slot = atomicAdd(count, 2);
    const double t = S.t[i], x = S.y[i], y = S.y[n + i], z = S.y[2 * n + i];
    pt_secondary p{i, 1, t, {x, y, z}, o.Ep, o.phip, o.thetap, o.polp, 0};
    pt_secondary e{i, 2, t, {x, y, z}, o.Ee, o.phie, o.thetae, o.pole, 0};
    out[slot] = p;
    out[slot + 1] = e;
}
#endif

/* ---- field and solid queries (TFieldManager::BField/EField, TGeometry::GetSolid) ---- */

/* field sum at n points and times, FP64 tables; dB row-major [n][3][3] */
template <int MODE, bool COOP>
__global__ void __launch_bounds__(256) fields_eval_kernel(const Params *__restrict__ Pp, int n, const double *pts, const double *t, double *B,
                                                          double *dB, double *V, double *E) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long nf = 0;
    double b[3], db[3][3], v, e[3];
    const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    const bool grad = dB != nullptr;       /* without dB: BField(x, y, z, t, B), value-only */
    field_B<MODE, COOP>(P, x, y, z, t[i], b, db, grad, nf);
    field_E<MODE>(P, x, y, z, t[i], &v, e, nf);
    for (int k = 0; k < 3; ++k) {
        B[3 * i + k] = b[k];
        E[3 * i + k] = e[k];
        if (grad)
            for (int j = 0; j < 3; ++j) dB[9 * i + 3 * k + j] = db[k][j];
    }
    V[i] = v;
}

/* TGeometry::GetSolid (src/geometry.cpp:151-175): the solids containing the point (+z ray
   parity), the highest ID among those not ignored at t; -1 if more than MAX_SOLIDS nest,
   -2 if the traversal reported an incomplete visit (the stackless traversal never does) */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) solid_at_kernel(Geom G, int n, const double *pts, const double *t, int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long sk = sk_set(0ull, 0, -1, 0);
    int ns = 1;
    int ovf;
    unsigned long long par = ray_parity(G, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], &ovf);
    if (ovf) { out[i] = -2; return; }
    for (int m = 0; m < G.n_solid; ++m) {
        if ((par >> m) & 1ull) {
            if (ns >= MAX_SOLIDS) { out[i] = -1; return; }
            sk = sk_set(sk, ns, m, solid_ignored(G, m, t[i]));
            ++ns;
        }
    }
    out[i] = (int)solid_ID(G, current_solid(G, sk, ns));
}
#endif

/* TTriangleMesh::Collision's intersection count of the segments p[i] -> q[i] (-1 coplanar
   overlap, -2 more than MAX_HITS): mode 0 as the tracking kernels count (clearance grid,
   cell candidate list, cell subtree of the BVH, BVH), mode 1 from the BVH alone, mode 2 the
   way mode 0 took (0 cleared by the grid, 1 cell list, 2 whole BVH, 3 cell subtree) */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) segment_hits_kernel(Geom G, int n, const double *p, const double *q, int mode,
                                                           int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}, b[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    long long tests = 0;
    if (mode == 0) {
        out[i] = count_hits_geom<true>(G, a, b, tests);
    } else if (mode == 1) {
        out[i] = seg_query_v(G, a[0], a[1], a[2], b[0], b[1], b[2], nullptr).x;
    } else {
        int cell[3];
        if (clear_of_walls(G, a, b, cell)) out[i] = 0;
        else if (cell_list(G, cell, a, b).y >= 0) out[i] = 1;
        else out[i] = cell_root(G, cell, a, b) != 0 ? 3 : 2;
    }
}
#endif

/* ---- work list in spatial order (counting sort into Morton-ordered buckets) ---- */

/* clears the bucket counts and the work-list counters (kernels, not memsets, so the
   launch sequence can be captured in a graph) */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) clear_buckets_kernel(int *bhist, int *wcount) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= NBUCKET) bhist[b] = 0;
    if (b < WCOUNT_LEN) wcount[b] = 0;
}
#endif

/* spreads the low BUCKET_BITS bits of v to every third bit */
PTD unsigned spread3(unsigned v) {
    unsigned r = 0;
#pragma unroll
    for (int k = 0; k < BUCKET_BITS; ++k) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}

/* bucket of each particle with steps to do: its cell of a 32^3 grid over the mesh
   bounding box (the clearance grid's box), in Morton order */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) bucket_count_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n) return;
    if (S.stop[i] == ID_UNKNOWN && (S.flags[i] & FL_STARTED) && S.budget[i] > 0) {
        const Geom &G = P.geom;
        const size_t n = (size_t)S.n;
        unsigned c[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double u = (S.y[a * n + i] - G.glo[a]) * G.ginv_h[a] * ((double)(1 << BUCKET_BITS) / G.gn[a]);
            /* NaN and negatives to bucket 0, beyond the box to the last bucket */
            c[a] = !(u >= 0.0) ? 0u : (u >= (double)((1 << BUCKET_BITS) - 1) ? (unsigned)((1 << BUCKET_BITS) - 1) : (unsigned)u);
        }
        const int key = (int)(spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2));
        S.pkey[i] = key;
        gatomic_add(S.bhist + key, 1);
        gatomic_add(S.wcount, 1);
    }
}
#endif

/* exclusive prefix sum of the NBUCKET counts in one block of 1024 threads */
#if PT_COMMON_TU
__global__ void __launch_bounds__(1024) bucket_scan_kernel(int *bhist) {
    constexpr int PER = NBUCKET / 1024;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    int local[PER];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) { local[k] = bhist[t * PER + k]; sum += local[k]; }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;                 /* exclusive prefix of this thread's range */
#pragma unroll
    for (int k = 0; k < PER; ++k) { bhist[t * PER + k] = run; run += local[k]; }
}
#endif

/* places each particle with steps to do at the next free slot of its bucket */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) bucket_scatter_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const State &S = P.st;
    if (i >= S.n) return;
    if (S.stop[i] == ID_UNKNOWN && (S.flags[i] & FL_STARTED) && S.budget[i] > 0) {
        int slot = gatomic_add(S.bhist + S.pkey[i], 1);
        S.work[slot] = i;
    }
}
#endif

/*
 * Collision kernel: the parked particles of the queue, compacted into full
 * waves, run the collision path (CheckHit's bisection, DoHit/OnHit) on their
 * pending chord piece and finish the rest of that step.
 */
/* four waves per SIMD (128 VGPRs, the rest spilled; four 36 KB blocks of staged candidates
   per CU): the extra waves hide the latency of the collision path's BVH walks, scratch and
   dense-output reads (cylinder, 1e6 neutrons, 64 steps in 2 rounds: 25.1 -> 23.6 ms per
   launch against three waves with five staged candidates; earlier, one wave 18.6 ms against
   three 12.6 ms) */
#ifndef PT_COLL_MIN_WAVES
#define PT_COLL_MIN_WAVES 4
#endif
static_assert(CAND_LANES == 256, "the collision kernel's blocks fill the candidate staging area");
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PT_COLL_MIN_WAVES))) collision_kernel(const Params *__restrict__ Pp, const int *__restrict__ queue) {
    const Params &P = *Pp;
    const int count = *P.st.qcount;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        Ctr c = {0, 0, 0};
        collision_path<MODE>(P, queue[j], c);
    }
}

/* ---- collision order: the queue counting-sorted by the candidate count of each parked
   chord piece (the particles are independent, so the order changes no result) ---- */
#if PT_COMMON_TU
__global__ void qclear_kernel(int *qhist) {
    if (threadIdx.x <= QKEYS) qhist[threadIdx.x] = 0;
}

/* key of each queue entry: the number of candidates CheckHit will collect for its piece
   (QKEYS - 1 when there are more than MAX_CAND, 0 for a spin-only piece of zero length) */
__global__ void __launch_bounds__(256) qkey_kernel(const Params *__restrict__ Pp) {
    const Params &P = *Pp;
    const State &S = P.st;
    const int count = *S.qcount;
    const size_t n = (size_t)S.n;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        const int i = S.queue[j];
        const double *pd = S.pend;
        int key = 0;
        if (pd[10 * n + i] > pd[i]) {
            double y1[9], y2[9];
            for (int k = 0; k < 9; ++k) { y1[k] = pd[(1 + k) * n + i]; y2[k] = pd[(11 + k) * n + i]; }
            double lo[3], hi[3];
            piece_cand_box(y1, y2, lo, hi);
            Cand C;
            cand_build(P.geom, lo, hi, C, false);
            key = C.n < 0 ? QKEYS - 1 : C.n;
        }
        S.qkey[j] = key;
        gatomic_add(S.qhist + key, 1);
    }
}

/* exclusive prefix sum of the key counts (QKEYS of them: one thread) */
__global__ void qscan_kernel(int *qhist) {
    int run = 0;
    for (int k = 0; k < QKEYS; ++k) {
        const int c = qhist[k];
        qhist[k] = run;
        run += c;
    }
}

__global__ void __launch_bounds__(256) qscatter_kernel(const Params *__restrict__ Pp) {
    const State &S = Pp->st;
    const int count = *S.qcount;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        const int slot = gatomic_add(S.qhist + S.qkey[j], 1);
        S.qsorted[slot] = S.queue[j];
    }
}
#endif

/* empties the collision queue (a kernel, not a memset, so launch loops can be captured) */
#if PT_COMMON_TU
__global__ void reset_queue_kernel(int *qcount) { *qcount = 0; }
#endif

/* step budget of every particle for the next round of launches */
#if PT_COMMON_TU
__global__ void __launch_bounds__(256) budget_kernel(int *budget, int n, int steps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) budget[i] = steps;
}
#endif

} // namespace ptd

/* ---- launchers (host side, called from pt_host.cpp) ---------------------- */

#if PT_MODE_TU >= 0
/* the kernels of one mode (this translation unit is built once per mode) */
template <int MODE>
hipError_t pt_launch_init_mode(const ptd::Params &P, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(ptd::init_kernel<MODE>, dim3(grid), dim3(256), 0, stream, P.self);
    return hipGetLastError();
}
template <int MODE>
hipError_t pt_launch_step_mode(const ptd::Params &P, int grid, int block, hipStream_t stream) {
    hipLaunchKernelGGL(ptd::step_kernel<MODE>, dim3(grid), dim3(block), 0, stream, P.self);
    return hipGetLastError();
}
template <int MODE>
hipError_t pt_launch_step_persistent_mode(const ptd::Params &P, int grid, int split, hipStream_t stream) {
    hipLaunchKernelGGL(ptd::step_kernel_persistent<MODE>, dim3(grid), dim3(256), 0, stream, P.self, split);
    return hipGetLastError();
}
template <int MODE>
hipError_t pt_launch_collision_mode(const ptd::Params &P, int grid, const int *queue, hipStream_t stream) {
    /* blocks of CAND_LANES threads: the size of the candidate staging area in LDS */
    hipLaunchKernelGGL(ptd::collision_kernel<MODE>, dim3(grid), dim3(ptd::CAND_LANES), 0, stream, P.self, queue);
    return hipGetLastError();
}
template <int MODE>
hipError_t pt_launch_fields_eval_mode(const ptd::Params &P, int coop, int n, const double *pts, const double *t, double *B,
                                      double *dB, double *V, double *E, hipStream_t stream) {
    if (coop)
        hipLaunchKernelGGL((ptd::fields_eval_kernel<MODE, true>), dim3((n + 255) / 256), dim3(256), 0, stream, P.self, n, pts,
                           t, B, dB, V, E);
    else
        hipLaunchKernelGGL((ptd::fields_eval_kernel<MODE, false>), dim3((n + 255) / 256), dim3(256), 0, stream, P.self, n,
                           pts, t, B, dB, V, E);
    return hipGetLastError();
}
/* PT_KERNEL_SET 0: the set-up and step kernels of the mode, 2: its persistent step kernel,
   1: its collision kernel and field queries (three translation units per mode) */
#if PT_KERNEL_SET == 0
template hipError_t pt_launch_init_mode<PT_MODE_TU>(const ptd::Params &, int, hipStream_t);
template hipError_t pt_launch_step_mode<PT_MODE_TU>(const ptd::Params &, int, int, hipStream_t);
#elif PT_KERNEL_SET == 2
template hipError_t pt_launch_step_persistent_mode<PT_MODE_TU>(const ptd::Params &, int, int, hipStream_t);
#else
template hipError_t pt_launch_collision_mode<PT_MODE_TU>(const ptd::Params &, int, const int *, hipStream_t);
template hipError_t pt_launch_fields_eval_mode<PT_MODE_TU>(const ptd::Params &, int, int, const double *, const double *,
                                                           double *, double *, double *, double *, hipStream_t);
#endif
#endif /* PT_MODE_TU >= 0 */

#if PT_COMMON_TU
/* dispatch on the mode to the translation unit that holds its kernels */
static hipError_t init_dispatch(const ptd::Params &P, int mode, int grid, hipStream_t s) {
    switch (mode) {
    case 0: return pt_launch_init_mode<0>(P, grid, s);
    case 1: return pt_launch_init_mode<1>(P, grid, s);
    case 2: return pt_launch_init_mode<2>(P, grid, s);
    default: return pt_launch_init_mode<3>(P, grid, s);
    }
}

hipError_t pt_launch_init(const ptd::Params &P, int mode, int steps, hipStream_t stream) {
    int grid = (P.st.n + 255) / 256;
    if (grid == 0) return hipSuccess;
    hipError_t e = init_dispatch(P, mode, grid, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ptd::budget_kernel, dim3(grid), dim3(256), 0, stream, P.st.budget, P.st.n, steps);
    return hipGetLastError();
}

hipError_t pt_launch_step(const ptd::Params &P, int mode, int block, hipStream_t stream) {
    int grid = (P.st.n + block - 1) / block;
    if (grid == 0) return hipSuccess;
    switch (mode) {
    case 0: return pt_launch_step_mode<0>(P, grid, block, stream);
    case 1: return pt_launch_step_mode<1>(P, grid, block, stream);
    case 2: return pt_launch_step_mode<2>(P, grid, block, stream);
    default: return pt_launch_step_mode<3>(P, grid, block, stream);
    }
}

hipError_t pt_launch_step_persistent(const ptd::Params &P, int mode, int grid, int spatial, hipStream_t stream) {
    if (P.st.n == 0) return hipSuccess;
    int wgrid = (P.st.n + 255) / 256;
    if (spatial) {
        static_assert(ptd::NBUCKET % 1024 == 0, "bucket scan layout");
        hipLaunchKernelGGL(ptd::clear_buckets_kernel, dim3(ptd::NBUCKET / 256 + 1), dim3(256), 0, stream, P.st.bhist,
                           P.st.wcount);
        hipLaunchKernelGGL(ptd::bucket_count_kernel, dim3(wgrid), dim3(256), 0, stream, P.self);
        hipLaunchKernelGGL(ptd::bucket_scan_kernel, dim3(1), dim3(1024), 0, stream, P.st.bhist);
        hipLaunchKernelGGL(ptd::bucket_scatter_kernel, dim3(wgrid), dim3(256), 0, stream, P.self);
    } else {
        hipLaunchKernelGGL(ptd::reset_work_kernel, dim3(1), dim3(1), 0, stream, P.st.wcount);
        hipLaunchKernelGGL(ptd::work_kernel, dim3(wgrid), dim3(256), 0, stream, P.self);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int split = spatial >= 2;
    switch (mode) {
    case 0: return pt_launch_step_persistent_mode<0>(P, grid, split, stream);
    case 1: return pt_launch_step_persistent_mode<1>(P, grid, split, stream);
    case 2: return pt_launch_step_persistent_mode<2>(P, grid, split, stream);
    default: return pt_launch_step_persistent_mode<3>(P, grid, split, stream);
    }
}

hipError_t pt_launch_fields_eval(const ptd::Params &P, int mode, int coop, int n, const double *pts, const double *t,
                                 double *B, double *dB, double *V, double *E, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    switch (mode) {
    case 0: return pt_launch_fields_eval_mode<0>(P, coop, n, pts, t, B, dB, V, E, stream);
    case 1: return pt_launch_fields_eval_mode<1>(P, coop, n, pts, t, B, dB, V, E, stream);
    case 2: return pt_launch_fields_eval_mode<2>(P, coop, n, pts, t, B, dB, V, E, stream);
    default: return pt_launch_fields_eval_mode<3>(P, coop, n, pts, t, B, dB, V, E, stream);
    }
}

hipError_t pt_launch_solid_at(const ptd::Geom &G, int n, const double *pts, const double *t, int *out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ptd::solid_at_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, G, n, pts, t, out);
    return hipGetLastError();
}

hipError_t pt_launch_segment_hits(const ptd::Geom &G, int n, const double *p, const double *q, int mode, int *out,
                                  hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ptd::segment_hits_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, G, n, p, q, mode, out);
    return hipGetLastError();
}

hipError_t pt_launch_decay(const ptd::Params &P, pt_beta_table T, pt_secondary *out, int *count, hipStream_t stream) {
    hipLaunchKernelGGL(ptd::reset_count_kernel, dim3(1), dim3(1), 0, stream, count);
    if (P.st.n > 0)
        hipLaunchKernelGGL(ptd::decay_kernel, dim3((P.st.n + 255) / 256), dim3(256), 0, stream, P.self, T, out, count);
    return hipGetLastError();
}

hipError_t pt_launch_collision(const ptd::Params &P, int mode, int coll_grid, int order, hipStream_t stream) {
    if (P.st.n == 0) return hipSuccess;
    hipError_t e;
    const int *queue = P.st.queue;
    if (order) {
        hipLaunchKernelGGL(ptd::qclear_kernel, dim3(1), dim3(64), 0, stream, P.st.qhist);
        hipLaunchKernelGGL(ptd::qkey_kernel, dim3(coll_grid), dim3(256), 0, stream, P.self);
        hipLaunchKernelGGL(ptd::qscan_kernel, dim3(1), dim3(1), 0, stream, P.st.qhist);
        hipLaunchKernelGGL(ptd::qscatter_kernel, dim3(coll_grid), dim3(256), 0, stream, P.self);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        queue = P.st.qsorted;
    }
    switch (mode) {
    case 0: e = pt_launch_collision_mode<0>(P, coll_grid, queue, stream); break;
    case 1: e = pt_launch_collision_mode<1>(P, coll_grid, queue, stream); break;
    case 2: e = pt_launch_collision_mode<2>(P, coll_grid, queue, stream); break;
    default: e = pt_launch_collision_mode<3>(P, coll_grid, queue, stream); break;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ptd::reset_queue_kernel, dim3(1), dim3(1), 0, stream, P.st.qcount);
    return hipGetLastError();
}
#endif /* PT_COMMON_TU */
