/*
 * pt_kernels.h -- launchers exported by the HIP translation units to the host
 * code of the library.
 */
#ifndef PT_KERNELS_H_
#define PT_KERNELS_H_

#include <hip/hip_runtime.h>

#include "../../include/pentrack_gpu.h"
#include "pt_decay.h"

namespace ptd { struct Params; struct Geom; }

/* mode: bit 0 mixed-precision field tables (PT_VARIANT_MIXED), bit 1 containers with scale formulas,
   analytic fields or conductors present */

/* static LDS of the step kernels: the dopri5 stage vectors k2..k5, 4 x 8 doubles for each of
   at most 256 threads, and the staging of the wave-cooperative coefficient gathers, 4 KB for
   each of at most 4 waves (must match stage_lds and coop_lds in pt_track.hip) */
#define PT_STEP_LDS_BYTES (4 * 8 * 256 * 8 + 4 * 32 * 128)

/* the mode-templated launches, each instantiated in the translation unit of its mode */
template <int MODE> hipError_t pt_launch_init_mode(const ptd::Params &P, int grid, hipStream_t stream);
template <int MODE> hipError_t pt_launch_step_mode(const ptd::Params &P, int grid, int block, hipStream_t stream);
template <int MODE> hipError_t pt_launch_step_persistent_mode(const ptd::Params &P, int grid, int split, hipStream_t stream);
template <int MODE> hipError_t pt_launch_collision_mode(const ptd::Params &P, int grid, const int *queue, hipStream_t stream);
template <int MODE>
hipError_t pt_launch_fields_eval_mode(const ptd::Params &P, int coop, int n, const double *pts, const double *t, double *B,
                                      double *dB, double *V, double *E, hipStream_t stream);

/* particle set-up (first call only) and the step budget of every particle */
hipError_t pt_launch_init(const ptd::Params &P, int mode, int steps, hipStream_t stream);
/* one round = the step kernel, then the collision kernel over the queue and the queue reset */
hipError_t pt_launch_step(const ptd::Params &P, int mode, int block, hipStream_t stream);
/* persistent variant of the step kernel: the work list (in index order, or with spatial != 0
   in Morton order of the particle positions), then `grid` blocks of 256 slots */
hipError_t pt_launch_step_persistent(const ptd::Params &P, int mode, int grid, int spatial, hipStream_t stream);
hipError_t pt_launch_collision(const ptd::Params &P, int mode, int coll_grid, int order, hipStream_t stream);
/* field sum at n points and times; mode: 1 mixed-precision tables | 2 containers with formulas or
   analytic fields; coop: the step kernels' cooperative coefficient gathers instead of per-lane loads */
hipError_t pt_launch_fields_eval(const ptd::Params &P, int mode, int coop, int n, const double *pts, const double *t,
                                 double *B, double *dB, double *V, double *E, hipStream_t stream);
/* TabField3::PreInterpol on the device: the seven spline-derivative tables der[1..7] of the
   samples der[0] (each [nx][ny][nz]; scratch: 2 nx ny nz doubles), then the 64 coefficients
   per cell of component comp (der_dev: the 8 table pointers in device memory, A: the 64 x 64
   matrix as doubles; device_order 1: the tracker's coefficient order) */
hipError_t pt_launch_tricubic_derivs(const double *const g[3], int nx, int ny, int nz, double *const der[8],
                                     double *scratch, hipStream_t stream);
hipError_t pt_launch_tricubic_coef(const double *xg, const double *yg, const double *zg, int nx, int ny, int nz,
                                   const double *const *der_dev, const double *A, int ncomp, int comp, int device_order,
                                   double *coef, hipStream_t stream);
/* the mixed-precision table (ptd::mix_cell_bytes(ncomp) per cell) of a map from its
   FP64 table in the tracker's order; *n_fp64 counts the blocks kept in FP64 (tol: bound on
   the FP32 remainder's gradient error, field unit per metre) */
hipError_t pt_launch_mixed_table(const double *xg, const double *yg, const double *zg, int nx, int ny, int nz,
                                 int ncomp, const double *coef, double tol, unsigned char *mix, int *n_fp64,
                                 hipStream_t stream);
/* ID of the solid with the highest priority at each point and time */
hipError_t pt_launch_solid_at(const ptd::Geom &G, int n, const double *pts, const double *t, int *out, hipStream_t stream);
hipError_t pt_launch_segment_hits(const ptd::Geom &G, int n, const double *p, const double *q, int mode, int *out,
                                  hipStream_t stream);
/* decay products of the batch's decayed neutrons into out (count reset first) */
hipError_t pt_launch_decay(const ptd::Params &P, pt_beta_table T, pt_secondary *out, int *count, hipStream_t stream);

#endif /* PT_KERNELS_H_ */
