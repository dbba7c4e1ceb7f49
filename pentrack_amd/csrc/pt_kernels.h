/*
 * pt_kernels.h -- launchers exported by the HIP translation units to the host
 * code of the library.
 */
#ifndef PT_KERNELS_H_
#define PT_KERNELS_H_

#include <hip/hip_runtime.h>

namespace ptd { struct Params; }

hipError_t pt_launch_track(const ptd::Params &P, int variant, int block, hipStream_t stream);

#endif /* PT_KERNELS_H_ */
