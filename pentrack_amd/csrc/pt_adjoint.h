/*
 * pt_adjoint.h -- device-side view of the differentiable RK4 path (the C ABI
 * structures of include/pentrack_gpu.h under the names the kernels use).
 */
#ifndef PT_ADJOINT_H_
#define PT_ADJOINT_H_

#include "../../include/pentrack_gpu.h"

#define PTA_MAX_MAPS PT_ADJ_MAX_MAPS

/* records msg as the calling thread's pt_last_error() and returns code (pt_host.cpp) */
int pt_fail(int code, const char *msg);

namespace pta {
typedef pt_adj_maps MapSet;
typedef pt_adj_params AdjParams;
} // namespace pta

#endif /* PT_ADJOINT_H_ */
