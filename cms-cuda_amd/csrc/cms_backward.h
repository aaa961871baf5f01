#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cms_internal.h"

namespace cms {

__global__ void k_check_dots(BatchView b, Settings s, Count cnD, const double4* dpos, const int4* dd1, const int4* dd2,
                             unsigned long long* nbad, double* maxdev);
__global__ void k_cms_adjoint(Count cnB, const int* blist, const int* bgrp_of, const int* bgrp, const int* bidx,
                              const double* bmin, const double4* dpos, const int* datom, const double* g_total,
                              const double* g_atom, double weight, double* adj);
__global__ void k_dot_grad(BatchView b, Settings s, Count cnD, const int4* dd1, const int4* dd2, const double* adj,
                           double* grad);
__global__ void k_frozen_eval(BatchView b, Settings s, const double4* coords, Count cnB, const int* blist,
                              const int4* dd1, const int4* dd2, double4* out);
__global__ void k_frozen_cms(Count cnB, const int* bgrp_of, const int* bgrp, const int* bidx, const double4* ev,
                             double weight, double* totals);

}  // namespace cms
