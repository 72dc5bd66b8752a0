// FP64 plane kernel of the local-potential transform with each line held by a
// quad of lanes (declarations).
#pragma once

#include "wh_fft.h"

namespace wh {

// Number of quad-lane plane variants for this plan (0: the grid is not cubic
// of a compiled size or its sphere is too wide).
int plane_q4_variants(const StickPlanDev& P);

// Variant v (0 <= v < plane_q4_variants(P)) of the FP64 plane step
// sticks_out = P_sphere F_12 V F_12^-1 sticks_in (same contract as the other
// plane kernels: items[i].a selects the input sticks, output item i;
// items[i].pot >= 0 selects a per-item potential vpot + pot * ldpot).
void plane_q4_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z64* sin, z64* sout,
                   const double* vpot, long ldpot, double scale, hipStream_t s);

}  // namespace wh
