// Plane kernel of the local-potential transform with register-resident lines
// (declarations).
#pragma once

#include "wh_fft.h"

namespace wh {

// Number of register-line plane variants for this plan in FP32 (fp32) or
// FP64 (0: the grid is not cubic of a compiled size -- FP64 lines stop at 64
// points -- or its sphere is too wide for them).
int plane_reg_variants(const StickPlanDev& P, bool fp32);

// Variant v (0 <= v < plane_reg_variants(P, true)) of the FP32 plane step
// sticks_out = P_sphere F_12 V F_12^-1 sticks_in (same contract as the
// LDS-pass plane kernels: items[i].a selects the input sticks, output item i;
// items[i].pot >= 0 selects a per-item potential vpot + pot * ldpot).
void plane_reg_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z32* sin, z32* sout,
                    const float* vpot, long ldpot, double scale, hipStream_t s);
// The same in FP64 (0 <= v < plane_reg_variants(P, false)).
void plane_reg_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z64* sin, z64* sout,
                    const double* vpot, long ldpot, double scale, hipStream_t s);

// Register-line z transforms of the FP32 operator (same contract as the
// size-specialised zinv/zfwd; zfwd in the pair modes only), available when
// the grid is cubic of a compiled size and the sphere spans |k| <= N/4.
bool zreg_available(const StickPlanDev& P);
void zinv_reg(const StickPlanDev& P, const FftItem* items, int n, const z32* psiA, long ldA, const z32* psiB,
              long ldB, z32* sticks, int mode, hipStream_t s);
void zfwd_reg(const StickPlanDev& P, const FftItem* items, int n, const z32* sticks, const z32* in, long ldin,
              z32* out, long ldout, const ZfwdArgs& A, hipStream_t s);

}  // namespace wh
