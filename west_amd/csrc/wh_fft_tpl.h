// Compile-time-size FFT kernels for cubic grids (declarations).
#pragma once

#include "wh_fft.h"

namespace wh {

struct TplKernels {
  void (*zinv)(const StickPlanDev&, const FftItem*, int, const z64*, long, const z64*, long, void*, int, hipStream_t);
  void (*zfwd)(const StickPlanDev&, const FftItem*, int, const void*, const z64*, long, z64*, long, const ZfwdArgs&,
               hipStream_t);
  // occ: workgroups per CU the plane kernel's LDS is sized for (1..max_occ)
  void (*vloc)(const StickPlanDev&, const FftItem*, int, const void*, void*, const double*, long, double, int,
               hipStream_t);
  void (*to_real)(const StickPlanDev&, const FftItem*, int, const void*, double*, long, hipStream_t);
  // density(P, npert, nband, sticks in, occ, sticks out, scratch, stream): scratch holds
  // density_scratch reals per perturbation (may be null when that is 0)
  void (*density)(const StickPlanDev&, int, int, const void*, const double*, void*, void*, hipStream_t);
  long density_scratch;   // 0: none needed; -1: this grid has no density kernel
  int max_occ;
};

// FP32 operator (sticks, vectors and potential in FP32)
struct TplKernels32 {
  void (*zinv)(const StickPlanDev&, const FftItem*, int, const z32*, long, const z32*, long, void*, int, hipStream_t);
  void (*zfwd)(const StickPlanDev&, const FftItem*, int, const void*, const z32*, long, z32*, long, const ZfwdArgs&,
               hipStream_t);
  void (*vloc)(const StickPlanDev&, const FftItem*, int, const void*, void*, const float*, long, double, int,
               hipStream_t);
  int max_occ;
};

// kernels specialised on N = nr1 = nr2 = nr3, or null when the grid is not
// one of the compiled sizes (the runtime-size kernels are used then)
const TplKernels* tpl_kernels_f64(const StickPlanDev& P);
const TplKernels32* tpl_kernels_f32(const StickPlanDev& P);

}  // namespace wh
