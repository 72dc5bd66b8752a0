// FP64 instantiations of the register-line plane kernel (wh_fft_reg.hip),
// compiled as a translation unit of their own.
#define WH_REG_FP64 1
#include "wh_fft_reg.hip"
