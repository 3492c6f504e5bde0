// complex128 instantiation of the register-engine matvec kernels
#define MARIE_REG_T double
#include "vie_mvp_reg.inc"
