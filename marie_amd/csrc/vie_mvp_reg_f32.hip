// complex64 instantiation of the register-engine matvec kernels
#define MARIE_REG_T float
#include "vie_mvp_reg.inc"
