/*
 * FP64 tile kernels with two W sets ((fT)).  The same source as
 * pt_kernel.hip, built without the VGPR-form MFMA option: with 256
 * accumulator registers these kernels need the AGPR file.
 */
#define PT_KERNEL_F64X2 1
#include "pt_kernel.hip"
