/*
 * Tile kernels built without the VGPR-form MFMA option: FP64 with two W
 * sets ((fT)) and the split forms.  The same source as pt_kernel.hip;
 * pt_launch_tiles there forwards these launches here.
 */
#define PT_KERNEL_AGPR 1
#include "pt_kernel.hip"
