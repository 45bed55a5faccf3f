// vs_fast_k17.hip -- the K = 17 forward and recording instantiations of the
// on-chip FP32 step kernel of vs_fast.hip, built as their own translation
// unit so that the Makefile can give them another instruction scheduler
// (-mllvm -amdgpu-sched-strategy=iterative-ilp).  Under the default
// scheduler the forward kernel of this K alone runs at the 128-VGPR limit
// and spills; under the other it compiles like its neighbours (about 18
// VALU and 6.4 LDS instructions per cell slot and step, no spills).
#define VS_FAST_SPLIT_K 17
#include "vs_fast.hip"
