// vs_fast_k19.hip -- the K = 19 forward and recording instantiations of the
// on-chip FP32 step kernel of vs_fast.hip, built as their own translation
// unit so that the Makefile can give them another instruction scheduler
// (-mllvm -amdgpu-sched-strategy=iterative-ilp).  Under the default
// scheduler this K alone gets a loop-carried register rotation of about
// 100 moves per step and spills; under the other it compiles like its
// neighbours (about 20 VALU per cell slot and step, no copies, no spills).
#define VS_FAST_SPLIT_K 19
#include "vs_fast.hip"
