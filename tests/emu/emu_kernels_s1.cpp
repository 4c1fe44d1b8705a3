// CPU emulation build of the product kernels, S=1 (test infrastructure).
#define CNS_S 1
#include "../../cudanavierstokes_amd/csrc/cns_kernels.hip"
