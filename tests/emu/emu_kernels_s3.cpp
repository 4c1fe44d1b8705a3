// CPU emulation build of the product kernels, S=3 (test infrastructure).
#define CNS_S 3
#include "../../cudanavierstokes_amd/csrc/cns_kernels.hip"
