// CPU emulation build of the product kernels, S=2 (test infrastructure).
#define CNS_S 2
#include "../../cudanavierstokes_amd/csrc/cns_kernels.hip"
