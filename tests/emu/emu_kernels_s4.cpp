// CPU emulation build of the product kernels, S=4 (test infrastructure).
#define CNS_S 4
#include "../../cudanavierstokes_amd/csrc/cns_kernels.hip"
