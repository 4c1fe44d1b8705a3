// CPU emulation build of the boundary-layer kernels, S=1 (test infrastructure).
#define CNS_S 1
#include "../../cudanavierstokes_amd/csrc/cns_kernels_bl.hip"
