// CPU emulation build of the boundary-layer kernels, S=3 (test infrastructure).
#define CNS_S 3
#include "../../cudanavierstokes_amd/csrc/cns_kernels_bl.hip"
