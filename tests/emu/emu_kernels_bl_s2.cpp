// CPU emulation build of the boundary-layer kernels, S=2 (test infrastructure).
#define CNS_S 2
#include "../../cudanavierstokes_amd/csrc/cns_kernels_bl.hip"
