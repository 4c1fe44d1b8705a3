// CPU emulation build of the boundary-layer kernels, S=4 (test infrastructure).
#define CNS_S 4
#include "../../cudanavierstokes_amd/csrc/cns_kernels_bl.hip"
