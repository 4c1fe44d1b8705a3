// CPU emulation build of the product host API (test infrastructure).
#include "../../cudanavierstokes_amd/csrc/cns_api.hip"
