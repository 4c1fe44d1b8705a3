// CPU emulation build: the HIP runtime header maps to the emulation layer.
#pragma once
#include "../hip_emu.hpp"
