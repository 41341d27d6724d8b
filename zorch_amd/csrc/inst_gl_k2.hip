// Proof instances, batch groups and server kinds with Goldilocks tables and Goldilocks challenges,
// 2 factor tables (one translation unit per field pair and factor count,
// so that the kernel templates compile in parallel).
#include "units.h"

ZS_UNIT(gl, GL, GL, 2)
