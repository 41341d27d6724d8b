// Proof instances, batch groups and server kinds with BabyBear tables and BabyBear challenges,
// 1 factor table (one translation unit per field pair and factor count,
// so that the kernel templates compile in parallel).
#include "units.h"

ZS_UNIT(bb, BB, BB, 1)
