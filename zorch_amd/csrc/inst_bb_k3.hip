// Proof instances, batch groups and server kinds with BabyBear tables and BabyBear challenges,
// 3 factor tables (one translation unit per field pair and factor count,
// so that the kernel templates compile in parallel).
#include "units.h"

ZS_UNIT(bb, BB, BB, 3)
