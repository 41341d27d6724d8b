// Proof instances, batch groups and server kinds with BabyBear tables and BabyBear^4 challenges,
// 2 factor tables (one translation unit per field pair and factor count,
// so that the kernel templates compile in parallel).
#include "units.h"

ZS_UNIT(bb_bb4, BB, BB4, 2)
