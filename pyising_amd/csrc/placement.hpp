// Where the workgroups of a dispatch run: the XCD of the calling workgroup, and the placement
// pattern of a device's dispatches (probed once per device).
#pragma once

#include "common.hpp"

__device__ __forceinline__ u32 xcc_id()
{
    return __builtin_amdgcn_s_getreg((15 << 11) | (0 << 6) | 20) & 15u;  // HW_REG_XCC_ID
}

// Placement of a dispatch's workgroups on the current device: 0 = every workgroup on one XCD,
// 8 = round-robin over 8 XCDs (workgroup i on XCD (i + r) % 8 for a starting XCD r that may
// change between dispatches), -1 = neither (defined in msc_fused.hip; probed once per device).
int fused_placement(int* out);
