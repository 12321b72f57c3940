// Host objects behind the opaque handles of the C ABI that more than one
// translation unit needs.
#pragma once

#include "cuphy_hip.h"

// LDPC decoder object: immutable after creation (decodes from several host
// threads on different streams are safe), owns no device memory.
struct cuphyLDPCDecoder {
    int device = 0;
    int cu_count = 0;
};
