// dns_ens_internal.h -- library-internal view of an ensemble's training pairs (used by
// the closure trainer); not part of the C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

struct dns_ens;

struct EnsPairView {
    const float *pairs;  // [max_pairs][count][4][n][n]
    int count;           // members in the group
    int stored;          // pairs held per member
    int n;               // coarse grid
};

int dns_ens_group_count(const dns_ens *e);
// 0 on success
int dns_ens_group_pairs(const dns_ens *e, int g, EnsPairView *out);
