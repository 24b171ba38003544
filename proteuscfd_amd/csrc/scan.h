// Fixed-order exclusive scan of per-pass counts (kernels_misc.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace pcfd {

// off[p] = cnt[0] + ... + cnt[p-1] for p <= n (off has n+1 entries), on `stream`.
void launch_scan_counts(hipStream_t stream, int n, const int* cnt, int* off);

}  // namespace pcfd
