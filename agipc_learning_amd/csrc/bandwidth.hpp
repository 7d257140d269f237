// Streaming-rate probes of the card (read-only stream, copy).
#pragma once
#include <hip/hip_runtime.h>

namespace masx
{
// Fastest of `reps` launches over a buffer of `bytes` (kind 0: read-only,
// counting the bytes read; kind 1: copy, counting bytes read + written), GB/s.
double stream_rate(long long bytes, int kind, int reps);
}  // namespace masx
