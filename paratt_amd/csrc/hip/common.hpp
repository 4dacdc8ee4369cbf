// common.hpp -- shared helpers for the gfx950 kernels of the TT library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ptt_hip.h"

namespace ptt {

extern hipStream_t g_stream;
extern hipStream_t g_launch;  // g_stream, or the side stream between side_begin / side_end
[[noreturn]] void fatal_hip(hipError_t e, const char *what, const char *file,
                            int line);

#define HIPCHECK(x)                                              \
    do {                                                         \
        hipError_t e_ = (x);                                     \
        if (e_ != hipSuccess) ptt::fatal_hip(e_, #x, __FILE__, __LINE__); \
    } while (0)

// Records HIP events around the launches issued in its lifetime when
// profiling is on; the elapsed time is attributed to `name`.
struct ProfScope {
    bool active;
    const char *nm;
    double flops, bytes;
    hipEvent_t a;
    explicit ProfScope(const char *name, double flops = 0.0, double bytes = 0.0);
    ~ProfScope();
};

// Launch on the current library stream; with profiling on, HIP events
// bracket the launch and its algorithmic flops / bytes are credited to `name`.
#define PTT_LAUNCH_FB(name, flops, bytes, kern, grid, block, shmem, ...) \
    do {                                                                \
        ptt::ProfScope ps_(name, flops, bytes);                         \
        hipLaunchKernelGGL(kern, grid, block, shmem, ptt::g_launch,     \
                           __VA_ARGS__);                                \
        HIPCHECK(hipGetLastError());                                    \
    } while (0)
#define PTT_LAUNCH(name, kern, grid, block, shmem, ...) \
    PTT_LAUNCH_FB(name, 0.0, 0.0, kern, grid, block, shmem, __VA_ARGS__)

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Grid-stride cap for streaming kernels: enough waves to fill 256 CUs.
inline unsigned stream_blocks(long n, int threads) {
    long b = (n + threads - 1) / threads;
    long cap = 256L * 16;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace ptt
