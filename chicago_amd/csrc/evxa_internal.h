// evxa_internal.h -- device-side job layout shared by the HIP kernels and the
// host runtime of libevxa.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evxa.h"

namespace evxa {

// Pictures live in HBM as int16 planes with a replicated-edge apron of
// kApron samples on every side, so every 6-tap read lands inside the buffer
// and equals the clamped read of interp.cpp:8-14.  Pointers address sample
// (0, 0); pitches are in elements and keep rows 128-byte aligned.
constexpr int kApron = 16;

struct DevPic {
    const int16_t *y;
    const int16_t *u;
    const int16_t *v;
    int32_t py;
    int32_t pc;
};

// One frame of work inside a launch.
struct DevJob {
    DevPic src;
    DevPic refs[EVXA_MAX_REFS];
    int32_t width, height, wb, hb;
    int32_t type, quality, n_refs, thresh;
    float lambda;
    uint32_t target_ids;               // 4 packed bytes
    const evxa_block_record *prev;     // temporal table or nullptr
    evxa_block_record *records;
    int16_t *coeffs;
    int16_t *levels;
    int32_t *progress;                 // hb counters: MBs finished per row
};

// Work-list entry: job index in the top 12 bits, MB index in the low 20.
constexpr int kJobShift = 20;
constexpr uint32_t kMbMask = (1u << kJobShift) - 1;
constexpr int kMaxJobs = 1 << (32 - kJobShift);

// Counter block of a launch: [0] next ticket, [1] error flags.
constexpr int kCounterTicket = 0;
constexpr int kCounterError = 1;
constexpr int kCounterWords = 16;

// Error flag bits.
constexpr int32_t kErrSearchFailed = 1;
constexpr int32_t kErrWaitTimeout = 2;

// Kernel launchers (evxa_kernels.hip).
hipError_t launch_reset(int32_t *words, int32_t n, hipStream_t st);
hipError_t launch_analyze(const DevJob *jobs, const uint32_t *tickets, int32_t n_tickets,
                          int32_t *counters, int32_t grid, hipStream_t st);
hipError_t launch_apron(int16_t *plane0, int32_t pitch, int32_t w, int32_t h, hipStream_t st);
hipError_t launch_interp_blocks(DevPic pic, int32_t w, int32_t h, int32_t n, const int32_t *params,
                                int16_t *out, hipStream_t st);
hipError_t launch_search_blocks(DevPic src, DevPic ref, int32_t w, int32_t h, int32_t quality,
                                float lambda, int32_t n, const int32_t *params,
                                evxa_search_result *out, hipStream_t st);
hipError_t launch_dct_quant_blocks(int32_t n, const int16_t *src, const int16_t *pred,
                                   const int32_t *quality, int16_t *coeffs, int16_t *levels,
                                   int32_t *qp, int32_t *variance, hipStream_t st);
// Grid size that keeps every wave of the analysis kernel resident.
int32_t analyze_resident_waves(int32_t device);

}  // namespace evxa
