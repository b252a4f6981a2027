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
constexpr int kApron = 24;

// Spread range records (see note_range): kRangeWays word pairs, kRangeStride
// words apart.
constexpr int kRangeWays = 64;
constexpr int kRangeStride = 16;
constexpr int kRangeWords = kRangeWays * kRangeStride;

struct DevPic {
    const int16_t *y;
    const int16_t *u;
    const int16_t *v;
    int32_t py;
    int32_t pc;
    // The picture's sample range word, max << 16 | (min & 0xFFFF) (zero
    // when never written; folded after upload, reconstruction and
    // filtering).  A source / reference pair spanning at most 2048 values
    // within [-8192, 8191] takes the packed 16-bit metric paths.
    const int32_t *wide;
};

// Outcome of one (MB, reference) search, handed to the wave that finishes
// the MB (32 bytes).  Three 64-bit words, each stored as one access with
// bit 63 set, so a reader that sees the bit sees the whole word:
//   w[0]: bits 0..15 best x, 16..31 best y, 32..62 MAD
//   w[1]: bits 0..31 RD cost (fp32 bits), 32..62 SAD
//   w[2]: bits 0..15 integer candidates, 16..23 axial and 24..31 diagonal
//         sub-pel directions, 32 sub-pel enable, 33 quarter, 34..36 sub-pel
//         direction index, 40..47 block type
// SAD and MAD are below 2^25 (256 luma or 128 chroma differences of int16
// samples).  The finishing wave clears the words it read, so they are zero
// between launches.
struct SearchRec {
    uint64_t w[4];
};
static_assert(sizeof(SearchRec) == 32, "SearchRec layout");

// A MB's current decision in sweep mode (64 bytes): the searches its
// record and prediction come from, written at every re-search and read by
// the final pass.  w[0..2]: the primary search (P: the selected reference;
// B: forward) and w[3..5]: the backward search (B), both in the SearchRec
// word layout; w[6]: bits 0..31 the selection SAD, 32..33 the selected
// reference, 34 bidirectional; w[7]: bits 0..31 integer candidates, 32..47
// axial and 48..63 diagonal sub-pel directions over all searches.
struct Decision {
    uint64_t w[8];
};
static_assert(sizeof(Decision) == 64, "Decision layout");

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
    SearchRec *searches;               // EVXA_MAX_REFS per MB
    int16_t *fpred;                    // MB wavefront, B frames: per MB the forward search's prediction
                                       // (EVXA_MB_SAMPLES), stored before its result; else nullptr
    uint64_t *mvw;                     // per MB: predictor word, zeroed per launch
    int32_t eager;                     // P with several references, alone in its launch:
                                       // every reference's wave tries the selection
    int32_t reach;                     // reach of the search window staged per ticket (1..15)
    // Sweep mode, per MB: the predictor words by sweep parity (2 per MB;
    // parity 0 is zero at the start of a launch, so sweep 0 can tell which
    // words are published), the key of the predictor the current decision
    // was searched with, the predictor the MB's searches of the current sweep
    // use (zero at the start), the last sweep the MB was listed for (zero at
    // the start) and the current decision.  `guess` is the table whose block
    // types and MVs stand in for the neighbours' records in sweep 0 (the
    // temporal table of a P frame), or nullptr for none.
    uint64_t *sw_words;
    uint32_t *sw_key;
    uint64_t *sw_pred;                 // bit 63 | the predictor (x low, y high) the MB's searches use
    uint32_t *sw_listed;
    Decision *sw_dec;
    const evxa_block_record *guess;
};

// Work-list entry: job (10 bits) | reference (2 bits) | MB index (20 bits).
constexpr int kMbBits = 20;
constexpr uint32_t kMbMask = (1u << kMbBits) - 1;
constexpr int kRefShift = 20;
constexpr int kJobShift = 22;
constexpr int kMaxJobs = 1 << (32 - kJobShift);

// Work lists.  The work of a launch is split into up to kLists lists, one
// per XCD of the MI355X: the dispatcher places workgroup w on XCD w % 8, so
// workgroup w takes its tickets from list w % n_lists and each XCD's L2 holds
// the pictures of its own share of rows.  Every list is ordered by
// anti-diagonal, which keeps the wavefront deadlock-free whatever the
// placement.  The work buffer starts with a header of kWorkHeader words:
// [0] n_lists, [1 + k] first ticket of list k (k = 0 .. n_lists, the last
// entry closes the final list); tickets follow.
constexpr int kLists = 8;
constexpr int kWorkHeader = 32;
// Sweep mode adds to the header: [kWorkFinal + k] first ticket of final-pass
// list k (k = 0 .. n_lists; the final pass has one ticket per MB), the
// number of sweep-0 lists that hold tickets and the sweep counter blocks
// available.
constexpr int kWorkFinal = 12;
constexpr int kWorkNonEmpty = 24;
constexpr int kWorkMaxSweeps = 25;

// Counter block of a launch: [1] error flags; the next-ticket counter of
// list k at kCounterList + k * kCounterListStride (own 256-byte block each),
// and in sweep mode, in the same block, its count of finished sweep-0
// tickets (+ kCounterDone, a line of its own).
constexpr int kCounterError = 1;
constexpr int kCounterList = 64;
constexpr int kCounterListStride = 64;
constexpr int kCounterDone = 32;
constexpr int kCounterCheck1 = 48;      // next chunk of sweep 1's check over the list's MBs
constexpr int kCounterWords = kCounterList + kLists * kCounterListStride;

// Sweep counters: block s (kSweepStride words) per sweep.  Sweeps s >= 1
// have a check phase (the MBs listed for the sweep, 64 per claim, one lane
// each; sweep 1 checks every MB, in the final pass's lists) and a search
// phase (the tickets of the MBs whose predictor changed).
// Both phases' lists are split into kSubLists sub-lists, filled by the
// workgroups whose index is k modulo kSubLists and taken by any; sub-list k's
// counters sit at k * kSubStride: [kSubChunkHead] next check chunk,
// [kSubCheckTail] MBs listed, [kSubSearchTail] search tickets,
// [kSubDesignated] MBs searched, and on a line of their own
// [kSubSearchHead] next search ticket, [kSubDone] MBs finished.  Then, each
// on its own 128-byte line: check phase finished (flag), check chunks
// finished, search sub-lists finished (sweep 0: static lists finished), the
// sweep's finished flag with its end time and (block 0) the launch's start
// time, and the MBs the sweep searched.
constexpr int kSubLists = 8;
constexpr int kSubStride = 64;
constexpr int kSubChunkHead = 0;
constexpr int kSubCheckTail = 1;
constexpr int kSubSearchTail = 2;
constexpr int kSubDesignated = 3;
constexpr int kSubSearchHead = 32;
constexpr int kSubDone = 33;
constexpr int kSwFlagC = kSubLists * kSubStride;   // 512
constexpr int kSwDoneC = kSwFlagC + 32;
constexpr int kSwSubDone = kSwFlagC + 64;
constexpr int kSwFlag = kSwFlagC + 96;
constexpr int kSwTime = kSwFlag + 2;    // s_memrealtime (100 MHz), 64-bit
constexpr int kSwStart = kSwFlag + 4;   // block 0: the first ticket's claim, 64-bit
constexpr int kSwSearched = kSwFlagC + 128;
constexpr int kSweepStride = kSwFlagC + 160;
constexpr int kCheckChunk = 64;         // listed MBs per check claim
constexpr int kMaxClaimBatch = 16;      // tickets per claim, at most

// Error flag bits.
constexpr int32_t kErrSearchFailed = 1;
constexpr int32_t kErrWaitTimeout = 2;
constexpr int32_t kErrSweepLimit = 4;

// How sweep 0 reads its neighbours' words: the guess table only (the
// deterministic Jacobi iteration), or also words published earlier in sweep
// 0, waiting up to a bounded time for them (a timing-dependent path to the
// same fixpoint).
constexpr int32_t kSweepReadsGuess = 1;
constexpr int32_t kSweepReadsFresh = 2;

// Sweep-mode resources of a launch beyond the counter block.  The dynamic
// lists come in two parities (sweep s uses s & 1), each of kSubLists
// sub-lists: check entries (one per listed MB, cap_check each) and search
// tickets (cap_search each).  Entries are zero between uses.
struct SweepArgs {
    uint64_t *check;
    uint64_t *search;
    int32_t cap_check;
    int32_t cap_search;
    int32_t max_sweeps;   // sweep counter blocks available
    int32_t *sctr;        // sweep counter blocks
    int32_t reads;        // kSweepReadsGuess or kSweepReadsFresh
    int32_t wait_ticks;   // fresh reads: longest wait, s_memrealtime ticks (10 ns)
    int32_t batch;        // tickets per claim (sweep 0, searches, final pass)
};

// Kernel launchers (evxa_kernels.hip).
hipError_t launch_reset(int32_t *words, int32_t n, hipStream_t st);
// `waves` (1 or 2) per workgroup; `grid` workgroups.
hipError_t launch_analyze(const DevJob *jobs, const uint32_t *work, int32_t *counters,
                          int32_t grid, int32_t waves, hipStream_t st);
// The same work in sweep mode (fixpoint of the MV predictor, then a final
// pass of prediction, DCT and quantisation per MB).
hipError_t launch_sweep(const DevJob *jobs, const uint32_t *work, int32_t *counters,
                        const SweepArgs &sa, int32_t grid, int32_t waves, int32_t n_final,
                        hipStream_t st);
// Reconstruction of the MBs of `job` (its records and prediction sources)
// from `levels` into the picture `dst`; grows the spread range record *wide
// (clear it first) by the stored samples.
hipError_t launch_recon(const DevJob &job, const int16_t *levels, const DevPic &dst,
                        int32_t *wide, hipStream_t st);
// The aprons of a picture's three planes (interior origins, luma w x h).
hipError_t launch_apron3(int16_t *y0, int16_t *u0, int16_t *v0, int32_t py, int32_t pc, int32_t w,
                         int32_t h, hipStream_t st);
// Folds a spread range record into a picture's range word.
hipError_t launch_fold_range(const int32_t *spread, int32_t *range, hipStream_t st);
// Grows the spread range record *wide by the samples of the w x h plane.
hipError_t launch_range(const int16_t *plane0, int32_t pitch, int32_t w, int32_t h, int32_t *wide,
                        hipStream_t st);
// The in-loop deblocking filter of a frame (luma w x h, MB descriptors in
// `records`) from the unfiltered planes src (pitches sp) into the picture
// planes dst (pitches dp, distinct memory, with their kApron aprons, which
// it refreshes); grows the spread range record *wide (clear it
// first) by the stored samples.  w and h are multiples of 16.
hipError_t launch_deblock(int16_t *const src[3], const int32_t sp[3], int16_t *const dst[3],
                          const int32_t dp[3], int32_t w, int32_t h,
                          const evxa_block_record *records, int32_t *wide, hipStream_t st);
// Copies a picture's three plane interiors (luma w x h) to packed planes at
// `out` (*wide is not used).
hipError_t launch_copy3(const int16_t *y, const int16_t *u, const int16_t *v, int32_t py, int32_t pc,
                        int16_t *out, int32_t w, int32_t h, int32_t *wide, hipStream_t st);
// SAO decisions of a frame (sao.cpp:378-590): the unfiltered reconstruction
// (packed luma and U planes), the frame's source picture, lambda; `scratch`
// holds 8 words per MB, `words` (one per MB) starts at zero; decisions go to
// `table`; counters[1] (zero at the start) gets error flags (1: a wait
// timed out, 2: an SAO chain deeper than three blocks).
hipError_t launch_sao_decide(const int16_t *rec_y, const int16_t *rec_u, const DevPic &src,
                             int32_t w, int32_t h, float lambda, uint32_t *scratch,
                             uint64_t *words, evxa_sao_info *table, int32_t *counters,
                             hipStream_t st);
// SAO of a frame per `table` from the unfiltered packed planes `rec` into
// the picture planes `dst` (pitches dp); grows the spread range record *wide
// (clear it first) by the stored samples and error flags into *error.
hipError_t launch_sao_apply(int16_t *const rec[3], int16_t *const dst[3], const int32_t dp[3],
                            int32_t w, int32_t h, const evxa_sao_info *table, int32_t *wide,
                            int32_t *error, hipStream_t st);
hipError_t launch_interp_blocks(DevPic pic, int32_t w, int32_t h, int32_t n, const int32_t *params,
                                int16_t *out, hipStream_t st);
hipError_t launch_search_blocks(DevPic src, DevPic ref, int32_t w, int32_t h, int32_t quality,
                                float lambda, int32_t n, const int32_t *params,
                                evxa_search_result *out, hipStream_t st);
// calculate_intra_prediction of n MBs of `src` over `pic` (the picture being
// reconstructed); params holds (x, y) per MB.
hipError_t launch_intra_blocks(DevPic src, DevPic pic, int32_t w, int32_t h, int32_t quality,
                               float lambda, int32_t n, const int32_t *params,
                               evxa_search_result *out, hipStream_t st);
hipError_t launch_dct_quant_blocks(int32_t n, const int16_t *src, const int16_t *pred,
                                   const int32_t *quality, int16_t *coeffs, int16_t *levels,
                                   int32_t *qp, int32_t *variance, hipStream_t st);
// Allocation-free analysis: bytes of the per-GPU scratch array, its device
// address (to be taken once per device, with that device current), the
// copy of a job descriptor into it and the wavefront work list of one frame
// (header + tickets, one list).
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;
hipError_t noalloc_scratch(void **ptr);
hipError_t launch_set_job(const DevJob &job, DevJob *dst, hipStream_t st);
hipError_t launch_wave_list(int32_t wb, int32_t hb, int32_t n_refs, uint32_t *work, hipStream_t st);
// One frame of an in-loop encode launch (encode_slice): `job` (type 0 I, 1
// P with one reference, 2 B; records, coefficients, levels and the per-MB
// words `mvw`, zero at the start), `pic` the picture being reconstructed
// (written in place, apron included) and `spread` its spread range record
// (cleared first).
struct EncJob {
    DevJob job;
    DevPic pic;
    int32_t *spread;
};
// Coefficient candidates each encode workgroup keeps in `coscratch`
// (kEncodeCands x EVXA_MB_SAMPLES int16 per workgroup).
constexpr int32_t kEncodeCands = 6;
// The encode of the frames in `jobs` (device memory): `order` lists
// n_tickets entries job << kMbBits | MB in dependency order; counters[0]
// (zero at the start) is the next ticket, counters[kCounterError] error
// flags.
hipError_t launch_encode(const EncJob *jobs, const uint32_t *order, int32_t n_tickets,
                         int32_t *counters, int16_t *coscratch, int32_t grid, hipStream_t st);
// Workgroups of the encode kernel that fit the device at once.
int32_t encode_resident_groups(int32_t device);
// Rows of MBs an encode MB may lead the row below it by (its wavefront lag).
constexpr int32_t kEncodeLag = 3;
// Workgroups of the analysis kernel (wavefront or sweep mode) with `waves`
// waves each that fit the device at once.
int32_t analyze_resident_groups(int32_t device, int32_t waves, bool sweep);

}  // namespace evxa
