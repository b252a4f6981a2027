/*
 * evxa.h -- C ABI of the MI355X (gfx950) inter-analysis path for the
 * Chicago/EVX3 video encoder: integer step search, 6-tap half/quarter-pel
 * refinement, 8x8 forward DCT, adaptive QP and inter quantisation, run as
 * hand-written HIP kernels on device-resident pictures.
 *
 * The reference has no FFI; its path sits behind encoder-internal C++ free
 * functions (namespace evx).  Each entry point below names the reference
 * interface it replaces (paths relative to the reference's src/).  The ABI
 * uses plain pointers, sizes and int status codes; no C++ exception crosses
 * it.  Status values equal the reference's evx_status codes (base.h:166-185).
 *
 * Threading: a context owns one HIP stream; calls on one context are
 * serialised by an internal mutex and execute in submission order.
 */
#ifndef EVXA_H
#define EVXA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: same values as base.h:166-185. */
#define EVXA_SUCCESS 0
#define EVXA_ERROR_INVALIDARG 1
#define EVXA_ERROR_NOTIMPL 2
#define EVXA_ERROR_OUTOFMEMORY 3
#define EVXA_ERROR_HARDWAREFAIL 5
#define EVXA_ERROR_CAPACITY_LIMIT 7
#define EVXA_ERROR_INVALID_RESOURCE 8
#define EVXA_ERROR_OPERATION_TIMEDOUT 9
#define EVXA_ERROR_EXECUTION_FAILURE 10

#define EVXA_FRAME_I 0
#define EVXA_FRAME_P 1
#define EVXA_FRAME_B 2
#define EVXA_MAX_REFS 4
#define EVXA_MB_SAMPLES 384 /* 16x16 Y + 8x8 U + 8x8 V, in that order */

/* Block type bits (types.h:72-91). */
#define EVXA_BLOCK_INTRA 0x1
#define EVXA_BLOCK_MOTION 0x2
#define EVXA_BLOCK_COPY 0x4

/* Per-macroblock analysis record, 64 bytes.  Fields up to `variance` carry
 * evx_block_desc (common.h:87-115); `sad` is the SAD the block selection
 * used (encode.cpp:194-208, 399-441); mad/cost are the primary search's
 * best_mad and fp32 best_cost bits (motion.cpp:53-66); *_b are the backward
 * search's for bidirectional blocks; n_int / n_axial / n_diag count the
 * integer candidates and sub-pel directions evaluated (work accounting);
 * flags bit0 = residual transformed, bit1 = bidirectional average. */
typedef struct evxa_block_record {
    uint8_t block_type;
    uint8_t prediction_mode;
    uint8_t prediction_target;
    uint8_t prediction_target_b;
    int16_t motion_x, motion_y;
    int16_t motion_x_b, motion_y_b;
    uint8_t sp_pred, sp_amount, sp_index, sp_pred_b;
    uint8_t sp_amount_b, sp_index_b, q_index, intra_mode;
    int16_t variance;
    int16_t reserved1;
    int32_t sad;
    int32_t mad;
    uint32_t cost_bits;
    int32_t sad_b, mad_b;
    uint32_t cost_b_bits;
    uint32_t n_int;
    uint16_t n_axial, n_diag;
    uint32_t flags;
    uint32_t reserved2;
} evxa_block_record;

/* Sample adaptive offset parameters of one plane of an MB (evx_sao_params,
 * sao.h:62-67, padded to 8 bytes).  type: 0 off, 1 edge offset, 2 band
 * offset; eo_class: 0 horizontal, 1 vertical, 2 135 degrees, 3 45 degrees;
 * band_position: first of the 4 bands of 8 levels that get offsets. */
typedef struct evxa_sao_params {
    uint8_t type, eo_class, band_position, reserved;
    int8_t offsets[4];
} evxa_sao_params;

/* The SAO entry of one MB (evx_sao_info, sao.h:69-74), 20 bytes. */
typedef struct evxa_sao_info {
    uint8_t merge_left, merge_above, reserved[2];
    evxa_sao_params luma, chroma;
} evxa_sao_info;

/* Result of one inter search (motion.cpp:567-679 plus its internal state). */
typedef struct evxa_search_result {
    int32_t best_x, best_y;
    int32_t sad, mad, ssd;
    uint32_t cost_bits;
    int32_t sp_enabled, sp_amount, sp_index;
    int32_t block_type;
    int32_t motion_x, motion_y;
    int32_t n_int, n_axial, n_diag;
} evxa_search_result;

/* One frame of analysis work on a stream. */
typedef struct evxa_frame_desc {
    int32_t type;                 /* EVXA_FRAME_P or EVXA_FRAME_B */
    int32_t quality;              /* frame QP 0..63 (evx3enc.cpp:64-82: q <= 16 maps 1:1) */
    int32_t src_slot;             /* picture slot holding the current frame */
    int32_t n_refs;               /* P: 1..4 (most recent first); B: 2 (fwd, bwd) */
    int32_t ref_slots[EVXA_MAX_REFS];
    uint8_t target_ids[EVXA_MAX_REFS]; /* written to prediction_target(_b) */
    int32_t temporal_mvp;         /* P: use the stream's temporal table if one exists */
    int32_t update_temporal;      /* P: this frame's table becomes the temporal table */
} evxa_frame_desc;

/* One frame of the in-loop encode (evxa_encode_frame). */
typedef struct evxa_encode_desc {
    int32_t type;                 /* EVXA_FRAME_I, EVXA_FRAME_P (one reference) or
                                     EVXA_FRAME_B (forward, backward) */
    int32_t quality;              /* frame QP 0..63 */
    int32_t src_slot;             /* the current frame */
    int32_t dst_slot;             /* the picture-buffer slot reconstructed into; it holds
                                     its previous content on entry, as in the reference */
    int32_t ref_slots[2];         /* P: [0]; B: [0] forward, [1] backward */
    uint8_t target_ids[2];        /* written to prediction_target(_b): the decoder's slot
                                     of each reference, so a B frame's two are equal only
                                     when ref_slots are (else EVXA_ERROR_INVALIDARG) */
    uint8_t reserved[2];
    int32_t temporal_mvp;         /* P: use the stream's temporal table if one exists */
    int32_t update_temporal;      /* I and P: this frame's records become the temporal
                                     table (the reference copies every non-B frame's
                                     block table, encode.cpp:990-997) */
} evxa_encode_desc;

typedef struct evxa_context evxa_context;
typedef struct evxa_stream evxa_stream;

/* ---- lifecycle (replaces evx_context set-up, common.cpp:106-202) ------ */

/* Create a context on `device_ordinal` that admits streams while their
 * device footprint stays within `hbm_budget_bytes` (0 = 90% of free HBM). */
int evxa_create(int32_t device_ordinal, uint64_t hbm_budget_bytes, evxa_context **out);
int evxa_destroy(evxa_context *ctx);
/* Last error message of the calling thread ("" if none). */
const char *evxa_last_error(void);
/* ABI version (major * 100 + minor). */
int32_t evxa_version(void);
/* Bytes of HBM a stream of this size would hold (admission accounting). */
uint64_t evxa_stream_bytes(int32_t width, int32_t height, int32_t n_slots);
/* HBM currently charged to streams of this context. */
uint64_t evxa_context_bytes(const evxa_context *ctx);
/* fp32 lambda of common.h:264-266 for `quality` 0..63, as the kernels use
 * it: the reference build's own values, compiled in as a table. */
float evxa_lambda(int32_t quality);

/* ---- streams and picture slots (replaces evx_cache_bank's
 *      prediction_cache, common.h:124-137) ---------------------------- */

/* width/height: luma size, multiples of 16 (frames aligned with zero rows,
 * image.cpp:89).  Fails with EVXA_ERROR_CAPACITY_LIMIT when the stream does
 * not fit the context's HBM budget. */
int evxa_stream_open(evxa_context *ctx, int32_t width, int32_t height, int32_t n_slots,
                     evxa_stream **out);
int evxa_stream_close(evxa_stream *s);
/* Copy a host picture (int16 planes, strides in elements) into `slot`. */
int evxa_upload_picture(evxa_stream *s, int32_t slot, const int16_t *y, const int16_t *u,
                        const int16_t *v, int32_t stride_y, int32_t stride_c);

/* Copy picture slot `slot` (luma w x h, chroma w/2 x h/2 interiors) to host
 * planes with the given strides (elements); blocks until done.  With
 * evxa_reconstruct it returns the decoded picture. */
int evxa_download_picture(evxa_stream *s, int32_t slot, int16_t *y, int16_t *u, int16_t *v,
                          int32_t stride_y, int32_t stride_c);
/* Forget the temporal MV table (next P frame has no temporal candidate). */
int evxa_stream_reset_temporal(evxa_stream *s);

/* ---- frame jobs (replaces the classify_block / classify_block_bidir /
 *      encode_block inter analysis, encode.cpp:168-248, 374-477, 481-739,
 *      without intra candidates, RDO and trellis) --------------------- */

/* Enqueue one frame: one kernel launch over the frame's MB wavefront. */
int evxa_analyze_frame(evxa_stream *s, const evxa_frame_desc *desc);
/* evxa_analyze_frame for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation): the same
 * outputs, from the MB wavefront, with the launch's scratch -- job
 * descriptor, work list, counters, predictor words -- in one per-GPU device
 * array of 1 MiB held for the life of the process, instead of buffers the
 * context allocates on first use.  It calls no allocating routine.  Frames
 * whose scratch does not fit (about 3840x2160 with 4 references is the
 * largest) fail with EVXA_ERROR_CAPACITY_LIMIT.  Calls on one GPU share the
 * array: calls that overlap on one GPU (from several contexts) race; calls
 * on one context are serialised.  No counterpart in the reference. */
int evxa_analyze_frame_noalloc(evxa_stream *s, const evxa_frame_desc *desc);
/* Enqueue frames of distinct streams as one fused persistent launch. */
int evxa_analyze_batch(evxa_context *ctx, evxa_stream *const *streams,
                       const evxa_frame_desc *descs, int32_t n);
/* Block until the context's queue is idle; reports device-side failures. */
int evxa_synchronize(evxa_context *ctx);
/* Block until every job queued so far that writes this stream's slots or
 * outputs (analysis alone, in a batch, a graph replay or a server launch;
 * reconstruction, deblocking, SAO) has completed, without waiting for other
 * streams' work: the wait(event) of a submitted job (the reference's frame
 * call returns only when its row threads have joined, encode.cpp:934-935).  Device-side
 * failures of the analysis are reported by evxa_synchronize,
 * evxa_read_results and server completions.  The stream must stay open
 * until the wait returns. */
int evxa_stream_wait(evxa_stream *s);
/* Non-blocking: *done = 1 when evxa_stream_wait would return at once. */
int evxa_stream_query(evxa_stream *s, int32_t *done);
/* Device time (ms) of the last analysis launch, from HIP events around it. */
int evxa_last_launch_ms(evxa_context *ctx, float *ms);
/* Waits for the queue, then reports the device times (ms, HIP events around
 * each analysis kernel) of up to `max` launches not reported before, oldest
 * first; *n_out receives the count.  Graph replays are not included. */
int evxa_launch_times(evxa_context *ctx, float *ms, int32_t max, int32_t *n_out);
/* Copy the last analysed frame's outputs to host memory (any may be NULL):
 * one record per MB in raster order, 384 coefficients and 384 pre-trellis
 * levels per MB (zero for copy blocks).  Blocks until they are available. */
int evxa_read_results(evxa_stream *s, evxa_block_record *records, int16_t *coeffs,
                      int16_t *levels);

/* Rebuild the last analysed frame of `s` into picture slot `dst_slot`, as
 * the reference decoder rebuilds inter MBs (decode_block, decode.cpp:31-433,
 * for the inter block types): the prediction each record names plus, for
 * transformed MBs, the levels inverse-quantised (quantize.cpp:427-449) and
 * inverse-transformed onto it (macroblock.h:290-295), stored as int16.
 * `levels`: n_mbs x 384 int16 in MB order on the host (for instance after
 * a trellis pass on the CPU), or NULL for the stream's own pre-trellis
 * levels.  The slot then serves as a reference like an uploaded picture,
 * so the decoded picture buffer stays on the device.  dst_slot must not be
 * a reference of that frame, and its references need distinct target ids.
 * After evxa_encode_frame, which reconstructs its frame itself, it fails
 * with EVXA_ERROR_INVALID_RESOURCE.
 * Replaces the encoder's per-MB reconstruction call, decode_block after
 * quantisation (encode.cpp:806, 918 in the row workers). */
int evxa_reconstruct(evxa_stream *s, int32_t dst_slot, const int16_t *levels);

/* The in-loop deblocking filter of picture slot `slot`, in place, with the
 * result of the reference's deblock_image_filter (deblock.cpp:222-306: luma
 * on 16x16 blocks, chroma on 8x8 at qp + 2, edges every 8 samples, strength
 * from the block types, thresholds from the averaged q_index).  `records`:
 * n_mbs block records on the host (block_type and q_index are read), or
 * NULL for the stream's last analysed frame.  Replaces the encoder's
 * deblock_image_filter call after the frame's reconstruction
 * (encode.cpp:968) and the decoder's (decode.cpp:583). */
int evxa_deblock(evxa_stream *s, int32_t slot, const evxa_block_record *records);

/* evxa_reconstruct followed by evxa_deblock with the frame's own records,
 * with the same result in `dst_slot`, in one call: the reconstruction is
 * written straight to the filter's input, which saves the picture copy
 * evxa_deblock makes.  The encoder's decode_block calls then its
 * deblock_image_filter (encode.cpp:806, 918, 968); the decoder's
 * decode_slice then deblock_image_filter (decode.cpp:466-583). */
int evxa_reconstruct_deblocked(evxa_stream *s, int32_t dst_slot, const int16_t *levels);

/* The encoder's sample adaptive offset of the frame in picture slot `slot`
 * (its deblocked reconstruction) against that frame's source picture in
 * `src_slot`, at `quality` (lambda of common.h:264-266): per MB the best
 * luma offsets and the best chroma offsets (chosen on U), band or edge,
 * by cost SSD + lambda * bits, then a merge with the left or upper MB's
 * parameters when that costs less, all in the reference's raster order;
 * then the offsets are applied to `slot` in place.  The n_mbs decisions
 * are copied to `table` when it is not NULL.  Replaces sao_encode_frame
 * (sao.cpp:378-590) after deblocking (encode.cpp:974). */
int evxa_sao(evxa_stream *s, int32_t slot, int32_t src_slot, int32_t quality,
             evxa_sao_info *table);

/* Apply a given SAO table (n_mbs entries on the host) to picture slot
 * `slot` in place, as the decoder's sao_filter_image does after deblocking
 * (sao.cpp:98-148; decode.cpp:589). */
int evxa_sao_filter(evxa_stream *s, int32_t slot, const evxa_sao_info *table);

/* The in-loop encode of one frame, as the reference's encode_slice
 * (encode.cpp:751-813) runs it: per MB in raster-order semantics, the mode
 * decision of classify_block (I and P, encode.cpp:168-370: the intra-frame
 * block search over the picture being reconstructed, motion.cpp:485-565;
 * for P the inter search with the temporal table; the directional intra
 * mode, intra_pred.cpp:217-326; RDO over the SAD winner, the plain intra
 * block and the three mode candidates, encode.cpp:128-166, 274-348) or of
 * classify_block_bidir (B, encode.cpp:372-476), encode_block with the
 * trellis (encode.cpp:479-748; quantize.cpp:454-464, 615-680), then
 * decode_block into `dst_slot` (decode.cpp:31-433).  One launch; MBs run in
 * a wavefront whose lag keeps every MB's reads of `dst_slot` equal to the
 * sequential encoder's.  Outputs: the reconstruction in `dst_slot` (before
 * deblocking and SAO, which evxa_deblock and evxa_sao then apply), and,
 * through evxa_read_results, one descriptor per MB (q_index and variance as
 * encode_block sets them; sad the SAD-stage best; cost_bits the RDO cost of
 * P frames; flags bit0 residual coded, bit1 bidirectional, bit2 encode),
 * the coefficients and the post-trellis levels (zero for copies).
 * dst_slot must differ from the source and the references; frames are at
 * least 48 x 48.  Replaces the encoder's encode_slice / encode_slice_wpp
 * call in engine_encode_frame (encode.cpp:958-962) and its row workers
 * (encode.cpp:817-941). */
int evxa_encode_frame(evxa_stream *s, const evxa_encode_desc *desc);
/* evxa_encode_frame for frames of `n` distinct streams in one launch: their
 * MBs interleave in one dependency order (each frame's wavefront runs as it
 * would alone), so the device holds the MBs of many frames at once.  Each
 * stream's outputs are those of evxa_encode_frame; every descriptor is
 * validated before anything is queued.  No counterpart in the reference
 * (its encoder codes one frame at a time); for the many-encoder case of
 * evxa_analyze_batch. */
int evxa_encode_batch(evxa_context *ctx, evxa_stream *const *streams,
                      const evxa_encode_desc *descs, int32_t n);

/* ---- graph-captured steps (hipGraph of a fixed batch composition) ---- */

/* Capture the launch of `n` frame descriptors into a replayable graph.
 * Replays re-run the same slots, qualities and temporal-table roles. */
int evxa_graph_create(evxa_context *ctx, evxa_stream *const *streams,
                      const evxa_frame_desc *descs, int32_t n, int32_t *graph_id);
int evxa_graph_launch(evxa_context *ctx, int32_t graph_id);
int evxa_graph_destroy(evxa_context *ctx, int32_t graph_id);

/* ---- per-block entries (parity tests and encoder drop-in) ------------ */

/* interpolate_subpixel_6tap (interp.cpp:166-202) at n integer bases of a
 * slot's picture; out receives n * 384 samples. */
int evxa_interp_blocks(evxa_stream *s, int32_t slot, int32_t n, const int32_t *xs,
                       const int32_t *ys, const int32_t *dir_x, const int32_t *dir_y,
                       const int32_t *quarter, int16_t *out);
/* perform_inter_search (motion.cpp:567-679) for n MBs of src_slot against
 * ref_slot, each with the predictor compute_mv_predictor would supply. */
int evxa_search_blocks(evxa_stream *s, int32_t src_slot, int32_t ref_slot, int32_t quality,
                       int32_t n, const int32_t *xs, const int32_t *ys, const int16_t *pred_x,
                       const int16_t *pred_y, evxa_search_result *out);
/* calculate_intra_prediction (motion.cpp:485-565) for n MBs: the intra-frame
 * block search of the source MB at (xs[k], ys[k]) of `src_slot` over the
 * picture in `pic_slot` as the picture being reconstructed (the
 * reference's own picture-buffer slot, motion.cpp:510-515: rows above and
 * columns left of the MB hold the reconstruction; the search never reads
 * below row y - 16 to the right of column x - 16).  Results as
 * evxa_search_result: block_type has EVXA_BLOCK_INTRA, best_x/best_y the
 * matched block in pic_slot. */
int evxa_intra_search_blocks(evxa_stream *s, int32_t src_slot, int32_t pic_slot, int32_t quality,
                             int32_t n, const int32_t *xs, const int32_t *ys,
                             evxa_search_result *out);
/* sub_transform_macroblock + query_block_quantization_parameter +
 * compute_block_variance2 + quantize_macroblock (inter) for n residual
 * pairs (macroblock.h:272-277, quantize.cpp:97-118, 372-397;
 * analysis.h:253-275).  src/pred/coeffs/levels hold n * 384 samples. */
int evxa_dct_quant_blocks(evxa_context *ctx, int32_t n, const int16_t *src,
                          const int16_t *pred, const int32_t *quality, int16_t *coeffs,
                          int16_t *levels, int32_t *qp, int32_t *variance);

/* ---- launch shape (no counterpart in the reference: its analysis runs on
 *      CPU threads, encode.cpp:817-941).  Set by the autotuner. ---------- */

/* How a launch resolves the MV predictor's dependency on the left, top and
 * top-right MBs' final records (compute_mv_predictor, motion.cpp:271-346).
 * Both give the reference's sequential records bit for bit.
 *   WAVEFRONT: MBs run in anti-diagonal order, each waiting for its
 *              neighbours' published results (one search per MB and
 *              reference; latency bound by the anti-diagonal chain).
 *   SWEEP:     Jacobi sweeps to the unique fixpoint of "each MB's record is
 *              its searches under its neighbours' records' predictor":
 *              sweep 0 searches every MB with the temporal table standing in
 *              for the neighbours, each later sweep searches again only MBs
 *              whose predictor changed, then a final pass runs prediction,
 *              DCT and quantisation once per MB.  No MB waits on another. */
#define EVXA_MODE_WAVEFRONT 1
#define EVXA_MODE_SWEEP 2

/* Order of sweep 0's tickets: each frame's MBs in raster order, or in the
 * wavefront's anti-diagonal order (with sweep_reads 2, most MBs then find
 * their neighbours' MVs decided, and a bounded wait does the rest). */
#define EVXA_SWEEP_ORDER_RASTER 1
#define EVXA_SWEEP_ORDER_DIAGONAL 2

typedef struct evxa_tuning {
    int32_t waves_per_search;     /* 1 or 2 waves search one (MB, reference) pair (0: keep) */
    int32_t max_groups;           /* workgroups per launch (0: all that fit the device) */
    int32_t lists;                /* work lists per launch, 1..8 (0: one per XCD for batches
                                     and sweeps) */
    int32_t cap_pct;              /* tickets per list and diagonal, % of its workgroups
                                     (wavefront, and sweep 0 in diagonal order; 0: 60) */
    int32_t mode;                 /* EVXA_MODE_WAVEFRONT or EVXA_MODE_SWEEP (0: keep) */
    int32_t sweep_reads;          /* sweep 0's neighbour MVs: 1 the temporal table's only (a
                                     deterministic iteration), 2 also those already decided
                                     in sweep 0 (fewer re-searches; 0: keep) */
    int32_t sweep_wait_us;        /* with sweep_reads 2: longest wait for a neighbour's
                                     sweep-0 MV, microseconds (-1: keep) */
    int32_t sweep_order;          /* EVXA_SWEEP_ORDER_* (0: keep) */
    int32_t claim_batch;          /* sweep mode: consecutive tickets a workgroup claims at
                                     once, 1..16 (default 1; 0: keep) */
    int32_t window_reach;         /* reach of the reference window each (MB, reference)
                                     search stages in LDS around the MB, 1..15 samples
                                     (default 15; 0: keep).  Levels whose candidates leave
                                     it load from HBM (step 16) or stage their own. */
} evxa_tuning;

int evxa_set_tuning(evxa_context *ctx, const evxa_tuning *t);
/* Current tuning; *resident_groups (optional) receives the workgroups of
 * the current shape that fit the device at once. */
int evxa_get_tuning(evxa_context *ctx, evxa_tuning *t, int32_t *resident_groups);

/* The sweeps of the context's last sweep-mode launch (blocks until it is
 * done): for sweep k < *n_sweeps (at most `max` reported), listed[k] the MBs
 * whose predictor it checked (sweep 0: all), searched[k] the MBs it searched
 * and, if done_us is not NULL, done_us[k] the microseconds from the launch's
 * first ticket to the end of that sweep's searches (the device's constant
 * 100 MHz clock; -1 for a sweep whose checks changed no predictor).
 * *n_sweeps is 0 after a wavefront launch.  Diagnostic; no counterpart in
 * the reference. */
int evxa_sweep_stats(evxa_context *ctx, int32_t *listed, int32_t *searched, float *done_us,
                     int32_t max, int32_t *n_sweeps);

/* ---- serving: continuous batching (no counterpart in the reference, whose
 *      encoder analyses one frame at a time inside engine_encode_frame,
 *      encode.cpp:945; this is the multi-stream request path) ------------ */

typedef struct evxa_server evxa_server;

/* One finished frame.  Times are evxa_clock() seconds: t_arrival as given at
 * submit, t_launch when its fused launch was enqueued, t_done when the
 * dispatcher saw the launch complete.  device_ms is the launch's kernel
 * time; batch_frames the number of frames fused into it. */
typedef struct evxa_completion {
    uint64_t tag;
    int32_t status;
    int32_t batch_frames;
    double t_arrival;
    double t_launch;
    double t_done;
    float device_ms;
    int32_t reserved;
} evxa_completion;

/* Monotonic clock (seconds) shared with the dispatcher. */
double evxa_clock(void);
/* Start a dispatcher with `max_inflight` (1..16) concurrent launches, each
 * fusing up to `max_frames` ready frames. */
int evxa_server_create(evxa_context *ctx, int32_t max_inflight, int32_t max_frames,
                       evxa_server **out);
/* Queue one frame (as for evxa_analyze_frame).  Frames of one stream run in
 * submission order, one at a time; t_arrival <= 0 means now. */
int evxa_server_submit(evxa_server *srv, evxa_stream *s, const evxa_frame_desc *desc,
                       uint64_t tag, double t_arrival);
/* Move up to `max` completions to `out` (non-blocking); *n_out receives the
 * count.  Returns the first dispatcher error, if any. */
int evxa_server_poll(evxa_server *srv, evxa_completion *out, int32_t max, int32_t *n_out);
/* Frames queued and frames in flight. */
int evxa_server_pending(evxa_server *srv, int32_t *queued, int32_t *inflight);
/* Run every queued frame to completion, then stop the dispatcher. */
int evxa_server_destroy(evxa_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* EVXA_H */
