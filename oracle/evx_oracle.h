/*
 * evx_oracle.h -- CPU restatement of the Chicago/EVX3 encoder's inter analysis
 * path (integer step search, 6-tap sub-pel refinement, 8x8 forward DCT,
 * adaptive QP and inter quantisation), written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (chicago_amd/libevxa.so) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference's src/ directory).
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - interpolation, SAD/MAD/SATD, Hadamard, DCT, variance2, QP, quantisation,
 *     lambda, MV predictor, inter search, bidir prediction and the composed
 *     frame drivers are compared against golden vectors written by the
 *     reference itself (oracle/_ref, tests/golden/make_golden.py) and, where
 *     oracle/_ref is built, against the reference on fresh random inputs;
 *   - closed forms and the reference's own exact unit-test cases
 *     (test_motion.cpp SAD/MAD/SATD/Hadamard/predictor, test_interp.cpp
 *     constant-in/constant-out, test_transform.cpp and test_quantize.cpp
 *     zero-in/zero-out) are restated as independent tests.
 */
#ifndef EVX_ORACLE_H
#define EVX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EVO_MB_SAMPLES 384           /* 16x16 Y + 8x8 U + 8x8 V */
#define EVO_MAX_REFS 4

/* Block type bits, types.h:72-91. */
#define EVO_BT_INTRA  0x1
#define EVO_BT_MOTION 0x2
#define EVO_BT_COPY   0x4

/* Directional intra modes, intra_pred.h:42-49: DC, vertical, horizontal,
 * diagonal down-left, diagonal down-right, vertical-right, horizontal-down;
 * 7 = none (a plain intra block, transformed without prediction). */
#define EVO_INTRA_NONE 7

/* A planar int16 4:2:0 picture.  width/height are the luma size (multiples
 * of 16, already aligned by the caller with zero rows, image.cpp:89);
 * chroma planes are width/2 x height/2.  Strides are in elements. */
typedef struct evo_planes {
    const int16_t *y;
    const int16_t *u;
    const int16_t *v;
    int32_t width;
    int32_t height;
    int32_t stride_y;
    int32_t stride_c;
} evo_planes;

/* Outcome of one perform_inter_search call (motion.cpp:567-679), plus the
 * candidate counters that price the work in algorithmic operations. */
typedef struct evo_search_result {
    int32_t best_x, best_y;
    int32_t sad, mad, ssd;
    uint32_t cost_bits;          /* fp32 bits of best_cost */
    int32_t sp_enabled, sp_amount, sp_index;
    int32_t block_type;          /* intra=0 | motion<<1 | copy<<2 */
    int32_t motion_x, motion_y;  /* best - (x, y) */
    int32_t n_int;               /* evaluate_motion_candidate calls */
    int32_t n_axial, n_diag;     /* sub-pel directions evaluated */
} evo_search_result;

/* Per-macroblock analysis record (64 bytes).  The fields up to `variance`
 * mirror evx_block_desc (common.h:87-115); the rest carry the search metrics
 * and work counters.  Layout is identical to evxa_block_record. */
typedef struct evo_record {
    uint8_t block_type;
    uint8_t prediction_mode;       /* 0 forward, 1 backward, 2 bi */
    uint8_t prediction_target;
    uint8_t prediction_target_b;
    int16_t motion_x, motion_y;
    int16_t motion_x_b, motion_y_b;
    uint8_t sp_pred, sp_amount, sp_index, sp_pred_b;
    uint8_t sp_amount_b, sp_index_b, q_index;
    uint8_t intra_mode;            /* intra_pred.h:42-49 (encode mode); 0 in analysis mode */
    int16_t variance;
    int16_t reserved1;
    int32_t sad;                   /* SAD used by the final selection */
    int32_t mad;                   /* primary search best_mad */
    uint32_t cost_bits;            /* primary search best_cost bits */
    int32_t sad_b, mad_b;          /* backward search (bidir only) */
    uint32_t cost_b_bits;
    uint32_t n_int;
    uint16_t n_axial, n_diag;
    uint32_t flags;                /* bit0: transformed, bit1: bidir average */
    uint32_t reserved2;
} evo_record;

/* One frame of analysis work. */
typedef struct evo_frame_job {
    int32_t type;                  /* 1 = P (1..4 refs), 2 = B (fwd, bwd) */
    int32_t quality;               /* frame QP (quality <= 16 maps 1:1) */
    evo_planes src;
    int32_t n_refs;                /* P: 1..4; B: 2 (0 fwd, 1 bwd) */
    evo_planes refs[EVO_MAX_REFS];
    uint8_t target_ids[EVO_MAX_REFS];  /* stored in prediction_target */
    const evo_record *prev;        /* temporal MVP table or NULL */
} evo_frame_job;

/* common.h:264-266 */
float evo_lambda(int32_t quality);

/* interp.cpp:166-202 (with 52-164, 18-41, 8-14): sub-pel MB prediction at
 * integer base (bx, by), direction (dir_x, dir_y) in {-1,0,1}^2 \ {0},
 * half (quarter = 0) or quarter pel.  out is 384 samples in MB order. */
void evo_interp_block(const evo_planes *ref, int32_t bx, int32_t by,
                      int32_t dir_x, int32_t dir_y, int32_t quarter,
                      int16_t out[EVO_MB_SAMPLES]);

/* create_macroblock (macroblock.h:82-88) copied into MB order. */
void evo_load_block(const evo_planes *p, int32_t x, int32_t y,
                    int16_t out[EVO_MB_SAMPLES]);

/* analysis.h:44-85 */
void evo_hadamard_4x4(const int16_t *src, int32_t stride, int32_t out[16]);
/* analysis.h:89-116, 119-132, 180-202 on MB-order blocks */
int32_t evo_block_satd(const int16_t *a, const int16_t *b);
int32_t evo_block_sad(const int16_t *a, const int16_t *b);
int32_t evo_block_mad(const int16_t *a, const int16_t *b);

/* motion.cpp:271-346 */
void evo_mv_predictor(const evo_record *table, uint32_t width_in_blocks,
                      uint32_t block_index, const evo_record *prev,
                      int16_t *pred_x, int16_t *pred_y);

/* motion.cpp:567-679 as a pure function of the predictor (pred_mx, pred_my)
 * that compute_mv_predictor would have produced.  src is the MB at (x, y). */
int32_t evo_inter_search(const int16_t src[EVO_MB_SAMPLES],
                         const evo_planes *ref, int32_t x, int32_t y,
                         int32_t quality, int16_t pred_mx, int16_t pred_my,
                         evo_search_result *res);

/* calculate_intra_prediction (motion.cpp:485-565): the intra-frame block
 * search of the MB at (x, y) over `pic`, the picture being reconstructed
 * (rows above and columns left of the MB hold reconstructed samples). */
int32_t evo_intra_search(const int16_t src[EVO_MB_SAMPLES], const evo_planes *pic,
                         int32_t x, int32_t y, int32_t quality, evo_search_result *res);

/* MB prediction for a search result (encode.cpp:696-723). */
void evo_predict_block(const evo_planes *ref, int32_t x, int32_t y,
                       const evo_search_result *r, int16_t out[EVO_MB_SAMPLES]);

/* motion.cpp:778-813: bidirectional average and its Y+U+V SAD. */
int32_t evo_bidir_average(const int16_t src[EVO_MB_SAMPLES],
                          const int16_t fwd[EVO_MB_SAMPLES],
                          const int16_t bwd[EVO_MB_SAMPLES],
                          int16_t avg[EVO_MB_SAMPLES]);

/* transform.cpp:264-284, 435-452 on one 8x8 block. */
void evo_sub_transform_8x8(const int16_t *src, int32_t src_pitch,
                           const int16_t *sub, int32_t sub_pitch,
                           int16_t *dst, int32_t dst_pitch);
/* macroblock.h:272-277 -> transform.cpp:584-594 on MB-order blocks. */
void evo_sub_transform_mb(const int16_t src[EVO_MB_SAMPLES],
                          const int16_t pred[EVO_MB_SAMPLES],
                          int16_t coeffs[EVO_MB_SAMPLES]);
/* analysis.h:253-275 over the 16x16 luma coefficients. */
int32_t evo_variance2(const int16_t coeffs[EVO_MB_SAMPLES]);
/* quantize.cpp:97-118 */
int32_t evo_block_qp(int32_t quality, const int16_t coeffs[EVO_MB_SAMPLES]);
/* quantize.cpp:372-397 with 217-252 */
void evo_quantize_inter_mb(int32_t qp, const int16_t coeffs[EVO_MB_SAMPLES],
                           int16_t levels[EVO_MB_SAMPLES]);

/* Analysis-mode frame driver: every MB in raster order runs the inter
 * searches of the job's references, selects a candidate with the rules of
 * encode.cpp:183-248 (P) or 386-443 (B) against an empty intra sentinel, and
 * for non-copy blocks forms the residual prediction (encode.cpp:569-739),
 * transforms, picks the QP and quantises (pre-trellis).  records gets one
 * entry per MB; coeffs/levels get 384 samples per MB (zero for copies).
 * Returns 0 on success, 10 (EVX_ERROR_EXECUTION_FAILURE) if a search fails. */
int32_t evo_analyze_frame(const evo_frame_job *job, evo_record *records,
                          int16_t *coeffs, int16_t *levels);

/* ---- Reconstruction of inter MBs (decode side) ---- */

/* quantize.cpp:427-449 (EVX_ENABLE_LINEAR_QUANTIZATION 0) with 322-332: the
 * inter inverse quantisation of an MB's levels, luma at qp and chroma at
 * min(qp + 2, 63); each value ((2 * level) * qm * qp) / 16 stored as int16. */
void evo_dequantize_inter_mb(int32_t qp, const int16_t levels[EVO_MB_SAMPLES],
                             int16_t coeffs[EVO_MB_SAMPLES]);

/* macroblock.h:290-295 -> transform.cpp:330-347, 396-433 and the 16x16 split
 * of 572-582: the inverse DCT of each 8x8 block added to the prediction,
 * stored as int16 (no clipping). */
void evo_inverse_transform_add_mb(const int16_t coeffs[EVO_MB_SAMPLES],
                                  const int16_t pred[EVO_MB_SAMPLES],
                                  int16_t out[EVO_MB_SAMPLES]);

/* decode.cpp:123-433 (the inter block types) for every MB of a frame the
 * job describes: the prediction each record names (its reference found by
 * prediction_target among job->target_ids; bidirectional records average
 * both), plus, for transformed (non-copy) MBs, the inverse-quantised,
 * inverse-transformed levels.  out planes: luma width x height with stride
 * width, chroma width/2 x height/2.  Returns 0, or 3 (EVX_ERROR_INVALIDARG)
 * if a record names a target the job does not hold. */
int32_t evo_reconstruct_frame(const evo_frame_job *job, const evo_record *records,
                              const int16_t *levels, int16_t *out_y, int16_t *out_u,
                              int16_t *out_v);

/* ---- In-loop deblocking (decode and encode side) ---- */

/* deblock.cpp:91-165: one line of 8 samples across an edge, v = p3 p2 p1 p0
 * q0 q1 q2 q3, filtered in place at the plane's effective qp (0..63) and
 * strength 0, 1 or 2. */
void evo_deblock_line(int16_t v[8], int32_t qp, int32_t strength, int32_t is_luma);

/* deblock.cpp:222-306 (encode.cpp:968): the deblocking filter of a frame in
 * the reference's raster order, in place.  records: one per MB (block_type
 * and q_index are read).  Planes are packed: luma stride width, chroma
 * width / 2. */
void evo_deblock_frame(const evo_record *records, int32_t width, int32_t height, int16_t *y,
                       int16_t *u, int16_t *v);

/* ---- Sample adaptive offset (decode and encode side) ---- */

#define EVO_SAO_OFF 0                /* sao.h:46-48 */
#define EVO_SAO_EDGE 1
#define EVO_SAO_BAND 2

/* evx_sao_params (sao.h:62-67), padded to 8 bytes. */
typedef struct evo_sao_params {
    uint8_t type;                  /* off, edge or band */
    uint8_t eo_class;              /* horizontal, vertical, 135, 45 degrees */
    uint8_t band_position;         /* first of 4 bands of 8 levels */
    uint8_t reserved;
    int8_t offsets[4];
} evo_sao_params;

/* evx_sao_info (sao.h:69-74), 20 bytes: one per MB. */
typedef struct evo_sao_info {
    uint8_t merge_left, merge_above, reserved[2];
    evo_sao_params luma, chroma;
} evo_sao_info;

/* sao.cpp:98-134 and sao_filter_image (decode.cpp:589): SAO of a frame per
 * `table`, in place, MB by MB in raster order.  Planes are packed. */
void evo_sao_filter_frame(const evo_sao_info *table, int32_t width, int32_t height, int16_t *y,
                          int16_t *u, int16_t *v);

/* sao_encode_frame (sao.cpp:378-590, encode.cpp:974): the SAO decisions of a
 * frame from the source and the (deblocked) reconstruction, written to
 * `table`, then applied to rec in place.  lambda: evo_lambda(quality). */
void evo_sao_encode_frame(const int16_t *src_y, const int16_t *src_u, const int16_t *src_v,
                          int16_t *rec_y, int16_t *rec_u, int16_t *rec_v, int32_t width,
                          int32_t height, float lambda, evo_sao_info *table);

/* ---- Mode decision and in-loop encode (the reference's encode_slice) ---- */

/* generate_intra_prediction (intra_pred.cpp:177-215) of `mode` for the MB at
 * (x, y) from the neighbouring samples of `pic` (128 where absent). */
void evo_intra_prediction(int32_t mode, const evo_planes *pic, int32_t x, int32_t y,
                          int16_t out[EVO_MB_SAMPLES]);
/* select_intra_mode_candidates (intra_pred.cpp:304-326). */
void evo_intra_mode_candidates(const evo_planes *pic, int32_t x, int32_t y, uint8_t cand[3]);
/* select_best_intra_mode (intra_pred.cpp:217-302, fast mode): the mode, and
 * its luma SAD in *sad. */
int32_t evo_select_intra_mode(const int16_t src[EVO_MB_SAMPLES], const evo_planes *pic,
                              int32_t x, int32_t y, int32_t *sad);
/* quantize_macroblock / inverse_quantize_macroblock (quantize.cpp:454-476):
 * intra matrices and DC scales for intra blocks without motion, inter
 * otherwise. */
void evo_quantize_mb(int32_t qp, int32_t block_type, const int16_t coeffs[EVO_MB_SAMPLES],
                     int16_t levels[EVO_MB_SAMPLES]);
void evo_dequantize_mb(int32_t qp, int32_t block_type, const int16_t levels[EVO_MB_SAMPLES],
                       int16_t coeffs[EVO_MB_SAMPLES]);
/* trellis_optimize_macroblock (quantize.cpp:615-680) in place on levels,
 * with the scan of the block's intra mode (scan.h:154-159). */
void evo_trellis_mb(int32_t qp, int32_t block_type, int32_t intra_mode, float lambda,
                    const int16_t coeffs[EVO_MB_SAMPLES], int16_t levels[EVO_MB_SAMPLES]);
/* estimate_block_bits (encode.cpp:39-110) of a descriptor and the luma of
 * `block`. */
uint32_t evo_block_bits(const evo_record *desc, const int16_t block[EVO_MB_SAMPLES]);

/* One frame of the in-loop encode. */
typedef struct evo_encode_job {
    int32_t type;                  /* 0 = I, 1 = P (one reference), 2 = B (fwd, bwd) */
    int32_t quality;
    evo_planes src;
    int32_t n_refs;                /* 0, 1 or 2, as the type says */
    evo_planes refs[2];
    uint8_t target_ids[2];         /* prediction_target of each reference */
    const evo_record *prev;        /* temporal MVP table (P only) or NULL */
} evo_encode_job;

/* encode_slice (encode.cpp:751-813) of one frame: per MB in raster order,
 * classify_block (P and I, encode.cpp:168-370, with RDO for P) or
 * classify_block_bidir (B, 372-476), encode_block with trellis (479-748),
 * and decode_block (decode.cpp:31-433) into the picture being reconstructed
 * (pic_*, packed: luma stride width, chroma width / 2), which holds the
 * previous content of its picture-buffer slot on entry.  records: the
 * descriptors (q_index and variance as encode_block sets them; sad the SAD
 * stage's best, cost_bits the RDO cost of P frames; flags bit0 residual
 * coded, bit1 bidirectional, bit2 encode mode); levels: 384 per MB after
 * trellis, zero for copies; coeffs (may be NULL): the transform
 * coefficients encode_block quantised, zero for copies.  Returns 0, 3 for a
 * bad job or 10 if a search fails. */
int32_t evo_encode_frame(const evo_encode_job *job, int16_t *pic_y, int16_t *pic_u,
                         int16_t *pic_v, evo_record *records, int16_t *levels, int16_t *coeffs);

#ifdef __cplusplus
}
#endif

#endif
