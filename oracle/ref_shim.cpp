// ref_shim.cpp -- extern "C" entry points over the reference's own inter
// analysis functions, compiled together with the reference sources into
// oracle/_ref/libevxref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used to write golden vectors, to check the C
// restatement (evx_oracle.c) on fresh inputs, and as the reference CPU
// baseline.  Nothing in the product loads it.
//
// The per-block functions call the reference directly.  The frame drivers
// compose reference calls the way the encoder does (encode.cpp:183-248 for
// P, 386-443 for B, 569-739 for the residual prediction, 726-728 for QP and
// quantisation) with the intra candidates replaced by an empty sentinel and
// trellis left out (levels are pre-trellis).
// <stdint.h> first: base.h includes it inside namespace evx (base.h:137).
#include <stdint.h>

#include "analysis.h"
#include "common.h"
#include "imageset.h"
#include "interp.h"
#include "intra_pred.h"
#include "macroblock.h"
#include "motion.h"
#include "quantize.h"
#include "sao.h"
#include "scan.h"
#include "transform.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "evx_oracle.h"

// deblock.cpp defines these at namespace scope without a header.
namespace evx {
evx_status deblock_image_filter(evx_block_desc *block_table, image_set *target_image);
void deblock_filter_values(int16 p3, int16 p2, int16 p1, int16 p0, int16 q0, int16 q1, int16 q2,
                           int16 q3, int16 *out_p2, int16 *out_p1, int16 *out_p0, int16 *out_q0,
                           int16 *out_q1, int16 *out_q2, uint8 average_qp, uint8 strength,
                           bool is_luma);
// serialize.cpp defines these at namespace scope without a header.
evx_status serialize_block_table_unified(uint32 block_count, uint32 width_in_blocks,
                                         evx_block_desc *block_table, bit_stream *feed_stream,
                                         entropy_coder *coder, evx_metadata_contexts *meta,
                                         const evx_block_desc *prev_block_table,
                                         bool is_bidir_frame, bit_stream *output);
evx_status serialize_sao_table(const evx_sao_info *sao_table, uint32 block_count,
                               uint32 width_in_blocks, bit_stream *feed_stream,
                               entropy_coder *coder, evx_metadata_contexts *meta,
                               bit_stream *output);
// encode.cpp defines the slice encoders at namespace scope without a header.
evx_status encode_slice(const evx_frame &frame, evx_context *context);
evx_status encode_slice_wpp(const evx_frame &frame, evx_context *context);
}  // namespace evx

using namespace evx;

namespace {

void fill_set(image_set *set, const evo_planes &p)
{
    set->initialize(EVX_IMAGE_FORMAT_R16S, (uint16)p.width, (uint16)p.height);
    int16 *y = reinterpret_cast<int16 *>(set->query_y_image()->query_data());
    int16 *u = reinterpret_cast<int16 *>(set->query_u_image()->query_data());
    int16 *v = reinterpret_cast<int16 *>(set->query_v_image()->query_data());
    for (int r = 0; r < p.height; r++)
        memcpy(y + r * p.width, p.y + r * p.stride_y, p.width * sizeof(int16));
    int cw = p.width >> 1, ch = p.height >> 1;
    for (int r = 0; r < ch; r++) {
        memcpy(u + r * cw, p.u + r * p.stride_c, cw * sizeof(int16));
        memcpy(v + r * cw, p.v + r * p.stride_c, cw * sizeof(int16));
    }
}

macroblock view(int16_t *buf)
{
    macroblock m;
    m.data_y = buf;
    m.data_u = buf + 256;
    m.data_v = buf + 320;
    m.stride = 16;
    return m;
}

void copy_out(const macroblock &m, int16_t *out)
{
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) out[r * 16 + c] = m.data_y[r * m.stride + c];
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            out[256 + r * 8 + c] = m.data_u[r * (m.stride >> 1) + c];
            out[320 + r * 8 + c] = m.data_v[r * (m.stride >> 1) + c];
        }
}

struct bank_holder {
    evx_cache_bank bank;
    bank_holder()
    {
        bank.transform_cache.initialize(EVX_IMAGE_FORMAT_R16S, 16, 16);
        bank.motion_cache.initialize(EVX_IMAGE_FORMAT_R16S, 16, 16);
        bank.staging_cache.initialize(EVX_IMAGE_FORMAT_R16S, 16, 16);
        create_macroblock(bank.transform_cache, 0, 0, &bank.transform_block);
        create_macroblock(bank.motion_cache, 0, 0, &bank.motion_block);
        create_macroblock(bank.staging_cache, 0, 0, &bank.staging_block);
    }
};

void desc_to_record(const evx_block_desc &d, evo_record *r)
{
    memset(r, 0, sizeof(*r));
    r->block_type = (uint8_t)d.block_type;
    r->prediction_mode = d.prediction_mode;
    r->prediction_target = d.prediction_target;
    r->prediction_target_b = d.prediction_target_b;
    r->motion_x = d.motion_x;
    r->motion_y = d.motion_y;
    r->motion_x_b = d.motion_x_b;
    r->motion_y_b = d.motion_y_b;
    r->sp_pred = d.sp_pred;
    r->sp_amount = d.sp_amount;
    r->sp_index = d.sp_index;
    r->sp_pred_b = d.sp_pred_b;
    r->sp_amount_b = d.sp_amount_b;
    r->sp_index_b = d.sp_index_b;
    r->q_index = d.q_index;
    r->variance = d.variance;
}

// Residual prediction for a chosen descriptor (encode.cpp:625-723), written
// into `out` (an MB-order buffer).
void form_prediction(evx_cache_bank *bank, const evx_block_desc &d, int32 x, int32 y,
                     int16_t *out)
{
    if (d.prediction_mode == 2) {
        macroblock fwd, bwd;
        create_macroblock(bank->prediction_cache[d.prediction_target], x + d.motion_x,
                          y + d.motion_y, &fwd);
        const macroblock *fr = &fwd;
        if (d.sp_pred) {
            int16 i, j;
            compute_motion_direction_from_frac_index(d.sp_index, &i, &j);
            interpolate_subpixel_6tap(bank->prediction_cache[d.prediction_target], x + d.motion_x,
                                      y + d.motion_y, i, j, d.sp_amount, &bank->motion_block);
            fr = &bank->motion_block;
        }
        create_macroblock(bank->prediction_cache[d.prediction_target_b], x + d.motion_x_b,
                          y + d.motion_y_b, &bwd);
        const macroblock *br = &bwd;
        if (d.sp_pred_b) {
            int16 i, j;
            compute_motion_direction_from_frac_index(d.sp_index_b, &i, &j);
            interpolate_subpixel_6tap(bank->prediction_cache[d.prediction_target_b],
                                      x + d.motion_x_b, y + d.motion_y_b, i, j, d.sp_amount_b,
                                      &bank->transform_block);
            br = &bank->transform_block;
        }
        for (int r = 0; r < 16; r++)
            for (int c = 0; c < 16; c++)
                out[r * 16 + c] = (int16)((fr->data_y[r * fr->stride + c] +
                                           br->data_y[r * br->stride + c] + 1) >> 1);
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                uint32 fs = fr->stride >> 1, bs = br->stride >> 1;
                out[256 + r * 8 + c] =
                    (int16)((fr->data_u[r * fs + c] + br->data_u[r * bs + c] + 1) >> 1);
                out[320 + r * 8 + c] =
                    (int16)((fr->data_v[r * fs + c] + br->data_v[r * bs + c] + 1) >> 1);
            }
        return;
    }
    macroblock beta;
    create_macroblock(bank->prediction_cache[d.prediction_target], x + d.motion_x,
                      y + d.motion_y, &beta);
    if (d.sp_pred) {
        int16 i, j;
        compute_motion_direction_from_frac_index(d.sp_index, &i, &j);
        interpolate_subpixel_6tap(bank->prediction_cache[d.prediction_target], x + d.motion_x,
                                  y + d.motion_y, i, j, d.sp_amount, &bank->motion_block);
        copy_out(bank->motion_block, out);
    } else {
        copy_out(beta, out);
    }
}

}  // namespace

extern "C" {

float evr_lambda(int32_t quality) { return compute_lambda((uint16)quality); }

void evr_interp_block(const evo_planes *ref, int32_t bx, int32_t by, int32_t dx, int32_t dy,
                      int32_t quarter, int16_t *out)
{
    image_set set;
    fill_set(&set, *ref);
    std::vector<int16_t> buf(384);
    macroblock m = view(buf.data());
    interpolate_subpixel_6tap(set, bx, by, (int16)dx, (int16)dy, quarter != 0, &m);
    memcpy(out, buf.data(), 384 * sizeof(int16_t));
}

// Batched interpolation over one picture: n positions/directions.
void evr_interp_many(const evo_planes *ref, int32_t n, const int32_t *bx, const int32_t *by,
                     const int32_t *dx, const int32_t *dy, const int32_t *quarter, int16_t *out)
{
    image_set set;
    fill_set(&set, *ref);
    std::vector<int16_t> buf(384);
    for (int32_t k = 0; k < n; k++) {
        macroblock m = view(buf.data());
        interpolate_subpixel_6tap(set, bx[k], by[k], (int16)dx[k], (int16)dy[k], quarter[k] != 0,
                                  &m);
        memcpy(out + (size_t)k * 384, buf.data(), 384 * sizeof(int16_t));
    }
}

void evr_hadamard_4x4(const int16_t *src, int32_t stride, int32_t *out)
{
    hadamard_4x4(src, stride, out);
}

void evr_block_metrics(const int16_t *a, const int16_t *b, int32_t *sad, int32_t *mad,
                       int32_t *satd)
{
    macroblock ma = view(const_cast<int16_t *>(a)), mb = view(const_cast<int16_t *>(b));
    *sad = compute_block_sad(ma, mb);
    *mad = compute_block_mad(ma, mb);
    *satd = compute_block_satd(ma, mb);
}

void evr_sub_transform_mb(const int16_t *src, const int16_t *pred, int16_t *coeffs)
{
    macroblock ms = view(const_cast<int16_t *>(src)), mp = view(const_cast<int16_t *>(pred));
    macroblock mc = view(coeffs);
    sub_transform_macroblock(ms, mp, &mc);
}

int32_t evr_variance2(const int16_t *coeffs)
{
    macroblock mc = view(const_cast<int16_t *>(coeffs));
    return compute_block_variance2(mc);
}

int32_t evr_block_qp(int32_t quality, const int16_t *coeffs)
{
    macroblock mc = view(const_cast<int16_t *>(coeffs));
    return query_block_quantization_parameter((uint8)quality, mc, EVX_BLOCK_INTER_MOTION_DELTA);
}

void evr_quantize_mb(int32_t qp, int32_t block_type, const int16_t *coeffs, int16_t *levels)
{
    macroblock mc = view(const_cast<int16_t *>(coeffs));
    macroblock ml = view(levels);
    quantize_macroblock((uint8)qp, (EVX_BLOCK_TYPE)block_type, mc, &ml);
}

void evr_mv_predictor(const evo_record *table, uint32_t wb, uint32_t n_blocks, uint32_t idx,
                      const evo_record *prev, int16_t *px, int16_t *py)
{
    std::vector<evx_block_desc> t(n_blocks), p(n_blocks);
    for (uint32_t k = 0; k < n_blocks; k++) {
        clear_block_desc(&t[k]);
        t[k].block_type = (EVX_BLOCK_TYPE)table[k].block_type;
        t[k].motion_x = table[k].motion_x;
        t[k].motion_y = table[k].motion_y;
        clear_block_desc(&p[k]);
        if (prev) {
            p[k].block_type = (EVX_BLOCK_TYPE)prev[k].block_type;
            p[k].motion_x = prev[k].motion_x;
            p[k].motion_y = prev[k].motion_y;
        }
    }
    compute_mv_predictor(t.data(), wb, idx, prev ? p.data() : nullptr, px, py);
}

// perform_inter_search with the predictor supplied through a one-neighbour
// block table (a single motion neighbour makes compute_mv_predictor return
// its vector, motion.cpp:331-335).
int32_t evr_inter_search(const int16_t *src, const evo_planes *ref, int32_t x, int32_t y,
                         int32_t quality, int16_t pmx, int16_t pmy, evo_record *out)
{
    bank_holder h;
    fill_set(&h.bank.prediction_cache[0], *ref);
    evx_block_desc table[2];
    clear_block_desc(&table[0]);
    clear_block_desc(&table[1]);
    if (pmx != 0 || pmy != 0) {
        table[0].block_type = EVX_BLOCK_INTER_MOTION_DELTA;
        table[0].motion_x = pmx;
        table[0].motion_y = pmy;
    } else {
        table[0].block_type = EVX_BLOCK_INTER_DELTA;
    }
    evx_frame frame;
    clear_frame(&frame);
    frame.type = EVX_FRAME_INTER;
    frame.quality = (uint16)quality;
    std::vector<int16_t> s(src, src + 384);
    macroblock ms = view(s.data());
    evx_block_desc d;
    int32 sad = perform_inter_search(frame, ms, x, y, &h.bank, 0, table, 2, 1, nullptr, &d);
    desc_to_record(d, out);
    out->sad = sad;
    return sad;
}

// calculate_intra_prediction (motion.cpp:485-565) with `pic` as the picture
// being reconstructed, the frame's own picture-buffer slot (motion.cpp:510-515).
int32_t evr_intra_search(const int16_t *src, const evo_planes *pic, int32_t x, int32_t y,
                         int32_t quality, evo_record *out)
{
    bank_holder h;
    fill_set(&h.bank.prediction_cache[0], *pic);
    evx_frame frame;
    clear_frame(&frame);
    frame.type = EVX_FRAME_INTER;
    frame.quality = (uint16)quality;
    frame.dpb_slot = 0;
    std::vector<int16_t> s(src, src + 384);
    macroblock ms = view(s.data());
    evx_block_desc d;
    int32 sad = calculate_intra_prediction(frame, ms, x, y, &h.bank, &d);
    desc_to_record(d, out);
    out->sad = sad;
    return sad;
}

int32_t evr_bidir(const int16_t *src, const evo_planes *fwd, const evo_planes *bwd, int32_t x,
                  int32_t y, int32_t quality, int16_t pmx, int16_t pmy, evo_record *out,
                  int16_t *avg)
{
    bank_holder h;
    fill_set(&h.bank.prediction_cache[0], *fwd);
    fill_set(&h.bank.prediction_cache[1], *bwd);
    evx_block_desc table[2];
    clear_block_desc(&table[0]);
    clear_block_desc(&table[1]);
    if (pmx != 0 || pmy != 0) {
        table[0].block_type = EVX_BLOCK_INTER_MOTION_DELTA;
        table[0].motion_x = pmx;
        table[0].motion_y = pmy;
    }
    evx_frame frame;
    clear_frame(&frame);
    frame.type = EVX_FRAME_BIDIR;
    frame.quality = (uint16)quality;
    std::vector<int16_t> s(src, src + 384);
    macroblock ms = view(s.data());
    evx_block_desc d;
    int32 sad = calculate_bidir_prediction(frame, ms, x, y, &h.bank, 0, 1, table, 2, 1, &d);
    desc_to_record(d, out);
    out->sad = sad;
    copy_out(h.bank.staging_block, avg);
    return sad;
}

}  // extern "C"

namespace {

// Per-frame state shared by the frame drivers: the source and reference
// pictures (filled once), the frame descriptor and the block table.
struct frame_state {
    const evo_frame_job *job;
    uint32 wb, hb;
    image_set src_set;
    evx_cache_bank refs;                  // prediction_cache[k] = reference k
    evx_frame frame;
    float lambda;
    std::vector<evx_block_desc> table, prev;
    const evx_block_desc *prev_ptr = nullptr;
    evo_record *records;
    int16_t *coeffs, *levels;

    frame_state(const evo_frame_job *j, evo_record *r, int16_t *c, int16_t *l)
        : job(j), records(r), coeffs(c), levels(l)
    {
        wb = j->src.width / 16;
        hb = j->src.height / 16;
        fill_set(&src_set, j->src);
        for (int32 k = 0; k < j->n_refs; k++) fill_set(&refs.prediction_cache[k], j->refs[k]);
        table.resize(wb * hb);
        if (j->type == 1 && j->prev) {
            prev.resize(wb * hb);
            for (uint32 k = 0; k < wb * hb; k++) {
                clear_block_desc(&prev[k]);
                prev[k].block_type = (EVX_BLOCK_TYPE)j->prev[k].block_type;
                prev[k].motion_x = j->prev[k].motion_x;
                prev[k].motion_y = j->prev[k].motion_y;
            }
            prev_ptr = prev.data();
        }
        clear_frame(&frame);
        frame.type = j->type == 1 ? EVX_FRAME_INTER : EVX_FRAME_BIDIR;
        frame.quality = (uint16)j->quality;
        lambda = compute_lambda(frame.quality);
    }
};

// One MB: the searches, selection, residual prediction, DCT, QP and
// quantisation, and its record.  `bank` holds the references (its own or
// borrowed) and the worker's scratch blocks.  Slot numbers are the
// reference indices; prediction_target is mapped to the caller's ids when
// the record is written out.
int32_t analyze_mb(frame_state &fs, evx_cache_bank *bank, uint32 bx, uint32 by)
{
    const evo_frame_job *job = fs.job;
    const uint32 wb = fs.wb;
    const uint32 idx = by * wb + bx;
    const int32 x = bx * 16, y = by * 16;
    macroblock src_block;
    create_macroblock(fs.src_set, x, y, &src_block);
    evx_block_desc best;
    int32 best_sad = EVX_MAX_INT32;
    bool best_copy = false;
    int32 sad_b = 0;
    if (job->type == 1) {
        for (int32 k = 0; k < job->n_refs; k++) {
            evx_block_desc d;
            int32 sad = perform_inter_search(fs.frame, src_block, x, y, bank, k, fs.table.data(),
                                             wb, idx, fs.prev_ptr, &d);
            d.prediction_target = (uint8)k;
            int32 bias = (int32)(fs.lambda * (float)((k + 1) * 2));
            bool copy = EVX_IS_COPY_BLOCK_TYPE(d.block_type);
            if (copy ^ best_copy) {
                if (copy) { best = d; best_sad = sad; best_copy = true; }
            } else if (sad + bias < best_sad) {
                best = d; best_sad = sad; best_copy = copy;
            }
        }
    } else {
        evx_block_desc fd, bd, bi;
        int32 fsad = perform_inter_search(fs.frame, src_block, x, y, bank, 0, fs.table.data(), wb,
                                          idx, nullptr, &fd);
        fd.prediction_mode = 0;
        fd.prediction_target = 0;
        if (fsad < best_sad) { best = fd; best_sad = fsad; }
        int32 bsad = perform_inter_search(fs.frame, src_block, x, y, bank, 1, fs.table.data(), wb,
                                          idx, nullptr, &bd);
        bd.prediction_mode = 1;
        bd.prediction_target = 1;
        if (bsad < best_sad) { best = bd; best_sad = bsad; }
        int32 bis = calculate_bidir_prediction(fs.frame, src_block, x, y, bank, 0, 1,
                                               fs.table.data(), wb, idx, &bi);
        int32 bias = (int32)(fs.lambda * 8.0f);
        if (bis + bias < best_sad) { best = bi; best_sad = bis; sad_b = bsad; }
    }
    if (best_sad == EVX_MAX_INT32) return 10;
    fs.table[idx] = best;

    int16_t pred[384];
    int16_t *c = fs.coeffs + (size_t)idx * 384, *l = fs.levels + (size_t)idx * 384;
    evx_block_desc &d = fs.table[idx];
    bool transformed = !EVX_IS_COPY_BLOCK_TYPE(d.block_type);
    if (transformed) {
        form_prediction(bank, d, x, y, pred);
        macroblock mp = view(pred), mc = view(c), ml = view(l);
        sub_transform_macroblock(src_block, mp, &mc);
        d.q_index = query_block_quantization_parameter((uint8)fs.frame.quality, mc, d.block_type);
        d.variance = compute_block_variance2(mc);
        quantize_macroblock(d.q_index, d.block_type, mc, &ml);
    } else {
        memset(c, 0, 384 * sizeof(int16_t));
        memset(l, 0, 384 * sizeof(int16_t));
    }
    evo_record *r = &fs.records[idx];
    desc_to_record(d, r);
    r->prediction_target = job->target_ids[d.prediction_target];
    if (d.prediction_mode == 2) r->prediction_target_b = job->target_ids[d.prediction_target_b];
    r->sad = best_sad;
    if (d.prediction_mode == 2) r->sad_b = sad_b;
    r->flags = (transformed ? 1u : 0u) | (d.prediction_mode == 2 ? 2u : 0u);
    return 0;
}

}  // namespace

extern "C" {

// Frame driver over reference calls; the same composition as
// evo_analyze_frame, in raster order on the calling thread.  mad/cost/
// counter fields are left zero (the reference does not expose them).
int32_t evr_analyze_frame(const evo_frame_job *job, evo_record *records, int16_t *coeffs,
                          int16_t *levels)
{
    frame_state fs(job, records, coeffs, levels);
    bank_holder bh;
    for (int32 k = 0; k < job->n_refs; k++)
        bh.bank.prediction_cache[k].borrow(fs.refs.prediction_cache[k]);
    for (uint32 by = 0; by < fs.hb; by++)
        for (uint32 bx = 0; bx < fs.wb; bx++) {
            const int32_t rc = analyze_mb(fs, &bh.bank, bx, by);
            if (rc) return rc;
        }
    return 0;
}

// The inter inverse quantisation of an MB (quantize.cpp:466-476 for a
// non-intra block type) and the inverse transform added to a prediction
// (macroblock.h:290-295), on MB-order buffers.
void evr_dequantize_inter_mb(int32_t qp, const int16_t *levels, int16_t *coeffs)
{
    std::vector<int16_t> in(levels, levels + 384);
    macroblock ms = view(in.data()), md = view(coeffs);
    inverse_quantize_macroblock((uint8)qp, EVX_BLOCK_INTER_MOTION_DELTA, ms, &md);
}

void evr_inverse_transform_add_mb(const int16_t *coeffs, const int16_t *pred, int16_t *out)
{
    std::vector<int16_t> c(coeffs, coeffs + 384), p(pred, pred + 384);
    macroblock mc = view(c.data()), mp = view(p.data()), mo = view(out);
    inverse_transform_add_macroblock(mc, mp, &mo);
}

// Reconstruction of an analysed frame the way decode_block treats the inter
// block types (decode.cpp:123-433): the prediction each record names, plus
// the inverse-quantised, inverse-transformed levels for non-copy blocks.
// Output planes are packed (luma stride width, chroma width / 2).
int32_t evr_reconstruct_frame(const evo_frame_job *job, const evo_record *records,
                              const int16_t *levels, int16_t *out_y, int16_t *out_u,
                              int16_t *out_v)
{
    const int32 w = job->src.width, h = job->src.height;
    const uint32 wb = w / 16, hb = h / 16;
    bank_holder bh;
    evx_cache_bank *bank = &bh.bank;
    for (int32 k = 0; k < job->n_refs; k++) fill_set(&bank->prediction_cache[k], job->refs[k]);
    auto slot_of = [&](uint8_t target) {
        for (int32 k = 0; k < job->n_refs; k++)
            if (job->target_ids[k] == target) return k;
        return -1;
    };
    std::vector<int16_t> pred(384), lv(384), co(384), rec(384);
    for (uint32 by = 0; by < hb; by++)
        for (uint32 bx = 0; bx < wb; bx++) {
            const uint32 idx = by * wb + bx;
            const int32 x = bx * 16, y = by * 16;
            const evo_record &r = records[idx];
            evx_block_desc d;
            clear_block_desc(&d);
            d.block_type = (EVX_BLOCK_TYPE)r.block_type;
            d.prediction_mode = r.prediction_mode;
            const int32 k = slot_of(r.prediction_target);
            const int32 kb = r.prediction_mode == 2 ? slot_of(r.prediction_target_b) : 0;
            if (k < 0 || kb < 0) return 3;
            d.prediction_target = (uint8)k;
            d.prediction_target_b = (uint8)kb;
            d.motion_x = r.motion_x; d.motion_y = r.motion_y;
            d.motion_x_b = r.motion_x_b; d.motion_y_b = r.motion_y_b;
            d.sp_pred = r.sp_pred; d.sp_amount = r.sp_amount; d.sp_index = r.sp_index;
            d.sp_pred_b = r.sp_pred_b; d.sp_amount_b = r.sp_amount_b; d.sp_index_b = r.sp_index_b;
            d.q_index = r.q_index;
            form_prediction(bank, d, x, y, pred.data());
            if (EVX_IS_COPY_BLOCK_TYPE(d.block_type)) {
                rec = pred;
            } else {
                memcpy(lv.data(), levels + (size_t)idx * 384, 384 * sizeof(int16_t));
                macroblock ml = view(lv.data()), mc = view(co.data()), mp = view(pred.data()),
                           mr = view(rec.data());
                inverse_quantize_macroblock(d.q_index, d.block_type, ml, &mc);
                inverse_transform_add_macroblock(mc, mp, &mr);
            }
            for (int j = 0; j < 16; j++)
                memcpy(out_y + (size_t)(y + j) * w + x, rec.data() + j * 16, 16 * sizeof(int16_t));
            for (int j = 0; j < 8; j++) {
                memcpy(out_u + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec.data() + 256 + j * 8,
                       8 * sizeof(int16_t));
                memcpy(out_v + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec.data() + 320 + j * 8,
                       8 * sizeof(int16_t));
            }
        }
    return 0;
}

// The same frame on `n_threads` threads in wavefront order, as the
// reference's encode_slice_wpp schedules rows (encode.cpp:817-941): workers
// claim rows, each with a private cache bank borrowing the shared pictures,
// and MB (bx, by) waits until row by - 1 has finished column
// min(bx + 1, wb - 1) -- the top-right neighbour the MV predictor reads
// (the reference waits for col + 4 because its intra search reads further).
// Results are identical to evr_analyze_frame's.
int32_t evr_analyze_frame_wpp(const evo_frame_job *job, evo_record *records, int16_t *coeffs,
                              int16_t *levels, int32_t n_threads)
{
    frame_state fs(job, records, coeffs, levels);
    const uint32 rows = fs.hb, cols = fs.wb;
    std::vector<std::atomic<int32>> progress(rows);
    for (auto &p : progress) p.store(-1, std::memory_order_relaxed);
    std::atomic<uint32> next_row(0);
    std::atomic<int32_t> error(0);
    auto worker = [&]() {
        bank_holder bh;
        for (int32 k = 0; k < job->n_refs; k++)
            bh.bank.prediction_cache[k].borrow(fs.refs.prediction_cache[k]);
        for (;;) {
            const uint32 row = next_row.fetch_add(1, std::memory_order_relaxed);
            if (row >= rows) return;
            for (uint32 col = 0; col < cols; col++) {
                if (row > 0) {
                    const int32 need = (int32)std::min(col + 1, cols - 1);
                    while (progress[row - 1].load(std::memory_order_acquire) < need) {
                        if (error.load(std::memory_order_relaxed)) return;
                        std::this_thread::yield();
                    }
                }
                const int32_t rc = analyze_mb(fs, &bh.bank, col, row);
                if (rc) { error.store(rc); return; }
                progress[row].store((int32)col, std::memory_order_release);
            }
        }
    };
    const int32_t n = std::max(1, std::min<int32_t>(n_threads, (int32_t)rows));
    std::vector<std::thread> pool;
    for (int32_t t = 1; t < n; t++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return error.load();
}


// deblock_filter_values (deblock.cpp:91-165) on one line v = p3 .. q3.
void evr_deblock_line(int16_t *v, int32_t qp, int32_t strength, int32_t is_luma)
{
    deblock_filter_values(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], &v[1], &v[2], &v[3],
                          &v[4], &v[5], &v[6], (uint8)qp, (uint8)strength, is_luma != 0);
}

// deblock_image_filter (deblock.cpp:298-306, encode.cpp:968) on a packed
// frame, in place; the block table carries each record's type and q_index.
void evr_deblock_frame(const evo_record *records, int32_t width, int32_t height, int16_t *y,
                       int16_t *u, int16_t *v)
{
    evo_planes p = {y, u, v, width, height, width, width / 2};
    image_set set;
    fill_set(&set, p);
    const size_t n = (size_t)(width / 16) * (height / 16);
    std::vector<evx_block_desc> table(n);
    for (size_t i = 0; i < n; i++) {
        clear_block_desc(&table[i]);
        table[i].block_type = (EVX_BLOCK_TYPE)records[i].block_type;
        table[i].q_index = records[i].q_index;
    }
    deblock_image_filter(table.data(), &set);
    const int16 *sy = reinterpret_cast<const int16 *>(set.query_y_image()->query_data());
    const int16 *su = reinterpret_cast<const int16 *>(set.query_u_image()->query_data());
    const int16 *sv = reinterpret_cast<const int16 *>(set.query_v_image()->query_data());
    memcpy(y, sy, (size_t)width * height * sizeof(int16_t));
    memcpy(u, su, (size_t)(width / 2) * (height / 2) * sizeof(int16_t));
    memcpy(v, sv, (size_t)(width / 2) * (height / 2) * sizeof(int16_t));
}

}  // extern "C"

namespace {

void to_ref(const evo_sao_params &a, evx_sao_params *b)
{
    b->type = a.type;
    b->eo_class = a.eo_class;
    b->band_position = a.band_position;
    for (int k = 0; k < 4; k++) b->offsets[k] = a.offsets[k];
}

void from_ref(const evx_sao_params &a, evo_sao_params *b)
{
    memset(b, 0, sizeof(*b));
    b->type = a.type;
    b->eo_class = a.eo_class;
    b->band_position = a.band_position;
    for (int k = 0; k < 4; k++) b->offsets[k] = a.offsets[k];
}

void copy_back(const image_set &set, int32_t width, int32_t height, int16_t *y, int16_t *u,
               int16_t *v)
{
    memcpy(y, set.query_y_image()->query_data(), (size_t)width * height * sizeof(int16_t));
    memcpy(u, set.query_u_image()->query_data(), (size_t)(width / 2) * (height / 2) * sizeof(int16_t));
    memcpy(v, set.query_v_image()->query_data(), (size_t)(width / 2) * (height / 2) * sizeof(int16_t));
}

// The arithmetic-coded metadata of a frame as serialize_slice writes it
// first (serialize.cpp:624-660): the unified block table, or the SAO table;
// the coder is finished and the stream padded to a byte.  Returns the byte
// count, or -1 when `cap` is too small.
template <class F>
int32_t coded_bytes(F &&write, uint8_t *out, int32_t cap)
{
    bit_stream feed(1 << 16), output((uint32)cap * 8 + 64);
    entropy_coder coder;
    coder.clear();
    evx_metadata_contexts meta;
    meta.clear();
    write(&feed, &coder, &meta, &output);
    coder.finish_encode(&output);
    while (output.query_write_index() & 7) output.write_bit(0);
    const int32_t n = (int32_t)(output.query_write_index() / 8);
    if (n > cap) return -1;
    memcpy(out, output.query_data(), (size_t)n);
    return n;
}

// Block descriptors from records: every field the wire format reads; the
// record's target ids become reference indexes through target_ids.
void records_to_descs(const evo_record *r, size_t n, const uint8_t *target_ids, int32_t n_ids,
                      std::vector<evx_block_desc> *out)
{
    auto index_of = [&](uint8_t id) -> uint8 {
        for (int32_t k = 0; k < n_ids; k++)
            if (target_ids[k] == id) return (uint8)k;
        return id;
    };
    out->resize(n);
    for (size_t i = 0; i < n; i++) {
        evx_block_desc &d = (*out)[i];
        clear_block_desc(&d);
        d.block_type = (EVX_BLOCK_TYPE)r[i].block_type;
        d.prediction_mode = r[i].prediction_mode;
        d.prediction_target = index_of(r[i].prediction_target);
        d.prediction_target_b = index_of(r[i].prediction_target_b);
        d.motion_x = r[i].motion_x;
        d.motion_y = r[i].motion_y;
        d.motion_x_b = r[i].motion_x_b;
        d.motion_y_b = r[i].motion_y_b;
        d.sp_pred = r[i].sp_pred;
        d.sp_amount = r[i].sp_amount;
        d.sp_index = r[i].sp_index;
        d.sp_pred_b = r[i].sp_pred_b;
        d.sp_amount_b = r[i].sp_amount_b;
        d.sp_index_b = r[i].sp_index_b;
        d.q_index = r[i].q_index;
        d.variance = r[i].variance;
    }
}

}  // namespace

extern "C" {

// sao_filter_image (sao.cpp:136-146) on a packed frame, in place.
void evr_sao_filter_frame(const evo_sao_info *table, int32_t width, int32_t height, int16_t *y,
                          int16_t *u, int16_t *v)
{
    evo_planes p = {y, u, v, width, height, width, width / 2};
    image_set set;
    fill_set(&set, p);
    const size_t n = (size_t)(width / 16) * (height / 16);
    std::vector<evx_sao_info> t(n);
    for (size_t i = 0; i < n; i++) {
        clear_sao_info(&t[i]);
        t[i].merge_left = table[i].merge_left != 0;
        t[i].merge_above = table[i].merge_above != 0;
        to_ref(table[i].luma, &t[i].luma);
        to_ref(table[i].chroma, &t[i].chroma);
    }
    sao_filter_image(t.data(), width / 16, height / 16, &set);
    copy_back(set, width, height, y, u, v);
}

// sao_encode_frame (sao.cpp:378-590) on packed frames: decisions into
// `table`, the reconstruction filtered in place.
void evr_sao_encode_frame(const int16_t *src_y, const int16_t *src_u, const int16_t *src_v,
                          int16_t *rec_y, int16_t *rec_u, int16_t *rec_v, int32_t width,
                          int32_t height, float lambda, evo_sao_info *table)
{
    evo_planes ps = {src_y, src_u, src_v, width, height, width, width / 2};
    evo_planes pr = {rec_y, rec_u, rec_v, width, height, width, width / 2};
    image_set src, rec;
    fill_set(&src, ps);
    fill_set(&rec, pr);
    const size_t n = (size_t)(width / 16) * (height / 16);
    std::vector<evx_sao_info> t(n);
    sao_encode_frame(src, &rec, t.data(), width / 16, height / 16, lambda);
    for (size_t i = 0; i < n; i++) {
        memset(&table[i], 0, sizeof(table[i]));
        table[i].merge_left = t[i].merge_left;
        table[i].merge_above = t[i].merge_above;
        from_ref(t[i].luma, &table[i].luma);
        from_ref(t[i].chroma, &table[i].chroma);
    }
    copy_back(rec, width, height, rec_y, rec_u, rec_v);
}


// The block table the reference's own analysis of `job` leaves
// (evx_block_desc, never converted), serialized as serialize_slice does.
int32_t evr_analyze_frame_serialized(const evo_frame_job *job, uint8_t *out, int32_t cap)
{
    const size_t n = (size_t)(job->src.width / 16) * (job->src.height / 16);
    std::vector<evo_record> rec(n);
    std::vector<int16_t> co(n * 384), lv(n * 384);
    frame_state fs(job, rec.data(), co.data(), lv.data());
    bank_holder bh;
    for (int32 k = 0; k < job->n_refs; k++)
        bh.bank.prediction_cache[k].borrow(fs.refs.prediction_cache[k]);
    for (uint32 by = 0; by < fs.hb; by++)
        for (uint32 bx = 0; bx < fs.wb; bx++) {
            const int32_t rc = analyze_mb(fs, &bh.bank, bx, by);
            if (rc) return -rc - 1000;
        }
    return coded_bytes([&](bit_stream *feed, entropy_coder *coder, evx_metadata_contexts *meta,
                           bit_stream *output) {
        serialize_block_table_unified((uint32)n, fs.wb, fs.table.data(), feed, coder, meta,
                                      fs.prev_ptr, job->type == 2, output);
    }, out, cap);
}

// The same serialization from records (the block type and motion of
// `prev`, when given, are the temporal MV predictor's table).
int32_t evr_serialize_records(const evo_record *records, int32_t wb, int32_t hb,
                              const evo_record *prev, int32_t is_bidir, const uint8_t *target_ids,
                              int32_t n_ids, uint8_t *out, int32_t cap)
{
    const size_t n = (size_t)wb * hb;
    std::vector<evx_block_desc> t, p;
    records_to_descs(records, n, target_ids, n_ids, &t);
    if (prev) records_to_descs(prev, n, target_ids, n_ids, &p);
    return coded_bytes([&](bit_stream *feed, entropy_coder *coder, evx_metadata_contexts *meta,
                           bit_stream *output) {
        serialize_block_table_unified((uint32)n, (uint32)wb, t.data(), feed, coder, meta,
                                      prev ? p.data() : nullptr, is_bidir != 0, output);
    }, out, cap);
}

// serialize_sao_table (serialize.cpp:564-622) of a table.
int32_t evr_serialize_sao(const evo_sao_info *table, int32_t wb, int32_t hb, uint8_t *out,
                          int32_t cap)
{
    const size_t n = (size_t)wb * hb;
    std::vector<evx_sao_info> t(n);
    for (size_t i = 0; i < n; i++) {
        clear_sao_info(&t[i]);
        t[i].merge_left = table[i].merge_left != 0;
        t[i].merge_above = table[i].merge_above != 0;
        to_ref(table[i].luma, &t[i].luma);
        to_ref(table[i].chroma, &t[i].chroma);
    }
    return coded_bytes([&](bit_stream *feed, entropy_coder *coder, evx_metadata_contexts *meta,
                           bit_stream *output) {
        serialize_sao_table(t.data(), (uint32)n, (uint32)wb, feed, coder, meta, output);
    }, out, cap);
}


// ---- Mode decision and in-loop encode: the reference's own functions ----

// generate_intra_prediction (intra_pred.cpp:177-215) into an MB-order buffer.
void evr_intra_prediction(int32_t mode, const evo_planes *pic, int32_t x, int32_t y,
                          int16_t *out)
{
    image_set set;
    fill_set(&set, *pic);
    std::vector<int16_t> buf(384);
    macroblock m = view(buf.data());
    generate_intra_prediction((uint8)mode, set, x, y, &m);
    copy_out(m, out);
}

// select_intra_mode_candidates (intra_pred.cpp:304-326).
void evr_intra_mode_candidates(const evo_planes *pic, int32_t x, int32_t y, uint8_t *cand)
{
    image_set set;
    fill_set(&set, *pic);
    select_intra_mode_candidates(set, x, y, cand);
}

// select_best_intra_mode (intra_pred.cpp:217-302).
int32_t evr_select_intra_mode(const int16_t *src, const evo_planes *pic, int32_t x, int32_t y,
                              int32_t *sad)
{
    image_set set;
    fill_set(&set, *pic);
    std::vector<int16_t> s(src, src + 384), scratch(384);
    macroblock ms = view(s.data()), mt = view(scratch.data());
    return select_best_intra_mode(ms, set, x, y, &mt, sad);
}

// inverse_quantize_macroblock (quantize.cpp:466-476); evr_quantize_mb above
// is its forward counterpart.
void evr_dequantize_mb(int32_t qp, int32_t block_type, const int16_t *levels, int16_t *coeffs)
{
    std::vector<int16_t> l(levels, levels + 384), c(384);
    macroblock ml = view(l.data()), mc = view(c.data());
    inverse_quantize_macroblock((uint8)qp, (EVX_BLOCK_TYPE)block_type, ml, &mc);
    copy_out(mc, coeffs);
}

// trellis_optimize_macroblock (quantize.cpp:615-680) with the scan of
// select_scan_table_8x8 (scan.h:154-159), levels in place.
void evr_trellis_mb(int32_t qp, int32_t block_type, int32_t intra_mode, float lambda,
                    const int16_t *coeffs, int16_t *levels)
{
    std::vector<int16_t> c(coeffs, coeffs + 384), l(levels, levels + 384);
    macroblock mc = view(c.data()), ml = view(l.data());
    trellis_optimize_macroblock((uint8)qp, (EVX_BLOCK_TYPE)block_type, mc,
                                select_scan_table_8x8((uint8)intra_mode), lambda, &ml);
    copy_out(ml, levels);
}

// encode_slice (encode.cpp:751-813) of one frame, run by the reference
// itself on a context set up as engine_encode_frame leaves it: the source in
// the input cache, the references in prediction slots 0 (forward) and 1
// (backward), the picture being reconstructed in slot 2 (its previous
// content from pic_*), the temporal table in prev_block_table.  threads > 1
// runs encode_slice_wpp (encode.cpp:817-941, on hardware_concurrency
// workers) instead.  Descriptors come back as records (reference slot
// numbers mapped to job->target_ids), levels per MB (zero for copies, whose
// output-cache area encode_block leaves untouched) and the reconstruction
// in pic_*.
int32_t evr_encode_frame(const evo_encode_job *job, int16_t *pic_y, int16_t *pic_u,
                         int16_t *pic_v, evo_record *records, int16_t *levels, int32_t threads)
{
    const int32_t w = job->src.width, h = job->src.height;
    const uint32 wb = (uint32)w / 16, hb = (uint32)h / 16, n = wb * hb;
    evx_context ctx;
    if (initialize_context((uint32)w, (uint32)h, &ctx) != EVX_SUCCESS) return 10;
    fill_set(&ctx.cache_bank.input_cache, job->src);
    for (int32 k = 0; k < job->n_refs && k < 2; k++)
        fill_set(&ctx.cache_bank.prediction_cache[k], job->refs[k]);
    evo_planes pic = {pic_y, pic_u, pic_v, w, h, w, w / 2};
    fill_set(&ctx.cache_bank.prediction_cache[2], pic);
    {
        image_set &o = ctx.cache_bank.output_cache;
        memset(o.query_y_image()->query_data(), 0, (size_t)o.query_y_image()->query_row_pitch() * h);
        memset(o.query_u_image()->query_data(), 0, (size_t)o.query_u_image()->query_row_pitch() * (h / 2));
        memset(o.query_v_image()->query_data(), 0, (size_t)o.query_v_image()->query_row_pitch() * (h / 2));
    }
    for (uint32 k = 0; k < n; k++) {
        clear_block_desc(&ctx.block_table[k]);
        clear_block_desc(&ctx.prev_block_table[k]);
        if (job->prev) {
            ctx.prev_block_table[k].block_type = (EVX_BLOCK_TYPE)job->prev[k].block_type;
            ctx.prev_block_table[k].motion_x = job->prev[k].motion_x;
            ctx.prev_block_table[k].motion_y = job->prev[k].motion_y;
        }
    }
    evx_frame frame;
    clear_frame(&frame);
    frame.type = (EVX_FRAME_TYPE)job->type;
    frame.quality = (uint16)job->quality;
    frame.dpb_slot = 2;
    frame.ref_slot_fwd = 0;
    frame.ref_slot_bwd = 1;
    const evx_status st = threads > 1 ? encode_slice_wpp(frame, &ctx) : encode_slice(frame, &ctx);
    if (st != EVX_SUCCESS) {
        clear_context(&ctx);
        return 10;
    }
    auto id_of = [&](uint8 slot) -> uint8_t {
        return slot < 2 && (int32)slot < job->n_refs ? job->target_ids[slot] : (uint8_t)(0xF0 | slot);
    };
    const image_set &o = ctx.cache_bank.output_cache;
    for (uint32 k = 0; k < n; k++) {
        const evx_block_desc &d = ctx.block_table[k];
        evo_record *r = &records[k];
        desc_to_record(d, r);
        r->intra_mode = d.intra_mode;
        if (!(d.block_type & 1)) {
            r->prediction_target = id_of(d.prediction_target);
            if (d.prediction_mode == 2) r->prediction_target_b = id_of(d.prediction_target_b);
        }
        macroblock m;
        create_macroblock(o, (int32)(k % wb) * 16, (int32)(k / wb) * 16, &m);
        copy_out(m, levels + (size_t)k * 384);
    }
    copy_back(ctx.cache_bank.prediction_cache[2], w, h, pic_y, pic_u, pic_v);
    clear_context(&ctx);
    return 0;
}

}  // extern "C"
