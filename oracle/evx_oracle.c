/*
 * evx_oracle.c -- CPU restatement of the Chicago/EVX3 inter analysis path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never part of the
 * product).  Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so that the
 * fp32 RD cost is a multiply followed by an add, as on the reference's x86
 * build with SSE arithmetic and no FMA contraction).
 *
 * Reference citations are path:line relative to the reference's src/.
 * Integer semantics follow the reference's C++ on x86-64: truncating
 * division, arithmetic right shift, int16 stores that truncate, int32 sums
 * that wrap (emulated with uint32 arithmetic where the sum can overflow).
 */
#include "evx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EVO_INT32_MAX 0x7FFFFFFF
#define EVO_MB 16
#define EVO_SAD_THRESHOLD (8 * 1024)        /* motion.cpp:24 */
#define EVO_SEARCH_RADIUS 16                /* motion.cpp:29 */
#define EVO_QUANT_SCALE 16                  /* quantize.cpp:10 */
#define EVO_MAX_QUANT_LEVELS 64             /* quantize.h:42 */
#define EVO_CHROMA_QP_OFFSET 2              /* config.h:68 */

/* 8x8 DCT-II basis scaled by 128: round(128 cos((2k+1) i pi / 16)), row i,
 * column k (xftables.h:57-67). */
static const int16_t kDct8[64] = {
    128,  128,  128,  128,  128,  128,  128,  128,
    126,  106,   71,   25,  -25,  -71, -106, -126,
    118,   49,  -49, -118, -118,  -49,   49,  118,
    106,  -25, -126,  -71,   71,  126,   25, -106,
     91,  -91,  -91,   91,   91,  -91,  -91,   91,
     71, -126,   25,  106, -106,  -25,  126,  -71,
     49, -118,  118,  -49,  -49,  118, -118,   49,
     25,  -71,  106, -126,  126, -106,   71,  -25};

/* Inter quantisation matrix, quantize.cpp:26-36. */
static const int16_t kInterQm[64] = {
    16, 17, 18, 19, 20, 21, 22, 23,
    17, 18, 19, 20, 21, 22, 23, 24,
    18, 19, 20, 21, 22, 23, 24, 25,
    19, 20, 21, 22, 23, 24, 26, 27,
    20, 21, 22, 23, 25, 26, 27, 28,
    21, 22, 23, 24, 26, 27, 28, 30,
    22, 23, 24, 26, 27, 28, 30, 31,
    23, 24, 25, 27, 28, 30, 31, 33};

/* Frequency deadzone weights, quantize.cpp:61-71. */
static const uint8_t kDeadzone[64] = {
    2, 2, 2, 3, 3, 3, 4, 4,
    2, 2, 3, 3, 3, 4, 4, 4,
    2, 3, 3, 3, 4, 4, 4, 4,
    3, 3, 3, 4, 4, 4, 4, 4,
    3, 3, 4, 4, 4, 4, 4, 4,
    3, 4, 4, 4, 4, 4, 4, 4,
    4, 4, 4, 4, 4, 4, 4, 4,
    4, 4, 4, 4, 4, 4, 4, 4};

/* ---------------------------------------------------------------- math -- */

/* evx_math.h:228-236: rounding division, halves away from zero, C `/`. */
static int32_t rounded_div(int32_t numer, int32_t denom)
{
    if (((uint32_t)numer & 0x80000000u) ^ ((uint32_t)denom & 0x80000000u))
        return (numer - denom / 2) / denom;
    return (numer + denom / 2) / denom;
}

/* evx_math.h:197-211 */
static int32_t abs32(int32_t v) { return v == INT32_MIN ? 0x7FFFFFFF : (v < 0 ? -v : v); }
static int16_t abs16(int16_t v) { return v == INT16_MIN ? 0x7FFF : (int16_t)(v < 0 ? -v : v); }

/* evx_math.h:149-154 */
static int16_t sign16(int16_t v) { return (int16_t)((v > 0) - (v < 0)); }

/* evx_math.h:69-138 (EVX_USE_FAST_32BIT_LOG2): floor(log2(v)), log2(0) = 0. */
static uint32_t floor_log2_u32(uint32_t v)
{
    uint32_t r = 0;
    while (v >>= 1) r++;
    return r;
}

static int32_t clip_i32(int32_t v, int32_t lo, int32_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* common.h:264-266: 0.85 * powf(2, (quality - 12) / 3) as the fp32 bits the
 * reference's own build computes (tests/golden/lambda.npz) for the frame
 * qualities 0..63 that set_quality produces (evx3enc.cpp:64-82), so that no
 * host's libm rounding enters the RD costs; other qualities are outside the
 * path and give NaN. */
static const uint32_t kLambdaBits[64] = {
    0x3D59999Au, 0x3D891454u, 0x3DACB591u, 0x3DD9999Au, 0x3E091454u, 0x3E2CB591u,
    0x3E59999Au, 0x3E891455u, 0x3EACB591u, 0x3ED9999Au, 0x3F091455u, 0x3F2CB591u,
    0x3F59999Au, 0x3F891455u, 0x3FACB591u, 0x3FD9999Au, 0x40091455u, 0x402CB591u,
    0x4059999Au, 0x40891454u, 0x40ACB591u, 0x40D9999Au, 0x41091454u, 0x412CB591u,
    0x4159999Au, 0x41891456u, 0x41ACB590u, 0x41D9999Au, 0x42091456u, 0x422CB590u,
    0x4259999Au, 0x42891456u, 0x42ACB590u, 0x42D9999Au, 0x43091456u, 0x432CB590u,
    0x4359999Au, 0x43891452u, 0x43ACB593u, 0x43D9999Au, 0x44091452u, 0x442CB593u,
    0x4459999Au, 0x44891452u, 0x44ACB593u, 0x44D9999Au, 0x45091452u, 0x452CB593u,
    0x4559999Au, 0x45891452u, 0x45ACB593u, 0x45D9999Au, 0x46091452u, 0x462CB593u,
    0x4659999Au, 0x46891452u, 0x46ACB593u, 0x46D9999Au, 0x47091452u, 0x472CB593u,
    0x4759999Au, 0x47891458u, 0x47ACB58Bu, 0x47D9999Au,
};

float evo_lambda(int32_t quality)
{
    float f;
    uint32_t b = quality >= 0 && quality < 64 ? kLambdaBits[quality] : 0x7FC00000u;
    memcpy(&f, &b, sizeof(f));
    return f;
}

/* -------------------------------------------------------- block access -- */

void evo_load_block(const evo_planes *p, int32_t x, int32_t y, int16_t out[EVO_MB_SAMPLES])
{
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++)
            out[r * 16 + c] = p->y[(y + r) * p->stride_y + x + c];
    int32_t cx = x >> 1, cy = y >> 1;
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            out[256 + r * 8 + c] = p->u[(cy + r) * p->stride_c + cx + c];
            out[320 + r * 8 + c] = p->v[(cy + r) * p->stride_c + cx + c];
        }
}

/* ------------------------------------------------------- interpolation -- */

/* interp.cpp:8-14 */
static int16_t sample_clamped(const int16_t *d, int32_t stride, int32_t x, int32_t y,
                              int32_t w, int32_t h)
{
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    return d[y * stride + x];
}

/* interp.cpp:18-28: taps [1,-5,20,20,-5,1], (sum+16)>>5, int16 store. */
static int16_t tap6_h(const int16_t *d, int32_t s, int32_t x, int32_t y, int32_t w, int32_t h)
{
    int32_t sum = sample_clamped(d, s, x - 2, y, w, h) - 5 * sample_clamped(d, s, x - 1, y, w, h) +
                  20 * sample_clamped(d, s, x, y, w, h) + 20 * sample_clamped(d, s, x + 1, y, w, h) -
                  5 * sample_clamped(d, s, x + 2, y, w, h) + sample_clamped(d, s, x + 3, y, w, h);
    return (int16_t)((sum + 16) >> 5);
}

/* interp.cpp:31-41 */
static int16_t tap6_v(const int16_t *d, int32_t s, int32_t x, int32_t y, int32_t w, int32_t h)
{
    int32_t sum = sample_clamped(d, s, x, y - 2, w, h) - 5 * sample_clamped(d, s, x, y - 1, w, h) +
                  20 * sample_clamped(d, s, x, y, w, h) + 20 * sample_clamped(d, s, x, y + 1, w, h) -
                  5 * sample_clamped(d, s, x, y + 2, w, h) + sample_clamped(d, s, x, y + 3, w, h);
    return (int16_t)((sum + 16) >> 5);
}

/* interp.cpp:52-164 for one plane; out has row pitch `bs`. */
static void interp_plane(const int16_t *d, int32_t s, int32_t w, int32_t h, int32_t bx,
                         int32_t by, int32_t bs, int32_t dx, int32_t dy, int32_t quarter,
                         int16_t *out)
{
    int32_t xo = dx > 0 ? 0 : -1;
    int32_t yo = dy > 0 ? 0 : -1;
    int16_t tmp[16 * 21];
    if (dx != 0 && dy != 0) {
        /* interp.cpp:125-131: H pass over rows -2 .. bs+2 (relative to yo). */
        for (int32_t r = -2; r < bs + 3; r++)
            for (int32_t c = 0; c < bs; c++)
                tmp[(r + 2) * bs + c] = tap6_h(d, s, bx + c + xo, by + r + yo, w, h);
    }
    for (int32_t r = 0; r < bs; r++)
        for (int32_t c = 0; c < bs; c++) {
            int16_t half;
            if (dx != 0 && dy == 0) {
                half = tap6_h(d, s, bx + c + xo, by + r, w, h);
            } else if (dx == 0 && dy != 0) {
                half = tap6_v(d, s, bx + c, by + r + yo, w, h);
            } else {
                /* interp.cpp:134-149 */
                const int16_t *t = tmp + r * bs + c;
                int32_t sum = t[0] - 5 * t[bs] + 20 * t[2 * bs] + 20 * t[3 * bs] - 5 * t[4 * bs] +
                              t[5 * bs];
                half = (int16_t)((sum + 16) >> 5);
            }
            if (quarter) {
                /* interp.cpp:74-80: average with the integer sample. */
                int16_t full = sample_clamped(d, s, bx + c, by + r, w, h);
                out[r * bs + c] = (int16_t)((full + half + 1) >> 1);
            } else {
                out[r * bs + c] = half;
            }
        }
}

void evo_interp_block(const evo_planes *ref, int32_t bx, int32_t by, int32_t dir_x,
                      int32_t dir_y, int32_t quarter, int16_t out[EVO_MB_SAMPLES])
{
    /* interp.cpp:186-201: luma 16x16, chroma 8x8 at (bx>>1, by>>1). */
    interp_plane(ref->y, ref->stride_y, ref->width, ref->height, bx, by, 16, dir_x, dir_y,
                 quarter, out);
    int32_t cw = ref->width >> 1, ch = ref->height >> 1;
    interp_plane(ref->u, ref->stride_c, cw, ch, bx >> 1, by >> 1, 8, dir_x, dir_y, quarter,
                 out + 256);
    interp_plane(ref->v, ref->stride_c, cw, ch, bx >> 1, by >> 1, 8, dir_x, dir_y, quarter,
                 out + 320);
}

/* ------------------------------------------------------------- metrics -- */

void evo_hadamard_4x4(const int16_t *src, int32_t stride, int32_t out[16])
{
    /* analysis.h:44-85 */
    int32_t t[16];
    for (int i = 0; i < 4; i++) {
        int32_t a = src[i * stride + 0], b = src[i * stride + 1];
        int32_t c = src[i * stride + 2], d = src[i * stride + 3];
        int32_t s0 = a + b, s1 = a - b, s2 = c + d, s3 = c - d;
        t[i * 4 + 0] = s0 + s2;
        t[i * 4 + 1] = s1 + s3;
        t[i * 4 + 2] = s0 - s2;
        t[i * 4 + 3] = s1 - s3;
    }
    for (int j = 0; j < 4; j++) {
        int32_t a = t[j], b = t[4 + j], c = t[8 + j], d = t[12 + j];
        int32_t s0 = a + b, s1 = a - b, s2 = c + d, s3 = c - d;
        out[j] = s0 + s2;
        out[4 + j] = s1 + s3;
        out[8 + j] = s0 - s2;
        out[12 + j] = s1 - s3;
    }
}

int32_t evo_block_satd(const int16_t *a, const int16_t *b)
{
    /* analysis.h:89-116: the 4x4 difference is stored as int16. */
    int32_t satd = 0;
    int16_t diff[16];
    int32_t coef[16];
    for (int by = 0; by < 16; by += 4)
        for (int bx = 0; bx < 16; bx += 4) {
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++)
                    diff[r * 4 + c] = (int16_t)(a[(by + r) * 16 + bx + c] - b[(by + r) * 16 + bx + c]);
            evo_hadamard_4x4(diff, 4, coef);
            for (int k = 0; k < 16; k++) satd += abs32(coef[k]);
        }
    return (satd + 1) >> 1;
}

int32_t evo_block_sad(const int16_t *a, const int16_t *b)
{
    /* analysis.h:119-132: luma only. */
    int32_t sad = 0;
    for (int k = 0; k < 256; k++) sad += abs32(a[k] - b[k]);
    return sad;
}

int32_t evo_block_mad(const int16_t *a, const int16_t *b)
{
    /* analysis.h:180-202: luma then U/V, running maximum. */
    int32_t mad = 0;
    for (int k = 0; k < 256; k++) {
        int32_t t = abs32(a[k] - b[k]);
        if (t > mad) mad = t;
    }
    for (int k = 0; k < 64; k++) {
        int32_t tu = abs32(a[256 + k] - b[256 + k]);
        int32_t tv = abs32(a[320 + k] - b[320 + k]);
        if (tu > mad) mad = tu;
        if (tv > mad) mad = tv;
    }
    return mad;
}

/* --------------------------------------------------------- mv predictor -- */

/* motion.cpp:263-269 */
static int16_t median3(int16_t a, int16_t b, int16_t c)
{
    if (a > b) { int16_t t = a; a = b; b = t; }
    if (b > c) b = c;
    if (a > b) b = a;
    return b;
}

void evo_mv_predictor(const evo_record *table, uint32_t wb, uint32_t idx, const evo_record *prev,
                      int16_t *pred_x, int16_t *pred_y)
{
    /* motion.cpp:271-346 */
    uint32_t bx = idx % wb, by = idx / wb;
    int16_t mx[3], my[3];
    int n = 0;
    if (bx > 0 && (table[idx - 1].block_type & EVO_BT_MOTION)) {
        mx[n] = table[idx - 1].motion_x; my[n] = table[idx - 1].motion_y; n++;
    }
    if (by > 0 && (table[idx - wb].block_type & EVO_BT_MOTION)) {
        mx[n] = table[idx - wb].motion_x; my[n] = table[idx - wb].motion_y; n++;
    }
    if (by > 0 && bx + 1 < wb && (table[idx - wb + 1].block_type & EVO_BT_MOTION)) {
        mx[n] = table[idx - wb + 1].motion_x; my[n] = table[idx - wb + 1].motion_y; n++;
    } else if (by > 0 && bx > 0 && (table[idx - wb - 1].block_type & EVO_BT_MOTION)) {
        mx[n] = table[idx - wb - 1].motion_x; my[n] = table[idx - wb - 1].motion_y; n++;
    }
    if (n < 3 && prev && (prev[idx].block_type & EVO_BT_MOTION)) {
        mx[n] = prev[idx].motion_x; my[n] = prev[idx].motion_y; n++;
    }
    if (n == 0) {
        *pred_x = 0; *pred_y = 0;
    } else if (n == 1) {
        *pred_x = mx[0]; *pred_y = my[0];
    } else if (n == 2) {
        *pred_x = (int16_t)((mx[0] + mx[1]) / 2);
        *pred_y = (int16_t)((my[0] + my[1]) / 2);
    } else {
        *pred_x = median3(mx[0], mx[1], mx[2]);
        *pred_y = median3(my[0], my[1], my[2]);
    }
}

/* --------------------------------------------------------- inter search -- */

typedef struct search_state {
    /* evx_motion_selection, motion.cpp:53-66 */
    int32_t best_x, best_y, best_sad, best_mad, best_ssd;
    float best_cost;
    int32_t sp_index, sp_amount, sp_enabled;
    /* evx_prediction_params, motion.cpp:43-51 */
    const evo_planes *ref;
    const int16_t *src;
    int32_t px, py, thresh;
    float lambda;
    int32_t n_int, n_axial, n_diag;
} search_state;

/* motion.cpp:125-178.  `counted` is 0 for the re-evaluated centre of a
 * search level, whose outcome is fixed (it is the level's starting best), so
 * the work counter n_int only prices candidates that can change the state. */
static void eval_candidate(search_state *s, int32_t cx, int32_t cy, int counted)
{
    int16_t blk[EVO_MB_SAMPLES];
    evo_load_block(s->ref, cx, cy, blk);
    s->n_int += counted;
    int32_t sad = evo_block_sad(s->src, blk);
    int32_t ssd = (cx - s->px) * (cx - s->px) + (cy - s->py) * (cy - s->py);
    int32_t mad = evo_block_mad(s->src, blk);
    float cost = (float)sad;
    if (s->lambda > 0.0f) {
        int16_t mvx = (int16_t)(cx - s->px), mvy = (int16_t)(cy - s->py);
        int32_t satd = evo_block_satd(s->src, blk);
        float bits = (float)(abs16(mvx) + abs16(mvy) + 1);   /* motion.cpp:120-123 */
        float lb = s->lambda * bits;                            /* no FMA contraction */
        cost = (float)satd + lb;
    }
    int accept;
    if (s->best_mad < s->thresh) {
        accept = mad < s->best_mad || (mad == s->best_mad && ssd < s->best_ssd);
    } else {
        /* motion.cpp:166-168, && binds tighter than || */
        accept = (cost < s->best_cost ||
                  ((cost == s->best_cost && ssd < s->best_ssd) && sad < EVO_SAD_THRESHOLD)) ||
                 mad < s->thresh;
    }
    if (accept) {
        s->best_x = cx; s->best_y = cy;
        s->best_sad = sad; s->best_ssd = ssd; s->best_mad = mad;
        s->best_cost = cost;
    }
}

/* motion.cpp:381-406 */
static void level_search(search_state *s, int32_t step)
{
    int32_t base_x = s->best_x, base_y = s->best_y;
    int32_t max_x = s->ref->width - EVO_MB, max_y = s->ref->height - EVO_MB;
    for (int32_t j = -step; j <= step; j += step)
        for (int32_t i = -step; i <= step; i += step) {
            int32_t cx = base_x + i, cy = base_y + j;
            if (cx < 0 || cx > max_x || cy < 0 || cy > max_y) continue;
            eval_candidate(s, cx, cy, i != 0 || j != 0);
            if (s->best_mad < s->thresh) return;
        }
}

/* motion.cpp:68-91 */
static int32_t frac_index(int32_t i, int32_t j)
{
    i++, j++;
    if (j == 0) return i;
    if (j == 1) return i == 0 ? 3 : 4;
    return i + 5;
}

/* motion.cpp:93-116 */
static void frac_direction(int32_t idx, int32_t *dx, int32_t *dy)
{
    if (idx <= 2) { *dy = -1; *dx = idx - 1; }
    else if (idx == 3) { *dx = -1; *dy = 0; }
    else if (idx == 4) { *dx = 1; *dy = 0; }
    else { *dy = 1; *dx = idx - 6; }
}

/* motion.cpp:180-261, one sub-pel candidate (half when quarter == 0). */
static void subpel_accept(search_state *s, int32_t sad, int32_t mad, int32_t i, int32_t j,
                          int32_t quarter)
{
    int accept;
    if (s->best_mad < s->thresh)
        accept = mad < s->best_mad;
    else
        accept = (sad < s->best_sad && sad < EVO_SAD_THRESHOLD) || mad < s->thresh;
    if (accept) {
        s->sp_enabled = 1;
        s->sp_amount = quarter;
        s->sp_index = frac_index(i, j);
        s->best_sad = sad;
        s->best_mad = mad;
    }
}

/* motion.cpp:450-483 */
static void subpel_search(search_state *s)
{
    int16_t blk[EVO_MB_SAMPLES];
    int32_t max_x = s->ref->width - EVO_MB, max_y = s->ref->height - EVO_MB;
    s->sp_index = 0; s->sp_amount = 0; s->sp_enabled = 0;
    for (int32_t j = -1; j <= 1; j++)
        for (int32_t i = -1; i <= 1; i++) {
            int32_t tx = s->best_x + i, ty = s->best_y + j;
            if (i == 0 && j == 0) continue;
            if (tx < 0 || tx > max_x || ty < 0 || ty > max_y) continue;
            if (i != 0 && j != 0) s->n_diag++; else s->n_axial++;
            for (int32_t q = 0; q < 2; q++) {
                evo_interp_block(s->ref, s->best_x, s->best_y, i, j, q, blk);
                int32_t sad = evo_block_sad(s->src, blk);
                int32_t mad = evo_block_mad(s->src, blk);
                subpel_accept(s, sad, mad, i, j, q);
            }
        }
}

int32_t evo_inter_search(const int16_t src[EVO_MB_SAMPLES], const evo_planes *ref, int32_t x,
                         int32_t y, int32_t quality, int16_t pred_mx, int16_t pred_my,
                         evo_search_result *res)
{
    /* motion.cpp:567-679 */
    search_state s;
    memset(&s, 0, sizeof(s));
    s.ref = ref; s.src = src; s.px = x; s.py = y;
    s.thresh = (quality >> 2) + 1;
    s.lambda = evo_lambda(quality);
    s.best_x = x; s.best_y = y;
    s.best_ssd = EVO_INT32_MAX;

    /* Initial candidate at (x, y), motion.cpp:600-611. */
    int16_t blk[EVO_MB_SAMPLES];
    evo_load_block(ref, x, y, blk);
    s.n_int = 1;
    s.best_sad = evo_block_sad(src, blk);
    s.best_mad = evo_block_mad(src, blk);
    s.best_cost = (float)s.best_sad;
    if (s.lambda > 0.0f) {
        float lb = s.lambda * 1.0f;
        s.best_cost = (float)evo_block_satd(src, blk) + lb;
    }

    /* Predictor candidate, motion.cpp:613-635. */
    if (pred_mx != 0 || pred_my != 0) {
        int32_t cx = x + pred_mx, cy = y + pred_my;
        int32_t max_x = ref->width - EVO_MB, max_y = ref->height - EVO_MB;
        if (cx < 0) cx = 0;
        if (cx > max_x) cx = max_x;
        if (cy < 0) cy = 0;
        if (cy > max_y) cy = max_y;
        eval_candidate(&s, cx, cy, 1);
    }

    /* motion.cpp:638-655 */
    if (s.best_mad >= s.thresh) {
        int32_t radius = EVO_SEARCH_RADIUS;
        if (s.best_sad < EVO_SAD_THRESHOLD / 4) radius = EVO_SEARCH_RADIUS / 4;
        else if (s.best_sad < EVO_SAD_THRESHOLD / 2) radius = EVO_SEARCH_RADIUS / 2;
        for (int32_t i = radius; i > 0; i >>= 1) {
            level_search(&s, i);
            if (s.best_mad < s.thresh) break;
        }
        subpel_search(&s);
    }

    /* motion.cpp:657-678 */
    int32_t bt = 0;
    if (s.best_x != x || s.best_y != y || s.sp_enabled) bt |= EVO_BT_MOTION;
    if (s.best_mad < s.thresh) bt |= EVO_BT_COPY;
    res->best_x = s.best_x; res->best_y = s.best_y;
    res->sad = s.best_sad; res->mad = s.best_mad; res->ssd = s.best_ssd;
    memcpy(&res->cost_bits, &s.best_cost, 4);
    res->sp_enabled = s.sp_enabled; res->sp_amount = s.sp_amount; res->sp_index = s.sp_index;
    res->block_type = bt;
    res->motion_x = s.best_x - x; res->motion_y = s.best_y - y;
    res->n_int = s.n_int; res->n_axial = s.n_axial; res->n_diag = s.n_diag;
    return s.best_sad;
}

/* ------------------------------------------------------- intra search -- */

/* The intra searches read the picture being reconstructed: a candidate
 * that reaches below row y - 16 and right of column x - 16 covers MBs not
 * yet reconstructed and is skipped (motion.cpp:361-365, 433-438). */
static int intra_unstable(const search_state *s, int32_t cx, int32_t cy)
{
    return cy > s->py - EVO_MB && cx > s->px - EVO_MB;
}

/* motion.cpp:348-379 */
static void intra_scan(search_state *s, int32_t left, int32_t top, int32_t right, int32_t bottom,
                       int32_t step)
{
    int32_t base_x = s->best_x, base_y = s->best_y;
    int32_t max_x = s->ref->width - EVO_MB, max_y = s->ref->height - EVO_MB;
    for (int32_t j = top; j <= bottom; j += step)
        for (int32_t i = left; i <= right; i += step) {
            int32_t cx = base_x + i, cy = base_y + j;
            if (intra_unstable(s, cx, cy)) continue;
            if (cx < 0 || cx > max_x || cy < 0 || cy > max_y) continue;
            eval_candidate(s, cx, cy, i != 0 || j != 0);
            if (s->best_mad < s->thresh) return;
        }
}

/* motion.cpp:408-448 */
static void intra_subpel_search(search_state *s)
{
    int16_t blk[EVO_MB_SAMPLES];
    int32_t max_x = s->ref->width - EVO_MB, max_y = s->ref->height - EVO_MB;
    s->sp_index = 0; s->sp_amount = 0; s->sp_enabled = 0;
    for (int32_t j = -1; j <= 1; j++)
        for (int32_t i = -1; i <= 1; i++) {
            int32_t tx = s->best_x + i, ty = s->best_y + j;
            if (i == 0 && j == 0) continue;
            if (intra_unstable(s, tx, ty)) continue;
            if (tx < 0 || tx > max_x || ty < 0 || ty > max_y) continue;
            if (i != 0 && j != 0) s->n_diag++; else s->n_axial++;
            for (int32_t q = 0; q < 2; q++) {
                evo_interp_block(s->ref, s->best_x, s->best_y, i, j, q, blk);
                int32_t sad = evo_block_sad(s->src, blk);
                int32_t mad = evo_block_mad(s->src, blk);
                subpel_accept(s, sad, mad, i, j, q);
            }
        }
}

int32_t evo_intra_search(const int16_t src[EVO_MB_SAMPLES], const evo_planes *pic, int32_t x,
                         int32_t y, int32_t quality, evo_search_result *res)
{
    /* motion.cpp:485-565 */
    search_state s;
    memset(&s, 0, sizeof(s));
    s.ref = pic; s.src = src; s.px = x; s.py = y;
    s.thresh = (quality >> 2) + 1;
    s.lambda = evo_lambda(quality);
    s.best_x = x; s.best_y = y;
    /* compute_block_sad(src) (analysis.h:134-145): luma only. */
    int32_t sad0 = 0;
    for (int32_t k = 0; k < 256; k++) sad0 += src[k] < 0 ? -src[k] : src[k];
    s.best_sad = sad0;
    s.best_mad = EVO_INT32_MAX;
    s.best_ssd = EVO_INT32_MAX;
    s.best_cost = (float)sad0;

    /* The triangle above and to the left: rows y - 32 .. y, columns
     * x - 16 .. x + 16, step 16 (motion.cpp:525-527). */
    intra_scan(&s, -EVO_SEARCH_RADIUS, -2 * EVO_SEARCH_RADIUS, EVO_SEARCH_RADIUS, 0, EVO_SEARCH_RADIUS);
    if (s.best_mad >= s.thresh) {
        for (int32_t i = EVO_SEARCH_RADIUS / 2; i > 0; i >>= 1) {
            intra_scan(&s, -i, -i, i, i, i);
            if (s.best_mad < s.thresh) break;
        }
    }
    /* motion.cpp:540: the sub-pel step runs whatever the integer scans found. */
    intra_subpel_search(&s);

    int32_t bt = EVO_BT_INTRA;
    if (s.best_x != x || s.best_y != y || s.sp_enabled) bt |= EVO_BT_MOTION;
    if (s.best_mad < s.thresh) bt |= EVO_BT_COPY;
    res->best_x = s.best_x; res->best_y = s.best_y;
    res->sad = s.best_sad; res->mad = s.best_mad; res->ssd = s.best_ssd;
    memcpy(&res->cost_bits, &s.best_cost, 4);
    res->sp_enabled = s.sp_enabled; res->sp_amount = s.sp_amount; res->sp_index = s.sp_index;
    res->block_type = bt;
    res->motion_x = s.best_x - x; res->motion_y = s.best_y - y;
    res->n_int = s.n_int; res->n_axial = s.n_axial; res->n_diag = s.n_diag;
    return s.best_sad;
}

void evo_predict_block(const evo_planes *ref, int32_t x, int32_t y, const evo_search_result *r,
                       int16_t out[EVO_MB_SAMPLES])
{
    /* encode.cpp:696-723 (and motion.cpp:729-776 for the bidir halves) */
    int32_t bx = x + r->motion_x, by = y + r->motion_y;
    if (r->sp_enabled) {
        int32_t dx, dy;
        frac_direction(r->sp_index, &dx, &dy);
        evo_interp_block(ref, bx, by, dx, dy, r->sp_amount, out);
    } else {
        evo_load_block(ref, bx, by, out);
    }
}

int32_t evo_bidir_average(const int16_t src[EVO_MB_SAMPLES], const int16_t fwd[EVO_MB_SAMPLES],
                          const int16_t bwd[EVO_MB_SAMPLES], int16_t avg[EVO_MB_SAMPLES])
{
    /* motion.cpp:778-813: SAD over Y, U and V against the int32 average. */
    int32_t sad = 0;
    for (int k = 0; k < EVO_MB_SAMPLES; k++) {
        int32_t a = (fwd[k] + bwd[k] + 1) >> 1;
        avg[k] = (int16_t)a;
        sad += abs32(src[k] - a);
    }
    return sad;
}

/* ---------------------------------------------------------- transform -- */

/* transform.cpp:264-284 */
static void dct_line(const int16_t *src, int32_t sp, int16_t *dst, int32_t dp)
{
    for (int i = 0; i < 8; i++) {
        int32_t t = 0;
        for (int k = 0; k < 8; k++) t += src[k * sp] * kDct8[i * 8 + k];
        t = (i == 0) ? (t * 45) / 128 : t / 2;
        t = rounded_div(t, 128);
        dst[i * dp] = (int16_t)t;
    }
}

void evo_sub_transform_8x8(const int16_t *src, int32_t src_pitch, const int16_t *sub,
                           int32_t sub_pitch, int16_t *dst, int32_t dst_pitch)
{
    /* transform.cpp:435-452 (difference stored as int16, transform.cpp:12-16) */
    int16_t diff[64], tmp[64];
    for (int j = 0; j < 8; j++) {
        for (int i = 0; i < 8; i++)
            diff[j * 8 + i] = (int16_t)(src[j * src_pitch + i] - sub[j * sub_pitch + i]);
        dct_line(diff + j * 8, 1, tmp + j * 8, 1);
    }
    for (int j = 0; j < 8; j++) dct_line(tmp + j, 8, dst + j, dst_pitch);
}

void evo_sub_transform_mb(const int16_t src[EVO_MB_SAMPLES], const int16_t pred[EVO_MB_SAMPLES],
                          int16_t coeffs[EVO_MB_SAMPLES])
{
    /* macroblock.h:272-277, transform.cpp:584-594 */
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        evo_sub_transform_8x8(src + off, 16, pred + off, 16, coeffs + off, 16);
    }
    evo_sub_transform_8x8(src + 256, 8, pred + 256, 8, coeffs + 256, 8);
    evo_sub_transform_8x8(src + 320, 8, pred + 320, 8, coeffs + 320, 8);
}

/* ------------------------------------------------------- quantisation -- */

int32_t evo_variance2(const int16_t coeffs[EVO_MB_SAMPLES])
{
    /* analysis.h:253-275; the int32 sums wrap on overflow. */
    int32_t sum = 0, count = 0;
    uint32_t sumsq = 0;
    for (int k = 1; k < 256; k++) {
        int32_t t = coeffs[k];
        if (t) {
            sum += t;
            sumsq += (uint32_t)(t * t);
            count++;
        }
    }
    if (count == 0) return 0;
    /* rounded_div(sum * sum, count): the square can only be negative after a
     * (undefined) int32 overflow, so the reference's optimised build drops
     * rounded_div's sign test and always rounds as for a non-negative
     * numerator: (sq + count / 2) / count on the wrapped square.  Pinned by
     * the golden vectors written by the reference's -O3 build. */
    int32_t sq = (int32_t)((uint32_t)sum * (uint32_t)sum);
    int32_t rd = (int32_t)((uint32_t)sq + (uint32_t)(count / 2)) / count;
    return (int32_t)(sumsq - (uint32_t)rd);
}

int32_t evo_block_qp(int32_t quality, const int16_t coeffs[EVO_MB_SAMPLES])
{
    /* quantize.cpp:97-118 */
    uint32_t variance = (uint32_t)evo_variance2(coeffs);
    int32_t index = clip_i32((int32_t)(floor_log2_u32(variance) >> 1), 1, EVO_MAX_QUANT_LEVELS - 1);
    int32_t q = (uint8_t)quality;
    int32_t delta = (int16_t)(index - q) >> 1;
    if (delta > 6) delta = 6;
    if (delta < -6) delta = -6;
    return clip_i32(q + delta, 1, EVO_MAX_QUANT_LEVELS - 1);
}

/* quantize.cpp:217-252 */
static void quantize_inter_8x8(int32_t qp, const int16_t *src, int32_t sp, int16_t *dst, int32_t dp)
{
    for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++) {
            int16_t sv = src[j * sp + k];
            int16_t qf = (int16_t)rounded_div(sv * EVO_QUANT_SCALE, kInterQm[j * 8 + k]);
            int32_t w = kDeadzone[j * 8 + k];
            if (qp >= 16) w = w + 1 < 6 ? w + 1 : 6;
            int16_t dz = (int16_t)((sign16(qf) * qp * w) >> 1);
            dst[j * dp + k] = (int16_t)rounded_div(qf - dz, qp << 1);
        }
}

void evo_quantize_inter_mb(int32_t qp, const int16_t coeffs[EVO_MB_SAMPLES],
                           int16_t levels[EVO_MB_SAMPLES])
{
    /* quantize.cpp:372-397 */
    int32_t cqp = qp + EVO_CHROMA_QP_OFFSET;
    if (cqp > EVO_MAX_QUANT_LEVELS - 1) cqp = EVO_MAX_QUANT_LEVELS - 1;
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        quantize_inter_8x8(qp, coeffs + off, 16, levels + off, 16);
    }
    quantize_inter_8x8(cqp, coeffs + 256, 8, levels + 256, 8);
    quantize_inter_8x8(cqp, coeffs + 320, 8, levels + 320, 8);
}

/* ------------------------------------------------------- frame driver -- */

static void fill_primary(evo_record *rec, const evo_search_result *r)
{
    rec->block_type = (uint8_t)r->block_type;
    rec->motion_x = (int16_t)r->motion_x;
    rec->motion_y = (int16_t)r->motion_y;
    rec->sp_pred = (uint8_t)r->sp_enabled;
    rec->sp_amount = (uint8_t)r->sp_amount;
    rec->sp_index = (uint8_t)r->sp_index;
    rec->mad = r->mad;
    rec->cost_bits = r->cost_bits;
}

int32_t evo_analyze_frame(const evo_frame_job *job, evo_record *records, int16_t *coeffs,
                          int16_t *levels)
{
    int32_t wb = job->src.width / EVO_MB, hb = job->src.height / EVO_MB;
    float lambda = evo_lambda(job->quality);
    int16_t src[EVO_MB_SAMPLES], pred[EVO_MB_SAMPLES], pb[EVO_MB_SAMPLES], tmp[EVO_MB_SAMPLES];
    evo_search_result r[EVO_MAX_REFS];

    for (int32_t by = 0; by < hb; by++)
        for (int32_t bx = 0; bx < wb; bx++) {
            uint32_t idx = (uint32_t)(by * wb + bx);
            int32_t x = bx * EVO_MB, y = by * EVO_MB;
            evo_record rec;
            memset(&rec, 0, sizeof(rec));
            evo_load_block(&job->src, x, y, src);

            /* The temporal candidate exists for P frames only (encode.cpp:393, 414). */
            int16_t pmx, pmy;
            evo_mv_predictor(records, (uint32_t)wb, idx, job->type == 1 ? job->prev : NULL, &pmx,
                             &pmy);

            int32_t best_sad = EVO_INT32_MAX, best = -1, bidir = 0;
            uint32_t n_int = 0, n_ax = 0, n_dg = 0;
            if (job->type == 1) {
                /* encode.cpp:215-247 against an empty sentinel (non-copy,
                 * SAD INT32_MAX); with one reference this is 183-210. */
                int32_t best_copy = 0;
                for (int32_t k = 0; k < job->n_refs; k++) {
                    int32_t sad = evo_inter_search(src, &job->refs[k], x, y, job->quality, pmx,
                                                   pmy, &r[k]);
                    n_int += r[k].n_int; n_ax += r[k].n_axial; n_dg += r[k].n_diag;
                    int32_t bias = (int32_t)(lambda * (float)((k + 1) * 2));
                    int32_t copy = (r[k].block_type & EVO_BT_COPY) != 0;
                    if (copy ^ best_copy) {
                        if (copy) { best = k; best_sad = sad; best_copy = 1; }
                    } else if (sad + bias < best_sad) {
                        best = k; best_sad = sad; best_copy = copy;
                    }
                }
                if (best < 0 || best_sad == EVO_INT32_MAX) return 10;
                fill_primary(&rec, &r[best]);
                rec.prediction_mode = 0;
                rec.prediction_target = job->target_ids[best];
                rec.sad = best_sad;
            } else {
                /* encode.cpp:386-443 without the intra candidates. */
                int32_t fs = evo_inter_search(src, &job->refs[0], x, y, job->quality, pmx, pmy, &r[0]);
                int32_t bs = evo_inter_search(src, &job->refs[1], x, y, job->quality, pmx, pmy, &r[1]);
                n_int = r[0].n_int + r[1].n_int;
                n_ax = r[0].n_axial + r[1].n_axial;
                n_dg = r[0].n_diag + r[1].n_diag;
                if (fs < best_sad) { best = 0; best_sad = fs; }
                if (bs < best_sad) { best = 1; best_sad = bs; }
                /* motion.cpp:702-840 reuses the two searches above. */
                evo_predict_block(&job->refs[0], x, y, &r[0], pred);
                evo_predict_block(&job->refs[1], x, y, &r[1], pb);
                int32_t bi_sad = evo_bidir_average(src, pred, pb, tmp);
                int32_t bias = (int32_t)(lambda * 8.0f);
                if (bi_sad + bias < best_sad) { bidir = 1; best_sad = bi_sad; }
                if (best_sad == EVO_INT32_MAX) return 10;
                if (bidir) {
                    fill_primary(&rec, &r[0]);
                    rec.block_type = EVO_BT_MOTION;   /* EVX_BLOCK_INTER_MOTION_DELTA */
                    rec.prediction_mode = 2;
                    rec.prediction_target = job->target_ids[0];
                    rec.prediction_target_b = job->target_ids[1];
                    rec.motion_x_b = (int16_t)r[1].motion_x;
                    rec.motion_y_b = (int16_t)r[1].motion_y;
                    rec.sp_pred_b = (uint8_t)r[1].sp_enabled;
                    rec.sp_amount_b = (uint8_t)r[1].sp_amount;
                    rec.sp_index_b = (uint8_t)r[1].sp_index;
                    rec.sad_b = r[1].sad;
                    rec.mad_b = r[1].mad;
                    rec.cost_b_bits = r[1].cost_bits;
                    memcpy(pred, tmp, sizeof(tmp));
                } else {
                    fill_primary(&rec, &r[best]);
                    rec.prediction_mode = (uint8_t)best;
                    rec.prediction_target = job->target_ids[best];
                }
                rec.sad = best_sad;
            }
            rec.n_int = n_int;
            rec.n_axial = (uint16_t)n_ax;
            rec.n_diag = (uint16_t)n_dg;

            int16_t *co = coeffs + (size_t)idx * EVO_MB_SAMPLES;
            int16_t *lv = levels + (size_t)idx * EVO_MB_SAMPLES;
            if (rec.block_type & EVO_BT_COPY) {
                /* encode.cpp:742-744: copies are neither transformed nor quantised. */
                memset(co, 0, EVO_MB_SAMPLES * sizeof(int16_t));
                memset(lv, 0, EVO_MB_SAMPLES * sizeof(int16_t));
            } else {
                if (!bidir) {
                    const evo_planes *ref = &job->refs[job->type == 1 ? best : best];
                    evo_predict_block(ref, x, y, &r[best], pred);
                }
                /* encode.cpp:717-728 */
                evo_sub_transform_mb(src, pred, co);
                int32_t qp = evo_block_qp(job->quality, co);
                rec.q_index = (uint8_t)qp;
                rec.variance = (int16_t)evo_variance2(co);
                evo_quantize_inter_mb(qp, co, lv);
                rec.flags = 1u | (bidir ? 2u : 0u);
            }
            records[idx] = rec;
        }
    return 0;
}

/* ----------------------------------------------------- reconstruction -- */

/* quantize.cpp:322-332 */
static void dequantize_inter_8x8(int32_t qp, const int16_t *src, int32_t sp, int16_t *dst,
                                 int32_t dp)
{
    for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++) {
            int32_t v = src[j * sp + k];
            dst[j * dp + k] = (int16_t)(((2 * v) * kInterQm[j * 8 + k] * qp) / EVO_QUANT_SCALE);
        }
}

void evo_dequantize_inter_mb(int32_t qp, const int16_t levels[EVO_MB_SAMPLES],
                             int16_t coeffs[EVO_MB_SAMPLES])
{
    /* quantize.cpp:427-449 */
    int32_t cqp = qp + EVO_CHROMA_QP_OFFSET;
    if (cqp > EVO_MAX_QUANT_LEVELS - 1) cqp = EVO_MAX_QUANT_LEVELS - 1;
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        dequantize_inter_8x8(qp, levels + off, 16, coeffs + off, 16);
    }
    dequantize_inter_8x8(cqp, levels + 256, 8, coeffs + 256, 8);
    dequantize_inter_8x8(cqp, levels + 320, 8, coeffs + 320, 8);
}

/* transform.cpp:330-347 (and 396-417 with the addend): sample i of a line
 * from its 8 frequencies, each term truncated on its own. */
static int32_t idct_sample(const int16_t *src, int32_t sp, int i)
{
    int32_t t = ((src[0] * kDct8[i]) * 45) / 128;
    for (int k = 1; k < 8; k++) t += (src[k * sp] * kDct8[k * 8 + i]) / 2;
    return rounded_div(t, 128);
}

static void inverse_transform_add_8x8(const int16_t *src, int32_t sp, const int16_t *add,
                                      int32_t ap, int16_t *dst, int32_t dp)
{
    /* transform.cpp:419-433: vertical lines into a scratch block, then
     * horizontal lines plus the addend. */
    int16_t tmp[64];
    for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++) tmp[i * 8 + j] = (int16_t)idct_sample(src + j, sp, i);
    for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++)
            dst[j * dp + i] = (int16_t)(idct_sample(tmp + j * 8, 1, i) + add[j * ap + i]);
}

void evo_inverse_transform_add_mb(const int16_t coeffs[EVO_MB_SAMPLES],
                                  const int16_t pred[EVO_MB_SAMPLES],
                                  int16_t out[EVO_MB_SAMPLES])
{
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        inverse_transform_add_8x8(coeffs + off, 16, pred + off, 16, out + off, 16);
    }
    inverse_transform_add_8x8(coeffs + 256, 8, pred + 256, 8, out + 256, 8);
    inverse_transform_add_8x8(coeffs + 320, 8, pred + 320, 8, out + 320, 8);
}

static int ref_of_target(const evo_frame_job *job, int32_t target)
{
    for (int k = 0; k < job->n_refs; k++)
        if (job->target_ids[k] == target) return k;
    return -1;
}

int32_t evo_reconstruct_frame(const evo_frame_job *job, const evo_record *records,
                              const int16_t *levels, int16_t *out_y, int16_t *out_u,
                              int16_t *out_v)
{
    int32_t w = job->src.width, h = job->src.height, wb = w / EVO_MB, hb = h / EVO_MB;
    int16_t pred[EVO_MB_SAMPLES], pb[EVO_MB_SAMPLES], co[EVO_MB_SAMPLES], rec[EVO_MB_SAMPLES];
    for (int32_t by = 0; by < hb; by++)
        for (int32_t bx = 0; bx < wb; bx++) {
            int32_t idx = by * wb + bx, x = bx * EVO_MB, y = by * EVO_MB;
            const evo_record *r = &records[idx];
            int k = ref_of_target(job, r->prediction_target);
            if (k < 0) return 3;
            evo_search_result s;
            memset(&s, 0, sizeof(s));
            s.motion_x = r->motion_x; s.motion_y = r->motion_y;
            s.sp_enabled = r->sp_pred; s.sp_amount = r->sp_amount; s.sp_index = r->sp_index;
            evo_predict_block(&job->refs[k], x, y, &s, pred);
            if (r->prediction_mode == 2) {
                /* decode.cpp:126-197, 272-330: average with the backward block. */
                int kb = ref_of_target(job, r->prediction_target_b);
                if (kb < 0) return 3;
                s.motion_x = r->motion_x_b; s.motion_y = r->motion_y_b;
                s.sp_enabled = r->sp_pred_b; s.sp_amount = r->sp_amount_b;
                s.sp_index = r->sp_index_b;
                evo_predict_block(&job->refs[kb], x, y, &s, pb);
                for (int i = 0; i < EVO_MB_SAMPLES; i++)
                    pred[i] = (int16_t)((pred[i] + pb[i] + 1) >> 1);
            }
            if (r->block_type & EVO_BT_COPY) {
                memcpy(rec, pred, sizeof(rec));        /* decode.cpp:123-269 */
            } else {
                evo_dequantize_inter_mb(r->q_index, levels + (size_t)idx * EVO_MB_SAMPLES, co);
                evo_inverse_transform_add_mb(co, pred, rec);
            }
            for (int j = 0; j < 16; j++)
                memcpy(out_y + (size_t)(y + j) * w + x, rec + j * 16, 16 * sizeof(int16_t));
            for (int j = 0; j < 8; j++) {
                memcpy(out_u + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec + 256 + j * 8, 8 * sizeof(int16_t));
                memcpy(out_v + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec + 320 + j * 8, 8 * sizeof(int16_t));
            }
        }
    return 0;
}

/* ------------------------------------------------------------ deblock -- */

#define EVO_DEBLOCK_STEP 8                  /* deblock.cpp:9 */
#define EVO_DEBLOCK_QP_OFFSET 2             /* config.h:74 */

/* deblock.cpp:13-35: edge thresholds by qp index. */
static const int16_t kDeblockAlpha[EVO_MAX_QUANT_LEVELS] = {
    0,   0,   0,   0,   0,   0,   0,   1,   1,   1,   2,   2,   3,   3,   4,   5,
    6,   7,   8,   9,   10,  12,  14,  16,  18,  20,  22,  24,  26,  29,  32,  35,
    38,  42,  46,  50,  54,  59,  64,  70,  76,  82,  89,  96,  104, 112, 120, 128,
    138, 148, 158, 170, 182, 196, 210, 224, 240, 255, 255, 255, 255, 255, 255, 255};
static const int16_t kDeblockBeta[EVO_MAX_QUANT_LEVELS] = {
    0,  0,  0,  0,  0,  0,  0,  0,  1,  1,  1,  1,  2,  2,  2,  3,
    3,  3,  4,  4,  4,  5,  5,  6,  6,  7,  7,  8,  8,  9,  10, 11,
    12, 13, 14, 15, 16, 17, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18,
    18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18};

/* deblock.cpp:91-165 */
void evo_deblock_line(int16_t v[8], int32_t qp, int32_t strength, int32_t is_luma)
{
    const int32_t p3 = v[0], p2 = v[1], p1 = v[2], p0 = v[3];
    const int32_t q0 = v[4], q1 = v[5], q2 = v[6], q3 = v[7];
    int32_t qi = qp;
    if (strength == 2 && qp > 0)
        qi = qp + EVO_DEBLOCK_QP_OFFSET < EVO_MAX_QUANT_LEVELS - 1 ? qp + EVO_DEBLOCK_QP_OFFSET
                                                                  : EVO_MAX_QUANT_LEVELS - 1;
    /* The differences are stored as int16 before the comparisons. */
    const int16_t d_pq = (int16_t)abs32(p0 - q0);
    const int16_t d_p = (int16_t)abs32(p1 - p0);
    const int16_t d_q = (int16_t)abs32(q1 - q0);
    if (d_pq >= kDeblockAlpha[qi] || d_p >= kDeblockBeta[qi] || d_q >= kDeblockBeta[qi]) return;
    if (strength == 2) {
        v[3] = (int16_t)rounded_div(p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1, 8);
        v[2] = (int16_t)rounded_div(p2 + p1 + p0 + q0, 4);
        v[4] = (int16_t)rounded_div(p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2, 8);
        v[5] = (int16_t)rounded_div(p0 + q0 + q1 + q2, 4);
        if (is_luma) {
            v[1] = (int16_t)rounded_div(2 * p3 + 3 * p2 + p1 + p0 + q0, 8);
            v[6] = (int16_t)rounded_div(2 * q3 + 3 * q2 + q1 + q0 + p0, 8);
        }
    } else if (strength == 1) {
        v[3] = (int16_t)rounded_div(((q0 + p0) * 4) + p1 - q1, 8);
        v[4] = (int16_t)rounded_div(((q0 + p0) * 4) + q1 - p1, 8);
        if (is_luma) {
            v[2] = (int16_t)rounded_div((p2 * 4) + (p0 * 2) + (q0 * 2), 8);
            v[5] = (int16_t)rounded_div((q2 * 4) + (q0 * 2) + (p0 * 2), 8);
        }
    }
}

/* deblock.cpp:57-89 with 167-220: strength (0, 1, 2) and average qp of the
 * edge between blocks a (left / top) and b; `internal` when both sides lie
 * in one block of the table. */
static int32_t deblock_edge(const evo_record *a, const evo_record *b, int32_t internal,
                            int32_t *avg_qp)
{
    const int32_t ca = (a->block_type & EVO_BT_COPY) != 0, cb = (b->block_type & EVO_BT_COPY) != 0;
    if (!ca && !cb) *avg_qp = (uint8_t)((a->q_index + b->q_index) >> 1);
    else if (!ca) *avg_qp = a->q_index;
    else if (!cb) *avg_qp = b->q_index;
    else *avg_qp = 0;
    if (ca && cb) return 0;
    if (ca ^ cb) return 1;
    return internal ? 1 : 2;
}

/* deblock.cpp:167-220: filter the 8 lines of one edge segment in place,
 * gathering each line's samples at stride `along` (1: across a vertical
 * edge; the row pitch: across a horizontal one) and stepping `next`. */
static void deblock_segment(int16_t *at, int32_t along, int32_t next, int32_t avg_qp,
                            int32_t strength, int32_t is_luma, int32_t qp_offset)
{
    int32_t qp = avg_qp + qp_offset;
    if (qp > EVO_MAX_QUANT_LEVELS - 1) qp = EVO_MAX_QUANT_LEVELS - 1;
    for (int32_t l = 0; l < EVO_DEBLOCK_STEP; l++, at += next) {
        int16_t v[8];
        for (int32_t t = 0; t < 8; t++) v[t] = at[(t - 4) * along];
        evo_deblock_line(v, qp, strength, is_luma);
        for (int32_t t = 1; t < 7; t++) at[(t - 4) * along] = v[t];
    }
}

/* deblock.cpp:222-276: one plane in the reference's order -- the vertical
 * edges of the first band of 8 rows, then per band its horizontal edges and,
 * after each horizontal segment from the second on, the vertical edge at the
 * same column.  Blocks of the table are `mb` samples square. */
static void deblock_plane(const evo_record *records, int32_t wb, int32_t mb, int32_t is_luma,
                          int16_t *d, int32_t w, int32_t h, int32_t qp_offset)
{
    int32_t avg, s;
#define EVO_BLK(i, j) (&records[(i) / mb + ((j) / mb) * wb])
    for (int32_t i = EVO_DEBLOCK_STEP; i < w; i += EVO_DEBLOCK_STEP) {
        s = deblock_edge(EVO_BLK(i - 1, 0), EVO_BLK(i, 0), (i - 1) / mb == i / mb, &avg);
        if (s) deblock_segment(d + i, 1, w, avg, s, is_luma, qp_offset);
    }
    for (int32_t j = EVO_DEBLOCK_STEP; j < h; j += EVO_DEBLOCK_STEP) {
        int16_t *row = d + (size_t)j * w;
        s = deblock_edge(EVO_BLK(0, j - 1), EVO_BLK(0, j), (j - 1) / mb == j / mb, &avg);
        if (s) deblock_segment(row, w, 1, avg, s, is_luma, qp_offset);
        for (int32_t i = EVO_DEBLOCK_STEP; i < w; i += EVO_DEBLOCK_STEP) {
            s = deblock_edge(EVO_BLK(i, j - 1), EVO_BLK(i, j), (j - 1) / mb == j / mb, &avg);
            if (s) deblock_segment(row + i, w, 1, avg, s, is_luma, qp_offset);
            s = deblock_edge(EVO_BLK(i - 1, j), EVO_BLK(i, j), (i - 1) / mb == i / mb, &avg);
            if (s) deblock_segment(row + i, 1, w, avg, s, is_luma, qp_offset);
        }
    }
#undef EVO_BLK
}

/* deblock.cpp:278-306 */
void evo_deblock_frame(const evo_record *records, int32_t width, int32_t height, int16_t *y,
                       int16_t *u, int16_t *v)
{
    const int32_t wb = width / EVO_MB;
    deblock_plane(records, wb, EVO_MB, 1, y, width, height, 0);
    deblock_plane(records, wb, EVO_MB / 2, 0, u, width / 2, height / 2, EVO_CHROMA_QP_OFFSET);
    deblock_plane(records, wb, EVO_MB / 2, 0, v, width / 2, height / 2, EVO_CHROMA_QP_OFFSET);
}

/* ---------------------------------------------------------------- SAO -- */

#define EVO_SAO_BAND_COUNT 32                /* sao.h:43-60 */
#define EVO_SAO_MAX_OFFSET 7
#define EVO_SAO_NUM_OFFSETS 4
#define EVO_SAO_CAT_NONE 4

/* sao.cpp:12-16 */
static int16_t sao_clamp(int32_t v) { return v < 0 ? 0 : (v > 255 ? 255 : (int16_t)v); }

/* sao.cpp:19-24: neighbours a and b of each edge class (dx_a, dy_a, dx_b, dy_b). */
static const int8_t kSaoNeighbours[4][4] = {
    {-1, 0, 1, 0}, {0, -1, 0, 1}, {-1, -1, 1, 1}, {1, -1, -1, 1}};

/* sao.cpp:26-37: valley 0, concave 1, convex 2, peak 3, none 4. */
static int32_t sao_category(int16_t pixel, int16_t a, int16_t b)
{
    const int32_t s = ((a > pixel) - (a < pixel)) + ((b > pixel) - (b < pixel));
    return s == 2 ? 0 : s == 1 ? 1 : s == -1 ? 2 : s == -2 ? 3 : EVO_SAO_CAT_NONE;
}

/* The category of sample (x, y) under edge class `cls` in plane d (w x h,
 * stride w), or none when a neighbour lies outside the plane. */
static int32_t sao_edge_category(const int16_t *d, int32_t w, int32_t h, int32_t x, int32_t y,
                                 int32_t cls)
{
    const int32_t ax = x + kSaoNeighbours[cls][0], ay = y + kSaoNeighbours[cls][1];
    const int32_t bx = x + kSaoNeighbours[cls][2], by = y + kSaoNeighbours[cls][3];
    if (ax < 0 || ax >= w || ay < 0 || ay >= h || bx < 0 || bx >= w || by < 0 || by >= h)
        return EVO_SAO_CAT_NONE;
    return sao_category(d[(size_t)y * w + x], d[(size_t)ay * w + ax], d[(size_t)by * w + bx]);
}

/* The corrected value of sample (x, y) of `rec` under params p (sao.cpp:39-53
 * and 55-96 for one sample, classified on `rec`). */
static int16_t sao_corrected(const int16_t *rec, int32_t w, int32_t h, int32_t x, int32_t y,
                             const evo_sao_params *p)
{
    const int16_t pixel = rec[(size_t)y * w + x];
    if (p->type == EVO_SAO_BAND) {
        const int32_t idx = (sao_clamp(pixel) >> 3) - (int32_t)p->band_position;
        if (idx >= 0 && idx < EVO_SAO_NUM_OFFSETS) return sao_clamp(pixel + p->offsets[idx]);
    } else if (p->type == EVO_SAO_EDGE) {
        const int32_t c = sao_edge_category(rec, w, h, x, y, p->eo_class);
        if (c < EVO_SAO_NUM_OFFSETS) return sao_clamp(pixel + p->offsets[c]);
    }
    return pixel;
}

/* sao.cpp:98-134: one plane in raster order, in place (blocks of `mb`
 * samples; edge categories of a block from the plane as it stands before the
 * block is written). */
static void sao_filter_plane(const evo_sao_info *table, int32_t wb, int32_t hb, int16_t *d,
                             int32_t w, int32_t h, int32_t mb, int32_t luma)
{
    int8_t cat[EVO_MB * EVO_MB];
    for (int32_t by = 0; by < hb; by++)
        for (int32_t bx = 0; bx < wb; bx++) {
            const int32_t idx = by * wb + bx;
            const evo_sao_info *info = &table[idx];
            const evo_sao_info *src = info;
            if (info->merge_left && bx > 0) src = &table[idx - 1];
            else if (info->merge_above && by > 0) src = &table[idx - wb];
            const evo_sao_params *p = luma ? &src->luma : &src->chroma;
            if (p->type == EVO_SAO_OFF) continue;
            const int32_t x0 = bx * mb, y0 = by * mb;
            const int32_t bw = x0 + mb > w ? w - x0 : mb, bh = y0 + mb > h ? h - y0 : mb;
            if (p->type == EVO_SAO_BAND) {
                for (int32_t j = y0; j < y0 + bh; j++)
                    for (int32_t i = x0; i < x0 + bw; i++)
                        d[(size_t)j * w + i] = sao_corrected(d, w, h, i, j, p);
            } else if (p->type == EVO_SAO_EDGE) {
                for (int32_t j = 0; j < bh; j++)
                    for (int32_t i = 0; i < bw; i++)
                        cat[j * bw + i] = (int8_t)sao_edge_category(d, w, h, x0 + i, y0 + j, p->eo_class);
                for (int32_t j = 0; j < bh; j++)
                    for (int32_t i = 0; i < bw; i++) {
                        const int32_t c = cat[j * bw + i];
                        int16_t *s = &d[(size_t)(y0 + j) * w + x0 + i];
                        if (c < EVO_SAO_NUM_OFFSETS) *s = sao_clamp(*s + p->offsets[c]);
                    }
            }
        }
}

/* sao.cpp:136-146 */
static uint32_t sao_bits(const evo_sao_params *p)
{
    if (p->type == EVO_SAO_OFF) return 2;
    uint32_t bits = 2 + (p->type == EVO_SAO_EDGE ? 2 : 5);
    for (int k = 0; k < EVO_SAO_NUM_OFFSETS; k++) bits += p->offsets[k] != 0 ? 4 : 1;
    return bits;
}

/* SSD of a block of `src` against `rec` corrected by p (sao.cpp:148-161 with
 * p off; the SSD loops of 183-239, 241-314 and 333-376). */
static int64_t sao_ssd(const int16_t *src, const int16_t *rec, int32_t w, int32_t h, int32_t x0,
                       int32_t y0, int32_t bw, int32_t bh, const evo_sao_params *p)
{
    int64_t ssd = 0;
    for (int32_t j = y0; j < y0 + bh; j++)
        for (int32_t i = x0; i < x0 + bw; i++) {
            const int32_t diff = (int32_t)src[(size_t)j * w + i] - sao_corrected(rec, w, h, i, j, p);
            ssd += diff * diff;
        }
    return ssd;
}

/* The offsets of a class from its sums and counts (sao.cpp:199-214,
 * 274-289): rounded mean of src - rec, clamped to +-7; 0 if no nonzero. */
static int32_t sao_offsets(const int32_t *sum, const int32_t *cnt, int8_t *off)
{
    int32_t nonzero = 0;
    for (int k = 0; k < EVO_SAO_NUM_OFFSETS; k++) {
        int32_t avg = 0;
        if (cnt[k] > 0) {
            avg = (sum[k] + (sum[k] >= 0 ? cnt[k] / 2 : -cnt[k] / 2)) / cnt[k];
            if (avg > EVO_SAO_MAX_OFFSET) avg = EVO_SAO_MAX_OFFSET;
            if (avg < -EVO_SAO_MAX_OFFSET) avg = -EVO_SAO_MAX_OFFSET;
        }
        off[k] = (int8_t)avg;
        nonzero |= avg != 0;
    }
    return nonzero;
}

static void sao_consider(const evo_sao_params *cand, int64_t ssd, float lambda,
                         evo_sao_params *best, float *best_cost)
{
    const float cost = (float)ssd + lambda * (float)sao_bits(cand);
    if (cost < *best_cost) {
        *best_cost = cost;
        *best = *cand;
    }
}

/* sao.cpp:316-331 with 163-239 (band) and 241-314 (edge): the best params of
 * one block of one plane. */
static void sao_rdo_block(const int16_t *src, const int16_t *rec, int32_t w, int32_t h,
                          int32_t x0, int32_t y0, int32_t bw, int32_t bh, float lambda,
                          evo_sao_params *out)
{
    memset(out, 0, sizeof(*out));
    float best = (float)sao_ssd(src, rec, w, h, x0, y0, bw, bh, out) + lambda * (float)sao_bits(out);
    int32_t bsum[EVO_SAO_BAND_COUNT] = {0}, bcnt[EVO_SAO_BAND_COUNT] = {0};
    for (int32_t j = y0; j < y0 + bh; j++)
        for (int32_t i = x0; i < x0 + bw; i++) {
            const int16_t pixel = rec[(size_t)j * w + i];
            const int32_t band = sao_clamp(pixel) >> 3;
            bsum[band] += (int32_t)src[(size_t)j * w + i] - pixel;
            bcnt[band]++;
        }
    for (int32_t bp = 0; bp <= EVO_SAO_BAND_COUNT - EVO_SAO_NUM_OFFSETS; bp++) {
        evo_sao_params c;
        memset(&c, 0, sizeof(c));
        c.type = EVO_SAO_BAND;
        c.band_position = (uint8_t)bp;
        if (!sao_offsets(bsum + bp, bcnt + bp, c.offsets)) continue;
        sao_consider(&c, sao_ssd(src, rec, w, h, x0, y0, bw, bh, &c), lambda, out, &best);
    }
    for (int32_t cls = 0; cls < 4; cls++) {
        int32_t csum[EVO_SAO_NUM_OFFSETS] = {0}, ccnt[EVO_SAO_NUM_OFFSETS] = {0};
        for (int32_t j = y0; j < y0 + bh; j++)
            for (int32_t i = x0; i < x0 + bw; i++) {
                const int32_t cat = sao_edge_category(rec, w, h, i, j, cls);
                if (cat < EVO_SAO_NUM_OFFSETS) {
                    csum[cat] += (int32_t)src[(size_t)j * w + i] - rec[(size_t)j * w + i];
                    ccnt[cat]++;
                }
            }
        evo_sao_params c;
        memset(&c, 0, sizeof(c));
        c.type = EVO_SAO_EDGE;
        c.eo_class = (uint8_t)cls;
        if (!sao_offsets(csum, ccnt, c.offsets)) continue;
        sao_consider(&c, sao_ssd(src, rec, w, h, x0, y0, bw, bh, &c), lambda, out, &best);
    }
}

/* sao.cpp:333-376 */
static float sao_merge_cost(const int16_t *src, const int16_t *rec, int32_t w, int32_t h,
                            int32_t x0, int32_t y0, int32_t bw, int32_t bh,
                            const evo_sao_params *p, float lambda)
{
    return (float)sao_ssd(src, rec, w, h, x0, y0, bw, bh, p) + lambda * 1.0f;
}

void evo_sao_filter_frame(const evo_sao_info *table, int32_t width, int32_t height, int16_t *y,
                          int16_t *u, int16_t *v)
{
    const int32_t wb = width / EVO_MB, hb = height / EVO_MB;
    sao_filter_plane(table, wb, hb, y, width, height, EVO_MB, 1);
    sao_filter_plane(table, wb, hb, u, width / 2, height / 2, EVO_MB / 2, 0);
    sao_filter_plane(table, wb, hb, v, width / 2, height / 2, EVO_MB / 2, 0);
}

/* sao.cpp:378-590 (EVX_ENABLE_SAO_CHROMA_MERGE 0): per MB in raster order
 * the best luma params, the best chroma params from the U plane, then a
 * merge with the left or upper MB's (final) params when its luma cost is
 * lower; finally the table is applied to rec in place. */
void evo_sao_encode_frame(const int16_t *src_y, const int16_t *src_u, const int16_t *src_v,
                          int16_t *rec_y, int16_t *rec_u, int16_t *rec_v, int32_t width,
                          int32_t height, float lambda, evo_sao_info *table)
{
    (void)src_v;
    const int32_t wb = width / EVO_MB, hb = height / EVO_MB, cw = width / 2, chh = height / 2;
    for (int32_t by = 0; by < hb; by++)
        for (int32_t bx = 0; bx < wb; bx++) {
            const int32_t idx = by * wb + bx;
            evo_sao_info *info = &table[idx];
            memset(info, 0, sizeof(*info));
            const int32_t lx0 = bx * EVO_MB, ly0 = by * EVO_MB;
            const int32_t lw = lx0 + EVO_MB > width ? width - lx0 : EVO_MB;
            const int32_t lh = ly0 + EVO_MB > height ? height - ly0 : EVO_MB;
            sao_rdo_block(src_y, rec_y, width, height, lx0, ly0, lw, lh, lambda, &info->luma);
            const int32_t cx0 = bx * (EVO_MB / 2), cy0 = by * (EVO_MB / 2);
            const int32_t cwb = cx0 + EVO_MB / 2 > cw ? cw - cx0 : EVO_MB / 2;
            const int32_t chb = cy0 + EVO_MB / 2 > chh ? chh - cy0 : EVO_MB / 2;
            sao_rdo_block(src_u, rec_u, cw, chh, cx0, cy0, cwb, chb, lambda, &info->chroma);
            float cur;
            if (info->luma.type == EVO_SAO_OFF) {
                cur = (float)sao_ssd(src_y, rec_y, width, height, lx0, ly0, lw, lh, &info->luma) +
                      lambda * 2.0f;
            } else {
                cur = sao_merge_cost(src_y, rec_y, width, height, lx0, ly0, lw, lh, &info->luma, lambda) -
                      lambda * 1.0f + lambda * (float)sao_bits(&info->luma);
            }
            if (bx > 0 && table[idx - 1].luma.type != EVO_SAO_OFF) {
                const float mc = sao_merge_cost(src_y, rec_y, width, height, lx0, ly0, lw, lh,
                                                &table[idx - 1].luma, lambda);
                if (mc < cur) {
                    info->merge_left = 1;
                    info->luma = table[idx - 1].luma;
                    info->chroma = table[idx - 1].chroma;
                    cur = mc;
                }
            }
            if (!info->merge_left && by > 0 && table[idx - wb].luma.type != EVO_SAO_OFF) {
                const float mc = sao_merge_cost(src_y, rec_y, width, height, lx0, ly0, lw, lh,
                                                &table[idx - wb].luma, lambda);
                if (mc < cur) {
                    info->merge_above = 1;
                    info->luma = table[idx - wb].luma;
                    info->chroma = table[idx - wb].chroma;
                }
            }
        }
    evo_sao_filter_frame(table, width, height, rec_y, rec_u, rec_v);
}

/* ===================================================================== */
/* Mode decision and the in-loop encode of a frame (encode.cpp:39-370,   */
/* 479-748, 751-813; decode.cpp:31-433; intra_pred.cpp; the trellis of   */
/* quantize.cpp:478-682).                                                */
/* ===================================================================== */

#define EVO_BT_INTRA_DEFAULT EVO_BT_INTRA   /* types.h:83 */

/* Intra quantisation matrix, quantize.cpp:13-23. */
static const int16_t kIntraQm[64] = {
     8, 10, 12, 15, 19, 23, 25, 27,
    10, 12, 14, 17, 21, 25, 27, 28,
    12, 14, 18, 21, 24, 26, 28, 30,
    15, 17, 21, 24, 26, 28, 30, 32,
    19, 21, 24, 26, 28, 30, 32, 35,
    23, 25, 26, 28, 30, 32, 35, 38,
    25, 27, 28, 30, 32, 35, 38, 41,
    27, 28, 30, 32, 35, 38, 41, 45};

/* Intra deadzone weights, quantize.cpp:42-52. */
static const uint8_t kIntraDeadzone[64] = {
    0, 1, 1, 1, 1, 1, 2, 2,
    1, 1, 1, 1, 1, 2, 2, 2,
    1, 1, 1, 1, 2, 2, 2, 2,
    1, 1, 1, 2, 2, 2, 2, 3,
    1, 1, 2, 2, 2, 2, 3, 3,
    1, 2, 2, 2, 2, 3, 3, 3,
    2, 2, 2, 2, 3, 3, 3, 3,
    2, 2, 2, 3, 3, 3, 3, 3};

/* evx_math.h:64 with the byte log2 of 69-87: floor(log2(n)) + 1 of the
 * value truncated to 8 bits, log2(0) = 0. */
static uint32_t required_bits_u8(uint32_t n)
{
    return floor_log2_u32(n & 0xFFu) + 1;
}

/* ------------------------------------------------ intra prediction -- */

/* intra_pred.cpp:12-52: the row above, the column left and the corner of
 * the n x n block at (cx, cy) of one plane, 128 where absent. */
static void intra_edges(const int16_t *d, int32_t stride, int32_t cx, int32_t cy, int32_t n,
                        int16_t *top, int16_t *left, int16_t *corner, int *has_top,
                        int *has_left)
{
    *has_top = cy > 0;
    *has_left = cx > 0;
    for (int32_t i = 0; i < n; i++) {
        top[i] = *has_top ? d[(cy - 1) * stride + cx + i] : 128;
        left[i] = *has_left ? d[(cy + i) * stride + cx - 1] : 128;
    }
    if (*has_top && *has_left) *corner = d[(cy - 1) * stride + cx - 1];
    else if (*has_top) *corner = top[0];
    else if (*has_left) *corner = left[0];
    else *corner = 128;
}

/* intra_pred.cpp:58-171: one plane's n x n prediction (out has stride n). */
static void intra_predict_plane(int32_t mode, const int16_t *top, const int16_t *left,
                                int16_t corner, int has_top, int has_left, int32_t n,
                                int16_t *out)
{
    int16_t ref[33];
    if (mode == 4 || mode == 5 || mode == 6) {
        /* left[n-1] .. left[0], corner, top[0] .. top[n-1] */
        for (int32_t i = 0; i < n; i++) ref[i] = left[n - 1 - i];
        ref[n] = corner;
        for (int32_t i = 0; i < n; i++) ref[n + 1 + i] = top[i];
    }
    for (int32_t y = 0; y < n; y++)
        for (int32_t x = 0; x < n; x++) {
            int32_t v;
            switch (mode) {
            case 0: { /* DC */
                int32_t sum = 0, count = 0;
                if (has_top) { for (int32_t i = 0; i < n; i++) sum += top[i]; count += n; }
                if (has_left) { for (int32_t i = 0; i < n; i++) sum += left[i]; count += n; }
                v = count > 0 ? (int16_t)((sum + count / 2) / count) : 128;
            } break;
            case 1: v = top[x]; break;            /* vertical */
            case 2: v = left[y]; break;           /* horizontal */
            case 3: {                             /* diagonal down-left */
                int32_t i0 = x + y;
                int32_t a = top[i0 < n - 1 ? i0 : n - 1];
                int32_t b = top[i0 + 1 < n - 1 ? i0 + 1 : n - 1];
                int32_t c = top[i0 + 2 < n - 1 ? i0 + 2 : n - 1];
                v = (a + 2 * b + c + 2) >> 2;
            } break;
            case 4: {                             /* diagonal down-right */
                int32_t d = x - y + n;
                v = (ref[d - 1] + 2 * ref[d] + ref[d + 1] + 2) >> 2;
            } break;
            case 5:                               /* vertical-right */
            case 6: {                             /* horizontal-down */
                int32_t base = mode == 5 ? x - (y >> 1) : (x >> 1) - y;
                int32_t odd = mode == 5 ? (y & 1) : (x & 1);
                int32_t d = base + n;
                if (d < 1) d = 1;
                if (d > 2 * n - 1) d = 2 * n - 1;
                v = odd ? (ref[d - 1] + 2 * ref[d] + ref[d + 1] + 2) >> 2
                        : (ref[d] + ref[d + 1] + 1) >> 1;
            } break;
            default: v = 128; break;
            }
            out[y * n + x] = (int16_t)v;
        }
}

void evo_intra_prediction(int32_t mode, const evo_planes *pic, int32_t x, int32_t y,
                          int16_t out[EVO_MB_SAMPLES])
{
    /* intra_pred.cpp:177-215 */
    int16_t top[16], left[16], corner;
    int has_top, has_left;
    intra_edges(pic->y, pic->stride_y, x, y, 16, top, left, &corner, &has_top, &has_left);
    intra_predict_plane(mode, top, left, corner, has_top, has_left, 16, out);
    intra_edges(pic->u, pic->stride_c, x >> 1, y >> 1, 8, top, left, &corner, &has_top, &has_left);
    intra_predict_plane(mode, top, left, corner, has_top, has_left, 8, out + 256);
    intra_edges(pic->v, pic->stride_c, x >> 1, y >> 1, 8, top, left, &corner, &has_top, &has_left);
    intra_predict_plane(mode, top, left, corner, has_top, has_left, 8, out + 320);
}

void evo_intra_mode_candidates(const evo_planes *pic, int32_t x, int32_t y, uint8_t cand[3])
{
    /* intra_pred.cpp:304-326 (also the candidate set of 227-268): luma edge
     * gradients pick a horizontal, vertical or diagonal family. */
    int16_t top[16], left[16], corner;
    int has_top, has_left;
    intra_edges(pic->y, pic->stride_y, x, y, 16, top, left, &corner, &has_top, &has_left);
    int32_t gh = 0, gv = 0;
    if (has_top)
        for (int i = 1; i < 16; i++) gh += abs32(top[i] - top[i - 1]);
    if (has_left)
        for (int i = 1; i < 16; i++) gv += abs32(left[i] - left[i - 1]);
    cand[0] = 0;
    if (gv > gh * 3 / 2) { cand[1] = 2; cand[2] = 6; }
    else if (gh > gv * 3 / 2) { cand[1] = 1; cand[2] = 5; }
    else { cand[1] = 3; cand[2] = 4; }
}

int32_t evo_select_intra_mode(const int16_t src[EVO_MB_SAMPLES], const evo_planes *pic,
                              int32_t x, int32_t y, int32_t *sad)
{
    /* intra_pred.cpp:217-302 (EVX_ENABLE_FAST_INTRA_MODE): the first of the
     * three candidates with the smallest luma SAD. */
    uint8_t cand[3];
    int16_t pred[EVO_MB_SAMPLES];
    evo_intra_mode_candidates(pic, x, y, cand);
    int32_t best_sad = EVO_INT32_MAX, best_mode = 0;
    for (int c = 0; c < 3; c++) {
        evo_intra_prediction(cand[c], pic, x, y, pred);
        int32_t s = evo_block_sad(src, pred);
        if (s < best_sad) { best_sad = s; best_mode = cand[c]; }
    }
    if (sad) *sad = best_sad;
    return best_mode;
}

/* ------------------------------------------ intra (de)quantisation -- */

/* quantize.cpp:74-95 */
static int32_t luma_dc_scale(int32_t qp)
{
    if (qp < 5) return (qp << 1) > 2 ? (qp << 1) : 2;
    if (qp < 9) return qp << 1;
    if (qp < 25) return qp + 8;
    return (qp << 1) - 16;
}

static int32_t chroma_dc_scale(int32_t qp)
{
    if (qp < 5) return (qp << 1) > 2 ? (qp << 1) : 2;
    if (qp < 25) return (qp + 13) >> 1;
    return qp - 6;
}

/* quantize.cpp:120-200 */
static void quantize_intra_8x8(int32_t qp, int32_t chroma, const int16_t *src, int32_t sp,
                               int16_t *dst, int32_t dp)
{
    for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++) {
            int16_t qf = (int16_t)rounded_div(src[j * sp + k] * EVO_QUANT_SCALE, kIntraQm[j * 8 + k]);
            if (qp >= 16 && !(k == 0 && j == 0)) {
                int16_t dz = (int16_t)((sign16(qf) * qp * kIntraDeadzone[j * 8 + k]) >> 1);
                qf = (int16_t)(qf - dz);
            }
            dst[j * dp + k] = (int16_t)rounded_div(qf, qp << 1);
        }
    dst[0] = (int16_t)rounded_div(src[0], chroma ? chroma_dc_scale(qp) : luma_dc_scale(qp));
}

/* quantize.cpp:271-301 */
static void dequantize_intra_8x8(int32_t qp, int32_t chroma, const int16_t *src, int32_t sp,
                                 int16_t *dst, int32_t dp)
{
    for (int j = 0; j < 8; j++)
        for (int k = 0; k < 8; k++)
            dst[j * dp + k] = (int16_t)((2 * src[j * sp + k] * kIntraQm[j * 8 + k] * qp) / EVO_QUANT_SCALE);
    dst[0] = (int16_t)(src[0] * (chroma ? chroma_dc_scale(qp) : luma_dc_scale(qp)));
}

static int is_intra_quant(int32_t bt)
{
    /* quantize.cpp:457, 469: intra matrices for intra blocks without motion. */
    return (bt & EVO_BT_INTRA) && !(bt & EVO_BT_MOTION);
}

void evo_quantize_mb(int32_t qp, int32_t block_type, const int16_t coeffs[EVO_MB_SAMPLES],
                     int16_t levels[EVO_MB_SAMPLES])
{
    /* quantize.cpp:454-464 with 345-370 */
    if (!is_intra_quant(block_type)) {
        evo_quantize_inter_mb(qp, coeffs, levels);
        return;
    }
    int32_t cqp = qp + EVO_CHROMA_QP_OFFSET;
    if (cqp > EVO_MAX_QUANT_LEVELS - 1) cqp = EVO_MAX_QUANT_LEVELS - 1;
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        quantize_intra_8x8(qp, 0, coeffs + off, 16, levels + off, 16);
    }
    quantize_intra_8x8(cqp, 1, coeffs + 256, 8, levels + 256, 8);
    quantize_intra_8x8(cqp, 1, coeffs + 320, 8, levels + 320, 8);
}

void evo_dequantize_mb(int32_t qp, int32_t block_type, const int16_t levels[EVO_MB_SAMPLES],
                       int16_t coeffs[EVO_MB_SAMPLES])
{
    /* quantize.cpp:466-476 with 399-424 */
    if (!is_intra_quant(block_type)) {
        evo_dequantize_inter_mb(qp, levels, coeffs);
        return;
    }
    int32_t cqp = qp + EVO_CHROMA_QP_OFFSET;
    if (cqp > EVO_MAX_QUANT_LEVELS - 1) cqp = EVO_MAX_QUANT_LEVELS - 1;
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        dequantize_intra_8x8(qp, 0, levels + off, 16, coeffs + off, 16);
    }
    dequantize_intra_8x8(cqp, 1, levels + 256, 8, coeffs + 256, 8);
    dequantize_intra_8x8(cqp, 1, levels + 320, 8, coeffs + 320, 8);
}

/* ------------------------------------------------------------ trellis -- */

/* scan.h:52-62 (zig-zag), 117-127 (rows), 137-147 (columns); the scan of a
 * block follows its intra mode (scan.h:154-159). */
static void scan_order(int32_t intra_mode, uint8_t order[64])
{
    if (intra_mode == 1 || intra_mode == 5) {
        for (int p = 0; p < 64; p++) order[p] = (uint8_t)p;
    } else if (intra_mode == 2 || intra_mode == 6) {
        for (int p = 0; p < 64; p++) order[p] = (uint8_t)((p & 7) * 8 + (p >> 3));
    } else {
        /* JPEG zig-zag: anti-diagonals d = r + c, odd ones walked downwards. */
        int p = 0;
        for (int d = 0; d < 15; d++)
            for (int t = 0; t <= d; t++) {
                int r = (d & 1) ? t : d - t;
                int c = d - r;
                if (r < 8 && c < 8) order[p++] = (uint8_t)(r * 8 + c);
            }
    }
}

/* quantize.cpp:481-514 */
static int16_t trellis_dequant(int16_t level, int32_t qp, int32_t pos, int intra, int chroma)
{
    if (level == 0) return 0;
    if (intra) {
        if (pos == 0) return (int16_t)(level * (chroma ? chroma_dc_scale(qp) : luma_dc_scale(qp)));
        return (int16_t)((2 * level * kIntraQm[pos] * qp) / EVO_QUANT_SCALE);
    }
    return (int16_t)((2 * level * kInterQm[pos] * qp) / EVO_QUANT_SCALE);
}

/* quantize.cpp:517-532 (EVX_ENABLE_SIGMAP_CODING) */
static uint32_t coeff_bits(int16_t level)
{
    if (level == 0) return 1;
    int32_t a = abs32(level);
    uint32_t v = (uint32_t)(a < 255 ? a : 255);
    uint32_t bits = 2;
    if (v > 1) bits += 1;
    if (v > 2) bits += 1 + 2 * required_bits_u8(v - 2);
    else bits += 1;
    return bits;
}

static float trellis_cost(int16_t coeff, int16_t level, int32_t qp, int32_t pos, int intra,
                          int chroma, float lambda)
{
    int32_t diff = (int32_t)coeff - (int32_t)trellis_dequant(level, qp, pos, intra, chroma);
    int32_t sq = (int32_t)((uint32_t)diff * (uint32_t)diff);
    float rate = lambda * (float)coeff_bits(level);
    return (float)sq + rate;
}

/* quantize.cpp:536-613 on one 8x8 block (coefficients and levels with
 * stride sp). */
static void trellis_8x8(int32_t qp, int intra, int chroma, const int16_t *co, int16_t *lv,
                        int32_t sp, const uint8_t *scan, float lambda)
{
#define AT(a, pos) (a)[((pos) >> 3) * sp + ((pos) & 7)]
    int32_t last = -1;
    for (int32_t p = 63; p >= 0; p--)
        if (AT(lv, scan[p]) != 0) { last = p; break; }
    if (last < 0) return;
    for (int32_t p = last; p >= 1; p--) {
        int32_t pos = scan[p];
        int16_t level = AT(lv, pos);
        if (level == 0) continue;
        float keep = trellis_cost(AT(co, pos), level, qp, pos, intra, chroma, lambda);
        float zero = trellis_cost(AT(co, pos), 0, qp, pos, intra, chroma, lambda);
        if (zero <= keep) AT(lv, pos) = 0;
        else break;
    }
    last = -1;
    for (int32_t p = 63; p >= 0; p--)
        if (AT(lv, scan[p]) != 0) { last = p; break; }
    if (last < 0) return;
    for (int32_t p = 1; p <= last; p++) {
        int32_t pos = scan[p];
        int16_t level = AT(lv, pos);
        if (level == 0) continue;
        int16_t reduced = (int16_t)(level - sign16(level));
        float keep = trellis_cost(AT(co, pos), level, qp, pos, intra, chroma, lambda);
        float red = trellis_cost(AT(co, pos), reduced, qp, pos, intra, chroma, lambda);
        if (red < keep) AT(lv, pos) = reduced;
    }
#undef AT
}

void evo_trellis_mb(int32_t qp, int32_t block_type, int32_t intra_mode, float lambda,
                    const int16_t coeffs[EVO_MB_SAMPLES], int16_t levels[EVO_MB_SAMPLES])
{
    /* quantize.cpp:615-680 */
    uint8_t scan[64];
    scan_order(intra_mode, scan);
    int intra = is_intra_quant(block_type);
    int32_t cqp = qp + EVO_CHROMA_QP_OFFSET;
    if (cqp > EVO_MAX_QUANT_LEVELS - 1) cqp = EVO_MAX_QUANT_LEVELS - 1;
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        trellis_8x8(qp, intra, 0, coeffs + off, levels + off, 16, scan, lambda);
    }
    trellis_8x8(cqp, intra, 1, coeffs + 256, levels + 256, 8, scan, lambda);
    trellis_8x8(cqp, intra, 1, coeffs + 320, levels + 320, 8, scan, lambda);
}

/* --------------------------------------------------------- rate model -- */

static uint32_t mv_bits(int32_t mx, int32_t my, int32_t sp_pred)
{
    /* encode.cpp:49-53, 65-69 */
    uint32_t bits = 2 * required_bits_u8((uint32_t)(abs32(mx) + 1)) +
                    2 * required_bits_u8((uint32_t)(abs32(my) + 1));
    return bits + (sp_pred ? 5 : 1);
}

uint32_t evo_block_bits(const evo_record *d, const int16_t block[EVO_MB_SAMPLES])
{
    /* encode.cpp:39-110 (EVX_ENABLE_SIGMAP_CODING, EVX_ENABLE_MPM) */
    int32_t bt = d->block_type;
    uint32_t bits = 3;
    if (!(bt & EVO_BT_INTRA)) bits += 2;
    if (bt & EVO_BT_MOTION) bits += mv_bits(d->motion_x, d->motion_y, d->sp_pred);
    if (bt & EVO_BT_COPY) return bits;
    if (bt == EVO_BT_INTRA_DEFAULT) bits += 3;
    bits += 2 * required_bits_u8((uint32_t)d->q_index + 1);
    for (int b = 0; b < 4; b++) {
        int off = (b >> 1) * 8 * 16 + (b & 1) * 8;
        uint32_t nz = 0, cb = 0;
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                int16_t v = block[off + r * 16 + c];
                if (v != 0) {
                    nz++;
                    cb += coeff_bits(v);
                }
            }
        bits += required_bits_u8(nz + 1) * 2 + 1;
        bits += cb;
    }
    return bits;
}

/* ------------------------------------------------- frame encode driver -- */

typedef struct encode_ctx {
    const evo_encode_job *job;
    const evo_planes *pic;         /* the picture being reconstructed */
    float lambda;
} encode_ctx;

/* The MB prediction a descriptor names (encode.cpp:485-723, decode.cpp:
 * 35-433): intra modes over the picture being reconstructed, intra motion
 * over the same picture, inter blocks over their references (averaged for
 * bidirectional records).  Returns 0 if the record names an unknown target. */
static int prediction_of(const encode_ctx *e, const evo_record *d, int32_t x, int32_t y,
                         int16_t out[EVO_MB_SAMPLES])
{
    const evo_encode_job *job = e->job;
    int32_t bt = d->block_type;
    if (bt == EVO_BT_INTRA_DEFAULT) {
        if (d->intra_mode != EVO_INTRA_NONE) evo_intra_prediction(d->intra_mode, e->pic, x, y, out);
        else memset(out, 0, EVO_MB_SAMPLES * sizeof(int16_t));
        return 1;
    }
    evo_search_result s;
    memset(&s, 0, sizeof(s));
    s.motion_x = d->motion_x; s.motion_y = d->motion_y;
    s.sp_enabled = d->sp_pred; s.sp_amount = d->sp_amount; s.sp_index = d->sp_index;
    if (bt & EVO_BT_INTRA) {
        evo_predict_block(e->pic, x, y, &s, out);
        return 1;
    }
    int k = -1, kb = -1;
    for (int i = 0; i < job->n_refs; i++) {
        if (job->target_ids[i] == d->prediction_target && k < 0) k = i;
        if (job->target_ids[i] == d->prediction_target_b && kb < 0) kb = i;
    }
    if (k < 0) return 0;
    evo_predict_block(&job->refs[k], x, y, &s, out);
    if (job->type == 2 && d->prediction_mode == 2) {
        int16_t pb[EVO_MB_SAMPLES];
        if (kb < 0) return 0;
        s.motion_x = d->motion_x_b; s.motion_y = d->motion_y_b;
        s.sp_enabled = d->sp_pred_b; s.sp_amount = d->sp_amount_b; s.sp_index = d->sp_index_b;
        evo_predict_block(&job->refs[kb], x, y, &s, pb);
        for (int i = 0; i < EVO_MB_SAMPLES; i++) out[i] = (int16_t)((out[i] + pb[i] + 1) >> 1);
    }
    return 1;
}

/* encode_block (encode.cpp:479-748): residual, QP, variance, quantisation
 * and trellis into lv; copies are left alone (lv zeroed here). */
static int encode_mb(const encode_ctx *e, const int16_t *src, evo_record *d, int32_t x,
                     int32_t y, int16_t co[EVO_MB_SAMPLES], int16_t lv[EVO_MB_SAMPLES])
{
    int16_t pred[EVO_MB_SAMPLES];
    if (d->block_type & EVO_BT_COPY) {
        memset(lv, 0, EVO_MB_SAMPLES * sizeof(int16_t));
        memset(co, 0, EVO_MB_SAMPLES * sizeof(int16_t));
        return 1;
    }
    if (!prediction_of(e, d, x, y, pred)) return 0;
    evo_sub_transform_mb(src, pred, co);
    int32_t qp = evo_block_qp(e->job->quality, co);
    d->q_index = (uint8_t)qp;
    d->variance = (int16_t)evo_variance2(co);
    evo_quantize_mb(qp, d->block_type, co, lv);
    evo_trellis_mb(qp, d->block_type, d->intra_mode, e->lambda, co, lv);
    return 1;
}

/* decode_block (decode.cpp:31-433): the reconstruction of one MB. */
static int decode_mb(const encode_ctx *e, const evo_record *d, int32_t x, int32_t y,
                     const int16_t lv[EVO_MB_SAMPLES], int16_t out[EVO_MB_SAMPLES])
{
    int16_t pred[EVO_MB_SAMPLES], co[EVO_MB_SAMPLES];
    if (!prediction_of(e, d, x, y, pred)) return 0;
    if (d->block_type & EVO_BT_COPY) {
        memcpy(out, pred, sizeof(pred));
        return 1;
    }
    evo_dequantize_mb(d->q_index, d->block_type, lv, co);
    evo_inverse_transform_add_mb(co, pred, out);
    return 1;
}

/* compute_rdo_cost (encode.cpp:128-166) of a candidate, which gets its
 * q_index and variance as encode_block sets them.  Distortion is the luma
 * SSD of the reconstruction.  The reference decodes into the buffer that
 * holds the candidate's levels (both are views of the staging cache at
 * (0, 0), encode.cpp:150-156), so the residual rate is estimated from the
 * reconstructed luma samples. */
static int rdo_cost(const encode_ctx *e, const int16_t *src, evo_record *d, int32_t x, int32_t y,
                    float *cost)
{
    int16_t co[EVO_MB_SAMPLES], lv[EVO_MB_SAMPLES], rec[EVO_MB_SAMPLES];
    if (!encode_mb(e, src, d, x, y, co, lv)) return 0;
    if (!decode_mb(e, d, x, y, lv, rec)) return 0;
    uint32_t ssd = 0;
    for (int k = 0; k < 256; k++) {
        int32_t t = src[k] - rec[k];
        ssd += (uint32_t)t * (uint32_t)t;
    }
    uint32_t rate = evo_block_bits(d, rec);
    float rt = e->lambda * (float)rate;
    *cost = (float)(int32_t)ssd + rt;
    return 1;
}

static void desc_clear(evo_record *d)
{
    /* clear_block_desc (common.cpp:78-94) */
    memset(d, 0, sizeof(*d));
    d->block_type = EVO_BT_INTRA_DEFAULT;
    d->intra_mode = EVO_INTRA_NONE;
}

static void desc_from_search(evo_record *d, const evo_search_result *r, uint8_t target)
{
    desc_clear(d);
    d->block_type = (uint8_t)r->block_type;
    d->prediction_target = target;
    d->motion_x = (int16_t)r->motion_x;
    d->motion_y = (int16_t)r->motion_y;
    d->sp_pred = (uint8_t)r->sp_enabled;
    d->sp_amount = (uint8_t)r->sp_amount;
    d->sp_index = (uint8_t)r->sp_index;
}

int32_t evo_encode_frame(const evo_encode_job *job, int16_t *pic_y, int16_t *pic_u,
                         int16_t *pic_v, evo_record *records, int16_t *levels, int16_t *coeffs)
{
    int32_t w = job->src.width, h = job->src.height, wb = w / EVO_MB, hb = h / EVO_MB;
    evo_planes pic = {pic_y, pic_u, pic_v, w, h, w, w / 2};
    encode_ctx e = {job, &pic, evo_lambda(job->quality)};
    if (job->type < 0 || job->type > 2 || job->n_refs != (job->type == 0 ? 0 : job->type))
        return 3;
    int16_t src[EVO_MB_SAMPLES], co[EVO_MB_SAMPLES], rec[EVO_MB_SAMPLES];
    int16_t pf[EVO_MB_SAMPLES], pb[EVO_MB_SAMPLES], avg[EVO_MB_SAMPLES];

    for (int32_t by = 0; by < hb; by++)
        for (int32_t bx = 0; bx < wb; bx++) {
            uint32_t idx = (uint32_t)(by * wb + bx);
            int32_t x = bx * EVO_MB, y = by * EVO_MB;
            evo_load_block(&job->src, x, y, src);
            evo_record best;
            evo_search_result ri, r[2];
            uint32_t n_int, n_ax, n_dg;

            /* calculate_intra_prediction (motion.cpp:485-565) over the
             * picture being reconstructed: every frame type starts with it
             * (encode.cpp:175, 362). */
            int32_t best_sad = evo_intra_search(src, &pic, x, y, job->quality, &ri);
            desc_from_search(&best, &ri, 0);
            n_int = ri.n_int; n_ax = ri.n_axial; n_dg = ri.n_diag;

            if (job->type == 1) {
                /* encode.cpp:183-210: one reference, with the temporal table. */
                int16_t pmx, pmy;
                evo_mv_predictor(records, (uint32_t)wb, idx, job->prev, &pmx, &pmy);
                int32_t sad = evo_inter_search(src, &job->refs[0], x, y, job->quality, pmx, pmy, &r[0]);
                n_int += r[0].n_int; n_ax += r[0].n_axial; n_dg += r[0].n_diag;
                int inter_copy = (r[0].block_type & EVO_BT_COPY) != 0;
                int best_copy = (best.block_type & EVO_BT_COPY) != 0;
                if (inter_copy ^ best_copy) {
                    if (inter_copy) { desc_from_search(&best, &r[0], job->target_ids[0]); best_sad = sad; }
                } else if (sad < best_sad) {
                    desc_from_search(&best, &r[0], job->target_ids[0]);
                    best_sad = sad;
                }
            } else if (job->type == 2) {
                /* encode.cpp:372-458: forward, backward, then the average of
                 * the same two searches (motion.cpp:702-840). */
                int16_t pmx, pmy;
                evo_mv_predictor(records, (uint32_t)wb, idx, NULL, &pmx, &pmy);
                int32_t fs = evo_inter_search(src, &job->refs[0], x, y, job->quality, pmx, pmy, &r[0]);
                int32_t bs = evo_inter_search(src, &job->refs[1], x, y, job->quality, pmx, pmy, &r[1]);
                n_int += r[0].n_int + r[1].n_int;
                n_ax += r[0].n_axial + r[1].n_axial;
                n_dg += r[0].n_diag + r[1].n_diag;
                if (fs < best_sad) {
                    desc_from_search(&best, &r[0], job->target_ids[0]);
                    best.prediction_mode = 0;
                    best_sad = fs;
                }
                if (bs < best_sad) {
                    desc_from_search(&best, &r[1], job->target_ids[1]);
                    best.prediction_mode = 1;
                    best_sad = bs;
                }
                evo_predict_block(&job->refs[0], x, y, &r[0], pf);
                evo_predict_block(&job->refs[1], x, y, &r[1], pb);
                int32_t bi_sad = evo_bidir_average(src, pf, pb, avg);
                if (bi_sad + (int32_t)(e.lambda * 8.0f) < best_sad) {
                    desc_clear(&best);
                    best.block_type = EVO_BT_MOTION;   /* EVX_BLOCK_INTER_MOTION_DELTA */
                    best.prediction_mode = 2;
                    best.prediction_target = job->target_ids[0];
                    best.motion_x = (int16_t)r[0].motion_x;
                    best.motion_y = (int16_t)r[0].motion_y;
                    best.sp_pred = (uint8_t)r[0].sp_enabled;
                    best.sp_amount = (uint8_t)r[0].sp_amount;
                    best.sp_index = (uint8_t)r[0].sp_index;
                    best.prediction_target_b = job->target_ids[1];
                    best.motion_x_b = (int16_t)r[1].motion_x;
                    best.motion_y_b = (int16_t)r[1].motion_y;
                    best.sp_pred_b = (uint8_t)r[1].sp_enabled;
                    best.sp_amount_b = (uint8_t)r[1].sp_amount;
                    best.sp_index_b = (uint8_t)r[1].sp_index;
                    best_sad = bi_sad;
                }
            }
            if (best_sad == EVO_INT32_MAX) return 10;

            /* The directional intra candidate (encode.cpp:254-272, 446-468). */
            int32_t dir_sad;
            int32_t mode = evo_select_intra_mode(src, &pic, x, y, &dir_sad);
            if (dir_sad < best_sad) {
                desc_clear(&best);
                best.intra_mode = (uint8_t)mode;
                best_sad = dir_sad;
            }

            float best_cost = 0.0f;
            if (job->type == 1) {
                /* RDO (encode.cpp:274-348, EVX_ENABLE_RDO_COPY_EVAL and
                 * EVX_ENABLE_RDO_INTRA_MODES): the SAD winner, then the plain
                 * intra block unless the winner is a directional one, then
                 * the three gradient-pruned modes; strictly lower cost wins. */
                if (!rdo_cost(&e, src, &best, x, y, &best_cost)) return 3;
                if (best.block_type != EVO_BT_INTRA_DEFAULT) {
                    evo_record cand;
                    float c;
                    desc_clear(&cand);
                    if (!rdo_cost(&e, src, &cand, x, y, &c)) return 3;
                    if (c < best_cost) { best = cand; best_cost = c; }
                }
                uint8_t modes[3];
                evo_intra_mode_candidates(&pic, x, y, modes);
                for (int c = 0; c < 3; c++) {
                    evo_record cand;
                    float cc;
                    desc_clear(&cand);
                    cand.intra_mode = modes[c];
                    if (!rdo_cost(&e, src, &cand, x, y, &cc)) return 3;
                    if (cc < best_cost) { best = cand; best_cost = cc; }
                }
            }

            /* encode_block, then decode_block into the picture
             * (encode.cpp:795-811). */
            int16_t *lv = levels + (size_t)idx * EVO_MB_SAMPLES;
            if (!encode_mb(&e, src, &best, x, y, co, lv)) return 3;
            if (coeffs) memcpy(coeffs + (size_t)idx * EVO_MB_SAMPLES, co, sizeof(co));
            if (!decode_mb(&e, &best, x, y, lv, rec)) return 3;
            for (int j = 0; j < 16; j++)
                memcpy(pic_y + (size_t)(y + j) * w + x, rec + j * 16, 16 * sizeof(int16_t));
            for (int j = 0; j < 8; j++) {
                memcpy(pic_u + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec + 256 + j * 8, 8 * sizeof(int16_t));
                memcpy(pic_v + (size_t)(y / 2 + j) * (w / 2) + x / 2, rec + 320 + j * 8, 8 * sizeof(int16_t));
            }

            best.sad = best_sad;
            memcpy(&best.cost_bits, &best_cost, 4);
            best.n_int = n_int;
            best.n_axial = (uint16_t)n_ax;
            best.n_diag = (uint16_t)n_dg;
            best.flags = ((best.block_type & EVO_BT_COPY) ? 0u : 1u) |
                         (best.prediction_mode == 2 ? 2u : 0u) | 4u;
            records[idx] = best;
        }
    return 0;
}
