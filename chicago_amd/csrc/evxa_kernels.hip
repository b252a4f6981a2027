// evxa_kernels.hip -- gfx950 kernels of the EVX3 inter analysis path.
//
// One wavefront (64 lanes, one workgroup) analyses one macroblock at a time.
// Candidate metrics run wide: the wave splits into four 16-lane groups, each
// group evaluating one candidate position with one lane per 4x4 luma block
// (plus 8 chroma samples), and DPP row reductions give each group's SAD,
// SATD and MAD.  The decision is the reference's ordered fold over the
// candidates (motion.cpp:125-178, 180-261), done on wave-uniform values.
//
// Work is scheduled by a persistent work list: waves take MB tickets in
// anti-diagonal order (t = bx + 2*by, all jobs of the launch interleaved) and
// wait until the MB's left and top-right neighbours are published, because
// the MV predictor reads their final records (motion.cpp:271-346).  A MB's
// record is published before its DCT/quantisation, which is off the
// dependency chain.
//
// All arithmetic is integer except the RD cost, which is an fp32 multiply
// followed by an fp32 add with round-to-nearest and no contraction
// (motion.cpp:144), as on the reference's x86 build.

#include "evxa_internal.h"

namespace evxa {
namespace {

constexpr int kSadThreshold = 8 * 1024;  // motion.cpp:24
constexpr int kSearchRadius = 16;        // motion.cpp:29
constexpr int kInt32Max = 0x7FFFFFFF;
constexpr int kSpinLimit = 1 << 22;

// xftables.h:57-67: round(128 cos((2k+1) i pi / 16)).
__constant__ int16_t c_dct8[64] = {
    128, 128,  128,  128,  128,  128,  128,  128,
    126, 106,   71,   25,  -25,  -71, -106, -126,
    118,  49,  -49, -118, -118,  -49,   49,  118,
    106, -25, -126,  -71,   71,  126,   25, -106,
     91, -91,  -91,   91,   91,  -91,  -91,   91,
     71, -126,  25,  106, -106,  -25,  126,  -71,
     49, -118, 118,  -49,  -49,  118, -118,   49,
     25,  -71, 106, -126,  126, -106,   71,  -25};

// quantize.cpp:26-36
__constant__ int16_t c_inter_qm[64] = {
    16, 17, 18, 19, 20, 21, 22, 23, 17, 18, 19, 20, 21, 22, 23, 24,
    18, 19, 20, 21, 22, 23, 24, 25, 19, 20, 21, 22, 23, 24, 26, 27,
    20, 21, 22, 23, 25, 26, 27, 28, 21, 22, 23, 24, 26, 27, 28, 30,
    22, 23, 24, 26, 27, 28, 30, 31, 23, 24, 25, 27, 28, 30, 31, 33};

// quantize.cpp:61-71
__constant__ uint8_t c_deadzone[64] = {
    2, 2, 2, 3, 3, 3, 4, 4, 2, 2, 3, 3, 3, 4, 4, 4, 2, 3, 3, 3, 4, 4, 4, 4,
    3, 3, 3, 4, 4, 4, 4, 4, 3, 3, 4, 4, 4, 4, 4, 4, 3, 4, 4, 4, 4, 4, 4, 4,
    4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <int CTRL>
__device__ __forceinline__ int dpp(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// Reductions over each 16-lane row (quad swaps, half-row and row mirrors).
// Sums wrap modulo 2^32 like the reference's int32 accumulators.
__device__ __forceinline__ int row_sum(int v)
{
    uint32_t u = (uint32_t)v;
    u += (uint32_t)dpp<0xB1>((int)u);
    u += (uint32_t)dpp<0x4E>((int)u);
    u += (uint32_t)dpp<0x141>((int)u);
    u += (uint32_t)dpp<0x140>((int)u);
    return (int)u;
}

__device__ __forceinline__ int row_max(int v)
{
    v = max(v, dpp<0xB1>(v));
    v = max(v, dpp<0x4E>(v));
    v = max(v, dpp<0x141>(v));
    v = max(v, dpp<0x140>(v));
    return v;
}

__device__ __forceinline__ int lane_value(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ int wave_sum(int v)
{
    v = row_sum(v);
    return (int)((uint32_t)lane_value(v, 0) + (uint32_t)lane_value(v, 16) +
                 (uint32_t)lane_value(v, 32) + (uint32_t)lane_value(v, 48));
}

__device__ __forceinline__ int abs16(int v)
{
    // evx_math.h:197-203 on an int16 value.
    int16_t s = (int16_t)v;
    return s == INT16_MIN ? 0x7FFF : (s < 0 ? -s : s);
}

// evx_math.h:228-236
__device__ __forceinline__ int rounded_div(int numer, int denom)
{
    if (((uint32_t)numer ^ (uint32_t)denom) & 0x80000000u) return (numer - denom / 2) / denom;
    return (numer + denom / 2) / denom;
}

// interp.cpp:18-41: six taps starting at p with spacing `step`.
__device__ __forceinline__ int tap6(const int16_t *p, int step)
{
    int s = p[0] - 5 * p[step] + 20 * p[2 * step] + 20 * p[3 * step] - 5 * p[4 * step] +
            p[5 * step];
    return (int16_t)((s + 16) >> 5);
}

// Half-pel sample of interp.cpp:52-164 at integer position (X, Y) of a plane
// (apron-backed, so no clamping is needed).
__device__ __forceinline__ int half_sample(const int16_t *plane, int pitch, int X, int Y, int dx,
                                           int dy)
{
    const int xo = dx > 0 ? 0 : -1;
    const int yo = dy > 0 ? 0 : -1;
    if (dy == 0) return tap6(plane + Y * pitch + X + xo - 2, 1);
    if (dx == 0) return tap6(plane + (Y + yo - 2) * pitch + X, pitch);
    const int16_t *p = plane + (Y + yo - 2) * pitch + X + xo - 2;
    int h0 = tap6(p, 1), h1 = tap6(p + pitch, 1), h2 = tap6(p + 2 * pitch, 1);
    int h3 = tap6(p + 3 * pitch, 1), h4 = tap6(p + 4 * pitch, 1), h5 = tap6(p + 5 * pitch, 1);
    int s = h0 - 5 * h1 + 20 * h2 + 20 * h3 - 5 * h4 + h5;
    return (int16_t)((s + 16) >> 5);
}

// Sample i (0..383, MB order) of the prediction at integer base (bx, by),
// optionally sub-pel (interp.cpp:166-202) or a plain copy (macroblock.h:82).
__device__ __forceinline__ int pred_sample(const DevPic &ref, int i, int bx, int by, int sp,
                                           int dx, int dy, int quarter)
{
    const int16_t *plane;
    int pitch, X, Y;
    if (i < 256) {
        plane = ref.y; pitch = ref.py;
        X = bx + (i & 15); Y = by + (i >> 4);
    } else {
        int k = (i - 256) & 63;
        plane = i < 320 ? ref.u : ref.v; pitch = ref.pc;
        X = (bx >> 1) + (k & 7); Y = (by >> 1) + (k >> 3);
    }
    int full = plane[Y * pitch + X];
    if (!sp) return full;
    int half = half_sample(plane, pitch, X, Y, dx, dy);
    return quarter ? (int16_t)((full + half + 1) >> 1) : half;
}

// motion.cpp:93-116
__device__ __forceinline__ void frac_direction(int idx, int &dx, int &dy)
{
    if (idx <= 2) { dy = -1; dx = idx - 1; }
    else if (idx == 3) { dx = -1; dy = 0; }
    else if (idx == 4) { dx = 1; dy = 0; }
    else { dy = 1; dx = idx - 6; }
}

// Direction k (0..7) of the sub-pel scan order (j outer, i inner, centre
// skipped; motion.cpp:464-473).  k equals the frac index of motion.cpp:68-91.
__device__ __forceinline__ void scan_direction(int k, int &i, int &j)
{
    int m = k < 4 ? k : k + 1;
    i = m % 3 - 1;
    j = m / 3 - 1;
}

// The k-th set bit (k = 0..3) of an 8-bit mask, or -1.
__device__ __forceinline__ int nth_bit(uint32_t mask, int k)
{
    for (int t = 0; t < k; t++) mask &= mask - 1;
    return mask ? __builtin_ctz(mask) : -1;
}

// Source MB samples held in registers.  Lane b = lane & 15 owns luma 4x4
// block (b & 3, b >> 2) and chroma row (b & 7), columns (b >> 3) * 4 .. +3.
struct Src {
    int16_t y[16];
    int16_t u[4];
    int16_t v[4];
    int ox, oy, cr, cc;
};

__device__ __forceinline__ void load_src(const DevPic &pic, int x, int y, Src &s)
{
    const int b = lane_id() & 15;
    s.ox = (b & 3) * 4; s.oy = (b >> 2) * 4;
    s.cr = b & 7; s.cc = (b >> 3) * 4;
    const int16_t *p = pic.y + (y + s.oy) * pic.py + x + s.ox;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) s.y[r * 4 + c] = p[r * pic.py + c];
    const int co = ((y >> 1) + s.cr) * pic.pc + (x >> 1) + s.cc;
#pragma unroll
    for (int k = 0; k < 4; k++) { s.u[k] = pic.u[co + k]; s.v[k] = pic.v[co + k]; }
}

// Per-lane part of an integer candidate (analysis.h:89-132, 180-202).
__device__ __forceinline__ void eval_int_lane(const DevPic &ref, const Src &s, int cx, int cy,
                                              int &sad, int &mad, int &satd)
{
    const int16_t *p = ref.y + (cy + s.oy) * ref.py + cx + s.ox;
    int d[16];
    sad = 0; mad = 0;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int dd = s.y[r * 4 + c] - p[r * ref.py + c];
            int a = dd < 0 ? -dd : dd;
            sad += a;
            mad = max(mad, a);
            d[r * 4 + c] = (int16_t)dd;   // analysis.h:102 stores the difference as int16
        }
    // analysis.h:44-85
    int t[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int a = d[r * 4], b = d[r * 4 + 1], c = d[r * 4 + 2], e = d[r * 4 + 3];
        int s0 = a + b, s1 = a - b, s2 = c + e, s3 = c - e;
        t[r * 4] = s0 + s2; t[r * 4 + 1] = s1 + s3; t[r * 4 + 2] = s0 - s2; t[r * 4 + 3] = s1 - s3;
    }
    satd = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int a = t[c], b = t[4 + c], cc = t[8 + c], e = t[12 + c];
        int s0 = a + b, s1 = a - b, s2 = cc + e, s3 = cc - e;
        satd += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
    const int co = ((cy >> 1) + s.cr) * ref.pc + (cx >> 1) + s.cc;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        mad = max(mad, abs(s.u[k] - ref.u[co + k]));
        mad = max(mad, abs(s.v[k] - ref.v[co + k]));
    }
}

// Per-lane part of one sub-pel direction: half and quarter SAD (luma) and
// MAD (luma + chroma), motion.cpp:180-261 with interp.cpp:52-164.
__device__ __forceinline__ void eval_subpel_lane(const DevPic &ref, const Src &s, int bx, int by,
                                                 int dx, int dy, int &sad_h, int &mad_h,
                                                 int &sad_q, int &mad_q)
{
    sad_h = 0; mad_h = 0; sad_q = 0; mad_q = 0;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int X = bx + s.ox + c, Y = by + s.oy + r;
            const int full = ref.y[Y * ref.py + X];
            const int half = half_sample(ref.y, ref.py, X, Y, dx, dy);
            const int quarter = (int16_t)((full + half + 1) >> 1);
            const int sv = s.y[r * 4 + c];
            const int ah = abs(sv - half), aq = abs(sv - quarter);
            sad_h += ah; mad_h = max(mad_h, ah);
            sad_q += aq; mad_q = max(mad_q, aq);
        }
    const int cbx = bx >> 1, cby = by >> 1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int X = cbx + s.cc + k, Y = cby + s.cr;
        int full = ref.u[Y * ref.pc + X];
        int half = half_sample(ref.u, ref.pc, X, Y, dx, dy);
        int quarter = (int16_t)((full + half + 1) >> 1);
        mad_h = max(mad_h, abs(s.u[k] - half));
        mad_q = max(mad_q, abs(s.u[k] - quarter));
        full = ref.v[Y * ref.pc + X];
        half = half_sample(ref.v, ref.pc, X, Y, dx, dy);
        quarter = (int16_t)((full + half + 1) >> 1);
        mad_h = max(mad_h, abs(s.v[k] - half));
        mad_q = max(mad_q, abs(s.v[k] - quarter));
    }
}

struct SearchOut {
    int best_x, best_y, sad, mad, ssd;
    float cost;
    int sp_en, sp_amt, sp_idx, btype;
    int n_int, n_ax, n_dg;
};

// perform_inter_search (motion.cpp:567-679) for the MB at (x, y) against
// `ref`, predictor (pmx, pmy) as compute_mv_predictor produced it.  All
// lanes of the wave call it with uniform arguments; it returns uniform
// values.
__device__ SearchOut inter_search(const DevPic &ref, int W, int H, const Src &s, int x, int y,
                                  int thresh, float lambda, int pmx, int pmy)
{
    const int g = lane_id() >> 4;
    const int max_x = W - 16, max_y = H - 16;
    SearchOut o;
    o.best_x = x; o.best_y = y;
    o.ssd = kInt32Max;
    o.sp_en = 0; o.sp_amt = 0; o.sp_idx = 0;
    o.n_ax = 0; o.n_dg = 0;

    // Initial candidate (motion.cpp:600-611) in group 0 and the predictor
    // candidate (motion.cpp:621-634, clamped into the picture) in group 1.
    const bool has_pred = pmx != 0 || pmy != 0;
    const int px = min(max(x + pmx, 0), max_x), py = min(max(y + pmy, 0), max_y);
    int sad, mad, satd;
    {
        const int cx = (g == 1 && has_pred) ? px : x;
        const int cy = (g == 1 && has_pred) ? py : y;
        eval_int_lane(ref, s, cx, cy, sad, mad, satd);
        sad = row_sum(sad); mad = row_max(mad); satd = row_sum(satd);
    }
    o.sad = lane_value(sad, 0);
    o.mad = lane_value(mad, 0);
    o.cost = (float)o.sad;
    if (lambda > 0.0f) {
        const int st = (lane_value(satd, 0) + 1) >> 1;
        o.cost = __fadd_rn(__int2float_rn(st), __fmul_rn(lambda, 1.0f));
    }
    o.n_int = 1;

    // evaluate_motion_candidate's acceptance (motion.cpp:125-178) for a
    // candidate whose RD cost is already known.
    auto fold_cost = [&](int cx, int cy, int csad, int cmad, float cost) {
        const int ssd = (cx - x) * (cx - x) + (cy - y) * (cy - y);
        bool accept;
        if (o.mad < thresh)
            accept = cmad < o.mad || (cmad == o.mad && ssd < o.ssd);
        else
            accept = (cost < o.cost || ((cost == o.cost && ssd < o.ssd) && csad < kSadThreshold)) ||
                     cmad < thresh;
        if (accept) {
            o.best_x = cx; o.best_y = cy;
            o.sad = csad; o.ssd = ssd; o.mad = cmad; o.cost = cost;
        }
    };
    auto fold = [&](int cx, int cy, int csad, int cmad, int csatd_raw) {
        float cost = (float)csad;
        if (lambda > 0.0f) {
            const int bits = abs16(cx - x) + abs16(cy - y) + 1;   // motion.cpp:120-123
            cost = __fadd_rn(__int2float_rn((csatd_raw + 1) >> 1),
                             __fmul_rn(lambda, __int2float_rn(bits)));
        }
        fold_cost(cx, cy, csad, cmad, cost);
    };

    if (has_pred) {
        fold(px, py, lane_value(sad, 16), lane_value(mad, 16), lane_value(satd, 16));
        o.n_int++;
    }

    if (o.mad >= thresh) {
        int radius = kSearchRadius;
        if (o.sad < kSadThreshold / 4) radius = kSearchRadius / 4;
        else if (o.sad < kSadThreshold / 2) radius = kSearchRadius / 2;

        bool copy = false;
        for (int step = radius; step > 0 && !copy; step >>= 1) {
            // perform_inter_motion_search (motion.cpp:381-406).  The centre of
            // each 3x3 level is the level's starting best, so its metrics are
            // known without evaluation; it is still folded at its place in
            // the scan (fifth), because the initial candidate carries
            // best_ssd = INT32_MAX and a tie on cost then moves best_ssd to
            // the centre's true distance.
            const int base_x = o.best_x, base_y = o.best_y;
            const int c_sad = o.sad, c_mad = o.mad;
            const float c_cost = o.cost;
            bool centre_done = false;
            uint32_t mask = 0;
            for (int k = 0; k < 8; k++) {
                int i, j;
                scan_direction(k, i, j);
                const int cx = base_x + i * step, cy = base_y + j * step;
                if (cx >= 0 && cx <= max_x && cy >= 0 && cy <= max_y) mask |= 1u << k;
            }
            while (mask && !copy) {
                const int k = nth_bit(mask, g);
                int i, j;
                scan_direction(k < 0 ? 0 : k, i, j);
                const int cx = k < 0 ? base_x : base_x + i * step;
                const int cy = k < 0 ? base_y : base_y + j * step;
                eval_int_lane(ref, s, cx, cy, sad, mad, satd);
                sad = row_sum(sad); mad = row_max(mad); satd = row_sum(satd);
                for (int gg = 0; gg < 4 && mask; gg++) {
                    const int kk = __builtin_ctz(mask);
                    mask &= mask - 1;
                    if (kk >= 4 && !centre_done) {
                        fold_cost(base_x, base_y, c_sad, c_mad, c_cost);
                        centre_done = true;
                    }
                    int ii, jj;
                    scan_direction(kk, ii, jj);
                    fold(base_x + ii * step, base_y + jj * step, lane_value(sad, gg * 16),
                         lane_value(mad, gg * 16), lane_value(satd, gg * 16));
                    o.n_int++;
                    if (o.mad < thresh) { copy = true; break; }   // motion.cpp:402-403
                }
            }
            if (!copy && !centre_done) fold_cost(base_x, base_y, c_sad, c_mad, c_cost);
        }

        // perform_inter_subpixel_motion_search (motion.cpp:450-483).
        uint32_t mask = 0;
        for (int k = 0; k < 8; k++) {
            int i, j;
            scan_direction(k, i, j);
            const int tx = o.best_x + i, ty = o.best_y + j;
            if (tx >= 0 && tx <= max_x && ty >= 0 && ty <= max_y) mask |= 1u << k;
        }
        while (mask) {
            const int k = nth_bit(mask, g);
            int i, j;
            scan_direction(k < 0 ? 0 : k, i, j);
            int sh, mh, sq, mq;
            eval_subpel_lane(ref, s, o.best_x, o.best_y, i, j, sh, mh, sq, mq);
            sh = row_sum(sh); mh = row_max(mh); sq = row_sum(sq); mq = row_max(mq);
            for (int gg = 0; gg < 4 && mask; gg++) {
                const int kk = __builtin_ctz(mask);
                mask &= mask - 1;
                int ii, jj;
                scan_direction(kk, ii, jj);
                if (ii != 0 && jj != 0) o.n_dg++; else o.n_ax++;
                const int csad[2] = {lane_value(sh, gg * 16), lane_value(sq, gg * 16)};
                const int cmad[2] = {lane_value(mh, gg * 16), lane_value(mq, gg * 16)};
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    bool accept;
                    if (o.mad < thresh)
                        accept = cmad[q] < o.mad;
                    else
                        accept = (csad[q] < o.sad && csad[q] < kSadThreshold) || cmad[q] < thresh;
                    if (accept) {
                        o.sp_en = 1; o.sp_amt = q; o.sp_idx = kk;
                        o.sad = csad[q]; o.mad = cmad[q];
                    }
                }
            }
        }
    }

    o.btype = 0;
    if (o.best_x != x || o.best_y != y || o.sp_en) o.btype |= EVXA_BLOCK_MOTION;
    if (o.mad < thresh) o.btype |= EVXA_BLOCK_COPY;
    return o;
}

// Prediction of a search result into LDS (encode.cpp:696-723).
__device__ __forceinline__ void predict_mb(const DevPic &ref, int x, int y, const SearchOut &r,
                                           int16_t *out)
{
    int dx = 0, dy = 0;
    if (r.sp_en) frac_direction(r.sp_idx, dx, dy);
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64)
        out[i] = (int16_t)pred_sample(ref, i, r.best_x, r.best_y, r.sp_en, dx, dy, r.sp_amt);
}

// transform.cpp:264-284 on a line of 8 samples.
__device__ __forceinline__ void dct_line(const int16_t *src, int sp, int16_t *dst, int dp)
{
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = src[k * sp];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int t = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) t += v[k] * c_dct8[i * 8 + k];
        t = (i == 0) ? (t * 45) / 128 : t / 2;
        dst[i * dp] = (int16_t)rounded_div(t, 128);
    }
}

// Origin and pitch of 8x8 block `blk` (0..3 luma quadrants, 4 U, 5 V) in
// MB order.
__device__ __forceinline__ void block_origin(int blk, int &origin, int &pitch)
{
    if (blk < 4) { origin = (blk >> 1) * 128 + (blk & 1) * 8; pitch = 16; }
    else { origin = 256 + (blk - 4) * 64; pitch = 8; }
}

// Residual, forward DCT, adaptive QP, variance and inter quantisation of one
// MB held in LDS (encode.cpp:717-728; transform.cpp:435-452, 584-594;
// quantize.cpp:97-118, 217-252, 372-397; analysis.h:253-275).
// res/tmp are LDS scratch; coefficients end in `res`.  Returns qp and
// variance2 (uniform); levels are written to `lv` (global).
__device__ void dct_quant(const int16_t *src, const int16_t *pred, int16_t *res, int16_t *tmp,
                          int quality, int16_t *co_out, int16_t *lv_out, int &qp_out,
                          int &var_out)
{
    const int lane = lane_id();
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) res[i] = (int16_t)(src[i] - pred[i]);
    __syncthreads();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int row = origin + (lane & 7) * pitch;
        dct_line(res + row, 1, tmp + row, 1);
    }
    __syncthreads();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int col = origin + (lane & 7);
        dct_line(tmp + col, pitch, res + col, pitch);
    }
    __syncthreads();

    // compute_block_variance2 over the 16x16 luma coefficients, (0,0) skipped;
    // the int32 sums wrap like the reference's.
    int sum = 0, cnt = 0;
    uint32_t sq = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = lane * 4 + k;
        const int t = res[i];
        if (i != 0 && t != 0) { sum += t; sq += (uint32_t)(t * t); cnt++; }
    }
    sum = wave_sum(sum);
    cnt = wave_sum(cnt);
    sq = (uint32_t)wave_sum((int)sq);
    int var2 = 0;
    if (cnt > 0) {
        // rounded_div(sum * sum, count) as the reference's optimised build
        // evaluates it: a square is taken as non-negative, so the wrapped
        // square always gets + count / 2 before the truncating division.
        const int sq2 = (int)((uint32_t)sum * (uint32_t)sum);
        const int rd = (int)((uint32_t)sq2 + (uint32_t)(cnt / 2)) / cnt;
        var2 = (int)(sq - (uint32_t)rd);
    }
    // query_block_quantization_parameter
    const uint32_t uv = (uint32_t)var2;
    const int lg = uv ? 31 - __builtin_clz(uv) : 0;
    const int index = min(max(lg >> 1, 1), 63);
    const int q8 = quality & 0xFF;
    int delta = ((int)(int16_t)(index - q8)) >> 1;
    delta = min(max(delta, -6), 6);
    const int qp = min(max(q8 + delta, 1), 63);
    const int cqp = min(qp + 2, 63);

    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        int pos, bq;
        if (i < 256) { pos = ((i >> 4) & 7) * 8 + (i & 7); bq = qp; }
        else { pos = (i - 256) & 63; bq = cqp; }
        const int sv = res[i];
        const int qf = (int16_t)rounded_div(sv * 16, c_inter_qm[pos]);
        int w = c_deadzone[pos];
        if (bq >= 16) w = min(w + 1, 6);
        const int sg = (qf > 0) - (qf < 0);
        const int dz = (int16_t)((sg * bq * w) >> 1);
        lv_out[i] = (int16_t)rounded_div(qf - dz, bq << 1);
        co_out[i] = (int16_t)sv;
    }
    qp_out = qp;
    var_out = var2;
}

__device__ __forceinline__ int16_t load_src_sample(const DevPic &pic, int x, int y, int i)
{
    if (i < 256) return pic.y[(y + (i >> 4)) * pic.py + x + (i & 15)];
    const int k = (i - 256) & 63;
    const int16_t *pl = i < 320 ? pic.u : pic.v;
    return pl[((y >> 1) + (k >> 3)) * pic.pc + (x >> 1) + (k & 7)];
}

__device__ __forceinline__ uint32_t ld_acquire(const int32_t *p)
{
    return (uint32_t)__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// compute_mv_predictor (motion.cpp:271-346) from published records.
__device__ void mv_predictor(const DevJob &job, int bx, int by, int &pmx, int &pmy)
{
    const int idx = by * job.wb + bx;
    int mx[3], my[3], n = 0;
    auto take = [&](const evxa_block_record *r, bool cond) -> bool {
        if (!cond) return false;
        const uint32_t *w = reinterpret_cast<const uint32_t *>(r);
        const uint32_t w0 = ld_relaxed(w);
        if (!(w0 & EVXA_BLOCK_MOTION)) return false;
        const uint32_t w1 = ld_relaxed(w + 1);
        mx[n] = (int16_t)(w1 & 0xFFFF);
        my[n] = (int16_t)(w1 >> 16);
        n++;
        return true;
    };
    take(job.records + idx - 1, bx > 0);
    take(job.records + idx - job.wb, by > 0);
    if (!take(job.records + idx - job.wb + 1, by > 0 && bx + 1 < job.wb))
        take(job.records + idx - job.wb - 1, by > 0 && bx > 0);
    if (n < 3 && job.prev) take(job.prev + idx, true);
    if (n == 0) { pmx = 0; pmy = 0; }
    else if (n == 1) { pmx = mx[0]; pmy = my[0]; }
    else if (n == 2) { pmx = (int16_t)((mx[0] + mx[1]) / 2); pmy = (int16_t)((my[0] + my[1]) / 2); }
    else {
        auto med = [](int a, int b, int c) {   // motion.cpp:263-269
            if (a > b) { int t = a; a = b; b = t; }
            if (b > c) b = c;
            if (a > b) b = a;
            return b;
        };
        pmx = med(mx[0], mx[1], mx[2]);
        pmy = med(my[0], my[1], my[2]);
    }
}

__device__ __forceinline__ void write_record(evxa_block_record *dst, const uint32_t (&w)[16])
{
    const int lane = lane_id();
    if (lane < 16) {
        uint32_t v = w[0];
#pragma unroll
        for (int k = 1; k < 16; k++) v = lane == k ? w[k] : v;
        reinterpret_cast<uint32_t *>(dst)[lane] = v;
    }
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

struct MbScratch {
    int16_t src[EVXA_MB_SAMPLES];
    int16_t pred[EVXA_MB_SAMPLES];
    int16_t pred_b[EVXA_MB_SAMPLES];
    int16_t res[EVXA_MB_SAMPLES];
    int16_t tmp[EVXA_MB_SAMPLES];
};

// Analyse one MB of a job: searches, selection, record publication, then
// prediction, DCT and quantisation.
__device__ void analyze_mb(const DevJob &job, int bx, int by, int32_t *counters, MbScratch &sh)
{
    const int lane = lane_id();
    const int x = bx * 16, y = by * 16;
    const int idx = by * job.wb + bx;
    const float lambda = job.lambda;

    Src s;
    load_src(job.src, x, y, s);
    int pmx, pmy;
    mv_predictor(job, bx, by, pmx, pmy);

    SearchOut best, other;
    int best_k = -1, best_sad = kInt32Max, bidir = 0;
    int n_int = 0, n_ax = 0, n_dg = 0;
    if (job.type == EVXA_FRAME_P) {
        // encode.cpp:215-247 against an empty (non-copy, INT32_MAX) sentinel.
        bool best_copy = false;
        for (int k = 0; k < job.n_refs; k++) {
            SearchOut r = inter_search(job.refs[k], job.width, job.height, s, x, y, job.thresh,
                                       lambda, pmx, pmy);
            n_int += r.n_int; n_ax += r.n_ax; n_dg += r.n_dg;
            const int bias = __float2int_rz(__fmul_rn(lambda, __int2float_rn((k + 1) * 2)));
            const bool copy = (r.btype & EVXA_BLOCK_COPY) != 0;
            if (copy != best_copy) {
                if (copy) { best = r; best_k = k; best_sad = r.sad; best_copy = true; }
            } else if (r.sad + bias < best_sad) {
                best = r; best_k = k; best_sad = r.sad; best_copy = copy;
            }
        }
    } else {
        // encode.cpp:386-443 without the intra candidates; the bidirectional
        // candidate (motion.cpp:702-840) reuses both searches.
        SearchOut f = inter_search(job.refs[0], job.width, job.height, s, x, y, job.thresh, lambda,
                                   pmx, pmy);
        SearchOut b = inter_search(job.refs[1], job.width, job.height, s, x, y, job.thresh, lambda,
                                   pmx, pmy);
        n_int = f.n_int + b.n_int; n_ax = f.n_ax + b.n_ax; n_dg = f.n_dg + b.n_dg;
        if (f.sad < best_sad) { best = f; best_k = 0; best_sad = f.sad; }
        if (b.sad < best_sad) { best = b; best_k = 1; best_sad = b.sad; }
        predict_mb(job.refs[0], x, y, f, sh.pred);
        predict_mb(job.refs[1], x, y, b, sh.pred_b);
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) sh.src[i] = load_src_sample(job.src, x, y, i);
        __syncthreads();
        int bsad = 0;
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
            const int a = (sh.pred[i] + sh.pred_b[i] + 1) >> 1;
            sh.tmp[i] = (int16_t)a;
            bsad += abs(sh.src[i] - a);
        }
        bsad = wave_sum(bsad);
        const int bias = __float2int_rz(__fmul_rn(lambda, 8.0f));
        if (bsad + bias < best_sad) {
            bidir = 1; best_sad = bsad; best = f; best_k = 0; other = b;
        }
        __syncthreads();
        if (!bidir && best_k == 1) {
            for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) sh.pred[i] = sh.pred_b[i];
        } else if (bidir) {
            for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) sh.pred[i] = sh.tmp[i];
        }
        __syncthreads();
    }

    if (best_k < 0 || best_sad == kInt32Max) {
        if (lane == 0) atomicOr(&counters[kCounterError], kErrSearchFailed);
        best_k = 0;
        best = SearchOut{x, y, best_sad, 0, 0, 0.0f, 0, 0, 0, EVXA_BLOCK_COPY, 0, 0, 0};
    }

    const uint32_t ids = job.target_ids;
    uint32_t w[16];
    const int btype = bidir ? EVXA_BLOCK_MOTION : best.btype;
    const int mode = bidir ? 2 : (job.type == EVXA_FRAME_P ? 0 : best_k);
    const int tgt = (ids >> (8 * best_k)) & 0xFF;
    const int tgt_b = bidir ? ((ids >> 8) & 0xFF) : 0;
    const int mvx = best.best_x - x, mvy = best.best_y - y;
    const int mvx_b = bidir ? other.best_x - x : 0, mvy_b = bidir ? other.best_y - y : 0;
    w[0] = (uint32_t)btype | ((uint32_t)mode << 8) | ((uint32_t)tgt << 16) | ((uint32_t)tgt_b << 24);
    w[1] = (uint32_t)(uint16_t)mvx | ((uint32_t)(uint16_t)mvy << 16);
    w[2] = (uint32_t)(uint16_t)mvx_b | ((uint32_t)(uint16_t)mvy_b << 16);
    w[3] = (uint32_t)best.sp_en | ((uint32_t)best.sp_amt << 8) | ((uint32_t)best.sp_idx << 16) |
           ((uint32_t)(bidir ? other.sp_en : 0) << 24);
    w[4] = bidir ? ((uint32_t)other.sp_amt | ((uint32_t)other.sp_idx << 8)) : 0u;
    w[5] = 0;
    w[6] = (uint32_t)best_sad;
    w[7] = (uint32_t)best.mad;
    w[8] = fbits(best.cost);
    w[9] = bidir ? (uint32_t)other.sad : 0u;
    w[10] = bidir ? (uint32_t)other.mad : 0u;
    w[11] = bidir ? fbits(other.cost) : 0u;
    w[12] = (uint32_t)n_int;
    w[13] = (uint32_t)n_ax | ((uint32_t)n_dg << 16);
    w[14] = 0;
    w[15] = 0;
    evxa_block_record *rec = job.records + idx;
    write_record(rec, w);

    // Publish: the record is visible before the row counter moves past it.
    __threadfence();
    __syncthreads();
    if (lane == 0)
        __hip_atomic_store(job.progress + by, bx + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

    int16_t *co = job.coeffs + (size_t)idx * EVXA_MB_SAMPLES;
    int16_t *lv = job.levels + (size_t)idx * EVXA_MB_SAMPLES;
    if (btype & EVXA_BLOCK_COPY) {
        // encode.cpp:742-744: copy blocks are not transformed.
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) { co[i] = 0; lv[i] = 0; }
        return;
    }
    if (job.type == EVXA_FRAME_P) {
        predict_mb(job.refs[best_k], x, y, best, sh.pred);
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) sh.src[i] = load_src_sample(job.src, x, y, i);
        __syncthreads();
    }
    int qp, var2;
    dct_quant(sh.src, sh.pred, sh.res, sh.tmp, job.quality, co, lv, qp, var2);
    w[4] |= (uint32_t)qp << 16;
    w[5] = (uint32_t)(uint16_t)(int16_t)var2;
    w[14] = 1u | (bidir ? 2u : 0u);
    write_record(rec, w);
    __syncthreads();
}

__global__ __launch_bounds__(64) void analyze_kernel(const DevJob *__restrict__ jobs,
                                                     const uint32_t *__restrict__ tickets,
                                                     int32_t n_tickets, int32_t *counters)
{
    __shared__ MbScratch sh;
    __shared__ int32_t s_ticket;
    const int lane = lane_id();
    for (;;) {
        if (lane == 0) s_ticket = atomicAdd(&counters[kCounterTicket], 1);
        __syncthreads();
        const int t = s_ticket;
        __syncthreads();
        if (t >= n_tickets) break;
        const uint32_t e = tickets[t];
        const DevJob &job = jobs[e >> kJobShift];
        const int mb = (int)(e & kMbMask);
        const int bx = mb % job.wb, by = mb / job.wb;
        // Wait for the left neighbour (same row) and the top-right one (row
        // above; the whole row above at the right edge).
        int spins = 0;
        bool timed_out = false;
        if (bx > 0) {
            while ((int)ld_acquire(job.progress + by) < bx) {
                if (++spins > kSpinLimit) { timed_out = true; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (by > 0) {
            const int need = min(bx + 2, job.wb);
            while ((int)ld_acquire(job.progress + by - 1) < need) {
                if (++spins > kSpinLimit) { timed_out = true; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (timed_out && lane == 0) atomicOr(&counters[kCounterError], kErrWaitTimeout);
        analyze_mb(job, bx, by, counters, sh);
    }
}

__global__ void reset_kernel(int32_t *words, int32_t n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        words[i] = 0;
}

// Replicate the plane's edge samples into its apron (equivalent to the
// clamped sampling of interp.cpp:8-14).
__global__ void apron_kernel(int16_t *plane0, int32_t pitch, int32_t w, int32_t h)
{
    const int W = w + 2 * kApron, H = h + 2 * kApron;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < W * H; i += gridDim.x * blockDim.x) {
        const int X = i % W - kApron, Y = i / W - kApron;
        if (X >= 0 && X < w && Y >= 0 && Y < h) continue;
        const int cx = min(max(X, 0), w - 1), cy = min(max(Y, 0), h - 1);
        plane0[Y * pitch + X] = plane0[cy * pitch + cx];
    }
}

// params: n x {x, y, dir_x, dir_y, quarter}
__global__ __launch_bounds__(64) void interp_blocks_kernel(DevPic pic, int32_t n,
                                                           const int32_t *params, int16_t *out)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const int32_t *p = params + b * 5;
    const int dx = p[2], dy = p[3];
    const int sp = (dx != 0 || dy != 0);
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64)
        out[(size_t)b * EVXA_MB_SAMPLES + i] =
            (int16_t)pred_sample(pic, i, p[0], p[1], sp, dx, dy, p[4]);
}

// params: n x {x, y, pred_x, pred_y}
__global__ __launch_bounds__(64) void search_blocks_kernel(DevPic src, DevPic ref, int32_t w,
                                                           int32_t h, int32_t quality,
                                                           float lambda, int32_t n,
                                                           const int32_t *params,
                                                           evxa_search_result *out)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const int32_t *p = params + b * 4;
    Src s;
    load_src(src, p[0], p[1], s);
    SearchOut o = inter_search(ref, w, h, s, p[0], p[1], (quality >> 2) + 1, lambda, p[2], p[3]);
    if (lane_id() == 0) {
        evxa_search_result r;
        r.best_x = o.best_x; r.best_y = o.best_y;
        r.sad = o.sad; r.mad = o.mad; r.ssd = o.ssd;
        r.cost_bits = fbits(o.cost);
        r.sp_enabled = o.sp_en; r.sp_amount = o.sp_amt; r.sp_index = o.sp_idx;
        r.block_type = o.btype;
        r.motion_x = o.best_x - p[0]; r.motion_y = o.best_y - p[1];
        r.n_int = o.n_int; r.n_axial = o.n_ax; r.n_diag = o.n_dg;
        out[b] = r;
    }
}

__global__ __launch_bounds__(64) void dct_quant_blocks_kernel(int32_t n, const int16_t *src,
                                                              const int16_t *pred,
                                                              const int32_t *quality,
                                                              int16_t *coeffs, int16_t *levels,
                                                              int32_t *qp, int32_t *variance)
{
    __shared__ MbScratch sh;
    const int b = blockIdx.x;
    if (b >= n) return;
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64) {
        sh.src[i] = src[(size_t)b * EVXA_MB_SAMPLES + i];
        sh.pred[i] = pred[(size_t)b * EVXA_MB_SAMPLES + i];
    }
    __syncthreads();
    int q, v;
    dct_quant(sh.src, sh.pred, sh.res, sh.tmp, quality[b], coeffs + (size_t)b * EVXA_MB_SAMPLES,
              levels + (size_t)b * EVXA_MB_SAMPLES, q, v);
    if (lane_id() == 0) { qp[b] = q; variance[b] = v; }
}

}  // namespace

hipError_t launch_reset(int32_t *words, int32_t n, hipStream_t st)
{
    const int blocks = min((n + 255) / 256, 1024);
    hipLaunchKernelGGL(reset_kernel, dim3(max(blocks, 1)), dim3(256), 0, st, words, n);
    return hipGetLastError();
}

hipError_t launch_analyze(const DevJob *jobs, const uint32_t *tickets, int32_t n_tickets,
                          int32_t *counters, int32_t grid, hipStream_t st)
{
    hipLaunchKernelGGL(analyze_kernel, dim3(grid), dim3(64), 0, st, jobs, tickets, n_tickets,
                       counters);
    return hipGetLastError();
}

hipError_t launch_apron(int16_t *plane0, int32_t pitch, int32_t w, int32_t h, hipStream_t st)
{
    const int total = (w + 2 * kApron) * (h + 2 * kApron);
    const int blocks = min((total + 255) / 256, 2048);
    hipLaunchKernelGGL(apron_kernel, dim3(blocks), dim3(256), 0, st, plane0, pitch, w, h);
    return hipGetLastError();
}

hipError_t launch_interp_blocks(DevPic pic, int32_t w, int32_t h, int32_t n, const int32_t *params,
                                int16_t *out, hipStream_t st)
{
    (void)w; (void)h;
    hipLaunchKernelGGL(interp_blocks_kernel, dim3(n), dim3(64), 0, st, pic, n, params, out);
    return hipGetLastError();
}

hipError_t launch_search_blocks(DevPic src, DevPic ref, int32_t w, int32_t h, int32_t quality,
                                float lambda, int32_t n, const int32_t *params,
                                evxa_search_result *out, hipStream_t st)
{
    hipLaunchKernelGGL(search_blocks_kernel, dim3(n), dim3(64), 0, st, src, ref, w, h, quality,
                       lambda, n, params, out);
    return hipGetLastError();
}

hipError_t launch_dct_quant_blocks(int32_t n, const int16_t *src, const int16_t *pred,
                                   const int32_t *quality, int16_t *coeffs, int16_t *levels,
                                   int32_t *qp, int32_t *variance, hipStream_t st)
{
    hipLaunchKernelGGL(dct_quant_blocks_kernel, dim3(n), dim3(64), 0, st, n, src, pred, quality,
                       coeffs, levels, qp, variance);
    return hipGetLastError();
}

int32_t analyze_resident_waves(int32_t device)
{
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        n_cu <= 0)
        n_cu = 256;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, analyze_kernel, 64, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    return n_cu * per_cu;
}

}  // namespace evxa
