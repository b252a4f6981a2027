// evxa_kernels.hip -- gfx950 kernels of the EVX3 inter analysis path.
//
// Work items are (macroblock, reference) searches, one wavefront (64 lanes,
// one workgroup) each.  Candidate metrics run wide: the wave splits into four
// 16-lane groups, each evaluating one candidate position with one lane per
// 4x4 luma block (plus 8 chroma samples), and DPP row reductions give each
// group's SAD, SATD and MAD.  The decision is the reference's ordered fold
// over the candidates (motion.cpp:125-178, 180-261) on wave-uniform values.
// Sub-pel refinement builds the horizontal, vertical and diagonal half-pel
// planes around the integer best once, in LDS, and reads all 16 candidates
// from them.
//
// A persistent work list hands out items in anti-diagonal order
// (t = bx + 2*by, all jobs of a launch interleaved, the references of a MB
// adjacent).  An item waits until the MB's left and top-right neighbours are
// published, because the MV predictor reads their final records
// (motion.cpp:271-346).  The wave that finishes a MB's last search selects
// among the references, publishes the record, then forms the residual
// prediction and runs DCT, adaptive QP and quantisation, which are off the
// dependency chain.
//
// Arithmetic is integer except the RD cost: an fp32 multiply followed by an
// fp32 add with round-to-nearest and no contraction (motion.cpp:144), as on
// the reference's x86 build.

#include "evxa_internal.h"

namespace evxa {
namespace {

constexpr int kSadThreshold = 8 * 1024;  // motion.cpp:24
constexpr int kSearchRadius = 16;        // motion.cpp:29
constexpr int kInt32Max = 0x7FFFFFFF;
constexpr int kSpinLimit = 1 << 19;
// Longest back-off between two polls of a dependency, in s_sleep(1) units
// (about 64 cycles each): a MB's wait is on the dependency chain, so the
// back-off stays short; every poll is one memory instruction of four lanes.
constexpr int kMaxNap = 8;

// xftables.h:57-67: round(128 cos((2k+1) i pi / 16)), columns 0..3 of each
// row; row i is symmetric (even i) or antisymmetric (odd i) about its middle,
// c[i][7 - k] = (-1)^i c[i][k].  Compile-time constants, so they fold into
// the multiplies.
constexpr int kDct8Half[8][4] = {
    {128, 128, 128, 128},
    {126, 106, 71, 25},
    {118, 49, -49, -118},
    {106, -25, -126, -71},
    {91, -91, -91, 91},
    {71, -126, 25, 106},
    {49, -118, 118, -49},
    {25, -71, 106, -126}};

// Quantiser tables: the inter matrix (quantize.cpp:26-36), the dead-zone
// widths (quantize.cpp:61-71) and, per matrix entry, the reciprocal
// multiplier ceil(2^32 / qm) that rdiv_magic divides by.
struct QuantTables {
    int16_t qm[64];
    uint8_t dz[64];
    uint32_t qm_magic[64];
};

constexpr QuantTables make_quant_tables()
{
    QuantTables t{
        {16, 17, 18, 19, 20, 21, 22, 23, 17, 18, 19, 20, 21, 22, 23, 24,
         18, 19, 20, 21, 22, 23, 24, 25, 19, 20, 21, 22, 23, 24, 26, 27,
         20, 21, 22, 23, 25, 26, 27, 28, 21, 22, 23, 24, 26, 27, 28, 30,
         22, 23, 24, 26, 27, 28, 30, 31, 23, 24, 25, 27, 28, 30, 31, 33},
        {2, 2, 2, 3, 3, 3, 4, 4, 2, 2, 3, 3, 3, 4, 4, 4, 2, 3, 3, 3, 4, 4, 4, 4,
         3, 3, 3, 4, 4, 4, 4, 4, 3, 3, 4, 4, 4, 4, 4, 4, 3, 4, 4, 4, 4, 4, 4, 4,
         4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4},
        {}};
    for (int i = 0; i < 64; i++) t.qm_magic[i] = 0xFFFFFFFFu / (uint32_t)t.qm[i] + 1u;
    return t;
}

__constant__ QuantTables c_quant = make_quant_tables();

// The lane index, opaque to the optimiser: lane-dependent constants are then
// recomputed where they are used instead of being hoisted out of the
// persistent kernels' loops, where they would stay live across a whole
// ticket (in the one-wave analysis kernel this took the VGPR count from 256,
// with spills to scratch, to 136, and the waves per SIMD from 2 to 3).
__device__ __forceinline__ int lane_id()
{
    int l = (int)(threadIdx.x & 63);
    asm volatile("" : "+v"(l));
    return l;
}
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Orders one wave's LDS accesses across its lanes: a wave's LDS
// instructions execute in order, so completing them suffices.  Used by the
// per-MB stages that a single wave runs (selection, prediction, DCT).
__device__ __forceinline__ void wave_sync()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);     // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

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

__device__ __forceinline__ int wave_min(int v)
{
    v = min(v, dpp<0xB1>(v));
    v = min(v, dpp<0x4E>(v));
    v = min(v, dpp<0x141>(v));
    v = min(v, dpp<0x140>(v));
    return min(min(lane_value(v, 0), lane_value(v, 16)), min(lane_value(v, 32), lane_value(v, 48)));
}

__device__ __forceinline__ int wave_max(int v)
{
    v = row_max(v);
    return max(max(lane_value(v, 0), lane_value(v, 16)), max(lane_value(v, 32), lane_value(v, 48)));
}

// A picture's sample range is gathered in a spread record: kRangeWays pairs
// of words, kRangeStride words apart (one cache line each), cleared to zero;
// pair k holds 32767 - min and max + 32768 over the samples of the waves
// whose workgroup index is k modulo kRangeWays, grown with atomicMax.  A
// wave first reads its pair and updates a word only when it would grow it
// (the words only grow, so a stale read costs an extra atomic but loses no
// update).  fold_range_kernel then folds the pairs into the picture's two
// words the analysis reads.  Called by every lane of a wave; lo / hi are
// the lane's extremes (INT_MAX / INT_MIN when it stored nothing).
__device__ __forceinline__ void note_range(int32_t *spread, int lo, int hi)
{
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane_id() == 0 && hi >= lo) {
        int32_t *pair = spread + ((blockIdx.x + blockIdx.y * gridDim.x) % kRangeWays) * kRangeStride;
        const int e0 = 32767 - lo, e1 = hi + 32768;
        if (e0 > __hip_atomic_load(&pair[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&pair[0], e0);
        if (e1 > __hip_atomic_load(&pair[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&pair[1], e1);
    }
}

// Whether a source / reference pair takes the packed int16 metric paths:
// their samples together span at most 2048 values within [-8192, 8191], so
// every difference is at most 2047 in magnitude and every half-pel sample
// and sum those paths form fits in int16.  A picture's range word holds its
// maximum in the high half and its minimum in the low half (zero for a
// picture never written).
__device__ __forceinline__ bool narrow_pair(const int32_t *a, const int32_t *b)
{
    const int32_t wa = *a, wb = *b;
    const int lo = min((int)(int16_t)(wa & 0xFFFF), (int)(int16_t)(wb & 0xFFFF));
    const int hi = max(wa >> 16, wb >> 16);
    return hi - lo <= 2047 && lo >= -8192 && hi <= 8191;
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

// evx_math.h:228-236 with a positive divisor d and |numer| + d/2 < 2^20:
// the truncating quotient of |numer| + d/2 is the high half of its product
// with m = ceil(2^32 / d) (the error term stays below one unit because
// (|numer| + d/2) * (m d - 2^32) < 2^32).
__device__ __forceinline__ int rdiv_magic(int numer, int d, uint32_t m)
{
    const uint32_t a = (uint32_t)(numer < 0 ? -numer : numer) + (uint32_t)(d >> 1);
    const int q = (int)__umulhi(a, m);
    return numer < 0 ? -q : q;
}

__device__ __forceinline__ uint32_t magic_of(int d) { return 0xFFFFFFFFu / (uint32_t)d + 1u; }

// interp.cpp:18-41: six taps starting at p with spacing `step`.
__device__ __forceinline__ int tap6(const int16_t *p, int step)
{
    int s = p[0] - 5 * p[step] + 20 * p[2 * step] + 20 * p[3 * step] - 5 * p[4 * step] +
            p[5 * step];
    return (int16_t)((s + 16) >> 5);
}

// motion.cpp:93-116
__device__ __forceinline__ void frac_direction(int idx, int &dx, int &dy)
{
    if (idx <= 2) { dy = -1; dx = idx - 1; }
    else if (idx == 3) { dx = -1; dy = 0; }
    else if (idx == 4) { dx = 1; dy = 0; }
    else { dy = 1; dx = idx - 6; }
}

// Packed 16-bit lanes: two int16 samples per 32-bit register, low half first.
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 pk(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ s16x2 pabs(s16x2 x) { return __builtin_elementwise_max(x, -x); }
__device__ __forceinline__ s16x2 pmax(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 pswap(s16x2 x) { return __builtin_shufflevector(x, x, 1, 0); }
// Low half + high half, both taken as unsigned 16-bit.
__device__ __forceinline__ uint32_t hsum(s16x2 x)
{
    return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, x), (u16x2)(1), 0u, false);
}
__device__ __forceinline__ int lo16(uint32_t x) { return (int16_t)(x & 0xFFFFu); }
__device__ __forceinline__ int hi16(uint32_t x) { return (int16_t)(x >> 16); }

// Samples of one MB-sized block in the lane layout: lane b = lane & 15 owns
// luma 4x4 block (b & 3, b >> 2) -- rows oy .. oy + 3, columns ox .. ox + 3,
// two packed pairs per row (y[2r], y[2r + 1]) -- and chroma row cr, columns
// cc .. cc + 3 of U and V (two pairs each).
struct Blk {
    uint32_t y[8];
    uint32_t u[2];
    uint32_t v[2];
};

struct Geo {
    int ox, oy, cr, cc;
};

__device__ __forceinline__ Geo lane_geometry()
{
    const int b = lane_id() & 15;
    Geo g;
    g.ox = (b & 3) * 4; g.oy = (b >> 2) * 4;
    g.cr = b & 7; g.cc = (b >> 3) * 4;
    return g;
}

// Four consecutive samples of a picture in global memory as one 8-byte
// access (the address only needs 2-byte alignment, which gfx950 serves as
// one access).  The pointer is taken as a global one, so the load is a
// global_load: a flat load would also count against the LDS wait counter,
// and every wait for an LDS read would then wait for it too.
struct __attribute__((packed, aligned(2))) Pair32 {
    uint32_t lo, hi;
};
typedef const __attribute__((address_space(1))) Pair32 *GlobalPair32;

__device__ __forceinline__ void load4(const int16_t *p, uint32_t &lo, uint32_t &hi)
{
    const GlobalPair32 q = (GlobalPair32)(p);
    lo = q->lo;
    hi = q->hi;
}

// The block at (x, y) of planes with pitches (py, pc) in the lane layout.
__device__ __forceinline__ void load_blk(const int16_t *py_, const int16_t *pu_, const int16_t *pv_,
                                         int pitch_y, int pitch_c, int x, int y, const Geo &g,
                                         Blk &b)
{
    const int16_t *p = py_ + (y + g.oy) * pitch_y + x + g.ox;
#pragma unroll
    for (int r = 0; r < 4; r++) load4(p + r * pitch_y, b.y[2 * r], b.y[2 * r + 1]);
    const int co = ((y >> 1) + g.cr) * pitch_c + (x >> 1) + g.cc;
    load4(pu_ + co, b.u[0], b.u[1]);
    load4(pv_ + co, b.v[0], b.v[1]);
}

__device__ __forceinline__ void load_cand(const DevPic &pic, const Geo &g, int x, int y, Blk &b)
{
    load_blk(pic.y, pic.u, pic.v, pic.py, pic.pc, x, y, g, b);
}

__device__ __forceinline__ void select_blk(bool take_b, const Blk &a, const Blk &b, Blk &out)
{
#pragma unroll
    for (int k = 0; k < 8; k++) out.y[k] = take_b ? b.y[k] : a.y[k];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        out.u[k] = take_b ? b.u[k] : a.u[k];
        out.v[k] = take_b ? b.v[k] : a.v[k];
    }
}

// Per-lane part of an integer candidate (analysis.h:89-132, 180-202) for a
// source / reference pair that narrow_pair accepts (together at most 2048
// distinct values within [-8192, 8191]): |differences| <= 2047, so the
// differences, their sums over a lane's 16 samples and the 4x4 Hadamard
// (intermediates <= 16 * 2047 = 32752) are exact in packed int16.  The
// Hadamard runs the vertical butterflies first (the set of coefficients does
// not depend on the pass order); the last horizontal stage uses
// |x + y| + |x - y| = 2 max(|x|, |y|).
__device__ __forceinline__ void metrics_narrow(const Blk &s, const Blk &r, int &sad, int &mad,
                                               int &satd)
{
    s16x2 d[8], a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        d[k] = pk(s.y[k]) - pk(r.y[k]);
        a[k] = pabs(d[k]);
    }
    const s16x2 sum = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    s16x2 mx = pmax(pmax(pmax(a[0], a[1]), pmax(a[2], a[3])), pmax(pmax(a[4], a[5]), pmax(a[6], a[7])));
#pragma unroll
    for (int k = 0; k < 2; k++) {
        mx = pmax(mx, pabs(pk(s.u[k]) - pk(r.u[k])));
        mx = pmax(mx, pabs(pk(s.v[k]) - pk(r.v[k])));
    }
    sad = (int)hsum(sum);
    mad = pmax(mx, pswap(mx)).x;
    s16x2 u[4][2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const s16x2 s0 = d[h] + d[2 + h], s1 = d[h] - d[2 + h];
        const s16x2 s2 = d[4 + h] + d[6 + h], s3 = d[4 + h] - d[6 + h];
        u[0][h] = s0 + s2; u[1][h] = s1 + s3; u[2][h] = s0 - s2; u[3][h] = s1 - s3;
    }
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const s16x2 aS = pabs(u[i][0] + u[i][1]), aD = pabs(u[i][0] - u[i][1]);
        acc += (uint32_t)(uint16_t)pmax(aS, pswap(aS)).x + (uint32_t)(uint16_t)pmax(aD, pswap(aD)).x;
    }
    satd = (int)(2 * acc);
}

// The same for general int16 samples, in 32-bit arithmetic as the reference:
// SAD and MAD on int32 differences, the Hadamard on differences stored as
// int16 (analysis.h:102).
__device__ __forceinline__ void metrics_wide(const Blk &s, const Blk &r, int &sad, int &mad,
                                             int &satd)
{
    int d[16];
    sad = 0; mad = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int d0 = lo16(s.y[k]) - lo16(r.y[k]), d1 = hi16(s.y[k]) - hi16(r.y[k]);
        const int a0 = d0 < 0 ? -d0 : d0, a1 = d1 < 0 ? -d1 : d1;
        sad += a0 + a1;
        mad = max(mad, max(a0, a1));
        d[2 * k] = (int16_t)d0;
        d[2 * k + 1] = (int16_t)d1;
    }
    // analysis.h:44-85
    int t[16];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        int a = d[rr * 4], b = d[rr * 4 + 1], c = d[rr * 4 + 2], e = d[rr * 4 + 3];
        int s0 = a + b, s1 = a - b, s2 = c + e, s3 = c - e;
        t[rr * 4] = s0 + s2; t[rr * 4 + 1] = s1 + s3; t[rr * 4 + 2] = s0 - s2; t[rr * 4 + 3] = s1 - s3;
    }
    satd = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int a = t[c], b = t[4 + c], cc = t[8 + c], e = t[12 + c];
        int s0 = a + b, s1 = a - b, s2 = cc + e, s3 = cc - e;
        satd += abs(s0 + s2) + abs(s1 + s3) + abs(s0 - s2) + abs(s1 - s3);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        mad = max(mad, max(abs(lo16(s.u[k]) - lo16(r.u[k])), abs(hi16(s.u[k]) - hi16(r.u[k]))));
        mad = max(mad, max(abs(lo16(s.v[k]) - lo16(r.v[k])), abs(hi16(s.v[k]) - hi16(r.v[k]))));
    }
}

// One integer candidate over the 16 lanes of a group: SAD, MAD and raw
// SATD (before the final (x + 1) >> 1) in every lane of the group.
__device__ __forceinline__ void cand_metrics(bool narrow, const Blk &s, const Blk &r, int &sad,
                                             int &mad, int &satd)
{
    if (narrow) metrics_narrow(s, r, sad, mad, satd);
    else metrics_wide(s, r, sad, mad, satd);
    sad = row_sum(sad);
    mad = row_max(mad);
    satd = row_sum(satd);
}

// Half-pel planes around an integer base (bx, by), interp.cpp:52-164 split
// into its separable passes.  For a block of size B (16 luma, 8 chroma):
//   h: horizontal halves between columns c and c+1, rows -3..B+2, cols -1..B-1
//   v: vertical halves between rows r and r+1, rows -1..B-1, cols 0..B-1
//   d: vertical 6-tap over h (the diagonal), rows -1..B-1, cols -1..B-1
// Every direction's half-pel block is a B x B window of one plane.
template <int B>
struct HalfPlanes {
    static constexpr int HR = B + 6, HC = B + 1;
    static constexpr int VR = B + 1, VC = B;
    static constexpr int DR = B + 1, DC = B + 1;
    // Each plane is padded to whole dwords plus one: blocks are read as
    // aligned dwords, up to one dword past their last sample.
    static constexpr int pad(int n) { return (n + 3) & ~1; }
    alignas(4) int16_t h[pad(HR * HC)];
    alignas(4) int16_t v[pad(VR * VC)];
    alignas(4) int16_t d[pad(DR * DC)];
};

// Reference window staged in LDS.  A window around a base position b with
// reach r covers the reads of any level whose candidates stay within b +- r
// and of the sub-pel refinement around any such position: luma
// [b - r - 3, b + r + 18] and chroma [((b - r) >> 1) - 3, ((b + r) >> 1) + 10]
// on each axis.  The largest window (r = 15) is staged around the MB itself
// while it waits for its neighbours; a level or the sub-pel step whose reads
// fall outside the current window stages a new one around its base.  Column
// origins are even so rows load as aligned dwords.
constexpr int kReachMax = 15;
constexpr int kWinLS = 54;    // luma row stride (>= 2 * 15 + 23, even)
constexpr int kWinLR = 52;    // luma rows (2 * 15 + 22)
constexpr int kWinCS = 34;    // chroma row stride (>= 15 + 16, even; 17 dwords spread rows over the LDS banks)
constexpr int kWinCR = 30;    // chroma rows (15 + 15)

struct Window {
    int16_t y[kWinLR * kWinLS];
    int16_t u[kWinCR * kWinCS];
    int16_t v[kWinCR * kWinCS];
};

// Per-slot results exchanged between the waves of a two-wave search:
// integer levels (double-buffered by level parity) and sub-pel directions.
struct Exchange {
    uint32_t cost[2][8];
    int32_t sad[2][8];
    int32_t mad[2][8];
    int32_t sub[8][4];        // half SAD, half MAD, quarter SAD, quarter MAD
};

struct SubpelLds {
    Window win;
    HalfPlanes<16> y;
    HalfPlanes<8> u;
    HalfPlanes<8> v;
    Exchange xch;
};

// Where the staged window sits in picture coordinates (wave-uniform): luma
// columns [x0, x0 + kWinLS) and rows [y0, y1), chroma columns
// [cx0, cx0 + kWinCS) and rows [cy0, cy1).
struct WinPos {
    int x0, y0, y1, cx0, cy0, cy1;
};

// Does the window hold luma [lx0, lx1) x [ly0, ly1) and chroma
// [cx0, cx1) x [cy0, cy1)?
__device__ __forceinline__ bool win_holds(const WinPos &w, int lx0, int lx1, int ly0, int ly1,
                                          int cx0, int cx1, int cy0, int cy1)
{
    return lx0 >= w.x0 && lx1 <= w.x0 + kWinLS && ly0 >= w.y0 && ly1 <= w.y1 && cx0 >= w.cx0 &&
           cx1 <= w.cx0 + kWinCS && cy0 >= w.cy0 && cy1 <= w.cy1;
}

// Reads of the candidates of a level with step `step` around (bx, by),
// clamped into the picture like the candidates themselves.
__device__ __forceinline__ bool win_holds_level(const WinPos &w, int bx, int by, int step, int max_x,
                                                int max_y)
{
    const int lx0 = max(bx - step, 0), lx1 = min(bx + step, max_x) + 16;
    const int ly0 = max(by - step, 0), ly1 = min(by + step, max_y) + 16;
    return win_holds(w, lx0, lx1, ly0, ly1, lx0 >> 1, ((lx1 - 16) >> 1) + 8, ly0 >> 1,
                     ((ly1 - 16) >> 1) + 8);
}

// Reads of the sub-pel refinement around (bx, by) (interp.cpp:52-164 with
// its 3-sample apron, luma 16 and chroma 8 samples wide).
__device__ __forceinline__ bool win_holds_subpel(const WinPos &w, int bx, int by)
{
    return win_holds(w, bx - 3, bx + 19, by - 3, by + 19, (bx >> 1) - 3, (bx >> 1) + 11,
                     (by >> 1) - 3, (by >> 1) + 11);
}

typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;

// Copies rows of `row_dw` dwords (row r of the source at g + r * pitch_dw)
// into LDS packed with stride row_dw, n_rows rows, straight from memory to
// LDS (global_load_lds_dword: lane i of the k-th load writes dword 64k + i).
// The W waves of the workgroup take every W-th load.  All loads are issued
// back to back; they complete under vmcnt(0).  AUX is the loads' cache
// policy (kLoadDevice: device scope, for pictures other workgroups of the
// same launch write).
constexpr int kLoadDevice = 16;     // sc1
template <int ROW_DW, int W, int AUX = 0>
__device__ __forceinline__ void dma_rows(const uint32_t *g, int pitch_dw, int n_rows, uint32_t *lds)
{
    const int lane = lane_id();
    const int w0 = W > 1 ? wave_id() : 0;
    const int n = ROW_DW * n_rows;
    const int first = 64 * w0 + lane;
    int r = first / ROW_DW, c = first - (first / ROW_DW) * ROW_DW;   // of dword 64k + lane
    constexpr int kStep = 64 * W;
    constexpr int kStepR = kStep / ROW_DW, kStepC = kStep - kStepR * ROW_DW;
    for (int k = w0; k * 64 < n; k += W) {
        if (k * 64 + lane < n)
            __builtin_amdgcn_global_load_lds((global_void *)(g + r * pitch_dw + c),
                                             (lds_void *)(lds + k * 64), 4, 0, AUX);
        r += kStepR;
        c += kStepC;
        if (c >= ROW_DW) { c -= ROW_DW; r++; }
    }
}

// Issue the copy of the window around (bx, by) with reach `reach` into LDS
// (call window_ready before reading it).  Luma rows are kWinLS samples and
// chroma rows kWinCS, always; the column origins are even and clamped so that
// a row never runs past the right apron of its own picture row.
template <int W>
__device__ WinPos stage_window(const DevPic &ref, int bx, int by, int reach, SubpelLds &sp)
{
    WinPos w;
    w.x0 = min((bx - reach - 3) & ~1, ref.py - kApron - kWinLS);
    w.y0 = by - reach - 3;
    w.y1 = w.y0 + 2 * reach + 22;
    w.cx0 = min((((bx - reach) >> 1) - 3) & ~1, ref.pc - kApron - kWinCS);
    w.cy0 = ((by - reach) >> 1) - 3;
    w.cy1 = ((by + reach) >> 1) + 11;
    dma_rows<kWinLS / 2, W>(reinterpret_cast<const uint32_t *>(ref.y + w.y0 * ref.py + w.x0),
                            ref.py / 2, w.y1 - w.y0, reinterpret_cast<uint32_t *>(sp.win.y));
    dma_rows<kWinCS / 2, W>(reinterpret_cast<const uint32_t *>(ref.u + w.cy0 * ref.pc + w.cx0),
                            ref.pc / 2, w.cy1 - w.cy0, reinterpret_cast<uint32_t *>(sp.win.u));
    dma_rows<kWinCS / 2, W>(reinterpret_cast<const uint32_t *>(ref.v + w.cy0 * ref.pc + w.cx0),
                            ref.pc / 2, w.cy1 - w.cy0, reinterpret_cast<uint32_t *>(sp.win.v));
    return w;
}

// Every wave's LDS copies have landed and are visible to the workgroup.
__device__ __forceinline__ void window_ready()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0)
    __syncthreads();
}

// The LDS ordering a search needs between its steps: the workgroup's, or,
// for a wave that searches alone inside a larger workgroup (SOLO), its own
// (a wave's LDS accesses complete in order, and its LDS copies have landed
// once its vmcnt is zero).
template <bool SOLO>
__device__ __forceinline__ void group_sync()
{
    if constexpr (SOLO) wave_sync();
    else __syncthreads();
}

template <bool SOLO>
__device__ __forceinline__ void window_ready()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0)
    group_sync<SOLO>();
}

// Four consecutive samples from sample index i of an LDS array with a
// dword-aligned base: three aligned dword reads and two funnel shifts, so
// no LDS access straddles a dword whatever the parity of i.
__device__ __forceinline__ void lds_load4(const int16_t *base, int i, uint32_t &lo, uint32_t &hi)
{
    const uint32_t *q = reinterpret_cast<const uint32_t *>(base) + (i >> 1);
    const uint32_t d0 = q[0], d1 = q[1], d2 = q[2];
    const uint32_t sh = (uint32_t)(i & 1) << 4;
    lo = __builtin_amdgcn_alignbit(d1, d0, sh);
    hi = __builtin_amdgcn_alignbit(d2, d1, sh);
}

// A candidate block from the staged window, in the lane layout (window
// rows have even strides, so a row's parity is its column's).
__device__ __forceinline__ void load_cand_win(const SubpelLds &sp, const WinPos &w, const Geo &g,
                                              int cx, int cy, Blk &b)
{
    const int ly = (cy - w.y0 + g.oy) * kWinLS + (cx - w.x0 + g.ox);
#pragma unroll
    for (int r = 0; r < 4; r++) lds_load4(sp.win.y, ly + r * kWinLS, b.y[2 * r], b.y[2 * r + 1]);
    const int lc = ((cy >> 1) - w.cy0 + g.cr) * kWinCS + ((cx >> 1) - w.cx0 + g.cc);
    lds_load4(sp.win.u, lc, b.u[0], b.u[1]);
    lds_load4(sp.win.v, lc, b.v[0], b.v[1]);
}

// Six-tap outputs of an R x C plane into dst (row-major, row stride C):
// output (r, c) filters src[r * sp + c + t * ts], t = 0..5 (interp.cpp:18-41).
// Thread t of the W waves computes outputs t, t + 64W, ...; positions
// advance by constants (no division in the loop), and each chunk of kChunk
// outputs issues all of its LDS loads before any arithmetic, so their
// latencies overlap.
template <int R, int C, int W>
__device__ __forceinline__ void filter_plane(const int16_t *src, int sp, int ts, int16_t *dst)
{
    constexpr int N = R * C;
    constexpr int NT = 64 * W;
    constexpr int K = (N + NT - 1) / NT;
    constexpr int kChunk = 4;
    const int tid = W > 1 ? (int)threadIdx.x : lane_id();
    const int r0 = tid / C, c0 = tid - (tid / C) * C;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += kChunk) {
        int v[kChunk][6];
#pragma unroll
        for (int j = 0; j < kChunk; j++) {
            const int k = k0 + j;
            if (k < K && (k < K - 1 || tid + NT * k < N)) {
                int c = c0 + (NT * k) % C, r = r0 + (NT * k) / C;
                if (c >= C) { c -= C; r++; }
                const int16_t *q = src + r * sp + c;
#pragma unroll
                for (int t = 0; t < 6; t++) v[j][t] = q[t * ts];
            }
        }
#pragma unroll
        for (int j = 0; j < kChunk; j++) {
            const int k = k0 + j;
            if (k < K && (k < K - 1 || tid + NT * k < N)) {
                const int sum = v[j][0] - 5 * v[j][1] + 20 * v[j][2] + 20 * v[j][3] -
                                5 * v[j][4] + v[j][5];
                dst[tid + NT * k] = (int16_t)((sum + 16) >> 5);
            }
        }
    }
}

// First pass: the h and v planes of one picture plane.  `plane` addresses
// the block base (bx, by) and has row pitch `pitch`.  h row r, column c is
// the horizontal half between columns c - 1 and c of row r - 3; v row r is
// the vertical half between rows r - 1 and r.
template <int B, int W>
__device__ __forceinline__ void build_hv(const int16_t *plane, int pitch, HalfPlanes<B> &hp)
{
    using P = HalfPlanes<B>;
    filter_plane<P::HR, P::HC, W>(plane - 3 * pitch - 3, pitch, 1, hp.h);
    filter_plane<P::VR, P::VC, W>(plane - 3 * pitch, pitch, pitch, hp.v);
}

// Second pass: the diagonal plane, a vertical six-tap over the h plane
// (plane row -1 starts at h row 0, i.e. picture row -3).
template <int B, int W>
__device__ __forceinline__ void build_d(HalfPlanes<B> &hp)
{
    using P = HalfPlanes<B>;
    filter_plane<P::DR, P::DC, W>(hp.h, P::HC, P::HC, hp.d);
}

// Where direction (dx, dy)'s half-pel block sits inside the planes.
struct PlaneView {
    const int16_t *arr;     // the plane (dword-aligned)
    int off, stride;        // the block's first sample and row stride in it
};

template <int B>
__device__ __forceinline__ PlaneView plane_view(const HalfPlanes<B> &hp, int dx, int dy)
{
    using P = HalfPlanes<B>;
    const int xo = dx > 0 ? 0 : -1, yo = dy > 0 ? 0 : -1;
    PlaneView pv;
    if (dy == 0) { pv.arr = hp.h; pv.off = 3 * P::HC + (xo + 1); pv.stride = P::HC; }
    else if (dx == 0) { pv.arr = hp.v; pv.off = (yo + 1) * P::VC; pv.stride = P::VC; }
    else { pv.arr = hp.d; pv.off = (yo + 1) * P::DC + (xo + 1); pv.stride = P::DC; }
    return pv;
}

// One direction's half-pel block from its plane views, in the lane layout.
__device__ __forceinline__ void load_view(const PlaneView &vy, const PlaneView &vu,
                                          const PlaneView &vv, const Geo &g, Blk &b)
{
    const int ly = vy.off + g.oy * vy.stride + g.ox;
#pragma unroll
    for (int r = 0; r < 4; r++) lds_load4(vy.arr, ly + r * vy.stride, b.y[2 * r], b.y[2 * r + 1]);
    lds_load4(vu.arr, vu.off + g.cr * vu.stride + g.cc, b.u[0], b.u[1]);
    lds_load4(vv.arr, vv.off + g.cr * vv.stride + g.cc, b.v[0], b.v[1]);
}

struct SearchOut {
    int best_x, best_y, sad, mad, ssd;
    float cost;
    int sp_en, sp_amt, sp_idx, btype;
    int n_int, n_ax, n_dg;
};

// Offsets of scan slot k (0..7) of a 3x3 level without its centre, in scan
// order (j outer, i inner; motion.cpp:388-389, 464-465).  Slot k is also the
// frac index of the sub-pel direction (motion.cpp:68-91).
__device__ __forceinline__ constexpr int slot_i(int k) { return (k < 4 ? k : k + 1) % 3 - 1; }
__device__ __forceinline__ constexpr int slot_j(int k) { return (k < 4 ? k : k + 1) / 3 - 1; }

// The same for a run-time k (0..7): the offsets + 1 as 2-bit fields of one
// constant, three scalar instructions instead of a division by 3.
constexpr uint32_t kSlotI = 0x9224u, kSlotJ = 0xA940u;
__device__ __forceinline__ int slot_i_rt(int k) { return (int)((kSlotI >> (2 * k)) & 3u) - 1; }
__device__ __forceinline__ int slot_j_rt(int k) { return (int)((kSlotJ >> (2 * k)) & 3u) - 1; }
static_assert(((kSlotI >> 0) & 3) == 0 && ((kSlotI >> 8) & 3) == 2 && ((kSlotI >> 12) & 3) == 1 &&
              ((kSlotJ >> 6) & 3) == 1 && ((kSlotJ >> 10) & 3) == 2 && ((kSlotJ >> 4) & 3) == 0,
              "slot offset tables");

// RD cost of a candidate (motion.cpp:135-145): fp32 lambda * bits, then
// + (float)SATD, no contraction.  Costs are non-negative, so their IEEE bits
// order like the values and the fold compares them as integers.
__device__ __forceinline__ uint32_t cost_bits(int sad, int satd_raw, int mvx, int mvy, float lambda)
{
    if (!(lambda > 0.0f)) return __float_as_uint(__int2float_rn(sad));
    const int bits = abs16(mvx) + abs16(mvy) + 1;   // motion.cpp:120-123
    return __float_as_uint(__fadd_rn(__int2float_rn((satd_raw + 1) >> 1),
                                     __fmul_rn(lambda, __int2float_rn(bits))));
}

// Bits 0, 16, 32 and 48 of a wave mask (lane 0 of each 16-lane group) as
// bits 0..3.
__device__ __forceinline__ uint32_t group_bits(uint64_t m)
{
    return (uint32_t)((m & 1u) | ((m >> 15) & 2u) | ((m >> 30) & 4u) | ((m >> 45) & 8u));
}

// SAD and MAD of this lane's samples against the half-pel block of one
// direction (its plane views) and against the quarter-pel block
// (full + half + 1) >> 1 (interp.cpp:166-202), per lane before reduction.
// For a pair narrow_pair accepts (range [lo, hi], hi - lo <= 2047, within
// [-8192, 8191]), a 6-tap pass overshoots [lo, hi] by at most 10/32 of its
// span (the negative taps' weight), 640, and the diagonal's second pass by
// 10/32 of 2047 + 2 * 640 more, 1040: every half-pel sample lies within
// 1680 of [lo, hi], in [-9872, 9871].  Differences to the source then stay
// within 2047 + 1680 = 3727 and full + half + 1 within [-18064, 18063], all
// exact in the packed int16 lanes.
__device__ __forceinline__ void subpel_metrics(bool narrow, const Blk &s, const Blk &full,
                                               const PlaneView &vy, const PlaneView &vu,
                                               const PlaneView &vv, const Geo &g, int &sh, int &mh,
                                               int &sq, int &mq)
{
    Blk h;
    load_view(vy, vu, vv, g, h);
    if (narrow) {
        const s16x2 one = (s16x2)(1);
        s16x2 sumh = (s16x2)(0), sumq = (s16x2)(0), mxh = (s16x2)(0), mxq = (s16x2)(0);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const s16x2 hv = pk(h.y[k]), qv = (pk(full.y[k]) + hv + one) >> one;
            const s16x2 ah = pabs(pk(s.y[k]) - hv), aq = pabs(pk(s.y[k]) - qv);
            sumh += ah; sumq += aq;
            mxh = pmax(mxh, ah); mxq = pmax(mxq, aq);
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const s16x2 hu = pk(h.u[k]), qu = (pk(full.u[k]) + hu + one) >> one;
            const s16x2 hw = pk(h.v[k]), qw = (pk(full.v[k]) + hw + one) >> one;
            mxh = pmax(mxh, pmax(pabs(pk(s.u[k]) - hu), pabs(pk(s.v[k]) - hw)));
            mxq = pmax(mxq, pmax(pabs(pk(s.u[k]) - qu), pabs(pk(s.v[k]) - qw)));
        }
        sh = (int)hsum(sumh);
        sq = (int)hsum(sumq);
        mh = pmax(mxh, pswap(mxh)).x;
        mq = pmax(mxq, pswap(mxq)).x;
        return;
    }
    sh = 0; mh = 0; sq = 0; mq = 0;
    auto one_sample = [&](int sv, int fv, int hv, bool luma) {
        const int qv = (int16_t)((fv + hv + 1) >> 1);
        const int ah = abs(sv - hv), aq = abs(sv - qv);
        if (luma) { sh += ah; sq += aq; }
        mh = max(mh, ah);
        mq = max(mq, aq);
    };
#pragma unroll
    for (int k = 0; k < 8; k++) {
        one_sample(lo16(s.y[k]), lo16(full.y[k]), lo16(h.y[k]), true);
        one_sample(hi16(s.y[k]), hi16(full.y[k]), hi16(h.y[k]), true);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        one_sample(lo16(s.u[k]), lo16(full.u[k]), lo16(h.u[k]), false);
        one_sample(hi16(s.u[k]), hi16(full.u[k]), hi16(h.u[k]), false);
        one_sample(lo16(s.v[k]), lo16(full.v[k]), lo16(h.v[k]), false);
        one_sample(hi16(s.v[k]), hi16(full.v[k]), hi16(h.v[k]), false);
    }
}

// The closed-form fold of one level (motion.cpp:388-405 over
// evaluate_motion_candidate), given lane k (k < 8) holding scan slot k's
// validity, cost bits, SAD and MAD.  Order: the level's starting best b0,
// slots 0..3, the centre (b0 again, with its true distance), slots 4..7.
//  * If a valid slot has MAD < T, the first such slot is accepted
//    unconditionally and the level ends there (copy state).
//  * Otherwise the best ends at the minimum cost m: the first element (in
//    order) at cost m, replaced by any later element at cost m with a
//    strictly smaller ssd and SAD < 8192.
struct Best {
    int x, y, sad, mad, ssd;
    uint32_t cost;
};

__device__ __forceinline__ bool fold_level(Best &b, int base_x, int base_y, int step, int x, int y,
                                           int thresh, uint32_t mask, bool valid, uint32_t cost,
                                           int sad, int mad, int &n_int)
{
    auto dist = [&](int cx, int cy) { return (cx - x) * (cx - x) + (cy - y) * (cy - y); };
    auto take = [&](int k) {
        b.x = base_x + slot_i_rt(k) * step;
        b.y = base_y + slot_j_rt(k) * step;
        b.sad = lane_value(sad, k);
        b.mad = lane_value(mad, k);
        b.ssd = dist(b.x, b.y);
    };
    const uint32_t copym = (uint32_t)__ballot(valid && mad < thresh) & 0xFFu;
    if (copym) {
        const int p = __builtin_ctz(copym);
        n_int += __builtin_popcount(mask & ((2u << p) - 1u));
        take(p);
        b.cost = (uint32_t)lane_value((int)cost, p);
        return true;                                   // motion.cpp:402-403
    }
    n_int += __builtin_popcount(mask);
    // Minimum over the slots in lanes 0..7 (quad swaps and the half-row
    // mirror reach all eight).
    uint32_t ml = valid ? cost : 0xFFFFFFFFu;
    ml = min(ml, (uint32_t)dpp<0xB1>((int)ml));
    ml = min(ml, (uint32_t)dpp<0x4E>((int)ml));
    ml = min(ml, (uint32_t)dpp<0x141>((int)ml));
    const uint32_t m = min(b.cost, (uint32_t)lane_value((int)ml, 0));
    const uint32_t eq = (uint32_t)__ballot(valid && cost == m) & 0xFFu;
    bool have = b.cost == m;
    b.cost = m;
    // The usual cases.  The level's start stays best alone: only the centre
    // (the start again) is visited, which can lower best_ssd to the true
    // distance.  Or exactly one slot reaches the minimum: it is taken (no
    // earlier element ties with it).
    if (eq == 0) {
        const int d = dist(base_x, base_y);
        if (d < b.ssd && b.sad < kSadThreshold) b.ssd = d;
        return false;
    }
    if (!have && (eq & (eq - 1u)) == 0) {
        take(__builtin_ctz(eq));
        return false;
    }
    // Ties: visit order as bits -- slots 0..3, the centre (bit 4), slots
    // 4..7.
    uint32_t seq = (eq & 0xFu) | (have ? 0x10u : 0u) | ((eq & 0xF0u) << 1);
    Best c = b;
    while (seq) {
        const int q = __builtin_ctz(seq);
        seq &= seq - 1;
        int ex = base_x, ey = base_y, esad = b.sad, emad = b.mad;
        if (q != 4) {
            const int k = q < 4 ? q : q - 1;
            ex = base_x + slot_i_rt(k) * step;
            ey = base_y + slot_j_rt(k) * step;
            esad = lane_value(sad, k);
            emad = lane_value(mad, k);
        }
        const int essd = dist(ex, ey);
        if (!have || (essd < c.ssd && esad < kSadThreshold)) {
            c.x = ex; c.y = ey; c.sad = esad; c.mad = emad; c.ssd = essd;
        }
        have = true;
    }
    b = c;
    return false;
}

// Cross-wave hand-off without cache-wide fences.  The MI355X has one L2 per
// XCD, and an agent-scope release or acquire fence writes back or
// invalidates a whole L2 (every coefficient and level store in flight).  The
// few words other waves read -- MB predictor words and search results -- are
// instead written and read with device-scope (sc1) accesses, which go to the
// coherent side of the L2s.  Each such word is one 64-bit access carrying
// its own published bit, so a reader needs no flag, no ordering between
// words and no wait on the writer's side.
// (Every such word lives in global memory: the accesses take the pointer as
// a global one, so they are global, not flat, instructions and do not count
// against the LDS wait counter.)
#define EVXA_GLOBAL(T, p) ((__attribute__((address_space(1))) T *)(p))

__device__ __forceinline__ uint64_t ld_relaxed64(const uint64_t *p)
{
    return __hip_atomic_load(EVXA_GLOBAL(const uint64_t, p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed64(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(EVXA_GLOBAL(uint64_t, p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t ld_relaxed32(const int32_t *p)
{
    return __hip_atomic_load(EVXA_GLOBAL(const int32_t, p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed32(int32_t *p, int32_t v)
{
    __hip_atomic_store(EVXA_GLOBAL(int32_t, p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every store of this wave has completed (device-scope stores are then
// visible to every agent-scope load).
__device__ __forceinline__ void stores_done()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// A MB's predictor word: bit 63 = published, bits 32..39 = the block type's
// motion flag (the only type bit the MV predictor reads), bits 0..31 = the
// integer MV (x low, y high; record word 1).  Zero until published.
__device__ __forceinline__ uint64_t pred_word(bool motion, uint32_t mv)
{
    return (1ull << 63) | ((uint64_t)(motion ? EVXA_BLOCK_MOTION : 0) << 32) | mv;
}

__device__ __forceinline__ uint32_t mv_word(int mvx, int mvy)
{
    return (uint32_t)(uint16_t)mvx | ((uint32_t)(uint16_t)mvy << 16);
}

// perform_inter_search (motion.cpp:567-679) for the MB at (x, y) against
// `ref` (a pw x ph picture), predictor (pmx, pmy) as compute_mv_predictor
// produced it.  All lanes of the NW waves call it with uniform arguments; it
// returns uniform values.  `s` is the source MB and `pre` the reference block
// at (x, y), both in the lane layout; `w` is the window already staged in
// `sp` (any position).  `early`, if given, is the MB's predictor word, to be
// published as soon as it is final (single-reference P frames).
//
// Each lane group (16 lanes) evaluates one candidate position at a time,
// one lane per 4x4 luma block plus 8 chroma samples.  With one wave, group
// g takes scan slots g and g + 4 of every level and sub-pel round; with two
// waves, group g of wave h takes slot 4h + g and the waves exchange their
// slots' results through LDS.  The ordered fold then runs on wave-uniform
// values, identically in every wave.
template <int NW, bool SOLO = false>
__device__ __forceinline__ SearchOut inter_search(const DevPic &ref, int pw, int ph, bool narrow, const Blk &s,
                                  const Geo &geo, const Blk &pre, WinPos w, int x, int y,
                                  int thresh, float lambda, int pmx, int pmy, SubpelLds &sp,
                                  uint64_t *early = nullptr)
{
    const int lane = lane_id();
    const int g = lane >> 4;
    const int sa = NW == 1 ? g : 4 * wave_id() + g;       // this group's (first) slot
    const int ia = slot_i(sa), ja = slot_j(sa), ib = slot_i(g + 4), jb = slot_j(g + 4);
    const int max_x = pw - 16, max_y = ph - 16;
    Best b;
    b.x = x; b.y = y; b.ssd = kInt32Max;
    int sp_en = 0, sp_amt = 0, sp_idx = 0, n_int = 1, n_ax = 0, n_dg = 0;
    auto dist = [&](int cx, int cy) { return (cx - x) * (cx - x) + (cy - y) * (cy - y); };

    // Initial candidate (motion.cpp:600-611) in group 0 and the predictor
    // candidate (motion.cpp:621-634, clamped into the picture) in group 1
    // (every wave evaluates both).
    const bool has_pred = pmx != 0 || pmy != 0;
    const int px = min(max(x + pmx, 0), max_x), py = min(max(y + pmy, 0), max_y);
    int sad, mad, satd;
    if (has_pred) {
        Blk rp, rb;
        if (win_holds_level(w, px, py, 0, max_x, max_y)) load_cand_win(sp, w, geo, px, py, rp);
        else load_cand(ref, geo, px, py, rp);
        select_blk(g == 1, pre, rp, rb);
        cand_metrics(narrow, s, rb, sad, mad, satd);
    } else {
        cand_metrics(narrow, s, pre, sad, mad, satd);
    }
    const uint32_t cst0 = g == 1 ? cost_bits(sad, satd, px - x, py - y, lambda)
                                 : cost_bits(sad, satd, 0, 0, lambda);
    b.sad = lane_value(sad, 0);
    b.mad = lane_value(mad, 0);
    b.cost = (uint32_t)lane_value((int)cst0, 0);

    if (has_pred) {
        // evaluate_motion_candidate's acceptance (motion.cpp:125-178), C
        // precedence: && binds tighter than ||.
        const int csad = lane_value(sad, 16), cmad = lane_value(mad, 16);
        const uint32_t ccost = (uint32_t)lane_value((int)cst0, 16);
        const int ssd = dist(px, py);
        bool accept;
        if (b.mad < thresh)
            accept = cmad < b.mad || (cmad == b.mad && ssd < b.ssd);
        else
            accept = (ccost < b.cost ||
                      ((ccost == b.cost && ssd < b.ssd) && csad < kSadThreshold)) ||
                     cmad < thresh;
        if (accept) {
            b.x = px; b.y = py;
            b.sad = csad; b.ssd = ssd; b.mad = cmad; b.cost = ccost;
        }
        n_int++;
    }

    if (b.mad >= thresh) {
        int radius = kSearchRadius;
        if (b.sad < kSadThreshold / 4) radius = kSearchRadius / 4;
        else if (b.sad < kSadThreshold / 2) radius = kSearchRadius / 2;

        bool copy = false;
        int parity = 0;
        for (int step = radius; step > 0 && !copy; step >>= 1) {
            // perform_inter_motion_search (motion.cpp:381-406).
            const int base_x = b.x, base_y = b.y;
            bool in_win = win_holds_level(w, base_x, base_y, step, max_x, max_y);
            if (!in_win && step <= 8) {
                w = stage_window<NW>(ref, base_x, base_y, 2 * step - 1, sp);
                window_ready<SOLO>();
                in_win = true;
            }
            uint32_t mask = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int cx = base_x + slot_i(k) * step, cy = base_y + slot_j(k) * step;
                if (cx >= 0 && cx <= max_x && cy >= 0 && cy <= max_y) mask |= 1u << k;
            }
            // This group's slot(s), clamped into the picture (an out-of-range
            // slot is evaluated at a safe position and never folded).
            const int ax = min(max(base_x + ia * step, 0), max_x);
            const int ay = min(max(base_y + ja * step, 0), max_y);
            Blk ra;
            if (in_win) load_cand_win(sp, w, geo, ax, ay, ra);
            else load_cand(ref, geo, ax, ay, ra);
            int sad_a, mad_a, satd_a;
            if constexpr (NW == 1) {
                const int bxx = min(max(base_x + ib * step, 0), max_x);
                const int byy = min(max(base_y + jb * step, 0), max_y);
                Blk rb;
                const bool two = (mask & 0xF0u) != 0;
                if (two) {
                    if (in_win) load_cand_win(sp, w, geo, bxx, byy, rb);
                    else load_cand(ref, geo, bxx, byy, rb);
                }
                cand_metrics(narrow, s, ra, sad_a, mad_a, satd_a);
                const uint32_t cst_a = cost_bits(sad_a, satd_a, ax - x, ay - y, lambda);
                int sad_b = 0, mad_b = 0, satd_b = 0;
                uint32_t cst_b = 0;
                if (two) {
                    cand_metrics(narrow, s, rb, sad_b, mad_b, satd_b);
                    cst_b = cost_bits(sad_b, satd_b, bxx - x, byy - y, lambda);
                }
                // Slot k's values into lane k (k < 8): slot k sits in lane
                // 16 * (k & 3), value a for k < 4 and b for k >= 4.
                const int from = ((lane & 3) << 4) << 2;             // byte address of that lane
                const bool hi = (lane & 4) != 0;
                auto pull = [&](int va, int vb) {
                    const int ra_ = __builtin_amdgcn_ds_bpermute(from, va);
                    const int rb_ = __builtin_amdgcn_ds_bpermute(from, vb);
                    return hi ? rb_ : ra_;
                };
                const uint32_t kc = (uint32_t)pull((int)cst_a, (int)cst_b);
                const int ks = pull(sad_a, sad_b);
                const int km = pull(mad_a, mad_b);
                const bool kv = lane < 8 && ((mask >> (lane & 7)) & 1u) != 0;
                copy = fold_level(b, base_x, base_y, step, x, y, thresh, mask, kv, kc, ks, km, n_int);
            } else {
                cand_metrics(narrow, s, ra, sad_a, mad_a, satd_a);
                const uint32_t cst_a = cost_bits(sad_a, satd_a, ax - x, ay - y, lambda);
                Exchange &xc = sp.xch;
                if ((lane & 15) == 0) {
                    xc.cost[parity][sa] = cst_a;
                    xc.sad[parity][sa] = sad_a;
                    xc.mad[parity][sa] = mad_a;
                }
                group_sync<SOLO>();
                const int k = lane & 7;
                const uint32_t kc = xc.cost[parity][k];
                const int ks = xc.sad[parity][k], km = xc.mad[parity][k];
                const bool kv = lane < 8 && ((mask >> k) & 1u) != 0;
                copy = fold_level(b, base_x, base_y, step, x, y, thresh, mask, kv, kc, ks, km, n_int);
                parity ^= 1;
            }
        }

        // Once the integer search has left the MB's own position, the block
        // type has its motion flag whatever the sub-pel step finds, and the
        // sub-pel step leaves the integer MV alone: with a single reference
        // that is the MB's final predictor word, so dependent MBs can start
        // during this MB's sub-pel round.
        if (early && (b.x != x || b.y != y) && (NW == 1 || wave_id() == 0) && lane == 0)
            st_relaxed64(early, pred_word(true, mv_word(b.x - x, b.y - y)));

        // perform_inter_subpixel_motion_search (motion.cpp:450-483): all
        // candidates interpolate around the integer best (motion.cpp:186, 226).
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int tx = b.x + slot_i(k), ty = b.y + slot_j(k);
            if (tx >= 0 && tx <= max_x && ty >= 0 && ty <= max_y) mask |= 1u << k;
        }
        if (mask) {
            const int bx = b.x, by = b.y;
            if (!win_holds_subpel(w, bx, by)) {
                w = stage_window<NW>(ref, bx, by, 0, sp);
                window_ready<SOLO>();
            }
            const int wlx = bx - w.x0, wly = by - w.y0;
            const int wcx = (bx >> 1) - w.cx0, wcy = (by >> 1) - w.cy0;
            build_hv<16, NW>(sp.win.y + wly * kWinLS + wlx, kWinLS, sp.y);
            build_hv<8, NW>(sp.win.u + wcy * kWinCS + wcx, kWinCS, sp.u);
            build_hv<8, NW>(sp.win.v + wcy * kWinCS + wcx, kWinCS, sp.v);
            Blk full;
            load_cand_win(sp, w, geo, bx, by, full);
            group_sync<SOLO>();
            build_d<16, NW>(sp.y);
            build_d<8, NW>(sp.u);
            build_d<8, NW>(sp.v);
            group_sync<SOLO>();
            // Half and quarter SAD/MAD of candidate j = 2 * direction +
            // (quarter) in lane j (j < 16).
            int jsad, jmad;
            const int j = lane & 15, jd = j >> 1, jq = j & 1;
            if constexpr (NW == 1) {
                int m[8];
#pragma unroll
                for (int slot = 0; slot < 2; slot++) {
                    const int di = slot ? ib : ia, dj = slot ? jb : ja;
                    int sh, mh, sq, mq;
                    subpel_metrics(narrow, s, full, plane_view<16>(sp.y, di, dj),
                                   plane_view<8>(sp.u, di, dj), plane_view<8>(sp.v, di, dj), geo,
                                   sh, mh, sq, mq);
                    m[slot * 4 + 0] = row_sum(sh);
                    m[slot * 4 + 1] = row_max(mh);
                    m[slot * 4 + 2] = row_sum(sq);
                    m[slot * 4 + 3] = row_max(mq);
                }
                // Direction d's values sit in lane 16 * (d & 3), slot d >> 2.
                const int from = ((jd & 3) << 4) << 2;
                int pulled[8];
#pragma unroll
                for (int i = 0; i < 8; i++) pulled[i] = __builtin_amdgcn_ds_bpermute(from, m[i]);
                const int base = (jd >= 4 ? 4 : 0) + (jq ? 2 : 0);
                jsad = base == 0 ? pulled[0] : base == 2 ? pulled[2] : base == 4 ? pulled[4] : pulled[6];
                jmad = base == 0 ? pulled[1] : base == 2 ? pulled[3] : base == 4 ? pulled[5] : pulled[7];
            } else {
                int sh, mh, sq, mq;
                subpel_metrics(narrow, s, full, plane_view<16>(sp.y, ia, ja),
                               plane_view<8>(sp.u, ia, ja), plane_view<8>(sp.v, ia, ja), geo, sh, mh,
                               sq, mq);
                sh = row_sum(sh); mh = row_max(mh); sq = row_sum(sq); mq = row_max(mq);
                if ((lane & 15) == 0) {
                    sp.xch.sub[sa][0] = sh; sp.xch.sub[sa][1] = mh;
                    sp.xch.sub[sa][2] = sq; sp.xch.sub[sa][3] = mq;
                }
                group_sync<SOLO>();
                jsad = sp.xch.sub[jd][2 * jq];
                jmad = sp.xch.sub[jd][2 * jq + 1];
            }
            // evaluate_subpel_motion_candidate's acceptance (motion.cpp:196-260)
            // over the 16 candidates in order (lane j holds candidate j), in
            // closed form; there is no early exit.
            //  * Copy state from the start (best MAD < T): a candidate is taken
            //    iff its MAD is below the running best's, so the result is the
            //    first candidate at the smallest MAD, if that is below the
            //    starting MAD.
            //  * Otherwise the first candidate j* with MAD < T is taken
            //    unconditionally and switches to copy state; from j* on the
            //    result is the first candidate at the smallest MAD.
            //  * Without such a candidate, one is taken iff its SAD is below
            //    the running best's and below 8192: the first candidate at
            //    the smallest such SAD, if that is below the starting SAD.
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const bool valid = ((mask >> k) & 1u) != 0;
                const bool diag = slot_i(k) != 0 && slot_j(k) != 0;
                n_dg += (valid && diag) ? 1 : 0;
                n_ax += (valid && !diag) ? 1 : 0;
            }
            const bool vj = lane < 16 && ((mask >> (lane >> 1)) & 1u) != 0;
            auto row_min = [](int v) {
                v = min(v, dpp<0xB1>(v));
                v = min(v, dpp<0x4E>(v));
                v = min(v, dpp<0x141>(v));
                v = min(v, dpp<0x140>(v));
                return lane_value(v, 0);
            };
            int win = -1;
            if (b.mad < thresh) {
                const int m2 = row_min(vj ? jmad : kInt32Max);
                if (m2 < b.mad) win = (int)__builtin_ctzll(__ballot(vj && jmad == m2));
            } else {
                const uint64_t cm = __ballot(vj && jmad < thresh);
                if (cm) {
                    const int js = (int)__builtin_ctzll(cm);
                    const bool after = vj && lane >= js;
                    const int m2 = row_min(after ? jmad : kInt32Max);
                    win = (int)__builtin_ctzll(__ballot(after && jmad == m2));
                } else {
                    const bool ok = vj && jsad < kSadThreshold;
                    const int m1 = row_min(ok ? jsad : kInt32Max);
                    if (m1 < b.sad) win = (int)__builtin_ctzll(__ballot(ok && jsad == m1));
                }
            }
            if (win >= 0) {
                sp_en = 1;
                sp_amt = win & 1;
                sp_idx = win >> 1;
                b.sad = lane_value(jsad, win);
                b.mad = lane_value(jmad, win);
            }
        }
        group_sync<SOLO>();
    }

    SearchOut o;
    o.best_x = b.x; o.best_y = b.y;
    o.sad = b.sad; o.mad = b.mad; o.ssd = b.ssd;
    o.cost = __uint_as_float(b.cost);
    o.sp_en = sp_en; o.sp_amt = sp_amt; o.sp_idx = sp_idx;
    o.n_int = n_int; o.n_ax = n_ax; o.n_dg = n_dg;
    o.btype = 0;
    if (b.x != x || b.y != y || sp_en) o.btype |= EVXA_BLOCK_MOTION;
    if (b.mad < thresh) o.btype |= EVXA_BLOCK_COPY;
    return o;
}

// Scratch of the MB-level stages (selection, prediction, DCT).
struct MbLds {
    int16_t src[EVXA_MB_SAMPLES];
    int16_t pred[EVXA_MB_SAMPLES];
    int16_t pred_b[EVXA_MB_SAMPLES];
    int16_t res[EVXA_MB_SAMPLES];
    int16_t tmp[16 * 21 + 2 * 8 * 13];   // diagonal H pass / DCT rows
};

union alignas(16) Shared {
    SubpelLds sp;
    MbLds mb;
};
// The MB arrays overlap only the staged window: a search's half-pel planes
// survive the selection that follows it.
static_assert(sizeof(MbLds) <= offsetof(SubpelLds, y), "MB arrays reach the half-pel planes");
static_assert(sizeof(MbLds) <= offsetof(SubpelLds, win) + offsetof(Window, u),
              "MB arrays reach the window's chroma area");

// A window of a picture staged in LDS: its planes (row strides py, pc) and
// where they start in the picture, luma (lx, ly) and chroma (cx, cy).  Reads
// index the arrays relative to those origins; every read stays inside.
struct WinView {
    const int16_t *y, *u, *v;
    int32_t py, pc;
    int32_t lx, ly, cx, cy;
};

// Where a picture view's plane arrays start, in picture coordinates: a
// device picture's pointers address sample (0, 0) itself.
__device__ __forceinline__ void view_origin(const DevPic &, int &lx, int &ly, int &cx, int &cy)
{
    lx = ly = cx = cy = 0;
}
__device__ __forceinline__ void view_origin(const WinView &w, int &lx, int &ly, int &cx, int &cy)
{
    lx = w.lx; ly = w.ly; cx = w.cx; cy = w.cy;
}

// Residual prediction for a search result into `out` (encode.cpp:696-723;
// interp.cpp:166-202 when sub-pel).  `tmp` holds the diagonal H pass.  Neither
// overlaps the picture or window read (restrict: the loads of later samples
// need not wait for the stores of earlier ones).
//
// `ref` is a device picture or a window of one staged in LDS (WinView).
template <class Pic>
__device__ void predict_mb(const Pic &ref, const SearchOut &r, int16_t *__restrict__ out,
                           int16_t *__restrict__ tmp)
{
    const int lane = lane_id();
    int olx, oly, ocx, ocy;
    view_origin(ref, olx, oly, ocx, ocy);
    const int bx = r.best_x - olx, by = r.best_y - oly;
    const int cbx = (r.best_x >> 1) - ocx, cby = (r.best_y >> 1) - ocy;
    int dx = 0, dy = 0;
    if (r.sp_en) frac_direction(r.sp_idx, dx, dy);
    const int xo = dx > 0 ? 0 : -1, yo = dy > 0 ? 0 : -1;
    auto plane_of = [&](int i, const int16_t *&pl, int &pitch, int &X, int &Y) {
        if (i < 256) { pl = ref.y; pitch = ref.py; X = bx + (i & 15); Y = by + (i >> 4); }
        else {
            const int k = (i - 256) & 63;
            pl = i < 320 ? ref.u : ref.v; pitch = ref.pc;
            X = cbx + (k & 7); Y = cby + (k >> 3);
        }
    };
    if (!r.sp_en) {
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
            const int16_t *pl; int pitch, X, Y;
            plane_of(i, pl, pitch, X, Y);
            out[i] = pl[Y * pitch + X];
        }
        wave_sync();
        return;
    }
    if (dx != 0 && dy != 0) {
        // interp.cpp:125-131: H pass over 21 (luma) / 13 (chroma) rows.
        for (int i = lane; i < 16 * 21 + 2 * 8 * 13; i += 64) {
            const int16_t *pl; int pitch, X, Y;
            if (i < 336) { pl = ref.y; pitch = ref.py; X = bx + (i & 15); Y = by + (i >> 4) - 2; }
            else {
                const int k = (i - 336) % 104;
                pl = i < 440 ? ref.u : ref.v; pitch = ref.pc;
                X = cbx + (k & 7); Y = cby + (k >> 3) - 2;
            }
            tmp[i] = (int16_t)tap6(pl + (Y + yo) * pitch + X + xo - 2, 1);
        }
        wave_sync();
    }
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        const int16_t *pl; int pitch, X, Y;
        plane_of(i, pl, pitch, X, Y);
        int half;
        if (dy == 0) half = tap6(pl + Y * pitch + X + xo - 2, 1);
        else if (dx == 0) half = tap6(pl + (Y + yo - 2) * pitch + X, pitch);
        else {
            // interp.cpp:134-149 over the H pass.
            const int16_t *t;
            int ts;
            if (i < 256) { t = tmp + (i >> 4) * 16 + (i & 15); ts = 16; }
            else {
                const int k = (i - 256) & 63;
                t = tmp + 336 + (i < 320 ? 0 : 104) + (k >> 3) * 8 + (k & 7); ts = 8;
            }
            half = tap6(t, ts);
        }
        const int full = pl[Y * pitch + X];
        out[i] = (int16_t)(r.sp_amt ? (int16_t)((full + half + 1) >> 1) : half);
    }
    wave_sync();
}

// predict_mb of a device picture over a copy of the region it reads, staged
// into LDS (`stage`, kPredStage samples) by one DMA round instead of rounds
// of 2-byte loads from the picture: luma rows and columns -3 .. +18 of the
// result's integer position, chroma -3 .. +10 of its half (the 6-tap reach
// in every direction), from even columns so that rows copy as dwords; the
// picture's 24-sample apron holds every row.
constexpr int kPredLS = 24, kPredLR = 22, kPredCS = 16, kPredCR = 14;
constexpr int kPredStage = kPredLS * kPredLR + 2 * kPredCS * kPredCR;
static_assert(kPredStage <= kWinCR * kWinCS, "a window's chroma plane holds the staging");
__device__ void predict_mb_lds(const DevPic &ref, const SearchOut &r, int16_t *out, int16_t *tmp,
                               int16_t *stage)
{
    const int lx = (r.best_x - 3) & ~1, ly = r.best_y - 3;
    const int cx = ((r.best_x >> 1) - 3) & ~1, cy = (r.best_y >> 1) - 3;
    int16_t *sy = stage, *su = stage + kPredLS * kPredLR, *sv = su + kPredCS * kPredCR;
    dma_rows<kPredLS / 2, 1>(reinterpret_cast<const uint32_t *>(ref.y + ly * ref.py + lx), ref.py / 2,
                             kPredLR, reinterpret_cast<uint32_t *>(sy));
    dma_rows<kPredCS / 2, 1>(reinterpret_cast<const uint32_t *>(ref.u + cy * ref.pc + cx), ref.pc / 2,
                             kPredCR, reinterpret_cast<uint32_t *>(su));
    dma_rows<kPredCS / 2, 1>(reinterpret_cast<const uint32_t *>(ref.v + cy * ref.pc + cx), ref.pc / 2,
                             kPredCR, reinterpret_cast<uint32_t *>(sv));
    __builtin_amdgcn_s_waitcnt(0x0F70);     // vmcnt(0): the copies have landed
    wave_sync();
    const WinView wv{sy, su, sv, kPredLS, kPredCS, lx, ly, cx, cy};
    predict_mb(wv, r, out, tmp);
}

// transform.cpp:264-284 on a line of 8 samples.  The 8-term products fold
// the table's symmetry: even rows take sums v[k] + v[7 - k], odd rows
// differences v[k] - v[7 - k] (the same integers; no term exceeds
// 8 * 65535 * 128 in magnitude, so nothing wraps).
__device__ __forceinline__ void dct_line(const int16_t *src, int sp, int16_t *dst, int dp)
{
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = src[k * sp];
    int e[4], o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { e[k] = v[k] + v[7 - k]; o[k] = v[k] - v[7 - k]; }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int *x = (i & 1) ? o : e;
        int t = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) t += x[k] * kDct8Half[i][k];
        t = (i == 0) ? (t * 45) / 128 : t / 2;
        // rounded_div(t, 128), evx_math.h:228-236
        t = t < 0 ? (t - 64) / 128 : (t + 64) / 128;
        dst[i * dp] = (int16_t)t;
    }
}

// Origin and pitch of 8x8 block `blk` (0..3 luma quadrants, 4 U, 5 V) in
// MB order.
__device__ __forceinline__ void block_origin(int blk, int &origin, int &pitch)
{
    if (blk < 4) { origin = (blk >> 1) * 128 + (blk & 1) * 8; pitch = 16; }
    else { origin = 256 + (blk - 4) * 64; pitch = 8; }
}

// Residual and forward DCT of one MB held in LDS (encode.cpp:717-728;
// transform.cpp:435-452, 584-594): the coefficients replace the residual
// in `res`.
__device__ __forceinline__ void fdct_mb(const int16_t *src, const int16_t *pred, int16_t *res,
                                        int16_t *tmp)
{
    const int lane = lane_id();
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) res[i] = (int16_t)(src[i] - pred[i]);
    wave_sync();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int row = origin + (lane & 7) * pitch;
        dct_line(res + row, 1, tmp + row, 1);
    }
    wave_sync();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int col = origin + (lane & 7);
        dct_line(tmp + col, pitch, res + col, pitch);
    }
    wave_sync();
}

// compute_block_variance2 and query_block_quantization_parameter over the
// coefficients in `res` (analysis.h:253-275; quantize.cpp:97-118).
__device__ __forceinline__ void mb_qp(const int16_t *res, int quality, int &qp_out, int &var_out)
{
    const int lane = lane_id();
    // compute_block_variance2 over the 16x16 luma coefficients, (0,0)
    // skipped; the int32 sums wrap like the reference's.
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
    qp_out = min(max(q8 + delta, 1), 63);
    var_out = var2;
}

// Residual, forward DCT, adaptive QP, variance and inter quantisation of one
// MB held in LDS (encode.cpp:717-728; transform.cpp:435-452, 584-594;
// quantize.cpp:97-118, 217-252, 372-397; analysis.h:253-275).
// Coefficients and levels go to global memory; returns qp and variance2.
__device__ void dct_quant(const int16_t *src, const int16_t *pred, int16_t *res, int16_t *tmp,
                          int quality, int16_t *co_out, int16_t *lv_out, int &qp_out,
                          int &var_out)
{
    const int lane = lane_id();
    fdct_mb(src, pred, res, tmp);
    int qp, var2;
    mb_qp(res, quality, qp, var2);
    const int cqp = min(qp + 2, 63);
    const uint32_t m_luma = magic_of(qp << 1), m_chroma = magic_of(cqp << 1);

    // One sample per lane per round, rolled: the rounds' divisions would
    // otherwise all be in flight at once and raise the kernel's register
    // count (and lower its occupancy).
#pragma unroll 1
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        int pos, bq;
        uint32_t mq;
        if (i < 256) { pos = ((i >> 4) & 7) * 8 + (i & 7); bq = qp; mq = m_luma; }
        else { pos = (i - 256) & 63; bq = cqp; mq = m_chroma; }
        const int sv = res[i];
        const int qm = c_quant.qm[pos];
        const int qf = (int16_t)rdiv_magic(sv * 16, qm, c_quant.qm_magic[pos]);
        int w = c_quant.dz[pos];
        if (bq >= 16) w = min(w + 1, 6);
        const int sg = (qf > 0) - (qf < 0);
        const int dz = (int16_t)((sg * bq * w) >> 1);
        lv_out[i] = (int16_t)rdiv_magic(qf - dz, bq << 1, mq);
        co_out[i] = (int16_t)sv;
    }
    qp_out = qp;
    var_out = var2;
}

// The source MB from the lanes that hold it (lanes 0..15 own the whole MB in
// the lane layout) into `dst` (16x16 Y, 8x8 U, 8x8 V, row-major).
__device__ __forceinline__ void store_src(const Blk &s, const Geo &g, int16_t *dst)
{
    if (lane_id() < 16) {
        uint32_t *d = reinterpret_cast<uint32_t *>(dst);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            d[((g.oy + r) * 16 + g.ox) / 2] = s.y[2 * r];
            d[((g.oy + r) * 16 + g.ox) / 2 + 1] = s.y[2 * r + 1];
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            d[(256 + g.cr * 8 + g.cc) / 2 + k] = s.u[k];
            d[(320 + g.cr * 8 + g.cc) / 2 + k] = s.v[k];
        }
    }
}

// compute_mv_predictor (motion.cpp:271-346) from the predictor words of
// the left, top, top-right and top-left MBs (0 where absent) and the first
// two words (type, MV) of the temporal table's record, combined in the
// reference's order.
__device__ void mv_predictor(const DevJob &job, int bx, int by, uint64_t wl, uint64_t wt,
                             uint64_t wtr, uint64_t wtl, uint32_t p0, uint32_t p1, int &pmx,
                             int &pmy)
{
    const bool has_l = bx > 0, has_t = by > 0, has_tr = by > 0 && bx + 1 < job.wb;
    const bool has_tl = by > 0 && bx > 0, has_p = job.prev != nullptr;
    auto motion = [](uint64_t w) { return ((uint32_t)(w >> 32) & EVXA_BLOCK_MOTION) != 0; };
    int mx[3] = {0, 0, 0}, my[3] = {0, 0, 0}, n = 0;
    auto take = [&](uint32_t w1) {
        const int vx = (int16_t)(w1 & 0xFFFF), vy = (int16_t)(w1 >> 16);
        if (n == 0) { mx[0] = vx; my[0] = vy; }
        else if (n == 1) { mx[1] = vx; my[1] = vy; }
        else { mx[2] = vx; my[2] = vy; }
        n++;
    };
    if (has_l && motion(wl)) take((uint32_t)wl);
    if (has_t && motion(wt)) take((uint32_t)wt);
    if (has_tr && motion(wtr)) take((uint32_t)wtr);
    else if (has_tl && motion(wtl)) take((uint32_t)wtl);
    if (n < 3 && has_p && (p0 & EVXA_BLOCK_MOTION)) take(p1);
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

// Lane k < n stores word k of `w` to dst[k] (one coalesced store of a
// record that no wave of the launch reads).
__device__ __forceinline__ void write_words(uint32_t *dst, const uint32_t *w, int n)
{
    const int lane = lane_id();
    if (lane < n) {
        uint32_t v = w[0];
        for (int k = 1; k < n; k++) v = lane == k ? w[k] : v;
        dst[lane] = v;
    }
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

// Word i (0..2) of a search result in the SearchRec layout, bit 63 set.
__device__ __forceinline__ uint64_t pack_search(const SearchOut &o, int i)
{
    const uint64_t pub = 1ull << 63;
    if (i == 0) {
        const uint32_t pos = (uint32_t)(uint16_t)o.best_x | ((uint32_t)(uint16_t)o.best_y << 16);
        return pub | ((uint64_t)(uint32_t)o.mad << 32) | pos;
    }
    if (i == 1) return pub | ((uint64_t)(uint32_t)o.sad << 32) | fbits(o.cost);
    const uint32_t cnt = (uint32_t)(o.n_int & 0xFFFF) | ((uint32_t)(o.n_ax & 0xFF) << 16) |
                         ((uint32_t)(o.n_dg & 0xFF) << 24);
    const uint32_t fl = (uint32_t)(o.sp_en & 1) | ((uint32_t)(o.sp_amt & 1) << 1) |
                        ((uint32_t)(o.sp_idx & 7) << 2) | ((uint32_t)(o.btype & 0xFF) << 8);
    return pub | ((uint64_t)fl << 32) | cnt;
}

// Lanes 0..2 store this search's three result words (SearchRec layout,
// bit 63 set) for the wave that finishes the MB.
__device__ __forceinline__ void publish_search(uint64_t *dst, const SearchOut &o)
{
    const int lane = lane_id();
    const uint64_t w = pack_search(o, lane < 3 ? lane : 0);
    if (lane < 3) st_relaxed64(dst + lane, w);
}

// Search result k of a MB, read wave-uniformly from the words in `sw`
// (lane 3k + i holds word i of reference k's SearchRec).
__device__ __forceinline__ SearchOut unpack_search(uint64_t sw, int k)
{
    auto lo = [&](int i) { return (uint32_t)lane_value((int)(uint32_t)sw, 3 * k + i); };
    auto hi = [&](int i) { return (uint32_t)lane_value((int)(uint32_t)(sw >> 32), 3 * k + i) & 0x7FFFFFFFu; };
    SearchOut o;
    const uint32_t pos = lo(0), cnt = lo(2), fl = hi(2);
    o.best_x = (int16_t)(pos & 0xFFFF);
    o.best_y = (int16_t)(pos >> 16);
    o.mad = (int)hi(0);
    o.sad = (int)hi(1);
    o.cost = __uint_as_float(lo(1));
    o.ssd = 0;
    o.sp_en = fl & 1; o.sp_amt = (fl >> 1) & 1; o.sp_idx = (fl >> 2) & 7;
    o.btype = (fl >> 8) & 0xFF;
    o.n_int = cnt & 0xFFFF; o.n_ax = (cnt >> 16) & 0xFF; o.n_dg = cnt >> 24;
    return o;
}

// encode.cpp:215-247 for a P MB, against an empty (non-copy, INT32_MAX)
// sentinel: reference k's result from result(k).  Returns the index chosen.
template <typename F>
__device__ __forceinline__ int select_p(int n_refs, float lambda, F result, SearchOut &best,
                                        int &best_sad)
{
    int best_k = -1;
    bool best_copy = false;
    best_sad = kInt32Max;
    for (int k = 0; k < n_refs; k++) {
        const SearchOut r = result(k);
        const int bias = __float2int_rz(__fmul_rn(lambda, __int2float_rn((k + 1) * 2)));
        const bool copy = (r.btype & EVXA_BLOCK_COPY) != 0;
        if (copy != best_copy) {
            if (copy) { best = r; best_k = k; best_sad = r.sad; best_copy = true; }
        } else if (r.sad + bias < best_sad) {
            best = r; best_k = k; best_sad = r.sad; best_copy = copy;
        }
    }
    return best_k;
}

// One look (no waiting) at all of a P MB's search results; if every one is
// there, select and publish the MB's predictor word (the value the
// finishing wave publishes as well).
__device__ __forceinline__ void try_publish(const DevJob &job, int mb, int x, int y)
{
    const int lane = lane_id();
    const int n_words = 3 * job.n_refs;
    const uint64_t *p = reinterpret_cast<const uint64_t *>(job.searches + (size_t)mb * EVXA_MAX_REFS) +
                        (lane / 3) * 4 + lane % 3;
    const uint64_t sw = lane < n_words ? ld_relaxed64(p) : 0;
    if (__ballot(lane < n_words && (sw >> 63) == 0) != 0) return;
    SearchOut best;
    int best_sad;
    if (select_p(job.n_refs, job.lambda, [&](int k) { return unpack_search(sw, k); }, best,
                 best_sad) < 0)
        return;
    if (lane == 0)
        st_relaxed64(job.mvw + mb, pred_word((best.btype & EVXA_BLOCK_MOTION) != 0,
                                             mv_word(best.best_x - x, best.best_y - y)));
}

// The results of the searches of a MB's first n_refs - 1 references, which
// their waves publish for the wave of the last one: lane l < 3 (n_refs - 1)
// polls word l % 3 of reference l / 3 (they were claimed earlier, so their
// waves are running), then clears what it read for the next search, and
// this wave's own words too if it published them (`own_published`).
__device__ uint64_t gather_searches(const DevJob &job, int idx, int32_t *counters, bool own_published)
{
    const int lane = lane_id();
    const int n_words = 3 * (job.n_refs - 1);
    uint64_t sw = 0;
    if (n_words > 0) {
        uint64_t *p = reinterpret_cast<uint64_t *>(job.searches + (size_t)idx * EVXA_MAX_REFS) +
                      (lane / 3) * 4 + lane % 3;
        int spins = 0, nap = 1;
        for (;;) {
            if (lane < n_words) sw = ld_relaxed64(p);
            if (__ballot(lane < n_words && (sw >> 63) == 0) == 0) break;
            if (++spins > kSpinLimit) {
                if (lane == 0) atomicOr(&counters[kCounterError], kErrWaitTimeout);
                break;
            }
            for (int k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(1);
            nap = min(nap * 2, kMaxNap);
        }
        if (lane < n_words) st_relaxed64(p, 0ull);
        if (own_published && lane >= n_words && lane < n_words + 3) st_relaxed64(p, 0ull);
    }
    return sw;
}

// A MB's selection (encode.cpp:183-248 for P, 386-443 for B) and the
// searches it rests on.
struct Sel {
    SearchOut best, other;      // other: the backward search of a bidirectional MB
    int best_k, best_sad, bidir;
    int n_int, n_ax, n_dg;
};

// The prediction of an inter search's result into `out` (predict_mb's
// samples): for sub-pel results the half-pel planes the search built around
// its integer best (averaged with the reference's integer block for
// quarter-pel), else the reference's integer block.
__device__ __forceinline__ void predict_inter_result(const DevPic &ref, const SearchOut &r,
                                                     const SubpelLds &sp, int16_t *out)
{
    int dx = 0, dy = 0;
    if (r.sp_en) frac_direction(r.sp_idx, dx, dy);
    const PlaneView vy = plane_view<16>(sp.y, dx, dy), vu = plane_view<8>(sp.u, dx, dy),
                    vv = plane_view<8>(sp.v, dx, dy);
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64) {
        const bool luma = i < 256;
        const int k = luma ? i : (i - 256) & 63;
        const int col = luma ? k & 15 : k & 7, row = luma ? k >> 4 : k >> 3;
        const int16_t *pl = luma ? ref.y : (i < 320 ? ref.u : ref.v);
        const int pitch = luma ? ref.py : ref.pc;
        const int X = (luma ? r.best_x : r.best_x >> 1) + col;
        const int Y = (luma ? r.best_y : r.best_y >> 1) + row;
        const int full = pl[Y * pitch + X];
        int v = full;
        if (r.sp_en) {
            const PlaneView &pv = luma ? vy : (i < 320 ? vu : vv);
            const int half = pv.arr[pv.off + row * pv.stride + col];
            v = r.sp_amt ? (int16_t)((full + half + 1) >> 1) : half;
        }
        out[i] = (int16_t)v;
    }
    wave_sync();
}

// The two predictions of a B MB (motion.cpp:729-776) into sh.pred and
// sh.pred_b; returns the SAD of their average against sh.src over Y, U and V
// (motion.cpp:779-813).  `search`: the LDS the backward search just used
// (the MB arrays overlap only the start of its window): its half-pel planes
// give the backward prediction without interpolating again, and its
// window's chroma area stages the forward prediction's reads.  `fpred`: the
// forward prediction as the forward search's wave stored it, else nullptr.
__device__ __forceinline__ int b_predictions(const DevJob &job, const SearchOut &f, const SearchOut &b,
                                             MbLds &sh, SubpelLds *search = nullptr,
                                             const int16_t *fpred = nullptr)
{
    const int lane = lane_id();
    if (search) {
        if (fpred) {
            const int32_t *src = reinterpret_cast<const int32_t *>(fpred);
            int32_t *dst = reinterpret_cast<int32_t *>(sh.pred);
#pragma unroll
            for (int k = 0; k < EVXA_MB_SAMPLES / 128; k++) dst[lane + 64 * k] = ld_relaxed32(src + lane + 64 * k);
            wave_sync();
        } else {
            predict_mb_lds(job.refs[0], f, sh.pred, sh.tmp, search->win.u);
        }
        predict_inter_result(job.refs[1], b, *search, sh.pred_b);
    } else {
        predict_mb(job.refs[0], f, sh.pred, sh.tmp);
        predict_mb(job.refs[1], b, sh.pred_b, sh.tmp);
    }
    int bsad = 0;
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        const int a = (sh.pred[i] + sh.pred_b[i] + 1) >> 1;
        bsad += abs(sh.src[i] - a);
    }
    return wave_sum(bsad);
}

// The B prediction the selection chose, into sh.pred (after b_predictions).
__device__ __forceinline__ void b_choose(const Sel &sel, MbLds &sh)
{
    const int lane = lane_id();
    wave_sync();
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        if (sel.bidir) sh.pred[i] = (int16_t)((sh.pred[i] + sh.pred_b[i] + 1) >> 1);
        else if (sel.best_k == 1) sh.pred[i] = sh.pred_b[i];
    }
    wave_sync();
}

// Selection among a MB's searches, reference k's result from result(k); the
// source MB is in sh.src.  For B frames the chosen prediction is left in
// sh.pred.
template <typename F>
__device__ Sel select_mb(const DevJob &job, int x, int y, MbLds &sh, F result, int32_t *counters,
                         SubpelLds *search = nullptr, const int16_t *fpred = nullptr)
{
    const int lane = lane_id();
    const float lambda = job.lambda;
    Sel sel;
    sel.other = SearchOut{x, y, 0, 0, 0, 0.0f, 0, 0, 0, 0, 0, 0, 0};
    sel.best_k = -1;
    sel.best_sad = kInt32Max;
    sel.bidir = 0;
    sel.n_int = 0; sel.n_ax = 0; sel.n_dg = 0;
    if (job.type == EVXA_FRAME_P) {
        // encode.cpp:215-247 against an empty (non-copy, INT32_MAX) sentinel
        // (the same rule as select_p, written out: the compiler schedules
        // this path better that way).
        bool best_copy = false;
        for (int k = 0; k < job.n_refs; k++) {
            const SearchOut r = result(k);
            sel.n_int += r.n_int; sel.n_ax += r.n_ax; sel.n_dg += r.n_dg;
            const int bias = __float2int_rz(__fmul_rn(lambda, __int2float_rn((k + 1) * 2)));
            const bool copy = (r.btype & EVXA_BLOCK_COPY) != 0;
            if (copy != best_copy) {
                if (copy) { sel.best = r; sel.best_k = k; sel.best_sad = r.sad; best_copy = true; }
            } else if (r.sad + bias < sel.best_sad) {
                sel.best = r; sel.best_k = k; sel.best_sad = r.sad; best_copy = copy;
            }
        }
    } else {
        // encode.cpp:386-443 without the intra candidates; the bidirectional
        // candidate (motion.cpp:702-840) reuses both searches.
        const SearchOut f = result(0);
        const SearchOut b = result(1);
        sel.n_int = f.n_int + b.n_int; sel.n_ax = f.n_ax + b.n_ax; sel.n_dg = f.n_dg + b.n_dg;
        if (f.sad < sel.best_sad) { sel.best = f; sel.best_k = 0; sel.best_sad = f.sad; }
        if (b.sad < sel.best_sad) { sel.best = b; sel.best_k = 1; sel.best_sad = b.sad; }
        const int bsad = b_predictions(job, f, b, sh, search, fpred);
        const int bias = __float2int_rz(__fmul_rn(lambda, 8.0f));
        if (bsad + bias < sel.best_sad) {
            sel.bidir = 1; sel.best_sad = bsad; sel.best = f; sel.best_k = 0; sel.other = b;
        }
        b_choose(sel, sh);
    }
    if (sel.best_k < 0 || sel.best_sad == kInt32Max) {
        if (lane == 0) atomicOr(&counters[kCounterError], kErrSearchFailed);
        sel.best_k = 0;
        sel.best = SearchOut{x, y, sel.best_sad, 0, 0, 0.0f, 0, 0, 0, EVXA_BLOCK_COPY, 0, 0, 0};
    }
    return sel;
}

// The block type and MV of a selection, as the MV predictor reads them.
__device__ __forceinline__ uint64_t sel_word(const Sel &sel, int x, int y)
{
    const int btype = sel.bidir ? EVXA_BLOCK_MOTION : sel.best.btype;
    return pred_word((btype & EVXA_BLOCK_MOTION) != 0,
                     mv_word(sel.best.best_x - x, sel.best.best_y - y));
}

// The record of a selected MB, then (for non-copy MBs) its residual
// prediction, DCT and quantisation; `publish`: also store the MB's predictor
// word for the MBs that wait on it.  sh.src holds the source MB, and for B
// frames sh.pred the chosen prediction.
__device__ void emit_mb(const DevJob &job, int bx, int by, MbLds &sh, const Sel &sel, bool publish)
{
    const int lane = lane_id();
    const int x = bx * 16, y = by * 16;
    const int idx = by * job.wb + bx;
    const SearchOut &best = sel.best, &other = sel.other;
    const int bidir = sel.bidir, best_k = sel.best_k;
    const uint32_t ids = job.target_ids;
    uint32_t w[16];
    const int btype = bidir ? EVXA_BLOCK_MOTION : best.btype;
    const int mode = bidir ? 2 : (job.type == EVXA_FRAME_P ? 0 : best_k);
    const int tgt = (ids >> (8 * best_k)) & 0xFF;
    const int tgt_b = bidir ? ((ids >> 8) & 0xFF) : 0;
    const int mvx = best.best_x - x, mvy = best.best_y - y;
    const int mvx_b = bidir ? other.best_x - x : 0, mvy_b = bidir ? other.best_y - y : 0;
    w[0] = (uint32_t)btype | ((uint32_t)mode << 8) | ((uint32_t)tgt << 16) | ((uint32_t)tgt_b << 24);
    w[1] = mv_word(mvx, mvy);
    w[2] = (uint32_t)(uint16_t)mvx_b | ((uint32_t)(uint16_t)mvy_b << 16);
    w[3] = (uint32_t)best.sp_en | ((uint32_t)best.sp_amt << 8) | ((uint32_t)best.sp_idx << 16) |
           ((uint32_t)(bidir ? other.sp_en : 0) << 24);
    w[4] = bidir ? ((uint32_t)other.sp_amt | ((uint32_t)other.sp_idx << 8)) : 0u;
    w[5] = 0;
    w[6] = (uint32_t)sel.best_sad;
    w[7] = (uint32_t)best.mad;
    w[8] = fbits(best.cost);
    w[9] = bidir ? (uint32_t)other.sad : 0u;
    w[10] = bidir ? (uint32_t)other.mad : 0u;
    w[11] = bidir ? fbits(other.cost) : 0u;
    w[12] = (uint32_t)sel.n_int;
    w[13] = (uint32_t)sel.n_ax | ((uint32_t)sel.n_dg << 16);
    w[14] = 0;
    w[15] = 0;
    uint32_t *rec = reinterpret_cast<uint32_t *>(job.records + idx);
    write_words(rec, w, 16);

    // Publish what the MV predictor of later MBs reads (block type and MV)
    // as one device-scope word: they poll it directly.
    if (publish && lane == 0)
        st_relaxed64(job.mvw + idx, pred_word((btype & EVXA_BLOCK_MOTION) != 0, w[1]));

    int16_t *co = job.coeffs + (size_t)idx * EVXA_MB_SAMPLES;
    int16_t *lv = job.levels + (size_t)idx * EVXA_MB_SAMPLES;
    if (btype & EVXA_BLOCK_COPY) {
        // encode.cpp:742-744: copy blocks are not transformed.
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) { co[i] = 0; lv[i] = 0; }
        wave_sync();
        return;
    }
    if (job.type == EVXA_FRAME_P) predict_mb(job.refs[best_k], best, sh.pred, sh.tmp);
    wave_sync();
    int qp, var2;
    dct_quant(sh.src, sh.pred, sh.res, sh.tmp, job.quality, co, lv, qp, var2);
    uint32_t w2[2];
    w2[0] = w[4] | ((uint32_t)qp << 16);
    w2[1] = (uint32_t)(uint16_t)(int16_t)var2;
    write_words(rec + 4, w2, 2);
    w2[0] = 1u | (bidir ? 2u : 0u);
    write_words(rec + 14, w2, 1);
    wave_sync();
}

// Selection among a MB's searches, record publication, residual prediction,
// DCT and quantisation; run by the wave that searched the MB's last
// reference, which passes its own result in `own`.  The results of the
// other references are polled from job.searches (they were claimed earlier,
// so their waves are running) and cleared for the next launch.
__device__ void finish_mb(const DevJob &job, int bx, int by, int32_t *counters, MbLds &sh,
                          const Blk &src, const Geo &geo, const SearchOut &own,
                          SubpelLds *search)
{
    const int x = bx * 16, y = by * 16;
    const int idx = by * job.wb + bx;
    const uint64_t sw = gather_searches(job, idx, counters, job.eager != 0);
    auto result = [&](int k) { return k == job.n_refs - 1 ? own : unpack_search(sw, k); };
    store_src(src, geo, sh.src);
    const Sel sel = select_mb(job, x, y, sh, result, counters, search,
                              job.fpred ? job.fpred + (size_t)idx * EVXA_MB_SAMPLES : nullptr);
    emit_mb(job, bx, by, sh, sel, true);
}

// The persistent analysis kernel: NW waves per workgroup search one
// (MB, reference) ticket together; wave 0 then finishes it.  Register
// budgets: the one-wave kernel may take 256 VGPRs (2 waves per SIMD); the
// two-wave kernel fits 168 without spills, so 3 waves share each SIMD.
template <int NW>
__global__ __launch_bounds__(64 * NW, 3) void analyze_kernel(const DevJob *__restrict__ jobs,
                                                          const uint32_t *__restrict__ work,
                                                          int32_t *counters)
{
    __shared__ Shared sh;
    __shared__ int32_t s_word;
    const int lane = lane_id();
    // This workgroup's list (see kLists): tickets [lo, hi) of the work buffer.
    const int list = (int)(blockIdx.x % work[0]);
    const int lo = (int)work[1 + list], hi = (int)work[2 + list];
    const uint32_t *tickets = work + kWorkHeader;
    int32_t *next = counters + kCounterList + list * kCounterListStride;
    for (;;) {
        if (threadIdx.x == 0) s_word = lo + atomicAdd(next, 1);
        __syncthreads();
        // LDS hands the value back in a vector register; broadcasting it makes
        // the work item, and all control derived from it, wave-uniform
        // (scalar registers and scalar branches).
        const int t = __builtin_amdgcn_readfirstlane(s_word);
        __syncthreads();
        if (t >= hi) break;
        const uint32_t e = tickets[t];
        const DevJob &job = jobs[e >> kJobShift];
        const int ref = (int)((e >> kRefShift) & 3);
        const int mb = (int)(e & kMbMask);
        const int bx = mb % job.wb, by = mb / job.wb;
        const int x = bx * 16, y = by * 16;

        // Load the source block and the initial candidate, and stage the
        // largest search window around the MB, while the neighbours finish.
        const DevPic &rp = job.refs[ref];
        // The metric path first: its two loads then overlap the staging.
        const bool narrow = __builtin_amdgcn_readfirstlane(narrow_pair(job.src.wide, rp.wide)) != 0;
        const Geo geo = lane_geometry();
        Blk s, pre;
        load_cand(job.src, geo, x, y, s);
        load_cand(rp, geo, x, y, pre);
        const WinPos win = stage_window<NW>(rp, x, y, job.reach, sh.sp);
        // The temporal table's record of this MB (written by an earlier launch).
        uint32_t p0 = 0, p1 = 0;
        if (job.prev) {
            const uint32_t *prv = reinterpret_cast<const uint32_t *>(job.prev) + (size_t)mb * 16;
            p0 = prv[0];
            p1 = prv[1];
        }

        // Wait until the predictor words of the left, top, top-right and
        // top-left MBs are published (the row above up to its end at the
        // right edge), polling with a short back-off.  Lanes 0..3 load one
        // word each; absent neighbours need nothing.  Every wave of the
        // workgroup reads the same published words.
        const bool has_l = bx > 0, has_t = by > 0, has_tr = has_t && bx + 1 < job.wb;
        const bool has_tl = has_t && has_l;
        const int nidx = lane == 0 ? (has_l ? mb - 1 : -1)
                       : lane == 1 ? (has_t ? mb - job.wb : -1)
                       : lane == 2 ? (has_tr ? mb - job.wb + 1 : -1)
                       : lane == 3 ? (has_tl ? mb - job.wb - 1 : -1) : -1;
        uint64_t word = 0;
        int spins = 0, nap = 1;
        bool timed_out = false;
        for (;;) {
            if (nidx >= 0) word = ld_relaxed64(job.mvw + nidx);
            const uint64_t missing = __ballot(nidx >= 0 && (word >> 63) == 0);
            if (missing == 0) break;
            if (++spins > kSpinLimit) { timed_out = true; break; }
            for (int k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(1);
            nap = min(nap * 2, kMaxNap);
        }
        if (timed_out && lane == 0) atomicOr(&counters[kCounterError], kErrWaitTimeout);
        auto lane_word = [&](int l) {
            return ((uint64_t)(uint32_t)lane_value((int)(word >> 32), l) << 32) |
                   (uint32_t)lane_value((int)(uint32_t)word, l);
        };
        int pmx, pmy;
        mv_predictor(job, bx, by, lane_word(0), lane_word(1), lane_word(2), lane_word(3), p0, p1,
                     pmx, pmy);
        pmx = __builtin_amdgcn_readfirstlane(pmx);
        pmy = __builtin_amdgcn_readfirstlane(pmy);
        window_ready();
        uint64_t *early = job.type == EVXA_FRAME_P && job.n_refs == 1 ? job.mvw + mb : nullptr;
        const SearchOut o = inter_search<NW>(rp, job.width, job.height, narrow, s, geo, pre, win,
                                             x, y, job.thresh, job.lambda, pmx, pmy, sh.sp, early);
        // Wave 0 finishes the ticket in LDS the search used; the other waves
        // go on to the next claim.
        if (NW > 1) __syncthreads();
        if (NW > 1 && wave_id() != 0) continue;
        uint64_t *own_words =
            reinterpret_cast<uint64_t *>(job.searches + (size_t)mb * EVXA_MAX_REFS + ref);
        if (ref < job.n_refs - 1) {
            // Hand the result to the wave of the MB's last reference (a later
            // ticket of the same list, so never one this wave waits on); in
            // B frames its prediction first, from this search's planes (the
            // MB arrays overlap only the window), complete before the result.
            if (job.fpred) {
                predict_inter_result(rp, o, sh.sp, sh.mb.pred);
                const int32_t *src = reinterpret_cast<const int32_t *>(sh.mb.pred);
                int32_t *dst = reinterpret_cast<int32_t *>(job.fpred + (size_t)mb * EVXA_MB_SAMPLES);
#pragma unroll
                for (int k = 0; k < EVXA_MB_SAMPLES / 128; k++) st_relaxed32(dst + lane + 64 * k, src[lane + 64 * k]);
                stores_done();
            }
            publish_search(own_words, o);
            if (job.eager) try_publish(job, mb, x, y);
            continue;
        }
        if (job.eager) {
            publish_search(own_words, o);
            try_publish(job, mb, x, y);
        }
        // A B MB's wave is the backward search's: its planes are still in LDS.
        finish_mb(job, bx, by, counters, sh.mb, s, geo, o,
                  job.type == EVXA_FRAME_B ? &sh.sp : nullptr);
    }
}

// ------------------------------------------------------------ sweep mode --
//
// perform_inter_search depends on the frame's block table only through
// compute_mv_predictor (motion.cpp:613-635), and a MB's predictor reads MBs
// before it in raster order: left, top, top-right or top-left
// (motion.cpp:271-346).  So the frame's records are the unique fixpoint of
// "every MB's record is its searches under the predictor of its neighbours'
// records": by induction along raster order, any table with that property
// equals the sequential result.  Sweep mode reaches the fixpoint by Jacobi
// sweeps, with no MB waiting on another:
//  * sweep 0 searches every (MB, reference) pair, with a guess table (the
//    temporal table, or none) standing in for the neighbours' records;
//  * sweep s >= 1 first checks the MBs listed for it, one lane each: it
//    recomputes the predictor from every MB's word (block type and MV) as of
//    sweep s - 1, and keeps the MB's decision if the predictor as the search
//    uses it -- absent, or the clamped candidate position -- is unchanged;
//    then it searches the others again, one (MB, reference) ticket per
//    workgroup;
//  * a MB whose word changes lists itself and the four MBs that read it for
//    the next sweep, and the iteration ends with a sweep that changes none.
// A MB at dependency depth d is final after sweep d, so at most the frame's
// anti-diagonal count plus one sweeps run.  A final pass (final_kernel, a
// launch of its own at high occupancy) then forms each MB's prediction, DCT
// and quantisation once, from its last decision.
//
// Words by sweep parity: sweep s reads W[(s - 1) & 1] and writes W[s & 1];
// after sweep s, W[s & 1] holds every MB's word as of sweep s (a listed MB
// that keeps its predictor copies its word over).  Nothing is read and
// written in the same sweep, so every reference of a MB computes the same
// predictor, and each sweep's lists and counts are deterministic.
//
// Phases follow each other without a grid barrier: a wave moves on once
// every item of the current phase has finished (a flag set by the wave that
// finishes the last one), and only running waves hold items, so the launch
// needs no co-residency.  The dynamic lists and their counters are split in
// kSubLists sub-lists (a workgroup fills the one of its index and takes from
// any), because device-scope atomics on one address serialise.  Cross-wave
// data (words, keys, predictors, decisions, list entries, counters) move as
// device-scope accesses, and every item's stores complete before it counts
// itself finished.

constexpr uint64_t kSweepWaitTicks = 100000000ull;   // 1 s of s_memrealtime (100 MHz)
constexpr int kSweepMaxNap = 4;                      // longest back-off, s_sleep(8) units
// Workgroups past the first kTailActive poll a phase's flag kSweepFarNap
// times less often, and sit out a phase with fewer items than their index:
// thousands of waves polling one address, or bumping one counter, queue in
// front of the few that still work.  One still sitting out after kJoinTicks
// (200 us) claims like the others, so every item is taken whichever
// workgroups run.
constexpr int kSweepFarNap = 16;
constexpr int kTailActive = 256;
constexpr uint64_t kJoinTicks = 20000;


__device__ __forceinline__ int32_t fetch_add(int32_t *p, int32_t v)
{
    return __hip_atomic_fetch_add(EVXA_GLOBAL(int32_t, p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// One lane waits until *flag is set, with back-off up to max_nap units;
// false after `limit` ticks.
__device__ bool wait_flag(const int32_t *flag, int max_nap = kSweepMaxNap,
                          uint64_t limit = kSweepWaitTicks)
{
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    int nap = 1;
    while (ld_relaxed32(flag) == 0) {
        if (__builtin_amdgcn_s_memrealtime() - t0 > limit) return false;
        for (int k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(8);
        nap = min(nap * 2, max_nap);
    }
    return true;
}

// The back-off of this workgroup's waits between phases.
__device__ __forceinline__ int phase_nap()
{
    return (int)blockIdx.x < kTailActive ? kSweepMaxNap : kSweepFarNap;
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t w, int l)
{
    return ((uint64_t)(uint32_t)lane_value((int)(w >> 32), l) << 32) |
           (uint32_t)lane_value((int)(uint32_t)w, l);
}

// Entry of a dynamic list: bit 63 set, the sweep in bits 32..46, the ticket
// (job, reference, MB) below.  The claiming wave clears it, so the lists are
// zero between uses.
__device__ __forceinline__ uint64_t list_entry(int sweep, uint32_t ticket)
{
    return (1ull << 63) | ((uint64_t)(sweep & 0x7FFF) << 32) | ticket;
}

// Reads and clears the entry of sweep s at *slot (written by a wave that
// counted its item finished before this phase began; the poll only guards
// the ordering of the two device-scope stores).  Returns false if none came.
__device__ __forceinline__ bool take_entry(uint64_t *slot, int s, uint32_t &ticket)
{
    const uint64_t want = list_entry(s, 0);
    uint64_t e = ld_relaxed64(slot);
    int spins = 0;
    while ((e & 0xFFFFFFFF00000000ull) != want && ++spins < kSpinLimit) {
        __builtin_amdgcn_s_sleep(1);
        e = ld_relaxed64(slot);
    }
    if ((e & 0xFFFFFFFF00000000ull) != want) return false;
    st_relaxed64(slot, 0ull);
    ticket = (uint32_t)e;
    return true;
}

__device__ __forceinline__ int32_t *sub_ctr(const SweepArgs &sa, int s, int k)
{
    return sa.sctr + (size_t)s * kSweepStride + k * kSubStride;
}

__device__ __forceinline__ uint64_t *check_slot(const SweepArgs &sa, int s, int k, int i)
{
    return sa.check + ((size_t)(s & 1) * kSubLists + k) * sa.cap_check + i;
}

__device__ __forceinline__ uint64_t *search_slot(const SweepArgs &sa, int s, int k, int i)
{
    return sa.search + ((size_t)(s & 1) * kSubLists + k) * sa.cap_search + i;
}

enum { kClaimSweep0 = 0, kClaimCheck = 1, kClaimSearch = 2, kClaimExit = 4 };

// Where a workgroup's claims stand (in LDS; only thread 0 writes it).
struct ClaimState {
    int phase;                  // 0: sweep 0, 1: check, 2: search, 3: done
    int s;                      // current sweep
    int sit_out;                // this workgroup claims nothing in the phase (unless it joins in)
    int total_chunks;           // check phase: chunks over all (sub-)lists
    int chunks[kSubLists];      // check phase: chunks per (sub-)list
    int tickets[kSubLists];     // search phase: tickets per sub-list
    int done0[kLists];          // sweep 0: this workgroup's finished tickets per list, not yet counted
    int b_kind, b_next, b_end, b_list;   // the claimed batch: kind, next and end item, list
    uint32_t b_tk[kMaxClaimBatch];       // its tickets, in the order they are run
};

// Thread 0: the check phase of sweep s, or the end of the sweeps if no MB
// is listed for it (the MB lists are final: sweep s - 1 has finished).
// Sweep 1 checks every MB, taking the final pass's static lists in chunks.
__device__ void enter_check(ClaimState &c, const uint32_t *work, const SweepArgs &sa, int s)
{
    int any = 0;
    c.total_chunks = 0;
    for (int k = 0; k < kSubLists; k++) {
        int n;
        if (s == 1)
            n = k < (int)work[0] ? (int)work[kWorkFinal + k + 1] - (int)work[kWorkFinal + k] : 0;
        else
            n = s < sa.max_sweeps ? min(ld_relaxed32(sub_ctr(sa, s, k) + kSubCheckTail), sa.cap_check) : 0;
        c.chunks[k] = (n + kCheckChunk - 1) / kCheckChunk;
        c.total_chunks += c.chunks[k];
        any += n;
    }
    c.s = s;
    c.phase = any > 0 ? 1 : 3;
    c.sit_out = (int)blockIdx.x >= max(kTailActive, c.total_chunks);
}

// Thread 0: the search phase of sweep c.s, or the end of the sweeps if its
// check phase changed no predictor (then no word changed either).
__device__ void enter_search(ClaimState &c, const SweepArgs &sa, int kb)
{
    int any = 0;
    for (int k = 0; k < kSubLists; k++) {
        c.tickets[k] = min(ld_relaxed32(sub_ctr(sa, c.s, k) + kSubSearchTail), sa.cap_search);
        any += c.tickets[k];
    }
    c.phase = any > 0 ? 2 : 3;
    c.sit_out = (int)blockIdx.x >= max(kTailActive, (any + kb - 1) / kb);
}

// Thread 0, once the static lists of sweep 0 are all taken: count the
// workgroup's finished tickets (kept per list in LDS, so that no ticket pays
// an atomic round trip) after its stores completed; the count that finishes
// the last list sets the sweep's flag.  (Wave 0 makes every global store of
// a ticket.)
__device__ void flush_sweep0(ClaimState &c, const uint32_t *work, int32_t *counters, const SweepArgs &sa)
{
    stores_done();
    for (int l = 0; l < (int)work[0]; l++) {
        const int n = c.done0[l];
        if (n == 0) continue;
        c.done0[l] = 0;
        const int size = (int)work[2 + l] - (int)work[1 + l];
        if (fetch_add(counters + kCounterList + l * kCounterListStride + kCounterDone, n) + n == size &&
            fetch_add(sa.sctr + kSwSubDone, 1) == (int)work[kWorkNonEmpty] - 1) {
            st_relaxed64(reinterpret_cast<uint64_t *>(sa.sctr + kSwTime), __builtin_amdgcn_s_memrealtime());
            st_relaxed32(sa.sctr + kSwFlag, 1);
        }
    }
}

// Thread 0: stage a claimed batch of n items from index t of its list (an
// item is a ticket, or for the check phase a chunk).  Tickets run in two
// passes: first those that never wait on another MB's decision (every
// reference but the last), then the designated ones (the last reference),
// which wait only on tickets of the first kind.  A designated ticket at the
// head of a batch would otherwise hold back the batch's tail, which the
// designated ticket of another wave's batch may wait on, and so on along a
// chain of waves.  Returns false if a dynamic entry never came.
__device__ bool stage_batch(ClaimState &c, const DevJob *jobs, const uint32_t *work, const SweepArgs &sa,
                            int kind, int t, int n, int list)
{
    c.b_kind = kind;
    c.b_list = list;
    c.b_next = 0;
    c.b_end = n;
    if (kind == kClaimCheck) {
        for (int i = 0; i < n; i++) c.b_tk[i] = (uint32_t)(t + i);
        return true;
    }
    uint32_t tk[kMaxClaimBatch];
    bool ok = true;
    for (int i = 0; i < n; i++) {
        if (kind == kClaimSearch) {
            if (!take_entry(search_slot(sa, c.s, list, t + i), c.s, tk[i])) {
                ok = false;
                n = i;
                break;
            }
        } else {
            tk[i] = work[kWorkHeader + t + i];
        }
    }
    int m = 0;
    for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i < n; i++) {
            const int ref = (int)((tk[i] >> kRefShift) & 3);
            const bool designated = ref == jobs[tk[i] >> kJobShift].n_refs - 1;
            if (designated == (pass == 1)) c.b_tk[m++] = tk[i];
        }
    c.b_end = m;
    return ok;
}

// Thread 0: the next item of the staged batch, if any, out = {kind, ticket
// or chunk, sweep, list}.
__device__ bool batch_item(ClaimState &c, uint32_t out[4])
{
    if (c.b_next >= c.b_end) return false;
    out[0] = (uint32_t)c.b_kind;
    out[1] = c.b_tk[c.b_next++];
    out[2] = c.b_kind == kClaimSearch || c.b_kind == kClaimCheck ? (uint32_t)c.s : 0u;
    out[3] = (uint32_t)c.b_list;
    return true;
}

// Thread 0: the workgroup's next item, out = {kind, ticket or chunk, sweep,
// list}.  Items are claimed in batches of sa.batch consecutive ones (a
// device-scope atomic on one address serialises, so a claim per item would
// bound the rate); a wave works through its batch in order, and since
// dependencies point to earlier items, the wave holding the earliest
// unfinished item never waits on another.  Sweep 0's static lists and the
// sub-lists are taken own first, then the others'; between phases it waits
// for the finished flag of the last.
__device__ __noinline__ void claim(ClaimState &c, const DevJob *jobs, const uint32_t *work, int32_t *counters,
                      const SweepArgs &sa, int home, uint32_t out[4])
{
    if (batch_item(c, out)) return;
    const int n_lists = (int)work[0];
    const int kb = min(max(sa.batch, 1), kMaxClaimBatch);
    for (;;) {
        // The final pass is a launch of its own (final_kernel).
        if (c.phase == 3) break;
        if (c.phase == 0) {
            const int h = home % n_lists;
            for (int k = 0; k < n_lists; k++) {
                const int l = h + k < n_lists ? h + k : h + k - n_lists;
                const int lo = (int)work[1 + l], hi = (int)work[2 + l];
                int32_t *next = counters + kCounterList + l * kCounterListStride;
                // Another list's counter is read before it is bumped (a
                // stealing wave should not add traffic to a busy counter);
                // the own list's is bumped directly.
                if (k > 0 && lo + ld_relaxed32(next) >= hi) continue;
                const int t = lo + fetch_add(next, kb);
                if (t < hi) {
                    if (l == 0 && t == lo)
                        st_relaxed64(reinterpret_cast<uint64_t *>(sa.sctr + kSwStart),
                                     __builtin_amdgcn_s_memrealtime());
                    stage_batch(c, jobs, work, sa, kClaimSweep0, t, min(t + kb, hi) - t, l);
                    batch_item(c, out);
                    return;
                }
            }
            flush_sweep0(c, work, counters, sa);
            if (!wait_flag(sa.sctr + kSwFlag, phase_nap())) break;
            enter_check(c, work, sa, 1);
            continue;
        }
        if (c.phase == 1) {
            const int32_t *flag = sa.sctr + (size_t)c.s * kSweepStride + kSwFlagC;
            if (c.sit_out) {
                c.sit_out = 0;
                if (wait_flag(flag, kSweepFarNap, kJoinTicks)) enter_search(c, sa, kb);
                continue;
            }
            for (int i = 0; i < kSubLists; i++) {
                const int k = (home + i) & (kSubLists - 1);
                if (c.chunks[k] == 0) continue;
                int32_t *next = c.s == 1 ? counters + kCounterList + k * kCounterListStride + kCounterCheck1
                                         : sub_ctr(sa, c.s, k) + kSubChunkHead;
                if (ld_relaxed32(next) >= c.chunks[k]) continue;
                const int j = fetch_add(next, 1);
                if (j < c.chunks[k]) {
                    stage_batch(c, jobs, work, sa, kClaimCheck, j, 1, k);
                    batch_item(c, out);
                    return;
                }
            }
            if (!wait_flag(flag, phase_nap())) break;
            enter_search(c, sa, kb);
            continue;
        }
        const int32_t *flag = sa.sctr + (size_t)c.s * kSweepStride + kSwFlag;
        if (c.sit_out) {
            c.sit_out = 0;
            if (wait_flag(flag, kSweepFarNap, kJoinTicks)) enter_check(c, work, sa, c.s + 1);
            continue;
        }
        for (int i = 0; i < kSubLists; i++) {
            const int k = (home + i) & (kSubLists - 1);
            if (c.tickets[k] == 0) continue;
            int32_t *next = sub_ctr(sa, c.s, k) + kSubSearchHead;
            if (i > 0 && ld_relaxed32(next) >= c.tickets[k]) continue;
            const int t = fetch_add(next, kb);
            if (t < c.tickets[k]) {
                if (!stage_batch(c, jobs, work, sa, kClaimSearch, t, min(t + kb, c.tickets[k]) - t, k))
                    atomicOr(&counters[kCounterError], kErrWaitTimeout);
                if (batch_item(c, out)) return;
                break;
            }
        }
        if (!wait_flag(flag, phase_nap())) break;
        enter_check(c, work, sa, c.s + 1);
    }
    if (c.phase != 3) atomicOr(&counters[kCounterError], kErrWaitTimeout);
    out[0] = kClaimExit;
}

// The word a MB of the guess table stands in with (no guess: no motion).
__device__ __forceinline__ uint64_t guess_word(const DevJob &job, int idx)
{
    if (!job.guess) return pred_word(false, 0);
    const uint32_t *g = reinterpret_cast<const uint32_t *>(job.guess + idx);
    const uint32_t g0 = g[0], g1 = g[1];
    return (g0 & EVXA_BLOCK_MOTION) ? pred_word(true, g1) : pred_word(false, 0);
}

// The words of the left, top, top-right and top-left MBs of (bx, by) and
// its own, as of sweep s - 1 >= 0, and its predictor from them and the
// temporal table (compute_mv_predictor, motion.cpp:271-346).  Per lane.
__device__ void sweep_predictor(const DevJob &job, int mb, int s, int &pmx, int &pmy,
                                uint64_t &own_old)
{
    const int bx = mb % job.wb, by = mb / job.wb;
    const size_t n_mb = (size_t)job.wb * job.hb;
    const bool has_l = bx > 0, has_t = by > 0, has_tr = has_t && bx + 1 < job.wb;
    const bool has_tl = has_t && has_l;
    const int idx[5] = {has_l ? mb - 1 : -1, has_t ? mb - job.wb : -1,
                        has_tr ? mb - job.wb + 1 : -1, has_tl ? mb - job.wb - 1 : -1, mb};
    uint64_t w[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        w[i] = pred_word(false, 0);
        if (idx[i] >= 0) w[i] = ld_relaxed64(job.sw_words + (size_t)((s - 1) & 1) * n_mb + idx[i]);
    }
    uint32_t p0 = 0, p1 = 0;
    if (job.prev) {
        const uint32_t *prv = reinterpret_cast<const uint32_t *>(job.prev) + (size_t)mb * 16;
        p0 = prv[0];
        p1 = prv[1];
    }
    mv_predictor(job, bx, by, w[0], w[1], w[2], w[3], p0, p1, pmx, pmy);
    own_old = w[4];
}

// The key of a predictor as the search uses it: absent, or the clamped
// candidate position (motion.cpp:613-635).
__device__ __forceinline__ uint32_t pred_key(const DevJob &job, int mb, int pmx, int pmy)
{
    if (pmx == 0 && pmy == 0) return 0xFFFFFFFFu;
    const int x = (mb % job.wb) * 16, y = (mb / job.wb) * 16;
    return (uint32_t)min(max(x + pmx, 0), job.width - 16) |
           ((uint32_t)min(max(y + pmy, 0), job.height - 16) << 16);
}

// A MB's word changed in sweep s: list it (for its carry-over) and the MBs
// whose predictor reads it -- right, bottom-left, bottom, bottom-right --
// for sweep s + 1, each once (the listed stamp), in sub-list `home`.
__device__ void push_dependants(const DevJob &job, int jidx, int mb, int s, const SweepArgs &sa,
                                int32_t *counters, int home)
{
    const int lane = lane_id();
    if (s + 1 >= sa.max_sweeps) {
        if (lane == 0) atomicOr(&counters[kCounterError], kErrSweepLimit);
        return;
    }
    const int bx = mb % job.wb, by = mb / job.wb;
    int m2 = -1;
    if (lane == 0) m2 = mb;
    else if (lane == 1 && bx + 1 < job.wb) m2 = mb + 1;
    else if (lane == 2 && bx > 0 && by + 1 < job.hb) m2 = mb + job.wb - 1;
    else if (lane == 3 && by + 1 < job.hb) m2 = mb + job.wb;
    else if (lane == 4 && bx + 1 < job.wb && by + 1 < job.hb) m2 = mb + job.wb + 1;
    bool need = false;
    if (m2 >= 0)
        need = __hip_atomic_fetch_max(reinterpret_cast<int32_t *>(job.sw_listed + m2), s + 1,
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < s + 1;
    const uint64_t mask = __ballot(need);
    const int cnt = __builtin_popcountll(mask);
    if (cnt == 0) return;
    int base = 0;
    if (lane == 0) base = fetch_add(sub_ctr(sa, s + 1, home) + kSubCheckTail, cnt);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base + cnt > sa.cap_check) {
        if (lane == 0) atomicOr(&counters[kCounterError], kErrSweepLimit);
        return;
    }
    if (need) {
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        st_relaxed64(check_slot(sa, s + 1, home, base + rank),
                     list_entry(s + 1, ((uint32_t)jidx << kJobShift) | (uint32_t)m2));
    }
}

// The MB's decision after its searches of sweep s (the designated wave):
// selection, decision record, word for the next sweep and, from sweep 1 on,
// its dependants listed if the word changed.  Sweep 0 also records the
// predictor's key.
__device__ void decide_mb(const DevJob &job, int jidx, int mb, int s, const SweepArgs &sa,
                          MbLds &sh, const Blk &src, const Geo &geo, const SearchOut &own,
                          uint64_t own_old, uint32_t key, int32_t *counters, int home,
                          SubpelLds *search)
{
    const int lane = lane_id();
    const int x = (mb % job.wb) * 16, y = (mb / job.wb) * 16;
    const size_t n_mb = (size_t)job.wb * job.hb;
    const uint64_t sw = gather_searches(job, mb, counters, false);
    auto result = [&](int k) { return k == job.n_refs - 1 ? own : unpack_search(sw, k); };
    store_src(src, geo, sh.src);
    const Sel sel = select_mb(job, x, y, sh, result, counters, search);
    const uint64_t w = sel_word(sel, x, y);
    // The decision (P: the selected search; B: both searches).
    const bool is_b = job.type != EVXA_FRAME_P;
    const SearchOut first = is_b ? result(0) : sel.best;
    const SearchOut second = is_b ? result(1) : sel.best;
    uint64_t dw;
    if (lane < 3) dw = pack_search(first, lane);
    else if (lane < 6) dw = is_b ? pack_search(second, lane - 3) : 0ull;
    else if (lane == 6)
        dw = (uint32_t)sel.best_sad | ((uint64_t)(sel.best_k & 3) << 32) | ((uint64_t)sel.bidir << 34);
    else
        dw = (uint32_t)sel.n_int | ((uint64_t)(uint16_t)sel.n_ax << 32) |
             ((uint64_t)(uint16_t)sel.n_dg << 48);
    if (lane < 8) st_relaxed64(&job.sw_dec[mb].w[lane], dw);
    if (lane == 0) {
        st_relaxed64(job.sw_words + (size_t)(s & 1) * n_mb + mb, w);
        if (s == 0) st_relaxed32(reinterpret_cast<int32_t *>(job.sw_key + mb), (int32_t)key);
    }
    // Sweep 1 checks every MB; later sweeps the MBs listed.
    if (s > 0 && w != own_old) push_dependants(job, jidx, mb, s, sa, counters, home);
}

// One check chunk of sweep s (one wave): up to kCheckChunk MBs listed in
// sub-list k, one lane each.  A MB whose predictor key is unchanged keeps
// its decision and carries its word over; the others get a new key and
// predictor and their search tickets (all references, the last one
// designated to decide) in sub-list `home`.  The wave that finishes the
// phase's last chunk clears sweep s + 2's counters and sets the phase flag.
__device__ void check_chunk(const DevJob *jobs, const uint32_t *work, int s, int k, int j,
                            int total_chunks, const SweepArgs &sa, int32_t *counters, int home)
{
    const int lane = lane_id();
    const int i = j * kCheckChunk + lane;
    uint32_t tk = 0;
    bool have = false;
    if (s == 1) {
        // Every MB: list k of the final pass.
        const int lo = (int)work[kWorkFinal + k], hi = (int)work[kWorkFinal + k + 1];
        if (lo + i < hi) {
            tk = work[kWorkHeader + lo + i];
            have = true;
        }
    } else {
        const int n = min(ld_relaxed32(sub_ctr(sa, s, k) + kSubCheckTail), sa.cap_check);
        if (i < n) {
            have = take_entry(check_slot(sa, s, k, i), s, tk);
            if (!have) atomicOr(&counters[kCounterError], kErrWaitTimeout);
        }
    }
    int nr = 0, mb = 0, jidx = 0;
    if (have) {
        jidx = (int)(tk >> kJobShift);
        mb = (int)(tk & kMbMask);
        const DevJob &job = jobs[jidx];
        int pmx, pmy;
        uint64_t own_old;
        sweep_predictor(job, mb, s, pmx, pmy, own_old);
        const uint32_t key = pred_key(job, mb, pmx, pmy);
        int32_t *key_p = reinterpret_cast<int32_t *>(job.sw_key + mb);
        if ((uint32_t)ld_relaxed32(key_p) == key) {
            st_relaxed64(job.sw_words + (size_t)(s & 1) * job.wb * job.hb + mb, own_old);
        } else {
            st_relaxed32(key_p, (int32_t)key);
            st_relaxed64(job.sw_pred + mb, (1ull << 63) | (uint32_t)(uint16_t)pmx |
                                               ((uint32_t)(uint16_t)pmy << 16));
            nr = job.n_refs;
        }
    }
    // Search tickets: an exclusive scan of the lanes' counts.
    int incl = nr;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    const int total = __shfl(incl, 63);
    const int n_mbs = __builtin_popcountll(__ballot(nr > 0));
    if (total > 0) {
        int base = 0;
        if (lane == 0) {
            base = fetch_add(sub_ctr(sa, s, home) + kSubSearchTail, total);
            fetch_add(sub_ctr(sa, s, home) + kSubDesignated, n_mbs);
            fetch_add(sa.sctr + (size_t)s * kSweepStride + kSwSearched, n_mbs);
        }
        base = __builtin_amdgcn_readfirstlane(base);
        if (base + total > sa.cap_search) {
            if (lane == 0) atomicOr(&counters[kCounterError], kErrSweepLimit);
        } else {
            for (int r = 0; r < nr; r++)
                st_relaxed64(search_slot(sa, s, home, base + incl - nr + r),
                             list_entry(s, ((uint32_t)jidx << kJobShift) | ((uint32_t)r << kRefShift) |
                                               (uint32_t)mb));
        }
    }
    stores_done();
    int last = 0;
    if (lane == 0)
        last = fetch_add(sa.sctr + (size_t)s * kSweepStride + kSwDoneC, 1) == total_chunks - 1;
    if (!__builtin_amdgcn_readfirstlane(last)) return;
    // Sweep s + 2's counters are untouched until sweep s + 1's searches.
    if (s + 2 < sa.max_sweeps) {
        int32_t *b2 = sa.sctr + (size_t)(s + 2) * kSweepStride;
        const int sub = lane / 6, f = lane % 6;
        const int off = f < 4 ? f : (f == 4 ? kSubSearchHead : kSubDone);
        if (lane < 6 * kSubLists) st_relaxed32(b2 + sub * kSubStride + off, 0);
        else if (lane == 48) st_relaxed32(b2 + kSwFlagC, 0);
        else if (lane == 49) st_relaxed32(b2 + kSwDoneC, 0);
        else if (lane == 50) st_relaxed32(b2 + kSwSubDone, 0);
        else if (lane == 51) st_relaxed32(b2 + kSwFlag, 0);
        else if (lane == 52) st_relaxed32(b2 + kSwSearched, 0);
        stores_done();
    }
    if (lane == 0) st_relaxed32(sa.sctr + (size_t)s * kSweepStride + kSwFlagC, 1);
}

// One (MB, reference) ticket of sweep s, from (sub-)list k.  In sweep 0 the
// wave of reference 0 forms the MB's predictor from the words of the left,
// top, top-right and top-left MBs -- words already published in sweep 0
// (fresh reads, waiting up to sa.wait_ticks for missing ones), else the
// guess table's -- and publishes it for the MB's other references, so that
// all of them search under the same one; in later sweeps the MB's check has
// set it.  The memory stage (source, initial candidate, window) is issued
// first and overlaps it.  The designated wave (last reference) decides the
// MB; in sweeps s >= 1 it then counts the MB finished, and the last MB of
// the last sub-list ends the sweep.  One function for both, so the search
// is compiled once.
template <int NW>
__device__ void search_ticket(const DevJob &job, int jidx, int ref, int mb, int s, int k,
                              const SweepArgs &sa, Shared &sh, int32_t *counters, int home,
                              uint32_t *s_pred)
{
    const int lane = lane_id();
    const int bx = mb % job.wb, by = mb / job.wb;
    const int x = bx * 16, y = by * 16;
    const size_t n_mb = (size_t)job.wb * job.hb;
    const bool designated = ref == job.n_refs - 1;
    const DevPic &rp = job.refs[ref];
    const bool narrow = __builtin_amdgcn_readfirstlane(narrow_pair(job.src.wide, rp.wide)) != 0;
    const Geo geo = lane_geometry();
    Blk src, pre;
    load_cand(job.src, geo, x, y, src);
    load_cand(rp, geo, x, y, pre);
    const WinPos win = stage_window<NW>(rp, x, y, job.reach, sh.sp);
    int pmx, pmy;
    if (s == 0 && ref == 0) {
        const bool has_l = bx > 0, has_t = by > 0, has_tr = has_t && bx + 1 < job.wb;
        const bool has_tl = has_t && has_l;
        const int nidx = lane == 0 ? (has_l ? mb - 1 : -1)
                       : lane == 1 ? (has_t ? mb - job.wb : -1)
                       : lane == 2 ? (has_tr ? mb - job.wb + 1 : -1)
                       : lane == 3 ? (has_tl ? mb - job.wb - 1 : -1) : -1;
        uint64_t word = 0;
        if (sa.reads == kSweepReadsFresh) {
            if (nidx >= 0) word = ld_relaxed64(job.sw_words + nidx);
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            int nap = 1;
            while (__ballot(nidx >= 0 && (word >> 63) == 0) != 0 &&
                   __builtin_amdgcn_s_memrealtime() - t0 < (uint64_t)sa.wait_ticks) {
                for (int q = 0; q < nap; q++) __builtin_amdgcn_s_sleep(1);
                nap = min(nap * 2, kMaxNap);
                if (nidx >= 0 && (word >> 63) == 0) word = ld_relaxed64(job.sw_words + nidx);
            }
        }
        if (nidx >= 0 && (word >> 63) == 0) word = guess_word(job, nidx);
        uint32_t p0 = 0, p1 = 0;
        if (job.prev) {
            const uint32_t *prv = reinterpret_cast<const uint32_t *>(job.prev) + (size_t)mb * 16;
            p0 = prv[0];
            p1 = prv[1];
        }
        mv_predictor(job, bx, by, lane_u64(word, 0), lane_u64(word, 1), lane_u64(word, 2),
                     lane_u64(word, 3), p0, p1, pmx, pmy);
        pmx = __builtin_amdgcn_readfirstlane(pmx);
        pmy = __builtin_amdgcn_readfirstlane(pmy);
        if (job.n_refs > 1 && (NW == 1 || wave_id() == 0) && lane == 0)
            st_relaxed64(job.sw_pred + mb, (1ull << 63) | (uint32_t)(uint16_t)pmx |
                                               ((uint32_t)(uint16_t)pmy << 16));
        if (NW > 1) {
            // The waves of the workgroup may have seen different words
            // published: wave 0's predictor is the MB's, handed over in LDS
            // across window_ready's barrier.
            if (wave_id() == 0 && lane == 0)
                *s_pred = (uint32_t)(uint16_t)pmx | ((uint32_t)(uint16_t)pmy << 16);
        }
    } else {
        // Set before this sweep (check), or by reference 0's wave, whose
        // ticket was claimed earlier, so its wave is running.
        uint64_t pw = ld_relaxed64(job.sw_pred + mb);
        int spins = 0;
        while ((pw >> 63) == 0 && ++spins < kSpinLimit) {
            __builtin_amdgcn_s_sleep(1);
            pw = ld_relaxed64(job.sw_pred + mb);
        }
        if ((pw >> 63) == 0 && lane == 0) atomicOr(&counters[kCounterError], kErrWaitTimeout);
        const uint32_t p = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pw);
        pmx = (int16_t)(p & 0xFFFF);
        pmy = (int16_t)(p >> 16);
    }
    uint64_t own_old = 0;
    if (s > 0 && designated) own_old = ld_relaxed64(job.sw_words + (size_t)((s - 1) & 1) * n_mb + mb);
    window_ready();
    if (NW > 1 && s == 0 && ref == 0) {
        const uint32_t p = (uint32_t)__builtin_amdgcn_readfirstlane((int)*s_pred);
        pmx = (int16_t)(p & 0xFFFF);
        pmy = (int16_t)(p >> 16);
    }
    const SearchOut o = inter_search<NW>(rp, job.width, job.height, narrow, src, geo, pre, win, x, y,
                                         job.thresh, job.lambda, pmx, pmy, sh.sp, nullptr);
    if (NW > 1) __syncthreads();             // both waves are done with the window
    if (NW > 1 && wave_id() != 0) return;
    if (!designated) {
        publish_search(reinterpret_cast<uint64_t *>(job.searches + (size_t)mb * EVXA_MAX_REFS + ref), o);
        return;
    }
    own_old = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(own_old >> 32)) << 32) |
              (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)own_old);
    decide_mb(job, jidx, mb, s, sa, sh.mb, src, geo, o, own_old,
              s == 0 ? pred_key(job, mb, pmx, pmy) : 0u, counters, home,
              job.type == EVXA_FRAME_B ? &sh.sp : nullptr);   // the backward search's planes
    // Sweep 0 searches every MB: its count is the launch's MB count, and no
    // ticket pays an atomic on one shared address for it.
    if (s == 0) return;
    stores_done();
    if (lane == 0) {
        int32_t *c = sub_ctr(sa, s, k);
        if (fetch_add(c + kSubDone, 1) == ld_relaxed32(c + kSubDesignated) - 1) {
            int nonempty = 0;
            for (int q = 0; q < kSubLists; q++) nonempty += ld_relaxed32(sub_ctr(sa, s, q) + kSubDesignated) > 0;
            int32_t *blk = sa.sctr + (size_t)s * kSweepStride;
            if (fetch_add(blk + kSwSubDone, 1) == nonempty - 1) {
                st_relaxed64(reinterpret_cast<uint64_t *>(blk + kSwTime), __builtin_amdgcn_s_memrealtime());
                st_relaxed32(blk + kSwFlag, 1);
            }
        }
    }
}

// Final pass: a MB's prediction, DCT, quantisation and record from its last
// decision (one wave).
__device__ void final_mb(const DevJob &job, int mb, MbLds &sh)
{
    const int lane = lane_id();
    const int bx = mb % job.wb, by = mb / job.wb;
    const int x = bx * 16, y = by * 16;
    const uint64_t dw = lane < 8 ? ld_relaxed64(&job.sw_dec[mb].w[lane]) : 0ull;
    const Geo geo = lane_geometry();
    Blk src;
    load_cand(job.src, geo, x, y, src);
    store_src(src, geo, sh.src);
    const uint64_t w6 = lane_u64(dw, 6), w7 = lane_u64(dw, 7);
    Sel sel;
    sel.best_sad = (int)(uint32_t)w6;
    sel.best_k = (int)((w6 >> 32) & 3);
    sel.bidir = (int)((w6 >> 34) & 1);
    sel.n_int = (int)(uint32_t)w7;
    sel.n_ax = (int)((w7 >> 32) & 0xFFFF);
    sel.n_dg = (int)(w7 >> 48);
    sel.other = SearchOut{x, y, 0, 0, 0, 0.0f, 0, 0, 0, 0, 0, 0, 0};
    const SearchOut a = unpack_search(dw, 0);
    if (job.type == EVXA_FRAME_P) {
        sel.best = a;
    } else {
        const SearchOut b = unpack_search(dw, 1);
        wave_sync();
        (void)b_predictions(job, a, b, sh);
        sel.best = (sel.bidir || sel.best_k == 0) ? a : b;
        if (sel.bidir) sel.other = b;
        b_choose(sel, sh);
    }
    emit_mb(job, bx, by, sh, sel, false);
}

// The persistent sweep-mode kernel: sweep 0, then check and search phases
// sweep after sweep (final_kernel follows it).  NW waves per workgroup
// search one (MB, reference) ticket together; wave 0 runs the per-MB
// stages.
template <int NW>
__global__ __launch_bounds__(64 * NW, 3) void sweep_kernel(const DevJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ work,
                                                        int32_t *counters, SweepArgs sa)
{
    __shared__ Shared sh;
    __shared__ ClaimState cs;
    __shared__ uint32_t s_claim[4];
    __shared__ uint32_t s_pred;
    const int home = (int)(blockIdx.x % kSubLists);
    if (threadIdx.x == 0) {
        cs.phase = 0;
        cs.s = 0;
        cs.sit_out = 0;
        cs.b_next = cs.b_end = 0;
        for (int l = 0; l < kLists; l++) cs.done0[l] = 0;
    }
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t o[4];
            claim(cs, jobs, work, counters, sa, (int)blockIdx.x, o);
            s_claim[0] = o[0]; s_claim[1] = o[1]; s_claim[2] = o[2]; s_claim[3] = o[3];
        }
        __syncthreads();
        const int kind = __builtin_amdgcn_readfirstlane((int)s_claim[0]);
        const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_claim[1]);
        const int sweep = __builtin_amdgcn_readfirstlane((int)s_claim[2]);
        const int list = __builtin_amdgcn_readfirstlane((int)s_claim[3]);
        const int total_chunks = __builtin_amdgcn_readfirstlane(cs.total_chunks);
        __syncthreads();
        if (kind == kClaimExit) break;
        if (kind == kClaimCheck) {
            if (NW == 1 || wave_id() == 0)
                check_chunk(jobs, work, sweep, list, (int)e, total_chunks, sa, counters, home);
            continue;
        }
        const int jidx = (int)(e >> kJobShift);
        const DevJob &job = jobs[jidx];
        const int mb = (int)(e & kMbMask);
        const int ref = (int)((e >> kRefShift) & 3);
        search_ticket<NW>(job, jidx, ref, mb, kind == kClaimSearch ? sweep : 0, list, sa, sh,
                          counters, home, &s_pred);
        if (kind == kClaimSweep0 && threadIdx.x == 0) cs.done0[list]++;
    }
}

// Sweep mode's final pass, after sweep_kernel on the same queue: one wave
// per MB of the final lists (every MB of the launch), kFinalMbs waves per
// workgroup, no dependencies.  Skipped when the sweeps reported an error
// (their decisions are then incomplete).
constexpr int kFinalMbs = 4;

__global__ __launch_bounds__(64 * kFinalMbs, 8) void final_kernel(const DevJob *__restrict__ jobs,
                                                              const uint32_t *__restrict__ work,
                                                              const int32_t *counters)
{
    __shared__ MbLds sh[kFinalMbs];
    if (__hip_atomic_load(&counters[kCounterError], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return;
    const int w = wave_id();
    const int lo = (int)work[kWorkFinal], hi = (int)work[kWorkFinal + work[0]];
    for (int t = lo + (int)blockIdx.x * kFinalMbs + w; t < hi; t += (int)gridDim.x * kFinalMbs) {
        const uint32_t e = work[kWorkHeader + t];
        final_mb(jobs[e >> kJobShift], (int)(e & kMbMask), sh[w]);
    }
}

// ------------------------------------------------------ reconstruction --

// The 8-point cosine table (xftables.h:57-67), row k = frequency, column i =
// sample, as the inverse transform indexes it.
constexpr int kDct8Full[8][8] = {
    {128, 128, 128, 128, 128, 128, 128, 128},
    {126, 106, 71, 25, -25, -71, -106, -126},
    {118, 49, -49, -118, -118, -49, 49, 118},
    {106, -25, -126, -71, 71, 126, 25, -106},
    {91, -91, -91, 91, 91, -91, -91, 91},
    {71, -126, 25, 106, -106, -25, 126, -71},
    {49, -118, 118, -49, -49, 118, -118, 49},
    {25, -71, 106, -126, 126, -106, 71, -25}};

// transform.cpp:330-347 (and 396-417 before the addend): sample i of an
// 8-point line from its frequencies v[0..7], every term truncated on its
// own, then rounded_div(., 128) (evx_math.h:228-236).
__device__ __forceinline__ int idct_sample(const int (&v)[8], int i)
{
    int t = ((v[0] * kDct8Full[0][i]) * 45) / 128;
#pragma unroll
    for (int k = 1; k < 8; k++) t += (v[k] * kDct8Full[k][i]) / 2;
    return t < 0 ? (t - 64) / 128 : (t + 64) / 128;
}

constexpr int kReconMbs = 4;           // MBs (one wave each) per workgroup

struct ReconLds {
    int16_t pred[EVXA_MB_SAMPLES];
    int16_t pred_b[EVXA_MB_SAMPLES];
    int16_t co[EVXA_MB_SAMPLES];
    int16_t tmp[16 * 21 + 2 * 8 * 13];  // diagonal H pass, then the vertical IDCT pass
};

// Reconstruction of the inter MBs of an analysed frame into `dst`, as the
// reference decoder rebuilds them (decode.cpp:123-433): the prediction the
// record names (the reference whose target id it carries; bidirectional
// records average both, decode.cpp:126-197, 272-330), plus, for transformed
// blocks, the levels inverse-quantised (quantize.cpp:322-332, 427-449) and
// inverse-transformed onto it (macroblock.h:290-295, transform.cpp:419-433,
// 572-582), stored as int16.  One wave per MB; MBs are independent.
__global__ __launch_bounds__(64 * kReconMbs) void recon_kernel(const DevJob job,
                                                               const int16_t *__restrict levels,
                                                               DevPic dst, int32_t *wide)
{
    __shared__ ReconLds lds[kReconMbs];
    const int lane = lane_id();
    const int mb = blockIdx.x * kReconMbs + wave_id();
    if (mb >= job.wb * job.hb) return;   // the whole wave; no workgroup barriers follow
    ReconLds &sh = lds[wave_id()];
    const int x = (mb % job.wb) * 16, y = (mb / job.wb) * 16;
    const uint32_t *rw = reinterpret_cast<const uint32_t *>(job.records + mb);
    const uint32_t w0 = rw[0], w1 = rw[1], w2 = rw[2], w3 = rw[3], w4 = rw[4];
    const int btype = w0 & 0xFF, mode = (w0 >> 8) & 0xFF;
    auto ref_of = [&](uint32_t target) {
        int k = 0;
        for (int r = job.n_refs - 1; r >= 0; r--)
            if (((job.target_ids >> (8 * r)) & 0xFF) == target) k = r;
        return k;
    };
    // Levels of a transformed MB: loads issued now, used after the prediction.
    const bool coded = (btype & EVXA_BLOCK_COPY) == 0;
    int16_t lvr[EVXA_MB_SAMPLES / 64];
    if (coded) {
        const int16_t *lv = levels + (size_t)mb * EVXA_MB_SAMPLES;
#pragma unroll
        for (int k = 0; k < EVXA_MB_SAMPLES / 64; k++) lvr[k] = lv[lane + 64 * k];
    }
    SearchOut o{};
    o.best_x = x + (int16_t)(w1 & 0xFFFF);
    o.best_y = y + (int16_t)(w1 >> 16);
    o.sp_en = w3 & 0xFF; o.sp_amt = (w3 >> 8) & 0xFF; o.sp_idx = (w3 >> 16) & 0xFF;
    predict_mb(job.refs[ref_of((w0 >> 16) & 0xFF)], o, sh.pred, sh.tmp);
    if (mode == 2) {
        SearchOut b{};
        b.best_x = x + (int16_t)(w2 & 0xFFFF);
        b.best_y = y + (int16_t)(w2 >> 16);
        b.sp_en = w3 >> 24; b.sp_amt = w4 & 0xFF; b.sp_idx = (w4 >> 8) & 0xFF;
        predict_mb(job.refs[ref_of(w0 >> 24)], b, sh.pred_b, sh.tmp);
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64)
            sh.pred[i] = (int16_t)((sh.pred[i] + sh.pred_b[i] + 1) >> 1);
        wave_sync();
    }
    if (coded) {
        // Inverse quantisation, luma at qp and chroma at min(qp + 2, 63).
        const int qp = (w4 >> 16) & 0xFF, cqp = min(qp + 2, 63);
#pragma unroll
        for (int k = 0; k < EVXA_MB_SAMPLES / 64; k++) {
            const int i = lane + 64 * k;
            const int pos = i < 256 ? ((i >> 4) & 7) * 8 + (i & 7) : (i - 256) & 63;
            const int q = i < 256 ? qp : cqp;
            sh.co[i] = (int16_t)(((2 * (int)lvr[k]) * c_quant.qm[pos] * q) / 16);
        }
        wave_sync();
        // Vertical pass: lane = (8x8 block, column) into tmp.
        if (lane < 48) {
            int origin, pitch;
            block_origin(lane >> 3, origin, pitch);
            const int c = origin + (lane & 7);
            int v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = sh.co[c + k * pitch];
#pragma unroll
            for (int i = 0; i < 8; i++) sh.tmp[c + i * pitch] = (int16_t)idct_sample(v, i);
        }
        wave_sync();
    }
    // Horizontal pass plus the prediction (or the prediction alone): lane =
    // (8x8 block, row), eight samples stored as one 16-byte row segment.
    int lo = 0x7FFFFFFF, hi = -0x7FFFFFFF - 1;   // extremes of the stored samples
    if (lane < 48) {
        const int blk = lane >> 3, j = lane & 7;
        int origin, pitch;
        block_origin(blk, origin, pitch);
        const int r = origin + j * pitch;
        int v[8];
        if (coded) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = sh.tmp[r + k];
        }
        uint32_t out[4];
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            int a = sh.pred[r + i], b2 = sh.pred[r + i + 1];
            if (coded) { a += idct_sample(v, i); b2 += idct_sample(v, i + 1); }
            const int sa = (int16_t)a, sb = (int16_t)b2;
            lo = min(lo, min(sa, sb));
            hi = max(hi, max(sa, sb));
            out[i / 2] = (uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)b2 << 16);
        }
        // DevPic views are read-only for the analysis; the destination slot
        // is this kernel's output.
        int16_t *row = const_cast<int16_t *>(
            blk < 4 ? dst.y + (size_t)(y + (blk >> 1) * 8 + j) * dst.py + x + (blk & 1) * 8
                    : (blk == 4 ? dst.u : dst.v) + (size_t)(y / 2 + j) * dst.pc + x / 2);
        uint4 q;
        q.x = out[0]; q.y = out[1]; q.z = out[2]; q.w = out[3];
        __builtin_memcpy(row, &q, 16);
    }
    note_range(wide, lo, hi);             // the picture's range record
}

__global__ void reset_kernel(int32_t *words, int32_t n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        words[i] = 0;
}

// Replicate the edge samples of a picture's three planes into their aprons
// (equivalent to the clamped sampling of interp.cpp:8-14), in one launch
// (blockIdx.y = plane).  Only apron positions are visited: the top and
// bottom bands at full padded width, then the left and right bands of the
// interior rows; every value is read from its clamped interior position.
__global__ void apron3_kernel(int16_t *y0, int16_t *u0, int16_t *v0, int32_t py, int32_t pc,
                              int32_t w, int32_t h)
{
    const int plane = blockIdx.y;
    int16_t *p0 = plane == 0 ? y0 : plane == 1 ? u0 : v0;
    const int pw = plane == 0 ? w : w / 2, ph = plane == 0 ? h : h / 2;
    const int pitch = plane == 0 ? py : pc;
    const int W = pw + 2 * kApron;
    const int band = kApron * W, side = kApron * ph;
    const int n = 2 * band + 2 * side;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int X, Y;
        if (i < 2 * band) {
            const int j = i < band ? i : i - band;
            X = j % W - kApron;
            Y = i < band ? j / W - kApron : ph + j / W;
        } else {
            const int j = i - 2 * band;
            const int k = j < side ? j : j - side;
            Y = k / kApron;
            X = j < side ? k % kApron - kApron : pw + k % kApron;
        }
        const int cx = min(max(X, 0), pw - 1), cy = min(max(Y, 0), ph - 1);
        p0[Y * pitch + X] = p0[cy * pitch + cx];
    }
}

// Folds a spread range record into the picture's range word, max << 16 |
// (min & 0xFFFF), or zero when no sample was recorded (one wave).
__global__ __launch_bounds__(64) void fold_range_kernel(const int32_t *spread, int32_t *range)
{
    const int l = lane_id();
    const int e0 = l < kRangeWays ? spread[l * kRangeStride] : 0;
    const int e1 = l < kRangeWays ? spread[l * kRangeStride + 1] : 0;
    const int lo = 32767 - wave_max(e0), hi = wave_max(e1) - 32768;
    if (l == 0) range[0] = hi < lo ? 0 : (int32_t)((uint32_t)hi << 16 | ((uint32_t)lo & 0xFFFF));
}

// Grows a picture's spread range record by the samples of one w x h plane.
__global__ void range_kernel(const int16_t *plane0, int32_t pitch, int32_t w, int32_t h,
                             int32_t *wide)
{
    int lo = 0x7FFFFFFF, hi = -0x7FFFFFFF - 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < w * h; i += gridDim.x * blockDim.x) {
        const int v = plane0[(i / w) * pitch + i % w];
        lo = min(lo, v);
        hi = max(hi, v);
    }
    note_range(wide, lo, hi);
}

// -------------------------------------------------------------- deblock --

// The in-loop deblocking filter (deblock.cpp:222-306) as a parallel stencil.
// The reference filters in place in raster order: the vertical edges of the
// first band of 8 rows, then per band the horizontal edge at row j = 8b and
// the vertical edges of rows j .. j+7.  A horizontal edge reads rows
// j-4 .. j+3 and writes j-3 .. j+2; a vertical edge reads and writes within
// its row.  Within a band the segments of one direction touch disjoint
// samples, and the horizontal edge at j reads rows j-4 .. j-1 after the
// vertical edges of band b-1 but before anything else has written them, so
// every sample's final value follows from three stages over the unfiltered
// picture:
//   1. vertical edges on rows 8b+3 .. 8b+7 (and all rows of band 0), which no
//      horizontal edge writes before them; row 8b+3 is also read unfiltered
//      by the horizontal edge at 8b, so its result is kept aside;
//   2. horizontal edges on the result;
//   3. vertical edges on rows 8b .. 8b+2 (b >= 1) after stage 2.
// A workgroup filters a tile of kDbTh x kDbTw samples from an LDS copy with
// a 4-sample apron on every side (the reach of one edge filter).
constexpr int kDbTw = 128, kDbTh = 32;
constexpr int kDbCols = kDbTw + 8, kDbRows = kDbTh + 8;
constexpr int kDbEdges = kDbTw / 8 + 1;         // vertical edges at C0, C0+8, ..., C0+kDbTw
constexpr int kDbThreads = 256;
constexpr int kDbBlkRows = 8, kDbBlkCols = 20;   // MBs (luma) or 8x8 blocks (chroma) around a tile

// deblock.cpp:13-35: edge thresholds by qp index.
constexpr uint8_t kDeblockAlpha[64] = {
    0,   0,   0,   0,   0,   0,   0,   1,   1,   1,   2,   2,   3,   3,   4,   5,
    6,   7,   8,   9,   10,  12,  14,  16,  18,  20,  22,  24,  26,  29,  32,  35,
    38,  42,  46,  50,  54,  59,  64,  70,  76,  82,  89,  96,  104, 112, 120, 128,
    138, 148, 158, 170, 182, 196, 210, 224, 240, 255, 255, 255, 255, 255, 255, 255};
constexpr uint8_t kDeblockBeta[64] = {
    0,  0,  0,  0,  0,  0,  0,  0,  1,  1,  1,  1,  2,  2,  2,  3,
    3,  3,  4,  4,  4,  5,  5,  6,  6,  7,  7,  8,  8,  9,  10, 11,
    12, 13, 14, 15, 16, 17, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18,
    18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18, 18};

struct DeblockLds {
    alignas(16) int16_t t[kDbRows][kDbCols];     // rows R0-4 .., columns C0-4 ..
    alignas(16) int16_t r3[kDbTh / 8][kDbCols];  // stage-1 results of each band's row 3
    uint16_t blk[kDbBlkRows][kDbBlkCols];        // copy << 8 | q_index
    uint16_t ab[64];                             // alpha | beta << 8
};

// deblock.cpp:91-165 on one line v = p3 .. q3 (strength 1 or 2), in place;
// `ab` holds the thresholds of the edge's qp index.
__device__ __forceinline__ void deblock_line(int (&v)[8], uint32_t ab, int strength, bool luma)
{
    const int p3 = v[0], p2 = v[1], p1 = v[2], p0 = v[3], q0 = v[4], q1 = v[5], q2 = v[6],
              q3 = v[7];
    const int alpha = ab & 0xFF, beta = ab >> 8;
    // The differences are stored as int16 before the comparisons.
    const int dpq = (int16_t)abs(p0 - q0), dp = (int16_t)abs(p1 - p0), dq = (int16_t)abs(q1 - q0);
    if (dpq >= alpha || dp >= beta || dq >= beta) return;
    // rounded_div (evx_math.h:228-236) by 8 and by 4 -- halves away from
    // zero, C division -- as floor shifts: (x + d/2 - [x < 0]) >> log2(d).
    auto rd8 = [](int x) { return (x + 4 + (x >> 31)) >> 3; };
    auto rd4 = [](int x) { return (x + 2 + (x >> 31)) >> 2; };
    if (strength == 2) {
        v[3] = (int16_t)rd8(p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1);
        v[2] = (int16_t)rd4(p2 + p1 + p0 + q0);
        v[4] = (int16_t)rd8(p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2);
        v[5] = (int16_t)rd4(p0 + q0 + q1 + q2);
        if (luma) {
            v[1] = (int16_t)rd8(2 * p3 + 3 * p2 + p1 + p0 + q0);
            v[6] = (int16_t)rd8(2 * q3 + 3 * q2 + q1 + q0 + p0);
        }
    } else {
        v[3] = (int16_t)rd8(((q0 + p0) * 4) + p1 - q1);
        v[4] = (int16_t)rd8(((q0 + p0) * 4) + q1 - p1);
        if (luma) {
            v[2] = (int16_t)rd8((p2 * 4) + (p0 * 2) + (q0 * 2));
            v[5] = (int16_t)rd8((q2 * 4) + (q0 * 2) + (p0 * 2));
        }
    }
}

// deblock.cpp:57-89 and 167-168: strength of the edge between blocks a (left
// or top) and b, and the thresholds of its qp index.
__device__ __forceinline__ int deblock_edge(const DeblockLds &sh, uint32_t a, uint32_t b,
                                            bool internal, int qp_offset, uint32_t &ab)
{
    const bool ca = a >> 8, cb = b >> 8;
    if (ca && cb) return 0;
    const int avg = !ca && !cb ? ((a & 0xFF) + (b & 0xFF)) >> 1 : !ca ? (a & 0xFF) : (b & 0xFF);
    const int s = ca != cb ? 1 : internal ? 1 : 2;
    const int qp = min(avg + qp_offset, 63);
    ab = sh.ab[s == 2 && qp > 0 ? min(qp + 2, 63) : qp];
    return s;
}

// The planes are tiled kDbTh x kDbTw; tiles of plane k start at first[k].
struct DeblockArgs {
    const int16_t *src[3];
    int16_t *dst[3];
    int32_t sp[3], dp[3];      // pitches (elements)
    int32_t first[4];          // first tile of each plane; first[3] = total
    int32_t w, h, wb, hb;      // luma size and MB grid
};

__device__ __forceinline__ void lds_line(const int16_t *p, int (&v)[8])
{
    const int4 q = *reinterpret_cast<const int4 *>(p);
    const int32_t d[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[2 * k] = (int16_t)(d[k] & 0xFFFF);
        v[2 * k + 1] = d[k] >> 16;
    }
}

__device__ __forceinline__ void lds_store_line(int16_t *p, const int (&v)[8])
{
    int4 q;
    q.x = (v[0] & 0xFFFF) | (v[1] << 16);
    q.y = (v[2] & 0xFFFF) | (v[3] << 16);
    q.z = (v[4] & 0xFFFF) | (v[5] << 16);
    q.w = (v[6] & 0xFFFF) | (v[7] << 16);
    *reinterpret_cast<int4 *>(p) = q;
}

__global__ __launch_bounds__(kDbThreads) void deblock_kernel(const DeblockArgs a,
                                                             const evxa_block_record *records,
                                                             int32_t *wide)
{
    __shared__ DeblockLds sh;
    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    const int pl = (tile >= a.first[1]) + (tile >= a.first[2]);
    const bool luma = pl == 0;
    const int pw = luma ? a.w : a.w / 2, ph = luma ? a.h : a.h / 2;
    // Blocks of the table are 1 << lg samples square; chroma thresholds use
    // qp + 2 (config.h:68).
    const int lg = luma ? 4 : 3, off = luma ? 0 : 2;
    const int tiles_x = (pw + kDbTw - 1) / kDbTw;
    const int t = tile - a.first[pl];
    const int C0 = (t % tiles_x) * kDbTw, R0 = (t / tiles_x) * kDbTh;
    const int16_t *src = a.src[pl];
    const int sp = a.sp[pl];

    // Unfiltered samples of the tile and its apron (zero outside the plane;
    // no edge reaches them), in chunks of 4 samples (w is a multiple of 8),
    // and the block descriptors around the tile.  Every global load is
    // issued before the first LDS store, so their latencies overlap.
    constexpr int kChunks = kDbRows * (kDbCols / 4);
    constexpr int kLoadSteps = (kChunks + kDbThreads - 1) / kDbThreads;
    static_assert(kDbBlkRows * kDbBlkCols <= kDbThreads, "one descriptor per thread");
    int2 ld[kLoadSteps];
#pragma unroll
    for (int k = 0; k < kLoadSteps; k++) {
        const int i = tid + k * kDbThreads;
        const int lr = i / (kDbCols / 4), q = i % (kDbCols / 4);
        const int r = R0 - 4 + lr, c = C0 - 4 + 4 * q;
        ld[k] = make_int2(0, 0);
        if (i < kChunks && r >= 0 && r < ph && c >= 0 && c < pw)
            ld[k] = *reinterpret_cast<const int2 *>(src + (size_t)r * sp + c);
    }
    const int bx0 = max(C0 - 4, 0) >> lg, by0 = max(R0 - 4, 0) >> lg;
    uint32_t bt = 0, bq = 0;
    {
        const int bx = bx0 + tid % kDbBlkCols, by = by0 + tid / kDbBlkCols;
        if (tid < kDbBlkRows * kDbBlkCols && bx < a.wb && by < a.hb) {
            const uint32_t *rw = reinterpret_cast<const uint32_t *>(records + by * a.wb + bx);
            bt = rw[0];   // block_type: byte 0
            bq = rw[4];   // q_index: byte 18
        }
    }
#pragma unroll
    for (int k = 0; k < kLoadSteps; k++) {
        const int i = tid + k * kDbThreads;
        if (i < kChunks)
            *reinterpret_cast<int2 *>(&sh.t[i / (kDbCols / 4)][4 * (i % (kDbCols / 4))]) = ld[k];
    }
    if (tid < kDbBlkRows * kDbBlkCols)
        sh.blk[tid / kDbBlkCols][tid % kDbBlkCols] =
            (uint16_t)(((bt & 4) ? 0x100u : 0u) | ((bq >> 16) & 0xFF));
    if (tid < 64) sh.ab[tid] = kDeblockAlpha[tid] | kDeblockBeta[tid] << 8;
    __syncthreads();

    auto blk = [&](int x, int y) -> uint32_t { return sh.blk[(y >> lg) - by0][(x >> lg) - bx0]; };
    const int inner = (1 << lg) - 1;   // an edge inside one block: position & inner != 0
    // The line of row r across vertical edge e (column C0 + 8e); false if
    // the edge is not filtered.
    auto vertical = [&](int r, int e, int (&v)[8]) -> bool {
        const int i = C0 + 8 * e;
        if (i < 8 || i >= pw) return false;
        uint32_t ab;
        const int s = deblock_edge(sh, blk(i - 1, r), blk(i, r), (i & inner) != 0, off, ab);
        if (!s) return false;
        deblock_line(v, ab, s, luma);
        return true;
    };

    // Stage 1: vertical edges of the rows no horizontal edge writes first:
    // the top apron (rows 4..7 of the band above), rows 3..7 of each band of
    // the tile, and rows 0..2 of the picture's first band.
    constexpr int kS1Rows = 4 + 5 * (kDbTh / 8) + 3;
    for (int i = tid; i < kS1Rows * kDbEdges; i += kDbThreads) {
        const int sr = i / kDbEdges, e = i % kDbEdges;
        const int lr = sr < 4 ? sr
                     : sr < kS1Rows - 3 ? 4 + 8 * ((sr - 4) / 5) + 3 + (sr - 4) % 5
                                        : 4 + sr - (kS1Rows - 3);
        const int r = R0 - 4 + lr;
        if (r < 0 || r >= ph || (r >= 8 && (r & 7) < 3)) continue;
        int v[8];
        lds_line(&sh.t[lr][8 * e], v);
        const bool f = vertical(r, e, v);
        if (r >= 8 && (r & 7) == 3) lds_store_line(&sh.r3[(r - R0) >> 3][8 * e], v);
        else if (f) lds_store_line(&sh.t[lr][8 * e], v);
    }
    __syncthreads();

    // Stage 2: horizontal edges, two columns per thread.
    for (int i = tid; i < (kDbTh / 8 + 1) * (kDbCols / 2); i += kDbThreads) {
        const int hb = i / (kDbCols / 2), cp = i % (kDbCols / 2);
        const int j = R0 + 8 * hb, c = C0 - 4 + 2 * cp;
        if (j < 8 || j >= ph || c < 0 || c >= pw) continue;
        uint32_t ab;
        const int s = deblock_edge(sh, blk(c, j - 1), blk(c, j), (j & inner) != 0, off, ab);
        if (!s) continue;
        int v0[8], v1[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int d = *reinterpret_cast<const int32_t *>(&sh.t[8 * hb + k][2 * cp]);
            v0[k] = (int16_t)(d & 0xFFFF);
            v1[k] = d >> 16;
        }
        deblock_line(v0, ab, s, luma);
        deblock_line(v1, ab, s, luma);
#pragma unroll
        for (int k = 1; k < 7; k++)
            *reinterpret_cast<int32_t *>(&sh.t[8 * hb + k][2 * cp]) = (v0[k] & 0xFFFF) | (v1[k] << 16);
    }
    __syncthreads();

    // Stage 3: vertical edges of rows 8b .. 8b+2 after the horizontal edges.
    for (int i = tid; i < (kDbTh / 8) * 3 * kDbEdges; i += kDbThreads) {
        const int k = i / kDbEdges, e = i % kDbEdges;
        const int r = R0 + 8 * (k / 3) + k % 3;
        if (r < 8 || r >= ph) continue;
        const int lr = r - R0 + 4;
        int v[8];
        lds_line(&sh.t[lr][8 * e], v);
        if (vertical(r, e, v)) lds_store_line(&sh.t[lr][8 * e], v);
    }
    __syncthreads();

    // The tile's samples, 8 per thread and step; rows 8b+3 from stage 1.
    int16_t *dst = a.dst[pl];
    const int dpitch = a.dp[pl];
    int vmin = 0x7FFFFFFF, vmax = -0x7FFFFFFF - 1;   // extremes of the stored samples
    for (int i = tid; i < kDbTh * (kDbTw / 8); i += kDbThreads) {
        const int rr = i / (kDbTw / 8), q = i % (kDbTw / 8);
        const int r = R0 + rr, c = C0 + 8 * q;
        if (r >= ph || c >= pw) continue;
        const int16_t *s = r >= 8 && (r & 7) == 3 ? &sh.r3[rr >> 3][8 * q + 4]
                                                  : &sh.t[rr + 4][8 * q + 4];
        const int2 lo = *reinterpret_cast<const int2 *>(s);
        const int2 hi = *reinterpret_cast<const int2 *>(s + 4);
        const int32_t d[4] = {lo.x, lo.y, hi.x, hi.y};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int x0 = (int16_t)(d[k] & 0xFFFF), x1 = d[k] >> 16;
            vmin = min(vmin, min(x0, x1));
            vmax = max(vmax, max(x0, x1));
        }
        *reinterpret_cast<int4 *>(dst + (size_t)r * dpitch + c) = make_int4(lo.x, lo.y, hi.x, hi.y);
    }
    note_range(wide, vmin, vmax);         // the picture's range record

    // The apron (kApron samples on every side, the clamped sampling of
    // interp.cpp:8-14): tiles on the plane's border replicate their final
    // edge samples, 8 per store, as apron3_kernel does for uploads.
    auto final_row = [&](int r) -> const int16_t * {   // LDS row of r, column C0 at [0]
        return r >= 8 && (r & 7) == 3 ? &sh.r3[(r - R0) >> 3][4] : &sh.t[r - R0 + 4][4];
    };
    auto fill8 = [&](int r, int c, int val) {
        const int d = (val & 0xFFFF) | (val << 16);
        *reinterpret_cast<int4 *>(dst + (ptrdiff_t)r * dpitch + c) = make_int4(d, d, d, d);
    };
    const bool left = C0 == 0, right = C0 + kDbTw >= pw, top = R0 == 0, bottom = R0 + kDbTh >= ph;
    const int rows_here = min(kDbTh, ph - R0), chunks_here = min(kDbTw, pw - C0) / 8;
    constexpr int kSide = kApron / 8;                    // 8-sample chunks per side row
    if (left || right) {
        for (int i = tid; i < 2 * rows_here * kSide; i += kDbThreads) {
            const bool rs = i >= rows_here * kSide;
            const int k = rs ? i - rows_here * kSide : i;
            if (rs ? !right : !left) continue;
            const int r = R0 + k / kSide;
            fill8(r, rs ? pw + 8 * (k % kSide) : -kApron + 8 * (k % kSide),
                  final_row(r)[rs ? pw - 1 - C0 : 0]);
        }
    }
    if (top || bottom) {
        const int cl = left ? kSide : 0, cr = right ? kSide : 0;
        const int per = cl + chunks_here + cr;           // chunks per apron row
        for (int i = tid; i < 2 * kApron * per; i += kDbThreads) {
            const bool bs = i >= kApron * per;
            const int k = bs ? i - kApron * per : i;
            if (bs ? !bottom : !top) continue;
            const int r = bs ? ph + k / per : -kApron + k / per;
            const int16_t *src_row = final_row(bs ? ph - 1 : 0);
            const int ch = k % per;
            if (ch < cl) {
                fill8(r, -kApron + 8 * ch, src_row[0]);
            } else if (ch < cl + chunks_here) {
                const int c = C0 + 8 * (ch - cl);
                const int2 lo = *reinterpret_cast<const int2 *>(src_row + (c - C0));
                const int2 hi = *reinterpret_cast<const int2 *>(src_row + (c - C0) + 4);
                *reinterpret_cast<int4 *>(dst + (ptrdiff_t)r * dpitch + c) =
                    make_int4(lo.x, lo.y, hi.x, hi.y);
            } else {
                fill8(r, pw + 8 * (ch - cl - chunks_here), src_row[pw - 1 - C0]);
            }
        }
    }
}

// Copies the interiors of a picture's three planes (luma w x h) into packed
// planes at `out` (luma, then the two chroma planes; 8 samples per thread
// and step, blockIdx.y = plane).
__global__ void copy3_kernel(const int16_t *y, const int16_t *u, const int16_t *v, int32_t py,
                             int32_t pc, int16_t *out, int32_t w, int32_t h, int32_t *wide)
{
    const int plane = blockIdx.y;
    const int pw = plane ? w / 2 : w, ph = plane ? h / 2 : h;
    const int16_t *src = plane == 0 ? y : plane == 1 ? u : v;
    const int pitch = plane ? pc : py;
    int16_t *o = out + (plane == 0 ? 0 : (size_t)w * h + (plane - 1) * (size_t)(w / 2) * (h / 2));
    const int per = pw / 8, n = per * ph;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r = i / per, c = 8 * (i % per);
        *reinterpret_cast<int4 *>(o + (size_t)r * pw + c) =
            *reinterpret_cast<const int4 *>(src + (size_t)r * pitch + c);
    }
}


// ------------------------------------------------------------------ SAO --

constexpr int kSaoOff = 0, kSaoEdge = 1, kSaoBand = 2;   // sao.h:46-48

// sao.cpp:12-16
__device__ __forceinline__ int sao_clamp(int v) { return min(max(v, 0), 255); }

// sao.cpp:19-24: neighbour a of each edge class; neighbour b mirrors it.
__device__ __forceinline__ void sao_neighbour(int cls, int &dx, int &dy)
{
    dx = cls == 3 ? 1 : cls == 1 ? 0 : -1;
    dy = cls == 0 ? 0 : -1;
}

// sao.cpp:26-37: valley 0, concave 1, convex 2, peak 3, none 4.
__device__ __forceinline__ int sao_category(int p, int a, int b)
{
    const int s = ((a > p) - (a < p)) + ((b > p) - (b < p));
    return s == 2 ? 0 : s == 1 ? 1 : s == -1 ? 2 : s == -2 ? 3 : 4;
}

struct SaoP {
    int type, cls, band;
    int off[4];
};

// Decided params in 25 bits: type | class << 2 | band << 4 | the four
// offsets (-7..7) as 4-bit two's complement from bit 9.
__device__ __forceinline__ uint32_t sao_pack(const SaoP &p)
{
    uint32_t v = (uint32_t)p.type | (uint32_t)p.cls << 2 | (uint32_t)p.band << 4;
#pragma unroll
    for (int k = 0; k < 4; k++) v |= ((uint32_t)p.off[k] & 0xF) << (9 + 4 * k);
    return v;
}

__device__ __forceinline__ SaoP sao_unpack(uint32_t v)
{
    SaoP p;
    p.type = v & 3;
    p.cls = (v >> 2) & 3;
    p.band = (v >> 4) & 31;
#pragma unroll
    for (int k = 0; k < 4; k++) p.off[k] = (int)(v << (32 - 13 - 4 * k)) >> 28;
    return p;
}

// An evxa_sao_params as two 32-bit words.
__device__ __forceinline__ SaoP sao_from_words(uint32_t w0, uint32_t w1)
{
    SaoP p;
    p.type = w0 & 0xFF;
    p.cls = (w0 >> 8) & 0xFF;
    p.band = (w0 >> 16) & 0xFF;
#pragma unroll
    for (int k = 0; k < 4; k++) p.off[k] = (int8_t)(w1 >> (8 * k));
    return p;
}

__device__ __forceinline__ void sao_to_words(const SaoP &p, uint32_t &w0, uint32_t &w1)
{
    w0 = (uint32_t)p.type | (uint32_t)p.cls << 8 | (uint32_t)p.band << 16;
    w1 = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) w1 |= ((uint32_t)p.off[k] & 0xFF) << (8 * k);
}

// Offset i (0..3) of p, by selects (no indexed private array).
__device__ __forceinline__ int sao_off(const SaoP &p, int i)
{
    return i == 0 ? p.off[0] : i == 1 ? p.off[1] : i == 2 ? p.off[2] : p.off[3];
}

// sao.cpp:136-146
__device__ __forceinline__ int sao_bits(const SaoP &p)
{
    if (p.type == kSaoOff) return 2;
    int bits = 2 + (p.type == kSaoEdge ? 2 : 5);
#pragma unroll
    for (int k = 0; k < 4; k++) bits += p.off[k] != 0 ? 4 : 1;
    return bits;
}

template <int CTRL>
__device__ __forceinline__ int64_t dpp64(int64_t v)
{
    const uint32_t lo = (uint32_t)dpp<CTRL>((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)dpp<CTRL>((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t lane_value64(int64_t v, int lane)
{
    const uint32_t lo = (uint32_t)lane_value((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)lane_value((int)(uint32_t)((uint64_t)v >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A 64-bit sum over the wave (row reductions as row_sum, then the 4 rows).
__device__ __forceinline__ int64_t wave_sum64(int64_t v)
{
    v += dpp64<0xB1>(v);
    v += dpp64<0x4E>(v);
    v += dpp64<0x141>(v);
    v += dpp64<0x140>(v);
    return lane_value64(v, 0) + lane_value64(v, 16) + lane_value64(v, 32) + lane_value64(v, 48);
}

struct SaoLds {
    int16_t ry[18][18];          // luma reconstruction, rows and columns -1 .. 16 of the MB
    int16_t ru[10][10];          // U, -1 .. 8
    int32_t boff[32];            // band offsets
    long long bd[32];            // SSD change per band under its offset
};

// The rounded mean of src - rec over a class, clamped to +-7 (sao.cpp:199-214).
__device__ __forceinline__ int sao_mean(int sum, int cnt)
{
    if (cnt <= 0) return 0;
    const int avg = (sum + (sum >= 0 ? cnt / 2 : -cnt / 2)) / cnt;
    return min(max(avg, -7), 7);
}

// The corrected value of sample (x, y) (window position r, c) under p, the
// window holding the unfiltered reconstruction (sao.cpp:333-376 per sample).
template <int N>
__device__ __forceinline__ int sao_corrected(const int16_t *win, int r, int c, int x, int y, int pw,
                                             int ph, const SaoP &p)
{
    constexpr int P = N + 2;
    const int v = win[(r + 1) * P + c + 1];
    if (p.type == kSaoBand) {
        const int idx = (sao_clamp(v) >> 3) - p.band;
        return idx >= 0 && idx < 4 ? sao_clamp(v + sao_off(p, idx)) : v;
    }
    if (p.type != kSaoEdge) return v;
    int dx, dy;
    sao_neighbour(p.cls, dx, dy);
    if (x + dx < 0 || x + dx >= pw || y + dy < 0 || y + dy >= ph || x - dx < 0 || x - dx >= pw ||
        y - dy < 0 || y - dy >= ph)
        return v;
    const int cat = sao_category(v, win[(r + 1 + dy) * P + c + 1 + dx], win[(r + 1 - dy) * P + c + 1 - dx]);
    return cat < 4 ? sao_clamp(v + sao_off(p, cat)) : v;
}

// SSD of an N x N block of src (sv: this lane's samples) against the
// reconstruction corrected by p.
template <int N>
__device__ __forceinline__ int64_t sao_block_ssd(const int16_t *win, const int (&sv)[N * N / 64],
                                                 int x0, int y0, int pw, int ph, const SaoP &p)
{
    constexpr int K = N * N / 64;
    int64_t ssd = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int idx = lane_id() * K + k, r = idx / N, c = idx % N;
        const int d = sv[k] - sao_corrected<N>(win, r, c, x0 + c, y0 + r, pw, ph, p);
        ssd += (int64_t)d * d;
    }
    return wave_sum64(ssd);
}

// sao.cpp:316-331 with 163-239 and 241-314: the best params of one N x N
// block.  Each class's sum and count of src - rec, its offset and the SSD
// change it brings are wave reductions: band by band over the bands
// present (the lowest unprocessed band of the wave each pass), then per
// edge class and category.  The 29 band and 4 edge candidates are then
// priced in the reference's order (ties keep the earlier); best_ssd and
// best_bits describe the winner.
template <int N>
__device__ SaoP sao_rdo(SaoLds &sh, const int16_t *win, const int (&sv)[N * N / 64], int x0,
                        int y0, int pw, int ph, float lambda, int64_t &best_ssd, int &best_bits)
{
    constexpr int K = N * N / 64, P = N + 2;
    const int lane = lane_id();
    if (lane < 32) {
        sh.boff[lane] = 0;
        sh.bd[lane] = 0;
    }
    int v[K], d[K], band[K], cat[K][4];
    int64_t base = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int idx = lane * K + k, r = idx / N, c = idx % N, x = x0 + c, y = y0 + r;
        v[k] = win[(r + 1) * P + c + 1];
        d[k] = sv[k] - v[k];
        base += (int64_t)d[k] * d[k];
        band[k] = sao_clamp(v[k]) >> 3;
#pragma unroll
        for (int cls = 0; cls < 4; cls++) {
            int dx, dy;
            sao_neighbour(cls, dx, dy);
            const bool in = x + dx >= 0 && x + dx < pw && y + dy >= 0 && y + dy < ph && x - dx >= 0 &&
                            x - dx < pw && y - dy >= 0 && y - dy < ph;
            cat[k][cls] = in ? sao_category(v[k], win[(r + 1 + dy) * P + c + 1 + dx],
                                            win[(r + 1 - dy) * P + c + 1 - dx])
                             : 4;
        }
    }
    base = wave_sum64(base);
    __syncthreads();
    // Band classes present, lowest first.
    uint32_t done = 0;
    for (;;) {
        int mine = 32;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (!((done >> k) & 1)) mine = min(mine, band[k]);
        const int b0 = wave_min(mine);
        if (b0 >= 32) break;
        int sum = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (!((done >> k) & 1) && band[k] == b0) {
                sum += d[k];
                cnt++;
            }
        const int off = sao_mean(wave_sum(sum), wave_sum(cnt));
        int64_t dl = 0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (!((done >> k) & 1) && band[k] == b0) {
                const int e = sv[k] - sao_clamp(v[k] + off);
                dl += (int64_t)e * e - (int64_t)d[k] * d[k];
                done |= 1u << k;
            }
        dl = wave_sum64(dl);
        if (lane == 0) {
            sh.boff[b0] = off;
            sh.bd[b0] = dl;
        }
    }
    // Edge classes.
    int eoff[4][4];
    int64_t ed[4];
#pragma unroll
    for (int cls = 0; cls < 4; cls++) {
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            int sum = 0, cnt = 0;
#pragma unroll
            for (int k = 0; k < K; k++)
                if (cat[k][cls] == ct) {
                    sum += d[k];
                    cnt++;
                }
            eoff[cls][ct] = sao_mean(wave_sum(sum), wave_sum(cnt));
        }
        int64_t dl = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int ct = cat[k][cls];
            if (ct < 4) {
                const int o = ct == 0 ? eoff[cls][0] : ct == 1 ? eoff[cls][1] : ct == 2 ? eoff[cls][2] : eoff[cls][3];
                const int e = sv[k] - sao_clamp(v[k] + o);
                dl += (int64_t)e * e - (int64_t)d[k] * d[k];
            }
        }
        ed[cls] = wave_sum64(dl);
    }
    __syncthreads();
    SaoP best{};
    float best_cost = (float)base + lambda * 2.0f;
    best_ssd = base;
    best_bits = 2;
    for (int bp = 0; bp <= 28; bp++) {
        SaoP c{};
        c.type = kSaoBand;
        c.band = bp;
        bool nz = false;
        int64_t ssd = base;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            c.off[j] = sh.boff[bp + j];
            nz = nz || c.off[j] != 0;
            ssd += sh.bd[bp + j];
        }
        if (!nz) continue;
        const int bits = sao_bits(c);
        const float cost = (float)ssd + lambda * (float)bits;
        if (cost < best_cost) {
            best_cost = cost;
            best = c;
            best_ssd = ssd;
            best_bits = bits;
        }
    }
#pragma unroll
    for (int cls = 0; cls < 4; cls++) {
        SaoP c{};
        c.type = kSaoEdge;
        c.cls = cls;
        bool nz = false;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            c.off[j] = eoff[cls][j];
            nz = nz || c.off[j] != 0;
        }
        if (!nz) continue;
        const int64_t ssd = base + ed[cls];
        const int bits = sao_bits(c);
        const float cost = (float)ssd + lambda * (float)bits;
        if (cost < best_cost) {
            best_cost = cost;
            best = c;
            best_ssd = ssd;
            best_bits = bits;
        }
    }
    return best;
}

struct SaoArgs {
    const int16_t *rec_y, *rec_u;   // unfiltered reconstruction, packed (pitch w, w / 2)
    const int16_t *src_y, *src_u;   // the frame's source planes
    int32_t spy, spc;               // their pitches
    int32_t w, h, wb, hb;
    float lambda;
    uint32_t *own_l, *own_c;        // per MB: its own best params (packed)
    float *cur;                     // per MB: the cost of its own luma params
    float *mc;                      // 5 planes of n: merge costs under neighbours' own luma params
    uint64_t *words;                // per MB: the decided params (zero until published)
    evxa_sao_info *table;
    int32_t *counters;              // [1] error flags
};

constexpr int kSaoErrWait = 1, kSaoErrDepth = 2;
constexpr uint32_t kSaoMask25 = (1u << 25) - 1;

// An MB's luma reconstruction window (rows and columns -1 .. 16, zero
// outside the plane) into LDS and this lane's 4 source samples.
__device__ __forceinline__ void sao_load_luma(SaoLds &sh, const SaoArgs &a, int bx, int by,
                                              int (&sy)[4])
{
    const int lane = lane_id();
    for (int i = lane; i < 18 * 18; i += 64) {
        const int X = bx * 16 + i % 18 - 1, Y = by * 16 + i / 18 - 1;
        sh.ry[i / 18][i % 18] = X >= 0 && X < a.w && Y >= 0 && Y < a.h ? a.rec_y[(size_t)Y * a.w + X] : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int idx = lane * 4 + k;
        sy[k] = a.src_y[(size_t)(by * 16 + idx / 16) * a.spy + bx * 16 + idx % 16];
    }
}

// The merge cost of an MB under luma params p (sao.cpp:333-376): its SSD
// plus lambda for the merge flag.
__device__ __forceinline__ float sao_merge_cost(const SaoLds &sh, const SaoArgs &a, int bx, int by,
                                                const int (&sy)[4], const SaoP &p)
{
    return (float)sao_block_ssd<16>(&sh.ry[0][0], sy, bx * 16, by * 16, a.w, a.h, p) + a.lambda * 1.0f;
}

// Stage 1 of the SAO decisions (sao.cpp:409-550, chroma chosen on U): each
// MB's own best luma and chroma params and the cost of its own luma params,
// one wave per MB, no dependency.
__global__ __launch_bounds__(64) void sao_own_kernel(const SaoArgs a)
{
    __shared__ SaoLds sh;
    const int lane = lane_id();
    const int mb = blockIdx.x, bx = mb % a.wb, by = mb / a.wb;
    const int cw = a.w / 2, chh = a.h / 2;
    int sy[4], su[1];
    sao_load_luma(sh, a, bx, by, sy);
    for (int i = lane; i < 10 * 10; i += 64) {
        const int X = bx * 8 + i % 10 - 1, Y = by * 8 + i / 10 - 1;
        sh.ru[i / 10][i % 10] = X >= 0 && X < cw && Y >= 0 && Y < chh ? a.rec_u[(size_t)Y * cw + X] : 0;
    }
    su[0] = a.src_u[(size_t)(by * 8 + lane / 8) * a.spc + bx * 8 + lane % 8];
    __syncthreads();
    int64_t lssd, cssd;
    int lbits, cbits;
    const SaoP own_l = sao_rdo<16>(sh, &sh.ry[0][0], sy, bx * 16, by * 16, a.w, a.h, a.lambda, lssd, lbits);
    const SaoP own_c = sao_rdo<8>(sh, &sh.ru[0][0], su, bx * 8, by * 8, cw, chh, a.lambda, cssd, cbits);
    // sao.cpp:538-550: the merge cost of the own params, less the merge
    // flag, plus their bits; the plain SSD and 2 bits when off.
    const float cur = own_l.type == kSaoOff
                          ? (float)lssd + a.lambda * 2.0f
                          : (((float)lssd + a.lambda * 1.0f) - a.lambda * 1.0f) + a.lambda * (float)lbits;
    if (lane == 0) {
        a.own_l[mb] = sao_pack(own_l);
        a.own_c[mb] = sao_pack(own_c);
        a.cur[mb] = cur;
    }
}

// Stage 2: each MB's merge costs under the own luma params of the MBs whose
// decided params a merge can hand it: left, upper, and the ones a merging
// neighbour takes them from (left of left, upper-left, upper of upper).
// Slots of `mc` per MB: 0 left, 1 upper, 2 left of left, 3 upper-left,
// 4 upper of upper (zero where absent or off).  One wave per MB.
__global__ __launch_bounds__(64) void sao_cost_kernel(const SaoArgs a)
{
    __shared__ SaoLds sh;
    const int mb = blockIdx.x, bx = mb % a.wb, by = mb / a.wb;
    int sy[4];
    sao_load_luma(sh, a, bx, by, sy);
    __syncthreads();
    const int src[5] = {bx > 0 ? mb - 1 : -1, by > 0 ? mb - a.wb : -1, bx > 1 ? mb - 2 : -1,
                        bx > 0 && by > 0 ? mb - a.wb - 1 : -1, by > 1 ? mb - 2 * a.wb : -1};
    float mc[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        mc[k] = 0.0f;
        if (src[k] >= 0) {
            const SaoP p = sao_unpack(a.own_l[src[k]]);
            if (p.type != kSaoOff) mc[k] = sao_merge_cost(sh, a, bx, by, sy, p);
        }
    }
    if (lane_id() < 5) {
        const int l = lane_id();
        a.mc[(size_t)l * a.wb * a.hb + mb] = l == 0 ? mc[0] : l == 1 ? mc[1] : l == 2 ? mc[2] : l == 3 ? mc[3] : mc[4];
    }
}

// Stage 3: the merges in the reference's raster order (sao.cpp:553-583).
// One wave per MB row walks it left to right in chunks of 64 columns: lane j
// holds column base + j's own params, cost and stage-2 merge costs, and the
// upper row's own luma and decided word, read per column with readlane; the
// left MB's decision stays in scalars.  The upper row's words (device-scope
// 64-bit accesses: bit 63 = published | merge_above << 51 | merge_left << 50
// | chroma params << 25 | luma params) are polled for the whole chunk at
// once.  A merge cost is taken from stage 2 when the params handed over are
// the own luma of an MB it was computed for, else computed here.
__global__ __launch_bounds__(64) void sao_merge_kernel(const SaoArgs a)
{
    __shared__ SaoLds sh;
    const int lane = lane_id(), by = blockIdx.x, wb = a.wb, n = a.wb * a.hb;
    auto rl = [](uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); };
    auto rlf = [](float v, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)); };
    uint32_t pl = 0, pc = 0;               // the left MB's decided params
    uint32_t ol1 = 0, ol2 = 0, al1 = 0;    // own luma of the left, left-of-left and upper-left MBs
    int spins = 0;
    bool timed_out = false;
    for (int cb = 0; cb < wb; cb += 64) {
        const int col = cb + lane, m = by * wb + col;
        const bool in = col < wb;
        const uint32_t ol = in ? a.own_l[m] : 0u, oc = in ? a.own_c[m] : 0u;
        const float cur0 = in ? a.cur[m] : 0.0f;
        float mcv[5];
#pragma unroll
        for (int k = 0; k < 5; k++) mcv[k] = in ? a.mc[(size_t)k * n + m] : 0.0f;
        const uint32_t al = in && by > 0 ? a.own_l[m - wb] : 0u;
        const uint32_t al2 = in && by > 1 ? a.own_l[m - 2 * wb] : 0u;
        uint64_t aw = 0;
        int avail = by > 0 ? cb : wb;      // upper words of this chunk known published below this column
        const int end = min(cb + 64, wb);
        for (int bx = cb; bx < end; bx++) {
            const int j = bx - cb, mb = by * wb + bx;
            if (bx >= avail) {
                int nap = 1;
                for (;;) {
                    // Lanes from j on poll their column's upper word.
                    if (in && lane >= j && (aw >> 63) == 0) aw = ld_relaxed64(a.words + m - wb);
                    const uint64_t pub = __ballot(in && (aw >> 63) != 0) >> j;
                    const int run = ~pub == 0 ? 64 - j : __builtin_ctzll(~pub);
                    if (run > 0) {
                        avail = bx + run;
                        break;
                    }
                    if (++spins > kSpinLimit) {
                        timed_out = true;
                        avail = wb;
                        break;
                    }
                    for (int k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(1);
                    nap = min(nap * 2, kMaxNap);
                }
            }
            float cur = rlf(cur0, j);
            uint32_t fl = rl(ol, j), fc = rl(oc, j);
            const uint32_t al_j = rl(al, j), al2_j = rl(al2, j);
            auto cost = [&](uint32_t p) -> float {
                if (bx > 0 && p == ol1) return rlf(mcv[0], j);
                if (by > 0 && p == al_j) return rlf(mcv[1], j);
                if (bx > 1 && p == ol2) return rlf(mcv[2], j);
                if (bx > 0 && by > 0 && p == al1) return rlf(mcv[3], j);
                if (by > 1 && p == al2_j) return rlf(mcv[4], j);
                int sy[4];
                __syncthreads();
                sao_load_luma(sh, a, bx, by, sy);
                __syncthreads();
                return sao_merge_cost(sh, a, bx, by, sy, sao_unpack(p));
            };
            int merge = 0;
            if (bx > 0 && (pl & 3) != kSaoOff) {
                const float mc = cost(pl);
                if (mc < cur) {
                    merge = 1;
                    fl = pl;
                    fc = pc;
                    cur = mc;
                }
            }
            if (merge == 0 && by > 0) {
                const uint64_t w = ((uint64_t)rl((uint32_t)(aw >> 32), j) << 32) | rl((uint32_t)aw, j);
                const uint32_t ul = (uint32_t)w & kSaoMask25;
                if ((ul & 3) != kSaoOff) {
                    const float mc = cost(ul);
                    if (mc < cur) {
                        merge = 2;
                        fl = ul;
                        fc = (uint32_t)(w >> 25) & kSaoMask25;
                    }
                }
            }
            if (lane == 0)
                st_relaxed64(a.words + mb, 1ull << 63 | (uint64_t)(merge == 2) << 51 |
                                               (uint64_t)(merge == 1) << 50 | (uint64_t)fc << 25 | fl);
            // The table entry (evxa_sao_info, five 32-bit words).
            uint32_t lw0, lw1, cw0, cw1;
            sao_to_words(sao_unpack(fl), lw0, lw1);
            sao_to_words(sao_unpack(fc), cw0, cw1);
            uint32_t *e = reinterpret_cast<uint32_t *>(a.table + mb);
            if (lane < 5)
                e[lane] = lane == 0 ? (uint32_t)(merge == 1) | (uint32_t)(merge == 2) << 8
                        : lane == 1 ? lw0 : lane == 2 ? lw1 : lane == 3 ? cw0 : cw1;
            pl = fl;
            pc = fc;
            ol2 = ol1;
            ol1 = rl(ol, j);
            al1 = al_j;
        }
    }
    if (timed_out && lane == 0) atomicOr(&a.counters[1], kSaoErrWait);
}

struct SaoApplyArgs {
    const int16_t *rec[3];     // unfiltered planes, packed
    int16_t *dst[3];
    int32_t dp[3];
    int32_t w, h, wb;
    const evxa_sao_info *table;
    int32_t *wide;
    int32_t *error;
};

// The params that apply to block (bx, by) (sao.cpp:111-119).
__device__ __forceinline__ SaoP sao_params_at(const evxa_sao_info *table, int wb, int bx, int by,
                                              bool luma)
{
    int m = by * wb + bx;
    const uint32_t f = reinterpret_cast<const uint32_t *>(table + m)[0];
    if ((f & 0xFF) && bx > 0) m -= 1;
    else if (((f >> 8) & 0xFF) && by > 0) m -= wb;
    const uint32_t *q = reinterpret_cast<const uint32_t *>(table + m) + (luma ? 1 : 3);
    return sao_from_words(q[0], q[1]);
}

// The final value of sample (x, y) of a plane after the reference's
// in-place raster pass (sao.cpp:98-134): an edge-offset block classifies a
// sample against neighbours as they stand when the block is processed --
// final values in blocks before it in raster order, unfiltered values
// elsewhere -- so such neighbours are evaluated the same way, recursively.
// Such a chain steps into an earlier block each time and can continue only
// from a block corner, which limits it to three steps; a deeper one would be
// flagged.
template <int D>
__device__ int sao_final(const SaoApplyArgs &a, int pl, int x, int y)
{
    const int lg = pl ? 3 : 4, pw = pl ? a.w / 2 : a.w, ph = pl ? a.h / 2 : a.h;
    const int16_t *rec = a.rec[pl];
    const int bx = x >> lg, by = y >> lg;
    const SaoP p = sao_params_at(a.table, a.wb, bx, by, pl == 0);
    const int v = rec[(size_t)y * pw + x];
    if (p.type == kSaoBand) {
        const int idx = (sao_clamp(v) >> 3) - p.band;
        return idx >= 0 && idx < 4 ? sao_clamp(v + sao_off(p, idx)) : v;
    }
    if (p.type != kSaoEdge) return v;
    int dx, dy;
    sao_neighbour(p.cls, dx, dy);
    const int ax = x + dx, ay = y + dy, cx = x - dx, cy = y - dy;
    if (ax < 0 || ax >= pw || ay < 0 || ay >= ph || cx < 0 || cx >= pw || cy < 0 || cy >= ph) return v;
    const int me = by * a.wb + bx;
    auto seen = [&](int qx, int qy) -> int {
        if ((qy >> lg) * a.wb + (qx >> lg) < me) {
            if constexpr (D < 3) return sao_final<D + 1>(a, pl, qx, qy);
            atomicOr(a.error, kSaoErrDepth);
        }
        return rec[(size_t)qy * pw + qx];
    };
    const int cat = sao_category(v, seen(ax, ay), seen(cx, cy));
    return cat < 4 ? sao_clamp(v + sao_off(p, cat)) : v;
}

// SAO of a frame per its table, from the unfiltered planes into the picture
// slot (blockIdx.y = plane), growing the slot's spread range record.
__global__ __launch_bounds__(256) void sao_apply_kernel(const SaoApplyArgs a)
{
    const int pl = blockIdx.y;
    const int pw = pl ? a.w / 2 : a.w, ph = pl ? a.h / 2 : a.h;
    int lo = 0x7FFFFFFF, hi = -0x7FFFFFFF - 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < pw * ph; i += gridDim.x * blockDim.x) {
        const int x = i % pw, y = i / pw;
        const int v = sao_final<0>(a, pl, x, y);
        lo = min(lo, v);
        hi = max(hi, v);
        a.dst[pl][(size_t)y * a.dp[pl] + x] = (int16_t)v;
    }
    note_range(a.wide, lo, hi);
}

// params: n x {x, y, dir_x, dir_y, quarter}
__global__ __launch_bounds__(64) void interp_blocks_kernel(DevPic pic, int32_t n,
                                                           const int32_t *params, int16_t *out)
{
    __shared__ MbLds sh;
    const int b = blockIdx.x;
    if (b >= n) return;
    const int32_t *p = params + b * 5;
    SearchOut r{};
    r.best_x = p[0]; r.best_y = p[1];
    r.sp_en = (p[2] != 0 || p[3] != 0);
    r.sp_amt = p[4];
    // frac index of the direction (motion.cpp:68-91)
    const int m = (p[3] + 1) * 3 + (p[2] + 1);
    r.sp_idx = m < 4 ? m : m - 1;
    predict_mb(pic, r, sh.pred, sh.tmp);
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64)
        out[(size_t)b * EVXA_MB_SAMPLES + i] = sh.pred[i];
}

// params: n x {x, y, pred_x, pred_y}
__global__ __launch_bounds__(64) void search_blocks_kernel(DevPic src, DevPic ref, int32_t w,
                                                           int32_t h, int32_t quality,
                                                           float lambda, int32_t n,
                                                           const int32_t *params,
                                                           evxa_search_result *out)
{
    __shared__ SubpelLds sp;
    const int b = blockIdx.x;
    if (b >= n) return;
    const int32_t *p = params + b * 4;
    const Geo geo = lane_geometry();
    Blk s, pre;
    load_cand(src, geo, p[0], p[1], s);
    load_cand(ref, geo, p[0], p[1], pre);
    const WinPos win = stage_window<1>(ref, p[0], p[1], kReachMax, sp);
    const bool narrow = narrow_pair(src.wide, ref.wide);
    window_ready();
    SearchOut o = inter_search<1>(ref, w, h, narrow, s, geo, pre, win, p[0], p[1], (quality >> 2) + 1,
                               lambda, p[2], p[3], sp);
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

// ---------------------------------------------------- intra block search --
//
// calculate_intra_prediction (motion.cpp:485-565): a block search of the MB
// at (x, y) in the picture being reconstructed, over the part already
// reconstructed -- a candidate that reaches below row y - 16 and right of
// column x - 16 is skipped (motion.cpp:361-365, 433-438).  The state starts
// from the block against zero (luma SAD, MAD INT32_MAX, cost the SAD); a
// 3x3 scan at step 16 over rows y - 32 .. y and columns x - 16 .. x + 16
// (motion.cpp:525-527) is refined by 3x3 levels at steps 8, 4, 2, 1 around
// the best until a copy (MAD < T) is found, then the sub-pel step runs in
// every case (motion.cpp:540).  Candidates are evaluated four at a time (one
// per 16-lane group) in 32-bit arithmetic and folded one by one in scan
// order with evaluate_motion_candidate's and
// evaluate_subpel_motion_candidate's acceptance (motion.cpp:125-261).

struct IntraState {
    int x, y, sad, mad, ssd;
    uint32_t cost;                 // fp32 bits (costs are >= 0)
    int n_int;
};

__device__ __forceinline__ bool intra_unstable(int cx, int cy, int x, int y)
{
    return cy > y - 16 && cx > x - 16;
}

// The block at (x, y) of a staged window, in the lane layout.
__device__ __forceinline__ void load_cand(const WinView &p, const Geo &g, int x, int y, Blk &b)
{
#pragma unroll
    for (int r = 0; r < 4; r++)
        lds_load4(p.y, (y - p.ly + g.oy + r) * p.py + x - p.lx + g.ox, b.y[2 * r], b.y[2 * r + 1]);
    const int co = ((y >> 1) - p.cy + g.cr) * p.pc + (x >> 1) - p.cx + g.cc;
    lds_load4(p.u, co, b.u[0], b.u[1]);
    lds_load4(p.v, co, b.v[0], b.v[1]);
}

// One 3x3 scan (perform_intra_motion_search, motion.cpp:348-379) of step
// `step` whose first position is (left, top) from the current best; true
// once the best is a copy (the scan stops there).  `pic` is the picture
// being reconstructed: a device picture, or a window of it in LDS.
template <class Pic>
__device__ bool intra_scan(const Pic &pic, int pw, int ph, const Blk &s, const Geo &geo, int x,
                           int y, int thresh, float lambda, int left, int top, int step,
                           IntraState &st)
{
    const int g = lane_id() >> 4;
    const int base_x = st.x, base_y = st.y, max_x = pw - 16, max_y = ph - 16;
    for (int c = 0; c < 3; c++) {
        const int p = 4 * c + g;
        const int cx = base_x + left + (p % 3) * step, cy = base_y + top + (p / 3) * step;
        const bool valid = p < 9 && !intra_unstable(cx, cy, x, y) && cx >= 0 && cx <= max_x &&
                           cy >= 0 && cy <= max_y;
        Blk r;
        load_cand(pic, geo, valid ? cx : x, valid ? cy : y, r);   // invalid slots read the MB's own position
        int sad, mad, satd;
        cand_metrics(false, s, r, sad, mad, satd);
        const uint32_t cst = cost_bits(sad, satd, cx - x, cy - y, lambda);
        const uint64_t vm = __ballot(valid);
        for (int k = 0; k < 4; k++) {
            if (((vm >> (16 * k)) & 1u) == 0) continue;
            const int pk = 4 * c + k;
            const int ik = left + (pk % 3) * step, jk = top + (pk / 3) * step;
            const int kx = base_x + ik, ky = base_y + jk;
            const int ksad = lane_value(sad, 16 * k), kmad = lane_value(mad, 16 * k);
            const uint32_t kcost = (uint32_t)lane_value((int)cst, 16 * k);
            const int ssd = (kx - x) * (kx - x) + (ky - y) * (ky - y);
            st.n_int += (ik != 0 || jk != 0) ? 1 : 0;
            bool take;
            if (st.mad < thresh)
                take = kmad < st.mad || (kmad == st.mad && ssd < st.ssd);
            else   // C precedence: && binds tighter than || (motion.cpp:166-168)
                take = (kcost < st.cost || ((kcost == st.cost && ssd < st.ssd) && ksad < kSadThreshold)) ||
                       kmad < thresh;
            if (take) {
                st.x = kx; st.y = ky;
                st.sad = ksad; st.mad = kmad; st.ssd = ssd; st.cost = kcost;
            }
            if (st.mad < thresh) return true;
        }
    }
    return false;
}

// A block of 384 samples in LDS (Y 16x16, U 8x8, V 8x8) in the lane layout.
__device__ __forceinline__ void load_mb_lds(const int16_t *mb, const Geo &g, Blk &b)
{
#pragma unroll
    for (int r = 0; r < 4; r++) lds_load4(mb, (g.oy + r) * 16 + g.ox, b.y[2 * r], b.y[2 * r + 1]);
    lds_load4(mb, 256 + g.cr * 8 + g.cc, b.u[0], b.u[1]);
    lds_load4(mb, 320 + g.cr * 8 + g.cc, b.v[0], b.v[1]);
}

__device__ __forceinline__ int frac_index_of(int i, int j)
{
    // compute_motion_frac_index_from_direction (motion.cpp:68-91)
    i++; j++;
    if (j == 0) return i;
    if (j == 1) return i == 0 ? 3 : 4;
    return i + 5;
}

template <class Pic>
__device__ SearchOut intra_search(const Pic &pic, int pw, int ph, const Blk &s, const Geo &geo,
                                  int x, int y, int thresh, float lambda, MbLds &sh)
{
    int a = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) a += abs(lo16(s.y[k])) + abs(hi16(s.y[k]));
    const int sad0 = lane_value(row_sum(a), 0);       // compute_block_sad(src), analysis.h:134-145
    IntraState st{x, y, sad0, kInt32Max, kInt32Max, __float_as_uint(__int2float_rn(sad0)), 0};
    intra_scan(pic, pw, ph, s, geo, x, y, thresh, lambda, -kSearchRadius, -2 * kSearchRadius,
               kSearchRadius, st);
    if (st.mad >= thresh)
        for (int step = kSearchRadius / 2; step > 0; step >>= 1)
            if (intra_scan(pic, pw, ph, s, geo, x, y, thresh, lambda, -step, -step, step, st)) break;

    // perform_intra_subpixel_motion_search (motion.cpp:408-448): half then
    // quarter of each direction around the integer best.
    int sp_en = 0, sp_amt = 0, sp_idx = 0, n_ax = 0, n_dg = 0;
    for (int j = -1; j <= 1; j++)
        for (int i = -1; i <= 1; i++) {
            const int tx = st.x + i, ty = st.y + j;
            if ((i == 0 && j == 0) || intra_unstable(tx, ty, x, y)) continue;
            if (tx < 0 || tx > pw - 16 || ty < 0 || ty > ph - 16) continue;
            if (i != 0 && j != 0) n_dg++; else n_ax++;
            for (int q = 0; q < 2; q++) {
                SearchOut r{st.x, st.y, 0, 0, 0, 0.0f, 1, q, frac_index_of(i, j), 0, 0, 0, 0};
                predict_mb(pic, r, sh.pred, sh.tmp);
                Blk c;
                load_mb_lds(sh.pred, geo, c);
                int sad, mad, satd;
                cand_metrics(false, s, c, sad, mad, satd);
                sad = lane_value(sad, 0);
                mad = lane_value(mad, 0);
                const bool take = st.mad < thresh ? mad < st.mad
                                                  : ((sad < st.sad && sad < kSadThreshold) || mad < thresh);
                if (take) {
                    sp_en = 1; sp_amt = q; sp_idx = frac_index_of(i, j);
                    st.sad = sad; st.mad = mad;
                }
                wave_sync();            // every lane has read sh.pred before the next block
            }
        }
    SearchOut o;
    o.best_x = st.x; o.best_y = st.y;
    o.sad = st.sad; o.mad = st.mad; o.ssd = st.ssd;
    o.cost = __uint_as_float(st.cost);
    o.sp_en = sp_en; o.sp_amt = sp_amt; o.sp_idx = sp_idx;
    o.btype = EVXA_BLOCK_INTRA;
    if (st.x != x || st.y != y || sp_en) o.btype |= EVXA_BLOCK_MOTION;
    if (st.mad < thresh) o.btype |= EVXA_BLOCK_COPY;
    o.n_int = st.n_int; o.n_ax = n_ax; o.n_dg = n_dg;
    return o;
}

// One wave per MB: intra searches of the MBs at params[2b], params[2b + 1]
// of `src` over `pic`.
__global__ __launch_bounds__(64) void intra_blocks_kernel(DevPic src, DevPic pic, int32_t w,
                                                          int32_t h, int32_t quality, float lambda,
                                                          int32_t n, const int32_t *params,
                                                          evxa_search_result *out)
{
    __shared__ MbLds sh;
    const int b = blockIdx.x;
    if (b >= n) return;
    const int x = params[2 * b], y = params[2 * b + 1];
    const Geo geo = lane_geometry();
    Blk s;
    load_cand(src, geo, x, y, s);
    const SearchOut o = intra_search(pic, w, h, s, geo, x, y, (quality >> 2) + 1, lambda, sh);
    if (lane_id() == 0) {
        evxa_search_result r;
        r.best_x = o.best_x; r.best_y = o.best_y;
        r.sad = o.sad; r.mad = o.mad; r.ssd = o.ssd;
        r.cost_bits = fbits(o.cost);
        r.sp_enabled = o.sp_en; r.sp_amount = o.sp_amt; r.sp_index = o.sp_idx;
        r.block_type = o.btype;
        r.motion_x = o.best_x - x; r.motion_y = o.best_y - y;
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
    __shared__ MbLds sh;
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

// ------------------------------------------------------- in-loop encode --
//
// encode_slice (encode.cpp:751-813) of one frame on the device: per MB the
// mode decision of classify_block (I and P, encode.cpp:168-370, with RDO
// for P) or classify_block_bidir (B, encode.cpp:372-476), encode_block with
// the trellis (encode.cpp:479-748; quantize.cpp:454-464, 615-680), and
// decode_block into the picture being reconstructed (decode.cpp:31-433).
//
// The intra-frame block search (motion.cpp:485-565) and the directional
// intra modes (intra_pred.cpp:177-326) read that picture while it is being
// written, so MBs run in a wavefront: MB (bx, by) starts once MB (bx - 1, by)
// and MB (bx + kEncodeLag, by - 1) are done (the reference's row workers keep
// the same kind of lag, encode.cpp:864-875).  An intra candidate lies on or
// above row y - 16, or on or left of column x - 16 (motion.cpp:361-365,
// 433-438), within 15 samples of a position of the first 3x3 scan of step
// 16, whose reach is column x + 16 and row y - 32; with the 6-tap apron its
// reads stay within columns x - 34 .. x + 49 and rows y - 50 .. y + 33.  So
// the MB reads nothing right of MB bx + 3 in the row above, and of the rows
// below only columns x - 34 .. x + 2, which no MB has written yet when the
// lag is 3 (row by + 1 has at most reached MB bx - 4).  Each MB therefore
// sees the picture exactly as the sequential encoder does: reconstructed
// samples above and to the left, the slot's previous content elsewhere.
// The frames of a batch interleave their MBs in one such order.
//
// Four waves per MB (encode_mb).  Pictures written in the launch move between workgroups
// (and XCDs) only through device-scope accesses: the reconstruction, and the
// apron of MBs on the picture border, are stored with device-scope stores
// that complete before the MB's predictor word is published, and a MB
// stages the window of the picture it reads (luma 84 x 84, chroma 46 x 45
// around it) into LDS with device-scope loads after its dependencies are
// published.  The window keeps the apron's clamped reads: its origin moves
// inside the picture's apron at the borders.

// The window holds every read of the intra search, the intra predictions
// and the intra block copies: a candidate (cx, cy) lies in [x - 31, x + 31] x
// [y - 47, y + 15], and its 6-tap reads reach 3 samples before and 18 after
// it (luma), 3 before and 10 after (chroma, at (cx >> 1, cy >> 1)).
constexpr int kIwLS = 84, kIwLR = 84;    // luma: columns x - 34 .. x + 49, rows y - 50 .. y + 33
constexpr int kIwCS = 46, kIwCR = 45;    // chroma: columns x/2 - 20 .. x/2 + 25, rows y/2 - 27 .. y/2 + 17
constexpr int kIntraNone = 7;            // intra_pred.h:42-49: no directional prediction
constexpr uint64_t kEncWaitTicks = 200000000ull;   // 2 s of s_memrealtime (100 MHz)

struct IntraWin {
    alignas(16) int16_t y[kIwLR * kIwLS + 4];   // + one dword read past the last sample
    alignas(16) int16_t u[kIwCR * kIwCS + 4];
    alignas(16) int16_t v[kIwCR * kIwCS + 4];
};

// The neighbour samples of the directional intra modes (intra_pred.cpp:
// 12-52): per plane the row above, the column left and the corner, 128
// where absent, and the DC value (intra_pred.cpp:58-75).
struct IntraEdges {
    int16_t top[3][16];
    int16_t left[3][16];
    int16_t corner[3];
    int16_t dc[3];
    int32_t has_top, has_left;
};

// The decision of a MB as a block descriptor (common.h:87-115).
struct EncDesc {
    int btype, pmode, tgt, tgt_b;
    int mvx, mvy, mvx_b, mvy_b;
    int sp, sa, si, sp_b, sa_b, si_b;
    int imode, qp, var;
};

// Intra matrix (quantize.cpp:13-23), its dead-zone weights (quantize.cpp:
// 42-52), the reciprocal multipliers of the matrix, and the zig-zag scan
// (scan.h:52-62: anti-diagonals, odd ones walked downwards).
struct IntraTables {
    int16_t qm[64];
    uint8_t dz[64];
    uint32_t qm_magic[64];
    uint8_t zigzag[64];
};

constexpr IntraTables make_intra_tables()
{
    IntraTables t{
        {8, 10, 12, 15, 19, 23, 25, 27, 10, 12, 14, 17, 21, 25, 27, 28,
         12, 14, 18, 21, 24, 26, 28, 30, 15, 17, 21, 24, 26, 28, 30, 32,
         19, 21, 24, 26, 28, 30, 32, 35, 23, 25, 26, 28, 30, 32, 35, 38,
         25, 27, 28, 30, 32, 35, 38, 41, 27, 28, 30, 32, 35, 38, 41, 45},
        {0, 1, 1, 1, 1, 1, 2, 2, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 2, 2, 2, 2,
         1, 1, 1, 2, 2, 2, 2, 3, 1, 1, 2, 2, 2, 2, 3, 3, 1, 2, 2, 2, 2, 3, 3, 3,
         2, 2, 2, 2, 3, 3, 3, 3, 2, 2, 2, 3, 3, 3, 3, 3},
        {},
        {}};
    for (int i = 0; i < 64; i++) t.qm_magic[i] = 0xFFFFFFFFu / (uint32_t)t.qm[i] + 1u;
    int p = 0;
    for (int d = 0; d < 15; d++)
        for (int k = 0; k <= d; k++) {
            const int r = (d & 1) ? k : d - k, c = d - r;
            if (r < 8 && c < 8) t.zigzag[p++] = (uint8_t)(r * 8 + c);
        }
    return t;
}

__constant__ IntraTables c_intra = make_intra_tables();

// The quantiser tables and the zig-zag scan as the encode's coding stage
// reads them: copied into LDS once per workgroup, since every lookup is
// indexed per lane inside the serial loops of the quantisation, the
// trellis and the inverse quantisation.
struct CodeTables {
    QuantTables inter;
    IntraTables intra;
};

__device__ __forceinline__ void load_code_tables(CodeTables &t)
{
    const uint32_t *a = reinterpret_cast<const uint32_t *>(&c_quant);
    const uint32_t *b = reinterpret_cast<const uint32_t *>(&c_intra);
    uint32_t *ta = reinterpret_cast<uint32_t *>(&t.inter);
    uint32_t *tb = reinterpret_cast<uint32_t *>(&t.intra);
    for (int i = threadIdx.x; i < (int)(sizeof(QuantTables) / 4); i += blockDim.x) ta[i] = a[i];
    for (int i = threadIdx.x; i < (int)(sizeof(IntraTables) / 4); i += blockDim.x) tb[i] = b[i];
}

// Waves of the encode's workgroup and the RDO's candidates.
constexpr int kEncWaves = 4;
// Coded candidates (slots): 0 the intra-frame block search's result, 1 the
// (forward) inter search's, 2 the backward search's (B) or the plain intra
// block (P), 3..5 the directional modes (P: the three candidates; I and B:
// the chosen one in 3).
constexpr int kEncCands = kEncodeCands;
constexpr int kSlotIntra = 0, kSlotInter = 1, kSlotPlain = 2, kSlotBwd = 2, kSlotMode = 3;
constexpr int kSlotBidir = 4;         // B: the average of both searches

// What the waves of a MB's workgroup exchange.
struct EncXchg {
    SearchOut intra, inter[2];
    int32_t dir_mode, dir_sad, cands, pmx, pmy;
    int32_t preds;                        // B: bit k set once search k's prediction is in LDS
    int32_t bsad;                         // B: SAD of the average of both predictions
    float cost[kEncCands];
    int32_t qp[kEncCands], var[kEncCands];
};

struct EncLds {
    IntraWin iw;                          // the picture being reconstructed, around the MB
    SubpelLds sp[2];                      // the inter searches (waves 1 and 2)
    HalfPlanes<16> hy;                    // the intra search's half-pel planes (wave 0)
    HalfPlanes<8> hu, hv;
    MbLds mb[kEncWaves];                  // per wave: source, predictions, residual, scratch
    IntraEdges edges;
    EncXchg x;
    CodeTables tables;
    int16_t lv[kEncCands][EVXA_MB_SAMPLES];   // per candidate: levels after the trellis,
    int16_t rec[kEncCands][EVXA_MB_SAMPLES];  // reconstruction (coefficients: EncArgs::coscratch)
};

struct EncArgs {
    const EncJob *jobs;         // the frames of the launch
    const uint32_t *order;      // tickets: job << kMbBits | MB, in dependency order
    int32_t *counters;          // [0] next ticket, [kCounterError] error flags
    int32_t n_tickets;
    int16_t *coscratch;         // per workgroup, kEncCands x 384 coefficients
};

// quantize.cpp:74-95: the DC scales of intra blocks.
__device__ __forceinline__ int luma_dc_scale(int qp)
{
    if (qp < 5) return max(qp << 1, 2);
    if (qp < 9) return qp << 1;
    if (qp < 25) return qp + 8;
    return (qp << 1) - 16;
}

__device__ __forceinline__ int chroma_dc_scale(int qp)
{
    if (qp < 5) return max(qp << 1, 2);
    if (qp < 25) return (qp + 13) >> 1;
    return qp - 6;
}

// evx_math.h:64 with the byte log2 of 69-87: floor(log2(n & 0xFF)) + 1,
// log2(0) = 0.
__device__ __forceinline__ int bits_u8(uint32_t n)
{
    n &= 0xFFu;
    return (n ? 31 - __builtin_clz(n) : 0) + 1;
}

// quantize.cpp:517-532 (EVX_ENABLE_SIGMAP_CODING): the bits of a level.
__device__ __forceinline__ int coeff_bits(int level)
{
    if (level == 0) return 1;
    const int v = min(abs(level), 255);
    return 2 + (v > 1 ? 1 : 0) + (v > 2 ? 1 + 2 * bits_u8((uint32_t)(v - 2)) : 1);
}

// Intra matrices for intra blocks without motion, inter otherwise
// (quantize.cpp:457, 469).
__device__ __forceinline__ bool intra_quant(int btype)
{
    return (btype & EVXA_BLOCK_INTRA) != 0 && (btype & EVXA_BLOCK_MOTION) == 0;
}

// Scan position p of an 8x8 block as a raster index, by the block's intra
// mode (scan.h:154-159): rows for vertical modes, columns for horizontal
// ones, zig-zag otherwise.
__device__ __forceinline__ int scan_pos(const CodeTables &t, int imode, int p)
{
    if (imode == 1 || imode == 5) return p;
    if (imode == 2 || imode == 6) return (p & 7) * 8 + (p >> 3);
    return t.intra.zigzag[p];
}

// Sample k of an MB in MB order: its 8x8 block's raster position and
// whether it is chroma.
__device__ __forceinline__ int mb_pos(int i, bool &chroma)
{
    chroma = i >= 256;
    return chroma ? (i - 256) & 63 : ((i >> 4) & 7) * 8 + (i & 7);
}

// Inverse quantisation of one level (quantize.cpp:271-301, 322-332; the
// trellis's own restatement, quantize.cpp:481-514, is the same arithmetic).
__device__ __forceinline__ int dequant(const CodeTables &t, int level, int pos, int bq, bool intra,
                                       bool chroma)
{
    if (intra && pos == 0) return (int16_t)(level * (chroma ? chroma_dc_scale(bq) : luma_dc_scale(bq)));
    const int qm = intra ? t.intra.qm[pos] : t.inter.qm[pos];
    return (int16_t)((2 * level * qm * bq) / 16);
}

// Quantisation of one coefficient: quantize_intra_block_8x8
// (quantize.cpp:120-200) or quantize_inter_block_8x8 (217-252).
__device__ __forceinline__ int quant(const CodeTables &t, int sv, int pos, int bq, uint32_t mq,
                                     bool intra, bool chroma, int dcs, uint32_t mdc)
{
    if (intra) {
        if (pos == 0) return (int16_t)rdiv_magic(sv, dcs, mdc);   // dcs: the DC scale, mdc its magic
        int qf = (int16_t)rdiv_magic(sv * 16, t.intra.qm[pos], t.intra.qm_magic[pos]);
        if (bq >= 16) {
            const int sg = (qf > 0) - (qf < 0);
            qf = (int16_t)(qf - (int16_t)((sg * bq * t.intra.dz[pos]) >> 1));
        }
        return (int16_t)rdiv_magic(qf, bq << 1, mq);
    }
    const int qf = (int16_t)rdiv_magic(sv * 16, t.inter.qm[pos], t.inter.qm_magic[pos]);
    int w = t.inter.dz[pos];
    if (bq >= 16) w = min(w + 1, 6);
    const int sg = (qf > 0) - (qf < 0);
    const int dz = (int16_t)((sg * bq * w) >> 1);
    return (int16_t)rdiv_magic(qf - dz, bq << 1, mq);
}

// trellis_optimize_macroblock (quantize.cpp:615-680) on the levels `lv` of
// coefficients `co`, one 8x8 block at a time, lane p holding scan position
// p.  The backward pass walks from the last nonzero level down to position
// 1, zeroing while zero costs no more than the level, and stops at the
// first that it keeps: so it keeps every nonzero level at or below the
// highest position whose level it would keep.  The forward pass then lowers
// each nonzero level (positions >= 1) by one when that costs less; its
// decisions do not depend on each other.
__device__ void trellis_mb(const CodeTables &t, int qp, bool intra, int imode, float lambda,
                           const int16_t *co, int16_t *lv)
{
    const int p = lane_id();
    const int cqp = min(qp + 2, 63);
    // Scan position p's raster position is the same in every block; so are
    // its matrix entry and whether it is the DC.
    const int pos = scan_pos(t, imode, p);
    const int qm = intra ? t.intra.qm[pos] : t.inter.qm[pos];
    const bool dc_scaled = intra && pos == 0;
    int c[6], level[6];
#pragma unroll
    for (int blk = 0; blk < 6; blk++) {
        int origin, pitch;
        block_origin(blk, origin, pitch);
        const int idx = origin + (pos >> 3) * pitch + (pos & 7);
        c[blk] = co[idx];
        level[blk] = lv[idx];
    }
#pragma unroll
    for (int blk = 0; blk < 6; blk++) {
        const bool chroma = blk >= 4;
        const int bq = chroma ? cqp : qp;
        const int dcs = chroma ? chroma_dc_scale(bq) : luma_dc_scale(bq);
        // quantize.cpp:481-514 (the inverse quantisation the costs use) and
        // 536-560.
        auto cost = [&](int lvl) {
            const int deq = lvl == 0 ? 0
                          : dc_scaled ? (int)(int16_t)(lvl * dcs)
                                      : (int)(int16_t)((2 * lvl * qm * bq) / 16);
            const int diff = c[blk] - deq;
            const int sq = (int)((uint32_t)diff * (uint32_t)diff);
            return __fadd_rn(__int2float_rn(sq), __fmul_rn(lambda, __int2float_rn(coeff_bits(lvl))));
        };
        int lvl = level[blk];
        const uint64_t nz = __ballot(lvl != 0);
        if (nz == 0) continue;
        const int last = 63 - __builtin_clzll(nz);
        const float keep = cost(lvl);
        const float zero = cost(0);
        const uint64_t kept = __ballot(lvl != 0 && p >= 1 && p <= last && !(zero <= keep));
        const int stop = kept ? 63 - __builtin_clzll(kept) : 0;
        if (p >= 1 && p > stop) lvl = 0;
        if (p >= 1 && lvl != 0) {
            const int reduced = (int16_t)(lvl - ((lvl > 0) - (lvl < 0)));
            if (cost(reduced) < keep) lvl = reduced;
        }
        if (lvl != level[blk]) {
            int origin, pitch;
            block_origin(blk, origin, pitch);
            lv[origin + (pos >> 3) * pitch + (pos & 7)] = (int16_t)lvl;
        }
    }
    wave_sync();
}

// The descriptor's residual coding (encode.cpp:717-739 with
// quantize_macroblock and the trellis) and its reconstruction
// (decode.cpp: inverse quantisation, inverse DCT added to the prediction):
// from sh.mb.src and the prediction in sh.mb.pred into co (global memory),
// lv and rec (LDS).
// Copies are not transformed (encode.cpp:742-744): zero coefficients and
// levels, the prediction as the reconstruction.
__device__ __forceinline__ void code_mb(const DevJob &job, const CodeTables &t, EncDesc &d,
                                        MbLds &mb, int16_t *co, int16_t *lv, int16_t *rec)
{
    const int lane = lane_id();
    if (d.btype & EVXA_BLOCK_COPY) {
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
            co[i] = 0;
            lv[i] = 0;
            rec[i] = mb.pred[i];
        }
        wave_sync();
        return;
    }
    fdct_mb(mb.src, mb.pred, mb.res, mb.tmp);
    int qp, var2;
    mb_qp(mb.res, job.quality, qp, var2);
    d.qp = qp;
    d.var = var2;
    const bool intra = intra_quant(d.btype);
    const int cqp = min(qp + 2, 63);
    const uint32_t m_luma = magic_of(qp << 1), m_chroma = magic_of(cqp << 1);
    const int dc_luma = luma_dc_scale(qp), dc_chroma = chroma_dc_scale(cqp);
    const uint32_t mdc_luma = magic_of(dc_luma), mdc_chroma = magic_of(dc_chroma);
#pragma unroll
    for (int k = 0; k < EVXA_MB_SAMPLES / 64; k++) {
        const int i = lane + 64 * k;
        bool chroma;
        const int pos = mb_pos(i, chroma);
        const int sv = mb.res[i];
        co[i] = (int16_t)sv;
        lv[i] = (int16_t)quant(t, sv, pos, chroma ? cqp : qp, chroma ? m_chroma : m_luma, intra,
                               chroma, chroma ? dc_chroma : dc_luma, chroma ? mdc_chroma : mdc_luma);
    }
    wave_sync();
    trellis_mb(t, qp, intra, d.imode, job.lambda, mb.res, lv);
    // Inverse quantisation into mb.res, then the inverse DCT: vertical
    // pass (lane = block column) into tmp, horizontal pass plus the
    // prediction (lane = block row) into rec, stored as int16.
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        bool chroma;
        const int pos = mb_pos(i, chroma);
        mb.res[i] = (int16_t)dequant(t, lv[i], pos, chroma ? cqp : qp, intra, chroma);
    }
    wave_sync();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int c = origin + (lane & 7);
        int v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = mb.res[c + k * pitch];
#pragma unroll
        for (int i = 0; i < 8; i++) mb.tmp[c + i * pitch] = (int16_t)idct_sample(v, i);
    }
    wave_sync();
    if (lane < 48) {
        int origin, pitch;
        block_origin(lane >> 3, origin, pitch);
        const int r = origin + (lane & 7) * pitch;
        int v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = mb.tmp[r + k];
#pragma unroll
        for (int i = 0; i < 8; i++) rec[r + i] = (int16_t)(mb.pred[r + i] + idct_sample(v, i));
    }
    wave_sync();
}

// estimate_block_bits (encode.cpp:39-110; EVX_ENABLE_SIGMAP_CODING,
// EVX_ENABLE_MPM) of a descriptor, the residual part counted on the luma of
// `blk`: the reference estimates it from the buffer that holds the
// candidate's reconstruction by then (encode.cpp:150-156).
__device__ uint32_t block_bits(const EncDesc &d, const int16_t *blk)
{
    uint32_t bits = 3;
    if (!(d.btype & EVXA_BLOCK_INTRA)) bits += 2;
    if (d.btype & EVXA_BLOCK_MOTION)
        bits += 2 * bits_u8((uint32_t)(abs(d.mvx) + 1)) + 2 * bits_u8((uint32_t)(abs(d.mvy) + 1)) +
                (d.sp ? 5 : 1);
    if (d.btype & EVXA_BLOCK_COPY) return bits;
    if (d.btype == EVXA_BLOCK_INTRA) bits += 3;
    bits += 2 * bits_u8((uint32_t)d.qp + 1);
    // Lane l holds luma samples 4l .. 4l + 3: row l >> 2, columns
    // (l & 3) * 4 .. + 3, in 8x8 block ((l >> 5) << 1) | ((l >> 1) & 1).
    const int lane = lane_id();
    int nz = 0, cb = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int v = blk[4 * lane + k];
        if (v != 0) {
            nz++;
            cb += coeff_bits(v);
        }
    }
    // nz (at most 64) and cb (at most 64 * 21) of a block packed in one
    // word: one reduction per block.
    const int mine = ((lane >> 5) << 1) | ((lane >> 1) & 1);
    const int packed = nz | (cb << 8);
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int v = wave_sum(mine == b ? packed : 0);
        const int n = v & 0xFF, c = v >> 8;
        bits += (uint32_t)(bits_u8((uint32_t)n + 1) * 2 + 1 + c);
    }
    return bits;
}

// Gathers the directional modes' neighbour samples of the MB at (x, y)
// from the window (intra_pred.cpp:12-52) and the DC values
// (intra_pred.cpp:58-75: truncating division, 128 without neighbours).
__device__ __forceinline__ void gather_edges(const WinView &wv, int x, int y, IntraEdges &e)
{
    const int lane = lane_id();
    const bool has_top = y > 0, has_left = x > 0;   // the chroma block's neighbours exist alike
    if (lane < 48) {
        const int pl = lane >> 4, i = lane & 15, n = pl ? 8 : 16;
        if (i < n) {
            const int16_t *d = pl == 0 ? wv.y : pl == 1 ? wv.u : wv.v;
            const int pitch = pl ? wv.pc : wv.py;
            const int cx = pl ? (x >> 1) - wv.cx : x - wv.lx, cy = pl ? (y >> 1) - wv.cy : y - wv.ly;
            e.top[pl][i] = has_top ? d[(cy - 1) * pitch + cx + i] : (int16_t)128;
            e.left[pl][i] = has_left ? d[(cy + i) * pitch + cx - 1] : (int16_t)128;
            if (i == 0) {
                int16_t c;
                if (has_top && has_left) c = d[(cy - 1) * pitch + cx - 1];
                else if (has_top) c = d[(cy - 1) * pitch + cx];
                else if (has_left) c = d[cy * pitch + cx - 1];
                else c = 128;
                e.corner[pl] = c;
            }
        }
    }
    if (lane == 0) {
        e.has_top = has_top;
        e.has_left = has_left;
    }
    wave_sync();
    if (lane < 3) {
        const int n = lane ? 8 : 16;
        int sum = 0, count = 0;
        if (has_top) { for (int i = 0; i < n; i++) sum += e.top[lane][i]; count += n; }
        if (has_left) { for (int i = 0; i < n; i++) sum += e.left[lane][i]; count += n; }
        e.dc[lane] = count > 0 ? (int16_t)((sum + count / 2) / count) : (int16_t)128;
    }
    wave_sync();
}

// generate_intra_prediction (intra_pred.cpp:58-215) of `mode` into `out`.
__device__ __forceinline__ void intra_predict(int mode, const IntraEdges &e, int16_t *out)
{
    const int lane = lane_id();
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        const int pl = i < 256 ? 0 : (i < 320 ? 1 : 2);
        const int n = pl ? 8 : 16;
        const int k = pl ? (i - 256) & 63 : i;
        const int x = pl ? k & 7 : k & 15, y = pl ? k >> 3 : k >> 4;
        const int16_t *top = e.top[pl], *left = e.left[pl];
        // left[n - 1] .. left[0], corner, top[0] .. top[n - 1]
        auto ref = [&](int j) {
            return j < n ? (int)left[n - 1 - j] : j == n ? (int)e.corner[pl] : (int)top[j - n - 1];
        };
        int v;
        switch (mode) {
        case 0: v = e.dc[pl]; break;
        case 1: v = top[x]; break;
        case 2: v = left[y]; break;
        case 3: {
            const int i0 = x + y;
            v = (top[min(i0, n - 1)] + 2 * top[min(i0 + 1, n - 1)] + top[min(i0 + 2, n - 1)] + 2) >> 2;
        } break;
        case 4: {
            const int d = x - y + n;
            v = (ref(d - 1) + 2 * ref(d) + ref(d + 1) + 2) >> 2;
        } break;
        case 5:
        case 6: {
            const int base = mode == 5 ? x - (y >> 1) : (x >> 1) - y;
            const int odd = mode == 5 ? (y & 1) : (x & 1);
            const int d = min(max(base + n, 1), 2 * n - 1);
            v = odd ? (ref(d - 1) + 2 * ref(d) + ref(d + 1) + 2) >> 2 : (ref(d) + ref(d + 1) + 1) >> 1;
        } break;
        default: v = 128; break;
        }
        out[i] = (int16_t)v;
    }
    wave_sync();
}

// select_intra_mode_candidates (intra_pred.cpp:304-326): DC plus the pair of
// the luma edge gradients' family, packed as three bytes.
__device__ __forceinline__ uint32_t intra_candidates(const IntraEdges &e)
{
    int gh = 0, gv = 0;
    if (e.has_top)
        for (int i = 1; i < 16; i++) gh += abs(e.top[0][i] - e.top[0][i - 1]);
    if (e.has_left)
        for (int i = 1; i < 16; i++) gv += abs(e.left[0][i] - e.left[0][i - 1]);
    if (gv > gh * 3 / 2) return 0u | (2u << 8) | (6u << 16);
    if (gh > gv * 3 / 2) return 0u | (1u << 8) | (5u << 16);
    return 0u | (3u << 8) | (4u << 16);
}

__device__ __forceinline__ int luma_sad(const int16_t *a, const int16_t *b)
{
    const int lane = lane_id();
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) s += abs(a[4 * lane + k] - b[4 * lane + k]);
    return wave_sum(s);
}

__device__ __forceinline__ EncDesc desc_clear()
{
    // clear_block_desc (common.cpp:78-94)
    EncDesc d{};
    d.btype = EVXA_BLOCK_INTRA;
    d.imode = kIntraNone;
    return d;
}

__device__ __forceinline__ EncDesc desc_of(const SearchOut &r, int x, int y, int tgt, int pmode)
{
    EncDesc d = desc_clear();
    d.btype = r.btype;
    d.pmode = pmode;
    d.tgt = tgt;
    d.mvx = r.best_x - x;
    d.mvy = r.best_y - y;
    d.sp = r.sp_en;
    d.sa = r.sp_amt;
    d.si = r.sp_idx;
    return d;
}

__device__ __forceinline__ SearchOut motion_of(int x, int y, int mvx, int mvy, int sp, int sa, int si)
{
    return SearchOut{x + mvx, y + mvy, 0, 0, 0, 0.0f, sp, sa, si, 0, 0, 0, 0};
}

// The prediction a descriptor names into sh.mb.pred (encode.cpp:485-723;
// decode.cpp:35-433): a directional mode over the window's neighbours, zero
// for a plain intra block, an intra block copy from the window, or its
// reference(s) for inter blocks (the first reference with the target id;
// bidirectional blocks average both).
__device__ __forceinline__ void predict_desc(const DevJob &job, const EncDesc &d, int x, int y,
                                             const WinView &wv, const IntraEdges &edges, MbLds &mb)
{
    const int lane = lane_id();
    if (d.btype == EVXA_BLOCK_INTRA) {
        if (d.imode != kIntraNone) {
            intra_predict(d.imode, edges, mb.pred);
        } else {
            for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) mb.pred[i] = 0;
            wave_sync();
        }
        return;
    }
    const SearchOut r = motion_of(x, y, d.mvx, d.mvy, d.sp, d.sa, d.si);
    if (d.btype & EVXA_BLOCK_INTRA) {
        predict_mb(wv, r, mb.pred, mb.tmp);
        return;
    }
    auto ref_of = [&](int target) {
        int k = 0;
        for (int i = job.n_refs - 1; i >= 0; i--)
            if ((int)((job.target_ids >> (8 * i)) & 0xFF) == target) k = i;
        return k;
    };
    predict_mb(job.refs[ref_of(d.tgt)], r, mb.pred, mb.tmp);
    if (job.type == EVXA_FRAME_B && d.pmode == 2) {
        const SearchOut b = motion_of(x, y, d.mvx_b, d.mvy_b, d.sp_b, d.sa_b, d.si_b);
        predict_mb(job.refs[ref_of(d.tgt_b)], b, mb.pred_b, mb.tmp);
        for (int i = lane; i < EVXA_MB_SAMPLES; i += 64)
            mb.pred[i] = (int16_t)((mb.pred[i] + mb.pred_b[i] + 1) >> 1);
        wave_sync();
    }
}

// The prediction of the intra-frame block search's result into `out`, from
// what the search built: the half-pel planes around its integer best (the
// same samples interpolate_subpixel_6tap forms, averaged with the integer
// block for quarter-pel) or the integer block of the window.
__device__ __forceinline__ void predict_intra_result(const WinView &wv, const SearchOut &r,
                                                     const HalfPlanes<16> &hy,
                                                     const HalfPlanes<8> &hu,
                                                     const HalfPlanes<8> &hv, int16_t *out)
{
    int dx = 0, dy = 0;
    if (r.sp_en) frac_direction(r.sp_idx, dx, dy);
    const PlaneView vy = plane_view<16>(hy, dx, dy), vu = plane_view<8>(hu, dx, dy),
                    vv = plane_view<8>(hv, dx, dy);
    for (int i = lane_id(); i < EVXA_MB_SAMPLES; i += 64) {
        const bool luma = i < 256;
        const int k = luma ? i : (i - 256) & 63;
        const int col = luma ? k & 15 : k & 7, row = luma ? k >> 4 : k >> 3;
        const int16_t *pl = luma ? wv.y : (i < 320 ? wv.u : wv.v);
        const int pitch = luma ? wv.py : wv.pc;
        const int X = (luma ? r.best_x - wv.lx : (r.best_x >> 1) - wv.cx) + col;
        const int Y = (luma ? r.best_y - wv.ly : (r.best_y >> 1) - wv.cy) + row;
        const int full = pl[Y * pitch + X];
        int v = full;
        if (r.sp_en) {
            const PlaneView &pv = luma ? vy : (i < 320 ? vu : vv);
            const int half = pv.arr[pv.off + row * pv.stride + col];
            v = r.sp_amt ? (int16_t)((full + half + 1) >> 1) : half;
        }
        out[i] = (int16_t)v;
    }
    wave_sync();
}


// The bidirectional descriptor of a B MB from its two searches
// (encode.cpp:446-458; motion.cpp:702-840).
__device__ __forceinline__ EncDesc bidir_desc(const SearchOut &f, const SearchOut &b, int x, int y,
                                              int tgt0, int tgt1)
{
    EncDesc d = desc_clear();
    d.btype = EVXA_BLOCK_MOTION;
    d.pmode = 2;
    d.tgt = tgt0;
    d.mvx = f.best_x - x; d.mvy = f.best_y - y;
    d.sp = f.sp_en; d.sa = f.sp_amt; d.si = f.sp_idx;
    d.tgt_b = tgt1;
    d.mvx_b = b.best_x - x; d.mvy_b = b.best_y - y;
    d.sp_b = b.sp_en; d.sa_b = b.sp_amt; d.si_b = b.sp_idx;
    return d;
}

// compute_rdo_cost (encode.cpp:128-166): the candidate coded and
// reconstructed into `slot`, luma SSD of the reconstruction (an int32 that
// wraps) plus lambda times its estimated bits, fp32 multiply then add.
// With `rdo` false only the coding and reconstruction (cost 0).
__device__ __forceinline__ float rdo_cost(const DevJob &job, const CodeTables &t, EncDesc &d, int x,
                                          int y, const WinView &wv, const IntraEdges &edges,
                                          MbLds &mb, int16_t *co, int16_t *lv, int16_t *rec,
                                          bool rdo, bool predicted = false)
{
    if (!predicted) predict_desc(job, d, x, y, wv, edges, mb);
    code_mb(job, t, d, mb, co, lv, rec);
    if (!rdo) return 0.0f;
    const int lane = lane_id();
    uint32_t ssd = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int t = mb.src[4 * lane + k] - rec[4 * lane + k];
        ssd += (uint32_t)t * (uint32_t)t;
    }
    ssd = (uint32_t)wave_sum((int)ssd);
    const uint32_t rate = block_bits(d, rec);
    return __fadd_rn(__int2float_rn((int)ssd), __fmul_rn(job.lambda, __uint2float_rn(rate)));
}

// Stores the MB's reconstruction into the picture (and the picture's apron
// next to it for MBs on the border) with device-scope stores, and grows
// the picture's range record by it.
__device__ __forceinline__ void store_reconstruction(const EncJob &ej, int bx, int by,
                                                     const int16_t *rec)
{
    const DevJob &job = ej.job;
    const DevPic &pic = ej.pic;
    const int lane = lane_id();
    const int x = bx * 16, y = by * 16;
    int lo = 0x7FFFFFFF, hi = -0x7FFFFFFF - 1;
    for (int i = lane; i < 192; i += 64) {
        int16_t *dst;
        const int16_t *s;
        if (i < 128) {
            const int r = i >> 3, c = (i & 7) * 2;
            dst = const_cast<int16_t *>(pic.y) + (y + r) * pic.py + x + c;
            s = rec + r * 16 + c;
        } else {
            const int j = i - 128, pl = j >> 5, k = j & 31, r = k >> 2, c = (k & 3) * 2;
            dst = const_cast<int16_t *>(pl ? pic.v : pic.u) + ((y >> 1) + r) * pic.pc + (x >> 1) + c;
            s = rec + 256 + 64 * pl + r * 8 + c;
        }
        const int v0 = s[0], v1 = s[1];
        lo = min(lo, min(v0, v1));
        hi = max(hi, max(v0, v1));
        st_relaxed32(reinterpret_cast<int32_t *>(dst),
                     (int32_t)((uint32_t)(uint16_t)v0 | ((uint32_t)(uint16_t)v1 << 16)));
    }
    note_range(ej.spread, lo, hi);
    if (bx > 0 && bx < job.wb - 1 && by > 0 && by < job.hb - 1) return;
    // The apron part this MB owns: its rows and columns extended to the
    // apron's edge on each border side it touches, the interior excluded;
    // every value is the clamped sample, which lies in this MB.
    for (int pl = 0; pl < 3; pl++) {
        const int n = pl ? 8 : 16;
        const int X0 = pl ? x >> 1 : x, Y0 = pl ? y >> 1 : y;
        const int PW = pl ? job.width >> 1 : job.width, PH = pl ? job.height >> 1 : job.height;
        const int pitch = pl ? pic.pc : pic.py;
        int16_t *base = const_cast<int16_t *>(pl == 0 ? pic.y : pl == 1 ? pic.u : pic.v);
        const int16_t *src = rec + (pl == 0 ? 0 : pl == 1 ? 256 : 320);
        const int xl = bx == 0 ? -kApron : X0, xh = bx == job.wb - 1 ? PW + kApron : X0 + n;
        const int yl = by == 0 ? -kApron : Y0, yh = by == job.hb - 1 ? PH + kApron : Y0 + n;
        const int nw = (xh - xl) >> 1;
        const int total = nw * (yh - yl);
        for (int i = lane; i < total; i += 64) {
            const int r = yl + i / nw, c = xl + 2 * (i % nw);
            if (r >= Y0 && r < Y0 + n && c >= X0 && c < X0 + n) continue;
            const int sr = min(max(r, 0), PH - 1) - Y0;
            const int c0 = min(max(c, 0), PW - 1) - X0, c1 = min(max(c + 1, 0), PW - 1) - X0;
            const int v0 = src[sr * n + c0], v1 = src[sr * n + c1];
            st_relaxed32(reinterpret_cast<int32_t *>(base + r * pitch + c),
                         (int32_t)((uint32_t)(uint16_t)v0 | ((uint32_t)(uint16_t)v1 << 16)));
        }
    }
}

// One 3x3 level of the intra-frame block search (motion.cpp:348-379) over
// the window, in two passes and a closed-form fold.  The level's positions
// in scan order without its centre are slots 0..7: for the first scan
// (first = true: columns x - 16 .. x + 16, rows y - 32 .. y, step 16) the
// row above the MB, the row of the MB above, then (x - 16, y) and
// (x + 16, y); for the refinements the 3x3 ring of step `step` around the
// best.  The centre is never an effective candidate: the first scan's is
// the MB itself, which is unstable, and a refinement's is the running best,
// whose re-evaluation leaves the state as it is (same cost and distance; a
// copy would have ended the search) or is the MB itself (nothing taken
// yet).  Group g evaluates slot g, then slot g + 4.  The fold is the
// sequential rule's closed form: the first valid slot with MAD < T ends the
// level in copy state; otherwise the running best ends at the minimum cost,
// the first slot at that cost replaced by later equal-cost slots with a
// strictly smaller distance and SAD < 8192 (motion.cpp:125-178).  Returns
// true on a copy.
__device__ __forceinline__ bool intra_level(const WinView &wv, int pw, int ph, const Blk &s,
                                            const Geo &geo, int x, int y, int thresh, float lambda,
                                            bool first, int step, IntraState &st)
{
    const int lane = lane_id();
    const int g = lane >> 4;
    auto off = [&](int k, int &i, int &j) {
        if (first) {
            i = k < 6 ? (k % 3 - 1) * 16 : (k == 6 ? -16 : 16);
            j = k < 6 ? (k / 3) * 16 - 32 : 0;
        } else {
            i = slot_i_rt(k) * step;
            j = slot_j_rt(k) * step;
        }
    };
    const int base_x = st.x, base_y = st.y, max_x = pw - 16, max_y = ph - 16;
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int i, j;
        off(k, i, j);
        const int cx = base_x + i, cy = base_y + j;
        if (!intra_unstable(cx, cy, x, y) && cx >= 0 && cx <= max_x && cy >= 0 && cy <= max_y)
            mask |= 1u << k;
    }
    if (mask == 0) return false;
    int m[2][4];
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        const int k = g + 4 * pass;
        int i, j;
        off(k, i, j);
        const bool valid = ((mask >> k) & 1u) != 0;
        const int cx = valid ? base_x + i : x, cy = valid ? base_y + j : y;
        Blk r;
        load_cand(wv, geo, cx, cy, r);
        int sad, mad, satd;
        cand_metrics(false, s, r, sad, mad, satd);
        m[pass][0] = sad;
        m[pass][1] = mad;
        m[pass][2] = (int)cost_bits(sad, satd, cx - x, cy - y, lambda);
    }
    // Slot k's values into lane k (k < 8): slot g of pass p sits in lane 16 g.
    const int from = ((lane & 3) << 4) << 2;
    const bool second = (lane & 4) != 0;
    int ks, km, kc;
    {
        const int s0 = __builtin_amdgcn_ds_bpermute(from, m[0][0]);
        const int s1 = __builtin_amdgcn_ds_bpermute(from, m[1][0]);
        const int d0 = __builtin_amdgcn_ds_bpermute(from, m[0][1]);
        const int d1 = __builtin_amdgcn_ds_bpermute(from, m[1][1]);
        const int c0 = __builtin_amdgcn_ds_bpermute(from, m[0][2]);
        const int c1 = __builtin_amdgcn_ds_bpermute(from, m[1][2]);
        ks = second ? s1 : s0;
        km = second ? d1 : d0;
        kc = second ? c1 : c0;
    }
    const bool kv = lane < 8 && ((mask >> (lane & 7)) & 1u) != 0;
    auto dist = [&](int cx, int cy) { return (cx - x) * (cx - x) + (cy - y) * (cy - y); };
    auto take = [&](int k) {
        int i, j;
        off(k, i, j);
        st.x = base_x + i;
        st.y = base_y + j;
        st.sad = lane_value(ks, k);
        st.mad = lane_value(km, k);
        st.ssd = dist(st.x, st.y);
        st.cost = (uint32_t)lane_value(kc, k);
    };
    const uint32_t copym = (uint32_t)__ballot(kv && km < thresh) & 0xFFu;
    if (copym) {
        const int p = __builtin_ctz(copym);
        st.n_int += __builtin_popcount(mask & ((2u << p) - 1u));
        take(p);
        return true;
    }
    st.n_int += __builtin_popcount(mask);
    uint32_t ml = kv ? (uint32_t)kc : 0xFFFFFFFFu;
    ml = min(ml, (uint32_t)dpp<0xB1>((int)ml));
    ml = min(ml, (uint32_t)dpp<0x4E>((int)ml));
    ml = min(ml, (uint32_t)dpp<0x141>((int)ml));
    const uint32_t mn = min(st.cost, (uint32_t)lane_value((int)ml, 0));
    uint32_t eq = (uint32_t)__ballot(kv && (uint32_t)kc == mn) & 0xFFu;
    if (eq == 0) return false;                     // the running best stays
    bool have = st.cost == mn;
    while (eq) {
        const int q = __builtin_ctz(eq);
        eq &= eq - 1;
        int i, j;
        off(q, i, j);
        const int ex = base_x + i, ey = base_y + j;
        const int esad = lane_value(ks, q), essd = dist(ex, ey);
        if (!have || (essd < st.ssd && esad < kSadThreshold)) {
            st.x = ex; st.y = ey;
            st.sad = esad;
            st.mad = lane_value(km, q);
            st.ssd = essd;
            st.cost = mn;
        }
        have = true;
    }
    return false;
}

// calculate_intra_prediction (motion.cpp:485-565) over the window of the
// picture being reconstructed, as intra_search, with the sub-pel step of
// perform_intra_subpixel_motion_search (motion.cpp:408-448) evaluated the
// way inter_search does its own: the h, v and d half-pel planes around the
// integer best are built once from the window, every lane group takes two
// of the 8 directions (half and quarter of each), and the 16 candidates are
// folded in their order in closed form.  The acceptance rule is the same as
// the inter sub-pel step's (a candidate is taken iff its MAD is below the
// running best's in copy state; otherwise iff its MAD is below T, or its
// SAD is below the running best's and below 8192); only directions whose
// target is stable and inside the picture take part.
__device__ __forceinline__ SearchOut intra_search_win(const WinView &wv, int pw, int ph,
                                                      const Blk &s, const Geo &geo, int x, int y,
                                                      int thresh, float lambda, HalfPlanes<16> &hy,
                                                      HalfPlanes<8> &hu, HalfPlanes<8> &hv)
{
    const int lane = lane_id();
    const int g = lane >> 4;
    int a = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) a += abs(lo16(s.y[k])) + abs(hi16(s.y[k]));
    const int sad0 = lane_value(row_sum(a), 0);       // compute_block_sad(src), analysis.h:134-145
    IntraState st{x, y, sad0, kInt32Max, kInt32Max, __float_as_uint(__int2float_rn(sad0)), 0};
    if (!intra_level(wv, pw, ph, s, geo, x, y, thresh, lambda, true, kSearchRadius, st))
        for (int step = kSearchRadius / 2; step > 0; step >>= 1)
            if (intra_level(wv, pw, ph, s, geo, x, y, thresh, lambda, false, step, st)) break;

    int sp_en = 0, sp_amt = 0, sp_idx = 0, n_ax = 0, n_dg = 0;
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int tx = st.x + slot_i(k), ty = st.y + slot_j(k);
        if (!intra_unstable(tx, ty, x, y) && tx >= 0 && tx <= pw - 16 && ty >= 0 && ty <= ph - 16)
            mask |= 1u << k;
    }
    if (mask) {
        const int bx = st.x, by = st.y;
        build_hv<16, 1>(wv.y + (by - wv.ly) * wv.py + (bx - wv.lx), wv.py, hy);
        const int co = ((by >> 1) - wv.cy) * wv.pc + ((bx >> 1) - wv.cx);
        build_hv<8, 1>(wv.u + co, wv.pc, hu);
        build_hv<8, 1>(wv.v + co, wv.pc, hv);
        Blk full;
        load_cand(wv, geo, bx, by, full);
        wave_sync();
        build_d<16, 1>(hy);
        build_d<8, 1>(hu);
        build_d<8, 1>(hv);
        wave_sync();
        // Half and quarter SAD/MAD of candidate j = 2 * direction + quarter
        // in lane j (j < 16): group g evaluates directions g and g + 4.
        const int ia = slot_i(g), ja = slot_j(g), ib = slot_i(g + 4), jb = slot_j(g + 4);
        int m[8];
#pragma unroll
        for (int slot = 0; slot < 2; slot++) {
            const int di = slot ? ib : ia, dj = slot ? jb : ja;
            int sh, mh, sq, mq;
            subpel_metrics(false, s, full, plane_view<16>(hy, di, dj), plane_view<8>(hu, di, dj),
                           plane_view<8>(hv, di, dj), geo, sh, mh, sq, mq);
            m[slot * 4 + 0] = row_sum(sh);
            m[slot * 4 + 1] = row_max(mh);
            m[slot * 4 + 2] = row_sum(sq);
            m[slot * 4 + 3] = row_max(mq);
        }
        const int j = lane & 15, jd = j >> 1, jq = j & 1;
        const int from = ((jd & 3) << 4) << 2;
        int pulled[8];
#pragma unroll
        for (int i = 0; i < 8; i++) pulled[i] = __builtin_amdgcn_ds_bpermute(from, m[i]);
        const int base = (jd >= 4 ? 4 : 0) + (jq ? 2 : 0);
        const int jsad = base == 0 ? pulled[0] : base == 2 ? pulled[2] : base == 4 ? pulled[4] : pulled[6];
        const int jmad = base == 0 ? pulled[1] : base == 2 ? pulled[3] : base == 4 ? pulled[5] : pulled[7];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool valid = ((mask >> k) & 1u) != 0;
            const bool diag = slot_i(k) != 0 && slot_j(k) != 0;
            n_dg += (valid && diag) ? 1 : 0;
            n_ax += (valid && !diag) ? 1 : 0;
        }
        const bool vj = lane < 16 && ((mask >> (lane >> 1)) & 1u) != 0;
        auto row_min = [](int v) {
            v = min(v, dpp<0xB1>(v));
            v = min(v, dpp<0x4E>(v));
            v = min(v, dpp<0x141>(v));
            v = min(v, dpp<0x140>(v));
            return lane_value(v, 0);
        };
        int win = -1;
        if (st.mad < thresh) {
            const int m2 = row_min(vj ? jmad : kInt32Max);
            if (m2 < st.mad) win = (int)__builtin_ctzll(__ballot(vj && jmad == m2));
        } else {
            const uint64_t cm = __ballot(vj && jmad < thresh);
            if (cm) {
                const int js = (int)__builtin_ctzll(cm);
                const bool after = vj && lane >= js;
                const int m2 = row_min(after ? jmad : kInt32Max);
                win = (int)__builtin_ctzll(__ballot(after && jmad == m2));
            } else {
                const bool ok = vj && jsad < kSadThreshold;
                const int m1 = row_min(ok ? jsad : kInt32Max);
                if (m1 < st.sad) win = (int)__builtin_ctzll(__ballot(ok && jsad == m1));
            }
        }
        if (win >= 0) {
            sp_en = 1;
            sp_amt = win & 1;
            sp_idx = win >> 1;
            st.sad = lane_value(jsad, win);
            st.mad = lane_value(jmad, win);
        }
        wave_sync();            // the planes are free again
    }
    SearchOut o;
    o.best_x = st.x; o.best_y = st.y;
    o.sad = st.sad; o.mad = st.mad; o.ssd = st.ssd;
    o.cost = __uint_as_float(st.cost);
    o.sp_en = sp_en; o.sp_amt = sp_amt; o.sp_idx = sp_idx;
    o.btype = EVXA_BLOCK_INTRA;
    if (st.x != x || st.y != y || sp_en) o.btype |= EVXA_BLOCK_MOTION;
    if (st.mad < thresh) o.btype |= EVXA_BLOCK_COPY;
    o.n_int = st.n_int; o.n_ax = n_ax; o.n_dg = n_dg;
    return o;
}

// One MB of the in-loop encode (see the section comment), by the four waves
// of the workgroup.  Wave 0 waits for the MB's dependencies; then all
// waves stage the window of the picture being reconstructed, and the
// independent parts run side by side, each followed by the coding of what
// it found: wave 0 the intra-frame block search, wave 1 the (forward) inter
// search, wave 2 the backward one (B) or, in P frames, the plain intra block
// and the third mode candidate, wave 3 the directional intra modes (P: the
// first two candidates; B: also the bidirectional average once both
// searches are in).  Every wave then makes the same selection from what
// they left in LDS; wave 0 takes the RDO choice among the coded candidates
// (P) or the selected one (I, B; a B frame's intra-frame search candidate
// is coded only then, if it won), stores the MB and publishes its word.
// Inlined into the kernel, so that `sh` is known to be LDS.
__device__ __forceinline__ void encode_mb(const EncArgs &a, const EncJob &ej, int mb, EncLds &sh)
{
    const DevJob &job = ej.job;
    const int lane = lane_id();
    const int wv_id = __builtin_amdgcn_readfirstlane(wave_id());
    const int bx = mb % job.wb, by = mb / job.wb;
    const int x = bx * 16, y = by * 16;
    const int W = job.width, H = job.height;
    const Geo geo = lane_geometry();
    MbLds &mine = sh.mb[wv_id];

    // While the dependencies finish: every wave loads the source block; the
    // searching waves load their reference's block at (x, y) and stage its
    // window.
    Blk s, pre;
    load_cand(job.src, geo, x, y, s);
    store_src(s, geo, mine.src);
    const int my_ref = wv_id - 1;                  // wave 1: reference 0, wave 2: reference 1
    const bool searcher = my_ref >= 0 && my_ref < job.n_refs;
    WinPos wk{};
    bool narrowk = false;
    if (searcher) {
        narrowk = __builtin_amdgcn_readfirstlane(narrow_pair(job.src.wide, job.refs[my_ref].wide)) != 0;
        load_cand(job.refs[my_ref], geo, x, y, pre);
        wk = stage_window<1>(job.refs[my_ref], x, y, kReachMax, sh.sp[my_ref]);
    }

    if (wv_id == 0) {
        uint32_t p0 = 0, p1 = 0;
        if (job.type == EVXA_FRAME_P && job.prev) {
            const uint32_t *prv = reinterpret_cast<const uint32_t *>(job.prev) + (size_t)mb * 16;
            p0 = prv[0];
            p1 = prv[1];
        }
        // Wait for MB (bx - 1, by) and MB (min(bx + lag, wb - 1), by - 1);
        // the rest of what this MB reads was done before them.  Lanes 0..3
        // poll the MV predictor's neighbours (left, top, top-right,
        // top-left), lane 4 the MB the lag waits for, so the words the
        // predictor reads are at hand when the wait ends.
        const bool has_l = bx > 0, has_t = by > 0, has_tr = has_t && bx + 1 < job.wb;
        const bool has_tl = has_t && has_l;
        const int nidx = lane == 0 ? (has_l ? mb - 1 : -1)
                       : lane == 1 ? (has_t ? mb - job.wb : -1)
                       : lane == 2 ? (has_tr ? mb - job.wb + 1 : -1)
                       : lane == 3 ? (has_tl ? mb - job.wb - 1 : -1)
                       : lane == 4 ? (has_t ? (by - 1) * job.wb + min(bx + kEncodeLag, job.wb - 1) : -1)
                       : -1;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        int nap = 1;
        bool timed_out = false;
        uint64_t word = 0;
        for (;;) {
            if (nidx >= 0 && (word >> 63) == 0) word = ld_relaxed64(job.mvw + nidx);
            if (__ballot(nidx >= 0 && (word >> 63) == 0) == 0) break;
            if (__builtin_amdgcn_s_memrealtime() - t0 > kEncWaitTicks) { timed_out = true; break; }
            for (int k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(1);
            nap = min(nap * 2, kMaxNap);
        }
        if (timed_out && lane == 0) atomicOr(&a.counters[kCounterError], kErrWaitTimeout);
        int pmx = 0, pmy = 0;
        if (job.type != 0)
            mv_predictor(job, bx, by, lane_u64(word, 0), lane_u64(word, 1), lane_u64(word, 2),
                         lane_u64(word, 3), p0, p1, pmx, pmy);
        if (lane == 0) {
            sh.x.pmx = pmx;
            sh.x.pmy = pmy;
            sh.x.cands = -1;                      // wave 3 publishes this MB's mode candidates
            sh.x.preds = 0;
        }
    }
    __syncthreads();                              // the dependencies are done

    // The window of the picture being reconstructed, origin moved inside the
    // apron at the borders (its reads are then the clamped ones); the four
    // waves share the copy.
    const int x0 = min(max(x - 34, -kApron), W + kApron - kIwLS);
    const int y0 = min(max(y - 50, -kApron), H + kApron - kIwLR);
    const int cx0 = min(max((x >> 1) - 20, -kApron), (W >> 1) + kApron - kIwCS);
    const int cy0 = min(max((y >> 1) - 27, -kApron), (H >> 1) + kApron - kIwCR);
    dma_rows<kIwLS / 2, kEncWaves, kLoadDevice>(
        reinterpret_cast<const uint32_t *>(ej.pic.y + y0 * ej.pic.py + x0), ej.pic.py / 2, kIwLR,
        reinterpret_cast<uint32_t *>(sh.iw.y));
    dma_rows<kIwCS / 2, kEncWaves, kLoadDevice>(
        reinterpret_cast<const uint32_t *>(ej.pic.u + cy0 * ej.pic.pc + cx0), ej.pic.pc / 2, kIwCR,
        reinterpret_cast<uint32_t *>(sh.iw.u));
    dma_rows<kIwCS / 2, kEncWaves, kLoadDevice>(
        reinterpret_cast<const uint32_t *>(ej.pic.v + cy0 * ej.pic.pc + cx0), ej.pic.pc / 2, kIwCR,
        reinterpret_cast<uint32_t *>(sh.iw.v));
    window_ready();
    const WinView wv{sh.iw.y, sh.iw.u, sh.iw.v, kIwLS, kIwCS, x0, y0, cx0, cy0};
    const int thresh = job.thresh;
    const float lambda = job.lambda;
    const int pmx = sh.x.pmx, pmy = sh.x.pmy;

    // The searches, each followed by the coding of its result, which is a
    // candidate of the decision whatever the others find (P frames' RDO
    // codes every candidate; I and B frames code the winner, which is one of
    // these unless it is the bidirectional average).  Costs are those of
    // compute_rdo_cost (encode.cpp:128-166); only P frames use them.
    const bool rdo = job.type == EVXA_FRAME_P;
    const int tgt0 = job.target_ids & 0xFF, tgt1 = (job.target_ids >> 8) & 0xFF;
    auto code_slot = [&](EncDesc d, int slot, bool predicted = false) {
        int16_t *co = a.coscratch + ((size_t)blockIdx.x * kEncCands + slot) * EVXA_MB_SAMPLES;
        const float cost = rdo_cost(job, sh.tables, d, x, y, wv, sh.edges, mine, co, sh.lv[slot],
                                    sh.rec[slot], rdo, predicted);
        if (lane == 0) {
            sh.x.cost[slot] = cost;
            sh.x.qp[slot] = d.qp;
            sh.x.var[slot] = d.var;
        }
    };
    // The intra-frame search's candidate, coded from the search's planes
    // (they stay in LDS until the MB is stored).
    auto code_intra = [&](const SearchOut &ri) {
        const EncDesc di = desc_of(ri, x, y, 0, 0);
        if (di.btype == EVXA_BLOCK_INTRA) {
            code_slot(di, kSlotIntra);                   // nothing found: the plain intra block
        } else {
            predict_intra_result(wv, ri, sh.hy, sh.hu, sh.hv, mine.pred);
            code_slot(di, kSlotIntra, true);
        }
    };
    // B frames code it only if it wins (about one MB in twenty): coded
    // speculatively it would make wave 0, which searches longest, the last
    // to reach the barrier.
    const bool defer_intra = job.type == EVXA_FRAME_B;
    if (wv_id == 0) {
        // calculate_intra_prediction, for every frame type (encode.cpp:175,
        // 362, 380).
        const SearchOut ri = intra_search_win(wv, W, H, s, geo, x, y, thresh, lambda, sh.hy,
                                              sh.hu, sh.hv);
        if (lane == 0) sh.x.intra = ri;
        if (!defer_intra) code_intra(ri);
    } else if (searcher) {
        const SearchOut r = inter_search<1, true>(job.refs[my_ref], W, H, narrowk, s, geo, pre, wk,
                                                  x, y, thresh, lambda, pmx, pmy, sh.sp[my_ref]);
        // The search's prediction, from its planes (the bidirectional
        // candidate of a B frame averages both searches', motion.cpp:729-776).
        predict_inter_result(job.refs[my_ref], r, sh.sp[my_ref], mine.pred_b);
        if (lane == 0) {
            sh.x.inter[my_ref] = r;
            __hip_atomic_fetch_or(&sh.x.preds, 1 << my_ref, __ATOMIC_RELEASE,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        const EncDesc d = desc_of(r, x, y, my_ref ? tgt1 : tgt0, my_ref);
        // Coded from that prediction when the descriptor names this search's
        // reference (the first with its target id, as decode_block finds it).
        const bool same_ref = my_ref == 0 || tgt1 != tgt0;
        if (same_ref) {
            for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) mine.pred[i] = mine.pred_b[i];
            wave_sync();
        }
        code_slot(d, my_ref ? kSlotBwd : kSlotInter, same_ref);
    } else if (wv_id == kEncWaves - 1) {
        // The directional intra candidate (encode.cpp:254-272, 446-468;
        // select_best_intra_mode, intra_pred.cpp:217-302: the first
        // candidate with the smallest luma SAD).
        gather_edges(wv, x, y, sh.edges);
        const uint32_t cands = intra_candidates(sh.edges);
        int dir_sad = kInt32Max, dir_mode = 0;
        for (int c = 0; c < 3; c++) {
            const int m = (cands >> (8 * c)) & 0xFF;
            intra_predict(m, sh.edges, mine.pred);
            const int sd = luma_sad(mine.src, mine.pred);
            if (sd < dir_sad) { dir_sad = sd; dir_mode = m; }
            wave_sync();
        }
        if (lane == 0) {
            sh.x.dir_sad = dir_sad;
            sh.x.dir_mode = dir_mode;
            __hip_atomic_store(&sh.x.cands, (int32_t)cands, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // P: the first two mode candidates; I and B: the chosen mode.
        for (int c = 0; c < (rdo ? 2 : 1); c++) {
            EncDesc d = desc_clear();
            d.imode = rdo ? (int)((cands >> (8 * c)) & 0xFF) : dir_mode;
            code_slot(d, kSlotMode + c);
        }
        if (job.type == EVXA_FRAME_B) {
            // B: the bidirectional candidate too, once both searches'
            // predictions are in LDS.
            int32_t ready;
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            while ((ready = __hip_atomic_load(&sh.x.preds, __ATOMIC_ACQUIRE,
                                              __HIP_MEMORY_SCOPE_WORKGROUP)) != 3) {
                if (__builtin_amdgcn_s_memrealtime() - t0 > kEncWaitTicks) {
                    if (lane == 0) atomicOr(&a.counters[kCounterError], kErrWaitTimeout);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            const EncDesc d = bidir_desc(sh.x.inter[0], sh.x.inter[1], x, y, tgt0, tgt1);
            const bool own = tgt1 != tgt0;   // else decode_block takes both from the first reference
            // The average of the searches' predictions and its SAD against
            // the source over Y, U and V (motion.cpp:778-813), which the
            // selection compares.
            int bsad = 0;
            for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
                const int av = (sh.mb[1].pred_b[i] + sh.mb[2].pred_b[i] + 1) >> 1;
                bsad += abs(mine.src[i] - av);
                if (own) mine.pred[i] = (int16_t)av;
            }
            bsad = wave_sum(bsad);
            if (lane == 0) sh.x.bsad = bsad;
            wave_sync();
            code_slot(d, kSlotBidir, own);
        }
    } else if (rdo) {
        // The other P candidates: the plain intra block (unused when the
        // winner is of the intra-default kind) and the third mode, which
        // needs wave 3's neighbour samples.
        code_slot(desc_clear(), kSlotPlain);
        int32_t cw;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while ((cw = __hip_atomic_load(&sh.x.cands, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) < 0) {
            if (__builtin_amdgcn_s_memrealtime() - t0 > kEncWaitTicks) {
                if (lane == 0) atomicOr(&a.counters[kCounterError], kErrWaitTimeout);
                cw = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        EncDesc d = desc_clear();
        d.imode = (int)(((uint32_t)__builtin_amdgcn_readfirstlane(cw) >> 16) & 0xFF);
        code_slot(d, kSlotMode + 2);
    }
    __syncthreads();

    // The selection, made alike by every wave from what the waves left.
    const SearchOut ri = sh.x.intra;
    EncDesc best = desc_of(ri, x, y, 0, 0);
    int best_sad = ri.sad, best_slot = kSlotIntra;
    int n_int = ri.n_int, n_ax = ri.n_ax, n_dg = ri.n_dg;
    if (job.type == EVXA_FRAME_P) {
        // encode.cpp:183-210: the copy state decides first, then the SAD.
        const SearchOut r = sh.x.inter[0];
        n_int += r.n_int; n_ax += r.n_ax; n_dg += r.n_dg;
        const bool inter_copy = (r.btype & EVXA_BLOCK_COPY) != 0;
        const bool best_copy = (best.btype & EVXA_BLOCK_COPY) != 0;
        if (inter_copy != best_copy) {
            if (inter_copy) { best = desc_of(r, x, y, tgt0, 0); best_sad = r.sad; best_slot = kSlotInter; }
        } else if (r.sad < best_sad) {
            best = desc_of(r, x, y, tgt0, 0);
            best_sad = r.sad;
            best_slot = kSlotInter;
        }
    } else if (job.type == EVXA_FRAME_B) {
        // encode.cpp:386-458: forward, backward, then their average.
        const SearchOut f = sh.x.inter[0], b = sh.x.inter[1];
        n_int += f.n_int + b.n_int;
        n_ax += f.n_ax + b.n_ax;
        n_dg += f.n_dg + b.n_dg;
        if (f.sad < best_sad) { best = desc_of(f, x, y, tgt0, 0); best_sad = f.sad; best_slot = kSlotInter; }
        if (b.sad < best_sad) { best = desc_of(b, x, y, tgt1, 1); best_sad = b.sad; best_slot = kSlotBwd; }
        const int bsad = sh.x.bsad;
        if (bsad + __float2int_rz(__fmul_rn(lambda, 8.0f)) < best_sad) {
            best = bidir_desc(f, b, x, y, tgt0, tgt1);
            best_sad = bsad;
            best_slot = kSlotBidir;
        }
    }
    if (best_sad == kInt32Max && wv_id == 0 && lane == 0)
        atomicOr(&a.counters[kCounterError], kErrSearchFailed);
    const uint32_t cands = (uint32_t)sh.x.cands;
    if (sh.x.dir_sad < best_sad) {
        best = desc_clear();
        best.imode = sh.x.dir_mode;
        best_sad = sh.x.dir_sad;
        best_slot = kSlotMode;
        if (rdo)
            for (int c = 0; c < 3; c++)
                if ((int)((cands >> (8 * c)) & 0xFF) == best.imode) { best_slot = kSlotMode + c; break; }
    }
    if (wv_id != 0) return;
    if (defer_intra && best_slot == kSlotIntra) {
        code_intra(ri);
        wave_sync();                              // lane 0's cost, QP and variance
    }

    // RDO for P frames (encode.cpp:274-348, EVX_ENABLE_RDO_COPY_EVAL and
    // EVX_ENABLE_RDO_INTRA_MODES): the SAD winner, the plain intra block
    // unless the winner is one of the intra-default kind, then the three
    // mode candidates; a later candidate wins only at a strictly lower cost.
    int slot = best_slot;
    float best_cost = 0.0f;
    if (rdo) {
        best_cost = sh.x.cost[best_slot];
        if (best.btype != EVXA_BLOCK_INTRA && sh.x.cost[kSlotPlain] < best_cost) {
            best_cost = sh.x.cost[kSlotPlain];
            slot = kSlotPlain;
            best = desc_clear();
        }
        for (int c = 0; c < 3; c++)
            if (sh.x.cost[kSlotMode + c] < best_cost) {
                best_cost = sh.x.cost[kSlotMode + c];
                slot = kSlotMode + c;
                best = desc_clear();
                best.imode = (cands >> (8 * c)) & 0xFF;
            }
    }
    best.qp = sh.x.qp[slot];
    best.var = sh.x.var[slot];

    // Outputs: the reconstruction into the picture, coefficients, levels,
    // the record; then the MB's word, once every store has completed.
    store_reconstruction(ej, bx, by, sh.rec[slot]);
    // The MB's word goes out once the reconstruction's stores have completed
    // (what later MBs read of this one); coefficients, levels and the record
    // are read after the launch.
    stores_done();
    if (lane == 0)
        st_relaxed64(job.mvw + mb, pred_word((best.btype & EVXA_BLOCK_MOTION) != 0,
                                             mv_word(best.mvx, best.mvy)));
    int16_t *co = job.coeffs + (size_t)mb * EVXA_MB_SAMPLES;
    int16_t *lv = job.levels + (size_t)mb * EVXA_MB_SAMPLES;
    const int16_t *cos = a.coscratch + ((size_t)blockIdx.x * kEncCands + slot) * EVXA_MB_SAMPLES;
    for (int i = lane; i < EVXA_MB_SAMPLES; i += 64) {
        co[i] = cos[i];
        lv[i] = sh.lv[slot][i];
    }
    uint32_t w[16];
    w[0] = (uint32_t)best.btype | ((uint32_t)best.pmode << 8) | ((uint32_t)best.tgt << 16) |
           ((uint32_t)best.tgt_b << 24);
    w[1] = mv_word(best.mvx, best.mvy);
    w[2] = mv_word(best.mvx_b, best.mvy_b);
    w[3] = (uint32_t)best.sp | ((uint32_t)best.sa << 8) | ((uint32_t)best.si << 16) |
           ((uint32_t)best.sp_b << 24);
    w[4] = (uint32_t)best.sa_b | ((uint32_t)best.si_b << 8) | ((uint32_t)best.qp << 16) |
           ((uint32_t)best.imode << 24);
    w[5] = (uint32_t)(uint16_t)(int16_t)best.var;
    w[6] = (uint32_t)best_sad;
    w[7] = 0;
    w[8] = fbits(best_cost);
    w[9] = 0; w[10] = 0; w[11] = 0;
    w[12] = (uint32_t)n_int;
    w[13] = (uint32_t)n_ax | ((uint32_t)n_dg << 16);
    w[14] = ((best.btype & EVXA_BLOCK_COPY) ? 0u : 1u) | (best.pmode == 2 ? 2u : 0u) | 4u;
    w[15] = 0;
    write_words(reinterpret_cast<uint32_t *>(job.records + mb), w, 16);
}

// The persistent in-loop encode kernel: a workgroup of four waves takes the
// next MB of the launch's dependency order (the MBs of all its frames) until
// none is left.  A MB waits only on
// MBs earlier in the order, which running workgroups hold, so the launch
// cannot deadlock whatever the number of resident workgroups.
__global__ __launch_bounds__(64 * kEncWaves) void encode_kernel(const EncArgs a)
{
    __shared__ EncLds sh;
    __shared__ int32_t s_word;
    load_code_tables(sh.tables);        // read after the claim's barrier
    for (;;) {
        if (threadIdx.x == 0) s_word = atomicAdd(&a.counters[0], 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_word);
        __syncthreads();
        if (t >= a.n_tickets) break;
        const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.order[t]);
        encode_mb(a, a.jobs[e >> kMbBits], (int)(e & kMbMask), sh);
    }
}


// ------------------------------------------------- allocation-free path --
//
// The scratch of an analysis launch -- job descriptor, work list, counters
// and predictor words -- for callers that cannot allocate: one array per
// GPU, alive for the process.  Calls that overlap on one GPU share it and
// race; every call clears what must start at zero.  Host code reaches it
// only through the address hipGetSymbolAddress gives (noalloc_scratch).
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Copies a job descriptor (a kernel argument) into the scratch.
__global__ void set_job_kernel(DevJob job, DevJob *dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = job;
}

// The work buffer of one frame as the wavefront lists it (one list,
// anti-diagonal order v = bx + 2 by, MBs of a diagonal by row, a MB's
// references adjacent): MB (bx, by) is preceded by the MBs (x, y) with
// x + 2 y < v -- min(max(v - 2 y, 0), wb) of them on row y -- and by those
// of diagonal v on rows y < by.  One thread per MB writes its tickets.
__global__ void wave_list_kernel(int32_t wb, int32_t hb, int32_t n_refs, uint32_t *work)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = wb * hb;
    if (i == 0) {
        work[0] = 1u;
        work[1] = 0u;
        work[2] = (uint32_t)(n * n_refs);
    }
    if (i >= n) return;
    const int bx = i % wb, by = i / wb, v = bx + 2 * by;
    int before = 0;
    for (int y = 0; y < hb; y++) {
        const int c = v - 2 * y;
        before += c <= 0 ? 0 : (c < wb ? c : wb);
        if (y < by && c >= 0 && c < wb) before++;
    }
    uint32_t *t = work + kWorkHeader + (size_t)before * n_refs;
    for (int r = 0; r < n_refs; r++) t[r] = ((uint32_t)r << kRefShift) | (uint32_t)i;
}

}  // namespace

hipError_t launch_reset(int32_t *words, int32_t n, hipStream_t st)
{
    const int blocks = min((n + 255) / 256, 1024);
    hipLaunchKernelGGL(reset_kernel, dim3(max(blocks, 1)), dim3(256), 0, st, words, n);
    return hipGetLastError();
}

hipError_t launch_analyze(const DevJob *jobs, const uint32_t *work, int32_t *counters,
                          int32_t grid, int32_t waves, hipStream_t st)
{
    if (waves == 2)
        hipLaunchKernelGGL(analyze_kernel<2>, dim3(grid), dim3(128), 0, st, jobs, work, counters);
    else
        hipLaunchKernelGGL(analyze_kernel<1>, dim3(grid), dim3(64), 0, st, jobs, work, counters);
    return hipGetLastError();
}

hipError_t launch_sweep(const DevJob *jobs, const uint32_t *work, int32_t *counters,
                        const SweepArgs &sa, int32_t grid, int32_t waves, int32_t n_final,
                        hipStream_t st)
{
    if (waves == 2)
        hipLaunchKernelGGL(sweep_kernel<2>, dim3(grid), dim3(128), 0, st, jobs, work, counters, sa);
    else
        hipLaunchKernelGGL(sweep_kernel<1>, dim3(grid), dim3(64), 0, st, jobs, work, counters, sa);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n_final <= 0) return e;
    // Enough workgroups for every MB at once up to a few rounds of the
    // device; each wave then strides over the list.
    const int32_t fgrid = std::min((n_final + kFinalMbs - 1) / kFinalMbs, 8192);
    hipLaunchKernelGGL(final_kernel, dim3(fgrid), dim3(64 * kFinalMbs), 0, st, jobs, work, counters);
    return hipGetLastError();
}

hipError_t launch_recon(const DevJob &job, const int16_t *levels, const DevPic &dst,
                        int32_t *wide, hipStream_t st)
{
    const int n = job.wb * job.hb;
    hipLaunchKernelGGL(recon_kernel, dim3((n + kReconMbs - 1) / kReconMbs), dim3(64 * kReconMbs),
                       0, st, job, levels, dst, wide);
    return hipGetLastError();
}

hipError_t launch_apron3(int16_t *y0, int16_t *u0, int16_t *v0, int32_t py, int32_t pc, int32_t w,
                         int32_t h, hipStream_t st)
{
    const int n = 2 * kApron * (w + 2 * kApron) + 2 * kApron * h;
    const int blocks = min((n + 255) / 256, 512);
    hipLaunchKernelGGL(apron3_kernel, dim3(blocks, 3), dim3(256), 0, st, y0, u0, v0, py, pc, w, h);
    return hipGetLastError();
}

hipError_t launch_fold_range(const int32_t *spread, int32_t *range, hipStream_t st)
{
    hipLaunchKernelGGL(fold_range_kernel, dim3(1), dim3(64), 0, st, spread, range);
    return hipGetLastError();
}

hipError_t launch_range(const int16_t *plane0, int32_t pitch, int32_t w, int32_t h, int32_t *wide,
                        hipStream_t st)
{
    const int blocks = min((w * h + 255) / 256, 1024);
    hipLaunchKernelGGL(range_kernel, dim3(max(blocks, 1)), dim3(256), 0, st, plane0, pitch, w, h,
                       wide);
    return hipGetLastError();
}

hipError_t launch_deblock(int16_t *const src[3], const int32_t sp[3], int16_t *const dst[3],
                          const int32_t dp[3], int32_t w, int32_t h,
                          const evxa_block_record *records, int32_t *wide, hipStream_t st)
{
    DeblockArgs a{};
    int32_t n = 0;
    for (int p = 0; p < 3; p++) {
        const int pw = p ? w / 2 : w, ph = p ? h / 2 : h;
        a.src[p] = src[p];
        a.dst[p] = dst[p];
        a.sp[p] = sp[p];
        a.dp[p] = dp[p];
        a.first[p] = n;
        n += ((pw + kDbTw - 1) / kDbTw) * ((ph + kDbTh - 1) / kDbTh);
    }
    a.first[3] = n;
    a.w = w;
    a.h = h;
    a.wb = w / 16;
    a.hb = h / 16;
    hipLaunchKernelGGL(deblock_kernel, dim3(n), dim3(kDbThreads), 0, st, a, records, wide);
    return hipGetLastError();
}

hipError_t launch_copy3(const int16_t *y, const int16_t *u, const int16_t *v, int32_t py, int32_t pc,
                        int16_t *out, int32_t w, int32_t h, int32_t *wide, hipStream_t st)
{
    const int blocks = min((w / 8 * h + 255) / 256, 1024);
    hipLaunchKernelGGL(copy3_kernel, dim3(blocks, 3), dim3(256), 0, st, y, u, v, py, pc, out, w, h,
                       wide);
    return hipGetLastError();
}

hipError_t launch_sao_decide(const int16_t *rec_y, const int16_t *rec_u, const DevPic &src,
                             int32_t w, int32_t h, float lambda, uint32_t *scratch,
                             uint64_t *words, evxa_sao_info *table, int32_t *counters,
                             hipStream_t st)
{
    SaoArgs a{};
    a.rec_y = rec_y;
    a.rec_u = rec_u;
    a.src_y = src.y;
    a.src_u = src.u;
    a.spy = src.py;
    a.spc = src.pc;
    a.w = w;
    a.h = h;
    a.wb = w / 16;
    a.hb = h / 16;
    a.lambda = lambda;
    const int n = a.wb * a.hb;
    a.own_l = scratch;
    a.own_c = scratch + n;
    a.cur = reinterpret_cast<float *>(scratch + 2 * n);
    a.mc = reinterpret_cast<float *>(scratch + 3 * n);
    a.words = words;
    a.table = table;
    a.counters = counters;
    hipLaunchKernelGGL(sao_own_kernel, dim3(n), dim3(64), 0, st, a);
    hipLaunchKernelGGL(sao_cost_kernel, dim3(n), dim3(64), 0, st, a);
    hipLaunchKernelGGL(sao_merge_kernel, dim3(a.hb), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sao_apply(int16_t *const rec[3], int16_t *const dst[3], const int32_t dp[3],
                            int32_t w, int32_t h, const evxa_sao_info *table, int32_t *wide,
                            int32_t *error, hipStream_t st)
{
    SaoApplyArgs a{};
    for (int p = 0; p < 3; p++) {
        a.rec[p] = rec[p];
        a.dst[p] = dst[p];
        a.dp[p] = dp[p];
    }
    a.w = w;
    a.h = h;
    a.wb = w / 16;
    a.table = table;
    a.wide = wide;
    a.error = error;
    const int blocks = min((w * h + 255) / 256, 2048);
    hipLaunchKernelGGL(sao_apply_kernel, dim3(blocks, 3), dim3(256), 0, st, a);
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

hipError_t launch_intra_blocks(DevPic src, DevPic pic, int32_t w, int32_t h, int32_t quality,
                               float lambda, int32_t n, const int32_t *params,
                               evxa_search_result *out, hipStream_t st)
{
    hipLaunchKernelGGL(intra_blocks_kernel, dim3(n), dim3(64), 0, st, src, pic, w, h, quality,
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

hipError_t noalloc_scratch(void **ptr)
{
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t launch_set_job(const DevJob &job, DevJob *dst, hipStream_t st)
{
    hipLaunchKernelGGL(set_job_kernel, dim3(1), dim3(64), 0, st, job, dst);
    return hipGetLastError();
}

hipError_t launch_wave_list(int32_t wb, int32_t hb, int32_t n_refs, uint32_t *work, hipStream_t st)
{
    const int n = wb * hb;
    hipLaunchKernelGGL(wave_list_kernel, dim3((n + 255) / 256), dim3(256), 0, st, wb, hb, n_refs, work);
    return hipGetLastError();
}

hipError_t launch_encode(const EncJob *jobs, const uint32_t *order, int32_t n_tickets,
                         int32_t *counters, int16_t *coscratch, int32_t grid, hipStream_t st)
{
    EncArgs a;
    a.jobs = jobs;
    a.order = order;
    a.counters = counters;
    a.n_tickets = n_tickets;
    a.coscratch = coscratch;
    hipLaunchKernelGGL(encode_kernel, dim3(grid), dim3(64 * kEncWaves), 0, st, a);
    return hipGetLastError();
}

int32_t encode_resident_groups(int32_t device)
{
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        n_cu <= 0)
        n_cu = 256;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, encode_kernel, 64 * kEncWaves, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return n_cu * per_cu;
}

int32_t analyze_resident_groups(int32_t device, int32_t waves, bool sweep)
{
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        n_cu <= 0)
        n_cu = 256;
    int per_cu = 0;
    hipError_t e;
    if (sweep)
        e = waves == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_kernel<2>, 128, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_kernel<1>, 64, 0);
    else
        e = waves == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, analyze_kernel<2>, 128, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, analyze_kernel<1>, 64, 0);
    if (e != hipSuccess || per_cu <= 0) per_cu = waves == 2 ? 2 : 4;
    return n_cu * per_cu;
}

}  // namespace evxa
