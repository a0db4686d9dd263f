// agg_kernels.hip -- fixed-pivot aggregation source term on gfx950.
//
// Reference semantics: src/aggregation_kernels.cuh:145-239 (per pair (j <= k):
// rate = beta N_j N_k, halved on the diagonal; death of rate at j and k; birth
// of rate split by Kumar-Ramkrishna weights between the pivots bracketing
// v = x_j + x_k, clipped to bin n-1 / bin 0 outside the grid).
//
// Design (MI355X-first, not the reference's one-thread-per-pair layout):
//
//  * Two launches over the upper triangle, cut into blocks of T = 64*U
//    columns.  The diagonal band -- for each block, the rows of the block
//    itself and of the block below, where the near pairs live -- runs as
//    square T x T tiles split into row slices, many small workgroups
//    scheduled by the hardware (agg_tile.hpp).  Everything below the band
//    runs as column strips split evenly over exactly the resident workgroups
//    (agg_strip.hpp).  Lanes own columns k (U each), waves own rows; row
//    factors are staged in LDS and read as wave-uniform broadcasts.
//
//  * The pair rate uses the separable form of the collision kernel
//    (pbe_device.hpp): R <= 5 multiply-adds per pair.
//
//  * "Far" pairs -- v = x_j + x_k < x_{k+1} -- are the vast majority at large
//    n; all of them land in the bracket (k, k+1), so a lane keeps that
//    column's births in registers: sum_j w rate (to k+1) and the lo part
//    sum_j (1-w) rate = colsum_k - sum_j w rate, where colsum_k is the
//    separable column sum of the rates.  A chunk of rows all far from the
//    wave's columns runs a 3-op-plus-rate inner loop with no compares.
//    Column n-1 joins them: all its pairs clip to bin n-1 (weight w = 0).
//
//  * "Near" pairs (the band) keep a monotone exact bracket per column (a
//    cursor) and ds_add_f64 their births into an LDS window of the tile's
//    target bins (global atomics beyond it).
//
//  * Death N_i sum_k beta_ik N_k is separable: rows lose sum_r A_r(j) TK_r
//    (TK_r: the column sums of the tile or block) and columns lose
//    sum_r B_r(k) TJ_r; a diagonal tile's rows lose sum_r A_r(i) TK_r, which
//    covers both sides of its own pairs.
//
//  * One flush per tile or visit of a block: the LDS window is added to rhs
//    with FP64 global atomics (zero entries skipped).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "agg_band.hpp"
#include "agg_strip.hpp"

namespace pbe {

template <int KT, int U>
__global__ __launch_bounds__(256, (U >= 8 ? 2 : 3))
void agg_band_kernel(AggArgs a, int S)
{
    __shared__ AggSmem<KT, U> sm;
    band_item<KT, U>(a, blockIdx.x, S, sm);
}

// Diagonal band of 256-column blocks as diagonal sweeps (agg_band.hpp):
// item = (block, offset segment).
template <int KT>
__global__ __launch_bounds__(256, 4)
void agg_sweep_kernel(AggArgs a, int nseg, BandSegs segs)
{
    __shared__ BandSmem<KT> sm;
    agg_band_sweep<KT>(a, blockIdx.x / nseg, blockIdx.x % nseg, segs, sm);
}

// Offset segments of equal estimated cost.  A wave walks every offset of its
// segment whatever share of its lanes holds a pair there, so a step (one
// offset) costs 1; a move of the bracket costs about 2 more.  In block 0 no
// column has a pair beyond offset T.  On a geometric grid of ratio r the
// bracket of x_{k-e} + x_k sits log_r(1 + r^-e) bins above k, so offsets
// [0, e) make log_r 2 - log_r(1 + r^-e) moves.  Without a ratio (a general
// grid) the segments are equal.
static BandSegs band_segments(int S, double inv_log_r)
{
    constexpr int T = kBandT, E = 2 * kBandT;
    constexpr double kMoveCost = 2.0;
    BandSegs sg;
    auto cost = [&](int e, bool corner) {
        const double fe = corner && e > T ? T : e;
        double c = fe;
        if (inv_log_r > 0.0) {
            const double lr = 1.0 / inv_log_r;
            c += kMoveCost * (std::log(2.0) - std::log1p(std::exp(-fe * lr))) * inv_log_r;
        }
        return c;
    };
    for (int corner = 0; corner < 2; ++corner) {
        short* out = corner ? sg.e0 : sg.e;
        const double total = cost(E, corner);
        const int EE = corner ? T : E;
        for (int s = 0; s <= S; ++s) {
            int e = s * EE / S;
            if (inv_log_r > 0.0 && s > 0 && s < S) {
                const double target = total * s / S;
                int lo = 0, hi = EE;   // smallest e with cost(e) >= target
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (cost(mid, corner) >= target) hi = mid; else lo = mid + 1;
                }
                e = lo;
            }
            out[s] = (short)e;
        }
        for (int s = S + 1; s <= kBandMaxSeg; ++s) out[s] = (short)E;
        // bracket offset at each stream's first pair (see BandSegs::d)
        short(*dd)[kBandStreams] = corner ? sg.d0 : sg.d;
        for (int s = 0; s < kBandMaxSeg; ++s) {
            const int e0 = out[s], e1 = out[s + 1];
            const int L = (e1 - e0 + kBandStreams - 1) / kBandStreams;
            for (int t = 0; t < kBandStreams; ++t) {
                const int eu = std::min(e0 + (t + 1) * L - 1, e1 - 1);
                short dl = -1;
                if (inv_log_r > 0.0 && s < S && eu >= e0 + t * L) {
                    const double f =
                        std::log1p(std::exp(-(double)eu / inv_log_r)) * inv_log_r;
                    if (f >= 0.0 && f < 30000.0) dl = (short)std::floor(f);
                }
                dd[s][t] = dl;
            }
        }
    }
    return sg;
}

// band_segments costs a few hundred log/exp evaluations (about 9 us of host
// time); a launch reuses the calling thread's last plan when S and the grid
// ratio are the same (compared bit for bit).
static const BandSegs& band_segments_cached(int S, double inv_log_r)
{
    thread_local int c_S = -1;
    thread_local double c_ilr = 0.0;
    thread_local BandSegs c_segs;
    if (S != c_S || std::memcmp(&inv_log_r, &c_ilr, sizeof(double)) != 0) {
        c_segs = band_segments(S, inv_log_r);
        c_S = S;
        c_ilr = inv_log_r;
    }
    return c_segs;
}

// Small problems (too few far chunks to fill the device once): one launch of
// every tile, the band slices first, then the far tiles whole in
// diagonal-major order, all scheduled by the hardware.
template <int KT, int U>
__global__ __launch_bounds__(256, (U >= 8 ? 2 : 3))
void agg_tiles_kernel(AggArgs a, int S, int D)
{
    __shared__ AggSmem<KT, U> sm;
    agg_tiles_item<KT, U>(a, blockIdx.x, S, D, sm);
}

// Column strips over the resident grid (with or without each block's
// diagonal band): one workgroup of StripWaves<U> waves per compute unit,
// each taking an equal share of the plan's cost line.
template <int KT, int U>
__global__ __launch_bounds__(kWave * StripWaves<U>::value, 1)
void agg_strip_kernel(AggArgs a, StripPlan plan)
{
    __shared__ StripSmem<KT, U> sm;
    const long long CT = strip_cost_total(plan), G = gridDim.x, g = blockIdx.x;
    int qa, ca, qb, cb;
    strip_chunk_at_cost(plan, CT * g / G, qa, ca);
    strip_chunk_at_cost(plan, CT * (g + 1) / G, qb, cb);
    // visits: blocks qa .. qe - 1 (the last one ends at chunk cb of block qb)
    if (qa > qb || (qa == qb && ca >= cb)) return;   // empty share
    const int qe = cb > 0 ? qb + 1 : qb;
    StripCols<KT, U> cols;
    cols.load(a, qa);
    for (int i = threadIdx.x; i < StripSmem<KT, U>::WIN; i += blockDim.x) sm.colwin[i] = 0.0;
    for (int q = qa, c = ca; q < qe; ++q, c = 0) {
        const int hi = (q == qb) ? cb : strip_nc(plan, q);
        agg_strip_visit<KT, U>(a, plan, q, c, hi, q + 1 < qe ? q + 1 : -1, cols, sm);
    }
}

// Mid-size problems (form 2): the band sweep items and strip shares of four
// waves in one launch, side by side (two 256-thread workgroups per CU), so
// that the two latency chains run together instead of one after the other
// (the sweep's births are bound by LDS reads, the strip's by FP64 issue).
// The sweep items come first in the grid; the strip shares are the rest, all
// resident at once.  Whether this beats two launches depends on n and on the
// grid's near-pair reach (measured: n = 16384, r = 1.02, Brownian+shear
// 56.0 us against 62.4; at r = 1.01 114 against 91), so the tuned table
// picks it.  MB workgroups are resident per CU: 2 (form 2; 197 VGPRs, one
// strip and one sweep wave per SIMD) or 3 (form 4; the compiler holds the
// kernel to 168 VGPRs, spilling a few, so that three waves share a SIMD:
// n = 16384, r = 1.02, Brownian+shear 52.6 us against 56.2, but slower at
// n <= 8192, where the spills weigh more than the third wave: n = 4096 18.1
// against 13.6 us).
constexpr int kFusedWaves = 4;
constexpr int kFusedMaxBlocks = 64;   // 256-column blocks (n <= 16384)

template <int KT>
union FusedSmem {
    BandSmem<KT> band;
    StripSmem<KT, kBandT / kWave, kFusedWaves> strip;
};

template <int KT, int MB>
__global__ __launch_bounds__(kWave * kFusedWaves, MB)
void agg_fused_kernel(AggArgs a, int nseg, BandSegs segs, StripPlan plan, int nsweep)
{
    constexpr int U = kBandT / kWave;
    __shared__ FusedSmem<KT> sm;
    if ((int)blockIdx.x < nsweep) {
        agg_band_sweep<KT>(a, blockIdx.x / nseg, blockIdx.x % nseg, segs, sm.band);
        return;
    }
    const long long CT = strip_cost_total(plan), G = gridDim.x - nsweep, g = blockIdx.x - nsweep;
    int qa, ca, qb, cb;
    strip_chunk_at_cost(plan, CT * g / G, qa, ca);
    strip_chunk_at_cost(plan, CT * (g + 1) / G, qb, cb);
    if (qa > qb || (qa == qb && ca >= cb)) return;   // empty share
    const int qe = cb > 0 ? qb + 1 : qb;
    StripCols<KT, U, kFusedWaves> cols;
    cols.load(a, qa);
    for (int i = threadIdx.x; i < StripSmem<KT, U, kFusedWaves>::WIN; i += blockDim.x)
        sm.strip.colwin[i] = 0.0;
    for (int q = qa, c = ca; q < qe; ++q, c = 0) {
        const int hi = (q == qb) ? cb : strip_nc(plan, q);
        agg_strip_visit<KT, U, kFusedWaves>(a, plan, q, c, hi, q + 1 < qe ? q + 1 : -1, cols,
                                             sm.strip);
    }
}

// Workgroups of the fused kernel resident on the current device at once.
template <int KT, int MB>
static int fused_resident_groups()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&agg_fused_kernel<KT, MB>), kWave * kFusedWaves,
            0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    v = cus * per_cu;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

// Workgroups of the strip kernel resident on the current device at once.
template <int KT, int U>
static int resident_groups()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&agg_strip_kernel<KT, U>),
            kWave * StripWaves<U>::value, 0) !=
            hipSuccess || per_cu <= 0)
        per_cu = 1;
    v = cus * per_cu;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

// Strip plan of an n-bin problem in blocks of T = 64 U columns (the rows
// below each block's diagonal band; `band`: the band's rows too).  Chunk
// costs on the workgroups' cost line: a zero-birth chunk 2, a far chunk 20,
// a near chunk 200 (a near row is a chain of dependent steps with a bracket
// move in most of them: on the MI355X it takes about ten times a far row), a
// diagonal chunk 30, a block's visit 100.  On a geometric grid the strip chunks within the
// near-pair reach log_r(1 / (r - 1)) bins below a block's band are the near
// ones; measured on the strip launch at r = 1.005: n = 16384 231 us with the
// near chunks charged 10, 107 us at 40; n = 65536 (512-column blocks)
// 1685 us at 10, 674 us at 40.
static StripPlan make_strip_plan(int n, int U, bool band, double inv_log_r)
{
    StripPlan plan;
    const int T = kWave * U;
    plan.U = U;
    plan.nb = (n + T - 1) / T;
    plan.band = band ? 1 : 0;
    plan.m_near = 0;
    if (inv_log_r > 0.0 && std::isfinite(inv_log_r)) {
        const double reach = -std::log(std::expm1(1.0 / inv_log_r)) * inv_log_r;
        if (reach > T && reach < 1e7) plan.m_near = (int)std::ceil((reach - T) / kStripRows);
    }
    plan.mq = plan.m_near / U + 1;
    // Zero-birth chunks: on a geometric grid the rows more than
    // log_r 2^53.5 bins below a block's first column lie below half an ulp
    // of its pivot (the kernel tests every row; this only prices them).
    // Such a chunk stages its rows and adds their deaths but runs no pair
    // loop: priced at a tenth of a far chunk, and only where such chunks are
    // at least 60 % of the far ones.  Measured on the MI355X (r = 1.02,
    // Brownian+shear, default launch shape): n = 16384 49.4 us unpriced,
    // 36.3 at a tenth, 36.0 at a fifth, 39.8 at two fifths (the strip launch
    // alone: 43.9 at a whole far chunk, 34.5 at a twentieth, 28.9 at a
    // tenth); n = 32768 154 -> 72 us; n = 12288 (74 % zero-birth) 41.6 ->
    // 30.8; n = 8192 (62 %) 22.2 -> 21.9-22.9; n = 6144 (51 %) 21.0 -> 22.1-
    // 23.5 and n = 4096 (30 %) 15.0 -> 17.6-18.3, where the chunks are too
    // few for the price to pay.
    plan.zb = 0;
    plan.m_zero = 0;
    plan.mz = 1;
    plan.cz = 2;
    if (inv_log_r > 0.0 && std::isfinite(inv_log_r)) {
        const double dz = 53.5 * std::log(2.0) * inv_log_r;
        if (dz < 1e7) {
            plan.zb = 1;
            plan.m_zero = std::max((int)std::ceil(dz / kStripRows), U + plan.m_near);
            plan.mz = plan.m_zero / U + 1;
            const long long zf = strip_sz_sum(plan, plan.nb), ff = strip_sf_sum(plan, plan.nb);
            if (!(ff > 0 && 10 * zf >= 6 * ff)) {
                plan.zb = 0;
                plan.m_zero = 0;
                plan.mz = 1;
            }
        }
    }
    plan.cf = 20;
    plan.cn = 200;
    plan.cd = 30;
    plan.cv = 100;
    plan.total = strip_total(plan);
    return plan;
}

// Which launch(es) an aggregation call makes at block width U (the tuned
// band split and form given): the fused band + strip launch, one launch of
// every tile, or a band launch followed by the strip launch.  Form 3 asks
// for the tile launch whatever the strips' share of the work; form 4 is the
// fused launch at three workgroups per CU.
enum class AggForm { Fused, Tiles, BandStrips };
struct AggLaunch {
    AggForm form;
    int mb;              // Fused: workgroups per CU the kernel is built for (2 or 3)
    int nseg;            // Fused: offset segments per block
    long long groups;    // Fused / BandStrips: resident workgroups of the strip part
    long long items;     // Tiles: work items; Fused: sweep items; BandStrips: band items
    int diags;           // Tiles: leading diagonals run as row slices
    StripPlan sep;       // the strips below each block's band
};

// Diagonals of T-column tiles that hold near pairs: a tile on diagonal d
// has offsets k - j >= (d - 1) T + 1, so with near-pair reach `reach` those
// with d <= 1 + (reach - 1) / T (at least the two band diagonals).
static int tile_band_diags(int T, double inv_log_r)
{
    const double reach = grid_reach(inv_log_r);
    if (!(reach > T)) return 2;
    const double d = 2.0 + std::floor((reach - 1.0) / T);
    return d < 64.0 ? (int)d : 64;
}

template <int KT, int U>
static AggLaunch plan_agg_t(int n, double inv_log_r, int split, int form)
{
    constexpr int T = AggShape<U>::T;
    const int nb = (n + T - 1) / T;
    AggLaunch L;
    L.sep = make_strip_plan(n, U, false, inv_log_r);
    L.nseg = split;
    L.mb = form == 4 ? 3 : 2;
    if ((form == 2 || form == 4) && T == kBandT && nb <= kFusedMaxBlocks && L.sep.total > 0) {
        // mid sizes: sweep items on at most half of the resident slots of
        // the fused kernel, strip shares on the rest (empty shares return)
        const long long Gf = L.mb == 3 ? fused_resident_groups<KT, 3>()
                                       : fused_resident_groups<KT, 2>();
        int Sf = split;
        while (Sf > 1 && (long long)nb * Sf * 2 > Gf) Sf /= 2;
        const long long nsweep = (long long)nb * Sf;
        if (nsweep * 2 <= Gf) {
            L.form = AggForm::Fused;
            L.nseg = Sf;
            L.groups = Gf;
            L.items = nsweep;
            return L;
        }
    }
    L.groups = form == 3 ? 0 : resident_groups<KT, U>();
    if (form == 3 || L.sep.total < L.groups) {
        // asked for, or too little far work for the strips: one launch of
        // all tiles
        const long long tiles = (long long)nb * (nb + 1) / 2;
        L.form = AggForm::Tiles;
        L.diags = tile_band_diags(T, inv_log_r);
        const long long bt = band_tiles_d(nb, L.diags);
        L.items = (long long)split * bt + tiles - bt;
        return L;
    }
    L.form = AggForm::BandStrips;
    L.items = (long long)split * band_tiles(nb);
    return L;
}

template <int KT, int U>
static hipError_t launch_agg_t(AggArgs args, int split, int form, hipStream_t stream)
{
    constexpr int T = AggShape<U>::T;
    args.nb = (args.n + T - 1) / T;
    const AggLaunch L = plan_agg_t<KT, U>(args.n, args.inv_log_r, split, form);
    if (L.items > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (L.form) {
    case AggForm::Fused:
        if (L.mb == 3)
            hipLaunchKernelGGL((agg_fused_kernel<KT, 3>), dim3((unsigned)L.groups),
                               dim3(kWave * kFusedWaves), 0, stream, args, L.nseg,
                               band_segments_cached(L.nseg, args.inv_log_r), L.sep,
                               (int)L.items);
        else
            hipLaunchKernelGGL((agg_fused_kernel<KT, 2>), dim3((unsigned)L.groups),
                               dim3(kWave * kFusedWaves), 0, stream, args, L.nseg,
                               band_segments_cached(L.nseg, args.inv_log_r), L.sep,
                               (int)L.items);
        return hipSuccess;
    case AggForm::Tiles:
        hipLaunchKernelGGL((agg_tiles_kernel<KT, U>), dim3((unsigned)L.items), dim3(256), 0,
                           stream, args, split, L.diags);
        return hipSuccess;
    case AggForm::BandStrips:
        break;
    }
    if (T == kBandT) {
        // band as diagonal sweeps: split = offset segments per block
        hipLaunchKernelGGL((agg_sweep_kernel<KT>), dim3((unsigned)(args.nb * split)), dim3(256),
                           0, stream, args, split, band_segments_cached(split, args.inv_log_r));
    } else {
        hipLaunchKernelGGL((agg_band_kernel<KT, U>), dim3((unsigned)L.items), dim3(256), 0,
                           stream, args, split);
    }
    hipLaunchKernelGGL((agg_strip_kernel<KT, U>), dim3((unsigned)L.groups),
                       dim3(kWave * StripWaves<U>::value), 0, stream, args, L.sep);
    return hipSuccess;
}

template <int U>
static hipError_t launch_agg_u(const AggArgs& args, int kernel, int split, int form,
                               hipStream_t stream)
{
    switch (kernel) {
    case 0: return launch_agg_t<0, U>(args, split, form, stream);
    case 1: return launch_agg_t<1, U>(args, split, form, stream);
    case 2: return launch_agg_t<2, U>(args, split, form, stream);
    case 3: return launch_agg_t<3, U>(args, split, form, stream);
    case 4: return launch_agg_t<4, U>(args, split, form, stream);
    case 5: return launch_agg_t<5, U>(args, split, form, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_aggregation(const double* N, const double* x, double* rhs,
                              const pbe_aggregation_params& p, int tile_u, int band_split,
                              int form, hipStream_t stream)
{
    AggArgs args;
    args.N = N;
    args.x = x;
    args.rhs = rhs;
    args.n = p.n;
    args.nb = 0;
    args.coef = KernelCoef{p.beta0, p.beta_br, p.beta_sh};
    args.log_x0 = p.log_x0;
    args.inv_log_r = p.inv_log_r;
    const int S = band_split_or_default(band_split);
    switch (tile_u) {
    case 1: return launch_agg_u<1>(args, p.kernel_type, S, form, stream);
    case 2: return launch_agg_u<2>(args, p.kernel_type, S, form, stream);
    case 8: return launch_agg_u<8>(args, p.kernel_type, S, form, stream);
    default: return launch_agg_u<4>(args, p.kernel_type, S, form, stream);
    }
}

template <int U>
static TilesShape tiles_shape_u(int n, int kernel, double inv_log_r, int split, int form)
{
    AggLaunch L;
    switch (kernel) {
    case 0: L = plan_agg_t<0, U>(n, inv_log_r, split, form); break;
    case 1: L = plan_agg_t<1, U>(n, inv_log_r, split, form); break;
    case 2: L = plan_agg_t<2, U>(n, inv_log_r, split, form); break;
    case 3: L = plan_agg_t<3, U>(n, inv_log_r, split, form); break;
    case 4: L = plan_agg_t<4, U>(n, inv_log_r, split, form); break;
    case 5: L = plan_agg_t<5, U>(n, inv_log_r, split, form); break;
    default: return TilesShape{-1, 2};
    }
    return L.form == AggForm::Tiles ? TilesShape{L.items, L.diags} : TilesShape{-1, 2};
}

TilesShape aggregation_tiles_shape(const pbe_aggregation_params& p, int tile_u, int band_split,
                                   int form)
{
    const int S = band_split_or_default(band_split);
    switch (tile_u) {
    case 1: return tiles_shape_u<1>(p.n, p.kernel_type, p.inv_log_r, S, form);
    case 2: return tiles_shape_u<2>(p.n, p.kernel_type, p.inv_log_r, S, form);
    case 4: return tiles_shape_u<4>(p.n, p.kernel_type, p.inv_log_r, S, form);
    case 8: return tiles_shape_u<8>(p.n, p.kernel_type, p.inv_log_r, S, form);
    default: return TilesShape{-1, 2};
    }
}

}  // namespace pbe
