// det_kernels.hip -- atomic-free, bitwise repeatable aggregation and breakage
// source terms (the gather mode; no counterpart in the reference, whose FP64
// atomics make the summation order vary from run to run).
//
// Every output bin is written by exactly one lane, and every sum is formed in
// an order fixed by n (and n_quad) alone, so two calls on the same inputs
// return identical bits.  Each contribution is the same expression as in the
// atomic kernels; only the order of the additions differs.
//
// Aggregation, three launches:
//   1. table: per bin A_r(i) = N_i a_r(i), B_r(i) = N_i phi_r(i), x_i,
//      1/(x_{i+1} - x_i), and per 64-bin block the column sums of B_r.
//   2. far partials: one wave per (64-column block, row segment); lane k sums
//      over the segment's rows j <= k whose pair is far (v = x_j + x_k <
//      x_{k+1}, or k = n-1 where every pair clips with weight 0) the rate and
//      the weighted rate, and writes its lo / hi parts to the segment's slot.
//   3. gather: lane i adds, in a fixed order, the far parts of columns i and
//      i-1 over all segments, the near pairs whose bracket is (i-1, i) or
//      (i, i+1) (enumerated column by column with two monotone row pointers),
//      for i = n-1 the clipped pairs (summed by its whole wave in a fixed
//      butterfly), and the separable death -sum_r A_r(i) TB_r.
//
// Breakage, two launches: a table of S_j N_j (the reference's guard applied),
// then lane i gathers its death and, per quadrature point, the parents whose
// fragment x_j t_q lands in (x_{i-1}, x_{i+1}) (binary search on the rounded
// fragment volume); the fragments clipped to bin 0 or bin n-1 are summed by
// the wave that owns that bin.
#include <algorithm>
#include <mutex>

#include "pbe_device.hpp"
#include "pbe_internal.hpp"

namespace pbe {

namespace {

constexpr int kDetBlock = 64;   // columns per far-partial item, bins per table block
constexpr int kDetGather = 64;  // threads per gather workgroup (one wave: many workgroups)
constexpr int kDetSplit = 4;    // lanes per gather target (its near columns interleaved)
constexpr int kDetTargets = kDetGather / kDetSplit;   // targets per gather workgroup
static_assert(kDetGather == kWave, "the gather's sums are wave butterflies");
constexpr int kDetXBelow = 512; // pivots below a gather workgroup's targets held in LDS
constexpr int kDetTabBelow = 256; // table rows below a gather workgroup's targets held in LDS

template <int KT>
struct DetLayout {
    static constexpr int R = Kern<KT>::R;
    static constexpr int TS = 2 * R + 2;   // table stride: A[R], B[R], x, 1/dx
};

// Row segments of the far partials: a function of n only.
constexpr int det_segments(int n)
{
    return (n + 1023) / 1024 < 1 ? 1 : ((n + 1023) / 1024 > 16 ? 16 : (n + 1023) / 1024);
}

// Workspace of the aggregation gather: the table (the stride of the widest
// kernel), the per-block column sums and the lo / hi far partials.
constexpr size_t det_agg_doubles(int n)
{
    return (size_t)n * DetLayout<5>::TS + (size_t)((n + kDetBlock - 1) / kDetBlock) * Kern<5>::R +
           2 * (size_t)det_segments(n) * (size_t)n;
}

struct DetAggArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    double* rhs;
    double* tab;      // [n][TS]
    double* tbq;      // [nbq][R]
    double* part;     // [2][S][n]: lo parts, then hi parts
    int n, nbq, S;
    KernelCoef coef;
};

template <int KT>
__global__ __launch_bounds__(64) void det_agg_table_kernel(DetAggArgs a)
{
    using L = DetLayout<KT>;
    constexpr int R = L::R, TS = L::TS;
    const int i = blockIdx.x * kDetBlock + threadIdx.x;
    double A[R], B[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        A[r] = 0.0;
        B[r] = 0.0;
    }
    if (i < a.n) {
        const double xi = a.x[i], Ni = a.N[i];
        double cb, icb;
        pivot_roots<KT>(xi, cb, icb);
        double f[R];
        Kern<KT>::row(A, xi, cb, icb, a.coef);
        Kern<KT>::col(f, xi, cb, icb);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            A[r] *= Ni;
            B[r] = Ni * f[r];
        }
        double* t = a.tab + (size_t)TS * i;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            t[r] = A[r];
            t[R + r] = B[r];
        }
        t[2 * R] = xi;
        t[2 * R + 1] = i < a.n - 1 ? 1.0 / (a.x[i + 1] - xi) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double s = wave_sum(B[r]);
        if (threadIdx.x == 0) a.tbq[(size_t)R * blockIdx.x + r] = s;
    }
}

// Far partials.  Item = (block q, segment s), heaviest blocks first.
template <int KT>
__global__ __launch_bounds__(64) void det_agg_far_kernel(DetAggArgs a)
{
    using L = DetLayout<KT>;
    constexpr int R = L::R, TS = L::TS;
    const int n = a.n, S = a.S;
    const int q = a.nbq - 1 - (int)blockIdx.x / S;
    const int s = (int)blockIdx.x % S;
    const int k = q * kDetBlock + (int)threadIdx.x;
    const int rq = min(n, (q + 1) * kDetBlock);   // rows j < rq (j <= k)
    const int r0 = (int)((long long)s * rq / S), r1 = (int)((long long)(s + 1) * rq / S);
    const bool valid = k < n;
    double B[R], xk = 0.0, xk1 = 0.0, idk = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) B[r] = 0.0;
    if (valid) {
        const double* t = a.tab + (size_t)TS * k;
#pragma unroll
        for (int r = 0; r < R; ++r) B[r] = t[R + r];
        xk = t[2 * R];
        idk = t[2 * R + 1];
        xk1 = k < n - 1 ? a.x[k + 1] : __builtin_inf();
    }
    // Rows are staged 64 at a time in LDS (one row's A_r and x per lane) and
    // read back as wave-uniform broadcasts, four rows per step with their own
    // accumulators (four independent FP64 chains per lane): row j goes to
    // accumulator (j - r0) % 4, and the four are combined in a fixed order,
    // so the sums depend only on n and the segment bounds.  A pair that is
    // not far adds zero.  A staged block whose rows all lie below the
    // column block and are far from every column (its last row is, since x
    // increases) takes the same steps without the masks: the sums are
    // bitwise those of the masked steps, which multiply by exactly 1.
    constexpr int FR = 4;
    constexpr int RS = R + 1;
    __shared__ double rows[kDetBlock][RS];
    double ar[FR], aw[FR];
#pragma unroll
    for (int f = 0; f < FR; ++f) {
        ar[f] = 0.0;
        aw[f] = 0.0;
    }
    auto row = [&](int j, const double* t, double& acr, double& acw) {
        double pr = t[0] * B[0];
#pragma unroll
        for (int r = 1; r < R; ++r) pr = fma(t[r], B[r], pr);
        const double v = t[R] + xk;
        const bool far = valid && j <= k && v < xk1;
        pr *= far ? (j == k ? 0.5 : 1.0) : 0.0;
        acr += pr;
        acw = fma(v - xk, pr, acw);
    };
    auto row_far = [&](const double* t, double& acr, double& acw) {
        double pr = t[0] * B[0];
#pragma unroll
        for (int r = 1; r < R; ++r) pr = fma(t[r], B[r], pr);
        const double v = t[R] + xk;
        acr += pr;
        acw = fma(v - xk, pr, acw);
    };
    for (int jb = r0; jb < r1; jb += kDetBlock) {
        const int nr = min(kDetBlock, r1 - jb);
        if ((int)threadIdx.x < nr) {
            const double* t = a.tab + (size_t)TS * (jb + threadIdx.x);
#pragma unroll
            for (int r = 0; r < R; ++r) rows[threadIdx.x][r] = t[r];
            rows[threadIdx.x][R] = t[2 * R];
        }
        __syncthreads();
        const bool allfar = jb + nr <= q * kDetBlock &&
                            __all(!valid || rows[nr - 1][R] + xk < xk1);
        int jj = 0;
        if (allfar) {
            for (; jj + FR <= nr; jj += FR) {
#pragma unroll
                for (int f = 0; f < FR; ++f) row_far(rows[jj + f], ar[f], aw[f]);
            }
            for (; jj < nr; ++jj) row_far(rows[jj], ar[jj % FR], aw[jj % FR]);
        } else {
            for (; jj + FR <= nr; jj += FR) {
#pragma unroll
                for (int f = 0; f < FR; ++f) row(jb + jj + f, rows[jj + f], ar[f], aw[f]);
            }
            for (; jj < nr; ++jj) row(jb + jj, rows[jj], ar[jj % FR], aw[jj % FR]);
        }
        __syncthreads();
    }
    const double accr = (ar[0] + ar[1]) + (ar[2] + ar[3]);
    const double accw = (aw[0] + aw[1]) + (aw[2] + aw[3]);
    if (valid) {
        const double hi = accw * idk;   // 0 for k = n-1: every pair clips there
        a.part[(size_t)s * n + k] = accr - hi;
        a.part[(size_t)(S + s) * n + k] = hi;
    }
}

// First index in [lo, hi) where pred holds (pred monotone false..true); hi if none.
template <class P>
__device__ __forceinline__ int first_true(int lo, int hi, P pred)
{
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (pred(mid)) hi = mid; else lo = mid + 1;
    }
    return lo;
}

template <int KT>
__device__ __forceinline__ double det_rate(const double* tab, int j, int k)
{
    using L = DetLayout<KT>;
    constexpr int R = L::R, TS = L::TS;
    const double* tj = tab + (size_t)TS * j;
    const double* tk = tab + (size_t)TS * k;
    double pr = tj[0] * tk[R];
#pragma unroll
    for (int r = 1; r < R; ++r) pr = fma(tj[r], tk[R + r], pr);
    return j == k ? 0.5 * pr : pr;
}

// Table reads of the gather: element c of bin i's row, from the workgroup's
// LDS copy of the rows [wlo, whi) (LDS = true; every row its pairs touch lies
// there) or from global memory.
template <int TS, bool LDS>
struct DetTabWin {
    const double* ts;    // LDS copy of tab rows [wlo, whi)
    const double* tab;
    int wlo;
    __device__ __forceinline__ double operator()(int i, int c) const
    {
        if constexpr (LDS) return lds_ld(&ts[(i - wlo) * TS + c]);
        else return glb_ld(&tab[(size_t)TS * i + c]);
    }
};

template <int KT, class TW>
__device__ __forceinline__ double det_rate_w(const TW& T, int j, int k)
{
    constexpr int R = DetLayout<KT>::R;
    double pr = T(j, 0) * T(k, R);
#pragma unroll
    for (int r = 1; r < R; ++r) pr = fma(T(j, r), T(k, R + r), pr);
    return j == k ? 0.5 * pr : pr;
}

// Pivot reads of the gather: from the workgroup's LDS copy of x[wlo, whi)
// (LDS = true; every index its pairs touch lies there) or from global memory.
template <bool LDS>
struct DetXWin {
    const double* xs;   // LDS copy of x[wlo, whi)
    const double* x;
    int wlo, whi;
    __device__ __forceinline__ double operator()(int j) const
    {
        if constexpr (LDS) return lds_ld(&xs[j - wlo]);
        else return glb_ld(&x[j]);
    }
};

// Near births of target i from pairs whose bracket is (Lb, Lb+1); upper = true
// takes the weight w of the upper pivot (Lb = i - 1), else 1 - w (Lb = i).
// Columns k = Lb - 1 - sub, Lb - 1 - sub - kDetSplit, ... (this lane's share
// of the target's columns, walked downward): the rows of a column whose sum
// lands in the bracket form a range [ja, jb) that only moves up as k falls.
template <int KT, class XW, class TW>
__device__ __forceinline__ double det_near(const XW& X, const TW& T, int Lb, bool upper,
                                           int sub)
{
    constexpr int R = DetLayout<KT>::R;
    const double xL = X(Lb), xL1 = X(Lb + 1);
    const double idL = T(Lb, 2 * R + 1);
    double acc = 0.0;
    int k = Lb - 1 - sub;
    if (k < 0 || X(k) + X(k) < xL) return 0.0;   // nor can any smaller column
    const double xk0 = X(k);
    int ja = first_true(X.wlo, k + 1, [&](int j) { return X(j) + xk0 >= xL; });
    int jb = first_true(ja, k + 1, [&](int j) { return X(j) + xk0 >= xL1; });
    for (; k >= 0 && ja <= k; k -= kDetSplit) {
        const double xk = X(k);
        while (ja <= k && X(ja) + xk < xL) ++ja;
        if (jb < ja) jb = ja;
        while (jb <= k && X(jb) + xk < xL1) ++jb;
        const int je = min(jb, k + 1);   // rows j <= k
        for (int j = ja; j < je; ++j) {
            const double pr = det_rate_w<KT>(T, j, k);
            const double w = (X(j) + xk - xL) * idL;
            acc += upper ? w * pr : pr - w * pr;
        }
    }
    return acc;
}

template <int KT>
__global__ __launch_bounds__(kDetGather) void det_agg_gather_kernel(DetAggArgs a)
{
    using L = DetLayout<KT>;
    constexpr int R = L::R, TS = L::TS;
    __shared__ double xs[kDetXBelow + kDetTargets + 1];
    __shared__ double ts[(kDetTabBelow + kDetTargets + 1) * TS];
    const int n = a.n, S = a.S;
    const int tid = threadIdx.x;
    // pivots of the neighbourhood: the targets' brackets and the rows and
    // columns of their near pairs (on a grid of ratio 1.01 or more, all of
    // them), and the table rows of the nearer part (ratio ~1.015 or more);
    // the barrier below orders these stores before the reads
    const int i0 = blockIdx.x * kDetTargets;
    const int wlo = max(0, i0 - kDetXBelow), whi = min(n, i0 + kDetTargets + 1);
    const int twlo = max(0, i0 - kDetTabBelow);
    for (int t = tid; t < whi - wlo; t += kDetGather) xs[t] = a.x[wlo + t];
    for (int t = tid; t < (whi - twlo) * TS; t += kDetGather)
        ts[t] = a.tab[(size_t)TS * twlo + t];
    // TB_r = sum of all column factors in a fixed order (each lane its
    // strided block totals in turn, then the wave's butterfly), the same in
    // every lane and every workgroup
    double TB[R];
    {
        double sb[R];
#pragma unroll
        for (int r = 0; r < R; ++r) sb[r] = 0.0;
        for (int b = tid; b < a.nbq; b += kDetGather)
#pragma unroll
            for (int r = 0; r < R; ++r) sb[r] += a.tbq[(size_t)R * b + r];
#pragma unroll
        for (int r = 0; r < R; ++r) TB[r] = wave_sum(sb[r]);
    }
    __syncthreads();   // the window's stores before its reads
    // lanes 4m .. 4m + 3 share target i0 + m
    const int i = i0 + tid / kDetSplit, sub = tid % kDetSplit;
    // The window serves every pivot read when, for each bracket pivot L of
    // the workgroup's targets (and n-1 for the clipped pairs), the smallest
    // row of a pair reaching x_L, from column L-1 where the rows reach
    // lowest, lies above wlo: x_wlo + x_{L-1} < x_L.  Columns and rows then
    // stay in [wlo, L].
    // The same test against the first table row held in LDS says whether
    // the table reads can come from there too (then so can the pivots).
    bool inwin = true, intab = true;
    {
        // lane t <= kDetTargets checks L = i0 - 1 + t: the pivots [i0 - 1, i0 + 15]
        const int L = min(i0 - 1 + min(tid, kDetTargets), n - 1);
        const bool clipwg = i0 + kDetTargets >= n - 1;
        if (wlo > 0) {
            inwin = xs[0] + a.x[L - 1] < a.x[L];
            if (clipwg) inwin = inwin && xs[0] + a.x[n - 2] < a.x[n - 1];
        }
        if (twlo > 0) {
            const double xt = xs[twlo - wlo];
            intab = xt + a.x[L - 1] < a.x[L];
            if (clipwg) intab = intab && xt + a.x[n - 2] < a.x[n - 1];
        }
    }
    inwin = __all(inwin);
    intab = __all(intab) && inwin;
    const DetXWin<true> XL{xs, a.x, wlo, whi};
    const DetXWin<false> XG{xs, a.x, 0, n};
    const DetTabWin<TS, true> TL{ts, a.tab, twlo};
    const DetTabWin<TS, false> TG{ts, a.tab, 0};

    // clipped pairs (v >= x_{n-1}, k <= n-2): the workgroup that owns bin n-1
    double clip = 0.0;
    if (i0 <= n - 1 && n - 1 < i0 + kDetTargets && n >= 2) {
        auto clipped = [&](const auto& X, const auto& T) {
            const double xl = X(n - 1);
            double c = 0.0;
            for (int k = n - 2 - tid; k >= X.wlo; k -= kDetGather) {
                const double xk = X(k);
                if (xk + xk < xl) break;   // smaller k cannot reach x_{n-1}
                const int jc = first_true(X.wlo, k + 1, [&](int j) { return X(j) + xk >= xl; });
                for (int j = jc; j <= k; ++j) c += det_rate_w<KT>(T, j, k);
            }
            return c;
        };
        clip = wave_sum(intab ? clipped(XL, TL) : inwin ? clipped(XL, TG) : clipped(XG, TG));
    }

    // near births: each lane its share of the target's columns, the four
    // shares summed in a fixed order ((s0 + s1) + (s2 + s3), the same value
    // in all four lanes)
    double nl = 0.0, nu = 0.0;
    auto near = [&](const auto& X, const auto& T) {
        if (i <= n - 2) nl = det_near<KT>(X, T, i, false, sub);
        if (i >= 1 && i - 1 <= n - 2) nu = det_near<KT>(X, T, i - 1, true, sub);
    };
    if (i < n) {
        if (intab) near(XL, TL);
        else if (inwin) near(XL, TG);
        else near(XG, TG);
    }
    nl += dpp_move<0xB1, 0xF>(nl);   // quad_perm 1,0,3,2
    nl += dpp_move<0x4E, 0xF>(nl);   // quad_perm 2,3,0,1
    nu += dpp_move<0xB1, 0xF>(nu);
    nu += dpp_move<0x4E, 0xF>(nu);
    if (i >= n || sub != 0) return;

    double acc = 0.0;
    for (int s = 0; s < S; ++s) acc += a.part[(size_t)s * n + i];
    if (i >= 1)
        for (int s = 0; s < S; ++s) acc += a.part[(size_t)(S + s) * n + i - 1];
    acc += nl;
    acc += nu;
    if (i == n - 1) acc += clip;
    const double* t = a.tab + (size_t)TS * i;
    double dth = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) dth = fma(t[r], TB[r], dth);
    a.rhs[i] += acc - dth;
}

template <int KT>
hipError_t det_agg_launch(const DetAggArgs& a, hipStream_t s)
{
    const int nbq = a.nbq;
    hipLaunchKernelGGL(det_agg_table_kernel<KT>, dim3((unsigned)nbq), dim3(64), 0, s, a);
    hipLaunchKernelGGL(det_agg_far_kernel<KT>, dim3((unsigned)(nbq * a.S)), dim3(64), 0, s, a);
    hipLaunchKernelGGL(det_agg_gather_kernel<KT>,
                       dim3((unsigned)((a.n + kDetTargets - 1) / kDetTargets)),
                       dim3(kDetGather), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------- breakage ---

struct DetBrkArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    const double* __restrict__ t_q;
    const double* __restrict__ bw_q;
    double* rhs;
    double* sn;       // [n]: S_j N_j of the guarded parents, else 0
    int n, nq;
    double S0, v_ref, alpha, v_min;
};

template <int SEL>
__global__ __launch_bounds__(256) void det_brk_table_kernel(DetBrkArgs a)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.n) return;
    const double xj = a.x[j], Nj = a.N[j];
    const double Sj = selection<SEL>(xj, a.S0, a.v_ref, a.alpha, a.v_min);
    // guard of src/breakage_kernels.cuh:131,195
    a.sn[j] = (Sj != 0.0 && Nj > 2.2e-308) ? Sj * Nj : 0.0;
}

__global__ __launch_bounds__(256) void det_brk_gather_kernel(DetBrkArgs a)
{
    const int n = a.n, nq = a.nq;
    const int tid = threadIdx.x, lane = tid & 63;
    const int i = blockIdx.x * 256 + tid;
    const double* x = a.x;
    const double x0 = x[0], xl = x[n - 1];

    // fragments clipped below x_0 (to bin 0) or at/above x_{n-1} (to bin n-1),
    // summed by the wave that owns the bin
    const int wave0 = blockIdx.x * 256 + (tid & ~63);
    double clip0 = 0.0, clipn = 0.0;
    if (wave0 == 0) {
        for (int q = 0; q < nq; ++q) {
            const double t = a.t_q[q], b = a.bw_q[q];
            const int J = first_true(0, n, [&](int j) { return !(x[j] * t < x0); });
            double s = 0.0;
            for (int j = lane; j < J; j += 64) s += (a.sn[j] * b);
            clip0 += s;
        }
        clip0 = wave_sum(clip0);
    }
    if (wave0 <= n - 1 && n - 1 < wave0 + 64) {
        for (int q = 0; q < nq; ++q) {
            const double t = a.t_q[q], b = a.bw_q[q];
            const int J = first_true(0, n, [&](int j) { return x[j] * t >= xl; });
            double s = 0.0;
            for (int j = J + lane; j < n; j += 64) s += (a.sn[j] * b);
            clipn += s;
        }
        clipn = wave_sum(clipn);
    }
    if (i >= n) return;

    double acc = -a.sn[i];   // death (src/breakage_kernels.cuh:132)
    if (n >= 2) {
        for (int q = 0; q < nq; ++q) {
            const double t = a.t_q[q], b = a.bw_q[q];
            // brackets (i, i+1): weight 1 - eta; (i-1, i): weight eta; only
            // unclipped fragments x_0 <= v < x_{n-1} (v = x_0 gives eta = 0,
            // the same as the reference's clip to bin 0)
            for (int side = 0; side < 2; ++side) {
                const int lo = i - side;
                if (lo < 0 || lo > n - 2) continue;
                const double xlo = x[lo], xhi = x[lo + 1];
                const int ja = first_true(0, n, [&](int j) { return x[j] * t >= xlo; });
                const int jb = first_true(ja, n, [&](int j) { return x[j] * t >= xhi; });
                for (int j = ja; j < jb; ++j) {
                    const double v = x[j] * t;
                    if (!(v < xl)) break;
                    const double contrib = a.sn[j] * b;
                    const double eta = (v - xlo) / (xhi - xlo);
                    acc += side ? eta * contrib : (1.0 - eta) * contrib;
                }
            }
        }
    }
    if (i == 0) acc += clip0;
    if (i == n - 1) acc += clipn;
    a.rhs[i] += acc;
}

template <int SEL>
void det_brk_launch(const DetBrkArgs& a, hipStream_t s)
{
    const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
    hipLaunchKernelGGL(det_brk_table_kernel<SEL>, g, b, 0, s, a);
    hipLaunchKernelGGL(det_brk_gather_kernel, g, b, 0, s, a);
}

}  // namespace

size_t det_aggregation_workspace_doubles(int n) { return det_agg_doubles(n); }

size_t det_breakage_workspace_doubles(int n) { return (size_t)n; }

hipError_t launch_aggregation_deterministic(const double* N, const double* x, double* rhs,
                                            const pbe_aggregation_params& p, double* ws,
                                            hipStream_t stream)
{
    DetAggArgs a;
    a.N = N;
    a.x = x;
    a.rhs = rhs;
    a.n = p.n;
    a.nbq = (p.n + kDetBlock - 1) / kDetBlock;
    a.S = det_segments(p.n);
    a.coef = KernelCoef{p.beta0, p.beta_br, p.beta_sh};
    // the table stride depends on the kernel; the layout below fits the largest
    const size_t tab_doubles = (size_t)p.n * DetLayout<5>::TS;
    a.tab = ws;
    a.tbq = ws + tab_doubles;
    a.part = a.tbq + (size_t)a.nbq * Kern<5>::R;
    switch (p.kernel_type) {
    case 0: return det_agg_launch<0>(a, stream);
    case 1: return det_agg_launch<1>(a, stream);
    case 2: return det_agg_launch<2>(a, stream);
    case 3: return det_agg_launch<3>(a, stream);
    case 4: return det_agg_launch<4>(a, stream);
    case 5: return det_agg_launch<5>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// The aggregation gather without a caller workspace, for callers that cannot
// allocate device memory (code under graph capture, serving loops).  Its
// workspace is `scratch`, a file-scope device array sized for n up to
// kDetNoAllocMaxN (23 MB; each device holds its own copy until the process
// ends).  Host code never uses the array's name as an address: the first call
// on a device looks the device's copy up with hipGetSymbolAddress and keeps
// the pointer per device.  Every call on a device shares its copy, so calls
// that overlap on one device (other streams, other threads) race.  The
// kernels overwrite every entry they read, so nothing needs clearing between
// calls.
// ---------------------------------------------------------------------------
constexpr size_t kDetNoAllocDoubles = det_agg_doubles(kDetNoAllocMaxN);
static __device__ __align__(256) double scratch[kDetNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_on_current_device(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

hipError_t launch_aggregation_deterministic_noalloc(const double* N, const double* x,
                                                    double* rhs,
                                                    const pbe_aggregation_params& p,
                                                    hipStream_t stream)
{
    if (p.n > kDetNoAllocMaxN) return hipErrorInvalidValue;
    double* ws = nullptr;
    hipError_t e = scratch_on_current_device(&ws);
    if (e != hipSuccess) return e;
    return launch_aggregation_deterministic(N, x, rhs, p, ws, stream);
}

hipError_t launch_breakage_deterministic(const double* N, const double* x, const double* t_q,
                                         const double* bw_q, double* rhs,
                                         const pbe_breakage_params& p, double* ws,
                                         hipStream_t stream)
{
    DetBrkArgs a;
    a.N = N;
    a.x = x;
    a.t_q = t_q;
    a.bw_q = bw_q;
    a.rhs = rhs;
    a.sn = ws;
    a.n = p.n;
    a.nq = p.n_quad;
    a.S0 = p.S0;
    a.v_ref = p.v_ref;
    a.alpha = p.alpha;
    a.v_min = p.v_min;
    switch (p.selection) {
    case 0: det_brk_launch<0>(a, stream); break;
    case 1: det_brk_launch<1>(a, stream); break;
    case 2: det_brk_launch<2>(a, stream); break;
    case 3: det_brk_launch<3>(a, stream); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pbe
