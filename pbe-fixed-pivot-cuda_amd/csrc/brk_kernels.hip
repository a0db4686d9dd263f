// brk_kernels.hip -- fixed-pivot breakage source term on gfx950.
//
// Reference semantics: src/breakage_kernels.cuh:115-133 (death
// rhs[j] -= S(x_j) N_j, skipped unless S_j != 0 and N_j > 2.2e-308) and
// :160-251 (birth: per parent j and quadrature point q, contrib =
// S_j N_j bw_q lands at v = x_j t_q, clipped to bin 0 for v <= x_0 and to bin
// n-1 for v >= x_{n-1}, otherwise split between the left bracket lo and lo+1
// with eta = (v - x_lo)/(x_{lo+1} - x_lo)).
//
// Design: one launch does death and birth.  A 256-thread workgroup owns P
// consecutive parents; S_j N_j is evaluated once per parent (one pow per
// parent for the power law, not one per quadrature point), the fragments of
// those parents land in the window of bins just below them, which is staged
// in LDS together with the pivots it covers.  Births and deaths ds_add_f64
// into the window; brackets are found by binary search in the LDS copy of x;
// the window is flushed once with FP64 global atomics.  Targets outside the
// window (t_q >= 1, or fragments below its reach) fall back to a search in
// global x and a global atomic.
#include "pbe_device.hpp"
#include "pbe_internal.hpp"

namespace pbe {

struct BrkArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    const double* __restrict__ t_q;
    const double* __restrict__ bw_q;
    double* rhs;
    int n;
    int nq;
    double S0, v_ref, alpha, v_min;
};

constexpr int kBrkParents = 256;   // parents per workgroup
constexpr int kBrkWin = 2048;      // bins in the LDS window (16 KB)

template <int SEL>
__global__ __launch_bounds__(256)
void brk_kernel(BrkArgs a)
{
    constexpr int P = kBrkParents, W = kBrkWin;
    __shared__ double sn[P];
    __shared__ double win[W];
    __shared__ double xw[W];

    const int tid = threadIdx.x;
    const int n = a.n, nq = a.nq;
    const int p0 = blockIdx.x * P;
    const int wbase = p0 + P - W;           // window = bins [wbase, wbase + W)
    const int wlo = wbase > 0 ? wbase : 0;  // first real bin of the window
    const int whi = min(wbase + W, n) - 1;  // last real bin of the window

    for (int i = tid; i < W; i += 256) {
        win[i] = 0.0;
        const int b = wbase + i;
        xw[i] = (b >= 0 && b < n) ? a.x[b] : 0.0;
    }
    __syncthreads();

    auto put = [&](int i, double c) {
        const int o = i - wbase;
        if (o >= 0 && o < W) lds_add(&win[o], c);
        else global_add(&a.rhs[i], c);
    };

    // death and S_j N_j, once per parent
    for (int i = tid; i < P; i += 256) {
        const int j = p0 + i;
        double s = 0.0;
        if (j < n) {
            const double xj = xw[j - wbase];
            const double Nj = a.N[j];
            const double Sj = selection<SEL>(xj, a.S0, a.v_ref, a.alpha, a.v_min);
            if (Sj != 0.0 && Nj > 2.2e-308) {
                s = Sj * Nj;
                put(j, -s);
            }
        }
        sn[i] = s;
    }
    __syncthreads();

    const double xfirst = a.x[0];
    const double xlast = a.x[n - 1];
    const double xwlo = xw[wlo - wbase];
    const double xwhi = xw[whi - wbase];
    const int items = P * nq;
    for (int it = tid; it < items; it += 256) {
        const int i = it / nq;
        const int q = it - i * nq;
        const int j = p0 + i;
        if (j >= n) continue;
        const double s = sn[i];
        if (s == 0.0) continue;
        const double c = s * a.bw_q[q];
        const double v = xw[j - wbase] * a.t_q[q];
        if (v <= xfirst) {
            put(0, c);
        } else if (v >= xlast) {
            put(n - 1, c);
        } else {
            int hi;
            if (v >= xwlo && v < xwhi) {
                // search the LDS copy: bracket inside [wlo, whi]
                int l = wlo - wbase, h = whi - wbase;
                while (h - l > 1) {
                    const int m = l + ((h - l) >> 1);
                    if (xw[m] <= v) l = m; else h = m;
                }
                hi = h + wbase;
            } else {
                hi = bracket_search(a.x, 0, n - 1, v);
            }
            const int lo = hi - 1;
            const double xl = a.x[lo], xh = a.x[hi];
            const double eta = (v - xl) / (xh - xl);
            put(lo, (1.0 - eta) * c);
            put(hi, eta * c);
        }
    }
    __syncthreads();

    for (int i = tid; i < W; i += 256) {
        const int b = wbase + i;
        const double v = win[i];
        if (b >= 0 && b < n && v != 0.0) global_add(&a.rhs[b], v);
    }
}

hipError_t launch_breakage(const double* N, const double* x, const double* t_q,
                           const double* bw_q, double* rhs,
                           const pbe_breakage_params& p, hipStream_t stream)
{
    BrkArgs args{N, x, t_q, bw_q, rhs, p.n, p.n_quad, p.S0, p.v_ref, p.alpha, p.v_min};
    const unsigned grid = (unsigned)((p.n + kBrkParents - 1) / kBrkParents);
    switch (p.selection) {
    case 0: hipLaunchKernelGGL(brk_kernel<0>, dim3(grid), dim3(256), 0, stream, args); break;
    case 1: hipLaunchKernelGGL(brk_kernel<1>, dim3(grid), dim3(256), 0, stream, args); break;
    case 2: hipLaunchKernelGGL(brk_kernel<2>, dim3(grid), dim3(256), 0, stream, args); break;
    case 3: hipLaunchKernelGGL(brk_kernel<3>, dim3(grid), dim3(256), 0, stream, args); break;
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

}  // namespace pbe
