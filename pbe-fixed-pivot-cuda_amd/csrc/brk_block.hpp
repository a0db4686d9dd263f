// brk_block.hpp -- breakage death + birth of one block of parents (device
// code shared by the single-instance and batched kernels); see brk_kernels.hip.
#pragma once

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

constexpr int kBrkQuadLds = 64;   // quadrature points held in LDS (more: read from global)

struct BrkSmem {
    double sn[kBrkParents];
    double win[kBrkWin];
    double xw[kBrkWin];
    double tq[kBrkQuadLds], bw[kBrkQuadLds];
};

// Death and birth of parents [p0, p0 + 256) of the instance described by `a`.
// All 256 threads call it.
template <int SEL>
__device__ __forceinline__ void brk_block_body(const BrkArgs& a, const int p0, BrkSmem& sm)
{
    constexpr int P = kBrkParents, W = kBrkWin;
    double* sn = sm.sn;
    double* win = sm.win;
    double* xw = sm.xw;

    const int tid = threadIdx.x;
    const int n = a.n, nq = a.nq;
    const int wbase = p0 + P - W;           // window = bins [wbase, wbase + W)
    const int wlo = wbase > 0 ? wbase : 0;  // first real bin of the window
    const int whi = min(wbase + W, n) - 1;  // last real bin of the window

    // the parents' populations are loaded with the window, one round trip
    double nj[(P + 255) / 256];
#pragma unroll
    for (int t = 0; t < (P + 255) / 256; ++t) {
        const int j = p0 + tid + 256 * t;
        nj[t] = j < n ? a.N[j] : 0.0;
    }
    for (int i = tid; i < W; i += 256) {
        win[i] = 0.0;
        const int b = wbase + i;
        xw[i] = (b >= 0 && b < n) ? a.x[b] : 0.0;
    }
    // the daughter table and the grid's end pivots come with the same round
    // trip as the window
    if (tid < nq && tid < kBrkQuadLds) {
        sm.tq[tid] = a.t_q[tid];
        sm.bw[tid] = a.bw_q[tid];
    }
    const double xfirst = a.x[0];
    const double xlast = a.x[n - 1];
    __syncthreads();

    auto put = [&](int i, double c) {
        const int o = i - wbase;
        if (o >= 0 && o < W) lds_add(&win[o], c);
        else global_add(&a.rhs[i], c);
    };

    // death and S_j N_j, once per parent
#pragma unroll
    for (int t = 0; t < (P + 255) / 256; ++t) {
        const int i = tid + 256 * t;
        const int j = p0 + i;
        double s = 0.0;
        if (i < P && j < n) {
            const double xj = xw[j - wbase];
            const double Nj = nj[t];
            const double Sj = selection<SEL>(xj, a.S0, a.v_ref, a.alpha, a.v_min);
            if (Sj != 0.0 && Nj > 2.2e-308) {
                s = Sj * Nj;
                put(j, -s);
            }
        }
        if (i < P) sn[i] = s;
    }
    __syncthreads();

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
        const double bwq = q < kBrkQuadLds ? lds_ld(&sm.bw[q]) : glb_ld(&a.bw_q[q]);
        const double tqq = q < kBrkQuadLds ? lds_ld(&sm.tq[q]) : glb_ld(&a.t_q[q]);
        const double c = s * bwq;
        const double v = xw[j - wbase] * tqq;
        if (v <= xfirst) {
            put(0, c);
        } else if (v >= xlast) {
            put(n - 1, c);
        } else {
            int hi;
            double xl, xh;
            if (v >= xwlo && v < xwhi) {
                // search the LDS copy: bracket inside [wlo, whi]
                int l = wlo - wbase, h = whi - wbase;
                while (h - l > 1) {
                    const int m = l + ((h - l) >> 1);
                    if (xw[m] <= v) l = m; else h = m;
                }
                hi = h + wbase;
                xl = xw[l];
                xh = xw[h];
            } else {
                hi = bracket_search(a.x, 0, n - 1, v);
                xl = a.x[hi - 1];
                xh = a.x[hi];
            }
            const int lo = hi - 1;
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


}  // namespace pbe
