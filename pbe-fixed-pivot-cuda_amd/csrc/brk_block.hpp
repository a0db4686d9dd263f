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
// All NT threads of the workgroup call it (NT a multiple of 64, at least 256:
// more threads give each fewer of the block's P * n_quad birth items).
template <int SEL, int NT = 256>
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
    double nj[(P + NT - 1) / NT];
#pragma unroll
    for (int t = 0; t < (P + NT - 1) / NT; ++t) {
        const int j = p0 + tid + NT * t;
        nj[t] = j < n ? a.N[j] : 0.0;
    }
    for (int i = tid; i < W; i += NT) {
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
    for (int t = 0; t < (P + NT - 1) / NT; ++t) {
        const int i = tid + NT * t;
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
    // A thread's items step by NT: for a table of m points dividing NT its
    // quadrature point stays the same and its parent moves up by NT / m.
    // On a grid of constant ratio the left bracket of x_j t_q then moves up
    // by as many bins, so the last bracket found, moved up by the parent
    // step, is checked first (two LDS reads) and the binary search runs only
    // when it misses; the check accepts only the exact bracket
    // x_l <= v < x_{l+1}, so any grid gives the search's result.
    int cur_q = -1, cur_j = 0, cur_l = 0;
    // Fragments clipped to bin 0 or bin n-1 are summed in registers and added
    // once per wave: as LDS atomics on one address they would serialise (a
    // uniform daughter table on a short grid clips about a thousand).
    double clip_lo = 0.0, clip_hi = 0.0;
    for (int it = tid; it < items; it += NT) {
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
            clip_lo += c;
        } else if (v >= xlast) {
            clip_hi += c;
        } else {
            int hi;
            double xl, xh;
            if (v >= xwlo && v < xwhi) {
                // the LDS copy holds the bracket, inside [wlo, whi]
                const int l0 = wlo - wbase, h0 = whi - wbase;
                int l = -1;
                if (q == cur_q) {
                    const int g = cur_l + (j - cur_j);
                    if (g >= l0 && g < h0 && xw[g] <= v && v < xw[g + 1]) l = g;
                }
                if (l < 0) {
                    int h = h0;
                    l = l0;
                    while (h - l > 1) {
                        const int m = l + ((h - l) >> 1);
                        if (xw[m] <= v) l = m; else h = m;
                    }
                }
                cur_q = q;
                cur_j = j;
                cur_l = l;
                hi = l + 1 + wbase;
                xl = xw[l];
                xh = xw[l + 1];
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
    // every lane of the wave is here (the loop above only skips items)
    clip_lo = wave_sum(clip_lo);
    clip_hi = wave_sum(clip_hi);
    if ((tid & (kWave - 1)) == 0) {
        if (clip_lo != 0.0) put(0, clip_lo);
        if (clip_hi != 0.0) put(n - 1, clip_hi);
    }
    __syncthreads();

    for (int i = tid; i < W; i += NT) {
        const int b = wbase + i;
        const double v = win[i];
        if (b >= 0 && b < n && v != 0.0) global_add(&a.rhs[b], v);
    }
}


}  // namespace pbe
