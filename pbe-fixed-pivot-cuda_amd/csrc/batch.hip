// batch.hip -- thousands of independent PBE instances of uneven n in one
// persistent launch (no counterpart in the reference, which evaluates one
// instance per call).
//
// A plan, built once on the host, lists every unit of work of every instance
// -- aggregation pair tiles (T = 64 U, U from the autotuned batched table for
// the instance's size, kernel and grid reach) and breakage parent blocks (256
// parents) -- sorts them by estimated cost (longest first) and uploads them
// with per-instance descriptors (size, offset into the packed arrays, own
// grid lookup and own coefficients).  The persistent kernels run a fixed grid
// sized to the device; workgroup b takes items b, b + G, b + 2G, ... of the
// sorted list, which balances the load without a device-side queue (and keeps
// the launch free of counters that would need resetting under graph replay).
// Each item runs exactly the single-instance tile / block code.  Instances
// with tiles of different widths go to different launches (one per width:
// the tile's LDS footprint grows with U, so a launch of mixed widths would
// run every item at the widest one's occupancy).
#include <algorithm>
#include <new>
#include <vector>

#include "agg_tile.hpp"
#include "batch_plan.hpp"
#include "brk_block.hpp"
#include "pbe_internal.hpp"

namespace pbe {

struct BatchArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    const double* __restrict__ t_q;
    const double* __restrict__ bw_q;
    double* rhs;
    const BatchInstDev* __restrict__ inst;
    const BatchItem* __restrict__ items;
    int nitems;
    int nq;
};

template <int KT, int U>
__global__ __launch_bounds__(256)
void agg_batch_kernel(BatchArgs b)
{
    constexpr int T = AggShape<U>::T;
    __shared__ AggSmem<KT, U> sm;
    for (int it = blockIdx.x; it < b.nitems; it += gridDim.x) {
        const BatchItem w = b.items[it];
        const BatchInstDev d = b.inst[w.inst];
        AggArgs a;
        a.N = b.N + d.off;
        a.x = b.x + d.off;
        a.rhs = b.rhs + d.off;
        a.n = d.n;
        a.nb = d.nb;
        a.coef = KernelCoef{d.b0, d.b_br, d.b_sh};
        a.log_x0 = d.log_x0;
        a.inv_log_r = d.inv_log_r;
        int p, dg;
        decode_tile(w.unit, d.nb, p, dg);
        constexpr int S = AggShape<U>::SPLIT, RW = AggShape<U>::RW;
        if (w.slice >= 0)
            agg_tile_body<KT, U>(a, p * T, (p + dg) * T, dg == 0, p * T + w.slice * (T / S),
                                 RW / S, sm);
        else
            agg_tile_body<KT, U>(a, p * T, (p + dg) * T, dg == 0, p * T, RW, sm);
        __syncthreads();
    }
}

template <int SEL>
__global__ __launch_bounds__(256)
void brk_batch_kernel(BatchArgs b)
{
    __shared__ BrkSmem sm;
    for (int it = blockIdx.x; it < b.nitems; it += gridDim.x) {
        const BatchItem w = b.items[it];
        const BatchInstDev d = b.inst[w.inst];
        BrkArgs a{b.N + d.off, b.x + d.off, b.t_q, b.bw_q, b.rhs + d.off, d.n, b.nq,
                  d.S0, d.v_ref, d.alpha, d.v_min};
        brk_block_body<SEL>(a, w.unit * kBrkParents, sm);
        __syncthreads();
    }
}

}  // namespace pbe


namespace pbe {

static double tile_cost(int n, int T, int t)
{
    const long long nb = (n + T - 1) / T;
    // decode as the device does
    long long d = 0;
    while (d + 1 < nb && (d + 1) * nb - (d + 1) * d / 2 <= t) ++d;
    const long long p = t - (d * nb - d * (d - 1) / 2);
    const long long j0 = p * T, k0 = (p + d) * T;
    const double rows = (double)std::min<long long>(T, n - j0);
    const double cols = (double)std::min<long long>(T, n - k0);
    if (d == 0) return 3.0 * rows * (rows + 1) * 0.5;   // near pairs cost several far ones
    if (d == 1) return 2.0 * rows * cols;              // near band spill
    return rows * cols;
}

static int batch_group(int u)
{
    for (int g = 0; g < kBatchGroups; ++g)
        if (kBatchTiles[g] == u) return g;
    return -1;
}

hipError_t batch_plan_create(pbe_batch_plan* out, const pbe_batch_instance* hi, int count,
                             int kernel_type, int selection, int n_quad, int tile_u)
{
    if (!out) return hipErrorInvalidValue;
    if (tile_u != 0 && batch_group(tile_u) < 0) return hipErrorInvalidValue;
    *out = nullptr;
    if (count <= 0 || !hi) return hipErrorInvalidValue;
    if (kernel_type < -1 || kernel_type > 5 || selection < -1 || selection > 3)
        return hipErrorInvalidValue;
    if (kernel_type < 0 && selection < 0) return hipErrorInvalidValue;
    if (selection >= 0 && n_quad <= 0) return hipErrorInvalidValue;

    std::vector<BatchInstDev> inst(count);
    long long total = 0;
    for (int i = 0; i < count; ++i) {
        const pbe_batch_instance& s = hi[i];
        if (s.n <= 0 || s.offset < 0) return hipErrorInvalidValue;
        total = std::max<long long>(total, (long long)s.offset + s.n);
        if (total > 0x7fffffffLL) return hipErrorInvalidValue;
        BatchInstDev& d = inst[i];
        d.n = s.n;
        d.off = s.offset;
        d.tile_u = tile_u != 0 ? tile_u
                 : kernel_type >= 0 ? batch_tile_u(s.n, kernel_type, s.inv_log_r) : 2;
        if (batch_group(d.tile_u) < 0) d.tile_u = 2;
        d.nb = (s.n + kWave * d.tile_u - 1) / (kWave * d.tile_u);
        d.log_x0 = s.log_x0;
        d.inv_log_r = s.inv_log_r;
        d.b0 = s.beta0;
        d.b_br = s.beta_br;
        d.b_sh = s.beta_sh;
        d.S0 = s.S0;
        d.v_ref = s.v_ref;
        d.alpha = s.alpha;
        d.v_min = s.v_min;
    }

    struct Costed { double c; BatchItem w; };
    std::vector<Costed> agg[kBatchGroups], brk;
    if (kernel_type >= 0) {
        for (int i = 0; i < count; ++i) {
            const long long nb = inst[i].nb;
            const long long tiles = nb * (nb + 1) / 2;
            if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
            const int T = kWave * inst[i].tile_u;
            std::vector<Costed>& list = agg[batch_group(inst[i].tile_u)];
            constexpr int S = AggShape<1>::SPLIT;
            static_assert(AggShape<8>::SPLIT == S, "one band split for every tile width");
            for (long long t = 0; t < tiles; ++t) {
                const double c = tile_cost(inst[i].n, T, (int)t);
                if (t < band_tiles(nb)) {
                    for (int q = 0; q < S; ++q) list.push_back({c / S, BatchItem{i, (int)t, q, 0}});
                } else {
                    list.push_back({c, BatchItem{i, (int)t, -1, 0}});
                }
            }
        }
    }
    if (selection >= 0) {
        for (int i = 0; i < count; ++i) {
            const int blocks = (inst[i].n + kBrkParents - 1) / kBrkParents;
            for (int b = 0; b < blocks; ++b)
                brk.push_back({(double)std::min(kBrkParents, inst[i].n - b * kBrkParents) * n_quad,
                               BatchItem{i, b, -1, 0}});
        }
    }
    std::vector<Costed> rows;
    for (int i = 0; i < count; ++i) {
        const int blocks = (inst[i].n + kWave - 1) / kWave;
        for (int b = 0; b < blocks; ++b)
            rows.push_back({(double)std::min(kWave, inst[i].n - b * kWave) * inst[i].n,
                            BatchItem{i, b, -1, 0}});
    }
    auto by_cost = [](const Costed& a, const Costed& b) {
        if (a.c != b.c) return a.c > b.c;
        if (a.w.inst != b.w.inst) return a.w.inst < b.w.inst;
        return a.w.unit < b.w.unit;
    };
    for (auto& list : agg) std::sort(list.begin(), list.end(), by_cost);
    std::sort(brk.begin(), brk.end(), by_cost);
    std::sort(rows.begin(), rows.end(), by_cost);

    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
    }

    auto* p = new (std::nothrow) pbe_batch_plan_s();
    if (!p) return hipErrorOutOfMemory;
    p->count = count;
    p->total = (int)total;
    p->kernel_type = kernel_type;
    p->selection = selection;
    p->n_quad = n_quad;
    for (int g = 0; g < kBatchGroups; ++g) {
        if (agg[g].size() > 0x7fffffffULL) {
            delete p;
            return hipErrorInvalidValue;
        }
        p->n_agg_items[g] = (int)agg[g].size();
        for (const Costed& c : agg[g]) p->agg_cost[g] += c.c;
    }
    p->n_brk_items = (int)brk.size();
    p->grid = cus * 4;

    std::vector<BatchItem> items;
    hipError_t e = hipMalloc(&p->d_inst, sizeof(BatchInstDev) * count);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_inst, inst.data(), sizeof(BatchInstDev) * count, hipMemcpyHostToDevice);
    for (int g = 0; g < kBatchGroups && e == hipSuccess; ++g) {
        if (agg[g].empty()) continue;
        items.clear();
        for (auto& c : agg[g]) items.push_back(c.w);
        e = hipMalloc(&p->d_agg_items[g], sizeof(BatchItem) * items.size());
        if (e == hipSuccess)
            e = hipMemcpy(p->d_agg_items[g], items.data(), sizeof(BatchItem) * items.size(),
                          hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !brk.empty()) {
        items.clear();
        for (auto& c : brk) items.push_back(c.w);
        e = hipMalloc(&p->d_brk_items, sizeof(BatchItem) * items.size());
        if (e == hipSuccess)
            e = hipMemcpy(p->d_brk_items, items.data(), sizeof(BatchItem) * items.size(),
                          hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !rows.empty()) {
        items.clear();
        for (auto& c : rows) items.push_back(c.w);
        p->n_row_items = (int)items.size();
        e = hipMalloc(&p->d_row_items, sizeof(BatchItem) * items.size());
        if (e == hipSuccess)
            e = hipMemcpy(p->d_row_items, items.data(), sizeof(BatchItem) * items.size(),
                          hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        (void)batch_plan_destroy(p);
        return e;
    }
    *out = p;
    return hipSuccess;
}

hipError_t batch_plan_destroy(pbe_batch_plan p)
{
    if (!p) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    for (void* q : {(void*)p->d_inst, (void*)p->d_agg_items[0], (void*)p->d_agg_items[1],
                    (void*)p->d_agg_items[2], (void*)p->d_agg_items[3], (void*)p->d_brk_items,
                    (void*)p->d_row_items}) {
        if (q) {
            hipError_t e2 = hipFree(q);
            if (e == hipSuccess) e = e2;
        }
    }
    delete p;
    return e;
}

int batch_plan_total(pbe_batch_plan p) { return p ? p->total : 0; }

bool batch_plan_needs_table(pbe_batch_plan p) { return p && p->selection >= 0; }

template <int U>
static hipError_t launch_agg_batch_u(int kernel, const BatchArgs& a, int grid, hipStream_t s)
{
    const dim3 g((unsigned)grid), b(256);
    switch (kernel) {
    case 0: hipLaunchKernelGGL((agg_batch_kernel<0, U>), g, b, 0, s, a); break;
    case 1: hipLaunchKernelGGL((agg_batch_kernel<1, U>), g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL((agg_batch_kernel<2, U>), g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL((agg_batch_kernel<3, U>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((agg_batch_kernel<4, U>), g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL((agg_batch_kernel<5, U>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

static hipError_t launch_agg_batch(int kernel, int u, const BatchArgs& a, int grid,
                                   hipStream_t s)
{
    switch (u) {
    case 1: return launch_agg_batch_u<1>(kernel, a, grid, s);
    case 2: return launch_agg_batch_u<2>(kernel, a, grid, s);
    case 4: return launch_agg_batch_u<4>(kernel, a, grid, s);
    case 8: return launch_agg_batch_u<8>(kernel, a, grid, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t enqueue_batched_rhs(pbe_batch_plan p, const double* N, const double* x,
                               const double* t_q, const double* bw_q, double* rhs,
                               hipStream_t s)
{
    BatchArgs a{N, x, t_q, bw_q, rhs, p->d_inst, nullptr, 0, p->n_quad};
    if (p->kernel_type >= 0) {
        // costliest group first
        int order[kBatchGroups] = {0, 1, 2, 3};
        std::stable_sort(order, order + kBatchGroups,
                         [p](int u, int v) { return p->agg_cost[u] > p->agg_cost[v]; });
        for (int g : order) {
            if (p->n_agg_items[g] == 0) continue;
            a.items = p->d_agg_items[g];
            a.nitems = p->n_agg_items[g];
            const hipError_t e = launch_agg_batch(p->kernel_type, kBatchTiles[g], a,
                                                  std::min(p->grid, a.nitems), s);
            if (e != hipSuccess) return e;
        }
    }
    if (p->selection >= 0 && p->n_brk_items > 0) {
        a.items = p->d_brk_items;
        a.nitems = p->n_brk_items;
        const dim3 g((unsigned)std::min(p->grid, p->n_brk_items)), b(256);
        switch (p->selection) {
        case 0: hipLaunchKernelGGL(brk_batch_kernel<0>, g, b, 0, s, a); break;
        case 1: hipLaunchKernelGGL(brk_batch_kernel<1>, g, b, 0, s, a); break;
        case 2: hipLaunchKernelGGL(brk_batch_kernel<2>, g, b, 0, s, a); break;
        case 3: hipLaunchKernelGGL(brk_batch_kernel<3>, g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace pbe
