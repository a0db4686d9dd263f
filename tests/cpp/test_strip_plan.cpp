// test_strip_plan.cpp -- host check of the strip launch's work split
// (agg_strip.hpp): for many sizes, block widths, band modes, near-chunk
// counts, zero-birth chunk counts and grid sizes, every chunk of every block goes to exactly one
// workgroup share, the cost line matches the sum of its chunk costs, and the
// waves' row ranges of each visit tile the visit's rows in order.  Prints
// one JSON line; exit status 0 when every check holds.  Runs without a GPU.
#include "agg_strip.hpp"

#include <cstdio>
#include <vector>

using namespace pbe;

static int fails = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c) && fails++ < 20) std::fprintf(stderr, "fail %s line %d\n", #c, __LINE__); \
    } while (0)

static long long chunk_cost(const StripPlan& s, int q, int c)
{
    if (c < strip_sz(s, q)) return s.cz;
    if (c < strip_sf(s, q)) return s.cf;
    if (c < strip_se(s, q)) return s.cn;
    return s.cd;
}

int main()
{
    long long cases = 0, chunks_checked = 0;
    const int ns[] = {300, 1024, 3000, 4096, 9000, 16384, 20000, 40007, 65536};
    const int us[] = {1, 2, 4, 8};
    const int gs[] = {1, 7, 256, 512, 1000};
    for (int n : ns)
        for (int U : us)
            for (int band = 0; band < 2; ++band)
                for (int m_near : {0, 3, 4, 9, 40})
                    for (int zc : {-1, 0, 1, 5, 30, 100})
                    for (int G : gs) {
                        StripPlan s;
                        s.U = U;
                        s.nb = (n + 64 * U - 1) / (64 * U);
                        s.band = band;
                        s.m_near = m_near;
                        s.mq = m_near / U + 1;
                        // zc < 0: no zero-birth chunks; else m_zero = U + m_near + zc
                        s.zb = zc >= 0;
                        s.m_zero = zc >= 0 ? U + m_near + zc : 0;
                        s.mz = s.m_zero / U + 1;
                        s.cz = 1;
                        s.cf = 20;
                        s.cn = 50;
                        s.cd = 30;
                        s.cv = 100;
                        s.total = strip_total(s);
                        ++cases;
                        // cost line against the sum of its parts
                        long long cs = 0, tot = 0;
                        for (int q = 0; q < s.nb; ++q) {
                            CHECK(strip_cost_start(s, q) == cs);
                            CHECK(strip_prefix(s, q) == tot);
                            const int nc = strip_nc(s, q);
                            if (nc > 0) cs += s.cv;
                            for (int c = 0; c < nc; ++c) cs += chunk_cost(s, q, c);
                            tot += nc;
                            CHECK(strip_sf(s, q) <= strip_se(s, q) && strip_se(s, q) <= nc);
                            CHECK(strip_sz(s, q) <= strip_sf(s, q));
                        }
                        CHECK(strip_cost_total(s) == cs);
                        CHECK(s.total == tot);
                        // shares: each chunk exactly once, in order
                        std::vector<int> seen(tot, 0);
                        std::vector<long long> first(s.nb + 1, 0);
                        for (int q = 0; q < s.nb; ++q) first[q + 1] = first[q] + strip_nc(s, q);
                        const long long CT = strip_cost_total(s);
                        for (int g = 0; g < G; ++g) {
                            int qa, ca, qb, cb;
                            strip_chunk_at_cost(s, CT * g / G, qa, ca);
                            strip_chunk_at_cost(s, CT * (g + 1) / G, qb, cb);
                            CHECK(qa <= s.nb && qb <= s.nb);
                            if (qa > qb || (qa == qb && ca >= cb)) continue;
                            const int qe = cb > 0 ? qb + 1 : qb;
                            for (int q = qa, c = ca; q < qe; ++q, c = 0) {
                                const int hi = (q == qb) ? cb : strip_nc(s, q);
                                CHECK(hi <= strip_nc(s, q));
                                for (int cc = c; cc < hi; ++cc) ++seen[first[q] + cc];
                                // the visit's rows split over 8 waves by cost
                                const int T = 64 * U, k0 = q * T;
                                const int rmax = n < k0 + T ? n : k0 + T;
                                const int r0 = c * 64, r1 = hi * 64 < rmax ? hi * 64 : rmax;
                                if (r0 >= r1) continue;
                                const VisitRows vr(s, q, r0, r1);
                                const int vc = vr.total();
                                int prev = r0;
                                for (int w = 0; w < 8; ++w) {
                                    const int lo = vr.row_at(vc * w / 8), hi2 = vr.row_at(vc * (w + 1) / 8);
                                    CHECK(lo == prev && hi2 >= lo);
                                    prev = hi2;
                                }
                                CHECK(prev == r1);
                                CHECK(vr.cost(r0) == 0 && vr.cost(r1) == vc);
                            }
                        }
                        for (long long t = 0; t < tot; ++t) CHECK(seen[t] == 1);
                        chunks_checked += tot;
                    }
    std::printf("{\"cases\": %lld, \"chunks\": %lld, \"failures\": %d}\n", cases, chunks_checked,
                fails);
    return fails ? 1 : 0;
}
