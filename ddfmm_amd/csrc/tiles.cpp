// GEMM tiling after planning: per-stage work accounting and the rebuild of the tiles of
// every logical GEMM stage under a tile-height limit (the autotuner's search space).
#include <algorithm>
#include <stdexcept>
#include <string>

#include "planner.hpp"

namespace ddfmm {

void stage_work(const HostPlan& pl, Stage& s) {
    double w = 0;
    s.exec = 0;
    if (s.kind == ST_GEMM) {
        // useful flops of each tile: its rows x (K, or the present K blocks) x columns
        for (int t = s.begin; t < s.end; ++t) {
            const GTile& tl = pl.tiles[t];
            const GGroup& g = pl.groups[tl.g];
            const double rows = std::min(16 * tl.mb, g.M - tl.m0);
            const int n1 = std::min(tl.n0 + GEMM_TN, g.ncols);
            double k = 0;
            if (g.kb == 0) {
                k = (double)g.K * (n1 - tl.n0);
            } else {
                const int nkb = g.K / g.kb;
                for (int n = tl.n0; n < n1; ++n)
                    for (int h = 0; h < nkb; ++h) k += pl.colB[g.bcol0 + (int64_t)n * nkb + h] >= 0 ? g.kb : 0;
            }
            w += 8.0 * rows * k;
            // executed: per segment (K, or each K block present in some column), every
            // wave with a column holding it runs ceil(len / 4) k steps of 4 * mb MFMAs
            const int nkb = g.kb ? g.K / g.kb : 1;
            const int len = g.kb ? g.kb : g.K;
            for (int h = 0; h < nkb; ++h) {
                int waves = 0;
                for (int wv = 0; wv < 4; ++wv) {
                    bool run = false;
                    for (int n = tl.n0 + 16 * wv; n < std::min(tl.n0 + 16 * wv + 16, g.ncols) && !run; ++n)
                        run = g.kb == 0 || pl.colB[g.bcol0 + (int64_t)n * nkb + h] >= 0;
                    waves += run;
                }
                s.exec += waves * (double)((len + 3) / 4) * 4 * tl.mb * 2048.0;
            }
        }
    } else if (s.kind == ST_KAPPLY) {
        for (int t = s.begin; t < s.end; ++t) {
            const KTask& k = pl.ktasks[t];
            for (int q = k.s0; q < k.s1; ++q)   // a mirrored source: both directions
                w += (double)k.count * pl.ksrcs[q].count * (pl.ksrcs[q].moff >= 0 ? 2 : 1);
        }
    } else if (s.kind == ST_FFTFWD) {
        w = s.end - s.begin;
    } else if (s.kind == ST_VLIST) {
        for (int t = s.begin; t < s.end; ++t) w += pl.vtrg[t].p1 - pl.vtrg[t].p0;
    } else if (s.kind == ST_REDUCE) {
        for (int t = s.begin; t < s.end; ++t) w += (double)pl.rtrg[t].cnt * pl.rtrg[t].M;
    }
    s.work = w;
}

void retile(HostPlan& pl, const std::vector<int>& mbmax) {
    if (mbmax.size() != pl.lgemm.size()) throw std::runtime_error("retile: one height limit per logical GEMM stage");
    for (int m : mbmax)
        if (m < 1 || m > GEMM_MB) throw std::runtime_error("retile: tile height limit out of range [1, 7]");
    std::vector<GTile> tiles;
    std::vector<Stage> stages;
    std::vector<char> done(pl.lgemm.size(), 0);
    for (const Stage& s : pl.stages) {
        if (s.kind != ST_GEMM) {
            stages.push_back(s);
            continue;
        }
        if (s.lstage < 0 || s.lstage >= (int)pl.lgemm.size()) throw std::runtime_error("retile: GEMM stage without logical stage");
        if (done[s.lstage]) continue;          // the other tile heights of a logical stage
        done[s.lstage] = 1;
        LogicalGemm& lg = pl.lgemm[s.lstage];
        lg.mbmax = mbmax[s.lstage];
        const int b = (int)tiles.size();
        for (int gi : lg.groups) {
            const GGroup& g = pl.groups[gi];
            const int mbt = (g.M + 15) / 16, nmt = (mbt + lg.mbmax - 1) / lg.mbmax;
            const int nt = (g.ncols + GEMM_TN - 1) / GEMM_TN;
            for (int a = 0, b0 = 0; a < nmt; ++a) {
                const int mb = (mbt - b0 + (nmt - a) - 1) / (nmt - a);
                for (int c = 0; c < nt; ++c) tiles.push_back(GTile{gi, 16 * b0, c * GEMM_TN, mb});
                b0 += mb;
            }
        }
        std::stable_sort(tiles.begin() + b, tiles.end(), [](const GTile& x, const GTile& y) { return x.mb < y.mb; });
        for (int t = b; t < (int)tiles.size();) {
            int u = t;
            while (u < (int)tiles.size() && tiles[u].mb == tiles[t].mb) ++u;
            Stage q = s;
            q.begin = t;
            q.end = u;
            q.mb = tiles[t].mb;
            stages.push_back(q);
            t = u;
        }
    }
    for (const GTile& t : tiles) {
        const GGroup& g = pl.groups[t.g];
        if (!(t.mb >= 1 && t.mb <= GEMM_MB && t.m0 % 16 == 0 && t.m0 < g.M && t.m0 + 16 * (t.mb - 1) < g.M &&
              t.n0 < g.ncols))
            throw std::runtime_error("retile: tile out of bounds");
    }
    pl.tiles.swap(tiles);
    pl.stages.swap(stages);
    pl.setup_stages.clear();
    for (Stage& s : pl.stages) stage_work(pl, s);
}

}  // namespace ddfmm
