// Flattening of Wave3d::eval (wave3d_eval.cpp:299-1040) into stage descriptors.
//
// Stage order (one launch each):
//   gather densities into leaf order                                  (eval 319-333)
//   LF upward per width, finest first: S2M, M2M, pinv V*diag(1/s)*U^H   (EvalUpwardLow 418-490)
//   HF upward per width, 1 -> max: U2U, E3, E2, E1                      (EvalUpwardHigh 773-854)
//   HF downward per width, max -> 1: M2L by (width, offset), E1^T,E2^T,E3^T, L2L
//                                                    (HighFrequencyM2L/L2L 876-993)
//   LF downward per width, coarsest first: FFT, V list, X list, pinv^T, L2L
//                                                    (EvalDownwardLow 493-583)
//   leaf kernel: U list, W list, X list (direct), L2T; scatter to input order
//
// The per-box zgemv calls of the reference become grouped GEMMs whose columns are
// (box, rhs) pairs; the translation matrices are either read from the operator
// library or generated once on the device (KGen) and kept resident.
#include "planner.hpp"

#include <algorithm>
#include <set>
#include <cstring>
#include <sstream>
#include <string>

namespace ddfmm {

namespace {

const double EPS = 1e-12;

uint64_t morton(const BoxKey& k) {
    uint64_t code = 0;
    for (int b = 0; b < 20; ++b)
        for (int d = 0; d < 3; ++d)
            code |= (uint64_t)((k.second[d] >> b) & 1) << (3 * b + (2 - d));
    return code;
}

std::vector<BoxKey> morton_sorted(std::vector<BoxKey> v) {
    std::stable_sort(v.begin(), v.end(), [](const BoxKey& a, const BoxKey& b) {
        return a.first != b.first ? a.first < b.first : morton(a) < morton(b);
    });
    return v;
}

// C = A * B for column-major complex matrices (setup-time products of solve factors)
CMat matmul(const CMat& A, const CMat& B) {
    if (A.n != B.m) throw Error("internal: factor size mismatch");
    CMat C;
    C.m = A.m; C.n = B.n;
    C.a.assign((size_t)C.m * C.n, cpx(0, 0));
    for (int j = 0; j < B.n; ++j)
        for (int k = 0; k < A.n; ++k) {
            const cpx b = B.a[k + (size_t)j * B.m];
            const cpx* a = &A.a[(size_t)k * A.m];
            cpx* c = &C.a[(size_t)j * C.m];
            for (int i = 0; i < A.m; ++i) c[i] += a[i] * b;
        }
    return C;
}

// Column order for a K-blocked GEMM whose columns skip absent K blocks (child masks): a
// wave of 16 columns multiplies the union of its columns' blocks, a tile of four waves
// loads the union of theirs.  Columns of one mask fill whole waves first (masks in
// Gray-code order); the remainders are packed greedily, the next column always the first
// one (by popcount, then Gray code) that adds the fewest new blocks to the wave; waves are
// then grouped into tiles the same way.  Deterministic for a given input order.
template <class K>
void pack_by_mask(std::vector<std::pair<int, K>>& cols) {
    auto gray = [](int m) { return m ^ (m >> 1); };
    std::map<int, std::vector<size_t>> by;     // gray code -> column indices in input order
    for (size_t i = 0; i < cols.size(); ++i) by[gray(cols[i].first)].push_back(i);
    std::vector<size_t> order, left;
    order.reserve(cols.size());
    for (auto& kv : by) {
        const size_t full = kv.second.size() / 16 * 16;
        order.insert(order.end(), kv.second.begin(), kv.second.begin() + full);
        left.insert(left.end(), kv.second.begin() + full, kv.second.end());
    }
    std::stable_sort(left.begin(), left.end(), [&](size_t a, size_t b) {
        return __builtin_popcount(cols[a].first) > __builtin_popcount(cols[b].first);
    });
    std::vector<char> used(left.size(), 0);
    for (size_t i = 0; i < left.size(); ++i) {
        if (used[i]) continue;
        used[i] = 1;
        order.push_back(left[i]);
        int u = cols[left[i]].first;
        for (int n = 1; n < 16; ++n) {
            int best = -1, bg = 99;
            for (size_t j = i + 1; j < left.size() && bg > 0; ++j) {
                if (used[j]) continue;
                const int g = __builtin_popcount(u | cols[left[j]].first) - __builtin_popcount(u);
                if (g < bg) { bg = g; best = (int)j; }
            }
            if (best < 0) break;
            used[best] = 1;
            order.push_back(left[best]);
            u |= cols[left[best]].first;
        }
    }
    // waves of 16 into tiles of 4 (GEMM_TN columns): a tile loops over the union of its
    // waves' blocks, so each tile takes the first free wave and the three that add the
    // fewest blocks to it; a partial wave (the last one) stays last
    std::vector<std::pair<size_t, int>> waves;      // (first position in order, union mask)
    for (size_t s = 0; s < order.size(); s += 16) {
        int u = 0;
        for (size_t j = s; j < std::min(order.size(), s + 16); ++j) u |= cols[order[j]].first;
        waves.push_back({s, u});
    }
    const bool partial = order.size() % 16 != 0;
    const size_t nfull = waves.size() - (partial ? 1 : 0);
    std::vector<size_t> wseq;
    std::vector<char> wused(nfull, 0);
    for (size_t i = 0; i < nfull; ++i) {
        if (wused[i]) continue;
        wused[i] = 1;
        wseq.push_back(i);
        int u = waves[i].second;
        for (int n = 1; n < GEMM_TN / 16; ++n) {
            int best = -1, bg = 99;
            for (size_t j = i + 1; j < nfull && bg > 0; ++j) {
                if (wused[j]) continue;
                const int g = __builtin_popcount(u | waves[j].second) - __builtin_popcount(u);
                if (g < bg) { bg = g; best = (int)j; }
            }
            if (best < 0) break;
            wused[best] = 1;
            wseq.push_back(best);
            u |= waves[best].second;
        }
    }
    if (partial) wseq.push_back(waves.size() - 1);
    std::vector<std::pair<int, K>> out;
    out.reserve(cols.size());
    for (size_t w : wseq)
        for (size_t j = waves[w].first; j < std::min(order.size(), waves[w].first + 16); ++j)
            out.push_back(cols[order[j]]);
    cols.swap(out);
}

// which independent sections run on the second stream
constexpr bool SECOND_STREAM_NEAR = true;   // near field (U list, direct W/X lists)
constexpr bool SECOND_STREAM_LF = true;     // FFT, V list, X list per LF width

// LF operators of one width (offsets into the point and operator tables)
struct LfW {
    const LowFreqEntry* le;
    int r;
    int64_t uep_p, ucp_p;          // kapply point offsets
    int64_t pinv, V, invs, UH, tab;   // opmat offsets (fused solve or its factors)
    int64_t ue2uc[8];              // generated
    int grid;
    bool has_children;
};

struct HfDir {          // one (width, direction) group
    double W;
    Index3 dir, srt, sgn, prm;
    const HighFreqEntry* he;
    int r1, r2, r1c;    // equivalent / check / child-equivalent sizes
    int64_t uep_g, ucp_g, uepc_g;   // generator point offsets (shuffled)
    int64_t ue2uc[8];   // generated matrices (relative offsets)
    int64_t de2dc[8];   // W = 1 only: G(uep_LF(W/2) + s_c, -ucp) (not a transpose of ue2uc)
    int64_t e1, e2, e3; // fused: e1 = E1 E2 E3 (r1 x r2); factored: E1, E2, E3
    // fused mode: solves folded into the translations (relative generated offsets)
    int64_t F = -1;     // F_c = (E1 E2 E3) ue2uc_c, r1 x r1c each, contiguous over c
    int64_t D[8];       // W = 1: D_c = de2dc_c (E1 E2 E3)^T, r1c x r1
    int64_t eT = -1;    // W = 1: (E1 E2 E3)^T (r2 x r1), uploaded
    std::vector<BoxKey> out, in;
};

// E1..E3 (or their product) of one canonical HF entry
struct EOff { int64_t e1, e2, eT; };
// one reduction target of the HF downward pass: its contributions in adjacent slots
struct RedT {
    int64_t out_off;
    int obuf, M, mode;
    int cnt = 0, next = 0;
    int64_t base = 0;
};

// W = 1 -> LF: incoming directions of one class, bundled by up to 8
struct Bundle {
    std::vector<const HfDir*> d;
    std::map<BoxKey, int> cols;
};

// HF M2L pairs of one (W, centre offset)
struct M2LGroup {
    const HfDir* h = nullptr;
    int64_t mat;
    std::vector<int64_t> b;
    std::vector<int> t;      // reduction target of each pair
};


struct Builder {
    const Params& prm;
    const Mlib& mlib;
    Tree tree;
    HostPlan& pl;
    int64_t gen_cursor = 0;     // generated part of opmat (appended after upload part)
    std::vector<std::pair<int64_t, int64_t>> gen_ranges;

    Builder(const Params& p, const Mlib& m, HostPlan& h) : prm(p), mlib(m), pl(h) {}

    // ---- helpers
    int64_t upload(const CMat& m) {
        int64_t off = (int64_t)pl.opmat.size();
        pl.opmat.insert(pl.opmat.end(), m.a.begin(), m.a.end());
        return off;
    }
    int64_t genpts(const std::vector<double>& xyz) {
        int64_t off = (int64_t)pl.genpts.size() / 3;
        pl.genpts.insert(pl.genpts.end(), xyz.begin(), xyz.end());
        return off;
    }
    static std::vector<double> rows(const DMat& d) {   // 3 x n column-major -> n x 3 rows
        std::vector<double> r((size_t)d.n * 3);
        for (int k = 0; k < d.n; ++k)
            for (int j = 0; j < 3; ++j) r[3 * (size_t)k + j] = d(j, k);
        return r;
    }
    int64_t add_pts(const DMat& d) {                   // into px/py/pz (kapply point table)
        int64_t off = (int64_t)pl.px.size();
        for (int k = 0; k < d.n; ++k) {
            pl.px.push_back(d(0, k));
            pl.py.push_back(d(1, k));
            pl.pz.push_back(d(2, k));
        }
        return off;
    }
    // B offsets: ncols entries (kb = 0, contiguous columns) or ncols x ceil(K/kb) block
    // offsets (kb > 0, -1 = zero block); C offsets: ncols entries
    int64_t reserve(int64_t n) {        // generated-part space written by a setup GEMM
        int64_t off = gen_cursor;
        gen_cursor += n;
        return off;
    }
    int add_group(GGroup g, const std::vector<int64_t>& bc, const std::vector<int64_t>& cc,
                  double* flops, int kb = 0, const std::vector<int64_t>& ablk = {}, bool count = true) {
        if (cc.empty()) return -1;
        if (kb > 0 && (g.K % kb != 0 || g.K / kb > 8))
            throw Error("internal: K must split into at most 8 equal blocks");
        g.kb = kb;
        g.acol0 = (int)pl.colA.size();
        pl.colA.insert(pl.colA.end(), ablk.begin(), ablk.end());
        g.ncols = (int)cc.size();
        g.col0 = (int)pl.colC.size();
        pl.colC.insert(pl.colC.end(), cc.begin(), cc.end());
        g.bcol0 = (int)pl.colB.size();
        pl.colB.insert(pl.colB.end(), bc.begin(), bc.end());
        int gi = (int)pl.groups.size();
        pl.groups.push_back(g);
        pl.group_flops.push_back(0);
        // rows in blocks of 16 spread evenly over ceil(blocks / GEMM_MB) tiles
        const int mbt = (g.M + 15) / 16, nmt = (mbt + GEMM_MB - 1) / GEMM_MB;
        const int nt = (g.ncols + GEMM_TN - 1) / GEMM_TN;
        for (int a = 0, b0 = 0; a < nmt; ++a) {
            const int mb = (mbt - b0 + (nmt - a) - 1) / (nmt - a);
            for (int b = 0; b < nt; ++b) pl.tiles.push_back(GTile{gi, 16 * b0, b * GEMM_TN, mb});
            b0 += mb;
        }
        double f = 8.0 * g.M * g.K * g.ncols;
        if (kb > 0) {                        // useful work: present blocks only
            int64_t present = 0;
            for (int64_t v : bc) present += v >= 0;
            f = 8.0 * g.M * kb * (double)present;
        }
        if (count) pl.stats.gemm_flops += f;
        pl.group_flops[gi] = f;
        if (flops) *flops += f;
        return gi;
    }
    // Stages are emitted in dependency order per section and merged at the end:
    //   0 gather | fork 0 | 1 (stream 1, unused) | 2 LF upward | fork 1 |
    //   3 FFT, V list, X list per LF width, near field (stream 1) | 4 HF upward |
    //   5 HF downward | join | 6 W=1 -> LF reduction, LF downward, leaf far field, scatter
    // (the near field waits for the LF upward pass: running beside it slows that short
    // chain more than it gains)
    int cur_sec = 0;
    std::vector<int> stage_sec;
    void stage(int kind, int b, int e, const char* name, int buf = -1, int64_t count = 0) {
        const size_t n0 = pl.stages.size();
        stage_emit(kind, b, e, name, buf, count);
        stage_sec.resize(pl.stages.size(), cur_sec);
        (void)n0;
    }
    void merge_sections() {
        std::vector<Stage> out;
        auto sync = [&](int kind, int ev, const char* nm) {
            Stage q;
            std::memset(&q, 0, sizeof(q));
            q.kind = kind; q.buf = ev;
            std::strncpy(q.name, nm, sizeof(q.name) - 1);
            out.push_back(q);
        };
        auto emit = [&](int sec) {
            for (size_t i = 0; i < pl.stages.size(); ++i)
                if (stage_sec[i] == sec) {
                    Stage q = pl.stages[i];
                    q.stream = ((sec == 1 && SECOND_STREAM_NEAR) || (sec == 3 && SECOND_STREAM_LF)) ? 1 : 0;
                    out.push_back(q);
                }
        };
        emit(0);
        sync(ST_FORK, 0, "fork");
        emit(1);
        emit(2);
        sync(ST_FORK, 1, "fork");
        emit(3);
        emit(4);
        emit(5);
        sync(ST_JOIN, 2, "join");
        emit(6);
        pl.stages.swap(out);
        stage_sec.clear();
    }
    // one launch per tile height (a kernel instance each): tiles sorted by height, the
    // tiles of a group kept together
    void push_gemm_stages(std::vector<Stage>& dst, const Stage& s) {
        std::stable_sort(pl.tiles.begin() + s.begin, pl.tiles.begin() + s.end,
                         [](const GTile& x, const GTile& y) { return x.mb < y.mb; });
        for (int t = s.begin; t < s.end;) {
            int u = t;
            while (u < s.end && pl.tiles[u].mb == pl.tiles[t].mb) ++u;
            if (pl.tiles[t].mb < 1 || pl.tiles[t].mb > GEMM_MB) throw Error("internal: tile height out of range");
            Stage q = s;
            q.begin = t; q.end = u; q.mb = pl.tiles[t].mb;
            dst.push_back(q);
            t = u;
        }
    }
    void stage_emit(int kind, int b, int e, const char* name, int buf, int64_t count) {
        if (kind != ST_ZERO && kind != ST_GATHER && kind != ST_SCATTER && e <= b) return;
        Stage s;
        std::memset(&s, 0, sizeof(s));
        s.kind = kind; s.begin = b; s.end = e; s.buf = buf; s.count = count; s.lstage = -1;
        std::strncpy(s.name, name, sizeof(s.name) - 1);
        // solves: LF pinv factors and HF E1/E2/E3 chains
        const std::string nm(name);
        s.prec = (nm.find("uc2ue") != std::string::npos || nm.find("dc2de") != std::string::npos ||
                  nm.find("hf_up_e") != std::string::npos || nm.find("hf_dn_e") != std::string::npos)
                     ? PREC_SOLVE : PREC_TRANSLATION;
        if (kind == ST_KAPPLY) {
            // most expensive tasks first: the blocks are dispatched in task order, round robin
            // over the XCDs, so the long leaves (up to 400 targets, up to 27 neighbours of
            // as many points) start early and the launch ends on short ones.  Every task
            // owns its target range, so the order does not change any result.
            // (a source evaluated for both directions costs about 5/4 of a one-way source)
            std::vector<std::pair<int64_t, int>> order;
            for (int i = b; i < e; ++i) {
                const KTask& t = pl.ktasks[i];
                int64_t c = 0;
                for (int q = t.s0; q < t.s1; ++q) {
                    c += pl.ksrcs[q].count * (pl.ksrcs[q].moff >= 0 ? 5 : 4);
                    if (pl.ksrcs[q].moff >= 0) s.mb = 1;
                }
                order.push_back({-c * t.count, i});
            }
            std::stable_sort(order.begin(), order.end(),
                             [](const std::pair<int64_t, int>& x, const std::pair<int64_t, int>& y) {
                                 return x.first < y.first;
                             });
            std::vector<KTask> sorted;
            for (auto& o : order) sorted.push_back(pl.ktasks[o.second]);
            std::copy(sorted.begin(), sorted.end(), pl.ktasks.begin() + b);
        }
        if (kind == ST_GEMM) {
            // the logical stage keeps its groups and its tiling, so that the tiles can be
            // rebuilt with another height limit (retile)
            LogicalGemm lg;
            for (int t = b; t < e; ++t)
                if (lg.groups.empty() || lg.groups.back() != pl.tiles[t].g) lg.groups.push_back(pl.tiles[t].g);
            std::strncpy(lg.name, name, sizeof(lg.name) - 1);
            resplit_rows(b, e);
            s.end = (int)pl.tiles.size();
            lg.mbmax = 0;
            for (int t = b; t < s.end; ++t) lg.mbmax = std::max(lg.mbmax, pl.tiles[t].mb);
            s.lstage = (int)pl.lgemm.size();
            pl.lgemm.push_back(lg);
            push_gemm_stages(pl.stages, s);
            return;
        }
        pl.stages.push_back(s);
    }
    // A GEMM stage whose launch (tiles x right-hand sides) has fewer than MIN_GEMM_TILES
    // tiles (an LF level's M2M over a few hundred boxes) would leave most CUs idle: its
    // groups' rows are split into proportionally more, shorter tiles (the tiles of the
    // range are the last ones added).
    static constexpr int MIN_GEMM_TILES = 2048;
    void resplit_rows(int b, int e) {
        if ((int64_t)(e - b) * std::max(1, prm.nrhs_max) >= MIN_GEMM_TILES || e != (int)pl.tiles.size() || e <= b)
            return;
        std::vector<int> gs;
        for (int t = b; t < e; ++t)
            if (gs.empty() || gs.back() != pl.tiles[t].g) gs.push_back(pl.tiles[t].g);
        const int64_t want = MIN_GEMM_TILES / std::max(1, prm.nrhs_max);
        const int f = (int)((want + (e - b) - 1) / (e - b));
        pl.tiles.resize(b);
        for (int gi : gs) {
            const GGroup& g = pl.groups[gi];
            const int mbt = (g.M + 15) / 16;
            const int nmt = std::min(mbt, (mbt + GEMM_MB - 1) / GEMM_MB * f);
            const int nt = (g.ncols + GEMM_TN - 1) / GEMM_TN;
            for (int a = 0, b0 = 0; a < nmt; ++a) {
                const int mb = (mbt - b0 + (nmt - a) - 1) / (nmt - a);
                for (int c = 0; c < nt; ++c) pl.tiles.push_back(GTile{gi, 16 * b0, c * GEMM_TN, mb});
                b0 += mb;
            }
        }
    }
    int64_t gen(int64_t toff, int m, const double ts[3], int tneg, int64_t soff, int n,
                const double ss[3]) {
        KGen k;
        std::memset(&k, 0, sizeof(k));
        k.toff = toff; k.soff = soff; k.m = m; k.n = n; k.tneg = tneg;
        for (int d = 0; d < 3; ++d) { k.tshift[d] = ts[d]; k.sshift[d] = ss[d]; }
        k.out = gen_cursor;       // relative; made absolute after the upload part is known
        gen_cursor += (int64_t)m * n;
        pl.kgen.push_back(k);
        return k.out;
    }
    static void shift_of(int ind, double W, double s[3]) {       // ShiftedPoint, mlib3d.cpp:22-30
        s[0] = (((ind & 4) >> 2) - 0.5) * W / 2;
        s[1] = (((ind & 2) >> 1) - 0.5) * W / 2;
        s[2] = ((ind & 1) - 0.5) * W / 2;
    }

    void run(const double* pos, int64_t N, bool want_dump);
    // ---- planning state shared by the phases of run()
    int P = 0, G = 0, G3 = 0;
    std::map<double, std::vector<BoxKey>> ldmap;                                  // LF boxes per width
    std::map<Index3, std::pair<std::vector<BoxKey>, std::vector<BoxKey>>> hdmap;  // HF out / in boxes per direction
    std::map<BoxKey, int64_t> lfidx, ptoff;
    std::vector<BoxKey> leaves;
    std::map<double, LfW> lfw;
    int NL = 0;
    int64_t nlf = 0;
    std::map<double, std::vector<HfDir>> hfl;
    std::map<std::pair<BoxKey, Index3>, int64_t> outslot, inslot, oscr, iscr;
    int64_t hf_up_size = 0, hf_dn_size = 0, scr_size = 0;
    std::map<std::pair<double, Index3>, EOff> hf_entry_off;
    std::vector<RedT> rts;
    std::map<double, std::pair<int, int>> hf_region;   // W -> target range of the level
    std::map<std::pair<BoxKey, Index3>, int> hf_rt;    // in-slot -> target
    std::map<BoxKey, int> lf_rt;                       // LF child of a width-1 box -> target
    std::vector<Bundle> w1b;
    int lf_r0 = -1, lf_r1 = 0;
    std::map<double, std::map<Index3, M2LGroup>> m2l;
    int64_t red_size = 0, up_n = 0;
    int64_t spec_max = 0, vtrg_max = 0;

    int npts(const BoxKey& k) const { return (int)tree.boxes.at(k).ptidx.size(); }
    int64_t lslot(const BoxKey& k) const { return lfidx.at(k) * NL; }
    int64_t G_(int64_t rel) const { return rel + up_n; }     // generated-table offset -> table offset
    int l2l_target(const HfDir& h, const BoxKey& ck);
    bool child_present(const BoxKey& ck) const;
    void assign(int r0, int r1);
    int64_t slot(int ri);
    void reduce_stage(int r0, int r1, const char* name);

    void init(const double* pos, int64_t N);
    void collect_boxes(const double* pos, int64_t N);
    void lf_operators();
    void hf_groups();
    void hf_reductions();
    void hf_m2l_pairs();
    void assign_slots();
    void operator_layout();
    void buffer_sizes();
    void stages_lf_upward();
    void stages_hf_upward();
    void stages_hf_downward();
    void stages_lf_downward();
    void stages_leaf();
    void finish(bool want_dump);

    // Every descriptor's element range must lie inside the buffer or table it indexes: the
    // kernels trust these offsets, so a planner mistake is caught here on the host.
    void validate() const {
        auto need = [](bool ok, const char* what) {
            if (!ok) throw Error(std::string("internal: descriptor out of bounds (") + what + ")");
        };
        const int64_t npt = (int64_t)pl.px.size();
        const int64_t G3 = (int64_t)(2 * pl.P) * (2 * pl.P) * (2 * pl.P);
        for (const Stage& st : pl.stages) {
            need(st.begin >= 0 && st.begin <= st.end, "stage range");
            if (st.kind == ST_GEMM) {
                need(st.end <= (int64_t)pl.tiles.size(), "gemm stage");
                for (int t = st.begin; t < st.end; ++t) need(pl.tiles[t].mb == st.mb, "gemm stage tile height");
            }
            if (st.kind == ST_KAPPLY) need(st.end <= (int64_t)pl.ktasks.size(), "kapply stage");
            if (st.kind == ST_FFTFWD) need(st.end <= (int64_t)pl.fsrc.size(), "fft stage");
            if (st.kind == ST_VLIST) {
                need(st.end <= (int64_t)pl.vtrg.size(), "vlist stage");
                need((int64_t)(st.end - st.begin) * G3 <= pl.bufsize[BUF_TSPEC], "vlist target spectra");
                for (int t = st.begin; t < st.end; ++t)
                    need(pl.vtrg[t].tab_off == pl.vtrg[st.begin].tab_off, "vlist stage table");
                // chunks tile the stage's targets in order; every pair's local index names
                // its own source in the chunk's union
                need(st.buf >= 0 && st.count >= 1 && st.buf + st.count <= (int64_t)pl.vchunk.size(), "vlist chunks");
                int next = st.begin;
                for (int64_t c = st.buf; c < st.buf + st.count; ++c) {
                    const VChunk& ch = pl.vchunk[c];
                    need(ch.t0 == next && ch.t0 < ch.t1 && ch.t1 <= st.end, "vlist chunk targets");
                    need(ch.u0 >= 0 && ch.u0 <= ch.u1 && ch.u1 - ch.u0 <= VCH_SRC &&
                             ch.u1 <= (int)pl.vsrc.size(), "vlist chunk sources");
                    need(ch.k0 >= 0 && ch.k0 % 8 == 0 && ch.k0 <= ch.k1 && ch.k1 - ch.k0 <= VCH_PK &&
                             ch.k1 <= (int)pl.vpk.size(), "vlist chunk pairs");
                    for (int u = ch.u0; u < ch.u1; ++u)
                        need(pl.vsrc[u] >= 0 && (int64_t)pl.vsrc[u] * G3 + G3 <= pl.bufsize[BUF_SPEC], "vlist chunk source");
                    for (int t = ch.t0; t < ch.t1; ++t) {
                        const VTarget& v = pl.vtrg[t];
                        const int n8 = (v.p1 - v.p0 + 7) & ~7;
                        need(v.k0 >= ch.k0 && v.k0 % 8 == 0 && v.k0 + n8 <= ch.k1, "vlist packed range");
                        for (int j = 0; j < n8; ++j) {
                            const int w = pl.vpk[v.k0 + j], l = w & 0xffff, o = w >> 16;
                            if (j >= v.p1 - v.p0) {
                                need(w == 0, "vlist packed padding");
                                continue;
                            }
                            const VPair& pr = pl.vpair[v.p0 + j];
                            need(l >= 1 && l < ch.u1 - ch.u0 && (int64_t)pl.vsrc[ch.u0 + l] * G3 == pr.spec_off &&
                                     o == pr.offidx, "vlist packed pair");
                        }
                    }
                    next = ch.t1;
                }
                need(next == st.end, "vlist chunks cover the stage");
            }
            if (st.kind == ST_REDUCE) need(st.end <= (int64_t)pl.rtrg.size(), "reduce stage");
            if (st.kind == ST_ZERO) need(st.buf >= 0 && st.buf < BUF_COUNT && st.count <= pl.bufsize[st.buf], "zero");
            if (st.kind == ST_FORK || st.kind == ST_JOIN) need(st.buf >= 0 && st.buf < STAGE_EVENTS, "sync event");
            need(st.stream == 0 || st.stream == 1, "stage stream");
        }
        for (const GTile& t : pl.tiles) {
            need(t.g >= 0 && t.g < (int)pl.groups.size(), "tile group");
            const GGroup& g = pl.groups[t.g];
            need(t.mb >= 1 && t.mb <= GEMM_MB && t.m0 >= 0 && t.m0 % 16 == 0 && t.m0 < g.M &&
                 t.m0 + 16 * (t.mb - 1) < g.M && t.n0 >= 0 && t.n0 < g.ncols, "tile index");
        }
        auto bsize = [&](int b) { return b == BUF_OPMAT ? pl.opmat_total : pl.bufsize[b]; };
        for (const Stage& st : pl.setup_stages) {
            need(st.kind == ST_GEMM && st.begin >= 0 && st.begin <= st.end && st.end <= (int64_t)pl.tiles.size(),
                 "setup stage");
            for (int t = st.begin; t < st.end; ++t) {
                need(pl.tiles[t].mb == st.mb, "setup tile height");
                const GGroup& g = pl.groups[pl.tiles[t].g];
                need(g.bbuf == BUF_OPMAT && g.cbuf == BUF_OPMAT, "setup buffers");
            }
        }
        for (const GGroup& g : pl.groups) {
            need(g.bbuf >= 0 && g.bbuf <= BUF_OPMAT && g.cbuf >= 0 && g.cbuf <= BUF_OPMAT, "gemm buffers");
            need(g.col0 >= 0 && g.col0 + (int64_t)g.ncols <= (int64_t)pl.colC.size(), "gemm C columns");
            for (int n = 0; n < g.ncols; ++n) {
                const int64_t c = pl.colC[g.col0 + n];
                need(c >= 0 && c + g.M <= bsize(g.cbuf), "gemm C range");
            }
            if (g.soff >= 0) need(g.soff + g.M <= pl.opmat_total, "gemm row scale");
            auto a_ok = [&](int64_t aoff, int kcols) {
                const int64_t last = g.transA ? aoff + (int64_t)(g.M - 1) * g.lda + kcols
                                              : aoff + (int64_t)(kcols - 1) * g.lda + g.M;
                return aoff >= 0 && last <= pl.opmat_total && g.lda >= (g.transA ? kcols : g.M);
            };
            if (g.kb == 0) {
                need(a_ok(g.aoff, g.K), "gemm A");
                need(g.bcol0 + (int64_t)g.ncols <= (int64_t)pl.colB.size(), "gemm B columns");
                for (int n = 0; n < g.ncols; ++n) {
                    const int64_t b = pl.colB[g.bcol0 + n];
                    need(b >= 0 && b + g.K <= bsize(g.bbuf), "gemm B range");
                }
            } else {
                const int nkb = g.K / g.kb;
                need(nkb * g.kb == g.K && nkb <= 8, "gemm K blocks");
                need(g.acol0 + (int64_t)nkb <= (int64_t)pl.colA.size(), "gemm A blocks");
                for (int h = 0; h < nkb; ++h) need(a_ok(pl.colA[g.acol0 + h], g.kb), "gemm A block");
                need(g.bcol0 + (int64_t)g.ncols * nkb <= (int64_t)pl.colB.size(), "gemm B blocks");
                for (int64_t e = 0; e < (int64_t)g.ncols * nkb; ++e) {
                    const int64_t b = pl.colB[g.bcol0 + e];
                    need(b >= -1 && b + g.kb <= bsize(g.bbuf), "gemm B block range");
                }
            }
        }
        for (const KTask& t : pl.ktasks) {
            need(t.obuf >= 0 && t.obuf < BUF_COUNT && t.ooff >= 0 && t.ooff + t.count <= pl.bufsize[t.obuf], "kapply out");
            need(t.off >= 0 && t.off + t.count <= npt, "kapply targets");
            need(t.s0 >= 0 && t.s0 <= t.s1 && t.s1 <= (int)pl.ksrcs.size(), "kapply sources");
            for (int q = t.s0; q < t.s1; ++q) {
                const KSrc& k = pl.ksrcs[q];
                need(k.off >= 0 && k.off + k.count <= npt, "kapply source points");
                need(k.vbuf >= 0 && k.vbuf < BUF_COUNT && k.voff >= 0 && k.voff + k.count <= pl.bufsize[k.vbuf],
                     "kapply source values");
                // a mirrored source: point sources read unscaled, and the task's own
                // densities are the mirror's operand
                need(k.moff == -1 || (k.moff >= 0 && k.moff + k.count <= pl.bufsize[BUF_NMIR] &&
                                      t.count <= MIRROR_MAX_TARGETS &&
                                      k.scale == 1.0 && k.c[0] == 0 && k.c[1] == 0 && k.c[2] == 0 &&
                                      t.obuf == BUF_VAL && t.ooff == t.off && t.scale == 1.0 &&
                                      t.c[0] == 0 && t.c[1] == 0 && t.c[2] == 0 &&
                                      t.off + t.count <= pl.bufsize[BUF_DEN]),
                     "kapply mirror slot");
            }
        }
        for (const FftSrc& f : pl.fsrc) {
            need(f.in_off >= 0 && f.in_off + pl.NL <= pl.bufsize[BUF_LF_UPEQ], "fft input");
            need(f.spec_off >= 0 && f.spec_off + G3 <= pl.bufsize[BUF_SPEC], "fft spectrum");
            need(f.grid >= 0 && f.grid + pl.NL <= (int64_t)pl.grid.size(), "fft grid");
        }
        for (const VTarget& v : pl.vtrg) {
            need(v.out_off >= 0 && v.out_off + pl.NL <= pl.bufsize[BUF_LF_CHK], "vlist out");
            need(v.tab_off >= 0 && v.tab_off + 343 * G3 <= pl.opmat_total, "vlist table");
            need(v.grid >= 0 && v.grid + pl.NL <= (int64_t)pl.grid.size(), "vlist grid");
            need(v.p0 >= 0 && v.p0 <= v.p1 && v.p1 <= (int)pl.vpair.size(), "vlist pairs");
            for (int q = v.p0; q < v.p1; ++q) {
                need(pl.vpair[q].spec_off >= 0 && pl.vpair[q].spec_off + G3 <= pl.bufsize[BUF_SPEC], "vlist source");
                need(pl.vpair[q].offidx >= 0 && pl.vpair[q].offidx < 343, "vlist offset");
            }
        }
        for (const Stage& st : pl.stages) {
            if (st.kind != ST_REDUCE) continue;
            need(st.buf == BUF_RED || st.buf == BUF_NMIR, "reduce input buffer");
            for (int q = st.begin; q < st.end; ++q) {
                const RTarget& r = pl.rtrg[q];
                need(r.obuf >= 0 && r.obuf < BUF_COUNT && r.out_off >= 0 && r.out_off + r.M <= pl.bufsize[r.obuf],
                     "reduce out");
                need(r.cnt >= 0 && r.red_off >= 0 && r.red_off + (int64_t)r.cnt * r.M <= pl.bufsize[st.buf],
                     "reduce in");
            }
        }
    }
};


}  // namespace

// The planning phases in order: each one reads the state the earlier ones left in the
// Builder (boxes and lists, operator offsets, slot maps, reduction targets).
void Builder::run(const double* pos, int64_t N, bool want_dump) {
    init(pos, N);
    collect_boxes(pos, N);
    lf_operators();
    hf_groups();
    hf_reductions();
    hf_m2l_pairs();
    assign_slots();
    operator_layout();
    buffer_sizes();
    stages_lf_upward();
    stages_hf_upward();
    stages_hf_downward();
    stages_lf_downward();
    stages_leaf();
    finish(want_dump);
}

int Builder::l2l_target(const HfDir& h, const BoxKey& ck) {
    if (h.W == 1.0) {
        auto it = lf_rt.find(ck);
        if (it != lf_rt.end()) return it->second;
        RedT r;
        r.out_off = lslot(ck); r.obuf = BUF_LF_CHK; r.M = NL; r.mode = 1;
        lf_rt[ck] = (int)rts.size();
        rts.push_back(r);
        return (int)rts.size() - 1;
    }
    auto it = hf_rt.find({ck, parent_dir(h.dir)});
    if (it == hf_rt.end()) throw Error("internal: child lacks inherited direction");
    return it->second;
}

bool Builder::child_present(const BoxKey& ck) const {
    auto bi = tree.boxes.find(ck);
    return bi != tree.boxes.end() && bi->second.has_pts();
}

void Builder::assign(int r0, int r1) {
    int64_t cur = 0;
    for (int i = r0; i < r1; ++i) {
        rts[i].base = cur;
        cur += (int64_t)rts[i].cnt * rts[i].M;
    }
    red_size = std::max(red_size, cur);
}

int64_t Builder::slot(int ri) {
    RedT& r = rts[ri];
    if (r.next >= r.cnt) throw Error("internal: reduction slot overflow");
    return r.base + (int64_t)(r.next++) * r.M;
}

void Builder::reduce_stage(int r0, int r1, const char* name) {
    int q0 = (int)pl.rtrg.size();
    for (int i = r0; i < r1; ++i) {
        const RedT& r = rts[i];
        if (r.next != r.cnt) throw Error("internal: reduction target missing contributions");
        if (r.mode == 1 && r.cnt == 0) continue;
        RTarget t{};
        t.out_off = r.out_off; t.red_off = r.base; t.cnt = r.cnt; t.M = r.M; t.obuf = r.obuf; t.mode = r.mode;
        pl.rtrg.push_back(t);
    }
    stage(ST_REDUCE, q0, (int)pl.rtrg.size(), name, BUF_RED);
}

void Builder::init(const double* pos, int64_t N) {
    tree.prm = prm;
    tree.pos = pos;
    tree.N = N;
    tree.build();
    pl.prm = prm;
    pl.N = N;
    pl.P = prm.ACCU == 1 ? 4 : (prm.ACCU == 2 ? 6 : 8);
    P = pl.P; G = 2 * P; G3 = G * G * G;
}

void Builder::collect_boxes(const double* pos, int64_t N) {
    auto& boxes = tree.boxes;
    // ---------------- live boxes, leaves, LF index (ConstructMaps, eval 252-295)
    for (auto& kv : boxes) {
        const Box& b = kv.second;
        if (!b.has_pts()) continue;
        double W = tree.width(kv.first);
        if (W < 1 - EPS) {
            ldmap[W].push_back(kv.first);
            lfidx[kv.first] = (int64_t)lfidx.size();
            if (b.terminal()) leaves.push_back(kv.first);
        } else {
            for (auto& d : b.outdir) hdmap[d].first.push_back(kv.first);
            for (auto& d : b.incdir) hdmap[d].second.push_back(kv.first);
        }
    }
    for (auto& k : leaves) {
        ptoff[k] = (int64_t)pl.perm.size();
        for (int id : boxes.at(k).ptidx) pl.perm.push_back(id);
    }
    if ((int64_t)pl.perm.size() != N) throw Error("internal: leaves do not cover all points");
    pl.px.resize(N); pl.py.resize(N); pl.pz.resize(N);
    for (int64_t i = 0; i < N; ++i) {
        pl.px[i] = pos[3 * (size_t)pl.perm[i]];
        pl.py[i] = pos[3 * (size_t)pl.perm[i] + 1];
        pl.pz[i] = pos[3 * (size_t)pl.perm[i] + 2];
    }
    pl.stats.npts = N;
    pl.stats.nleaf = (int64_t)leaves.size();
    pl.stats.nlf = (int64_t)lfidx.size();
    pl.stats.nlevels_lf = (int)ldmap.size();
}

void Builder::lf_operators() {
    auto& boxes = tree.boxes;
    // ---------------- LF operators per width
    for (auto& kv : ldmap) {
        double W = kv.first;
        LfW w;
        w.le = &mlib.low(W);
        const LowFreqEntry& le = *w.le;
        if (NL == 0) NL = le.uep.n;
        if (le.uep.n != NL || le.ucp.n != NL) throw Error("LF surfaces must have equal sizes");
        w.r = le.uc2ue[1].m;
        w.uep_p = add_pts(le.uep);
        w.ucp_p = add_pts(le.ucp);
        if (prm.fuse_solves) {
            // check -> equivalent solve V * diag(1/s) * U^H as one FP64 matrix (NL x NL)
            CMat Vs = le.uc2ue[0];
            for (int j = 0; j < Vs.n; ++j)
                for (int i = 0; i < Vs.m; ++i) Vs.a[i + (size_t)j * Vs.m] *= le.uc2ue[1].a[j];
            w.pinv = upload(matmul(Vs, le.uc2ue[2]));
        } else {
            // factored as the reference applies it (EvalUpwardLow, wave3d_eval.cpp:467-478)
            w.V = upload(le.uc2ue[0]);
            w.invs = upload(le.uc2ue[1]);
            w.UH = upload(le.uc2ue[2]);
        }
        w.tab = (int64_t)pl.opmat.size();
        pl.opmat.resize(pl.opmat.size() + (size_t)343 * G3, cpx(0, 0));
        for (auto& t : le.ue2dc) {
            if (t.second.m != G || t.second.n != G || t.second.p != G)
                throw Error("LF translation tensor has wrong size");
            int idx = (t.first[0] + 3) + 7 * (t.first[1] + 3) + 49 * (t.first[2] + 3);
            std::copy(t.second.a.begin(), t.second.a.end(), pl.opmat.begin() + w.tab + (int64_t)idx * G3);
        }
        w.grid = (int)pl.grid.size();
        double step = W / (P - 1);
        for (int k = 0; k < NL; ++k) {
            int g[3];
            for (int d = 0; d < 3; ++d) g[d] = int(std::round((le.uep(d, k) + W / 2) / step)) + P;
            pl.grid.push_back(g[0] + G * g[1] + G * G * g[2]);
        }
        w.has_children = false;
        for (auto& k : kv.second)
            if (!boxes.at(k).terminal()) w.has_children = true;
        for (int c = 0; c < 8; ++c) w.ue2uc[c] = -1;
        if (w.has_children) {
            const LowFreqEntry& lc = mlib.low(W / 2);
            int64_t tg = genpts(rows(le.ucp)), sg = genpts(rows(lc.uep));
            double zero[3] = {0, 0, 0};
            for (int c = 0; c < 8; ++c) {
                double s[3];
                shift_of(c, W, s);
                w.ue2uc[c] = gen(tg, NL, zero, 0, sg, lc.uep.n, s);
            }
        }
        lfw[W] = w;
    }
    pl.NL = NL;
    nlf = (int64_t)lfidx.size();
}

void Builder::hf_groups() {
    // ---------------- HF groups per (W, dir)
    for (auto& kv : hdmap) {
        const Index3& d = kv.first;
        double W = double(d.linfty() / prm.NPQ);
        HfDir h;
        h.W = W;
        h.dir = d;
        sort_dir(d, h.srt, h.sgn, h.prm);
        h.he = &mlib.high(W, h.srt);
        h.r1 = h.he->uep.n;
        h.r2 = h.he->ucp.n;
        h.out = kv.second.first;
        h.in = kv.second.second;
        hfl[W].push_back(h);
    }
    pl.stats.nlevels_hf = (int)hfl.size();
    double maxW = hfl.empty() ? 0 : hfl.rbegin()->first;
    pl.stats.max_hf_width = (int)maxW;
    // Out-slots (directional upward equivalents) of every level stay live from the upward
    // pass to the M2L of their level.  In-slots (downward check values) of width W are
    // written by the level's reduction and last read by its L2L, which precedes the next
    // finer level's reduction on the same stream, so every level's in-slots start at 0:
    // BUF_HF_DN holds one level (the level-by-level schedule of the reference's
    // LIMITED_MEMORY build, wave3d_eval.cpp:106-186, applied to the downward data).
    for (auto& lv : hfl) {
        double W = lv.first;
        int64_t lvl_o = 0, lvl_i = 0, lvl_dn = 0;
        for (auto& h : lv.second) {
            int rm = std::max(h.r1, h.r2);
            for (auto& k : h.out) {
                outslot[{k, h.dir}] = hf_up_size; hf_up_size += h.r1;
                oscr[{k, h.dir}] = lvl_o; lvl_o += rm;
            }
            for (auto& k : h.in) {
                inslot[{k, h.dir}] = lvl_dn; lvl_dn += h.r1;
                iscr[{k, h.dir}] = lvl_i; lvl_i += rm;
            }
            hf_dn_size = std::max(hf_dn_size, lvl_dn);
            pl.stats.nhf_out += (int64_t)h.out.size();
            pl.stats.nhf_in += (int64_t)h.in.size();
            // operators: E1, E2, E3 uploaded once per canonical entry
            auto key = std::make_pair(W, h.srt);
            auto it = hf_entry_off.find(key);
            if (it == hf_entry_off.end()) {
                EOff o{-1, -1, -1};
                if (prm.fuse_solves) {
                    // E1 * E2 * E3 (r1 x r2) as one FP64 matrix, and at W = 1 its transpose
                    const CMat Ef = matmul(matmul(h.he->uc2ue[0], h.he->uc2ue[1]), h.he->uc2ue[2]);
                    o.e1 = o.e2 = upload(Ef);
                    if (W == 1.0) {
                        CMat T;
                        T.m = Ef.n; T.n = Ef.m;
                        T.a.resize(Ef.a.size());
                        for (int j = 0; j < Ef.n; ++j)
                            for (int i = 0; i < Ef.m; ++i) T.a[j + (size_t)i * T.m] = Ef.a[i + (size_t)j * Ef.m];
                        o.eT = upload(T);
                    }
                } else {
                    o.e1 = upload(h.he->uc2ue[0]);
                    o.e2 = upload(h.he->uc2ue[1]);
                    upload(h.he->uc2ue[2]);
                }
                it = hf_entry_off.emplace(key, o).first;
            }
            h.e1 = it->second.e1;
            h.e2 = it->second.e2;
            h.e3 = h.e2 + (int64_t)h.r1 * h.r2;
            h.eT = it->second.eT;
            // shuffled point sets and child kernels (UpwardHighFetch, mlib3d.cpp:176-224)
            std::vector<double> uep = Mlib::shuffled(h.prm, h.sgn, h.he->uep);
            std::vector<double> ucp = Mlib::shuffled(h.prm, h.sgn, h.he->ucp);
            h.uep_g = genpts(uep);
            h.ucp_g = genpts(ucp);
            std::vector<double> uepc;
            if (W == 1.0) {
                uepc = rows(mlib.low(0.5).uep);
            } else {
                Index3 pd = parent_dir(h.dir), s2, g2, p2;
                sort_dir(pd, s2, g2, p2);
                uepc = Mlib::shuffled(p2, g2, mlib.high(W / 2, s2).uep);
            }
            h.r1c = (int)uepc.size() / 3;
            h.uepc_g = genpts(uepc);
            double zero[3] = {0, 0, 0};
            for (int c = 0; c < 8; ++c) {       // contiguous: one r2 x 8 r1c matrix
                double s[3];
                shift_of(c, W, s);
                h.ue2uc[c] = gen(h.ucp_g, h.r2, zero, 0, h.uepc_g, h.r1c, s);
            }
            for (int c = 0; c < 8; ++c) {
                double s[3];
                shift_of(c, W, s);
                h.de2dc[c] = -1;
                if (W == 1.0) {
                    // DownwardHighFetch at W = 1 (mlib3d.cpp:262-265): child check points are
                    // the LF equivalent surface, not negated; de2dc(c) = G(uepc + s_c, -ucp),
                    // evaluated as G((-uepc - s_c) - ucp).
                    double ms[3] = {-s[0], -s[1], -s[2]};
                    h.de2dc[c] = gen(h.uepc_g, h.r1c, ms, 1, h.ucp_g, h.r2, zero);
                }
            }
            if (prm.fuse_solves) {
                h.F = reserve((int64_t)h.r1 * 8 * h.r1c);
                for (int c = 0; c < 8; ++c) h.D[c] = W == 1.0 ? reserve((int64_t)h.r1c * h.r1) : -1;
            }
        }
        scr_size = std::max(scr_size, std::max(lvl_o, lvl_i));
    }
}

void Builder::hf_reductions() {
    auto& boxes = tree.boxes;
    // ---------------- reduction targets of the HF downward pass
    // Every in-slot (box, direction) of a level receives its M2L contributions (one per
    // far-field source box) and its L2L contributions (one per direction of the parent that
    // refines this direction); the LF children of width-1 boxes receive one L2L contribution
    // per incoming direction of the parent.  Each contribution is written by a plain store
    // into its own slot of BUF_RED (the slots of one target adjacent), then one reduction per
    // level sums them into the target: no atomics, fixed summation order.
    for (auto lt = hfl.rbegin(); lt != hfl.rend(); ++lt) {
        int r0 = (int)rts.size();
        for (auto& h : lt->second)
            for (auto& t : h.in) {
                RedT r;
                r.out_off = inslot.at({t, h.dir}); r.obuf = BUF_HF_DN; r.M = h.r1; r.mode = 0;
                hf_rt[{t, h.dir}] = (int)rts.size();
                rts.push_back(r);
            }
        hf_region[lt->first] = {r0, (int)rts.size()};
    }
    // L2L contributions, enumerated in the order the L2L groups are emitted below
    // Width 1 -> LF children: the incoming directions of one class (equal solve size, so
    // equal K blocks) are bundled by up to 8 into one K-blocked GEMM per (bundle, child
    // position) whose columns are the boxes with any of the bundle's directions: one LF
    // check contribution per bundle instead of one per direction.  Columns: the union of
    // the bundle's in-lists, each with its direction-presence mask.
    if (hfl.count(1.0)) {
        std::map<int, std::vector<const HfDir*>> cls;
        for (auto& h : hfl.at(1.0)) cls[prm.fuse_solves ? h.r1 : h.r2].push_back(&h);
        for (auto& kv : cls)
            for (size_t s = 0; s < kv.second.size(); s += 8) {
                Bundle b;
                for (size_t j = s; j < std::min(kv.second.size(), s + 8); ++j) {
                    const int bit = 1 << (int)b.d.size();
                    b.d.push_back(kv.second[j]);
                    for (auto& k : kv.second[j]->in) b.cols[k] |= bit;
                }
                w1b.push_back(std::move(b));
            }
    }
    for (auto lt = hfl.rbegin(); lt != hfl.rend(); ++lt) {
        if (lt->first == 1.0) {
            lf_r0 = (int)rts.size();
            for (auto& b : w1b)
                for (int c = 0; c < 8; ++c)
                    for (auto& km : b.cols) {
                        BoxKey ck = Tree::child_key(km.first, c);
                        if (child_present(ck)) rts[l2l_target(*b.d[0], ck)].cnt++;
                    }
            continue;
        }
        for (auto& h : lt->second)
            for (int c = 0; c < 8; ++c)
                for (auto& k : h.in) {
                    BoxKey ck = Tree::child_key(k, c);
                    if (child_present(ck)) rts[l2l_target(h, ck)].cnt++;
                }
    }
    lf_r1 = (int)rts.size();
}

void Builder::hf_m2l_pairs() {
    auto& boxes = tree.boxes;
    // ---------------- HF M2L pairs grouped by (W, offset) (HighFrequencyM2L, eval 876-928)
    for (auto& lv : hfl) {
        double W = lv.first;
        auto& groups = m2l[W];
        for (auto& h : lv.second) {
            for (auto& t : h.in) {
                const Box& tb = boxes.at(t);
                auto f = tb.fnde.find(h.dir);
                if (f == tb.fnde.end()) continue;
                double ct[3];
                tree.center(t, ct);
                for (auto& s : f->second) {
                    double cs[3];
                    tree.center(s, cs);
                    Index3 off;
                    for (int d = 0; d < 3; ++d) off[d] = int(std::round((ct[d] - cs[d]) / W));
                    auto os = outslot.find({s, h.dir});
                    if (os == outslot.end()) throw Error("internal: M2L source without equivalent density");
                    auto& g = groups[off];
                    if (!g.h) {
                        g.h = &h;
                        double ts[3] = {off[0] * W, off[1] * W, off[2] * W}, zero[3] = {0, 0, 0};
                        g.mat = gen(h.uep_g, h.r1, ts, 1, h.uep_g, h.r1, zero);
                    } else if (!(g.h->dir == h.dir)) {
                        throw Error("internal: one M2L offset maps to two directions");
                    }
                    g.b.push_back(os->second);
                    const int ri = hf_rt.at({t, h.dir});
                    g.t.push_back(ri);
                    rts[ri].cnt++;
                }
            }
        }
    }
}

void Builder::assign_slots() {
    // slots: each region (one HF level, or the LF children of width 1) starts at 0
    for (auto& rg : hf_region) assign(rg.second.first, rg.second.second);
    if (lf_r0 >= 0) assign(lf_r0, lf_r1);
}

void Builder::operator_layout() {
    // ---------------- operator table layout: uploaded part, then generated part
    up_n = (int64_t)pl.opmat.size();
    pl.opmat_total = up_n + gen_cursor;
    for (auto& k : pl.kgen) k.out += up_n;
    pl.stats.opmat_elems = pl.opmat_total;
    pl.stats.opmat_generated = gen_cursor;

    // ---------------- fused mode: solves folded into the translations, by FP64 GEMMs on
    // the operator table at setup (after the generation of the child kernels):
    //   upward    F_c = (E1 E2 E3) ue2uc_c: child equivalents -> parent equivalents, and
    //             (W > 1) transposed for the downward pass: de2dc_c (E1 E2 E3)^T = F_{7-c}^T
    //   W = 1     D_c = de2dc_c (E1 E2 E3)^T: in-slot check values -> LF child check values
    if (prm.fuse_solves) {
        const int g0 = (int)pl.tiles.size();
        for (auto& lv : hfl)
            for (auto& h : lv.second) {
                std::vector<int64_t> bc, cc;
                for (int j = 0; j < 8 * h.r1c; ++j) {
                    bc.push_back(G_(h.ue2uc[0]) + (int64_t)j * h.r2);
                    cc.push_back(G_(h.F) + (int64_t)j * h.r1);
                }
                GGroup g{};
                g.aoff = h.e1; g.soff = -1; g.M = h.r1; g.K = h.r2; g.lda = h.r1; g.transA = 0; g.mode = 0;
                g.bbuf = BUF_OPMAT; g.cbuf = BUF_OPMAT;
                add_group(g, bc, cc, nullptr, 0, {}, false);
                if (lv.first != 1.0) continue;
                for (int c = 0; c < 8; ++c) {
                    std::vector<int64_t> bd, cd;
                    for (int j = 0; j < h.r1; ++j) {
                        bd.push_back(h.eT + (int64_t)j * h.r2);
                        cd.push_back(G_(h.D[c]) + (int64_t)j * h.r1c);
                    }
                    GGroup q{};
                    q.aoff = G_(h.de2dc[c]); q.soff = -1; q.M = h.r1c; q.K = h.r2; q.lda = h.r1c; q.transA = 0;
                    q.mode = 0; q.bbuf = BUF_OPMAT; q.cbuf = BUF_OPMAT;
                    add_group(q, bd, cd, nullptr, 0, {}, false);
                }
            }
        if ((int)pl.tiles.size() > g0) {
            Stage s;
            std::memset(&s, 0, sizeof(s));
            s.kind = ST_GEMM; s.begin = g0; s.end = (int)pl.tiles.size(); s.prec = PREC_SOLVE; s.lstage = -1;
            std::strncpy(s.name, "setup_fuse", sizeof(s.name) - 1);
            push_gemm_stages(pl.setup_stages, s);
        }
    }
}

void Builder::buffer_sizes() {
    // ---------------- buffers (elements per rhs)
    pl.bufsize[BUF_DEN] = pl.N;
    pl.bufsize[BUF_VAL] = pl.N;
    pl.bufsize[BUF_LF_UPEQ] = nlf * NL;
    pl.bufsize[BUF_LF_CHK] = nlf * NL;
    pl.bufsize[BUF_LF_TMP] = prm.fuse_solves ? 0 : nlf * NL;
    pl.bufsize[BUF_LF_DNEQ] = nlf * NL;
    pl.bufsize[BUF_HF_UP] = hf_up_size;
    pl.bufsize[BUF_HF_DN] = hf_dn_size;
    pl.bufsize[BUF_HF_A] = prm.fuse_solves ? 0 : scr_size;
    pl.bufsize[BUF_HF_B] = prm.fuse_solves ? 0 : scr_size;
    pl.bufsize[BUF_RED] = red_size;
}

void Builder::stages_lf_upward() {
    auto& boxes = tree.boxes;
    // ================= stages
    cur_sec = 0;
    stage(ST_GATHER, 0, 0, "gather");
    cur_sec = 2;

    // ---- LF upward (W ascending = finest first), EvalUpwardLow 418-490
    for (auto& kv : ldmap) {
        double W = kv.first;
        LfW& w = lfw[W];
        int t0 = (int)pl.ktasks.size();
        for (auto& k : kv.second) {
            const Box& b = boxes.at(k);
            if (!b.terminal()) continue;
            KTask t{};
            t.off = w.ucp_p; t.count = NL; t.scale = 1.0;
            tree.center(k, t.c);
            t.obuf = BUF_LF_CHK; t.ooff = lslot(k); t.mode = 0;
            t.s0 = (int)pl.ksrcs.size();
            KSrc s{};
            s.off = ptoff.at(k); s.count = npts(k); s.scale = 1.0; s.vbuf = BUF_DEN; s.voff = s.off;
            pl.ksrcs.push_back(s);
            t.s1 = (int)pl.ksrcs.size();
            pl.ktasks.push_back(t);
            pl.stats.s2m_interactions += (int64_t)NL * npts(k);
        }
        stage(ST_KAPPLY, t0, (int)pl.ktasks.size(), "lf_s2m");
        int g0 = (int)pl.tiles.size();
        if (w.has_children) {
            // one GEMM: [ue2uc_0 .. ue2uc_7] (NL x 8 NL) times the children stacked along K;
            // columns ordered by child-occupancy mask (pack_by_mask) so that a wave's union
            // of present children (the K blocks it multiplies) stays small
            std::vector<std::pair<int, BoxKey>> cols;
            for (auto& k : kv.second) {
                if (boxes.at(k).terminal()) continue;
                int mask = 0;
                for (int c = 0; c < 8; ++c)
                    if (lfidx.count(Tree::child_key(k, c))) mask |= 1 << c;
                cols.push_back({mask, k});
            }
            pack_by_mask(cols);
            std::vector<int64_t> bb, cc, ab;
            for (auto& mk : cols) {
                for (int c = 0; c < 8; ++c) {
                    auto it = lfidx.find(Tree::child_key(mk.second, c));
                    bb.push_back(it == lfidx.end() ? -1 : it->second * NL);
                }
                cc.push_back(lslot(mk.second));
            }
            for (int c = 0; c < 8; ++c) ab.push_back(G_(w.ue2uc[c]));
            GGroup g{};
            g.aoff = -1; g.soff = -1; g.M = NL; g.K = 8 * NL; g.lda = NL; g.transA = 0;
            g.bbuf = BUF_LF_UPEQ; g.cbuf = BUF_LF_CHK; g.mode = 0;
            add_group(g, bb, cc, nullptr, NL, ab);
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_m2m");
        std::vector<int64_t> all;
        for (auto& k : kv.second) all.push_back(lslot(k));
        g0 = (int)pl.tiles.size();
        GGroup g{};
        if (prm.fuse_solves) {
            g.aoff = w.pinv; g.soff = -1; g.M = NL; g.K = NL; g.lda = NL; g.transA = 0;
            g.bbuf = BUF_LF_CHK; g.cbuf = BUF_LF_UPEQ; g.mode = 0;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_uc2ue");
        } else {
            g.aoff = w.UH; g.soff = w.invs; g.M = w.r; g.K = NL; g.lda = w.r; g.transA = 0;
            g.bbuf = BUF_LF_CHK; g.cbuf = BUF_LF_TMP; g.mode = 0;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_uc2ue_1");
            g0 = (int)pl.tiles.size();
            g.aoff = w.V; g.soff = -1; g.M = NL; g.K = w.r; g.lda = NL;
            g.bbuf = BUF_LF_TMP; g.cbuf = BUF_LF_UPEQ;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_uc2ue_2");
        }
    }
    stage(ST_ZERO, 0, 0, "zero_lfchk", BUF_LF_CHK, pl.bufsize[BUF_LF_CHK]);
    cur_sec = 4;
}

void Builder::stages_hf_upward() {
    auto& boxes = tree.boxes;
    // ---- HF upward (W = 1, 2, ...), EvalUpwardHigh 773-854
    for (auto& lv : hfl) {
        double W = lv.first;
        int g0 = (int)pl.tiles.size();
        for (auto& h : lv.second) {
            Index3 pd = W == 1.0 ? Index3() : parent_dir(h.dir);
            // one GEMM per direction: [ue2uc_0 .. ue2uc_7] (r2 x 8 r1c), children stacked
            // along K, columns ordered by child-occupancy mask (pack_by_mask)
            std::vector<std::pair<int, BoxKey>> cols;
            for (auto& k : h.out) {
                int mask = 0;
                for (int c = 0; c < 8; ++c) {
                    auto bi = boxes.find(Tree::child_key(k, c));
                    if (bi != boxes.end() && bi->second.has_pts()) mask |= 1 << c;
                }
                cols.push_back({mask, k});
            }
            pack_by_mask(cols);
            std::vector<int64_t> bb, cc, ab;
            for (auto& mk : cols) {
                for (int c = 0; c < 8; ++c) {
                    if (!(mk.first & (1 << c))) { bb.push_back(-1); continue; }
                    BoxKey ck = Tree::child_key(mk.second, c);
                    bb.push_back(W == 1.0 ? lslot(ck) : outslot.at({ck, pd}));
                }
                cc.push_back(oscr.at({mk.second, h.dir}));
            }
            GGroup g{};
            g.aoff = -1; g.soff = -1; g.K = 8 * h.r1c; g.transA = 0; g.mode = 0;
            g.bbuf = W == 1.0 ? BUF_LF_UPEQ : BUF_HF_UP;
            if (prm.fuse_solves) {
                // equivalents directly: F_c (r1 x r1c) into the out-slots
                for (int c = 0; c < 8; ++c) ab.push_back(G_(h.F) + (int64_t)c * h.r1 * h.r1c);
                for (size_t j = 0; j < cols.size(); ++j) cc[j] = outslot.at({cols[j].second, h.dir});
                g.M = h.r1; g.lda = h.r1; g.cbuf = BUF_HF_UP;
            } else {
                for (int c = 0; c < 8; ++c) ab.push_back(G_(h.ue2uc[c]));
                g.M = h.r2; g.lda = h.r2; g.cbuf = BUF_HF_A;
            }
            add_group(g, bb, cc, nullptr, h.r1c, ab);
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_u2u");
        if (!prm.fuse_solves) {
            // E3, then E2, then E1 (EvalUpwardHigh, wave3d_eval.cpp:833-852)
            const char* nm[3] = {"hf_up_e3", "hf_up_e2", "hf_up_e1"};
            for (int step = 0; step < 3; ++step) {
                g0 = (int)pl.tiles.size();
                for (auto& h : lv.second) {
                    std::vector<int64_t> bc, cc;
                    for (auto& k : h.out) {
                        bc.push_back(oscr.at({k, h.dir}));
                        cc.push_back(step == 2 ? outslot.at({k, h.dir}) : oscr.at({k, h.dir}));
                    }
                    GGroup g{};
                    g.soff = -1; g.transA = 0; g.mode = 0;
                    if (step == 0) { g.aoff = h.e3; g.M = h.r2; g.K = h.r2; g.lda = h.r2; g.bbuf = BUF_HF_A; g.cbuf = BUF_HF_B; }
                    if (step == 1) { g.aoff = h.e2; g.M = h.r1; g.K = h.r2; g.lda = h.r1; g.bbuf = BUF_HF_B; g.cbuf = BUF_HF_A; }
                    if (step == 2) { g.aoff = h.e1; g.M = h.r1; g.K = h.r1; g.lda = h.r1; g.bbuf = BUF_HF_A; g.cbuf = BUF_HF_UP; }
                    add_group(g, bc, cc, nullptr);
                }
                stage(ST_GEMM, g0, (int)pl.tiles.size(), nm[step]);
            }
        }
    }
}

void Builder::stages_hf_downward() {
    // ---- HF downward (W = max .. 1), EvalDownwardHigh 996-1040
    cur_sec = 5;
    for (auto lt = hfl.rbegin(); lt != hfl.rend(); ++lt) {
        double W = lt->first;
        int g0 = (int)pl.tiles.size();
        for (auto& og : m2l[W]) {
            M2LGroup& mg = og.second;
            GGroup g{};
            g.aoff = G_(mg.mat); g.soff = -1; g.M = mg.h->r1; g.K = mg.h->r1; g.lda = mg.h->r1; g.transA = 0;
            g.bbuf = BUF_HF_UP; g.cbuf = BUF_RED; g.mode = 0;
            std::vector<int64_t> cc;
            for (int ri : mg.t) cc.push_back(slot(ri));
            add_group(g, mg.b, cc, &pl.stats.m2l_flops);
            pl.stats.m2l_pairs += (int64_t)mg.b.size();
            pl.stats.m2l_groups++;
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_m2l");
        reduce_stage(hf_region.at(W).first, hf_region.at(W).second, "hf_reduce");
        if (!prm.fuse_solves) {
            // E1^T, then E2^T, then E3^T (HighFrequencyL2L, wave3d_eval.cpp:941-949)
            const char* nm[3] = {"hf_dn_e1t", "hf_dn_e2t", "hf_dn_e3t"};
            for (int step = 0; step < 3; ++step) {
                g0 = (int)pl.tiles.size();
                for (auto& h : lt->second) {
                    std::vector<int64_t> bc, cc;
                    for (auto& k : h.in) {
                        bc.push_back(step == 0 ? inslot.at({k, h.dir}) : iscr.at({k, h.dir}));
                        cc.push_back(iscr.at({k, h.dir}));
                    }
                    GGroup g{};
                    g.soff = -1; g.transA = 1; g.mode = 0;
                    if (step == 0) { g.aoff = h.e1; g.M = h.r1; g.K = h.r1; g.lda = h.r1; g.bbuf = BUF_HF_DN; g.cbuf = BUF_HF_A; }
                    if (step == 1) { g.aoff = h.e2; g.M = h.r2; g.K = h.r1; g.lda = h.r1; g.bbuf = BUF_HF_A; g.cbuf = BUF_HF_B; }
                    if (step == 2) { g.aoff = h.e3; g.M = h.r2; g.K = h.r2; g.lda = h.r2; g.bbuf = BUF_HF_B; g.cbuf = BUF_HF_A; }
                    add_group(g, bc, cc, nullptr);
                }
                stage(ST_GEMM, g0, (int)pl.tiles.size(), nm[step]);
            }
        }
        g0 = (int)pl.tiles.size();
        if (W == 1.0) {
            // bundled: C(LF child c) = sum_b A_b(c) x(box, dir_b), A_b(c) = D_c (fused) or
            // de2dc_c (factored) of direction b, K blocks of the bundle's solve size
            for (auto& b : w1b) {
                const int kb = prm.fuse_solves ? b.d[0]->r1 : b.d[0]->r2;
                for (int c = 0; c < 8; ++c) {
                    std::vector<std::pair<int, BoxKey>> cols;
                    for (auto& km : b.cols)
                        if (child_present(Tree::child_key(km.first, c))) cols.push_back({km.second, km.first});
                    if (cols.empty()) continue;
                    pack_by_mask(cols);
                    std::vector<int64_t> bb, cc, ab;
                    for (auto& mk : cols) {
                        for (size_t j = 0; j < b.d.size(); ++j) {
                            if (!(mk.first & (1 << j))) { bb.push_back(-1); continue; }
                            bb.push_back(prm.fuse_solves ? inslot.at({mk.second, b.d[j]->dir})
                                                         : iscr.at({mk.second, b.d[j]->dir}));
                        }
                        cc.push_back(slot(l2l_target(*b.d[0], Tree::child_key(mk.second, c))));
                    }
                    for (auto* h : b.d) ab.push_back(prm.fuse_solves ? G_(h->D[c]) : G_(h->de2dc[c]));
                    GGroup g{};
                    g.aoff = -1; g.soff = -1; g.M = b.d[0]->r1c; g.K = (int)b.d.size() * kb; g.lda = b.d[0]->r1c;
                    g.transA = 0; g.bbuf = prm.fuse_solves ? BUF_HF_DN : BUF_HF_A; g.cbuf = BUF_RED; g.mode = 0;
                    add_group(g, bb, cc, nullptr, kb, ab);
                }
            }
        }
        for (auto& h : lt->second) {
            if (W == 1.0) break;
            for (int c = 0; c < 8; ++c) {
                std::vector<int64_t> bc, cc;
                for (auto& k : h.in) {
                    BoxKey ck = Tree::child_key(k, c);
                    if (!child_present(ck)) continue;
                    bc.push_back(prm.fuse_solves ? inslot.at({k, h.dir}) : iscr.at({k, h.dir}));
                    cc.push_back(slot(l2l_target(h, ck)));
                }
                // W > 1: de2dc(c) = G(-uepc + s_c, -ucp) = ue2uc(7 - c)^T (mlib3d.cpp:227-280);
                // W = 1: explicit matrix (r1c x r2).  Fused mode: composed with the downward
                // solve (E1 E2 E3)^T, from the in-slot check values: F_{7-c}^T, or D_c at W = 1
                GGroup g{};
                if (prm.fuse_solves) {
                    if (W == 1.0) {
                        g.aoff = G_(h.D[c]); g.M = h.r1c; g.K = h.r1; g.lda = h.r1c; g.transA = 0;
                    } else {
                        g.aoff = G_(h.F) + (int64_t)(7 - c) * h.r1 * h.r1c;
                        g.M = h.r1c; g.K = h.r1; g.lda = h.r1; g.transA = 1;
                    }
                    g.bbuf = BUF_HF_DN;
                } else {
                    if (W == 1.0) {
                        g.aoff = G_(h.de2dc[c]); g.M = h.r1c; g.K = h.r2; g.lda = h.r1c; g.transA = 0;
                    } else {
                        g.aoff = G_(h.ue2uc[7 - c]); g.M = h.r1c; g.K = h.r2; g.lda = h.r2; g.transA = 1;
                    }
                    g.bbuf = BUF_HF_A;
                }
                g.soff = -1;
                g.cbuf = BUF_RED; g.mode = 0;
                add_group(g, bc, cc, nullptr);
            }
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_l2l");
        if (W == 1.0 && lf_r0 >= 0) {
            cur_sec = 6;
            reduce_stage(lf_r0, lf_r1, "lf_reduce");
            cur_sec = 5;
        }
    }
}

void Builder::stages_lf_downward() {
    auto& boxes = tree.boxes;
    // ---- LF downward (W descending), EvalDownwardLow 493-583
    for (auto lt = ldmap.rbegin(); lt != ldmap.rend(); ++lt) {
        double W = lt->first;
        LfW& w = lfw[W];
        // V list: spectra of this level's V sources
        std::map<BoxKey, int64_t> specidx;
        int f0 = (int)pl.fsrc.size();
        for (auto& k : lt->second) {
            if (boxes.at(k).fftnum == 0) continue;
            FftSrc f{};
            f.in_off = lslot(k);
            f.spec_off = (int64_t)specidx.size() * G3;
            f.grid = w.grid;
            specidx[k] = (int64_t)specidx.size();
            pl.fsrc.push_back(f);
        }
        spec_max = std::max(spec_max, (int64_t)specidx.size());
        cur_sec = 3;
        stage(ST_FFTFWD, f0, (int)pl.fsrc.size(), "lf_fft");
        int v0 = (int)pl.vtrg.size();
        for (auto& k : morton_sorted(lt->second)) {
            const Box& b = boxes.at(k);
            if (b.V.empty()) continue;
            double ct[3];
            tree.center(k, ct);
            VTarget v{};
            v.out_off = lslot(k);
            v.tab_off = w.tab;
            v.grid = w.grid;
            v.p0 = (int)pl.vpair.size();
            for (auto& s : b.V) {
                double cs[3];
                tree.center(s, cs);
                int o[3];
                for (int d = 0; d < 3; ++d) o[d] = int(std::round((ct[d] - cs[d]) / W));
                VPair p{};
                p.spec_off = specidx.at(s) * G3;
                p.offidx = (o[0] + 3) + 7 * (o[1] + 3) + 49 * (o[2] + 3);
                pl.vpair.push_back(p);
            }
            v.p1 = (int)pl.vpair.size();
            pl.vtrg.push_back(v);
            pl.stats.v_pairs += (int64_t)b.V.size();
        }
        pl.stats.v_targets += (int64_t)pl.vtrg.size() - v0;
        vtrg_max = std::max(vtrg_max, (int64_t)pl.vtrg.size() - v0);
        pl.stats.v_sources += (int64_t)specidx.size();
        // chunks of Morton-consecutive targets, closed when the union of their sources
        // would exceed VCH_SRC rows or the chunk holds VCH_TRG targets; row 0 of every
        // union is a zero row (the kernel stages zeros there) that pads partial pair batches
        const int c0 = (int)pl.vchunk.size();
        std::vector<int> lpk(pl.vpair.size());      // packed pair q (chunk-local row, offset)
        {
            std::unordered_map<int64_t, int> loc;
            VChunk ch{};
            auto open = [&](int t0) {
                ch.t0 = t0;
                ch.u0 = (int)pl.vsrc.size();
                pl.vsrc.push_back(0);
                loc.clear();
            };
            auto close = [&](int t1) {
                ch.t1 = t1;
                ch.u1 = (int)pl.vsrc.size();
                pl.vchunk.push_back(ch);
            };
            open(v0);
            int npk = 0;                                   // padded pair words of the open chunk
            for (int i = v0; i < (int)pl.vtrg.size(); ++i) {
                const VTarget& v = pl.vtrg[i];
                int add = 0;
                for (int q = v.p0; q < v.p1; ++q) add += !loc.count(pl.vpair[q].spec_off / G3);
                const int n8 = (v.p1 - v.p0 + 7) & ~7;
                if (n8 > VCH_PK) throw Error("internal: V list longer than a chunk's pair buffer");
                if (i > ch.t0 && (1 + (int)loc.size() + add > VCH_SRC || i - ch.t0 >= VCH_TRG || npk + n8 > VCH_PK)) {
                    close(i);
                    open(i);
                    npk = 0;
                }
                npk += n8;
                for (int q = v.p0; q < v.p1; ++q) {
                    const int64_t s = pl.vpair[q].spec_off / G3;
                    auto it = loc.find(s);
                    int l;
                    if (it == loc.end()) {
                        l = 1 + (int)loc.size();
                        loc.emplace(s, l);
                        pl.vsrc.push_back((int)s);
                    } else {
                        l = it->second;
                    }
                    lpk[q] = l | (pl.vpair[q].offidx << 16);
                }
            }
            if ((int)pl.vtrg.size() > ch.t0) close((int)pl.vtrg.size());
            else pl.vsrc.pop_back();
        }
        // within a chunk, targets by descending pair count: the targets that share a wave
        // run loops of about the same length (a target's slot follows its new position)
        for (int c = c0; c < (int)pl.vchunk.size(); ++c)
            std::stable_sort(pl.vtrg.begin() + pl.vchunk[c].t0, pl.vtrg.begin() + pl.vchunk[c].t1,
                             [](const VTarget& a, const VTarget& b) { return a.p1 - a.p0 > b.p1 - b.p0; });
        for (int i = v0; i < (int)pl.vtrg.size(); ++i) {
            VTarget& v = pl.vtrg[i];
            v.k0 = (int)pl.vpk.size();
            for (int q = v.p0; q < v.p1; ++q) pl.vpk.push_back(lpk[q]);
            while (pl.vpk.size() % 8) pl.vpk.push_back(0);
        }
        for (int c = c0; c < (int)pl.vchunk.size(); ++c) {
            VChunk& ch = pl.vchunk[c];
            const VTarget& last = pl.vtrg[ch.t1 - 1];
            ch.k0 = pl.vtrg[ch.t0].k0;
            ch.k1 = last.k0 + ((last.p1 - last.p0 + 7) & ~7);
        }
        stage(ST_VLIST, v0, (int)pl.vtrg.size(), "lf_vlist", c0, (int64_t)pl.vchunk.size() - c0);
        // X list into the check potential
        int t0 = (int)pl.ktasks.size();
        for (auto& k : lt->second) {
            const Box& b = boxes.at(k);
            if (b.X.empty()) continue;
            if (b.terminal() && npts(k) < NL) continue;     // handled directly in the leaf kernel
            KTask t{};
            t.off = w.uep_p; t.count = NL; t.scale = 1.0;
            tree.center(k, t.c);
            t.obuf = BUF_LF_CHK; t.ooff = lslot(k); t.mode = 1;
            t.s0 = (int)pl.ksrcs.size();
            for (auto& x : b.X) {
                KSrc s{};
                s.off = ptoff.at(x); s.count = npts(x); s.scale = 1.0; s.vbuf = BUF_DEN; s.voff = s.off;
                pl.ksrcs.push_back(s);
                pl.stats.x2chk_interactions += (int64_t)NL * npts(x);
            }
            t.s1 = (int)pl.ksrcs.size();
            pl.ktasks.push_back(t);
        }
        stage(ST_KAPPLY, t0, (int)pl.ktasks.size(), "lf_x2chk");
        cur_sec = 6;
        std::vector<int64_t> all;
        for (auto& k : lt->second) all.push_back(lslot(k));
        int g0 = (int)pl.tiles.size();
        GGroup g{};
        if (prm.fuse_solves) {
            g.aoff = w.pinv; g.soff = -1; g.M = NL; g.K = NL; g.lda = NL; g.transA = 1;
            g.bbuf = BUF_LF_CHK; g.cbuf = BUF_LF_DNEQ; g.mode = 0;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_dc2de");
        } else {
            // transpose of the factored solve (DownwardLowFetch, mlib3d.cpp:146-151)
            g.aoff = w.V; g.soff = w.invs; g.M = w.r; g.K = NL; g.lda = NL; g.transA = 1;
            g.bbuf = BUF_LF_CHK; g.cbuf = BUF_LF_TMP; g.mode = 0;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_dc2de_1");
            g0 = (int)pl.tiles.size();
            g.aoff = w.UH; g.soff = -1; g.M = NL; g.K = w.r; g.lda = w.r; g.transA = 1;
            g.bbuf = BUF_LF_TMP; g.cbuf = BUF_LF_DNEQ;
            add_group(g, all, all, nullptr);
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_dc2de_2");
        }
        if (w.has_children) {
            g0 = (int)pl.tiles.size();
            for (int c = 0; c < 8; ++c) {
                std::vector<int64_t> bc, cc;
                for (auto& k : lt->second) {
                    if (boxes.at(k).terminal()) continue;
                    BoxKey ck = Tree::child_key(k, c);
                    auto it = lfidx.find(ck);
                    if (it == lfidx.end()) continue;
                    bc.push_back(lslot(k));
                    cc.push_back(it->second * NL);
                }
                // LF de2dc(c) = G(uep_{W/2} + s_c, ucp_W) = ue2uc(c)^T
                GGroup gg{};
                gg.aoff = G_(w.ue2uc[c]); gg.soff = -1; gg.M = NL; gg.K = NL; gg.lda = NL; gg.transA = 1;
                gg.bbuf = BUF_LF_DNEQ; gg.cbuf = BUF_LF_CHK; gg.mode = 1;
                add_group(gg, bc, cc, nullptr);
            }
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "lf_l2l");
        }
    }
    pl.bufsize[BUF_SPEC] = spec_max * G3;
    pl.bufsize[BUF_TSPEC] = vtrg_max * G3;
}

void Builder::stages_leaf() {
    auto& boxes = tree.boxes;
    // ---- leaf kernel (eval 550-560, 585-718), in two launches: the near field (U list,
    // W and X lists evaluated directly from point densities) needs only the densities and
    // runs on the second stream from the start; the far part (W list through source
    // equivalent densities, L2T) is added at the end
    // Near-field pairs evaluated once for both directions (prm.near_pairs): leaves that are
    // each in the other's near list exactly once and the smaller of which (the lower task
    // index on a tie) has at most MIRROR_MAX_TARGETS points.  That leaf evaluates the pair in
    // one target pass; larger tasks keep their source tiles shared by all waves (a mirrored
    // pass of a large leaf would re-read every source per 64 targets).  Evaluator i writes
    // the mirror product into slot (j, rank of i among j's evaluators); task j skips i; a
    // reduction adds j's slots to its potentials in that order.
    const std::vector<BoxKey> order = morton_sorted(leaves);
    std::map<BoxKey, int> leaf_index;
    for (int i = 0; i < (int)order.size(); ++i) leaf_index[order[i]] = i;
    auto near_list = [&](const BoxKey& k) {
        const Box& b = boxes.at(k);
        std::vector<BoxKey> out(b.U.begin(), b.U.end());
        for (auto& s : b.W)
            if (boxes.at(s).terminal() && npts(s) < NL) out.push_back(s);
        if (npts(k) < NL) out.insert(out.end(), b.X.begin(), b.X.end());
        return out;
    };
    std::map<std::pair<int, int>, int> pair_count;
    std::vector<std::vector<int>> mirror_in(order.size());   // j -> its evaluators, ascending
    std::vector<int64_t> mirror_base(order.size(), 0);
    auto evaluates = [&](int i, int j) {      // i evaluates the mutual pair (i, j)
        const int64_t ni = npts(order[i]), nj = npts(order[j]);
        return ni <= MIRROR_MAX_TARGETS && (ni < nj || (ni == nj && i < j));
    };
    if (prm.near_pairs) {
        for (int i = 0; i < (int)order.size(); ++i)
            for (auto& s : near_list(order[i])) {
                auto it = leaf_index.find(s);
                if (it != leaf_index.end()) ++pair_count[{i, it->second}];
            }
        for (auto& pc : pair_count) {      // in (i, j) order: each mirror_in[j] ascends
            const int i = pc.first.first, j = pc.first.second;
            if (i != j && pc.second == 1 && evaluates(i, j)) {
                auto r = pair_count.find({j, i});
                if (r != pair_count.end() && r->second == 1) mirror_in[j].push_back(i);
            }
        }
        int64_t cur = 0;
        for (int j = 0; j < (int)order.size(); ++j) {
            mirror_base[j] = cur;
            cur += (int64_t)mirror_in[j].size() * npts(order[j]);
            pl.stats.near_mirrored += (int64_t)mirror_in[j].size();
        }
        pl.bufsize[BUF_NMIR] = cur;
    }
    auto mirrored = [&](int i, int j) {      // (i, j) is a mutual pair evaluated once
        if (!prm.near_pairs || i == j) return false;
        const int e = evaluates(i, j) ? i : evaluates(j, i) ? j : -1, o = e == i ? j : i;
        if (e < 0) return false;
        const auto& in = mirror_in[o];
        return std::binary_search(in.begin(), in.end(), e);
    };
    auto leaf_tasks = [&](bool near) {
        int t0 = (int)pl.ktasks.size();
        for (auto& k : order) {
            const Box& b = boxes.at(k);
            double W = tree.width(k);
            LfW& w = lfw[W];
            KTask t{};
            t.off = ptoff.at(k); t.count = npts(k); t.scale = 1.0;
            t.obuf = BUF_VAL; t.ooff = t.off; t.mode = near ? 0 : 1;
            t.s0 = (int)pl.ksrcs.size();
            auto direct = [&](const BoxKey& s) {
                KSrc q{};
                q.off = ptoff.at(s); q.count = npts(s); q.scale = 1.0; q.vbuf = BUF_DEN; q.voff = q.off;
                pl.ksrcs.push_back(q);
                pl.stats.near_interactions += (int64_t)npts(k) * npts(s);
            };
            if (near) {
                const int i = leaf_index.at(k);
                for (auto& s : near_list(k)) {
                    auto it = leaf_index.find(s);
                    const int j = it == leaf_index.end() ? -1 : it->second;
                    if (j >= 0 && mirrored(i, j)) {
                        if (!evaluates(i, j)) {               // task j evaluates the pair
                            pl.stats.near_interactions += (int64_t)npts(k) * npts(s);
                            continue;
                        }
                        const auto& in = mirror_in[j];
                        const int64_t rank = std::lower_bound(in.begin(), in.end(), i) - in.begin();
                        direct(s);
                        pl.ksrcs.back().moff = mirror_base[j] + rank * npts(s);
                    } else {
                        direct(s);
                    }
                }
            } else {
                for (auto& s : b.W) {
                    const Box& sb = boxes.at(s);
                    if (sb.terminal() && npts(s) < NL) continue;
                    KSrc q{};
                    q.off = w.uep_p; q.count = NL; q.scale = tree.width(s) / W;
                    tree.center(s, q.c);
                    q.vbuf = BUF_LF_UPEQ; q.voff = lslot(s);
                    pl.ksrcs.push_back(q);
                    pl.stats.leaf_op_interactions += (int64_t)npts(k) * NL;
                }
                KSrc q{};
                q.off = w.ucp_p; q.count = NL; q.scale = 1.0;
                tree.center(k, q.c);
                q.vbuf = BUF_LF_DNEQ; q.voff = lslot(k);
                pl.ksrcs.push_back(q);
                pl.stats.leaf_op_interactions += (int64_t)npts(k) * NL;
            }
            t.s1 = (int)pl.ksrcs.size();
            pl.ktasks.push_back(t);
        }
        if (near && prm.near_pairs) {
            // two launches: the tasks that evaluate mirrored pairs (kernel instance with the
            // mirror path and its registers) and the others (the leaner instance)
            std::stable_partition(pl.ktasks.begin() + t0, pl.ktasks.end(), [&](const KTask& t) {
                for (int q = t.s0; q < t.s1; ++q)
                    if (pl.ksrcs[q].moff >= 0) return false;
                return true;
            });
            int t1 = t0;
            while (t1 < (int)pl.ktasks.size()) {
                const KTask& t = pl.ktasks[t1];
                bool mir = false;
                for (int q = t.s0; q < t.s1 && !mir; ++q) mir = pl.ksrcs[q].moff >= 0;
                if (mir) break;
                ++t1;
            }
            stage(ST_KAPPLY, t0, t1, "leaf_near");
            stage(ST_KAPPLY, t1, (int)pl.ktasks.size(), "leaf_near_pairs");
        } else {
            stage(ST_KAPPLY, t0, (int)pl.ktasks.size(), near ? "leaf_near" : "leaf_far");
        }
    };
    cur_sec = 3;
    leaf_tasks(true);
    if (prm.near_pairs) {      // the mirror products into their leaves, sources in task order
        const int q0 = (int)pl.rtrg.size();
        for (int j = 0; j < (int)order.size(); ++j) {
            if (mirror_in[j].empty()) continue;
            RTarget t{};
            t.out_off = ptoff.at(order[j]); t.red_off = mirror_base[j]; t.cnt = (int)mirror_in[j].size();
            t.M = (int)npts(order[j]); t.obuf = BUF_VAL; t.mode = 1;
            pl.rtrg.push_back(t);
        }
        stage(ST_REDUCE, q0, (int)pl.rtrg.size(), "near_mirror", BUF_NMIR);
    }
    cur_sec = 6;
    leaf_tasks(false);
    stage(ST_SCATTER, 0, 0, "scatter");
    merge_sections();
}

void Builder::finish(bool want_dump) {
    auto& boxes = tree.boxes;
    // ---------------- per-stage algorithmic work (per RHS)
    for (auto& st : pl.stages) stage_work(pl, st);

    validate();

    // ---------------- canonical dump of the lists (integer parity with the oracle)
    if (want_dump) {
        std::ostringstream os;
        auto key = [&](const BoxKey& k) {
            os << k.first << ':' << k.second[0] << ',' << k.second[1] << ',' << k.second[2];
        };
        auto dir = [&](const Index3& d) { os << d[0] << ',' << d[1] << ',' << d[2]; };
        for (auto& kv : boxes) {
            const Box& b = kv.second;
            if (!b.has_pts()) continue;
            key(kv.first);
            os << " t" << b.tag << " n" << b.ptidx.size() << " f" << b.fftnum;
            const std::vector<BoxKey>* L[5] = {&b.U, &b.V, &b.W, &b.X, &b.E};
            const char* nm = "UVWXE";
            for (int i = 0; i < 5; ++i) {
                os << ' ' << nm[i] << '[';
                for (auto& k : *L[i]) { key(k); os << ' '; }
                os << ']';
            }
            for (auto& f : b.fnde) {
                os << " F("; dir(f.first); os << ")[";
                for (auto& k : f.second) { key(k); os << ' '; }
                os << ']';
            }
            os << " I[";
            for (auto& d : b.incdir) { dir(d); os << ' '; }
            os << "] O[";
            for (auto& d : b.outdir) { dir(d); os << ' '; }
            os << "]\n";
        }
        pl.dump = os.str();
    }
}

void build_plan(const Params& prm, const double* pos, int64_t N, const Mlib& mlib,
                HostPlan& plan, bool want_dump) {
    Builder b(prm, mlib, plan);
    b.run(pos, N, want_dump);
}

}  // namespace ddfmm
