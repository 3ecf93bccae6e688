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
    int add_group(GGroup g, const std::vector<int64_t>& bc, const std::vector<int64_t>& cc,
                  double* flops, int kb = 0, const std::vector<int64_t>& ablk = {}) {
        if (cc.empty()) return -1;
        if (kb > 0 && (kb < 32 || (g.K + kb - 1) / kb > 8))
            throw Error("internal: K blocks must hold at least 32 rows, at most 8 blocks");
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
        int mt = (g.M + GEMM_TM - 1) / GEMM_TM, nt = (g.ncols + GEMM_TN - 1) / GEMM_TN;
        for (int a = 0; a < mt; ++a)
            for (int b = 0; b < nt; ++b) pl.tiles.push_back(GTile{gi, a, b, 0});
        double f = 8.0 * g.M * g.K * g.ncols;
        if (kb > 0) {                        // useful work: present blocks only
            int64_t present = 0;
            for (int64_t v : bc) present += v >= 0;
            f = 8.0 * g.M * kb * (double)present;
        }
        pl.stats.gemm_flops += f;
        pl.group_flops[gi] = f;
        if (flops) *flops += f;
        return gi;
    }
    void stage(int kind, int b, int e, const char* name, int buf = -1, int64_t count = 0) {
        if (kind != ST_ZERO && kind != ST_GATHER && kind != ST_SCATTER && e <= b) return;
        Stage s;
        std::memset(&s, 0, sizeof(s));
        s.kind = kind; s.begin = b; s.end = e; s.buf = buf; s.count = count;
        std::strncpy(s.name, name, sizeof(s.name) - 1);
        // solves: LF pinv factors and HF E1/E2/E3 chains
        const std::string nm(name);
        s.prec = (nm.find("uc2ue") != std::string::npos || nm.find("dc2de") != std::string::npos ||
                  nm.find("hf_up_e") != std::string::npos || nm.find("hf_dn_e") != std::string::npos)
                     ? PREC_SOLVE : PREC_TRANSLATION;
        pl.stages.push_back(s);
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
    std::vector<BoxKey> out, in;
};

}  // namespace

void Builder::run(const double* pos, int64_t N, bool want_dump) {
    tree.prm = prm;
    tree.pos = pos;
    tree.N = N;
    tree.build();
    pl.prm = prm;
    pl.N = N;
    pl.P = prm.ACCU == 1 ? 4 : (prm.ACCU == 2 ? 6 : 8);
    const int P = pl.P, G = 2 * P, G3 = G * G * G;
    auto& boxes = tree.boxes;

    // ---------------- live boxes, leaves, LF index (ConstructMaps, eval 252-295)
    std::map<double, std::vector<BoxKey>> ldmap;
    std::map<Index3, std::pair<std::vector<BoxKey>, std::vector<BoxKey>>> hdmap;
    std::map<BoxKey, int64_t> lfidx, ptoff;
    std::vector<BoxKey> leaves;
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
    auto npts = [&](const BoxKey& k) { return (int)boxes.at(k).ptidx.size(); };
    pl.stats.npts = N;
    pl.stats.nleaf = (int64_t)leaves.size();
    pl.stats.nlf = (int64_t)lfidx.size();
    pl.stats.nlevels_lf = (int)ldmap.size();

    // ---------------- LF operators per width
    struct LfW {
        const LowFreqEntry* le;
        int r;
        int64_t uep_p, ucp_p;          // kapply point offsets
        int64_t pinv, V, invs, UH, tab;   // opmat offsets (fused solve or its factors)
        int64_t ue2uc[8];              // generated
        int grid;
        bool has_children;
    };
    std::map<double, LfW> lfw;
    int NL = 0;
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
    const int64_t nlf = (int64_t)lfidx.size();
    auto lslot = [&](const BoxKey& k) { return lfidx.at(k) * NL; };

    // ---------------- HF groups per (W, dir)
    std::map<double, std::vector<HfDir>> hfl;
    std::map<std::pair<BoxKey, Index3>, int64_t> outslot, inslot, oscr, iscr;
    int64_t hf_up_size = 0, hf_dn_size = 0, scr_size = 0;
    std::map<std::pair<double, Index3>, std::pair<int64_t, int64_t>> hf_entry_off;   // (E1..E3) cache
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
    for (auto& lv : hfl) {
        double W = lv.first;
        int64_t lvl_o = 0, lvl_i = 0;
        for (auto& h : lv.second) {
            int rm = std::max(h.r1, h.r2);
            for (auto& k : h.out) {
                outslot[{k, h.dir}] = hf_up_size; hf_up_size += h.r1;
                oscr[{k, h.dir}] = lvl_o; lvl_o += rm;
            }
            for (auto& k : h.in) {
                inslot[{k, h.dir}] = hf_dn_size; hf_dn_size += h.r1;
                iscr[{k, h.dir}] = lvl_i; lvl_i += rm;
            }
            pl.stats.nhf_out += (int64_t)h.out.size();
            pl.stats.nhf_in += (int64_t)h.in.size();
            // operators: E1, E2, E3 uploaded once per canonical entry
            auto key = std::make_pair(W, h.srt);
            auto it = hf_entry_off.find(key);
            if (it == hf_entry_off.end()) {
                int64_t e1, e2;
                if (prm.fuse_solves) {
                    // E1 * E2 * E3 (r1 x r2) as one FP64 matrix
                    e1 = e2 = upload(matmul(matmul(h.he->uc2ue[0], h.he->uc2ue[1]), h.he->uc2ue[2]));
                } else {
                    e1 = upload(h.he->uc2ue[0]);
                    e2 = upload(h.he->uc2ue[1]);
                    upload(h.he->uc2ue[2]);
                }
                it = hf_entry_off.emplace(key, std::make_pair(e1, e2)).first;
            }
            h.e1 = it->second.first;
            h.e2 = it->second.second;
            h.e3 = h.e2 + (int64_t)h.r1 * h.r2;
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
        }
        scr_size = std::max(scr_size, std::max(lvl_o, lvl_i));
    }

    // ---------------- HF M2L pairs grouped by (W, offset) (HighFrequencyM2L, eval 876-928)
    struct M2LGroup {
        const HfDir* h;
        int64_t mat;
        std::vector<int64_t> b, c;
    };
    std::map<double, std::map<Index3, M2LGroup>> m2l;
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
                    g.c.push_back(inslot.at({t, h.dir}));
                }
            }
        }
    }

    // ---------------- operator table layout: uploaded part, then generated part
    const int64_t up_n = (int64_t)pl.opmat.size();
    pl.opmat_total = up_n + gen_cursor;
    for (auto& k : pl.kgen) k.out += up_n;
    auto G_ = [&](int64_t rel) { return rel + up_n; };
    pl.stats.opmat_elems = pl.opmat_total;
    pl.stats.opmat_generated = gen_cursor;

    // ---------------- buffers (elements per rhs)
    pl.bufsize[BUF_DEN] = N;
    pl.bufsize[BUF_VAL] = N;
    pl.bufsize[BUF_LF_UPEQ] = nlf * NL;
    pl.bufsize[BUF_LF_CHK] = nlf * NL;
    pl.bufsize[BUF_LF_TMP] = prm.fuse_solves ? 0 : nlf * NL;
    pl.bufsize[BUF_LF_DNEQ] = nlf * NL;
    pl.bufsize[BUF_HF_UP] = hf_up_size;
    pl.bufsize[BUF_HF_DN] = hf_dn_size;
    pl.bufsize[BUF_HF_A] = scr_size;
    pl.bufsize[BUF_HF_B] = prm.fuse_solves ? 0 : scr_size;

    // ================= stages
    stage(ST_GATHER, 0, 0, "gather");
    stage(ST_ZERO, 0, 0, "zero_hfdn", BUF_HF_DN, pl.bufsize[BUF_HF_DN]);

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
            // columns sorted by child-occupancy mask so that a tile's union of present
            // children (the K blocks it multiplies) stays small
            std::vector<std::pair<int, BoxKey>> cols;
            for (auto& k : kv.second) {
                if (boxes.at(k).terminal()) continue;
                int mask = 0;
                for (int c = 0; c < 8; ++c)
                    if (lfidx.count(Tree::child_key(k, c))) mask |= 1 << c;
                cols.push_back({mask, k});
            }
            std::stable_sort(cols.begin(), cols.end(),
                             [](const std::pair<int, BoxKey>& a, const std::pair<int, BoxKey>& b) { return a.first < b.first; });
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

    // ---- HF upward (W = 1, 2, ...), EvalUpwardHigh 773-854
    for (auto& lv : hfl) {
        double W = lv.first;
        int g0 = (int)pl.tiles.size();
        for (auto& h : lv.second) {
            Index3 pd = W == 1.0 ? Index3() : parent_dir(h.dir);
            // one GEMM per direction: [ue2uc_0 .. ue2uc_7] (r2 x 8 r1c), children stacked
            // along K, columns sorted by child-occupancy mask (see the LF M2M above)
            std::vector<std::pair<int, BoxKey>> cols;
            for (auto& k : h.out) {
                int mask = 0;
                for (int c = 0; c < 8; ++c) {
                    auto bi = boxes.find(Tree::child_key(k, c));
                    if (bi != boxes.end() && bi->second.has_pts()) mask |= 1 << c;
                }
                cols.push_back({mask, k});
            }
            std::stable_sort(cols.begin(), cols.end(),
                             [](const std::pair<int, BoxKey>& a, const std::pair<int, BoxKey>& b) { return a.first < b.first; });
            std::vector<int64_t> bb, cc, ab;
            for (auto& mk : cols) {
                for (int c = 0; c < 8; ++c) {
                    if (!(mk.first & (1 << c))) { bb.push_back(-1); continue; }
                    BoxKey ck = Tree::child_key(mk.second, c);
                    bb.push_back(W == 1.0 ? lslot(ck) : outslot.at({ck, pd}));
                }
                cc.push_back(oscr.at({mk.second, h.dir}));
            }
            for (int c = 0; c < 8; ++c) ab.push_back(G_(h.ue2uc[c]));
            GGroup g{};
            g.aoff = -1; g.soff = -1; g.M = h.r2; g.K = 8 * h.r1c; g.lda = h.r2; g.transA = 0;
            g.bbuf = W == 1.0 ? BUF_LF_UPEQ : BUF_HF_UP; g.cbuf = BUF_HF_A; g.mode = 0;
            add_group(g, bb, cc, nullptr, h.r1c, ab);
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_u2u");
        if (prm.fuse_solves) {
            g0 = (int)pl.tiles.size();
            for (auto& h : lv.second) {
                std::vector<int64_t> bc, cc;
                for (auto& k : h.out) {
                    bc.push_back(oscr.at({k, h.dir}));
                    cc.push_back(outslot.at({k, h.dir}));
                }
                GGroup g{};
                g.aoff = h.e1; g.soff = -1; g.M = h.r1; g.K = h.r2; g.lda = h.r1; g.transA = 0; g.mode = 0;
                g.bbuf = BUF_HF_A; g.cbuf = BUF_HF_UP;
                add_group(g, bc, cc, nullptr);
            }
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_up_e");
        } else {
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

    // ---- HF downward (W = max .. 1), EvalDownwardHigh 996-1040
    for (auto lt = hfl.rbegin(); lt != hfl.rend(); ++lt) {
        double W = lt->first;
        int g0 = (int)pl.tiles.size();
        for (auto& og : m2l[W]) {
            M2LGroup& mg = og.second;
            GGroup g{};
            g.aoff = G_(mg.mat); g.soff = -1; g.M = mg.h->r1; g.K = mg.h->r1; g.lda = mg.h->r1; g.transA = 0;
            g.bbuf = BUF_HF_UP; g.cbuf = BUF_HF_DN; g.mode = 2;
            add_group(g, mg.b, mg.c, &pl.stats.m2l_flops);
            pl.stats.m2l_pairs += (int64_t)mg.b.size();
            pl.stats.m2l_groups++;
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_m2l");
        if (prm.fuse_solves) {
            g0 = (int)pl.tiles.size();
            for (auto& h : lt->second) {
                std::vector<int64_t> bc, cc;
                for (auto& k : h.in) {
                    bc.push_back(inslot.at({k, h.dir}));
                    cc.push_back(iscr.at({k, h.dir}));
                }
                // (E1 E2 E3)^T = E3^T E2^T E1^T (DownwardHighFetch transposes, mlib3d.cpp:247-252)
                GGroup g{};
                g.aoff = h.e1; g.soff = -1; g.M = h.r2; g.K = h.r1; g.lda = h.r1; g.transA = 1; g.mode = 0;
                g.bbuf = BUF_HF_DN; g.cbuf = BUF_HF_A;
                add_group(g, bc, cc, nullptr);
            }
            stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_dn_et");
        } else {
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
        for (auto& h : lt->second) {
            Index3 pd = W == 1.0 ? Index3() : parent_dir(h.dir);
            for (int c = 0; c < 8; ++c) {
                std::vector<int64_t> bc, cc;
                for (auto& k : h.in) {
                    BoxKey ck = Tree::child_key(k, c);
                    auto bi = boxes.find(ck);
                    if (bi == boxes.end() || !bi->second.has_pts()) continue;
                    bc.push_back(iscr.at({k, h.dir}));
                    if (W == 1.0) cc.push_back(lslot(ck));
                    else {
                        auto it = inslot.find({ck, pd});
                        if (it == inslot.end()) throw Error("internal: child lacks inherited direction");
                        cc.push_back(it->second);
                    }
                }
                // W > 1: de2dc(c) = G(-uepc + s_c, -ucp) = ue2uc(7 - c)^T (mlib3d.cpp:227-280);
                // W = 1: explicit matrix (r1c x r2)
                GGroup g{};
                if (W == 1.0) {
                    g.aoff = G_(h.de2dc[c]); g.M = h.r1c; g.K = h.r2; g.lda = h.r1c; g.transA = 0;
                } else {
                    g.aoff = G_(h.ue2uc[7 - c]); g.M = h.r1c; g.K = h.r2; g.lda = h.r2; g.transA = 1;
                }
                g.soff = -1;
                g.bbuf = BUF_HF_A; g.cbuf = W == 1.0 ? BUF_LF_CHK : BUF_HF_DN; g.mode = 2;
                add_group(g, bc, cc, nullptr);
            }
        }
        stage(ST_GEMM, g0, (int)pl.tiles.size(), "hf_l2l");
    }

    // ---- LF downward (W descending), EvalDownwardLow 493-583
    int64_t spec_max = 0;
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
        pl.stats.v_sources += (int64_t)specidx.size();
        stage(ST_VLIST, v0, (int)pl.vtrg.size(), "lf_vlist");
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

    // ---- leaf kernel: U list, W list, X list (direct), L2T (eval 550-560, 585-718)
    int t0 = (int)pl.ktasks.size();
    for (auto& k : morton_sorted(leaves)) {
        const Box& b = boxes.at(k);
        double W = tree.width(k);
        LfW& w = lfw[W];
        KTask t{};
        t.off = ptoff.at(k); t.count = npts(k); t.scale = 1.0;
        t.obuf = BUF_VAL; t.ooff = t.off; t.mode = 0;
        t.s0 = (int)pl.ksrcs.size();
        auto direct = [&](const BoxKey& s) {
            KSrc q{};
            q.off = ptoff.at(s); q.count = npts(s); q.scale = 1.0; q.vbuf = BUF_DEN; q.voff = q.off;
            pl.ksrcs.push_back(q);
            pl.stats.near_interactions += (int64_t)npts(k) * npts(s);
        };
        for (auto& s : b.U) direct(s);
        for (auto& s : b.W) {
            const Box& sb = boxes.at(s);
            if (sb.terminal() && npts(s) < NL) {
                direct(s);
            } else {
                KSrc q{};
                q.off = w.uep_p; q.count = NL; q.scale = tree.width(s) / W;
                tree.center(s, q.c);
                q.vbuf = BUF_LF_UPEQ; q.voff = lslot(s);
                pl.ksrcs.push_back(q);
                pl.stats.leaf_op_interactions += (int64_t)npts(k) * NL;
            }
        }
        if (npts(k) < NL)
            for (auto& s : b.X) direct(s);
        KSrc q{};
        q.off = w.ucp_p; q.count = NL; q.scale = 1.0;
        tree.center(k, q.c);
        q.vbuf = BUF_LF_DNEQ; q.voff = lslot(k);
        pl.ksrcs.push_back(q);
        pl.stats.leaf_op_interactions += (int64_t)npts(k) * NL;
        t.s1 = (int)pl.ksrcs.size();
        pl.ktasks.push_back(t);
    }
    stage(ST_KAPPLY, t0, (int)pl.ktasks.size(), "leaf_near_l2t");
    stage(ST_SCATTER, 0, 0, "scatter");

    // ---------------- per-stage algorithmic work (per RHS)
    for (auto& s : pl.stages) {
        double w = 0;
        if (s.kind == ST_GEMM) {
            int last = -1;
            for (int t = s.begin; t < s.end; ++t) {
                int gi = pl.tiles[t].g;
                if (gi == last) continue;
                last = gi;
                w += pl.group_flops[gi];
            }
        } else if (s.kind == ST_KAPPLY) {
            for (int t = s.begin; t < s.end; ++t) {
                const KTask& k = pl.ktasks[t];
                for (int q = k.s0; q < k.s1; ++q) w += (double)k.count * pl.ksrcs[q].count;
            }
        } else if (s.kind == ST_FFTFWD) {
            w = s.end - s.begin;
        } else if (s.kind == ST_VLIST) {
            for (int t = s.begin; t < s.end; ++t) w += pl.vtrg[t].p1 - pl.vtrg[t].p0;
        }
        s.work = w;
    }

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
