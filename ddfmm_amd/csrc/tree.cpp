// Octree, LF (U/V/W/X) and HF (near/far by direction) lists and direction sets.
// Follows wave3d_setup.cpp:100-477 and wave3d.cpp:43-142 for a single rank that owns
// every cell (geomprtn all zero).
#include "host.hpp"

#include <algorithm>
#include <deque>
#include <queue>

namespace ddfmm {

// ---------------------------------------------------------------- direction algebra
Index3 nml2dir(const double n[3], double W, int NPQ) {
    int C = NPQ * int(std::round(W));
    int B = C / 2;
    int midx = 0;
    double mval = std::abs(n[0]);
    for (int d = 1; d < 3; ++d) {
        if (mval < std::abs(n[d])) { midx = d; mval = std::abs(n[d]); }
    }
    double val[3];
    for (int d = 0; d < 3; ++d) val[d] = std::atan(n[d] / mval);
    double Ang = (M_PI / 2) / C;
    Index3 res;
    for (int d = 0; d < 3; ++d) {
        int t = int(std::floor(val[d] / Ang));
        t = std::min(std::max(t, -B), B - 1);
        res[d] = 2 * t + 1;
    }
    res[midx] = C * int(std::round(val[midx]));
    return res;
}

Index3 parent_dir(const Index3& dir) {
    int C = dir.linfty();
    int B = C / 2;
    int midx = -1;
    for (int d = 0; d < 3; ++d)
        if (std::abs(dir[d]) == C) midx = d;
    if (midx < 0) throw Error("parent_dir: not a direction");
    Index3 res;
    for (int d = 0; d < 3; ++d) {
        int r = (dir[d] + C - 1) / 2;
        res[d] = 2 * (r / 2) + 1 - B;
    }
    res[midx] = dir[midx] / 2;
    return res;
}

std::vector<Index3> child_dirs(const Index3& dir) {
    int C = dir.linfty();
    std::vector<int> oidx;
    for (int d = 0; d < 3; ++d)
        if (std::abs(dir[d]) != C) oidx.push_back(d);
    std::vector<Index3> res;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            Index3 t(2 * dir[0], 2 * dir[1], 2 * dir[2]);
            t[oidx[0]] += 2 * a - 1;
            t[oidx[1]] += 2 * b - 1;
            res.push_back(t);
        }
    return res;
}

void sort_dir(Index3 val, Index3& srt, Index3& sgn, Index3& prm) {
    for (int d = 0; d < 3; ++d) {
        sgn[d] = val[d] < 0 ? -1 : 1;
        val[d] = std::abs(val[d]);
    }
    int upp = val.linfty() + 1;
    for (int d = 0; d < 3; ++d) {
        int minidx = d, minval = val[d];
        for (int e = 0; e < 3; ++e)
            if (val[e] < minval) { minval = val[e]; minidx = e; }
        srt[d] = minval;
        prm[d] = minidx;
        val[minidx] = upp;
    }
}

// ---------------------------------------------------------------- tree
void Tree::center(const BoxKey& k, double c[3]) const {
    int tmp = 1 << k.first;
    for (int d = 0; d < 3; ++d)
        c[d] = prm.ctr[d] - prm.K / 2 + (k.second[d] + 0.5) / tmp * prm.K;
}

bool Tree::adjacent(const BoxKey& me, const BoxKey& yo) {
    int md = std::max(me.first, yo.first);
    int sm = 1 << (md - me.first), sy = 1 << (md - yo.first);
    int rad = sm + sy;
    int mx = 0;
    for (int d = 0; d < 3; ++d) {
        int dif = std::abs((2 * me.second[d] + 1) * sm - (2 * yo.second[d] + 1) * sy);
        if (dif > rad) return false;
        mx = std::max(mx, dif);
    }
    return mx == rad;
}

bool Tree::find(const BoxKey& want, BoxKey& res) const {
    res = want;
    int lc = cell_level();
    while (res.first != lc) {
        if (boxes.count(res)) return true;
        res = parent_key(res);
    }
    return boxes.count(res) != 0;
}

void Tree::build() {
    if (prm.numC < 1 || (prm.numC & (prm.numC - 1)))
        throw Error("numC must be a power of two");
    const double K = prm.K;
    const int numC = prm.numC, lvlC = cell_level();
    if (K / numC < 1 - 1e-12) throw Error("cell boxes must have width >= 1");
    std::vector<std::vector<int>> cells((size_t)numC * numC * numC);
    for (long p = 0; p < N; ++p) {
        int idx[3];
        for (int d = 0; d < 3; ++d) {
            idx[d] = (int)std::floor(numC * ((pos[3 * p + d] - prm.ctr[d] + K / 2) / K));
            if (idx[d] < 0 || idx[d] >= numC)
                throw Error("point " + std::to_string(p) + " lies outside the computational domain");
        }
        cells[((size_t)idx[0] * numC + idx[1]) * numC + idx[2]].push_back((int)p);
    }
    std::deque<std::pair<BoxKey, Box>> q;
    for (int a = 0; a < numC; ++a)
        for (int b = 0; b < numC; ++b)
            for (int c = 0; c < numC; ++c) {
                Box bx;
                bx.ptidx.swap(cells[((size_t)a * numC + b) * numC + c]);
                q.emplace_back(BoxKey(lvlC, Index3(a, b, c)), std::move(bx));
            }
    const int ul = unitlevel();
    while (!q.empty()) {
        BoxKey key = q.front().first;
        Box box = std::move(q.front().second);
        q.pop_front();
        size_t np = box.ptidx.size();
        if (np > 0) box.tag |= TAG_PTS;
        bool action = (key.first <= ul && np > 0) ||
                      ((int)np > prm.ptsmax && key.first < prm.maxlevel - 1);
        if (action) {
            double cc[3];
            center(key, cc);
            Box ch[8];
            for (int id : box.ptidx) {
                const double* p = pos + 3 * (size_t)id;
                int ind = (p[0] >= cc[0]) * 4 + (p[1] >= cc[1]) * 2 + (p[2] >= cc[2]);
                ch[ind].ptidx.push_back(id);
            }
            for (int ind = 0; ind < 8; ++ind) q.emplace_back(child_key(key, ind), std::move(ch[ind]));
            box.ptidx.clear();
            box.ptidx.shrink_to_fit();
        } else {
            box.tag |= TAG_TERMINAL;
        }
        boxes.emplace(key, std::move(box));
    }
    const double eps = 1e-12;
    for (auto& kv : boxes) {
        if (!kv.second.has_pts()) continue;
        if (width(kv.first) < 1 - eps) callowlist(kv.first, kv.second);
        else calhghlist(kv.first, kv.second);
    }
    for (auto& kv : boxes)
        if (kv.second.has_pts())
            for (auto& nk : kv.second.V) boxes.at(nk).fftnum++;
    for (auto& kv : boxes) {
        const BoxKey& key = kv.first;
        Box& box = kv.second;
        double W = width(key);
        if (!(W > 1 - eps && box.has_pts())) continue;
        if (key.first != lvlC) {
            const Box& par = boxes.at(parent_key(key));
            for (auto& d : par.outdir) box.outdir.insert(parent_dir(d));
            for (auto& d : par.incdir) box.incdir.insert(parent_dir(d));
        }
        double cc[3];
        center(key, cc);
        for (auto& f : box.fnde)
            for (auto& ok : f.second) {
                double oc[3], t[3];
                center(ok, oc);
                for (int d = 0; d < 3; ++d) t[d] = oc[d] - cc[d];
                double l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                for (int d = 0; d < 3; ++d) t[d] /= l;
                box.outdir.insert(nml2dir(t, W, prm.NPQ));
                for (int d = 0; d < 3; ++d) t[d] = cc[d] - oc[d];
                l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                for (int d = 0; d < 3; ++d) t[d] /= l;
                box.incdir.insert(nml2dir(t, W, prm.NPQ));
            }
    }
}

void Tree::callowlist(const BoxKey& key, Box& box) {
    if (key.first == cell_level()) throw Error("cell-level boxes must not be low-frequency");
    std::set<BoxKey> U, V, Wl, X;
    BoxKey par = parent_key(key);
    int L = 1 << key.first;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(2 * par.second[d] - 2, 0);
        hi[d] = std::min(2 * par.second[d] + 4, L);
    }
    const bool term = box.terminal();
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                if (i == key.second[0] && j == key.second[1] && k == key.second[2]) continue;
                BoxKey res;
                if (!find(BoxKey(key.first, Index3(i, j, k)), res)) continue;
                const Box& rb = boxes.at(res);
                bool adj = adjacent(res, key);
                if (res.first < key.first && rb.has_pts()) {
                    if (!adj) X.insert(res);
                    else if (term) U.insert(res);
                }
                if (res.first == key.first) {
                    if (!adj) {
                        for (int d = 0; d < 3; ++d)
                            if (std::abs(res.second[d] - key.second[d]) > 3)
                                throw Error("V-list offset exceeds 3");
                        if (rb.has_pts()) V.insert(res);
                    } else if (term) {
                        std::queue<BoxKey> rest;
                        rest.push(res);
                        while (!rest.empty()) {
                            BoxKey fk = rest.front();
                            rest.pop();
                            const Box& fb = boxes.at(fk);
                            bool a2 = adjacent(fk, key);
                            if (!a2 && fb.has_pts()) Wl.insert(fk);
                            if (a2 && fb.terminal() && fb.has_pts()) U.insert(fk);
                            if (a2 && !fb.terminal())
                                for (int ind = 0; ind < 8; ++ind) rest.push(child_key(fk, ind));
                        }
                    }
                }
            }
    if (term && box.has_pts()) U.insert(key);
    box.U.assign(U.begin(), U.end());
    box.V.assign(V.begin(), V.end());
    box.W.assign(Wl.begin(), Wl.end());
    box.X.assign(X.begin(), X.end());
}

void Tree::calhghlist(const BoxKey& key, Box& box) {
    double cc[3];
    center(key, cc);
    double W = width(key);
    double thr = W * W + W - 1e-12;
    auto test = [&](const BoxKey& ok) {
        const Box& ob = boxes.at(ok);
        if (!ob.has_pts()) return;
        double oc[3], diff[3];
        center(ok, oc);
        for (int d = 0; d < 3; ++d) diff[d] = cc[d] - oc[d];
        double l = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
        if (l >= thr) {
            double n[3] = {diff[0] / l, diff[1] / l, diff[2] / l};
            box.fnde[nml2dir(n, W, prm.NPQ)].push_back(ok);
        } else {
            box.E.push_back(ok);
        }
    };
    if (key.first == cell_level()) {
        for (auto it = boxes.begin(); it != boxes.end() && it->first.first == cell_level(); ++it)
            test(it->first);
    } else {
        const Box& par = boxes.at(parent_key(key));
        for (auto& tk : par.E)
            for (int ind = 0; ind < 8; ++ind) test(child_key(tk, ind));
    }
}

}  // namespace ddfmm
