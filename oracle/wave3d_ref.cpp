// ORACLE (test infrastructure only): a compiled FP64 restatement of the reference's
// directional FMM (Wave3d::setup + Wave3d::eval), multi-threaded with OpenMP.
//
// Only tests and the CPU-baseline leg of bench.py load it, as the checker and as the
// timed CPU path; the product never does.  It restates oracle/wave3d_ref.py (the numpy
// oracle) function by function, with the same reference citations:
//   Kernel3d::kernel               kernel3d.cpp:26-107 (r^2 < 1e-16 => r := 1, k = 2*pi)
//   nml2dir / ParentDir / ChildDir wave3d.cpp:43-119, BoxCenter 131-142
//   setup_tree                     wave3d_setup.cpp:100-299
//   setup_tree_callowlist          wave3d_setup.cpp:302-390 (U/V/W/X lists)
//   setup_tree_calhghlist          wave3d_setup.cpp:393-443 (far lists by direction)
//   setup_tree_find / adjacent     wave3d_setup.cpp:446-477
//   Mlib3d fetches                 mlib3d.cpp:106-357 (symmetry sort/shuffle, child kernels)
//   eval                           wave3d_eval.cpp:299-415 with EvalUpwardLow 418-490,
//                                  EvalDownwardLow 493-583, U/V/X/W lists 585-718,
//                                  EvalUpwardHigh 773-854, HighFrequencyM2L 876-928,
//                                  HighFrequencyL2L 931-993, EvalDownwardHigh 996-1040
// Schedule: the reference walks the HF direction tree recursively (wave3d_eval.cpp:721-751)
// and the LF levels box by box; here every level is one parallel loop (upward: finest
// first; downward: coarsest first), and a level's contributions to the next finer level
// are pulled by each receiving (box, direction), so no two threads add into one vector.
// The sums are the same terms in another order (differences at rounding level).  The
// V list uses the FFT-diagonal form of the reference (FFTW there, a separable 12-point
// DFT here).  One right-hand side per evaluation, like the reference.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cpx;
const double TPK = 2.0 * M_PI;
const double MINDIF2 = 1e-16;
enum { PTS = 1, TERMINAL = 2 };

struct Key {
    int l, x, y, z;
    bool operator<(const Key& o) const {
        if (l != o.l) return l < o.l;
        if (x != o.x) return x < o.x;
        if (y != o.y) return y < o.y;
        return z < o.z;
    }
    bool operator==(const Key& o) const { return l == o.l && x == o.x && y == o.y && z == o.z; }
};
struct Dir {
    int v[3];
    bool operator<(const Dir& o) const {
        for (int d = 0; d < 3; ++d)
            if (v[d] != o.v[d]) return v[d] < o.v[d];
        return false;
    }
    bool operator==(const Dir& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};

Key child_key(const Key& k, int ind) {
    return {k.l + 1, 2 * k.x + ((ind & 4) >> 2), 2 * k.y + ((ind & 2) >> 1), 2 * k.z + (ind & 1)};
}
Key parent_key(const Key& k) { return {k.l - 1, k.x >> 1, k.y >> 1, k.z >> 1}; }

double cround(double x) { return x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5); }

// wave3d.cpp:43-71
Dir nml2dir(const double n[3], double W, int NPQ) {
    const int C = NPQ * (int)std::lround(W), B = C / 2;
    int midx = 0;
    double mval = std::fabs(n[0]);
    for (int d = 1; d < 3; ++d)
        if (mval < std::fabs(n[d])) {
            midx = d;
            mval = std::fabs(n[d]);
        }
    double val[3];
    for (int d = 0; d < 3; ++d) val[d] = std::atan(n[d] / mval);
    const double Ang = (M_PI / 2) / C;
    Dir r;
    for (int d = 0; d < 3; ++d) {
        int t = (int)std::floor(val[d] / Ang);
        t = std::min(std::max(t, -B), B - 1);
        r.v[d] = 2 * t + 1;
    }
    r.v[midx] = C * (int)cround(val[midx]);
    return r;
}

// wave3d.cpp:74-95 (C integer division)
Dir parent_dir(const Dir& d) {
    const int C = std::max(std::abs(d.v[0]), std::max(std::abs(d.v[1]), std::abs(d.v[2])));
    const int B = C / 2;
    int midx = -1;
    for (int k = 0; k < 3; ++k)
        if (std::abs(d.v[k]) == C) midx = k;
    Dir r;
    for (int k = 0; k < 3; ++k) {
        int t = (d.v[k] + C - 1) / 2;
        r.v[k] = 2 * (t / 2) + 1 - B;
    }
    r.v[midx] = d.v[midx] / 2;
    return r;
}

// wave3d.cpp:98-119
std::vector<Dir> child_dirs(const Dir& d) {
    const int C = std::max(std::abs(d.v[0]), std::max(std::abs(d.v[1]), std::abs(d.v[2])));
    int oidx[2], no = 0;
    for (int k = 0; k < 3; ++k)
        if (std::abs(d.v[k]) != C) oidx[no++] = k;
    std::vector<Dir> out;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            Dir t;
            for (int k = 0; k < 3; ++k) t.v[k] = 2 * d.v[k];
            t.v[oidx[0]] += 2 * a - 1;
            t.v[oidx[1]] += 2 * b - 1;
            out.push_back(t);
        }
    return out;
}

// mlib3d.cpp:328-357
void sort_dir(const Dir& val, Dir& srt, int sgn[3], int prm[3]) {
    int v[3];
    for (int d = 0; d < 3; ++d) {
        sgn[d] = val.v[d] >= 0 ? 1 : -1;
        v[d] = std::abs(val.v[d]);
    }
    const int upp = std::max(v[0], std::max(v[1], v[2])) + 1;
    for (int d = 0; d < 3; ++d) {
        int minidx = d, minval = v[d];
        for (int e = 0; e < 3; ++e)
            if (v[e] < minval) {
                minval = v[e];
                minidx = e;
            }
        srt.v[d] = minval;
        prm[d] = minidx;
        v[minidx] = upp;
    }
}

// mlib3d.cpp:308-325: rows (n x 3) of the shuffled point set
std::vector<double> shuffle_pts(const int prm[3], const int sgn[3], const std::vector<double>& rows) {
    const size_t n = rows.size() / 3;
    std::vector<double> out(rows.size());
    for (size_t j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) out[3 * j + prm[d]] = rows[3 * j + d];
    for (size_t j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) out[3 * j + d] *= sgn[d];
    return out;
}

void shifted_point(int ind, double W, double s[3]) {
    const int a = (ind & 4) >> 2, b = (ind & 2) >> 1, c = ind & 1;
    s[0] = (a - 0.5) * W / 2;
    s[1] = (b - 0.5) * W / 2;
    s[2] = (c - 0.5) * W / 2;
}

inline cpx green(double dx, double dy, double dz, int ktype) {
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < MINDIF2) r2 = 1.0;
    const double r = std::sqrt(r2), kr = r * TPK;
    const double c = std::cos(kr), s = std::sin(kr);
    if (ktype == 0) {
        const double ir = 1.0 / r;
        return cpx(c * ir, s * ir);
    }
    return cpx(c, s);
}

// Sources as separate coordinate and density arrays (the point sums vectorise over j).
struct Src {
    std::vector<double> x, y, z, re, im;
    int n() const { return (int)x.size(); }
    void add(double px, double py, double pz, cpx v) {
        x.push_back(px);
        y.push_back(py);
        z.push_back(pz);
        re.push_back(v.real());
        im.push_back(v.imag());
    }
};

// sum_j G(t, s_j) v_j, the kernel of green() written out for vectorisation
template <int KT>
cpx ksum_t(double tx, double ty, double tz, const Src& s) {
    const double* sx = s.x.data();
    const double* sy = s.y.data();
    const double* sz = s.z.data();
    const double* vr = s.re.data();
    const double* vi = s.im.data();
    const int n = s.n();
    double ar = 0, ai = 0;
#pragma omp simd reduction(+ : ar, ai)
    for (int j = 0; j < n; ++j) {
        const double dx = tx - sx[j], dy = ty - sy[j], dz = tz - sz[j];
        double r2 = dx * dx + dy * dy + dz * dz;
        r2 = r2 < MINDIF2 ? 1.0 : r2;
        const double r = std::sqrt(r2), kr = r * TPK;
        double c = std::cos(kr), sn = std::sin(kr);
        if (KT == 0) {
            const double ir = 1.0 / r;
            c *= ir;
            sn *= ir;
        }
        ar += c * vr[j] - sn * vi[j];
        ai += c * vi[j] + sn * vr[j];
    }
    return cpx(ar, ai);
}
inline cpx ksum(double tx, double ty, double tz, const Src& s, int kt) {
    return kt == 0 ? ksum_t<0>(tx, ty, tz, s) : ksum_t<1>(tx, ty, tz, s);
}

// dense kernel matrix G(t_i + ts, s_j + ss) (m x n row-major); t, s: point rows
std::vector<cpx> kmat(const std::vector<double>& t, const double ts[3], const std::vector<double>& s,
                      const double ss[3], int ktype) {
    const size_t m = t.size() / 3, n = s.size() / 3;
    std::vector<cpx> a(m * n);
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j)
            a[i * n + j] = green((t[3 * i] + ts[0]) - (s[3 * j] + ss[0]), (t[3 * i + 1] + ts[1]) - (s[3 * j + 1] + ss[1]),
                                 (t[3 * i + 2] + ts[2]) - (s[3 * j + 2] + ss[2]), ktype);
    return a;
}

// y (+)= A x, A m x n row-major
void gemv(const cpx* A, int m, int n, const cpx* x, cpx* y, bool add) {
    for (int i = 0; i < m; ++i) {
        cpx acc = 0;
        const cpx* a = A + (size_t)i * n;
        for (int j = 0; j < n; ++j) acc += a[j] * x[j];
        y[i] = add ? y[i] + acc : acc;
    }
}
// y = A^T x, A m x n row-major (transpose, not conjugate)
void gemvT(const cpx* A, int m, int n, const cpx* x, cpx* y) {
    for (int j = 0; j < n; ++j) y[j] = 0;
    for (int i = 0; i < m; ++i) {
        const cpx* a = A + (size_t)i * n;
        const cpx xi = x[i];
        for (int j = 0; j < n; ++j) y[j] += a[j] * xi;
    }
}

struct CMat {
    int m = 0, n = 0;
    std::vector<cpx> a;
};

struct LfOp {
    std::vector<double> uep, ucp;   // rows
    CMat V, invs, UH;
    std::map<Dir, std::vector<cpx>> ue2dc;   // offset -> (2P)^3 tensor, C order
};
struct HfOp {
    std::vector<double> uep, ucp;
    CMat E1, E2, E3;
};

struct Box {
    int tag = 0;
    std::vector<int> pts;
    std::vector<Key> U, V, W, X, E;
    std::map<Dir, std::vector<Key>> fnde;
    std::set<Dir> incdir, outdir;
    int fftnum = 0;
};

struct Oracle {
    std::vector<double> pos;
    int64_t N = 0;
    double K = 1, ctr[3] = {0, 0, 0};
    int ktype = 0, P = 6, NPQ = 4, ptsmax = 100, maxlevel = 12, numC = 1;
    std::map<double, LfOp> ld;
    std::map<double, std::map<Dir, HfOp>> hd;
    std::map<Key, Box> boxes;

    int unitlevel() const { return (int)std::lround(std::log(K) / std::log(2.0)); }
    int cell_level() const { return (int)std::lround(std::log((double)numC) / std::log(2.0)); }
    double width(const Key& k) const { return K / (double)(1 << k.l); }
    void center(const Key& k, double c[3]) const {
        const double t = (double)(1 << k.l);
        const int idx[3] = {k.x, k.y, k.z};
        for (int d = 0; d < 3; ++d) c[d] = ctr[d] - K / 2 + (idx[d] + 0.5) / t * K;
    }
    bool has(const Key& k) const {
        auto it = boxes.find(k);
        return it != boxes.end() && (it->second.tag & PTS);
    }

    // wave3d_setup.cpp:100-299
    void setup() {
        const int lvlC = cell_level(), ul = unitlevel();
        std::map<Key, std::vector<int>> cells;
        for (int64_t p = 0; p < N; ++p) {
            int ix[3];
            for (int d = 0; d < 3; ++d) {
                ix[d] = (int)std::floor(numC * ((pos[3 * p + d] - ctr[d] + K / 2) / K));
                if (ix[d] < 0 || ix[d] >= numC) throw std::runtime_error("point outside the computational domain");
            }
            cells[{lvlC, ix[0], ix[1], ix[2]}].push_back((int)p);
        }
        std::vector<std::pair<Key, std::vector<int>>> q;
        for (int a = 0; a < numC; ++a)
            for (int b = 0; b < numC; ++b)
                for (int c = 0; c < numC; ++c) {
                    Key k{lvlC, a, b, c};
                    auto it = cells.find(k);
                    q.push_back({k, it == cells.end() ? std::vector<int>() : it->second});
                }
        for (size_t h = 0; h < q.size(); ++h) {
            Key key = q[h].first;
            std::vector<int> pts = std::move(q[h].second);
            Box box;
            const int np = (int)pts.size();
            if (np > 0) box.tag |= PTS;
            const bool action = (key.l <= ul && np > 0) || (np > ptsmax && key.l < maxlevel - 1);
            if (action) {
                double cc[3];
                center(key, cc);
                std::vector<int> ch[8];
                for (int p : pts) {
                    const int code = (pos[3 * p] >= cc[0] ? 4 : 0) + (pos[3 * p + 1] >= cc[1] ? 2 : 0) +
                                     (pos[3 * p + 2] >= cc[2] ? 1 : 0);
                    ch[code].push_back(p);
                }
                for (int ind = 0; ind < 8; ++ind) q.push_back({child_key(key, ind), std::move(ch[ind])});
            } else {
                box.tag |= TERMINAL;
                box.pts = std::move(pts);
            }
            boxes[key] = std::move(box);
        }
        const double eps = 1e-12;
        for (auto& kv : boxes) {
            if (!(kv.second.tag & PTS)) continue;
            if (width(kv.first) < 1 - eps) callowlist(kv.first, kv.second);
            else calhghlist(kv.first, kv.second);
        }
        for (auto& kv : boxes)
            if (kv.second.tag & PTS)
                for (auto& nk : kv.second.V) boxes.at(nk).fftnum++;
        for (auto& kv : boxes) {
            const Key& key = kv.first;
            Box& box = kv.second;
            const double W = width(key);
            if (!(W > 1 - eps && (box.tag & PTS))) continue;
            if (key.l != lvlC) {
                const Box& par = boxes.at(parent_key(key));
                for (auto& d : par.outdir) box.outdir.insert(parent_dir(d));
                for (auto& d : par.incdir) box.incdir.insert(parent_dir(d));
            }
            double cc[3];
            center(key, cc);
            for (auto& fl : box.fnde)
                for (auto& ok : fl.second) {
                    double oc[3], t[3];
                    center(ok, oc);
                    for (int d = 0; d < 3; ++d) t[d] = oc[d] - cc[d];
                    double nr = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    for (int d = 0; d < 3; ++d) t[d] /= nr;
                    box.outdir.insert(nml2dir(t, W, NPQ));
                    for (int d = 0; d < 3; ++d) t[d] = cc[d] - oc[d];
                    nr = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                    for (int d = 0; d < 3; ++d) t[d] /= nr;
                    box.incdir.insert(nml2dir(t, W, NPQ));
                }
        }
    }

    // wave3d_setup.cpp:446-461
    bool find(const Key& key, Key& res) const {
        const int lvlC = cell_level();
        Key k = key;
        while (k.l != lvlC) {
            if (boxes.count(k)) {
                res = k;
                return true;
            }
            k = parent_key(k);
        }
        res = k;
        return boxes.count(k) > 0;
    }
    // wave3d_setup.cpp:464-477
    static bool adjacent(const Key& me, const Key& yo) {
        const int md = std::max(me.l, yo.l);
        const int sm = 1 << (md - me.l), sy = 1 << (md - yo.l);
        const int rad = sm + sy;
        const int a[3] = {me.x, me.y, me.z}, b[3] = {yo.x, yo.y, yo.z};
        int mx = 0;
        for (int d = 0; d < 3; ++d) {
            const int dif = std::abs((2 * a[d] + 1) * sm - (2 * b[d] + 1) * sy);
            if (dif > rad) return false;
            mx = std::max(mx, dif);
        }
        return mx == rad;
    }

    // wave3d_setup.cpp:302-390
    void callowlist(const Key& key, Box& box) {
        const int lvlC = cell_level();
        if (key.l == lvlC) throw std::runtime_error("cell-level boxes must have width >= 1");
        std::set<Key> U, V, Wl, X;
        const Key par = parent_key(key);
        const int L = 1 << key.l;
        const int pc[3] = {par.x, par.y, par.z};
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::max(2 * pc[d] - 2, 0);
            hi[d] = std::min(2 * pc[d] + 4, L);
        }
        const bool term = box.tag & TERMINAL;
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    if (i == key.x && j == key.y && k == key.z) continue;
                    Key res;
                    if (!find({key.l, i, j, k}, res)) continue;
                    const Box& rb = boxes.at(res);
                    const bool adj = adjacent(res, key);
                    if (res.l < key.l && (rb.tag & PTS)) {
                        if (!adj) X.insert(res);
                        else if (term) U.insert(res);
                    }
                    if (res.l == key.l) {
                        if (!adj) {
                            if (rb.tag & PTS) V.insert(res);
                        } else if (term) {
                            std::vector<Key> rest{res};
                            for (size_t h = 0; h < rest.size(); ++h) {
                                const Key fk = rest[h];
                                const Box& fb = boxes.at(fk);
                                const bool a2 = adjacent(fk, key);
                                if (!a2 && (fb.tag & PTS)) Wl.insert(fk);
                                if (a2 && (fb.tag & TERMINAL) && (fb.tag & PTS)) U.insert(fk);
                                if (a2 && !(fb.tag & TERMINAL))
                                    for (int ind = 0; ind < 8; ++ind) rest.push_back(child_key(fk, ind));
                            }
                        }
                    }
                }
        if (term && (box.tag & PTS)) U.insert(key);
        box.U.assign(U.begin(), U.end());
        box.V.assign(V.begin(), V.end());
        box.W.assign(Wl.begin(), Wl.end());
        box.X.assign(X.begin(), X.end());
    }

    // wave3d_setup.cpp:393-443
    void calhghlist(const Key& key, Box& box) {
        double cc[3];
        center(key, cc);
        const double W = width(key);
        const double thr = W * W + W - 1e-12;
        const int lvlC = cell_level();
        std::vector<Key> cands;
        if (key.l == lvlC) {
            for (auto& kv : boxes)
                if (kv.first.l == lvlC) cands.push_back(kv.first);
        } else {
            for (auto& tk : boxes.at(parent_key(key)).E)
                for (int ind = 0; ind < 8; ++ind) cands.push_back(child_key(tk, ind));
        }
        for (auto& ok : cands) {
            auto it = boxes.find(ok);
            if (it == boxes.end() || !(it->second.tag & PTS)) continue;
            double oc[3], diff[3];
            center(ok, oc);
            for (int d = 0; d < 3; ++d) diff[d] = cc[d] - oc[d];
            const double nrm = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
            if (nrm >= thr) {
                for (int d = 0; d < 3; ++d) diff[d] /= nrm;
                box.fnde[nml2dir(diff, W, NPQ)].push_back(ok);
            } else {
                box.E.push_back(ok);
            }
        }
    }

    // -- operator fetch (mlib3d.cpp:106-280)
    const LfOp& lf(double W) const {
        auto it = ld.find(W);
        if (it == ld.end()) throw std::runtime_error("LF operator width missing");
        return it->second;
    }
    struct HfPts {
        std::vector<double> uep, ucp;
        const HfOp* op;
    };
    HfPts hf(double W, const Dir& d) const {
        Dir srt;
        int sgn[3], prm[3];
        sort_dir(d, srt, sgn, prm);
        auto it = hd.find(W);
        if (it == hd.end()) throw std::runtime_error("HF operator width missing");
        auto jt = it->second.find(srt);
        if (jt == it->second.end()) throw std::runtime_error("HF operator direction missing");
        return {shuffle_pts(prm, sgn, jt->second.uep), shuffle_pts(prm, sgn, jt->second.ucp), &jt->second};
    }

    std::vector<cpx> eval(const cpx* den);
};

// ---------------------------------------------------------------- 12^3 DFT (numpy fftn convention)
struct Dft {
    int G;
    std::vector<cpx> w;   // w[k * G + n] = exp(-2 pi i k n / G)
    explicit Dft(int g) : G(g), w((size_t)g * g) {
        for (int k = 0; k < G; ++k)
            for (int n = 0; n < G; ++n) w[(size_t)k * G + n] = std::polar(1.0, -2.0 * M_PI * ((k * n) % G) / G);
    }
    // in-place 3-D transform of a C-order G^3 array; inverse: conjugate twiddles, / G^3
    void run(std::vector<cpx>& a, bool inverse) const {
        std::vector<cpx> line(G), out(G);
        const int s[3] = {G * G, G, 1};
        for (int ax = 0; ax < 3; ++ax) {
            const int o1 = s[(ax + 1) % 3], o2 = s[(ax + 2) % 3], st = s[ax];
            for (int i = 0; i < G; ++i)
                for (int j = 0; j < G; ++j) {
                    const int base = i * o1 + j * o2;
                    for (int n = 0; n < G; ++n) line[n] = a[base + n * st];
                    for (int k = 0; k < G; ++k) {
                        cpx acc = 0;
                        for (int n = 0; n < G; ++n) acc += (inverse ? std::conj(w[(size_t)k * G + n]) : w[(size_t)k * G + n]) * line[n];
                        out[k] = acc;
                    }
                    for (int k = 0; k < G; ++k) a[base + k * st] = out[k];
                }
        }
        if (inverse) {
            const double sc = 1.0 / ((double)G * G * G);
            for (auto& v : a) v *= sc;
        }
    }
};

static double wall() {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0;
#endif
}

std::vector<cpx> Oracle::eval(const cpx* den) {
    const bool trace = std::getenv("WR_TRACE") != nullptr;
    double tsec = wall();
    auto mark = [&](const char* what) {
        if (!trace) return;
        const double t = wall();
        std::fprintf(stderr, "[wave3d_ref] %s %.3f s\n", what, t - tsec);
        tsec = t;
    };
    const double eps = 1e-12;
    const int kt = ktype;
    // live boxes and their slots
    std::vector<Key> live;
    std::map<Key, int> id;
    for (auto& kv : boxes)
        if (kv.second.tag & PTS) {
            id[kv.first] = (int)live.size();
            live.push_back(kv.first);
        }
    const int nb = (int)live.size();
    std::vector<std::vector<double>> extpos(nb);   // leaf points (rows)
    std::vector<Src> extsrc(nb);                    // leaf points with their densities
    for (int b = 0; b < nb; ++b) {
        const Box& bx = boxes.at(live[b]);
        if (!(bx.tag & TERMINAL)) continue;
        for (int p : bx.pts) {
            for (int d = 0; d < 3; ++d) extpos[b].push_back(pos[3 * p + d]);
            extsrc[b].add(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], den[p]);
        }
    }
    std::map<double, std::vector<int>> ldmap;       // W -> boxes
    std::map<Dir, std::pair<std::vector<int>, std::vector<int>>> hdmap;
    for (int b = 0; b < nb; ++b) {
        const double W = width(live[b]);
        const Box& bx = boxes.at(live[b]);
        if (W < 1 - eps) {
            ldmap[W].push_back(b);
        } else {
            for (auto& d : bx.outdir) hdmap[d].first.push_back(b);
            for (auto& d : bx.incdir) hdmap[d].second.push_back(b);
        }
    }
    std::vector<std::vector<cpx>> upeq(nb), dnchk(nb);
    const double zero3[3] = {0, 0, 0};

    // ---- LF upward, finest first (wave3d_eval.cpp:418-490)
    for (auto& lv : ldmap) {
        const double W = lv.first;
        const LfOp& op = lf(W);
        const int nc = (int)op.ucp.size() / 3;
        std::vector<std::vector<cpx>> ue2uc(8);
        bool has_children = false;
        for (int b : lv.second)
            if (!(boxes.at(live[b]).tag & TERMINAL)) has_children = true;
        if (has_children) {
            const LfOp& oc = lf(W / 2);
            for (int ind = 0; ind < 8; ++ind) {
                double s[3];
                shifted_point(ind, W, s);
                ue2uc[ind] = kmat(op.ucp, zero3, oc.uep, s, kt);
            }
        }
        const std::vector<int>& bl = lv.second;
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t h = 0; h < bl.size(); ++h) {
            const int b = bl[h];
            const Key& k = live[b];
            const Box& bx = boxes.at(k);
            double c[3];
            center(k, c);
            std::vector<cpx> chk(nc, 0.0);
            if (bx.tag & TERMINAL) {
                for (int i = 0; i < nc; ++i)
                    chk[i] = ksum(op.ucp[3 * i] + c[0], op.ucp[3 * i + 1] + c[1], op.ucp[3 * i + 2] + c[2], extsrc[b], kt);
            } else {
                for (int ind = 0; ind < 8; ++ind) {
                    auto it = id.find(child_key(k, ind));
                    if (it == id.end()) continue;
                    const std::vector<cpx>& x = upeq[it->second];
                    gemv(ue2uc[ind].data(), nc, (int)x.size(), x.data(), chk.data(), true);
                }
            }
            std::vector<cpx> t(op.UH.m);
            gemv(op.UH.a.data(), op.UH.m, op.UH.n, chk.data(), t.data(), false);
            for (int i = 0; i < op.UH.m; ++i) t[i] *= op.invs.a[i];
            upeq[b].resize(op.V.m);
            gemv(op.V.a.data(), op.V.m, op.V.n, t.data(), upeq[b].data(), false);
        }
    }

    mark("LF upward");
    // ---- HF: one slot per (box, direction) of the out / in lists
    std::map<std::pair<int, Dir>, int> oslot, islot;
    std::map<double, std::vector<Dir>> hfl;    // W -> directions
    for (auto& kv : hdmap) {
        const Dir& d = kv.first;
        const int C = std::max(std::abs(d.v[0]), std::max(std::abs(d.v[1]), std::abs(d.v[2])));
        hfl[(double)(C / NPQ)].push_back(d);
        for (int b : kv.second.first) oslot.emplace(std::make_pair(b, d), (int)oslot.size());
        for (int b : kv.second.second) islot.emplace(std::make_pair(b, d), (int)islot.size());
    }
    std::vector<std::vector<cpx>> dirup(oslot.size()), dirdn(islot.size()), dneq(islot.size());

    // HF upward, W ascending (EvalUpwardHigh, wave3d_eval.cpp:773-854)
    for (auto& lv : hfl) {
        const double W = lv.first;
        const std::vector<Dir>& dirs = lv.second;
        std::vector<std::pair<int, int>> work;   // (direction index, box)
        for (int di = 0; di < (int)dirs.size(); ++di)
            for (int b : hdmap.at(dirs[di]).first) work.push_back({di, b});
        std::vector<HfPts> pts(dirs.size());
        std::vector<std::vector<std::vector<cpx>>> ue2uc(dirs.size());
#pragma omp parallel for schedule(dynamic, 1)
        for (int di = 0; di < (int)dirs.size(); ++di) {
            const Dir& d = dirs[di];
            pts[di] = hf(W, d);
            std::vector<double> uepc = W == 1.0 ? lf(0.5).uep : hf(W / 2, parent_dir(d)).uep;
            ue2uc[di].resize(8);
            for (int ind = 0; ind < 8; ++ind) {
                double s[3];
                shifted_point(ind, W, s);
                ue2uc[di][ind] = kmat(pts[di].ucp, zero3, uepc, s, kt);
            }
        }
#pragma omp parallel for schedule(dynamic, 8)
        for (size_t h = 0; h < work.size(); ++h) {
            const int di = work[h].first, b = work[h].second;
            const Dir& d = dirs[di];
            const HfOp& op = *pts[di].op;
            const int r2 = (int)pts[di].ucp.size() / 3;
            std::vector<cpx> chk(r2, 0.0);
            const Key& k = live[b];
            Dir pd{};
            if (W != 1.0) pd = parent_dir(d);
            for (int ind = 0; ind < 8; ++ind) {
                auto it = id.find(child_key(k, ind));
                if (it == id.end()) continue;
                const std::vector<cpx>* x;
                if (W == 1.0) {
                    x = &upeq[it->second];
                } else {
                    auto os = oslot.find({it->second, pd});
                    if (os == oslot.end()) throw std::runtime_error("child lacks the parent direction");
                    x = &dirup[os->second];
                }
                gemv(ue2uc[di][ind].data(), r2, (int)x->size(), x->data(), chk.data(), true);
            }
            std::vector<cpx> t3(op.E3.m), t2(op.E2.m), out(op.E1.m);
            gemv(op.E3.a.data(), op.E3.m, op.E3.n, chk.data(), t3.data(), false);
            gemv(op.E2.a.data(), op.E2.m, op.E2.n, t3.data(), t2.data(), false);
            gemv(op.E1.a.data(), op.E1.m, op.E1.n, t2.data(), out.data(), false);
            dirup[oslot.at({b, d})] = std::move(out);
        }
    }

    // HF downward, W descending (EvalDownwardHigh 996-1040, HighFrequencyM2L 876-928,
    // HighFrequencyL2L 931-993); a level's L2L contributions are pulled by each child
    // (box, parent direction), in the order of the parent's sorted in-directions
    for (auto lt = hfl.rbegin(); lt != hfl.rend(); ++lt) {
        const double W = lt->first;
        const std::vector<Dir>& dirs = lt->second;
        std::map<Dir, int> dix;
        for (int di = 0; di < (int)dirs.size(); ++di) dix[dirs[di]] = di;
        std::vector<HfPts> pts(dirs.size());
        // M2L matrices per (direction, centre offset)
        std::map<std::pair<int, Dir>, int> mid;
        std::vector<std::pair<int, Dir>> mkeys;
        std::vector<std::pair<int, int>> work;
        for (int di = 0; di < (int)dirs.size(); ++di) {
            pts[di] = hf(W, dirs[di]);
            for (int b : hdmap.at(dirs[di]).second) {
                work.push_back({di, b});
                const Box& bx = boxes.at(live[b]);
                auto f = bx.fnde.find(dirs[di]);
                if (f == bx.fnde.end()) continue;
                double c[3];
                center(live[b], c);
                for (auto& sk : f->second) {
                    double cs[3];
                    center(sk, cs);
                    Dir off;
                    for (int q = 0; q < 3; ++q) off.v[q] = (int)std::nearbyint((c[q] - cs[q]) / W);
                    if (mid.emplace(std::make_pair(di, off), (int)mkeys.size()).second) mkeys.push_back({di, off});
                }
            }
        }
        std::vector<std::vector<cpx>> mats(mkeys.size());
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t h = 0; h < mkeys.size(); ++h) {
            const int di = mkeys[h].first;
            const Dir& off = mkeys[h].second;
            std::vector<double> dcp(pts[di].uep.size());
            for (size_t e = 0; e < dcp.size(); ++e) dcp[e] = -pts[di].uep[e];
            const double ts[3] = {off.v[0] * W, off.v[1] * W, off.v[2] * W};
            mats[h] = kmat(dcp, ts, pts[di].uep, zero3, kt);
        }
#pragma omp parallel for schedule(dynamic, 8)
        for (size_t h = 0; h < work.size(); ++h) {
            const int di = work[h].first, b = work[h].second;
            const Dir& d = dirs[di];
            const HfOp& op = *pts[di].op;
            const int r1 = (int)pts[di].uep.size() / 3;
            const int si = islot.at({b, d});
            std::vector<cpx>& dcv = dirdn[si];
            if (dcv.empty()) dcv.assign(r1, 0.0);
            const Box& bx = boxes.at(live[b]);
            auto f = bx.fnde.find(d);
            if (f != bx.fnde.end()) {
                double c[3];
                center(live[b], c);
                for (auto& sk : f->second) {
                    double cs[3];
                    center(sk, cs);
                    Dir off;
                    for (int q = 0; q < 3; ++q) off.v[q] = (int)std::nearbyint((c[q] - cs[q]) / W);
                    auto os = oslot.find({id.at(sk), d});
                    if (os == oslot.end()) continue;   // a source without this direction adds zero
                    gemv(mats[mid.at({di, off})].data(), r1, r1, dirup[os->second].data(), dcv.data(), true);
                }
            }
            std::vector<cpx> t1(op.E1.n), t2(op.E2.n), out(op.E3.n);
            gemvT(op.E1.a.data(), op.E1.m, op.E1.n, dcv.data(), t1.data());
            gemvT(op.E2.a.data(), op.E2.m, op.E2.n, t1.data(), t2.data());
            gemvT(op.E3.a.data(), op.E3.m, op.E3.n, t2.data(), out.data());
            dneq[si] = std::move(out);
        }
        // child kernels de2dc per direction
        std::vector<std::vector<std::vector<cpx>>> de2dc(dirs.size());
#pragma omp parallel for schedule(dynamic, 1)
        for (int di = 0; di < (int)dirs.size(); ++di) {
            std::vector<double> dep(pts[di].ucp.size()), dcpc;
            for (size_t e = 0; e < dep.size(); ++e) dep[e] = -pts[di].ucp[e];
            if (W == 1.0) {
                dcpc = lf(0.5).uep;
            } else {
                dcpc = hf(W / 2, parent_dir(dirs[di])).uep;
                for (auto& v : dcpc) v = -v;
            }
            de2dc[di].resize(8);
            for (int ind = 0; ind < 8; ++ind) {
                double s[3];
                shifted_point(ind, W, s);
                de2dc[di][ind] = kmat(dcpc, s, dep, zero3, kt);
            }
        }
        // pull: every child box of a box with in-directions at W
        std::vector<int> parents;
        {
            std::set<int> ps;
            for (auto& w : work) ps.insert(w.second);
            parents.assign(ps.begin(), ps.end());
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t h = 0; h < parents.size(); ++h) {
            const int b = parents[h];
            const Box& bx = boxes.at(live[b]);
            for (int ind = 0; ind < 8; ++ind) {
                auto it = id.find(child_key(live[b], ind));
                if (it == id.end()) continue;
                const int cb = it->second;
                for (auto& d : bx.incdir) {
                    auto dj = dix.find(d);
                    if (dj == dix.end()) continue;
                    const int di = dj->second;
                    const std::vector<cpx>& x = dneq[islot.at({b, d})];
                    const std::vector<cpx>& A = de2dc[di][ind];
                    const int m = (int)A.size() / (int)x.size();
                    std::vector<cpx>* y;
                    if (W == 1.0) {
                        y = &dnchk[cb];
                    } else {
                        y = &dirdn[islot.at({cb, parent_dir(d)})];
                    }
                    if (y->empty()) y->assign(m, 0.0);
                    gemv(A.data(), m, (int)x.size(), x.data(), y->data(), true);
                }
            }
        }
    }

    mark("HF passes");
    // ---- LF downward, coarsest first (EvalDownwardLow, wave3d_eval.cpp:493-583)
    std::vector<std::vector<cpx>> extval(nb);
    const int G = 2 * P;
    const Dft dft(G);
    for (auto lt = ldmap.rbegin(); lt != ldmap.rend(); ++lt) {
        const double W = lt->first;
        const LfOp& op = lf(W);
        const std::vector<double>& uep = op.uep;   // dcp of the downward pass
        const std::vector<double>& dep = op.ucp;
        const int ne = (int)uep.size() / 3, nchk = ne, nde = (int)dep.size() / 3;
        const double step = W / (P - 1);
        std::vector<int> gi(ne);
        for (int e = 0; e < ne; ++e) {
            int g3[3];
            for (int d = 0; d < 3; ++d) g3[d] = (int)std::nearbyint((uep[3 * e + d] + W / 2) / step) + P;
            gi[e] = (g3[0] * G + g3[1]) * G + g3[2];
        }
        // spectra of this level's V sources
        std::map<int, int> sidx;
        std::vector<int> srcs;
        for (int b : lt->second)
            for (auto& nk : boxes.at(live[b]).V) {
                const int s = id.at(nk);
                if (sidx.emplace(s, (int)srcs.size()).second) srcs.push_back(s);
            }
        std::vector<std::vector<cpx>> spec(srcs.size());
#pragma omp parallel for schedule(dynamic, 4)
        for (size_t h = 0; h < srcs.size(); ++h) {
            std::vector<cpx> g((size_t)G * G * G, 0.0);
            const std::vector<cpx>& u = upeq[srcs[h]];
            for (int e = 0; e < ne; ++e) g[gi[e]] = u[e];
            dft.run(g, false);
            spec[h] = std::move(g);
        }
        std::vector<std::vector<cpx>> de2dc(8);
        bool has_children = false;
        for (int b : lt->second)
            if (!(boxes.at(live[b]).tag & TERMINAL)) has_children = true;
        if (has_children) {
            const LfOp& oc = lf(W / 2);
            for (int ind = 0; ind < 8; ++ind) {
                double s[3];
                shifted_point(ind, W, s);
                de2dc[ind] = kmat(oc.uep, s, dep, zero3, kt);
            }
        }
        const std::vector<int>& bl = lt->second;
#pragma omp parallel for schedule(dynamic, 2)
        for (size_t h = 0; h < bl.size(); ++h) {
            const int b = bl[h];
            const Key& k = live[b];
            const Box& bx = boxes.at(k);
            double c[3];
            center(k, c);
            std::vector<cpx> chk = dnchk[b];
            if (chk.empty()) chk.assign(nchk, 0.0);
            const bool term = bx.tag & TERMINAL;
            const std::vector<double>& tp = extpos[b];
            const int nt = extsrc[b].n();
            std::vector<cpx> ev(nt, 0.0);
            auto direct = [&](const Src& src) {
                for (int i = 0; i < nt; ++i) ev[i] += ksum(tp[3 * i], tp[3 * i + 1], tp[3 * i + 2], src, kt);
            };
            for (auto& nk : bx.U) direct(extsrc[id.at(nk)]);
            if (!bx.V.empty()) {
                std::vector<cpx> acc((size_t)G * G * G, 0.0);
                for (auto& nk : bx.V) {
                    double cn[3];
                    center(nk, cn);
                    Dir off;
                    for (int d = 0; d < 3; ++d) off.v[d] = (int)cround((c[d] - cn[d]) / W);
                    auto tt = op.ue2dc.find(off);
                    if (tt == op.ue2dc.end()) throw std::runtime_error("V-list offset without a tensor");
                    const std::vector<cpx>& s = spec[sidx.at(id.at(nk))];
                    for (size_t e = 0; e < acc.size(); ++e) acc[e] += s[e] * tt->second[e];
                }
                dft.run(acc, true);
                for (int e = 0; e < ne; ++e) chk[e] += acc[gi[e]];
            }
            for (auto& nk : bx.W) {
                const int s = id.at(nk);
                const Box& nbx = boxes.at(nk);
                if ((nbx.tag & TERMINAL) && (int)nbx.pts.size() < ne) {
                    direct(extsrc[s]);
                } else {
                    const double coef = width(nk) / W;
                    double cn[3];
                    center(nk, cn);
                    Src sp;
                    for (int e = 0; e < ne; ++e)
                        sp.add(coef * uep[3 * e] + cn[0], coef * uep[3 * e + 1] + cn[1], coef * uep[3 * e + 2] + cn[2],
                               upeq[s][e]);
                    direct(sp);
                }
            }
            for (auto& nk : bx.X) {
                const int s = id.at(nk);
                if (term && nt < nchk) {
                    direct(extsrc[s]);
                } else {
                    for (int i = 0; i < nchk; ++i)
                        chk[i] += ksum(uep[3 * i] + c[0], uep[3 * i + 1] + c[1], uep[3 * i + 2] + c[2], extsrc[s], kt);
                }
            }
            // dneq = UH^T ((V^T chk) .* invs)
            std::vector<cpx> t(op.V.n), dq(op.UH.n);
            gemvT(op.V.a.data(), op.V.m, op.V.n, chk.data(), t.data());
            for (int i = 0; i < op.V.n; ++i) t[i] *= op.invs.a[i];
            gemvT(op.UH.a.data(), op.UH.m, op.UH.n, t.data(), dq.data());
            if (term) {
                Src sp;
                for (int e = 0; e < nde; ++e) sp.add(dep[3 * e] + c[0], dep[3 * e + 1] + c[1], dep[3 * e + 2] + c[2], dq[e]);
                direct(sp);
                extval[b] = std::move(ev);
            } else {
                for (int ind = 0; ind < 8; ++ind) {
                    auto it = id.find(child_key(k, ind));
                    if (it == id.end()) continue;
                    std::vector<cpx>& y = dnchk[it->second];
                    const int m = (int)de2dc[ind].size() / nde;
                    if (y.empty()) y.assign(m, 0.0);
                    gemv(de2dc[ind].data(), m, nde, dq.data(), y.data(), true);
                }
            }
        }
    }
    mark("LF downward and leaves");
    std::vector<cpx> val(N, 0.0);
    for (int b = 0; b < nb; ++b) {
        const Box& bx = boxes.at(live[b]);
        if (!(bx.tag & TERMINAL)) continue;
        for (size_t i = 0; i < bx.pts.size(); ++i) val[bx.pts[i]] = extval[b][i];
    }
    return val;
}

thread_local std::string g_err;

CMat cmat(int m, int n, const double* a) {
    CMat c;
    c.m = m;
    c.n = n;
    c.a.assign((const cpx*)a, (const cpx*)a + (size_t)m * n);
    return c;
}

}  // namespace

extern "C" {

const char* wr_error(void) { return g_err.c_str(); }

void* wr_create(const double* pos, int64_t n, double K, const double* ctr, int ktype, int P, int NPQ, int ptsmax,
                int maxlevel, int numC) {
    auto* o = new Oracle();
    o->pos.assign(pos, pos + 3 * n);
    o->N = n;
    o->K = K;
    for (int d = 0; d < 3; ++d) o->ctr[d] = ctr[d];
    o->ktype = ktype;
    o->P = P;
    o->NPQ = NPQ;
    o->ptsmax = ptsmax;
    o->maxlevel = maxlevel;
    o->numC = numC;
    return o;
}

void wr_free(void* h) { delete (Oracle*)h; }

// LF entry of width W: uep (ne x 3 rows), ucp (nc x 3 rows), the pseudo-inverse factors
// V (ne x r), invs (r), UH (r x nc) row-major complex, and noff tensors ((2P)^3, C order)
// at offsets offs (noff x 3).
int wr_add_lf(void* h, double W, int ne, const double* uep, int nc, const double* ucp, int vm, int vn,
              const double* V, const double* invs, int um, int un, const double* UH, int noff, const int* offs,
              const double* tens) {
    auto* o = (Oracle*)h;
    LfOp op;
    op.uep.assign(uep, uep + 3 * ne);
    op.ucp.assign(ucp, ucp + 3 * nc);
    op.V = cmat(vm, vn, V);
    op.invs = cmat(vn, 1, invs);
    op.UH = cmat(um, un, UH);
    const size_t G3 = (size_t)(2 * o->P) * (2 * o->P) * (2 * o->P);
    for (int t = 0; t < noff; ++t) {
        Dir d{{offs[3 * t], offs[3 * t + 1], offs[3 * t + 2]}};
        op.ue2dc[d].assign((const cpx*)tens + t * G3, (const cpx*)tens + (t + 1) * G3);
    }
    o->ld[W] = std::move(op);
    return 0;
}

// HF entry of width W for the sorted direction srt: uep (r1 x 3), ucp (r2 x 3), E1..E3
// row-major complex with their shapes.
int wr_add_hf(void* h, double W, const int* srt, int r1, const double* uep, int r2, const double* ucp, int m1,
              int n1, const double* E1, int m2, int n2, const double* E2, int m3, int n3, const double* E3) {
    auto* o = (Oracle*)h;
    HfOp op;
    op.uep.assign(uep, uep + 3 * r1);
    op.ucp.assign(ucp, ucp + 3 * r2);
    op.E1 = cmat(m1, n1, E1);
    op.E2 = cmat(m2, n2, E2);
    op.E3 = cmat(m3, n3, E3);
    Dir d{{srt[0], srt[1], srt[2]}};
    o->hd[W][d] = std::move(op);
    return 0;
}

int wr_setup(void* h) {
    try {
        ((Oracle*)h)->setup();
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}

// counts: boxes with points, leaves, LF boxes, near-field interactions (U list)
int wr_stats(void* h, int64_t* out) {
    auto* o = (Oracle*)h;
    int64_t live = 0, leaves = 0, lfb = 0, near = 0;
    for (auto& kv : o->boxes) {
        const Box& b = kv.second;
        if (!(b.tag & PTS)) continue;
        ++live;
        if (b.tag & TERMINAL) ++leaves;
        if (o->width(kv.first) < 1 - 1e-12) ++lfb;
        for (auto& nk : b.U) near += (int64_t)b.pts.size() * (int64_t)o->boxes.at(nk).pts.size();
    }
    out[0] = live;
    out[1] = leaves;
    out[2] = lfb;
    out[3] = near;
    return 0;
}

// one right-hand side: den, val N complex; threads 0 = the OpenMP default
int wr_eval(void* h, const double* den, double* val, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    try {
        std::vector<cpx> v = ((Oracle*)h)->eval((const cpx*)den);
        std::memcpy(val, v.data(), v.size() * sizeof(cpx));
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}

}  // extern "C"
