// Host-side structures of the directional FMM: octree, interaction lists, direction
// sets and operator library.  Restates Wave3d::setup (wave3d_setup.cpp:21-547),
// the direction algebra of wave3d.cpp:43-158 and the operator fetch of mlib3d.cpp:22-357.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace ddfmm {

using cpx = std::complex<double>;

struct Index3 {
    int v[3] = {0, 0, 0};
    Index3() = default;
    Index3(int a, int b, int c) { v[0] = a; v[1] = b; v[2] = c; }
    int& operator[](int i) { return v[i]; }
    int operator[](int i) const { return v[i]; }
    int linfty() const { return std::max(std::abs(v[0]), std::max(std::abs(v[1]), std::abs(v[2]))); }
    bool operator<(const Index3& o) const {            // lexicographic (vec3t.hpp:98-107)
        for (int i = 0; i < 3; i++) {
            if (v[i] < o.v[i]) return true;
            if (v[i] > o.v[i]) return false;
        }
        return false;
    }
    bool operator==(const Index3& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};

using BoxKey = std::pair<int, Index3>;   // (level, index in level), map order = reference order

enum { TAG_PTS = 1, TAG_TERMINAL = 2 };

struct Box {
    int tag = 0;
    std::vector<int> ptidx;                 // input-order point ids (leaves)
    std::vector<BoxKey> U, V, W, X, E;      // LF lists; E = HF near list
    std::map<Index3, std::vector<BoxKey>> fnde;  // HF far list keyed by direction (target - source)
    std::set<Index3> incdir, outdir;
    int fftnum = 0;
    bool has_pts() const { return tag & TAG_PTS; }
    bool terminal() const { return tag & TAG_TERMINAL; }
};

struct Params {
    double K = 64;
    double ctr[3] = {0, 0, 0};
    int ACCU = 2, NPQ = 4, ptsmax = 100, maxlevel = 12, numC = 1, ktype = 0;
    bool fuse_solves = false;   // apply check->equivalent solves as single precomputed matrices
    int nrhs_max = 1;           // right-hand sides an evaluation may carry (GEMM launch sizing)
    bool near_pairs = false;    // near field: evaluate each pair of leaves that are in each
                                // other's near lists once, for both directions
};

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

// ---------------------------------------------------------------- direction algebra
Index3 nml2dir(const double n[3], double W, int NPQ);   // wave3d.cpp:43-71
Index3 parent_dir(const Index3& d);                     // wave3d.cpp:74-95 / mlib3d.cpp:283-303
std::vector<Index3> child_dirs(const Index3& d);        // wave3d.cpp:98-119
void sort_dir(Index3 val, Index3& srt, Index3& sgn, Index3& prm);  // mlib3d.cpp:328-357

// ---------------------------------------------------------------- tree
class Tree {
public:
    Params prm;
    const double* pos = nullptr;   // N x 3 row-major
    long N = 0;
    std::map<BoxKey, Box> boxes;

    void build();                  // setup_tree (wave3d_setup.cpp:100-299)

    int unitlevel() const { return int(std::round(std::log(prm.K) / std::log(2.0))); }
    int cell_level() const { return int(std::round(std::log(double(prm.numC)) / std::log(2.0))); }
    double width(const BoxKey& k) const { return prm.K / double(1 << k.first); }
    void center(const BoxKey& k, double c[3]) const;
    static BoxKey parent_key(const BoxKey& k) {
        return BoxKey(k.first - 1, Index3(k.second[0] / 2, k.second[1] / 2, k.second[2] / 2));
    }
    static BoxKey child_key(const BoxKey& k, int ind) {
        return BoxKey(k.first + 1, Index3(2 * k.second[0] + ((ind & 4) >> 2),
                                          2 * k.second[1] + ((ind & 2) >> 1),
                                          2 * k.second[2] + (ind & 1)));
    }
    static bool adjacent(const BoxKey& me, const BoxKey& yo);   // wave3d_setup.cpp:463-477

private:
    bool find(const BoxKey& want, BoxKey& res) const;           // wave3d_setup.cpp:446-460
    void callowlist(const BoxKey& key, Box& box);               // wave3d_setup.cpp:302-390
    void calhghlist(const BoxKey& key, Box& box);               // wave3d_setup.cpp:393-443
};

// ---------------------------------------------------------------- operator library
struct DMat {              // column-major real matrix (DblNumMat)
    int m = 0, n = 0;
    std::vector<double> a;
    double operator()(int i, int j) const { return a[i + (size_t)j * m]; }
};
struct CMat {              // column-major complex matrix (CpxNumMat / CpxNumTns)
    int m = 0, n = 0, p = 1;
    std::vector<cpx> a;
};

struct LowFreqEntry {       // mlib3d.hpp:27-39
    DMat uep, ucp;           // 3 x n
    std::vector<CMat> uc2ue; // {V, 1/s (r x 1), U^H}
    std::map<Index3, CMat> ue2dc;   // key = offset in [-3,3]^3 \ [-1,1]^3
};
struct HighFreqEntry {      // mlib3d.hpp:41-52
    DMat uep, ucp;
    std::vector<CMat> uc2ue; // {E1, E2, E3}
};

class Mlib {
public:
    std::map<double, LowFreqEntry> ld;
    std::map<double, std::map<Index3, HighFreqEntry>> hd;
    void load(const std::string& ldname, const std::string& hdname);  // Mlib3d::setup
    const LowFreqEntry& low(double W) const;
    const HighFreqEntry& high(double W, const Index3& canon) const;
    // directional points of an actual direction: rows (n x 3) after the symmetry
    // shuffle of mlib3d.cpp:308-325
    static std::vector<double> shuffled(const Index3& prm, const Index3& sgn, const DMat& pts);
};

}  // namespace ddfmm
