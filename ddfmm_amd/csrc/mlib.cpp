// Operator-library reader (Mlib3d::setup, mlib3d.cpp:61-103) for the binary layout of
// serialize.hpp:24-711, and the symmetry shuffle of mlib3d.cpp:308-325.
#include "host.hpp"

#include <cstring>
#include <fstream>
#include <iterator>

namespace ddfmm {

namespace {
struct Reader {
    std::vector<char> buf;
    size_t p = 0;
    explicit Reader(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error("cannot open operator file " + path);
        buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    void need(size_t n) {
        if (p + n > buf.size()) throw Error("truncated operator file");
    }
    int i32() { need(4); int v; std::memcpy(&v, &buf[p], 4); p += 4; return v; }
    double f64() { need(8); double v; std::memcpy(&v, &buf[p], 8); p += 8; return v; }
    DMat dmat() {
        DMat m;
        m.m = i32(); m.n = i32();
        size_t n = (size_t)m.m * m.n;
        need(8 * n);
        m.a.resize(n);
        std::memcpy(m.a.data(), &buf[p], 8 * n);
        p += 8 * n;
        return m;
    }
    CMat cmat(int dims) {
        CMat m;
        m.m = i32(); m.n = i32(); m.p = dims == 3 ? i32() : 1;
        size_t n = (size_t)m.m * m.n * m.p;
        need(16 * n);
        m.a.resize(n);
        std::memcpy((void*)m.a.data(), &buf[p], 16 * n);
        p += 16 * n;
        return m;
    }
};
}  // namespace

void Mlib::load(const std::string& ldname, const std::string& hdname) {
    {
        Reader r(ldname);
        int nw = r.i32();
        for (int w = 0; w < nw; ++w) {
            double W = r.f64();
            LowFreqEntry e;
            e.uep = r.dmat();
            e.ucp = r.dmat();
            int nm = r.i32();
            for (int k = 0; k < nm; ++k) e.uc2ue.push_back(r.cmat(2));
            int m = r.i32(), n = r.i32(), q = r.i32();
            for (int c = 0; c < q; ++c)
                for (int b = 0; b < n; ++b)
                    for (int a = 0; a < m; ++a) {
                        CMat t = r.cmat(3);
                        if (t.m * t.n * t.p > 0) e.ue2dc[Index3(a - 3, b - 3, c - 3)] = std::move(t);
                    }
            if (e.uc2ue.size() != 3) throw Error("LF entry must hold 3 factors");
            ld[W] = std::move(e);
        }
    }
    {
        Reader r(hdname);
        int nw = r.i32();
        for (int w = 0; w < nw; ++w) {
            double W = r.f64();
            int nd = r.i32();
            auto& inner = hd[W];
            for (int k = 0; k < nd; ++k) {
                Index3 key;
                key[0] = r.i32(); key[1] = r.i32(); key[2] = r.i32();
                HighFreqEntry e;
                e.uep = r.dmat();
                e.ucp = r.dmat();
                int nm = r.i32();
                for (int j = 0; j < nm; ++j) e.uc2ue.push_back(r.cmat(2));
                if (e.uc2ue.size() != 3) throw Error("HF entry must hold 3 factors");
                inner[key] = std::move(e);
            }
        }
    }
}

const LowFreqEntry& Mlib::low(double W) const {
    auto it = ld.find(W);
    if (it == ld.end()) throw Error("operator library lacks low-frequency width " + std::to_string(W));
    return it->second;
}

const HighFreqEntry& Mlib::high(double W, const Index3& c) const {
    auto it = hd.find(W);
    if (it == hd.end()) throw Error("operator library lacks high-frequency width " + std::to_string(W));
    auto jt = it->second.find(c);
    if (jt == it->second.end())
        throw Error("operator library lacks direction (" + std::to_string(c[0]) + "," +
                    std::to_string(c[1]) + "," + std::to_string(c[2]) + ") at width " + std::to_string(W));
    return jt->second;
}

std::vector<double> Mlib::shuffled(const Index3& prm, const Index3& sgn, const DMat& pts) {
    std::vector<double> out((size_t)pts.n * 3);
    for (int k = 0; k < pts.n; ++k)
        for (int d = 0; d < 3; ++d) out[3 * (size_t)k + prm[d]] = pts(d, k);
    for (int k = 0; k < pts.n; ++k)
        for (int d = 0; d < 3; ++d) out[3 * (size_t)k + d] *= sgn[d];
    return out;
}

}  // namespace ddfmm
