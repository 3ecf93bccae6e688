// Builds against include/ddfmm_wave3d.hpp exactly as a caller of the reference's
// Wave3d/Mlib3d/Kernel3d would (tt.cpp:103-151 wiring, the reference's containers), on a
// host-only context (-wave3d_device -1): setup plans the tree and lists, eval must report
// that no device is set up; the containers round-trip through the reference file format.
#include <cmath>
#include <cstdio>
#include <random>

#include "ddfmm_wave3d.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::vector<int> all(1, 1);
    std::map<std::string, std::string> opts = {
        {"-mlib3d_ldname", argv[1]}, {"-mlib3d_hdname", argv[2]},
        {"-wave3d_ACCU", "2"}, {"-wave3d_NPQ", "4"}, {"-wave3d_K", "4"}, {"-wave3d_ptsmax", "100"},
        {"-wave3d_maxlevel", "12"}, {"-wave3d_device", "-1"}};
    ParVec<int, Point3, PtPrtn> pos;
    ParVec<int, cpx, PtPrtn> den, val;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    for (int k = 0; k < 3000; ++k) {
        double x = nd(rng), y = nd(rng), z = nd(rng), r = std::sqrt(x * x + y * y + z * z);
        Point3 p(2 * (1 - 1e-12) * x / r, 2 * (1 - 1e-12) * y / r, 2 * (1 - 1e-12) * z / r);
        cpx d(nd(rng), nd(rng));
        pos.insert(k, p);
        den.insert(k, d);
    }
    pos.prtn().ownerinfo() = {0, 3000};
    den.prtn().ownerinfo() = {0, 3000};
    val = den;
    // reference file format round trip
    std::ostringstream os;
    serialize(pos, os, all);
    std::istringstream is(os.str());
    ParVec<int, Point3, PtPrtn> back;
    deserialize(back, is, all);
    if (back.lclmap().size() != 3000 || back.access(17)(2) != pos.access(17)(2) || back.prtn().owner(2999) != 0) {
        std::printf("round trip failed\n");
        return 1;
    }
    Kernel3d knl(KERNEL_HELM);
    DblNumMat trg(3, 2), src(3, 3), nor;
    CpxNumMat mat;
    for (int j = 0; j < 2; ++j) for (int d = 0; d < 3; ++d) trg(d, j) = pos.access(j)(d);
    for (int j = 0; j < 3; ++j) for (int d = 0; d < 3; ++d) src(d, j) = pos.access(10 + j)(d);
    std::printf("kernel matrix shape %d x %d\n", mat.m(), mat.n());
    Mlib3d mlib("mlib3d_");
    mlib.kernel() = knl;
    mlib.setup(opts);
    std::printf("mlib widths %zu %zu\n", mlib.w2ldmap().size(), mlib.w2hdmap().size());
    IntNumTns geomprtn(1, 1, 1);
    Wave3d wave("wave3d_");
    wave.posptr() = &pos;
    wave.kernel() = knl;
    wave.mlibptr() = &mlib;
    wave.geomprtn() = geomprtn;
    wave.setup(opts);
    ddfmm_stats st;
    ddfmm::check_rc(ddfmm_get_stats(wave.handle(), &st));
    std::printf("npts %lld leaves %lld m2l_pairs %lld\n", (long long)st.npts, (long long)st.nleaf,
                (long long)st.m2l_pairs);
    try {
        wave.eval(den, val);
        std::printf("unexpected: eval succeeded without a device\n");
        return 1;
    } catch (std::exception* e) {
        std::printf("eval error: %s\n", e->what());
        delete e;
    }
    return 0;
}
