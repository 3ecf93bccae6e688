// Builds against include/ddfmm_wave3d.hpp exactly as a caller of the reference's
// Wave3d/Mlib3d/Kernel3d would (tt.cpp:103-151 wiring), on a host-only context
// (device -1): setup plans the tree and lists, eval must report that no device is set up.
#include <cmath>
#include <cstdio>
#include <random>

#include "ddfmm_wave3d.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::map<std::string, std::string> opts = {
        {"-mlib3d_ldname", argv[1]}, {"-mlib3d_hdname", argv[2]},
        {"-wave3d_ACCU", "2"}, {"-wave3d_NPQ", "4"}, {"-wave3d_K", "4"}, {"-wave3d_ptsmax", "100"},
        {"-wave3d_maxlevel", "12"}};
    std::map<int, ddfmm::Point3> pos;
    std::map<int, ddfmm::cpx> den, val;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    for (int k = 0; k < 3000; ++k) {
        double x = nd(rng), y = nd(rng), z = nd(rng), r = std::sqrt(x * x + y * y + z * z);
        pos[k] = {2 * (1 - 1e-12) * x / r, 2 * (1 - 1e-12) * y / r, 2 * (1 - 1e-12) * z / r};
        den[k] = ddfmm::cpx(nd(rng), nd(rng));
        val[k] = den[k];
    }
    ddfmm::Kernel3d knl(ddfmm::KERNEL_HELM);
    ddfmm::Mlib3d mlib("mlib3d_");
    mlib.kernel() = knl;
    mlib.setup(opts);
    ddfmm::Wave3d wave("wave3d_", -1);
    wave.posptr() = &pos;
    wave.kernel() = knl;
    wave.mlibptr() = &mlib;
    wave.geomprtn() = 1;
    wave.setup(opts);
    ddfmm_stats st;
    ddfmm::check_rc(ddfmm_get_stats(wave.handle(), &st));
    std::printf("npts %lld leaves %lld m2l_pairs %lld\n", (long long)st.npts, (long long)st.nleaf,
                (long long)st.m2l_pairs);
    try {
        wave.eval(den, val);
        std::printf("unexpected: eval succeeded without a device\n");
        return 1;
    } catch (const ddfmm::Error& e) {
        std::printf("eval error: %s\n", e.what());
    }
    return 0;
}
