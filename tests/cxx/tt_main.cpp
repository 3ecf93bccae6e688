// The reference's tt driver (tt.cpp:43-210) written against include/ddfmm_wave3d.hpp:
// the same options, file reads and writes, setup / eval / check calls and RESULT block,
// with the MPI start-up and shutdown removed (one process drives one GPU).
#include <chrono>
#include <cstdio>
#include <ctime>
#include <exception>
#include <iostream>

#include "ddfmm_wave3d.hpp"

#define CLOCK_DIFF_SECS(ck1, ck0) (double(ck1 - ck0) / CLOCKS_PER_SEC)

int optionsCreate(int argc, char** argv, std::map<std::string, std::string>& options) {
    options.clear();
    for (int k = 1; k + 1 < argc; k += 2) options[std::string(argv[k])] = std::string(argv[k + 1]);
    return 0;
}

std::string findOption(std::map<std::string, std::string>& opts, std::string option) {
    auto mi = opts.find(option);
    if (mi == opts.end()) {
        std::cerr << "Missing option " << option << std::endl;
        return "";
    }
    return mi->second;
}

int main(int argc, char** argv) {
    try {
        std::vector<int> all(1, 1);
        std::map<std::string, std::string> opts;
        optionsCreate(argc, argv, opts);
        std::string opt = findOption(opts, "-dataroot");
        if (!opt.empty()) ddfmm::data_root() = opt;

        // 1. read data
        ParVec<int, Point3, PtPrtn> pos;
        opt = findOption(opts, "-posfile");
        if (opt.empty()) return 0;
        std::istringstream piss;
        SAFE_FUNC_EVAL(SeparateRead(opt, piss));
        SAFE_FUNC_EVAL(deserialize(pos, piss, all));
        std::vector<int>& tmpinfo = pos.prtn().ownerinfo();
        int numpts = tmpinfo[tmpinfo.size() - 1];
        std::cerr << "Total number of points: " << numpts << std::endl;
        ParVec<int, cpx, PtPrtn> den;
        opt = findOption(opts, "-denfile");
        if (opt.empty()) return 0;
        std::istringstream diss;
        SAFE_FUNC_EVAL(SeparateRead(opt, diss));
        SAFE_FUNC_EVAL(deserialize(den, diss, all));
        ParVec<int, cpx, PtPrtn> val;
        val = den;
        Kernel3d kernel(KERNEL_HELM);
        auto mi = opts.find("-kernel");
        if (mi != opts.end()) {
            std::istringstream ss(mi->second);
            int type;
            ss >> type;
            kernel.type() = type;
        }
        Mlib3d mlib("mlib3d_");
        mlib.kernel() = kernel;
        mlib.NPQ() = 4;
        SAFE_FUNC_EVAL(mlib.setup(opts));
        std::cerr << "Done reading mlib " << mlib.w2ldmap().size() << " " << mlib.w2hdmap().size() << std::endl;
        IntNumTns geomprtn;
        opt = findOption(opts, "-geomprtn");
        if (opt.empty()) return 0;
        std::istringstream giss;
        SAFE_FUNC_EVAL(SharedRead(opt, giss));
        SAFE_FUNC_EVAL(deserialize(geomprtn, giss, all));
        Wave3d wave("wave3d_");
        wave.posptr() = &pos;
        wave.kernel() = kernel;
        wave.mlibptr() = &mlib;
        wave.geomprtn() = geomprtn;

        // 2. setup
        auto t0 = std::chrono::steady_clock::now();
        SAFE_FUNC_EVAL(wave.setup(opts));
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "wave setup used " << std::chrono::duration<double>(t1 - t0).count() << "secs " << std::endl;

        // 3. eval
        t0 = std::chrono::steady_clock::now();
        SAFE_FUNC_EVAL(wave.eval(den, val));
        t1 = std::chrono::steady_clock::now();
        double time_eval = std::chrono::duration<double>(t1 - t0).count();
        std::cout << "wave eval used " << time_eval << "secs " << std::endl;
        std::ostringstream oss;
        SAFE_FUNC_EVAL(serialize(val, oss, all));
        opt = findOption(opts, "-valfile");
        if (opt.empty()) return 0;
        SAFE_FUNC_EVAL(SeparateWrite(opt, oss));

        // 4. check
        IntNumVec chkkeyvec;
        opt = findOption(opts, "-chkfile");
        if (opt.empty()) return 0;
        std::istringstream iss;
        SAFE_FUNC_EVAL(SharedRead(opt, iss));
        SAFE_FUNC_EVAL(deserialize(chkkeyvec, iss, all));
        int numchk = chkkeyvec.m();
        double relerr;
        clock_t ck0 = clock();
        SAFE_FUNC_EVAL(wave.check(den, val, chkkeyvec, relerr));
        clock_t ck1 = clock();

        // 5. output results
        double time_drct = CLOCK_DIFF_SECS(ck1, ck0) * numpts / double(numchk);
        printf("----------------------\n");
        printf("RESULT\n");
        printf("K  %.2e\n", wave.K());
        printf("Ta %.2e\n", time_eval);
        printf("Td %.2e\n", time_drct);
        printf("Rt %.2e\n", time_drct / time_eval);
        printf("Ea %.6e\n", relerr);
        printf("----------------------\n");
    } catch (std::exception* e) {
        std::cerr << "caught error: " << e->what() << std::endl;
        std::printf("error: %s\n", e->what());
        return 1;
    }
    return 0;
}
