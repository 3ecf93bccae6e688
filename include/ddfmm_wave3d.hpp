// Header-only C++ drop-in for the reference's Wave3d / Kernel3d / Mlib3d surface
// (include/wave3d.hpp:219-265, include/kernel3d.hpp:29-45, include/mlib3d.hpp:58-99)
// with the reference's signatures and containers (ddfmm_numeric.hpp), implemented over
// the C ABI of ddfmm.h.  A caller written against the reference keeps its code
// (tt.cpp:64-197):
//
//     ParVec<int, Point3, PtPrtn> pos;  deserialize(pos, piss, all);
//     ParVec<int, cpx, PtPrtn> den, val; ... val = den;
//     Kernel3d kernel(KERNEL_HELM);
//     Mlib3d mlib("mlib3d_");  mlib.kernel() = kernel;  mlib.NPQ() = 4;  mlib.setup(opts);
//     IntNumTns geomprtn;  deserialize(geomprtn, giss, all);
//     Wave3d wave("wave3d_");
//     wave.posptr() = &pos;  wave.kernel() = kernel;  wave.mlibptr() = &mlib;
//     wave.geomprtn() = geomprtn;
//     wave.setup(opts);  wave.eval(den, val);  wave.check(den, val, chkkeyvec, relerr);
//
// One process drives one GPU (the device of -wave3d_device, default 0); the cell
// partition geomprtn only fixes the cell level (its edge length), as all cells are this
// rank's.  Beyond the reference: -wave3d_precision (0 FP64, 1 mixed, 2 FP64 with fused
// solves, default 0), -wave3d_hbm_budget (GiB of device memory the context may hold; the
// LIMITED_MEMORY build's memory bound, see ddfmm_set_hbm_budget), eval_multi for several
// density vectors in one evaluation, and
// handle() for the C ABI.  Every eval is a fresh evaluation (the reference's is
// single-shot, wave3d_eval.cpp:518-521).  Errors throw a heap-allocated ddfmm::Error
// (the reference throws `new std::exception()`, commoninc.hpp:52-53).
#pragma once

#include <string>

#include "ddfmm.h"
#include "ddfmm_numeric.hpp"

namespace ddfmm {

enum { KERNEL_HELM = DDFMM_KERNEL_HELM, KERNEL_EXPR = DDFMM_KERNEL_EXPR };
enum { WAVE3D_PTS = 1, WAVE3D_TERMINAL = 2 };

inline void check_rc(int rc) {
    if (rc != 0) raise(ddfmm_last_error());
}

inline std::string find_opt(const std::map<std::string, std::string>& opts, const std::string& k) {
    auto it = opts.find(k);
    return it == opts.end() ? std::string() : it->second;
}

class Kernel3d {
public:
    Kernel3d(int t = KERNEL_HELM) : _type(t) {}
    int& type() { return _type; }
    int dim() const { return 3; }
    int sdof() const { return 1; }
    int tdof() const { return 1; }
    // mat(i, j) = G(trgpos(:, i), srcpos(:, j)), k = 2*pi; points are the columns of the
    // 3 x M / 3 x N matrices, mat is M x N column-major; srcnor is unused (kernel3d.cpp:26-107)
    int kernel(const DblNumMat& trgpos, const DblNumMat& srcpos, const DblNumMat& srcnor, CpxNumMat& mat) {
        (void)srcnor;
        if (trgpos.m() != 3 || srcpos.m() != 3) raise("Kernel3d::kernel: positions must be 3 x n");
        if (mat.m() != trgpos.n() || mat.n() != srcpos.n()) mat.resize(trgpos.n(), srcpos.n());
        if (trgpos.n() == 0 || srcpos.n() == 0) return 0;
        check_rc(ddfmm_kernel_matrix(_type, trgpos.data(), trgpos.n(), srcpos.data(), srcpos.n(),
                                     reinterpret_cast<double*>(mat.data())));
        return 0;
    }

private:
    int _type;
};

// The operator data of a width stays inside the library (device tables); the entries
// of w2ldmap / w2hdmap mark which widths the two files hold.
class LowFreqEntry {
public:
    DblNumMat _uep, _ucp;
    NumVec<CpxNumMat> _uc2ue;
    NumTns<CpxNumTns> _ue2dc;
    DblNumMat& uep() { return _uep; }
    DblNumMat& ucp() { return _ucp; }
    NumVec<CpxNumMat>& uc2ue() { return _uc2ue; }
    NumTns<CpxNumTns>& ue2dc() { return _ue2dc; }
};
class HghFreqDirEntry {
public:
    DblNumMat _uep, _ucp;
    NumVec<CpxNumMat> _uc2ue;
    DblNumMat& uep() { return _uep; }
    DblNumMat& ucp() { return _ucp; }
    NumVec<CpxNumMat>& uc2ue() { return _uc2ue; }
};

class Mlib3d {
public:
    Mlib3d(const std::string& prefix) : _prefix(prefix) {}
    ~Mlib3d() {
        if (_h) ddfmm_mlib_free(_h);
    }
    Mlib3d(const Mlib3d&) = delete;
    Mlib3d& operator=(const Mlib3d&) = delete;
    Kernel3d& kernel() { return _kernel; }
    int& NPQ() { return _NPQ; }
    std::string& ldname() { return _ldname; }
    std::string& hdname() { return _hdname; }
    std::map<double, LowFreqEntry>& w2ldmap() { return _w2ldmap; }
    std::map<double, std::map<Index3, HghFreqDirEntry>>& w2hdmap() { return _w2hdmap; }
    // Mlib3d::setup (mlib3d.cpp:61-103): option keys -<prefix>NPQ/ldname/hdname; the
    // files are read once into a library handle that every Wave3d set up from it shares
    int setup(std::map<std::string, std::string>& opts) {
        std::string v = find_opt(opts, "-" + _prefix + "NPQ");
        if (!v.empty()) _NPQ = std::stoi(v);
        v = find_opt(opts, "-" + _prefix + "ldname");
        if (!v.empty()) _ldname = v;
        v = find_opt(opts, "-" + _prefix + "hdname");
        if (!v.empty()) _hdname = v;
        if (_h) {
            ddfmm_mlib_free(_h);
            _h = nullptr;
        }
        check_rc(ddfmm_mlib_load(_ldname.c_str(), _hdname.c_str(), &_h));
        double lw[64], hw[64];
        int nl = 0, nh = 0;
        check_rc(ddfmm_mlib_widths(_h, 64, lw, &nl, hw, &nh));
        _w2ldmap.clear();
        _w2hdmap.clear();
        for (int i = 0; i < nl && i < 64; ++i) _w2ldmap[lw[i]];
        for (int i = 0; i < nh && i < 64; ++i) _w2hdmap[hw[i]];
        return 0;
    }
    const ddfmm_mlib* handle() const { return _h; }

private:
    std::string _prefix;
    Kernel3d _kernel;
    int _NPQ = 4;
    std::string _ldname, _hdname;
    ddfmm_mlib* _h = nullptr;
    std::map<double, LowFreqEntry> _w2ldmap;
    std::map<double, std::map<Index3, HghFreqDirEntry>> _w2hdmap;
};

class Wave3d {
public:
    Wave3d(const std::string& prefix) : _prefix(prefix) {}
    ~Wave3d() {
        if (_ctx) ddfmm_destroy(_ctx);
    }
    Wave3d(const Wave3d&) = delete;
    Wave3d& operator=(const Wave3d&) = delete;

    // member access (wave3d.hpp:246-255)
    ParVec<int, Point3, PtPrtn>*& posptr() { return _posptr; }
    Kernel3d& kernel() { return _kernel; }
    int& ACCU() { return _ACCU; }
    int& NPQ() { return _NPQ; }
    Mlib3d*& mlibptr() { return _mlibptr; }
    IntNumTns& geomprtn() { return _geomprtn; }
    double& K() { return _K; }
    Point3& ctr() { return _ctr; }
    int& ptsmax() { return _ptsmax; }
    int& maxlevel() { return _maxlevel; }
    int& nrhs_max() { return _nrhs_max; }
    int& precision() { return _precision; }
    int& device() { return _device; }
    double& hbm_budget() { return _hbm_budget; }

    // Wave3d::setup (wave3d_setup.cpp:21-97); option keys -<prefix>ACCU/NPQ/K/ctr/ptsmax/
    // maxlevel (+ precision, nrhs_max, device, hbm_budget)
    int setup(std::map<std::string, std::string>& opts) {
        auto get = [&](const char* k) { return find_opt(opts, "-" + _prefix + k); };
        std::string v;
        if (!(v = get("ACCU")).empty()) _ACCU = std::stoi(v);
        if (!(v = get("NPQ")).empty()) _NPQ = std::stoi(v);
        if (!(v = get("K")).empty()) _K = std::stod(v);
        if (!(v = get("ctr")).empty()) {
            std::istringstream ss(v);
            ss >> _ctr(0) >> _ctr(1) >> _ctr(2);
        }
        if (!(v = get("ptsmax")).empty()) _ptsmax = std::stoi(v);
        if (!(v = get("maxlevel")).empty()) _maxlevel = std::stoi(v);
        if (!(v = get("precision")).empty()) _precision = std::stoi(v);
        if (!(v = get("nrhs_max")).empty()) _nrhs_max = std::stoi(v);
        if (!(v = get("device")).empty()) _device = std::stoi(v);
        if (!(v = get("hbm_budget")).empty()) _hbm_budget = std::stod(v);
        if (!_posptr || !_mlibptr) raise("Wave3d::setup: posptr and mlibptr must be set");
        if (_maxlevel <= 0) raise("Wave3d::setup: maxlevel must be given");
        if (!_mlibptr->handle()) raise("Wave3d::setup: Mlib3d::setup has not run");
        _keys.clear();
        std::vector<double> xyz;
        for (auto& kv : _posptr->lclmap()) {
            _keys.push_back(kv.first);
            for (int d = 0; d < 3; ++d) xyz.push_back(kv.second(d));
        }
        ddfmm_params p{};
        p.K = _K;
        for (int d = 0; d < 3; ++d) p.ctr[d] = _ctr(d);
        p.ACCU = _ACCU;
        p.NPQ = _NPQ;
        p.ptsmax = _ptsmax;
        p.maxlevel = _maxlevel;
        p.numC = _geomprtn.m() > 0 ? _geomprtn.m() : 1;
        p.kernel = _kernel.type();
        p.nrhs_max = _nrhs_max;
        p.precision = _precision;
        if (_ctx) {
            ddfmm_destroy(_ctx);
            _ctx = nullptr;
        }
        check_rc(ddfmm_create(_device, &_ctx));
        check_rc(ddfmm_set_hbm_budget(_ctx, (int64_t)(_hbm_budget * 1073741824.0)));
        check_rc(ddfmm_setup_mlib(_ctx, xyz.data(), (int64_t)_keys.size(), &p, _mlibptr->handle()));
        return 0;
    }

    // Wave3d::eval (wave3d_eval.cpp:299-415): val must hold an entry per point (tt.cpp:97-98);
    // entries are assigned, not accumulated
    int eval(ParVec<int, cpx, PtPrtn>& den, ParVec<int, cpx, PtPrtn>& val) {
        std::vector<ParVec<int, cpx, PtPrtn>*> d{&den}, v{&val};
        return eval_multi(d, v);
    }

    // several density vectors in one evaluation (nrhs <= nrhs_max)
    int eval_multi(std::vector<ParVec<int, cpx, PtPrtn>*>& dens, std::vector<ParVec<int, cpx, PtPrtn>*>& vals) {
        if (!_ctx) raise("Wave3d::eval: setup has not run");
        if (dens.size() != vals.size() || dens.empty()) raise("Wave3d::eval: density / potential count mismatch");
        const size_t n = _keys.size(), nr = dens.size();
        std::vector<cpx> d(n * nr), v(n * nr);
        for (size_t r = 0; r < nr; ++r)
            for (size_t i = 0; i < n; ++i) d[r * n + i] = dens[r]->access(_keys[i]);
        check_rc(ddfmm_eval(_ctx, reinterpret_cast<const double*>(d.data()), reinterpret_cast<double*>(v.data()),
                            (int)nr));
        for (size_t r = 0; r < nr; ++r)
            for (size_t i = 0; i < n; ++i) vals[r]->access(_keys[i]) = v[r * n + i];
        return 0;
    }

    // Wave3d::check (wave3d_check.cpp:22-100): relerr = ||direct - fmm|| / ||fmm|| at the
    // keys of chkkeyvec, the direct sum in FP64 on the device
    int check(ParVec<int, cpx, PtPrtn>& den, ParVec<int, cpx, PtPrtn>& val, IntNumVec& chkkeyvec, double& relerr) {
        std::vector<double> src, trg;
        std::vector<cpx> d;
        for (auto& kv : _posptr->lclmap()) {
            for (int k = 0; k < 3; ++k) src.push_back(kv.second(k));
            d.push_back(den.access(kv.first));
        }
        const int nc = chkkeyvec.m();
        for (int i = 0; i < nc; ++i) {
            const Point3& p = _posptr->access(chkkeyvec(i));
            for (int k = 0; k < 3; ++k) trg.push_back(p(k));
        }
        std::vector<cpx> dir((size_t)nc);
        check_rc(ddfmm_direct_sum(_kernel.type(), trg.data(), nc, src.data(), (int)d.size(),
                                  reinterpret_cast<const double*>(d.data()), 1, reinterpret_cast<double*>(dir.data())));
        double en = 0, tn = 0;
        for (int i = 0; i < nc; ++i) {
            const cpx f = val.access(chkkeyvec(i));
            en += std::norm(dir[(size_t)i] - f);
            tn += std::norm(f);
        }
        relerr = std::sqrt(en / tn);
        return 0;
    }

    ddfmm_ctx* handle() { return _ctx; }

private:
    std::string _prefix;
    int _device = 0;
    ddfmm_ctx* _ctx = nullptr;
    ParVec<int, Point3, PtPrtn>* _posptr = nullptr;
    Kernel3d _kernel;
    Mlib3d* _mlibptr = nullptr;
    IntNumTns _geomprtn;
    // defaults of Wave3d::Wave3d (wave3d.cpp:24-26); maxlevel has none there
    int _ACCU = 1, _NPQ = 4, _ptsmax = 100, _maxlevel = 0, _nrhs_max = 1, _precision = DDFMM_FP64;
    double _K = 64;
    double _hbm_budget = 0;
    Point3 _ctr{0.0, 0.0, 0.0};
    std::vector<int> _keys;
};

}  // namespace ddfmm

#ifndef DDFMM_NO_GLOBAL_NAMES
// status checks of commoninc.hpp:52-53: a nonzero status or a false condition throws a
// heap-allocated exception
#ifndef SAFE_FUNC_EVAL
#define SAFE_FUNC_EVAL(fun)                                                              \
    {                                                                                     \
        int ierr_ = (fun);                                                                \
        if (ierr_ != 0) throw new ddfmm::Error(std::string("call failed: ") + #fun);      \
    }
#endif
#ifndef CHECK_TRUE
#define CHECK_TRUE(expr)                                                                  \
    {                                                                                     \
        if ((expr) == 0) throw new ddfmm::Error(std::string("check failed: ") + #expr);   \
    }
#endif
// the reference declares these at global scope
using ddfmm::cpx;
using ddfmm::CpxNumMat;
using ddfmm::CpxNumTns;
using ddfmm::CpxNumVec;
using ddfmm::DblNumMat;
using ddfmm::DblNumTns;
using ddfmm::DblNumVec;
using ddfmm::deserialize;
using ddfmm::HghFreqDirEntry;
using ddfmm::Index3;
using ddfmm::IntNumMat;
using ddfmm::IntNumTns;
using ddfmm::IntNumVec;
using ddfmm::Kernel3d;
using ddfmm::LowFreqEntry;
using ddfmm::Mlib3d;
using ddfmm::NumMat;
using ddfmm::NumTns;
using ddfmm::NumVec;
using ddfmm::ParVec;
using ddfmm::Point3;
using ddfmm::PtPrtn;
using ddfmm::SeparateRead;
using ddfmm::SeparateWrite;
using ddfmm::serialize;
using ddfmm::SharedRead;
using ddfmm::Vec3T;
using ddfmm::Wave3d;
using ddfmm::KERNEL_EXPR;
using ddfmm::KERNEL_HELM;
#endif
