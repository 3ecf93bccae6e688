// Header-only C++ drop-in for the reference's Wave3d / Kernel3d / Mlib3d surface
// (include/wave3d.hpp:219-265, include/kernel3d.hpp:29-45, include/mlib3d.hpp:58-99),
// implemented over the C ABI of ddfmm.h.  A caller written against the reference keeps
// its wiring (tt.cpp:103-151):
//
//     ddfmm::Kernel3d knl(ddfmm::KERNEL_HELM);
//     ddfmm::Mlib3d mlib("mlib3d_");  mlib.kernel() = knl;  mlib.setup(opts);
//     ddfmm::Wave3d wave("wave3d_");
//     wave.posptr() = &pos;  wave.kernel() = knl;  wave.mlibptr() = &mlib;
//     wave.geomprtn() = numC;            // cells per edge (one device owns all cells)
//     wave.setup(opts);
//     wave.eval(den, val);
//     wave.check(den, val, chkkeys, relerr);
//
// Containers: positions are a map point-id -> Point3 and densities/potentials a map
// point-id -> cpx (the single-rank content of ParVec<int,Point3,PtPrtn> and
// ParVec<int,cpx,PtPrtn>, parvec.hpp:25-88).  Errors throw ddfmm::Error carrying the
// library's message (the reference throws `new std::exception()`, commoninc.hpp:52-53).
#pragma once

#include <array>
#include <complex>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ddfmm.h"

namespace ddfmm {

using cpx = std::complex<double>;
using Point3 = std::array<double, 3>;
enum { KERNEL_HELM = DDFMM_KERNEL_HELM, KERNEL_EXPR = DDFMM_KERNEL_EXPR };

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

inline void check_rc(int rc) {
    if (rc != 0) throw Error(ddfmm_last_error());
}

class Kernel3d {
public:
    explicit Kernel3d(int type = KERNEL_HELM) : _type(type) {}
    int& type() { return _type; }
    // trg: M points, src: N points; out: M x N column-major (Kernel3d::kernel, kernel3d.cpp:26-107)
    int kernel(const std::vector<Point3>& trg, const std::vector<Point3>& src, std::vector<cpx>& out) const {
        out.assign(trg.size() * src.size(), cpx(0, 0));
        check_rc(ddfmm_kernel_matrix(_type, trg.empty() ? nullptr : trg[0].data(), (int)trg.size(),
                                     src.empty() ? nullptr : src[0].data(), (int)src.size(),
                                     reinterpret_cast<double*>(out.data())));
        return 0;
    }

private:
    int _type;
};

inline std::string find_opt(const std::map<std::string, std::string>& opts, const std::string& k) {
    auto it = opts.find(k);
    return it == opts.end() ? std::string() : it->second;
}

class Mlib3d {
public:
    explicit Mlib3d(const std::string& prefix) : _prefix(prefix) {}
    Kernel3d& kernel() { return _kernel; }
    int& NPQ() { return _NPQ; }
    std::string& ldname() { return _ldname; }
    std::string& hdname() { return _hdname; }
    // Mlib3d::setup (mlib3d.cpp:61-103): option keys -<prefix>NPQ/ldname/hdname
    int setup(const std::map<std::string, std::string>& opts) {
        std::string v = find_opt(opts, "-" + _prefix + "NPQ");
        if (!v.empty()) _NPQ = std::stoi(v);
        v = find_opt(opts, "-" + _prefix + "ldname");
        if (!v.empty()) _ldname = v;
        v = find_opt(opts, "-" + _prefix + "hdname");
        if (!v.empty()) _hdname = v;
        return 0;
    }

private:
    std::string _prefix;
    Kernel3d _kernel;
    int _NPQ = 4;
    std::string _ldname, _hdname;
};

class Wave3d {
public:
    explicit Wave3d(const std::string& prefix, int device = 0) : _prefix(prefix), _device(device) {}
    ~Wave3d() {
        if (_ctx) ddfmm_destroy(_ctx);
    }
    Wave3d(const Wave3d&) = delete;
    Wave3d& operator=(const Wave3d&) = delete;

    std::map<int, Point3>*& posptr() { return _posptr; }
    Kernel3d& kernel() { return _kernel; }
    int& ACCU() { return _ACCU; }
    int& NPQ() { return _NPQ; }
    Mlib3d*& mlibptr() { return _mlibptr; }
    int& geomprtn() { return _numC; }
    double& K() { return _K; }
    Point3& ctr() { return _ctr; }
    int& ptsmax() { return _ptsmax; }
    int& maxlevel() { return _maxlevel; }
    int& nrhs_max() { return _nrhs_max; }
    int& precision() { return _precision; }

    // Wave3d::setup (wave3d_setup.cpp:21-97); option keys -<prefix>ACCU/NPQ/K/ctr/ptsmax/maxlevel
    int setup(const std::map<std::string, std::string>& opts) {
        auto get = [&](const char* k) { return find_opt(opts, "-" + _prefix + k); };
        std::string v;
        if (!(v = get("ACCU")).empty()) _ACCU = std::stoi(v);
        if (!(v = get("NPQ")).empty()) _NPQ = std::stoi(v);
        if (!(v = get("K")).empty()) _K = std::stod(v);
        if (!(v = get("ctr")).empty()) { std::istringstream ss(v); ss >> _ctr[0] >> _ctr[1] >> _ctr[2]; }
        if (!(v = get("ptsmax")).empty()) _ptsmax = std::stoi(v);
        if (!(v = get("maxlevel")).empty()) _maxlevel = std::stoi(v);
        if (!_posptr || !_mlibptr) throw Error("posptr and mlibptr must be set before setup");
        if (_maxlevel <= 0) throw Error("maxlevel must be given");
        _keys.clear();
        std::vector<double> xyz;
        for (auto& kv : *_posptr) {
            _keys.push_back(kv.first);
            xyz.insert(xyz.end(), kv.second.begin(), kv.second.end());
        }
        ddfmm_params p{};
        p.K = _K;
        for (int d = 0; d < 3; ++d) p.ctr[d] = _ctr[d];
        p.ACCU = _ACCU; p.NPQ = _NPQ; p.ptsmax = _ptsmax; p.maxlevel = _maxlevel; p.numC = _numC;
        p.kernel = _kernel.type(); p.nrhs_max = _nrhs_max; p.precision = _precision;
        if (_ctx) { ddfmm_destroy(_ctx); _ctx = nullptr; }
        check_rc(ddfmm_create(_device, &_ctx));
        check_rc(ddfmm_setup(_ctx, xyz.data(), (int64_t)_keys.size(), &p, _mlibptr->ldname().c_str(),
                             _mlibptr->hdname().c_str()));
        return 0;
    }

    // Wave3d::eval (wave3d_eval.cpp:299-415): val must hold an entry per point (tt.cpp:97-98);
    // entries are assigned, not accumulated.
    int eval(const std::map<int, cpx>& den, std::map<int, cpx>& val) {
        std::vector<cpx> d(_keys.size()), v(_keys.size());
        for (size_t i = 0; i < _keys.size(); ++i) {
            auto it = den.find(_keys[i]);
            if (it == den.end()) throw Error("density missing for point " + std::to_string(_keys[i]));
            d[i] = it->second;
        }
        check_rc(ddfmm_eval(_ctx, reinterpret_cast<const double*>(d.data()), reinterpret_cast<double*>(v.data()), 1));
        for (size_t i = 0; i < _keys.size(); ++i) val[_keys[i]] = v[i];
        return 0;
    }

    // Wave3d::check (wave3d_check.cpp:22-100): relerr = ||direct - fmm|| / ||fmm|| at chkkeys
    int check(const std::map<int, cpx>& den, const std::map<int, cpx>& val, const std::vector<int>& chkkeys,
              double& relerr) {
        std::vector<double> src, trg;
        std::vector<cpx> d;
        for (auto& kv : *_posptr) {
            src.insert(src.end(), kv.second.begin(), kv.second.end());
            d.push_back(den.at(kv.first));
        }
        for (int k : chkkeys) {
            const Point3& p = _posptr->at(k);
            trg.insert(trg.end(), p.begin(), p.end());
        }
        std::vector<cpx> dir(chkkeys.size());
        check_rc(ddfmm_direct_sum(_kernel.type(), trg.data(), (int)chkkeys.size(), src.data(), (int)d.size(),
                                  reinterpret_cast<const double*>(d.data()), 1, reinterpret_cast<double*>(dir.data())));
        double en = 0, tn = 0;
        for (size_t i = 0; i < chkkeys.size(); ++i) {
            cpx f = val.at(chkkeys[i]);
            en += std::norm(dir[i] - f);
            tn += std::norm(f);
        }
        relerr = std::sqrt(en / tn);
        return 0;
    }

    ddfmm_ctx* handle() { return _ctx; }

private:
    std::string _prefix;
    int _device;
    ddfmm_ctx* _ctx = nullptr;
    std::map<int, Point3>* _posptr = nullptr;
    Kernel3d _kernel;
    Mlib3d* _mlibptr = nullptr;
    int _ACCU = 1, _NPQ = 4, _ptsmax = 100, _maxlevel = 0, _numC = 1, _nrhs_max = 1, _precision = DDFMM_FP64;
    double _K = 64;
    Point3 _ctr{{0, 0, 0}};
    std::vector<int> _keys;
};

}  // namespace ddfmm
