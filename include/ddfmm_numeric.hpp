// Containers, keyed vectors and the binary file format of the reference's C++ surface,
// for callers that include this library's headers in place of the reference's:
//
//   cpx, Vec3T / Point3 / Index3               commoninc.hpp:47, vec3t.hpp:26-190
//   NumVec / NumMat / NumTns (+ Int/Dbl/Cpx)    numvec.hpp, nummat.hpp, numtns.hpp
//                                               (column-major: (i, j) at i + j*m,
//                                                (i, j, k) at i + j*m + k*m*n)
//   PtPrtn                                      wave3d.hpp:40-57 (owner by prefix array)
//   ParVec<Key, Data, Partition>                parvec.hpp:25-88, on one rank: every key
//                                               is local, the exchange calls have nothing
//                                               to move and return 0
//   serialize / deserialize                     serialize.hpp (native-endian binary:
//                                               int32 counts, then the elements)
//   SeparateRead / SeparateWrite / SharedRead   parallel.cpp:24-99, one rank:
//                                               data/<name>_0_1 and data/<name>
//
// Errors follow the reference's convention (commoninc.hpp:52-53): functions return an
// int status and failures throw a heap-allocated exception (here ddfmm::Error, a
// std::exception whose what() carries the message).
#pragma once

#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ddfmm {

using cpx = std::complex<double>;

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

[[noreturn]] inline void raise(const std::string& msg) { throw new Error(msg); }

// ---------------------------------------------------------------- Vec3T
template <class F>
class Vec3T {
public:
    F _v[3];
    Vec3T() : _v{F(0), F(0), F(0)} {}
    explicit Vec3T(F a) : _v{a, a, a} {}
    Vec3T(F a, F b, F c) : _v{a, b, c} {}
    explicit Vec3T(const F* p) : _v{p[0], p[1], p[2]} {}
    F& operator()(int i) { return _v[i]; }
    const F& operator()(int i) const { return _v[i]; }
    F& operator[](int i) { return _v[i]; }
    const F& operator[](int i) const { return _v[i]; }
    F* array() { return _v; }
    const F* array() const { return _v; }
    Vec3T& operator+=(const Vec3T& o) { for (int d = 0; d < 3; ++d) _v[d] += o._v[d]; return *this; }
    Vec3T& operator-=(const Vec3T& o) { for (int d = 0; d < 3; ++d) _v[d] -= o._v[d]; return *this; }
    Vec3T& operator*=(F s) { for (int d = 0; d < 3; ++d) _v[d] *= s; return *this; }
    Vec3T& operator/=(F s) { for (int d = 0; d < 3; ++d) _v[d] /= s; return *this; }
    F l1() const { F s = 0; for (int d = 0; d < 3; ++d) s += std::abs(_v[d]); return s; }
    F linfty() const { F s = 0; for (int d = 0; d < 3; ++d) s = std::max(s, F(std::abs(_v[d]))); return s; }
    double l2() const { double s = 0; for (int d = 0; d < 3; ++d) s += double(_v[d]) * double(_v[d]); return std::sqrt(s); }
};
// lexicographic order (the order of std::map keys, vec3t.hpp:98-107)
template <class F> bool operator<(const Vec3T<F>& a, const Vec3T<F>& b) {
    for (int d = 0; d < 3; ++d) {
        if (a(d) < b(d)) return true;
        if (b(d) < a(d)) return false;
    }
    return false;
}
template <class F> bool operator==(const Vec3T<F>& a, const Vec3T<F>& b) {
    return a(0) == b(0) && a(1) == b(1) && a(2) == b(2);
}
template <class F> bool operator!=(const Vec3T<F>& a, const Vec3T<F>& b) { return !(a == b); }
template <class F> bool operator>(const Vec3T<F>& a, const Vec3T<F>& b) { return b < a; }
template <class F> bool operator<=(const Vec3T<F>& a, const Vec3T<F>& b) { return !(b < a); }
template <class F> bool operator>=(const Vec3T<F>& a, const Vec3T<F>& b) { return !(a < b); }
template <class F> Vec3T<F> operator-(const Vec3T<F>& a) { return Vec3T<F>(-a(0), -a(1), -a(2)); }
template <class F> Vec3T<F> operator+(Vec3T<F> a, const Vec3T<F>& b) { return a += b; }
template <class F> Vec3T<F> operator-(Vec3T<F> a, const Vec3T<F>& b) { return a -= b; }
template <class F> Vec3T<F> operator*(F s, Vec3T<F> a) { return a *= s; }
template <class F> Vec3T<F> operator*(Vec3T<F> a, F s) { return a *= s; }
template <class F> Vec3T<F> operator/(Vec3T<F> a, F s) { return a /= s; }
template <class F> F operator*(const Vec3T<F>& a, const Vec3T<F>& b) { return a(0) * b(0) + a(1) * b(1) + a(2) * b(2); }
template <class F> F dot(const Vec3T<F>& a, const Vec3T<F>& b) { return a * b; }
template <class F> std::ostream& operator<<(std::ostream& os, const Vec3T<F>& a) {
    return os << a(0) << " " << a(1) << " " << a(2);
}
using Point3 = Vec3T<double>;
using Index3 = Vec3T<int>;

// ---------------------------------------------------------------- NumVec / NumMat / NumTns
// Dense arrays that either own their storage or view a caller's buffer (owndata false).
template <class F>
class NumStore {
protected:
    std::vector<F> _own;
    F* _data = nullptr;
    bool _owndata = true;
    void attach(int64_t n, bool owndata, F* data) {
        _owndata = owndata;
        if (owndata) {
            _own.assign((size_t)std::max<int64_t>(n, 0), F());
            if (data)
                for (int64_t i = 0; i < n; ++i) _own[(size_t)i] = data[i];
            _data = _own.empty() ? nullptr : _own.data();
        } else {
            _own.clear();
            _data = data;
        }
    }
    void copy_from(const NumStore& o, int64_t n) {
        if (o._owndata) attach(n, true, o._data);
        else attach(n, false, o._data);
    }
};

template <class F>
class NumVec : public NumStore<F> {
public:
    explicit NumVec(int m = 0) : _m(m) { this->attach(m, true, nullptr); }
    NumVec(int m, bool owndata, F* data) : _m(m) { this->attach(m, owndata, data); }
    NumVec(const NumVec& o) : NumStore<F>(), _m(o._m) { this->copy_from(o, o._m); }
    NumVec& operator=(const NumVec& o) {
        if (this != &o) { _m = o._m; this->copy_from(o, o._m); }
        return *this;
    }
    void resize(int m) {
        if (!this->_owndata) raise("NumVec::resize on a view");
        if (m != _m) { _m = m; this->attach(m, true, nullptr); }
    }
    F& operator()(int i) { assert(i >= 0 && i < _m); return this->_data[i]; }
    const F& operator()(int i) const { assert(i >= 0 && i < _m); return this->_data[i]; }
    F* data() const { return this->_data; }
    int m() const { return _m; }

private:
    int _m;
};

template <class F>
class NumMat : public NumStore<F> {
public:
    NumMat(int m = 0, int n = 0) : _m(m), _n(n) { this->attach((int64_t)m * n, true, nullptr); }
    NumMat(int m, int n, bool owndata, F* data) : _m(m), _n(n) { this->attach((int64_t)m * n, owndata, data); }
    NumMat(const NumMat& o) : NumStore<F>(), _m(o._m), _n(o._n) { this->copy_from(o, (int64_t)o._m * o._n); }
    NumMat& operator=(const NumMat& o) {
        if (this != &o) { _m = o._m; _n = o._n; this->copy_from(o, (int64_t)_m * _n); }
        return *this;
    }
    void resize(int m, int n) {
        if (!this->_owndata) raise("NumMat::resize on a view");
        if (m != _m || n != _n) { _m = m; _n = n; this->attach((int64_t)m * n, true, nullptr); }
    }
    F& operator()(int i, int j) { assert(i >= 0 && i < _m && j >= 0 && j < _n); return this->_data[i + (int64_t)j * _m]; }
    const F& operator()(int i, int j) const {
        assert(i >= 0 && i < _m && j >= 0 && j < _n);
        return this->_data[i + (int64_t)j * _m];
    }
    F* clmdata(int j) { return this->_data + (int64_t)j * _m; }
    F* data() const { return this->_data; }
    int m() const { return _m; }
    int n() const { return _n; }

private:
    int _m, _n;
};

template <class F>
class NumTns : public NumStore<F> {
public:
    NumTns(int m = 0, int n = 0, int p = 0) : _m(m), _n(n), _p(p) { this->attach((int64_t)m * n * p, true, nullptr); }
    NumTns(int m, int n, int p, bool owndata, F* data) : _m(m), _n(n), _p(p) {
        this->attach((int64_t)m * n * p, owndata, data);
    }
    NumTns(const NumTns& o) : NumStore<F>(), _m(o._m), _n(o._n), _p(o._p) {
        this->copy_from(o, (int64_t)o._m * o._n * o._p);
    }
    NumTns& operator=(const NumTns& o) {
        if (this != &o) { _m = o._m; _n = o._n; _p = o._p; this->copy_from(o, (int64_t)_m * _n * _p); }
        return *this;
    }
    void resize(int m, int n, int p) {
        if (!this->_owndata) raise("NumTns::resize on a view");
        if (m != _m || n != _n || p != _p) { _m = m; _n = n; _p = p; this->attach((int64_t)m * n * p, true, nullptr); }
    }
    F& operator()(int i, int j, int k) { return this->_data[i + (int64_t)_m * (j + (int64_t)_n * k)]; }
    const F& operator()(int i, int j, int k) const { return this->_data[i + (int64_t)_m * (j + (int64_t)_n * k)]; }
    F* data() const { return this->_data; }
    int m() const { return _m; }
    int n() const { return _n; }
    int p() const { return _p; }

private:
    int _m, _n, _p;
};

template <class F> void setvalue(NumVec<F>& v, F x) { for (int i = 0; i < v.m(); ++i) v(i) = x; }
template <class F> void setvalue(NumMat<F>& a, F x) {
    for (int j = 0; j < a.n(); ++j) for (int i = 0; i < a.m(); ++i) a(i, j) = x;
}
template <class F> void setvalue(NumTns<F>& t, F x) {
    for (int k = 0; k < t.p(); ++k) for (int j = 0; j < t.n(); ++j) for (int i = 0; i < t.m(); ++i) t(i, j, k) = x;
}
template <class F> double energy(const NumVec<F>& v) {
    double s = 0;
    for (int i = 0; i < v.m(); ++i) s += std::norm(cpx(v(i)));
    return s;
}
template <class F> double energy(const NumMat<F>& a) {
    double s = 0;
    for (int j = 0; j < a.n(); ++j) for (int i = 0; i < a.m(); ++i) s += std::norm(cpx(a(i, j)));
    return s;
}
template <class F> std::ostream& operator<<(std::ostream& os, const NumVec<F>& v) {
    os << v.m() << std::endl;
    for (int i = 0; i < v.m(); ++i) os << " " << v(i) << std::endl;
    return os;
}
template <class F> std::ostream& operator<<(std::ostream& os, const NumMat<F>& a) {
    os << a.m() << " " << a.n() << std::endl;
    for (int i = 0; i < a.m(); ++i) {
        for (int j = 0; j < a.n(); ++j) os << " " << a(i, j);
        os << std::endl;
    }
    return os;
}
template <class F> std::ostream& operator<<(std::ostream& os, const NumTns<F>& t) {
    os << t.m() << " " << t.n() << " " << t.p() << std::endl;
    for (int i = 0; i < t.m(); ++i) {
        for (int j = 0; j < t.n(); ++j) {
            for (int k = 0; k < t.p(); ++k) os << " " << t(i, j, k);
            os << std::endl;
        }
        os << std::endl;
    }
    return os;
}

using IntNumVec = NumVec<int>;
using DblNumVec = NumVec<double>;
using CpxNumVec = NumVec<cpx>;
using IntNumMat = NumMat<int>;
using DblNumMat = NumMat<double>;
using CpxNumMat = NumMat<cpx>;
using IntNumTns = NumTns<int>;
using DblNumTns = NumTns<double>;
using CpxNumTns = NumTns<cpx>;

// ---------------------------------------------------------------- PtPrtn / ParVec
class PtPrtn {
public:
    std::vector<int> _ownerinfo;     // prefix array of point ids per rank; last entry = N
    std::vector<int>& ownerinfo() { return _ownerinfo; }
    int owner(int key) const {
        if (_ownerinfo.empty() || key >= _ownerinfo.back()) raise("PtPrtn::owner: key out of range");
        int r = 0;
        while (r + 1 < (int)_ownerinfo.size() && _ownerinfo[r + 1] <= key) ++r;
        return r;
    }
};

template <class Key, class Data, class Partition>
class ParVec {
public:
    std::map<Key, Data> _lclmap;
    Partition _prtn;
    std::map<Key, Data>& lclmap() { return _lclmap; }
    Partition& prtn() { return _prtn; }
    int insert(Key key, Data& dat) {
        _lclmap[key] = dat;
        return 0;
    }
    Data& access(Key key) {
        auto it = _lclmap.find(key);
        if (it == _lclmap.end()) raise("ParVec::access: key not present");
        return it->second;
    }
    // (found, entry); when absent the reference is to a placeholder owned by the vector
    std::pair<bool, Data&> contains(Key key) {
        auto it = _lclmap.find(key);
        if (it != _lclmap.end()) return std::pair<bool, Data&>(true, it->second);
        _absent = Data();
        return std::pair<bool, Data&>(false, _absent);
    }
    // exchanges between ranks: with one rank every key is local (parvec.hpp:44-65)
    int getBegin(int (*)(Key, Data&, std::vector<int>&), const std::vector<int>&) { return 0; }
    int getBegin(std::vector<Key>&, const std::vector<int>&) { return 0; }
    int getEnd(const std::vector<int>&) { return 0; }
    int putBegin(std::vector<Key>&, const std::vector<int>&) { return 0; }
    int putEnd(const std::vector<int>&) { return 0; }
    int expand(std::vector<Key>&) { return 0; }
    int discard(std::vector<Key>&) { return 0; }
    int kbytes_received() { return 0; }
    int kbytes_sent() { return 0; }

private:
    Data _absent{};
};

// ---------------------------------------------------------------- serialize / deserialize
namespace io {
template <class T> void put(std::ostream& os, const T& v) { os.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <class T> void get(std::istream& is, T& v) {
    is.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!is) raise("deserialize: unexpected end of data");
}
}  // namespace io

using Mask = std::vector<int>;
inline int serialize(const int& v, std::ostream& os, const Mask&) { io::put(os, v); return 0; }
inline int deserialize(int& v, std::istream& is, const Mask&) { io::get(is, v); return 0; }
inline int serialize(const double& v, std::ostream& os, const Mask&) { io::put(os, v); return 0; }
inline int deserialize(double& v, std::istream& is, const Mask&) { io::get(is, v); return 0; }
inline int serialize(const cpx& v, std::ostream& os, const Mask&) { io::put(os, v); return 0; }
inline int deserialize(cpx& v, std::istream& is, const Mask&) { io::get(is, v); return 0; }
template <class F> int serialize(const Vec3T<F>& v, std::ostream& os, const Mask&) {
    for (int d = 0; d < 3; ++d) io::put(os, v(d));
    return 0;
}
template <class F> int deserialize(Vec3T<F>& v, std::istream& is, const Mask&) {
    for (int d = 0; d < 3; ++d) io::get(is, v(d));
    return 0;
}
template <class T> int serialize(const std::vector<T>& v, std::ostream& os, const Mask& m);
template <class T> int deserialize(std::vector<T>& v, std::istream& is, const Mask& m);
template <class T, class S> int serialize(const std::map<T, S>& v, std::ostream& os, const Mask& m);
template <class T, class S> int deserialize(std::map<T, S>& v, std::istream& is, const Mask& m);
template <class T, class S> int serialize(const std::pair<T, S>& v, std::ostream& os, const Mask& m) {
    serialize(v.first, os, m);
    serialize(v.second, os, m);
    return 0;
}
template <class T, class S> int deserialize(std::pair<T, S>& v, std::istream& is, const Mask& m) {
    deserialize(v.first, is, m);
    deserialize(v.second, is, m);
    return 0;
}
template <class T> int serialize(const std::vector<T>& v, std::ostream& os, const Mask& m) {
    io::put(os, (int)v.size());
    for (const T& x : v) serialize(x, os, m);
    return 0;
}
template <class T> int deserialize(std::vector<T>& v, std::istream& is, const Mask& m) {
    int n = 0;
    io::get(is, n);
    v.resize((size_t)n);
    for (T& x : v) deserialize(x, is, m);
    return 0;
}
template <class T> int serialize(const std::set<T>& v, std::ostream& os, const Mask& m) {
    io::put(os, (int)v.size());
    for (const T& x : v) serialize(x, os, m);
    return 0;
}
template <class T> int deserialize(std::set<T>& v, std::istream& is, const Mask& m) {
    int n = 0;
    io::get(is, n);
    v.clear();
    for (int k = 0; k < n; ++k) {
        T x;
        deserialize(x, is, m);
        v.insert(x);
    }
    return 0;
}
template <class T, class S> int serialize(const std::map<T, S>& v, std::ostream& os, const Mask& m) {
    io::put(os, (int)v.size());
    for (const auto& kv : v) {
        serialize(kv.first, os, m);
        serialize(kv.second, os, m);
    }
    return 0;
}
template <class T, class S> int deserialize(std::map<T, S>& v, std::istream& is, const Mask& m) {
    int n = 0;
    io::get(is, n);
    v.clear();
    for (int k = 0; k < n; ++k) {
        T key;
        S val;
        deserialize(key, is, m);
        deserialize(val, is, m);
        v[key] = val;
    }
    return 0;
}
// dense arrays: their extents, then the elements in storage order
template <class F> int serialize(const NumVec<F>& v, std::ostream& os, const Mask&) {
    io::put(os, v.m());
    os.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)sizeof(F) * v.m());
    return 0;
}
template <class F> int deserialize(NumVec<F>& v, std::istream& is, const Mask&) {
    int m = 0;
    io::get(is, m);
    v.resize(m);
    is.read(reinterpret_cast<char*>(v.data()), (std::streamsize)sizeof(F) * m);
    if (!is) raise("deserialize: unexpected end of data");
    return 0;
}
template <class F> int serialize(const NumMat<F>& a, std::ostream& os, const Mask&) {
    io::put(os, a.m());
    io::put(os, a.n());
    os.write(reinterpret_cast<const char*>(a.data()), (std::streamsize)sizeof(F) * a.m() * a.n());
    return 0;
}
template <class F> int deserialize(NumMat<F>& a, std::istream& is, const Mask&) {
    int m = 0, n = 0;
    io::get(is, m);
    io::get(is, n);
    a.resize(m, n);
    is.read(reinterpret_cast<char*>(a.data()), (std::streamsize)sizeof(F) * m * n);
    if (!is) raise("deserialize: unexpected end of data");
    return 0;
}
template <class F> int serialize(const NumTns<F>& t, std::ostream& os, const Mask&) {
    io::put(os, t.m());
    io::put(os, t.n());
    io::put(os, t.p());
    os.write(reinterpret_cast<const char*>(t.data()), (std::streamsize)sizeof(F) * t.m() * t.n() * t.p());
    return 0;
}
template <class F> int deserialize(NumTns<F>& t, std::istream& is, const Mask&) {
    int m = 0, n = 0, p = 0;
    io::get(is, m);
    io::get(is, n);
    io::get(is, p);
    t.resize(m, n, p);
    is.read(reinterpret_cast<char*>(t.data()), (std::streamsize)sizeof(F) * m * n * p);
    if (!is) raise("deserialize: unexpected end of data");
    return 0;
}
inline int serialize(const PtPrtn& p, std::ostream& os, const Mask& m) { return serialize(p._ownerinfo, os, m); }
inline int deserialize(PtPrtn& p, std::istream& is, const Mask& m) { return deserialize(p._ownerinfo, is, m); }
template <class K, class D, class P> int serialize(const ParVec<K, D, P>& v, std::ostream& os, const Mask& m) {
    serialize(v._lclmap, os, m);
    serialize(v._prtn, os, m);
    return 0;
}
template <class K, class D, class P> int deserialize(ParVec<K, D, P>& v, std::istream& is, const Mask& m) {
    deserialize(v._lclmap, is, m);
    deserialize(v._prtn, is, m);
    return 0;
}

// ---------------------------------------------------------------- data files (one rank)
inline std::string& data_root() {
    static std::string root = "data";
    return root;
}
inline int read_whole(const std::string& path, std::istringstream& is) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::cerr << "failed to open input file stream: " << path << std::endl;
        return -1;
    }
    std::ostringstream ss;
    ss << f.rdbuf();
    is.str(ss.str());
    return 0;
}
inline int SeparateRead(std::string name, std::istringstream& is) { return read_whole(data_root() + "/" + name + "_0_1", is); }
inline int SharedRead(std::string name, std::istringstream& is) { return read_whole(data_root() + "/" + name, is); }
inline int SeparateWrite(std::string name, std::ostringstream& os) {
    const std::string path = data_root() + "/" + name + "_0_1";
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        std::cerr << "failed to open output file stream (" << path << ")" << std::endl;
        return -1;
    }
    const std::string s = os.str();
    f.write(s.data(), (std::streamsize)s.size());
    return f ? 0 : -1;
}

}  // namespace ddfmm
