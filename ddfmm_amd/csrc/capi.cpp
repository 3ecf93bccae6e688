// C ABI (include/ddfmm.h): context lifetime, setup (host planning + device upload +
// translation-matrix generation) and evaluation (stage sequence on one HIP stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ddfmm.h"
#include "device.hpp"
#include "planner.hpp"

using namespace ddfmm;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct HipFail : std::runtime_error {
    explicit HipFail(const std::string& s) : std::runtime_error(s) {}
};

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw HipFail(std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)
// Parsed operator libraries shared by every context that uses them: a handle holds a
// reference, and ddfmm_setup reuses a library that a live handle (or a concurrent setup)
// still holds instead of reading the files again.
std::mutex g_mlib_mu;
std::map<std::pair<std::string, std::string>, std::weak_ptr<const Mlib>> g_mlib_cache;

std::shared_ptr<const Mlib> load_mlib(const std::string& ld, const std::string& hd) {
    std::lock_guard<std::mutex> lock(g_mlib_mu);
    auto key = std::make_pair(ld, hd);
    auto it = g_mlib_cache.find(key);
    if (it != g_mlib_cache.end())
        if (auto sp = it->second.lock()) return sp;
    auto m = std::make_shared<Mlib>();
    m->load(ld, hd);
    std::shared_ptr<const Mlib> c = m;
    g_mlib_cache[key] = c;
    return c;
}
}  // namespace

struct ddfmm_mlib {
    std::shared_ptr<const Mlib> lib;
};

struct ddfmm_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;                 // second stream of an evaluation (near field, LF lists)
    hipEvent_t sync_ev[STAGE_EVENTS] = {};     // fork/join events between the two streams
    hipEvent_t done_ev = nullptr;              // end of the last evaluation (on its stream)
    bool done_recorded = false;
    bool use_graph = false;                    // replay a captured graph of the stage sequence
    std::map<int, hipGraphExec_t> graphs;      // one per right-hand-side count
    size_t tiles_bytes = 0;                    // device bytes of the GEMM tile table
    Params prm;
    int nrhs_max = 1, precision = 0;
    int64_t hbm_budget = 0;                    // bytes; 0 = no budget
    int slab = 1;                              // right-hand sides per pass (work buffers sized for it)
    int64_t fixed_bytes = 0, rhs_bytes = 0, setup_transient = 0;
    HostPlan hp;
    DevPlan dp;
    std::vector<void*> allocs;
    int64_t device_bytes = 0;
    bool ready = false;
    bool timing = false;
    bool serial = false;                       // timing with every stage on one stream
    std::vector<hipEvent_t> ev;
    std::vector<double> stage_ms;

    template <class T>
    T* alloc(int64_t n) {
        if (n <= 0) return nullptr;
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, (size_t)n * sizeof(T)));
        allocs.push_back(p);
        device_bytes += n * (int64_t)sizeof(T);
        return (T*)p;
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>((int64_t)v.size());
        if (p) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    void release() {
        drop_graphs();
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
        device_bytes = 0;
        ready = false;
    }
};

// Device bytes of a set-up context: `fixed` for the point, operator and descriptor
// tables (kept for the context's life), `per_rhs` for the work buffers and the staging
// of one right-hand side, `transient` for the generator tables that live only during
// setup (freed before the work buffers are allocated).
template <class T>
static int64_t vbytes(const std::vector<T>& v) { return (int64_t)v.size() * (int64_t)sizeof(T); }
static void plan_bytes(const HostPlan& h, int precision, int64_t& fixed, int64_t& per_rhs, int64_t& transient) {
    const bool mixed = precision == DDFMM_MIXED || precision == DDFMM_MIXED_SPLIT;
    fixed = vbytes(h.px) + vbytes(h.py) + vbytes(h.pz) + 2 * vbytes(h.perm) + vbytes(h.grid) +
            h.opmat_total * (int64_t)sizeof(double2) + (mixed ? h.opmat_total * (int64_t)sizeof(float2) : 0) +
            (precision == DDFMM_MIXED_SPLIT ? h.opmat_total * 16 : 0) +
            vbytes(h.ktasks) + vbytes(h.ksrcs) + vbytes(h.groups) + vbytes(h.tiles) + vbytes(h.colB) +
            vbytes(h.colC) + vbytes(h.colA) + vbytes(h.fsrc) + vbytes(h.vtrg) + vbytes(h.vchunk) +
            vbytes(h.vsrc) + vbytes(h.vpk) + vbytes(h.rtrg) + (int64_t)sizeof(IoPtrs);
    const int64_t esz = mixed ? (int64_t)sizeof(float2) : (int64_t)sizeof(double2);
    per_rhs = 2 * h.N * (int64_t)sizeof(double2);
    for (int b = 0; b < BUF_COUNT; ++b) per_rhs += h.bufsize[b] * esz;
    transient = vbytes(h.genpts) + vbytes(h.kgen);
}

// The kernel-apply flavour is fixed per context (each right-hand side then has the same
// bits alone and in any batch): the matrix-core kernel in the mixed modes from four
// right-hand sides up, the vector-ALU kernel otherwise.
static bool kapply_on_mfma(int precision, int nrhs_max) {
    return (precision == DDFMM_MIXED || precision == DDFMM_MIXED_SPLIT) && nrhs_max >= 4;
}

static std::string gib(int64_t b) {
    char t[32];
    std::snprintf(t, sizeof(t), "%.3f GiB", (double)b / (1 << 30));
    return t;
}

extern "C" {

const char* ddfmm_last_error(void) { return g_err.c_str(); }

int ddfmm_create(int device, ddfmm_ctx** out) {
    if (!out) return fail(DDFMM_EINVAL, "null output pointer");
    auto* c = new ddfmm_ctx();
    c->device = device;
    if (device >= 0) {
        try {
            HIPCHK(hipSetDevice(device));
            HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            HIPCHK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
            for (auto& e : c->sync_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming));
        } catch (const std::exception& e) {
            delete c;
            return fail(DDFMM_EDEVICE, e.what());
        }
    }
    *out = c;
    return DDFMM_OK;
}

int ddfmm_destroy(ddfmm_ctx* c) {
    if (!c) return DDFMM_OK;
    if (c->device >= 0) {
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->aux) (void)hipStreamSynchronize(c->aux);
        c->release();
        for (auto e : c->sync_ev)
            if (e) (void)hipEventDestroy(e);
        if (c->done_ev) (void)hipEventDestroy(c->done_ev);
        if (c->aux) (void)hipStreamDestroy(c->aux);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return DDFMM_OK;
}

static int setup_common(ddfmm_ctx* c, const double* pos, int64_t n, const ddfmm_params* p,
                        std::shared_ptr<const Mlib> shared, const char* ldfile, const char* hdfile);

int ddfmm_setup(ddfmm_ctx* c, const double* pos, int64_t n, const ddfmm_params* p,
                const char* ldfile, const char* hdfile) {
    if (!c || !pos || !p || !ldfile || !hdfile) return fail(DDFMM_EINVAL, "null argument");
    return setup_common(c, pos, n, p, nullptr, ldfile, hdfile);
}

int ddfmm_setup_mlib(ddfmm_ctx* c, const double* pos, int64_t n, const ddfmm_params* p, const ddfmm_mlib* m) {
    if (!c || !pos || !p || !m || !m->lib) return fail(DDFMM_EINVAL, "null argument");
    return setup_common(c, pos, n, p, m->lib, nullptr, nullptr);
}

int ddfmm_mlib_load(const char* ldfile, const char* hdfile, ddfmm_mlib** out) {
    if (!ldfile || !hdfile || !out) return fail(DDFMM_EINVAL, "null argument");
    try {
        auto* m = new ddfmm_mlib();
        try {
            m->lib = load_mlib(ldfile, hdfile);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    } catch (const std::exception& e) {
        return fail(DDFMM_ESETUP, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_mlib_free(ddfmm_mlib* m) {
    delete m;
    return DDFMM_OK;
}

int ddfmm_mlib_widths(const ddfmm_mlib* m, int cap, double* ld_widths, int* nld, double* hd_widths, int* nhd) {
    if (!m || !m->lib || !nld || !nhd) return fail(DDFMM_EINVAL, "null argument");
    *nld = (int)m->lib->ld.size();
    *nhd = (int)m->lib->hd.size();
    int i = 0;
    for (auto& kv : m->lib->ld)
        if (ld_widths && i < cap) ld_widths[i++] = kv.first;
    i = 0;
    for (auto& kv : m->lib->hd)
        if (hd_widths && i < cap) hd_widths[i++] = kv.first;
    return DDFMM_OK;
}

static int setup_common(ddfmm_ctx* c, const double* pos, int64_t n, const ddfmm_params* p,
                        std::shared_ptr<const Mlib> shared, const char* ldfile, const char* hdfile) {
    if (n <= 0 || n > (int64_t)INT32_MAX) return fail(DDFMM_EINVAL, "point count out of range");
    if (p->kernel != 0 && p->kernel != 1) return fail(DDFMM_EINVAL, "Unknown kernel type " + std::to_string(p->kernel));
    if (p->ACCU < 1 || p->ACCU > 3) return fail(DDFMM_EINVAL, "ACCU must be 1, 2 or 3");
    if (p->ACCU != 2) return fail(DDFMM_EINVAL, "this build evaluates ACCU=2 (P=6) only");
    if (p->nrhs_max < 1) return fail(DDFMM_EINVAL, "nrhs_max must be >= 1");
    if (p->precision != DDFMM_FP64 && p->precision != DDFMM_MIXED && p->precision != DDFMM_FP64_FUSED &&
        p->precision != DDFMM_MIXED_SPLIT)
        return fail(DDFMM_EINVAL, "unsupported precision mode");
    double lk = std::log2(p->K);
    if (!(p->K >= 1) || std::abs(lk - std::round(lk)) > 1e-12) return fail(DDFMM_EINVAL, "K must be a power of two >= 1");
    try {
        if (c->device >= 0) {
            HIPCHK(hipSetDevice(c->device));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->release();
        c->hp = HostPlan();
        Params& q = c->prm;
        q.K = p->K;
        for (int d = 0; d < 3; ++d) q.ctr[d] = p->ctr[d];
        q.ACCU = p->ACCU; q.NPQ = p->NPQ; q.ptsmax = p->ptsmax; q.maxlevel = p->maxlevel;
        q.numC = p->numC; q.ktype = p->kernel;
        q.fuse_solves = p->precision != DDFMM_FP64;
        q.nrhs_max = p->nrhs_max;
        // the matrix-core kernel apply (mixed modes, four or more right-hand sides)
        // evaluates near-field leaf pairs of a small leaf once for both directions; with
        // leaves of up to 128 points most pairs qualify (with larger leaves few do, and the
        // reduction of the mirror products costs more than it saves)
        q.near_pairs = kapply_on_mfma(p->precision, p->nrhs_max) && p->ptsmax <= 2 * MIRROR_MAX_TARGETS;
        c->nrhs_max = p->nrhs_max;
        c->precision = p->precision;
        std::shared_ptr<const Mlib> mlib = shared ? shared : load_mlib(ldfile, hdfile);
        build_plan(q, pos, n, *mlib, c->hp, c->device < 0);
    } catch (const HipFail& e) {
        return fail(DDFMM_EDEVICE, e.what());
    } catch (const std::exception& e) {
        return fail(DDFMM_ESETUP, e.what());
    }
    // HBM budget: the work buffers hold `slab` right-hand sides; an evaluation of more
    // runs in passes of at most `slab` (each right-hand side is independent, so the
    // results are the same bits as in one pass).
    plan_bytes(c->hp, c->precision, c->fixed_bytes, c->rhs_bytes, c->setup_transient);
    c->slab = c->nrhs_max;
    if (c->hbm_budget > 0) {
        const int64_t need = c->fixed_bytes + std::max(c->rhs_bytes, c->setup_transient);
        if (need > c->hbm_budget) {
            c->ready = false;
            return fail(DDFMM_ENOMEM, "HBM budget " + gib(c->hbm_budget) + " below the minimum " + gib(need) +
                                          " (tables " + gib(c->fixed_bytes) + ", one right-hand side " +
                                          gib(c->rhs_bytes) + ")");
        }
        c->slab = (int)std::min<int64_t>(c->nrhs_max, (c->hbm_budget - c->fixed_bytes) / c->rhs_bytes);
    }
    if (c->device < 0) {
        c->ready = true;
        return DDFMM_OK;
    }
    try {
        HostPlan& h = c->hp;
        DevPlan& d = c->dp;
        d = DevPlan();
        d.N = h.N; d.P = h.P; d.NL = h.NL; d.ktype = c->prm.ktype;
        // MIXED and MIXED_SPLIT: FP32 vectors; FP64 and FP64_FUSED: FP64 arithmetic
        d.mixed = c->precision == DDFMM_MIXED || c->precision == DDFMM_MIXED_SPLIT ? 1 : 0;
        d.kmfma = kapply_on_mfma(c->precision, c->nrhs_max) ? 1 : 0;
        const int64_t esz = d.mixed ? (int64_t)sizeof(float2) : (int64_t)sizeof(double2);
        d.px = c->upload(h.px); d.py = c->upload(h.py); d.pz = c->upload(h.pz);
        d.perm = c->upload(h.perm);
        {
            std::vector<int> iperm(h.perm.size());
            for (size_t i = 0; i < h.perm.size(); ++i) iperm[h.perm[i]] = (int)i;
            d.iperm = c->upload(iperm);
        }
        d.grid = c->upload(h.grid);
        d.opmat = c->alloc<double2>(h.opmat_total);
        if (!h.opmat.empty())
            HIPCHK(hipMemcpy(d.opmat, h.opmat.data(), h.opmat.size() * sizeof(double2), hipMemcpyHostToDevice));
        d.ktasks = c->upload(h.ktasks); d.ksrcs = c->upload(h.ksrcs);
        d.groups = c->upload(h.groups); d.tiles = c->upload(h.tiles);
        c->tiles_bytes = h.tiles.size() * sizeof(GTile);
        d.colB = c->upload(h.colB); d.colC = c->upload(h.colC); d.colA = c->upload(h.colA);
        d.fsrc = c->upload(h.fsrc); d.vtrg = c->upload(h.vtrg);
        d.vchunk = c->upload(h.vchunk); d.vsrc = c->upload(h.vsrc); d.vpk = c->upload(h.vpk);
        d.rtrg = c->upload(h.rtrg);
        d.io = c->alloc<IoPtrs>(1);
        // translation matrices generated on the device (temporary generator tables)
        if (!h.kgen.empty()) {
            double* gp = nullptr;
            KGen* jobs = nullptr;
            HIPCHK(hipMalloc(&gp, h.genpts.size() * sizeof(double)));
            HIPCHK(hipMalloc(&jobs, h.kgen.size() * sizeof(KGen)));
            HIPCHK(hipMemcpy(gp, h.genpts.data(), h.genpts.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(jobs, h.kgen.data(), h.kgen.size() * sizeof(KGen), hipMemcpyHostToDevice));
            HIPCHK(launch_kgen(jobs, (int)h.kgen.size(), gp, d.opmat, d.ktype, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(gp));
            HIPCHK(hipFree(jobs));
        }
        // fused mode: solves folded into the translations by FP64 GEMMs on the table
        if (!h.setup_stages.empty()) {
            d.bt.p[BUF_OPMAT] = d.opmat;
            d.bt.stride[BUF_OPMAT] = 0;
            for (const Stage& s : h.setup_stages) HIPCHK(launch_setup_gemm(d, s, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (d.mixed) {
            d.opmat32 = c->alloc<float2>(h.opmat_total);
            HIPCHK(launch_to_f32(d.opmat, d.opmat32, h.opmat_total, c->stream));
            if (c->precision == DDFMM_MIXED_SPLIT) {    // translations as split-BF16 products
                d.opmatbf = (void*)c->alloc<double2>(h.opmat_total);
                HIPCHK(launch_to_bf16s(d.opmat, d.opmatbf, h.opmat_total, c->stream));
            }
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        // work buffers, after the setup-time tables are gone
        for (int b = 0; b < BUF_COUNT; ++b) {
            d.bt.stride[b] = h.bufsize[b];
            d.bt.p[b] = esz == 8 ? (void*)c->alloc<float2>(h.bufsize[b] * c->slab)
                                 : (void*)c->alloc<double2>(h.bufsize[b] * c->slab);
        }
        d.den_io = c->alloc<double2>(h.N * c->slab);
        d.val_io = c->alloc<double2>(h.N * c->slab);
        // host copies no longer needed on the device path
        std::vector<cpx>().swap(h.opmat);
        std::vector<double>().swap(h.genpts);
        std::vector<KGen>().swap(h.kgen);
        c->ready = true;
    } catch (const HipFail& e) {
        c->release();
        return fail(DDFMM_EDEVICE, e.what());
    } catch (const std::exception& e) {
        c->release();
        return fail(DDFMM_ENOMEM, e.what());
    }
    return DDFMM_OK;
}

// Stages run on the evaluation stream `st` or on the context's second stream; ST_FORK
// makes the second stream wait for the work issued so far on `st`, ST_JOIN the reverse,
// so that the caller's stream sees a complete evaluation.  With timing on, each stage is
// bracketed by events on its own stream.
static void issue_stages(ddfmm_ctx* c, int nrhs, hipStream_t st, bool timing, bool serial = false) {
    const auto& stages = c->hp.stages;
    for (size_t i = 0; i < stages.size(); ++i) {
        const Stage& s = stages[i];
        if (serial && (s.kind == ST_FORK || s.kind == ST_JOIN)) continue;
        if (s.kind == ST_FORK) {
            HIPCHK(hipEventRecord(c->sync_ev[s.buf], st));
            HIPCHK(hipStreamWaitEvent(c->aux, c->sync_ev[s.buf], 0));
            continue;
        }
        if (s.kind == ST_JOIN) {
            HIPCHK(hipEventRecord(c->sync_ev[s.buf], c->aux));
            HIPCHK(hipStreamWaitEvent(st, c->sync_ev[s.buf], 0));
            continue;
        }
        hipStream_t ss = (s.stream && !serial) ? c->aux : st;
        if (timing) HIPCHK(hipEventRecord(c->ev[2 * i], ss));
        hipError_t e = launch_stage(c->dp, s, nrhs, ss);
        if (e != hipSuccess) throw HipFail(std::string("stage ") + s.name + ": " + hipGetErrorString(e));
        if (timing) HIPCHK(hipEventRecord(c->ev[2 * i + 1], ss));
    }
}

// The stage sequence for nrhs captured once from the context's stream (the second stream
// joins the capture through the fork events) and instantiated; replays read the caller's
// buffers through DevPlan::io.
static hipGraphExec_t graph_for(ddfmm_ctx* c, int nrhs) {
    auto it = c->graphs.find(nrhs);
    if (it != c->graphs.end()) return it->second;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        issue_stages(c, nrhs, c->stream, false);
    } catch (...) {
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw HipFail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->graphs[nrhs] = x;
    return x;
}

// One evaluation on `st` reading den and writing val (device buffers).  Evaluations of
// one context share its work buffers, so each one first waits for the end of the
// previous one (which may have been issued on another stream).
static int run_eval(ddfmm_ctx* c, const double2* den, double2* val, int nrhs, hipStream_t st) {
    const auto& stages = c->hp.stages;
    if (c->done_recorded) HIPCHK(hipStreamWaitEvent(st, c->done_ev, 0));
    HIPCHK(launch_set_io(c->dp.io, den, val, st));
    if (c->use_graph && !c->timing) {
        hipGraphExec_t x = graph_for(c, nrhs);
        HIPCHK(hipGraphLaunch(x, st));
    } else {
        if (c->timing && c->ev.size() != 2 * stages.size()) {
            for (auto e : c->ev) (void)hipEventDestroy(e);
            c->ev.assign(2 * stages.size(), nullptr);
            for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
        }
        issue_stages(c, nrhs, st, c->timing, c->serial);
    }
    HIPCHK(hipEventRecord(c->done_ev, st));
    c->done_recorded = true;
    if (c->timing) {
        HIPCHK(hipStreamSynchronize(st));
        c->stage_ms.assign(stages.size(), 0.0);
        for (size_t i = 0; i < stages.size(); ++i) {
            if (stages[i].kind == ST_FORK || stages[i].kind == ST_JOIN) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
            c->stage_ms[i] = ms;
        }
    }
    return DDFMM_OK;
}

int ddfmm_eval(ddfmm_ctx* c, const double* den, double* val, int nrhs) {
    if (!c || !den || !val) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready || c->device < 0) return fail(DDFMM_ESTATE, "context not set up on a device");
    if (nrhs < 1 || nrhs > c->nrhs_max) return fail(DDFMM_EINVAL, "nrhs out of range [1, nrhs_max]");
    try {
        HIPCHK(hipSetDevice(c->device));
        for (int r0 = 0; r0 < nrhs; r0 += c->slab) {
            const int nr = std::min(c->slab, nrhs - r0);
            const size_t bytes = (size_t)c->hp.N * nr * sizeof(double2);
            const size_t off = (size_t)c->hp.N * r0 * 2;
            if (c->done_recorded) HIPCHK(hipStreamWaitEvent(c->stream, c->done_ev, 0));
            HIPCHK(hipMemcpyAsync(c->dp.den_io, den + off, bytes, hipMemcpyHostToDevice, c->stream));
            run_eval(c, c->dp.den_io, c->dp.val_io, nr, c->stream);
            HIPCHK(hipMemcpyAsync(val + off, c->dp.val_io, bytes, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const std::exception& e) {
        return fail(DDFMM_EDEVICE, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_eval_device(ddfmm_ctx* c, const void* den_dev, void* val_dev, int nrhs, void* stream) {
    if (!c || !den_dev || !val_dev) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready || c->device < 0) return fail(DDFMM_ESTATE, "context not set up on a device");
    if (nrhs < 1 || nrhs > c->nrhs_max) return fail(DDFMM_EINVAL, "nrhs out of range [1, nrhs_max]");
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    try {
        HIPCHK(hipSetDevice(c->device));
        for (int r0 = 0; r0 < nrhs; r0 += c->slab)
            run_eval(c, (const double2*)den_dev + c->hp.N * r0, (double2*)val_dev + c->hp.N * r0,
                     std::min(c->slab, nrhs - r0), st);
    } catch (const std::exception& e) {
        return fail(DDFMM_EDEVICE, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_set_graph(ddfmm_ctx* c, int on) {
    if (!c) return fail(DDFMM_EINVAL, "null argument");
    if (!on && c->device >= 0) {
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        c->drop_graphs();
    }
    c->use_graph = on != 0;
    return DDFMM_OK;
}

// ---------------------------------------------------------------- GEMM tile autotuning
// Upload the host tiles after a rebuild: the device table is replaced (the context's
// stream is idle), the captured graphs are dropped and the stage times reset.
static void upload_tiles(ddfmm_ctx* c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->aux));
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->dp.tiles);
    if (it != c->allocs.end()) {
        HIPCHK(hipFree(*it));
        c->allocs.erase(it);
        c->device_bytes -= (int64_t)c->tiles_bytes;
    }
    c->fixed_bytes -= (int64_t)c->tiles_bytes;
    c->dp.tiles = c->upload(c->hp.tiles);
    c->tiles_bytes = c->hp.tiles.size() * sizeof(GTile);
    c->fixed_bytes += (int64_t)c->tiles_bytes;
    c->drop_graphs();
    c->ev.clear();
    c->stage_ms.clear();
}

// Per logical GEMM stage, the milliseconds of its launches in a stage-by-stage evaluation
// on one stream (no overlap with the second stream's work), best of reps.
static std::vector<double> time_logical(ddfmm_ctx* c, int nrhs, int reps) {
    const auto& stages = c->hp.stages;
    std::vector<double> best(c->hp.lgemm.size(), 1e30);
    if (c->ev.size() != 2 * stages.size()) {
        for (auto e : c->ev) (void)hipEventDestroy(e);
        c->ev.assign(2 * stages.size(), nullptr);
        for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    }
    HIPCHK(launch_set_io(c->dp.io, c->dp.den_io, c->dp.val_io, c->stream));
    for (int r = 0; r <= reps; ++r) {            // the first run warms up
        issue_stages(c, nrhs, c->stream, true, true);
        HIPCHK(hipStreamSynchronize(c->stream));
        if (r == 0) continue;
        std::vector<double> t(c->hp.lgemm.size(), 0.0);
        for (size_t i = 0; i < stages.size(); ++i) {
            if (stages[i].kind != ST_GEMM) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
            t[stages[i].lstage] += ms;
        }
        for (size_t l = 0; l < t.size(); ++l) best[l] = std::min(best[l], t[l]);
    }
    return best;
}

int ddfmm_autotune(ddfmm_ctx* c, int nrhs, int reps, int cap, int* mbmax, int* count, double* ms_before,
                   double* ms_after) {
    if (!c || !count) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready || c->device < 0) return fail(DDFMM_ESTATE, "context not set up on a device");
    if (nrhs < 1 || nrhs > c->nrhs_max) return fail(DDFMM_EINVAL, "nrhs out of range [1, nrhs_max]");
    nrhs = std::min(nrhs, c->slab);
    reps = std::max(1, reps);
    const int nl = (int)c->hp.lgemm.size();
    *count = nl;
    try {
        HIPCHK(hipSetDevice(c->device));
        if (c->done_recorded) HIPCHK(hipStreamWaitEvent(c->stream, c->done_ev, 0));
        HIPCHK(hipMemsetAsync(c->dp.den_io, 0, (size_t)c->hp.N * nrhs * sizeof(double2), c->stream));
        std::vector<int> cur(nl);
        for (int l = 0; l < nl; ++l) cur[l] = c->hp.lgemm[l].mbmax;
        // the planner's tiling (height limit GEMM_MB, rows split further for small stages)
        std::vector<double> t0 = time_logical(c, nrhs, reps);
        std::vector<int> best = cur;
        std::vector<double> bt = t0;
        for (int m = 1; m <= GEMM_MB; ++m) {
            retile(c->hp, std::vector<int>(nl, m));
            upload_tiles(c);
            std::vector<double> t = time_logical(c, nrhs, reps);
            for (int l = 0; l < nl; ++l)
                if (t[l] < bt[l]) {
                    bt[l] = t[l];
                    best[l] = m;
                }
        }
        retile(c->hp, best);
        upload_tiles(c);
        std::vector<double> t1 = time_logical(c, nrhs, reps);
        double a = 0, b = 0;
        for (int l = 0; l < nl; ++l) {
            a += t0[l];
            b += t1[l];
            if (mbmax && l < cap) mbmax[l] = best[l];
        }
        if (ms_before) *ms_before = a;
        if (ms_after) *ms_after = b;
    } catch (const std::exception& e) {
        return fail(DDFMM_EDEVICE, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_set_tiles(ddfmm_ctx* c, int count, const int* mbmax) {
    if (!c || !mbmax) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready) return fail(DDFMM_ESTATE, "context not set up");
    if (count != (int)c->hp.lgemm.size())
        return fail(DDFMM_EINVAL, "tile table has " + std::to_string(count) + " entries, the plan has " +
                                      std::to_string(c->hp.lgemm.size()) + " GEMM stages");
    try {
        retile(c->hp, std::vector<int>(mbmax, mbmax + count));
        if (c->device >= 0) {
            HIPCHK(hipSetDevice(c->device));
            upload_tiles(c);
        }
    } catch (const HipFail& e) {
        return fail(DDFMM_EDEVICE, e.what());
    } catch (const std::exception& e) {
        return fail(DDFMM_EINVAL, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_get_tiles(ddfmm_ctx* c, int cap, int* mbmax, char* names /* cap x 24 */, int* count) {
    if (!c || !count) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready) return fail(DDFMM_ESTATE, "context not set up");
    *count = (int)c->hp.lgemm.size();
    for (int l = 0; l < *count && l < cap; ++l) {
        if (mbmax) mbmax[l] = c->hp.lgemm[l].mbmax;
        if (names) std::memcpy(names + 24 * l, c->hp.lgemm[l].name, 24);
    }
    return DDFMM_OK;
}

int ddfmm_direct_sum(int kernel, const double* trg, int m, const double* src, int n, const double* den, int nrhs,
                     double* out) {
    if (!trg || !src || !den || !out || m < 0 || n < 0 || nrhs < 1) return fail(DDFMM_EINVAL, "bad argument");
    if (kernel != 0 && kernel != 1) return fail(DDFMM_EINVAL, "Unknown kernel type " + std::to_string(kernel));
    if (m == 0) return DDFMM_OK;
    double *dt = nullptr, *ds = nullptr;
    double2 *dd = nullptr, *dout = nullptr, *part = nullptr;
    int rc = DDFMM_OK;
    try {
        const DirectSplit sp = direct_split(m, n, nrhs);
        HIPCHK(hipMalloc(&dt, (size_t)m * 3 * sizeof(double)));
        HIPCHK(hipMalloc(&ds, (size_t)std::max(n, 1) * 3 * sizeof(double)));
        HIPCHK(hipMalloc(&dd, (size_t)std::max(n, 1) * nrhs * sizeof(double2)));
        HIPCHK(hipMalloc(&dout, (size_t)m * nrhs * sizeof(double2)));
        if (sp.part_elems > 0) HIPCHK(hipMalloc(&part, (size_t)sp.part_elems * sizeof(double2)));
        HIPCHK(hipMemcpy(dt, trg, (size_t)m * 3 * sizeof(double), hipMemcpyHostToDevice));
        if (n > 0) {
            HIPCHK(hipMemcpy(ds, src, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dd, den, (size_t)n * nrhs * sizeof(double2), hipMemcpyHostToDevice));
        }
        HIPCHK(launch_direct(dt, m, ds, n, dd, nrhs, dout, part, kernel, nullptr));
        HIPCHK(hipMemcpy(out, dout, (size_t)m * nrhs * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (const std::exception& e) {
        rc = fail(DDFMM_EDEVICE, e.what());
    }
    (void)hipFree(dt); (void)hipFree(ds); (void)hipFree(dd); (void)hipFree(dout); (void)hipFree(part);
    return rc;
}

int ddfmm_direct_sum_noalloc(int kernel, const double* trg_dev, int m, const double* src_dev, int n,
                             const double* den_dev, int nrhs, double* out_dev, void* stream) {
    if (!trg_dev || !src_dev || !den_dev || !out_dev || m < 0 || n < 0 || nrhs < 1)
        return fail(DDFMM_EINVAL, "bad argument");
    if (kernel != 0 && kernel != 1) return fail(DDFMM_EINVAL, "Unknown kernel type " + std::to_string(kernel));
    if (m == 0) return DDFMM_OK;
    const DirectSplit sp = direct_split(m, n, nrhs);
    if (sp.part_elems * (int64_t)sizeof(double2) > DIRECT_SCRATCH_BYTES)
        return fail(DDFMM_EINVAL, "direct sum needs " + std::to_string(sp.part_elems * 16) +
                                      " bytes of partial sums, more than the static " +
                                      std::to_string(DIRECT_SCRATCH_BYTES) + "-byte scratch");
    try {
        double2* part = nullptr;
        if (sp.part_elems > 0) HIPCHK(direct_static_scratch(&part));
        HIPCHK(launch_direct(trg_dev, m, src_dev, n, (const double2*)den_dev, nrhs, (double2*)out_dev, part, kernel,
                             (hipStream_t)stream));
    } catch (const std::exception& e) {
        return fail(DDFMM_EDEVICE, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_kernel_matrix(int kernel, const double* trg, int m, const double* src, int n, double* out) {
    if (!trg || !src || !out || m < 0 || n < 0) return fail(DDFMM_EINVAL, "bad argument");
    if (kernel != 0 && kernel != 1) return fail(DDFMM_EINVAL, "Unknown kernel type " + std::to_string(kernel));
    if ((int64_t)m * n == 0) return DDFMM_OK;
    double *dt = nullptr, *ds = nullptr;
    double2* dout = nullptr;
    int rc = DDFMM_OK;
    try {
        HIPCHK(hipMalloc(&dt, (size_t)m * 3 * sizeof(double)));
        HIPCHK(hipMalloc(&ds, (size_t)n * 3 * sizeof(double)));
        HIPCHK(hipMalloc(&dout, (size_t)m * n * sizeof(double2)));
        HIPCHK(hipMemcpy(dt, trg, (size_t)m * 3 * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ds, src, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(launch_kmatrix(dt, m, ds, n, dout, kernel, 0, nullptr));
        HIPCHK(hipMemcpy(out, dout, (size_t)m * n * sizeof(double2), hipMemcpyDeviceToHost));
    } catch (const std::exception& e) {
        rc = fail(DDFMM_EDEVICE, e.what());
    }
    (void)hipFree(dt); (void)hipFree(ds); (void)hipFree(dout);
    return rc;
}

int ddfmm_get_stats(ddfmm_ctx* c, ddfmm_stats* s) {
    if (!c || !s) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready) return fail(DDFMM_ESTATE, "context not set up");
    const PlanStats& p = c->hp.stats;
    std::memset(s, 0, sizeof(*s));
    s->npts = p.npts; s->nleaf = p.nleaf; s->nlf = p.nlf; s->nhf_out = p.nhf_out; s->nhf_in = p.nhf_in;
    s->near_interactions = p.near_interactions; s->leaf_op_interactions = p.leaf_op_interactions;
    s->s2m_interactions = p.s2m_interactions; s->x2chk_interactions = p.x2chk_interactions;
    s->v_pairs = p.v_pairs; s->v_targets = p.v_targets; s->v_sources = p.v_sources;
    s->m2l_pairs = p.m2l_pairs; s->m2l_groups = p.m2l_groups;
    s->m2l_flops = p.m2l_flops; s->gemm_flops = p.gemm_flops;
    s->opmat_elems = p.opmat_elems; s->device_bytes = c->device_bytes;
    s->nlevels_lf = p.nlevels_lf; s->nlevels_hf = p.nlevels_hf; s->max_hf_width = p.max_hf_width;
    s->nstages = (int)c->hp.stages.size();
    return DDFMM_OK;
}

int ddfmm_dump_lists(ddfmm_ctx* c, char* buf, int64_t cap, int64_t* len) {
    if (!c || !len) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready) return fail(DDFMM_ESTATE, "context not set up");
    *len = (int64_t)c->hp.dump.size();
    if (buf && cap > 0) std::memcpy(buf, c->hp.dump.data(), (size_t)std::min<int64_t>(cap, *len));
    return DDFMM_OK;
}

int ddfmm_read_buffer(ddfmm_ctx* c, int buf, int rhs, int64_t off, int64_t count, double* out) {
    if (!c || !out) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready || c->device < 0) return fail(DDFMM_ESTATE, "context not set up on a device");
    if (buf < 0 || buf >= BUF_COUNT || rhs < 0 || rhs >= c->slab || off < 0 || count < 0 ||
        off + count > c->hp.bufsize[buf])
        return fail(DDFMM_EINVAL, "buffer range out of bounds");
    if (count == 0) return DDFMM_OK;
    try {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->dp.mixed) {
            std::vector<float2> tmp((size_t)count);
            HIPCHK(hipMemcpy(tmp.data(), (float2*)c->dp.bt.p[buf] + (int64_t)rhs * c->hp.bufsize[buf] + off,
                             (size_t)count * sizeof(float2), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < count; ++i) { out[2 * i] = tmp[i].x; out[2 * i + 1] = tmp[i].y; }
        } else {
            HIPCHK(hipMemcpy(out, (double2*)c->dp.bt.p[buf] + (int64_t)rhs * c->hp.bufsize[buf] + off,
                             (size_t)count * sizeof(double2), hipMemcpyDeviceToHost));
        }
    } catch (const std::exception& e) {
        return fail(DDFMM_EDEVICE, e.what());
    }
    return DDFMM_OK;
}

int ddfmm_set_hbm_budget(ddfmm_ctx* c, int64_t bytes) {
    if (!c) return fail(DDFMM_EINVAL, "null argument");
    if (bytes < 0) return fail(DDFMM_EINVAL, "negative HBM budget");
    c->hbm_budget = bytes;
    return DDFMM_OK;
}

int ddfmm_memory(ddfmm_ctx* c, int64_t* fixed_bytes, int64_t* per_rhs_bytes, int64_t* setup_bytes, int* slab) {
    if (!c) return fail(DDFMM_EINVAL, "null argument");
    if (!c->ready) return fail(DDFMM_ESTATE, "context not set up");
    if (fixed_bytes) *fixed_bytes = c->fixed_bytes;
    if (per_rhs_bytes) *per_rhs_bytes = c->rhs_bytes;
    if (setup_bytes) *setup_bytes = c->setup_transient;
    if (slab) *slab = c->slab;
    return DDFMM_OK;
}

int ddfmm_set_timing(ddfmm_ctx* c, int on) {
    if (!c) return fail(DDFMM_EINVAL, "null argument");
    c->timing = on != 0;
    c->serial = on == 2;
    return DDFMM_OK;
}

int ddfmm_stage_times(ddfmm_ctx* c, int cap, char* names, double* ms, double* work, int* kinds, int* count) {
    if (!c || !count) return fail(DDFMM_EINVAL, "null argument");
    const auto& st = c->hp.stages;
    *count = (int)st.size();
    for (int i = 0; i < (int)st.size() && i < cap; ++i) {
        if (names) std::memcpy(names + 24 * i, st[i].name, 24);
        if (ms) ms[i] = i < (int)c->stage_ms.size() ? c->stage_ms[i] : 0.0;
        if (work) work[i] = st[i].work;
        if (kinds) kinds[i] = st[i].kind + 16 * st[i].prec;
    }
    return DDFMM_OK;
}

int ddfmm_stage_exec(ddfmm_ctx* c, int cap, double* exec, int* count) {
    if (!c || !count) return fail(DDFMM_EINVAL, "null argument");
    const auto& st = c->hp.stages;
    *count = (int)st.size();
    for (int i = 0; i < (int)st.size() && i < cap; ++i)
        if (exec) exec[i] = st[i].exec;
    return DDFMM_OK;
}

}  // extern "C"
