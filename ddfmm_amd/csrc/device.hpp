// Device-side view of an evaluation plan and the kernel launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "plan.hpp"

namespace ddfmm {

struct BufTable {
    void* p[BUF_COUNT + 1];        // double2 (FP64 mode) or float2 (mixed mode) elements;
                                   // p[BUF_OPMAT]: the FP64 operator table (setup GEMMs)
    int64_t stride[BUF_COUNT + 1]; // elements per rhs slab
};

// The caller's density and potential buffers of the current evaluation, read by the
// gather and scatter kernels from device memory, so that a captured graph of the stage
// sequence serves every call (set before each evaluation by launch_set_io).
struct IoPtrs {
    const double2* den;            // [rhs][N] input order (gather source)
    double2* val;                  // [rhs][N] input order (scatter target)
};

struct DevPlan {
    int64_t N = 0;
    int P = 0, NL = 0, ktype = 0;
    int mixed = 0;                 // 1: FP32 vectors/translations, FP64 solves
    int kmfma = 0;                 // mixed and nrhs_max >= 4: kernel applies on the matrix cores
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    int* perm = nullptr;           // sorted position -> input index
    int* iperm = nullptr;          // input index -> sorted position
    int* grid = nullptr;
    double2* opmat = nullptr;
    float2* opmat32 = nullptr;     // FP32 copy of the operator table (mixed mode)
    void* opmatbf = nullptr;       // split-BF16 copy (16 bytes per entry; mixed-split mode)
    KTask* ktasks = nullptr;
    KSrc* ksrcs = nullptr;
    GGroup* groups = nullptr;
    GTile* tiles = nullptr;
    int64_t *colB = nullptr, *colC = nullptr, *colA = nullptr;
    FftSrc* fsrc = nullptr;
    VTarget* vtrg = nullptr;
    VChunk* vchunk = nullptr;
    int* vsrc = nullptr;
    int* vpk = nullptr;
    RTarget* rtrg = nullptr;
    BufTable bt{};
    double2* den_io = nullptr;     // context-owned [rhs][N] staging of the host-buffer entry point
    double2* val_io = nullptr;
    IoPtrs* io = nullptr;          // device copy of the current evaluation's buffers
};

hipError_t launch_stage(const DevPlan& d, const Stage& s, int nrhs, hipStream_t st);
hipError_t launch_set_io(IoPtrs* io, const double2* den, double2* val, hipStream_t st);
hipError_t launch_setup_gemm(const DevPlan& d, const Stage& s, hipStream_t st);   // FP64, on the operator table
hipError_t launch_kgen(const KGen* jobs, int njobs, const double* gp, double2* opmat, int ktype, hipStream_t st);
// Direct sum split over S source ranges (chunk sources each, RC right-hand sides per
// block); with S > 1 the ranges write part (S x nrhs x M complex, part_elems) and a
// fixed-order reduction adds them into out.  The split depends on (M, Ns, nrhs) only, so
// every entry point that uses it sums in the same order.
struct DirectSplit {
    int S = 1, chunk = 0, rc = 1;
    int64_t part_elems = 0;
};
constexpr int64_t DIRECT_SCRATCH_BYTES = int64_t(32) << 20;   // bound of part_elems * 16 bytes
DirectSplit direct_split(int M, int Ns, int nrhs);
hipError_t launch_direct(const double* trg, int M, const double* src, int Ns, const double2* den, int nrhs,
                         double2* out, double2* part, int ktype, hipStream_t st);
// The static partial-sum array of the no-allocation direct sum on the current device.
hipError_t direct_static_scratch(double2** out);
hipError_t launch_kmatrix(const double* trg, int M, const double* src, int Ns, void* out, int ktype, int f32,
                          hipStream_t st);
hipError_t launch_to_f32(const double2* in, float2* out, int64_t n, hipStream_t st);
hipError_t launch_to_bf16s(const double2* in, void* out, int64_t n, hipStream_t st);

}  // namespace ddfmm
