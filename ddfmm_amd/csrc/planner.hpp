// Host planner: builds the tree (Tree), fetches operators (Mlib) and flattens every
// stage of Wave3d::eval into descriptor tables (HostPlan) for the device.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "host.hpp"
#include "plan.hpp"

namespace ddfmm {

struct PlanStats {
    int64_t npts = 0, nleaf = 0, nlf = 0, nhf_out = 0, nhf_in = 0;
    int64_t near_interactions = 0;      // sum over U/W/X-direct pairs of n_t * n_s (leaf kernel, point sources)
    int64_t near_mirrored = 0;          // near-field leaf pairs evaluated once for both directions
    int64_t leaf_op_interactions = 0;   // leaf kernel interactions with operator point sets (L2T, W via upeq)
    int64_t s2m_interactions = 0, x2chk_interactions = 0;
    int64_t v_pairs = 0, v_targets = 0, v_sources = 0;
    int64_t m2l_pairs = 0, m2l_groups = 0;
    double m2l_flops = 0, gemm_flops = 0;   // useful complex-GEMM flops (8*M*K*N) per RHS
    int64_t opmat_elems = 0, opmat_generated = 0;
    int nlevels_lf = 0, nlevels_hf = 0, max_hf_width = 0;
};

// A planned GEMM stage before it is cut into one launch per tile height: its groups, in
// order, and the largest tile height its tiling uses (GEMM_MB, or less where rows were
// split for a small stage).
struct LogicalGemm {
    std::vector<int> groups;
    int mbmax = GEMM_MB;
    char name[24] = {0};
};

struct HostPlan {
    Params prm;
    int64_t N = 0;
    int P = 0, NL = 0;
    std::vector<int> perm;                 // sorted position -> input index
    std::vector<double> px, py, pz;        // leaf-sorted points then LF operator point sets
    std::vector<cpx> opmat;                // uploaded operator tables (prefix of the table)
    int64_t opmat_total = 0;               // uploaded + generated
    std::vector<int> grid;                 // FFT grid index tables
    std::vector<double> genpts;            // kernel-generation point table (xyz rows)
    std::vector<KGen> kgen;
    std::vector<KTask> ktasks;
    std::vector<KSrc> ksrcs;
    std::vector<GGroup> groups;
    std::vector<double> group_flops;       // useful flops per RHS of each group
    std::vector<GTile> tiles;
    std::vector<int64_t> colB, colC, colA;
    std::vector<FftSrc> fsrc;
    std::vector<VTarget> vtrg;
    std::vector<VPair> vpair;
    std::vector<VChunk> vchunk;
    std::vector<int> vsrc;                  // chunk source unions (spectrum indices)
    std::vector<int> vpk;                   // packed pairs per target, padded to multiples of 8
    std::vector<RTarget> rtrg;
    std::vector<Stage> stages;
    std::vector<Stage> setup_stages;       // FP64 GEMMs on the operator table, run once after generation
    std::vector<LogicalGemm> lgemm;        // logical GEMM stages of `stages` (Stage::lstage)
    int64_t bufsize[BUF_COUNT] = {0};      // elements per rhs
    PlanStats stats;
    std::string dump;                      // canonical text of the lists (debug/parity)
};

void build_plan(const Params& prm, const double* pos, int64_t N, const Mlib& mlib,
                HostPlan& plan, bool want_dump);

// Algorithmic work per RHS of a stage (Stage::work) and, for GEMM stages, the MFMA flops
// the kernel executes under its tiling (Stage::exec).
void stage_work(const HostPlan& plan, Stage& s);

// Rebuild the tiles of every logical GEMM stage with at most mbmax[l] row blocks of 16 per
// tile (1..GEMM_MB; a group's rows spread evenly over ceil(blocks / mbmax) tiles) and the
// GEMM launches of `stages` (one per tile height), after setup: the setup-time GEMMs'
// tiles are dropped.  Other stages keep their order.
void retile(HostPlan& plan, const std::vector<int>& mbmax);

}  // namespace ddfmm
