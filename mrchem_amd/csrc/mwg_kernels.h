// Kernel launchers and the device-side tables they consume.
#pragma once

#include <hip/hip_runtime.h>

#include "mwg_core.h"
#include "mwg_dev.h"

namespace mwg {

constexpr double GAUSS_CUT_D = GAUSS_CUT;

struct DevBasis {
    int K;
    const double *U;     // K x K
    const double *qx;    // K
    const double *qw;    // K
    const double *phiq;  // K x K, [q][i]
    const double *W;     // 2K x 2K
};

struct ProjGeom {
    double lo[3];
    double sfac;
    int n0;
};

struct WorkItem {
    int job;
    int node;    // index in the job's output tree
    int l[3];    // translation relative to the root corner
    int parent;  // the node's parent in the output tree (-1 on the roots)
};

struct JobTab {
    DHash inHash;
    DHash genHash;
    const double *inCoef;
    const double *inNorm;
    const double *operBlocks;  // this depth's block array of the job's operator
    // split-FP16 copies (mixed precision): each K x K operator block and each K^3 input
    // component as rows of f16 hi and lo parts (layouts at split16_block_uint4), a
    // power-of-two exponent per block
    const uint4 *operBlocks16;
    const int *operExp16;
    const uint4 *inCoef16;
    const int *inExp16;
    const double *operNorms;   // this depth's norm array
    double *outCoef;
    double *outNorm;
    double thr;
    double route;  // mixed mode: slices with a bound above this run in FP64
    const int *pre;   // [ntab + 1] candidate prefix over terms (union boxes ub), device-built
    const int *pre0;  // [ntab + 1] same over the ft = 0 union boxes ub0
    int tab0, ntab;
    int lmax[3];
    int depth;
    int ubox[3];  // union of ub over the table's terms (device-built)
    // refinement-predictor features of the nodes' children (see k_child_features)
    double feat_tnorm;   // output norm estimate at the start of the iteration
    double feat_innorm;  // input tree norm
    double prec;
    float feat_mu;       // log10(mu + 1e-3)
    float feat_cell;     // log10 of the children's cell width
    int n0;              // root scale
};

// Device-built contraction schedule of one refinement iteration (sizes the host reads
// back with the output norms; overflow set when a buffer capacity was too small or the
// generated-node pool did not hold every generated node, and then no kernel of the
// schedule runs).
struct TaskSummary {
    long long ntuples, nslices;
    int ntasks, nstasks, nreds, nslots, nnslots;
    int overflow;  // bit 0: a capacity below, bit 1: generated-node pool
    int gen_count, gen_overflow;  // generated-node counter and table-overflow flag, for the host
    int schunk;                   // slices per MFMA task used by the schedule (mixed)
};
struct TaskCaps {
    long long tuples, slices;
    int tasks, stasks, reds, slots, nnslots;
};

// one Gaussian term of one operator level, as the screening-box kernel reads it
struct LevTerm {
    double w;        // weight (unit coordinates)
    long long blk;   // block offset (doubles) in the level block array
    long long nrm;   // norm offset (doubles) in the level norm array
    long long env;   // offset of the term's 4 x (band + 1) norm envelopes in the level envelope array
    double maxn[4];  // largest block norm per type over the band
    int band;
    int active;
};

// screening-box request for one (job, depth) table
struct BoxJob {
    const LevTerm *terms;
    const double *env;
    double thr;   // tuple threshold of the job at this depth (0 on the roots)
    double F[8];  // largest input component norm per component at this depth
    int nterms;
    int depth;
    int tab0;     // first TermTab of this table
    int jt;       // index of the JobTab that receives the routing bound
    int pre_off;  // first entry of this table's prefix arrays (tab0 + jt)
};

// the candidate walk of an output node is split over up to ENUM_MAX_PARTS workgroups so
// that a work list of nw nodes launches about ENUM_BLOCKS of them (host and device agree)
constexpr int ENUM_BLOCKS = 512, ENUM_MAX_PARTS = 16;
__host__ __device__ inline int enum_parts(int nw) {
    if (nw <= 0) return 1;
    const int p = (ENUM_BLOCKS + nw - 1) / nw;
    return p < 1 ? 1 : (p > ENUM_MAX_PARTS ? ENUM_MAX_PARTS : p);
}

struct GenState {
    uint64_t *keys;
    int *job;
    int *count;
    int cap;
    int *overflow;
    double *pool;  // pcap x K^3
    double *norm;  // pcap
    int pcap;      // entries allocated in pool and norm (<= cap)
    long long *tabcnt;  // per job table: [4 t + 3] counts the generated nodes it inserted
};

hipError_t launch_tblock(hipStream_t s, int n, const int *item_term, const long long *item_delta, const double *b,
                         double pre, const DevBasis &B, const double *gx, const double *gw, double *out);
hipError_t launch_tblock_mom(hipStream_t s, int n, const int *item_term, const long long *item_delta,
                             const double *b, double pre, int K, const double *ccoef, const double *gx,
                             const double *gw, double *out);
hipError_t launch_nsblocks(hipStream_t s, int n, const long long *item_delta, const int *item_row0,
                           const int *item_tlo, const int *item_thi, const double *t1, const double *W, int K,
                           double *blocks, double *norms, double *norms2 = nullptr);
hipError_t launch_norms(hipStream_t s, int n, const int *nodes, const double *coef, double *norm, int K3);
hipError_t launch_dot(hipStream_t s, int n, const int *ia, const int *ib, const int *c0, const double *A,
                      const double *B, int K3, double *part);
// sum over node pairs of a . b (part: nblocks doubles of block sums, res: the total)
hipError_t launch_dot_pairs(hipStream_t s, int n, const unsigned long long *pairs, const double *A,
                            const double *B, int K3, int nblocks, double *part, double *res);
hipError_t launch_scale(hipStream_t s, size_t n, double *x, double c);
hipError_t launch_project_gauss(hipStream_t s, int nw, const int *work, const int *depth, const int *lr,
                                const ProjGeom &geo, const DevBasis &B, int nterm, const double *coef,
                                const double *alpha, const double *center, const double *poly, double *out);
hipError_t launch_project_hyd(hipStream_t s, int nw, const int *work, const int *depth, const int *lr,
                              const ProjGeom &geo, const DevBasis &B, int nterm, const double *coef, const int *q,
                              double Z, const double *origin, double *out);
// the device loop's start of an iteration, done by the screening-box kernels (LoopJob etc.
// are declared below): thresholds from the jobs' running norms instead of the tables'
// (table jb = job jb), table counts cleared, work counters and contraction stamps reset,
// generated-node range started
struct LoopJob;
struct LoopCtl;
struct LoopInit {
    const LoopJob *lj;  // null: the host loop (thresholds from the tables)
    long long *tabcnt;
    LoopCtl *ctl;
    unsigned long long *ts;
    const int *gen_count;
};
// TermTab[tab0 + t] for every (table, term): per-combination screening boxes from the
// monotone norm envelopes, union boxes, and the table's FP64 routing bound
hipError_t launch_job_boxes(hipStream_t s, int nbj, int max_terms, const BoxJob *bj, TermTab *tabs, double *tbmax,
                            JobTab *jobs, int *pre, int *pre0, int *gen_overflow, const LoopInit &li = LoopInit{});
// Device work counts: the work-list kernels below take an optional nwp, the work list's
// node count on the device (written by an earlier kernel of the stream); nw is then only
// an upper bound for the launch, and the kernels loop over the nodes on a grid of at most
// DEVCOUNT_GRID workgroups, taking them from the zeroed counter *queue when one is given.
constexpr int DEVCOUNT_GRID = 2048;
hipError_t launch_gen_mark(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs, const TermTab *tabs,
                           const GenState &gs, const int *nwp = nullptr, int *queue = nullptr);
// generated nodes first .. min(device count, cap) - 1 (cap = entries allocated in the pool;
// firstp: first read on the device)
hipError_t launch_gen_compute(hipStream_t s, int first, int cap, const JobTab *jobs, const GenState &gs,
                              const double *W, int K, const int *firstp = nullptr);
// per tuple segment (output node, component): first tuple, count, first FP64 task, first
// partial slot (-1: its single task writes the output directly)
struct SegRec {
    long long start;
    int count;
    int task0;
    int slot0;
    int pad_stask0;  // on component 0 of a node: its first slice task (mixed)
};
// nparts workgroups share each output node's candidates (counts and offsets per
// (node, segment, part))
hipError_t launch_enum(hipStream_t s, int mode, int nw, int nparts, const WorkItem *work, const JobTab *jobs,
                       const TermTab *tabs, const GenState &gs, int *thread_off, int *seg_total,
                       const long long *seg_ofs, Tuple *out, const TaskSummary *sum, const SegRec *segs,
                       const int *nwp = nullptr, int *queue = nullptr);
hipError_t launch_work_norms(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs, int K3,
                             double *compact, const int *nwp = nullptr, unsigned long long *ts_end = nullptr);
// the device clock into *t (one thread; marks the end of the kernels before it on the stream)
hipError_t launch_stamp(hipStream_t s, unsigned long long *t);
struct TransformItem {
    double *coef;
    double *norm;  // component norms of the tree (null: not updated)
    int parent;
    int first_child;
};
hipError_t launch_transform(hipStream_t s, int mode, int n, const TransformItem *items, const double *M, int K);
// one wavefront task: a chunk of the tuples of segment (g, gt)
struct ContractTask {
    long long start;
    int count;
    int g;     // work index
    int gt;
    int slot;  // partial slot, -1 = write the output node directly
};
struct ReduceItem {
    int g, gt, slot0, nslots;  // K^3 slots of FP64 tuple tasks
    int nslot0, nnslots;       // 8 K^3 node slots of mixed-precision slice tasks
};
struct SliceTask {
    long long start;
    int count;
    int g;
    int slot;  // node slot (8 K^3 doubles)
};
hipError_t launch_enum_mixed(hipStream_t s, int mode, int K, int nw, int nparts, const WorkItem *work,
                             const JobTab *jobs,
                             const TermTab *tabs, const GenState &gs, int *thread_off, int *seg_total,
                             const long long *seg_ofs, Tuple *out, Slice *sout, const TaskSummary *sum,
                             const SegRec *segs, const long long *gslice, const int *nwp = nullptr,
                             int *queue = nullptr);
// Split-FP16 copies of FP64 blocks (mwg_mfma.hip): block b (values src[b bvals ..]) as rows
// of K values, scaled by 2^exp[b] so that the block's largest magnitude lies in [0.5, 1)
// (exp 0 for a zero block); x 2^e = hi + lo up to about 2^-23 of it.  K = 8: each row
// [hi 8 | lo 8] (two uint4).  K = 10, role 0 (operator blocks, A operands): each row
// [hi 10 | hi 10 | lo 10 | 0 0] (four uint4); role 1 (input nodes, B operands): each row
// [hi 10 | lo 10] (40 bytes).  With count, only blocks first .. min(*count, first + n) - 1;
// with firstp, first is read on the device and the blocks are first .. min(*count, n) - 1.
// the same for the blocks of several arrays (e.g. the input trees of a batch) in one launch:
// item i's blocks are first[i] .. first[i + 1] - 1 of the concatenation (first[nitems] = total)
struct Split16Item {
    const double *src;
    uint4 *dst;
    int *exp;
};
hipError_t launch_split16_multi(hipStream_t s, int nitems, const Split16Item *items, const long long *first,
                                long long total, int bvals, int K, int role);
hipError_t launch_split16(hipStream_t s, int n, int first, const int *count, int bvals, int K, const double *src,
                          uint4 *dst, int *exp, int role = 0, const int *firstp = nullptr);
// uint4 per block of bvals values of a split-FP16 copy
inline size_t split16_block_uint4(int K, int role, int bvals) {
    return K == 8 ? (size_t)bvals / 4 : role == 0 ? (size_t)bvals / 10 * 4 : (size_t)bvals / 10 * 40 / 16;
}
// FP32 / split-FP16 MFMA slice contraction (K = 8 or 10; mwg_mfma.hip).  A task is a run of
// slices of one output node; mfma_tasks_per_block(K) tasks share a 256-thread block.  Each
// task adds the surviving tuples of its slices to tabcnt[4 t + 2] of its job table t.
// genPool16 / genExp16: the generated nodes' split-FP16 copies.
int mfma_tasks_per_block(int K);
hipError_t launch_contract_mfma(hipStream_t s, int K, int ntask, const TaskSummary *sum, int grid,
                                const SliceTask *tasks, const WorkItem *work, const JobTab *jobs,
                                const TermTab *tabs, const Slice *slices, const double *genPool, double *nodeslots,
                                long long *tabcnt, const uint4 *genPool16, const int *genExp16);
struct MulItem {
    int node;          // output node
    int depth;
    int l[3];
    int asrc, asteps;  // input a: node (steps = 0) or leaf ancestor and levels to refine
    int bsrc, bsteps;
};
hipError_t launch_multiply(hipStream_t s, int n, const MulItem *items, const ProjGeom &geo, const DevBasis &B,
                           double coef, const double *acoef, const double *bcoef, double *out);
hipError_t launch_add(hipStream_t s, int n, const MulItem *items, const DevBasis &B, double ca, const double *acoef,
                      double cb, const double *bcoef, double *out);
// from the per-segment counts (nseg = 8: plain FP64, 9: mixed): segment offsets, FP64
// tasks (full chunks first), MFMA slice tasks, reduce items, slot numbering.  wchunk:
// slices per MFMA task, or (<= 0) about -wchunk tasks over the iteration's slices.
hipError_t launch_build_tasks(hipStream_t s, int nw, int nseg, int nparts, int chunk, int wchunk, const int *tot,
                              TaskCaps caps, const int *gen_count, int gen_pcap, long long *ofs, ContractTask *tasks,
                              SliceTask *stasks, ReduceItem *reds, TaskSummary *sum, SegRec *segs, int *segcnt,
                              long long *gslice, const int *gen_overflow, const WorkItem *work, long long *tabcnt,
                              int ntab, const int *nwp = nullptr);
// ---- device-resident refinement (mwg_kernels.hip)
// The plain method's refinement loop without a host round trip: every job of the batch has
// one job table per iteration (table index = job index, all work nodes of iteration d at
// depth d); after the contraction and the output norms, k_loop_split takes the split test
// and appends the children to the output trees, k_loop_emit writes the next work list.
constexpr int LOOP_MAX_DEPTH = 32, LOOP_MAX_JOBS = 4096;
struct LoopJob {
    int *depth, *parent, *child, *lr;  // output tree topology on the device (lr: 3 per node), cap nodes
    int n, cap;                        // nodes so far, capacity
    int w0, nwk;                       // this iteration's work nodes [w0, w0 + nwk) of the work list
    int splits, pad;                   // nodes split by this iteration
    double sNorm, wNorm, gsq;          // running output norm (gsq < 0: none yet)
    double prec;
    long long tuples, tuples_fp64, slices, gen;  // statistics (as mwg_apply_stats)
};
// what the host reads back after each iteration
struct LoopIter {
    int nw_next;   // work nodes of the next iteration (0: done, or stopped by an overflow)
    int overflow;  // sticky: bit 0 schedule capacity, 1 generated-node pool, 2 node capacity,
                   // 3 work-list capacity, 4 generated-node table, 5 transform-item capacity
};
struct LoopCtl {
    int nw[2];     // work list sizes: iteration d reads nw[d & 1], writes nw[(d + 1) & 1]
    int queue[4];  // work counters of the iteration's k_gen_mark, k_enum count and fill (zeroed)
    int genFirst;  // generated-node counter at the start of the iteration
    int overflow;  // sticky (LoopIter bits)
    double splitfac[LOOP_MAX_DEPTH];  // 2^(-(n0 + d + 1) / 2) of the split test at depth d
    LoopIter iter[LOOP_MAX_DEPTH];
    int tofs[LOOP_MAX_DEPTH + 1];     // first transform item of the split nodes of depth d
    unsigned long long ts[LOOP_MAX_DEPTH][2];  // contraction start and end, device clock
};
// the roots (nr per job, translations rootlr, the same for every job) in the topology and
// the first work list
hipError_t launch_loop_init(hipStream_t s, int njobs, int nr, const LoopJob *lj, const int *rootlr, WorkItem *work);
// per job (one workgroup): running norm, split test of its work nodes, children appended
// (cidx[i]: first child of work node i, -1 if not split)
hipError_t launch_loop_split(hipStream_t s, int njobs, LoopJob *lj, const WorkItem *work, const double *wnorm,
                             int max_depth, LoopCtl *ctl, int *cidx, const long long *tabcnt,
                             unsigned long long *ts_end = nullptr);
// the next work list (and its size in *nw_next and in ctl->iter[d] and host_iter[d], pinned
// host memory the host reads once the iteration is done), and the split nodes as transform
// items of depth d (from ctl->tofs[d], capacity tcap)
hipError_t launch_loop_emit(hipStream_t s, int njobs, LoopJob *lj, const WorkItem *work, const int *nwp,
                            const int *cidx, WorkItem *next, int *nw_next, int wcap, const TaskSummary *sum,
                            LoopCtl *ctl, int d, const JobTab *jt, TransformItem *titems, int tcap,
                            LoopIter *host_iter);

// refinement-predictor features of the 8 children of every work node (row i*8 + a, a the
// child index), from the work nodes' norms just computed, their parents' norms and the
// input tree around each child
hipError_t launch_child_features(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs,
                                 const double *wnorms, float *feat);
// with sum: the task count is read on the device; grid (blocks) covers the capacity
hipError_t launch_contract_fp64(hipStream_t s, int K, int ntask, const TaskSummary *sum, int grid,
                                const ContractTask *tasks, const WorkItem *work, const JobTab *jobs,
                                const TermTab *tabs, const Tuple *tuples, const double *genPool, double *partials,
                                unsigned long long *ts = nullptr);
hipError_t launch_move_groups(hipStream_t s, int npairs, const int *from_to, double *coef, double *norm, int nb);
hipError_t launch_reduce_partials(hipStream_t s, int n, const TaskSummary *sum, int grid, const ReduceItem *items,
                                  const WorkItem *work, const JobTab *jobs, const double *partials,
                                  const double *nodeslots, int K3, double *wnorm = nullptr);

}  // namespace mwg
