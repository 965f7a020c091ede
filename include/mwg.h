/* mwg.h — C ABI of the MI355X multiwavelet Green's-function apply (libmwgreen.so).
 *
 * Drop-in boundary for the hot path of MRChem: mrcpp::apply(prec, out, oper, inp)
 * with mrcpp::PoissonOperator / mrcpp::HelmholtzOperator, as called from
 *   src/scf_solver/HelmholtzVector.cpp:171-181           (Helmholtz, per orbital)
 *   src/qmoperators/two_electron/CoulombPotential.cpp:85  (Poisson on the density)
 *   src/qmoperators/two_electron/ExchangePotential.cpp:259,318 (pair Poissons)
 * The reference interface is a set of C++ templates in namespace mrcpp (MRCPP is an
 * external, non-vendored library, .gitmodules:4-6); this header is the C ABI the
 * C++ facade in include/mwg_mrcpp.hpp (and any FFI) binds to.
 *
 * Conventions
 *  - Every call returns 0 on success and a non-zero status on failure; the message
 *    is available from mwg_last_error(ctx).  No exception crosses the ABI.
 *  - The caller owns every handle and destroys it explicitly.  Operators are
 *    borrowed by mwg_apply (ExchangePotential.cpp:19 "does not take ownership").
 *  - One context per host thread.  Calls are ordered on the context's HIP stream;
 *    functions that return host data synchronise that stream.
 *  - Trees live in device memory (HBM).  Import/export use host buffers in the
 *    documented packed layout: nodes in parent-before-children order, complete
 *    sibling groups of 8, translations relative to the root corner at the node's
 *    depth, and 8*(k+1)^3 FP64 coefficients per node, component-major
 *    (component bit 0 = x, bit 1 = y, bit 2 = z; 0 = scaling), within a component
 *    index (iz*(k+1) + iy)*(k+1) + ix.
 */
#ifndef MWG_H
#define MWG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mwg_ctx_s *mwg_ctx;
typedef struct mwg_mra_s *mwg_mra;
typedef struct mwg_tree_s *mwg_tree;
typedef struct mwg_oper_s *mwg_oper;
typedef struct mwg_predictor_s *mwg_predictor;

/* status codes */
#define MWG_OK 0
#define MWG_ERR_ARG 1
#define MWG_ERR_HIP 2
#define MWG_ERR_STATE 3
#define MWG_ERR_RANGE 4

/* apply flags */
#define MWG_APPLY_FP64 0u     /* plain method: FP64 contractions, norm-test refinement */
#define MWG_APPLY_MIXED 1u    /* MFMA slice contractions with split-FP16 operands (hi + lo, 22
                                 significand bits) on f16 MFMA: k = 7 FP32 sums over <= 64 slices,
                                 FP64 beyond; k = 9 every slice's FP32 result summed in FP64.
                                 Slices with large bounds run as FP64 tuples.  k = 7 and k = 9 only
                                 (MWG_ERR_ARG at other orders) */
#define MWG_APPLY_PREDICT 2u  /* learned refinement predictor (norm test stays the judge; the output
                                 tree can differ from the plain one, see mwg_predictor_create) */

/* basis types (src/mrenv.cpp:62-69: "I" Interpolating, "L" Legendre) */
#define MWG_BASIS_LEGENDRE 0
#define MWG_BASIS_INTERPOLATING 1

typedef struct mwg_apply_stats {
    int64_t nodes_out;       /* nodes of the output tree (output nodes of this apply) */
    int64_t tuples;          /* surviving (g, f, ft, gt, term) tuples contracted (FP64 + MFMA slices) */
    int64_t tuples_fp64;     /* of which contracted in FP64 (all in plain mode; mixed mode routes
                                slices with large bounds to FP64 tuples) */
    int64_t gen_nodes;       /* generated (scaling-only) input nodes of this apply */
    int64_t iters;           /* refinement iterations */
    int64_t slices;          /* MFMA slices (g, f, ft, term: all 8 gt at once), mixed mode */
    int64_t predicted_splits;/* output nodes computed ahead of their parent's norm test (predictor) */
    int64_t predicted_wasted;/* of which discarded because the norm test did not split the parent */
    double t_apply;          /* seconds: work-list build + contraction + refine + transforms */
    double t_contract;       /* seconds spent in contraction kernels (device clock: first contraction
                                workgroup to last reduction workgroup, summed over the iterations) */
    double flops;            /* flops issued: FP64 tuples * 6 (k+1)^4 + MFMA flops of the slices
                                (f16 16x16x32 MFMAs as issued, padding included: k = 7 28 per slice,
                                2*16*16*32*28, a skipped z half counted; k = 9 90, 2*16*16*32*90) */
    int64_t device_loop;     /* 1: the refinement loop ran on the device (split tests and work lists
                                without a host round trip); 0: on the host (first batch on a context,
                                predictor, traces, maxIter, or a capacity the earlier batches did not
                                reach) */
} mwg_apply_stats;

/* ---- context (mrenv / device) ------------------------------------------------ */
/* Handles may be destroyed in any order: an MRA or predictor keeps its context, and a
 * tree or operator its MRA, until the last of them is destroyed. */
int mwg_ctx_create(int device, mwg_ctx *out);
int mwg_ctx_destroy(mwg_ctx ctx);
const char *mwg_last_error(mwg_ctx ctx);
int mwg_sync(mwg_ctx ctx);
/* contraction task granularity (no reference counterpart: MRCPP threads over output
 * nodes with OpenMP): at most this many tuples per FP64 wave task (1..4096; an iteration
 * with fewer than 4096 tasks' worth of tuples uses smaller tasks, down to 16) and MFMA slices per
 * workgroup task (0..4096, 0 = automatic; k = 7 uses at most 256, 128 when automatic;
 * k = 9 sizes automatic tasks from each refinement iteration's slice total, 16 to 128
 * slices, and raises a task size, automatic or set, where the FP64 node slots of an
 * iteration would exceed 8 GiB).  Results agree to FP64 rounding
 * for any setting on the FP64 path and at k = 9, and to the rounding of the MFMA
 * slices' FP32 sums at k = 7; mrchem_amd.autotune picks the fastest on a sample
 * workload. */
int mwg_ctx_set_tuning(mwg_ctx ctx, int fp64_chunk, int slice_chunk);
int mwg_ctx_get_tuning(mwg_ctx ctx, int *fp64_chunk, int *slice_chunk);
/* return the device memory the context keeps for reuse (blocks of destroyed trees and
 * operators, up to half the device memory) to the device, after the context's work has
 * finished; *released (may be null) gets the bytes.  No reference counterpart: for callers
 * that share the GPU with other allocators (as torch.cuda.empty_cache). */
int mwg_ctx_trim(mwg_ctx ctx, int64_t *released);

/* ---- MRA: BoundingBox + scaling basis + max depth
 * replaces mrcpp::MultiResolutionAnalysis<3>(BoundingBox<3>(min_scale, corner, boxes), basis, max_depth)
 * (src/mrenv.cpp:46-72, tests/unit_tests.cpp:28-44); scaling_factor stretches the box. */
int mwg_mra_create(mwg_ctx ctx, int root_scale, const int *corner, const int *boxes, double scaling_factor,
                   int order, int basis_type, int max_depth, mwg_mra *out);
int mwg_mra_destroy(mwg_mra mra);

/* ---- function trees: replaces mrcpp::FunctionTree<3>(MRA) (QMFunction.cpp:68-77) */
int mwg_tree_create(mwg_mra mra, mwg_tree *out);
int mwg_tree_destroy(mwg_tree tree);
int mwg_tree_nnodes(mwg_tree tree, int64_t *n);              /* FunctionTree::getNNodes */
int mwg_tree_sqnorm(mwg_tree tree, double *out);             /* FunctionTree::getSquareNorm */
int mwg_tree_maxdepth(mwg_tree tree, int *out);
int mwg_tree_mra(mwg_tree tree, mwg_mra *out);              /* FunctionTree::getMRA (borrowed handle) */
int mwg_mra_order(mwg_mra mra, int *order);                  /* MultiResolutionAnalysis::getOrder */
/* import refuses (MWG_ERR_ARG) input that breaks the packed layout: the roots first in
 * x-fastest order, then complete groups of 8 siblings in child order, each after its
 * parent; depths in [0, max_depth]; translations inside the box; no node twice */
int mwg_tree_import(mwg_tree tree, int64_t n, const int *depth, const int64_t *lr, const double *coefs);
/* the same check without a context or device (boxes: the MRA's root boxes); msg (may be
 * null) receives the first violation.  No reference counterpart: MRCPP reads its own
 * tree files only. */
int mwg_tree_layout_check(const int *boxes, int max_depth, int64_t n, const int *depth, const int64_t *lr,
                          char *msg, int msg_len);
int mwg_tree_export(mwg_tree tree, int *depth, int64_t *lr, double *coefs);
int mwg_tree_rescale(mwg_tree tree, double c);               /* FunctionTree::rescale (HelmholtzVector.cpp:81) */
int mwg_tree_dot(mwg_tree a, mwg_tree b, double *out);       /* mrcpp::dot */
/* mwg_tree_dot without device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Node pairs and partial sums live in a static 8 MiB
 * device array per GPU (at most 1048576 node pairs, the roots plus every deeper node of a
 * that b also holds; MWG_ERR_RANGE beyond).  Every call on one GPU shares that array:
 * calls that overlap on one GPU race.  Returns the same value as mwg_tree_dot up to the
 * order of the FP64 summation. */
int mwg_tree_dot_noalloc(mwg_tree a, mwg_tree b, double *out);
int mwg_tree_eval(mwg_tree tree, int64_t n, const double *xyz, double *out);

/* ---- MRChem orbital files: replaces Orbital::saveOrbital / loadOrbital
 * (src/qmfunctions/Orbital.cpp:118-190) and Density::saveDensity / loadDensity
 * (src/qmfunctions/Density.cpp:55-125).  "<prefix>.meta" holds the raw FunctionData
 * (ComplexFunction.h:33-37) and, for an orbital, OrbitalData (Orbital.h:51-56): 12 or 36
 * bytes in the x86-64 layout (padding written as zeros); the parts go to
 * "<prefix>_re.tree" and "<prefix>_im.tree". */
typedef struct mwg_orbital_meta {
    int32_t real_size, imag_size; /* tree chunks of each part (0: the part is absent) */
    int32_t is_shared;
    int32_t rank_id, spin, occ;   /* spin: 0 paired, 1 alpha, 2 beta (src/mrchem.h:41) */
    double error;
} mwg_orbital_meta;
int mwg_meta_write(const char *path, const mwg_orbital_meta *meta, int orbital); /* orbital: 36 bytes, else 12 */
int mwg_meta_read(const char *path, mwg_orbital_meta *meta, int *orbital);
/* FunctionTree::saveTree(name) / loadTree(name) (Orbital.cpp:138, 180): the file
 * "<name>.tree" in this library's format (MRCPP's is not part of the reference): a 96-byte
 * header (magic "MWGTREE1", order, basis, root scale, max depth, corner[3], boxes[3],
 * scaling factor, node count, chunk nodes (1024), chunk count, record bytes), then per
 * node in the import layout int32 depth, int32 0, int64 l[3], 8 (k+1)^3 doubles.  Load
 * refuses a file written for another MRA.  nchunks (may be null) receives the chunk
 * count, the .meta's real_size or imag_size. */
int mwg_tree_save(mwg_tree tree, const char *name, int *nchunks);
int mwg_tree_load(mwg_tree tree, const char *name);
/* the same file without a device, from and into host arrays in the import layout; read
 * with null arrays returns the header (node count) only */
typedef struct mwg_tree_file_info {
    int32_t order, basis, root_scale, max_depth;
    int32_t corner[3], boxes[3];
    double scaling_factor;
    int64_t nnodes;
    int32_t chunk_nodes, nchunks; /* set by the writer */
} mwg_tree_file_info;
int mwg_tree_file_write(const char *name, const mwg_tree_file_info *info, const int *depth, const int64_t *lr,
                        const double *coefs);
int mwg_tree_file_read(const char *name, mwg_tree_file_info *info, int *depth, int64_t *lr, double *coefs);

/* ---- projection of analytic functions: replaces mrcpp::project (qmfunction_utils.cpp:88-120)
 * Gaussian sum: term i = coef[i] * prod_d poly[15i+5d+e] (x_d-c_d)^e * exp(-alpha[i] |x-c|^2). */
int mwg_project_gauss(double prec, mwg_tree out, int n, const double *coef, const double *alpha,
                      const double *center, const double *poly);
/* hydrogen orbitals (src/analyticfunctions/HydrogenFunction.cpp:34-165) and products of two:
 * term i = coef[i] * H_{q[6i..6i+2]} (* H_{q[6i+3..6i+5]} when q[6i+3] > 0), (n,l,m). */
int mwg_project_hydrogen(double prec, mwg_tree out, int n, const double *coef, const int *q, double Z,
                         const double *origin);

/* ---- pointwise product: replaces mrcpp::multiply(prec, out, coef, a, b) as used by
 * qmfunction::multiply (src/qmfunctions/qmfunction_utils.cpp:141-263) for the exchange pair
 * products (ExchangePotential.cpp:312, 329, 335), V*phi and the density.  out = coef*a*b on
 * the union grid of a and b, refined by the wavelet-norm split test. */
int mwg_multiply(double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b);
/* n independent products out[i] = coef[i]*a[i]*b[i] at prec[i]: no single reference
 * counterpart (MRChem loops over qmfunction::multiply for the pairs of
 * ExchangePotential.cpp:312-335); each output is identical to its mwg_multiply, with the
 * launches of all products merged and one host round trip per refinement level. */
int mwg_multiply_batch(int n, const double *prec, mwg_tree *out, const double *coef, mwg_tree *a, mwg_tree *b);

/* ---- linear combination: replaces mrcpp::add(prec, out, ca, a, cb, b) as used by
 * QMFunction::add / qmfunction::linear_combination (src/qmfunctions/QMFunction.cpp:157-200):
 * out = ca*a + cb*b on the union grid of a and b (exact in the compressed representation). */
int mwg_add(mwg_tree out, double ca, mwg_tree a, double cb, mwg_tree b);

/* ---- operators: replace new mrcpp::PoissonOperator(MRA, prec) (SCFDriver.cpp:279) and
 * mrcpp::HelmholtzOperator(MRA, mu, prec) (HelmholtzVector.cpp:171).  Kernels 1/r and
 * exp(-mu r)/r (the -1/(2 pi) rescale stays with the caller, HelmholtzVector.cpp:81,135).
 * Creation fits the Gaussian expansion (host); the non-standard-form blocks of each scale
 * are computed on the device by the first apply that needs them, for all operators of the
 * batch together, and kept until the operator is destroyed. */
int mwg_oper_poisson(mwg_mra mra, double prec, mwg_oper *out);
int mwg_oper_helmholtz(mwg_mra mra, double mu, double prec, mwg_oper *out);
int mwg_oper_destroy(mwg_oper oper);
int mwg_oper_nterms(mwg_oper oper, int *n);
int mwg_oper_terms(mwg_oper oper, double *exponents, double *weights);
/* NS blocks of one term at one depth (0 .. the MRA's max depth) for translations [dlo, dhi]:
 * out[((delta-dlo)*4 + type)*(k+1)^2 + a*(k+1) + i], type = 2*gt_bit + ft_bit (T, C, B, A) */
int mwg_oper_blocks(mwg_oper oper, int term, int depth, int64_t dlo, int64_t dhi, double *out);

/* ---- apply: replaces mrcpp::apply(prec, out, oper, inp[, maxIter])
 * (HelmholtzVector.cpp:176,180; CoulombPotential.cpp:85; ExchangePotential.cpp:137,259,318).
 * `out` must be empty (no nodes) — it is built adaptively; `inp` is not modified. */
int mwg_apply(double prec, mwg_tree out, mwg_oper oper, mwg_tree inp, int max_iter, unsigned flags,
              mwg_apply_stats *stats);
/* n independent applies (exchange pair Poissons, per-orbital Helmholtz) in one batch:
 * one work list and one contraction launch per refinement level for all of them.  After a
 * first batch on the context, the refinement levels (split tests, children, work lists) run
 * on the device without host round trips, with capacities taken from the earlier batches;
 * a batch that exceeds them runs with the host loop instead (same results;
 * stats.device_loop tells which; MWG_HOST_LOOP=1 in the environment forces the host loop). */
int mwg_apply_batch(int n, const double *prec, mwg_tree *out, mwg_oper *oper, mwg_tree *inp, int max_iter,
                    unsigned flags, mwg_apply_stats *stats);

/* ---- learned refinement predictor (no counterpart in the reference: MRCPP refines one level
 * per iteration, TreeBuilder/WaveletAdaptor).  A 2 x 64 tanh MLP (FP32) trained on refinement
 * traces predicts which freshly created output nodes the norm test will split; with
 * MWG_APPLY_PREDICT their children are computed in the same iteration and kept only when the
 * norm test (still run on every node) splits the parent.  Features: see mrchem_amd/csrc/mwg_predict.h. */
int mwg_predictor_create(mwg_ctx ctx, uint64_t seed, mwg_predictor *out);
int mwg_predictor_destroy(mwg_predictor pred);
int mwg_predictor_sizes(int *nparams, int *nfeatures);
int mwg_predictor_get(mwg_predictor pred, float *params, float *mean, float *sdev);
int mwg_predictor_set(mwg_predictor pred, const float *params, const float *mean, const float *sdev);
/* mini-batch Adam on (X[n][nfeatures], y[n] in {0,1}); standardisation from X; mean BCE before/after */
int mwg_predictor_train(mwg_predictor pred, int64_t n, const float *X, const float *y, int epochs, int batch,
                        float lr, uint64_t seed, double *loss_before, double *loss_after);
/* gradient of the mean BCE over all n samples with the current parameters and standardisation */
int mwg_predictor_gradient(mwg_predictor pred, int64_t n, const float *X, const float *y, float *grad);
int mwg_predictor_infer(mwg_predictor pred, int64_t n, const float *X, float *prob);
int mwg_ctx_set_predictor(mwg_ctx ctx, mwg_predictor pred, float threshold);
/* refinement traces: features of every node created by a split and its own split decision */
int mwg_trace_enable(mwg_ctx ctx, int on);
int mwg_trace_size(mwg_ctx ctx, int64_t *n);
int mwg_trace_get(mwg_ctx ctx, float *X, float *y);

#ifdef __cplusplus
}
#endif

#endif /* MWG_H */
