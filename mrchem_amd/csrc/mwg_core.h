// Host-side numerics of the MI355X multiwavelet apply: scaling bases, two-scale
// filters, Gaussian-sum kernel fit and the per-level operator bookkeeping that
// decides which NS blocks the device builds and which tuples the device screens.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mwg {

constexpr int MAXK = 10;  // k <= 9

// Gauss-Legendre rule on [0,1], ascending nodes.
void gauss_legendre01(int q, double *x, double *w);

struct Basis {
    int k = 0, K = 0, K2 = 0, K3 = 0, type = 0;  // type 0 Legendre, 1 Interpolating
    std::vector<double> U;       // K x K: phi^B_i = sum_j U[i][j] sqrt(2j+1) P_j(2x-1)
    std::vector<double> qx, qw;  // K-point rule on [0,1]
    std::vector<double> phiq;    // [q][i] = phi^B_i(qx[q])
    std::vector<double> W;       // 2K x 2K two-scale matrix [[H0 H1][G0 G1]]
    // cross-correlation C_ij(z) = int phi_i(u) phi_j(u - z) du on the halves z in [-1,0]
    // (h = 0) and [0,1] (h = 1), expanded in orthonormal shifted Legendre polynomials
    // L_p(t), t = z + 1 - h in [0,1], p < 2K: ccoef[(h*K*K + i*K + j)*2K + p]
    std::vector<double> ccoef;
    void phi(double x, double *out) const;
};
void build_basis(Basis &B, int k, int type);

struct KernelFit {
    std::vector<double> a, w;  // exp(-a r^2) terms (physical units)
    double rmin = 0, rmax = 0, eps = 0, h = 0, err = 0;
};
// 1/r (mu = 0) or exp(-mu r)/r as a Gaussian sum with max |fit - K| r <= eps on [rmin, rmax]
KernelFit fit_kernel(double mu, double eps, double rmin, double rmax);

// node key: depth (6 bits) | 3 x 19-bit translations relative to the root corner
inline uint64_t node_key(int depth, int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)depth << 57) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}
inline uint32_t key_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb53fe1a85ec3ULL;
    k ^= k >> 33;
    return (uint32_t)k;
}

// screening constants shared with the device code
constexpr double SCREEN_REL = 0.1;  // tuple threshold = SCREEN_REL * prec * ||g|| / sqrt(n_terms)
constexpr double ACT_REL = 1e-4;    // term active at depth > 0 iff max ||C,B,A|| > ACT_REL * prec * max ||T||
constexpr double GAUSS_CUT = 46.0;  // exp(-46) ~ 1e-20: translation band of a Gaussian

// split rule: ||w|| > prec * ||tree|| * 2^{-(n+1)/2}
bool split_check(int scale, double wnorm, double prec, double tnorm);

// The packed tree layout of import/export (include/mwg.h): the roots first, x fastest,
// then complete groups of 8 siblings in child order (x & 1 | (y & 1) << 1 | (z & 1) << 2),
// each group after its parent; depths in [0, max_depth], translations inside the box at
// the node's depth, no node twice.  Returns an empty string or the first violation.
std::string check_tree_layout(const int boxes[3], int max_depth, int64_t n, const int *depth, const int64_t *lr);

// ---- MRChem orbital files (src/qmfunctions/Orbital.cpp:118-190, ComplexFunction.h:33-37,
// Orbital.h:51-56): the .meta file is the raw FunctionData {int real_size, int imag_size,
// bool is_shared} (12 bytes with padding) followed, for an orbital, by OrbitalData {int
// rank_id, int spin, int occ, double error} (24 bytes with padding), little-endian x86-64
// layout.  A density's .meta holds the FunctionData only (Density.cpp:55-67).
struct OrbitalMeta {
    int32_t real_size = 0, imag_size = 0;  // tree chunks of each part (0: part absent)
    int32_t is_shared = 0;
    int32_t rank_id = 0, spin = 0, occ = 0;  // spin: 0 paired, 1 alpha, 2 beta
    double error = 0.0;
};
std::string write_meta(const std::string &path, const OrbitalMeta &m, bool orbital);
// reads 12 (density) or 36 (orbital) bytes; `orbital` tells which
std::string read_meta(const std::string &path, OrbitalMeta &m, bool &orbital);

// Tree files ("<name>.tree", this library's own format; MRCPP's FunctionTree::saveTree
// format is not part of the reference): a 96-byte header, then the nodes in the packed
// import layout, each as int32 depth, int32 0, int64 l[3], double coefficients[8 K^3]
// (the 8 components, each K^3 values with x fastest), in chunks of `chunk_nodes` nodes.
struct TreeFileInfo {
    int32_t order = 0, basis = 0, root_scale = 0, max_depth = 0;
    int32_t corner[3] = {0, 0, 0}, boxes[3] = {0, 0, 0};
    double scaling_factor = 1.0;
    int64_t nnodes = 0;
    int32_t chunk_nodes = 0, nchunks = 0;
};
constexpr int TREE_FILE_CHUNK = 1024;  // nodes per chunk written
std::string write_tree_file(const std::string &path, TreeFileInfo info, const int *depth, const int64_t *lr,
                            const double *coefs);
// header only (depth, lr, coefs null) or the whole file into arrays sized from the header
std::string read_tree_file(const std::string &path, TreeFileInfo &info, int *depth, int64_t *lr, double *coefs);

}  // namespace mwg
