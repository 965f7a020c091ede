// Host-side numerics: quadrature, scaling bases, two-scale filters, kernel fit.
//
// Scaling bases follow MRCPP's two choices as selected by src/mrenv.cpp:62-69
// (Legendre "L", Interpolating "I"); the wavelet rows of the two-scale matrix are
// any orthonormal complement (norms, and hence every screening and refinement
// decision, do not depend on that choice).  The kernel fit writes 1/r and
// exp(-mu r)/r as (2/sqrt(pi)) int exp(-r^2 e^{2s} - mu^2 e^{-2s}/4 + s) ds and
// applies the trapezoidal rule in s (Harrison et al., J. Chem. Phys. 121, 11587
// (2004)); terms that are constant over the box to eps^0.4 are merged.
#include "mwg_core.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

namespace mwg {

// Gauss-Legendre rule on [0,1] in extended precision: Newton on P_q from the asymptotic
// guess for each root, ascending nodes.
static void gauss_legendre01_ld(int q, long double *x, long double *w) {
    for (int i = 0; i < q; i++) {
        long double t = std::cos(M_PI * (i + 0.75L) / (q + 0.5L));
        long double p0 = 1, p1 = t, dp = 1;
        for (int it = 0; it < 100; it++) {
            p0 = 1.0L;
            p1 = t;
            for (int n = 2; n <= q; n++) {
                long double p2 = ((2.0L * n - 1.0L) * t * p1 - (n - 1.0L) * p0) / n;
                p0 = p1;
                p1 = p2;
            }
            dp = q * (t * p1 - p0) / (t * t - 1.0L);
            long double dt = p1 / dp;
            t -= dt;
            if (std::fabs((double)dt) < 1e-19) break;
        }
        p0 = 1.0L;
        p1 = t;
        for (int n = 2; n <= q; n++) {
            long double p2 = ((2.0L * n - 1.0L) * t * p1 - (n - 1.0L) * p0) / n;
            p0 = p1;
            p1 = p2;
        }
        dp = q * (t * p1 - p0) / (t * t - 1.0L);
        x[q - 1 - i] = 0.5L * (1.0L + t);
        w[q - 1 - i] = 1.0L / ((1.0L - t * t) * dp * dp);
    }
}

void gauss_legendre01(int q, double *x, double *w) {
    std::vector<long double> xl(q), wl(q);
    gauss_legendre01_ld(q, xl.data(), wl.data());
    for (int i = 0; i < q; i++) {
        x[i] = (double)xl[i];
        w[i] = (double)wl[i];
    }
}

// sqrt(2n+1) P_n(2x-1), n < K, in extended precision
static void legendre_scaled_ld(int K, long double x, long double *out) {
    long double t = 2.0L * x - 1.0L, p0 = 1.0L, p1 = t;
    out[0] = 1.0L;
    if (K > 1) out[1] = std::sqrt(3.0L) * t;
    for (int n = 2; n < K; n++) {
        long double p2 = ((2.0L * n - 1.0L) * t * p1 - (n - 1.0L) * p0) / n;
        p0 = p1;
        p1 = p2;
        out[n] = std::sqrt(2.0L * n + 1.0L) * p2;
    }
}

static void legendre_scaled(int K, double x, double *out) {
    double t = 2.0 * x - 1.0, p0 = 1.0, p1 = t;
    out[0] = 1.0;
    if (K > 1) out[1] = std::sqrt(3.0) * t;
    for (int n = 2; n < K; n++) {
        double p2 = ((2.0 * n - 1.0) * t * p1 - (n - 1.0) * p0) / n;
        p0 = p1;
        p1 = p2;
        out[n] = std::sqrt(2.0 * n + 1.0) * p2;
    }
}

void Basis::phi(double x, double *out) const {
    double l[MAXK];
    legendre_scaled(K, x, l);
    for (int i = 0; i < K; i++) {
        double s = 0.0;
        for (int j = 0; j < K; j++) s += U[i * K + j] * l[j];
        out[i] = s;
    }
}

void build_basis(Basis &B, int k, int type) {
    B.k = k;
    B.K = k + 1;
    B.K2 = B.K * B.K;
    B.K3 = B.K2 * B.K;
    B.type = type;
    const int K = B.K, N = 2 * K;
    // The basis tables and the two-scale matrix are built in extended precision and
    // rounded once: the wavelet rows below amplify errors in the scaling rows by up to
    // ~1e6 at k = 9 (the flipped rows are nearly dependent modulo the scaling rows).
    std::vector<long double> xl(K), wl(K), Ul(K * K, 0.0L);
    gauss_legendre01_ld(K, xl.data(), wl.data());
    if (type == 0) {
        for (int i = 0; i < K; i++) Ul[i * K + i] = 1.0L;
    } else {
        long double l[MAXK];
        for (int j = 0; j < K; j++) {
            legendre_scaled_ld(K, xl[j], l);
            for (int i = 0; i < K; i++) Ul[j * K + i] = std::sqrt(wl[j]) * l[i];
        }
    }
    auto phil = [&](long double x, long double *out) {
        long double l[MAXK];
        legendre_scaled_ld(K, x, l);
        for (int i = 0; i < K; i++) {
            long double acc = 0.0L;
            for (int j = 0; j < K; j++) acc += Ul[i * K + j] * l[j];
            out[i] = acc;
        }
    };
    B.qx.assign(K, 0.0);
    B.qw.assign(K, 0.0);
    B.U.assign(K * K, 0.0);
    B.phiq.assign(K * K, 0.0);
    for (int q = 0; q < K; q++) {
        B.qx[q] = (double)xl[q];
        B.qw[q] = (double)wl[q];
    }
    for (int i = 0; i < K * K; i++) B.U[i] = (double)Ul[i];
    for (int q = 0; q < K; q++) {
        long double v[MAXK];
        phil(xl[q], v);
        for (int i = 0; i < K; i++) B.phiq[q * K + i] = (double)v[i];
    }

    std::vector<long double> Wl(N * N, 0.0L);
    long double pa[MAXK], pb[MAXK];
    for (int a = 0; a < 2; a++)
        for (int q = 0; q < K; q++) {
            phil(0.5L * (xl[q] + a), pa);
            phil(xl[q], pb);
            for (int i = 0; i < K; i++)
                for (int j = 0; j < K; j++) Wl[i * N + a * K + j] += wl[q] * pa[i] * pb[j] / std::sqrt(2.0L);
        }
    // wavelet rows: orthonormal complement built from the sign-flipped scaling rows
    for (int j = 0; j < K; j++) {
        long double v[2 * MAXK];
        for (int c = 0; c < K; c++) {
            v[c] = Wl[j * N + c];
            v[K + c] = -Wl[j * N + K + c];
        }
        for (int pass = 0; pass < 2; pass++)
            for (int r = 0; r < K + j; r++) {
                long double d = 0.0L;
                for (int c = 0; c < N; c++) d += v[c] * Wl[r * N + c];
                for (int c = 0; c < N; c++) v[c] -= d * Wl[r * N + c];
            }
        long double nrm = 0.0L;
        for (int c = 0; c < N; c++) nrm += v[c] * v[c];
        nrm = std::sqrt(nrm);
        for (int c = 0; c < N; c++) Wl[(K + j) * N + c] = v[c] / nrm;
    }
    B.W.assign(N * N, 0.0);
    for (int i = 0; i < N * N; i++) B.W[i] = (double)Wl[i];

    // cross-correlation in shifted Legendre polynomials (exact: degree <= 2K-1 per half)
    const int P = 2 * K;
    std::vector<double> tx(P), tw(P);
    gauss_legendre01(P, tx.data(), tw.data());
    B.ccoef.assign((size_t)2 * K * K * P, 0.0);
    std::vector<long double> cz(K * K);
    double pu[MAXK], pv[MAXK], L[2 * MAXK];
    for (int h = 0; h < 2; h++)
        for (int q = 0; q < P; q++) {
            const double t = tx[q];
            const double z = t - 1.0 + h;
            // C_ij(z) by the K-point rule on the overlap of the two cells
            const double lo = std::max(0.0, z), hi = std::min(1.0, 1.0 + z);
            std::fill(cz.begin(), cz.end(), 0.0L);
            for (int r = 0; r < K; r++) {
                const double u = lo + (hi - lo) * B.qx[r];
                const long double wr = (long double)(hi - lo) * B.qw[r];
                B.phi(u, pu);
                B.phi(u - z, pv);
                for (int i = 0; i < K; i++)
                    for (int j = 0; j < K; j++) cz[i * K + j] += wr * pu[i] * pv[j];
            }
            // orthonormal shifted Legendre values at t
            const double x = 2.0 * t - 1.0;
            double p0 = 1.0, p1 = x;
            L[0] = 1.0;
            if (P > 1) L[1] = std::sqrt(3.0) * x;
            for (int n = 2; n < P; n++) {
                const double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
                p0 = p1;
                p1 = p2;
                L[n] = std::sqrt(2.0 * n + 1.0) * p2;
            }
            for (int ij = 0; ij < K * K; ij++)
                for (int n = 0; n < P; n++)
                    B.ccoef[((size_t)h * K * K + ij) * P + n] += (double)(tw[q] * cz[ij]) * L[n];
        }
}

// max over a log grid of r in [rmin, rmax] of r |sum_i w_i exp(-a_i r^2) - exp(-mu r) / r|;
// returns early (with a value above stop) once a point exceeds stop
static double fit_error(const KernelFit &F, const std::vector<double> &r, const std::vector<double> &g, double stop) {
    double err = 0.0;
    for (size_t j = 0; j < r.size(); j++) {
        double f = 0.0;
        for (size_t i = 0; i < F.a.size(); i++) f += F.w[i] * std::exp(-F.a[i] * r[j] * r[j]);
        err = std::max(err, std::fabs(f - g[j]) * r[j]);
        if (err > stop) return err;
    }
    return err;
}

KernelFit fit_kernel(double mu, double eps, double rmin, double rmax) {
    const double L = std::log(10.0 / eps);
    const double s_hi = std::log(std::sqrt(L) / rmin);
    double s_lo = std::log(eps * std::sqrt(M_PI) / (20.0 * rmax));
    if (mu > 0.0) s_lo = std::max(s_lo, -0.5 * std::log(4.0 * L / (mu * mu)));
    const double tau = std::pow(eps, 0.4);
    KernelFit F;
    F.rmin = rmin;
    F.rmax = rmax;
    F.eps = eps;
    const int J = 400;
    std::vector<double> rg(J), gg(J);
    for (int j = 0; j < J; j++) {
        rg[j] = rmin * std::pow(rmax / rmin, j / (double)(J - 1));
        gg[j] = std::exp(-mu * rg[j]) / rg[j];
    }
    for (int m = 0; m < 400; m++) {
        const double h0 = std::pow(0.95, m);
        const int N = (int)std::ceil((s_hi - s_lo) / h0) + 1;
        const double h = (s_hi - s_lo) / (N - 1);
        std::vector<double> a, w;
        double wsum = 0.0, awsum = 0.0;
        int merged = 0;
        for (int i = 0; i < N; i++) {
            const double s = s_lo + h * i;
            const double ai = std::exp(2.0 * s);
            const double wi = 2.0 / std::sqrt(M_PI) * h * std::exp(s) * std::exp(-mu * mu * std::exp(-2.0 * s) / 4.0);
            if (ai * rmax * rmax < tau) {
                wsum += wi;
                awsum += wi * ai;
                merged++;
            } else {
                a.push_back(ai);
                w.push_back(wi);
            }
        }
        std::vector<double> a1, w1;
        if (merged > 0 && wsum > 0.0) {
            a1.push_back(awsum / wsum);
            w1.push_back(wsum);
        }
        a1.insert(a1.end(), a.begin(), a.end());
        w1.insert(w1.end(), w.begin(), w.end());
        F.a.clear();
        F.w.clear();
        for (size_t i = 0; i < a1.size(); i++) {
            const double rs = std::min(std::max(1.0 / std::sqrt(2.0 * a1[i]), rmin), rmax);
            if (w1[i] * rs * std::exp(-a1[i] * rs * rs) >= eps * 1e-3) {
                F.a.push_back(a1[i]);
                F.w.push_back(w1[i]);
            }
        }
        F.h = h;
        F.err = fit_error(F, rg, gg, eps);
        if (F.err <= eps) return F;
    }
    F.err = fit_error(F, rg, gg, INFINITY);
    return F;
}

bool split_check(int scale, double wnorm, double prec, double tnorm) {
    const double thr = std::max(2.0 * 2.220446049250313e-16, prec * tnorm * std::pow(2.0, -0.5 * (scale + 1)));
    return wnorm > thr;
}

std::string check_tree_layout(const int boxes[3], int max_depth, int64_t n, const int *depth, const int64_t *lr) {
    const int64_t nr = (int64_t)boxes[0] * boxes[1] * boxes[2];
    if (n < nr || (n - nr) % 8 != 0) return "expected the roots followed by complete sibling groups of 8";
    if (max_depth < 0 || max_depth > 30) return "max depth outside [0, 30]";
    std::unordered_map<uint64_t, int64_t> index;
    index.reserve((size_t)n * 2);
    std::vector<char> has_children;
    has_children.reserve((size_t)n);
    int64_t group_parent = -1;
    for (int64_t i = 0; i < n; i++) {
        const std::string at = "node " + std::to_string(i) + ": ";
        const int d = depth[i];
        if (d < 0 || d > max_depth) return at + "depth outside [0, max depth]";
        for (int q = 0; q < 3; q++)
            if (lr[3 * i + q] < 0 || lr[3 * i + q] >= ((int64_t)boxes[q] << d)) return at + "translation outside the box";
        const int64_t x = lr[3 * i], y = lr[3 * i + 1], z = lr[3 * i + 2];
        const uint64_t key = node_key(d, x, y, z);
        if (index.count(key)) return at + "duplicate node";
        if (i < nr) {
            const int64_t ex = i % boxes[0], ey = (i / boxes[0]) % boxes[1], ez = i / ((int64_t)boxes[0] * boxes[1]);
            if (d != 0 || x != ex || y != ey || z != ez) return at + "the roots must come first, in x-fastest order";
        } else {
            const int a = (int)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
            if (d == 0 || a != (int)((i - nr) % 8)) return at + "sibling groups must be complete and in child order";
            const auto it = index.find(node_key(d - 1, x >> 1, y >> 1, z >> 1));
            if (it == index.end()) return at + "node before its parent";
            if (a == 0) {
                if (has_children[(size_t)it->second]) return at + "parent already has children";
                has_children[(size_t)it->second] = 1;
                group_parent = it->second;
            } else if (it->second != group_parent) {
                return at + "sibling groups must be complete and in child order";
            }
        }
        index.emplace(key, i);
        has_children.push_back(0);
    }
    return std::string();
}

// ------------------------------------------------------------------ orbital files

namespace {
struct FileCloser {
    void operator()(FILE *f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;
constexpr char TREE_MAGIC[8] = {'M', 'W', 'G', 'T', 'R', 'E', 'E', '1'};
constexpr size_t TREE_HEADER = 96;

template <class T>
void put(unsigned char *b, size_t off, T v) {
    std::memcpy(b + off, &v, sizeof(T));
}
template <class T>
T get(const unsigned char *b, size_t off) {
    T v;
    std::memcpy(&v, b + off, sizeof(T));
    return v;
}
}  // namespace

// byte layout of FunctionData + OrbitalData on x86-64 (the padding is written as zeros)
std::string write_meta(const std::string &path, const OrbitalMeta &m, bool orbital) {
    unsigned char b[36] = {};
    put<int32_t>(b, 0, m.real_size);
    put<int32_t>(b, 4, m.imag_size);
    b[8] = m.is_shared ? 1 : 0;
    put<int32_t>(b, 12, m.rank_id);
    put<int32_t>(b, 16, m.spin);
    put<int32_t>(b, 20, m.occ);
    put<double>(b, 28, m.error);
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) return "cannot open " + path + " for writing";
    const size_t n = orbital ? 36 : 12;
    if (std::fwrite(b, 1, n, f.get()) != n) return "short write to " + path;
    return std::string();
}

std::string read_meta(const std::string &path, OrbitalMeta &m, bool &orbital) {
    unsigned char b[37];
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) return "cannot open " + path;
    const size_t n = std::fread(b, 1, sizeof(b), f.get());
    if (n != 12 && n != 36) return path + ": a .meta file holds 12 or 36 bytes";
    m = OrbitalMeta{};
    m.real_size = get<int32_t>(b, 0);
    m.imag_size = get<int32_t>(b, 4);
    if (b[8] > 1) return path + ": is_shared is not a bool";
    m.is_shared = b[8];
    orbital = n == 36;
    if (orbital) {
        m.rank_id = get<int32_t>(b, 12);
        m.spin = get<int32_t>(b, 16);
        m.occ = get<int32_t>(b, 20);
        m.error = get<double>(b, 28);
    }
    if (m.real_size < 0 || m.imag_size < 0) return path + ": negative tree size";
    return std::string();
}

std::string write_tree_file(const std::string &path, TreeFileInfo info, const int *depth, const int64_t *lr,
                            const double *coefs) {
    if (info.order < 1 || info.order > MAXK - 1 || info.nnodes < 0) return "tree file: bad order or node count";
    const int K = info.order + 1;
    const size_t rec = 32 + sizeof(double) * 8 * (size_t)K * K * K;
    info.chunk_nodes = TREE_FILE_CHUNK;
    info.nchunks = (int32_t)((info.nnodes + TREE_FILE_CHUNK - 1) / TREE_FILE_CHUNK);
    unsigned char h[TREE_HEADER] = {};
    std::memcpy(h, TREE_MAGIC, 8);
    put<int32_t>(h, 8, info.order);
    put<int32_t>(h, 12, info.basis);
    put<int32_t>(h, 16, info.root_scale);
    put<int32_t>(h, 20, info.max_depth);
    for (int d = 0; d < 3; d++) {
        put<int32_t>(h, 24 + 4 * d, info.corner[d]);
        put<int32_t>(h, 36 + 4 * d, info.boxes[d]);
    }
    put<double>(h, 48, info.scaling_factor);
    put<int64_t>(h, 56, info.nnodes);
    put<int32_t>(h, 64, info.chunk_nodes);
    put<int32_t>(h, 68, info.nchunks);
    put<int64_t>(h, 72, (int64_t)rec);
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) return "cannot open " + path + " for writing";
    if (std::fwrite(h, 1, TREE_HEADER, f.get()) != TREE_HEADER) return "short write to " + path;
    std::vector<unsigned char> r(rec);
    for (int64_t i = 0; i < info.nnodes; i++) {
        std::memset(r.data(), 0, 32);
        put<int32_t>(r.data(), 0, depth[i]);
        for (int q = 0; q < 3; q++) put<int64_t>(r.data(), 8 + 8 * q, lr[3 * i + q]);
        std::memcpy(r.data() + 32, coefs + (size_t)i * 8 * K * K * K, rec - 32);
        if (std::fwrite(r.data(), 1, rec, f.get()) != rec) return "short write to " + path;
    }
    return std::string();
}

std::string read_tree_file(const std::string &path, TreeFileInfo &info, int *depth, int64_t *lr, double *coefs) {
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) return "cannot open " + path;
    unsigned char h[TREE_HEADER];
    if (std::fread(h, 1, TREE_HEADER, f.get()) != TREE_HEADER) return path + ": truncated header";
    if (std::memcmp(h, TREE_MAGIC, 8) != 0) return path + ": not a tree file";
    TreeFileInfo t;
    t.order = get<int32_t>(h, 8);
    t.basis = get<int32_t>(h, 12);
    t.root_scale = get<int32_t>(h, 16);
    t.max_depth = get<int32_t>(h, 20);
    for (int d = 0; d < 3; d++) {
        t.corner[d] = get<int32_t>(h, 24 + 4 * d);
        t.boxes[d] = get<int32_t>(h, 36 + 4 * d);
    }
    t.scaling_factor = get<double>(h, 48);
    t.nnodes = get<int64_t>(h, 56);
    t.chunk_nodes = get<int32_t>(h, 64);
    t.nchunks = get<int32_t>(h, 68);
    const int64_t rec = get<int64_t>(h, 72);
    if (t.order < 1 || t.order > MAXK - 1 || t.nnodes < 0 || t.chunk_nodes <= 0) return path + ": bad header";
    const int K = t.order + 1;
    if (rec != (int64_t)(32 + sizeof(double) * 8 * (size_t)K * K * K)) return path + ": bad record size";
    if (t.nchunks != (int32_t)((t.nnodes + t.chunk_nodes - 1) / t.chunk_nodes)) return path + ": bad chunk count";
    info = t;
    if (!depth && !lr && !coefs) return std::string();
    std::vector<unsigned char> r((size_t)rec);
    for (int64_t i = 0; i < t.nnodes; i++) {
        if (std::fread(r.data(), 1, (size_t)rec, f.get()) != (size_t)rec) return path + ": truncated node records";
        if (depth) depth[i] = get<int32_t>(r.data(), 0);
        if (lr)
            for (int q = 0; q < 3; q++) lr[3 * i + q] = get<int64_t>(r.data(), 8 + 8 * q);
        if (coefs) std::memcpy(coefs + (size_t)i * 8 * K * K * K, r.data() + 32, (size_t)rec - 32);
    }
    unsigned char extra;
    if (std::fread(&extra, 1, 1, f.get()) != 0) return path + ": trailing bytes after the node records";
    return std::string();
}

}  // namespace mwg
