// HIP kernels of the MI355X multiwavelet apply (gfx950, wave64).
//
// Coefficient layout of a node: 8 components x K^3 FP64 (K = k+1), component bit
// 0 = x, 1 = y, 2 = z (0 = scaling), index (iz*K + iy)*K + ix inside a component.
// Two-scale transforms work on the interleaved (2K)^3 tensor whose per-dimension
// index is bit*K + i; each 1D pass multiplies every line of 2K values in place.
#include "mwg_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <climits>
#include <cmath>

namespace mwg {

// ----------------------------------------------------------------- device basis

__device__ inline void d_phi(const DevBasis &B, double x, double *out) {
    const int K = B.K;
    double l[MAXK];
    double t = 2.0 * x - 1.0, p0 = 1.0, p1 = t;
    l[0] = 1.0;
    if (K > 1) l[1] = sqrt(3.0) * t;
    for (int n = 2; n < K; n++) {
        double p2 = ((2.0 * n - 1.0) * t * p1 - (n - 1.0) * p0) / n;
        p0 = p1;
        p1 = p2;
        l[n] = sqrt(2.0 * n + 1.0) * p2;
    }
    for (int i = 0; i < K; i++) {
        double s = 0.0;
        for (int j = 0; j < K; j++) s += B.U[i * K + j] * l[j];
        out[i] = s;
    }
}

// global-address-space views: uniform loads through these become scalar loads
typedef const double __attribute__((address_space(1))) gdouble;
typedef const float __attribute__((address_space(1))) gfloat;

__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum over the block of one value per thread, same order every call (all threads get it)
__device__ inline double block_sum_256(double v, double *wsum) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    return (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// ----------------------------------------------------------------- operator blocks

// Scaling-scaling block of one unit Gaussian exp(-b (delta + u - v)^2) at cell scale m:
// t[i][j] = pre * int_{-1}^{1} exp(-b (delta+z)^2) C_ij(z) dz,
// C_ij(z) = int phi_i(u) phi_j(u - z) du (K-point rule on the overlap of the two cells),
// outer integral by 24-point Gauss-Legendre panels of width <= 0.5/sqrt(b) around the peak.
__global__ void k_tblock(int n, const int *item_term, const long long *item_delta, const double *bvals, double pre,
                         DevBasis B, const double *gx, const double *gw, double *out) {
    const int it = blockIdx.x;
    if (it >= n) return;
    const int K = B.K, K2 = K * K, tid = threadIdx.x;
    __shared__ double pu[MAXK * MAXK], pv[MAXK * MAXK], wq[MAXK];
    const double b = bvals[item_term[it]];
    const double delta = (double)item_delta[it];
    const double zc = -delta;
    const double R = sqrt(GAUSS_CUT_D / b);
    const double pw = (b > 4.0) ? 0.5 / sqrt(b) : 0.25;
    double acc = 0.0;
    for (int half = 0; half < 2; half++) {
        const double lo = half == 0 ? -1.0 : 0.0, hi = half == 0 ? 0.0 : 1.0;
        const double a0 = fmax(lo, zc - R), a1 = fmin(hi, zc + R);
        if (a1 <= a0) continue;
        int np = (int)ceil((a1 - a0) / pw);
        if (np < 1) np = 1;
        const double ph = (a1 - a0) / np;
        for (int p = 0; p < np; p++) {
            const double p0 = a0 + p * ph;
            for (int q = 0; q < 24; q++) {
                const double z = p0 + ph * gx[q];
                const double g = exp(-b * (delta + z) * (delta + z)) * ph * gw[q];
                if (g == 0.0) continue;
                const double ulo = fmax(0.0, z), uhi = fmin(1.0, 1.0 + z);
                __syncthreads();
                if (tid < K) {
                    double v[MAXK];
                    d_phi(B, ulo + (uhi - ulo) * B.qx[tid], v);
                    for (int i = 0; i < K; i++) pu[tid * K + i] = v[i];
                    wq[tid] = (uhi - ulo) * B.qw[tid];
                } else if (tid < 2 * K) {
                    const int q2 = tid - K;
                    double v[MAXK];
                    d_phi(B, ulo + (uhi - ulo) * B.qx[q2] - z, v);
                    for (int i = 0; i < K; i++) pv[q2 * K + i] = v[i];
                }
                __syncthreads();
                if (tid < K2) {
                    const int i = tid / K, j = tid % K;
                    double c = 0.0;
                    for (int qq = 0; qq < K; qq++) c += wq[qq] * pu[qq * K + i] * pv[qq * K + j];
                    acc += g * c;
                }
            }
        }
    }
    if (tid < K2) out[(size_t)it * K2 + tid] = acc * pre;
}


// Same block by moments: t[i][j] = pre * sum_h sum_p ccoef[h][i][j][p] M_h[p] with
// M_h[p] = int_half exp(-b (delta+z)^2) L_p(t(z)) dz on the same Gauss-Legendre panels.
// One wave per block: lanes take quadrature points independently and keep the 2 x 2K
// moments in registers; one wave reduction, then the K^2 entries.
template <int K>
__global__ void __launch_bounds__(64) k_tblock_mom(int n, const int *item_term, const long long *item_delta,
                                                   const double *bvals, double pre, const double *ccoef,
                                                   const double *gx, const double *gw, double *out) {
    const int it = blockIdx.x;
    if (it >= n) return;
    constexpr int P = 2 * K, K2 = K * K;
    __shared__ double Ms[2 * P];
    const double b = bvals[item_term[it]];
    const double delta = (double)item_delta[it];
    const double zc = -delta;
    const double R = sqrt(GAUSS_CUT_D / b);
    const double pw = (b > 4.0) ? 0.5 / sqrt(b) : 0.25;
    double M[2][P];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int p = 0; p < P; p++) M[h][p] = 0.0;
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const double lo = half == 0 ? -1.0 : 0.0, hi = half == 0 ? 0.0 : 1.0;
        const double a0 = fmax(lo, zc - R), a1 = fmin(hi, zc + R);
        if (a1 <= a0) continue;
        int np = (int)ceil((a1 - a0) / pw);
        if (np < 1) np = 1;
        const double ph = (a1 - a0) / np;
        for (int e = threadIdx.x; e < np * 24; e += 64) {
            const int pnl = e / 24, q = e % 24;
            const double z = a0 + pnl * ph + ph * gx[q];
            const double g = exp(-b * (delta + z) * (delta + z)) * ph * gw[q];
            const double x = 2.0 * (z - lo) - 1.0;
            double p0 = 1.0, p1 = x;
            M[half][0] += g;
            if (P > 1) M[half][1] += g * sqrt(3.0) * x;
#pragma unroll
            for (int k = 2; k < P; k++) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
                M[half][k] += g * sqrt(2.0 * k + 1.0) * p2;
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int p = 0; p < P; p++) {
            double v = M[h][p];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (threadIdx.x == 0) Ms[h * P + p] = v;
        }
    __syncthreads();
    for (int e = threadIdx.x; e < K2; e += 64) {
        double t = 0.0;
        for (int h = 0; h < 2; h++) {
            const double *c = ccoef + ((size_t)h * K2 + e) * P;
            for (int p = 0; p < P; p++) t += c[p] * Ms[h * P + p];
        }
        out[(size_t)it * K2 + e] = t * pre;
    }
}

// NS blocks from the child-scale scaling blocks: R = W M W^T with
// M = [[t(2d), t(2d-1)], [t(2d+1), t(2d)]]; type = 2*gt_bit + ft_bit.  The 4 block norms of
// each translation go to norms and, when given, to norms2.
__global__ void k_nsblocks(int n, const long long *item_delta, const int *item_row0, const int *item_tlo,
                           const int *item_thi, const double *t1, const double *W, int K, double *blocks,
                           double *norms, double *norms2) {
    const int it = blockIdx.x;
    if (it >= n) return;
    const int N = 2 * K, K2 = K * K, tid = threadIdx.x, nt = blockDim.x;
    __shared__ double Ms[4 * MAXK * MAXK], R1[4 * MAXK * MAXK], Ws[4 * MAXK * MAXK];
    __shared__ double red[4][256];
    const long long d = item_delta[it];
    const int row0 = item_row0[it], tlo = item_tlo[it], thi = item_thi[it];
    for (int e = tid; e < N * N; e += nt) Ws[e] = W[e];
    for (int e = tid; e < N * N; e += nt) {
        const int r = e / N, c = e % N;
        const int a = r / K, i = r % K, bb = c / K, j = c % K;
        const long long dd = 2 * d + a - bb;
        double v = 0.0;
        if (dd >= tlo && dd <= thi) v = t1[(size_t)(row0 + (dd - tlo)) * K2 + i * K + j];
        Ms[e] = v;
    }
    __syncthreads();
    for (int e = tid; e < N * N; e += nt) {
        const int i = e / N, j = e % N;
        double s = 0.0;
        for (int l = 0; l < N; l++) s += Ws[i * N + l] * Ms[l * N + j];
        R1[e] = s;
    }
    __syncthreads();
    double part[4] = {0, 0, 0, 0};
    for (int e = tid; e < N * N; e += nt) {
        const int i = e / N, j = e % N;
        double s = 0.0;
        for (int l = 0; l < N; l++) s += R1[i * N + l] * Ws[j * N + l];
        const int gb = i / K, a = i % K, fb = j / K, i2 = j % K;
        const int type = 2 * gb + fb;
        blocks[((size_t)it * 4 + type) * K2 + a * K + i2] = s;
        part[type] += s * s;
    }
    for (int q = 0; q < 4; q++) red[q][tid] = part[q];
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int q = 0; q < 4; q++) red[q][tid] += red[q][tid + s];
        __syncthreads();
    }
    if (tid < 4) {
        const double v = sqrt(red[tid][0]);
        norms[(size_t)it * 4 + tid] = v;
        if (norms2) norms2[(size_t)it * 4 + tid] = v;
    }
}

// ----------------------------------------------------------------- tensor line passes

// (2K)^3 node tensor in LDS with the x rows padded to N + 1 doubles: x-lines (stride 1)
// of neighbouring lanes then start 2(N+1) banks apart instead of 2N (no conflicts)
template <int K>
struct PadT {
    static constexpr int N = 2 * K, P = N + 1, PL = N * P, SIZE = N * PL;
};

template <int K>
__device__ inline int padded_index(int c, int e) {
    constexpr int N = 2 * K, P = N + 1;
    const int x = e % K, y = (e / K) % K, z = e / (K * K);
    const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
    return ((bz * K + z) * N + (by * K + y)) * P + bx * K + x;
}

// in-place N x N filter on every line of the padded tensor along dim (0 = x)
template <int N>
__device__ inline void line_pass_padded(double *T, const double *M, int dim) {
    constexpr int P = N + 1, N2 = N * N;
    for (int line = threadIdx.x; line < N2; line += blockDim.x) {
        const int u = line % N, v = line / N;
        int base, stride;
        if (dim == 0) {
            base = line * P;
            stride = 1;
        } else if (dim == 1) {
            base = v * N * P + u;
            stride = P;
        } else {
            base = v * P + u;
            stride = N * P;
        }
        double in[N];
#pragma unroll
        for (int q = 0; q < N; q++) in[q] = T[base + q * stride];
        // (output rows two at a time: a full unroll would keep all N^2 matrix entries live)
#pragma unroll 2
        for (int p = 0; p < N; p++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < N; q++) acc += M[p * N + q] * in[q];
            T[base + p * stride] = acc;
        }
    }
}

// component norms of listed nodes (one wave per node)
__global__ void k_norms(int n, const int *nodes, const double *coef, double *norm, int K3) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const int node = nodes ? nodes[i] : i;
    const double *p = coef + (size_t)node * 8 * K3;
    for (int c = 0; c < 8; c++) {
        double s = 0.0;
        for (int e = threadIdx.x; e < K3; e += 64) s += p[c * K3 + e] * p[c * K3 + e];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (threadIdx.x == 0) norm[(size_t)node * 8 + c] = sqrt(s);
    }
}

// per-pair partial dot products: sum over components [c0, 8)
__global__ void k_dot(int n, const int *ia, const int *ib, const int *c0, const double *A, const double *B, int K3,
                      double *part) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const double *a = A + (size_t)ia[i] * 8 * K3;
    const double *b = B + (size_t)ib[i] * 8 * K3;
    double s = 0.0;
    for (int e = c0[i] * K3 + threadIdx.x; e < 8 * K3; e += 64) s += a[e] * b[e];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) part[i] = s;
}

// node-pair dot products over a fixed grid, one wave per pair (a pair word holds the a node
// in bits 0..30, a flag to skip component 0 in bit 31 and the b node in bits 32..63); every
// block writes its sum to part[blockIdx.x], so part needs no clearing
__global__ void __launch_bounds__(256) k_dot_pairs(int n, const unsigned long long *pairs, const double *A,
                                                   const double *B, int K3, double *part) {
    __shared__ double red[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int i = (int)blockIdx.x * 4 + wave; i < n; i += (int)gridDim.x * 4) {
        const unsigned long long p = pairs[i];
        const int ia = (int)(p & 0x7fffffffu), c0 = (int)((p >> 31) & 1u), ib = (int)(p >> 32);
        const double *a = A + (size_t)ia * 8 * K3;
        const double *b = B + (size_t)ib * 8 * K3;
        for (int e = c0 * K3 + lane; e < 8 * K3; e += 64) s += a[e] * b[e];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// one wave: *res = sum of part[0 .. n)
__global__ void __launch_bounds__(64) k_sum_parts(int n, const double *part, double *res) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) s += part[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) *res = s;
}

__global__ void k_scale(size_t n, double *x, double c) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= c;
}

// ----------------------------------------------------------------- projection

// Gaussian sums are separable: the children's scaling tensor is a sum of outer
// products of 2K-vectors (one per dimension), and so is its compression.
template <int K>
__global__ void __launch_bounds__(256) k_project_gauss(int nw, const int *work, const int *depth, const int *lr,
                                                       ProjGeom geo, DevBasis B, int nterm, const double *coef,
                                                       const double *alpha, const double *center,
                                                       const double *poly, double *out) {
    constexpr int N = 2 * K, K3 = K * K * K;
    constexpr int LPER = (N * N + 255) / 256;  // x-lines (y, z) per thread
    constexpr int CH = 8;
    __shared__ double P[CH][3][N], WP[CH][3][N], Ws[N * N], cf[CH];
    const int node = work[blockIdx.x];
    const int d = depth[node];
    const double hc = geo.sfac * ldexp(1.0, -(geo.n0 + d + 1));
    const double pre = sqrt(hc);
    const int tid = threadIdx.x;
    for (int e = tid; e < N * N; e += 256) Ws[e] = B.W[e];
    double acc[LPER][N];
#pragma unroll
    for (int r = 0; r < LPER; r++)
#pragma unroll
        for (int x = 0; x < N; x++) acc[r][x] = 0.0;
    for (int t0 = 0; t0 < nterm; t0 += CH) {
        __syncthreads();
        for (int e = tid; e < CH * 3 * N; e += 256) {
            const int tt = e / (3 * N), dim = (e / N) % 3, p = e % N;
            const int bit = p / K, i = p % K;
            const int t = t0 + tt;
            double s = 0.0;
            if (t < nterm) {
                const double x0 = geo.lo[dim] + hc * (double)(2 * lr[3 * node + dim] + bit);
                const double c = center[3 * t + dim];
                for (int q = 0; q < K; q++) {
                    const double u = x0 + hc * B.qx[q] - c;
                    double pv = 0.0, up = 1.0;
                    for (int ee = 0; ee < 5; ee++) {
                        pv += poly[15 * t + 5 * dim + ee] * up;
                        up *= u;
                    }
                    s += B.phiq[q * K + i] * (pv * exp(-alpha[t] * u * u) * B.qw[q]);
                }
                s *= pre;
            }
            P[tt][dim][p] = s;
            if (e < CH) cf[e] = (t0 + e < nterm) ? coef[t0 + e] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < CH * 3 * N; e += 256) {
            const int tt = e / (3 * N), dim = (e / N) % 3, p = e % N;
            double s = 0.0;
            for (int q = 0; q < N; q++) s += Ws[p * N + q] * P[tt][dim][q];
            WP[tt][dim][p] = s;
        }
        __syncthreads();
        // the thread's x-lines accumulate the outer products term by term
#pragma unroll 1
        for (int tt = 0; tt < CH; tt++) {
#pragma unroll
            for (int r = 0; r < LPER; r++) {
                const int line = tid + 256 * r;
                if (line < N * N) {
                    const int py = line % N, pz = line / N;
                    const double yz = cf[tt] * WP[tt][2][pz] * WP[tt][1][py];
#pragma unroll
                    for (int x = 0; x < N; x++) acc[r][x] += yz * WP[tt][0][x];
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < LPER; r++) {
        const int line = tid + 256 * r;
        if (line < N * N) {
            const int py = line % N, pz = line / N;
#pragma unroll
            for (int px = 0; px < N; px++) {
                const int c = (px / K) | ((py / K) << 1) | ((pz / K) << 2);
                const int idx = ((pz % K) * K + (py % K)) * K + (px % K);
                out[(size_t)node * 8 * K3 + c * K3 + idx] = acc[r][px];
            }
        }
    }
    (void)nw;
}

__device__ double d_hydrogen_orb(int N, int L, int M, double Z, double x, double y, double z) {
    const double r = sqrt(x * x + y * y + z * z);
    double c0;
    if (N == 1 && L == 0) c0 = 2.0;
    else if (N == 2 && L == 0) c0 = 1.0 / (2.0 * sqrt(2.0));
    else if (N == 3 && L == 0) c0 = 1.0 / (9.0 * sqrt(3.0));
    else if (N == 2 && L == 1) c0 = 1.0 / (2.0 * sqrt(6.0));
    else if (N == 3 && L == 1) c0 = 1.0 / (9.0 * sqrt(6.0));
    else if (N == 3 && L == 2) c0 = 1.0 / (9.0 * sqrt(30.0));
    else return 0.0;
    c0 *= pow(Z, 1.5);
    const double rho = 2.0 * Z / N * r;
    double poly;
    if (N == 1 && L == 0) poly = 1.0;
    else if (N == 2 && L == 0) poly = 2.0 - rho;
    else if (N == 3 && L == 0) poly = 6.0 - 6.0 * rho + rho * rho;
    else if (N == 2 && L == 1) poly = rho;
    else if (N == 3 && L == 1) poly = rho * (4.0 - rho);
    else poly = rho * rho;
    const double Rr = c0 * poly * exp(-rho / 2.0);
    double ang, ca;
    if (L == 0) {
        ca = 1.0;
        ang = 1.0;
    } else if (L == 1) {
        ca = sqrt(3.0);
        const double q = (M == 0) ? x : (M == 1) ? y : z;
        ang = (r > 0.0) ? q / r : 0.0;
    } else {
        const double r2 = r * r;
        ca = (M <= 2) ? sqrt(15.0) : (M == 3 ? sqrt(15.0 / 4.0) : sqrt(5.0 / 4.0));
        if (r2 == 0.0) ang = 0.0;
        else if (M == 0) ang = x * y / r2;
        else if (M == 1) ang = x * z / r2;
        else if (M == 2) ang = y * z / r2;
        else if (M == 3) ang = (x * x - y * y) / r2;
        else ang = (2.0 * z * z - x * x - y * y) / r2;
    }
    return Rr * ca / (2.0 * sqrt(M_PI)) * ang;
}

// hydrogen functions are not separable: values at the K^3 quadrature points of
// each child, then three 1D quadrature passes, then compression
template <int K>
__global__ void __launch_bounds__(256) k_project_hyd(int nw, const int *work, const int *depth, const int *lr,
                                                     ProjGeom geo, DevBasis B, int nterm, const double *coef,
                                                     const int *q, double Z, const double *origin, double *out) {
    using L = PadT<K>;
    constexpr int N = L::N, P = L::P, N3 = N * N * N, K3 = K * K * K;
    extern __shared__ double smem[];
    double *T = smem;
    double *Ws = smem + L::SIZE;
    double *Qm = Ws + N * N;  // N x N block-diagonal quadrature matrix: phi_i(x_q) w_q * pre
    const int node = work[blockIdx.x];
    const int d = depth[node];
    const double hc = geo.sfac * ldexp(1.0, -(geo.n0 + d + 1));
    const double pre = sqrt(hc);
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        Ws[e] = B.W[e];
        const int p = e / N, qq = e % N;
        Qm[e] = ((p / K) == (qq / K)) ? B.phiq[(qq % K) * K + (p % K)] * B.qw[qq % K] * pre : 0.0;
    }
    for (int e = threadIdx.x; e < N3; e += blockDim.x) {
        const int px = e % N, py = (e / N) % N, pz = e / (N * N);
        const double x = geo.lo[0] + hc * ((double)(2 * lr[3 * node] + px / K) + B.qx[px % K]) - origin[0];
        const double y = geo.lo[1] + hc * ((double)(2 * lr[3 * node + 1] + py / K) + B.qx[py % K]) - origin[1];
        const double z = geo.lo[2] + hc * ((double)(2 * lr[3 * node + 2] + pz / K) + B.qx[pz % K]) - origin[2];
        double v = 0.0;
        for (int t = 0; t < nterm; t++) {
            double a = d_hydrogen_orb(q[6 * t], q[6 * t + 1], q[6 * t + 2], Z, x, y, z);
            if (q[6 * t + 3] > 0) a *= d_hydrogen_orb(q[6 * t + 3], q[6 * t + 4], q[6 * t + 5], Z, x, y, z);
            v += coef[t] * a;
        }
        T[(pz * N + py) * P + px] = v;
    }
    __syncthreads();
    for (int dim = 0; dim < 3; dim++) {
        line_pass_padded<N>(T, Qm, dim);
        __syncthreads();
    }
    for (int dim = 0; dim < 3; dim++) {
        line_pass_padded<N>(T, Ws, dim);
        __syncthreads();
    }
    for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x)
        out[(size_t)node * 8 * K3 + e] = T[padded_index<K>(e / K3, e % K3)];
    (void)nw;
}

// ----------------------------------------------------------------- apply: generated nodes

__device__ inline bool in_box(const short *b, int dx, int dy, int dz) {
    return b[0] >= 0 && abs(dx) <= b[0] && abs(dy) <= b[1] && abs(dz) <= b[2];
}

// bit c (= gt * 8 + ft) set iff the screening box of combo c holds (dx, dy, dz)
__device__ inline unsigned long long combo_mask(const TermTab &tb, int dx, int dy, int dz) {
    const int ax = abs(dx), ay = abs(dy), az = abs(dz);
    if (ax < BM && ay < BM && az < BM) return tb.bm[0][ax] & tb.bm[1][ay] & tb.bm[2][az];
    unsigned long long m = 0;
    for (int c = 0; c < 64; c++)
        if (in_box(tb.box[c], dx, dy, dz)) m |= 1ull << c;
    return m;
}
constexpr unsigned long long FT0_COMBOS = 0x0101010101010101ull;  // the combos with ft = 0

// mark the input nodes that the output nodes need at this depth but the input
// tree does not have (scaling-only "generated" nodes)
// translations cached per output node in LDS (the union box over all terms, cells)
constexpr int FCACHE = 4096;

__device__ inline void gen_insert(const JobTab &J, const GenState &gs, int job, uint64_t key) {
    const unsigned long long kk = (unsigned long long)key + 1ull;
    uint32_t i = dkey_hash(key) & J.genHash.mask;
    for (uint32_t probe = 0; probe <= J.genHash.mask; probe++) {
        unsigned long long old = atomicCAS(&J.genHash.keys[i], 0ull, kk);
        if (old == 0ull) {
            const int idx = atomicAdd(gs.count, 1);
            atomicAdd((unsigned long long *)&gs.tabcnt[4 * job + 3], 1ull);
            if (idx < gs.cap) {
                gs.keys[idx] = key;
                gs.job[idx] = job;
            } else {
                atomicOr(gs.overflow, 1);
            }
            J.genHash.vals[i] = idx;
            return;
        }
        if (old == kk) return;
        i = (i + 1) & J.genHash.mask;
        if (probe == J.genHash.mask) atomicOr(gs.overflow, 2);
    }
}

// Walks the (term, delta) candidates of the ft = 0 boxes; the translations that some
// term needs are first collected in an LDS bitmap over the job's union box, so that
// each input position is probed once (candidates of different terms share them).
__device__ __forceinline__ void gen_mark_node(int item, const WorkItem *work, const JobTab *jobs,
                                              const TermTab *tabs, const GenState &gs) {
    const WorkItem wi = work[item];
    const JobTab &J = jobs[wi.job];
    __shared__ int start[257];
    __shared__ unsigned need[FCACHE / 32];
    for (int t = threadIdx.x; t <= J.ntab; t += blockDim.x) start[t] = J.pre0[t];
    const int ux = J.ubox[0], uy = J.ubox[1], uz = J.ubox[2];
    const int nx = 2 * ux + 1, ny = 2 * uy + 1, nz = 2 * uz + 1;
    const bool cached = ux >= 0 && nx * ny * nz <= FCACHE;
    for (int q = threadIdx.x; q < FCACHE / 32; q += blockDim.x) need[q] = 0u;
    __syncthreads();
    const int total = start[J.ntab];
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        int lo = 0, hi = J.ntab - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const TermTab &tb = tabs[J.tab0 + lo];
        const int c = k - start[lo];
        const int vx = tb.ub0[0], vy = tb.ub0[1], vz = tb.ub0[2];
        const int dx = c % (2 * vx + 1) - vx;
        const int dy = (c / (2 * vx + 1)) % (2 * vy + 1) - vy;
        const int dz = c / ((2 * vx + 1) * (2 * vy + 1)) - vz;
        const int fx = wi.l[0] - dx, fy = wi.l[1] - dy, fz = wi.l[2] - dz;
        if (fx < 0 || fy < 0 || fz < 0 || fx >= J.lmax[0] || fy >= J.lmax[1] || fz >= J.lmax[2]) continue;
        if (!(combo_mask(tb, dx, dy, dz) & FT0_COMBOS)) continue;
        if (cached) {
            const int cell = ((dz + uz) * ny + (dy + uy)) * nx + (dx + ux);
            atomicOr(&need[cell >> 5], 1u << (cell & 31));
            continue;
        }
        const uint64_t key = dnode_key(J.depth, fx, fy, fz);
        if (dhash_find(J.inHash, key) < 0) gen_insert(J, gs, wi.job, key);
    }
    if (!cached) return;
    __syncthreads();
    for (int cell = threadIdx.x; cell < nx * ny * nz; cell += blockDim.x) {
        if (!((need[cell >> 5] >> (cell & 31)) & 1u)) continue;
        const int dx = cell % nx - ux, dy = (cell / nx) % ny - uy, dz = cell / (nx * ny) - uz;
        const uint64_t key = dnode_key(J.depth, wi.l[0] - dx, wi.l[1] - dy, wi.l[2] - dz);
        if (dhash_find(J.inHash, key) < 0) gen_insert(J, gs, wi.job, key);
    }
}

// Items 0 .. total - 1 of a workgroup-per-item kernel: item blockIdx.x first, then
// grid-stride, or (queue) the items beyond the grid taken one at a time from a zeroed
// counter, so that workgroups that drew cheap items take more of them.  Workgroups without
// an item leave at once; the others leave once the counter has passed total.
template <class F>
__device__ __forceinline__ void for_items(int total, int *queue, F f) {
    if ((int)blockIdx.x >= total) return;
    f((int)blockIdx.x);
    if (total <= (int)gridDim.x) return;
    __syncthreads();  // (LDS reused by the next item)
    if (!queue) {
        for (int b = blockIdx.x + gridDim.x; b < total; b += gridDim.x) {
            f(b);
            __syncthreads();
        }
        return;
    }
    __shared__ int s_item;
    while (true) {
        if (threadIdx.x == 0) s_item = (int)gridDim.x + atomicAdd(queue, 1);
        __syncthreads();
        const int b = s_item;
        __syncthreads();
        if (b >= total) break;
        f(b);
        __syncthreads();
    }
}

// one workgroup per work node (nwp: the node count on the device)
__global__ void __launch_bounds__(256) k_gen_mark(int nw, const int *nwp, int *queue, const WorkItem *work,
                                                  const JobTab *jobs, const TermTab *tabs, GenState gs) {
    const int n = nwp ? *nwp : nw;
    for_items(n, queue, [&](int i) { gen_mark_node(i, work, jobs, tabs, gs); });
}

// Refines the node tensor in T (padded (2K)^3, all 8 components of an ancestor) down
// `steps` levels along the path bits of l (bit steps-1 first): the first step uses
// the scaling and wavelet parts (2K -> K per line), the others the scaling part only.
// The scaling block of the descendant is left at T[(z N + y) P + x], x, y, z < K.
template <int K, class IT>
__device__ void descend_padded(double *T, const double *Ws, const IT *l, int steps) {
    using L = PadT<K>;
    constexpr int N = L::N, P = L::P;
    // first step: child a of the (s, d) parent; line (2K) -> K values, s_j = sum_p W[p][a*K+j] in_p
    {
        const int sh = steps - 1;
        const int ax = (int)((l[0] >> sh) & 1), ay = (int)((l[1] >> sh) & 1), az = (int)((l[2] >> sh) & 1);
        // x: lines (z, y) over all N x N
        for (int line = threadIdx.x; line < N * N; line += blockDim.x) {
            const int base = line * P;
            double in[N];
#pragma unroll
            for (int p = 0; p < N; p++) in[p] = T[base + p];
#pragma unroll
            for (int j = 0; j < K; j++) {
                double acc = 0.0;
#pragma unroll
                for (int p = 0; p < N; p++) acc += Ws[p * N + ax * K + j] * in[p];
                T[base + j] = acc;
            }
        }
        __syncthreads();
        // y: lines (z, x < K)
        for (int line = threadIdx.x; line < N * K; line += blockDim.x) {
            const int z = line / K, x = line % K;
            const int base = z * N * P + x;
            double in[N];
#pragma unroll
            for (int p = 0; p < N; p++) in[p] = T[base + p * P];
#pragma unroll
            for (int j = 0; j < K; j++) {
                double acc = 0.0;
#pragma unroll
                for (int p = 0; p < N; p++) acc += Ws[p * N + ay * K + j] * in[p];
                T[base + j * P] = acc;
            }
        }
        __syncthreads();
        // z: lines (y < K, x < K)
        for (int line = threadIdx.x; line < K * K; line += blockDim.x) {
            const int y = line / K, x = line % K;
            const int base = y * P + x;
            double in[N];
#pragma unroll
            for (int p = 0; p < N; p++) in[p] = T[base + p * N * P];
#pragma unroll
            for (int j = 0; j < K; j++) {
                double acc = 0.0;
#pragma unroll
                for (int p = 0; p < N; p++) acc += Ws[p * N + az * K + j] * in[p];
                T[base + j * N * P] = acc;
            }
        }
        __syncthreads();
    }
    // further steps: scaling-only descent, K x K filter per dimension
    for (int sh = steps - 2; sh >= 0; sh--) {
        const int a3[3] = {(int)((l[0] >> sh) & 1), (int)((l[1] >> sh) & 1), (int)((l[2] >> sh) & 1)};
        for (int dim = 0; dim < 3; dim++) {
            for (int line = threadIdx.x; line < K * K; line += blockDim.x) {
                const int u = line % K, v = line / K;
                int base, stride;
                if (dim == 0) {
                    base = (v * N + u) * P;
                    stride = 1;
                } else if (dim == 1) {
                    base = v * N * P + u;
                    stride = P;
                } else {
                    base = v * P + u;
                    stride = N * P;
                }
                double in[K];
#pragma unroll
                for (int p = 0; p < K; p++) in[p] = T[base + p * stride];
#pragma unroll
                for (int j = 0; j < K; j++) {
                    double acc = 0.0;
#pragma unroll
                    for (int p = 0; p < K; p++) acc += Ws[p * N + a3[dim] * K + j] * in[p];
                    T[base + j * stride] = acc;
                }
            }
            __syncthreads();
        }
    }
}

// scaling block of a generated node: reconstruct from the leaf ancestor (using its
// wavelets for the first step), then refine with the low-pass filters only
template <int K>
__device__ void gen_one(int g, const JobTab *jobs, const GenState &gs, double *T, const double *Ws) {
    using L = PadT<K>;
    constexpr int N = L::N, P = L::P, K3 = K * K * K;
    const uint64_t key = gs.keys[g];
    const JobTab &J = jobs[gs.job[g]];
    const int d = (int)(key >> 57);
    const int64_t l[3] = {(int64_t)((key >> 38) & 0x7FFFF), (int64_t)((key >> 19) & 0x7FFFF), (int64_t)(key & 0x7FFFF)};
    // the deepest ancestor present in the input tree: one thread per level probes it
    __shared__ int s_anc, s_ad;
    if (threadIdx.x == 0) {
        s_anc = -1;
        s_ad = -1;
    }
    __syncthreads();
    const int my_dd = d - 1 - (int)threadIdx.x;
    int my_a = -1;
    if (my_dd >= 0) {
        const int sh = d - my_dd;
        my_a = dhash_find(J.inHash, dnode_key(my_dd, l[0] >> sh, l[1] >> sh, l[2] >> sh));
        if (my_a >= 0) atomicMax(&s_ad, my_dd);
    }
    __syncthreads();
    if (my_a >= 0 && my_dd == s_ad) s_anc = my_a;
    __syncthreads();
    const int anc = s_anc, ad = s_ad;
    double *outp = gs.pool + (size_t)g * K3;
    if (anc < 0) {
        for (int e = threadIdx.x; e < K3; e += blockDim.x) outp[e] = 0.0;
        if (threadIdx.x == 0) gs.norm[g] = 0.0;
        __syncthreads();
        return;
    }
    const double *src = J.inCoef + (size_t)anc * 8 * K3;
    for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x) T[padded_index<K>(e / K3, e % K3)] = src[e];
    __syncthreads();
    descend_padded<K>(T, Ws, l, d - ad);
    __shared__ double wsum[4];
    double s2 = 0.0;
    for (int e = threadIdx.x; e < K3; e += blockDim.x) {
        const int x = e % K, y = (e / K) % K, z = e / (K * K);
        const double v = T[(z * N + y) * P + x];
        outp[e] = v;
        s2 += v * v;
    }
    s2 = block_sum_256(s2, wsum);
    if (threadIdx.x == 0) gs.norm[g] = sqrt(s2);
    __syncthreads();
}

// Generated nodes first .. min(count, cap) - 1, count read from the device counter
// that k_gen_mark advanced (no host round trip); grid-stride over the range.
template <int K>
__global__ void __launch_bounds__(256) k_gen_compute(int first, const int *firstp, int cap, const JobTab *jobs,
                                                     GenState gs, const double *W) {
    using L = PadT<K>;
    constexpr int N = L::N;
    extern __shared__ double smem[];
    double *T = smem;
    double *Ws = smem + L::SIZE;
    if (firstp) first = *firstp;
    const int end = min(*(volatile const int *)gs.count, cap);
    if (first + (int)blockIdx.x >= end) return;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) Ws[e] = W[e];
    __syncthreads();
    for (int g = first + blockIdx.x; g < end; g += gridDim.x) gen_one<K>(g, jobs, gs, T, Ws);
}


// Per output node, the input node (>= 0), generated node (-(g+1)) or nothing
// (INT_MIN) at every translation of the union box over all terms, cached in LDS so
// that each term's loop reads the cache instead of probing the hash tables again.
struct FCache {
    int ux, uy, uz;  // half-widths, ux < 0: no cache (box too large)
};

__device__ inline int probe_f(const JobTab &J, const GenState &gs, int fx, int fy, int fz) {
    if (fx < 0 || fy < 0 || fz < 0 || fx >= J.lmax[0] || fy >= J.lmax[1] || fz >= J.lmax[2]) return INT_MIN;
    const uint64_t key = dnode_key(J.depth, fx, fy, fz);
    const int f = dhash_find(J.inHash, key);
    if (f >= 0) return f;
    const int gi = dhash_find(J.genHash, key);
    return gi >= 0 ? -(gi + 1) : INT_MIN;
}

__device__ inline FCache fill_fcache(int *cache, const WorkItem &wi, const JobTab &J, const TermTab *tabs,
                                     const GenState &gs) {
    const int ux = J.ubox[0], uy = J.ubox[1], uz = J.ubox[2];
    FCache fc{ux, uy, uz};
    (void)tabs;
    if (ux < 0) return fc;
    const int nx = 2 * ux + 1, ny = 2 * uy + 1, nz = 2 * uz + 1;
    if (nx * ny * nz > FCACHE) {
        fc.ux = -1;
        return fc;
    }
    for (int c = threadIdx.x; c < nx * ny * nz; c += blockDim.x) {
        const int dx = c % nx - ux, dy = (c / nx) % ny - uy, dz = c / (nx * ny) - uz;
        cache[c] = probe_f(J, gs, wi.l[0] - dx, wi.l[1] - dy, wi.l[2] - dz);
    }
    __syncthreads();
    return fc;
}

__device__ inline int lookup_f(const int *cache, const FCache &fc, const WorkItem &wi, const JobTab &J,
                               const GenState &gs, int dx, int dy, int dz) {
    if (fc.ux >= 0)
        return cache[((dz + fc.uz) * (2 * fc.uy + 1) + (dy + fc.uy)) * (2 * fc.ux + 1) + (dx + fc.ux)];
    return probe_f(J, gs, wi.l[0] - dx, wi.l[1] - dy, wi.l[2] - dz);
}


// The screening of one candidate (term, delta, input node): bit gt of surv[ft] is set when
// the combination is inside its screening box (bit 8 gt + ft of cm) and its bound
// (((|w| ||X_tx||) ||Y_ty||) ||Z_tz||) ||f_ft|| exceeds thr; maxb[ft] is the largest such
// bound.  The products are formed in that order, as in the oracle's screening, so the
// decisions are the same to the bit.  Unrolled and free of branches: the lanes of a wave
// hold different candidates.
template <bool MAXB>
__device__ inline void screen_candidate(double aw, const double *nx, const double *ny, const double *nz,
                                        const double *fn, unsigned long long cm, double thr, unsigned (&surv)[8],
                                        double (&maxb)[8]) {
    double pxy[4][4];
#pragma unroll
    for (int tx = 0; tx < 4; tx++) {
        const double ax = aw * nx[tx];
#pragma unroll
        for (int ty = 0; ty < 4; ty++) pxy[tx][ty] = ax * ny[ty];
    }
#pragma unroll
    for (int ft = 0; ft < 8; ft++) {
        unsigned sv = 0u;
        double mb = 0.0;
#pragma unroll
        for (int gt = 0; gt < 8; gt++) {
            const int tx = 2 * (gt & 1) + (ft & 1);
            const int ty = 2 * ((gt >> 1) & 1) + ((ft >> 1) & 1);
            const int tz = 2 * ((gt >> 2) & 1) + ((ft >> 2) & 1);
            const double b = (pxy[tx][ty] * nz[tz]) * fn[ft];
            const bool ok = ((cm >> (8 * gt + ft)) & 1ull) && b > thr;
            sv |= (unsigned)ok << gt;
            if (MAXB) mb = ok ? fmax(mb, b) : mb;
        }
        surv[ft] = sv;
        maxb[ft] = mb;
    }
}

// Flattened candidate walk: every (term, delta) cell of the job's screened boxes is
// one candidate, spread over all threads of the block (term prefix sums in LDS).
// For each candidate with an input node the visitor gets the 12 operator norms
// (4 block types per dimension) and the input component norms.
struct CandTerms {
    int start[257];
};

__device__ inline void term_prefix(CandTerms &ct, const JobTab &J, const TermTab *tabs) {
    for (int t = threadIdx.x; t <= J.ntab; t += blockDim.x) ct.start[t] = J.pre[t];
    __syncthreads();
    (void)tabs;
}

template <class Visit>
__device__ inline void walk_candidates(const CandTerms &ct, const int *fcache, const FCache &fc, const WorkItem &wi,
                                       const JobTab &J, const TermTab *tabs, const GenState &gs, int part, int nparts,
                                       Visit &&visit) {
    const int total = ct.start[J.ntab];
    // part of nparts workgroups that share this output node: a contiguous candidate range
    const int k0 = (int)((long long)total * part / nparts), k1 = (int)((long long)total * (part + 1) / nparts);
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        int lo = 0, hi = J.ntab - 1;  // last t with start[t] <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ct.start[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int t = lo;
        const TermTab &tb = tabs[J.tab0 + t];
        const int c = k - ct.start[t];
        const int ux = tb.ub[0], uy = tb.ub[1], uz = tb.ub[2];
        const int dx = c % (2 * ux + 1) - ux;
        const int dy = (c / (2 * ux + 1)) % (2 * uy + 1) - uy;
        const int dz = c / ((2 * ux + 1) * (2 * uy + 1)) - uz;
        const int f = lookup_f(fcache, fc, wi, J, gs, dx, dy, dz);
        if (f == INT_MIN) continue;
        double fn[8];
        int nft = 8;
        if (f >= 0) {
            for (int q = 0; q < 8; q++) fn[q] = J.inNorm[(size_t)f * 8 + q];
        } else {
            // an entry past the allocated pool (its count is redone after the pool grows)
            fn[0] = -f - 1 < gs.pcap ? gs.norm[-f - 1] : 0.0;
            for (int q = 1; q < 8; q++) fn[q] = 0.0;  // (no wavelet part: no bound above thr >= 0)
            nft = 1;
        }
        const double *nrm = J.operNorms + tb.nrm;
        double nx[4], ny[4], nz[4];
        for (int q = 0; q < 4; q++) {
            nx[q] = nrm[(dx + tb.band) * 4 + q];
            ny[q] = nrm[(dy + tb.band) * 4 + q];
            nz[q] = nrm[(dz + tb.band) * 4 + q];
        }
        visit(t, tb, dx, dy, dz, f, fn, nft, nx, ny, nz);
    }
}

// ----------------------------------------------------------------- apply: screening boxes

// One wavefront per (table, term); lane = combination gt*8+ft.  For the bound
// |w| F_ft prod_d ||X_d|| of the combination to exceed thr somewhere, each dimension's
// translation must stay inside the largest m with env_type[m] > thr / (|w| F_ft prod of
// the other two dimensions' largest norms); env is non-increasing in m.
__device__ inline int wave_max_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wave_max_double(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr int ENV_LDS = 512;  // envelope entries per type staged in LDS
// the tuple threshold of a device-loop job: the host loop's expression,
// SCREEN_REL prec sqrt(||g||^2 / terms), 0 before any norm
__device__ inline double loop_thr(const LoopJob &L, int nterms) {
    const double gsq = L.gsq;
    return gsq > 0.0 ? SCREEN_REL * L.prec * sqrt(gsq / (double)nterms) : 0.0;
}

__global__ void __launch_bounds__(64) k_job_boxes(int nbj, const BoxJob *bjs, TermTab *tabs, double *tbmax,
                                                  int max_terms, LoopInit li) {
    const int jb = blockIdx.x, t = blockIdx.y;
    if (jb >= nbj) return;
    const BoxJob &B = bjs[jb];
    if (t >= B.nterms) return;
    const double thr = li.lj ? loop_thr(li.lj[jb], B.nterms) : B.thr;
    const LevTerm &lt = B.terms[t];  // (read in place: maxn is indexed by the lane's types)
    TermTab *tb = tabs + B.tab0 + t;
    const int c = threadIdx.x, gt = c >> 3, ft = c & 7;
    // the term's norm envelopes (4 types x (band + 1)) staged in LDS for the searches
    __shared__ double senv[4 * ENV_LDS];
    const int ne = 4 * (lt.band + 1);
    const bool env_lds = ne <= 4 * ENV_LDS;
    if (env_lds)
        for (int e = c; e < ne; e += 64) senv[e] = B.env[lt.env + e];
    __syncthreads();
    const double *envb = env_lds ? senv : B.env + lt.env;
    int m3[3] = {-1, -1, -1};
    double full = 0.0;
    bool ok = false;
    if (lt.active && !(B.depth > 0 && c == 0)) {
        const double F = B.F[ft];
        if (F > 0.0) {
            int ty[3];
#pragma unroll
            for (int q = 0; q < 3; q++) ty[q] = 2 * ((gt >> q) & 1) + ((ft >> q) & 1);
            const double base = fabs(lt.w) * F;
            const double f3 = base * lt.maxn[ty[0]] * lt.maxn[ty[1]] * lt.maxn[ty[2]];
            if (f3 > thr) {
                full = f3;
                ok = true;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    double other = base;
#pragma unroll
                    for (int q2 = 0; q2 < 3; q2++)
                        if (q2 != q) other *= lt.maxn[ty[q2]];
                    const double tau = thr / other;
                    const double *env = envb + (long long)ty[q] * (lt.band + 1);
                    int m = -1;
                    if (env[0] > tau) {
                        int lo = 0, hi = lt.band;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) / 2;
                            if (env[mid] > tau) lo = mid;
                            else hi = mid - 1;
                        }
                        m = lo;
                    }
                    if (m < 0) ok = false;
                    m3[q] = m;
                }
            }
        }
    }
    int u[3], u0[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const short v = ok ? (short)m3[q] : (short)-1;
        tb->box[c][q] = v;
        u[q] = wave_max_int(v);
        u0[q] = wave_max_int(ft == 0 ? v : -1);
        for (int m = 0; m < BM; m++) {
            const unsigned long long b = __ballot(ok && m3[q] >= m);
            if (c == 0) tb->bm[q][m] = b;
        }
    }
    const double bmax = wave_max_double(full);
    if (c == 0) {
        tb->w = lt.w;
        tb->blk = lt.blk;
        tb->nrm = lt.nrm;
        tb->band = lt.band;
        for (int q = 0; q < 3; q++) {
            tb->ub[q] = (short)u[q];
            tb->ub0[q] = (short)u0[q];
        }
        tbmax[(size_t)jb * max_terms + t] = bmax;
    }
}

// Per table, one wave: the FP64 routing bound (mixed mode: slices whose bound exceeds
// it run in FP64; above the roots a fixed multiple of the screening threshold, on the
// roots (no screening) a fraction of the largest bound of the table), the union box of
// the terms, and the prefix sums over terms of the candidate counts of the union boxes
// ub and ub0 (the work-list and generated-node kernels split their walks by these).
__device__ inline int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

__global__ void __launch_bounds__(64) k_job_route(int nbj, const BoxJob *bjs, const TermTab *tabs,
                                                  const double *tbmax, int max_terms, JobTab *jobs, int *pre,
                                                  int *pre0, int *gen_overflow, LoopInit li) {
    const int jb = blockIdx.x;
    if (jb >= nbj) return;
    if (jb == 0 && threadIdx.x == 0) *gen_overflow = 0;  // the iteration's generated-table flag
    const BoxJob &B = bjs[jb];
    const int lane = threadIdx.x;
    const double thr = li.lj ? loop_thr(li.lj[jb], B.nterms) : B.thr;
    if (li.lj) {  // the device loop's iteration start (before any kernel that uses these)
        if (lane == 0) jobs[B.jt].thr = thr;
        if (lane < 4) li.tabcnt[4 * (size_t)jb + lane] = 0;
        if (jb == 0 && lane == 0) {
            li.ctl->genFirst = *(volatile const int *)li.gen_count;
            for (int q = 0; q < 4; q++) li.ctl->queue[q] = 0;
            li.ts[0] = li.ts[1] = 0ull;
        }
    }
    double m = 0.0;
    int ux = -1, uy = -1, uz = -1, carry = 0, carry0 = 0;
    for (int base = 0; base < B.nterms; base += 64) {
        const int t = base + lane;
        int v = 0, v0 = 0;
        if (t < B.nterms) {
            m = fmax(m, tbmax[(size_t)jb * max_terms + t]);
            const TermTab &tb = tabs[B.tab0 + t];
            if (tb.ub[0] >= 0) {
                v = (2 * tb.ub[0] + 1) * (2 * tb.ub[1] + 1) * (2 * tb.ub[2] + 1);
                ux = max(ux, (int)tb.ub[0]);
                uy = max(uy, (int)tb.ub[1]);
                uz = max(uz, (int)tb.ub[2]);
            }
            if (tb.ub0[0] >= 0) v0 = (2 * tb.ub0[0] + 1) * (2 * tb.ub0[1] + 1) * (2 * tb.ub0[2] + 1);
        }
        const int incl = wave_incl_scan(v), incl0 = wave_incl_scan(v0);
        if (t < B.nterms) {
            pre[B.pre_off + t] = carry + incl - v;
            pre0[B.pre_off + t] = carry0 + incl0 - v0;
        }
        carry += __shfl(incl, 63, 64);
        carry0 += __shfl(incl0, 63, 64);
    }
    m = wave_max_double(m);
    ux = wave_max_int(ux);
    uy = wave_max_int(uy);
    uz = wave_max_int(uz);
    if (lane == 0) {
        pre[B.pre_off + B.nterms] = carry;
        pre0[B.pre_off + B.nterms] = carry0;
        JobTab &J = jobs[B.jt];
        J.route = thr > 0.0 ? ROUTE_FP64 * thr : ROUTE_ROOT * m;
        J.ubox[0] = ux;
        J.ubox[1] = uy;
        J.ubox[2] = uz;
    }
}

// ----------------------------------------------------------------- apply: work list

// Enumerates, for one output node, every (term, delta, ft, gt) candidate inside the
// screened boxes and keeps those whose bound |w| prod_d ||X_d|| ||f_ft|| exceeds the
// job threshold.  Tuples are laid out by segment (g, gt):
//   mode 0 counts per (thread, gt), scans over the block and writes per-thread
//          offsets inside each segment and the segment totals;
//   mode 1 writes every tuple at the segment's start (from the schedule) + the part's
//          offset inside the segment + thread offset + running index.
template <int MODE>
__device__ __forceinline__ void enum_node(int g, int part, int nparts, const WorkItem *work, const JobTab *jobs,
                                          const TermTab *tabs, const GenState &gs, int *thread_off,
                                          int *seg_total, const long long *seg_ofs, Tuple *out,
                                          const SegRec *segs) {
    const WorkItem wi = work[g];
    const JobTab &J = jobs[wi.job];
    __shared__ int scan[8][256];
    __shared__ int fcache[FCACHE];
    __shared__ CandTerms ct;
    const FCache fc = fill_fcache(fcache, wi, J, tabs, gs);
    term_prefix(ct, J, tabs);
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long base[8];
    if (MODE == 1)
        for (int q = 0; q < 8; q++) {
            const size_t sp = ((size_t)g * 8 + q) * nparts + part;  // (segment, part)
            base[q] = segs[(size_t)g * 8 + q].start + (nparts > 1 ? seg_ofs[sp] : 0) +
                      thread_off[sp * 256 + threadIdx.x];
        }
    walk_candidates(ct, fcache, fc, wi, J, tabs, gs, part, nparts,
                    [&](int t, const TermTab &tb, int dx, int dy, int dz, int f, const double *fn, int nft,
                        const double *nx, const double *ny, const double *nz) {
                        const double aw = fabs(tb.w);
                        const unsigned long long cm = combo_mask(tb, dx, dy, dz);
                        unsigned surv[8];
                        double maxb[8];
                        screen_candidate<false>(aw, nx, ny, nz, fn, cm, J.thr, surv, maxb);
                        (void)nft;
                        if (MODE == 0) {
#pragma unroll
                            for (int ft = 0; ft < 8; ft++)
#pragma unroll
                                for (int gt = 0; gt < 8; gt++) cnt[gt] += (surv[ft] >> gt) & 1u;
                        } else {
                            // per gt (static index of base and cnt) its surviving ft, in order
#pragma unroll
                            for (int gt = 0; gt < 8; gt++) {
                                unsigned sg = 0u;
#pragma unroll
                                for (int ft = 0; ft < 8; ft++) sg |= ((surv[ft] >> gt) & 1u) << ft;
                                for (unsigned m = sg; m; m &= m - 1) {
                                    const int ft = __builtin_ctz(m);
                                    Tuple tp;
                                    tp.g = g;
                                    tp.f = f;
                                    tp.term = (uint8_t)t;
                                    tp.combo = (uint8_t)(gt * 8 + ft);
                                    tp.dx = (int16_t)dx;
                                    tp.dy = (int16_t)dy;
                                    tp.dz = (int16_t)dz;
                                    out[base[gt] + cnt[gt]] = tp;
                                    cnt[gt]++;
                                }
                            }
                        }
                    });
    if (MODE == 0) {
        for (int q = 0; q < 8; q++) scan[q][threadIdx.x] = cnt[q];
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            int v[8];
            for (int q = 0; q < 8; q++) v[q] = threadIdx.x >= off ? scan[q][threadIdx.x - off] : 0;
            __syncthreads();
            for (int q = 0; q < 8; q++) scan[q][threadIdx.x] += v[q];
            __syncthreads();
        }
        for (int q = 0; q < 8; q++)
            thread_off[(((size_t)g * 8 + q) * nparts + part) * 256 + threadIdx.x] = scan[q][threadIdx.x] - cnt[q];
        if (threadIdx.x == 255)
            for (int q = 0; q < 8; q++) seg_total[((size_t)g * 8 + q) * nparts + part] = scan[q][255];
    }
}

// one workgroup per (work node, part) (grid-stride; nwp: the node count on the device, the
// parts then follow from it as on the host)
template <int MODE>
__global__ void __launch_bounds__(256) k_enum(int nw, const int *nwp, int *queue, int nparts, const WorkItem *work,
                                              const JobTab *jobs, const TermTab *tabs, GenState gs, int *thread_off,
                                              int *seg_total, const long long *seg_ofs, Tuple *out,
                                              const TaskSummary *sum, const SegRec *segs) {
    if (MODE == 1 && sum->overflow) return;
    const int n = nwp ? *nwp : nw, np = nwp ? enum_parts(n) : nparts;
    for_items(n * np, queue, [&](int b) {
        enum_node<MODE>(b / np, b % np, np, work, jobs, tabs, gs, thread_off, seg_total, seg_ofs, out, segs);
    });
}

// ----------------------------------------------------------------- apply: predictor features

// One thread per (work node, child).  Features (norms relative to split thresholds, so
// that one network serves several precisions; thresholds use the output-norm estimate
// at the start of the iteration):
//  0 log10(||w_parent|| / split threshold of the parent)
//  1 log10(||s_parent|| / split threshold of the child)
//  2 log10(||w_parent|| / ||w_grandparent||) (0 below the roots)
//  3 log10(||w_parent|| / ||s_parent||)    4 input tree has the child's cell (0/1)
//  5 log10(||w_input(cell)|| ||g|| / ||input|| / split threshold of the child), -20 absent
//  6 the input node of the cell has children (0/1)
//  7 as 5 for the largest of the 26 neighbouring cells, -20 when none
//  8 log10(mu + 1e-3)                      9 log10(cell width of the child)
__device__ inline float feat_lg(double v) {
    return (float)fmin(20.0, fmax(-20.0, log10(fmax(v, 1e-300))));
}

__device__ inline double input_wnorm(const JobTab &J, int i) {
    double a = 0.0;
    for (int q = 1; q < 8; q++) a += J.inNorm[(size_t)i * 8 + q] * J.inNorm[(size_t)i * 8 + q];
    return sqrt(a);
}

// 32 lanes per child: lanes 0..25 look up the 26 same-level neighbours of the child
// in the input tree, lane 26 the child itself, lane 27 its first child, lane 28 the
// parent's wavelet norm; the features are then gathered with shuffles in the half-wave
__global__ void __launch_bounds__(256) k_child_features(int nw, const WorkItem *work, const JobTab *jobs,
                                                        const double *wnorms, float *feat) {
    const int tt = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 5);
    const int l = threadIdx.x & 31;
    const bool valid = tt < nw * 8;
    const int t = valid ? tt : nw * 8 - 1;
    const int i = t >> 3, a = t & 7;
    const WorkItem wi = work[i];
    const JobTab &J = jobs[wi.job];
    const int d = J.depth + 1;
    const int cx = 2 * wi.l[0] + (a & 1), cy = 2 * wi.l[1] + ((a >> 1) & 1), cz = 2 * wi.l[2] + ((a >> 2) & 1);
    double v = 0.0;  // lanes < 26: neighbour wavelet norm; 26: own norm (-1: absent); 28: parent's
    int found = 0;   // lanes 26 and 27: node present in the input tree
    if (l < 26) {
        const int c = l < 13 ? l : l + 1;  // 27 cells without the centre
        const int x = cx + c % 3 - 1, y = cy + (c / 3) % 3 - 1, z = cz + c / 9 - 1;
        if (x >= 0 && y >= 0 && z >= 0 && x < 2 * J.lmax[0] && y < 2 * J.lmax[1] && z < 2 * J.lmax[2]) {
            const int j = dhash_find(J.inHash, dnode_key(d, x, y, z));
            if (j >= 0) v = input_wnorm(J, j);
        }
    } else if (l == 26) {
        const int i0 = dhash_find(J.inHash, dnode_key(d, cx, cy, cz));
        found = i0 >= 0;
        v = found ? input_wnorm(J, i0) : 0.0;
    } else if (l == 27) {
        found = dhash_find(J.inHash, dnode_key(d + 1, 2 * cx, 2 * cy, 2 * cz)) >= 0;
    } else if (l == 28 && wi.parent >= 0) {
        double wpp = 0.0;
        for (int q = 1; q < 8; q++) {
            const double u = J.outNorm[(size_t)wi.parent * 8 + q];
            wpp += u * u;
        }
        v = sqrt(wpp);
    }
    double nb = l < 26 ? v : 0.0;
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) nb = fmax(nb, __shfl_xor(nb, o, 32));
    const double own = __shfl(v, 26, 32), wpp = __shfl(v, 28, 32);
    const int has = __shfl(found, 26, 32), has2 = __shfl(found, 27, 32);
    if (!valid || l != 0) return;
    const double *pn = wnorms + (size_t)i * 8;
    double w2 = 0.0;
    for (int q = 1; q < 8; q++) w2 += pn[q] * pn[q];
    const double wn = sqrt(w2), sn = pn[0];
    const int np = J.n0 + J.depth;
    const double eps2 = 2.0 * 2.220446049250313e-16;
    const double thr = fmax(eps2, J.prec * J.feat_tnorm * exp2(-0.5 * (np + 1)));
    const double thr_c = fmax(eps2, J.prec * J.feat_tnorm * exp2(-0.5 * (np + 2)));
    const double in_to_out = J.feat_tnorm / fmax(J.feat_innorm, 1e-300);
    float *f = feat + (size_t)t * 10;
    f[0] = feat_lg(wn / thr);
    f[1] = feat_lg(sn / thr_c);
    f[2] = wpp > 0.0 ? feat_lg(wn / wpp) : 0.0f;
    f[3] = feat_lg(wn / fmax(fabs(sn), 1e-300));
    f[4] = has ? 1.0f : 0.0f;
    f[5] = has ? feat_lg(own * in_to_out / thr_c) : -20.0f;
    f[6] = (has && has2) ? 1.0f : 0.0f;
    f[7] = nb > 0.0 ? feat_lg(nb * in_to_out / thr_c) : -20.0f;
    f[8] = J.feat_mu;
    f[9] = J.feat_cell;
}

hipError_t launch_child_features(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs,
                                 const double *wnorms, float *feat) {
    if (nw <= 0) return hipSuccess;
    const int n = nw * 8;
    hipLaunchKernelGGL(k_child_features, dim3((n + 7) / 8), dim3(256), 0, s, nw, work, jobs, wnorms, feat);
    return hipGetLastError();
}

// ----------------------------------------------------------------- apply: contraction schedule

__device__ inline long long wave_incl_scan_ll(long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// exclusive prefixes of NV values at once over the 1024 threads of the block (one
// exchange through LDS for all of them); tot[k] gets the sums
template <int NV>
__device__ inline void block_excl_scan_1024_n(long long (&v)[NV], long long (*wsum)[NV], long long (&tot)[NV]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    long long incl[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) incl[k] = wave_incl_scan_ll(v[k]);
    __syncthreads();
    if (lane == 63)
#pragma unroll
        for (int k = 0; k < NV; k++) wsum[wave][k] = incl[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        long long acc = 0;
        for (int w = 0; w < 16; w++) {
            const long long x = wsum[w][k];
            wsum[w][k] = acc;
            acc += x;
        }
        wsum[16][k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        tot[k] = wsum[16][k];
        v[k] = wsum[wave][k] + incl[k] - v[k];
    }
}


// One workgroup.  Phase 0 (several parts per node): segment counts and the parts'
// offsets inside their segment.  Phase 1: per thread (a range of output nodes) the
// totals of tuples, slices, FP64 tasks, slots, slice tasks and reduce items; one block
// scan of the six gives each thread's offsets.  Phase 2: per segment (node, component)
// its record (start, count, first task, first slot), reduce items, and the tasks of
// segments with few tasks; segments (and nodes' slice runs) with many tasks go to a list
// in LDS.  Phase 3: one wave per listed segment writes its tasks.  Every task has a fixed
// position, so the schedule does not depend on the order of the list.
constexpr int BT_FEW = 32, BT_LIST = 2048, BT_TABS = 256;
constexpr long long FP64_TASKS = 4096;  // FP64 tasks an iteration aims at (task size permitting)

// (one workgroup: the pointers never alias, so that loads of later iterations can be
// issued ahead of the stores of earlier ones)
__global__ void __launch_bounds__(1024)
    k_build_tasks(int nw, const int *nwp, int nseg, int nparts, int chunk, int wchunk, const int *__restrict__ tot,
                  TaskCaps caps,
                  const int *__restrict__ gen_count, int gen_pcap, long long *__restrict__ ofs,
                  ContractTask *__restrict__ tasks, SliceTask *__restrict__ stasks, ReduceItem *__restrict__ reds,
                  TaskSummary *__restrict__ sum, SegRec *__restrict__ segs, int *__restrict__ segcnt,
                  long long *__restrict__ gslice, const int *__restrict__ gen_overflow,
                  const WorkItem *__restrict__ work, long long *__restrict__ tabcnt, int ntab) {
    __shared__ long long wsum6[17][6];
    __shared__ int big[BT_LIST];     // segment ids (>= 0) or -(node + 1) for slice runs
    __shared__ int nbig;
    if (nwp) {  // the work list's size on the device (its parts as the enumeration derived them)
        nw = *nwp;
        nparts = enum_parts(nw);
    }
    const bool mixed = nseg == 9;
    // FP64 tasks of a tuple segment with c tuples (plain: always one, even empty, so that
    // every output component is written; mixed: the reduce item writes the component)
    auto ntasks_of = [&](int c) { return mixed ? (c + chunk - 1) / chunk : (c <= chunk ? 1 : (c + chunk - 1) / chunk); };
    auto nslots_of = [&](int c) { return mixed ? (c + chunk - 1) / chunk : (c <= chunk ? 0 : (c + chunk - 1) / chunk); };
    if (threadIdx.x == 0) nbig = 0;
    // per job table: tuples and slices of its output nodes (statistics for the host),
    // summed in LDS for the first BT_TABS tables, in global memory beyond
    __shared__ unsigned long long lcnt[2 * BT_TABS];
    // (entries 4 t + 3 hold the generated-node counts of the iteration: kept)
    for (int q = threadIdx.x; q < 3 * ntab; q += blockDim.x) tabcnt[4 * (q / 3) + q % 3] = 0;
    for (int q = threadIdx.x; q < 2 * BT_TABS; q += blockDim.x) lcnt[q] = 0;
    // phase 0: segment counts (sum over the parts) and per-part offsets inside the segment
    // (relative to the segment's start)
    // (one part per node: the counts are the totals and every part offset is 0, which
    // the fill kernels know)
    const int nsg = nw * nseg;  // (work lists stay far below 2^31 / 9 nodes)
    if (nparts > 1) {
#pragma unroll 4
        for (int sg = threadIdx.x; sg < nsg; sg += blockDim.x) {
            int c = 0;
            for (int p = 0; p < nparts; p++) {
                ofs[sg * nparts + p] = c;
                c += tot[sg * nparts + p];
            }
            segcnt[sg] = c;
        }
    }
    const int *sc = nparts > 1 ? segcnt : tot;
    __syncthreads();
    // phase 1: this thread's totals (a contiguous range of output nodes)
    const int g0 = (int)((long long)nw * threadIdx.x / 1024), g1 = (int)((long long)nw * (threadIdx.x + 1) / 1024);
    {
        // task sizes from the iteration's totals.  FP64 tasks: at most chunk tuples, fewer
        // (down to 16) when the iteration has fewer than FP64_TASKS x chunk tuples, so that a
        // coarse iteration still spreads over the GPU.  Slices per MFMA task when automatic
        // (wchunk <= 0): about -wchunk tasks, at least 16 and at most 128 slices each (larger
        // tasks leave too few to balance the k = 9 slices over the CUs)
        long long tu1 = 0, sl1 = 0;
        for (int g = g0; g < g1; g++) {
            for (int q = 0; q < 8; q++) tu1 += sc[(size_t)g * nseg + q];
            if (mixed) sl1 += sc[(size_t)g * 9 + 8];
        }
        long long v2[2] = {tu1, sl1}, t2[2];
        block_excl_scan_1024_n<2>(v2, (long long(*)[2])wsum6, t2);
        chunk = (int)min((long long)chunk, max(16LL, (t2[0] + FP64_TASKS - 1) / FP64_TASKS));
        if (mixed && wchunk <= 0) {
            const long long want = (t2[1] + (-wchunk) - 1) / (-wchunk);
            wchunk = (int)min(128LL, max(16LL, want));
        }
        __syncthreads();
    }
    long long tup = 0, sl = 0, nt = 0, nst = 0, slots = 0, nred = 0;
    for (int g = g0; g < g1; g++) {
        if (mixed) {
            const int cs = sc[(size_t)g * 9 + 8];
            sl += cs;
            nst += (cs + wchunk - 1) / wchunk;
        }
        for (int q = 0; q < 8; q++) {
            const int c = sc[(size_t)g * nseg + q];
            tup += c;
            nt += ntasks_of(c);
            slots += nslots_of(c);
            if (!mixed && c > chunk) nred++;
        }
    }
    long long pv[6] = {tup, sl, nt, nst, slots, nred}, pt[6];
    block_excl_scan_1024_n<6>(pv, wsum6, pt);
    long long btup = pv[0], bsl = pv[1], bnt = pv[2], bnst = pv[3], bslots = pv[4], bnred = pv[5];
    long long Ttup = pt[0], Tsl = pt[1], Tnt = pt[2], Tnst = pt[3], Tslots = pt[4], Tnred = pt[5];
    if (mixed) Tnred = 8LL * nw;
    int overflow = 0;
    if (Ttup > caps.tuples || Tsl > caps.slices || Tnt > caps.tasks || Tnst > caps.stasks || Tnred > caps.reds ||
        Tslots > caps.slots || Tnst > caps.nnslots)
        overflow |= 1;
    if (*(volatile const int *)gen_count > gen_pcap) overflow |= 2;
    if (threadIdx.x == 0) {
        sum->ntuples = Ttup;
        sum->nslices = Tsl;
        sum->ntasks = (int)Tnt;
        sum->nstasks = (int)Tnst;
        sum->nreds = (int)Tnred;
        sum->nslots = (int)Tslots;
        sum->nnslots = (int)Tnst;
        sum->overflow = overflow;
        sum->gen_count = *(volatile const int *)gen_count;
        sum->gen_overflow = *(volatile const int *)gen_overflow;
        sum->schunk = mixed ? wchunk : 0;
    }
    if (overflow) return;
    auto emit_task = [&](const SegRec &r, int sgi, int h) {
        const int cnt = r.slot0 < 0 ? r.count : min(chunk, r.count - h * chunk);
        tasks[r.task0 + h] = ContractTask{r.start + (long long)h * chunk, cnt, sgi >> 3, sgi & 7,
                                          r.slot0 < 0 ? -1 : r.slot0 + h};
    };
    auto emit_stask = [&](int g, int h) {
        const int cs = sc[(size_t)g * 9 + 8], nn0 = segs[(size_t)g * 8].pad_stask0;
        stasks[nn0 + h] = SliceTask{gslice[g] + (long long)h * wchunk, min(wchunk, cs - h * wchunk), g, nn0 + h};
    };
    // phase 2: segment records, reduce items, tasks of small segments
    int run_jt = -1;  // table statistics: one add per run of nodes of the same table
    unsigned long long run_t = 0, run_s = 0;
    auto flush_run = [&]() {
        if (run_jt < 0) return;
        if (run_jt < BT_TABS) {
            atomicAdd(&lcnt[2 * run_jt], run_t);
            atomicAdd(&lcnt[2 * run_jt + 1], run_s);
        } else {
            atomicAdd((unsigned long long *)&tabcnt[4 * run_jt], run_t);
            atomicAdd((unsigned long long *)&tabcnt[4 * run_jt + 1], run_s);
        }
    };
    for (int g = g0; g < g1; g++) {
        {
            const int jt = work[g].job;
            if (jt != run_jt) {
                flush_run();
                run_jt = jt;
                run_t = run_s = 0;
            }
            for (int q = 0; q < 8; q++) run_t += (unsigned long long)sc[(size_t)g * nseg + q];
            if (mixed) run_s += (unsigned long long)sc[(size_t)g * 9 + 8];
        }
        int nn0 = (int)bnst, nns = 0;
        if (mixed) {
            const int cs = sc[(size_t)g * 9 + 8];
            nns = (cs + wchunk - 1) / wchunk;
            gslice[g] = bsl;
            segs[(size_t)g * 8].pad_stask0 = nn0;  // first slice task of node g
            int slot_ = -1;
            if (nns > BT_FEW) {
                slot_ = atomicAdd(&nbig, 1);
                if (slot_ < BT_LIST) big[slot_] = -(g + 1);
            }
            if (slot_ < 0 || slot_ >= BT_LIST)
                for (int h = 0; h < nns; h++)
                    stasks[nn0 + h] = SliceTask{bsl + (long long)h * wchunk, min(wchunk, cs - h * wchunk), g, nn0 + h};
            bnst += nns;
            bsl += cs;
        }
        for (int q = 0; q < 8; q++) {
            const int c = sc[(size_t)g * nseg + q];
            const int k = ntasks_of(c), ns = nslots_of(c);
            const int sgi = g * 8 + q;
            const SegRec r{btup, c, (int)bnt, ns ? (int)bslots : -1, 0};
            segs[sgi].start = r.start;
            segs[sgi].count = r.count;
            segs[sgi].task0 = r.task0;
            segs[sgi].slot0 = r.slot0;
            if (mixed) reds[(size_t)g * 8 + q] = ReduceItem{g, q, (int)bslots, ns, nn0, nns};
            else if (ns) reds[bnred++] = ReduceItem{g, q, (int)bslots, ns, 0, 0};
            int slot_ = -1;
            if (k > BT_FEW) {
                slot_ = atomicAdd(&nbig, 1);
                if (slot_ < BT_LIST) big[slot_] = sgi;
            }
            if (slot_ < 0 || slot_ >= BT_LIST)
                for (int h = 0; h < k; h++) emit_task(r, sgi, h);
            bnt += k;
            bslots += ns;
            btup += c;
        }
    }
    flush_run();
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * min(ntab, BT_TABS); q += blockDim.x)
        tabcnt[4 * (q >> 1) + (q & 1)] = (long long)lcnt[q];
    // phase 3: the listed segments, one wave per segment with the lanes over its tasks
    // (the fill kernels add the segment starts to the per-part offsets themselves)
    const int nb = min(nbig, BT_LIST);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int b = wave; b < nb; b += 16) {
        const int id = big[b];
        if (id >= 0) {
            const SegRec r = segs[id];
            const int k = ntasks_of(r.count);
            for (int h = lane; h < k; h += 64) emit_task(r, id, h);
        } else {
            const int g = -id - 1;
            const int nns = (sc[(size_t)g * 9 + 8] + wchunk - 1) / wchunk;
            for (int h = lane; h < nns; h += 64) emit_stask(g, h);
        }
    }
}

// ----------------------------------------------------------------- apply: FP64 contraction

// One wavefront per task = a chunk of the tuples of one segment (g, gt):
//   out_gt[c][b][a] += w * sum X[a][i] Y[b][j] Z[c][l] f_ft[l][j][i]
// Pass 1 keeps a row of f per lane (X by scalar loads), passes 2 and 3 exchange
// through the wave's LDS slice.  The result goes to the output node when the
// segment has one chunk, else to a partial slot summed by k_reduce_partials.
template <int K>
__device__ inline void contract_fp64_task(const ContractTask task, const WorkItem *work, const JobTab *jobs,
                                          const TermTab *tabs, const Tuple *tuples, const double *genPool,
                                          double *partials, double *B, int lane) {
    constexpr int K2 = K * K, K3 = K2 * K;
    constexpr int NP = (K2 + 63) / 64;
    const WorkItem wi = work[task.g];
    const JobTab &J = jobs[wi.job];
    const int gt = task.gt;
    const gdouble *blocks = (const gdouble *)J.operBlocks;
    const gdouble *inCoef = (const gdouble *)J.inCoef;
    const gdouble *gen = (const gdouble *)genPool;
    double acc[NP][K];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int c = 0; c < K; c++) acc[p][c] = 0.0;
    for (int it = 0; it < task.count; it++) {
        const Tuple tp = tuples[task.start + it];
        const int ft = tp.combo & 7;
        const TermTab &tb = tabs[J.tab0 + tp.term];
        const int tx = 2 * (gt & 1) + (ft & 1);
        const int ty = 2 * ((gt >> 1) & 1) + ((ft >> 1) & 1);
        const int tz = 2 * ((gt >> 2) & 1) + ((ft >> 2) & 1);
        const long long band = tb.band, blk = tb.blk;
        const gdouble *X = blocks + blk + ((tp.dx + band) * 4 + tx) * K2;
        const gdouble *Y = blocks + blk + ((tp.dy + band) * 4 + ty) * K2;
        const gdouble *Z = blocks + blk + ((tp.dz + band) * 4 + tz) * K2;
        const gdouble *f = tp.f >= 0 ? inCoef + (size_t)tp.f * 8 * K3 + ft * K3 : gen + (size_t)(-tp.f - 1) * K3;
        const double w = tb.w;
        // The operator rows come through scalar loads.  Up to K = 8 a whole K x K block fits
        // the scalar registers and the loops unroll fully; beyond, two rows (pass 3: two
        // columns) are live at a time — the same sums in the same order.
        constexpr bool ROWWISE = K > 8;
        constexpr int RU = ROWWISE ? 2 : K;
        for (int r = lane; r < K2; r += 64) {
            double fr[K];
#pragma unroll
            for (int i = 0; i < K; i++) fr[i] = f[r * K + i];
#pragma unroll RU
            for (int a = 0; a < K; a++) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < K; i++) s += X[a * K + i] * fr[i];
                B[r * K + a] = s;
            }
        }
        wave_sync();
        for (int col = lane; col < K2; col += 64) {
            const int l = col / K, a = col % K;
            double tc[K];
#pragma unroll
            for (int j = 0; j < K; j++) tc[j] = B[(l * K + j) * K + a];
#pragma unroll RU
            for (int b = 0; b < K; b++) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < K; j++) s += Y[b * K + j] * tc[j];
                B[(l * K + b) * K + a] = s;
            }
        }
        wave_sync();
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int ba = lane + 64 * p;
            if (ba < K2) {
                if constexpr (ROWWISE) {
                    double s[K];
#pragma unroll
                    for (int c = 0; c < K; c++) s[c] = 0.0;
#pragma unroll 2
                    for (int l = 0; l < K; l++) {
                        const double t = B[l * K2 + ba];
#pragma unroll
                        for (int c = 0; c < K; c++) s[c] += Z[c * K + l] * t;
                    }
#pragma unroll
                    for (int c = 0; c < K; c++) acc[p][c] += w * s[c];
                } else {
                    double tc[K];
#pragma unroll
                    for (int l = 0; l < K; l++) tc[l] = B[l * K2 + ba];
#pragma unroll
                    for (int c = 0; c < K; c++) {
                        double s = 0.0;
#pragma unroll
                        for (int l = 0; l < K; l++) s += Z[c * K + l] * tc[l];
                        acc[p][c] += w * s;
                    }
                }
            }
        }
        wave_sync();
    }
    double *dst = task.slot >= 0 ? partials + (size_t)task.slot * K3
                                 : J.outCoef + (size_t)wi.node * 8 * K3 + gt * K3;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const int ba = lane + 64 * p;
        if (ba < K2) {
#pragma unroll
            for (int c = 0; c < K; c++) dst[c * K2 + ba] = acc[p][c];
        }
    }
}

// The task count comes from the device-built schedule (sum) when given; the grid
// covers the schedule's capacity and waves past the count exit.
template <int K>
__global__ void __launch_bounds__(256) k_contract_fp64(int ntask_arg, const TaskSummary *sum,
                                                       const ContractTask *tasks, const WorkItem *work,
                                                       const JobTab *jobs, const TermTab *tabs,
                                                       const Tuple *tuples, const double *genPool,
                                                       double *partials, unsigned long long *ts) {
    // (no stores or clock reads on the way to the task loop: the operator rows are read
    // through scalar loads, which the compiler only uses for memory nothing before them may
    // have changed.  With ts, block 0, dispatched first, only stamps the contraction's start
    // and the tasks start at block 1.)
    constexpr int K3 = K * K * K;
    __shared__ double buf[4][K3];
    if (ts && blockIdx.x == 0) {
        if (threadIdx.x == 0) *ts = (unsigned long long)wall_clock64();
        return;
    }
    const int wave = wave_id(), lane = threadIdx.x & 63;
    const int ntask = sum ? (sum->overflow ? 0 : sum->ntasks) : ntask_arg;
    const int ti = ((int)blockIdx.x - (ts ? 1 : 0)) * 4 + wave;
    if (ti >= ntask) return;
    contract_fp64_task<K>(tasks[ti], work, jobs, tabs, tuples, genPool, partials, buf[wave], lane);
}


// sums, per segment (g, gt), the FP64 tuple slots and the gt part of the mixed
// node slots in slot order (deterministic)
// Sums the partial results of an output component in a fixed order.  With wnorm
// (mixed path: every component of every work node has an item) the component's norm
// goes to the tree and to the compact per-work-node array as well.
__global__ void __launch_bounds__(256) k_reduce_partials(int n_arg, const TaskSummary *sum, const ReduceItem *items,
                                                         const WorkItem *work, const JobTab *jobs,
                                                         const double *partials, const double *nodeslots, int K3,
                                                         double *wnorm) {
    __shared__ double wsum[4];
    const int n = sum ? (sum->overflow ? 0 : sum->nreds) : n_arg;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const ReduceItem it = items[i];
        const WorkItem wi = work[it.g];
        const JobTab &J = jobs[wi.job];
        double *dst = J.outCoef + (size_t)wi.node * 8 * K3 + it.gt * K3;
        double ss = 0.0;
        for (int e = threadIdx.x; e < K3; e += blockDim.x) {
            double s = 0.0;
            for (int q = 0; q < it.nslots; q++) s += partials[(size_t)(it.slot0 + q) * K3 + e];
            for (int q = 0; q < it.nnslots; q++) s += nodeslots[((size_t)(it.nslot0 + q) * 8 + it.gt) * K3 + e];
            dst[e] = s;
            ss += s * s;
        }
        if (wnorm) {
            ss = block_sum_256(ss, wsum);
            if (threadIdx.x == 0) {
                const double v = sqrt(ss);
                J.outNorm[(size_t)wi.node * 8 + it.gt] = v;
                wnorm[(size_t)it.g * 8 + it.gt] = v;
            }
        }
    }
}


// ----------------------------------------------------------------- apply: mixed-precision work list

// Same candidates and bounds as k_enum.  A candidate (term, delta, ft) whose
// surviving combos all stay below ROUTE_FP64 * thr becomes one slice (all allowed
// gt, split-FP16 MFMA); otherwise its surviving combos become FP64 tuples.  Counters per
// thread: 8 tuple segments (per gt) + 1 slice segment; layout of thread_off and
// seg_total: [g][9][...].
template <int MODE>
__device__ __forceinline__ void enum_mixed_node(int g, int part, int nparts, int K2, const WorkItem *work,
                                                const JobTab *jobs, const TermTab *tabs, const GenState &gs,
                                                int *thread_off, int *seg_total, const long long *seg_ofs,
                                                Tuple *out, Slice *sout, const SegRec *segs,
                                                const long long *gslice) {
    const WorkItem wi = work[g];
    const JobTab &J = jobs[wi.job];
    __shared__ int scan[9][256];
    __shared__ int fcache[FCACHE];
    __shared__ CandTerms ct;
    const FCache fc = fill_fcache(fcache, wi, J, tabs, gs);
    term_prefix(ct, J, tabs);
    int cnt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    long long base[9];
    if (MODE == 1)
        for (int q = 0; q < 9; q++) {
            const size_t sp = ((size_t)g * 9 + q) * nparts + part;  // (segment, part)
            base[q] = (q < 8 ? segs[(size_t)g * 8 + q].start : gslice[g]) + (nparts > 1 ? seg_ofs[sp] : 0) +
                      thread_off[sp * 256 + threadIdx.x];
        }
    walk_candidates(ct, fcache, fc, wi, J, tabs, gs, part, nparts,
                    [&](int t, const TermTab &tb, int dx, int dy, int dz, int f, const double *fn, int nft,
                        const double *nx, const double *ny, const double *nz) {
                        const double aw = fabs(tb.w);
                        const unsigned long long cm = combo_mask(tb, dx, dy, dz);
                        unsigned survs[8];
                        double maxbs[8];
                        screen_candidate<true>(aw, nx, ny, nz, fn, cm, J.thr, survs, maxbs);
                        (void)nft;  // (survs[ft] = 0 beyond the node's components)
#pragma unroll
                        for (int ft = 0; ft < 8; ft++) {
                            const int surv = (int)survs[ft];
                            const double maxb = maxbs[ft];
                            if (!surv) continue;
                            if (maxb > J.route) {
#pragma unroll
                                for (int gt = 0; gt < 8; gt++) {
                                    if (!((surv >> gt) & 1)) continue;
                                    if (MODE == 1) {
                                        Tuple tp;
                                        tp.g = g;
                                        tp.f = f;
                                        tp.term = (uint8_t)t;
                                        tp.combo = (uint8_t)(gt * 8 + ft);
                                        tp.dx = (int16_t)dx;
                                        tp.dy = (int16_t)dy;
                                        tp.dz = (int16_t)dz;
                                        out[base[gt] + cnt[gt]] = tp;
                                    }
                                    cnt[gt]++;
                                }
                            } else {
                                if (MODE == 1) {
                                    // the record as two 16-byte stores (layout of Slice): weight,
                                    // f, FP32 block offsets of the three dimensions, ft, and the
                                    // surviving components only (the same tuples as the plain
                                    // work list, FP32 in place of FP64)
                                    const unsigned long long wb = (unsigned long long)__double_as_longlong(tb.w);
                                    const int xo = (int)(tb.blk + ((dx + tb.band) * 4 + (ft & 1)) * K2);
                                    const int yo = (int)(tb.blk + ((dy + tb.band) * 4 + ((ft >> 1) & 1)) * K2);
                                    const int zo = (int)(tb.blk + ((dz + tb.band) * 4 + ((ft >> 2) & 1)) * K2);
                                    uint4 *dst = (uint4 *)(sout + base[8] + cnt[8]);
                                    dst[0] = make_uint4((unsigned)wb, (unsigned)(wb >> 32), (unsigned)f, (unsigned)xo);
                                    dst[1] = make_uint4((unsigned)yo, (unsigned)zo, (unsigned)ft | ((unsigned)surv << 8), 0u);
                                }
                                cnt[8]++;
                            }
                        }
                    });
    if (MODE == 0) {
        for (int q = 0; q < 9; q++) scan[q][threadIdx.x] = cnt[q];
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            int v[9];
            for (int q = 0; q < 9; q++) v[q] = threadIdx.x >= off ? scan[q][threadIdx.x - off] : 0;
            __syncthreads();
            for (int q = 0; q < 9; q++) scan[q][threadIdx.x] += v[q];
            __syncthreads();
        }
        for (int q = 0; q < 9; q++)
            thread_off[(((size_t)g * 9 + q) * nparts + part) * 256 + threadIdx.x] = scan[q][threadIdx.x] - cnt[q];
        if (threadIdx.x == 255)
            for (int q = 0; q < 9; q++) seg_total[((size_t)g * 9 + q) * nparts + part] = scan[q][255];
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_enum_mixed(int nw, const int *nwp, int *queue, int nparts, int K2,
                                                    const WorkItem *work, const JobTab *jobs, const TermTab *tabs,
                                                    GenState gs, int *thread_off, int *seg_total,
                                                    const long long *seg_ofs, Tuple *out, Slice *sout,
                                                    const TaskSummary *sum, const SegRec *segs,
                                                    const long long *gslice) {
    if (MODE == 1 && sum->overflow) return;
    const int n = nwp ? *nwp : nw, np = nwp ? enum_parts(n) : nparts;
    for_items(n * np, queue, [&](int b) {
        enum_mixed_node<MODE>(b / np, b % np, np, K2, work, jobs, tabs, gs, thread_off, seg_total, seg_ofs, out,
                              sout, segs, gslice);
    });
}


// norms of the work nodes of all jobs, into their trees and into a compact array
__device__ __forceinline__ void work_norms_node(int i, const WorkItem *work, const JobTab *jobs, int K3,
                                                double *compact) {
    const WorkItem wi = work[i];
    const JobTab &J = jobs[wi.job];
    const double *p = J.outCoef + (size_t)wi.node * 8 * K3;
    for (int c = 0; c < 8; c++) {
        double s = 0.0;
        for (int e = threadIdx.x; e < K3; e += 64) s += p[c * K3 + e] * p[c * K3 + e];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (threadIdx.x == 0) {
            const double v = sqrt(s);
            J.outNorm[(size_t)wi.node * 8 + c] = v;
            compact[(size_t)i * 8 + c] = v;
        }
    }
}

__global__ void k_work_norms(int nw, const int *nwp, const WorkItem *work, const JobTab *jobs, int K3,
                             double *compact, unsigned long long *ts_end) {
    // (the contraction's end on the device clock: this kernel starts once the reduction is done)
    if (ts_end && blockIdx.x == 0 && threadIdx.x == 0) *ts_end = (unsigned long long)wall_clock64();
    const int n = nwp ? *nwp : nw;
    for (int i = blockIdx.x; i < n; i += gridDim.x) work_norms_node(i, work, jobs, K3, compact);
}

// ----------------------------------------------------------------- apply: device-resident refinement

__global__ void __launch_bounds__(256) k_loop_init(int njobs, int nr, const LoopJob *lj, const int *rootlr,
                                                   WorkItem *work) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < njobs * nr; i += gridDim.x * blockDim.x) {
        const int j = i / nr, r = i % nr;
        const LoopJob &L = lj[j];
        WorkItem wi;
        wi.job = j;
        wi.node = r;
        wi.parent = -1;
        L.depth[r] = 0;
        L.parent[r] = -1;
        L.child[r] = -1;
        for (int q = 0; q < 3; q++) {
            wi.l[q] = rootlr[3 * r + q];
            L.lr[3 * r + q] = wi.l[q];
        }
        work[i] = wi;
    }
}

// One workgroup per job.  The running norm adds the squared norms of the job's work nodes
// (scaling parts on the roots, wavelet parts everywhere; per thread over a contiguous run,
// then a fixed reduction tree); the split test is the host's (split_check with the factor
// 2^(-(n0 + d + 1) / 2) the host computed); the split nodes' children follow the node count in
// work-list order, 8 each, as the host appends them.
__global__ void __launch_bounds__(256) k_loop_split(int njobs, LoopJob *lj, const WorkItem *work, const double *wn,
                                                    int max_depth, LoopCtl *ctl, int *cidx, const long long *tabcnt,
                                                    unsigned long long *ts_end) {
    const int j = blockIdx.x, tid = threadIdx.x;
    if (ts_end && j == 0 && tid == 0) *ts_end = (unsigned long long)wall_clock64();
    if (j >= njobs) return;
    LoopJob &L = lj[j];
    __shared__ double rs[256], rw[256];
    __shared__ int sc[256];
    __shared__ double s_tn;
    __shared__ int s_n;
    if (tid == 0) {
        L.tuples += tabcnt[4 * (size_t)j] + tabcnt[4 * (size_t)j + 2];
        L.tuples_fp64 += tabcnt[4 * (size_t)j];
        L.slices += tabcnt[4 * (size_t)j + 1];
        L.gen += tabcnt[4 * (size_t)j + 3];
    }
    const int w0 = L.w0, nwk = L.nwk;
    if (nwk <= 0) {
        if (tid == 0) L.splits = 0;
        return;
    }
    const int d = L.depth[work[w0].node];
    const int per = (nwk + 255) / 256, i0 = min(nwk, tid * per), i1 = min(nwk, i0 + per);
    double s = 0.0, w = 0.0;
    for (int i = i0; i < i1; i++) {
        const double *p = wn + (size_t)(w0 + i) * 8;
        if (d == 0) s += p[0] * p[0];
        for (int c = 1; c < 8; c++) w += p[c] * p[c];
    }
    rs[tid] = s;
    rw[tid] = w;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            rs[tid] += rs[tid + off];
            rw[tid] += rw[tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        L.sNorm += rs[0];
        L.wNorm += rw[0];
        L.gsq = L.sNorm + L.wNorm;
        s_tn = sqrt(L.gsq);
        s_n = L.n;
    }
    __syncthreads();
    const double thr = fmax(2.0 * 2.220446049250313e-16, L.prec * s_tn * ctl->splitfac[d]);
    const bool can = d < max_depth;
    auto splits_at = [&](int i) {
        const double *p = wn + (size_t)(w0 + i) * 8;
        double w2 = 0.0;
        for (int c = 1; c < 8; c++) w2 += p[c] * p[c];
        return can && sqrt(w2) > thr;
    };
    int cnt = 0;
    for (int i = i0; i < i1; i++) cnt += splits_at(i) ? 1 : 0;
    sc[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = tid >= off ? sc[tid - off] : 0;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    const int S = sc[255], n0 = s_n;
    const bool fits = (long long)n0 + 8LL * S <= (long long)L.cap;
    int r = sc[tid] - cnt;
    for (int i = i0; i < i1; i++) {
        const bool sp = fits && splits_at(i);
        cidx[w0 + i] = sp ? n0 + 8 * r : -1;
        if (!sp) continue;
        const WorkItem wi = work[w0 + i];
        const int c0 = n0 + 8 * r;
        L.child[wi.node] = c0;
        for (int a = 0; a < 8; a++) {
            const size_t c = (size_t)c0 + a;
            L.depth[c] = d + 1;
            L.parent[c] = wi.node;
            L.child[c] = -1;
            L.lr[3 * c] = 2 * wi.l[0] + (a & 1);
            L.lr[3 * c + 1] = 2 * wi.l[1] + ((a >> 1) & 1);
            L.lr[3 * c + 2] = 2 * wi.l[2] + ((a >> 2) & 1);
        }
        r++;
    }
    if (tid == 0) {
        if (!fits) atomicOr(&ctl->overflow, 4);
        L.n = fits ? n0 + 8 * S : n0;
        L.splits = fits ? S : 0;
    }
}

// The next work list: job j's children start at 8 x (split nodes of the jobs before j),
// its split nodes' children in work-list order (the order of the host loop).  Every
// workgroup forms the same prefix over the jobs; workgroup 0 also records the jobs' ranges,
// the list's size and the iteration's read-back entry.
__global__ void __launch_bounds__(256) k_loop_emit(int njobs, LoopJob *lj, const WorkItem *work, const int *nwp,
                                                   const int *cidx, WorkItem *next, int *nw_next, int wcap,
                                                   const TaskSummary *sum, LoopCtl *ctl, int d, const JobTab *jt,
                                                   TransformItem *titems, int tcap, LoopIter *host_iter) {
    __shared__ int base[LOOP_MAX_JOBS];
    __shared__ int sc[256];
    const int tid = threadIdx.x;
    const int per = (njobs + 255) / 256, j0 = min(njobs, tid * per), j1 = min(njobs, j0 + per);
    int c = 0;
    for (int j = j0; j < j1; j++) c += lj[j].splits;
    sc[tid] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = tid >= off ? sc[tid - off] : 0;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    int b = sc[tid] - c;
    for (int j = j0; j < j1; j++) {
        base[j] = b;
        b += lj[j].splits;
    }
    const long long total = sc[255];
    __syncthreads();
    // (bits of earlier kernels; the ones added here follow from the same inputs in every workgroup)
    const int t0 = ctl->tofs[d];
    int ovf = *(volatile int *)&ctl->overflow;
    if (sum->overflow) ovf |= sum->overflow & 3;
    if (sum->gen_overflow) ovf |= 16;
    if (8 * total > (long long)wcap) ovf |= 8;
    if ((long long)t0 + total > (long long)tcap) ovf |= 32;
    if (blockIdx.x == 0) {
        for (int j = tid; j < njobs; j += 256) {
            lj[j].w0 = 8 * base[j];
            lj[j].nwk = ovf ? 0 : 8 * lj[j].splits;
        }
        if (tid == 0) {
            const int nn = ovf ? 0 : (int)(8 * total);
            ctl->overflow = ovf;
            *nw_next = nn;
            ctl->iter[d].nw_next = nn;
            ctl->iter[d].overflow = ovf;
            host_iter[d].nw_next = nn;
            host_iter[d].overflow = ovf;
            ctl->tofs[d + 1] = t0 + (ovf ? 0 : (int)total);
        }
    }
    if (ovf) return;
    const int n = *nwp;
    for (int i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        const int c0 = cidx[i];
        if (c0 < 0) continue;
        const WorkItem wi = work[i];
        const int j = wi.job;
        const int first = lj[j].n - 8 * lj[j].splits;  // the job's node count before this iteration
        const int rank = (c0 - first) >> 3;
        const int p = 8 * (base[j] + rank);
        titems[t0 + base[j] + rank] = TransformItem{jt[j].outCoef, jt[j].outNorm, wi.node, c0};
        for (int a = 0; a < 8; a++) {
            WorkItem ch;
            ch.job = j;
            ch.node = c0 + a;
            ch.l[0] = 2 * wi.l[0] + (a & 1);
            ch.l[1] = 2 * wi.l[1] + ((a >> 1) & 1);
            ch.l[2] = 2 * wi.l[2] + ((a >> 2) & 1);
            ch.parent = wi.node;
            next[p + a] = ch;
        }
    }
}

// batched two-scale transforms over (tree, parent) items:
// mode 0 compress children -> parent (and the parent's 8 component norms),
// mode 1 add reconstruct(parent) to the children's scaling parts (and their norms)
template <int K>
__global__ void __launch_bounds__(256) k_transform(int mode, int n, const TransformItem *items, const double *M) {
    using L = PadT<K>;
    constexpr int N = L::N, K3 = K * K * K;
    extern __shared__ double smem[];
    __shared__ double wsum[4][8];
    double *T = smem;
    double *Ms = smem + L::SIZE;
    const TransformItem it = items[blockIdx.x];
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) Ms[e] = M[e];
    if (mode == 0) {
        for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x) {
            const int a = e / K3, r = e % K3;
            T[padded_index<K>(a, r)] = it.coef[(size_t)(it.first_child + a) * 8 * K3 + r];
        }
    } else {
        for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x)
            T[padded_index<K>(e / K3, e % K3)] = it.coef[(size_t)it.parent * 8 * K3 + e];
    }
    __syncthreads();
    for (int dim = 0; dim < 3; dim++) {
        line_pass_padded<N>(T, Ms, dim);
        __syncthreads();
    }
    double ss[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        ss[c] = 0.0;
        if (mode == 0) {
            double *dst = it.coef + (size_t)it.parent * 8 * K3 + c * K3;
            for (int r = threadIdx.x; r < K3; r += blockDim.x) {
                const double v = T[padded_index<K>(c, r)];
                dst[r] = v;
                ss[c] += v * v;
            }
        } else {
            double *dst = it.coef + (size_t)(it.first_child + c) * 8 * K3;
            for (int r = threadIdx.x; r < K3; r += blockDim.x) {
                const double v = dst[r] + T[padded_index<K>(c, r)];
                dst[r] = v;
                ss[c] += v * v;
            }
        }
    }
    // the 8 sums over the block at once (fixed order: lanes by butterfly, then waves 0..3)
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) ss[c] += __shfl_xor(ss[c], o, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 8; c++) wsum[threadIdx.x >> 6][c] = ss[c];
    __syncthreads();
    if (threadIdx.x < 8 && it.norm) {
        const int c = threadIdx.x;
        const double tot = (wsum[0][c] + wsum[1][c]) + (wsum[2][c] + wsum[3][c]);
        if (mode == 0) it.norm[(size_t)it.parent * 8 + c] = sqrt(tot);
        else it.norm[(size_t)(it.first_child + c) * 8] = sqrt(tot);
    }
    (void)n;
}


// ----------------------------------------------------------------- tree compaction

// node groups (8 siblings) moved inside a tree: pair i copies the group at node
// from_to[2i] to node from_to[2i+1], one block per node of the group (nb coefficient
// doubles and 8 norms per node); no source of a pair is the destination of another
__global__ void __launch_bounds__(256) k_move_groups(int npairs, const int *from_to, double *coef, double *norm,
                                                     int nb) {
    const int i = blockIdx.x, a = blockIdx.y;
    if (i >= npairs) return;
    const size_t f = (size_t)from_to[2 * i] + a, t = (size_t)from_to[2 * i + 1] + a;
    for (int e = threadIdx.x; e < nb; e += blockDim.x) coef[t * nb + e] = coef[f * nb + e];
    if (threadIdx.x < 8) norm[t * 8 + threadIdx.x] = norm[f * 8 + threadIdx.x];
}

hipError_t launch_move_groups(hipStream_t s, int npairs, const int *from_to, double *coef, double *norm, int nb) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_move_groups, dim3(npairs, 8), dim3(256), 0, s, npairs, from_to, coef, norm, nb);
    return hipGetLastError();
}

// ----------------------------------------------------------------- pointwise multiplication

// Loads the representation of an input tree at an output node into the padded node
// tensor T: the node's 8 components when it exists (steps = 0), else the scaling block
// refined from the leaf ancestor `src` over `steps` levels (path bits from l) with zero
// wavelet components.
template <int K>
__device__ void load_rep_padded(double *T, const double *coef, int src, int steps, const int *l, const double *Ws) {
    using L = PadT<K>;
    constexpr int N = L::N, P = L::P, K3 = K * K * K;
    const double *p = coef + (size_t)src * 8 * K3;
    for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x) T[padded_index<K>(e / K3, e % K3)] = p[e];
    __syncthreads();
    if (steps == 0) return;
    descend_padded<K>(T, Ws, l, steps);
    for (int e = threadIdx.x; e < N * N * N; e += blockDim.x) {
        const int x = e % N, y = (e / N) % N, z = e / (N * N);
        if (x >= K || y >= K || z >= K) T[(z * N + y) * P + x] = 0.0;
    }
    __syncthreads();
}

// out = coef * a * b at the listed output nodes: both inputs to values at the K^3
// quadrature points of the 8 children, product, quadrature projection, compression
template <int K>
__global__ void __launch_bounds__(256) k_multiply(int n, const MulItem *items, ProjGeom geo, DevBasis B, double coef,
                                                  const double *acoef, const double *bcoef, double *out) {
    using L = PadT<K>;
    constexpr int N = L::N, P = L::P, N2 = N * N, K3 = K * K * K;
    extern __shared__ double smem[];
    double *Ta = smem;
    double *Tb = Ta + L::SIZE;
    double *Ws = Tb + L::SIZE;
    double *WTs = Ws + N2;
    double *Vs = WTs + N2;
    double *Ps = Vs + N2;
    const MulItem it = items[blockIdx.x];
    for (int e = threadIdx.x; e < N2; e += blockDim.x) {
        const int r = e / N, c = e % N;
        Ws[e] = B.W[e];
        WTs[e] = B.W[c * N + r];
        const bool same = (r / K) == (c / K);
        Vs[e] = same ? B.phiq[(r % K) * K + (c % K)] : 0.0;  // values: row = point q, col = function i
        Ps[e] = same ? B.phiq[(c % K) * K + (r % K)] : 0.0;  // coefficients: row = function i, col = point q
    }
    __syncthreads();
    load_rep_padded<K>(Ta, acoef, it.asrc, it.asteps, it.l, Ws);
    load_rep_padded<K>(Tb, bcoef, it.bsrc, it.bsteps, it.l, Ws);
    for (int dim = 0; dim < 3; dim++) {  // reconstruct to the children's scaling coefficients
        line_pass_padded<N>(Ta, WTs, dim);
        line_pass_padded<N>(Tb, WTs, dim);
        __syncthreads();
    }
    for (int dim = 0; dim < 3; dim++) {  // values at the quadrature points
        line_pass_padded<N>(Ta, Vs, dim);
        line_pass_padded<N>(Tb, Vs, dim);
        __syncthreads();
    }
    const double h = geo.sfac * ldexp(1.0, -(geo.n0 + it.depth + 1));
    const double scale = coef / (h * sqrt(h));
    for (int e = threadIdx.x; e < N * N2; e += blockDim.x) {
        const int x = e % N, y = (e / N) % N, z = e / N2;
        const int i = (z * N + y) * P + x;
        Ta[i] = Ta[i] * Tb[i] * scale * B.qw[x % K] * B.qw[y % K] * B.qw[z % K];
    }
    __syncthreads();
    for (int dim = 0; dim < 3; dim++) {  // projection onto the children's scaling functions
        line_pass_padded<N>(Ta, Ps, dim);
        __syncthreads();
    }
    for (int dim = 0; dim < 3; dim++) {  // compression to the node's 8 components
        line_pass_padded<N>(Ta, Ws, dim);
        __syncthreads();
    }
    for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x)
        out[(size_t)it.node * 8 * K3 + e] = Ta[padded_index<K>(e / K3, e % K3)];
    (void)n;
}

// out = ca * a + cb * b at the listed output nodes
template <int K>
__global__ void __launch_bounds__(256) k_add(int n, const MulItem *items, DevBasis B, double ca,
                                             const double *acoef, double cb, const double *bcoef, double *out) {
    using L = PadT<K>;
    constexpr int N = L::N, K3 = K * K * K;
    extern __shared__ double smem[];
    double *Ta = smem;
    double *Tb = Ta + L::SIZE;
    double *Ws = Tb + L::SIZE;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) Ws[e] = B.W[e];
    __syncthreads();
    const MulItem it = items[blockIdx.x];
    load_rep_padded<K>(Ta, acoef, it.asrc, it.asteps, it.l, Ws);
    load_rep_padded<K>(Tb, bcoef, it.bsrc, it.bsteps, it.l, Ws);
    for (int e = threadIdx.x; e < 8 * K3; e += blockDim.x) {
        const int idx = padded_index<K>(e / K3, e % K3);
        out[(size_t)it.node * 8 * K3 + e] = ca * Ta[idx] + cb * Tb[idx];
    }
    (void)n;
}

// ----------------------------------------------------------------- launchers

// kernels with more than 64 KiB of dynamic LDS must say so before their launch
template <class Kern>
static hipError_t allow_smem(Kern *k, size_t sm) {
    if (sm <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
}

#define MWG_DISPATCH_K(K, CALL)                                       \
    switch (K) {                                                      \
        case 4: { constexpr int KK = 4; CALL; } break;                \
        case 5: { constexpr int KK = 5; CALL; } break;                \
        case 6: { constexpr int KK = 6; CALL; } break;                \
        case 7: { constexpr int KK = 7; CALL; } break;                \
        case 8: { constexpr int KK = 8; CALL; } break;                \
        case 9: { constexpr int KK = 9; CALL; } break;                \
        case 10: { constexpr int KK = 10; CALL; } break;              \
        default: return hipErrorInvalidValue;                         \
    }

hipError_t launch_tblock(hipStream_t s, int n, const int *item_term, const long long *item_delta, const double *b,
                         double pre, const DevBasis &B, const double *gx, const double *gw, double *out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tblock, dim3(n), dim3(128), 0, s, n, item_term, item_delta, b, pre, B, gx, gw, out);
    return hipGetLastError();
}

hipError_t launch_tblock_mom(hipStream_t s, int n, const int *item_term, const long long *item_delta,
                             const double *b, double pre, int K, const double *ccoef, const double *gx,
                             const double *gw, double *out) {
    if (n <= 0) return hipSuccess;
    MWG_DISPATCH_K(K, hipLaunchKernelGGL(k_tblock_mom<KK>, dim3(n), dim3(64), 0, s, n, item_term, item_delta, b, pre,
                                         ccoef, gx, gw, out));
    return hipGetLastError();
}

hipError_t launch_nsblocks(hipStream_t s, int n, const long long *item_delta, const int *item_row0,
                           const int *item_tlo, const int *item_thi, const double *t1, const double *W, int K,
                           double *blocks, double *norms, double *norms2) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nsblocks, dim3(n), dim3(256), 0, s, n, item_delta, item_row0, item_tlo, item_thi, t1, W, K,
                       blocks, norms, norms2);
    return hipGetLastError();
}

static size_t tensor_smem(int K) { return (size_t)(8 * K * K * K + 4 * K * K) * sizeof(double); }

hipError_t launch_norms(hipStream_t s, int n, const int *nodes, const double *coef, double *norm, int K3) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_norms, dim3(n), dim3(64), 0, s, n, nodes, coef, norm, K3);
    return hipGetLastError();
}

hipError_t launch_dot_pairs(hipStream_t s, int n, const unsigned long long *pairs, const double *A,
                            const double *B, int K3, int nblocks, double *part, double *res) {
    if (n < 0 || nblocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dot_pairs, dim3(nblocks), dim3(256), 0, s, n, pairs, A, B, K3, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, s, nblocks, part, res);
    return hipGetLastError();
}

hipError_t launch_dot(hipStream_t s, int n, const int *ia, const int *ib, const int *c0, const double *A,
                      const double *B, int K3, double *part) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dot, dim3(n), dim3(64), 0, s, n, ia, ib, c0, A, B, K3, part);
    return hipGetLastError();
}

hipError_t launch_scale(hipStream_t s, size_t n, double *x, double c) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, c);
    return hipGetLastError();
}

hipError_t launch_project_gauss(hipStream_t s, int nw, const int *work, const int *depth, const int *lr,
                                const ProjGeom &geo, const DevBasis &B, int nterm, const double *coef,
                                const double *alpha, const double *center, const double *poly, double *out) {
    if (nw <= 0) return hipSuccess;
    MWG_DISPATCH_K(B.K, hipLaunchKernelGGL(k_project_gauss<KK>, dim3(nw), dim3(256), 0, s, nw, work, depth, lr, geo,
                                           B, nterm, coef, alpha, center, poly, out));
    return hipGetLastError();
}

hipError_t launch_project_hyd(hipStream_t s, int nw, const int *work, const int *depth, const int *lr,
                              const ProjGeom &geo, const DevBasis &B, int nterm, const double *coef, const int *q,
                              double Z, const double *origin, double *out) {
    if (nw <= 0) return hipSuccess;
    const int K = B.K, N = 2 * K;
    const size_t sm = (size_t)(N * N * (N + 1) + 2 * N * N) * sizeof(double);
    MWG_DISPATCH_K(K, {
        hipError_t e = allow_smem(k_project_hyd<KK>, sm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_project_hyd<KK>, dim3(nw), dim3(256), sm, s, nw, work, depth, lr, geo, B, nterm, coef,
                           q, Z, origin, out);
    });
    return hipGetLastError();
}

hipError_t launch_job_boxes(hipStream_t s, int nbj, int max_terms, const BoxJob *bj, TermTab *tabs, double *tbmax,
                            JobTab *jobs, int *pre, int *pre0, int *gen_overflow, const LoopInit &li) {
    if (nbj <= 0 || max_terms <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_job_boxes, dim3(nbj, max_terms), dim3(64), 0, s, nbj, bj, tabs, tbmax, max_terms, li);
    hipLaunchKernelGGL(k_job_route, dim3(nbj), dim3(64), 0, s, nbj, bj, tabs, tbmax, max_terms, jobs, pre, pre0,
                       gen_overflow, li);
    return hipGetLastError();
}

hipError_t launch_gen_mark(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs, const TermTab *tabs,
                           const GenState &gs, const int *nwp, int *queue) {
    if (nw <= 0) return hipSuccess;
    const int nb = nwp ? std::min(nw, DEVCOUNT_GRID) : nw;
    hipLaunchKernelGGL(k_gen_mark, dim3(nb), dim3(256), 0, s, nw, nwp, queue, work, jobs, tabs, gs);
    return hipGetLastError();
}

// padded node tensor + one (2K)^2 filter matrix
static size_t padded_smem(int K) { return (size_t)(2 * K * 2 * K * (2 * K + 1) + 4 * K * K) * sizeof(double); }

hipError_t launch_gen_compute(hipStream_t s, int first, int cap, const JobTab *jobs, const GenState &gs,
                              const double *W, int K, const int *firstp) {
    if (cap <= first) return hipSuccess;
    const size_t sm = padded_smem(K);
    const int nb = std::min(cap - first, 2048);
    MWG_DISPATCH_K(K, {
        hipError_t e = allow_smem(k_gen_compute<KK>, sm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_gen_compute<KK>, dim3(nb), dim3(256), sm, s, first, firstp, cap, jobs, gs, W);
    });
    return hipGetLastError();
}

hipError_t launch_enum(hipStream_t s, int mode, int nw, int nparts, const WorkItem *work, const JobTab *jobs,
                       const TermTab *tabs, const GenState &gs, int *thread_off, int *seg_total,
                       const long long *seg_ofs, Tuple *out, const TaskSummary *sum, const SegRec *segs,
                       const int *nwp, int *queue) {
    if (nw <= 0) return hipSuccess;
    const unsigned nb = nwp ? (unsigned)DEVCOUNT_GRID : (unsigned)nw * (unsigned)nparts;
    if (mode == 0)
        hipLaunchKernelGGL(k_enum<0>, dim3(nb), dim3(256), 0, s, nw, nwp, queue, nparts, work, jobs, tabs, gs,
                           thread_off, seg_total, seg_ofs, out, sum, segs);
    else
        hipLaunchKernelGGL(k_enum<1>, dim3(nb), dim3(256), 0, s, nw, nwp, queue, nparts, work, jobs, tabs, gs,
                           thread_off, seg_total, seg_ofs, out, sum, segs);
    return hipGetLastError();
}

hipError_t launch_work_norms(hipStream_t s, int nw, const WorkItem *work, const JobTab *jobs, int K3,
                             double *compact, const int *nwp, unsigned long long *ts_end) {
    if (nw <= 0) return hipSuccess;
    const int nb = nwp ? std::min(nw, DEVCOUNT_GRID) : nw;
    hipLaunchKernelGGL(k_work_norms, dim3(nb), dim3(64), 0, s, nw, nwp, work, jobs, K3, compact, ts_end);
    return hipGetLastError();
}

__global__ void k_stamp(unsigned long long *t) { *t = (unsigned long long)wall_clock64(); }

hipError_t launch_stamp(hipStream_t s, unsigned long long *t) {
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_loop_init(hipStream_t s, int njobs, int nr, const LoopJob *lj, const int *rootlr, WorkItem *work) {
    const int n = njobs * nr;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_loop_init, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, s, njobs, nr, lj, rootlr,
                       work);
    return hipGetLastError();
}

hipError_t launch_loop_split(hipStream_t s, int njobs, LoopJob *lj, const WorkItem *work, const double *wnorm,
                             int max_depth, LoopCtl *ctl, int *cidx, const long long *tabcnt,
                             unsigned long long *ts_end) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_loop_split, dim3(njobs), dim3(256), 0, s, njobs, lj, work, wnorm, max_depth, ctl, cidx,
                       tabcnt, ts_end);
    return hipGetLastError();
}

hipError_t launch_loop_emit(hipStream_t s, int njobs, LoopJob *lj, const WorkItem *work, const int *nwp,
                            const int *cidx, WorkItem *next, int *nw_next, int wcap, const TaskSummary *sum,
                            LoopCtl *ctl, int d, const JobTab *jt, TransformItem *titems, int tcap,
                            LoopIter *host_iter) {
    if (njobs <= 0 || njobs > LOOP_MAX_JOBS || d < 0 || d >= LOOP_MAX_DEPTH) return hipErrorInvalidValue;
    const int nb = std::max(1, std::min((wcap + 255) / 256, 1024));
    hipLaunchKernelGGL(k_loop_emit, dim3(nb), dim3(256), 0, s, njobs, lj, work, nwp, cidx, next, nw_next, wcap, sum,
                       ctl, d, jt, titems, tcap, host_iter);
    return hipGetLastError();
}

hipError_t launch_transform(hipStream_t s, int mode, int n, const TransformItem *items, const double *M, int K) {
    if (n <= 0) return hipSuccess;
    const size_t sm = padded_smem(K);
    MWG_DISPATCH_K(K, {
        hipError_t e = allow_smem(k_transform<KK>, sm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_transform<KK>, dim3(n), dim3(256), sm, s, mode, n, items, M);
    });
    return hipGetLastError();
}

hipError_t launch_build_tasks(hipStream_t s, int nw, int nseg, int nparts, int chunk, int wchunk, const int *tot,
                              TaskCaps caps, const int *gen_count, int gen_pcap, long long *ofs, ContractTask *tasks,
                              SliceTask *stasks, ReduceItem *reds, TaskSummary *sum, SegRec *segs, int *segcnt,
                              long long *gslice, const int *gen_overflow, const WorkItem *work, long long *tabcnt,
                              int ntab, const int *nwp) {
    hipLaunchKernelGGL(k_build_tasks, dim3(1), dim3(1024), 0, s, nw, nwp, nseg, nparts, chunk, wchunk, tot, caps,
                       gen_count, gen_pcap, ofs, tasks, stasks, reds, sum, segs, segcnt, gslice, gen_overflow, work,
                       tabcnt, ntab);
    return hipGetLastError();
}

hipError_t launch_contract_fp64(hipStream_t s, int K, int ntask, const TaskSummary *sum, int grid,
                                const ContractTask *tasks, const WorkItem *work, const JobTab *jobs,
                                const TermTab *tabs, const Tuple *tuples, const double *genPool, double *partials,
                                unsigned long long *ts) {
    const int nb = sum ? grid : (ntask + 3) / 4;
    if (nb <= 0) return hipSuccess;
    MWG_DISPATCH_K(K, hipLaunchKernelGGL(k_contract_fp64<KK>, dim3(nb + (ts ? 1 : 0)), dim3(256), 0, s, ntask, sum,
                                         tasks, work, jobs, tabs, tuples, genPool, partials, ts));
    return hipGetLastError();
}

hipError_t launch_reduce_partials(hipStream_t s, int n, const TaskSummary *sum, int grid, const ReduceItem *items,
                                  const WorkItem *work, const JobTab *jobs, const double *partials,
                                  const double *nodeslots, int K3, double *wnorm) {
    const int nb = sum ? grid : n;
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3(nb), dim3(256), 0, s, n, sum, items, work, jobs, partials, nodeslots,
                       K3, wnorm);
    return hipGetLastError();
}

hipError_t launch_enum_mixed(hipStream_t s, int mode, int K, int nw, int nparts, const WorkItem *work, const JobTab *jobs,
                             const TermTab *tabs, const GenState &gs, int *thread_off, int *seg_total,
                             const long long *seg_ofs, Tuple *out, Slice *sout, const TaskSummary *sum,
                             const SegRec *segs, const long long *gslice, const int *nwp, int *queue) {
    if (nw <= 0) return hipSuccess;
    const unsigned nb = nwp ? (unsigned)DEVCOUNT_GRID : (unsigned)nw * (unsigned)nparts;
    if (mode == 0)
        hipLaunchKernelGGL(k_enum_mixed<0>, dim3(nb), dim3(256), 0, s, nw, nwp, queue, nparts, K * K, work, jobs,
                           tabs, gs, thread_off, seg_total, seg_ofs, out, sout, sum, segs, gslice);
    else
        hipLaunchKernelGGL(k_enum_mixed<1>, dim3(nb), dim3(256), 0, s, nw, nwp, queue, nparts, K * K, work, jobs,
                           tabs, gs, thread_off, seg_total, seg_ofs, out, sout, sum, segs, gslice);
    return hipGetLastError();
}

hipError_t launch_multiply(hipStream_t s, int n, const MulItem *items, const ProjGeom &geo, const DevBasis &B,
                           double coef, const double *acoef, const double *bcoef, double *out) {
    if (n <= 0) return hipSuccess;
    const int K = B.K, N = 2 * K;
    const size_t sm = (size_t)(2 * N * N * (N + 1) + 4 * N * N) * sizeof(double);
    MWG_DISPATCH_K(K, {
        hipError_t e = allow_smem(k_multiply<KK>, sm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_multiply<KK>, dim3(n), dim3(256), sm, s, n, items, geo, B, coef, acoef, bcoef, out);
    });
    return hipGetLastError();
}

hipError_t launch_add(hipStream_t s, int n, const MulItem *items, const DevBasis &B, double ca, const double *acoef,
                      double cb, const double *bcoef, double *out) {
    if (n <= 0) return hipSuccess;
    const int K = B.K, N = 2 * K;
    const size_t sm = (size_t)(2 * N * N * (N + 1) + N * N) * sizeof(double);
    MWG_DISPATCH_K(K, {
        hipError_t e = allow_smem(k_add<KK>, sm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_add<KK>, dim3(n), dim3(256), sm, s, n, items, B, ca, acoef, cb, bcoef, out);
    });
    return hipGetLastError();
}

}  // namespace mwg
