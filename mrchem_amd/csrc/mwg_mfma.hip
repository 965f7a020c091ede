// FP32 MFMA slice contractions of the mixed-precision apply (gfx950, wave64).
//
// A slice (g, f, ft, term) produces all 8 output components gt of node g from input
// component ft of node f: out_gt[c][b][a] += w * sum Z[c][l] Y[b][j] X[a][i] f[l][j][i],
// with X = X_{gt_x, ft_x} etc.  Per dimension the two output choices gt_d = 0, 1 are
// stacked into one 2K x K operator X' (row K gt_d + a), so a slice is three GEMMs.
// Products and the sums inside a slice are FP32 (v_mfma_f32_16x16x4_f32).  The slice
// result, scaled by the term weight and masked to the components whose tuples survive
// the screening, is summed over the task's slices (k = 7: FP32 over at most 16 slices
// per wave, then FP64 over the waves; k = 9: FP64 per slice), and a task's sum goes
// to one node slot (8 K^3 FP64) that k_reduce_partials adds, in slot order, to the
// node's FP64 tuple partials.  A task is a run of slices of one output node, one task
// per 256-thread workgroup.
//
// Instruction layout (v_mfma_f32_16x16x4_f32, D = A B + C): lane L supplies
// A[L % 16][L / 16] and B[L / 16][L % 16]; D[i][j] is register i % 4 of lane
// j + 16 (i / 4).
#include <hip/hip_ext.h>

#include "mwg_kernels.h"

namespace mwg {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef const double __attribute__((address_space(1))) gdbl;
typedef const float __attribute__((address_space(1))) gflt;

__device__ inline int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// orders a wave's LDS writes before its later LDS reads
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a slice record through a wave-uniform pointer (scalar loads of its two 16-byte halves)
__device__ inline Slice slice_at(const Slice *p) {
    const uint4 *q = (const uint4 *)p;
    const uint4 a = q[0], b = q[1];
    Slice s;
    s.w = __hiloint2double((int)a.y, (int)a.x);
    s.f = (int)a.z;
    s.xo = (int)a.w;
    s.yo = (int)b.x;
    s.zo = (int)b.y;
    s.ft = (uint8_t)(b.z & 0xffu);
    s.mask = (uint8_t)((b.z >> 8) & 0xffu);
    return s;
}

__device__ inline const gdbl *slice_input(const Slice &sl, const double *inCoef, const double *genPool, int K3) {
    return sl.f >= 0 ? (const gdbl *)inCoef + (size_t)sl.f * 8 * K3 + sl.ft * K3
                     : (const gdbl *)genPool + (size_t)(-sl.f - 1) * K3;
}

__device__ inline floatx4 mfma(float a, float b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline floatx4 mfma0(float a, float b) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
}

// ------------------------------------------------------------------ K = 8 (k = 7)
//
// One workgroup of 4 waves per task = a chunk of the slices of one output node, the
// waves taking every fourth slice.  A slice
// (f, ft, term, delta) produces all 8 output components with three GEMMs on
// v_mfma_f32_16x16x4_f32, M = 16 = (gt_d, a) stacked [X_{0,ft_d}; X_{1,ft_d}],
// K = 8, N = 64, 128, 256 (the dimensions already contracted).  Lane layout of
// the instruction: A[i][k] from lane i%16 + 16 k, B[k][j] from lane j%16 + 16 k,
// D[4 (lane/16) + r][lane%16] in register r.  Intermediates rotate through LDS
// (rows padded to 68 and 132 floats: conflict-free for the access patterns).
// Each slice result is scaled by the term weight (zero for components the slice
// does not reach) and summed in FP32 over the wave's slices (at most
// MAX_SLICE_CHUNK), then in FP64 over the waves and over the node's tasks.
// operands of one slice in registers: A fragments of the three passes and the
// pass-x B fragments of f, as loaded (FP64: converted to FP32 only when the slice
// is contracted, so that the loads for a later slice stay in flight meanwhile)
struct SliceRaw {
    float ax[2], ay[2], az[2];
    double fb[4][2];
    double w;
    int mask;
};

struct SliceOps {
    float ax[2], ay[2], az[2];
    float fb[4][2];
    double w;
    int mask;
};

__device__ inline SliceRaw load_slice_raw(const Slice &sl, const float *B32, const double *inCoef,
                                          const double *genPool, int m, int q) {
    constexpr int K = 8, K3 = 512;
    const int gxm = m >> 3, am = m & 7;
    SliceRaw o;
    const gflt *px = (const gflt *)B32 + sl.xo + gxm * 128 + am * K;
    const gflt *py = (const gflt *)B32 + sl.yo + gxm * 128 + am * K;
    const gflt *pz = (const gflt *)B32 + sl.zo + gxm * 128 + am * K;
#pragma unroll
    for (int s2 = 0; s2 < 2; s2++) {
        o.ax[s2] = px[4 * s2 + q];
        o.ay[s2] = py[4 * s2 + q];
        o.az[s2] = pz[4 * s2 + q];
    }
    const gdbl *f = sl.f >= 0 ? (const gdbl *)inCoef + (size_t)sl.f * 8 * K3 + sl.ft * K3
                                 : (const gdbl *)genPool + (size_t)(-sl.f - 1) * K3;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) o.fb[t][s2] = f[(16 * t + m) * K + 4 * s2 + q];
    o.w = sl.w;
    o.mask = sl.mask;
    return o;
}

__device__ inline SliceOps slice_ops(const SliceRaw &r) {
    SliceOps o;
#pragma unroll
    for (int s2 = 0; s2 < 2; s2++) {
        o.ax[s2] = r.ax[s2];
        o.ay[s2] = r.ay[s2];
        o.az[s2] = r.az[s2];
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) o.fb[t][s2] = (float)r.fb[t][s2];
    o.w = r.w;
    o.mask = r.mask;
    return o;
}

__device__ inline void mfma8_task(const SliceTask task, const WorkItem *work, const JobTab *jobs, const Slice *slices,
                                  const double *genPool, double *nodeslots, float *lds, long long *tabcnt) {
    constexpr int K = 8, K3 = 512;
    constexpr int S1 = 68, S2 = 132;
    const int wave = wave_index(), lane = threadIdx.x & 63;
    const WorkItem wi = work[task.g];
    const JobTab &J = jobs[wi.job];
    float *L1 = lds + wave * 16 * S1;
    float *L2 = lds + 4 * 16 * S1 + wave * 16 * S2;
    const int m = lane & 15, q = lane >> 4;
    const int gxm = m >> 3, am = m & 7;
    float acc[16][4];
#pragma unroll
    for (int t = 0; t < 16; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = 0.f;
    // the four waves take every fourth slice of the task; the next slice's operands
    // are loaded while the current one is contracted
    // software pipeline over the wave's slices it = wave, wave + 4, ...: the record of
    // slice it + 8 and the operands of slice it + 4 are loaded while slice it is contracted
    // (indices past the end are clamped to the last slice: the loads stay unconditional,
    // so that the waits for them fall where their data are used)
    const Slice *srec = slices + task.start;
    const int last = task.count - 1;  // a task holds at least one slice
    SliceRaw craw = load_slice_raw(slice_at(srec + min(wave, last)), J.operBlocks32, J.inCoef, genPool, m, q);
    Slice rnext = slice_at(srec + min(wave + 4, last));
    int ntup = 0;
    for (int it = wave; it < task.count; it += 4) {
        ntup += __popc((unsigned)craw.mask);
        const SliceRaw nraw = load_slice_raw(rnext, J.operBlocks32, J.inCoef, genPool, m, q);
        rnext = slice_at(srec + min(it + 8, last));
        const SliceOps cur = slice_ops(craw);
        // pass x: C1[(gx,a)][n = l*8 + j] = sum_i X[a][i] f[n*8 + i]
        // (tiles in groups of 4 independent MFMA chains, operands read before the MFMAs)
        {
            floatx4 c[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.ax[0], cur.fb[t][0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.ax[1], cur.fb[t][1], c[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) L1[(4 * q + r) * S1 + 16 * t + m] = c[t][r];
        }
        wave_sync();
        // pass y: C2[(gy,b)][n2 = l*16 + (gx,a)] = sum_j Y[b][j] C1[(gx,a)][l*8 + j]
#pragma unroll
        for (int g4 = 0; g4 < 2; g4++) {
            float b[4][2];
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int s2 = 0; s2 < 2; s2++) b[t][s2] = L1[m * S1 + (4 * g4 + t) * 8 + 4 * s2 + q];
            floatx4 c[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.ay[0], b[t][0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.ay[1], b[t][1], c[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) L2[(4 * q + r) * S2 + 16 * (4 * g4 + t) + m] = c[t][r];
        }
        wave_sync();
        // pass z: C3[(gz,c)][n3 = (gy,b)*16 + (gx,a)] = sum_l Z[c][l] C2[(gy,b)][l*16 + (gx,a)]
        // Output component gt = gx | gy << 1 | gz << 2 with gx = m >> 3 (the lane), gy = t3 >> 3
        // and gz = (4 q + r) >> 3 = q >> 1 (the lane): two weights per lane and slice, masked
        // by the slice's surviving components.
        const int gtl = gxm | ((q >> 1) << 2), gth = gtl | 2;
        const float wlow = ((cur.mask >> gtl) & 1) ? (float)cur.w : 0.f;
        const float whigh = ((cur.mask >> gth) & 1) ? (float)cur.w : 0.f;
        // tiles t3 < 8 carry gy = 0 (gt in {0,1,4,5}: mask bits 0x33), t3 >= 8 gy = 1 (0xCC);
        // a half without a surviving component is skipped (the mask is uniform per wave)
        const bool need_lo = (cur.mask & 0x33) != 0, need_hi = (cur.mask & 0xCC) != 0;
#pragma unroll
        for (int g4 = 0; g4 < 4; g4++) {
            if (g4 < 2 ? !need_lo : !need_hi) continue;
            float b[4][2];
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int s2 = 0; s2 < 2; s2++) b[t][s2] = L2[(4 * g4 + t) * S2 + (4 * s2 + q) * 16 + m];
            floatx4 c[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.az[0], b[t][0], floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) c[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur.az[1], b[t][1], c[t], 0, 0, 0);
            const float wm = g4 < 2 ? wlow : whigh;
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[4 * g4 + t][r] = fmaf(wm, c[t][r], acc[4 * g4 + t][r]);
        }
        wave_sync();
        craw = nraw;
    }
    if (lane == 0) atomicAdd((unsigned long long *)&tabcnt[4 * wi.job + 2], (unsigned long long)ntup);
    // FP64 sum over the waves in LDS, wave 0 first, then 1, 2, 3 (one 32 KB buffer)
    double *R = (double *)lds;
    __syncthreads();
    for (int src = 0; src < 4; src++) {
        if (wave == src) {
#pragma unroll
            for (int t3 = 0; t3 < 16; t3++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    double *e = &R[(t3 * 4 + r) * 64 + lane];
                    *e = src == 0 ? (double)acc[t3][r] : *e + (double)acc[t3][r];
                }
        }
        __syncthreads();
    }
    {
        double *dst = nodeslots + (size_t)task.slot * 8 * K3;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int t3 = 4 * wave + u;
            const int gy = t3 >> 3, b = t3 & 7;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m3 = 4 * q + r;
                const int gz = m3 >> 3, cc = m3 & 7;
                const int gt = gxm | (gy << 1) | (gz << 2);
                dst[gt * K3 + (cc * K + b) * K + am] = R[(t3 * 4 + r) * 64 + lane];
            }
        }
    }
}

__global__ void __launch_bounds__(256)
k_contract_mfma8(int ntask_arg, const TaskSummary *sum,
                                                        const SliceTask *tasks, const WorkItem *work,
                                                        const JobTab *jobs, const Slice *slices,
                                                        const double *genPool, double *nodeslots,
                                                        long long *tabcnt) {
    constexpr int S1 = 68, S2 = 132;
    // per-wave operand staging; reused for the cross-wave reduction at the end
    __shared__ __align__(16) float lds[4 * 16 * S1 + 4 * 16 * S2];
    const int ntask = sum ? (sum->overflow ? 0 : sum->nstasks) : ntask_arg;
    if ((int)blockIdx.x >= ntask) return;
    mfma8_task(tasks[blockIdx.x], work, jobs, slices, genPool, nodeslots, lds, tabcnt);
}


// ------------------------------------------------------------------ K = 10 (k = 9)
//
// One workgroup (4 waves) per task; the waves share every slice.  The stacked
// operators have 2K = 20 rows: M-tile 0 holds rows 0..15, M-tile 1 rows 16..19 at
// A-operand rows 0, 4, 8, 12, so that its results sit in register 0 of every lane
// (the other A rows are zero).  The contraction length 10 is three K-chunks of 4 (the
// last half zero).  Operands and pass results are staged in LDS with the contracted
// index innermost, padded to 12 with zeros (row stride 12 floats: the 16 columns of an
// N-tile fall in distinct banks):
//   F [n = 10 z + y][x]          f of the slice (double-buffered)
//   L1[n2 = 20 z + a'][y]        pass x result, B operand of pass y
//   L2[n3 = 20 b' + a'][z]       pass y result, B operand of pass z
// Per pass a wave takes the N-tiles wave + 4t of both M-tiles (7, 13 and 25 N-tiles),
// reads all their B fragments from LDS, then issues their MFMAs, then stores (or, in
// pass z, adds them in FP64 to its accumulators: 4 per lane and M-tile-0 tile, 1 per
// M-tile-1 tile).  Slice records are read 64 at a time into VGPRs (one per lane).
constexpr int M10_LD = 12;
constexpr int M10_F = 100 * M10_LD, M10_L1 = 200 * M10_LD, M10_L2 = 400 * M10_LD;

// f of a slice, four values per thread (1000 = 256 x 4 minus 24)
__device__ inline void load_f10(double (&fv)[4], const Slice &sl, const double *inCoef, const double *genPool) {
    const gdbl *f = slice_input(sl, inCoef, genPool, 1000);
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = threadIdx.x + 256 * u;
        fv[u] = e < 1000 ? f[e] : 0.0;
    }
}
__device__ inline void store_f10(float *F, const double (&fv)[4]) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = threadIdx.x + 256 * u;
        if (e < 1000) F[(e / 10) * M10_LD + e % 10] = (float)fv[u];
    }
}

// slice records of 64 consecutive slices, one per lane (two 16-byte halves)
struct RecBatch {
    uint4 a, b;
    int base;
};
__device__ inline void rec_load(RecBatch &rb, const Slice *srec, int base, int count, int lane) {
    const uint4 *p = (const uint4 *)(srec + min(base + lane, count - 1));
    rb.a = p[0];
    rb.b = p[1];
    rb.base = base;
}
__device__ inline Slice rec_get(const RecBatch &rb, int j) {
    const int l = j - rb.base;
    Slice s;
    s.w = __hiloint2double((int)__builtin_amdgcn_readlane(rb.a.y, l), (int)__builtin_amdgcn_readlane(rb.a.x, l));
    s.f = (int)__builtin_amdgcn_readlane(rb.a.z, l);
    s.xo = (int)__builtin_amdgcn_readlane(rb.a.w, l);
    s.yo = (int)__builtin_amdgcn_readlane(rb.b.x, l);
    s.zo = (int)__builtin_amdgcn_readlane(rb.b.y, l);
    const unsigned z = __builtin_amdgcn_readlane(rb.b.z, l);
    s.ft = (uint8_t)(z & 0xffu);
    s.mask = (uint8_t)((z >> 8) & 0xffu);
    return s;
}
struct Ops10c {
    float x[2][3], y[2][3], z[2][3];
};
__device__ inline Ops10c load_ops10c(const Slice &sl, const float *B32, int m, int q) {
    constexpr int K = 10, K2 = 100;
    Ops10c o;
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const int row = mt == 0 ? m : ((m & 3) == 0 ? 16 + (m >> 2) : 99);
        const bool rv = row < 20;
        const int roff = rv ? (row / 10) * 2 * K2 + (row % 10) * K : 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int col = 4 * c + q;
            const bool ok = rv && col < K;
            o.x[mt][c] = ok ? ((const gflt *)B32)[sl.xo + roff + col] : 0.f;
            o.y[mt][c] = ok ? ((const gflt *)B32)[sl.yo + roff + col] : 0.f;
            o.z[mt][c] = ok ? ((const gflt *)B32)[sl.zo + roff + col] : 0.f;
        }
    }
    return o;
}
// tiles t = 0..NT-1 of one M-tile: N-tile nt = wave + 4 t (< ntn), B fragments from LDS first
template <int NT>
__device__ inline void m10c_tiles(const float (&A)[3], const float *Lsrc, int nvalid_cols, int ntn, int wave, int m,
                                  int q, floatx4 (&d)[NT]) {
    float b[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int col = 16 * (wave + 4 * t) + m;
        const bool ok = wave + 4 * t < ntn && col < nvalid_cols;
#pragma unroll
        for (int c = 0; c < 3; c++) b[t][c] = ok ? Lsrc[col * M10_LD + 4 * c + q] : 0.f;
    }
#pragma unroll
    for (int t = 0; t < NT; t++) d[t] = mfma0(A[0], b[t][0]);
#pragma unroll
    for (int c = 1; c < 3; c++)
#pragma unroll
        for (int t = 0; t < NT; t++) d[t] = mfma(A[c], b[t][c], d[t]);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_contract_mfma10(int ntask_arg, const TaskSummary *sum, const SliceTask *tasks, const WorkItem *work,
                  const JobTab *jobs, const Slice *slices, const double *genPool, double *nodeslots,
                  long long *tabcnt) {
    constexpr int K3 = 1000;
    __shared__ __align__(16) float F[2][M10_F];
    __shared__ __align__(16) float L1[M10_L1];
    __shared__ __align__(16) float L2[M10_L2];
    const int ntask = sum ? (sum->overflow ? 0 : sum->nstasks) : ntask_arg;
    if ((int)blockIdx.x >= ntask) return;
    const SliceTask task = tasks[blockIdx.x];
    const int wave = wave_index(), lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
    const WorkItem wi = work[task.g];
    const JobTab &J = jobs[wi.job];
    for (int e = threadIdx.x; e < 2 * M10_F; e += 256) (&F[0][0])[e] = 0.f;
    for (int e = threadIdx.x; e < M10_L1; e += 256) L1[e] = 0.f;
    for (int e = threadIdx.x; e < M10_L2; e += 256) L2[e] = 0.f;
    double acc0[7][4], acc1[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
        acc1[t] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; r++) acc0[t][r] = 0.0;
    }
    const Slice *srec = slices + task.start;
    const int last = task.count - 1;
    Slice cs = slice_at(srec);
    double fv[4];
    load_f10(fv, cs, J.inCoef, genPool);
    __syncthreads();
    store_f10(F[0], fv);
    __syncthreads();
    Ops10c cur = load_ops10c(cs, J.operBlocks32, m, q);
    RecBatch rb;
    rec_load(rb, srec, 0, task.count, lane);
    Slice ns = rec_get(rb, min(1, last));
    int ntup = 0;
    for (int it = 0; it < task.count; it++) {
        float *Fc = F[it & 1];
        const int j2 = min(it + 2, last);
        if (j2 >= rb.base + 64) rec_load(rb, srec, j2, task.count, lane);
        const Slice n2 = rec_get(rb, j2);
        load_f10(fv, ns, J.inCoef, genPool);
        const Ops10c nxt = load_ops10c(ns, J.operBlocks32, m, q);
        ntup += __popc((unsigned)cs.mask);
        {   // pass x: M-tile 0 (rows a' = 4q + r), M-tile 1 (a' = 16 + q in register 0); 7 N-tiles
            floatx4 d0[2], d1[2];
            m10c_tiles<2>(cur.x[0], Fc, 100, 7, wave, m, q, d0);
            m10c_tiles<2>(cur.x[1], Fc, 100, 7, wave, m, q, d1);
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int n = 16 * (wave + 4 * t) + m;
                if (wave + 4 * t < 7 && n < 100) {
                    const int z = n / 10, y = n % 10;
#pragma unroll
                    for (int r = 0; r < 4; r++) L1[(20 * z + 4 * q + r) * M10_LD + y] = d0[t][r];
                    L1[(20 * z + 16 + q) * M10_LD + y] = d1[t][0];
                }
            }
        }
        __syncthreads();
        store_f10(F[(it + 1) & 1], fv);
        {   // pass y: 13 N-tiles (n2 = 20 z + a')
            floatx4 d0[4], d1[4];
            m10c_tiles<4>(cur.y[0], L1, 200, 13, wave, m, q, d0);
            m10c_tiles<4>(cur.y[1], L1, 200, 13, wave, m, q, d1);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int n2 = 16 * (wave + 4 * t) + m;
                if (wave + 4 * t < 13 && n2 < 200) {
                    const int z = n2 / 20, ap = n2 % 20;
#pragma unroll
                    for (int r = 0; r < 4; r++) L2[(20 * (4 * q + r) + ap) * M10_LD + z] = d0[t][r];
                    L2[(20 * (16 + q) + ap) * M10_LD + z] = d1[t][0];
                }
            }
        }
        __syncthreads();
        {   // pass z: 25 N-tiles (n3 = 20 b' + a'), FP64 accumulation
            const int mk = cs.mask;
            floatx4 d0[7];
            m10c_tiles<7>(cur.z[0], L2, 400, 25, wave, m, q, d0);
#pragma unroll
            for (int t = 0; t < 7; t++) {
                const int n3 = 16 * (wave + 4 * t) + m;
                const int gxy = (n3 % 20 >= 10 ? 1 : 0) | (n3 / 20 >= 10 ? 2 : 0);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int gt = gxy | (4 * q + r >= 10 ? 4 : 0);
                    const double ws = ((mk >> gt) & 1) ? cs.w : 0.0;
                    acc0[t][r] = fma((double)d0[t][r], ws, acc0[t][r]);
                }
            }
            floatx4 d1[7];
            m10c_tiles<7>(cur.z[1], L2, 400, 25, wave, m, q, d1);
#pragma unroll
            for (int t = 0; t < 7; t++) {
                const int n3 = 16 * (wave + 4 * t) + m;
                const int gt = (n3 % 20 >= 10 ? 1 : 0) | (n3 / 20 >= 10 ? 2 : 0) | 4;
                const double ws = ((mk >> gt) & 1) ? cs.w : 0.0;
                acc1[t] = fma((double)d1[t][0], ws, acc1[t]);
            }
        }
        cs = ns;
        ns = n2;
        cur = nxt;
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long *)&tabcnt[4 * wi.job + 2], (unsigned long long)ntup);
    double *dst = nodeslots + (size_t)task.slot * 8 * K3;
#pragma unroll
    for (int t = 0; t < 7; t++) {
        if (wave + 4 * t >= 25) break;
        const int n3 = 16 * (wave + 4 * t) + m;
        const int bp = n3 / 20, ap = n3 % 20;
        const int gxy = (ap >= 10 ? 1 : 0) | (bp >= 10 ? 2 : 0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int cp = 4 * q + r;
            const int gt = gxy | (cp >= 10 ? 4 : 0);
            dst[gt * K3 + ((cp % 10) * 10 + bp % 10) * 10 + ap % 10] = acc0[t][r];
        }
        const int cp = 16 + q;
        dst[(gxy | 4) * K3 + ((cp % 10) * 10 + bp % 10) * 10 + ap % 10] = acc1[t];
    }
}

}  // namespace

// one task per 256-thread block (its four waves share the task's slices)
int mfma_tasks_per_block(int K) {
    (void)K;
    return 1;
}

hipError_t launch_contract_mfma(hipStream_t s, int K, int ntask, const TaskSummary *sum, int grid,
                                const SliceTask *tasks, const WorkItem *work, const JobTab *jobs,
                                const TermTab *tabs, const Slice *slices, const double *genPool, double *nodeslots,
                                long long *tabcnt) {
    (void)tabs;
    if (K != 8 && K != 10) return hipErrorInvalidValue;
    const int tpb = mfma_tasks_per_block(K);
    const int nb = sum ? grid : (ntask + tpb - 1) / tpb;
    if (nb <= 0) return hipSuccess;
    if (K == 8)
        hipLaunchKernelGGL(k_contract_mfma8, dim3(nb), dim3(256), 0, s, ntask, sum, tasks, work, jobs, slices,
                           genPool, nodeslots, tabcnt);
    else
        hipLaunchKernelGGL(k_contract_mfma10, dim3(nb), dim3(256), 0, s, ntask, sum, tasks, work, jobs, slices,
                           genPool, nodeslots, tabcnt);
    return hipGetLastError();
}

}  // namespace mwg
