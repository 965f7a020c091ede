// MFMA slice contractions of the mixed-precision apply (gfx950, wave64).
//
// A slice (g, f, ft, term) produces all 8 output components gt of node g from input
// component ft of node f: out_gt[c][b][a] += w * sum Z[c][l] Y[b][j] X[a][i] f[l][j][i],
// with X = X_{gt_x, ft_x} etc.  Per dimension the two output choices gt_d = 0, 1 are
// stacked into one 2K x K operator X' (row K gt_d + a), so a slice is three GEMMs.
// k = 7: operands split into two FP16 halves each, products exact and sums FP32
// (v_mfma_f32_16x16x32_f16); k = 9: FP32 operands and sums (v_mfma_f32_16x16x4_f32).  The
// slice result, scaled by the term weight and masked to the components whose tuples
// survive the screening, is summed over the task's slices (k = 7: FP32 over at most 16
// slices per wave, then FP64 over the waves; k = 9: FP64 per slice), and a task's sum goes
// to one node slot (8 K^3 FP64) that k_reduce_partials adds, in slot order, to the
// node's FP64 tuple partials.  A task is a run of slices of one output node, one task
// per 256-thread workgroup.
//
// Instruction layout of v_mfma_f32_16x16x4_f32 (D = A B + C): lane L supplies
// A[L % 16][L / 16] and B[L / 16][L % 16]; D[i][j] is register i % 4 of lane j + 16 (i / 4).
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

// a slice record through a wave-uniform pointer: scalar loads of its two 16-byte halves
// (through the constant address space: the records are written by an earlier kernel and
// only read here, and the scalar cache starts each kernel clean)
typedef unsigned u32x8 __attribute__((ext_vector_type(8)));
typedef const u32x8 __attribute__((address_space(4))) cu32x8;
__device__ inline Slice slice_at(const Slice *base, int idx) {
    const u32x8 r = *(cu32x8 *)(base + __builtin_amdgcn_readfirstlane(idx));
    const uint4 a = make_uint4(r[0], r[1], r[2], r[3]), b = make_uint4(r[4], r[5], r[6], r[7]);
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

// ------------------------------------------------------------------ split-FP16 copies
//
// x 2^e = hi + lo: hi keeps the leading 11 significand bits of x 2^e (exact in f16 in its
// normal range), lo = x 2^e - hi is exact in FP32 and rounded to nearest in f16, so hi + lo
// carries about 22 bits.  Below the f16 normal range (2^-14 of the block's largest
// magnitude, which is scaled into [0.5, 1)) both parts round to the f16 subnormal grid,
// an absolute error of at most 2^-25 of the block's scale; v_mfma_f32_16x16x32_f16 keeps
// subnormal operands and forms every product exactly in FP32.

typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));

__device__ inline unsigned pack_rn(float a, float b) {
    const halfx2 h = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(unsigned, h);
}
// 2 K values (times the power of two sc) -> hi and lo halves (packed pairs)
template <int N>
__device__ inline void split_pairs(const float (&v)[N], float sc, unsigned (&hi)[N / 2], unsigned (&lo)[N / 2]) {
#pragma unroll
    for (int j = 0; j < N / 2; j++) {
        const float x0 = v[2 * j] * sc, x1 = v[2 * j + 1] * sc;
        const float a = __uint_as_float(__float_as_uint(x0) & 0xFFFFE000u);
        const float b = __uint_as_float(__float_as_uint(x1) & 0xFFFFE000u);
        hi[j] = pack_rn(a, b);
        lo[j] = pack_rn(x0 - a, x1 - b);
    }
}

// The same split in four instructions per pair, the power of two sc folded in:
// hi = f16(v sc) (v_fma_mix with addend 0), lo = f16(v sc - hi) (v_fma_mix with the f16 hi
// as negated addend; v sc is exact and the difference is formed exactly before its one
// rounding).  Both halves of a pair go into one register (mixlo / mixhi).
__device__ inline void split_pair_sc(float v0, float v1, float sc, unsigned &hi, unsigned &lo) {
    unsigned h, l;
    asm volatile("v_fma_mixlo_f16 %0, %1, %3, 0\n\t"
                 "v_fma_mixhi_f16 %0, %2, %3, 0"
                 : "=&v"(h)
                 : "v"(v0), "v"(v1), "v"(sc));
    asm volatile("v_fma_mixlo_f16 %0, %1, %3, -%4 op_sel_hi:[0,0,1]\n\t"
                 "v_fma_mixhi_f16 %0, %2, %3, -%4 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
                 : "=&v"(l)
                 : "v"(v0), "v"(v1), "v"(sc), "v"(h));
    hi = h;
    lo = l;
}

// largest value over the wave of a non-negative v (compared as unsigned bit patterns, which
// order non-negative floats like their values): across the lane halves and rows by
// permlane swaps (each returns the pair of values that meet), then within each row of 16
// by DPP rotations
__device__ inline float wave_max(float v) {
    unsigned u = __float_as_uint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = max((unsigned)a[0], (unsigned)a[1]);
    const auto b = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = max((unsigned)b[0], (unsigned)b[1]);
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x128, 0xf, 0xf, true));  // row_ror:8
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x124, 0xf, 0xf, true));  // row_ror:4
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x122, 0xf, 0xf, true));  // row_ror:2
    u = max(u, (unsigned)__builtin_amdgcn_mov_dpp((int)u, 0x121, 0xf, 0xf, true));  // row_ror:1
    return __uint_as_float(u);
}
// s with (largest magnitude) 2^s in [0.5, 1); 0 for an all-zero tensor
__device__ inline int scale_exp(float mx) { return mx > 0.f ? -(ilogbf(mx) + 1) : 0; }

// one wave per block; rows of K = 8 values (16 halves: hi | lo = two uint4)
__global__ void __launch_bounds__(256) k_split16(int n, int first, const int *count, int bvals, const double *src,
                                                 uint4 *dst, int *exp) {
    const int b = first + (int)blockIdx.x * 4 + wave_index(), lane = threadIdx.x & 63;
    const int end = count ? min(*(volatile const int *)count, first + n) : first + n;
    if (b >= end) return;
    const double *v = src + (size_t)b * bvals;
    double mx = 0.0;
    for (int e = lane; e < bvals; e += 64) mx = fmax(mx, fabs(v[e]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int ex = mx > 0.0 ? -(ilogb(mx) + 1) : 0;
    if (lane == 0) exp[b] = ex;
    const int rows = bvals / 8;
    for (int r = lane; r < rows; r += 64) {
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = (float)ldexp(v[r * 8 + j], ex);
        unsigned hi[4], lo[4];
        split_pairs<8>(x, 1.f, hi, lo);
        dst[((size_t)b * rows + r) * 2] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
        dst[((size_t)b * rows + r) * 2 + 1] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    }
}

// ------------------------------------------------------------------ K = 8 (k = 7)
//
// Split-FP16 slice contraction.  One v_mfma_f32_16x16x32_f16 per 16 x 16 output tile holds
// a whole K = 8 contraction with both operands split: its K = 32 is four groups of 8 (the
// lane group kg = lane / 16 of each operand), (A_hi, B_hi), (A_hi, B_lo), (A_lo, B_hi),
// (A_lo, B_lo).  Lane L supplies A[L % 16][8 kg .. 8 kg + 7] and B[8 kg .. 8 kg + 7][L % 16]
// as 8 halves: an A-role fragment is the hi row for kg 0, 1 and the lo row for kg 2, 3, a
// B-role fragment hi for kg 0, 2 and lo for kg 1, 3.  D[i][j] is register i % 4 of lane
// j + 16 (i / 4).
//  pass x  T1 = f X'^T: A = f (tile (yh, zq): row i holds y = 4 yh + i % 4, z = 4 zq + i / 4),
//          B = X' (rows a' = 8 gt_x + a).  Lane a' + 16 q then holds y = 0..7 of the columns
//          (z = 4 zq + q, a') across the tiles yh = 0, 1.
//  pass y  T2[z][b'][a'] = Y' T1: A = Y' (rows b' = 8 gt_y + b), B from LDS (per column
//          (z, a') the row [hi 8 | lo 8] of its y values, written by the lane that holds them).
//          Lane a' + 16 q then holds z = 0..7 of the columns (b' = 4 q + r, a').
//  pass z  out[c'][b'][a'] = Z' T2: A = Z', B from LDS rows [hi 8 | lo 8] per column (b', a').
// Output element (c' = 4 q + r, b', a' = lane % 16) of N-tile b': gt = (a' >> 3) |
// (b' >> 3) << 1 | (q >> 1) << 2.  The operators and inputs carry power-of-two exponents
// per block (x 2^e = hi + lo): the weight of output component gt absorbs 2^-(e_x(gt_x) +
// e_y(gt_y) + e_z(gt_z) + e_f), and that of the intermediates T1 and T2, each scaled per
// slice by 2^s1, 2^s2 to a largest magnitude in [0.5, 1) before it is split, 2^-(s1 + s2).
// Each wave sums its slices (every fourth of the task) in FP32, then the four wave sums are
// added in FP64 in LDS in a fixed order.
__device__ inline floatx4 mfmah(uint4 a, uint4 b) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8, a), __builtin_bit_cast(halfx8, b),
                                                 floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
}

struct OpsH8 {
    uint4 bx, ay, az;  // X' (B role), Y', Z' (A role) fragments of this lane
    uint4 f[4];        // f (A role) per tile (yh, zq)
    int exz;           // exponents of this lane's X' block (gt_x) + Z' block (gt_z) + f
    int ey0, ey1;      // the two Y' blocks
    double w;
    int mask;
};

// the job's split-FP16 arrays, read from the job table once per task (kept in registers:
// the LDS fences of the slice loop would otherwise make the compiler reload them per slice)
struct Arr16 {
    const uint4 *oper;
    const int *operExp;
    const uint4 *in;
    const int *inExp;
};

// global-address-space loads (not FLAT): the compiler then counts them exactly and waits
// only for the slice being contracted, not for the next slice's operands already in flight
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) gu32x4;
typedef const int __attribute__((address_space(4))) cint;
__device__ inline uint4 ld16(const uint4 *p) {
    const u32x4 v = *(gu32x4 *)p;
    return make_uint4(v[0], v[1], v[2], v[3]);
}
// a wave-uniform exponent (scalar load; written by an earlier kernel, read-only here)
__device__ inline int ld_exp(const int *p) { return *(cint *)p; }

__device__ inline OpsH8 load_opsh8(const Slice &sl, const Arr16 &J, const uint4 *gen16, const int *genExp,
                                   int m, int q) {
    OpsH8 o;
    const int bxb = sl.xo / 64, byb = sl.yo / 64, bzb = sl.zo / 64;  // block indices (gt_d = 0)
    const int rx = (m >> 3) * 2, row = m & 7;
    o.bx = ld16(J.oper + ((bxb + rx) * 8 + row) * 2 + (q & 1));
    o.ay = ld16(J.oper + ((byb + rx) * 8 + row) * 2 + (q >> 1));
    o.az = ld16(J.oper + ((bzb + rx) * 8 + row) * 2 + (q >> 1));
    const uint4 *fb;
    int ef;
    if (sl.f >= 0) {
        fb = J.in + ((size_t)sl.f * 8 + sl.ft) * 128;
        ef = ld_exp(J.inExp + (size_t)sl.f * 8 + sl.ft);
    } else {
        fb = gen16 + (size_t)(-sl.f - 1) * 128;
        ef = ld_exp(genExp + (-sl.f - 1));
    }
#pragma unroll
    for (int tl = 0; tl < 4; tl++) {
        const int y = 4 * (tl & 1) + (m & 3), z = 4 * (tl >> 1) + (m >> 2);
        o.f[tl] = ld16(fb + (z * 8 + y) * 2 + (q >> 1));
    }
    // block exponents: uniform loads, then this lane's gt_x (rx) and gt_z (q >> 1)
    const int ex0 = ld_exp(J.operExp + bxb), ex1 = ld_exp(J.operExp + bxb + 2);
    const int ez0 = ld_exp(J.operExp + bzb), ez1 = ld_exp(J.operExp + bzb + 2);
    o.exz = (rx ? ex1 : ex0) + ((q >> 1) ? ez1 : ez0) + ef;
    o.ey0 = ld_exp(J.operExp + byb);
    o.ey1 = ld_exp(J.operExp + byb + 2);
    o.w = sl.w;
    o.mask = sl.mask;
    return o;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_contract_mfma8(int ntask_arg, const TaskSummary *sum,
                                                        const SliceTask *tasks, const WorkItem *work,
                                                        const JobTab *jobs, const Slice *slices,
                                                        const uint4 *gen16, const int *genExp, double *nodeslots,
                                                        long long *tabcnt) {
    constexpr int K3 = 512;
    // per wave: L1 (8 z x 16 a' rows, 4 KB) | L2 (16 b' x 16 a' rows, 8 KB); reused at the end
    // for the FP64 sum over the waves (32 KB)
    __shared__ __align__(16) uint4 lds[4 * 768];
    const int ntask = sum ? (sum->overflow ? 0 : sum->nstasks) : ntask_arg;
    if ((int)blockIdx.x >= ntask) return;
    const SliceTask task = tasks[blockIdx.x];
    const int wave = wave_index(), lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
    // hi and lo in separate planes: a lane's 16-byte writes are contiguous across the wave
    uint4 *L1H = lds + wave * 768, *L1L = L1H + 128, *L2H = L1H + 256, *L2L = L2H + 256;
    const uint4 *L1r = (q & 1) ? L1L : L1H, *L2r = (q & 1) ? L2L : L2H;  // B role: hi for kg 0, 2
    const int gx = m >> 3, gz = q >> 1;
    const WorkItem wi = work[task.g];
    const Arr16 J{jobs[wi.job].operBlocks16, jobs[wi.job].operExp16, jobs[wi.job].inCoef16, jobs[wi.job].inExp16};
    float acc[16][4];
#pragma unroll
    for (int t = 0; t < 16; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = 0.f;
    // the waves take every fourth slice; the next slice's operands and the record of the
    // one after are loaded while a slice is contracted (indices clamped to the last slice)
    const Slice *srec = slices + task.start;
    const int last = task.count - 1;  // a task holds at least one slice
    int ntup = 0;
    // one slice: passes x, y, z and the weighted FP32 accumulation
    auto contract = [&](const OpsH8 &cur) {
            ntup += __popc((unsigned)cur.mask);
            floatx4 t1[4];
#pragma unroll
            for (int tl = 0; tl < 4; tl++) t1[tl] = mfmah(cur.f[tl], cur.bx);
            float m1 = 0.f;
#pragma unroll
            for (int tl = 0; tl < 4; tl++)
#pragma unroll
                for (int r = 0; r < 4; r++) m1 = fmaxf(m1, fabsf(t1[tl][r]));
            const int s1 = scale_exp(wave_max(m1));
            const float sc1 = ldexpf(1.f, s1);
#pragma unroll
            for (int zq = 0; zq < 2; zq++) {
                float v[8];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    v[r] = t1[2 * zq][r];
                    v[4 + r] = t1[2 * zq + 1][r];
                }
                unsigned hi[4], lo[4];
#pragma unroll
                for (int j = 0; j < 4; j++) split_pair_sc(v[2 * j], v[2 * j + 1], sc1, hi[j], lo[j]);
                const int z = 4 * zq + q;
                L1H[z * 16 + m] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
                L1L[z * 16 + m] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
            }
            wave_sync();
            floatx4 t2[8];
#pragma unroll
            for (int z = 0; z < 8; z++) t2[z] = mfmah(cur.ay, L1r[z * 16 + m]);
            float m2 = 0.f;
#pragma unroll
            for (int z = 0; z < 8; z++)
#pragma unroll
                for (int r = 0; r < 4; r++) m2 = fmaxf(m2, fabsf(t2[z][r]));
            const int s2 = scale_exp(wave_max(m2));
            const float sc2 = ldexpf(1.f, s2);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float v[8];
#pragma unroll
                for (int z = 0; z < 8; z++) v[z] = t2[z][r];
                unsigned hi[4], lo[4];
#pragma unroll
                for (int j = 0; j < 4; j++) split_pair_sc(v[2 * j], v[2 * j + 1], sc2, hi[j], lo[j]);
                const int bp = 4 * q + r;
                L2H[bp * 16 + m] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
                L2L[bp * 16 + m] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
            }
            wave_sync();
            // pass z: a half (gt_y) without a surviving component is skipped (uniform mask)
            const int mk = cur.mask;
            const float wlo = ((mk >> (gx | (gz << 2))) & 1) ? (float)ldexp(cur.w, -(cur.exz + cur.ey0 + s1 + s2)) : 0.f;
            const float whi = ((mk >> (gx | 2 | (gz << 2))) & 1) ? (float)ldexp(cur.w, -(cur.exz + cur.ey1 + s1 + s2)) : 0.f;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (h == 0 ? !(mk & 0x33) : !(mk & 0xCC)) continue;
                floatx4 d[8];
#pragma unroll
                for (int t = 0; t < 8; t++) d[t] = mfmah(cur.az, L2r[(8 * h + t) * 16 + m]);
                const float ws = h == 0 ? wlo : whi;
#pragma unroll
                for (int t = 0; t < 8; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[8 * h + t][r] = fmaf(ws, d[t][r], acc[8 * h + t][r]);
            }
            wave_sync();
    };
    // two operand sets in turn (no register copies between slices)
    OpsH8 opA = load_opsh8(slice_at(srec, min(wave, last)), J, gen16, genExp, m, q), opB;
    Slice rnext = slice_at(srec, min(wave + 4, last));
    for (int it = wave; it < task.count;) {
        opB = load_opsh8(rnext, J, gen16, genExp, m, q);
        rnext = slice_at(srec, min(it + 8, last));
        contract(opA);
        it += 4;
        if (it >= task.count) break;
        opA = load_opsh8(rnext, J, gen16, genExp, m, q);
        rnext = slice_at(srec, min(it + 8, last));
        contract(opB);
        it += 4;
    }
    if (lane == 0) atomicAdd((unsigned long long *)&tabcnt[4 * wi.job + 2], (unsigned long long)ntup);
    // FP64 sum over the waves in LDS, wave 0 first, then 1, 2, 3
    double *R = (double *)lds;
    __syncthreads();
    for (int src = 0; src < 4; src++) {
        if (wave == src) {
#pragma unroll
            for (int t = 0; t < 16; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    double *e = &R[(t * 4 + r) * 64 + lane];
                    *e = src == 0 ? (double)acc[t][r] : *e + (double)acc[t][r];
                }
        }
        __syncthreads();
    }
    double *dst = nodeslots + (size_t)task.slot * 8 * K3;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int bt = 4 * wave + u;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int gt = gx | ((bt >> 3) << 1) | (gz << 2);
            const int c = 4 * (q & 1) + r, b = bt & 7, a = m & 7;
            dst[gt * K3 + (c * 8 + b) * 8 + a] = R[(bt * 4 + r) * 64 + lane];
        }
    }
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
    Slice cs = slice_at(srec, 0);
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
                                long long *tabcnt, const uint4 *genPool16, const int *genExp16) {
    (void)tabs;
    if (K != 8 && K != 10) return hipErrorInvalidValue;
    const int nb = sum ? grid : ntask;
    if (nb <= 0) return hipSuccess;
    if (K == 8)
        hipLaunchKernelGGL(k_contract_mfma8, dim3(nb), dim3(256), 0, s, ntask, sum, tasks, work, jobs, slices,
                           genPool16, genExp16, nodeslots, tabcnt);
    else
        hipLaunchKernelGGL(k_contract_mfma10, dim3(nb), dim3(256), 0, s, ntask, sum, tasks, work, jobs, slices,
                           genPool, nodeslots, tabcnt);
    return hipGetLastError();
}

hipError_t launch_split16(hipStream_t s, int n, int first, const int *count, int bvals, int K, const double *src,
                          uint4 *dst, int *exp) {
    if (K != 8) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split16, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, first, count, bvals, src, dst,
                       exp);
    return hipGetLastError();
}

}  // namespace mwg
