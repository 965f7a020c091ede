// MFMA slice contractions of the mixed-precision apply (gfx950, wave64).
//
// A slice (g, f, ft, term) produces all 8 output components gt of node g from input
// component ft of node f: out_gt[c][b][a] += w * sum Z[c][l] Y[b][j] X[a][i] f[l][j][i],
// with X = X_{gt_x, ft_x} etc.  Per dimension the two output choices gt_d = 0, 1 are
// stacked into one 2K x K operator X' (row K gt_d + a), so a slice is three GEMMs.  The
// operands are split into FP16 hi + lo halves (about 22 significand bits) and multiplied on
// v_mfma_f32_16x16x32_f16, whose 32-long contraction holds the partial products hi hi,
// hi lo, lo hi (and at k = 7 lo lo) of one row pair; products are exact, sums FP32.  The
// slice result, scaled by the term weight and masked to the components whose tuples
// survive the screening, is summed over the task's slices (k = 7: FP32 over at most 64
// slices per wave, then FP64 over the waves; k = 9: FP64 per slice), and a task's sum goes
// to one node slot (8 K^3 FP64) that k_reduce_partials adds, in slot order, to the
// node's FP64 tuple partials.  A task is a run of slices of one output node, one task
// per 256-thread workgroup.
//
// Instruction layout of v_mfma_f32_16x16x32_f16 (D = A B + C): lane L supplies A[L % 16]
// [8 (L / 16) .. 8 (L / 16) + 7] and B[8 (L / 16) .. + 7][L % 16]; D[i][j] is register i % 4
// of lane j + 16 (i / 4).
#include <hip/hip_ext.h>

#include "mwg_kernels.h"

namespace mwg {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef const double __attribute__((address_space(1))) gdbl;

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
__device__ __forceinline__ void split16_block(int b, int bvals, const double *src, uint4 *dst, int *exp) {
    const int lane = threadIdx.x & 63;
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

// blocks first .. end - 1, one wave each, end = min(*count, first + n) (grid-stride).  With
// firstp, first is read from the device and n is the absolute end (the pool's capacity).
__global__ void __launch_bounds__(256) k_split16(int n, int first, const int *firstp, const int *count, int bvals,
                                                 const double *src, uint4 *dst, int *exp) {
    const int lim = firstp ? n : first + n;
    if (firstp) first = *firstp;
    const int end = count ? min(*(volatile const int *)count, lim) : lim;
    for (int b = first + (int)blockIdx.x * 4 + wave_index(); b < end; b += (int)gridDim.x * 4)
        split16_block(b, bvals, src, dst, exp);
}

// K = 10: one wave per block, one row of 10 values per lane -> 32 halves (four uint4)
__device__ __forceinline__ void split16_10_block(int b, int bvals, int role, const double *src, uint4 *dst,
                                                 int *exp) {
    const int lane = threadIdx.x & 63;
    const double *v = src + (size_t)b * bvals;
    double mx = 0.0;
    for (int e = lane; e < bvals; e += 64) mx = fmax(mx, fabs(v[e]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int ex = mx > 0.0 ? -(ilogb(mx) + 1) : 0;
    if (lane == 0) exp[b] = ex;
    const int rows = bvals / 10;
    for (int r = lane; r < rows; r += 64) {
        _Float16 hi[10], lo[10];
#pragma unroll
        for (int j = 0; j < 10; j++) {
            const float x = (float)ldexp(v[r * 10 + j], ex);
            hi[j] = (_Float16)x;
            lo[j] = (_Float16)(x - (float)hi[j]);
        }
        if (role == 0) {  // [hi | hi | lo | 0 0]: four uint4
            _Float16 h[32];
#pragma unroll
            for (int j = 0; j < 10; j++) {
                h[j] = hi[j];
                h[10 + j] = hi[j];
                h[20 + j] = lo[j];
            }
            h[30] = h[31] = (_Float16)0.f;
            unsigned d[16];
#pragma unroll
            for (int q = 0; q < 16; q++) d[q] = __builtin_bit_cast(unsigned, halfx2{h[2 * q], h[2 * q + 1]});
            uint4 *o = dst + ((size_t)b * rows + r) * 4;
#pragma unroll
            for (int q = 0; q < 4; q++) o[q] = make_uint4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
        } else {  // [hi | lo]: five uint2 (the B fragments are assembled from it)
            unsigned d[10];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                d[q] = __builtin_bit_cast(unsigned, halfx2{hi[2 * q], hi[2 * q + 1]});
                d[5 + q] = __builtin_bit_cast(unsigned, halfx2{lo[2 * q], lo[2 * q + 1]});
            }
            uint2 *o = (uint2 *)dst + ((size_t)b * rows + r) * 5;
#pragma unroll
            for (int q = 0; q < 5; q++) o[q] = make_uint2(d[2 * q], d[2 * q + 1]);
        }
    }
}

__global__ void __launch_bounds__(256) k_split16_10(int n, int first, const int *firstp, const int *count, int bvals,
                                                    int role, const double *src, uint4 *dst, int *exp) {
    const int lim = firstp ? n : first + n;
    if (firstp) first = *firstp;
    const int end = count ? min(*(volatile const int *)count, lim) : lim;
    for (int b = first + (int)blockIdx.x * 4 + wave_index(); b < end; b += (int)gridDim.x * 4)
        split16_10_block(b, bvals, role, src, dst, exp);
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
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef const u32x2 __attribute__((address_space(1))) gu32x2;
__device__ inline uint2 ld8(const uint2 *p) {
    const u32x2 v = *(gu32x2 *)p;
    return make_uint2(v[0], v[1]);
}
// a wave-uniform exponent (scalar load; written by an earlier kernel, read-only here)
__device__ inline int ld_exp(const int *p) { return *(cint *)p; }

// a lane's fragment offsets (uint4 units) inside an operator block pair and an f block: the
// same for every slice, so that a load is a uniform base (scalar registers) plus a lane offset
struct LaneOffs8 {
    unsigned bx, ayz, f[4];
};
__device__ inline LaneOffs8 lane_offs8(int m, int q) {
    const int rx = (m >> 3) * 2, row = m & 7;
    LaneOffs8 L;
    L.bx = (unsigned)((rx * 8 + row) * 2 + (q & 1));
    L.ayz = (unsigned)((rx * 8 + row) * 2 + (q >> 1));
#pragma unroll
    for (int tl = 0; tl < 4; tl++) {
        const int y = 4 * (tl & 1) + (m & 3), z = 4 * (tl >> 1) + (m >> 2);
        L.f[tl] = (unsigned)((z * 8 + y) * 2 + (q >> 1));
    }
    return L;
}

__device__ inline OpsH8 load_opsh8(const Slice &sl, const Arr16 &J, const uint4 *gen16, const int *genExp,
                                   int m, int q, const LaneOffs8 &L) {
    OpsH8 o;
    const int bxb = sl.xo / 64, byb = sl.yo / 64, bzb = sl.zo / 64;  // block indices (gt_d = 0)
    const int rx = (m >> 3) * 2;
    o.bx = ld16(J.oper + (size_t)bxb * 16 + L.bx);
    o.ay = ld16(J.oper + (size_t)byb * 16 + L.ayz);
    o.az = ld16(J.oper + (size_t)bzb * 16 + L.ayz);
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
    for (int tl = 0; tl < 4; tl++) o.f[tl] = ld16(fb + L.f[tl]);
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
    const LaneOffs8 LO = lane_offs8(m, q);
    OpsH8 opA = load_opsh8(slice_at(srec, min(wave, last)), J, gen16, genExp, m, q, LO), opB;
    Slice rnext = slice_at(srec, min(wave + 4, last));
    for (int it = wave; it < task.count;) {
        opB = load_opsh8(rnext, J, gen16, genExp, m, q, LO);
        rnext = slice_at(srec, min(it + 8, last));
        contract(opA);
        it += 4;
        if (it >= task.count) break;
        opA = load_opsh8(rnext, J, gen16, genExp, m, q, LO);
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

// ------------------------------------------------------------------ K = 10 (k = 9), split-FP16
//
// One v_mfma_f32_16x16x32_f16 holds a contraction of length 10 with split operands: K-slots
// 0..9 carry (A_hi, B_hi), 10..19 (A_hi, B_lo), 20..29 (A_lo, B_hi), 30..31 zero (the lo lo
// products lie below the split's own rounding).  Lane L supplies the 8 K-slots 8 kg ..
// 8 kg + 7 (kg = L / 16): operator rows (A operands) are stored as [hi | hi | lo | 0 0], 32
// halves, so that a lane's fragment is the 16-byte chunk kg of its row; input rows (B
// operands, [hi | lo | hi | 0 0] in K-slot order) are stored as [hi | lo] and a lane
// assembles its chunk from two 8-byte loads.
// The stacked operators have 20 rows (the gt = 0 and gt = 1 blocks): M-tile 0 holds rows
// 0..15, M-tile 1 rows 16..19 at A rows 0, 4, 8, 12 (its results in register 0 of every
// lane).  Per pass, N-tile t of the B operand goes to wave t % 4:
//   pass x  T1[a'][(z, y)]    A = X', B = f rows (z, y)                   7 N-tiles
//   pass y  T2[b'][(z, a')]   A = Y', B = T1 columns (z, a') from LDS      13 N-tiles
//   pass z  out[c'][(b', a')] A = Z', B = T2 columns (b', a') from LDS     25 N-tiles
// The intermediates go to LDS as one dword per value (f16 hi in the low half, lo in the high
// half), 18 dwords per column: the 10 values, the first 6 again, 2 unused (the stride spreads
// the 16 columns of an N-tile over distinct banks).  Lane group kg reads its 8 K-slots as 8
// consecutive dwords from position 0, 8, 6, 4 (kg = 0..3) and packs the hi or lo halves with
// four v_perm.  In block-exponent units |T1| < 10 and |T2| < 10 |T1|: T1 is scaled by 2^10
// and T2 (of the scaled T1) by 2^-1, which keeps both below the f16 maximum 65504 and their
// lo halves (2^-11 of a value) in the f16 normal range for all but the smallest values.
// Each wave accumulates its pass-z tiles over the task in FP64 with the weight of the output
// component: w 2^-(e_x + e_y + e_z + e_f + 9), zero when the screening dropped the component.  Two barriers per slice order the
// passes; one copy of each intermediate suffices (a pass rewrites a buffer only after every
// wave has passed the barrier that follows its last read).
constexpr int H10_COL = 18;
constexpr int H10_P1 = 208 * H10_COL, H10_P2 = 400 * H10_COL;

struct OpsH10 {
    uint4 x0, x1, y0, y1, z0, z1;  // A fragments of X', Y', Z' (M-tiles 0 and 1)
    uint4 f[2];                    // B fragments of f for the pass-x N-tiles wave, wave + 4
};

// a lane's fragment offsets inside an operator block pair (uint4 units: M-tile 0 row m, block
// gt = m >= 10; M-tile 1 rows 16..19 at rows 0, 4, 8, 12) and inside an f block (uint2 units,
// the pass-x rows of N-tiles wave and wave + 4): the same for every slice, so that a load is a
// uniform base (scalar registers) plus a lane offset
struct LaneOffs10 {
    unsigned op0, op1, f1[2], f2[2];
};
__device__ inline LaneOffs10 lane_offs10(int wave, int m, int kg) {
    LaneOffs10 L;
    L.op0 = (unsigned)((((m >= 10 ? 2 : 0)) * 10 + m % 10) * 4 + kg);
    L.op1 = (unsigned)((2 * 10 + 6 + (m >> 2)) * 4 + kg);
    // f rows [hi 10 | lo 10] (40 bytes, five uint2); lane group kg takes the dwords of its
    // K-slots [hi | lo | hi | 0 0]: kg 0 bytes 0-15, kg 1 16-31, kg 2 32-39 then 0-7, kg 3
    // 8-23 with its last dword zero
    const int o1 = kg == 0 ? 0 : kg == 1 ? 2 : kg == 2 ? 4 : 1, o2 = kg == 0 ? 1 : kg == 1 ? 3 : kg == 2 ? 0 : 2;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int r = min(16 * (wave + 4 * t) + m, 99);  // (rows beyond 99: results discarded)
        L.f1[t] = (unsigned)(r * 5 + o1);
        L.f2[t] = (unsigned)(r * 5 + o2);
    }
    return L;
}

// the A fragments of one operator (M-tile 1: rows 16..19 at rows 0, 4, 8, 12, zero elsewhere)
__device__ inline void load_oper_h10(const uint4 *oper, int blk, int m, const LaneOffs10 &L, uint4 &t0, uint4 &t1) {
    const uint4 *b = oper + (size_t)blk * 40;
    t0 = ld16(b + L.op0);
    t1 = (m & 3) == 0 ? ld16(b + L.op1) : make_uint4(0u, 0u, 0u, 0u);
}

__device__ inline OpsH10 load_opsh10(const Slice &sl, const Arr16 &J, const uint4 *gen16, int m, int kg,
                                     const LaneOffs10 &L) {
    OpsH10 o;
    load_oper_h10(J.oper, sl.xo / 100, m, L, o.x0, o.x1);
    load_oper_h10(J.oper, sl.yo / 100, m, L, o.y0, o.y1);
    load_oper_h10(J.oper, sl.zo / 100, m, L, o.z0, o.z1);
    const uint2 *fb = (const uint2 *)(sl.f >= 0 ? J.in + ((size_t)sl.f * 8 + sl.ft) * 250
                                                : gen16 + (size_t)(-sl.f - 1) * 250);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint2 u = ld8(fb + L.f1[t]), v = ld8(fb + L.f2[t]);
        o.f[t] = make_uint4(u.x, u.y, v.x, kg == 3 ? 0u : v.y);
    }
    return o;
}

// v = hi + lo as one dword (hi in the low half)
// v sc (sc a power of two) as one dword: hi = f16(v sc) in the low half, lo = f16(v sc - hi) in
// the high half, in two v_fma_mix instructions (the product is exact, the difference is
// formed exactly before its one rounding: the same bits as converting, subtracting and
// converting again)
__device__ inline unsigned pack_hl(float v, float sc) {
    unsigned d;
    asm volatile("v_fma_mixlo_f16 %0, %1, %2, 0\n\t"
                 "v_fma_mixhi_f16 %0, %1, %2, -%0 op_sel_hi:[0,0,1]"
                 : "=&v"(d)
                 : "v"(v), "v"(sc));
    return d;
}
// one value (times sc) to its LDS column (and to the duplicate tail for positions below 6)
__device__ inline void put_h10(unsigned *P, int col, int pos, float v, float sc) {
    const unsigned d = pack_hl(v, sc);
    P[col * H10_COL + pos] = d;
    if (pos < 6) P[col * H10_COL + 10 + pos] = d;
}
// the B fragment of column col for lane group kg (sel: the lane's v_perm selectors)
__device__ inline uint4 get_h10(const unsigned *P, int col, int start, const unsigned (&sel)[4]) {
    const unsigned *c = P + col * H10_COL + start;
    unsigned d[8];
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const uint2 v = *(const uint2 *)(c + 2 * p);
        d[2 * p] = v.x;
        d[2 * p + 1] = v.y;
    }
    return make_uint4(__builtin_amdgcn_perm(d[1], d[0], sel[0]), __builtin_amdgcn_perm(d[3], d[2], sel[1]),
                      __builtin_amdgcn_perm(d[5], d[4], sel[2]), __builtin_amdgcn_perm(d[7], d[6], sel[3]));
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_contract_mfma10h(int ntask_arg, const TaskSummary *sum, const SliceTask *tasks, const WorkItem *work,
                   const JobTab *jobs, const Slice *slices, const uint4 *gen16, const int *genExp,
                   double *nodeslots, long long *tabcnt) {
    constexpr int K3 = 1000;
    __shared__ __align__(16) unsigned P1[H10_P1];
    __shared__ __align__(16) unsigned P2[H10_P2];
    __shared__ double Wl[8];
    const int ntask = sum ? (sum->overflow ? 0 : sum->nstasks) : ntask_arg;
    if ((int)blockIdx.x >= ntask) return;
    const SliceTask task = tasks[blockIdx.x];
    const int wave = wave_index(), lane = threadIdx.x & 63, m = lane & 15, kg = lane >> 4;
    const WorkItem wi = work[task.g];
    const Arr16 J{jobs[wi.job].operBlocks16, jobs[wi.job].operExp16, jobs[wi.job].inCoef16, jobs[wi.job].inExp16};
    // this lane's K-slot window in an LDS column and the halves it takes (hi = bytes 0-1,
    // lo = bytes 2-3 of a dword; v_perm byte codes: first source bytes 4-7, second 0-3)
    constexpr unsigned HH = 0x05040100u, LL = 0x07060302u, ZZ = 0x0c0c0c0cu;
    const int start = kg == 0 ? 0 : kg == 1 ? 8 : kg == 2 ? 6 : 4;
    unsigned sel[4];
    sel[0] = kg == 2 ? LL : HH;
    sel[1] = (kg == 1 || kg == 2) ? LL : HH;
    sel[2] = kg == 1 ? LL : HH;
    sel[3] = kg == 1 ? LL : kg == 3 ? ZZ : HH;
    double acc0[7][4], acc1[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
        acc1[t] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; r++) acc0[t][r] = 0.0;
    }
    const Slice *srec = slices + task.start;
    const int last = task.count - 1;
    int ntup = 0;
    auto contract = [&](const OpsH10 &cur, const Slice &cs) {
        ntup += __popc((unsigned)cs.mask);
        // pass x
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int tx = wave + 4 * t;
            if (tx >= 7) break;
            const floatx4 d0 = mfmah(cur.x0, cur.f[t]), d1 = mfmah(cur.x1, cur.f[t]);
            const int r = 16 * tx + m;
            if (r < 100) {
                const int z = r / 10, y = r % 10;
#pragma unroll
                for (int i = 0; i < 4; i++) put_h10(P1, 20 * z + 4 * kg + i, y, d0[i], 1024.f);
                put_h10(P1, 20 * z + 16 + kg, y, d1[0], 1024.f);
            }
        }
        __syncthreads();
        if (wave == 0 && lane < 8) {  // the weights of the 8 output components of this slice
            const int gt = lane, gx = gt & 1, gy = (gt >> 1) & 1, gz = gt >> 2;
            const int e = ld_exp(J.operExp + cs.xo / 100 + 2 * gx) + ld_exp(J.operExp + cs.yo / 100 + 2 * gy) +
                          ld_exp(J.operExp + cs.zo / 100 + 2 * gz) +
                          (cs.f >= 0 ? ld_exp(J.inExp + (size_t)cs.f * 8 + cs.ft) : ld_exp(genExp + (-cs.f - 1)));
            Wl[gt] = ((cs.mask >> gt) & 1) ? ldexp(cs.w, -(e + 9)) : 0.0;
        }
        // pass y
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int ty = wave + 4 * t;
            if (ty >= 13) break;
            const int n = 16 * ty + m;
            const uint4 b = get_h10(P1, n, start, sel);
            const floatx4 d0 = mfmah(cur.y0, b), d1 = mfmah(cur.y1, b);
            if (n < 200) {
                const int z = n / 20, ap = n % 20;
#pragma unroll
                for (int i = 0; i < 4; i++) put_h10(P2, 20 * (4 * kg + i) + ap, z, d0[i], 0.5f);
                put_h10(P2, 20 * (16 + kg) + ap, z, d1[0], 0.5f);
            }
        }
        __syncthreads();
        // pass z: FP64 accumulation with the component weights
#pragma unroll
        for (int t = 0; t < 7; t++) {
            const int tz = wave + 4 * t;
            if (tz >= 25) break;
            const int n = 16 * tz + m;
            const uint4 b = get_h10(P2, n, start, sel);
            const floatx4 d0 = mfmah(cur.z0, b), d1 = mfmah(cur.z1, b);
            const int gxy = (n % 20 >= 10 ? 1 : 0) | (n / 20 >= 10 ? 2 : 0);
            const double wlo = Wl[gxy], whi = Wl[gxy + 4];
#pragma unroll
            for (int i = 0; i < 4; i++) acc0[t][i] = fma((double)d0[i], 4 * kg + i >= 10 ? whi : wlo, acc0[t][i]);
            acc1[t] = fma((double)d1[0], whi, acc1[t]);
        }
    };
    // two operand sets in turn; the next slice's operands and the record after it are
    // loaded while a slice is contracted
    const LaneOffs10 LO = lane_offs10(wave, m, kg);
    Slice cA = slice_at(srec, 0), cB;
    OpsH10 opA = load_opsh10(cA, J, gen16, m, kg, LO), opB;
    Slice rnext = slice_at(srec, min(1, last));
    for (int it = 0; it < task.count;) {
        cB = rnext;
        opB = load_opsh10(cB, J, gen16, m, kg, LO);
        rnext = slice_at(srec, min(it + 2, last));
        contract(opA, cA);
        if (++it >= task.count) break;
        cA = rnext;
        opA = load_opsh10(cA, J, gen16, m, kg, LO);
        rnext = slice_at(srec, min(it + 2, last));
        contract(opB, cB);
        ++it;
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long *)&tabcnt[4 * wi.job + 2], (unsigned long long)ntup);
    double *dst = nodeslots + (size_t)task.slot * 8 * K3;
#pragma unroll
    for (int t = 0; t < 7; t++) {
        if (wave + 4 * t >= 25) break;
        const int n = 16 * (wave + 4 * t) + m;
        const int bp = n / 20, ap = n % 20;
        const int gxy = (ap >= 10 ? 1 : 0) | (bp >= 10 ? 2 : 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int cp = 4 * kg + i;
            dst[(gxy | (cp >= 10 ? 4 : 0)) * K3 + ((cp % 10) * 10 + bp % 10) * 10 + ap % 10] = acc0[t][i];
        }
        const int cp = 16 + kg;
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
        hipLaunchKernelGGL(k_contract_mfma10h, dim3(nb), dim3(256), 0, s, ntask, sum, tasks, work, jobs, slices,
                           genPool16, genExp16, nodeslots, tabcnt);
    return hipGetLastError();
}

// the blocks of several arrays, one wave per block over their concatenation (item i's blocks
// are first[i] .. first[i + 1] - 1, its own blocks 0 ..; first[nitems] = the total)
template <int K>
__global__ void __launch_bounds__(256) k_split16_multi(int nitems, const Split16Item *items, const long long *first,
                                                       int bvals, int role) {
    const long long total = first[nitems];
    for (long long b = (long long)blockIdx.x * 4 + wave_index(); b < total; b += (long long)gridDim.x * 4) {
        int lo = 0, hi = nitems - 1;  // the item holding block b
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (first[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        const Split16Item it = items[lo];
        const int lb = (int)(b - first[lo]);
        if (K == 8) split16_block(lb, bvals, it.src, it.dst, it.exp);
        else split16_10_block(lb, bvals, role, it.src, it.dst, it.exp);
    }
}

hipError_t launch_split16_multi(hipStream_t s, int nitems, const Split16Item *items, const long long *first,
                                long long total, int bvals, int K, int role) {
    if (K != 8 && K != 10) return hipErrorInvalidValue;
    if (nitems <= 0 || total <= 0) return hipSuccess;
    const unsigned nb = (unsigned)std::min<long long>((total + 3) / 4, 8192);
    if (K == 8)
        hipLaunchKernelGGL(k_split16_multi<8>, dim3(nb), dim3(256), 0, s, nitems, items, first, bvals, role);
    else
        hipLaunchKernelGGL(k_split16_multi<10>, dim3(nb), dim3(256), 0, s, nitems, items, first, bvals, role);
    return hipGetLastError();
}

hipError_t launch_split16(hipStream_t s, int n, int first, const int *count, int bvals, int K, const double *src,
                          uint4 *dst, int *exp, int role, const int *firstp) {
    if (K != 8 && K != 10) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    // (a device first: at most 2048 workgroups, grid-stride)
    const unsigned nb = (unsigned)std::min((n + 3) / 4, firstp ? 2048 : INT_MAX);
    if (K == 8)
        hipLaunchKernelGGL(k_split16, dim3(nb), dim3(256), 0, s, n, first, firstp, count, bvals, src, dst, exp);
    else
        hipLaunchKernelGGL(k_split16_10, dim3(nb), dim3(256), 0, s, n, first, firstp, count, bvals, role, src, dst,
                           exp);
    return hipGetLastError();
}

}  // namespace mwg
