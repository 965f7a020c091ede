// Device-side data layouts shared by the HIP kernels and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mwg {

// open-addressing hash: keys[] holds key+1 (0 = empty), vals[] the node index
struct DHash {
    unsigned long long *keys = nullptr;
    int *vals = nullptr;
    uint32_t mask = 0;
};

__host__ __device__ inline uint32_t dkey_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb53fe1a85ec3ULL;
    k ^= k >> 33;
    return (uint32_t)k;
}
__host__ __device__ inline uint64_t dnode_key(int depth, int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)depth << 57) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}

__device__ inline int dhash_find(const DHash &h, uint64_t key) {
    const unsigned long long kk = (unsigned long long)key + 1ull;
    uint32_t i = dkey_hash(key) & h.mask;
    for (uint32_t probe = 0; probe <= h.mask; probe++) {
        unsigned long long v = h.keys[i];
        if (v == kk) return h.vals[i];
        if (v == 0ull) return -1;
        i = (i + 1) & h.mask;
    }
    return -1;
}

// one active Gaussian term at the current depth, as the work-list build sees it
constexpr int BM = 16;  // translations |d| < BM have precomputed combination masks
struct TermTab {
    double w;          // weight (unit-coordinate, includes scaling factor^3)
    long long blk;     // offset (doubles) of delta=-band, type 0 block in the level block array
    long long nrm;     // offset (doubles) of delta=-band, type 0 norm in the level norm array
    int band;
    short ub[3];       // union box half-widths over all combos (-1: term unused)
    short ub0[3];      // union box over combos with ft = 0
    short box[64][3];  // per combo (gt*8+ft) half-widths, -1 = combo screened out entirely
    // bm[q][m]: bit c set iff combo c is active and box[c][q] >= m; the combos whose box
    // holds delta are bm[0][|dx|] & bm[1][|dy|] & bm[2][|dz|]
    unsigned long long bm[3][BM];
};

// one surviving (g, f, ft, gt, term) tuple
struct Tuple {
    int g;          // index into the work list
    int f;          // input node index (>= 0) or -(generated index + 1)
    uint8_t term;   // index into the TermTab array
    uint8_t combo;  // gt * 8 + ft
    int16_t dx, dy, dz;
};
static_assert(sizeof(Tuple) == 16, "tuple record is 16 bytes");

// one (g, f, ft, term) slice of the mixed-precision apply: all 8 output
// components at once (mask: which gt are accumulated)
// one MFMA slice (g, f, ft, term): every output component gt of node g from component ft
// of input node f, with the operator blocks already resolved: xo, yo, zo are the offsets
// (in block values) of the (gt_d = 0, ft_d) block of each dimension in the level's block
// array (the gt_d = 1 block follows 2 (k+1)^2 values later)
struct Slice {
    double w;      // term weight
    int f;         // input node index (>= 0) or -(generated index + 1)
    int xo, yo, zo;
    uint8_t ft;
    uint8_t mask;  // output components to accumulate (bit gt)
    uint8_t pad[6];
};
static_assert(sizeof(Slice) == 32, "slice record is 32 bytes");

// a slice is contracted in FP64 (as tuples) when its largest bound exceeds
// ROUTE_FP64 = 2^21 times the screening threshold thr: the FP32 rounding of a slice
// (~8 * 2^-24 of its bound) then stays below thr, the size of every tuple the
// screening drops
#ifndef MWG_ROUTE_FP64
#define MWG_ROUTE_FP64 2097152.0
#endif
constexpr double ROUTE_FP64 = MWG_ROUTE_FP64;
// slices per MFMA task.  k = 9: each slice result is added to the task's FP64
// accumulators, so the length of a task changes only the summation grouping; it sets
// how many node slots (8 (k+1)^3 FP64 each) an iteration writes and reads back.
// k = 7: a task's four waves take every fourth slice and each sums its slices in FP32
// before the FP64 sum over waves, so its run is kept short: the FP32 running sum of n
// slices adds at most (n - 1) 2^-24 of their summed bounds.  Measured on the 5-orbital
// Helmholtz batch at prec 1e-6, FP32 runs of 16, 32 and 64 slices per wave give the same
// error against the FP64 apply (1.37e-8, 1.38e-8, 1.43e-8 of ||out||; the split of the
// operands dominates); 32 per wave (128 per task) is the fastest and the default.
constexpr int MAX_SLICE_CHUNK = 4096;
constexpr int MFMA8_TASK_SLICES = 256;  // k = 7 limit (64 slices per wave in FP32)
constexpr int MFMA8_TASK_DEFAULT = 128;
// on the root level (no screening threshold yet) slices above ROUTE_ROOT times the
// largest possible bound run in FP64; FP32 rounding of the others stays ~1e-12 of it,
// small also for outputs far below that bound (pair densities of orthogonal orbitals)
constexpr double ROUTE_ROOT = 1e-5;

}  // namespace mwg
