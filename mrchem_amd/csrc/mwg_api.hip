// Host orchestration and C ABI of libmwgreen.so.
//
// mrcpp::apply(prec, out, oper, inp) is restated as an adaptive build whose
// every numerical step runs on the GPU; the host keeps only the tree topology
// (node keys, parent/child links) and 8 norms per node, and decides refinement:
//   per level: screened boxes (host, from operator norm envelopes)
//            -> generated input nodes (device mark + compute)
//            -> work list of (g, f, ft, gt, term) tuples (device count + fill)
//            -> contraction (device) -> norms (device) -> split test (host)
//   after the last level: additive top-down push of the coarse scaling parts,
//   bottom-up compression and norms (device).
// Operators (Gaussian-sum fit on the host, NS blocks on the device) are built
// level by level as the output trees reach new depths.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mwg.h"
#include "mwg_core.h"
#include "mwg_kernels.h"
#include "mwg_predict.h"

namespace mwg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
// the generated-node tables of an apply were too small (the batch runs again with larger ones)
struct GenOverflow : Error {
    GenOverflow() : Error(MWG_ERR_RANGE, "generated-node table overflow") {}
};

#define HC(x)                                                                                         \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw Error(MWG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// device allocation with one retry after the block caches have been emptied
void flush_device_caches();
inline hipError_t dev_malloc(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        flush_device_caches();
        e = hipMalloc(p, bytes);
    }
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        throw Error(MWG_ERR_HIP, "device allocation of " + std::to_string(bytes >> 20) + " MiB failed (" +
                                     std::to_string(fr >> 20) + " of " + std::to_string(tot >> 20) +
                                     " MiB free after emptying the block caches)");
    }
    return e;
}

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t m) {  // discard contents
        if (m <= n && p) return;
        if (p) HC(hipFree(p));
        p = nullptr;
        n = 0;
        HC(dev_malloc((void **)&p, std::max<size_t>(m, 1) * sizeof(T)));
        n = std::max<size_t>(m, 1);
    }
    void grow_keep(size_t m, hipStream_t s) {
        if (m <= n && p) return;
        size_t nn = std::max<size_t>(m, 2 * n);
        T *q = nullptr;
        HC(dev_malloc((void **)&q, std::max<size_t>(nn, 1) * sizeof(T)));
        if (p) {
            HC(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
            HC(hipStreamSynchronize(s));
            HC(hipFree(p));
        }
        p = q;
        n = std::max<size_t>(nn, 1);
    }
    // ownership transfer (device block caches)
    T *release() {
        T *q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    void adopt(T *q, size_t m) {
        if (p) HC(hipFree(p));
        p = q;
        n = m;
    }
    void upload(const T *h, size_t m, hipStream_t s) {
        alloc(m);
        if (m) HC(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
        HC(hipStreamSynchronize(s));
    }
    void download(T *h, size_t m, hipStream_t s) const {
        if (m) HC(hipMemcpyAsync(h, p, m * sizeof(T), hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
    }
};

// ------------------------------------------------------------------ handles

// pinned host staging for asynchronous uploads: data is copied into the ring and
// sent with hipMemcpyAsync; the ring is reused only after the stream has synchronised
struct Stage {
    char *h = nullptr;
    size_t cap = 0, off = 0;
    ~Stage() {
        if (h) (void)hipHostFree(h);
    }
};

// Device blocks released by trees (coefficients, norms), kept for reuse by later trees
// of the same context.  All work of a context is ordered on its one stream, so a block
// handed out again is only touched after every earlier use of it.
struct DevCache;
static std::mutex g_cache_mu;
static std::vector<DevCache *> g_caches;

struct DevCache {
    std::multimap<size_t, void *> blocks;  // bytes -> block
    size_t bytes = 0;
    // bytes kept for reuse: half of the device memory (at least 8 GiB), set at the first
    // return (with the context's device current).  (A quarter made a 560-apply k = 9 batch
    // free and allocate again ~80 blocks per step, each free waiting for the device.)  Freed trees come back here
    // instead of hipFree (which waits for the device); an allocation that fails empties
    // every cache first (dev_malloc).
    size_t LIMIT = 0;
    long long evicted = 0;  // blocks freed because the cache was full (MWG_TIMING reports it)
    DevCache() {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_caches.push_back(this);
    }
    void flush() {
        for (auto &b : blocks) (void)hipFree(b.second);
        blocks.clear();
        bytes = 0;
    }
    void *get(size_t want, size_t &got) {
        auto it = blocks.lower_bound(want);  // a cached block at most an eighth (+1 MiB) larger
        if (it != blocks.end() && it->first <= want + want / 8 + ((size_t)1 << 20)) {
            got = it->first;
            void *q = it->second;
            bytes -= it->first;
            blocks.erase(it);
            return q;
        }
        void *q = nullptr;
        HC(dev_malloc(&q, std::max<size_t>(want, 8)));
        got = std::max<size_t>(want, 8);
        return q;
    }
    void put(void *q, size_t nbytes) {
        if (!q) return;
        if (!LIMIT) {
            size_t fr = 0, tot = 0;
            LIMIT = (size_t)8 << 30;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess) LIMIT = std::max(LIMIT, tot / 2);
            else (void)hipGetLastError();
        }
        blocks.insert({nbytes, q});
        bytes += nbytes;
        while (bytes > LIMIT && !blocks.empty()) {  // drop the largest blocks first
            auto it = std::prev(blocks.end());
            bytes -= it->first;
            (void)hipFree(it->second);
            blocks.erase(it);
            evicted++;
        }
    }
    ~DevCache() {
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            g_caches.erase(std::remove(g_caches.begin(), g_caches.end(), this), g_caches.end());
        }
        flush();
    }
};

// cached blocks are only reused on their own context's stream; emptying a cache first
// waits for the device so that no queued work still uses a block being freed
void flush_device_caches() {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    (void)hipDeviceSynchronize();
    for (DevCache *c : g_caches) c->flush();
}

// pinned host area for device-to-host reads (plain DMA, no staging through pageable memory)
struct Readback {
    char *h = nullptr;
    size_t cap = 0, off = 0;
    ~Readback() {
        if (h) (void)hipHostFree(h);
    }
    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        if (h) HC(hipHostFree(h));
        h = nullptr;
        cap = std::max(bytes, (size_t)1 << 20);
        HC(hipHostMalloc((void **)&h, cap, hipHostMallocDefault));
    }
};

// scratch of the operator level builder: item lists, child-scale blocks, norms
struct NSScratch {
    DBuf<double> t1, blk, nrm;
    DBuf<char> lists;
};

struct Ctx {
    int device = 0;
    std::atomic<int> refs{1};  // the owner's handle, plus one per MRA and predictor made on it
    DevCache cache;
    Readback rb;
    Stage stage;
    NSScratch ns;
    DBuf<char> s16items;  // item list of the batched split-FP16 conversion of an apply's inputs
    hipStream_t stream = nullptr;
    double wall_khz = 0.0;  // device wall clock (contraction timestamps), ticks per ms
    // device-resident refinement: high-water marks of the applies run on this context (work
    // list size, generated nodes, output/input node ratio, raised FP64 and slice task sizes),
    // which size its buffers, and its buffers, read-back entries and events
    int hw_nw = 0, hw_gen = 0;
    // generated-node table size of the applies (4 << gen_shift entries per input node, at
    // least 2^(14 + gen_shift) per job): raised when a batch overflowed the tables
    int gen_shift = 0;
    int hw_chunk[LOOP_MAX_DEPTH] = {}, hw_schunk[LOOP_MAX_DEPTH] = {};  // per refinement iteration
    double hw_ratio = 0.0;
    std::vector<int> last_in, last_out;  // input and output node counts of the last batch, per job
    DBuf<WorkItem> lwork[2];
    DBuf<int> lcidx, ltopo, lroots;
    DBuf<LoopJob> ljobs;
    DBuf<LoopCtl> lctl;
    DBuf<char> ltab;
    LoopIter *liter_h = nullptr, *liter_d = nullptr;  // pinned (coherent), and its device address
    hipEvent_t evIter[LOOP_MAX_DEPTH] = {}, evTopo = nullptr;
    std::string err;
    DBuf<double> gx, gw;  // 24-point Gauss-Legendre (operator quadrature)
    void *dot_scratch = nullptr;  // this device's allocation-free dot scratch (device address)
    // apply scratch
    DBuf<TermTab> tabs;
    DBuf<double> tbmax;
    DBuf<int> tpre, tpre0;
    // per-iteration read-back block, written by the kernels and copied to the host in
    // one transfer: schedule summary | per-table counts | work norms | child
    // probabilities | child features
    DBuf<char> rbd;
    DBuf<int> moves;  // tree compaction: (from, to) node-group pairs
    DBuf<char> iterbuf;  // per-iteration tables (work items, job tables, box requests)
    DBuf<SegRec> segs;
    DBuf<int> segcnt;        // per work-list segment: count over all parts
    DBuf<long long> gslice;  // per output node: first slice
    TaskCaps caps{1 << 20, 1 << 18, 1 << 15, 1 << 13, 1 << 15, 1 << 15, 1 << 13};
    DBuf<int> thread_off, gtotal;
    DBuf<long long> gofs;
    DBuf<Tuple> tuples;
    DBuf<uint64_t> genKeys;
    DBuf<int> genJob, genCount, genOverflow;
    DBuf<double> genPool, genNorm;
    DBuf<uint4> genPool16;  // split-FP16 copy of the generated nodes (mixed precision)
    DBuf<int> genExp16;
    DBuf<unsigned long long> genHashKeys;
    DBuf<int> genHashVals;
    DBuf<TransformItem> titems;
    DBuf<MulItem> mitems;  // batched products: items and work lists of all jobs
    DBuf<int> mwork;
    DBuf<ContractTask> tasks;
    DBuf<ReduceItem> reds;
    DBuf<double> partials, nodeslots;
    DBuf<Slice> slices;
    DBuf<SliceTask> stasks;
    // refinement traces (features, split label) and the learned predictor
    bool tracing = false;
    std::vector<float> trace_X, trace_y;
    struct PredH *pred = nullptr;
    float pred_thr = 0.5f;
    int chunk = 64;        // tuples per FP64 contraction task
    int slice_chunk = 0;   // slices per MFMA task (at most MAX_SLICE_CHUNK; 0: from each iteration's total)
};

// Wait for the context's stream by polling: the refinement loop waits once per
// iteration for a few hundred microseconds of GPU work, and a blocking wait adds its
// wake-up latency to every iteration.
static void stream_wait(Ctx &C) {
    while (true) {
        const hipError_t e = hipStreamQuery(C.stream);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) HC(e);
    }
}

// asynchronous upload through the context's staging ring
static void stage_copy(Ctx &C, void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (C.stage.off + need > C.stage.cap) {
        HC(hipStreamSynchronize(C.stream));
        C.stage.off = 0;
        if (need > C.stage.cap) {
            if (C.stage.h) HC(hipHostFree(C.stage.h));
            C.stage.h = nullptr;
            const size_t cap = std::max<size_t>(need, (size_t)64 << 20);
            HC(hipHostMalloc((void **)&C.stage.h, cap, hipHostMallocDefault));
            C.stage.cap = cap;
        }
    }
    char *p = C.stage.h + C.stage.off;
    std::memcpy(p, src, bytes);
    C.stage.off += need;
    HC(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, C.stream));
}

template <class T>
static void upload_async(Ctx &C, DBuf<T> &d, const T *src, size_t n) {
    d.alloc(n);
    if (n) stage_copy(C, d.p, src, n * sizeof(T));
}

// read back (synchronises the stream; the staging ring is free again afterwards)
template <class T>
static void download_sync(Ctx &C, const DBuf<T> &d, T *dst, size_t n) {
    if (n) HC(hipMemcpyAsync(dst, d.p, n * sizeof(T), hipMemcpyDeviceToHost, C.stream));
    stream_wait(C);
    C.stage.off = 0;
}

struct MraH {
    Ctx *ctx = nullptr;
    std::atomic<int> refs{1};  // the owner's handle, plus one per tree and operator made on it
    int n0 = 0, corner[3] = {0, 0, 0}, boxes[3] = {1, 1, 1}, maxDepth = 16;
    double sfac = 1.0;
    Basis B;
    DBuf<double> U, qx, qw, phiq, W, WT, ccoef;
    double lo(int d) const { return sfac * std::ldexp((double)corner[d], -n0); }
    double width(int d) const { return sfac * std::ldexp((double)boxes[d], -n0); }
    DevBasis dev() const { return DevBasis{B.K, U.p, qx.p, qw.p, phiq.p, W.p}; }
    ProjGeom geom() const { return ProjGeom{{lo(0), lo(1), lo(2)}, sfac, n0}; }
};

// open-addressing node index (slot key = node key + 1, 0 = empty; load at most 1/2):
// the layout of the device tables, so that build_hash uploads it as it is
struct NodeIndex {
    std::vector<unsigned long long> keys;
    std::vector<int> vals;
    size_t count = 0;
    void clear() {
        keys.clear();
        vals.clear();
        count = 0;
    }
    uint32_t mask() const { return (uint32_t)keys.size() - 1u; }
    void put(uint64_t k, int v) {
        if (2 * (count + 1) > keys.size()) rehash(std::max<size_t>(1024, 2 * keys.size()));
        const unsigned long long kk = (unsigned long long)k + 1ull;
        uint32_t h = dkey_hash(k) & mask();
        while (keys[h] != 0ull && keys[h] != kk) h = (h + 1) & mask();
        if (keys[h] == 0ull) {
            keys[h] = kk;
            count++;
        }
        vals[h] = v;
    }
    int get(uint64_t k) const {
        if (keys.empty()) return -1;
        const unsigned long long kk = (unsigned long long)k + 1ull;
        uint32_t h = dkey_hash(k) & mask();
        while (true) {  // an empty slot always exists (load <= 1/2)
            const unsigned long long v = keys[h];
            if (v == kk) return vals[h];
            if (v == 0ull) return -1;
            h = (h + 1) & mask();
        }
    }
    void rehash(size_t cap) {
        std::vector<unsigned long long> ok;
        std::vector<int> ov;
        ok.swap(keys);
        ov.swap(vals);
        keys.assign(cap, 0ull);
        vals.assign(cap, -1);
        count = 0;
        for (size_t i = 0; i < ok.size(); i++)
            if (ok[i]) put(ok[i] - 1ull, ov[i]);
    }
};

struct TreeH {
    MraH *mra = nullptr;
    std::vector<int> depth, parent, child, lr;  // lr: 3 per node
    NodeIndex index;
    std::vector<int> roots;
    std::vector<double> cnorm;  // 8 per node (host copy)
    double sqnorm = 0.0;
    DBuf<double> coef, norm;
    DBuf<unsigned long long> hkeys;
    DBuf<int> hvals;
    uint32_t hmask = 0;
    bool hvalid = false;
    // split-FP16 copy of the coefficients for the mixed-precision apply, made when the tree
    // is an input and dropped at the end of that apply (ver guards against stale copies)
    DBuf<uint4> coef16;
    DBuf<int> exp16;
    uint64_t ver = 1, ver16 = 0;

    int K3() const { return mra->B.K3; }
    int n() const { return (int)depth.size(); }
    void clear() {
        ver++;
        depth.clear();
        parent.clear();
        child.clear();
        lr.clear();
        index.clear();
        roots.clear();
        cnorm.clear();
        sqnorm = 0.0;
        hvalid = false;
    }
    int add(int d, int x, int y, int z, int par) {
        const int id = n();
        depth.push_back(d);
        parent.push_back(par);
        child.push_back(-1);
        lr.push_back(x);
        lr.push_back(y);
        lr.push_back(z);
        cnorm.resize(8 * (size_t)(id + 1), 0.0);
        index.put(node_key(d, x, y, z), id);
        hvalid = false;
        return id;
    }
    // host capacity for m nodes (no reallocation while the tree grows to m)
    void reserve_host(size_t m) {
        depth.reserve(m);
        parent.reserve(m);
        child.reserve(m);
        lr.reserve(3 * m);
        cnorm.reserve(8 * m);
        size_t cap = 1024;
        while (cap < 2 * m + 2) cap <<= 1;
        if (index.keys.size() < cap) index.rehash(cap);
    }
    void make_roots() {
        clear();
        for (int z = 0; z < mra->boxes[2]; z++)
            for (int y = 0; y < mra->boxes[1]; y++)
                for (int x = 0; x < mra->boxes[0]; x++) roots.push_back(add(0, x, y, z, -1));
    }
    void split(int i) {
        const int d = depth[i], x = lr[3 * i], y = lr[3 * i + 1], z = lr[3 * i + 2];
        const int first = n();
        for (int a = 0; a < 8; a++) add(d + 1, 2 * x + (a & 1), 2 * y + ((a >> 1) & 1), 2 * z + ((a >> 2) & 1), i);
        child[i] = first;
    }
    int find(int d, int x, int y, int z) const {
        return index.get(node_key(d, x, y, z));
    }
    // the topology of m nodes built elsewhere (the device refinement loop), the roots first
    void set_topology(int m, const int *dep, const int *par, const int *chi, const int *l3) {
        const std::vector<int> r = roots;
        clear();
        roots = r;
        reserve_host(m);
        depth.assign(dep, dep + m);
        parent.assign(par, par + m);
        child.assign(chi, chi + m);
        lr.assign(l3, l3 + 3 * (size_t)m);
        cnorm.assign(8 * (size_t)m, 0.0);
        for (int i = 0; i < m; i++) index.put(node_key(depth[i], lr[3 * i], lr[3 * i + 1], lr[3 * i + 2]), i);
    }
    // grow the coefficient and norm arrays through the context's block cache (the old
    // contents are copied on the stream; no synchronisation).  Growth is by half again, and
    // fit_dev trims a finished tree, so a tree holds at most about 1.5x its nodes' memory.
    void grow_cached(DBuf<double> &b, size_t m, size_t valid) {
        if (m <= b.n && b.p) return;
        Ctx &C = *mra->ctx;
        const size_t want = std::max<size_t>(m, b.n + b.n / 2);
        size_t got = 0;
        double *q = (double *)C.cache.get(want * sizeof(double), got);
        if (b.p) {
            const size_t keep = std::min(valid, b.n);  // (only the tree's nodes hold data)
            if (keep) HC(hipMemcpyAsync(q, b.p, keep * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
            const size_t old = b.n * sizeof(double);
            C.cache.put(b.release(), old);
        }
        b.adopt(q, got / sizeof(double));
    }
    void ensure_dev(size_t nodes) {
        grow_cached(coef, nodes * 8 * (size_t)K3(), (size_t)n() * 8 * (size_t)K3());
        grow_cached(norm, nodes * 8, (size_t)n() * 8);
    }
    // a finished tree: move the coefficients to a block of their size when the capacity
    // exceeds it by more than half (a copy on the stream; the old block goes to the cache)
    void fit_dev() {
        const size_t need = std::max<size_t>((size_t)n() * 8 * (size_t)K3(), 1);
        if (!coef.p || coef.n <= need + need / 2) return;
        Ctx &C = *mra->ctx;
        size_t got = 0;
        double *q = (double *)C.cache.get(need * sizeof(double), got);
        HC(hipMemcpyAsync(q, coef.p, need * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
        const size_t old = coef.n * sizeof(double);
        C.cache.put(coef.release(), old);
        coef.adopt(q, got / sizeof(double));
    }
    ~TreeH() {
        if (!mra) return;
        Ctx &C = *mra->ctx;
        const size_t cb = coef.n * sizeof(double), nb = norm.n * sizeof(double);
        C.cache.put(coef.release(), cb);
        C.cache.put(norm.release(), nb);
    }
    void build_hash() {
        if (hvalid) return;
        if (index.keys.empty()) index.rehash(1024);
        hipStream_t s = mra->ctx->stream;
        hkeys.upload(index.keys.data(), index.keys.size(), s);
        hvals.upload(index.vals.data(), index.vals.size(), s);
        hmask = index.mask();
        hvalid = true;
    }
    // the split-FP16 arrays, sized for the tree (through the block cache, where drop16
    // returns them)
    void alloc16() {
        const int K = mra->B.K;
        const size_t nb = (size_t)n() * 8;
        Ctx &C = *mra->ctx;
        const size_t need = nb * split16_block_uint4(K, 1, K3());
        size_t got = 0;
        if (coef16.n < need || !coef16.p) {
            if (coef16.p) C.cache.put(coef16.release(), coef16.n * sizeof(uint4));
            coef16.adopt((uint4 *)C.cache.get(need * sizeof(uint4), got), got / sizeof(uint4));
        }
        if (exp16.n < nb || !exp16.p) {
            if (exp16.p) C.cache.put(exp16.release(), exp16.n * sizeof(int));
            exp16.adopt((int *)C.cache.get(nb * sizeof(int), got), got / sizeof(int));
        }
    }
    void ensure16() {
        if (ver16 == ver) return;
        alloc16();
        const int K = mra->B.K;
        HC(launch_split16(mra->ctx->stream, n() * 8, 0, nullptr, K3(), K, coef.p, coef16.p, exp16.p, 1));
        ver16 = ver;
    }
    // the split-FP16 copy back to the block cache (made again on the next mixed apply)
    void drop16() {
        Ctx &C = *mra->ctx;
        const size_t cb = coef16.n * sizeof(uint4), eb = exp16.n * sizeof(int);
        if (coef16.p) C.cache.put(coef16.release(), cb);
        if (exp16.p) C.cache.put(exp16.release(), eb);
        ver16 = 0;
    }
    DHash dhash() const { return DHash{hkeys.p, hvals.p, hmask}; }
    void refresh_norms() {  // device norms -> host, square norm
        hipStream_t s = mra->ctx->stream;
        HC(launch_norms(s, n(), nullptr, coef.p, norm.p, K3()));
        cnorm.resize(8 * (size_t)n());
        norm.download(cnorm.data(), 8 * (size_t)n(), s);
        sqnorm_from_cnorm();
    }
    void sqnorm_from_cnorm() {  // compressed tree: root scaling parts + every wavelet part
        double sq = 0.0;
        for (int r : roots) sq += cnorm[8 * r] * cnorm[8 * r];
        for (int i = 0; i < n(); i++)
            for (int c = 1; c < 8; c++) sq += cnorm[8 * i + c] * cnorm[8 * i + c];
        sqnorm = sq;
    }
    // per-level lists of internal nodes (parents of complete sibling groups)
    std::vector<std::vector<int>> internal_by_depth() const {
        std::vector<std::vector<int>> L;
        for (int i = 0; i < n(); i++)
            if (child[i] >= 0) {
                if ((int)L.size() <= depth[i]) L.resize(depth[i] + 1);
                L[depth[i]].push_back(i);
            }
        return L;
    }
};

struct TermLevel {
    bool active = false;
    int band = -1;
    long long blk = 0, nrm = 0;  // offsets (doubles) into the level arrays
    std::vector<double> norms;   // (2 band + 1) * 4
    std::vector<double> env;     // 4 * (band + 1)
    double maxn[4] = {0, 0, 0, 0};
};

// one device allocation holding the arrays of several levels of an operator (those built
// together), from the context's block cache and returned there with the last of them (an
// operator keeps its MRA, and so the context, alive)
struct DevArena {
    Ctx *ctx;
    char *p = nullptr;
    size_t got = 0, used = 0;
    DevArena(Ctx &C, size_t bytes) : ctx(&C) { p = (char *)C.cache.get(std::max<size_t>(bytes, 256), got); }
    DevArena(const DevArena &) = delete;
    DevArena &operator=(const DevArena &) = delete;
    ~DevArena() { ctx->cache.put(p, got); }
    template <class T>
    T *take(size_t n, size_t align = 256) {
        T *q = (T *)(p + used);
        used += (n * sizeof(T) + align - 1) & ~(align - 1);
        return q;
    }
};
template <class T>
static size_t arena_bytes(size_t n) {
    return (n * sizeof(T) + 255) & ~(size_t)255;
}

struct LevelH {
    std::vector<TermLevel> t;
    std::shared_ptr<DevArena> mem, mem16, small;  // device arrays below, shared with other levels
    double *blocks = nullptr, *norms = nullptr;
    size_t nblocks = 0;       // doubles in blocks (4 K^2 per translation row)
    size_t nnorms = 0;        // doubles in norms (4 per row)
    LevTerm *dterms = nullptr;  // per-term screening data for the device box builder
    double *denv = nullptr;     // concatenated norm envelopes of the active terms
    uint4 *blocks16 = nullptr;  // split-FP16 copy (operator rows, exponent per block) for the mixed apply
    int *exp16 = nullptr;
    bool has16 = false;
};

struct OperH {
    MraH *mra = nullptr;
    double mu = 0.0, prec = 0.0;
    KernelFit fit;
    std::vector<double> beta, weff;
    std::vector<std::unique_ptr<LevelH>> lev;

    int nterms() const { return (int)beta.size(); }
    bool has(int d) const { return (int)lev.size() > d && lev[d]; }
};

// NS blocks of lists of (term, translation range) requests, one list per depth (NSBatch),
// for one or more operators of one MRA.  b1[term] is the term's exponent at the child scale
// of the batch's depth.  Rows (one per translation, 4 K^2 doubles: types by (gt, ft) bit)
// are numbered in request order (rows, plus the total at the end); the requests from
// groups[i].r0 up to the next group's r0 write their rows to groups[i].blocks (row 0 = the
// group's first row), and their norms to groups[i].norms when given; the 4 norms of every
// row also go to nrm.  The item lists of all batches travel in one upload; the child-scale
// blocks are computed in chunks of whole groups that fit a bounded scratch.  Asynchronous on
// the context's stream: the caller synchronises (by downloading the norms).
struct NSRequest {
    int term;
    long long dlo, dhi;
};
struct NSGroup {
    size_t r0;
    double *blocks;
    double *norms;
};
struct NSBatch {
    int d = 0;
    std::vector<double> b1;
    std::vector<NSRequest> req;
    std::vector<NSGroup> groups;
    double *nrm = nullptr;
    std::vector<long long> rows;
};

static void compute_ns(MraH &M, const std::vector<NSBatch *> &bs) {
    Ctx &C = *M.ctx;
    NSScratch &S = C.ns;
    const Basis &B = M.B;
    const int K2 = B.K2;
    const long long boxmax = std::max({M.boxes[0], M.boxes[1], M.boxes[2]});
    const long long cap = std::max<long long>(1, ((long long)256 << 20) / (8 * K2));
    struct Plan {
        std::vector<long long> lo1, hi1, trow;
        std::vector<size_t> chunk0;
        size_t b1off = 0, toff = 0, noff = 0;
    };
    std::vector<Plan> pl(bs.size());
    size_t nb1 = 0, nt = 0, nn = 0;
    long long tmax = 0;
    for (size_t b = 0; b < bs.size(); b++) {
        NSBatch &X = *bs[b];
        Plan &P = pl[b];
        const size_t nr = X.req.size(), ng = X.groups.size();
        const long long lmax1 = (boxmax << (X.d + 1)) - 1;
        X.rows.assign(nr + 1, 0);
        P.lo1.resize(nr);
        P.hi1.resize(nr);
        P.trow.assign(nr + 1, 0);
        for (size_t r = 0; r < nr; r++) {
            const double bb = X.b1[X.req[r].term];
            const long long band1 = std::min<long long>(lmax1, (long long)std::ceil(1.0 + std::sqrt(GAUSS_CUT / bb)));
            P.lo1[r] = std::max<long long>(-band1, 2 * X.req[r].dlo - 1);
            P.hi1[r] = std::min<long long>(band1, 2 * X.req[r].dhi + 1);
            X.rows[r + 1] = X.rows[r] + (X.req[r].dhi - X.req[r].dlo + 1);
            P.trow[r + 1] = P.trow[r] + (P.hi1[r] - P.lo1[r] + 1);
        }
        if (X.rows[nr] >= ((long long)1 << 31) || P.trow[nr] >= ((long long)1 << 31))
            throw Error(MWG_ERR_RANGE, "operator level too large");
        // chunks of groups whose child-scale rows fit the scratch (at least one group each)
        auto gend = [&](size_t g) { return g + 1 < ng ? X.groups[g + 1].r0 : nr; };
        P.chunk0 = {0};
        for (size_t g = 0; g < ng;) {
            const size_t c0 = g;
            while (g < ng && (g == c0 || P.trow[gend(g)] - P.trow[X.groups[c0].r0] <= cap)) g++;
            P.chunk0.push_back(g);
            tmax = std::max(tmax, P.trow[gend(g - 1)] - P.trow[X.groups[c0].r0]);
        }
        P.b1off = nb1;
        P.toff = nt;
        P.noff = nn;
        nb1 += X.b1.size();
        nt += (size_t)P.trow[nr];
        nn += (size_t)X.rows[nr];
    }
    // item lists of all batches; child-scale rows count from their chunk's first row
    std::vector<double> b1(nb1);
    std::vector<int> tterm(nt), row0(nn), tlo(nn), thi(nn);
    std::vector<long long> tdelta(nt), ndelta(nn);
    for (size_t b = 0; b < bs.size(); b++) {
        const NSBatch &X = *bs[b];
        const Plan &P = pl[b];
        const size_t nr = X.req.size(), ng = X.groups.size();
        auto gend = [&](size_t g) { return g + 1 < ng ? X.groups[g + 1].r0 : nr; };
        std::copy(X.b1.begin(), X.b1.end(), b1.begin() + P.b1off);
        for (size_t c = 0; c + 1 < P.chunk0.size(); c++) {
            const long long base = P.trow[X.groups[P.chunk0[c]].r0];
            for (size_t r = X.groups[P.chunk0[c]].r0; r < gend(P.chunk0[c + 1] - 1); r++) {
                for (long long q = P.lo1[r]; q <= P.hi1[r]; q++) {
                    const size_t i = P.toff + (size_t)(P.trow[r] + q - P.lo1[r]);
                    tterm[i] = (int)P.b1off + X.req[r].term;
                    tdelta[i] = q;
                }
                for (long long q = X.req[r].dlo; q <= X.req[r].dhi; q++) {
                    const size_t i = P.noff + (size_t)(X.rows[r] + q - X.req[r].dlo);
                    ndelta[i] = q;
                    row0[i] = (int)(P.trow[r] - base);
                    tlo[i] = (int)P.lo1[r];
                    thi[i] = (int)P.hi1[r];
                }
            }
        }
    }
    size_t off = 0;
    auto sec = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_b1 = sec(nb1 * sizeof(double)), o_tt = sec(nt * sizeof(int)), o_td = sec(nt * sizeof(long long)),
                 o_nd = sec(nn * sizeof(long long)), o_r0 = sec(nn * sizeof(int)), o_lo = sec(nn * sizeof(int)),
                 o_hi = sec(nn * sizeof(int));
    std::vector<char> pk(std::max<size_t>(off, 1));
    std::memcpy(pk.data() + o_b1, b1.data(), nb1 * sizeof(double));
    std::memcpy(pk.data() + o_tt, tterm.data(), nt * sizeof(int));
    std::memcpy(pk.data() + o_td, tdelta.data(), nt * sizeof(long long));
    std::memcpy(pk.data() + o_nd, ndelta.data(), nn * sizeof(long long));
    std::memcpy(pk.data() + o_r0, row0.data(), nn * sizeof(int));
    std::memcpy(pk.data() + o_lo, tlo.data(), nn * sizeof(int));
    std::memcpy(pk.data() + o_hi, thi.data(), nn * sizeof(int));
    S.lists.alloc(pk.size());
    stage_copy(C, S.lists.p, pk.data(), off);
    const char *L = S.lists.p;
    const double *db1 = (const double *)(L + o_b1);
    const int *dtt = (const int *)(L + o_tt), *dr0 = (const int *)(L + o_r0), *dlo = (const int *)(L + o_lo),
              *dhi = (const int *)(L + o_hi);
    const long long *dtd = (const long long *)(L + o_td), *dnd = (const long long *)(L + o_nd);
    S.t1.alloc((size_t)std::max<long long>(tmax, 1) * K2);
    hipStream_t s = C.stream;
    for (size_t b = 0; b < bs.size(); b++) {
        const NSBatch &X = *bs[b];
        const Plan &P = pl[b];
        const size_t nr = X.req.size(), ng = X.groups.size();
        auto gend = [&](size_t g) { return g + 1 < ng ? X.groups[g + 1].r0 : nr; };
        const double pre1 = std::ldexp(1.0, -(M.n0 + X.d + 1));
        for (size_t c = 0; c + 1 < P.chunk0.size(); c++) {
            const long long t0 = P.trow[X.groups[P.chunk0[c]].r0], nt1 = P.trow[gend(P.chunk0[c + 1] - 1)] - t0;
            HC(launch_tblock_mom(s, (int)nt1, dtt + P.toff + t0, dtd + P.toff + t0, db1, pre1, B.K, M.ccoef.p, C.gx.p,
                                 C.gw.p, S.t1.p));
            for (size_t g = P.chunk0[c]; g < P.chunk0[c + 1]; g++) {
                const long long f0 = X.rows[X.groups[g].r0], n = X.rows[gend(g)] - f0;
                const size_t i0 = P.noff + (size_t)f0;
                if (n > 0)
                    HC(launch_nsblocks(s, (int)n, dnd + i0, dr0 + i0, dlo + i0, dhi + i0, S.t1.p, M.W.p, B.K,
                                       X.groups[g].blocks, X.nrm + 4 * f0, X.groups[g].norms));
            }
        }
    }
}

static long long ns_rows(const std::vector<NSRequest> &req) {
    long long n = 0;
    for (const NSRequest &r : req) n += r.dhi - r.dlo + 1;
    return n;
}

// the levels of one depth being built for some operators of an MRA
struct LevelJob {
    int d = 0;
    std::vector<OperH *> ops;
    std::vector<int> toff;         // first term of each operator in the batches' b1
    std::vector<long long> band;   // translation bands
    NSBatch near, full;            // near-diagonal blocks of every term; full bands of the active ones
    std::vector<size_t> req0;      // first full-band request of each operator
    std::vector<long long> nenv;   // envelope doubles of each operator's level
    std::vector<std::unique_ptr<LevelH>> L;
};

// Levels of operators of one MRA at one or more depths, built together: one round trip for
// the near-diagonal blocks that decide which terms are active at each depth, one for the full
// bands of the active terms, whatever the number of operators and depths.  Each operator's
// new blocks and norms (and with16: their split-FP16 copies, one conversion launch per
// operator) share one allocation, the blocks of its levels adjacent; the per-term data of
// all new levels share another, sent in one upload.
static void build_levels(MraH &M, std::vector<LevelJob> &lj, bool with16) {
    Ctx &C = *M.ctx;
    const int K2 = M.B.K2;
    const long long boxmax = std::max({M.boxes[0], M.boxes[1], M.boxes[2]});
    // phase 1: near-diagonal blocks decide which terms are active
    long long n1tot = 0, n1max = 0;
    std::vector<long long> nrm0(lj.size());
    for (size_t j = 0; j < lj.size(); j++) {
        LevelJob &J = lj[j];
        const int d = J.d, m1 = M.n0 + d + 1;
        const long long lmax1 = (boxmax << (d + 1)) - 1;
        const size_t no = J.ops.size();
        J.toff.assign(no + 1, 0);
        for (size_t o = 0; o < no; o++) J.toff[o + 1] = J.toff[o] + J.ops[o]->nterms();
        J.near.d = J.full.d = d;
        J.near.b1.resize(J.toff[no]);
        J.band.resize(J.toff[no]);
        J.near.req.resize(J.toff[no]);
        for (size_t o = 0; o < no; o++)
            for (int t = 0; t < J.ops[o]->nterms(); t++) {
                const int g = J.toff[o] + t;
                const double b1 = J.ops[o]->beta[t] * std::ldexp(1.0, -2 * m1);
                J.near.b1[g] = b1;
                const long long band1 = std::min<long long>(lmax1, (long long)std::ceil(1.0 + std::sqrt(GAUSS_CUT / b1)));
                J.band[g] = std::min<long long>((boxmax << d) - 1, (band1 + 2) / 2);
                const long long nb = std::min<long long>(J.band[g], 1);
                J.near.req[g] = {g, -nb, nb};
            }
        J.full.b1 = J.near.b1;
        const long long n1 = ns_rows(J.near.req);
        nrm0[j] = n1tot;
        n1tot += n1;
        n1max = std::max(n1max, n1);
    }
    C.ns.blk.alloc((size_t)std::max<long long>(n1max, 1) * 4 * K2);
    C.ns.nrm.alloc((size_t)std::max<long long>(n1tot, 1) * 4);
    std::vector<NSBatch *> batches;
    for (size_t j = 0; j < lj.size(); j++) {
        lj[j].near.groups = {NSGroup{0, C.ns.blk.p, nullptr}};
        lj[j].near.nrm = C.ns.nrm.p + 4 * nrm0[j];
        batches.push_back(&lj[j].near);
    }
    compute_ns(M, batches);
    std::vector<double> nn((size_t)n1tot * 4);
    C.ns.nrm.download(nn.data(), nn.size(), C.stream);
    // activity; sizes of the new levels
    long long n2tot = 0;
    struct OpMem {
        OperH *op;
        size_t bytes = 0, nb16 = 0;  // blocks and norms; K x K blocks
        std::shared_ptr<DevArena> mem;
        double *blocks = nullptr;  // first block of its new levels
        uint4 *b16 = nullptr;
        int *e16 = nullptr;
    };
    std::vector<OpMem> om;
    {
        size_t nops = 0;  // (an upper bound: references into om stay valid)
        for (const LevelJob &J : lj) nops += J.ops.size();
        om.reserve(nops);
    }
    auto op_mem = [&](OperH *op) -> OpMem & {
        for (auto &x : om)
            if (x.op == op) return x;
        om.push_back(OpMem{op});
        return om.back();
    };
    const int K = M.B.K;
    const size_t per16 = with16 ? split16_block_uint4(K, 0, K * K) : 0;
    size_t small = 0;  // per-term data of all new levels
    for (size_t j = 0; j < lj.size(); j++) {
        LevelJob &J = lj[j];
        const size_t no = J.ops.size();
        J.L.resize(no);
        J.req0.assign(no + 1, 0);
        J.nenv.assign(no, 0);
        for (size_t o = 0; o < no; o++) {
            J.L[o] = std::make_unique<LevelH>();
            J.L[o]->t.resize(J.ops[o]->nterms());
            J.req0[o] = J.full.req.size();
            long long nrow = 0;
            for (int t = 0; t < J.ops[o]->nterms(); t++) {
                const int g = J.toff[o] + t;
                double tn = 0.0, wn = 0.0;
                for (long long q = 0; q <= J.near.req[g].dhi - J.near.req[g].dlo; q++) {
                    const double *p = &nn[(nrm0[j] + J.near.rows[g] + q) * 4];
                    tn = std::max(tn, p[0]);
                    wn = std::max({wn, p[1], p[2], p[3]});
                }
                const bool active = (J.d == 0) || (wn > ACT_REL * J.ops[o]->prec * tn);
                J.L[o]->t[t].active = active;
                if (active) {
                    J.L[o]->t[t].band = (int)J.band[g];
                    J.full.req.push_back({g, -J.band[g], J.band[g]});
                    nrow += 2 * J.band[g] + 1;
                    J.nenv[o] += 4 * (J.band[g] + 1);
                }
            }
            // slice records address the level's FP32 block array with 32-bit offsets
            if (nrow * 4 * K2 >= ((long long)1 << 31)) throw Error(MWG_ERR_RANGE, "operator level too large");
            J.L[o]->nblocks = (size_t)nrow * 4 * K2;
            J.L[o]->nnorms = (size_t)nrow * 4;
            OpMem &X = op_mem(J.ops[o]);
            X.bytes += (J.L[o]->nblocks + J.L[o]->nnorms) * sizeof(double);
            X.nb16 += J.L[o]->nblocks / K2;
            small += arena_bytes<LevTerm>(J.ops[o]->nterms()) + arena_bytes<double>((size_t)std::max<long long>(J.nenv[o], 1));
        }
        J.req0[no] = J.full.req.size();
        nrm0[j] = n2tot;
        n2tot += ns_rows(J.full.req);
    }
    // each operator: the blocks of all its new levels, then their norms (rows of 32 bytes:
    // adjacent arrays stay aligned), then the split-FP16 copies
    for (OpMem &X : om) {
        const size_t b16 = with16 ? arena_bytes<uint4>(X.nb16 * per16) + arena_bytes<int>(X.nb16) : 0;
        X.mem = std::make_shared<DevArena>(C, X.bytes + 256 + b16);
    }
    auto small_mem = std::make_shared<DevArena>(C, small);
    for (int pass = 0; pass < 2; pass++)
        for (LevelJob &J : lj)
            for (size_t o = 0; o < J.ops.size(); o++) {
                LevelH &LV = *J.L[o];
                OpMem &X = op_mem(J.ops[o]);
                LV.mem = X.mem;
                if (pass == 0) {
                    LV.blocks = LV.mem->take<double>(LV.nblocks, 8);
                    if (!X.blocks) X.blocks = LV.blocks;
                } else {
                    LV.norms = LV.mem->take<double>(LV.nnorms, 8);
                }
            }
    if (with16)
        for (OpMem &X : om) {
            X.mem->used = (X.mem->used + 255) & ~(size_t)255;
            X.b16 = X.mem->take<uint4>(X.nb16 * per16);
            X.e16 = X.mem->take<int>(X.nb16);
            size_t at = 0;
            for (LevelJob &J : lj)
                for (size_t o = 0; o < J.ops.size(); o++)
                    if (J.ops[o] == X.op) {
                        LevelH &LV = *J.L[o];
                        LV.blocks16 = X.b16 + at * per16;
                        LV.exp16 = X.e16 + at;
                        LV.has16 = true;
                        at += LV.nblocks / K2;
                    }
        }
    // phase 2: full bands of the active terms, straight into the level arrays
    batches.clear();
    if (n2tot > 0) C.ns.nrm.alloc((size_t)n2tot * 4);
    for (size_t j = 0; j < lj.size(); j++) {
        LevelJob &J = lj[j];
        if (J.full.req.empty()) continue;
        for (size_t o = 0; o < J.ops.size(); o++)
            if (J.L[o]->nblocks) J.full.groups.push_back({J.req0[o], J.L[o]->blocks, J.L[o]->norms});
        J.full.nrm = C.ns.nrm.p + 4 * nrm0[j];
        batches.push_back(&J.full);
    }
    if (!batches.empty()) compute_ns(M, batches);
    if (with16)
        for (const OpMem &X : om)
            if (X.nb16) HC(launch_split16(C.stream, (int)X.nb16, 0, nullptr, K2, K, X.blocks, X.b16, X.e16, 0));
    std::vector<double> hn((size_t)n2tot * 4);
    if (n2tot > 0) C.ns.nrm.download(hn.data(), hn.size(), C.stream);
    // host-side term data, and the device copy of the per-term data in one upload
    std::vector<char> img(std::max<size_t>(small, 1));
    for (size_t j = 0; j < lj.size(); j++) {
        LevelJob &J = lj[j];
        for (size_t o = 0; o < J.ops.size(); o++) {
            LevelH &LV = *J.L[o];
            if (J.req0[o + 1] > J.req0[o]) {
                const long long f0 = J.full.rows[J.req0[o]];
                for (size_t r = J.req0[o]; r < J.req0[o + 1]; r++) {
                    TermLevel &tl = LV.t[J.full.req[r].term - J.toff[o]];
                    const int b = tl.band;
                    const long long h0 = (nrm0[j] + J.full.rows[r]) * 4;
                    tl.blk = (J.full.rows[r] - f0) * 4 * (long long)K2;
                    tl.nrm = (J.full.rows[r] - f0) * 4;
                    tl.norms.assign(hn.begin() + h0, hn.begin() + h0 + (2 * b + 1) * 4);
                    for (int q = 0; q < 2 * b + 1; q++)
                        for (int type = 0; type < 4; type++)
                            tl.maxn[type] = std::max(tl.maxn[type], tl.norms[q * 4 + type]);
                    tl.env.assign(4 * (b + 1), 0.0);
                    for (int type = 0; type < 4; type++) {
                        double run = 0.0;
                        for (int m = b; m >= 0; m--) {
                            run = std::max(run, std::max(tl.norms[(b + m) * 4 + type], tl.norms[(b - m) * 4 + type]));
                            tl.env[type * (b + 1) + m] = run;
                        }
                    }
                }
            }
            const int nt = J.ops[o]->nterms();
            std::vector<LevTerm> lts(nt);
            std::vector<double> envs;
            for (int t = 0; t < nt; t++) {
                const TermLevel &tl = LV.t[t];
                LevTerm &x = lts[t];
                std::memset(&x, 0, sizeof(x));
                x.w = J.ops[o]->weff[t];
                x.active = tl.active ? 1 : 0;
                x.band = tl.band;
                if (!tl.active) continue;
                x.blk = tl.blk;
                x.nrm = tl.nrm;
                x.env = (long long)envs.size();
                for (int q = 0; q < 4; q++) x.maxn[q] = tl.maxn[q];
                envs.insert(envs.end(), tl.env.begin(), tl.env.end());
            }
            if (envs.empty()) envs.push_back(0.0);
            LV.dterms = small_mem->take<LevTerm>(nt);
            LV.denv = small_mem->take<double>(envs.size());
            std::memcpy(img.data() + ((char *)LV.dterms - small_mem->p), lts.data(), nt * sizeof(LevTerm));
            std::memcpy(img.data() + ((char *)LV.denv - small_mem->p), envs.data(), envs.size() * sizeof(double));
            LV.small = small_mem;
        }
    }
    stage_copy(C, small_mem->p, img.data(), small);
    for (LevelJob &J : lj)
        for (size_t o = 0; o < J.ops.size(); o++) {
            if ((int)J.ops[o]->lev.size() <= J.d) J.ops[o]->lev.resize(J.d + 1);
            J.ops[o]->lev[J.d] = std::move(J.L[o]);
        }
}

// Levels 0..depth of each (operator, depth) pair: every missing level of the operators of
// one MRA is built in one build_levels call (with16: with their split-FP16 copies)
static void ensure_levels(const std::vector<std::pair<OperH *, int>> &need, bool with16 = false) {
    std::vector<MraH *> mras;
    for (const auto &x : need)
        if (std::find(mras.begin(), mras.end(), x.first->mra) == mras.end()) mras.push_back(x.first->mra);
    for (MraH *m : mras) {
        int dmax = -1;
        for (const auto &x : need)
            if (x.first->mra == m) dmax = std::max(dmax, x.second);
        std::vector<LevelJob> lj;
        for (int q = dmax; q >= 0; q--) {
            LevelJob J;
            J.d = q;
            for (const auto &x : need)
                if (x.first->mra == m && q <= x.second && !x.first->has(q) &&
                    std::find(J.ops.begin(), J.ops.end(), x.first) == J.ops.end())
                    J.ops.push_back(x.first);
            if (!J.ops.empty()) lj.push_back(std::move(J));
        }
        if (!lj.empty()) build_levels(*m, lj, with16);
    }
}

// Split-FP16 copies (mixed apply) of levels 0..depth of each (operator, depth) pair, whose
// levels exist: one allocation per operator for all its levels that lack one, and one
// conversion launch per run of levels whose blocks are adjacent in memory
static void ensure_levels16(Ctx &C, const std::vector<std::pair<OperH *, int>> &need, int K) {
    std::vector<std::pair<OperH *, int>> ops;
    for (const auto &x : need) {
        auto it = std::find_if(ops.begin(), ops.end(), [&](const std::pair<OperH *, int> &y) { return y.first == x.first; });
        if (it == ops.end())
            ops.push_back(x);
        else
            it->second = std::max(it->second, x.second);
    }
    const size_t per = split16_block_uint4(K, 0, K * K), KK = (size_t)K * K;
    for (const auto &x : ops) {
        OperH &O = *x.first;
        std::vector<LevelH *> todo;
        for (int q = 0; q <= x.second; q++) {
            LevelH &LV = *O.lev[q];
            if (LV.has16) continue;
            if (LV.nblocks)
                todo.push_back(&LV);
            else
                LV.has16 = true;
        }
        if (todo.empty()) continue;
        std::sort(todo.begin(), todo.end(), [](const LevelH *a, const LevelH *b) { return a->blocks < b->blocks; });
        size_t nb = 0;
        for (const LevelH *LV : todo) nb += LV->nblocks / KK;
        auto A = std::make_shared<DevArena>(C, arena_bytes<uint4>(nb * per) + arena_bytes<int>(nb));
        uint4 *b16 = A->take<uint4>(nb * per);
        int *e16 = A->take<int>(nb);
        size_t at = 0;
        for (size_t i = 0; i < todo.size();) {
            size_t j = i, run = 0;
            while (j < todo.size() && (j == i || todo[j]->blocks == todo[j - 1]->blocks + todo[j - 1]->nblocks)) {
                LevelH &LV = *todo[j];
                LV.mem16 = A;
                LV.blocks16 = b16 + (at + run) * per;
                LV.exp16 = e16 + at + run;
                LV.has16 = true;
                run += LV.nblocks / KK;
                j++;
            }
            HC(launch_split16(C.stream, (int)run, 0, nullptr, (int)KK, K, todo[i]->blocks, b16 + at * per, e16 + at, 0));
            at += run;
            i = j;
        }
    }
}

static void build_oper(OperH &O, MraH *mra, double mu, double prec) {
    O.mra = mra;
    O.mu = mu;
    O.prec = prec;
    const double eps = prec / 10.0;
    const double rmin = 0.1 * std::sqrt(eps);
    double w = 0.0;
    for (int d = 0; d < 3; d++) w = std::max(w, mra->width(d));
    O.fit = fit_kernel(mu, eps, rmin, std::sqrt(3.0) * w);
    const int nt = (int)O.fit.a.size();
    if (nt > 255) throw Error(MWG_ERR_RANGE, "kernel fit needs more than 255 terms");
    const double s = mra->sfac;
    O.beta.resize(nt);
    O.weff.resize(nt);
    for (int t = 0; t < nt; t++) {
        O.beta[t] = O.fit.a[t] * s * s;
        O.weff[t] = O.fit.w[t] * s * s * s;
    }
}

// ------------------------------------------------------------------ projection

template <class LaunchFn>
static void project_adaptive(TreeH &T, double prec, LaunchFn launch, const std::vector<double> *gauss_alpha,
                             const std::vector<double> *gauss_center) {
    MraH &M = *T.mra;
    Ctx &C = *M.ctx;
    hipStream_t s = C.stream;
    T.make_roots();
    std::vector<int> work = T.roots;
    double sNorm = 0.0, wNorm = 0.0;
    int iter = 0;
    DBuf<int> dwork, ddepth, dlr;
    while (!work.empty()) {
        T.ensure_dev(T.n());
        dwork.upload(work.data(), work.size(), s);
        ddepth.upload(T.depth.data(), T.depth.size(), s);
        dlr.upload(T.lr.data(), T.lr.size(), s);
        launch((int)work.size(), dwork.p, ddepth.p, dlr.p, T.coef.p);
        HC(launch_norms(s, (int)work.size(), dwork.p, T.coef.p, T.norm.p, T.K3()));
        std::vector<double> nrm(8 * (size_t)T.n());
        T.norm.download(nrm.data(), nrm.size(), s);
        for (int i : work) {
            std::memcpy(&T.cnorm[8 * (size_t)i], &nrm[8 * (size_t)i], 8 * sizeof(double));
            if (iter == 0) sNorm += nrm[8 * i] * nrm[8 * i];
            for (int c = 1; c < 8; c++) wNorm += nrm[8 * i + c] * nrm[8 * i + c];
        }
        const double tnorm = std::sqrt(sNorm + wNorm);
        std::vector<int> next;
        for (int i : work) {
            const int d = T.depth[i];
            if (d >= M.maxDepth) continue;
            double wn = 0.0;
            for (int c = 1; c < 8; c++) wn += T.cnorm[8 * i + c] * T.cnorm[8 * i + c];
            bool must = false;
            if (gauss_alpha) {
                // a Gaussian narrower than the quadrature spacing is invisible to the norm test
                const double h = M.sfac * std::ldexp(1.0, -(M.n0 + d));
                const double hc = 0.5 * h;
                for (size_t t = 0; t < gauss_alpha->size() && !must; t++) {
                    const double two_sigma = 2.0 / std::sqrt(2.0 * (*gauss_alpha)[t]);
                    if (hc <= M.B.K * two_sigma) continue;
                    double d2 = 0.0;
                    for (int q = 0; q < 3; q++) {
                        const double lo = M.lo(q) + h * (double)T.lr[3 * i + q], hi = lo + h;
                        const double c = (*gauss_center)[3 * t + q];
                        const double dd = c < lo ? lo - c : (c > hi ? c - hi : 0.0);
                        d2 += dd * dd;
                    }
                    if ((*gauss_alpha)[t] * d2 <= 40.0) must = true;
                }
            }
            if (split_check(M.n0 + d, std::sqrt(wn), prec, tnorm) || must) {
                T.split(i);
                for (int a = 0; a < 8; a++) next.push_back(T.child[i] + a);
            }
        }
        work.swap(next);
        iter++;
    }
    // bottom-up compression of the internal nodes
    auto L = T.internal_by_depth();
    std::vector<TransformItem> items;
    for (int d = (int)L.size() - 1; d >= 0; d--) {
        items.clear();
        for (int p : L[d]) items.push_back({T.coef.p, nullptr, p, T.child[p]});
        C.titems.upload(items.data(), items.size(), s);
        HC(launch_transform(s, 0, (int)items.size(), C.titems.p, M.W.p, M.B.K));
    }
    T.refresh_norms();
    T.fit_dev();
}

// ------------------------------------------------------------------ apply

struct JobH {
    double prec;
    TreeH *out;
    OperH *op;
    TreeH *in;
    std::vector<std::array<double, 8>> fmax;
    double sNorm = 0.0, wNorm = 0.0, gsq = -1.0;
    std::vector<int> work;
    std::vector<int> wpos;  // work-list index of each node of this iteration (-1: none)
    int64_t tuples = 0, tuples_fp64 = 0, slices = 0, gen = 0;
    size_t genOff = 0;
    uint32_t genMask = 0;
    std::vector<char> spec, alive, has_feat;  // per output node
    std::vector<float> feat;                  // per output node, MLP_FEATURES each
};

// features of a freshly created child c of node p (see mwg_predict.h)
// ------------------------------------------------------------------ learned predictor

struct PredH {
    Ctx *ctx = nullptr;
    DBuf<float> P, m, v, G, mean, sdev;
    DBuf<float> X, Y, prob, Xn, H1, H2, dZ1, dZ2, dZ3;
    DBuf<float> Xq;  // feature rows of inference queries
    DBuf<int> idx;
    int adam_t = 0;
};

// probabilities for n feature rows (pinned staging both ways, one polling wait)
static void predictor_infer(PredH &Pr, int n, const float *X, float *prob) {
    Ctx &C = *Pr.ctx;
    upload_async(C, Pr.Xq, X, (size_t)n * MLP_FEATURES);
    Pr.prob.alloc(n);
    HC(launch_mlp_forward(C.stream, n, nullptr, Pr.Xq.p, Pr.mean.p, Pr.sdev.p, Pr.P.p, Pr.prob.p, nullptr, nullptr,
                          nullptr));
    C.rb.reserve((size_t)n * sizeof(float));
    HC(hipMemcpyAsync(C.rb.h, Pr.prob.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, C.stream));
    stream_wait(C);
    C.stage.off = 0;
    std::memcpy(prob, C.rb.h, (size_t)n * sizeof(float));
}

// forward + backward over samples idx[0..n) of the resident set (idx null: all); gradient in Pr.G
static float predictor_grad(PredH &Pr, int n, const int *didx) {
    Ctx &C = *Pr.ctx;
    hipStream_t s = C.stream;
    Pr.prob.alloc(n);
    Pr.Xn.alloc((size_t)n * MLP_FEATURES);
    Pr.H1.alloc((size_t)n * MLP_HIDDEN);
    Pr.H2.alloc((size_t)n * MLP_HIDDEN);
    Pr.dZ1.alloc((size_t)n * MLP_HIDDEN);
    Pr.dZ2.alloc((size_t)n * MLP_HIDDEN);
    Pr.dZ3.alloc(n);
    HC(launch_mlp_forward(s, n, didx, Pr.X.p, Pr.mean.p, Pr.sdev.p, Pr.P.p, Pr.prob.p, Pr.Xn.p, Pr.H1.p, Pr.H2.p));
    HC(launch_mlp_backward(s, n, didx, Pr.Y.p, Pr.P.p, Pr.prob.p, Pr.H1.p, Pr.H2.p, Pr.dZ1.p, Pr.dZ2.p, Pr.dZ3.p));
    HC(launch_mlp_grad(s, n, Pr.Xn.p, Pr.H1.p, Pr.H2.p, Pr.dZ1.p, Pr.dZ2.p, Pr.dZ3.p, Pr.G.p));
    return 0.0f;
}

// mean binary cross-entropy of the resident set (host-side from the device probabilities)
static double predictor_loss(PredH &Pr, int n, const float *y) {
    Ctx &C = *Pr.ctx;
    Pr.prob.alloc(n);
    HC(launch_mlp_forward(C.stream, n, nullptr, Pr.X.p, Pr.mean.p, Pr.sdev.p, Pr.P.p, Pr.prob.p, nullptr, nullptr,
                          nullptr));
    std::vector<float> pr(n);
    Pr.prob.download(pr.data(), n, C.stream);
    double L = 0.0;
    for (int i = 0; i < n; i++) {
        const double p = std::min(std::max((double)pr[i], 1e-7), 1.0 - 1e-7);
        L -= y[i] > 0.5f ? std::log(p) : std::log(1.0 - p);
    }
    return L / std::max(n, 1);
}

// ------------------------------------------------------------------ compaction


// removes the nodes with alive == 0 (whole sibling groups already unlinked from their parents)
// Remove discarded speculative children.  Below the roots the node array is a sequence
// of sibling groups of 8 (one per split), and a discarded group is a group of leaves, so
// the array is compacted by moving the last live group into each hole: O(discarded)
// host work and one 8-node device copy per move.
static void compact_tree(TreeH &T, const std::vector<char> &alive) {
    Ctx &C = *T.mra->ctx;
    const size_t nb = (size_t)8 * T.K3();  // doubles per node
    const int nroots = (int)T.roots.size();
    int n = T.n();
    auto key_of = [&](int i) { return node_key(T.depth[i], T.lr[3 * i], T.lr[3 * i + 1], T.lr[3 * i + 2]); };
    std::vector<int> holes;
    for (int i = nroots; i < n; i += 8)
        if (!alive[i]) holes.push_back(i);
    std::vector<int> moves;  // (from, to) pairs, copied on the device by one launch
    auto move_group = [&](int from, int to) {
        const int p = T.parent[from];
        if (p >= 0) T.child[p] = to;
        for (int a = 0; a < 8; a++) {
            const int f = from + a, t = to + a;
            T.depth[t] = T.depth[f];
            T.parent[t] = T.parent[f];
            T.child[t] = T.child[f];
            for (int q = 0; q < 3; q++) T.lr[3 * t + q] = T.lr[3 * f + q];
            std::memcpy(&T.cnorm[8 * (size_t)t], &T.cnorm[8 * (size_t)f], 8 * sizeof(double));
            if (T.child[t] >= 0)
                for (int b = 0; b < 8; b++) T.parent[T.child[t] + b] = t;
        }
        moves.push_back(from);
        moves.push_back(to);
    };
    std::vector<char> dead_at((size_t)n, 0);
    for (int h : holes) dead_at[h] = 1;
    size_t hi = 0;
    while (hi < holes.size()) {
        while (n > nroots && dead_at[n - 8]) n -= 8;  // discarded groups at the end just go
        const int h = holes[hi];
        if (h >= n) break;
        move_group(n - 8, h);
        dead_at[h] = 0;
        n -= 8;
        hi++;
    }
    if (!moves.empty()) {
        upload_async(C, C.moves, moves.data(), moves.size());
        HC(launch_move_groups(C.stream, (int)moves.size() / 2, C.moves.p, T.coef.p, T.norm.p, (int)nb));
    }
    T.depth.resize(n);
    T.parent.resize(n);
    T.child.resize(n);
    T.lr.resize(3 * (size_t)n);
    T.cnorm.resize(8 * (size_t)n);
    T.index.clear();
    for (int i = 0; i < n; i++) T.index.put(key_of(i), i);
    T.hvalid = false;
}

struct PhaseTimer {
    bool on = false;
    std::vector<std::pair<std::string, double>> acc;
    std::chrono::steady_clock::time_point t0;
    void start() { t0 = std::chrono::steady_clock::now(); }
    void lap(const char *name) {
        if (!on) return;
        auto t1 = std::chrono::steady_clock::now();
        const double dt = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
        for (auto &p : acc)
            if (p.first == name) {
                p.second += dt;
                return;
            }
        acc.push_back({name, dt});
    }
    void report() {
        if (!on) return;
        for (auto &p : acc) fprintf(stderr, "[mwg timing] %-14s %8.3f ms\n", p.first.c_str(), 1e3 * p.second);
    }
};

// f(j) for every job j on up to 16 host threads (host-only work on distinct output
// trees; the inputs are only read); the first exception is rethrown
template <class F>
static void for_each_job(size_t n, F f) {
    const size_t nt = std::min<size_t>({n, (size_t)16, (size_t)std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        for (size_t j = 0; j < n; j++) f(j);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            try {
                for (size_t j = next++; j < n; j = next++) f(j);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto &x : th) x.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// milliseconds between two device-clock stamps of a contraction (0 when they are missing)
static float clock_ms(const Ctx &C, unsigned long long start, unsigned long long end) {
    if (start == 0ull || end == 0ull || end < start) return 0.0f;
    return (float)((double)(end - start) / C.wall_khz);
}

// MFMA flops one slice issues (v_mfma_f32_16x16x32_f16 with split operands, counted as
// issued, padding included): K = 8 28 MFMAs, K = 10 90 (two M-tiles of 7 + 13 + 25 N-tiles)
static double slice_flops(int K) {
    if (K == 8) return 2.0 * 16 * 16 * 32 * 28;  // split-FP16: 4 + 8 + 16 MFMAs of 16x16x32
    if (K == 10) return 2.0 * 16 * 16 * 32 * 90;  // split-FP16: 2 x (7 + 13 + 25) MFMAs of 16x16x32
    return 0.0;
}

// device memory for the FP64 partial slots of one refinement iteration
constexpr double SLOT_BUDGET = 8.0 * (1ull << 30);

// gen_shift: the generated-node tables hold 4 << gen_shift entries per input node, and at
// least 2^(14 + gen_shift)
static void apply_jobs(std::vector<JobH> &jobs, int max_iter, unsigned flags, std::vector<mwg_apply_stats> &stats,
                       int gen_shift) {
    if (jobs.empty()) return;
    PhaseTimer PT;
    PT.on = std::getenv("MWG_TIMING") != nullptr;
    PT.start();
    MraH &M = *jobs[0].in->mra;
    Ctx &C = *M.ctx;
    hipStream_t s = C.stream;
    const int K = M.B.K, K3 = M.B.K3;
    if (flags & ~(MWG_APPLY_MIXED | MWG_APPLY_PREDICT)) throw Error(MWG_ERR_ARG, "unknown apply flags");
    const bool mixed = (flags & MWG_APPLY_MIXED) != 0;
    if (mixed && K != 8 && K != 10)
        throw Error(MWG_ERR_ARG, "MWG_APPLY_MIXED is built for orders k = 7 and k = 9 only (this MRA has k = " +
                                     std::to_string(K - 1) + ")");
    auto t_start = std::chrono::steady_clock::now();
    float t_contract_ms = 0.0f;
    const bool use_pred = (flags & MWG_APPLY_PREDICT) && C.pred != nullptr;
    // the device-resident refinement loop (see device_loop below) is tried when a batch ran
    // on this context before; it reserves the output trees itself
    const bool dl_plan = !use_pred && !C.tracing && max_iter < 0 && C.hw_nw > 0 && jobs.size() <= (size_t)LOOP_MAX_JOBS &&
                         M.maxDepth < LOOP_MAX_DEPTH && std::getenv("MWG_HOST_LOOP") == nullptr;
    // inputs: device hash, norm maxima per depth (host work per job, on threads for large
    // batches)
    size_t genTotal = 0, in_nodes = 0;
    std::vector<std::pair<OperH *, int>> need;
    for (JobH &J : jobs) {
        if (J.in->mra != &M || J.out->mra != &M || J.op->mra != &M)
            throw Error(MWG_ERR_ARG, "all trees and operators of an apply must share one MRA");
        if (J.in->n() == 0) throw Error(MWG_ERR_STATE, "input tree is empty");
        J.in->build_hash();
        in_nodes += J.in->n();
    }
    std::vector<int> dmax_in(jobs.size(), 0);
    auto scan_input = [&](size_t j) {
        JobH &J = jobs[j];
        J.fmax.assign(M.maxDepth + 2, std::array<double, 8>{0, 0, 0, 0, 0, 0, 0, 0});
        std::vector<double> leafmax(M.maxDepth + 2, 0.0);
        int dm = 0;
        for (int i = 0; i < J.in->n(); i++) {
            const int d = J.in->depth[i];
            dm = std::max(dm, d);
            for (int c = 0; c < 8; c++) J.fmax[d][c] = std::max(J.fmax[d][c], J.in->cnorm[8 * (size_t)i + c]);
            if (J.in->child[i] < 0) {
                double fn = 0.0;
                for (int c = 0; c < 8; c++) fn += J.in->cnorm[8 * (size_t)i + c] * J.in->cnorm[8 * (size_t)i + c];
                leafmax[d] = std::max(leafmax[d], std::sqrt(fn));
            }
        }
        double run = 0.0;
        for (int d = 0; d < M.maxDepth + 2; d++) {
            J.fmax[d][0] = std::max(J.fmax[d][0], run);
            run = std::max(run, leafmax[d]);
        }
        dmax_in[j] = dm;
    };
    if (in_nodes < 65536)  // (threads cost more than they save on small batches)
        for (size_t j = 0; j < jobs.size(); j++) scan_input(j);
    else
        for_each_job(jobs.size(), scan_input);
    for (size_t j = 0; j < jobs.size(); j++) {
        JobH &J = jobs[j];
        // output trees reach at most a level or so below their inputs: build one more
        need.push_back({J.op, std::min(dmax_in[j] + 1, M.maxDepth)});
        J.out->make_roots();
        // output trees end between a third of the input's size (pair Poissons) and about its
        // size (Helmholtz): reserve half of it once (beyond, the arrays grow geometrically:
        // an allocation and a copy on the stream each time)
        const size_t est = std::max<size_t>(64, (size_t)J.out->n() + J.in->n() / 2);
        if (!dl_plan) J.out->ensure_dev(est);
        J.out->reserve_host(est);
        J.work = J.out->roots;
        size_t cap = (size_t)1 << (14 + gen_shift);
        while (cap < ((size_t)4 << gen_shift) * (size_t)J.in->n()) cap <<= 1;
        if (cap > ((size_t)1 << 31)) throw Error(MWG_ERR_RANGE, "generated-node table beyond 2^31 entries");
        J.genOff = genTotal;
        J.genMask = (uint32_t)(cap - 1);
        genTotal += cap;
    }
    if (genTotal > ((size_t)1 << 31)) throw Error(MWG_ERR_RANGE, "generated-node tables beyond 2^31 entries");
    // operator levels down to the inputs' depths, all operators at once (deeper ones on demand)
    PT.lap("inputs");
    ensure_levels(need, mixed);
    if (mixed) {
        ensure_levels16(C, need, K);
        // split-FP16 copies of all inputs that lack one: a launch per group of 64 trees (the
        // GPU converts a group while the host takes the next group's arrays from the cache)
        std::vector<TreeH *> todo;
        for (JobH &J : jobs)
            if (J.in->ver16 != J.in->ver && std::find(todo.begin(), todo.end(), J.in) == todo.end())
                todo.push_back(J.in);
        if (!todo.empty()) {
            constexpr size_t GROUP = 64;
            const size_t ng = (todo.size() + GROUP - 1) / GROUP;
            const size_t o_first = (GROUP * sizeof(Split16Item) + 255) & ~(size_t)255;
            const size_t gbytes = (o_first + (GROUP + 1) * sizeof(long long) + 255) & ~(size_t)255;
            C.s16items.alloc(ng * gbytes);
            for (size_t g = 0; g < ng; g++) {
                const size_t i0 = g * GROUP, i1 = std::min(todo.size(), i0 + GROUP);
                std::vector<char> blob(gbytes, 0);
                Split16Item *items = (Split16Item *)blob.data();
                long long *first = (long long *)(blob.data() + o_first);
                first[0] = 0;
                for (size_t i = i0; i < i1; i++) {
                    TreeH &T = *todo[i];
                    T.alloc16();
                    items[i - i0] = Split16Item{T.coef.p, T.coef16.p, T.exp16.p};
                    first[i - i0 + 1] = first[i - i0] + (long long)T.n() * 8;
                    T.ver16 = T.ver;
                }
                char *dst = C.s16items.p + g * gbytes;
                stage_copy(C, dst, blob.data(), gbytes);
                HC(launch_split16_multi(s, (int)(i1 - i0), (const Split16Item *)dst, (const long long *)(dst + o_first),
                                        first[i1 - i0], K3, K, 1));
            }
        }
    }
    PT.lap("operator levels");
    // generated-node tables
    C.genHashKeys.alloc(genTotal);
    C.genHashVals.alloc(genTotal);
    HC(hipMemsetAsync(C.genHashKeys.p, 0, genTotal * sizeof(unsigned long long), s));
    const int genCap = (int)std::min<size_t>(genTotal / 2, (size_t)1 << 24);
    C.genKeys.alloc(genCap);
    C.genJob.alloc(genCap);
    C.genCount.alloc(1);
    C.genOverflow.alloc(1);
    HC(hipMemsetAsync(C.genCount.p, 0, sizeof(int), s));
    int genDone = 0, genPoolCap = 0, genLastNew = 0;
    PT.lap("setup");

    int iter = 0;
    int64_t predicted = 0, wasted = 0;
    for (JobH &J : jobs) {
        J.spec.assign(J.out->n(), 0);
        J.alive.assign(J.out->n(), 1);
        if (C.tracing || use_pred) J.feat.assign((size_t)J.out->n() * MLP_FEATURES, 0.0f);
        J.has_feat.assign(J.out->n(), 0);
    }
    // job table (and screening-box request) of job J at depth d, appended to jt and bjobs:
    // operator level, input and output arrays, thresholds from the running norm
    auto add_table = [&](JobH &J, int d, std::vector<JobTab> &jt, std::vector<BoxJob> &bjobs, int &ntabs,
                         int &max_terms) {
        if (!J.op->has(d)) ensure_levels({{J.op, d}}, mixed);
        JobTab T;
        std::memset(&T, 0, sizeof(T));
        LevelH &LV = *J.op->lev[d];
        const int nt = J.op->nterms();
        T.thr = J.gsq > 0.0 ? SCREEN_REL * J.prec * std::sqrt(J.gsq / nt) : 0.0;
        T.tab0 = ntabs;
        T.ntab = nt;
        BoxJob B;
        std::memset(&B, 0, sizeof(B));
        B.terms = LV.dterms;
        B.env = LV.denv;
        B.thr = T.thr;
        for (int q = 0; q < 8; q++) B.F[q] = J.fmax[d][q];
        B.nterms = nt;
        B.depth = d;
        B.tab0 = ntabs;
        B.jt = (int)jt.size();
        B.pre_off = ntabs + (int)jt.size();
        bjobs.push_back(B);
        ntabs += nt;
        max_terms = std::max(max_terms, nt);
        T.inHash = J.in->dhash();
        T.genHash = DHash{C.genHashKeys.p + J.genOff, C.genHashVals.p + J.genOff, J.genMask};
        T.inCoef = J.in->coef.p;
        T.inNorm = J.in->norm.p;
        T.operBlocks = LV.blocks;
        if (mixed) {
            if (!LV.has16) ensure_levels16(C, {{J.op, d}}, K);
            J.in->ensure16();
            T.operBlocks16 = LV.blocks16;
            T.operExp16 = LV.exp16;
            T.inCoef16 = J.in->coef16.p;
            T.inExp16 = J.in->exp16.p;
        }
        T.operNorms = LV.norms;
        T.outCoef = J.out->coef.p;
        T.outNorm = J.out->norm.p;
        for (int q = 0; q < 3; q++) T.lmax[q] = M.boxes[q] << d;
        T.depth = d;
        T.feat_innorm = std::sqrt(std::max(J.in->sqnorm, 0.0));
        T.feat_tnorm = J.gsq > 0.0 ? std::sqrt(J.gsq) : T.feat_innorm;  // first iteration: input norm
        T.prec = J.prec;
        T.feat_mu = (float)std::log10(J.op->mu + 1e-3);
        T.feat_cell = (float)std::log10(M.width(0) / M.boxes[0] * std::ldexp(1.0, -(d + 1)));
        T.n0 = M.n0;
        jt.push_back(T);
    };
    // ---- device-resident refinement.  The plain split test (no predictor, no traces, no
    // iteration limit) runs on the device (k_loop_split, k_loop_emit), so an iteration is
    // queued without waiting for the one before: the host reads each iteration's work-list
    // size one iteration behind, only to stop queueing.  Capacities (work list, output
    // nodes, generated nodes, schedule) come from the high-water marks of earlier applies on
    // this context; when one is exceeded the batch runs again with the host loop below,
    // which raises them.  The output trees are those of the host loop (same split test in
    // the same order; the running norm sums the same squares in another order).
    const int njobs = (int)jobs.size();
    bool dl_done = false;

    auto device_loop = [&]() -> bool {
        const int nr = (int)jobs[0].out->roots.size();
        const int WCAP = std::max(C.hw_nw + C.hw_nw / 4 + 64, njobs * nr);
        // output node capacity per job: a batch like the last one (same size; SCF iterations,
        // repeated steps) gets 1.25x the output of the job at the same position, scaled by
        // the growth of its input; otherwise the largest output/input ratio seen, at most 2
        const bool like_last = C.last_out.size() == (size_t)njobs;
        const double ratio = std::min(2.0, std::max(0.5, 1.25 * C.hw_ratio));
        std::vector<int> caps_j(njobs);
        double need_bytes = 0.0;
        for (int j = 0; j < njobs; j++) {
            const int in_n = jobs[j].in->n();
            const double est = like_last ? 1.25 * C.last_out[j] * std::max(1.0, (double)in_n / std::max(1, C.last_in[j]))
                                         : ratio * in_n;
            caps_j[j] = (int)std::min<double>((nr + std::ceil(est) + 64 + 7), (double)(INT_MAX / 16)) & ~7;
            need_bytes += std::max(0.0, ((double)caps_j[j] * 8 * K3 - (double)jobs[j].out->coef.n) * sizeof(double));
        }
        {
            // the reservation must fit comfortably (else the host loop, which grows as it goes)
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
            if (need_bytes > 0.6 * ((double)fr + (double)C.cache.bytes)) return false;
        }
        std::vector<LoopJob> lj(njobs);
        size_t topo = 0;
        for (int j = 0; j < njobs; j++) {
            const int cap = caps_j[j];
            jobs[j].out->ensure_dev((size_t)cap);
            std::memset(&lj[j], 0, sizeof(LoopJob));
            lj[j].cap = cap;
            lj[j].n = nr;
            lj[j].w0 = j * nr;
            lj[j].nwk = nr;
            lj[j].gsq = -1.0;
            lj[j].prec = jobs[j].prec;
            topo += (size_t)cap;
        }
        C.ltopo.alloc(6 * topo);
        {
            size_t off = 0;
            for (int j = 0; j < njobs; j++) {
                int *base = C.ltopo.p + 6 * off;
                const size_t cap = (size_t)lj[j].cap;
                lj[j].depth = base;
                lj[j].parent = base + cap;
                lj[j].child = base + 2 * cap;
                lj[j].lr = base + 3 * cap;
                off += cap;
            }
        }
        const int nseg = mixed ? 9 : 8;
        const size_t eb = (size_t)ENUM_BLOCKS + (size_t)WCAP;  // >= work nodes x parts for any list <= WCAP
        C.thread_off.alloc(eb * nseg * 256);
        C.gtotal.alloc(eb * nseg);
        C.gofs.alloc(eb * nseg);
        C.segs.alloc((size_t)WCAP * 8);
        C.segcnt.alloc((size_t)WCAP * nseg);
        C.gslice.alloc((size_t)WCAP);
        C.lwork[0].alloc((size_t)WCAP);
        C.lwork[1].alloc((size_t)WCAP);
        C.lcidx.alloc((size_t)WCAP);
        C.ljobs.alloc((size_t)njobs);
        C.lctl.alloc(1);
        {
            const int want = std::min(genCap, C.hw_gen + C.hw_gen / 4 + 4096);
            if (genPoolCap < want) {
                C.genPool.grow_keep((size_t)want * K3, s);
                C.genNorm.grow_keep((size_t)want, s);
                if (mixed) {
                    C.genPool16.grow_keep((size_t)want * split16_block_uint4(K, 1, K3), s);
                    C.genExp16.grow_keep((size_t)want, s);
                }
                genPoolCap = want;
            }
        }
        // task sizes per iteration: as the host loop's, raised where it had to raise them
        auto chunk_dl = [&](int d) { return std::max(C.chunk, C.hw_chunk[d]); };
        auto schunk_dl = [&](int d) {
            return K == 8 ? std::min(C.slice_chunk > 0 ? C.slice_chunk : MFMA8_TASK_DEFAULT, MFMA8_TASK_SLICES)
                          : (C.slice_chunk > 0 ? std::max(C.slice_chunk, C.hw_schunk[d])
                                               : (C.hw_schunk[d] > 0 ? C.hw_schunk[d] : -2048));
        };
        C.tuples.alloc((size_t)std::max<long long>(C.caps.tuples, 1));
        C.tasks.alloc((size_t)std::max(C.caps.tasks, 1));
        C.reds.alloc((size_t)std::max(C.caps.reds, 1));
        C.partials.alloc((size_t)std::max(C.caps.slots, 1) * K3);
        if (mixed) {
            C.slices.alloc((size_t)std::max<long long>(C.caps.slices, 1));
            C.stasks.alloc((size_t)std::max(C.caps.stasks, 1));
            C.nodeslots.alloc((size_t)std::max(C.caps.nnslots, 1) * 8 * K3);
        }
        // schedule summary | table counts | work-node norms
        const size_t o_tab = 256;
        const size_t o_wn = (o_tab + 4 * (size_t)njobs * sizeof(long long) + 255) & ~(size_t)255;
        C.rbd.alloc(o_wn + (size_t)WCAP * 8 * sizeof(double));
        TaskSummary *dsum = (TaskSummary *)C.rbd.p;
        long long *dtab = (long long *)(C.rbd.p + o_tab);
        double *dwn = (double *)(C.rbd.p + o_wn);
        if (!C.liter_h) {
            HC(hipHostMalloc((void **)&C.liter_h, sizeof(LoopIter) * LOOP_MAX_DEPTH,
                             hipHostMallocMapped | hipHostMallocCoherent));
            HC(hipHostGetDevicePointer((void **)&C.liter_d, C.liter_h, 0));
        }
        if (!C.evTopo) HC(hipEventCreateWithFlags(&C.evTopo, hipEventDisableTiming));
        const int tcap = (int)std::min<size_t>(topo / 8 + 8, (size_t)INT_MAX);  // internal nodes <= nodes / 8
        C.titems.alloc((size_t)tcap);
        for (int q = 0; q < LOOP_MAX_DEPTH; q++)
            if (!C.evIter[q]) HC(hipEventCreateWithFlags(&C.evIter[q], hipEventDisableTiming));
        {
            LoopCtl ctl;
            std::memset(&ctl, 0, sizeof(ctl));
            ctl.nw[0] = njobs * nr;
            for (int d = 0; d < LOOP_MAX_DEPTH; d++) ctl.splitfac[d] = std::pow(2.0, -0.5 * (M.n0 + d + 1));
            upload_async(C, C.lctl, &ctl, 1);
        }
        upload_async(C, C.ljobs, lj.data(), lj.size());
        upload_async(C, C.lroots, jobs[0].out->lr.data(), 3 * (size_t)nr);
        HC(launch_loop_init(s, njobs, nr, C.ljobs.p, C.lroots.p, C.lwork[0].p));
        char *ctlb = (char *)C.lctl.p;
        int *gfirst = (int *)(ctlb + offsetof(LoopCtl, genFirst));
        int *queue = (int *)(ctlb + offsetof(LoopCtl, queue));
        for (int j = 0; j < njobs; j++) jobs[j].work.clear();
        PT.lap("loop setup");
        // job tables of every depth in a region of their own: the depths whose operator levels
        // exist go up in one upload before the first iteration, deeper ones as they are queued
        int ntabs = 0, max_terms = 0;
        for (const JobH &J : jobs) {
            ntabs += J.op->nterms();
            max_terms = std::max(max_terms, J.op->nterms());
        }
        C.tpre.alloc((size_t)ntabs + njobs + 1);
        C.tpre0.alloc((size_t)ntabs + njobs + 1);
        C.tabs.alloc((size_t)std::max(ntabs, 1));
        C.tbmax.alloc((size_t)std::max(njobs * max_terms, 1));
        const size_t o_bj = ((size_t)njobs * sizeof(JobTab) + 255) & ~(size_t)255;
        const size_t tstride = (o_bj + (size_t)njobs * sizeof(BoxJob) + 255) & ~(size_t)255;
        C.ltab.alloc((size_t)(M.maxDepth + 1) * tstride);
        auto fill_tables = [&](int d, char *dst) {
            std::vector<JobTab> jt;
            std::vector<BoxJob> bj;
            int nt = 0, mt = 0;
            jt.reserve(njobs);
            bj.reserve(njobs);
            for (int j = 0; j < njobs; j++) add_table(jobs[j], d, jt, bj, nt, mt);
            for (int q = 0; q < njobs; q++) {
                jt[q].pre = C.tpre.p + jt[q].tab0 + q;
                jt[q].pre0 = C.tpre0.p + jt[q].tab0 + q;
            }
            std::memcpy(dst, jt.data(), (size_t)njobs * sizeof(JobTab));
            std::memcpy(dst + o_bj, bj.data(), (size_t)njobs * sizeof(BoxJob));
        };
        int dpre = -1;  // levels 0..dpre exist for every job
        while (dpre < M.maxDepth) {
            bool all = true;
            for (const JobH &J : jobs) all = all && J.op->has(dpre + 1);
            if (!all) break;
            dpre++;
        }
        if (dpre >= 0) {
            std::vector<char> blob((size_t)(dpre + 1) * tstride, 0);
            for (int d = 0; d <= dpre; d++) fill_tables(d, blob.data() + (size_t)d * tstride);
            stage_copy(C, C.ltab.p, blob.data(), blob.size());
        }
        auto enqueue = [&](int d) {
            if (d > dpre) {
                std::vector<std::pair<OperH *, int>> deeper;
                for (int j = 0; j < njobs; j++)
                    if (!jobs[j].op->has(d)) deeper.push_back({jobs[j].op, d});
                if (!deeper.empty()) ensure_levels(deeper, mixed);
                std::vector<char> blob(tstride, 0);
                fill_tables(d, blob.data());
                stage_copy(C, C.ltab.p + (size_t)d * tstride, blob.data(), tstride);
            }
            JobTab *djt = (JobTab *)(C.ltab.p + (size_t)d * tstride);
            BoxJob *dbj = (BoxJob *)(C.ltab.p + (size_t)d * tstride + o_bj);
            const int cur = d & 1;
            const int *nwp = (const int *)(ctlb + offsetof(LoopCtl, nw)) + cur;
            int *nwnext = (int *)(ctlb + offsetof(LoopCtl, nw)) + (cur ^ 1);
            const WorkItem *dw = C.lwork[cur].p;
            unsigned long long *dts = (unsigned long long *)(ctlb + offsetof(LoopCtl, ts)) + 2 * d;
            const LoopInit li{C.ljobs.p, dtab, C.lctl.p, dts, C.genCount.p};
            HC(launch_job_boxes(s, njobs, max_terms, dbj, C.tabs.p, C.tbmax.p, djt, C.tpre.p, C.tpre0.p,
                                C.genOverflow.p, li));
            GenState gs{C.genKeys.p, C.genJob.p, C.genCount.p, genCap, C.genOverflow.p, C.genPool.p, C.genNorm.p,
                        genPoolCap, dtab};
            HC(launch_gen_mark(s, WCAP, dw, djt, C.tabs.p, gs, nwp, queue));
            HC(launch_gen_compute(s, 0, genPoolCap, djt, gs, M.W.p, K, gfirst));
            if (mixed)
                HC(launch_split16(s, genPoolCap, 0, C.genCount.p, K3, K, C.genPool.p, C.genPool16.p, C.genExp16.p, 1,
                                  gfirst));
            if (!mixed)
                HC(launch_enum(s, 0, WCAP, 0, dw, djt, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, nullptr, nullptr,
                               nullptr, nullptr, nwp, queue + 1));
            else
                HC(launch_enum_mixed(s, 0, K, WCAP, 0, dw, djt, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, nullptr,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, nwp, queue + 1));
            HC(launch_build_tasks(s, WCAP, nseg, 0, chunk_dl(d), schunk_dl(d), C.gtotal.p, C.caps, C.genCount.p, genPoolCap,
                                  C.gofs.p, C.tasks.p, C.stasks.p, C.reds.p, dsum, C.segs.p, C.segcnt.p, C.gslice.p,
                                  C.genOverflow.p, dw, dtab, njobs, nwp));
            if (!mixed)
                HC(launch_enum(s, 1, WCAP, 0, dw, djt, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, C.gofs.p, C.tuples.p,
                               dsum, C.segs.p, nwp, queue + 2));
            else
                HC(launch_enum_mixed(s, 1, K, WCAP, 0, dw, djt, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, C.gofs.p,
                                     C.tuples.p, C.slices.p, dsum, C.segs.p, C.gslice.p, nwp, queue + 2));
            const int grid64 = std::max(1, (C.caps.tasks + 3) / 4);
            const int gridred = std::min(4096, std::max(1, C.caps.reds));
            HC(launch_contract_fp64(s, K, 0, dsum, grid64, C.tasks.p, dw, djt, C.tabs.p, C.tuples.p, C.genPool.p,
                                    C.partials.p, dts));
            if (mixed) {
                const int tpb = mfma_tasks_per_block(K);
                const int gridmf = std::max(1, (C.caps.stasks + tpb - 1) / tpb);
                HC(launch_contract_mfma(s, K, 0, dsum, gridmf, C.stasks.p, dw, djt, C.tabs.p, C.slices.p, C.genPool.p,
                                        C.nodeslots.p, dtab, C.genPool16.p, C.genExp16.p));
            }
            HC(launch_reduce_partials(s, 0, dsum, gridred, C.reds.p, dw, djt, C.partials.p,
                                      mixed ? C.nodeslots.p : nullptr, K3, mixed ? dwn : nullptr));
            // (the kernel after the reduction stamps the contraction's end)
            if (!mixed) HC(launch_work_norms(s, WCAP, dw, djt, K3, dwn, nwp, dts + 1));
            HC(launch_loop_split(s, njobs, C.ljobs.p, dw, dwn, M.maxDepth, C.lctl.p, C.lcidx.p, dtab,
                                 mixed ? dts + 1 : nullptr));
            HC(launch_loop_emit(s, njobs, C.ljobs.p, dw, nwp, C.lcidx.p, C.lwork[cur ^ 1].p, nwnext, WCAP, dsum,
                                C.lctl.p, d, djt, C.titems.p, tcap, C.liter_d));
            HC(hipEventRecord(C.evIter[d], s));  // (k_loop_emit wrote the host's entry d)
        };
        int d = 0, enq = 0, maxnw = njobs * nr;
        bool ok = true;
        enqueue(0);
        enq = 1;
        while (true) {
            if (d + 1 <= M.maxDepth) {  // the next iteration, before this one is known to have children
                enqueue(d + 1);
                enq = d + 2;
            }
            hipError_t e;
            while ((e = hipEventQuery(C.evIter[d])) == hipErrorNotReady) {
            }
            HC(e);
            LoopIter it;
            it.nw_next = *(volatile int *)&C.liter_h[d].nw_next;
            it.overflow = *(volatile int *)&C.liter_h[d].overflow;
            iter = d + 1;
            if (it.overflow) {
                ok = false;
                break;
            }
            maxnw = std::max(maxnw, it.nw_next);
            if (it.nw_next == 0 || d + 1 >= enq) break;
            d++;
        }
        if (!ok) {
            stream_wait(C);
            C.stage.off = 0;
            return false;
        }
        PT.lap("loop iterations");
        // behind the last (empty) iteration: the topology read-back, then the transforms over
        // the device's item lists (the host rebuilds its trees while the GPU transforms)
        const size_t tb = 6 * topo * sizeof(int), jb = (size_t)njobs * sizeof(LoopJob);
        const size_t sb = sizeof(unsigned long long) * 2 * LOOP_MAX_DEPTH;
        C.rb.reserve(tb + jb + sb + 64);
        HC(hipMemcpyAsync(C.rb.h, C.ltopo.p, tb, hipMemcpyDeviceToHost, s));
        HC(hipMemcpyAsync(C.rb.h + tb, C.ljobs.p, jb, hipMemcpyDeviceToHost, s));
        HC(hipMemcpyAsync(C.rb.h + tb + jb, ctlb + offsetof(LoopCtl, ts), sb, hipMemcpyDeviceToHost, s));
        HC(hipMemcpyAsync(C.rb.h + tb + jb + sb, C.genCount.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HC(hipEventRecord(C.evTopo, s));
        {
            std::vector<int> tn(d + 1), to(d + 2, 0);
            for (int q = 0; q <= d; q++) {
                tn[q] = C.liter_h[q].nw_next / 8;
                to[q + 1] = to[q] + tn[q];
            }
            for (int q = 0; q <= d; q++) HC(launch_transform(s, 1, tn[q], C.titems.p + to[q], M.WT.p, K));
            for (int q = d; q >= 0; q--) HC(launch_transform(s, 0, tn[q], C.titems.p + to[q], M.W.p, K));
        }
        {
            hipError_t e;
            while ((e = hipEventQuery(C.evTopo)) == hipErrorNotReady) {
            }
            HC(e);
        }
        C.stage.off = 0;
        const LoopJob *hl = (const LoopJob *)(C.rb.h + tb);
        if (enq - 1 > d && C.liter_h[enq - 1].overflow)  // (an iteration without work cannot overflow)
            throw Error(MWG_ERR_HIP, "device refinement loop: overflow in an empty iteration");
        int genNow = 0;
        std::memcpy(&genNow, C.rb.h + tb + jb + sb, sizeof(int));
        {
            const unsigned long long *hts = (const unsigned long long *)(C.rb.h + tb + jb);
            for (int q = 0; q < enq; q++) t_contract_ms += clock_ms(C, hts[2 * q], hts[2 * q + 1]);
        }
        std::vector<size_t> toff(njobs + 1, 0);
        for (int j = 0; j < njobs; j++) toff[j + 1] = toff[j] + (size_t)lj[j].cap;
        size_t nodes = 0;
        for (int j = 0; j < njobs; j++) nodes += (size_t)hl[j].n;
        auto rebuild = [&](size_t j) {
            const int *base = (const int *)C.rb.h + 6 * toff[j];
            const size_t cap = (size_t)lj[j].cap;
            JobH &J = jobs[j];
            J.out->set_topology(hl[j].n, base, base + cap, base + 2 * cap, base + 3 * cap);
            J.tuples = hl[j].tuples;
            J.tuples_fp64 = hl[j].tuples_fp64;
            J.slices = hl[j].slices;
            J.gen = hl[j].gen;
            J.alive.assign(J.out->n(), 1);
        };
        if (nodes < 65536)  // (threads cost more than they save on small batches)
            for (int j = 0; j < njobs; j++) rebuild((size_t)j);
        else
            for_each_job((size_t)njobs, rebuild);
        C.hw_nw = std::max(C.hw_nw, maxnw);
        C.hw_gen = std::max(C.hw_gen, genNow);
        for (JobH &J : jobs) C.hw_ratio = std::max(C.hw_ratio, (double)J.out->n() / std::max(1, J.in->n()));
        genDone = genNow;
        PT.lap("loop read-back");
        return true;
    };
    if (dl_plan) {
        dl_done = device_loop();
        if (!dl_done) {
            // a capacity was too small: the host loop from the roots, from a clean state
            for (JobH &J : jobs) {
                J.out->make_roots();
                J.work = J.out->roots;
                J.sNorm = J.wNorm = 0.0;
                J.gsq = -1.0;
                J.tuples = J.tuples_fp64 = J.slices = J.gen = 0;
                J.spec.assign(J.out->n(), 0);
                J.alive.assign(J.out->n(), 1);
                J.has_feat.assign(J.out->n(), 0);
            }
            HC(hipMemsetAsync(C.genHashKeys.p, 0, genTotal * sizeof(unsigned long long), s));
            HC(hipMemsetAsync(C.genCount.p, 0, sizeof(int), s));
            genDone = genLastNew = 0;
            iter = 0;
            t_contract_ms = 0.0f;
        }
    }
    while (!dl_done) {
        std::vector<int> active;
        for (int j = 0; j < (int)jobs.size(); j++)
            if (!jobs[j].work.empty()) active.push_back(j);
        if (active.empty()) break;
        // one job table per (job, depth) present in the frontier
        std::vector<BoxJob> bjobs;
        int ntabs = 0, max_terms = 0;
        std::vector<JobTab> jt;
        std::vector<int> jt_job;
        std::vector<WorkItem> work;
        for (int j : active) jobs[j].out->ensure_dev(jobs[j].out->n());
        PT.lap("ensure_dev");
        {
            std::vector<std::pair<OperH *, int>> deeper;
            for (int j : active)
                for (int g : jobs[j].work)
                    if (!jobs[j].op->has(jobs[j].out->depth[g])) deeper.push_back({jobs[j].op, jobs[j].out->depth[g]});
            if (!deeper.empty()) ensure_levels(deeper, mixed);
        }
        for (int j : active) {
            JobH &J = jobs[j];
            std::vector<int> depths;
            for (int g : J.work) depths.push_back(J.out->depth[g]);
            std::sort(depths.begin(), depths.end());
            depths.erase(std::unique(depths.begin(), depths.end()), depths.end());
            std::vector<int> tab_of(M.maxDepth + 2, -1);
            for (int d : depths) {
                tab_of[d] = (int)jt.size();
                add_table(J, d, jt, bjobs, ntabs, max_terms);
                jt_job.push_back(j);
            }
            J.wpos.assign(J.out->n(), -1);
            for (int g : J.work) {
                J.wpos[g] = (int)work.size();
                work.push_back({tab_of[J.out->depth[g]], g, {J.out->lr[3 * g], J.out->lr[3 * g + 1], J.out->lr[3 * g + 2]},
                                J.out->parent[g]});
            }
        }
        const int nw = (int)work.size();
        C.hw_nw = std::max(C.hw_nw, nw);
        PT.lap("opers+boxes");
        // per-table prefix arrays (ntab + 1 entries each), filled on the device
        C.tpre.alloc((size_t)ntabs + jt.size() + 1);
        C.tpre0.alloc((size_t)ntabs + jt.size() + 1);
        for (size_t q = 0; q < jt.size(); q++) {
            jt[q].pre = C.tpre.p + jt[q].tab0 + q;
            jt[q].pre0 = C.tpre0.p + jt[q].tab0 + q;
        }
        // the iteration's tables in one upload: work items | job tables | box requests
        const size_t o_jobs = (work.size() * sizeof(WorkItem) + 255) & ~(size_t)255;
        const size_t o_bj = o_jobs + ((jt.size() * sizeof(JobTab) + 255) & ~(size_t)255);
        const size_t tbytes = o_bj + bjobs.size() * sizeof(BoxJob);
        {
            std::vector<char> blob(tbytes);
            std::memcpy(blob.data(), work.data(), work.size() * sizeof(WorkItem));
            std::memcpy(blob.data() + o_jobs, jt.data(), jt.size() * sizeof(JobTab));
            std::memcpy(blob.data() + o_bj, bjobs.data(), bjobs.size() * sizeof(BoxJob));
            upload_async(C, C.iterbuf, blob.data(), blob.size());
        }
        PT.lap("tables+upload");
        WorkItem *dwork = (WorkItem *)C.iterbuf.p;
        JobTab *djobs = (JobTab *)(C.iterbuf.p + o_jobs);
        BoxJob *dbjobs = (BoxJob *)(C.iterbuf.p + o_bj);
        // the iteration's read-back block: schedule summary | per-table counts (FP64 tuples,
        // slices, slice tuples, generated nodes) | work norms | child probabilities | child
        // features.  The table counts start at zero (the generated-node counts are added
        // from the marking onwards).
        const bool want_feat = use_pred || C.tracing;
        const size_t o_tab = 256;
        const size_t o_wn = (o_tab + 4 * jt.size() * sizeof(long long) + 255) & ~(size_t)255;
        const size_t o_pr = (o_wn + (size_t)nw * 8 * sizeof(double) + 255) & ~(size_t)255;
        const size_t o_ft = (o_pr + (use_pred ? (size_t)nw * 8 * sizeof(float) : 0) + 255) & ~(size_t)255;
        const size_t rb_dev = o_ft + (want_feat ? (size_t)nw * 8 * MLP_FEATURES * sizeof(float) : 0);
        const size_t rb_host = C.tracing ? rb_dev : o_ft;  // the features only when tracing
        C.rbd.alloc(rb_dev);
        TaskSummary *dsum = (TaskSummary *)C.rbd.p;
        long long *dtab = (long long *)(C.rbd.p + o_tab);
        double *dwn = (double *)(C.rbd.p + o_wn);
        float *dprob = (float *)(C.rbd.p + o_pr), *dfeat = (float *)(C.rbd.p + o_ft);
        // (contraction timestamps at byte 192 of the header, cleared with the table counts)
        unsigned long long *dts = (unsigned long long *)(C.rbd.p + 192);
        static_assert(sizeof(TaskSummary) <= 192, "read-back header layout");
        HC(hipMemsetAsync(C.rbd.p + 192, 0, (o_tab - 192) + 4 * jt.size() * sizeof(long long), s));
        C.tabs.alloc((size_t)std::max(ntabs, 1));
        C.tbmax.alloc((size_t)std::max<size_t>(bjobs.size() * max_terms, 1));
        HC(launch_job_boxes(s, (int)bjobs.size(), max_terms, dbjobs, C.tabs.p, C.tbmax.p, djobs, C.tpre.p,
                            C.tpre0.p, C.genOverflow.p));
        PT.lap("launch boxes");
        // generated input nodes: marked, then computed for the range the device counter
        // reached (read back with the work-list counts below, not before)
        {
            const int want = std::min(genCap, genDone + std::max(4096, 2 * genLastNew + 1024));
            if (genPoolCap < want) {
                C.genPool.grow_keep((size_t)want * K3, s);
                C.genNorm.grow_keep((size_t)want, s);
                if (mixed) {
                    C.genPool16.grow_keep((size_t)want * split16_block_uint4(K, 1, K3), s);
                    C.genExp16.grow_keep((size_t)want, s);
                }
                genPoolCap = want;
            }
        }
        GenState gs{C.genKeys.p, C.genJob.p, C.genCount.p, genCap, C.genOverflow.p, C.genPool.p, C.genNorm.p,
                    genPoolCap, dtab};
        HC(launch_gen_mark(s, nw, dwork, djobs, C.tabs.p, gs));
        HC(launch_gen_compute(s, genDone, genPoolCap, djobs, gs, M.W.p, K));
        if (mixed)
            HC(launch_split16(s, genPoolCap - genDone, genDone, C.genCount.p, K3, K, C.genPool.p, C.genPool16.p,
                              C.genExp16.p, 1));
        PT.lap("upload+genmark");
        // Work list and contraction without a host round trip: counts per segment
        // (nseg per output node), a device-built schedule (offsets, tasks, reduce items),
        // the tuple/slice fill, the contractions and the output norms; the host reads
        // the norms, the schedule sizes and the counts back once.  When a capacity (or
        // the generated-node pool) was too small, nothing of the schedule ran: grow and
        // run it again.
        const int nseg = mixed ? 9 : 8;
        // few output nodes (the roots): several workgroups share each node's candidates
        const int nparts = enum_parts(nw);
        C.thread_off.alloc((size_t)nw * nseg * nparts * 256);
        C.gtotal.alloc((size_t)nw * nseg * nparts);
        C.gofs.alloc((size_t)nw * nseg * nparts);
        C.segs.alloc((size_t)nw * 8);
        C.segcnt.alloc((size_t)nw * nseg);
        C.gslice.alloc((size_t)nw);
        auto launch_count = [&]() {
            if (!mixed)
                HC(launch_enum(s, 0, nw, nparts, dwork, djobs, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, nullptr,
                               nullptr, nullptr, nullptr));
            else
                HC(launch_enum_mixed(s, 0, K, nw, nparts, dwork, djobs, C.tabs.p, gs, C.thread_off.p, C.gtotal.p,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
        };
        auto ensure_caps = [&](const TaskSummary *need) {
            auto grow = [](long long have, long long want) {
                return want <= have ? have : std::max<long long>(want + want / 4, 1024);
            };
            if (need) {
                C.caps.tuples = grow(C.caps.tuples, need->ntuples);
                C.caps.slices = grow(C.caps.slices, need->nslices);
                C.caps.tasks = (int)grow(C.caps.tasks, need->ntasks);
                C.caps.stasks = (int)grow(C.caps.stasks, need->nstasks);
                C.caps.reds = (int)grow(C.caps.reds, need->nreds);
                C.caps.slots = (int)grow(C.caps.slots, need->nslots);
                C.caps.nnslots = (int)grow(C.caps.nnslots, need->nnslots);
            }
            C.tuples.alloc((size_t)std::max<long long>(C.caps.tuples, 1));
            C.tasks.alloc((size_t)std::max(C.caps.tasks, 1));
            C.reds.alloc((size_t)std::max(C.caps.reds, 1));
            C.partials.alloc((size_t)std::max(C.caps.slots, 1) * K3);
            if (mixed) {
                C.slices.alloc((size_t)std::max<long long>(C.caps.slices, 1));
                C.stasks.alloc((size_t)std::max(C.caps.stasks, 1));
                C.nodeslots.alloc((size_t)std::max(C.caps.nnslots, 1) * 8 * K3);
            }
        };
        std::vector<double> wn((size_t)nw * 8);
        std::vector<long long> tabcnt(4 * jt.size());
        std::vector<float> cprob(use_pred ? (size_t)nw * 8 : 0);
        std::vector<float> cfeat(C.tracing ? (size_t)nw * 8 * MLP_FEATURES : 0);
        TaskSummary hsum;
        unsigned long long hts[2] = {0ull, 0ull};
        int chunk_it = C.chunk;  // FP64 task size of this iteration (raised when partial slots would not fit)
        // slices per MFMA task (a workgroup task).  k = 7: at most MFMA8_TASK_SLICES (the
        // waves sum a quarter of them each in FP32), MFMA8_TASK_DEFAULT when automatic;
        // k = 9: about 2048 tasks per iteration when automatic (16 to 128 slices each),
        // raised when node slots would not fit
        int schunk_it = K == 8 ? std::min(C.slice_chunk > 0 ? C.slice_chunk : MFMA8_TASK_DEFAULT, MFMA8_TASK_SLICES)
                               : (C.slice_chunk > 0 ? C.slice_chunk : -2048);
        auto launch_schedule = [&]() {
            ensure_caps(nullptr);
            HC(launch_build_tasks(s, nw, nseg, nparts, chunk_it, schunk_it, C.gtotal.p, C.caps, C.genCount.p,
                                  genPoolCap, C.gofs.p, C.tasks.p, C.stasks.p, C.reds.p, dsum, C.segs.p,
                                  C.segcnt.p, C.gslice.p, C.genOverflow.p, dwork, dtab, (int)jt.size()));
            if (!mixed)
                HC(launch_enum(s, 1, nw, nparts, dwork, djobs, C.tabs.p, gs, C.thread_off.p, C.gtotal.p, C.gofs.p,
                               C.tuples.p, dsum, C.segs.p));
            else
                HC(launch_enum_mixed(s, 1, K, nw, nparts, dwork, djobs, C.tabs.p, gs, C.thread_off.p, C.gtotal.p,
                                     C.gofs.p, C.tuples.p, C.slices.p, dsum, C.segs.p, C.gslice.p));
            const int grid64 = std::max(1, (C.caps.tasks + 3) / 4);
            const int gridred = std::min(4096, std::max(1, C.caps.reds));
            // contraction time: the device clock at the first workgroup of the first and the
            // last workgroup of the last contraction kernel
            HC(launch_contract_fp64(s, K, 0, dsum, grid64, C.tasks.p, dwork, djobs, C.tabs.p, C.tuples.p,
                                    C.genPool.p, C.partials.p, dts));
            if (mixed) {
                const int tpb = mfma_tasks_per_block(K);
                const int gridmf = std::max(1, (C.caps.stasks + tpb - 1) / tpb);
                HC(launch_contract_mfma(s, K, 0, dsum, gridmf, C.stasks.p, dwork, djobs, C.tabs.p, C.slices.p,
                                        C.genPool.p, C.nodeslots.p, dtab, C.genPool16.p, C.genExp16.p));
            }
            // (mixed: the reduce items cover every output component and write the norms)
            HC(launch_reduce_partials(s, 0, dsum, gridred, C.reds.p, dwork, djobs, C.partials.p,
                                      mixed ? C.nodeslots.p : nullptr, K3, mixed ? dwn : nullptr));
            // (the contraction's end: stamped by the next kernel on the stream)
            if (!mixed) HC(launch_work_norms(s, nw, dwork, djobs, K3, dwn, nullptr, dts + 1));
            else HC(launch_stamp(s, dts + 1));
            // refinement predictor: features of every work node's children and their split
            // probabilities, read back with the norms (no extra round trip)
            if (want_feat) {
                HC(launch_child_features(s, nw, dwork, djobs, dwn, dfeat));
                if (use_pred)
                    HC(launch_mlp_forward(s, nw * 8, nullptr, dfeat, C.pred->mean.p, C.pred->sdev.p, C.pred->P.p,
                                          dprob, nullptr, nullptr, nullptr));
            }
            C.rb.reserve(rb_host);
            char *h = C.rb.h;
            HC(hipMemcpyAsync(h, C.rbd.p, rb_host, hipMemcpyDeviceToHost, s));
            PT.lap("launch sched");
            stream_wait(C);
            PT.lap("gpu wait");
            C.stage.off = 0;
            std::memcpy(&hsum, h, sizeof(TaskSummary));
            std::memcpy(hts, h + 192, sizeof(hts));
            std::memcpy(tabcnt.data(), h + o_tab, tabcnt.size() * sizeof(long long));
            std::memcpy(wn.data(), h + o_wn, wn.size() * sizeof(double));
            if (use_pred) std::memcpy(cprob.data(), h + o_pr, cprob.size() * sizeof(float));
            if (C.tracing) std::memcpy(cfeat.data(), h + o_ft, cfeat.size() * sizeof(float));
            if (hsum.gen_overflow) throw GenOverflow();
        };
        launch_count();
        PT.lap("enum_count");
        launch_schedule();
        // A capacity (or the generated-node pool) was too small: grow and run again.  A
        // partial-slot need beyond SLOT_BUDGET raises this iteration's FP64 task size
        // instead (fewer, larger tasks: only the summation grouping changes).
        for (int attempt = 0; hsum.overflow; attempt++) {
            if (attempt == 12) throw Error(MWG_ERR_RANGE, "work-list schedule sizing failed");
            if (hsum.overflow & 2) {
                // finish the generated nodes past the pool, then count again
                const int genNow = hsum.gen_count;
                const int want = std::min(genCap, genNow + std::max(4096, genNow / 2));
                C.genPool.grow_keep((size_t)want * K3, s);
                C.genNorm.grow_keep((size_t)want, s);
                gs.pool = C.genPool.p;
                gs.norm = C.genNorm.p;
                gs.pcap = want;
                HC(launch_gen_compute(s, genPoolCap, want, djobs, gs, M.W.p, K));
                if (mixed) {
                    C.genPool16.grow_keep((size_t)want * split16_block_uint4(K, 1, K3), s);
                    C.genExp16.grow_keep((size_t)want, s);
                    HC(launch_split16(s, want - genPoolCap, genPoolCap, C.genCount.p, K3, K, C.genPool.p,
                                      C.genPool16.p, C.genExp16.p, 1));
                }
                genPoolCap = want;
                launch_count();
            }
            TaskSummary need = hsum;
            if ((double)need.nslots * K3 * sizeof(double) > SLOT_BUDGET && chunk_it < 4096) {
                // twice the task size; the slot capacity stays, the next schedule tells
                chunk_it *= 2;
                need.nslots = 0;
            }
            if (K != 8 && (double)need.nnslots * 8 * K3 * sizeof(double) > SLOT_BUDGET &&
                hsum.schunk < MAX_SLICE_CHUNK) {
                schunk_it = 2 * hsum.schunk;  // the same for the MFMA tasks' node slots
                need.nnslots = 0;
                need.nstasks = 0;
            }
            ensure_caps(&need);
            HC(hipMemsetAsync(dts, 0, 2 * sizeof(unsigned long long), s));  // (time the last attempt only)
            launch_schedule();
        }
        genLastNew = hsum.gen_count - genDone;
        genDone = hsum.gen_count;
        if (iter < LOOP_MAX_DEPTH) {
            C.hw_chunk[iter] = std::max(C.hw_chunk[iter], chunk_it);
            // (a set slice_chunk is raised the same way when the node slots would not fit)
            if (K != 8 && schunk_it > 0) C.hw_schunk[iter] = std::max(C.hw_schunk[iter], schunk_it);
        }
        if (PT.on) {
            // candidates walked by the work-list kernels: per work node, its table's prefix total
            std::vector<int> pre((size_t)ntabs + jt.size() + 1);
            C.tpre.download(pre.data(), pre.size(), s);
            long long cand = 0;
            for (int i = 0; i < nw; i++) {
                const JobTab &T = jt[work[i].job];
                cand += pre[(size_t)T.tab0 + work[i].job + T.ntab];
            }
            fprintf(stderr,
                    "[mwg timing] iteration %d: %d work nodes, %lld candidates, %lld tuples, %lld slices, %d generated "
                    "(pool %d)\n",
                    iter, nw, cand, hsum.ntuples, hsum.nslices, genLastNew, genPoolCap);
        }
        for (size_t q = 0; q < jt.size(); q++) {
            JobH &J = jobs[jt_job[q]];
            J.tuples += tabcnt[4 * q] + tabcnt[4 * q + 2];
            J.tuples_fp64 += tabcnt[4 * q];
            J.slices += tabcnt[4 * q + 1];
            J.gen += tabcnt[4 * q + 3];
        }
        PT.lap("worklist+launch");
        t_contract_ms += clock_ms(C, hts[0], hts[1]);
        PT.lap("contract+norms");
        for (int i = 0; i < nw; i++) {
            JobH &J = jobs[jt_job[work[i].job]];
            std::memcpy(&J.out->cnorm[8 * (size_t)work[i].node], &wn[8 * (size_t)i], 8 * sizeof(double));
        }
        // split tests level by level (the plain method's order): the running norm of a
        // level includes every computed node of that level and above; speculative
        // children are kept only when the norm test splits their parent
        std::vector<int> cand_job, cand_node;
        std::vector<float> cand_prob;
        // without traces or the predictor (shared lists) the jobs' tests are independent:
        // large batches spread them over host threads
        const bool par_split = !C.tracing && !use_pred && active.size() >= 16;
        auto split_job = [&](int j) {
            JobH &J = jobs[j];
            std::vector<int> cur = J.work;
            bool one_depth = true;
            for (size_t q = 1; q < cur.size() && one_depth; q++)
                one_depth = J.out->depth[cur[q]] == J.out->depth[cur[0]];
            if (!one_depth)
                std::stable_sort(cur.begin(), cur.end(),
                                 [&](int a, int b) { return J.out->depth[a] < J.out->depth[b]; });
            std::vector<int> next;
            size_t q0 = 0;
            while (q0 < cur.size()) {
                const int d = J.out->depth[cur[q0]];
                size_t q1 = q0;
                while (q1 < cur.size() && J.out->depth[cur[q1]] == d) q1++;
                for (size_t q = q0; q < q1; q++) {
                    const int g = cur[q];
                    if (!J.alive[g]) continue;
                    const double *p = &J.out->cnorm[8 * (size_t)g];
                    if (d == 0) J.sNorm += p[0] * p[0];
                    for (int c = 1; c < 8; c++) J.wNorm += p[c] * p[c];
                }
                J.gsq = J.sNorm + J.wNorm;
                const double tnorm = std::sqrt(J.gsq);
                for (size_t q = q0; q < q1; q++) {
                    const int g = cur[q];
                    if (!J.alive[g]) continue;
                    double w2 = 0.0;
                    const double *p = &J.out->cnorm[8 * (size_t)g];
                    for (int c = 1; c < 8; c++) w2 += p[c] * p[c];
                    const bool can = d < M.maxDepth && !(max_iter >= 0 && d >= max_iter);
                    const bool split = can && split_check(M.n0 + d, std::sqrt(w2), J.prec, tnorm);
                    if (C.tracing && J.has_feat[g]) {
                        C.trace_X.insert(C.trace_X.end(), &J.feat[(size_t)g * MLP_FEATURES],
                                         &J.feat[(size_t)g * MLP_FEATURES] + MLP_FEATURES);
                        C.trace_y.push_back(split ? 1.0f : 0.0f);
                    }
                    if (J.spec[g]) {
                        J.spec[g] = 0;
                        if (!split) {
                            for (int a = 0; a < 8; a++) J.alive[J.out->child[g] + a] = 0;
                            J.out->child[g] = -1;
                            wasted += 8;
                        }
                        continue;
                    }
                    if (split) {
                        J.out->split(g);
                        J.spec.resize(J.out->n(), 0);
                        J.alive.resize(J.out->n(), 1);
                        J.has_feat.resize(J.out->n(), 0);
                        const int c0 = J.out->child[g], wi = J.wpos[g];
                        if (C.tracing) {
                            J.feat.resize((size_t)J.out->n() * MLP_FEATURES, 0.0f);
                            std::memcpy(&J.feat[(size_t)c0 * MLP_FEATURES], &cfeat[(size_t)wi * 8 * MLP_FEATURES],
                                        8 * MLP_FEATURES * sizeof(float));
                            for (int a = 0; a < 8; a++) J.has_feat[c0 + a] = 1;
                        }
                        for (int a = 0; a < 8; a++) {
                            next.push_back(c0 + a);
                            if (use_pred) {
                                cand_job.push_back(j);
                                cand_node.push_back(c0 + a);
                                cand_prob.push_back(cprob[(size_t)wi * 8 + a]);
                            }
                        }
                    }
                }
                q0 = q1;
            }
            J.work.swap(next);
        };
        PT.lap("norms to trees");
        if (par_split)
            for_each_job(active.size(), [&](size_t q) { split_job(active[q]); });
        else
            for (int j : active) split_job(j);
        PT.lap("split tests");
        // learned lookahead: nodes predicted to split get their children computed now
        if (use_pred && !cand_node.empty()) {
            const int nc = (int)cand_node.size();
            for (int i = 0; i < nc; i++) {
                if (cand_prob[i] < C.pred_thr) continue;
                JobH &J = jobs[cand_job[i]];
                const int c = cand_node[i];
                const int d = J.out->depth[c];
                if (d >= M.maxDepth || (max_iter >= 0 && d >= max_iter)) continue;
                J.out->split(c);
                J.spec.resize(J.out->n(), 0);
                J.alive.resize(J.out->n(), 1);
                J.has_feat.resize(J.out->n(), 0);
                J.spec[c] = 1;
                for (int a = 0; a < 8; a++) J.work.push_back(J.out->child[c] + a);
                predicted += 8;
            }
        }
        PT.lap("split");
        iter++;
    }
    if (!dl_done) {
        C.hw_gen = std::max(C.hw_gen, genDone);
        for (JobH &J : jobs) C.hw_ratio = std::max(C.hw_ratio, (double)J.out->n() / std::max(1, J.in->n()));
    }
    C.last_in.resize(jobs.size());
    C.last_out.resize(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) {
        C.last_in[j] = jobs[j].in->n();
        C.last_out[j] = jobs[j].out->n();
    }
    // drop the speculative children that the norm test rejected
    for (JobH &J : jobs) {
        bool any_dead = false;
        for (char a : J.alive)
            if (!a) any_dead = true;
        if (any_dead) compact_tree(*J.out, J.alive);
    }
    // top-down additive transform, then bottom-up compression, for all output trees (the
    // device loop has queued them already, over the item lists it wrote)
    if (!dl_done) {
        std::vector<std::vector<std::vector<int>>> L(jobs.size());
        int maxd = 0;
        for (size_t j = 0; j < jobs.size(); j++) {
            L[j] = jobs[j].out->internal_by_depth();
            maxd = std::max(maxd, (int)L[j].size());
        }
        // one item list for all levels (uploaded once); the same items serve both directions.
        // The transforms also write the norms they change: the children's scaling norms
        // (top-down; final for leaves) and all 8 norms of every parent (bottom-up)
        std::vector<TransformItem> items;
        std::vector<size_t> lofs(maxd + 1, 0);
        for (int d = 0; d < maxd; d++) {
            lofs[d] = items.size();
            for (size_t j = 0; j < jobs.size(); j++)
                if (d < (int)L[j].size())
                    for (int p : L[j][d])
                        items.push_back({jobs[j].out->coef.p, jobs[j].out->norm.p, p, jobs[j].out->child[p]});
        }
        lofs[maxd] = items.size();
        PT.lap("transform prep");
        if (!items.empty()) {
            upload_async(C, C.titems, items.data(), items.size());
            for (int d = 0; d < maxd; d++)
                HC(launch_transform(s, 1, (int)(lofs[d + 1] - lofs[d]), C.titems.p + lofs[d], M.WT.p, K));
            for (int d = maxd - 1; d >= 0; d--)
                HC(launch_transform(s, 0, (int)(lofs[d + 1] - lofs[d]), C.titems.p + lofs[d], M.W.p, K));
        }
        PT.lap("transform launch");
    }
    {
        size_t total = 0;
        for (JobH &J : jobs) total += 8 * (size_t)J.out->n();
        C.rb.reserve(total * sizeof(double));
        double *h = (double *)C.rb.h;
        size_t off = 0;
        for (JobH &J : jobs) {
            const size_t m = 8 * (size_t)J.out->n();
            if (m) HC(hipMemcpyAsync(h + off, J.out->norm.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
            off += m;
        }
        stream_wait(C);
        C.stage.off = 0;
        off = 0;
        for (JobH &J : jobs) {
            const size_t m = 8 * (size_t)J.out->n();
            J.out->cnorm.assign(h + off, h + off + m);
            off += m;
            J.out->sqnorm_from_cnorm();
        }
    }

    for (JobH &J : jobs) J.out->fit_dev();
    PT.lap("transforms");
    // the inputs' split-FP16 copies live for one apply (an SCF iteration brings new inputs;
    // at k = 9 a copy is 5/8 of the FP64 tree)
    if (mixed) {
        HC(hipStreamSynchronize(s));
        for (JobH &J : jobs) J.in->drop16();
    }
    PT.report();
    if (PT.on)
        fprintf(stderr, "[mwg timing] block cache: %.2f GiB kept, %lld blocks freed so far\n",
                (double)C.cache.bytes / (1 << 30), C.cache.evicted);
    const double t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    stats.assign(jobs.size(), mwg_apply_stats{});
    const double fl = 6.0 * K * K * K * K;
    for (size_t j = 0; j < jobs.size(); j++) {
        mwg_apply_stats &st = stats[j];
        st.nodes_out = jobs[j].out->n();
        st.tuples = jobs[j].tuples;
        st.tuples_fp64 = jobs[j].tuples_fp64;
        st.slices = jobs[j].slices;
        st.gen_nodes = jobs[j].gen;
        st.iters = iter;
        st.predicted_splits = predicted;
        st.predicted_wasted = wasted;
        st.t_apply = t_total;
        st.t_contract = t_contract_ms * 1e-3;
        st.device_loop = dl_done ? 1 : 0;
        st.flops = fl * (double)jobs[j].tuples_fp64 + slice_flops(K) * (double)jobs[j].slices;
    }
}


// ------------------------------------------------------------------ multiplication

// the node of T at (d, l) (steps 0) or its leaf ancestor and the levels between
static void locate(const TreeH &T, int d, const int *l, int &src, int &steps) {
    for (int dd = d; dd >= 0; dd--) {
        const int sh = d - dd;
        const int i = T.find(dd, l[0] >> sh, l[1] >> sh, l[2] >> sh);
        if (i >= 0) {
            src = i;
            steps = sh;
            return;
        }
    }
    throw Error(MWG_ERR_STATE, "input tree does not cover the box");
}

static void multiply_trees(TreeH &out, double prec, double coef, TreeH &A, TreeH &B) {
    MraH &M = *A.mra;
    Ctx &C = *M.ctx;
    hipStream_t s = C.stream;
    if (A.n() == 0 || B.n() == 0) throw Error(MWG_ERR_STATE, "multiply: empty input tree");
    out.make_roots();
    // union grid
    std::vector<int> order(out.roots.begin(), out.roots.end());
    for (size_t w = 0; w < order.size(); w++) {
        const int i = order[w];
        const int d = out.depth[i], x = out.lr[3 * i], y = out.lr[3 * i + 1], z = out.lr[3 * i + 2];
        const int ia = A.find(d, x, y, z), ib = B.find(d, x, y, z);
        const bool sp = (ia >= 0 && A.child[ia] >= 0) || (ib >= 0 && B.child[ib] >= 0);
        if (sp && d < M.maxDepth) {
            out.split(i);
            for (int a = 0; a < 8; a++) order.push_back(out.child[i] + a);
        }
    }
    std::vector<int> work;
    for (int i = 0; i < out.n(); i++) work.push_back(i);
    double sNorm = 0.0, wNorm = 0.0;
    bool first = true;
    DBuf<MulItem> ditems;
    DBuf<int> dwork;
    while (!work.empty()) {
        out.ensure_dev(out.n());
        std::vector<MulItem> items(work.size());
        for (size_t w = 0; w < work.size(); w++) {
            const int i = work[w];
            MulItem &it = items[w];
            it.node = i;
            it.depth = out.depth[i];
            for (int q = 0; q < 3; q++) it.l[q] = out.lr[3 * i + q];
            locate(A, it.depth, it.l, it.asrc, it.asteps);
            locate(B, it.depth, it.l, it.bsrc, it.bsteps);
        }
        ditems.upload(items.data(), items.size(), s);
        HC(launch_multiply(s, (int)items.size(), ditems.p, M.geom(), M.dev(), coef, A.coef.p, B.coef.p, out.coef.p));
        dwork.upload(work.data(), work.size(), s);
        HC(launch_norms(s, (int)work.size(), dwork.p, out.coef.p, out.norm.p, out.K3()));
        std::vector<double> nrm(8 * (size_t)out.n());
        out.norm.download(nrm.data(), nrm.size(), s);
        for (int i : work) {
            std::memcpy(&out.cnorm[8 * (size_t)i], &nrm[8 * (size_t)i], 8 * sizeof(double));
            if (first && out.depth[i] == 0) sNorm += nrm[8 * i] * nrm[8 * i];
            for (int c = 1; c < 8; c++) wNorm += nrm[8 * i + c] * nrm[8 * i + c];
        }
        first = false;
        const double tnorm = std::sqrt(sNorm + wNorm);
        std::vector<int> next;
        for (int i : work) {
            if (out.child[i] >= 0) continue;
            const int d = out.depth[i];
            if (d >= M.maxDepth) continue;
            double wn = 0.0;
            for (int c = 1; c < 8; c++) wn += out.cnorm[8 * i + c] * out.cnorm[8 * i + c];
            if (split_check(M.n0 + d, std::sqrt(wn), prec, tnorm)) {
                out.split(i);
                for (int a = 0; a < 8; a++) next.push_back(out.child[i] + a);
            }
        }
        work.swap(next);
    }
    auto L = out.internal_by_depth();
    std::vector<TransformItem> titems;
    for (int d = (int)L.size() - 1; d >= 0; d--) {
        titems.clear();
        for (int p : L[d]) titems.push_back({out.coef.p, nullptr, p, out.child[p]});
        C.titems.upload(titems.data(), titems.size(), s);
        HC(launch_transform(s, 0, (int)titems.size(), C.titems.p, M.W.p, M.B.K));
    }
    out.refresh_norms();
    out.fit_dev();
}

// Many independent products as one: the same items, kernels and split tests as
// multiply_trees for every job, with the launches of all jobs queued back to back and
// one host round trip per refinement iteration (uploads through the staging ring, one
// read-back of all jobs' norms).  Each output is identical to its single product.
struct MulJob {
    TreeH *out, *a, *b;
    double prec, coef;
    std::vector<int> work;
    double sNorm = 0.0, wNorm = 0.0;
    bool first = true;
};

static void multiply_batch(std::vector<MulJob> &jobs) {
    if (jobs.empty()) return;
    MraH &M = *jobs[0].a->mra;
    Ctx &C = *M.ctx;
    hipStream_t s = C.stream;
    for (MulJob &J : jobs)
        if (J.a->n() == 0 || J.b->n() == 0) throw Error(MWG_ERR_STATE, "multiply: empty input tree");
    for_each_job(jobs.size(), [&](size_t j) {
        MulJob &J = jobs[j];
        TreeH &out = *J.out, &A = *J.a, &B = *J.b;
        out.make_roots();
        std::vector<int> order(out.roots.begin(), out.roots.end());
        for (size_t w = 0; w < order.size(); w++) {
            const int i = order[w];
            const int d = out.depth[i], x = out.lr[3 * i], y = out.lr[3 * i + 1], z = out.lr[3 * i + 2];
            const int ia = A.find(d, x, y, z), ib = B.find(d, x, y, z);
            const bool sp = (ia >= 0 && A.child[ia] >= 0) || (ib >= 0 && B.child[ib] >= 0);
            if (sp && d < M.maxDepth) {
                out.split(i);
                for (int a = 0; a < 8; a++) order.push_back(out.child[i] + a);
            }
        }
        J.work.resize(out.n());
        for (int i = 0; i < out.n(); i++) J.work[i] = i;
    });
    std::vector<MulItem> items;
    std::vector<int> wl;
    std::vector<size_t> ofs(jobs.size() + 1);
    std::vector<std::vector<MulItem>> jitems(jobs.size());
    while (true) {
        for (MulJob &J : jobs)
            if (!J.work.empty()) J.out->ensure_dev(J.out->n());
        for_each_job(jobs.size(), [&](size_t j) {
            MulJob &J = jobs[j];
            std::vector<MulItem> &v = jitems[j];
            v.resize(J.work.size());
            for (size_t w = 0; w < J.work.size(); w++) {
                const int i = J.work[w];
                MulItem &it = v[w];
                it.node = i;
                it.depth = J.out->depth[i];
                for (int q = 0; q < 3; q++) it.l[q] = J.out->lr[3 * i + q];
                locate(*J.a, it.depth, it.l, it.asrc, it.asteps);
                locate(*J.b, it.depth, it.l, it.bsrc, it.bsteps);
            }
        });
        items.clear();
        wl.clear();
        for (size_t j = 0; j < jobs.size(); j++) {
            ofs[j] = items.size();
            items.insert(items.end(), jitems[j].begin(), jitems[j].end());
            wl.insert(wl.end(), jobs[j].work.begin(), jobs[j].work.end());
        }
        ofs[jobs.size()] = items.size();
        if (items.empty()) break;
        upload_async(C, C.mitems, items.data(), items.size());
        upload_async(C, C.mwork, wl.data(), wl.size());
        size_t total = 0;
        for (size_t j = 0; j < jobs.size(); j++) {
            const int nj = (int)(ofs[j + 1] - ofs[j]);
            if (!nj) continue;
            MulJob &J = jobs[j];
            HC(launch_multiply(s, nj, C.mitems.p + ofs[j], M.geom(), M.dev(), J.coef, J.a->coef.p, J.b->coef.p,
                               J.out->coef.p));
            HC(launch_norms(s, nj, C.mwork.p + ofs[j], J.out->coef.p, J.out->norm.p, J.out->K3()));
            total += 8 * (size_t)J.out->n();
        }
        C.rb.reserve(total * sizeof(double));
        double *h = (double *)C.rb.h;
        size_t off = 0;
        for (size_t j = 0; j < jobs.size(); j++) {
            if (ofs[j + 1] == ofs[j]) continue;
            const size_t m = 8 * (size_t)jobs[j].out->n();
            HC(hipMemcpyAsync(h + off, jobs[j].out->norm.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
            off += m;
        }
        stream_wait(C);
        C.stage.off = 0;
        std::vector<size_t> hofs(jobs.size(), 0);
        off = 0;
        for (size_t j = 0; j < jobs.size(); j++) {
            if (ofs[j + 1] == ofs[j]) continue;
            hofs[j] = off;
            off += 8 * (size_t)jobs[j].out->n();
        }
        for_each_job(jobs.size(), [&](size_t j) {
            if (ofs[j + 1] == ofs[j]) return;
            MulJob &J = jobs[j];
            TreeH &out = *J.out;
            const double *nrm = h + hofs[j];
            for (int i : J.work) {
                std::memcpy(&out.cnorm[8 * (size_t)i], &nrm[8 * (size_t)i], 8 * sizeof(double));
                if (J.first && out.depth[i] == 0) J.sNorm += nrm[8 * i] * nrm[8 * i];
                for (int c = 1; c < 8; c++) J.wNorm += nrm[8 * i + c] * nrm[8 * i + c];
            }
            J.first = false;
            const double tnorm = std::sqrt(J.sNorm + J.wNorm);
            std::vector<int> next;
            for (int i : J.work) {
                if (out.child[i] >= 0) continue;
                const int d = out.depth[i];
                if (d >= M.maxDepth) continue;
                double wn = 0.0;
                for (int c = 1; c < 8; c++) wn += out.cnorm[8 * i + c] * out.cnorm[8 * i + c];
                if (split_check(M.n0 + d, std::sqrt(wn), J.prec, tnorm)) {
                    out.split(i);
                    for (int a = 0; a < 8; a++) next.push_back(out.child[i] + a);
                }
            }
            J.work.swap(next);
        });
    }
    // bottom-up compression, all jobs' parents of a depth in one launch
    std::vector<std::vector<std::vector<int>>> L(jobs.size());
    int maxd = 0;
    for (size_t j = 0; j < jobs.size(); j++) {
        L[j] = jobs[j].out->internal_by_depth();
        maxd = std::max(maxd, (int)L[j].size());
    }
    std::vector<TransformItem> titems;
    std::vector<size_t> lofs(maxd + 1, 0);
    for (int d = 0; d < maxd; d++) {
        lofs[d] = titems.size();
        for (size_t j = 0; j < jobs.size(); j++)
            if (d < (int)L[j].size())
                for (int p : L[j][d]) titems.push_back({jobs[j].out->coef.p, nullptr, p, jobs[j].out->child[p]});
    }
    lofs[maxd] = titems.size();
    if (!titems.empty()) {
        upload_async(C, C.titems, titems.data(), titems.size());
        for (int d = maxd - 1; d >= 0; d--)
            HC(launch_transform(s, 0, (int)(lofs[d + 1] - lofs[d]), C.titems.p + lofs[d], M.W.p, M.B.K));
    }
    // every node's norms (as refresh_norms), one read-back for all jobs
    size_t total = 0;
    for (MulJob &J : jobs) {
        HC(launch_norms(s, J.out->n(), nullptr, J.out->coef.p, J.out->norm.p, J.out->K3()));
        total += 8 * (size_t)J.out->n();
    }
    C.rb.reserve(total * sizeof(double));
    double *h = (double *)C.rb.h;
    size_t off = 0;
    for (MulJob &J : jobs) {
        const size_t m = 8 * (size_t)J.out->n();
        HC(hipMemcpyAsync(h + off, J.out->norm.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
        off += m;
    }
    stream_wait(C);
    C.stage.off = 0;
    off = 0;
    for (MulJob &J : jobs) {
        const size_t m = 8 * (size_t)J.out->n();
        J.out->cnorm.assign(h + off, h + off + m);
        off += m;
        J.out->sqnorm_from_cnorm();
        J.out->fit_dev();
    }
}

// out = ca * a + cb * b on the union grid (mrcpp::add)
static void add_trees(TreeH &out, double ca, TreeH &A, double cb, TreeH &B) {
    MraH &M = *A.mra;
    Ctx &C = *M.ctx;
    hipStream_t s = C.stream;
    if (A.n() == 0 || B.n() == 0) throw Error(MWG_ERR_STATE, "add: empty input tree");
    out.make_roots();
    std::vector<int> order(out.roots.begin(), out.roots.end());
    for (size_t w = 0; w < order.size(); w++) {
        const int i = order[w];
        const int d = out.depth[i], x = out.lr[3 * i], y = out.lr[3 * i + 1], z = out.lr[3 * i + 2];
        const int ia = A.find(d, x, y, z), ib = B.find(d, x, y, z);
        if ((ia >= 0 && A.child[ia] >= 0) || (ib >= 0 && B.child[ib] >= 0)) {
            out.split(i);
            for (int a = 0; a < 8; a++) order.push_back(out.child[i] + a);
        }
    }
    out.ensure_dev(out.n());
    std::vector<MulItem> items(out.n());
    for (int i = 0; i < out.n(); i++) {
        MulItem &it = items[i];
        it.node = i;
        it.depth = out.depth[i];
        for (int q = 0; q < 3; q++) it.l[q] = out.lr[3 * i + q];
        locate(A, it.depth, it.l, it.asrc, it.asteps);
        locate(B, it.depth, it.l, it.bsrc, it.bsteps);
    }
    DBuf<MulItem> ditems;
    ditems.upload(items.data(), items.size(), s);
    HC(launch_add(s, (int)items.size(), ditems.p, M.dev(), ca, A.coef.p, cb, B.coef.p, out.coef.p));
    out.refresh_norms();
}

}  // namespace mwg

// ================================================================== C ABI

using namespace mwg;

struct mwg_ctx_s : Ctx {};
struct mwg_mra_s : MraH {};
struct mwg_tree_s : TreeH {};
struct mwg_oper_s : OperH {};
struct mwg_predictor_s : PredH {};

// Handles may be destroyed in any order (a garbage collector finalises them in no
// particular order): a tree or operator holds a reference to its MRA, an MRA or a
// predictor one to its context, and each object is freed with its last reference.
static void release_ctx(Ctx *c) {
    if (!c || c->refs.fetch_sub(1) != 1) return;
    (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->evIter)
        if (e) (void)hipEventDestroy(e);
    if (c->evTopo) (void)hipEventDestroy(c->evTopo);
    if (c->liter_h) (void)hipHostFree(c->liter_h);
    hipStream_t s = c->stream;
    delete static_cast<mwg_ctx_s *>(c);
    if (s) (void)hipStreamDestroy(s);
}
static void release_mra(MraH *m) {
    if (!m || m->refs.fetch_sub(1) != 1) return;
    Ctx *c = m->ctx;
    delete static_cast<mwg_mra_s *>(m);
    release_ctx(c);
}

static thread_local std::string g_err;

template <class F>
static int guarded(Ctx *ctx, F f) {
    try {
        f();
        return MWG_OK;
    } catch (const Error &e) {
        (ctx ? ctx->err : g_err) = e.what();
        return e.code;
    } catch (const std::exception &e) {
        (ctx ? ctx->err : g_err) = e.what();
        return MWG_ERR_STATE;
    }
}

extern "C" {

// ------------------------------------------------------------------ allocation-free dot scratch
//
// mwg_tree_dot_noalloc stages its node pairs and per-block sums in one static device array
// per GPU: DOT_NOALLOC_MAX_PAIRS pair words, then DOT_NOALLOC_BLOCKS partial sums, then
// the total (8 MiB + 4 KiB).  Every call on a GPU shares that GPU's array.  Host code uses
// the address hipGetSymbolAddress returns for the current device, never the array's name.
constexpr int DOT_NOALLOC_MAX_PAIRS = 1 << 20;
constexpr int DOT_NOALLOC_BLOCKS = 512;
static __device__ __align__(256) unsigned long long scratch[DOT_NOALLOC_MAX_PAIRS + DOT_NOALLOC_BLOCKS + 1];

// the scratch address on `device` (current), looked up once per device
static void *dot_scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(MWG_ERR_HIP, "hipGetSymbolAddress returned no address for the dot scratch");
    addr[device] = p;
    return p;
}

int mwg_ctx_create(int device, mwg_ctx *out) {
    if (!out) return MWG_ERR_ARG;
    auto *c = new mwg_ctx_s();
    int rc = guarded(nullptr, [&] {
        c->device = device;
        HC(hipSetDevice(device));
        HC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        {
            int khz = 0;
            HC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
            if (khz <= 0) throw Error(MWG_ERR_HIP, "device wall clock rate unavailable");
            c->wall_khz = (double)khz;
        }
        c->dot_scratch = dot_scratch_address(device);
        double x[24], w[24];
        gauss_legendre01(24, x, w);
        c->gx.upload(x, 24, c->stream);
        c->gw.upload(w, 24, c->stream);
    });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return MWG_OK;
}

int mwg_ctx_destroy(mwg_ctx ctx) {
    release_ctx(ctx);
    return MWG_OK;
}

const char *mwg_last_error(mwg_ctx ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int mwg_sync(mwg_ctx ctx) {
    if (!ctx) return MWG_ERR_ARG;
    return guarded(ctx, [&] { HC(hipStreamSynchronize(ctx->stream)); });
}

int mwg_mra_create(mwg_ctx ctx, int root_scale, const int *corner, const int *boxes, double scaling_factor, int order,
                   int basis_type, int max_depth, mwg_mra *out) {
    if (!ctx || !corner || !boxes || !out) return MWG_ERR_ARG;
    auto *m = new mwg_mra_s();
    int rc = guarded(ctx, [&] {
        if (order < 3 || order > 9) throw Error(MWG_ERR_ARG, "order must be in [3, 9]");
        if (basis_type != MWG_BASIS_LEGENDRE && basis_type != MWG_BASIS_INTERPOLATING)
            throw Error(MWG_ERR_ARG, "basis type must be Legendre or Interpolating");
        if (!(scaling_factor > 0.0)) throw Error(MWG_ERR_ARG, "scaling factor must be positive");
        for (int d = 0; d < 3; d++)
            if (boxes[d] < 1) throw Error(MWG_ERR_ARG, "boxes must be >= 1");
        int bmax = std::max({boxes[0], boxes[1], boxes[2]});
        if (max_depth < 1 || max_depth > 30 || ((int64_t)bmax << max_depth) > (1 << 19))
            throw Error(MWG_ERR_RANGE, "max depth too large for 19-bit translations");
        m->ctx = ctx;
        m->n0 = root_scale;
        for (int d = 0; d < 3; d++) {
            m->corner[d] = corner[d];
            m->boxes[d] = boxes[d];
        }
        m->sfac = scaling_factor;
        m->maxDepth = max_depth;
        build_basis(m->B, order, basis_type);
        hipStream_t s = ctx->stream;
        m->U.upload(m->B.U.data(), m->B.U.size(), s);
        m->qx.upload(m->B.qx.data(), m->B.qx.size(), s);
        m->qw.upload(m->B.qw.data(), m->B.qw.size(), s);
        m->phiq.upload(m->B.phiq.data(), m->B.phiq.size(), s);
        m->W.upload(m->B.W.data(), m->B.W.size(), s);
        const int N = 2 * m->B.K;
        std::vector<double> WT(N * N);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) WT[i * N + j] = m->B.W[j * N + i];
        m->WT.upload(WT.data(), WT.size(), s);
        m->ccoef.upload(m->B.ccoef.data(), m->B.ccoef.size(), s);
    });
    if (rc) {
        delete m;
        return rc;
    }
    ctx->refs++;
    *out = m;
    return MWG_OK;
}

int mwg_mra_destroy(mwg_mra mra) {
    release_mra(mra);
    return MWG_OK;
}

int mwg_tree_create(mwg_mra mra, mwg_tree *out) {
    if (!mra || !out) return MWG_ERR_ARG;
    auto *t = new mwg_tree_s();
    t->mra = mra;
    mra->refs++;
    *out = t;
    return MWG_OK;
}
int mwg_tree_destroy(mwg_tree tree) {
    MraH *m = tree ? tree->mra : nullptr;
    delete tree;
    release_mra(m);
    return MWG_OK;
}
int mwg_tree_nnodes(mwg_tree tree, int64_t *n) {
    if (!tree || !n) return MWG_ERR_ARG;
    *n = tree->n();
    return MWG_OK;
}
int mwg_tree_sqnorm(mwg_tree tree, double *out) {
    if (!tree || !out) return MWG_ERR_ARG;
    *out = tree->sqnorm;
    return MWG_OK;
}
int mwg_tree_maxdepth(mwg_tree tree, int *out) {
    if (!tree || !out) return MWG_ERR_ARG;
    int m = 0;
    for (int d : tree->depth) m = std::max(m, d);
    *out = m;
    return MWG_OK;
}

int mwg_tree_mra(mwg_tree tree, mwg_mra *out) {
    if (!tree || !out) return MWG_ERR_ARG;
    *out = static_cast<mwg_mra>(tree->mra);  // every MRA is created as an mwg_mra_s
    return MWG_OK;
}

int mwg_mra_order(mwg_mra mra, int *order) {
    if (!mra || !order) return MWG_ERR_ARG;
    *order = mra->B.k;
    return MWG_OK;
}

int mwg_tree_import(mwg_tree tree, int64_t n, const int *depth, const int64_t *lr, const double *coefs) {
    if (!tree || n < 0 || (n > 0 && (!depth || !lr || !coefs))) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    return guarded(C, [&] {
        TreeH &T = *tree;
        T.clear();
        const MraH &M = *T.mra;
        const std::string bad = check_tree_layout(M.boxes, M.maxDepth, n, depth, lr);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, "tree import: " + bad);
        for (int64_t i = 0; i < n; i++) {
            const int d = depth[i], x = (int)lr[3 * i], y = (int)lr[3 * i + 1], z = (int)lr[3 * i + 2];
            const int par = d > 0 ? T.find(d - 1, x >> 1, y >> 1, z >> 1) : -1;
            const int id = T.add(d, x, y, z, par);
            if (par < 0) T.roots.push_back(id);
            else if (((x & 1) | (y & 1) | (z & 1)) == 0) T.child[par] = id;
        }
        T.ensure_dev(std::max<int64_t>(n, 1));
        HC(hipMemcpyAsync(T.coef.p, coefs, (size_t)n * 8 * T.K3() * sizeof(double), hipMemcpyHostToDevice,
                          C->stream));
        T.refresh_norms();
    });
}

int mwg_tree_layout_check(const int *boxes, int max_depth, int64_t n, const int *depth, const int64_t *lr,
                          char *msg, int msg_len) {
    if (!boxes || n < 0 || (n > 0 && (!depth || !lr))) return MWG_ERR_ARG;
    for (int q = 0; q < 3; q++)
        if (boxes[q] < 1) return MWG_ERR_ARG;
    const std::string bad = check_tree_layout(boxes, max_depth, n, depth, lr);
    if (msg && msg_len > 0) {
        std::strncpy(msg, bad.c_str(), (size_t)msg_len - 1);
        msg[msg_len - 1] = 0;
    }
    return bad.empty() ? MWG_OK : MWG_ERR_ARG;
}

int mwg_tree_export(mwg_tree tree, int *depth, int64_t *lr, double *coefs) {
    if (!tree) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    return guarded(C, [&] {
        TreeH &T = *tree;
        for (int i = 0; i < T.n(); i++) {
            if (depth) depth[i] = T.depth[i];
            if (lr)
                for (int q = 0; q < 3; q++) lr[3 * i + q] = T.lr[3 * i + q];
        }
        if (coefs) T.coef.download(coefs, (size_t)T.n() * 8 * T.K3(), C->stream);
    });
}

// ------------------------------------------------------------------ orbital files

static OrbitalMeta meta_in(const mwg_orbital_meta &m) {
    OrbitalMeta o;
    o.real_size = m.real_size;
    o.imag_size = m.imag_size;
    o.is_shared = m.is_shared;
    o.rank_id = m.rank_id;
    o.spin = m.spin;
    o.occ = m.occ;
    o.error = m.error;
    return o;
}

int mwg_meta_write(const char *path, const mwg_orbital_meta *meta, int orbital) {
    if (!path || !meta) return MWG_ERR_ARG;
    return guarded(nullptr, [&] {
        const std::string bad = write_meta(path, meta_in(*meta), orbital != 0);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
    });
}

int mwg_meta_read(const char *path, mwg_orbital_meta *meta, int *orbital) {
    if (!path || !meta) return MWG_ERR_ARG;
    return guarded(nullptr, [&] {
        OrbitalMeta o;
        bool orb = false;
        const std::string bad = read_meta(path, o, orb);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
        *meta = mwg_orbital_meta{o.real_size, o.imag_size, o.is_shared, o.rank_id, o.spin, o.occ, o.error};
        if (orbital) *orbital = orb ? 1 : 0;
    });
}

static TreeFileInfo info_in(const mwg_tree_file_info &i) {
    TreeFileInfo t;
    t.order = i.order;
    t.basis = i.basis;
    t.root_scale = i.root_scale;
    t.max_depth = i.max_depth;
    for (int d = 0; d < 3; d++) {
        t.corner[d] = i.corner[d];
        t.boxes[d] = i.boxes[d];
    }
    t.scaling_factor = i.scaling_factor;
    t.nnodes = i.nnodes;
    return t;
}

static mwg_tree_file_info info_out(const TreeFileInfo &t) {
    mwg_tree_file_info i{};
    i.order = t.order;
    i.basis = t.basis;
    i.root_scale = t.root_scale;
    i.max_depth = t.max_depth;
    for (int d = 0; d < 3; d++) {
        i.corner[d] = t.corner[d];
        i.boxes[d] = t.boxes[d];
    }
    i.scaling_factor = t.scaling_factor;
    i.nnodes = t.nnodes;
    i.chunk_nodes = t.chunk_nodes;
    i.nchunks = t.nchunks;
    return i;
}

int mwg_tree_file_write(const char *name, const mwg_tree_file_info *info, const int *depth, const int64_t *lr,
                        const double *coefs) {
    if (!name || !info || info->nnodes < 0 || (info->nnodes > 0 && (!depth || !lr || !coefs))) return MWG_ERR_ARG;
    return guarded(nullptr, [&] {
        const TreeFileInfo t = info_in(*info);
        for (int q = 0; q < 3; q++)
            if (t.boxes[q] < 1) throw Error(MWG_ERR_ARG, "tree file: boxes must be positive");
        const std::string lay = check_tree_layout(t.boxes, t.max_depth, t.nnodes, depth, lr);
        if (!lay.empty()) throw Error(MWG_ERR_ARG, "tree file: " + lay);
        const std::string bad = write_tree_file(std::string(name) + ".tree", t, depth, lr, coefs);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
    });
}

int mwg_tree_file_read(const char *name, mwg_tree_file_info *info, int *depth, int64_t *lr, double *coefs) {
    if (!name || !info) return MWG_ERR_ARG;
    return guarded(nullptr, [&] {
        TreeFileInfo t;
        const std::string bad = read_tree_file(std::string(name) + ".tree", t, depth, lr, coefs);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
        *info = info_out(t);
    });
}

int mwg_tree_save(mwg_tree tree, const char *name, int *nchunks) {
    if (!tree || !name) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    return guarded(C, [&] {
        const TreeH &T = *tree;
        const MraH &M = *T.mra;
        TreeFileInfo t;
        t.order = M.B.k;
        t.basis = M.B.type;
        t.root_scale = M.n0;
        t.max_depth = M.maxDepth;
        for (int d = 0; d < 3; d++) {
            t.corner[d] = M.corner[d];
            t.boxes[d] = M.boxes[d];
        }
        t.scaling_factor = M.sfac;
        t.nnodes = T.n();
        std::vector<int64_t> lr(T.lr.begin(), T.lr.end());
        std::vector<double> coefs((size_t)T.n() * 8 * T.K3());
        if (T.n()) T.coef.download(coefs.data(), coefs.size(), C->stream);
        const std::string bad = write_tree_file(std::string(name) + ".tree", t, T.depth.data(), lr.data(),
                                                coefs.data());
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
        if (nchunks) *nchunks = (int)((t.nnodes + TREE_FILE_CHUNK - 1) / TREE_FILE_CHUNK);
    });
}

int mwg_tree_load(mwg_tree tree, const char *name) {
    if (!tree || !name) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    std::vector<int> depth;
    std::vector<int64_t> lr;
    std::vector<double> coefs;
    const int rc = guarded(C, [&] {
        const MraH &M = *tree->mra;
        TreeFileInfo t;
        const std::string path = std::string(name) + ".tree";
        std::string bad = read_tree_file(path, t, nullptr, nullptr, nullptr);
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
        bool same = t.order == M.B.k && t.basis == M.B.type && t.root_scale == M.n0 && t.max_depth == M.maxDepth &&
                    t.scaling_factor == M.sfac;
        for (int d = 0; d < 3; d++) same = same && t.corner[d] == M.corner[d] && t.boxes[d] == M.boxes[d];
        if (!same) throw Error(MWG_ERR_ARG, path + ": written for another MRA (order, basis, box or depth)");
        depth.resize((size_t)t.nnodes);
        lr.resize(3 * (size_t)t.nnodes);
        coefs.resize((size_t)t.nnodes * 8 * M.B.K3);
        bad = read_tree_file(path, t, depth.data(), lr.data(), coefs.data());
        if (!bad.empty()) throw Error(MWG_ERR_ARG, bad);
    });
    if (rc) return rc;
    return mwg_tree_import(tree, (int64_t)depth.size(), depth.data(), lr.data(), coefs.data());
}

int mwg_tree_rescale(mwg_tree tree, double c) {
    if (!tree) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    return guarded(C, [&] {
        TreeH &T = *tree;
        T.ver++;
        HC(launch_scale(C->stream, (size_t)T.n() * 8 * T.K3(), T.coef.p, c));
        HC(launch_scale(C->stream, (size_t)T.n() * 8, T.norm.p, std::fabs(c)));
        for (double &v : T.cnorm) v *= std::fabs(c);
        T.sqnorm *= c * c;
        HC(hipStreamSynchronize(C->stream));
    });
}

int mwg_tree_dot(mwg_tree a, mwg_tree b, double *out) {
    if (!a || !b || !out) return MWG_ERR_ARG;
    if (a->mra != b->mra) return MWG_ERR_ARG;
    Ctx *C = a->mra->ctx;
    return guarded(C, [&] {
        if (a->roots.empty() || b->roots.empty()) {  // an empty tree is the zero function
            *out = 0.0;
            return;
        }
        std::vector<int> ia, ib, c0;
        for (size_t r = 0; r < a->roots.size(); r++) {
            ia.push_back(a->roots[r]);
            ib.push_back(b->roots[r]);
            c0.push_back(0);
        }
        for (int i = 0; i < a->n(); i++) {
            if (a->depth[i] == 0) continue;
            const int j = b->find(a->depth[i], a->lr[3 * i], a->lr[3 * i + 1], a->lr[3 * i + 2]);
            if (j < 0) continue;
            ia.push_back(i);
            ib.push_back(j);
            c0.push_back(1);
        }
        hipStream_t s = C->stream;
        DBuf<int> dia, dib, dc0;
        DBuf<double> part;
        dia.upload(ia.data(), ia.size(), s);
        dib.upload(ib.data(), ib.size(), s);
        dc0.upload(c0.data(), c0.size(), s);
        part.alloc(ia.size());
        HC(launch_dot(s, (int)ia.size(), dia.p, dib.p, dc0.p, a->coef.p, b->coef.p, a->K3(), part.p));
        std::vector<double> hp(ia.size());
        part.download(hp.data(), hp.size(), s);
        double sum = 0.0;
        for (double v : hp) sum += v;
        *out = sum;
    });
}

int mwg_tree_dot_noalloc(mwg_tree a, mwg_tree b, double *out) {
    if (!a || !b || !out) return MWG_ERR_ARG;
    if (a->mra != b->mra) return MWG_ERR_ARG;
    Ctx *C = a->mra->ctx;
    return guarded(C, [&] {
        if (!C->dot_scratch) throw Error(MWG_ERR_STATE, "dot scratch not set up on this device");
        unsigned long long *pairs = (unsigned long long *)C->dot_scratch;
        double *part = (double *)(pairs + DOT_NOALLOC_MAX_PAIRS);
        double *res = part + DOT_NOALLOC_BLOCKS;
        // the same pairs as mwg_tree_dot: root with root (all 8 components), then every
        // deeper node of a that b holds too (wavelet components only); counted first so an
        // oversized call leaves the scratch untouched
        const TreeH &A = *a, &B = *b;
        if (A.roots.empty() || B.roots.empty()) {  // an empty tree is the zero function
            *out = 0.0;
            return;
        }
        auto match = [&](int i) {
            return B.find(A.depth[i], A.lr[3 * i], A.lr[3 * i + 1], A.lr[3 * i + 2]);
        };
        int64_t np = (int64_t)A.roots.size();
        for (int i = 0; i < A.n(); i++)
            if (A.depth[i] > 0 && match(i) >= 0) np++;
        if (np > DOT_NOALLOC_MAX_PAIRS)
            throw Error(MWG_ERR_RANGE, "mwg_tree_dot_noalloc: " + std::to_string(np) +
                                           " node pairs exceed the scratch limit of " +
                                           std::to_string(DOT_NOALLOC_MAX_PAIRS));
        hipStream_t s = C->stream;
        // staged through a host stack buffer, copied (and waited for) whenever it fills
        constexpr int STAGE = 4096;
        unsigned long long buf[STAGE];
        int nb = 0;
        int64_t off = 0;
        auto flush = [&]() {
            if (!nb) return;
            HC(hipMemcpyAsync(pairs + off, buf, (size_t)nb * sizeof(unsigned long long), hipMemcpyHostToDevice, s));
            HC(hipStreamSynchronize(s));
            off += nb;
            nb = 0;
        };
        auto put = [&](int ia, int ib, int skip0) {
            buf[nb++] = (unsigned long long)(unsigned)ia | ((unsigned long long)skip0 << 31) |
                        ((unsigned long long)(unsigned)ib << 32);
            if (nb == STAGE) flush();
        };
        for (size_t r = 0; r < A.roots.size(); r++) put(A.roots[r], B.roots[r], 0);
        for (int i = 0; i < A.n(); i++) {
            if (A.depth[i] == 0) continue;
            const int j = match(i);
            if (j >= 0) put(i, j, 1);
        }
        flush();
        HC(launch_dot_pairs(s, (int)np, pairs, A.coef.p, B.coef.p, A.K3(), DOT_NOALLOC_BLOCKS, part, res));
        double h = 0.0;
        HC(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        *out = h;
    });
}

int mwg_tree_eval(mwg_tree tree, int64_t n, const double *xyz, double *out) {
    if (!tree || n < 0 || (n > 0 && (!xyz || !out))) return MWG_ERR_ARG;
    Ctx *C = tree->mra->ctx;
    return guarded(C, [&] {
        TreeH &T = *tree;
        const MraH &M = *T.mra;
        const Basis &B = M.B;
        const int K = B.K, K3 = B.K3, N = 2 * K;
        std::vector<double> host((size_t)T.n() * 8 * K3);
        T.coef.download(host.data(), host.size(), C->stream);
        std::vector<double> WT(N * N);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) WT[i * N + j] = B.W[j * N + i];
        for (int64_t pi = 0; pi < n; pi++) {
            const double *p = &xyz[3 * pi];
            double u0[3];
            bool inside = true;
            for (int d = 0; d < 3; d++) {
                u0[d] = (p[d] - M.lo(d)) / (M.sfac * std::ldexp(1.0, -M.n0));
                if (u0[d] < 0.0 || u0[d] >= M.boxes[d]) inside = false;
            }
            if (!inside) {
                out[pi] = 0.0;
                continue;
            }
            int node = T.find(0, (int)u0[0], (int)u0[1], (int)u0[2]);
            int depth = 0;
            while (node >= 0 && T.child[node] >= 0) {
                depth++;
                int a = 0;
                for (int d = 0; d < 3; d++) a |= (int)(((int64_t)std::ldexp(u0[d], depth)) & 1) << d;
                node = T.child[node] + a;
            }
            if (node < 0) {
                out[pi] = 0.0;
                continue;
            }
            // reconstruct the child of the leaf that holds the point, then evaluate
            const int cd = depth + 1;
            int a = 0;
            double uc[3];
            for (int d = 0; d < 3; d++) {
                const double u = std::ldexp(u0[d], cd);
                const int64_t li = (int64_t)u;
                uc[d] = u - (double)li;
                a |= (int)(li & 1) << d;
            }
            const double *cf = &host[(size_t)node * 8 * K3];
            double px[MAXK], py[MAXK], pz[MAXK];
            B.phi(uc[0], px);
            B.phi(uc[1], py);
            B.phi(uc[2], pz);
            // child a scaling: s[z][y][x] = sum_c sum_{i,j,l} W[c_x K+i][a_x K+x] ... comp_c
            double acc = 0.0;
            for (int c = 0; c < 8; c++) {
                const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
                double vx[MAXK], vy[MAXK], vz[MAXK];
                for (int i = 0; i < K; i++) {
                    double sx = 0, sy = 0, sz = 0;
                    for (int x = 0; x < K; x++) {
                        sx += B.W[(bx * K + i) * N + (a & 1) * K + x] * px[x];
                        sy += B.W[(by * K + i) * N + ((a >> 1) & 1) * K + x] * py[x];
                        sz += B.W[(bz * K + i) * N + ((a >> 2) & 1) * K + x] * pz[x];
                    }
                    vx[i] = sx;
                    vy[i] = sy;
                    vz[i] = sz;
                }
                const double *cc = cf + c * K3;
                for (int z = 0; z < K; z++)
                    for (int y = 0; y < K; y++) {
                        const double yz = vz[z] * vy[y];
                        for (int x = 0; x < K; x++) acc += cc[(z * K + y) * K + x] * yz * vx[x];
                    }
            }
            const double h = M.sfac * std::ldexp(1.0, -(M.n0 + cd));
            out[pi] = acc / std::pow(h, 1.5);
        }
        (void)WT;
    });
}

int mwg_project_gauss(double prec, mwg_tree out, int n, const double *coef, const double *alpha,
                      const double *center, const double *poly) {
    if (!out || n < 0 || (n > 0 && (!coef || !alpha || !center || !poly))) return MWG_ERR_ARG;
    Ctx *C = out->mra->ctx;
    return guarded(C, [&] {
        hipStream_t s = C->stream;
        DBuf<double> dc, da, dce, dp;
        dc.upload(coef, n, s);
        da.upload(alpha, n, s);
        dce.upload(center, 3 * (size_t)n, s);
        dp.upload(poly, 15 * (size_t)n, s);
        std::vector<double> va(alpha, alpha + n), vc(center, center + 3 * (size_t)n);
        MraH &M = *out->mra;
        project_adaptive(
            *out, prec,
            [&](int nw, const int *work, const int *depth, const int *lr, double *dst) {
                HC(launch_project_gauss(s, nw, work, depth, lr, M.geom(), M.dev(), n, dc.p, da.p, dce.p, dp.p, dst));
            },
            &va, &vc);
    });
}

int mwg_project_hydrogen(double prec, mwg_tree out, int n, const double *coef, const int *q, double Z,
                         const double *origin) {
    if (!out || n < 0 || (n > 0 && (!coef || !q || !origin))) return MWG_ERR_ARG;
    Ctx *C = out->mra->ctx;
    return guarded(C, [&] {
        hipStream_t s = C->stream;
        DBuf<double> dc, dorig;
        DBuf<int> dq;
        dc.upload(coef, n, s);
        dq.upload(q, 6 * (size_t)n, s);
        dorig.upload(origin, 3, s);
        MraH &M = *out->mra;
        project_adaptive(
            *out, prec,
            [&](int nw, const int *work, const int *depth, const int *lr, double *dst) {
                HC(launch_project_hyd(s, nw, work, depth, lr, M.geom(), M.dev(), n, dc.p, dq.p, Z, dorig.p, dst));
            },
            nullptr, nullptr);
    });
}

static int make_oper(mwg_mra mra, double mu, double prec, mwg_oper *out) {
    if (!mra || !out) return MWG_ERR_ARG;
    auto *o = new mwg_oper_s();
    int rc = guarded(mra->ctx, [&] {
        if (!(prec > 0.0)) throw Error(MWG_ERR_ARG, "operator precision must be positive");
        if (mu < 0.0) throw Error(MWG_ERR_ARG, "Helmholtz mu must be non-negative");
        build_oper(*o, mra, mu, prec);
    });
    if (rc) {
        delete o;
        return rc;
    }
    mra->refs++;
    *out = o;
    return MWG_OK;
}

int mwg_oper_poisson(mwg_mra mra, double prec, mwg_oper *out) { return make_oper(mra, 0.0, prec, out); }
int mwg_oper_helmholtz(mwg_mra mra, double mu, double prec, mwg_oper *out) {
    if (!(mu > 0.0)) {
        if (mra) mra->ctx->err = "Helmholtz mu must be positive (mu = sqrt(-2 epsilon), epsilon < 0)";
        return MWG_ERR_ARG;
    }
    return make_oper(mra, mu, prec, out);
}
int mwg_oper_destroy(mwg_oper oper) {
    MraH *m = oper ? oper->mra : nullptr;
    delete oper;
    release_mra(m);
    return MWG_OK;
}
int mwg_oper_nterms(mwg_oper oper, int *n) {
    if (!oper || !n) return MWG_ERR_ARG;
    *n = oper->nterms();
    return MWG_OK;
}
int mwg_oper_terms(mwg_oper oper, double *exponents, double *weights) {
    if (!oper) return MWG_ERR_ARG;
    for (int t = 0; t < oper->nterms(); t++) {
        if (exponents) exponents[t] = oper->fit.a[t];
        if (weights) weights[t] = oper->fit.w[t];
    }
    return MWG_OK;
}
int mwg_oper_blocks(mwg_oper oper, int term, int depth, int64_t dlo, int64_t dhi, double *out) {
    if (!oper || !out || term < 0 || term >= oper->nterms() || depth < 0 || depth > oper->mra->maxDepth || dhi < dlo)
        return MWG_ERR_ARG;
    Ctx *C = oper->mra->ctx;
    return guarded(C, [&] {
        MraH &M = *oper->mra;
        DBuf<double> b, nrm;
        b.alloc((size_t)(dhi - dlo + 1) * 4 * M.B.K2);
        nrm.alloc((size_t)(dhi - dlo + 1) * 4);
        NSBatch X;
        X.d = depth;
        X.b1 = {oper->beta[term] * std::ldexp(1.0, -2 * (M.n0 + depth + 1))};
        X.req = {NSRequest{0, dlo, dhi}};
        X.groups = {NSGroup{0, b.p, nullptr}};
        X.nrm = nrm.p;
        compute_ns(M, {&X});
        b.download(out, (size_t)(dhi - dlo + 1) * 4 * M.B.K2, C->stream);
    });
}

int mwg_multiply(double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b) {
    if (!out || !a || !b) return MWG_ERR_ARG;
    if (a->mra != b->mra || out->mra != a->mra) return MWG_ERR_ARG;
    Ctx *C = a->mra->ctx;
    return guarded(C, [&] {
        if (out == a || out == b) throw Error(MWG_ERR_ARG, "multiply: output must differ from the inputs");
        if (!(prec > 0.0)) throw Error(MWG_ERR_ARG, "multiply: precision must be positive");
        multiply_trees(*out, prec, coef, *a, *b);
    });
}

int mwg_multiply_batch(int n, const double *prec, mwg_tree *out, const double *coef, mwg_tree *a, mwg_tree *b) {
    if (n <= 0 || !prec || !out || !coef || !a || !b) return MWG_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!out[i] || !a[i] || !b[i]) return MWG_ERR_ARG;
    MraH *M = a[0]->mra;
    for (int i = 0; i < n; i++)
        if (a[i]->mra != M || b[i]->mra != M || out[i]->mra != M) return MWG_ERR_ARG;
    Ctx *C = M->ctx;
    return guarded(C, [&] {
        std::vector<MulJob> jobs(n);
        for (int i = 0; i < n; i++) {
            if (!(prec[i] > 0.0)) throw Error(MWG_ERR_ARG, "multiply: precision must be positive");
            for (int j = 0; j < n; j++)
                if (out[i] == a[j] || out[i] == b[j] || (j != i && out[i] == out[j]))
                    throw Error(MWG_ERR_ARG, "multiply: outputs must differ from the inputs and each other");
            jobs[i].out = out[i];
            jobs[i].a = a[i];
            jobs[i].b = b[i];
            jobs[i].prec = prec[i];
            jobs[i].coef = coef[i];
        }
        multiply_batch(jobs);
    });
}

int mwg_add(mwg_tree out, double ca, mwg_tree a, double cb, mwg_tree b) {
    if (!out || !a || !b) return MWG_ERR_ARG;
    if (a->mra != b->mra || out->mra != a->mra) return MWG_ERR_ARG;
    Ctx *C = a->mra->ctx;
    return guarded(C, [&] {
        if (out == a || out == b) throw Error(MWG_ERR_ARG, "add: output must differ from the inputs");
        add_trees(*out, ca, *a, cb, *b);
    });
}

int mwg_apply_batch(int n, const double *prec, mwg_tree *out, mwg_oper *oper, mwg_tree *inp, int max_iter,
                    unsigned flags, mwg_apply_stats *stats) {
    if (n <= 0 || !prec || !out || !oper || !inp) return MWG_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!out[i] || !oper[i] || !inp[i]) return MWG_ERR_ARG;
    Ctx *C = inp[0]->mra->ctx;
    return guarded(C, [&] {
        if (flags & ~(MWG_APPLY_MIXED | MWG_APPLY_PREDICT)) throw Error(MWG_ERR_ARG, "unknown apply flags");
        if ((flags & MWG_APPLY_PREDICT) && !C->pred)
            throw Error(MWG_ERR_ARG, "MWG_APPLY_PREDICT needs a predictor installed with mwg_ctx_set_predictor");
        for (int i = 0; i < n; i++) {
            if (!(prec[i] > 0.0)) throw Error(MWG_ERR_ARG, "apply precision must be positive");
            for (int j = 0; j < n; j++)
                if (out[i] == inp[j] || (j != i && out[i] == out[j]))
                    throw Error(MWG_ERR_ARG, "output trees must be distinct from inputs and each other");
        }
        std::vector<mwg_apply_stats> st;
        // The generated-node tables are sized from the inputs (4 entries per input node); an
        // output that refines far below a coarse input can need more: run the batch again
        // from fresh output roots with 8x, 64x and 512x larger tables; later batches on the
        // context start at the size that sufficed.
        for (int gen_shift = C->gen_shift;; gen_shift += 3) {
            std::vector<JobH> jobs(n);
            for (int i = 0; i < n; i++) {
                jobs[i].prec = prec[i];
                jobs[i].out = out[i];
                jobs[i].op = oper[i];
                jobs[i].in = inp[i];
            }
            st.clear();
            try {
                apply_jobs(jobs, max_iter, flags, st, gen_shift);
                C->gen_shift = gen_shift;
                break;
            } catch (const GenOverflow &) {
                if (gen_shift >= 9) throw;
            }
        }
        if (stats)
            for (int i = 0; i < n; i++) stats[i] = st[i];
    });
}

int mwg_apply(double prec, mwg_tree out, mwg_oper oper, mwg_tree inp, int max_iter, unsigned flags,
              mwg_apply_stats *stats) {
    return mwg_apply_batch(1, &prec, &out, &oper, &inp, max_iter, flags, stats);
}


static uint64_t splitmix(uint64_t &st) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float urand(uint64_t &st) { return (float)((splitmix(st) >> 40) * (1.0 / 16777216.0)); }

int mwg_predictor_sizes(int *nparams, int *nfeatures) {
    if (nparams) *nparams = MLP_PARAMS;
    if (nfeatures) *nfeatures = MLP_FEATURES;
    return MWG_OK;
}

int mwg_predictor_create(mwg_ctx ctx, uint64_t seed, mwg_predictor *out) {
    if (!ctx || !out) return MWG_ERR_ARG;
    auto *p = new mwg_predictor_s();
    int rc = guarded(ctx, [&] {
        p->ctx = ctx;
        std::vector<float> P(MLP_PARAMS, 0.0f);
        uint64_t st = seed;
        const int H = MLP_HIDDEN, NF = MLP_FEATURES;
        const float a1 = std::sqrt(6.0f / (NF + H)), a2 = std::sqrt(6.0f / (2 * H)), a3 = std::sqrt(6.0f / (H + 1));
        size_t o = 0;
        for (int i = 0; i < H * NF; i++) P[o++] = a1 * (2.0f * urand(st) - 1.0f);
        o += H;
        for (int i = 0; i < H * H; i++) P[o++] = a2 * (2.0f * urand(st) - 1.0f);
        o += H;
        for (int i = 0; i < H; i++) P[o++] = a3 * (2.0f * urand(st) - 1.0f);
        std::vector<float> mean(NF, 0.0f), sdev(NF, 1.0f), zero(MLP_PARAMS, 0.0f);
        p->P.upload(P.data(), P.size(), ctx->stream);
        p->m.upload(zero.data(), zero.size(), ctx->stream);
        p->v.upload(zero.data(), zero.size(), ctx->stream);
        p->G.alloc(MLP_PARAMS);
        p->mean.upload(mean.data(), NF, ctx->stream);
        p->sdev.upload(sdev.data(), NF, ctx->stream);
    });
    if (rc) {
        delete p;
        return rc;
    }
    ctx->refs++;
    *out = p;
    return MWG_OK;
}

int mwg_predictor_destroy(mwg_predictor pred) {
    Ctx *c = pred ? pred->ctx : nullptr;
    if (c && c->pred == pred) c->pred = nullptr;
    delete pred;
    release_ctx(c);
    return MWG_OK;
}

int mwg_predictor_get(mwg_predictor pred, float *params, float *mean, float *sdev) {
    if (!pred) return MWG_ERR_ARG;
    return guarded(pred->ctx, [&] {
        if (params) pred->P.download(params, MLP_PARAMS, pred->ctx->stream);
        if (mean) pred->mean.download(mean, MLP_FEATURES, pred->ctx->stream);
        if (sdev) pred->sdev.download(sdev, MLP_FEATURES, pred->ctx->stream);
    });
}

int mwg_predictor_set(mwg_predictor pred, const float *params, const float *mean, const float *sdev) {
    if (!pred) return MWG_ERR_ARG;
    return guarded(pred->ctx, [&] {
        if (params) pred->P.upload(params, MLP_PARAMS, pred->ctx->stream);
        if (mean) pred->mean.upload(mean, MLP_FEATURES, pred->ctx->stream);
        if (sdev) pred->sdev.upload(sdev, MLP_FEATURES, pred->ctx->stream);
    });
}

int mwg_predictor_gradient(mwg_predictor pred, int64_t n, const float *X, const float *y, float *grad) {
    if (!pred || n <= 0 || !X || !y || !grad || n > (1 << 26)) return MWG_ERR_ARG;
    return guarded(pred->ctx, [&] {
        hipStream_t s = pred->ctx->stream;
        pred->X.upload(X, (size_t)n * MLP_FEATURES, s);
        pred->Y.upload(y, (size_t)n, s);
        predictor_grad(*pred, (int)n, nullptr);
        pred->G.download(grad, MLP_PARAMS, s);
    });
}

int mwg_predictor_infer(mwg_predictor pred, int64_t n, const float *X, float *prob) {
    if (!pred || n < 0 || (n > 0 && (!X || !prob)) || n > (1 << 26)) return MWG_ERR_ARG;
    return guarded(pred->ctx, [&] {
        if (n) predictor_infer(*pred, (int)n, X, prob);
    });
}

int mwg_predictor_train(mwg_predictor pred, int64_t n, const float *X, const float *y, int epochs, int batch,
                        float lr, uint64_t seed, double *loss_before, double *loss_after) {
    if (!pred || n <= 0 || !X || !y || epochs < 0 || batch <= 0 || n > (1 << 26)) return MWG_ERR_ARG;
    return guarded(pred->ctx, [&] {
        hipStream_t s = pred->ctx->stream;
        const int NF = MLP_FEATURES;
        std::vector<float> mean(NF, 0.0f), sdev(NF, 0.0f);
        for (int64_t i = 0; i < n; i++)
            for (int f = 0; f < NF; f++) mean[f] += X[i * NF + f];
        for (int f = 0; f < NF; f++) mean[f] /= (float)n;
        for (int64_t i = 0; i < n; i++)
            for (int f = 0; f < NF; f++) sdev[f] += (X[i * NF + f] - mean[f]) * (X[i * NF + f] - mean[f]);
        // floor: a feature (nearly) constant in the training set must not blow up on new data
        for (int f = 0; f < NF; f++) sdev[f] = std::max(std::sqrt(sdev[f] / (float)n), 0.1f);
        pred->mean.upload(mean.data(), NF, s);
        pred->sdev.upload(sdev.data(), NF, s);
        pred->X.upload(X, (size_t)n * NF, s);
        pred->Y.upload(y, (size_t)n, s);
        if (loss_before) *loss_before = predictor_loss(*pred, (int)n, y);
        std::vector<int> order(n);
        for (int64_t i = 0; i < n; i++) order[i] = (int)i;
        uint64_t st = seed;
        for (int e = 0; e < epochs; e++) {
            for (int64_t i = n - 1; i > 0; i--) std::swap(order[i], order[splitmix(st) % (uint64_t)(i + 1)]);
            pred->idx.upload(order.data(), order.size(), s);
            for (int64_t b0 = 0; b0 < n; b0 += batch) {
                const int nb = (int)std::min<int64_t>(batch, n - b0);
                predictor_grad(*pred, nb, pred->idx.p + b0);
                HC(launch_adam(s, pred->P.p, pred->G.p, pred->m.p, pred->v.p, lr, 0.9f, 0.999f, 1e-8f, ++pred->adam_t));
            }
        }
        HC(hipStreamSynchronize(s));
        if (loss_after) *loss_after = predictor_loss(*pred, (int)n, y);
    });
}

int mwg_ctx_set_tuning(mwg_ctx ctx, int fp64_chunk, int slice_chunk) {
    if (!ctx) return MWG_ERR_ARG;
    if (fp64_chunk < 1 || fp64_chunk > 4096 || slice_chunk < 0 || slice_chunk > MAX_SLICE_CHUNK) {
        ctx->err = "mwg_ctx_set_tuning: fp64_chunk must be in [1, 4096] and slice_chunk in [0, 4096]";
        return MWG_ERR_ARG;
    }
    ctx->chunk = fp64_chunk;
    ctx->slice_chunk = slice_chunk;
    return MWG_OK;
}

int mwg_ctx_get_tuning(mwg_ctx ctx, int *fp64_chunk, int *slice_chunk) {
    if (!ctx || !fp64_chunk || !slice_chunk) return MWG_ERR_ARG;
    *fp64_chunk = ctx->chunk;
    *slice_chunk = ctx->slice_chunk;
    return MWG_OK;
}

int mwg_ctx_trim(mwg_ctx ctx, int64_t *released) {
    if (!ctx) return MWG_ERR_ARG;
    return guarded(ctx, [&] {
        HC(hipStreamSynchronize(ctx->stream));
        const int64_t b = (int64_t)ctx->cache.bytes;
        ctx->cache.flush();
        if (released) *released = b;
    });
}

int mwg_ctx_set_predictor(mwg_ctx ctx, mwg_predictor pred, float threshold) {
    if (!ctx) return MWG_ERR_ARG;
    if (pred && pred->ctx != ctx) return MWG_ERR_ARG;
    ctx->pred = pred;
    ctx->pred_thr = threshold;
    return MWG_OK;
}

int mwg_trace_enable(mwg_ctx ctx, int on) {
    if (!ctx) return MWG_ERR_ARG;
    ctx->tracing = on != 0;
    ctx->trace_X.clear();
    ctx->trace_y.clear();
    return MWG_OK;
}

int mwg_trace_size(mwg_ctx ctx, int64_t *n) {
    if (!ctx || !n) return MWG_ERR_ARG;
    *n = (int64_t)ctx->trace_y.size();
    return MWG_OK;
}

int mwg_trace_get(mwg_ctx ctx, float *X, float *y) {
    if (!ctx) return MWG_ERR_ARG;
    if (X) std::memcpy(X, ctx->trace_X.data(), ctx->trace_X.size() * sizeof(float));
    if (y) std::memcpy(y, ctx->trace_y.data(), ctx->trace_y.size() * sizeof(float));
    return MWG_OK;
}

}  // extern "C"
