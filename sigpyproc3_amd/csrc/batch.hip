// Batched DM-trial sweep: many plans (uneven nbits, nchans, nsamps and DM lists) run by one
// persistent kernel launch over a device work queue.
//
// The queue is a sequence of segments. A round is a set of instances whose stage-1 copies G
// fit one G pool; each round has a stage-1 segment B_r (items: B_TJ words x BATCH_BCW packed
// bytes of one instance) and a stage-2 segment S_r (items: D DM rows x ntt time tiles of one
// instance). With R rounds the order is
//   B_0, B_1, S_0, S_1, B_2, B_3, S_2, S_3, B_4, ...
// and two pools alternate: S_r waits until every item of B_r is complete, B_r (r >= 2) until
// every item of S_{r-2} is (it overwrites the pool S_{r-2} read). Each wait is on the segment
// two places earlier, which has drained while the segment between ran, so blocks do not idle
// at segment changes (with B_2 right after S_0 every block waited for S_0's last items). One
// round needs no pool reuse. Every segment has 8 sub-queues, one per XCD (block b runs on XCD b % 8); a
// sub-queue takes runs of consecutive time tiles (stage 2) or time blocks (stage 1) of every
// instance, so the blocks that share G windows or staged rows share that XCD's L2. A block
// drains its own sub-queue, then takes items from the others, then moves to the next
// segment. Within a sub-queue the round's instances come in decreasing cost per item.
//
// Waits cannot deadlock: a block claims an item of a segment only after every sub-queue of
// the earlier segments was empty, so each item a wait depends on was claimed by a running
// block; that block finishes it without waiting on anything later in the queue, then finds
// the segment empty and publishes its count. Counts are published once per block and
// segment: stage-1 counts with an agent-scope release (the G stores become visible to every
// XCD), stage-2 counts relaxed (only the completion of their reads matters). A wait that
// polls SPIN_LIMIT times sets an error flag and skips its segment's items, so every wave
// exits. The last block out clears the queue counters for the next launch.
#include "spp_sweep_dev.hpp"
#include "spp_sweep_plan.hpp"

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

namespace {

constexpr int BQ = 8;                       // sub-queues per segment (one per XCD)
constexpr int BT_RUN = 4;                   // consecutive time-tile groups per sub-queue run
constexpr int BB_RUN = 32;                  // consecutive stage-1 time blocks per sub-queue run
constexpr int BATCH_BCW = 32;               // packed bytes per row of a stage-1 item
constexpr int BATCH_LOADS = 16;             // stage-1 staging loads in flight per thread
constexpr uint32_t SPIN_LIMIT = 1u << 24;   // polls (about 0.2 us each) before a wait fails
constexpr int NSHAPES = 5;

struct BInst {
    const uint8_t* packed;
    uint32_t* G;
    float* plane;
    const uint32_t* relw4;
    const int32_t* basew;
    const int32_t* tile_dm0;
    const int32_t* tile_ndm;
    const int32_t* tile_spread;
    const int32_t* row_len;
    int64_t nsamps, nchans, gpitch, glen, rpitch4, ld, ntiles_dm, ntiles_t, nby;
    int32_t nbits, big, use32, aligned;
};

// Instance descriptors are read through the constant address space: the loads are scalar and
// invariant, so the compiler may repeat them where a value is used instead of holding it in
// registers across an item (as it does with kernel arguments).
typedef __attribute__((address_space(4))) const BInst CInst;

// Read-only plan tables in the constant address space (scalar loads for uniform indices).
typedef __attribute__((address_space(4))) const uint32_t cu32_t;
typedef __attribute__((address_space(4))) const int32_t ci32_t;

// A pointer from a descriptor, marked as global memory (a pointer loaded from memory is
// otherwise accessed with flat instructions).
template <typename T>
__device__ __forceinline__ T* glob(T* p) {
    typedef __attribute__((address_space(1))) T gT;
    return (T*)(gT*)p;
}

struct BSeg {
    int32_t kind;   // 0: stage-1 items, 1: stage-2 items
    int32_t dep;    // segment whose items must all be complete first (-1: none)
    int32_t inst0;  // the round's instances are rinst[inst0, inst0 + ninst), by decreasing cost
    int32_t ninst;
    int64_t qoff;   // sub-queue prefix table: qpref[qoff + x * (ninst + 1) + k], k <= ninst
    int64_t total;  // items in the segment
};

// The queue tables of a bound batch (device memory; one kernel argument).
struct BArgs {
    const BInst* insts;
    const BSeg* segs;
    const int32_t* rinst;
    const int64_t* qpref;
    uint32_t* ctr;
    int32_t nseg;
    int32_t ntt;
};

// Entries of the counter block: [0] blocks exited, [1] error flags, then the completed items
// of each segment, then the heads of its BQ sub-queues.
constexpr int CTR_EXIT = 0, CTR_ERR = 1, CTR_DONE = 2;

// Items of sub-queue x among n runs-partitioned units (unit u belongs to sub-queue
// (u / run) % BQ).
__host__ __device__ inline int64_t units_of(int64_t n, int64_t run, int x) {
    const int64_t cyc = run * BQ;
    const int64_t rem = n % cyc - static_cast<int64_t>(x) * run;
    return (n / cyc) * run + std::min<int64_t>(std::max<int64_t>(rem, 0), run);
}

// The i-th unit of sub-queue x.
__device__ inline int64_t unit_at(int64_t i, int64_t run, int x) {
    return ((i / run) * BQ + x) * run + i % run;
}

// Largest k in [0, n) with pref[k] <= l (pref non-decreasing, pref[n] > l).
__device__ inline int find_inst(const int64_t* __restrict__ pref, int n, int64_t l) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= l) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The thread index, opaque to the optimiser: values a unit derives from it (lane addresses,
// offsets) are then computed inside the unit instead of being hoisted out of the persistent
// loop, where they would stay live across every path of the kernel (and be spilled).
__device__ __forceinline__ int opaque_tid() {
    int t = static_cast<int>(threadIdx.x);
    asm volatile("" : "+v"(t));
    return t;
}

template <int NBITS, bool BIG, int NT>
__device__ __forceinline__ void build_item(CInst& in, int64_t j0, int64_t b0, uint8_t* raw) {
    if (in.aligned)
        build_tile<NBITS, BIG, true, BATCH_BCW, BATCH_LOADS, NT>(glob(in.packed), in.nsamps, in.nchans, glob(in.G),
                                                                 in.gpitch, in.glen, j0, b0, raw, opaque_tid());
    else
        build_tile<NBITS, BIG, false, BATCH_BCW, BATCH_LOADS, NT>(glob(in.packed), in.nsamps, in.nchans, glob(in.G),
                                                                  in.gpitch, in.glen, j0, b0, raw, opaque_tid());
}

// Stage-1 unit: words [j0, j0 + B_TJ) of the packed row bytes [b0, b0 + BATCH_BCW).
template <int NT>
__device__ __forceinline__ void run_build(CInst& in, int64_t j0, int64_t b0, uint8_t* raw) {
    switch (in.nbits * 2 + in.big) {
        case 2: build_item<1, false, NT>(in, j0, b0, raw); break;
        case 3: build_item<1, true, NT>(in, j0, b0, raw); break;
        case 4: build_item<2, false, NT>(in, j0, b0, raw); break;
        case 5: build_item<2, true, NT>(in, j0, b0, raw); break;
        case 8: build_item<4, false, NT>(in, j0, b0, raw); break;
        case 9: build_item<4, true, NT>(in, j0, b0, raw); break;
        default: build_item<8, false, NT>(in, j0, b0, raw); break;
    }
}

// Block-uniform values read from LDS (readfirstlane: scalar registers).
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(v)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(v >> 32)));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Queue state of a block, kept in LDS and advanced by its thread 0 (next_item), which also
// decodes each unit; the other threads read the unit after a barrier. No queue state stays
// in registers across a unit, which keeps the stage-2 code near the register budget of the
// one-plan kernel.
struct BlockQueue {
    int64_t a, b;    // the unit: stage 1 (j0, b0); stage 2 (DM tile, time tile)
    int64_t tt_end;  // stage 2: end of the current item's time tiles
    int32_t sub;     // stage 2: sub-tile of the DM tile
    int32_t kind, inst;
    int32_t seg;     // current segment
    int32_t q;       // sub-queues of this segment already found empty (from the block's own)
    int32_t ndone;   // items of this segment the block completed
    int32_t pending; // an item was handed out and has not been counted yet
    int32_t ready;   // the segment's dependency has been waited for
    int32_t skip;    // the wait failed: the segment's items are counted but not run
    int32_t acquire; // a wait just ended: every wave acquires before reading
    int32_t end;     // every segment is empty
};

// Thread 0: the next time tile of the current item; or else publish the completed items of
// exhausted segments, claim the next item (its own XCD's sub-queue first), wait for the
// segment's dependency once, and decode the item's first unit. nsub: DM sub-tiles per tile.
__device__ __forceinline__ void next_item(const BArgs* A, BlockQueue* Q, int nsub) {
    Q->acquire = 0;
    if (Q->pending && Q->kind == 1 && Q->b + 1 < Q->tt_end) {
        ++Q->b;
        return;
    }
    uint32_t* done = A->ctr + CTR_DONE;
    uint32_t* heads = A->ctr + CTR_DONE + A->nseg;
    const int x0 = static_cast<int>(blockIdx.x % BQ);
    if (Q->pending) {  // the item handed out last has been run
        ++Q->ndone;
        Q->pending = 0;
    }
    while (Q->seg < A->nseg) {
        const BSeg sg = A->segs[Q->seg];
        int64_t item = -1;
        while (Q->q < BQ) {
            const int x = (x0 + Q->q) % BQ;
            const int64_t n = A->qpref[sg.qoff + static_cast<int64_t>(x) * (sg.ninst + 1) + sg.ninst];
            if (n > 0) {
                const uint32_t h = atomicAdd(&heads[Q->seg * BQ + x], 1u);
                if (h < n) {
                    item = h;
                    break;
                }
            }
            ++Q->q;
        }
        if (item < 0) {
            if (Q->ndone > 0) {
                const uint32_t nd = static_cast<uint32_t>(Q->ndone);
                if (sg.kind == 0)
                    __hip_atomic_fetch_add(&done[Q->seg], nd, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                else
                    __hip_atomic_fetch_add(&done[Q->seg], nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            ++Q->seg;
            Q->q = 0;
            Q->ndone = 0;
            Q->ready = 0;
            Q->skip = 0;
            continue;
        }
        if (!Q->ready) {
            if (sg.dep >= 0) {
                const uint32_t need = static_cast<uint32_t>(A->segs[sg.dep].total);
                uint32_t polls = 0;
                while (__hip_atomic_load(&done[sg.dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    __builtin_amdgcn_s_sleep(8);
                    if (++polls >= SPIN_LIMIT) {
                        atomicOr(&A->ctr[CTR_ERR], 1u);
                        Q->skip = 1;
                        break;
                    }
                }
                Q->acquire = 1;
            }
            Q->ready = 1;
        }
        const int x = (x0 + Q->q) % BQ;
        const int64_t* pref = A->qpref + sg.qoff + static_cast<int64_t>(x) * (sg.ninst + 1);
        const int k = find_inst(pref, sg.ninst, item);
        const int inst = A->rinst[sg.inst0 + k];
        const BInst* in = A->insts + inst;
        const int64_t li = item - pref[k];
        if (sg.kind == 0) {
            const int64_t nby = in->nby;
            Q->a = unit_at(li / nby, BB_RUN, x) * B_TJ;
            Q->b = (li % nby) * BATCH_BCW;
        } else {
            const int64_t per = in->ntiles_dm * nsub;
            const int64_t grp = unit_at(li / per, BT_RUN, x);
            const int64_t r = li % per;
            Q->a = r / nsub;
            Q->sub = static_cast<int32_t>(r % nsub);
            Q->b = grp * A->ntt;
            Q->tt_end = std::min<int64_t>(Q->b + A->ntt, in->ntiles_t);
        }
        Q->kind = sg.kind;
        Q->inst = inst;
        Q->pending = 1;
        Q->end = 0;
        return;
    }
    Q->end = 1;
}

template <int K, int WAVES, int CC, int MINW>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(MINW))) void k_batch(const BArgs* A) {
    static_assert(WAVES * 64 == B_THREADS || WAVES * 64 == 2 * B_THREADS,
                  "stage-1 items take the block's threads (B_THREADS or twice that)");
    constexpr int SWEEP_WORDS = 2 * CC * SW_LW;
    constexpr int BUILD_WORDS = (build_lds_bytes<1, BATCH_BCW>() + 3) / 4;  // IL = 8 stages the most
    __shared__ uint32_t lds[SWEEP_WORDS > BUILD_WORDS ? SWEEP_WORDS : BUILD_WORDS];
    __shared__ BlockQueue Q;
    const int tid = threadIdx.x;
    if (tid == 0) {
        Q.seg = 0;
        Q.q = 0;
        Q.ndone = 0;
        Q.pending = 0;
        Q.ready = 0;
        Q.skip = 0;
    }
    for (;;) {
        if (tid == 0) next_item(A, &Q, SW_DT / (K * WAVES));
        __syncthreads();
        if (uni(Q.end)) break;
        if (uni(Q.acquire)) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (!uni(Q.skip)) {
            CInst& in = *(CInst*)(A->insts + uni(Q.inst));
            if (uni(Q.kind) == 0) {
                run_build<WAVES * 64>(in, uni(Q.a), uni(Q.b), reinterpret_cast<uint8_t*>(lds));
            } else {
                // Stage-2 unit: sub-tile sub of DM tile dt over time tile tt, the one-plan
                // kernel's body for the instance's field width (included per case, so that it
                // compiles as part of this function).
#define SPP_TILE_EXIT break
#define SPP_TILE_TID opaque_tid()
                switch (in.nbits * 2 + in.use32) {
                    case 2: {
                        constexpr int NBITS = 1;
                        constexpr bool USE32 = false;
#include "spp_batch_case.inc"
                    } break;
                    case 3: {
                        constexpr int NBITS = 1;
                        constexpr bool USE32 = true;
#include "spp_batch_case.inc"
                    } break;
                    case 4: {
                        constexpr int NBITS = 2;
                        constexpr bool USE32 = false;
#include "spp_batch_case.inc"
                    } break;
                    case 5: {
                        constexpr int NBITS = 2;
                        constexpr bool USE32 = true;
#include "spp_batch_case.inc"
                    } break;
                    case 8: {
                        constexpr int NBITS = 4;
                        constexpr bool USE32 = false;
#include "spp_batch_case.inc"
                    } break;
                    case 9: {
                        constexpr int NBITS = 4;
                        constexpr bool USE32 = true;
#include "spp_batch_case.inc"
                    } break;
                    case 16: {
                        constexpr int NBITS = 8;
                        constexpr bool USE32 = false;
#include "spp_batch_case.inc"
                    } break;
                    case 17: {
                        constexpr int NBITS = 8;
                        constexpr bool USE32 = true;
#include "spp_batch_case.inc"
                    } break;
                    default: break;
                }
#undef SPP_TILE_TID
#undef SPP_TILE_EXIT
            }
        }
        __syncthreads();  // the next unit restages the LDS this one read; Q is rewritten
    }
    if (tid == 0) {
        uint32_t* ctr = A->ctr;
        const uint32_t e = atomicAdd(&ctr[CTR_EXIT], 1u);
        if (e == gridDim.x - 1) {
            // Every other block has exited: clear the counters for the next launch.
            for (int i = CTR_DONE; i < CTR_DONE + A->nseg * (BQ + 1); ++i) ctr[i] = 0;
            ctr[CTR_EXIT] = 0;
        }
    }
}

// Block shapes (DM rows per wave, waves, channels per stage, min waves per SIMD): every shape
// takes 1/2/4/8-bit data (CC <= 16 keeps 4-bit stage sums in byte lanes).
template <int S>
struct Shape;
template <> struct Shape<0> { static constexpr int K = 2, W = 8, CC = 12, MINW = 6; };
template <> struct Shape<1> { static constexpr int K = 2, W = 8, CC = 16, MINW = 4; };
template <> struct Shape<2> { static constexpr int K = 1, W = 8, CC = 8, MINW = 8; };
template <> struct Shape<3> { static constexpr int K = 2, W = 8, CC = 8, MINW = 6; };
// One DM row per wave in 16-wave blocks (as the one-plan kernel's nibble default): 64 KB of LDS,
// two blocks per CU at 8 waves per SIMD.
template <> struct Shape<4> { static constexpr int K = 1, W = 16, CC = 16, MINW = 8; };

template <int S>
int launch_shape(const BArgs* args, int64_t grid, hipStream_t st) {
    using Sh = Shape<S>;
    hipLaunchKernelGGL((k_batch<Sh::K, Sh::W, Sh::CC, Sh::MINW>), dim3(static_cast<unsigned>(grid)), dim3(Sh::W * 64),
                       0, st, args);
    SPP_LAUNCHED("k_batch");
    return SPP_OK;
}

template <int S>
int blocks_per_cu(int* n) {
    using Sh = Shape<S>;
    SPP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_batch<Sh::K, Sh::W, Sh::CC, Sh::MINW>, Sh::W * 64, 0));
    return SPP_OK;
}

int shape_rows(int s) {
    switch (s) {
        case 0: return Shape<0>::K * Shape<0>::W;
        case 1: return Shape<1>::K * Shape<1>::W;
        case 2: return Shape<2>::K * Shape<2>::W;
        case 3: return Shape<3>::K * Shape<3>::W;
        default: return Shape<4>::K * Shape<4>::W;
    }
}

int shape_cc(int s) {
    switch (s) {
        case 0: return Shape<0>::CC;
        case 1: return Shape<1>::CC;
        case 2: return Shape<2>::CC;
        case 3: return Shape<3>::CC;
        default: return Shape<4>::CC;
    }
}

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

int64_t g_bytes(const spp_sweep_plan* p) { return align256(4 * p->nchans * p->gpitch); }

int64_t tables_bytes(const spp_sweep_plan* p) {
    const int64_t nt = static_cast<int64_t>(p->tile_dm0.size());
    return align256(p->ndm * p->rpitch) + align256(4 * nt * p->nchans) + 3 * align256(4 * nt) +
           align256(4 * p->ndm);
}

}  // namespace

struct spp_batch {
    std::vector<const spp_sweep_plan*> plans;
    int shape = 4;  // 16 waves of one DM row (Shape<4>): the fastest on the mixed 64-instance batch
    int ntt = 1;
    int64_t g_budget = 0;
    int nrounds = 0, npools = 1;
    int64_t pool_bytes = 0;
    std::vector<int32_t> round_of;   // per instance
    std::vector<int64_t> g_off;      // per instance: G offset inside its pool
    std::vector<int64_t> tab_off;    // per instance: tables offset in the workspace
    int64_t off_args = 0, off_ctr = 0, off_insts = 0, off_segs = 0, off_rinst = 0, off_qpref = 0, off_pools = 0;
    int64_t nseg = 0, ctr_words = 0, qpref_len = 0, ws_bytes = 0;
    const void* uploaded_to = nullptr;
    const void* bound_to = nullptr;
    int64_t grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    ~spp_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

int64_t ntiles_t_of(const spp_sweep_plan* p) {
    return spp::ceil_div(p->max_len, int64_t(p->nbits <= 2 ? 8 : 4) * SW_S);
}

// Cost of one stage-2 item relative to others (channels x samples summed per word).
double item_cost(const spp_sweep_plan* p) {
    return static_cast<double>(p->nchans) * (p->nbits <= 2 ? 2.0 : 1.0);
}

// Rounds: every instance's G in one pool when the budget allows, otherwise first-fit
// decreasing into rounds of half the budget (two pools alternate).
int plan_rounds(spp_batch* b) {
    const int n = static_cast<int>(b->plans.size());
    int64_t total = 0, largest = 0;
    for (const spp_sweep_plan* p : b->plans) {
        total += g_bytes(p);
        largest = std::max(largest, g_bytes(p));
    }
    b->round_of.assign(n, 0);
    b->g_off.assign(n, 0);
    if (b->g_budget <= 0 || total <= b->g_budget) {
        b->nrounds = 1;
        b->npools = 1;
        b->pool_bytes = total;
        int64_t off = 0;
        for (int i = 0; i < n; ++i) {
            b->g_off[i] = off;
            off += g_bytes(b->plans[i]);
        }
        return SPP_OK;
    }
    const int64_t cap = (b->g_budget / 2) & ~int64_t(255);
    SPP_REQUIRE(largest <= cap,
                "G budget %lld too small: an instance's stage-1 copy needs %lld bytes and two pools "
                "of at least that size are needed",
                static_cast<long long>(b->g_budget), static_cast<long long>(largest));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return g_bytes(b->plans[a]) > g_bytes(b->plans[c]); });
    std::vector<int64_t> fill;
    for (int i : order) {
        const int64_t need = g_bytes(b->plans[i]);
        size_t r = 0;
        while (r < fill.size() && fill[r] + need > cap) ++r;
        if (r == fill.size()) fill.push_back(0);
        b->round_of[i] = static_cast<int32_t>(r);
        b->g_off[i] = fill[r];
        fill[r] += need;
    }
    b->nrounds = static_cast<int>(fill.size());
    b->npools = 2;
    b->pool_bytes = cap;
    return SPP_OK;
}

struct QueueTables {
    std::vector<BSeg> segs;
    std::vector<int32_t> rinst;
    std::vector<int64_t> qpref;
};

QueueTables queue_tables(const spp_batch* b) {
    QueueTables t;
    const int R = b->nrounds;
    const int nsub = SW_DT / shape_rows(b->shape);
    // The round's instances by decreasing stage-2 cost per item.
    std::vector<int32_t> first(R + 1, 0);
    for (int r = 0; r < R; ++r) {
        std::vector<int32_t> inst;
        for (size_t i = 0; i < b->plans.size(); ++i)
            if (b->round_of[i] == r) inst.push_back(static_cast<int32_t>(i));
        std::stable_sort(inst.begin(), inst.end(),
                         [&](int32_t a, int32_t c) { return item_cost(b->plans[a]) > item_cost(b->plans[c]); });
        first[r] = static_cast<int32_t>(t.rinst.size());
        t.rinst.insert(t.rinst.end(), inst.begin(), inst.end());
    }
    first[R] = static_cast<int32_t>(t.rinst.size());
    std::vector<int> seg_b(R), seg_s(R);
    auto add = [&](int kind, int r) {
        BSeg s;
        s.kind = kind;
        s.dep = -1;
        s.inst0 = first[r];
        s.ninst = first[r + 1] - first[r];
        s.qoff = static_cast<int64_t>(t.qpref.size());
        s.total = 0;
        for (int x = 0; x < BQ; ++x) {
            int64_t acc = 0;
            for (int k = 0; k < s.ninst; ++k) {
                t.qpref.push_back(acc);
                const spp_sweep_plan* p = b->plans[t.rinst[s.inst0 + k]];
                if (kind == 0) {
                    const int64_t ntb = spp::ceil_div(p->gpitch, int64_t(B_TJ));
                    acc += units_of(ntb, BB_RUN, x) * spp::ceil_div(p->nchans * p->nbits / 8, int64_t(BATCH_BCW));
                } else {
                    const int64_t ngrp = spp::ceil_div(ntiles_t_of(p), int64_t(b->ntt));
                    acc += units_of(ngrp, BT_RUN, x) * static_cast<int64_t>(p->tile_dm0.size()) * nsub;
                }
            }
            t.qpref.push_back(acc);
            s.total += acc;
        }
        (kind == 0 ? seg_b : seg_s)[r] = static_cast<int>(t.segs.size());
        t.segs.push_back(s);
    };
    add(0, 0);
    if (R > 1) add(0, 1);
    for (int r = 0; r < R; r += 2) {
        add(1, r);
        if (r + 1 < R) add(1, r + 1);
        if (r + 2 < R) add(0, r + 2);
        if (r + 3 < R) add(0, r + 3);
    }
    for (int r = 0; r < R; ++r) {
        t.segs[seg_s[r]].dep = seg_b[r];
        if (r >= 2) t.segs[seg_b[r]].dep = seg_s[r - 2];
    }
    return t;
}

int layout(spp_batch* b) {
    const int n = static_cast<int>(b->plans.size());
    const QueueTables t = queue_tables(b);
    b->nseg = static_cast<int64_t>(t.segs.size());
    b->qpref_len = static_cast<int64_t>(t.qpref.size());
    b->ctr_words = CTR_DONE + b->nseg * (BQ + 1);
    for (const BSeg& s : t.segs)
        SPP_REQUIRE(s.total < (int64_t(1) << 31), "a batch segment holds %lld items (limit 2^31)",
                    static_cast<long long>(s.total));
    int64_t off = 0;
    b->off_args = off; off = align256(off + static_cast<int64_t>(sizeof(BArgs)));
    b->off_ctr = off; off = align256(off + 4 * b->ctr_words);
    b->off_insts = off; off = align256(off + static_cast<int64_t>(sizeof(BInst)) * n);
    // room for the queue tables of any shape and time-tile grouping: the segment count is
    // fixed by the rounds, the prefix table by the instance count
    b->off_segs = off; off = align256(off + static_cast<int64_t>(sizeof(BSeg)) * b->nseg);
    b->off_rinst = off; off = align256(off + 4 * n);
    b->off_qpref = off; off = align256(off + 8 * b->qpref_len);
    b->tab_off.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        b->tab_off[i] = off;
        off += tables_bytes(b->plans[i]);
    }
    b->off_pools = off;
    off += b->npools * b->pool_bytes;
    b->ws_bytes = off;
    return SPP_OK;
}

int launch(spp_batch* b, void* workspace, hipStream_t st) {
    const auto* args = reinterpret_cast<const BArgs*>(static_cast<char*>(workspace) + b->off_args);
    switch (b->shape) {
        case 0: return launch_shape<0>(args, b->grid, st);
        case 1: return launch_shape<1>(args, b->grid, st);
        case 2: return launch_shape<2>(args, b->grid, st);
        case 3: return launch_shape<3>(args, b->grid, st);
        default: return launch_shape<4>(args, b->grid, st);
    }
}

int resident_grid(int shape, int64_t* grid) {
    int per_cu = 0;
    int rc;
    switch (shape) {
        case 0: rc = blocks_per_cu<0>(&per_cu); break;
        case 1: rc = blocks_per_cu<1>(&per_cu); break;
        case 2: rc = blocks_per_cu<2>(&per_cu); break;
        case 3: rc = blocks_per_cu<3>(&per_cu); break;
        default: rc = blocks_per_cu<4>(&per_cu); break;
    }
    if (rc != SPP_OK) return rc;
    const int64_t n = static_cast<int64_t>(std::max(per_cu, 1)) * spp::device_cu_count();
    *grid = std::max<int64_t>(BQ, (n / BQ) * BQ);
    return SPP_OK;
}

int check_bound(const spp_batch* b, const void* workspace, int64_t workspace_bytes) {
    SPP_REQUIRE(b != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace_bytes >= b->ws_bytes, "workspace too small");
    SPP_REQUIRE(workspace == b->uploaded_to && workspace == b->bound_to,
                "workspace is not the buffer the batch was uploaded to and bound in");
    return SPP_OK;
}

}  // namespace

extern "C" int spp_batch_create(const spp_sweep_plan* const* plans, int32_t n, int64_t g_budget_bytes,
                                spp_batch** batch) {
    spp::clear_error();
    SPP_REQUIRE(batch != nullptr, "null batch pointer");
    *batch = nullptr;
    SPP_REQUIRE(plans != nullptr && n > 0, "a batch needs at least one plan");
    SPP_REQUIRE(n <= 65536, "at most 65536 plans per batch");
    for (int i = 0; i < n; ++i) SPP_REQUIRE(plans[i] != nullptr, "plan %d is null", i);
    spp_batch* b = new (std::nothrow) spp_batch();
    SPP_REQUIRE(b != nullptr, "out of host memory");
    b->plans.assign(plans, plans + n);
    b->g_budget = g_budget_bytes;
    int rc = plan_rounds(b);
    if (rc == SPP_OK) rc = layout(b);
    if (rc != SPP_OK) {
        delete b;
        return rc;
    }
    *batch = b;
    return SPP_OK;
}

extern "C" int spp_batch_destroy(spp_batch* batch) {
    delete batch;
    return SPP_OK;
}

extern "C" int spp_batch_shape_count(void) { return NSHAPES; }

extern "C" int spp_batch_set_shape(spp_batch* b, int shape, int32_t time_tiles_per_item) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr, "null batch");
    SPP_REQUIRE(shape >= 0 && shape < NSHAPES, "shape %d outside [0, %d)", shape, NSHAPES);
    SPP_REQUIRE(time_tiles_per_item >= 1 && time_tiles_per_item <= 64, "time tiles per item must be in [1, 64]");
    b->shape = shape;
    b->ntt = time_tiles_per_item;
    b->bound_to = nullptr;  // the queue tables change: bind again
    return layout(b);
}

extern "C" int spp_batch_get_info(const spp_batch* b, spp_batch_info* info) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr && info != nullptr, "null argument");
    const QueueTables t = queue_tables(b);
    info->workspace_bytes = b->ws_bytes;
    info->pool_bytes = b->pool_bytes;
    info->rounds = b->nrounds;
    info->segments = static_cast<int32_t>(b->nseg);
    info->shape = b->shape;
    info->time_tiles_per_item = b->ntt;
    info->dm_rows_per_item = shape_rows(b->shape);
    info->chans_per_stage = shape_cc(b->shape);
    info->build_items = 0;
    info->sweep_items = 0;
    for (const BSeg& s : t.segs) (s.kind == 0 ? info->build_items : info->sweep_items) += s.total;
    return SPP_OK;
}

extern "C" int spp_batch_rounds(const spp_batch* b, int32_t* round_of_host) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr && round_of_host != nullptr, "null argument");
    std::copy(b->round_of.begin(), b->round_of.end(), round_of_host);
    return SPP_OK;
}

extern "C" int spp_batch_upload(spp_batch* b, void* workspace, int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr && workspace != nullptr, "null argument");
    SPP_REQUIRE(workspace_bytes >= b->ws_bytes, "workspace too small: %lld < %lld bytes",
                static_cast<long long>(workspace_bytes), static_cast<long long>(b->ws_bytes));
    hipStream_t st = spp::as_stream(stream);
    char* ws = static_cast<char*>(workspace);
    for (size_t i = 0; i < b->plans.size(); ++i) {
        const spp_sweep_plan* p = b->plans[i];
        const int64_t nt = static_cast<int64_t>(p->tile_dm0.size());
        char* t = ws + b->tab_off[i];
        SPP_HIP(hipMemcpyAsync(t, p->relw.data(), p->ndm * p->rpitch, hipMemcpyHostToDevice, st));
        t += align256(p->ndm * p->rpitch);
        SPP_HIP(hipMemcpyAsync(t, p->base.data(), 4 * nt * p->nchans, hipMemcpyHostToDevice, st));
        t += align256(4 * nt * p->nchans);
        SPP_HIP(hipMemcpyAsync(t, p->tile_dm0.data(), 4 * nt, hipMemcpyHostToDevice, st));
        t += align256(4 * nt);
        SPP_HIP(hipMemcpyAsync(t, p->tile_ndm.data(), 4 * nt, hipMemcpyHostToDevice, st));
        t += align256(4 * nt);
        SPP_HIP(hipMemcpyAsync(t, p->tile_spread.data(), 4 * nt, hipMemcpyHostToDevice, st));
        t += align256(4 * nt);
        SPP_HIP(hipMemcpyAsync(t, p->row_len.data(), 4 * p->ndm, hipMemcpyHostToDevice, st));
    }
    SPP_HIP(hipStreamSynchronize(st));
    b->uploaded_to = workspace;
    b->bound_to = nullptr;
    return SPP_OK;
}

extern "C" int spp_batch_bind(spp_batch* b, const uint8_t* const* packed, float* const* planes,
                              const int64_t* ld, void* workspace, int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr && packed != nullptr && planes != nullptr && ld != nullptr && workspace != nullptr,
                "null argument");
    SPP_REQUIRE(workspace == b->uploaded_to, "bind needs the workspace the batch was uploaded to");
    SPP_REQUIRE(workspace_bytes >= b->ws_bytes, "workspace too small");
    const int n = static_cast<int>(b->plans.size());
    char* ws = static_cast<char*>(workspace);
    std::vector<BInst> insts(n);
    for (int i = 0; i < n; ++i) {
        const spp_sweep_plan* p = b->plans[i];
        SPP_REQUIRE(packed[i] != nullptr && planes[i] != nullptr, "instance %d: null stream or plane", i);
        SPP_REQUIRE(ld[i] >= p->max_len, "instance %d: ld %lld < longest row %lld", i,
                    static_cast<long long>(ld[i]), static_cast<long long>(p->max_len));
        const int64_t nt = static_cast<int64_t>(p->tile_dm0.size());
        char* t = ws + b->tab_off[i];
        BInst& d = insts[i];
        d.packed = packed[i];
        d.G = reinterpret_cast<uint32_t*>(ws + b->off_pools + b->round_of[i] % b->npools * b->pool_bytes + b->g_off[i]);
        d.plane = planes[i];
        d.relw4 = reinterpret_cast<const uint32_t*>(t);
        t += align256(p->ndm * p->rpitch);
        d.basew = reinterpret_cast<const int32_t*>(t);
        t += align256(4 * nt * p->nchans);
        d.tile_dm0 = reinterpret_cast<const int32_t*>(t);
        t += align256(4 * nt);
        d.tile_ndm = reinterpret_cast<const int32_t*>(t);
        t += align256(4 * nt);
        d.tile_spread = reinterpret_cast<const int32_t*>(t);
        t += align256(4 * nt);
        d.row_len = reinterpret_cast<const int32_t*>(t);
        d.nsamps = p->nsamps;
        d.nchans = p->nchans;
        d.gpitch = p->gpitch;
        d.glen = p->gpitch;
        d.rpitch4 = p->rpitch / 4;
        d.ld = ld[i];
        d.ntiles_dm = nt;
        d.ntiles_t = ntiles_t_of(p);
        d.nby = spp::ceil_div(p->nchans * p->nbits / 8, int64_t(BATCH_BCW));
        d.nbits = p->nbits;
        d.big = p->big;
        d.use32 = p->use32 ? 1 : 0;
        d.aligned = ((p->nchans * p->nbits / 8) % 4 == 0 && reinterpret_cast<uintptr_t>(packed[i]) % 4 == 0) ? 1 : 0;
    }
    const QueueTables t = queue_tables(b);
    SPP_REQUIRE(static_cast<int64_t>(t.segs.size()) == b->nseg && static_cast<int64_t>(t.qpref.size()) == b->qpref_len,
                "internal: queue tables do not match the layout");
    int rc = resident_grid(b->shape, &b->grid);
    if (rc != SPP_OK) return rc;
    BArgs args;
    args.insts = reinterpret_cast<const BInst*>(ws + b->off_insts);
    args.segs = reinterpret_cast<const BSeg*>(ws + b->off_segs);
    args.rinst = reinterpret_cast<const int32_t*>(ws + b->off_rinst);
    args.qpref = reinterpret_cast<const int64_t*>(ws + b->off_qpref);
    args.ctr = reinterpret_cast<uint32_t*>(ws + b->off_ctr);
    args.nseg = static_cast<int32_t>(b->nseg);
    args.ntt = b->ntt;
    hipStream_t st = spp::as_stream(stream);
    SPP_HIP(hipMemcpyAsync(ws + b->off_args, &args, sizeof(BArgs), hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemsetAsync(ws + b->off_ctr, 0, 4 * b->ctr_words, st));
    SPP_HIP(hipMemcpyAsync(ws + b->off_insts, insts.data(), sizeof(BInst) * n, hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + b->off_segs, t.segs.data(), sizeof(BSeg) * t.segs.size(), hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + b->off_rinst, t.rinst.data(), 4 * t.rinst.size(), hipMemcpyHostToDevice, st));
    SPP_HIP(hipMemcpyAsync(ws + b->off_qpref, t.qpref.data(), 8 * t.qpref.size(), hipMemcpyHostToDevice, st));
    SPP_HIP(hipStreamSynchronize(st));
    b->bound_to = workspace;
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    return SPP_OK;
}

extern "C" int spp_batch_run(spp_batch* b, void* workspace, int64_t workspace_bytes, void* stream) {
    spp::clear_error();
    int rc = check_bound(b, workspace, workspace_bytes);
    if (rc) return rc;
    return launch(b, workspace, spp::as_stream(stream));
}

extern "C" int spp_batch_capture(spp_batch* b, void* workspace, int64_t workspace_bytes) {
    spp::clear_error();
    int rc = check_bound(b, workspace, workspace_bytes);
    if (rc) return rc;
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    hipStream_t cs = nullptr;
    SPP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = launch(b, workspace, cs);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (rc == SPP_OK && e2 == hipSuccess) {
            b->graph = g;
            e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        } else {
            if (g) (void)hipGraphDestroy(g);
            if (rc == SPP_OK) e = e2;
        }
    }
    (void)hipStreamDestroy(cs);
    if (rc != SPP_OK) return rc;
    SPP_HIP(e);
    return SPP_OK;
}

extern "C" int spp_batch_replay(spp_batch* b, void* stream) {
    spp::clear_error();
    SPP_REQUIRE(b != nullptr, "null batch");
    SPP_REQUIRE(b->gexec != nullptr, "no captured graph (spp_batch_capture after the last bind)");
    SPP_HIP(hipGraphLaunch(b->gexec, spp::as_stream(stream)));
    return SPP_OK;
}

extern "C" int spp_batch_errors(spp_batch* b, void* workspace, int64_t workspace_bytes, void* stream,
                                int32_t* flags) {
    spp::clear_error();
    int rc = check_bound(b, workspace, workspace_bytes);
    if (rc) return rc;
    SPP_REQUIRE(flags != nullptr, "null argument");
    hipStream_t st = spp::as_stream(stream);
    uint32_t* err = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + b->off_ctr) + CTR_ERR;
    uint32_t h = 0;
    SPP_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, st));
    SPP_HIP(hipStreamSynchronize(st));
    if (h != 0) {
        SPP_HIP(hipMemsetAsync(err, 0, 4, st));
        SPP_HIP(hipStreamSynchronize(st));
    }
    *flags = static_cast<int32_t>(h);
    return SPP_OK;
}
