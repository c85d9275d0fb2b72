// batch.hip -- independent replicas (a pH / salt titration grid) advanced
// together.  No counterpart in the reference, which runs one system per
// process.
//
// Every launch advances all replicas: workgroup r of the chain launch runs the
// chain of replica r (the persistent single-workgroup chain of mc.hip, so each
// replica's accept/reject sequence is exactly its serial one), and the two
// refresh launches cover the potential-grid tiles and planes of every replica
// through block-offset tables (the direct refresh's tiles, or the k-space
// refresh's z pencils and planes, per replica by the rule of its serial
// refresh, so each replica's potential grid is bit-identical to its serial
// run's).  A sweep gives each replica as many trials as
// it has titratable sites; the (chain, refresh) launch pairs of a sweep are
// captured once in a hipGraph and replayed until every replica has finished
// its sweep.  The autotuner times the speculation window (waves per trial),
// the pending-change budget between refreshes and the graph length on the
// batch itself and keeps the fastest.
//
// Persistent form (every replica on the direct refresh): one cooperative launch
// per sweep.  Workgroup r < R runs replica r's chain to its sweep target without
// waiting for the other replicas: whenever its pending budget fills it posts a
// refresh request (a device-scope release of a per-replica request count) and
// folds the request's tiles itself, tile by tile, while the other workgroups of
// the launch take tiles of posted requests as well; a tile is handed out by a
// compare-and-swap on the replica's tile counter, whose bound moves with the
// request count, and each finished tile is counted with a release.  The chain
// resumes once every tile of its request is counted, so its potential grid and
// accept/reject sequence are those of its serial run.  A chain never waits for
// a tile no workgroup has taken (it takes them itself), and every wait is
// bounded by a wall-clock limit that stops the launch with an error status.
#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>

#include "mc_host.hpp"
#include "mc_kernels.hpp"

namespace phb {

struct RefreshView {
    const McDev* D;                      // device view of the replica
    cplx<double>* kg;
    const cplx<double> *twx, *twy, *twz;
    const double* eta;
    FftPlan px, py, pz;
    int tx, nx, ny;
    double* phi;
    RefreshGate gate;
    const int* tiles;                    // direct refresh: the replica's listed tiles
};

// Largest r with start[r] <= b (start ascending, R entries).
__device__ __forceinline__ int replica_of(const int* __restrict__ start, int R, int b) {
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int P, typename PoolT, bool LDS_POOLS>
__global__ __launch_bounds__(MC_THREADS) void k_chain_batch(const McDev* __restrict__ views) {
    const McDev D = views[blockIdx.x];
    chain_body<P, PoolT, LDS_POOLS>(D);
}

template <int P>
__global__ __launch_bounds__(512) void k_kspace_batch(const RefreshView* __restrict__ v, const int* __restrict__ start,
                                                      int R) {
    const int r = replica_of(start, R, blockIdx.x);
    const RefreshView& w = v[r];
    refresh_kspace_tile<P>(*w.D, w.kg, w.twx, w.twy, w.twz, w.eta, w.pz, w.tx, blockIdx.x - start[r]);
}

template <int P>
__global__ __launch_bounds__(RD_THREADS) void k_direct_batch(const RefreshView* __restrict__ v,
                                                             const int* __restrict__ start, int R) {
    const int r = replica_of(start, R, blockIdx.x);
    const RefreshView& w = v[r];
    refresh_direct_tile<P>(*w.D, w.phi, nullptr, w.tiles[blockIdx.x - start[r]]);
}

__global__ __launch_bounds__(512) void k_c2r_batch(const RefreshView* __restrict__ v, const int* __restrict__ start,
                                                   int R) {
    const int r = replica_of(start, R, blockIdx.x);
    const RefreshView& w = v[r];
    fft_xy_c2r_plane<double>(w.kg, w.phi, w.nx, w.ny, w.px, w.py, w.twx, w.twy, 1, w.gate, blockIdx.x - start[r],
                             start[r + 1] - start[r]);
}

// ---- persistent sweep
struct PersistView {
    double* phi;
    const int* tiles;   // the replica's listed direct-refresh tiles
    int nt;     // how many
    int prio;   // the replica helpers look at k-th: most trials per sweep first
};

// layout of the sweep's synchronisation words (zeroed before each launch)
struct SweepSync {
    unsigned long long* req;    // [R] refresh requests posted
    unsigned long long* next;   // [R] tiles handed out (all requests, in order)
    unsigned long long* done;   // [R] tiles folded
    unsigned long long* fin;    // replicas finished
    unsigned long long* abort;  // a wait ran past the limit
    unsigned long long* tfin;   // [R] wall-clock ticks from the launch's start to the replica's finish
    unsigned long long* twait;  // [R] ticks the replica's chain spent on its refreshes
};

// Polling reads are relaxed device-scope atomic loads (coherent across the XCDs'
// L2s without invalidating them); the acquire is one fence once the awaited
// value is seen, so that spinning workgroups do not keep emptying the caches
// the chains work from.
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void acquire_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Up to `chunk` consecutive tiles of the posted requests of one replica (tiles
// below `goal`, never past the end of a request): the first tile's index within
// its request, the count in *cnt; -1 when none is left.
__device__ __forceinline__ int grab_tiles(unsigned long long* next, unsigned long long goal, int nt, int chunk,
                                          int* cnt) {
    unsigned long long t = __hip_atomic_load(next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (t < goal) {
        const unsigned long long end = (t / (unsigned long long)nt + 1ull) * (unsigned long long)nt;
        const unsigned long long c = min(min((unsigned long long)chunk, goal - t), end - t);
        if (__hip_atomic_compare_exchange_strong(next, &t, t + c, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            *cnt = (int)c;
            return (int)(t % (unsigned long long)nt);
        }
    }
    return -1;
}

// tiles handed out per grab: a request's tiles go out in about 32 grabs
__device__ __forceinline__ int tile_chunk(int nt) { return max(1, nt / 32); }

__device__ __forceinline__ bool past_limit(long long t0, long long limit) {
    return (long long)__builtin_amdgcn_s_memrealtime() - t0 > limit;
}

// One sweep of every replica (see the file comment).  limit: wall-clock ticks
// (100 MHz) any wait may last before the launch stops with status 6.
template <int P, typename PoolT, bool LDS_POOLS>
__global__ __launch_bounds__(MC_THREADS) void k_sweep_persistent(const McDev* __restrict__ views,
                                                                 const PersistView* __restrict__ pv, int R,
                                                                 SweepSync sy, long long limit) {
    __shared__ int s_cmd[4];
    const int tid = threadIdx.x;
    const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
    if ((int)blockIdx.x < R) {   // ---- chain of replica r
        const int r = blockIdx.x;
        const McDev D = views[r];
        const int nt = pv[r].nt;
        unsigned long long goal = 0;
        bool failed = false;
        for (;;) {
            chain_body<P, PoolT, LDS_POOLS>(D);
            __syncthreads();
            if (tid == 0) {
                const McCtl* C = D.ctl;
                const bool fin = C->trials_done >= C->trials_target || C->status != 0;
                const bool pend = pending_unfolded(D) > 0;
                s_cmd[0] = (fin ? 1 : 0) | (pend ? 2 : 0);
            }
            __syncthreads();
            const int st = s_cmd[0];
            __syncthreads();
            if (st & 2) {   // post the request, then fold tiles until all of them are counted
                const long long tw0 = (long long)__builtin_amdgcn_s_memrealtime();
                goal += (unsigned long long)nt;
                if (tid == 0) __hip_atomic_fetch_add(sy.req + r, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                for (;;) {
                    if (tid == 0) {
                        int cnt = 0;
                        int cmd = grab_tiles(sy.next + r, goal, nt, tile_chunk(nt), &cnt);   // >= 0: tiles to fold
                        s_cmd[2] = cnt;
                        if (cmd < 0) {
                            if (ld_rlx(sy.done + r) >= goal) { acquire_fence(); cmd = -1; }   // every tile counted
                            else if (ld_rlx(sy.abort) != 0ull || past_limit(t0, limit)) cmd = -3;
                            else cmd = -2;   // tiles still being folded elsewhere
                        }
                        s_cmd[1] = cmd;
                    }
                    __syncthreads();
                    const int cmd = s_cmd[1], cnt = s_cmd[2];
                    if (cmd >= 0) {
                        for (int q = 0; q < cnt; ++q) {
                            refresh_direct_tile<P>(D, pv[r].phi, nullptr, pv[r].tiles[cmd + q]);
                            __syncthreads();
                        }
                        if (tid == 0)
                            __hip_atomic_fetch_add(sy.done + r, (unsigned long long)cnt, __ATOMIC_RELEASE,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    } else if (cmd == -2) {
                        __builtin_amdgcn_s_sleep(2);
                    }
                    __syncthreads();
                    if (cmd == -1) break;
                    if (cmd == -3) { failed = true; break; }
                }
                if (tid == 0) sy.twait[r] += (unsigned long long)((long long)__builtin_amdgcn_s_memrealtime() - tw0);
            }
            if (failed || (st & 1)) break;
            if (!(st & 2)) { failed = true; break; }   // neither finished nor pending: no progress possible
        }
        if (tid == 0) {
            sy.tfin[r] = (unsigned long long)((long long)__builtin_amdgcn_s_memrealtime() - t0);
            if (failed) {
                D.ctl->status = 6;
                __hip_atomic_store(sy.abort, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_fetch_add(sy.fin, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    // ---- helper: tiles of any replica's posted requests, until every chain has finished;
    // the replicas with the longest sweeps first, so that the last chain to finish
    // waits least for its refreshes
    for (;;) {
        if (tid == 0) {
            int rr = -1, tt = -1, ex = 0, cnt = 0;
            for (int k = 0; k < R && tt < 0; ++k) {
                const int r = pv[k].prio;
                const int nt = pv[r].nt;
                const unsigned long long goal = ld_rlx(sy.req + r) * (unsigned long long)nt;
                tt = grab_tiles(sy.next + r, goal, nt, tile_chunk(nt), &cnt);
                if (tt >= 0) rr = r;
            }
            if (tt >= 0) acquire_fence();   // the request's pending list and grid, as posted
            else if (ld_rlx(sy.fin) >= (unsigned long long)R || ld_rlx(sy.abort) != 0ull || past_limit(t0, limit))
                ex = 1;
            s_cmd[0] = rr; s_cmd[1] = tt; s_cmd[2] = ex; s_cmd[3] = cnt;
        }
        __syncthreads();
        const int rr = s_cmd[0], tt = s_cmd[1], ex = s_cmd[2], cnt = s_cmd[3];
        if (ex) return;
        if (tt >= 0) {
            for (int q = 0; q < cnt; ++q) {
                refresh_direct_tile<P>(views[rr], pv[rr].phi, nullptr, pv[rr].tiles[tt + q]);
                __syncthreads();
            }
            if (tid == 0)
                __hip_atomic_fetch_add(sy.done + rr, (unsigned long long)cnt, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __builtin_amdgcn_s_sleep(16);
        }
        __syncthreads();
    }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

// Next sweep: each replica's target moves `add[r]` trials past where it stands.
__global__ void k_batch_targets(const McDev* __restrict__ views, int R, const long long* __restrict__ add) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    McCtl* C = views[r].ctl;
    C->trials_target = C->trials_done + add[r];
}

// Per replica: finished (target reached or stopped), trials done, accepted moves.
__global__ void k_batch_status(const McDev* __restrict__ views, int R, long long* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const McCtl* C = views[r].ctl;
    out[3 * r] = (C->trials_done >= C->trials_target || C->status != 0) ? 1 : 0;
    out[3 * r + 1] = C->trials_done;
    out[3 * r + 2] = C->accepted;
}

__global__ void k_batch_tuning(const McDev* __restrict__ views, int R, int pending_max, long long launch_cap) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) {
        views[r].ctl->pending_max = pending_max;
        views[r].ctl->launch_cap = launch_cap;
    }
}

}  // namespace phb

using namespace phb;

struct phb_batch_s {
    std::vector<phb_mc_s*> mcs;
    int R = 0, P = 5;
    bool lds_pools = false;
    size_t pool_lds = 0, klds = 0, plds = 0, dlds = 0;
    int kblocks = 0, pblocks = 0, dblocks = 0, nk = 0, nd = 0;   // k-space / direct refresh replicas
    DevBuf<McDev> views;
    std::vector<RefreshView> hviews;
    DevBuf<RefreshView> rv_k, rv_d;
    DevBuf<int> kstart, pstart, dstart;
    DevBuf<long long> add, status;
    std::vector<long long> h_status;
    phb_batch_desc d{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long replays = 0, launches = 0;
    // persistent form: one cooperative launch per sweep
    bool persistent = false;
    DevBuf<PersistView> pv;
    DevBuf<unsigned long long> syncw;   // 5R + 2 words (SweepSync)
    int pgrid = 0;                      // workgroups of the launch (R chains, the rest helpers)
    size_t sweep_lds = 0;               // its dynamic LDS
    double limit_s = 30.0;
    ~phb_batch_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

template <int P>
void enqueue_pairs_p(phb_batch_s* B, int pairs) {
    for (int k = 0; k < pairs; ++k) {
        if (B->lds_pools)
            hipLaunchKernelGGL((k_chain_batch<P, unsigned short, true>), dim3(B->R), dim3(MC_THREADS), B->pool_lds,
                               stream(), B->views.p);
        else
            hipLaunchKernelGGL((k_chain_batch<P, int, false>), dim3(B->R), dim3(MC_THREADS), 0, stream(), B->views.p);
        if (B->nd > 0)
            hipLaunchKernelGGL(k_direct_batch<P>, dim3(B->dblocks), dim3(RD_THREADS), B->dlds, stream(), B->rv_d.p,
                               B->dstart.p, B->nd);
        if (B->nk > 0) {
            hipLaunchKernelGGL(k_kspace_batch<P>, dim3(B->kblocks), dim3(512), B->klds, stream(), B->rv_k.p,
                               B->kstart.p, B->nk);
            hipLaunchKernelGGL(k_c2r_batch, dim3(B->pblocks), dim3(512), B->plds, stream(), B->rv_k.p, B->pstart.p,
                               B->nk);
        }
    }
}

void enqueue_pairs(phb_batch_s* B, int pairs) {
    switch (B->P) {
        case 3: enqueue_pairs_p<3>(B, pairs); break;
        case 4: enqueue_pairs_p<4>(B, pairs); break;
        case 5: enqueue_pairs_p<5>(B, pairs); break;
        case 6: enqueue_pairs_p<6>(B, pairs); break;
        case 7: enqueue_pairs_p<7>(B, pairs); break;
        default: enqueue_pairs_p<8>(B, pairs); break;
    }
}

template <int P>
void set_lds_limits_p(phb_batch_s* B) {
    if (B->lds_pools)
        PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_chain_batch<P, unsigned short, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)B->pool_lds));
    PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_kspace_batch<P>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)B->klds));
    PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_c2r_batch), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)B->plds));
    PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_direct_batch<P>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max<size_t>(B->dlds, 1)));
}

void set_lds_limits(phb_batch_s* B) {
    switch (B->P) {
        case 3: set_lds_limits_p<3>(B); break;
        case 4: set_lds_limits_p<4>(B); break;
        case 5: set_lds_limits_p<5>(B); break;
        case 6: set_lds_limits_p<6>(B); break;
        case 7: set_lds_limits_p<7>(B); break;
        default: set_lds_limits_p<8>(B); break;
    }
}

void drop_graph(phb_batch_s* B) {
    if (B->exec) PHB_HIP(hipGraphExecDestroy(B->exec));
    if (B->graph) PHB_HIP(hipGraphDestroy(B->graph));
    B->exec = nullptr;
    B->graph = nullptr;
}

void capture(phb_batch_s* B) {
    drop_graph(B);
    sync();
    PHB_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    enqueue_pairs(B, B->d.graph_pairs);
    PHB_HIP(hipStreamEndCapture(stream(), &B->graph));
    PHB_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
}

template <int P>
const void* sweep_kernel_p(bool lds_pools) {
    return lds_pools ? reinterpret_cast<const void*>(k_sweep_persistent<P, unsigned short, true>)
                     : reinterpret_cast<const void*>(k_sweep_persistent<P, int, false>);
}

const void* sweep_kernel(const phb_batch_s* B) {
    switch (B->P) {
        case 3: return sweep_kernel_p<3>(B->lds_pools);
        case 4: return sweep_kernel_p<4>(B->lds_pools);
        case 5: return sweep_kernel_p<5>(B->lds_pools);
        case 6: return sweep_kernel_p<6>(B->lds_pools);
        case 7: return sweep_kernel_p<7>(B->lds_pools);
        default: return sweep_kernel_p<8>(B->lds_pools);
    }
}

// Empty when the persistent form can run with the current tuning; else the reason.
const char* persistent_blocker(phb_batch_s* B) {
    for (phb_mc_s* M : B->mcs)
        if (!use_direct_refresh(M))
            return "phb_batch_mode: the persistent sweep needs every replica on the direct refresh (grid x extent "
                   "at most 64, pending budget at most 16)";
    const void* f = sweep_kernel(B);
    const size_t lds = std::max(B->lds_pools ? B->pool_lds : (size_t)0, refresh_direct_lds(B->P));
    PHB_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, dev = 0, ncu = 0;
    PHB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, MC_THREADS, lds));
    PHB_HIP(hipGetDevice(&dev));
    PHB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (per_cu * ncu < B->R)
        return "phb_batch_mode: more replicas than workgroups the device holds at once";
    B->sweep_lds = lds;
    B->pgrid = per_cu * ncu;
    return "";
}

void setup_persistent(phb_batch_s* B) {
    std::vector<PersistView> v(B->R);
    std::vector<int> order(B->R);
    for (int r = 0; r < B->R; ++r) {
        phb_mc_s* M = B->mcs[r];
        v[r].phi = M->phi.p;
        v[r].tiles = M->rd_tiles.p;
        v[r].nt = M->rd_ntiles;
        require(M->rd_tiles.p != nullptr && (size_t)M->rd_ntiles <= M->rd_tiles.n,
                "phb_batch_mode: tile list out of step with its count", PHB_ERR_INTERNAL);
        require(M->applied.n >= (size_t)v[r].nt, "phb_batch_mode: refresh gate smaller than the tile count",
                PHB_ERR_INTERNAL);
        order[r] = r;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->mcs[a]->nsites > B->mcs[b]->nsites; });
    for (int k = 0; k < B->R; ++k) v[k].prio = order[k];
    sync();
    B->pv.alloc(B->R);
    B->pv.upload(v.data(), B->R);
    B->syncw.alloc(5 * (size_t)B->R + 2);
    sync();
}

void launch_persistent(phb_batch_s* B) {
    const int R = B->R;
    hipLaunchKernelGGL(k_zero_u64, dim3(ceil_div(5 * R + 2, 256)), dim3(256), 0, stream(), B->syncw.p, 5 * R + 2);
    PHB_HIP(hipGetLastError());
    SweepSync sy;
    sy.req = B->syncw.p;
    sy.next = sy.req + R;
    sy.done = sy.next + R;
    sy.fin = sy.done + R;
    sy.abort = sy.fin + 1;
    sy.tfin = sy.abort + 1;
    sy.twait = sy.tfin + R;
    const McDev* views = B->views.p;
    const PersistView* pv = B->pv.p;
    int Rv = R;
    long long limit = (long long)(B->limit_s * 1e8);   // s_memrealtime runs at 100 MHz
    void* args[] = {&views, &pv, &Rv, &sy, &limit};
    PHB_HIP(hipLaunchCooperativeKernel(sweep_kernel(B), dim3(B->pgrid), dim3(MC_THREADS), args,
                                       (unsigned)B->sweep_lds, stream()));
}

// reads the per-replica status block; true when every replica finished its sweep
bool read_status(phb_batch_s* B) {
    hipLaunchKernelGGL(k_batch_status, dim3(ceil_div(B->R, 64)), dim3(64), 0, stream(), B->views.p, B->R, B->status.p);
    PHB_HIP(hipGetLastError());
    B->status.download(B->h_status.data(), 3 * (size_t)B->R);
    sync();
    for (int r = 0; r < B->R; ++r)
        if (!B->h_status[3 * r]) return false;
    return true;
}

void upload_views(phb_batch_s* B) {
    std::vector<McDev> v(B->R);
    for (int r = 0; r < B->R; ++r) v[r] = B->mcs[r]->dev;
    B->views.upload(v.data(), B->R);
    sync();
}

// one sweep of every replica; returns the replays (or launch groups) it took
long long sweep(phb_batch_s* B) {
    std::vector<long long> add(B->R);
    for (int r = 0; r < B->R; ++r) add[r] = B->mcs[r]->nsites;
    B->add.upload(add.data(), B->R);
    hipLaunchKernelGGL(k_batch_targets, dim3(ceil_div(B->R, 64)), dim3(64), 0, stream(), B->views.p, B->R, B->add.p);
    PHB_HIP(hipGetLastError());
    long long n = 0;
    if (B->persistent) {   // one launch: every replica to its target, refreshes inside
        launch_persistent(B);
        B->launches += 1;
        n = 1;
        const bool all = read_status(B);
        for (int r = 0; r < B->R; ++r)
            require(read_ctl(B->mcs[r]).status != 6, "phb_batch_run: a persistent sweep ran past its time limit",
                    PHB_ERR_INTERNAL);
        require(all, "phb_batch_run: a persistent sweep ended before every replica reached its target",
                PHB_ERR_INTERNAL);
    } else {
        for (;;) {
            if (B->d.use_graph) PHB_HIP(hipGraphLaunch(B->exec, stream()));
            else enqueue_pairs(B, B->d.graph_pairs);
            PHB_HIP(hipGetLastError());
            B->launches += B->d.graph_pairs;
            ++n;
            if (read_status(B)) break;
            require(n < 1000000, "phb_batch_run: a sweep made no progress", PHB_ERR_INTERNAL);
        }
    }
    for (int r = 0; r < B->R; ++r) {
        const McCtl c = read_ctl(B->mcs[r]);
        require(c.status != 2, "phb_batch_run: engine stream rolled back past its kept block", PHB_ERR_INTERNAL);
    }
    // drift reset of replicas past their recompute point (between sweeps, outside the graph)
    for (int r = 0; r < B->R; ++r) {
        phb_mc_s* M = B->mcs[r];
        if (M->next_recompute > 0 && B->h_status[3 * r + 1] >= M->next_recompute) {
            full_recompute(M);
            while (M->next_recompute <= B->h_status[3 * r + 1]) M->next_recompute += M->d.recompute_interval;
        }
    }
    B->replays += n;
    return n;
}

// Splits the replicas between the direct and the k-space refresh launches by
// the rule their serial refresh follows (which depends on pending_max).
void build_refresh_tables(phb_batch_s* B) {
    std::vector<RefreshView> vk, vd;
    std::vector<int> ks{0}, ps{0}, ds{0};
    for (int r = 0; r < B->R; ++r) {
        phb_mc_s* M = B->mcs[r];
        // the replica's current tile list (a new tile shape reallocates it)
        B->hviews[r].tiles = M->rd_tiles.p;
        if (use_direct_refresh(M)) {
            require(M->rd_tiles.p != nullptr && M->rd_ntiles > 0 && (size_t)M->rd_ntiles <= M->rd_tiles.n,
                    "batch refresh tables: tile list out of step with its count", PHB_ERR_INTERNAL);
            vd.push_back(B->hviews[r]);
            ds.push_back(ds.back() + M->rd_ntiles);
        } else {
            vk.push_back(B->hviews[r]);
            ks.push_back(ks.back() + M->pme->z_blocks);
            ps.push_back(ps.back() + M->pme->nz);
        }
    }
    B->nd = (int)vd.size();
    B->nk = (int)vk.size();
    B->dblocks = ds.back();
    B->kblocks = ks.back();
    B->pblocks = ps.back();
    sync();   // the tables may be in use by launches still queued
    B->rv_d.alloc(std::max<size_t>(vd.size(), 1));
    B->rv_k.alloc(std::max<size_t>(vk.size(), 1));
    B->dstart.alloc(ds.size());
    B->kstart.alloc(ks.size());
    B->pstart.alloc(ps.size());
    if (!vd.empty()) B->rv_d.upload(vd.data(), vd.size());
    if (!vk.empty()) B->rv_k.upload(vk.data(), vk.size());
    B->dstart.upload(ds.data(), ds.size());
    B->kstart.upload(ks.data(), ks.size());
    B->pstart.upload(ps.data(), ps.size());
    sync();
}

void apply_tuning(phb_batch_s* B, int W, int pending_max, int pairs, long long launch_cap, bool persistent, int ty = 0,
                  int tz = 0) {
    if (persistent) launch_cap = 0;   // a persistent chain runs to its sweep target
    for (phb_mc_s* M : B->mcs) {
        if (ty > 0) set_refresh_tile(M, ty, tz);
        M->dev.W = W;
        M->d.pending_max = pending_max;
        M->d.spec_window = W;
    }
    upload_views(B);
    if (B->d.use_graph) drop_graph(B);   // the graph holds the old tables
    build_refresh_tables(B);
    hipLaunchKernelGGL(k_batch_tuning, dim3(ceil_div(B->R, 64)), dim3(64), 0, stream(), B->views.p, B->R, pending_max,
                       launch_cap);
    PHB_HIP(hipGetLastError());
    B->d.graph_pairs = pairs;
    B->d.spec_window = W;
    B->d.pending_max = pending_max;
    B->d.launch_cap = launch_cap;
    B->persistent = false;
    if (persistent) {
        const char* why = persistent_blocker(B);
        require(why[0] == 0, why, PHB_ERR_UNSUPPORTED);
        setup_persistent(B);
        B->persistent = true;
    } else if (B->d.use_graph) {
        capture(B);
    }
    B->d.persistent = B->persistent ? 1 : 0;
    sync();
}

}  // namespace

extern "C" int phb_batch_create(const phb_mc* mcs, int n, const phb_batch_desc* d, phb_batch* out) {
    return guard([&] {
        require(mcs && n > 0 && out, "phb_batch_create: bad arguments");
        auto B = std::make_unique<phb_batch_s>();
        B->d = d ? *d : phb_batch_desc{};
        if (B->d.graph_pairs <= 0) B->d.graph_pairs = 4;
        B->R = n;
        B->mcs.assign(mcs, mcs + n);
        for (phb_mc_s* M : B->mcs) require(M != nullptr, "phb_batch_create: null replica");
        B->P = B->mcs[0]->dev.P;
        bool lds = true;
        std::vector<RefreshView> rv(n);
        for (int r = 0; r < n; ++r) {
            phb_mc_s* M = B->mcs[r];
            phb_pme_s* Pm = M->pme;
            require(M->dev.P == B->P, "phb_batch_create: every replica needs the same spline order", PHB_ERR_UNSUPPORTED);
            require(M->d.precision == PHB_FP64, "phb_batch_create: batched replicas run the FP64 form",
                    PHB_ERR_UNSUPPORTED);
            for (int q = 0; q < r; ++q) require(B->mcs[q] != M, "phb_batch_create: a replica appears twice");
            require(Pm->z_tx <= 16, "phb_batch_create: z tile wider than 16", PHB_ERR_INTERNAL);
            require(!Pm->split, "phb_batch_create: every replica's xy plane must fit one workgroup (the batched "
                                "refresh transforms whole planes)", PHB_ERR_UNSUPPORTED);
            lds = lds && M->lds_pools;
            B->pool_lds = std::max(B->pool_lds, M->pool_lds_bytes);
            B->klds = std::max(B->klds, Pm->z_lds(sizeof(cplx<double>)) + (size_t)REFRESH_BATCH * Pm->nz * sizeof(cplx<double>));
            B->plds = std::max(B->plds, Pm->plane_lds(sizeof(cplx<double>)));
            if (refresh_direct_allowed(M)) B->dlds = std::max(B->dlds, refresh_direct_lds(B->P));
        }
        B->lds_pools = lds && B->pool_lds <= 96 * 1024;
        require(B->klds <= 160 * 1024 && B->plds <= 160 * 1024, "phb_batch_create: refresh tiles exceed LDS",
                PHB_ERR_UNSUPPORTED);
        B->views.alloc(n);
        upload_views(B.get());
        for (int r = 0; r < n; ++r) {
            phb_mc_s* M = B->mcs[r];
            phb_pme_s* Pm = M->pme;
            RefreshView& w = rv[r];
            w.D = B->views.p + r;
            w.kg = Pm->kg64.p; w.twx = Pm->twx64.p; w.twy = Pm->twy64.p; w.twz = Pm->twz64.p; w.eta = Pm->conv_eta64();
            w.px = Pm->px; w.py = Pm->py; w.pz = Pm->pz;
            w.tx = Pm->z_tx; w.nx = Pm->nx; w.ny = Pm->ny;
            w.phi = M->phi.p;
            w.gate.pending_n = &M->ctl.p->pending_n;
            w.gate.seq = &M->ctl.p->pend_seq;
            w.gate.applied = M->applied.p;
            w.tiles = M->rd_tiles.p;
            require(Pm->kg64.p != nullptr, "phb_batch_create: replicas need FP64 plans", PHB_ERR_UNSUPPORTED);
        }
        // replicas sharing one PME plan would share its k-space buffer
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < r; ++q)
                require(B->mcs[q]->pme != B->mcs[r]->pme, "phb_batch_create: replicas must not share a PME plan");
        B->hviews = rv;
        build_refresh_tables(B.get());
        B->add.alloc(n);
        B->status.alloc(3 * (size_t)n);
        B->h_status.assign(3 * (size_t)n, 0);
        set_lds_limits(B.get());
        B->d.spec_window = B->mcs[0]->dev.W;
        B->d.pending_max = B->mcs[0]->d.pending_max;
        // every replica's potential grid must hold its accepted changes before the first launch
        for (phb_mc_s* M : B->mcs) refresh(M);
        const bool want_persistent = B->d.persistent != 0;
        B->d.persistent = 0;
        if (want_persistent)
            apply_tuning(B.get(), B->d.spec_window, B->d.pending_max, B->d.graph_pairs, 0, true);
        else if (B->d.use_graph)
            capture(B.get());
        sync();
        *out = B.release();
        return PHB_OK;
    });
}

extern "C" int phb_batch_destroy(phb_batch b) {
    return guard([&] { delete b; return PHB_OK; });
}

extern "C" int phb_batch_run(phb_batch B, int sweeps, phb_batch_stats* st) {
    return guard([&] {
        require(B && sweeps >= 0, "phb_batch_run: bad arguments");
        const auto t0 = std::chrono::steady_clock::now();
        const long long l0 = B->launches, r0 = B->replays;
        read_status(B);
        long long a0 = 0;
        for (int r = 0; r < B->R; ++r) a0 += B->h_status[3 * r + 1];
        for (int s = 0; s < sweeps; ++s) sweep(B);
        sync();
        if (st) {
            read_status(B);
            st->attempted = 0; st->accepted = 0;
            for (int r = 0; r < B->R; ++r) {
                st->attempted += B->h_status[3 * r + 1];
                st->accepted += B->h_status[3 * r + 2];
            }
            st->launches = B->launches - l0;
            st->replays = B->replays - r0;
            st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            st->trials_per_s = st->ms > 0.0 ? (double)(st->attempted - a0) / (st->ms * 1e-3) : 0.0;
            st->spec_window = B->d.spec_window;
            st->pending_max = B->d.pending_max;
            st->graph_pairs = B->d.graph_pairs;
            st->launch_cap = B->d.launch_cap;
            st->persistent = B->persistent ? 1 : 0;
            st->refresh_ty = B->mcs.empty() ? 0 : B->mcs[0]->dev.rd_ty;
            st->refresh_tz = B->mcs.empty() ? 0 : B->mcs[0]->dev.rd_tz;
        }
        return PHB_OK;
    });
}

extern "C" int phb_batch_tune(phb_batch B, int spec_window, int pending_max, int graph_pairs, long long launch_cap) {
    return guard([&] {
        require(B, "phb_batch_tune: null batch");
        require(spec_window == 4 || spec_window == 8 || spec_window == 16, "phb_batch_tune: window must be 4, 8 or 16");
        require(pending_max >= 2 && pending_max <= MC_KMAX && pending_max % 2 == 0,
                "phb_batch_tune: pending budget must be even, in [2, 64]");
        require(graph_pairs >= 1 && launch_cap >= 0, "phb_batch_tune: bad graph length or launch cap");
        apply_tuning(B, spec_window, pending_max, graph_pairs, launch_cap, B->persistent);
        return PHB_OK;
    });
}

extern "C" int phb_batch_replica_times(phb_batch B, double* run_ms, double* refresh_ms) {
    return guard([&] {
        require(B && run_ms && refresh_ms, "phb_batch_replica_times: bad arguments");
        require(B->persistent && B->syncw.p != nullptr, "phb_batch_replica_times: no persistent sweep has run",
                PHB_ERR_STATE);
        const int R = B->R;
        std::vector<unsigned long long> w(5 * (size_t)R + 2);
        sync();
        B->syncw.download(w.data(), w.size());
        sync();
        for (int r = 0; r < R; ++r) {   // 100 MHz ticks
            run_ms[r] = (double)w[3 * (size_t)R + 2 + r] * 1e-5;
            refresh_ms[r] = (double)w[4 * (size_t)R + 2 + r] * 1e-5;
        }
        return PHB_OK;
    });
}

extern "C" int phb_batch_set_refresh_tile(phb_batch B, int ty, int tz) {
    return guard([&] {
        require(B, "phb_batch_set_refresh_tile: null batch");
        require(refresh_tile_ok(ty, tz), "phb_batch_set_refresh_tile: 8 or 4 rows along y by 4 or 2 planes along z");
        apply_tuning(B, B->d.spec_window, B->d.pending_max, B->d.graph_pairs, B->d.launch_cap, B->persistent, ty, tz);
        return PHB_OK;
    });
}

extern "C" int phb_batch_mode(phb_batch B, int persistent) {
    return guard([&] {
        require(B, "phb_batch_mode: null batch");
        apply_tuning(B, B->d.spec_window, B->d.pending_max, B->d.graph_pairs, B->d.launch_cap, persistent != 0);
        return PHB_OK;
    });
}

extern "C" int phb_batch_autotune(phb_batch B, int sweeps_per_candidate, phb_batch_stats* best) {
    return guard([&] {
        require(B && sweeps_per_candidate > 0, "phb_batch_autotune: bad arguments");
        // The replicas move towards equilibrium while the candidates run (fewer accepted moves,
        // faster sweeps), so later candidates would look faster: warm-up sweeps first, and the
        // best graph shape and the best persistent shape timed again back to back before the
        // form is chosen.
        // stage 1: window x budget x graph length (graph form), window x budget (persistent form,
        // where every replica qualifies); stage 2: the per-launch trial cap on the best graph
        // shape; stage 3: the two forms' best shapes again; stage 4: the refresh tile shape.
        struct Cand { int W, pm, pairs; long long cap; bool pers; int ty, tz; };
        const int spc = sweeps_per_candidate;
        auto rate_of = [&](const Cand& c, int sweeps) {
            apply_tuning(B, c.W, c.pm, c.pairs, c.cap, c.pers, c.ty, c.tz);
            read_status(B);
            long long t0 = 0;
            for (int r = 0; r < B->R; ++r) t0 += B->h_status[3 * r + 1];
            const auto c0 = std::chrono::steady_clock::now();
            for (int s = 0; s < sweeps; ++s) sweep(B);
            sync();
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
            read_status(B);
            long long t1 = 0;
            for (int r = 0; r < B->R; ++r) t1 += B->h_status[3 * r + 1];
            return (double)(t1 - t0) / std::max(dt, 1e-9);
        };
        const int ty0 = B->mcs.empty() ? RD_TY : B->mcs[0]->dev.rd_ty, tz0 = B->mcs.empty() ? RD_TZ : B->mcs[0]->dev.rd_tz;
        for (int s = 0; s < 2 * spc; ++s) sweep(B);   // warm-up at the current shape
        sync();
        const bool pers_ok = B->persistent || [&] {
            for (phb_mc_s* M : B->mcs)
                if (!refresh_direct_allowed(M) || M->d.refresh == PHB_REFRESH_KSPACE) return false;
            return true;
        }();
        Cand bg{8, 8, 2, 0, false, ty0, tz0}, bp{8, 8, 2, 0, true, ty0, tz0};
        double rg = -1.0, rp = -1.0;
        for (int W : {4, 8, 16})
            for (int pm : {2, 4, 8, 16})   // 2: a refresh after every accepted move
                for (int pairs : {2, 4, 8}) {
                    const Cand c{W, pm, pairs, 0, false, ty0, tz0};
                    const double r = rate_of(c, spc);
                    if (r > rg) { rg = r; bg = c; }
                }
        if (pers_ok)
            for (int W : {4, 8, 16})
                for (int pm : {2, 4, 8, 16}) {
                    const Cand c{W, pm, B->d.graph_pairs, 0, true, ty0, tz0};
                    const double r = rate_of(c, spc);
                    if (r > rp) { rp = r; bp = c; }
                }
        for (long long cap : {16LL, 32LL, 64LL, 128LL, 256LL}) {
            Cand c = bg;
            c.cap = cap;
            const double r = rate_of(c, spc);
            if (r > rg) { rg = r; bg = c; }
        }
        Cand bc = bg;
        double best_rate = rate_of(bg, 2 * spc);
        if (pers_ok) {
            const double r = rate_of(bp, 2 * spc);
            if (r > best_rate) { best_rate = r; bc = bp; }
        }
        const bool tiles_ok = [&] {   // the tile shape matters where a replica refreshes directly
            for (phb_mc_s* M : B->mcs)
                if (use_direct_refresh(M)) return true;
            return false;
        }();
        if (tiles_ok) {
            const int shapes[4][2] = {{8, 4}, {8, 2}, {4, 4}, {4, 2}};
            double rt = -1.0;
            Cand bt = bc;
            for (const auto& t : shapes) {
                Cand c = bc;
                c.ty = t[0];
                c.tz = t[1];
                const double r = rate_of(c, spc);
                if (r > rt) { rt = r; bt = c; }
            }
            bc = bt;
            best_rate = rt;
        }
        apply_tuning(B, bc.W, bc.pm, bc.pairs, bc.cap, bc.pers, bc.ty, bc.tz);
        if (best) {
            *best = phb_batch_stats{};
            best->spec_window = bc.W;
            best->pending_max = bc.pm;
            best->graph_pairs = bc.pairs;
            best->launch_cap = bc.cap;
            best->persistent = bc.pers ? 1 : 0;
            best->refresh_ty = bc.ty;
            best->refresh_tz = bc.tz;
            best->trials_per_s = best_rate;
        }
        return PHB_OK;
    });
}
