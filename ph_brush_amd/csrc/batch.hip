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
    chain_body<P, PoolT, LDS_POOLS>(views[blockIdx.x]);
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
    refresh_direct_tile<P>(*w.D, w.phi, nullptr, blockIdx.x - start[r]);
}

__global__ __launch_bounds__(512) void k_c2r_batch(const RefreshView* __restrict__ v, const int* __restrict__ start,
                                                   int R) {
    const int r = replica_of(start, R, blockIdx.x);
    const RefreshView& w = v[r];
    fft_xy_c2r_plane<double>(w.kg, w.phi, w.nx, w.ny, w.px, w.py, w.twx, w.twy, 1, w.gate, blockIdx.x - start[r],
                             start[r + 1] - start[r]);
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
    for (;;) {
        if (B->d.use_graph) PHB_HIP(hipGraphLaunch(B->exec, stream()));
        else enqueue_pairs(B, B->d.graph_pairs);
        PHB_HIP(hipGetLastError());
        B->launches += B->d.graph_pairs;
        ++n;
        if (read_status(B)) break;
        require(n < 1000000, "phb_batch_run: a sweep made no progress", PHB_ERR_INTERNAL);
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
        if (use_direct_refresh(M)) {
            vd.push_back(B->hviews[r]);
            ds.push_back(ds.back() + refresh_direct_blocks(M->pme->ny, M->pme->nz));
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

void apply_tuning(phb_batch_s* B, int W, int pending_max, int pairs, long long launch_cap) {
    for (phb_mc_s* M : B->mcs) {
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
    if (B->d.use_graph) capture(B);
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
        if (B->d.use_graph) capture(B.get());
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
        apply_tuning(B, spec_window, pending_max, graph_pairs, launch_cap);
        return PHB_OK;
    });
}

extern "C" int phb_batch_autotune(phb_batch B, int sweeps_per_candidate, phb_batch_stats* best) {
    return guard([&] {
        require(B && sweeps_per_candidate > 0, "phb_batch_autotune: bad arguments");
        // stage 1: window x budget x graph length without a launch cap; stage 2: the
        // per-launch trial cap (a launch waits for its slowest replica) with the best of stage 1
        struct Cand { int W, pm, pairs; long long cap; };
        std::vector<Cand> cands;
        for (int W : {4, 8, 16})
            for (int pm : {4, 8, 16})
                for (int pairs : {2, 4, 8}) cands.push_back({W, pm, pairs, 0});
        double best_rate = -1.0;
        Cand bc = cands[0];
        const size_t n1 = cands.size();
        for (size_t ci = 0; ci < cands.size(); ++ci) {
            const Cand c = cands[ci];
            apply_tuning(B, c.W, c.pm, c.pairs, c.cap);
            read_status(B);
            long long t0 = 0;
            for (int r = 0; r < B->R; ++r) t0 += B->h_status[3 * r + 1];
            const auto c0 = std::chrono::steady_clock::now();
            for (int s = 0; s < sweeps_per_candidate; ++s) sweep(B);
            sync();
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
            read_status(B);
            long long t1 = 0;
            for (int r = 0; r < B->R; ++r) t1 += B->h_status[3 * r + 1];
            const double rate = (double)(t1 - t0) / std::max(dt, 1e-9);
            if (rate > best_rate) { best_rate = rate; bc = c; }
            if (ci + 1 == n1)   // stage 2: the caps, on the best shape of stage 1
                for (long long cap : {16LL, 32LL, 64LL, 128LL, 256LL}) cands.push_back({bc.W, bc.pm, bc.pairs, cap});
        }
        apply_tuning(B, bc.W, bc.pm, bc.pairs, bc.cap);
        if (best) {
            *best = phb_batch_stats{};
            best->spec_window = bc.W;
            best->pending_max = bc.pm;
            best->graph_pairs = bc.pairs;
            best->launch_cap = bc.cap;
            best->trials_per_s = best_rate;
        }
        return PHB_OK;
    });
}
