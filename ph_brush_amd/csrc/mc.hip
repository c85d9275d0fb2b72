// mc.hip -- device-resident protonation / ion-swap Monte Carlo.
//
// Replaces the trial loop of phase_gc_mu_control (MC/src/mc.cu:379-503) at a
// fixed mu_eff.  The reference recomputes the total energy (8+ launches, 9+
// host syncs) per trial; here one persistent workgroup runs the chain:
//
//   wave 0 (all 64 lanes redundantly, lane 0 writes):
//       device MT19937 (wave-cooperative twist), libstdc++ Lemire draws and
//       generate_canonical<double,53>, direction and pool picks, Metropolis-
//       Hastings, pool bookkeeping;
//   all 16 waves: the exact Delta-E of the move
//       - LJ and real-space Ewald over the two moved atoms' full neighbour rows
//         (precomputed elF*erfc(alpha r)/r and r^2 per entry);
//       - reciprocal space from the quadratic form E = 0.5 Q^T G Q on the mesh:
//           dE = sum_a dq_a W_a.phi + sum_{a,b} w_ab W_a^T G W_b
//         where phi = G Q is a maintained potential grid, G is applied through
//         the mesh kernel g = IFFT(eta) with separable stencil correlations
//         (9x9x9 lookups per atom pair at order 5), and moves accepted since
//         the last refresh of phi enter as extra pairs;
//       - self, net-charge and slab terms from O(1) updates of Q, sum q^2,
//         Mz and sum q z^2.
//
// After pending_max accepted atom changes the chain kernel exits; the refresh
// scatters the pending charge changes, runs the LDS FFT, multiplies by eta
// and accumulates IFFT into phi.  Every recompute_interval trials the host
// recomputes everything from scratch (drift reset), like the reference's full
// recompute.
#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <random>
#include <sstream>
#include <utility>
#include <vector>

#include "mc_host.hpp"
#include "mc_kernels.hpp"

namespace phb {

template <int P, typename PoolT, bool LDS_POOLS, typename T = double>
__global__ __launch_bounds__(MC_THREADS) void k_mc_chain(McDev D) {
    chain_body<P, PoolT, LDS_POOLS, T>(D);
}

// Delta-E of candidate moves, one block per move, no acceptance.
template <int P, typename T = double>
__global__ __launch_bounds__(MC_THREADS) void k_mc_probe(McDev D, int n, const int* __restrict__ mi,
                                                         const int* __restrict__ mj, const int* __restrict__ mf,
                                                         double* __restrict__ out) {
    __shared__ TrialT<P> s_t;
    __shared__ PendT<P> s_pd;
    __shared__ double s_wsum[16];
    const McCtl* C = D.ctl;
    const int m = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (m >= n) return;
    const int pend_n = pending_unfolded(D);
    for (int k = wave; k < pend_n; k += 16) stage_pending<P>(D, s_pd, k, C->pend_atom[k], C->pend_dq[k], lane, 64);
    if (tid == 0) stage_trial<P>(D, s_t, mi[m], mj[m], mf[m]);
    __syncthreads();
    stage_trial_data<P>(D, s_t, tid, blockDim.x);
    __syncthreads();
    double part = wave_sum(trial_partial<P, T>(D, s_t, s_pd, pend_n, tid, blockDim.x));
    if (lane == 0) s_wsum[wave] = part;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += s_wsum[w];
        out[m] = s + correction_delta<P>(D, s_t, C->Q, C->Q2, C->Mz, C->Qz2);
    }
}

// ------------------------------------------------------------ refresh
// Folds the pending charge changes into phi without a charge grid: their
// structure factor is separable per atom,
//   dQ^(k) = sum_p dq_p Fx_p(kx) Fy_p(ky) Fz_p(kz),  F(k) = sum_u w[u] e^{-2 pi i k (s+u)/n},
// so each z-pencil tile builds eta * dQ^ in LDS and runs the inverse z pass;
// k_fft_xy_c2r then accumulates the inverse plane pass into phi.  Both read
// the same gate (pending and not yet folded); the plane pass's last workgroup
// marks the changes folded, so no workgroup of either pass sees the flag move.
template <int P>
__global__ __launch_bounds__(512) void k_refresh_kspace(McDev D, cplx<double>* __restrict__ kg, const cplx<double>* __restrict__ twx,
                                 const cplx<double>* __restrict__ twy, const cplx<double>* __restrict__ twz,
                                 const double* __restrict__ eta, FftPlan pz, int tx) {
    refresh_kspace_tile<P>(D, kg, twx, twy, twz, eta, pz, tx, blockIdx.x);
}

// Direct refresh (refresh_direct_tile): one workgroup per listed (y, z) tile of full x rows,
// one instantiation per tile shape.
template <int P, int TY, int TZ>
__global__ __launch_bounds__(RD_THREADS) void k_refresh_direct(McDev D, double* __restrict__ phi,
                                                               float* __restrict__ phi32, const int* __restrict__ tiles) {
    refresh_direct_tile_fixed<P, TY, TZ>(D, phi, phi32, tiles[blockIdx.x]);
}

template <int P, int TY, int TZ>
void launch_refresh_direct(phb_mc_s* M) {
    static bool lds_set = false;   // per instantiation
    if (!lds_set) {
        PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_refresh_direct<P, TY, TZ>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)refresh_direct_lds(P)));
        lds_set = true;
    }
    hipLaunchKernelGGL((k_refresh_direct<P, TY, TZ>), dim3(M->rd_ntiles), dim3(RD_THREADS), refresh_direct_lds(P),
                       stream(), M->dev, M->phi.p, M->d.precision == PHB_FP32 ? M->phi32.p : nullptr, M->rd_tiles.p);
}

__global__ void k_reset_pending(McCtl* C) {
    if (threadIdx.x == 0) C->pending_n = 0;
}

__global__ void k_set_target(McCtl* C, long long target) {
    if (threadIdx.x == 0) C->trials_target = target;
}

// After a full recompute: res = [lj, real, recip, Q, Q2, Mz, Qz2]
__global__ void k_finish_recompute(McDev D, const double* __restrict__ res, double e_self_coef_unused) {
    (void)e_self_coef_unused;
    if (threadIdx.x != 0) return;
    McCtl* C = D.ctl;
    const double Q = res[3], Q2 = res[4], Mz = res[5], Qz2 = res[6];
    const double e_self = -D.self_coef * Q2;
    const double e_q = D.alpha > 0.0 ? -(M_PI * D.elF) * (Q * Q) / (2.0 * D.vol * D.alpha * D.alpha) : 0.0;
    const double E = res[0] + (res[1] + res[2] + e_self + e_q + slab_energy(D, Mz, Q, Qz2));
    if (C->E_recomputed == C->E_recomputed)  // not the first recompute
        C->max_drift = fmax(C->max_drift, fabs(C->E - E));
    C->E = E; C->E_recomputed = E;
    C->Q = Q; C->Q2 = Q2; C->Mz = Mz; C->Qz2 = Qz2;
    C->pending_n = 0;
}

// ------------------------------------------------------------ setup kernels
// Full rows from the half list, one wave per half row (lanes over its entries).
// Row i of the full list holds i's own half row first, then the atoms whose half
// rows hold i (in arrival order; rows are sorted afterwards).
__global__ __launch_bounds__(256) void k_full_degree(int N, const int* __restrict__ hh, const int* __restrict__ hl,
                                                     int* __restrict__ deg) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int s = hh[i], e = hh[i + 1];
    if (lane == 0) atomicAdd(&deg[i], e - s);
    for (int k = s + lane; k < e; k += 64) atomicAdd(&deg[hl[k]], 1);
}

__global__ __launch_bounds__(256) void k_full_fill(int N, const int* __restrict__ hh, const int* __restrict__ hl,
                                                   const int* __restrict__ fh, int* __restrict__ cursor,
                                                   int* __restrict__ tmp) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int s = hh[i], e = hh[i + 1];
    for (int k = s + lane; k < e; k += 64) {
        const int j = hl[k];
        tmp[fh[i] + (k - s)] = j;
        tmp[fh[j] + (hh[j + 1] - hh[j]) + atomicAdd(&cursor[j], 1)] = i;
    }
}

__global__ __launch_bounds__(256) void k_full_coeffs(int N, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, double Lx, double Ly, double Lz, double elF, double alpha,
                              const int* __restrict__ fh, const int* __restrict__ fn, double* __restrict__ kreal,
                              double* __restrict__ r2o) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);   // one wave per row
    if (i >= N) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    for (int e = fh[i] + lane; e < fh[i + 1]; e += 64) {
        const int j = fn[e];
        const double dx = min_image_exact(xi - x[j], Lx), dy = min_image_exact(yi - y[j], Ly),
                     dz = min_image_exact(zi - z[j], Lz);
        const double r2 = r2_exact(dx, dy, dz), r = sqrt(r2);
        r2o[e] = r2;
        kreal[e] = elF * (erfc(alpha * r) / r);
    }
}

__global__ void k_stencils(int N, const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, GridGeom g, int* __restrict__ sbase, double* __restrict__ sw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int i0[3];
    double w[3][8];
    stencil_of(x[i], y[i], z[i], g, i0, w);
    const int P = g.order, off = stencil_offset(P);
    sbase[3 * i + 0] = wrap(i0[0] - off, g.nx);
    sbase[3 * i + 1] = wrap(i0[1] - off, g.ny);
    sbase[3 * i + 2] = wrap(i0[2] - off, g.nz);
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k < P; ++k) sw[(size_t)i * 3 * P + ax * P + k] = w[ax][k];
}

// W_a^T G W_a for every atom: correlation of the stencil with itself against g.
__global__ void k_self_terms(int N, int P, int nx, int ny, int nz, const double* __restrict__ sw,
                             const double* __restrict__ gk, double* __restrict__ selfw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int CW = 2 * P - 1;
    double C[3][MC_CW];
    for (int ax = 0; ax < 3; ++ax) {
        const double* w = sw + (size_t)i * 3 * P + ax * P;
        for (int dd = 0; dd < CW; ++dd) {
            const int d = dd - (P - 1);
            double c = 0.0;
            for (int u = max(0, d); u < min(P, P + d); ++u) c += w[u] * w[u - d];
            C[ax][dd] = c;
        }
    }
    double s = 0.0;
    for (int dz = 0; dz < CW; ++dz)
        for (int dy = 0; dy < CW; ++dy) {
            const double* row = gk + ((size_t)wrap(dz - (P - 1), nz) * ny + wrap(dy - (P - 1), ny)) * nx;
            double sx = 0.0;
            for (int dx = 0; dx < CW; ++dx) sx += C[0][dx] * row[wrap(dx - (P - 1), nx)];
            s += C[2][dz] * C[1][dy] * sx;
        }
    selfw[i] = s;
}

__global__ void k_fold_g(int nx, int ny, int nz, const double* __restrict__ gk, double* __restrict__ gf) {
    const int fnx = nx / 2 + 1, fny = ny / 2 + 1, fnz = nz / 2 + 1;
    const long long tot = (long long)fnx * fny * fnz;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < tot; k += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(k % fnx), y = (int)((k / fnx) % fny), z = (int)(k / ((long long)fnx * fny));
        gf[k] = gk[((size_t)z * ny + y) * nx + x];
    }
}

__global__ void k_cast_f32(const double* __restrict__ in, float* __restrict__ out, long long n) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        out[k] = (float)in[k];
}

__global__ void k_delta_grid(double* __restrict__ g) {
    if (threadIdx.x == 0 && blockIdx.x == 0) g[0] = 1.0;
}

}  // namespace phb

using namespace phb;

namespace phb {

int launch_batch(const McCtl& c, long long remaining, int pending_max) {
    // a launch ends after pending_max / 2 accepted moves or at the target: expected
    // launches = accepted moves expected in the remaining trials / moves per launch, + 1
    const double acc = c.trials_done > 0 ? (double)c.accepted / (double)c.trials_done : 0.05;
    const double per = std::max(1, pending_max / 2);
    const double n = std::ceil(std::max(acc, 1e-3) * (double)remaining * 1.25 / per) + 1.0;
    return (int)std::max(1.0, std::min(n, 64.0));
}

void cast_f32(const double* in, float* out, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cast_f32, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, stream(),
                       in, out, n);
    PHB_HIP(hipGetLastError());
}

// the FP32 form reads an FP32 copy of the potential grid
void sync_phi32(phb_mc_s* M) {
    if (M->d.precision == PHB_FP32) cast_f32(M->phi.p, M->phi32.p, (long long)M->phi.n);
}

long long build_full_rows(int N, const phb_csr_s* half, DevBuf<int>& head, DevBuf<int>& nbr) {
    hipStream_t st = stream();
    DevBuf<int> deg(N), cursor(N);
    PHB_HIP(hipMemsetAsync(deg.p, 0, N * sizeof(int), st));
    PHB_HIP(hipMemsetAsync(cursor.p, 0, N * sizeof(int), st));
    head.alloc(N + 1);
    hipLaunchKernelGGL(k_full_degree, dim3(ceil_div(N, 4)), dim3(256), 0, st, N, half->head.p, half->list.p, deg.p);
    exclusive_scan(deg.p, head.p, N);
    const long long nfull = 2 * half->nnz;
    nbr.alloc(std::max<long long>(nfull, 1));
    if (nfull > 0) {
        DevBuf<int> tmp(nfull);
        hipLaunchKernelGGL(k_full_fill, dim3(ceil_div(N, 4)), dim3(256), 0, st, N, half->head.p, half->list.p, head.p,
                           cursor.p, tmp.p);
        sort_rows(N, head.p, tmp.p, nbr.p);
        PHB_HIP(hipGetLastError());
        sync();
    }
    return nfull;
}

McCtl read_ctl(phb_mc_s* M) {
    McCtl c;
    PHB_HIP(hipMemcpyAsync(&c, M->ctl.p, sizeof(McCtl), hipMemcpyDeviceToHost, stream()));
    sync();
    if (M->profiling) M->collect();
    return c;
}

void full_recompute(phb_mc_s* M) {
    phb_pme_s* P = M->pme;
    size_t e0 = M->ev_used;
    if (M->profiling) M->event();
    PmeKernels::launch_spread(P, M->atoms);
    PmeKernels::launch_fft_buf(P, P->grid64.p, M->phi.p, 3, true, false);
    double* res = P->result.p;
    reduce_partials(P->partials.p, P->z_blocks, 1, 1, res + 2);
    const int nbp = pair_blocks(M->atoms->n);
    PmeKernels::launch_pair(M->atoms, M->params, M->neigh, &M->box, 1 | 4, P->alpha, P->elF, PHB_FP64, P->partials.p,
                            nbp);
    reduce_partials(P->partials.p, nbp, 2, 2, res + 0);
    PmeKernels::launch_moments(M->atoms, P->partials.p, nbp);
    reduce_partials(P->partials.p, nbp, 4, 4, res + 3);
    hipLaunchKernelGGL(k_finish_recompute, dim3(1), dim3(64), 0, stream(), M->dev, res, 0.0);
    PHB_HIP(hipGetLastError());
    sync_phi32(M);
    if (M->profiling) { M->event(); M->marks.push_back({e0, 1}); }
    ++M->recomputes;
}

// Direct refresh tiles of the current shape: the potential is read only at atoms'
// stencil points (trial atoms, W . phi, the exact stage), and positions are fixed,
// so a tile that no atom's stencil reaches (the vacuum of a slab geometry) is never
// read and is left as is.  Tile 0 is always listed: its gate entry tells the chains
// that a refresh has run.
void build_rd_tiles(phb_mc_s* M) {
    const phb_pme_s* pme = M->pme;
    const int N = M->atoms->n, P = pme->order, ty = M->dev.rd_ty, tz = M->dev.rd_tz;
    const int nty = (pme->ny + ty - 1) / ty, ntz = (pme->nz + tz - 1) / tz;
    std::vector<int> sb(3 * (size_t)N);
    if (N > 0) M->sbase.download(sb.data(), 3 * (size_t)N);
    sync();
    std::vector<char> need((size_t)nty * ntz, 0);
    need[0] = 1;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < P; ++c) {
            const int tzi = ((sb[3 * (size_t)i + 2] + c) % pme->nz) / tz;
            for (int b = 0; b < P; ++b) need[(size_t)tzi * nty + ((sb[3 * (size_t)i + 1] + b) % pme->ny) / ty] = 1;
        }
    std::vector<int> tiles;
    for (int t = 0; t < nty * ntz; ++t)
        if (need[t]) tiles.push_back(t);
    // every listed tile has its gate entry (sized for the smallest tiles)
    require((size_t)nty * ntz <= M->applied.n, "refresh tiles beyond the gate array", PHB_ERR_INTERNAL);
    M->rd_ntiles = (int)tiles.size();
    M->rd_tiles.alloc(tiles.size());
    M->rd_tiles.upload(tiles.data(), tiles.size());
    sync();
}

bool refresh_direct_allowed(const phb_mc_s* M) {
    return M->pme->nx <= RD_NXMAX && refresh_direct_lds(M->pme->order) <= 64 * 1024;
}

bool use_direct_refresh(const phb_mc_s* M) {
    if (M->d.refresh == PHB_REFRESH_KSPACE) return false;
    if (M->d.refresh == PHB_REFRESH_DIRECT) return true;
    return refresh_direct_allowed(M) && M->d.pending_max <= 16;
}

template <int P>
void refresh_p(phb_mc_s* M) {
    phb_pme_s* Pm = M->pme;
    if (use_direct_refresh(M)) {
        const int ty = M->dev.rd_ty, tz = M->dev.rd_tz;
        if (ty == 8 && tz == 4) launch_refresh_direct<P, 8, 4>(M);
        else if (ty == 8) launch_refresh_direct<P, 8, 2>(M);
        else if (tz == 4) launch_refresh_direct<P, 4, 4>(M);
        else launch_refresh_direct<P, 4, 2>(M);
        PHB_HIP(hipGetLastError());
        ++M->refreshes;
        return;
    }
    require(Pm->z_tx <= 16, "refresh: z tile wider than 16", PHB_ERR_INTERNAL);
    const size_t lds = Pm->z_lds(sizeof(cplx<double>)) + (size_t)REFRESH_BATCH * Pm->nz * sizeof(cplx<double>);
    require(lds <= 160 * 1024, "refresh: z pencil and structure factors exceed LDS", PHB_ERR_UNSUPPORTED);
    require(M->applied.n >= (size_t)Pm->plane_gate_blocks(),
            "refresh: the plan's plane transform changed after the MC object was created", PHB_ERR_STATE);
    hipLaunchKernelGGL(k_refresh_kspace<P>, dim3(Pm->z_blocks), dim3(Pm->z_threads), lds, stream(), M->dev, Pm->kg64.p,
                       Pm->twx64.p, Pm->twy64.p, Pm->twz64.p, Pm->conv_eta64(), Pm->pz, Pm->z_tx);
    RefreshGate g;
    g.pending_n = &M->ctl.p->pending_n;
    g.seq = &M->ctl.p->pend_seq;
    g.applied = M->applied.p;
    PmeKernels::launch_c2r_buf(Pm, M->phi.p, true, g);
    PHB_HIP(hipGetLastError());
    sync_phi32(M);
    ++M->refreshes;
}

void refresh(phb_mc_s* M) {
    switch (M->dev.P) {
        case 3: refresh_p<3>(M); break;
        case 4: refresh_p<4>(M); break;
        case 5: refresh_p<5>(M); break;
        case 6: refresh_p<6>(M); break;
        case 7: refresh_p<7>(M); break;
        default: refresh_p<8>(M); break;
    }
}

template <int P, typename T = double>
void launch_chain_p(phb_mc_s* M) {
    if (M->lds_pools) {
        auto kern = k_mc_chain<P, unsigned short, true, T>;
        PHB_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)M->pool_lds_bytes));
        hipLaunchKernelGGL(kern, dim3(1), dim3(MC_THREADS), M->pool_lds_bytes, stream(), M->dev);
    } else {
        hipLaunchKernelGGL((k_mc_chain<P, int, false, T>), dim3(1), dim3(MC_THREADS), 0, stream(), M->dev);
    }
}

static void launch_chain(phb_mc_s* M) {
    if (M->d.precision == PHB_FP32) {   // FP32 form: orders 4 and 5 (checked at creation)
        if (M->dev.P == 4) launch_chain_p<4, float>(M);
        else launch_chain_p<5, float>(M);
        return;
    }
    switch (M->dev.P) {
        case 3: launch_chain_p<3>(M); break;
        case 4: launch_chain_p<4>(M); break;
        case 5: launch_chain_p<5>(M); break;
        case 6: launch_chain_p<6>(M); break;
        case 7: launch_chain_p<7>(M); break;
        default: launch_chain_p<8>(M); break;
    }
}

void launch_probe(phb_mc_s* M, int n, const int* i, const int* j, const int* f, double* out) {
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(n), dim3(MC_THREADS), 0, stream(), M->dev, n, i, j, f, out);
    };
    if (M->d.precision == PHB_FP32) {
        if (M->dev.P == 4) go(k_mc_probe<4, float>);
        else go(k_mc_probe<5, float>);
        return;
    }
    switch (M->dev.P) {
        case 3: go(k_mc_probe<3>); break;
        case 4: go(k_mc_probe<4>); break;
        case 5: go(k_mc_probe<5>); break;
        case 6: go(k_mc_probe<6>); break;
        case 7: go(k_mc_probe<7>); break;
        default: go(k_mc_probe<8>); break;
    }
}

void fill_stats(phb_mc_s* M, const McCtl& c, phb_mc_stats* s) {
    s->attempted = c.trials_done;
    s->accepted = c.accepted;
    s->accepted_forward = c.acc_fwd;
    s->accepted_backward = c.acc_bwd;
    s->nprot = c.nprot;
    s->nsites = M->nsites;
    s->status = c.status;
    s->E = c.E;
    s->E_recomputed = c.E_recomputed;
    s->max_drift = c.max_drift;
    s->refreshes = M->refreshes;
    s->recomputes = M->recomputes;
    s->launches = M->launches;
    s->chain_ms = M->chain_ms;
    s->refresh_ms = M->refresh_ms;
    s->recompute_ms = M->recompute_ms;
}

}  // namespace phb

extern "C" int phb_mc_create(phb_atoms a, phb_params p, phb_csr neigh, phb_pme pme, const phb_box* box,
                             const phb_mc_desc* d, unsigned seed, phb_mc* out) {
    return guard([&] {
        require(a && p && neigh && box && d && out, "phb_mc_create: null argument");
        require(p->d.coulombtype == PHB_COUL_PME || p->d.coulombtype == PHB_COUL_CUT,
                "phb_mc_create: coulombtype must be pme or coul_cut", PHB_ERR_UNSUPPORTED);
        // coul_cut (total_energy's other branch, MC/src/mc.cu:52-56): the chain runs on a plan of its
        // own with alpha 0 and no reciprocal part; a plan passed in is not used
        std::unique_ptr<phb_pme_s> cut_plan;
        if (p->d.coulombtype == PHB_COUL_CUT) {
            cut_plan.reset(pme_create_cut(p->d, *box));
            pme = cut_plan.get();
        }
        require(pme != nullptr, "phb_mc_create: null PME plan");
        require(d->precision == PHB_FP64 || d->precision == PHB_FP32, "phb_mc_create: precision must be PHB_FP64 or PHB_FP32",
                PHB_ERR_UNSUPPORTED);
        require(d->precision == PHB_FP64 || pme->order == 4 || pme->order == 5,
                "phb_mc_create: the FP32 form supports spline orders 4 and 5", PHB_ERR_UNSUPPORTED);
        require(pme->precision == PHB_FP64, "phb_mc_create: the PME plan must be FP64 (it keeps the FP64 potential "
                                            "from which the FP32 form's copy is made)", PHB_ERR_UNSUPPORTED);
        require(neigh->rows == a->n, "phb_mc_create: neighbour list rows != natoms");
        require(a->max_type < p->d.ntypes, "phb_mc_create: atom type id outside the LJ table");
        const int T = p->d.ntypes;
        for (int t : {d->nh_n_type, d->nh_p_type, d->w_type, d->cl_type})
            require(t >= 1 && t <= T, "phb_mc_create: MC type id must be between 1 and ntypes");  // MC/src/io.cpp:289-299
        const int pm = d->pending_max > 0 ? d->pending_max : 32;
        require(pm >= 2 && pm <= MC_KMAX, "phb_mc_create: pending_max must be in [2, 64]");
        require(d->record_cap >= 0, "phb_mc_create: record_cap must be >= 0");
        require(d->refresh == PHB_REFRESH_AUTO || d->refresh == PHB_REFRESH_DIRECT || d->refresh == PHB_REFRESH_KSPACE,
                "phb_mc_create: refresh must be PHB_REFRESH_AUTO, _DIRECT or _KSPACE");
        const int sw_ = d->spec_window > 0 ? d->spec_window : 16;
        require(sw_ == 1 || sw_ == 2 || sw_ == 4 || sw_ == 8 || sw_ == 16,
                "phb_mc_create: spec_window must be 1, 2, 4, 8 or 16");
        require(std::min(pme->nx, std::min(pme->ny, pme->nz)) >= 2 * pme->order - 1,
                "phb_mc_create: every grid dimension must be >= 2*order-1", PHB_ERR_UNSUPPORTED);
        auto M = std::make_unique<phb_mc_s>();
        M->atoms = a; M->params = p; M->neigh = neigh; M->pme = pme; M->box = *box; M->d = *d;
        M->own_pme = cut_plan.release();
        require(d->refresh != PHB_REFRESH_DIRECT || refresh_direct_allowed(M.get()),
                "phb_mc_create: the direct refresh needs nx <= 64", PHB_ERR_UNSUPPORTED);
        M->d.pending_max = pm;
        const int N = a->n;
        // ---- pools (build_plan, MC/src/mc.cu:67-100 and :332-341)
        std::vector<double> hq(N), hz(N);
        std::vector<int> ht(N);
        a->q.download(hq.data(), N); a->z.download(hz.data(), N); a->type.download(ht.data(), N);
        sync();
        std::vector<int> nh_n, nh_p, w, cl;
        for (int i = 0; i < N; ++i) {
            const int t = ht[i] + 1;
            if (t == d->nh_n_type) nh_n.push_back(i);
            else if (t == d->nh_p_type) nh_p.push_back(i);
            else if (t == d->w_type) w.push_back(i);
        }
        require(!nh_n.empty() || !nh_p.empty(), "[plan] No NH sites found.", PHB_ERR_STATE);  // MC/src/mc.cu:84
        double z_max_site = 0.0;
        for (int i : nh_n) z_max_site = std::max(z_max_site, hz[i]);
        for (int i : nh_p) z_max_site = std::max(z_max_site, hz[i]);
        const double z_cut = z_max_site + d->w_z_extra;
        M->z_cut = z_cut;
        { std::vector<int> f; for (int j : w) if (hz[j] < z_cut) f.push_back(j); w.swap(f); }
        for (int j = 0; j < N; ++j) if (ht[j] + 1 == d->cl_type && hz[j] < z_cut) cl.push_back(j);
        M->nsites = (int)(nh_n.size() + nh_p.size());
        M->nions = (int)(w.size() + cl.size());
        {
            // shared layout: nh_n up from 0, nh_p down from nsites-1, w up from nsites, cl down from the end
            const int ns = M->nsites, ni = M->nions;
            std::vector<int> h(std::max(ns + ni, 1), -1);
            for (size_t k = 0; k < nh_n.size(); ++k) h[k] = nh_n[k];
            for (size_t k = 0; k < nh_p.size(); ++k) h[ns - 1 - k] = nh_p[k];
            for (size_t k = 0; k < w.size(); ++k) h[ns + k] = w[k];
            for (size_t k = 0; k < cl.size(); ++k) h[ns + ni - 1 - k] = cl[k];
            M->pbuf.alloc(h.size());
            M->pbuf.upload(h.data(), h.size());
            M->pool_lds_bytes = ((size_t)(ns + ni) * sizeof(unsigned short) + 15) / 16 * 16;
            M->lds_pools = N <= 65535 && M->pool_lds_bytes <= 96 * 1024;
        }
        // ---- full rows + coefficients
        hipStream_t st = stream();
        const long long nfull = build_full_rows(N, neigh, M->fhead, M->fnbr);
        M->kreal.alloc(std::max<long long>(nfull, 1));
        M->r2.alloc(std::max<long long>(nfull, 1));
        if (nfull > 0) {
            hipLaunchKernelGGL(k_full_coeffs, dim3(ceil_div(N, 4)), dim3(256), 0, st, N, a->x.p, a->y.p, a->z.p,
                               box->Lx, box->Ly, box->Lz, pme->elF, pme->alpha, M->fhead.p, M->fnbr.p, M->kreal.p,
                               M->r2.p);
            PHB_HIP(hipGetLastError());
            sync();
        }
        // ---- stencils, mesh kernel, potential
        const int P = pme->order;
        M->sbase.alloc(3 * (size_t)N);
        M->sw.alloc(3 * (size_t)P * N);
        hipLaunchKernelGGL(k_stencils, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, a->x.p, a->y.p, a->z.p, pme->geom,
                           M->sbase.p, M->sw.p);
        const size_t G = (size_t)pme->nx * pme->ny * pme->nz;
        M->phi.alloc(G); M->gk.alloc(G); M->dgrid.alloc(G);
        PHB_HIP(hipMemsetAsync(M->dgrid.p, 0, G * sizeof(double), st));
        hipLaunchKernelGGL(k_delta_grid, dim3(1), dim3(64), 0, st, M->dgrid.p);
        PmeKernels::launch_fft_buf(pme, M->dgrid.p, M->gk.p, 2, true, false);
        sync();
        M->dgrid.release();
        const size_t GF = (size_t)(pme->nx / 2 + 1) * (pme->ny / 2 + 1) * (pme->nz / 2 + 1);
        M->gf.alloc(GF);
        hipLaunchKernelGGL(k_fold_g, dim3(std::min<size_t>((GF + 255) / 256, 4096)), dim3(256), 0, st, pme->nx, pme->ny,
                           pme->nz, M->gk.p, M->gf.p);
        M->selfw.alloc(N);
        hipLaunchKernelGGL(k_self_terms, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, P, pme->nx, pme->ny, pme->nz,
                           M->sw.p, M->gk.p, M->selfw.p);
        // ---- FP32 form: FP32 copies of the real-space coefficients, mesh kernel and potential
        if (d->precision == PHB_FP32) {
            M->kreal32.alloc(std::max<long long>(nfull, 1));
            M->gf32.alloc(GF);
            M->phi32.alloc(G);
            cast_f32(M->kreal.p, M->kreal32.p, nfull);
            cast_f32(M->gf.p, M->gf32.p, (long long)GF);
        }
        // ---- records
        const long long rc = d->record_cap;
        M->rec_fwd.alloc(std::max<long long>(rc, 1)); M->rec_acc.alloc(std::max<long long>(rc, 1));
        M->rec_i.alloc(std::max<long long>(rc, 1)); M->rec_j.alloc(std::max<long long>(rc, 1));
        M->rec_dE.alloc(std::max<long long>(rc, 1)); M->rec_u.alloc(std::max<long long>(rc, 1));
        // ---- control block with the seeded engine (std::mt19937(seed) state via its stream form)
        McCtl c{};
        c.trials_target = 0; c.trials_done = 0; c.rec_cap = rc;
        c.pending_max = pm; c.status = 0; c.nprot = (int)nh_p.size();
        c.sizes[0] = (int)nh_n.size(); c.sizes[1] = (int)nh_p.size(); c.sizes[2] = (int)w.size(); c.sizes[3] = (int)cl.size();
        c.E_recomputed = std::numeric_limits<double>::quiet_NaN();
        c.max_drift = 0.0;
        {
            std::mt19937 eng(seed);
            std::stringstream ss;
            ss << eng;
            for (int k = 0; k < 624; ++k) ss >> c.mt[k];
            ss >> c.mti;
        }
        c.nprot_target = -1;
        M->ctl.alloc(1);
        M->ctl.upload(&c, 1);
        {   // one entry per workgroup of either refresh (direct tiles, or the inverse plane pass)
            const size_t na = (size_t)std::max(refresh_direct_blocks(pme->ny, pme->nz, 4, 2), pme->plane_gate_blocks());
            M->applied.alloc(na);
            PHB_HIP(hipMemsetAsync(M->applied.p, 0, na * sizeof(unsigned), stream()));
        }
        M->dev.rd_ty = RD_TY;
        M->dev.rd_tz = RD_TZ;
        build_rd_tiles(M.get());
        // ---- device view
        McDev& D = M->dev;
        D.N = N; D.z = a->z.p; D.q = a->q.p; D.type = a->type.p;
        D.fhead = M->fhead.p; D.fnbr = M->fnbr.p; D.kreal = M->kreal.p; D.r2 = M->r2.p;
        D.sbase = M->sbase.p; D.sw = M->sw.p; D.phi = M->phi.p; D.gk = M->gk.p; D.selfw = M->selfw.p;
        D.gf = M->gf.p; D.fnx = pme->nx / 2 + 1; D.fny = pme->ny / 2 + 1;
        D.kreal32 = M->kreal32.p; D.gf32 = M->gf32.p; D.phi32 = M->phi32.p;
        D.nx = pme->nx; D.ny = pme->ny; D.nz = pme->nz; D.P = P;
        D.T = T; D.eps = p->d_eps.p; D.sig = p->d_sig.p; D.shift = p->d_shift.p;
        D.use_shift = p->d.lj_shift ? 1 : 0; D.use_cut = p->lj_cut_active ? 1 : 0; D.lj_rc2 = p->lj_rc * p->lj_rc;
        D.pbuf = M->pbuf.p; D.nsites = M->nsites; D.nions = M->nions;
        D.W = d->spec_window > 0 ? d->spec_window : 16;
        D.t_nhn = d->nh_n_type - 1; D.t_nhp = d->nh_p_type - 1; D.t_w = d->w_type - 1; D.t_cl = d->cl_type - 1;
        D.beta = d->beta; D.mu = d->mu_eff; D.elF = pme->elF; D.alpha = pme->alpha; D.vol = pme->vol;
        D.Lz = box->Lz; D.use3dc = pme->use3dc ? 1 : 0;
        D.self_coef = pme->elF * pme->alpha / std::sqrt(M_PI);
        D.ctl = M->ctl.p;
        D.applied = M->applied.p;
        D.rec_fwd = M->rec_fwd.p; D.rec_i = M->rec_i.p; D.rec_j = M->rec_j.p; D.rec_acc = M->rec_acc.p;
        D.rec_dE = M->rec_dE.p; D.rec_u = M->rec_u.p;
        // ---- initial full recompute (phi, E, moments)
        full_recompute(M.get());
        M->recomputes = 0;
        M->next_recompute = d->recompute_interval > 0 ? d->recompute_interval : -1;
        sync();
        *out = M.release();
        return PHB_OK;
    });
}

extern "C" int phb_mc_destroy(phb_mc M) {
    return guard([&] { delete M; return PHB_OK; });
}

namespace phb {

void run_plain(phb_mc_s* M, long long ntrials) {
    McCtl c = read_ctl(M);
    const long long target = c.trials_done + ntrials;
    long long done = c.trials_done;
    int status = c.status;
    while (done < target && status == 0) {
        long long chunk_end = target;
        if (M->next_recompute > 0) chunk_end = std::min(chunk_end, M->next_recompute);
        hipLaunchKernelGGL(k_set_target, dim3(1), dim3(64), 0, stream(), M->ctl.p, chunk_end);
        // each launch runs until the chunk ends or pending_max changes accumulate;
        // batch size from the trials one launch has been covering
        while (done < chunk_end && status == 0) {
            const int batch = launch_batch(c, chunk_end - done, M->d.pending_max);
            const long long before = done;
            for (int b = 0; b < batch; ++b) {
                const size_t e0 = M->ev_used;
                if (M->profiling) M->event();
                launch_chain(M);
                ++M->launches;
                if (M->profiling) M->event();
                refresh(M);
                if (M->profiling) { M->event(); M->marks.push_back({e0, 0}); }
            }
            PHB_HIP(hipGetLastError());
            c = read_ctl(M);
            require(c.trials_done > done || c.status != 0 || c.trials_done >= chunk_end,
                    "phb_mc_run: the chain made no progress", PHB_ERR_INTERNAL);
            require(c.status != 2, "phb_mc_run: engine stream rolled back past its kept block", PHB_ERR_INTERNAL);
            done = c.trials_done;
            status = c.status;
            M->trials_seen += done - before;
        }
        if (done >= chunk_end && M->next_recompute > 0 && chunk_end == M->next_recompute) {
            full_recompute(M);
            M->next_recompute += M->d.recompute_interval;
        }
    }
    sync();
}

}  // namespace phb

extern "C" int phb_mc_run(phb_mc M, long long ntrials, phb_mc_stats* out) {
    return guard([&] {
        require(M != nullptr && ntrials >= 0, "phb_mc_run: bad arguments");
        run_plain(M, ntrials);
        if (out) fill_stats(M, read_ctl(M), out);
        return PHB_OK;
    });
}

extern "C" int phb_mc_stats_get(phb_mc M, phb_mc_stats* out) {
    return guard([&] {
        require(M && out, "phb_mc_stats_get: null");
        fill_stats(M, read_ctl(M), out);
        return PHB_OK;
    });
}

extern "C" int phb_mc_record(phb_mc M, long long n, signed char* fwd, int* i, int* j, signed char* acc, double* dE,
                             double* u) {
    return guard([&] {
        require(M != nullptr && n >= 0, "phb_mc_record: bad arguments");
        require(n <= M->d.record_cap, "phb_mc_record: n exceeds record_cap");
        if (n == 0) return PHB_OK;
        if (fwd) M->rec_fwd.download(fwd, n);
        if (i) M->rec_i.download(i, n);
        if (j) M->rec_j.download(j, n);
        if (acc) M->rec_acc.download(acc, n);
        if (dE) M->rec_dE.download(dE, n);
        if (u) M->rec_u.download(u, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_mc_delta_e(phb_mc M, int n, const int* i, const int* j, const int* fwd, double* dE) {
    return guard([&] {
        require(M && n >= 0 && (n == 0 || (i && j && fwd && dE)), "phb_mc_delta_e: bad arguments");
        if (n == 0) return PHB_OK;
        for (int k = 0; k < n; ++k)
            require(i[k] >= 0 && i[k] < M->atoms->n && j[k] >= 0 && j[k] < M->atoms->n && i[k] != j[k],
                    "phb_mc_delta_e: atom index out of range");
        DevBuf<int> di(n), dj(n), df(n);
        DevBuf<double> dout(n);
        di.upload(i, n); dj.upload(j, n); df.upload(fwd, n);
        launch_probe(M, n, di.p, dj.p, df.p, dout.p);
        PHB_HIP(hipGetLastError());
        dout.download(dE, n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_mc_recompute(phb_mc M, double* E) {
    return guard([&] {
        require(M != nullptr, "phb_mc_recompute: null");
        full_recompute(M);
        McCtl c = read_ctl(M);
        if (E) *E = c.E;
        return PHB_OK;
    });
}

extern "C" int phb_mc_set_profiling(phb_mc M, int on) {
    return guard([&] {
        require(M != nullptr, "phb_mc_set_profiling: null");
        require(on >= 0 && on <= 2, "phb_mc_set_profiling: on must be 0, 1 or 2");
        sync();
        if (M->profiling) M->collect();
        M->profiling = on != 0;
        M->phase_clock = on == 1;
        M->chain_ms = M->refresh_ms = M->recompute_ms = 0.0;
        PHB_HIP(hipMemsetAsync(M->ctl.p->da_prof, 0, sizeof(M->ctl.p->da_prof), stream()));
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_mc_set_rng_state(phb_mc M, const unsigned* state) {
    return guard([&] {
        require(M && state, "phb_mc_set_rng_state: null");
        require(state[624] <= 624u, "phb_mc_set_rng_state: engine index beyond 624");
        sync();
        McCtl c = read_ctl(M);
        for (int k = 0; k < 624; ++k) c.mt[k] = state[k];
        c.mti = (int)state[624];
        M->ctl.upload(&c, 1);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_mc_set_mu(phb_mc M, double mu_eff) {
    return guard([&] {
        require(M, "phb_mc_set_mu: null");
        require(std::isfinite(mu_eff), "phb_mc_set_mu: mu_eff must be finite");
        M->dev.mu = mu_eff;
        M->d.mu_eff = mu_eff;
        return PHB_OK;
    });
}

namespace phb {

__global__ void k_fill_u32(unsigned* p, size_t n, unsigned v) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

void set_refresh_tile(phb_mc_s* M, int ty, int tz) {
    require(refresh_tile_ok(ty, tz), "refresh tile: 8 or 4 rows along y by 4 or 2 planes along z");
    if (ty == M->dev.rd_ty && tz == M->dev.rd_tz) return;
    refresh(M);   // the pending changes into phi with the current tiles
    McCtl c = read_ctl(M);
    M->dev.rd_ty = ty;
    M->dev.rd_tz = tz;
    build_rd_tiles(M);
    // phi holds every pending change: every gate entry of the new tiling says so
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((M->applied.n + 255) / 256, 1024)), dim3(256), 0,
                       stream(), M->applied.p, M->applied.n, c.pend_seq);
    PHB_HIP(hipGetLastError());
    sync();
}

}  // namespace phb

extern "C" int phb_mc_set_refresh_tile(phb_mc M, int ty, int tz) {
    return guard([&] {
        require(M, "phb_mc_set_refresh_tile: null");
        set_refresh_tile(M, ty, tz);
        return PHB_OK;
    });
}

extern "C" int phb_mc_rng_state(phb_mc M, unsigned* state) {
    return guard([&] {
        require(M && state, "phb_mc_rng_state: null");
        McCtl c = read_ctl(M);
        for (int k = 0; k < 624; ++k) state[k] = c.mt[k];
        state[624] = (unsigned)c.mti;
        return PHB_OK;
    });
}
