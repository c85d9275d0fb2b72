// phase2.hip -- the phase-2 move rule of phase2_swap_jitter (MC/src/mc.cu:159-315)
// on the device: the 4-point move (NH_P -> NH_N, NH_N -> NH_P, Cl -> W, W -> Cl)
// or, when no 4-point move is possible, the W-Cl-only move (Cl -> W, W -> Cl);
// accept if dE <= 0 or exp(clamp(-beta dE, +-40)) > u (u drawn only when dE > 0);
// pools erase(find) + push_back; stop after `patience` trials without a new
// best energy (best improves when E < best - best_eps).
//
// Exact Delta-E of a K-atom change (K = 4 or 2), applied in the order of the
// proposal's draws: each changed atom's full row with the earlier changed atoms
// at their new identity; the reciprocal part Sum_k dq_k W_k.phi + 1/2 Sum_k
// dq_k^2 s_k + Sum_{k<l} dq_k dq_l W_k^T G W_l (phi holds every accepted change:
// a launch ends at its accepted move and the refresh runs before the next one);
// the net-charge and slab terms from the running moments.
//
// Speculative rejection as in the plain chain: wave 0 draws W trials assuming
// every one is rejected (a rejected trial has dE > 0 and draws its uniform),
// all waves evaluate them, the first acceptance or the trial that exhausts the
// patience ends the round, and the stream position rolls back behind the last
// trial decided.  The sequence equals the sequential chain's.
#include <algorithm>
#include <memory>
#include <vector>

#include "mc_host.hpp"

namespace phb {

constexpr int P2_THREADS = 1024;

template <int P>
struct Trial4 {
    int atom[4], pool[4], kpos[4];
    int t_old[4], t_new[4];
    double q_old[4], q_new[4], z[4];
    int sb[4][3], hrow[4], nrow[4];
    double w[4][3][P];
    double self[4];
    int pos_u;                 // stream position of the trial's uniform
    double u;
};

struct P2Dev {
    int K;                     // 4: four-point move, 2: W-Cl only
    int patience;
    double best_eps;
    long long rec_cap;
    long long rec_base;        // trial index of record 0 (records rec_base .. rec_base + rec_cap - 1 are kept)
    int* rec_atoms;            // K per trial (4 slots)
    signed char* rec_acc;
    double* rec_dE;
};

// pools of the K atoms in draw order, and the new identity of each
__device__ __forceinline__ void p2_move(const McDev& D, int K, int k, int& pool, int& tnew, double& qnew) {
    if (K == 4) {
        const int pools[4] = {1, 0, 3, 2};
        pool = pools[k];
        tnew = k == 0 ? D.t_nhn : k == 1 ? D.t_nhp : k == 2 ? D.t_w : D.t_cl;
        qnew = k == 1 ? 1.0 : k == 3 ? -1.0 : 0.0;
    } else {
        pool = k == 0 ? 3 : 2;
        tnew = k == 0 ? D.t_w : D.t_cl;
        qnew = k == 0 ? 0.0 : -1.0;
    }
}

template <int P>
__device__ __forceinline__ void stage_trial4_data(const McDev& D, Trial4<P>& t, int K, int lt, int nthr) {
    const int NW = K * 3 * P;
    for (int k = lt; k < NW + 9 * K; k += nthr) {
        if (k < NW) {
            const int a = k / (3 * P), r = k - a * 3 * P;
            (&t.w[a][0][0])[r] = D.sw[(size_t)t.atom[a] * 3 * P + r];
        } else {
            const int q = k - NW, a = q / 9, r = q - a * 9, atom = t.atom[a];
            if (r < 3) t.sb[a][r] = D.sbase[3 * atom + r];
            else if (r == 3) t.hrow[a] = D.fhead[atom];
            else if (r == 4) t.nrow[a] = D.fhead[atom + 1] - D.fhead[atom];
            else if (r == 5) t.self[a] = D.selfw[atom];
            else if (r == 6) t.t_old[a] = D.type[atom];
            else if (r == 7) t.q_old[a] = D.q[atom];
            else t.z[a] = D.z[atom];
        }
    }
}

// This thread's share (items lt, lt + nt, ...) of the pair, stencil and
// mesh-kernel-window sums of a K-atom trial.
template <int P>
__device__ double trial4_partial(const McDev& D, const Trial4<P>& t, int K, int lt, int nt) {
    constexpr int CW = 2 * P - 1, PP = P * P;
    int rstart[5];
    rstart[0] = 0;
    for (int k = 0; k < K; ++k) rstart[k + 1] = rstart[k] + t.nrow[k];
    const int nA = rstart[K], nB = nA + K * PP, npair = K * (K - 1) / 2;
    const int total = nB + npair * CW * CW;
    double acc = 0.0;
    for (int it = lt; it < total; it += nt) {
        if (it < nA) {
            int k = 0;
            while (it >= rstart[k + 1]) ++k;
            const int e = t.hrow[k] + (it - rstart[k]);
            const int n = D.fnbr[e];
            double qn = D.q[n];
            int tn = D.type[n];
            for (int l = 0; l < k; ++l)
                if (n == t.atom[l]) { qn = t.q_new[l]; tn = t.t_new[l]; }
            double v = (t.q_new[k] - t.q_old[k]) * qn * D.kreal[e];
            if (t.t_old[k] != t.t_new[k]) {
                const double r2 = D.r2[e];
                v += lj_pair<double>(D, t.t_new[k], tn, r2) - lj_pair<double>(D, t.t_old[k], tn, r2);
            }
            acc += v;
        } else if (it < nB) {
            const int s = it - nA, k = s / PP, r = s - k * PP, wz = r / P, wy = r - wz * P;
            const int* sb = t.sb[k];
            const int gz = wrap1(sb[2] + wz, D.nz), gy = wrap1(sb[1] + wy, D.ny);
            const double* row = D.phi + ((size_t)gz * D.ny + gy) * D.nx;
            double sx = 0.0;
#pragma unroll
            for (int u = 0; u < P; ++u) sx += t.w[k][0][u] * row[wrap1(sb[0] + u, D.nx)];
            acc += (t.q_new[k] - t.q_old[k]) * t.w[k][2][wz] * t.w[k][1][wy] * sx;
        } else {
            const int s = it - nB, pr = s / (CW * CW), r = s - pr * (CW * CW), dz = r / CW, dy = r - dz * CW;
            int ka = 0, kb = 1, c = pr;   // pair index -> (ka, kb), ka < kb
            for (ka = 0; ka < K - 1; ++ka) {
                const int m = K - 1 - ka;
                if (c < m) { kb = ka + 1 + c; break; }
                c -= m;
            }
            const int* sba = t.sb[ka];
            const int* sbb = t.sb[kb];
            const double(*wa)[P] = t.w[ka];
            const double(*wb)[P] = t.w[kb];
            const int x0 = wrap1(sba[0] - sbb[0], D.nx) - (P - 1);
            const int gy = fold(wrap1(wrap1(sba[1] - sbb[1], D.ny) - (P - 1) + dy, D.ny), D.ny);
            const int gz = fold(wrap1(wrap1(sba[2] - sbb[2], D.nz) - (P - 1) + dz, D.nz), D.nz);
            const double* row = D.gf + ((size_t)gz * D.fny + gy) * D.fnx;
            double sx = 0.0;
#pragma unroll
            for (int d = 0; d < CW; ++d) sx += corr<P>(wa[0], wb[0], d) * row[fold(wrap1(x0 + d, D.nx), D.nx)];
            acc += (t.q_new[ka] - t.q_old[ka]) * (t.q_new[kb] - t.q_old[kb]) * corr<P>(wa[2], wb[2], dz) *
                   corr<P>(wa[1], wb[1], dy) * sx;
        }
    }
    return acc;
}

// Net-charge, self and slab terms of a K-atom change, plus the mesh self terms.
template <int P>
__device__ double correction4(const McDev& D, const Trial4<P>& t, int K, double Q, double Mz, double Qz2) {
    double dQ = 0.0, dQ2 = 0.0, dMz = 0.0, dQz2 = 0.0, mesh = 0.0;
    for (int k = 0; k < K; ++k) {
        const double dq = t.q_new[k] - t.q_old[k];
        dQ += dq;
        dQ2 += t.q_new[k] * t.q_new[k] - t.q_old[k] * t.q_old[k];
        dMz += dq * t.z[k];
        dQz2 += dq * t.z[k] * t.z[k];
        mesh += 0.5 * dq * dq * t.self[k];
    }
    const double Qn = Q + dQ, Mzn = Mz + dMz, Qz2n = Qz2 + dQz2;
    const double e_self = -D.self_coef * dQ2;
    const double e_q = D.alpha > 0.0 ? -(M_PI * D.elF) / (2.0 * D.vol * D.alpha * D.alpha) * (Qn * Qn - Q * Q) : 0.0;
    return e_self + e_q + (slab_energy(D, Mzn, Qn, Qz2n) - slab_energy(D, Mz, Q, Qz2)) + mesh;
}

template <int P>
__global__ __launch_bounds__(P2_THREADS) void k_mc_phase2(McDev D, P2Dev Z) {
    __shared__ RngLds s_rng;
    __shared__ Trial4<P> s_tr[16];
    __shared__ ChainState s_st;
    __shared__ double s_wsum[16];
    __shared__ int s_ntr, s_stop, s_acc_k, s_since;
    __shared__ double s_best;
    __shared__ int s_er_pool[4], s_er_pos[4], s_er_n[4], s_pb_pool[4], s_pb_val[4], s_pb_n[4];
    McCtl* C = D.ctl;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = Z.K, W = D.W, wpt = 16 / W;
    const long long target = C->p2_target;
    if (C->p2_done >= target || C->status != 0 || C->p2_stop != 0) return;
    if (pending_unfolded(D) > 0) {   // the per-move Delta-E assumes phi holds every change
        if (tid == 0) C->status = 3;
        return;
    }
    int* pbuf = D.pbuf;
    for (int k = tid; k < 624; k += blockDim.x) {
        s_rng.cur[k] = C->mt[k];
        s_rng.out[0][k] = mt_temper(C->mt[k]);
    }
    if (tid == 0) {
        s_st.done = C->p2_done;
        s_st.acc_n = C->p2_acc;
        for (int k = 0; k < 4; ++k) s_st.sz[k] = C->sizes[k];
        s_st.status = 0; s_st.blk = 0; s_st.pos = C->mti; s_st.pend_n = 0;
        s_st.E = C->E; s_st.Q = C->Q; s_st.Q2 = C->Q2; s_st.Mz = C->Mz; s_st.Qz2 = C->Qz2;
        s_since = C->p2_since; s_best = C->p2_best;
        s_stop = 0;
    }
    __syncthreads();
    for (;;) {
        // ---- 1. proposals (wave 0)
        if (wave == 0) {
            int blk = s_st.blk, pos = s_st.pos;
            const long long done = s_st.done;
            int ntr = 0, stop = 0;
            if (done >= target || s_st.pend_n > 0 || s_since >= Z.patience) stop = 1;
            else {
                const long long room = target - done;
                ntr = (int)(room < W ? room : W);
                const int per = K + 2;
                rng_ensure(s_rng, blk, pos + ntr * per, lane);
                bool ok = true;
                int at[4] = {0, 0, 0, 0}, kp[4] = {0, 0, 0, 0};
                double u = 0.0;
                int pu = 0;
                if (lane < ntr) {
                    const int q = pos + lane * per;
                    for (int k = 0; k < K; ++k) {
                        int pool, tnew;
                        double qnew;
                        p2_move(D, K, k, pool, tnew, qnew);
                        unsigned a;
                        ok = index_at(s_rng, q + k, (unsigned)s_st.sz[pool], a) && ok;
                        kp[k] = (int)a;
                        at[k] = pbuf[pool_slot(D, pool, (int)a)];
                    }
                    pu = q + K;
                    u = canonical_at(s_rng, pu);
                }
                if (__all(ok)) {
                    if (lane < ntr) {
                        Trial4<P>& t = s_tr[lane];
                        for (int k = 0; k < K; ++k) {
                            t.atom[k] = at[k]; t.kpos[k] = kp[k];
                            p2_move(D, K, k, t.pool[k], t.t_new[k], t.q_new[k]);
                        }
                        t.pos_u = pu; t.u = u;
                    }
                    pos += ntr * per;
                } else {
                    for (int m = 0; m < ntr; ++m) {
                        int at2[4], kp2[4];
                        for (int k = 0; k < K; ++k) {
                            int pool, tnew;
                            double qnew;
                            p2_move(D, K, k, pool, tnew, qnew);
                            kp2[k] = (int)draw_index(s_rng, blk, pos, lane, (unsigned)s_st.sz[pool]);
                            at2[k] = pbuf[pool_slot(D, pool, kp2[k])];
                        }
                        const int p2 = pos;
                        const double u2 = draw_canonical(s_rng, blk, pos, lane);
                        if (lane == 0) {
                            Trial4<P>& t = s_tr[m];
                            for (int k = 0; k < K; ++k) {
                                t.atom[k] = at2[k]; t.kpos[k] = kp2[k];
                                p2_move(D, K, k, t.pool[k], t.t_new[k], t.q_new[k]);
                            }
                            t.pos_u = p2; t.u = u2;
                        }
                    }
                }
            }
            if (lane == 0) { s_ntr = ntr; s_stop = stop; s_st.blk = blk; s_st.pos = pos; }
        }
        __syncthreads();
        if (s_stop) break;
        const int ntr = s_ntr;
        // ---- 2. Delta-E of every trial (wpt waves per trial)
        const int kt = wave / wpt, lt = (wave - kt * wpt) * 64 + lane, nt = wpt * 64;
        if (kt < ntr) stage_trial4_data<P>(D, s_tr[kt], K, lt, nt);
        __syncthreads();
        double part = 0.0;
        if (kt < ntr) part = trial4_partial<P>(D, s_tr[kt], K, lt, nt);
        part = wave_sum(part);
        if (lane == 0) s_wsum[wave] = part;
        __syncthreads();
        // ---- 3. decisions (wave 0): the first acceptance, or the trial after which the
        //         patience is exhausted, ends the round
        if (wave == 0) {
            const long long done = s_st.done;
            const int kk = lane;
            const double Q = s_st.Q, Mz = s_st.Mz, Qz2 = s_st.Qz2;
            bool acc = false;
            double dE = 0.0;
            if (kk < ntr) {
                const Trial4<P>& t = s_tr[kk];
                double dsum = 0.0;
                for (int w = kk * wpt; w < (kk + 1) * wpt; ++w) dsum += s_wsum[w];
                dE = dsum + correction4<P>(D, t, K, Q, Mz, Qz2);
                double expo;
                {
#pragma clang fp contract(off)
                    expo = fmin(fmax((-D.beta) * dE, -40.0), 40.0);
                }
                acc = dE <= 0.0 || exp(expo) > t.u;
            }
            const unsigned long long msk = __ballot(acc);
            const int first = msk ? __ffsll((long long)msk) - 1 : ntr;
            const int since0 = s_since;
            const int kstop = Z.patience - since0 - 1;   // last trial if every trial is rejected
            const bool pat_first = kstop < first && kstop < ntr;
            const int last = pat_first ? kstop : (first < ntr ? first : ntr - 1);
            const long long ri = done + kk - Z.rec_base;
            if (kk <= last && kk < ntr && ri >= 0 && ri < Z.rec_cap) {
                const Trial4<P>& t = s_tr[kk];
                for (int k = 0; k < 4; ++k) Z.rec_atoms[4 * ri + k] = k < K ? t.atom[k] : -1;
                Z.rec_acc[ri] = acc ? 1 : 0;
                Z.rec_dE[ri] = dE;
            }
            if (!pat_first && first < ntr && kk == first) {
                // accepted: state update; the stream resumes right after this trial's draws
                const Trial4<P>& t = s_tr[kk];
                const int pos = t.pos_u + (dE <= 0.0 ? 0 : 2);
                if (pos < s_st.blk * 624 - 624) s_st.status = 2;
                s_st.pos = pos;
                double dQ = 0.0, dQ2 = 0.0, dMz = 0.0, dQz2 = 0.0;
                for (int k = 0; k < K; ++k) {
                    const double dq = t.q_new[k] - t.q_old[k];
                    dQ += dq; dQ2 += t.q_new[k] * t.q_new[k] - t.q_old[k] * t.q_old[k];
                    dMz += dq * t.z[k]; dQz2 += dq * t.z[k] * t.z[k];
                    D.q[t.atom[k]] = t.q_new[k]; D.type[t.atom[k]] = t.t_new[k];
                    // erase from its pool at the drawn position; push onto its partner's pool
                    // (each pool loses one atom and gains one, so erasing all first gives the
                    // reference's order)
                    s_er_pool[k] = t.pool[k]; s_er_pos[k] = t.kpos[k]; s_er_n[k] = s_st.sz[t.pool[k]];
                    s_pb_pool[k] = t.pool[k ^ 1]; s_pb_val[k] = t.atom[k]; s_pb_n[k] = s_st.sz[t.pool[k ^ 1]] - 1;
                }
                s_st.E += dE;
                s_st.Q = Q + dQ; s_st.Q2 = s_st.Q2 + dQ2; s_st.Mz = Mz + dMz; s_st.Qz2 = Qz2 + dQz2;
                ++s_st.acc_n;
                if (s_st.E < s_best - Z.best_eps) { s_best = s_st.E; s_since = 0; }
                else s_since = since0 + first + 1;
                s_st.pend_n = K;
                s_acc_k = first;
            }
            if (lane == 0) {
                if (pat_first || first >= ntr) {
                    s_acc_k = -1;
                    const int nproc = pat_first ? kstop + 1 : ntr;
                    s_since = since0 + nproc;
                    if (pat_first) {   // the stream stops right after the last trial decided
                        const int pos = s_tr[kstop].pos_u + 2;
                        if (pos < s_st.blk * 624 - 624) s_st.status = 2;
                        s_st.pos = pos;
                    }
                    s_st.done = done + nproc;
                } else {
                    s_st.done = done + first + 1;
                }
            }
        }
        __syncthreads();
        if (s_acc_k >= 0) {
            // ---- 4. pools: the K erasures (disjoint pool regions) as one shift pass, then the pushes
            int L[5];
            L[0] = 0;
            for (int k = 0; k < K; ++k) L[k + 1] = L[k] + (s_er_n[k] - s_er_pos[k] - 1);
            constexpr int M = 8;
            for (int base = 0; base < L[K]; base += M * P2_THREADS) {
                int v[M], dst[M];
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const int e = base + tid + m * P2_THREADS;
                    dst[m] = -1;
                    v[m] = 0;
                    if (e < L[K]) {
                        int k = 0;
                        while (e >= L[k + 1]) ++k;
                        const int idx = s_er_pos[k] + 1 + (e - L[k]);
                        v[m] = pbuf[pool_slot(D, s_er_pool[k], idx)];
                        dst[m] = pool_slot(D, s_er_pool[k], idx - 1);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int m = 0; m < M; ++m)
                    if (dst[m] >= 0) pbuf[dst[m]] = v[m];
                __syncthreads();
            }
            if (tid < K) pbuf[pool_slot(D, s_pb_pool[tid], s_pb_n[tid])] = s_pb_val[tid];
            __syncthreads();
        }
    }
    // write back
    const int acc_k = s_acc_k;
    if (wave == 0) {
        const int blk = s_st.blk, pos = s_st.pos;
        const bool in_cur = pos >= blk * 624;
        const unsigned* st = in_cur ? s_rng.cur : s_rng.prev;
        for (int k = lane; k < 624; k += 64) C->mt[k] = st[k];
        if (lane == 0) {
            C->mti = in_cur ? pos - blk * 624 : pos - (blk - 1) * 624;
            C->p2_done = s_st.done;
            C->p2_acc = s_st.acc_n;
            C->p2_since = s_since;
            C->p2_best = s_best;
            if (s_since >= Z.patience) C->p2_stop = 1;
            if (s_st.status != 0) C->status = s_st.status;
            C->E = s_st.E; C->Q = s_st.Q; C->Q2 = s_st.Q2; C->Mz = s_st.Mz; C->Qz2 = s_st.Qz2;
            if (s_st.pend_n > 0) {
                const Trial4<P>& t = s_tr[acc_k];
                for (int k = 0; k < K; ++k) {
                    C->pend_atom[k] = t.atom[k];
                    C->pend_dq[k] = t.q_new[k] - t.q_old[k];
                }
                C->pending_n = K;
                C->pend_seq = C->pend_seq + 1u;   // a new list: no workgroup has folded it
            }
        }
    }
}

__global__ void k_p2_target(McCtl* C, long long target) {
    if (threadIdx.x == 0) C->p2_target = target;
}

__global__ void k_p2_reset(McCtl* C, const int* __restrict__ sizes) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) C->sizes[k] = sizes[k];
        C->p2_done = 0; C->p2_target = 0; C->p2_acc = 0;
        C->p2_since = 0; C->p2_stop = 0;
        C->p2_best = C->E;
    }
}

}  // namespace phb

using namespace phb;

namespace {

struct P2State {
    int K = 0;                 // 4, 2, or 0 (no valid proposals)
    int patience = 0;
    double best_eps = 1e-9;
    long long rec_cap = 0, rec_base = 0;
    DevBuf<int> rec_atoms;
    DevBuf<signed char> rec_acc;
    DevBuf<double> rec_dE;
    DevBuf<int> sizes;
    long long next_recompute = 0;
};

P2State* p2_of(phb_mc_s* M) {
    require(M->p2 != nullptr, "phase 2: phb_phase2_begin has not been called", PHB_ERR_STATE);
    return static_cast<P2State*>(M->p2);
}

void launch_p2(phb_mc_s* M, const P2State* S) {
    P2Dev Z{S->K, S->patience, S->best_eps, S->rec_cap, S->rec_base, S->rec_atoms.p, S->rec_acc.p, S->rec_dE.p};
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(P2_THREADS), 0, stream(), M->dev, Z);
    };
    switch (M->dev.P) {
        case 3: go(k_mc_phase2<3>); break;
        case 4: go(k_mc_phase2<4>); break;
        case 5: go(k_mc_phase2<5>); break;
        case 6: go(k_mc_phase2<6>); break;
        case 7: go(k_mc_phase2<7>); break;
        default: go(k_mc_phase2<8>); break;
    }
    PHB_HIP(hipGetLastError());
}

void fill_p2(phb_mc_s* M, const P2State* S, phb_phase2_stats* st) {
    const McCtl c = read_ctl(M);
    st->attempted = c.p2_done;
    st->accepted = c.p2_acc;
    st->mode = S->K;
    st->stop = S->K == 0 ? 2 : (c.p2_stop ? 1 : 0);
    st->since_best = c.p2_since;
    st->E = c.E;
    st->best_E = c.p2_best;
    st->status = c.status;
}

}  // namespace

extern "C" int phb_phase2_set_pools(phb_mc M, const int sizes[4], const int* lists) {
    return guard([&] {
        require(M && sizes, "phb_phase2_set_pools: null argument");
        long long total = 0;
        for (int k = 0; k < 4; ++k) {
            require(sizes[k] >= 0, "phb_phase2_set_pools: negative pool size");
            total += sizes[k];
        }
        require(total == 0 || lists != nullptr, "phb_phase2_set_pools: null lists");
        const int N = M->atoms->n;
        for (long long k = 0; k < total; ++k)
            require(lists[k] >= 0 && lists[k] < N, "phb_phase2_set_pools: atom index out of range");
        require(sizes[0] + sizes[1] <= M->nsites && sizes[2] + sizes[3] <= M->nions,
                "phb_phase2_set_pools: pools larger than the chain's site and ion pools", PHB_ERR_STATE);
        long long o = 0;
        for (int k = 0; k < 4; ++k) {
            M->handover[k].assign(lists + o, lists + o + sizes[k]);
            o += sizes[k];
        }
        M->has_handover = true;
        return PHB_OK;
    });
}

extern "C" int phb_phase2_begin(phb_mc M, const phb_phase2_desc* d, phb_phase2_stats* out) {
    return guard([&] {
        require(M && d, "phb_phase2_begin: null argument");
        require(d->patience > 0, "phb_phase2_begin: patience must be > 0");
        require(d->record_cap >= 0, "phb_phase2_begin: record_cap must be >= 0");
        if (!M->p2) {
            M->p2 = new P2State();
            M->p2_free = [](void* p) { delete static_cast<P2State*>(p); };
        }
        P2State* S = static_cast<P2State*>(M->p2);
        refresh(M);   // phi must hold every accepted change
        const int N = M->atoms->n;
        std::vector<int> ht(N);
        std::vector<double> hz(N);
        M->atoms->type.download(ht.data(), N);
        M->atoms->z.download(hz.data(), N);
        sync();
        std::vector<int> nh_n, nh_p, w, cl;
        if (M->has_handover) {
            // the pools phase 1 handed over (W and Cl below the pool z cut, MC/src/mc.cu:177-184)
            nh_n = M->handover[0];
            nh_p = M->handover[1];
            for (int j : M->handover[2]) if (hz[j] < M->z_cut) w.push_back(j);
            for (int j : M->handover[3]) if (hz[j] < M->z_cut) cl.push_back(j);
            require((int)(nh_n.size() + nh_p.size()) <= M->nsites && (int)(w.size() + cl.size()) <= M->nions,
                    "phb_phase2_begin: hand-over pools larger than the chain's pools", PHB_ERR_STATE);
        } else {
            // pools in ascending atom order from the current identities (MC/src/mc.cu:480-494)
            for (int i = 0; i < N; ++i) {
                const int t = ht[i] + 1;
                if (t == M->d.nh_p_type) nh_p.push_back(i);
                else if (t == M->d.nh_n_type) nh_n.push_back(i);
                else if (t == M->d.w_type) { if (hz[i] < M->z_cut) w.push_back(i); }
                else if (t == M->d.cl_type) { if (hz[i] < M->z_cut) cl.push_back(i); }
            }
            require((int)(nh_n.size() + nh_p.size()) == M->nsites && (int)(w.size() + cl.size()) == M->nions,
                    "phb_phase2_begin: pool membership differs from the chain's pools", PHB_ERR_STATE);
        }
        const int ns = M->nsites, ni = M->nions;
        std::vector<int> h(std::max(ns + ni, 1), -1);
        for (size_t k = 0; k < nh_n.size(); ++k) h[k] = nh_n[k];
        for (size_t k = 0; k < nh_p.size(); ++k) h[ns - 1 - k] = nh_p[k];
        for (size_t k = 0; k < w.size(); ++k) h[ns + k] = w[k];
        for (size_t k = 0; k < cl.size(); ++k) h[ns + ni - 1 - k] = cl[k];
        M->pbuf.upload(h.data(), h.size());
        const int sz[4] = {(int)nh_n.size(), (int)nh_p.size(), (int)w.size(), (int)cl.size()};
        S->sizes.alloc(4);
        S->sizes.upload(sz, 4);
        // mode (MC/src/mc.cu:193-207): 4-point unless nh_n is empty or no 4-point move exists
        const bool can_four = sz[1] > 0 && sz[0] > 0 && sz[2] > 0 && sz[3] > 0, can_wc = sz[2] > 0 && sz[3] > 0;
        bool prefer_wc = sz[0] == 0;
        if (!prefer_wc && !can_four && can_wc) prefer_wc = true;
        if ((!prefer_wc && !can_four && !can_wc) || (prefer_wc && !can_wc)) S->K = 0;
        else S->K = prefer_wc ? 2 : 4;
        S->patience = d->patience;
        S->best_eps = d->best_eps;
        S->rec_cap = d->record_cap;
        S->rec_base = 0;
        S->rec_atoms.alloc(std::max<long long>(4 * d->record_cap, 1));
        S->rec_acc.alloc(std::max<long long>(d->record_cap, 1));
        S->rec_dE.alloc(std::max<long long>(d->record_cap, 1));
        hipLaunchKernelGGL(k_p2_reset, dim3(1), dim3(64), 0, stream(), M->ctl.p, S->sizes.p);
        PHB_HIP(hipGetLastError());
        S->next_recompute = M->d.recompute_interval > 0 ? M->d.recompute_interval : 0;
        sync();
        if (out) fill_p2(M, S, out);
        return PHB_OK;
    });
}

extern "C" int phb_phase2_run(phb_mc M, long long max_attempts, phb_phase2_stats* out) {
    return guard([&] {
        require(M && max_attempts >= 0, "phb_phase2_run: bad arguments");
        P2State* S = p2_of(M);
        McCtl c = read_ctl(M);
        if (S->K != 0 && !c.p2_stop && c.status == 0) {
            refresh(M);
            const long long target = c.p2_done + max_attempts;
            long long done = c.p2_done;
            while (done < target && c.status == 0 && !c.p2_stop) {
                long long chunk_end = target;
                if (S->next_recompute > 0) chunk_end = std::min(chunk_end, S->next_recompute);
                hipLaunchKernelGGL(k_p2_target, dim3(1), dim3(64), 0, stream(), M->ctl.p, chunk_end);
                while (done < chunk_end && c.status == 0 && !c.p2_stop) {
                    // a launch ends at its accepted move; batch size from the acceptance rate
                    const double acc = done > 0 ? (double)c.p2_acc / (double)done : 0.05;
                    const int batch = (int)std::max(1.0, std::min(std::ceil(std::max(acc, 1e-3) *
                                                                            (double)(chunk_end - done) * 1.25) + 1.0,
                                                                  64.0));
                    const long long before = done;
                    for (int b = 0; b < batch; ++b) {
                        launch_p2(M, S);
                        ++M->launches;
                        refresh(M);
                    }
                    c = read_ctl(M);
                    require(c.status != 2, "phb_phase2_run: engine stream rolled back past its kept block",
                            PHB_ERR_INTERNAL);
                    require(c.status != 3, "phb_phase2_run: accepted changes pending at a launch", PHB_ERR_INTERNAL);
                    require(c.p2_done > before || c.p2_stop || c.p2_done >= chunk_end,
                            "phb_phase2_run: the chain made no progress", PHB_ERR_INTERNAL);
                    done = c.p2_done;
                }
                if (done >= chunk_end && S->next_recompute > 0 && chunk_end == S->next_recompute) {
                    full_recompute(M);
                    S->next_recompute += M->d.recompute_interval;
                    c = read_ctl(M);
                }
            }
        }
        sync();
        if (out) fill_p2(M, S, out);
        return PHB_OK;
    });
}

extern "C" int phb_phase2_record_from(phb_mc M, long long first) {
    return guard([&] {
        require(M && first >= 0, "phb_phase2_record_from: bad arguments");
        P2State* S = p2_of(M);
        S->rec_base = first;
        return PHB_OK;
    });
}

extern "C" int phb_phase2_record(phb_mc M, long long n, int* atoms, signed char* acc, double* dE) {
    return guard([&] {
        require(M && n >= 0, "phb_phase2_record: bad arguments");
        P2State* S = p2_of(M);
        require(n <= S->rec_cap, "phb_phase2_record: n exceeds record_cap");
        if (n == 0) return PHB_OK;
        if (atoms) S->rec_atoms.download(atoms, 4 * (size_t)n);
        if (acc) S->rec_acc.download(acc, (size_t)n);
        if (dE) S->rec_dE.download(dE, (size_t)n);
        sync();
        return PHB_OK;
    });
}
