// mc_kernels.hpp -- bodies of the chain and potential-refresh kernels, shared
// by the single-replica launches (mc.hip) and the batched replicas (batch.hip).
#pragma once

#include "mc_dev.hpp"

namespace phb {

// One persistent workgroup runs the chain of replica D (see mc.hip).
template <int P, typename PoolT, bool LDS_POOLS, typename T = double>
__device__ __forceinline__ void chain_body(const McDev& D) {
    __shared__ RngLds s_rng;
    __shared__ TrialT<P> s_tr[16];
    __shared__ PendT<P> s_pd;
    __shared__ DecisionLds s_dc;
    __shared__ ChainState s_st;
    __shared__ double s_wsum[16];
    extern __shared__ __align__(16) unsigned char s_dyn[];
    McCtl* C = D.ctl;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = D.W, wpt = 16 / W;   // waves per trial
    const long long target = C->trials_target, rec_cap = C->rec_cap;
    // a launch queued beyond the target (batches are sized ahead) leaves at once
    if (C->trials_done >= target || C->status != 0) return;
    const int pend_max = C->pending_max, ntg = C->nprot_target;
    // the launch's own stop: the target, or launch_cap trials from here (a launch
    // boundary never changes the sequence: the next launch continues the stream)
    const long long stop_at = C->launch_cap > 0 && C->trials_done + C->launch_cap < target
                                  ? C->trials_done + C->launch_cap : target;
    PoolT* pbuf = LDS_POOLS ? reinterpret_cast<PoolT*>(s_dyn) : reinterpret_cast<PoolT*>(D.pbuf);
    if (LDS_POOLS)
        for (int k = tid; k < D.nsites + D.nions; k += blockDim.x) pbuf[k] = (PoolT)D.pbuf[k];
    for (int k = tid; k < 624; k += blockDim.x) {
        s_rng.cur[k] = C->mt[k];
        s_rng.out[0][k] = mt_temper(C->mt[k]);
    }
    const int pn0 = pending_unfolded(D);
    for (int k = wave; k < pn0; k += 16) stage_pending<P>(D, s_pd, k, C->pend_atom[k], C->pend_dq[k], lane, 64);
    if (tid == 0) {
        s_st.done = C->trials_done;
        s_st.acc_n = C->accepted; s_st.acc_f = C->acc_fwd; s_st.acc_b = C->acc_bwd;
        for (int k = 0; k < 4; ++k) s_st.sz[k] = C->sizes[k];
        s_st.status = C->status; s_st.nprot = C->nprot;
        s_st.blk = 0; s_st.pos = C->mti; s_st.pend_n = pn0;
        s_st.E = C->E; s_st.Q = C->Q; s_st.Q2 = C->Q2; s_st.Mz = C->Mz; s_st.Qz2 = C->Qz2;
    }
    __syncthreads();
    for (;;) {
        // ---- 1. proposals (wave 0).  Without a Lemire rejection every trial uses a
        // fixed number of engine outputs, so lane k draws trial k at its own
        // position; a round that meets a rejection is redrawn serially.
        if (wave == 0) {
            int blk = s_st.blk, pos = s_st.pos, status = s_st.status;
            const long long done = s_st.done;
            const int sz0 = s_st.sz[0], sz1 = s_st.sz[1], sz2 = s_st.sz[2], sz3 = s_st.sz[3];
            int ntr = 0, stop = 0;
            // the target comes first: when the trial that reaches it is also the last
            // of the launch's trials, the run still ends there (status 4), as the
            // reference returns from inside its block loop (MC/src/mc.cu:456-468)
            if (status != 0) stop = 1;
            else if (ntg >= 0 && s_st.nprot == ntg) { status = 4; stop = 1; }
            else if (done >= stop_at || s_st.pend_n + 2 > pend_max) stop = 1;
            else {
                const long long wf = (long long)sz0 * sz2, wb = (long long)sz1 * sz3;
                if (wf == 0 && wb == 0) { status = 1; stop = 1; }
                else {
                    const long long room = stop_at - done;
                    ntr = (int)(room < W ? room : W);
                    const double sum = (double)(wf + wb);
                    const bool both = wf > 0 && wb > 0;
                    const int per = (both ? 2 : 0) + 4;
                    rng_ensure(s_rng, blk, pos + ntr * per, lane);
                    bool ok = true;
                    int fwd = 0, ii = 0, jj = 0, pp = 0;
                    double u = 0.0;
                    if (lane < ntr) {
                        int q = pos + lane * per;
                        if (both) { fwd = canonical_at(s_rng, q) < (double)wf / sum ? 1 : 0; q += 2; }
                        else fwd = wf > 0 ? 1 : 0;
                        unsigned a1, a2;
                        ok = index_at(s_rng, q, (unsigned)(fwd ? sz0 : sz1), a1);
                        ok = index_at(s_rng, q + 1, (unsigned)(fwd ? sz2 : sz3), a2) && ok;
                        ii = (int)pbuf[pool_slot(D, fwd ? 0 : 1, (int)a1)];
                        jj = (int)pbuf[pool_slot(D, fwd ? 2 : 3, (int)a2)];
                        pp = q + 2;
                        u = canonical_at(s_rng, pp);
                    }
                    if (__all(ok)) {
                        if (lane < ntr) {
                            s_tr[lane].pos_prop = pp;
                            s_tr[lane].u = u;
                            stage_trial<P>(D, s_tr[lane], ii, jj, fwd);
                        }
                        pos += ntr * per;
                    } else {
                        for (int k = 0; k < ntr; ++k) {
                            int f2;
                            if (both) f2 = draw_canonical(s_rng, blk, pos, lane) < (double)wf / sum ? 1 : 0;
                            else f2 = wf > 0 ? 1 : 0;
                            int i2, j2;
                            if (f2) {
                                i2 = (int)pbuf[pool_slot(D, 0, (int)draw_index(s_rng, blk, pos, lane, (unsigned)sz0))];
                                j2 = (int)pbuf[pool_slot(D, 2, (int)draw_index(s_rng, blk, pos, lane, (unsigned)sz2))];
                            } else {
                                i2 = (int)pbuf[pool_slot(D, 1, (int)draw_index(s_rng, blk, pos, lane, (unsigned)sz1))];
                                j2 = (int)pbuf[pool_slot(D, 3, (int)draw_index(s_rng, blk, pos, lane, (unsigned)sz3))];
                            }
                            const int p2 = pos;
                            const double u2 = draw_canonical(s_rng, blk, pos, lane);
                            if (lane == 0) {
                                s_tr[k].pos_prop = p2;
                                s_tr[k].u = u2;
                                stage_trial<P>(D, s_tr[k], i2, j2, f2);
                            }
                        }
                    }
                }
            }
            if (lane == 0) {
                s_dc.ntr = ntr; s_dc.stop = stop;
                s_st.blk = blk; s_st.pos = pos; s_st.status = status;
            }
        }
        __syncthreads();
        if (s_dc.stop) break;
        const int ntr = s_dc.ntr, pend_n = s_st.pend_n;
        // ---- 2. stage per-trial data, then evaluate Delta-E (wpt waves per trial)
        const int k = wave / wpt, lt = (wave - k * wpt) * 64 + lane, nt = wpt * 64;
        if (k < ntr) stage_trial_data<P>(D, s_tr[k], lt, nt);
        __syncthreads();
        double part = 0.0;
        if (k < ntr) part = trial_partial<P, T>(D, s_tr[k], s_pd, pend_n, lt, nt);
        part = wave_sum(part);
        if (lane == 0) s_wsum[wave] = part;
        __syncthreads();
        // ---- 3. decisions (wave 0): lane k evaluates trial k against the round's
        // starting state; the first accepted trial ends the round.
        if (wave == 0) {
            const long long done = s_st.done;
            const long long a = s_st.sz[0], b = s_st.sz[1], c = s_st.sz[2], d = s_st.sz[3];
            const long long Sb = a * c + b * d;
            const double Q = s_st.Q, Q2 = s_st.Q2, Mz = s_st.Mz, Qz2 = s_st.Qz2;
            const int kk = lane;
            bool acc = false;
            double dE = 0.0, expo = 0.0, u = __longlong_as_double(0x7FF8000000000000LL);
            if (kk < ntr) {
                const TrialT<P>& t = s_tr[kk];
                double dsum = 0.0;
                for (int w = kk * wpt; w < (kk + 1) * wpt; ++w) dsum += s_wsum[w];
                dE = dsum + correction_delta<P>(D, t, Q, Q2, Mz, Qz2);
                const int fwd = t.fwd;
                const long long Sa = fwd ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
                expo = mh_exponent(D.beta, dE, fwd ? 1 : -1, D.mu, Sb, Sa);
                acc = expo >= 0.0;
                if (!acc) { u = t.u; acc = exp(expo) > u; }
            }
            const unsigned long long m = __ballot(acc);
            const int first = m ? __ffsll((long long)m) - 1 : ntr;
            const int nproc = first < ntr ? first + 1 : ntr;
            if (kk < nproc && done + kk < rec_cap) {
                const TrialT<P>& t = s_tr[kk];
                D.rec_fwd[done + kk] = (signed char)t.fwd; D.rec_i[done + kk] = t.i; D.rec_j[done + kk] = t.j;
                D.rec_acc[done + kk] = acc ? 1 : 0; D.rec_dE[done + kk] = dE; D.rec_u[done + kk] = u;
            }
            if (first < ntr && kk == first) {
                // accepted: state update, roll the stream back behind this trial's own draws
                const TrialT<P>& t = s_tr[kk];
                const int fwd = t.fwd;
                const int pos = t.pos_prop + (expo >= 0.0 ? 0 : 2);
                if (pos < s_st.blk * 624 - 624) s_st.status = 2;   // rollback beyond the kept block
                s_st.pos = pos;
                const double dqi = t.qi_new - t.qi_old, dqj = t.qj_new - t.qj_old;
                s_st.E += dE;
                s_st.Q2 = Q2 + ((t.qi_new * t.qi_new - t.qi_old * t.qi_old) + (t.qj_new * t.qj_new - t.qj_old * t.qj_old));
                s_st.Q = Q + (dqi + dqj);
                s_st.Mz = Mz + (dqi * t.zi + dqj * t.zj);
                s_st.Qz2 = Qz2 + (dqi * t.zi * t.zi + dqj * t.zj * t.zj);
                ++s_st.acc_n;
                if (fwd) { ++s_st.acc_f; ++s_st.nprot; } else { ++s_st.acc_b; --s_st.nprot; }
                D.q[t.i] = t.qi_new; D.type[t.i] = t.ti_new;
                D.q[t.j] = t.qj_new; D.type[t.j] = t.tj_new;
                // fwd: nh_n -> nh_p, w -> cl; bwd: nh_p -> nh_n, cl -> w
                const int ps = fwd ? 0 : 1, pdd = fwd ? 1 : 0, is = fwd ? 2 : 3, id = fwd ? 3 : 2;
                s_dc.er_pool[0] = ps; s_dc.er_val[0] = t.i; s_dc.er_n[0] = s_st.sz[ps];
                s_dc.er_pool[1] = is; s_dc.er_val[1] = t.j; s_dc.er_n[1] = s_st.sz[is];
                s_dc.pb_pool[0] = pdd; s_dc.pb_val[0] = t.i; s_dc.pb_n[0] = s_st.sz[pdd];
                s_dc.pb_pool[1] = id; s_dc.pb_val[1] = t.j; s_dc.pb_n[1] = s_st.sz[id];
                if (fwd) { --s_st.sz[0]; ++s_st.sz[1]; --s_st.sz[2]; ++s_st.sz[3]; }
                else { ++s_st.sz[0]; --s_st.sz[1]; ++s_st.sz[2]; --s_st.sz[3]; }
            }
            if (lane == 0) { s_dc.acc_k = first < ntr ? first : -1; s_st.done = done + nproc; }
        }
        __syncthreads();
        // ---- 4. accepted move: pending list and pools (all threads)
        const int acc_k = s_dc.acc_k;
        if (acc_k >= 0) {
            const TrialT<P>& t = s_tr[acc_k];
            if (wave == 1) stage_pending<P>(D, s_pd, pend_n, t.i, t.qi_new - t.qi_old, lane, 64);
            if (wave == 2) stage_pending<P>(D, s_pd, pend_n + 1, t.j, t.qj_new - t.qj_old, lane, 64);
            pool_update<PoolT>(D, pbuf, s_dc);
            if (tid == 0) s_st.pend_n = pend_n + 2;
            __syncthreads();
        }
    }
    // write back
    if (LDS_POOLS)
        for (int k = tid; k < D.nsites + D.nions; k += blockDim.x) D.pbuf[k] = (int)pbuf[k];
    const int pend_n = s_st.pend_n;
    if (wave == 0) {
        const int blk = s_st.blk, pos = s_st.pos;
        const bool in_cur = pos >= blk * 624;
        const unsigned* st = in_cur ? s_rng.cur : s_rng.prev;
        for (int k = lane; k < 624; k += 64) C->mt[k] = st[k];
        if (lane == 0) {
            C->mti = in_cur ? pos - blk * 624 : pos - (blk - 1) * 624;
            for (int k = 0; k < 4; ++k) C->sizes[k] = s_st.sz[k];
            C->trials_done = s_st.done;
            C->accepted = s_st.acc_n; C->acc_fwd = s_st.acc_f; C->acc_bwd = s_st.acc_b;
            if (pend_n > pn0) C->pend_seq = C->pend_seq + 1u;   // a new list: no workgroup has folded it
            C->pending_n = pend_n; C->status = s_st.status; C->nprot = s_st.nprot;
            C->E = s_st.E; C->Q = s_st.Q; C->Q2 = s_st.Q2; C->Mz = s_st.Mz; C->Qz2 = s_st.Qz2;
        }
    }
    for (int k = tid; k < pend_n; k += blockDim.x) { C->pend_atom[k] = s_pd.atom[k]; C->pend_dq[k] = s_pd.dq[k]; }
}


// Folds the pending charge changes of replica D into its potential grid, tile
// blk of the z pencils (see k_refresh_kspace in mc.hip).
template <int P>
__device__ __forceinline__ void refresh_kspace_tile(const McDev& D, cplx<double>* __restrict__ kg,
                                                    const cplx<double>* __restrict__ twx,
                                                    const cplx<double>* __restrict__ twy,
                                                    const cplx<double>* __restrict__ twz, const double* __restrict__ eta,
                                                    const FftPlan& pz, int tx, int blk) {
    const McCtl* C = D.ctl;
    const int np = pending_unfolded(D);   // the plane pass, which records the list, runs after this one
    if (np == 0) return;
    constexpr int NB = REFRESH_BATCH;   // pending atoms per batch
    __shared__ int s_sb[NB][3];
    __shared__ double s_w[NB][3][P];
    __shared__ double s_dq[NB];
    __shared__ cplx<double> s_fxy[NB][16];   // dq_p * Fx_p(kx0 + c) * Fy_p(ky), c < tx <= 16
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = D.nx, ny = D.ny, nz = D.nz, nkx = nx / 2 + 1, hs = nz + 1;
    cplx<double>* stw = reinterpret_cast<cplx<double>*>(smem_raw);
    cplx<double>* Cc = stw + nz;
    cplx<double>* s_fz = Cc + (size_t)tx * hs;        // [NB][nz]: Fz_p(kz), the same for every tile
    const int ntile = (nkx + tx - 1) / tx;
    const int ky = blk / ntile, kx0 = (blk - ky * ntile) * tx, cnt = min(tx, nkx - kx0);
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < nz; k += bd) stw[k] = twz[k];
    for (int pb = 0; pb < np; pb += NB) {
        const int nb = min(NB, np - pb);
        for (int r = tid; r < nb * (3 * P + 4); r += bd) {
            const int p = r / (3 * P + 4), q = r - p * (3 * P + 4), atom = C->pend_atom[pb + p];
            if (q < 3 * P) (&s_w[p][0][0])[q] = D.sw[(size_t)atom * 3 * P + q];
            else if (q < 3 * P + 3) s_sb[p][q - 3 * P] = D.sbase[3 * atom + (q - 3 * P)];
            else s_dq[p] = C->pend_dq[pb + p];
        }
        __syncthreads();
        // separable structure factors: dq Fx Fy per column, Fz per kz
        for (int r = tid; r < nb * cnt; r += bd) {
            const int p = r / cnt, c = r - p * cnt, kx = kx0 + c;
            cplx<double> fx = {0.0, 0.0}, fy = {0.0, 0.0};
            int ix = (kx * s_sb[p][0]) % nx, iy = (ky * s_sb[p][1]) % ny;   // k*(s+u) mod n, stepped by k
#pragma unroll
            for (int u = 0; u < P; ++u) {
                fx = cadd(fx, cscale(twx[ix], s_w[p][0][u]));
                fy = cadd(fy, cscale(twy[iy], s_w[p][1][u]));
                ix += kx; ix -= ix >= nx ? nx : 0;
                iy += ky; iy -= iy >= ny ? ny : 0;
            }
            s_fxy[p][c] = cscale(cmul(fx, fy), s_dq[p]);
        }
        for (int r = tid; r < nb * nz; r += bd) {
            const int p = r / nz, kz = r - p * nz;
            cplx<double> fz = {0.0, 0.0};
            int iz = (kz * s_sb[p][2]) % nz;
#pragma unroll
            for (int u = 0; u < P; ++u) {
                fz = cadd(fz, cscale(stw[iz], s_w[p][2][u]));
                iz += kz; iz -= iz >= nz ? nz : 0;
            }
            s_fz[(size_t)p * nz + kz] = fz;
        }
        __syncthreads();
        // kx fastest (the influence table is read coalesced below)
        for (int idx = tid; idx < cnt * nz; idx += bd) {
            const int kz = idx / cnt, c = idx - kz * cnt;
            cplx<double> v = pb == 0 ? cplx<double>{0.0, 0.0} : Cc[c * hs + kz];
            for (int p = 0; p < nb; ++p) v = cadd(v, cmul(s_fxy[p][c], s_fz[(size_t)p * nz + kz]));
            Cc[c * hs + kz] = v;
        }
        __syncthreads();
    }
    // multiply by the influence function; every load of a chunk issued first
    for (int base = tid; base < cnt * nz; base += 8 * bd) {
        double et[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * bd;
            et[u] = 0.0;
            if (idx < cnt * nz) {
                const int kz = idx / cnt, c = idx - kz * cnt;
                et[u] = eta[((size_t)kz * ny + ky) * nkx + kx0 + c];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * bd;
            if (idx < cnt * nz) {
                const int kz = idx / cnt, c = idx - kz * cnt;
                const cplx<double> v = Cc[c * hs + kz];
                Cc[c * hs + kz] = {v.x * et[u], v.y * et[u]};
            }
        }
    }
    __syncthreads();
    lds_fft(Cc, cnt, hs, pz, stw, +1);
    for (int idx = tid; idx < nz * cnt; idx += bd) {
        const int kz = idx / cnt, c = idx - kz * cnt;
        kg[((size_t)kz * ny + ky) * nkx + kx0 + c] = Cc[c * hs + kz];
    }
}

// ------------------------------------------------------------ direct refresh
// phi(r) += sum_p dq_p sum_{u,v,w} wx_p[u] wy_p[v] wz_p[w] G(r - s_p - (u, v, w)):
// the pending charge changes convolved with the real-space mesh kernel G (the
// potential of a unit mesh charge, read from its copy folded to |offset| per
// axis, which stays in L2).  Equal to the k-space refresh up to rounding (both
// are the same circular convolution), in one pass over the potential grid.
//
// Workgroup blk owns the full x rows of ty y by tz z mesh points (the MC object's
// tile shape, D.rd_ty x D.rd_tz, at most RD_TY x RD_TZ); lane =
// x (nx <= 64), so every row index is uniform across a wave.  RD_GROUPS groups
// of four waves take the pending atoms p = g, g + RD_GROUPS, ...: per atom a z
// pass over the halo rows read into registers, a y pass in place in LDS (rows
// of pitch 64) and an x pass that applies the x shift, accumulated in
// registers.  The group sums are added in group order, the grid read (at the
// start) and written once.  Each workgroup records the pending list's sequence
// number in its own applied[] entry: no state is shared between workgroups (a
// device-wide fence or a shared counter would cost more than the convolution).
// A workgroup wider than RD_THREADS (the persistent sweep's) runs a tile with its
// first RD_THREADS threads; the others only meet the barriers, so the
// arithmetic is the same as in the RD_THREADS launches.  Register arrays and LDS
// are sized for the largest tile; smaller tiles leave the rest unused.  Each shape
// is its own instantiation (the loops and register arrays are sized at compile time).
constexpr int RD_TY = 8, RD_TZ = 4, RD_GROUPS = 2, RD_THREADS = 256 * RD_GROUPS, RD_NXMAX = 64, RD_PITCH = 64;

__host__ __device__ constexpr size_t refresh_direct_lds(int P) {
    return (size_t)RD_GROUPS * RD_TZ * (RD_TY + P - 1) * RD_PITCH * sizeof(double);
}

__host__ __device__ constexpr int refresh_direct_blocks(int ny, int nz, int ty = RD_TY, int tz = RD_TZ) {
    return ((ny + ty - 1) / ty) * ((nz + tz - 1) / tz);
}

// tile shapes a refresh can take (rows along y, planes along z)
__host__ __device__ constexpr bool refresh_tile_ok(int ty, int tz) {
    return (ty == 8 || ty == 4) && (tz == 4 || tz == 2);
}

template <int P, int TY, int TZ>
__device__ __forceinline__ void refresh_direct_tile_shape(const McDev& D, double* __restrict__ phi, float* __restrict__ phi32,
                                                          int blk, int (*s_sb)[3], double (*s_w)[3 * P],
                                                          double* s_dq) {
    const McCtl* C = D.ctl;
    const unsigned seq = C->pend_seq;
    const int np = D.applied[blk] == seq ? 0 : C->pending_n;
    if (np == 0) return;
    constexpr int HY = TY + P - 1, HZ = TZ + P - 1;
    constexpr int RPW = (HY + 3) / 4, KPW = (TZ + 3) / 4, OPW = (TY * TZ + 3) / 4;
    constexpr int ABUF = TZ * HY * RD_PITCH;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* A_all = reinterpret_cast<double*>(smem_raw);   // RD_GROUPS x [TZ][HY][RD_PITCH]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wg = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wg >> 2, wv = wg & 3;
    const bool active = tid < RD_THREADS;   // groups grp >= RD_GROUPS only meet the barriers
    double* A = A_all + (active ? grp : 0) * ABUF;
    const int nx = D.nx, ny = D.ny, nz = D.nz, fnx = D.fnx, fny = D.fny;
    const int nty = (ny + TY - 1) / TY;
    const int tz = blk / nty, y0 = (blk - tz * nty) * TY, z0 = tz * TZ;
    const bool xin = lane < nx;
    const int fx = xin ? fold(lane, nx) : 0;
    for (int e = tid; active && e < np * (3 * P + 4); e += RD_THREADS) {
        const int p = e / (3 * P + 4), q = e - p * (3 * P + 4), atom = C->pend_atom[p];
        if (q < 3 * P) s_w[p][q] = D.sw[(size_t)atom * 3 * P + q];
        else if (q < 3 * P + 3) s_sb[p][q - 3 * P] = D.sbase[3 * atom + (q - 3 * P)];
        else s_dq[p] = C->pend_dq[p];
    }
    double old[OPW];   // group 0 holds the tile's grid values from the start
#pragma unroll
    for (int i = 0; i < OPW; ++i) {
        const int r = wv * OPW + i, k = r / TY, j = r - k * TY;
        const bool ok = grp == 0 && xin && r < TY * TZ && y0 + j < ny && z0 + k < nz;
        old[i] = ok ? phi[((size_t)(z0 + k) * ny + y0 + j) * nx + lane] : 0.0;
    }
    double acc[OPW];
#pragma unroll
    for (int i = 0; i < OPW; ++i) acc[i] = 0.0;
    __syncthreads();
    const int rounds = (np + RD_GROUPS - 1) / RD_GROUPS;
    for (int rd = 0; rd < rounds; ++rd) {
        const int p = rd * RD_GROUPS + grp;
        const bool act = active && p < np;
        const int pp = act ? p : 0;
        const int sx = __builtin_amdgcn_readfirstlane(s_sb[pp][0]);
        const int sy = __builtin_amdgcn_readfirstlane(s_sb[pp][1]);
        const int sz = __builtin_amdgcn_readfirstlane(s_sb[pp][2]);
        const double* w = s_w[pp];
        if (act) {   // z pass: halo row yy holds G at y index y0 - sy - (P-1) + yy, x index = lane
            const int zb = wrap2(z0 - sz - (P - 1), nz);
            int zoff[HZ];
#pragma unroll
            for (int zz = 0; zz < HZ; ++zz) {
                int gz = zb + zz;
                gz -= gz >= nz ? nz : 0;
                gz -= gz >= nz ? nz : 0;
                zoff[zz] = fold(gz, nz) * fny * fnx;
            }
            double g[RPW][HZ];
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int yy = wv + 4 * r;
                const double* row = D.gf + (size_t)fold(wrap2(y0 - sy - (P - 1) + yy, ny), ny) * fnx + fx;
#pragma unroll
                for (int zz = 0; zz < HZ; ++zz) g[r][zz] = yy < HY ? row[zoff[zz]] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int yy = wv + 4 * r;
                if (yy < HY) {
#pragma unroll
                    for (int k = 0; k < TZ; ++k) {
                        double a = 0.0;
#pragma unroll
                        for (int q = 0; q < P; ++q) a += w[2 * P + q] * g[r][k + P - 1 - q];
                        A[(k * HY + yy) * RD_PITCH + lane] = a;
                    }
                }
            }
        }
        __syncthreads();
        if (act) {   // y pass, in place: column (k, lane)
#pragma unroll
            for (int c = 0; c < KPW; ++c) {
                const int k = wv + 4 * c;
                if (k < TZ) {
                    double v[HY];
#pragma unroll
                    for (int yy = 0; yy < HY; ++yy) v[yy] = A[(k * HY + yy) * RD_PITCH + lane];
#pragma unroll
                    for (int j = 0; j < TY; ++j) {
                        double b = 0.0;
#pragma unroll
                        for (int u = 0; u < P; ++u) b += w[P + u] * v[j + P - 1 - u];
                        A[(k * HY + j) * RD_PITCH + lane] = b;
                    }
                }
            }
        }
        __syncthreads();
        if (act) {   // x pass with the x shift
            const double dq = s_dq[pp];
            const int xb = xin ? wrap2(lane - sx, nx) : 0;
            int xo[P];
            double wx[P];
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int xi = xb - u;
                xo[u] = xi + (xi < 0 ? nx : 0);
                wx[u] = w[u];
            }
#pragma unroll
            for (int i = 0; i < OPW; ++i) {
                const int r = wv * OPW + i;
                if (r < TY * TZ) {
                    const int k = r / TY, j = r - k * TY;
                    const double* row = A + (k * HY + j) * RD_PITCH;
                    double s = 0.0;
#pragma unroll
                    for (int u = 0; u < P; ++u) s += wx[u] * row[xo[u]];
                    acc[i] += dq * s;
                }
            }
        }
        __syncthreads();
    }
    // groups 1.. hand their sums to group 0 through LDS, added in group order
    for (int g2 = 1; g2 < RD_GROUPS; ++g2) {
        if (grp == g2) {
#pragma unroll
            for (int i = 0; i < OPW; ++i) A_all[(wv * OPW + i) * RD_PITCH + lane] = acc[i];
        }
        __syncthreads();
        if (grp == 0) {
#pragma unroll
            for (int i = 0; i < OPW; ++i) acc[i] += A_all[(wv * OPW + i) * RD_PITCH + lane];
        }
        __syncthreads();
    }
    if (grp == 0) {
#pragma unroll
        for (int i = 0; i < OPW; ++i) {
            const int r = wv * OPW + i, k = r / TY, j = r - k * TY;
            if (xin && r < TY * TZ && y0 + j < ny && z0 + k < nz) {
                const size_t ix = ((size_t)(z0 + k) * ny + y0 + j) * nx + lane;
                const double v = old[i] + acc[i];
                phi[ix] = v;
                if (phi32) phi32[ix] = (float)v;
            }
        }
    }
    __syncthreads();
    if (tid == 0) D.applied[blk] = seq;
}

// The tile of the object's shape (D.rd_ty x D.rd_tz; uniform across a launch).
template <int P>
__device__ __forceinline__ void refresh_direct_tile(const McDev& D, double* __restrict__ phi, float* __restrict__ phi32,
                                                    int blk) {
    __shared__ int s_sb[MC_KMAX][3];
    __shared__ double s_w[MC_KMAX][3 * P];
    __shared__ double s_dq[MC_KMAX];
    if (D.rd_ty == 8) {
        if (D.rd_tz == 4) refresh_direct_tile_shape<P, 8, 4>(D, phi, phi32, blk, s_sb, s_w, s_dq);
        else refresh_direct_tile_shape<P, 8, 2>(D, phi, phi32, blk, s_sb, s_w, s_dq);
    } else {
        if (D.rd_tz == 4) refresh_direct_tile_shape<P, 4, 4>(D, phi, phi32, blk, s_sb, s_w, s_dq);
        else refresh_direct_tile_shape<P, 4, 2>(D, phi, phi32, blk, s_sb, s_w, s_dq);
    }
}

// One shape only (a launch whose shape the host knows).
template <int P, int TY, int TZ>
__device__ __forceinline__ void refresh_direct_tile_fixed(const McDev& D, double* __restrict__ phi,
                                                          float* __restrict__ phi32, int blk) {
    __shared__ int s_sb[MC_KMAX][3];
    __shared__ double s_w[MC_KMAX][3 * P];
    __shared__ double s_dq[MC_KMAX];
    refresh_direct_tile_shape<P, TY, TZ>(D, phi, phi32, blk, s_sb, s_w, s_dq);
}

}  // namespace phb
