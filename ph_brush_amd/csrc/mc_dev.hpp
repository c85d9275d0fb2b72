// mc_dev.hpp -- device side of the protonation / ion-swap chain shared by
// the plain chain (mc.hip) and the delayed-acceptance chain (da.hip):
// control block, device view, MT19937 output stream, pool layout and the
// exact per-move Delta-E pieces.
#pragma once

#include <cmath>

#include "device.hpp"
#include "handles.hpp"
#include "pme.hpp"

namespace phb {

constexpr int MC_THREADS = 1024;
constexpr int MC_KMAX = 64;                     // pending atom changes
constexpr int MC_PMAX = 8;                      // spline order
constexpr int MC_CW = 2 * MC_PMAX - 1;          // correlation width
constexpr int REFRESH_BATCH = 8;                // pending atoms per structure-factor batch of the refresh

struct McCtl {
    long long trials_target, trials_done;
    long long accepted, acc_fwd, acc_bwd;
    long long rec_cap;
    int pending_n, pending_max, status, nprot;
    unsigned pend_seq;            // bumped by a chain launch that leaves a new pending list
    int pad_;
    int sizes[4];                 // nh_n, nh_p, w, cl
    double E, Q, Q2, Mz, Qz2;
    double E_recomputed, max_drift;
    unsigned mt[624];
    int mti;
    int pend_atom[MC_KMAX];
    double pend_dq[MC_KMAX];
    long long da_stage1, da_rounds;   // delayed-acceptance chain: stage-1 acceptances, rounds
    long long da_prof[8];             // delayed-acceptance phase times (100 MHz ticks) when profiling
    // phase-2 (4-point / W-Cl) chain: trials done and target, accepted, patience counter, stop flag, best E
    int nprot_target;                 // >= 0: the chain stops (status 4) once nprot equals it
    long long p2_done, p2_target, p2_acc;
    int p2_since, p2_stop;
    double p2_best;
    long long launch_cap;             // > 0: a chain launch stops after this many trials (batched replicas)
};

struct McDev {
    int N;
    const double *z;
    double* q;
    int* type;
    const int *fhead, *fnbr;
    const double *kreal, *r2;
    const int* sbase;     // 3 per atom: first stencil index per axis, wrapped
    const double* sw;     // 3*P per atom: weights x[P], y[P], z[P]
    const double *phi, *gk;
    const double* gf;     // g folded to |offset| per axis: (nz/2+1) x (ny/2+1) x (nx/2+1)
    const float *kreal32, *phi32, *gf32;   // FP32 copies for the FP32 form (null otherwise)
    int fnx, fny;         // nx/2+1, ny/2+1
    const double* selfw;  // per atom W_a^T G W_a (positions are fixed)
    int nx, ny, nz, P;
    int rd_ty, rd_tz;     // direct-refresh tile: rows along y, planes along z
    int T;
    const double *eps, *sig, *shift;
    int use_shift, use_cut;
    double lj_rc2;
    int* pbuf;            // pools, nsites + nions slots (layout at pool_slot)
    int nsites, nions;
    int W;                // speculative trials per round (1, 2, 4, 8 or 16)
    int t_nhn, t_nhp, t_w, t_cl;   // 0-based
    double beta, mu, elF, alpha, vol, Lz, self_coef;
    int use3dc;
    McCtl* ctl;
    // per refresh workgroup: the pending list (pend_seq) it has folded into phi;
    // entry 0 equal to pend_seq means the whole grid holds the pending changes
    unsigned* applied;
    signed char* rec_fwd;
    int *rec_i, *rec_j;
    signed char* rec_acc;
    double *rec_dE, *rec_u;
};

// Pending changes not yet in phi (a refresh kernel runs all its workgroups, so
// after one has run, entry 0 stands for every workgroup).
__device__ __forceinline__ int pending_unfolded(const McDev& D) {
    const McCtl* C = D.ctl;
    return D.applied[0] == C->pend_seq ? 0 : C->pending_n;
}

// ------------------------------------------------------------ MT19937 stream
constexpr unsigned MT_UPPER = 0x80000000u, MT_LOWER = 0x7fffffffu, MT_A = 0x9908b0dfu;

__device__ __forceinline__ unsigned mt_mix(unsigned a, unsigned b, unsigned c) {
    unsigned y = (a & MT_UPPER) | (b & MT_LOWER);
    return c ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
}

__device__ __forceinline__ unsigned mt_temper(unsigned y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// The standard sequential twist split into dependency phases
// [0,227) <- old; [227,454) <- phase 1; [454,623) <- phase 2; 623 <- new mt[0], mt[396].
// Executed by the 64 lanes of one wave; each phase computes into registers first.
__device__ void mt_twist_wave(unsigned* mt, int lane) {
    unsigned v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = lane + 64 * k;
        if (i < 227) v[k] = mt_mix(mt[i], mt[i + 1], mt[i + 397]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = lane + 64 * k;
        if (i < 227) mt[i] = v[k];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = 227 + lane + 64 * k;
        if (i < 454) v[k] = mt_mix(mt[i], mt[i + 1], mt[i - 227]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = 227 + lane + 64 * k;
        if (i < 454) mt[i] = v[k];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int i = 454 + lane + 64 * k;
        if (i < 623) v[k] = mt_mix(mt[i], mt[i + 1], mt[i - 227]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int i = 454 + lane + 64 * k;
        if (i < 623) mt[i] = v[k];
    }
    __builtin_amdgcn_wave_barrier();
    unsigned last = mt_mix(mt[623], mt[0], mt[396]);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) mt[623] = last;
    __builtin_amdgcn_wave_barrier();
}

// The engine's output stream addressed by position, so that speculative draws
// can be rolled back: block b (positions [624b, 624b+624)) is kept tempered in
// out[b & 1]; cur holds the raw state of the newest block, prev the one before.
struct RngLds {
    unsigned cur[624], prev[624];
    unsigned out[2][624];
};

// All lanes of wave 0 call with identical (blk, pos).
__device__ __forceinline__ unsigned rng_draw(RngLds& R, int& blk, int& pos, int lane) {
    if (pos >= (blk + 1) * 624) {
        for (int k = lane; k < 624; k += 64) R.prev[k] = R.cur[k];
        __builtin_amdgcn_wave_barrier();
        mt_twist_wave(R.cur, lane);
        const int h = (blk + 1) & 1;
        for (int k = lane; k < 624; k += 64) R.out[h][k] = mt_temper(R.cur[k]);
        __builtin_amdgcn_wave_barrier();
        ++blk;
    }
    const unsigned y = R.out[(pos / 624) & 1][pos % 624];
    ++pos;
    return y;
}

// Make positions up to pos_end (exclusive) readable; needs pos_end - pos < 624.
__device__ __forceinline__ void rng_ensure(RngLds& R, int& blk, int pos_end, int lane) {
    if (pos_end > (blk + 1) * 624) {
        for (int k = lane; k < 624; k += 64) R.prev[k] = R.cur[k];
        __builtin_amdgcn_wave_barrier();
        mt_twist_wave(R.cur, lane);
        const int h = (blk + 1) & 1;
        for (int k = lane; k < 624; k += 64) R.out[h][k] = mt_temper(R.cur[k]);
        __builtin_amdgcn_wave_barrier();
        ++blk;
    }
}

__device__ __forceinline__ unsigned rng_at(const RngLds& R, int q) { return R.out[(q / 624) & 1][q % 624]; }

__device__ __forceinline__ double canonical_at(const RngLds& R, int q) {
#pragma clang fp contract(off)
    const double g1 = (double)rng_at(R, q), g2 = (double)rng_at(R, q + 1);
    double r = (g1 + g2 * 4294967296.0) / 18446744073709551616.0;
    if (r >= 1.0) r = __longlong_as_double(0x3FEFFFFFFFFFFFFFLL);
    return r;
}

// Lemire draw at a fixed position; false when the rejection branch would need more draws.
__device__ __forceinline__ bool index_at(const RngLds& R, int q, unsigned n, unsigned& out) {
    const unsigned long long prod = (unsigned long long)rng_at(R, q) * n;
    const unsigned low = (unsigned)prod;
    out = (unsigned)(prod >> 32);
    return !(low < n && low < (0u - n) % n);
}

// libstdc++ uniform_int_distribution<int>(0, n-1) on a 32-bit engine: Lemire's
// nearly divisionless method with a 64-bit product.
__device__ __forceinline__ unsigned draw_index(RngLds& R, int& blk, int& pos, int lane, unsigned n) {
    unsigned long long prod = (unsigned long long)rng_draw(R, blk, pos, lane) * n;
    unsigned low = (unsigned)prod;
    if (low < n) {
        const unsigned thr = (0u - n) % n;
        while (low < thr) {
            prod = (unsigned long long)rng_draw(R, blk, pos, lane) * n;
            low = (unsigned)prod;
        }
    }
    return (unsigned)(prod >> 32);
}

// generate_canonical<double,53>: (g1 + g2*2^32) / 2^64, clamped below 1.
__device__ __forceinline__ double draw_canonical(RngLds& R, int& blk, int& pos, int lane) {
#pragma clang fp contract(off)
    const double g1 = (double)rng_draw(R, blk, pos, lane);
    const double g2 = (double)rng_draw(R, blk, pos, lane);
    double s = g1 + g2 * 4294967296.0;
    double r = s / 18446744073709551616.0;
    if (r >= 1.0) r = __longlong_as_double(0x3FEFFFFFFFFFFFFFLL);
    return r;
}

// ------------------------------------------------------------ pools
// The four pools share one buffer of nsites + nions slots: nh_n grows up from
// slot 0 and nh_p down from slot nsites-1; w grows up from nsites and cl down
// from nsites+nions-1.  Logical order is the reference's std::vector order.
__device__ __forceinline__ int pool_slot(const McDev& D, int pool, int k) {
    switch (pool) {
        case 0: return k;
        case 1: return D.nsites - 1 - k;
        case 2: return D.nsites + k;
        default: return D.nsites + D.nions - 1 - k;
    }
}

// ------------------------------------------------------------ Delta-E
template <typename T = double>
__device__ __forceinline__ T lj_pair(const McDev& D, int t, int tn, double r2) {
    if (D.use_cut && r2 > D.lj_rc2) return T(0);
    const int k = t * D.T + tn;
    const T eps = (T)D.eps[k], sig = (T)D.sig[k];
    if (eps <= T(0) || sig <= T(0)) return T(0);
    const T r = sqrt((T)r2);
    const T sr = sig / r, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    T e = T(4) * eps * (sr12 - sr6);
    if (D.use_shift) e -= (T)D.shift[k];
    return e;
}

// The per-move arrays in the value type T of a Delta-E evaluation.
template <typename T>
struct McArrays;
template <>
struct McArrays<double> {
    const double *kreal, *phi, *gf;
    __device__ explicit McArrays(const McDev& D) : kreal(D.kreal), phi(D.phi), gf(D.gf) {}
};
template <>
struct McArrays<float> {
    const float *kreal, *phi, *gf;
    __device__ explicit McArrays(const McDev& D) : kreal(D.kreal32), phi(D.phi32), gf(D.gf32) {}
};

// One speculative trial (LDS).
template <int P>
struct TrialT {
    int i, j, fwd, pos_prop;
    int ti_old, ti_new, tj_old, tj_new;
    double qi_old, qi_new, qj_old, qj_new;
    double zi, zj;
    int sb[2][3];            // first stencil index per axis of i and j
    int hrow[2], nrow[2];    // neighbour-row start and length of i and j
    double w[2][3][P];       // spline weights of i and j
    double self[2];          // W^T G W of i and j
    double u;                // uniform pre-drawn for the Metropolis test
};

template <int P>
struct PendT {               // atom changes since the last refresh of phi
    int atom[MC_KMAX];
    double dq[MC_KMAX];
    int sb[MC_KMAX][3];
    double w[MC_KMAX][3][P];
};

__device__ __forceinline__ int wrap1(int x, int n) {  // x in (-n, 2n)
    x += x < 0 ? n : 0;
    return x >= n ? x - n : x;
}

__device__ __forceinline__ int wrap2(int x, int n) {  // x in [-2n, 3n)
    x += x < 0 ? n : 0;
    x += x < 0 ? n : 0;
    x -= x >= n ? n : 0;
    return x >= n ? x - n : x;
}

__device__ __forceinline__ int fold(int x, int n) {   // x in [0, n) -> periodic |x|
    return 2 * x <= n ? x : n - x;
}

// Stencil correlation sum_{u - v = d} wa[u] wb[v] for d = dd - (P-1), in T.
template <int P, typename T = double>
__device__ __forceinline__ T corr(const double* wa, const double* wb, int dd) {
    const int d = dd - (P - 1);
    T c = T(0);
#pragma unroll
    for (int u = 0; u < P; ++u) {
        const int v = u - d;
        if (v >= 0 && v < P) c += (T)wa[u] * (T)wb[v];
    }
    return c;
}

// This thread's share (items lt, lt+nt, ...) of the pair, real-space and
// reciprocal Delta-E of trial t against phi plus the pending changes; `win`
// (optional) also receives the share of the mesh-kernel pair items.
//   items [0, Ri+Rj):       neighbour-row entries (atom i first, then j with i new)
//   items [.., + 2P^2):     one (z, y) row of dq_a W_a . phi
//   items [.., + (1+2np)CW): one dz plane of w_ab sum_d C(d) g(delta + d) for the
//                            pair (i,j) and the pairs (i|j, pending atom)
// The arithmetic of every item runs in T (per-thread sums in T as well); the
// caller reduces the per-thread results in FP64.
template <int P, typename T = double>
__device__ double trial_partial(const McDev& D, const TrialT<P>& t, const PendT<P>& pd, int pend_n, int lt, int nt,
                                double* win = nullptr, double* kij = nullptr) {
    constexpr int CW = 2 * P - 1, PP = P * P;
    const McArrays<T> arr(D);
    const T dqi = (T)(t.qi_new - t.qi_old), dqj = (T)(t.qj_new - t.qj_old);
    const int Ri = t.nrow[0], nA = Ri + t.nrow[1], nB = nA + 2 * PP;
    const int total = nB + (1 + 2 * pend_n) * CW * CW;
    T acc = T(0);
    for (int it = lt; it < total; it += nt) {
        if (it < nA) {
            const bool first = it < Ri;
            const int e = first ? t.hrow[0] + it : t.hrow[1] + (it - Ri);
            const int n = D.fnbr[e];
            double qn;
            int tn;
            if (kij && first && n == t.j) *kij = D.kreal[e];   // the (i, j) entry, when the pair is listed
            if (!first && n == t.i) { qn = t.qi_new; tn = t.ti_new; }
            else { qn = D.q[n]; tn = D.type[n]; }
            T v = (first ? dqi : dqj) * (T)qn * arr.kreal[e];
            const int to = first ? t.ti_old : t.tj_old, tw = first ? t.ti_new : t.tj_new;
            if (to != tw) {
                const double r2 = D.r2[e];
                v += lj_pair<T>(D, tw, tn, r2) - lj_pair<T>(D, to, tn, r2);
            }
            acc += v;
        } else if (it < nB) {
            const int s = it - nA, which = s / PP, r = s - which * PP, wz = r / P, wy = r - wz * P;
            const int* sb = t.sb[which];
            const int gz = wrap1(sb[2] + wz, D.nz), gy = wrap1(sb[1] + wy, D.ny);
            const T* row = arr.phi + ((size_t)gz * D.ny + gy) * D.nx;
            T sx = T(0);
#pragma unroll
            for (int u = 0; u < P; ++u) sx += (T)t.w[which][0][u] * row[wrap1(sb[0] + u, D.nx)];
            acc += (which ? dqj : dqi) * (T)t.w[which][2][wz] * (T)t.w[which][1][wy] * sx;
        } else {
            const int s = it - nB, pr = s / (CW * CW), r = s - pr * (CW * CW), dz = r / CW, dy = r - dz * CW;
            int ta;
            const int* sbb;
            const double(*wb)[P];
            T wgt;
            if (pr == 0) { ta = 0; sbb = t.sb[1]; wb = t.w[1]; wgt = dqi * dqj; }
            else {
                const int m = (pr - 1) >> 1;
                ta = (pr - 1) & 1;
                sbb = pd.sb[m]; wb = pd.w[m];
                wgt = (ta ? dqj : dqi) * (T)pd.dq[m];
            }
            const int* sba = t.sb[ta];
            const double(*wa)[P] = t.w[ta];
            // g is even along each axis: read the folded table at |offset|
            const int x0 = wrap1(sba[0] - sbb[0], D.nx) - (P - 1);
            const int gy = fold(wrap1(wrap1(sba[1] - sbb[1], D.ny) - (P - 1) + dy, D.ny), D.ny);
            const int gz = fold(wrap1(wrap1(sba[2] - sbb[2], D.nz) - (P - 1) + dz, D.nz), D.nz);
            const T* row = arr.gf + ((size_t)gz * D.fny + gy) * D.fnx;
            T sx = T(0);
#pragma unroll
            for (int d = 0; d < CW; ++d) sx += corr<P, T>(wa[0], wb[0], d) * row[fold(wrap1(x0 + d, D.nx), D.nx)];
            const T v = wgt * corr<P, T>(wa[2], wb[2], dz) * corr<P, T>(wa[1], wb[1], dy) * sx;
            acc += v;
            if (win) *win += (double)v;   // the mesh-kernel (pair) part on its own
        }
    }
    return (double)acc;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double slab_energy(const McDev& D, double Mz, double Q, double Qz2) {
    if (!D.use3dc) return 0.0;
    const double dc = (2.0 * M_PI / D.vol) * D.elF;
    double e = dc * Mz * Mz;
    if (fabs(Q) > 1e-4) e -= dc * Q * (Qz2 + Q * (D.Lz * D.Lz) / 12.0);
    return e;
}

// O(1) self, net-charge and slab terms of a move of charges qi -> qin, qj -> qjn
// (MC/src/pme.cu:425-447).
__device__ __forceinline__ double corr_move(const McDev& D, double qi, double qin, double qj, double qjn, double zi,
                                            double zj, double Q, double Mz, double Qz2) {
    const double dqi = qin - qi, dqj = qjn - qj;
    const double dQ2 = (qin * qin - qi * qi) + (qjn * qjn - qj * qj);
    const double Qn = Q + dqi + dqj;
    const double Mzn = Mz + dqi * zi + dqj * zj;
    const double Qz2n = Qz2 + dqi * zi * zi + dqj * zj * zj;
    const double e_self = -D.self_coef * dQ2;
    // alpha = 0 is the cut-off Coulomb form: no Ewald net-charge term
    const double e_q = D.alpha > 0.0 ? -(M_PI * D.elF) / (2.0 * D.vol * D.alpha * D.alpha) * (Qn * Qn - Q * Q) : 0.0;
    return e_self + e_q + (slab_energy(D, Mzn, Qn, Qz2n) - slab_energy(D, Mz, Q, Qz2));
}

// ... plus the mesh self terms 0.5 dq^2 W^T G W of the two atoms.
template <int P>
__device__ double correction_delta(const McDev& D, const TrialT<P>& t, double Q, double Q2, double Mz, double Qz2) {
    (void)Q2;
    const double dqi = t.qi_new - t.qi_old, dqj = t.qj_new - t.qj_old;
    return corr_move(D, t.qi_old, t.qi_new, t.qj_old, t.qj_new, t.zi, t.zj, Q, Mz, Qz2) +
           0.5 * dqi * dqi * t.self[0] + 0.5 * dqj * dqj * t.self[1];
}

// Metropolis exponent exactly as the reference forms it (MC/src/mc.cu:434-439).
__device__ __forceinline__ double mh_exponent(double beta, double dE, int s, double mu, long long Sb, long long Sa) {
#pragma clang fp contract(off)
    double logH = 0.0;
    if (Sb > 0 && Sa > 0) logH = log((double)Sb) - log((double)Sa);
    double expo = (double)((-beta) * (dE - (double)s * mu)) + logH;
    return fmin(fmax(expo, -40.0), 40.0);
}

// The move of a trial from its (i, j, direction) (any one thread; no memory reads).
template <int P>
__device__ __forceinline__ void stage_trial(const McDev& D, TrialT<P>& t, int i, int j, int fwd) {
    t.i = i; t.j = j; t.fwd = fwd;
    t.ti_new = fwd ? D.t_nhp : D.t_nhn; t.qi_new = fwd ? 1.0 : 0.0;
    t.tj_new = fwd ? D.t_cl : D.t_w; t.qj_new = fwd ? -1.0 : 0.0;
}

// Current identities, positions, stencils and rows of the trial atoms, loaded
// in parallel by `nthr` threads from `lt`.
template <int P>
__device__ __forceinline__ void stage_trial_data(const McDev& D, TrialT<P>& t, int lt, int nthr) {
    constexpr int NW = 2 * 3 * P;
    for (int k = lt; k < NW + 18; k += nthr) {
        if (k < NW) {
            const int a = k / (3 * P), r = k - a * 3 * P;
            const int atom = a ? t.j : t.i;
            (&t.w[a][0][0])[r] = D.sw[(size_t)atom * 3 * P + r];
        } else {
            const int q = k - NW;
            const int a = q / 9, r = q - a * 9;
            const int atom = a ? t.j : t.i;
            if (r < 3) t.sb[a][r] = D.sbase[3 * atom + r];
            else if (r == 3) t.hrow[a] = D.fhead[atom];
            else if (r == 4) t.nrow[a] = D.fhead[atom + 1] - D.fhead[atom];
            else if (r == 5) t.self[a] = D.selfw[atom];
            else if (r == 6) { if (a) t.tj_old = D.type[atom]; else t.ti_old = D.type[atom]; }
            else if (r == 7) { if (a) t.qj_old = D.q[atom]; else t.qi_old = D.q[atom]; }
            else { if (a) t.zj = D.z[atom]; else t.zi = D.z[atom]; }
        }
    }
}

template <int P>
__device__ __forceinline__ void stage_pending(const McDev& D, PendT<P>& pd, int k, int atom, double dq, int lt, int nthr) {
    for (int r = lt; r < 3 * P + 3; r += nthr) {
        if (r < 3 * P) (&pd.w[k][0][0])[r] = D.sw[(size_t)atom * 3 * P + r];
        else pd.sb[k][r - 3 * P] = D.sbase[3 * atom + (r - 3 * P)];
    }
    if (lt == 0) { pd.atom[k] = atom; pd.dq[k] = dq; }
}

// ------------------------------------------------------------ chain kernel
struct DecisionLds {
    int ntr, stop, acc_k;                    // trials in this round, stop flag, accepted trial (-1 none)
    int er_pool[2], er_val[2], er_n[2];      // pool erase: pool id, value, size before
    int pb_pool[2], pb_val[2], pb_n[2];      // pool push: pool id, value, size before
    int pos[2];
};

template <typename PoolT>
__device__ void pool_update(const McDev& D, PoolT* pbuf, DecisionLds& t) {
    const int tid = threadIdx.x, bd = blockDim.x;
    if (tid < 2) t.pos[tid] = -1;
    __syncthreads();
    for (int s = 0; s < 2; ++s)
        for (int k = tid; k < t.er_n[s]; k += bd)
            if ((int)pbuf[pool_slot(D, t.er_pool[s], k)] == t.er_val[s]) t.pos[s] = k;
    __syncthreads();
    constexpr int M = 8;
    for (int s = 0; s < 2; ++s) {
        const int pool = t.er_pool[s], n = t.er_n[s], pos = t.pos[s];
        if (pos < 0) continue;  // cannot happen for a consistent state; never shift from -1
        for (int base = pos + 1; base < n; base += M * bd) {
            PoolT v[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int k = base + tid + m * bd;
                v[m] = k < n ? pbuf[pool_slot(D, pool, k)] : PoolT(0);
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int k = base + tid + m * bd;
                if (k < n) pbuf[pool_slot(D, pool, k - 1)] = v[m];
            }
            __syncthreads();
        }
    }
    if (tid < 2) pbuf[pool_slot(D, t.pb_pool[tid], t.pb_n[tid])] = (PoolT)t.pb_val[tid];
    __syncthreads();
}

// pool_update with the erased values' positions already known (t.pos, from the
// indices their proposal drew: the pools hold each atom once and do not change
// between that proposal and its acceptance).  The two erasures shift disjoint
// pool regions, so both shifts run as one range, every load of a chunk issued
// before its stores.
template <typename PoolT>
__device__ void pool_update_at(const McDev& D, PoolT* pbuf, const DecisionLds& t) {
    const int tid = threadIdx.x, bd = blockDim.x;
    const int p0 = t.er_pool[0], p1 = t.er_pool[1];
    const int a0 = t.pos[0] + 1, a1 = t.pos[1] + 1;
    const int L0 = t.er_n[0] - a0, L = L0 + (t.er_n[1] - a1);
    constexpr int M = 16;
    for (int base = 0; base < L; base += M * bd) {
        PoolT v[M];
        int dst[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int e = base + tid + m * bd;
            dst[m] = -1;
            v[m] = PoolT(0);
            if (e < L) {
                const bool first = e < L0;
                const int pool = first ? p0 : p1, k = first ? a0 + e : a1 + (e - L0);
                v[m] = pbuf[pool_slot(D, pool, k)];
                dst[m] = pool_slot(D, pool, k - 1);
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (dst[m] >= 0) pbuf[dst[m]] = v[m];
        __syncthreads();
    }
    if (tid < 2) pbuf[pool_slot(D, t.pb_pool[tid], t.pb_n[tid])] = (PoolT)t.pb_val[tid];
    __syncthreads();
}

// One persistent workgroup runs the chain.  Each round, wave 0 draws up to W
// proposals assuming every one is rejected (a rejected trial always draws its
// Metropolis uniform), all waves evaluate their Delta-E concurrently (16/W
// waves per trial), then wave 0 decides them in order; at the first accepted
// trial the later ones are discarded and the engine's stream position rolls
// back to just after the accepted trial's own draws.  The decisions are
// therefore exactly those of the sequential chain.
struct ChainState {            // chain scalars, read and written by wave 0 only inside its sections
    long long done, acc_n, acc_f, acc_b;
    int sz[4];
    int status, nprot, blk, pos, pend_n;
    double E, Q, Q2, Mz, Qz2;
};


}  // namespace phb
