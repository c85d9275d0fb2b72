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

#include "device.hpp"
#include "handles.hpp"
#include "pme.hpp"

namespace phb {

constexpr int MC_THREADS = 1024;
constexpr int MC_KMAX = 64;                     // pending atom changes
constexpr int MC_PAIRS = 1 + 2 * MC_KMAX;       // (i,j), (i|j, pending)
constexpr int MC_PMAX = 8;                      // spline order
constexpr int MC_CW = 2 * MC_PMAX - 1;          // correlation width

struct McCtl {
    long long trials_target, trials_done;
    long long accepted, acc_fwd, acc_bwd;
    long long rec_cap;
    int pending_n, pending_max, status, nprot;
    int sizes[4];                 // nh_n, nh_p, w, cl
    double E, Q, Q2, Mz, Qz2;
    double E_recomputed, max_drift;
    unsigned mt[624];
    int mti;
    int pend_atom[MC_KMAX];
    double pend_dq[MC_KMAX];
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
    const double* selfw;  // per atom W_a^T G W_a (positions are fixed)
    int nx, ny, nz, P;
    int T;
    const double *eps, *sig, *shift;
    int use_shift, use_cut;
    double lj_rc2;
    int* pool[4];
    int t_nhn, t_nhp, t_w, t_cl;   // 0-based
    double beta, mu, elF, alpha, vol, Lz, self_coef;
    int use3dc;
    McCtl* ctl;
    signed char* rec_fwd;
    int *rec_i, *rec_j;
    signed char* rec_acc;
    double *rec_dE, *rec_u;
};

// ------------------------------------------------------------ MT19937
constexpr unsigned MT_UPPER = 0x80000000u, MT_LOWER = 0x7fffffffu, MT_A = 0x9908b0dfu;

__device__ __forceinline__ unsigned mt_mix(unsigned a, unsigned b, unsigned c) {
    unsigned y = (a & MT_UPPER) | (b & MT_LOWER);
    return c ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
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

// All 64 lanes call with identical mti (uniform control flow).
__device__ __forceinline__ unsigned mt_next(unsigned* mt, int& mti, int lane) {
    if (mti >= 624) { mt_twist_wave(mt, lane); mti = 0; }
    unsigned y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// libstdc++ uniform_int_distribution<int>(0, n-1) on a 32-bit engine: Lemire's
// nearly divisionless method with a 64-bit product.
__device__ __forceinline__ unsigned draw_index(unsigned* mt, int& mti, int lane, unsigned n) {
    unsigned long long prod = (unsigned long long)mt_next(mt, mti, lane) * n;
    unsigned low = (unsigned)prod;
    if (low < n) {
        const unsigned thr = (0u - n) % n;
        while (low < thr) {
            prod = (unsigned long long)mt_next(mt, mti, lane) * n;
            low = (unsigned)prod;
        }
    }
    return (unsigned)(prod >> 32);
}

// generate_canonical<double,53>: (g1 + g2*2^32) / 2^64, clamped below 1.
__device__ __forceinline__ double draw_canonical(unsigned* mt, int& mti, int lane) {
#pragma clang fp contract(off)
    const double g1 = (double)mt_next(mt, mti, lane);
    const double g2 = (double)mt_next(mt, mti, lane);
    double s = g1 + g2 * 4294967296.0;
    double r = s / 18446744073709551616.0;
    if (r >= 1.0) r = __longlong_as_double(0x3FEFFFFFFFFFFFFFLL);
    return r;
}

// ------------------------------------------------------------ Delta-E
struct MoveInfo {
    int i, j;
    int ti_old, ti_new, tj_old, tj_new;
    double qi_old, qi_new, qj_old, qj_new;
};

__device__ __forceinline__ double lj_pair(const McDev& D, int t, int tn, double r2) {
    if (D.use_cut && r2 > D.lj_rc2) return 0.0;
    const int k = t * D.T + tn;
    const double eps = D.eps[k], sig = D.sig[k];
    if (eps <= 0.0 || sig <= 0.0) return 0.0;
    const double r = sqrt(r2);
    const double sr = sig / r, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
    double e = 4.0 * eps * (sr12 - sr6);
    if (D.use_shift) e -= D.shift[k];
    return e;
}

struct PairLds {
    int npairs;
    int a[MC_PAIRS], b[MC_PAIRS];
    double w[MC_PAIRS];
    int dlt[MC_PAIRS][3];            // (base_a - base_b) mod n per axis
    double C[MC_PAIRS][3][MC_CW];    // stencil correlations per axis
};

// Block-wide Delta-E of the move in `mv` (LDS) against the current state,
// with the pending changes in pend_atom/pend_dq.  Returns the sum of pair,
// real-space and reciprocal terms in thread 0 (O(1) corrections excluded).
__device__ double block_delta_e(const McDev& D, const MoveInfo& mv, const int* pend_atom, const double* pend_dq,
                                int pend_n, PairLds& PL, double* sred) {
    const int tid = threadIdx.x, bd = blockDim.x;
    const int P = D.P, P3 = 3 * P, CW = 2 * P - 1;
    const double dqi = mv.qi_new - mv.qi_old, dqj = mv.qj_new - mv.qj_old;
    // pair list (the (i,i) and (j,j) self pairs use the precomputed W^T G W)
    if (tid == 0) {
        int n = 0;
        PL.a[n] = mv.i; PL.b[n] = mv.j; PL.w[n++] = dqi * dqj;
        for (int k = 0; k < pend_n; ++k) {
            PL.a[n] = mv.i; PL.b[n] = pend_atom[k]; PL.w[n++] = dqi * pend_dq[k];
            PL.a[n] = mv.j; PL.b[n] = pend_atom[k]; PL.w[n++] = dqj * pend_dq[k];
        }
        PL.npairs = n;
    }
    __syncthreads();
    const int npairs = PL.npairs;
    // stencil correlations C_axis(d) = sum_{u-v=d} w_a[u] w_b[v], d in [-(P-1), P-1]
    for (int it = tid; it < npairs * 3 * CW; it += bd) {
        const int pr = it / (3 * CW), rem = it - pr * 3 * CW, ax = rem / CW, d = rem - ax * CW - (P - 1);
        const double* wa = D.sw + (size_t)PL.a[pr] * P3 + ax * P;
        const double* wb = D.sw + (size_t)PL.b[pr] * P3 + ax * P;
        double c = 0.0;
        for (int u = max(0, d); u < min(P, P + d); ++u) c += wa[u] * wb[u - d];
        PL.C[pr][ax][d + P - 1] = c;
        if (d == 0) {
            const int n = ax == 0 ? D.nx : (ax == 1 ? D.ny : D.nz);
            PL.dlt[pr][ax] = wrap(D.sbase[3 * PL.a[pr] + ax] - D.sbase[3 * PL.b[pr] + ax], n);
        }
    }
    __syncthreads();
    const int hi = D.fhead[mv.i], Ri = D.fhead[mv.i + 1] - hi;
    const int hj = D.fhead[mv.j], Rj = D.fhead[mv.j + 1] - hj;
    const int nS = 2 * P * P;
    const int nX = npairs * CW * CW;
    const int total = Ri + Rj + nS + nX;
    double acc = tid == 0 ? 0.5 * dqi * dqi * D.selfw[mv.i] + 0.5 * dqj * dqj * D.selfw[mv.j] : 0.0;
    for (int it = tid; it < total; it += bd) {
        if (it < Ri + Rj) {
            // neighbour rows: atom i first (j still old), then atom j (i already new)
            const bool first = it < Ri;
            const int e = first ? hi + it : hj + (it - Ri);
            const int n = D.fnbr[e];
            double qn;
            int tn;
            if (!first && n == mv.i) { qn = mv.qi_new; tn = mv.ti_new; }
            else { qn = D.q[n]; tn = D.type[n]; }
            const double dq = first ? dqi : dqj;
            const int to = first ? mv.ti_old : mv.tj_old, tnw = first ? mv.ti_new : mv.tj_new;
            double v = dq * qn * D.kreal[e];
            if (to != tnw) {
                const double r2 = D.r2[e];
                v += lj_pair(D, tnw, tn, r2) - lj_pair(D, to, tn, r2);
            }
            acc += v;
        } else if (it < Ri + Rj + nS) {
            // dq_a * W_a . phi, one (z, y) stencil row per item
            const int s = it - Ri - Rj, which = s / (P * P), r = s - which * P * P, wz = r / P, wy = r - wz * P;
            const int a = which ? mv.j : mv.i;
            const double dq = which ? dqj : dqi;
            const int* sb = D.sbase + 3 * a;
            const double* w = D.sw + (size_t)a * P3;
            const int gz = wrap(sb[2] + wz, D.nz), gy = wrap(sb[1] + wy, D.ny);
            const double* row = D.phi + ((size_t)gz * D.ny + gy) * D.nx;
            double sx = 0.0;
            for (int u = 0; u < P; ++u) sx += w[u] * row[wrap(sb[0] + u, D.nx)];
            acc += dq * w[2 * P + wz] * w[P + wy] * sx;
        } else {
            // w_ab * sum_d C(d) g(delta + d), one (dz, dy) row per item
            const int s = it - Ri - Rj - nS, pr = s / (CW * CW), r = s - pr * CW * CW, dz = r / CW, dy = r - dz * CW;
            const int gz = wrap(PL.dlt[pr][2] + dz - (P - 1), D.nz), gy = wrap(PL.dlt[pr][1] + dy - (P - 1), D.ny);
            const double* row = D.gk + ((size_t)gz * D.ny + gy) * D.nx;
            const int x0 = PL.dlt[pr][0] - (P - 1);
            double sx = 0.0;
            for (int dx = 0; dx < CW; ++dx) sx += PL.C[pr][0][dx] * row[wrap(x0 + dx, D.nx)];
            acc += PL.w[pr] * PL.C[pr][2][dz] * PL.C[pr][1][dy] * sx;
        }
    }
    return block_sum(acc, sred);
}

__device__ __forceinline__ double slab_energy(const McDev& D, double Mz, double Q, double Qz2) {
    if (!D.use3dc) return 0.0;
    const double dc = (2.0 * M_PI / D.vol) * D.elF;
    double e = dc * Mz * Mz;
    if (fabs(Q) > 1e-4) e -= dc * Q * (Qz2 + Q * (D.Lz * D.Lz) / 12.0);
    return e;
}

// O(1) self, net-charge and slab terms of a move (MC/src/pme.cu:425-447).
__device__ double correction_delta(const McDev& D, const MoveInfo& mv, double Q, double Q2, double Mz, double Qz2) {
    const double dqi = mv.qi_new - mv.qi_old, dqj = mv.qj_new - mv.qj_old;
    const double zi = D.z[mv.i], zj = D.z[mv.j];
    const double dQ2 = (mv.qi_new * mv.qi_new - mv.qi_old * mv.qi_old) + (mv.qj_new * mv.qj_new - mv.qj_old * mv.qj_old);
    const double Qn = Q + dqi + dqj;
    const double Mzn = Mz + dqi * zi + dqj * zj;
    const double Qz2n = Qz2 + dqi * zi * zi + dqj * zj * zj;
    const double e_self = -D.self_coef * dQ2;
    const double e_q = -(M_PI * D.elF) / (2.0 * D.vol * D.alpha * D.alpha) * (Qn * Qn - Q * Q);
    return e_self + e_q + (slab_energy(D, Mzn, Qn, Qz2n) - slab_energy(D, Mz, Q, Qz2));
}

// Metropolis exponent exactly as the reference forms it (MC/src/mc.cu:434-439).
__device__ __forceinline__ double mh_exponent(double beta, double dE, int s, double mu, long long Sb, long long Sa) {
#pragma clang fp contract(off)
    double logH = 0.0;
    if (Sb > 0 && Sa > 0) logH = log((double)Sb) - log((double)Sa);
    double expo = (double)((-beta) * (dE - (double)s * mu)) + logH;
    return fmin(fmax(expo, -40.0), 40.0);
}

// ------------------------------------------------------------ chain kernel
struct TrialLds {
    MoveInfo mv;
    int stop, accepted, fwd;
    int er_pool[2], er_val[2], er_n[2];     // erase value from pool (size before erase)
    int pb_pool[2], pb_val[2], pb_n[2];     // push value onto pool (size before push)
    int pos[2];
    double dsum;
};

__device__ void pool_update(const McDev& D, TrialLds& t) {
    const int tid = threadIdx.x, bd = blockDim.x;
    if (tid < 2) t.pos[tid] = -1;
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
        const int* pool = D.pool[t.er_pool[s]];
        for (int k = tid; k < t.er_n[s]; k += bd)
            if (pool[k] == t.er_val[s]) t.pos[s] = k;
    }
    __syncthreads();
    // shift left by one after the erased slot; chunks of 32 elements per thread
    constexpr int M = 32;
    for (int s = 0; s < 2; ++s) {
        int* pool = D.pool[t.er_pool[s]];
        const int n = t.er_n[s], pos = t.pos[s];
        if (pos < 0) continue;  // not found: cannot happen for a consistent state; never shift from -1
        for (int base = pos + 1; base < n; base += M * bd) {
            int v[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int k = base + tid + m * bd;
                v[m] = k < n ? pool[k] : 0;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int k = base + tid + m * bd;
                if (k < n) pool[k - 1] = v[m];
            }
            __syncthreads();
        }
    }
    if (tid < 2) D.pool[t.pb_pool[tid]][t.pb_n[tid]] = t.pb_val[tid];
    __syncthreads();
}

__global__ __launch_bounds__(MC_THREADS) void k_mc_chain(McDev D) {
    __shared__ unsigned s_mt[624];
    __shared__ TrialLds s_t;
    __shared__ PairLds s_pl;
    __shared__ double s_red[16];
    __shared__ int s_pend_atom[MC_KMAX];
    __shared__ double s_pend_dq[MC_KMAX];
    McCtl* C = D.ctl;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < 624; k += blockDim.x) s_mt[k] = C->mt[k];
    for (int k = tid; k < MC_KMAX; k += blockDim.x) { s_pend_atom[k] = C->pend_atom[k]; s_pend_dq[k] = C->pend_dq[k]; }
    // wave-0 state (identical in every lane)
    int mti = C->mti;
    int sz[4] = {C->sizes[0], C->sizes[1], C->sizes[2], C->sizes[3]};
    long long done = C->trials_done;
    const long long target = C->trials_target, rec_cap = C->rec_cap;
    long long acc_n = C->accepted, acc_f = C->acc_fwd, acc_b = C->acc_bwd;
    int pend_n = C->pending_n, status = C->status, nprot = C->nprot;
    const int pend_max = C->pending_max;
    double E = C->E, Q = C->Q, Q2 = C->Q2, Mz = C->Mz, Qz2 = C->Qz2;
    __syncthreads();
    long long Sb = 0;
    for (;;) {
        if (wave == 0) {
            int stop = (done >= target || pend_n + 2 > pend_max || status != 0) ? 1 : 0;
            int fwd = 0;
            if (!stop) {
                const long long a = sz[0], b = sz[1], c = sz[2], d = sz[3];
                const long long wf = a * c, wb = b * d;
                Sb = wf + wb;
                if (wf > 0 && wb > 0) {
                    const double sum = (double)(wf + wb);
                    fwd = draw_canonical(s_mt, mti, lane) < (double)wf / sum ? 1 : 0;
                } else if (wf > 0) fwd = 1;
                else if (wb > 0) fwd = 0;
                else { status = 1; stop = 1; }
            }
            if (!stop) {
                int i, j;
                if (fwd) {
                    i = D.pool[0][draw_index(s_mt, mti, lane, (unsigned)sz[0])];
                    j = D.pool[2][draw_index(s_mt, mti, lane, (unsigned)sz[2])];
                } else {
                    i = D.pool[1][draw_index(s_mt, mti, lane, (unsigned)sz[1])];
                    j = D.pool[3][draw_index(s_mt, mti, lane, (unsigned)sz[3])];
                }
                if (lane == 0) {
                    MoveInfo& mv = s_t.mv;
                    mv.i = i; mv.j = j;
                    mv.ti_old = D.type[i]; mv.tj_old = D.type[j];
                    mv.qi_old = D.q[i]; mv.qj_old = D.q[j];
                    mv.ti_new = fwd ? D.t_nhp : D.t_nhn; mv.qi_new = fwd ? 1.0 : 0.0;
                    mv.tj_new = fwd ? D.t_cl : D.t_w; mv.qj_new = fwd ? -1.0 : 0.0;
                    s_t.fwd = fwd;
                }
            }
            if (lane == 0) s_t.stop = stop;
        }
        __syncthreads();
        if (s_t.stop) break;
        const double dsum = block_delta_e(D, s_t.mv, s_pend_atom, s_pend_dq, pend_n, s_pl, s_red);
        if (tid == 0) s_t.dsum = dsum;
        __syncthreads();
        if (wave == 0) {
            const MoveInfo mv = s_t.mv;
            const int fwd = s_t.fwd;
            const double dE = s_t.dsum + correction_delta(D, mv, Q, Q2, Mz, Qz2);
            const long long a = sz[0], b = sz[1], c = sz[2], d = sz[3];
            const long long Sa = fwd ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
            const double expo = mh_exponent(D.beta, dE, fwd ? 1 : -1, D.mu, Sb, Sa);
            bool acc = expo >= 0.0;
            double u = __longlong_as_double(0x7FF8000000000000LL);
            if (!acc) {
                u = draw_canonical(s_mt, mti, lane);
                acc = exp(expo) > u;
            }
            if (lane == 0 && done < rec_cap) {
                D.rec_fwd[done] = (signed char)fwd; D.rec_i[done] = mv.i; D.rec_j[done] = mv.j;
                D.rec_acc[done] = acc ? 1 : 0; D.rec_dE[done] = dE; D.rec_u[done] = u;
            }
            ++done;
            if (acc) {
                const double dqi = mv.qi_new - mv.qi_old, dqj = mv.qj_new - mv.qj_old;
                const double zi = D.z[mv.i], zj = D.z[mv.j];
                E += dE;
                Q2 += (mv.qi_new * mv.qi_new - mv.qi_old * mv.qi_old) + (mv.qj_new * mv.qj_new - mv.qj_old * mv.qj_old);
                Q += dqi + dqj;
                Mz += dqi * zi + dqj * zj;
                Qz2 += dqi * zi * zi + dqj * zj * zj;
                ++acc_n;
                if (fwd) { ++acc_f; ++nprot; } else { ++acc_b; --nprot; }
                if (lane == 0) {
                    D.q[mv.i] = mv.qi_new; D.type[mv.i] = mv.ti_new;
                    D.q[mv.j] = mv.qj_new; D.type[mv.j] = mv.tj_new;
                    s_pend_atom[pend_n] = mv.i; s_pend_dq[pend_n] = dqi;
                    s_pend_atom[pend_n + 1] = mv.j; s_pend_dq[pend_n + 1] = dqj;
                    // fwd: nh_n -> nh_p, w -> cl; bwd: nh_p -> nh_n, cl -> w
                    const int ps = fwd ? 0 : 1, pd = fwd ? 1 : 0, is = fwd ? 2 : 3, id = fwd ? 3 : 2;
                    s_t.er_pool[0] = ps; s_t.er_val[0] = mv.i; s_t.er_n[0] = sz[ps];
                    s_t.er_pool[1] = is; s_t.er_val[1] = mv.j; s_t.er_n[1] = sz[is];
                    s_t.pb_pool[0] = pd; s_t.pb_val[0] = mv.i; s_t.pb_n[0] = sz[pd];
                    s_t.pb_pool[1] = id; s_t.pb_val[1] = mv.j; s_t.pb_n[1] = sz[id];
                }
                pend_n += 2;
                if (fwd) { --sz[0]; ++sz[1]; --sz[2]; ++sz[3]; }
                else { ++sz[0]; --sz[1]; ++sz[2]; --sz[3]; }
            }
            if (lane == 0) s_t.accepted = acc ? 1 : 0;
        }
        __syncthreads();
        if (s_t.accepted) pool_update(D, s_t);
    }
    // write back
    for (int k = tid; k < 624; k += blockDim.x) C->mt[k] = s_mt[k];
    for (int k = tid; k < MC_KMAX; k += blockDim.x) { C->pend_atom[k] = s_pend_atom[k]; C->pend_dq[k] = s_pend_dq[k]; }
    if (tid == 0) {
        C->mti = mti;
        for (int k = 0; k < 4; ++k) C->sizes[k] = sz[k];
        C->trials_done = done;
        C->accepted = acc_n; C->acc_fwd = acc_f; C->acc_bwd = acc_b;
        C->pending_n = pend_n; C->status = status; C->nprot = nprot;
        C->E = E; C->Q = Q; C->Q2 = Q2; C->Mz = Mz; C->Qz2 = Qz2;
    }
}

// Delta-E of candidate moves, one block per move, no acceptance.
__global__ __launch_bounds__(MC_THREADS) void k_mc_probe(McDev D, int n, const int* __restrict__ mi,
                                                         const int* __restrict__ mj, const int* __restrict__ mf,
                                                         double* __restrict__ out) {
    __shared__ MoveInfo s_mv;
    __shared__ PairLds s_pl;
    __shared__ double s_red[16];
    __shared__ int s_pend_atom[MC_KMAX];
    __shared__ double s_pend_dq[MC_KMAX];
    const McCtl* C = D.ctl;
    const int m = blockIdx.x;
    if (m >= n) return;
    for (int k = threadIdx.x; k < MC_KMAX; k += blockDim.x) { s_pend_atom[k] = C->pend_atom[k]; s_pend_dq[k] = C->pend_dq[k]; }
    if (threadIdx.x == 0) {
        const int i = mi[m], j = mj[m], f = mf[m];
        s_mv.i = i; s_mv.j = j;
        s_mv.ti_old = D.type[i]; s_mv.tj_old = D.type[j];
        s_mv.qi_old = D.q[i]; s_mv.qj_old = D.q[j];
        s_mv.ti_new = f ? D.t_nhp : D.t_nhn; s_mv.qi_new = f ? 1.0 : 0.0;
        s_mv.tj_new = f ? D.t_cl : D.t_w; s_mv.qj_new = f ? -1.0 : 0.0;
    }
    __syncthreads();
    const double dsum = block_delta_e(D, s_mv, s_pend_atom, s_pend_dq, C->pending_n, s_pl, s_red);
    if (threadIdx.x == 0) out[m] = dsum + correction_delta(D, s_mv, C->Q, C->Q2, C->Mz, C->Qz2);
}

// ------------------------------------------------------------ refresh
__global__ void k_scatter_pending(McDev D, double* __restrict__ grid, int clear) {
    const McCtl* C = D.ctl;
    const int np = C->pending_n, P = D.P, P3 = P * P * P;
    for (int it = threadIdx.x; it < np * P3; it += blockDim.x) {
        const int k = it / P3, r = it - k * P3, wz = r / (P * P), wy = (r / P) % P, wx = r % P;
        const int a = C->pend_atom[k];
        const int* sb = D.sbase + 3 * a;
        const double* w = D.sw + (size_t)a * 3 * P;
        const size_t g = ((size_t)wrap(sb[2] + wz, D.nz) * D.ny + wrap(sb[1] + wy, D.ny)) * D.nx + wrap(sb[0] + wx, D.nx);
        if (clear) grid[g] = 0.0;
        else unsafeAtomicAdd(grid + g, C->pend_dq[k] * w[wx] * w[P + wy] * w[2 * P + wz]);
    }
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
    const double e_q = -(M_PI * D.elF) * (Q * Q) / (2.0 * D.vol * D.alpha * D.alpha);
    const double E = res[0] + (res[1] + res[2] + e_self + e_q + slab_energy(D, Mz, Q, Qz2));
    if (C->E_recomputed == C->E_recomputed)  // not the first recompute
        C->max_drift = fmax(C->max_drift, fabs(C->E - E));
    C->E = E; C->E_recomputed = E;
    C->Q = Q; C->Q2 = Q2; C->Mz = Mz; C->Qz2 = Qz2;
    C->pending_n = 0;
}

// ------------------------------------------------------------ setup kernels
__global__ void k_full_degree(int N, const int* __restrict__ hh, const int* __restrict__ hl, int* __restrict__ deg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = hh[i], e = hh[i + 1];
    atomicAdd(&deg[i], e - s);
    for (int k = s; k < e; ++k) atomicAdd(&deg[hl[k]], 1);
}

__global__ void k_full_fill(int N, const int* __restrict__ hh, const int* __restrict__ hl, const int* __restrict__ fh,
                            int* __restrict__ cursor, int* __restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int k = hh[i]; k < hh[i + 1]; ++k) {
        const int j = hl[k];
        tmp[fh[i] + atomicAdd(&cursor[i], 1)] = j;
        tmp[fh[j] + atomicAdd(&cursor[j], 1)] = i;
    }
}

__global__ __launch_bounds__(256) void k_full_sort(int N, const int* __restrict__ fh, const int* __restrict__ tmp,
                                                   int* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= N) return;
    const int s = fh[i], e = fh[i + 1];
    for (int a = s + lane; a < e; a += 64) {
        const int v = tmp[a];
        int rank = 0;
        for (int b = s; b < e; ++b) rank += tmp[b] < v ? 1 : 0;
        out[s + rank] = v;
    }
}

__global__ void k_full_coeffs(int N, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, double Lx, double Ly, double Lz, double elF, double alpha,
                              const int* __restrict__ fh, const int* __restrict__ fn, double* __restrict__ kreal,
                              double* __restrict__ r2o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int e = fh[i]; e < fh[i + 1]; ++e) {
        const int j = fn[e];
        const double dx = min_image_exact(x[i] - x[j], Lx), dy = min_image_exact(y[i] - y[j], Ly),
                     dz = min_image_exact(z[i] - z[j], Lz);
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

__global__ void k_delta_grid(double* __restrict__ g) {
    if (threadIdx.x == 0 && blockIdx.x == 0) g[0] = 1.0;
}

}  // namespace phb

using namespace phb;

// ----------------------------------------------------------------------- host
struct phb_mc_s {
    phb_atoms_s* atoms = nullptr;
    phb_params_s* params = nullptr;
    phb_csr_s* neigh = nullptr;
    phb_pme_s* pme = nullptr;
    phb_box box{};
    phb_mc_desc d{};
    int nsites = 0, nions = 0;
    DevBuf<int> fhead, fnbr, sbase;
    DevBuf<double> kreal, r2, sw, phi, gk, dgrid, selfw;
    DevBuf<int> pools[4];
    DevBuf<McCtl> ctl;
    DevBuf<signed char> rec_fwd, rec_acc;
    DevBuf<int> rec_i, rec_j;
    DevBuf<double> rec_dE, rec_u;
    McDev dev{};
    long long refreshes = 0, recomputes = 0, launches = 0, next_recompute = 0;
    // profiling: event triples around (chain launch, refresh) pairs and recomputes
    bool profiling = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<std::pair<size_t, int>> marks;   // (first event index, kind: 0 chain+refresh, 1 recompute)
    double chain_ms = 0, refresh_ms = 0, recompute_ms = 0;
    ~phb_mc_s() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    hipEvent_t event() {
        if (ev_used == ev.size()) {
            hipEvent_t e;
            PHB_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        hipEvent_t e = ev[ev_used++];
        PHB_HIP(hipEventRecord(e, stream()));
        return e;
    }
    // after a stream sync: fold recorded intervals into the totals
    void collect() {
        for (auto& m : marks) {
            float a = 0, b = 0;
            if (m.second == 0) {
                PHB_HIP(hipEventElapsedTime(&a, ev[m.first], ev[m.first + 1]));
                PHB_HIP(hipEventElapsedTime(&b, ev[m.first + 1], ev[m.first + 2]));
                chain_ms += a;
                refresh_ms += b;
            } else {
                PHB_HIP(hipEventElapsedTime(&a, ev[m.first], ev[m.first + 1]));
                recompute_ms += a;
            }
        }
        marks.clear();
        ev_used = 0;
    }
};

namespace {

McCtl read_ctl(phb_mc_s* M) {
    McCtl c;
    PHB_HIP(hipMemcpyAsync(&c, M->ctl.p, sizeof(McCtl), hipMemcpyDeviceToHost, stream()));
    sync();
    if (M->profiling) M->collect();
    return c;
}

// Full recompute: phi = G Q from scratch, total energy, moments, E reset.
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
    if (M->profiling) { M->event(); M->marks.push_back({e0, 1}); }
    ++M->recomputes;
}

void refresh(phb_mc_s* M) {
    phb_pme_s* P = M->pme;
    const int* gate = &M->ctl.p->pending_n;
    hipLaunchKernelGGL(k_scatter_pending, dim3(1), dim3(1024), 0, stream(), M->dev, M->dgrid.p, 0);
    PmeKernels::launch_fft_buf(P, M->dgrid.p, M->phi.p, 2, false, false, gate);
    hipLaunchKernelGGL(k_scatter_pending, dim3(1), dim3(1024), 0, stream(), M->dev, M->dgrid.p, 1);
    PmeKernels::launch_c2r_buf(P, M->phi.p, true, gate);
    hipLaunchKernelGGL(k_reset_pending, dim3(1), dim3(64), 0, stream(), M->ctl.p);
    PHB_HIP(hipGetLastError());
    ++M->refreshes;
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

}  // namespace

extern "C" int phb_mc_create(phb_atoms a, phb_params p, phb_csr neigh, phb_pme pme, const phb_box* box,
                             const phb_mc_desc* d, unsigned seed, phb_mc* out) {
    return guard([&] {
        require(a && p && neigh && pme && box && d && out, "phb_mc_create: null argument");
        require(p->d.coulombtype == PHB_COUL_PME, "phb_mc_create: the device chain needs coulombtype=pme",
                PHB_ERR_UNSUPPORTED);
        require(d->precision == PHB_FP64, "phb_mc_create: only PHB_FP64 is supported by the chain", PHB_ERR_UNSUPPORTED);
        require(pme->precision == PHB_FP64, "phb_mc_create: PME plan must be FP64", PHB_ERR_UNSUPPORTED);
        require(neigh->rows == a->n, "phb_mc_create: neighbour list rows != natoms");
        require(pme->nx % 2 == 0, "phb_mc_create: nx must be even (C2R potential grid)", PHB_ERR_UNSUPPORTED);
        require(a->max_type < p->d.ntypes, "phb_mc_create: atom type id outside the LJ table");
        const int T = p->d.ntypes;
        for (int t : {d->nh_n_type, d->nh_p_type, d->w_type, d->cl_type})
            require(t >= 1 && t <= T, "phb_mc_create: MC type id must be between 1 and ntypes");  // MC/src/io.cpp:289-299
        const int pm = d->pending_max > 0 ? d->pending_max : 32;
        require(pm >= 2 && pm <= MC_KMAX, "phb_mc_create: pending_max must be in [2, 64]");
        require(d->record_cap >= 0, "phb_mc_create: record_cap must be >= 0");
        auto M = std::make_unique<phb_mc_s>();
        M->atoms = a; M->params = p; M->neigh = neigh; M->pme = pme; M->box = *box; M->d = *d;
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
        { std::vector<int> f; for (int j : w) if (hz[j] < z_cut) f.push_back(j); w.swap(f); }
        for (int j = 0; j < N; ++j) if (ht[j] + 1 == d->cl_type && hz[j] < z_cut) cl.push_back(j);
        M->nsites = (int)(nh_n.size() + nh_p.size());
        M->nions = (int)(w.size() + cl.size());
        const std::vector<int>* pv[4] = {&nh_n, &nh_p, &w, &cl};
        for (int k = 0; k < 4; ++k) {
            const int cap = k < 2 ? M->nsites : M->nions;
            M->pools[k].alloc(std::max(cap, 1));
            M->pools[k].upload(pv[k]->data(), pv[k]->size());
        }
        // ---- full rows + coefficients
        hipStream_t st = stream();
        DevBuf<int> deg(N), cursor(N);
        PHB_HIP(hipMemsetAsync(deg.p, 0, N * sizeof(int), st));
        PHB_HIP(hipMemsetAsync(cursor.p, 0, N * sizeof(int), st));
        M->fhead.alloc(N + 1);
        hipLaunchKernelGGL(k_full_degree, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, neigh->head.p, neigh->list.p, deg.p);
        exclusive_scan(deg.p, M->fhead.p, N);
        const long long nfull = 2 * neigh->nnz;
        M->fnbr.alloc(std::max<long long>(nfull, 1));
        M->kreal.alloc(std::max<long long>(nfull, 1));
        M->r2.alloc(std::max<long long>(nfull, 1));
        if (nfull > 0) {
            DevBuf<int> tmp(nfull);
            hipLaunchKernelGGL(k_full_fill, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, neigh->head.p, neigh->list.p,
                               M->fhead.p, cursor.p, tmp.p);
            hipLaunchKernelGGL(k_full_sort, dim3(ceil_div(N, 4)), dim3(256), 0, st, N, M->fhead.p, tmp.p, M->fnbr.p);
            hipLaunchKernelGGL(k_full_coeffs, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, a->x.p, a->y.p, a->z.p,
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
        PHB_HIP(hipMemsetAsync(M->dgrid.p, 0, G * sizeof(double), st));
        M->selfw.alloc(N);
        hipLaunchKernelGGL(k_self_terms, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, P, pme->nx, pme->ny, pme->nz,
                           M->sw.p, M->gk.p, M->selfw.p);
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
        M->ctl.alloc(1);
        M->ctl.upload(&c, 1);
        // ---- device view
        McDev& D = M->dev;
        D.N = N; D.z = a->z.p; D.q = a->q.p; D.type = a->type.p;
        D.fhead = M->fhead.p; D.fnbr = M->fnbr.p; D.kreal = M->kreal.p; D.r2 = M->r2.p;
        D.sbase = M->sbase.p; D.sw = M->sw.p; D.phi = M->phi.p; D.gk = M->gk.p; D.selfw = M->selfw.p;
        D.nx = pme->nx; D.ny = pme->ny; D.nz = pme->nz; D.P = P;
        D.T = T; D.eps = p->d_eps.p; D.sig = p->d_sig.p; D.shift = p->d_shift.p;
        D.use_shift = p->d.lj_shift ? 1 : 0; D.use_cut = p->lj_cut_active ? 1 : 0; D.lj_rc2 = p->lj_rc * p->lj_rc;
        for (int k = 0; k < 4; ++k) D.pool[k] = M->pools[k].p;
        D.t_nhn = d->nh_n_type - 1; D.t_nhp = d->nh_p_type - 1; D.t_w = d->w_type - 1; D.t_cl = d->cl_type - 1;
        D.beta = d->beta; D.mu = d->mu_eff; D.elF = pme->elF; D.alpha = pme->alpha; D.vol = pme->vol;
        D.Lz = box->Lz; D.use3dc = pme->use3dc ? 1 : 0;
        D.self_coef = pme->elF * pme->alpha / std::sqrt(M_PI);
        D.ctl = M->ctl.p;
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

extern "C" int phb_mc_run(phb_mc M, long long ntrials, phb_mc_stats* out) {
    return guard([&] {
        require(M != nullptr && ntrials >= 0, "phb_mc_run: bad arguments");
        McCtl c = read_ctl(M);
        const long long target = c.trials_done + ntrials;
        long long done = c.trials_done;
        int status = c.status;
        while (done < target && status == 0) {
            long long chunk_end = target;
            if (M->next_recompute > 0) chunk_end = std::min(chunk_end, M->next_recompute);
            hipLaunchKernelGGL(k_set_target, dim3(1), dim3(64), 0, stream(), M->ctl.p, chunk_end);
            // each launch runs until the chunk ends or pending_max changes accumulate
            int batch = 4;
            while (done < chunk_end && status == 0) {
                for (int b = 0; b < batch; ++b) {
                    const size_t e0 = M->ev_used;
                    if (M->profiling) M->event();
                    hipLaunchKernelGGL(k_mc_chain, dim3(1), dim3(MC_THREADS), 0, stream(), M->dev);
                    ++M->launches;
                    if (M->profiling) M->event();
                    refresh(M);
                    if (M->profiling) { M->event(); M->marks.push_back({e0, 0}); }
                }
                PHB_HIP(hipGetLastError());
                c = read_ctl(M);
                require(c.trials_done > done || c.status != 0 || c.trials_done >= chunk_end,
                        "phb_mc_run: the chain made no progress", PHB_ERR_INTERNAL);
                done = c.trials_done;
                status = c.status;
                batch = std::min(batch * 2, 64);
            }
            if (done >= chunk_end && M->next_recompute > 0 && chunk_end == M->next_recompute) {
                full_recompute(M);
                M->next_recompute += M->d.recompute_interval;
            }
        }
        sync();
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
        hipLaunchKernelGGL(k_mc_probe, dim3(n), dim3(MC_THREADS), 0, stream(), M->dev, n, di.p, dj.p, df.p, dout.p);
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
        sync();
        if (M->profiling) M->collect();
        M->profiling = on != 0;
        M->chain_ms = M->refresh_ms = M->recompute_ms = 0.0;
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
