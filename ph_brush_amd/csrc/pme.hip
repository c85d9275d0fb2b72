// pme.hip -- short-range pair energies, SPME (spread, hand-written LDS FFT,
// influence-function convolution, reciprocal energy) and the Ewald self,
// net-charge and slab corrections.
//
// Replaces compute_lj_sr_energy_gpu / compute_coul_sr_energy_gpu
// (MC/src/energy_gpu.cu:14-135) and pmeCreate / pmeEnergy (MC/src/pme.cu:291-457).
//
// FFT layout: the real grid is x-fastest [nz][ny][nx]; the half spectrum is
// [nz][ny][nx/2+1].  k_fft_xy_r2c transforms one z-plane per workgroup
// entirely in LDS (pairs of real rows packed as one complex row for the x
// pass, untangled, then the y pass); k_fft_z loads a tile of kx-columns of
// one ky, transforms along z in LDS, applies the influence function (energy
// and/or convolution) and optionally transforms back; k_fft_xy_c2r is the
// inverse plane pass.  Every energy is reduced per workgroup into FP64
// partials and then summed in a fixed order (deterministic).
#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "device.hpp"
#include "handles.hpp"

#include "pme.hpp"

namespace phb {

// ------------------------------------------------------------ pair energy
enum { PAIR_LJ = 1, PAIR_COUL_CUT = 2, PAIR_EWALD = 4 };

struct PairArgs {
    const double *x, *y, *z, *q;
    const int* type;
    const int *head, *list;
    int N, T;
    double Lx, Ly, Lz;
    const double *eps, *sig, *shift;
    int use_shift, use_cut;
    double lj_rc2, kf, alpha;
};

constexpr int PAIR_LANES = 16;   // lanes per neighbour row in the pair sums (power of 2)
constexpr int SPT_MAX = 8;        // largest spread tile edge (grid points)
constexpr int SP_THREADS = 512;   // threads of a spread tile's workgroup (one per atom)
constexpr int SP_SCAN = 1024;     // threads of the one-block tile scan

// PAIR_LANES lanes per row i (grid-stride over rows); FP64 per-lane sums
// (T = double) or FP32 row sums (T = float) folded into FP64, block partials in FP64.

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_pair_energy(PairArgs A, double* __restrict__ partials) {
    __shared__ double sred[16];
    double e_lj = 0.0, e_c = 0.0;
    // PAIR_LANES lanes per row (lane l takes the entries l, l + PAIR_LANES, ... of
    // the row): enough independent gathers in flight at every N; the sums keep a
    // fixed order
    const int l = threadIdx.x & (PAIR_LANES - 1);
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) / PAIR_LANES; i < A.N;
         i += (gridDim.x * blockDim.x) / PAIR_LANES) {
        const double xi = A.x[i], yi = A.y[i], zi = A.z[i], qi = A.q[i];
        const int ti = A.type[i];
        T row_lj = 0, row_c = 0;
        for (int k = A.head[i] + l, ke = A.head[i + 1]; k < ke; k += PAIR_LANES) {
            const int j = A.list[k];
            const double dx = min_image_exact(xi - A.x[j], A.Lx);
            const double dy = min_image_exact(yi - A.y[j], A.Ly);
            const double dz = min_image_exact(zi - A.z[j], A.Lz);
            const double r2d = r2_exact(dx, dy, dz);
            const T r2 = (T)r2d;
            const T r = sqrt(r2);
            if (MODE & PAIR_LJ) {
                if (!A.use_cut || r2d <= A.lj_rc2) {
                    const int tj = A.type[j];
                    const T eps = (T)A.eps[ti * A.T + tj], sig = (T)A.sig[ti * A.T + tj];
                    if (eps > T(0) && sig > T(0)) {
                        const T sr = sig / r, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
                        T e = T(4.0) * eps * (sr12 - sr6);
                        if (A.use_shift) e -= (T)A.shift[ti * A.T + tj];
                        row_lj += e;
                    }
                }
            }
            if (MODE & (PAIR_COUL_CUT | PAIR_EWALD)) {
                const T qq = (T)(qi * A.q[j]);
                if (qq != T(0)) {
                    if (MODE & PAIR_COUL_CUT) row_c += (T)A.kf * qq * (T(1.0) / r);
                    else row_c += (T)A.kf * qq * (erfc((T)A.alpha * r) / r);
                }
            }
        }
        e_lj += (double)row_lj;
        e_c += (double)row_c;
    }
    double s0 = block_sum(e_lj, sred);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = s0;
    double s1 = block_sum(e_c, sred);
    if (threadIdx.x == 0) partials[2 * blockIdx.x + 1] = s1;
}

// ------------------------------------------------------------ corrections
// Q, sum q^2, Mz = sum q z, sum q z^2 (MC/src/pme.cu:427-443).
__global__ __launch_bounds__(256) void k_charge_moments(const double* __restrict__ q, const double* __restrict__ z,
                                                        int N, double* __restrict__ partials) {
    __shared__ double sred[16];
    double s[4] = {0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double qi = q[i], zi = z[i];
        s[0] += qi; s[1] += qi * qi; s[2] += qi * zi; s[3] += qi * zi * zi;
    }
    for (int c = 0; c < 4; ++c) {
        double v = block_sum(s[c], sred);
        if (threadIdx.x == 0) partials[4 * blockIdx.x + c] = v;
    }
}

// ------------------------------------------------------------ spread
template <typename T>
__global__ void k_fill(T* __restrict__ p, size_t n, T v) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

// Order-p B-spline spreading (MC/src/pme.cu:140-172 at p = 4) through
// LDS-privatised tiles, summed exactly in 64-bit fixed point.  The grid is cut
// into tiles of SPT_MAX points a side (fewer along an axis shorter than that);
// an atom belongs to the tile that holds its stencil origin
// (the lowest grid index its stencil touches along each axis), and the tile's
// workgroup sums its atoms' P^3 contributions into an LDS copy of the
// (tx + P - 1)(ty + P - 1)(tz + P - 1) points they reach, then adds the non-zero
// points to the grid with global integer atomics.  Every contribution is
// rounded once to units of 2^-S, S = 62 - ceil(log2(max|q| * fullest tile *
// tiles that reach a point)), so no partial sum can overflow and every sum is
// exact: the grid does not depend on the order in which atoms or tiles arrive
// and is the same bit for bit on every run.  Atoms with q = 0 contribute
// nothing and are skipped.  Five kernels: bin (tile counts, each atom's slot,
// max|q|), one-block scan of the counts (tile offsets, S; clears the counts and
// max|q| for the next call), fill (tile lists), tile sums, and the conversion
// to T that also clears the fixed-point grid.

// stencil origin of an atom along each axis (wrapped), as stencil_of places it
__device__ __forceinline__ void spread_origin(double x, double y, double z, const GridGeom& g, int o[3]) {
    const int off = stencil_offset(g.order);
    o[0] = wrap((int)floor((x / g.Lx) * g.nx) - off, g.nx);
    o[1] = wrap((int)floor((y / g.Ly) * g.ny) - off, g.ny);
    o[2] = wrap((int)floor((z / g.Lz) * g.nz) - off, g.nz);
}

struct SpreadTiles {
    int tx, ty, tz;      // tile edges (grid points)
    int tnx, tny, tnz;   // tiles along each axis
    __device__ __forceinline__ int tile_of(const int o[3]) const {
        return ((o[2] / tz) * tny + o[1] / ty) * tnx + o[0] / tx;
    }
};

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int d = 32; d > 0; d >>= 1) {
        const unsigned long long u = __shfl_xor(v, d);
        v = u > v ? u : v;
    }
    return v;
}

// bound[0]: bits of max|q| (non-negative doubles order as their bit patterns),
// bound[2]: S of the call.  Lanes of a wave that share a tile take their slots
// with one atomic (consecutive atoms are bonded neighbours, mostly in one tile).
__global__ __launch_bounds__(256) void k_spread_bin(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ q, int N,
                                                    GridGeom g, SpreadTiles st, int* __restrict__ count,
                                                    int* __restrict__ slot, unsigned long long* __restrict__ bound) {
    __shared__ unsigned long long s_q[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
    int t = -1;
    double aq = 0.0;
    if (i < N) {
        const double qi = q[i];
        if (qi != 0.0) {
            int o[3];
            spread_origin(x[i], y[i], z[i], g, o);
            t = st.tile_of(o);
            aq = fabs(qi);
        }
    }
    // group the wave's lanes by tile (one ballot per distinct tile), then every
    // group's leader takes the group's slots with one atomic, all at once
    int leader = lane, rank = 0, gsize = 0;
    unsigned long long todo = __ballot(t >= 0);
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        const int tl = __shfl(t, l);
        const unsigned long long m = __ballot(t == tl);
        if (t == tl) { leader = l; rank = __popcll(m & ((1ull << lane) - 1ull)); }
        if (lane == l) gsize = __popcll(m);
        todo &= ~m;
    }
    int base = 0;
    if (gsize > 0) base = atomicAdd(&count[t], gsize);
    base = __shfl(base, leader);
    if (t >= 0) slot[i] = base + rank;
    const unsigned long long qb = wave_max_u64((unsigned long long)__double_as_longlong(aq));
    if (lane == 0) s_q[threadIdx.x >> 6] = qb;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long v = s_q[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = s_q[w] > v ? s_q[w] : v;
        if (v > 0ull) atomicMax(&bound[0], v);
    }
}

// One block: exclusive scan of the tile counts into off[0..nt] (each thread a
// run of consecutive tiles), the fullest tile, and S; clears the counts and
// max|q| for the next call.
__global__ __launch_bounds__(SP_SCAN) void k_spread_scan(int* __restrict__ count, int nt, int* __restrict__ off,
                                                         double reach, unsigned long long* __restrict__ bound) {
    __shared__ int s_sum[SP_SCAN], s_max[SP_SCAN];
    const int tid = threadIdx.x, per = (nt + SP_SCAN - 1) / SP_SCAN;
    const int b = min(nt, tid * per), e = min(nt, b + per);
    int sum = 0, mx = 0;
    for (int k = b; k < e; ++k) {
        const int c = count[k];
        sum += c;
        mx = max(mx, c);
    }
    s_sum[tid] = sum;
    s_max[tid] = mx;
    __syncthreads();
    for (int d = 1; d < SP_SCAN; d <<= 1) {   // inclusive Hillis-Steele scan and running max
        const int vs = tid >= d ? s_sum[tid - d] : 0, vm = tid >= d ? s_max[tid - d] : 0;
        __syncthreads();
        s_sum[tid] += vs;
        s_max[tid] = max(s_max[tid], vm);
        __syncthreads();
    }
    int run = s_sum[tid] - sum;
    for (int k = b; k < e; ++k) {
        const int c = count[k];
        off[k] = run;
        run += c;
        count[k] = 0;
    }
    if (tid == SP_SCAN - 1) {
        off[nt] = s_sum[tid];
        int ex;
        (void)frexp(__longlong_as_double((long long)bound[0]) * (double)s_max[tid] * reach, &ex);
        bound[2] = (unsigned long long)(long long)(62 - ex);
        bound[0] = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_spread_fill(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ q, int N,
                                                     GridGeom g, SpreadTiles st, const int* __restrict__ off,
                                                     const int* __restrict__ slot, int* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || q[i] == 0.0) return;
    int o[3];
    spread_origin(x[i], y[i], z[i], g, o);
    list[off[st.tile_of(o)] + slot[i]] = i;
}

// One workgroup per tile, one thread per atom (weights in registers).
template <typename T, int P>
__global__ __launch_bounds__(SP_THREADS) void k_spread_tile(const double* __restrict__ x, const double* __restrict__ y,
                                                           const double* __restrict__ z, const double* __restrict__ q,
                                                           GridGeom g, SpreadTiles st, const int* __restrict__ off,
                                                           const int* __restrict__ list,
                                                           const unsigned long long* __restrict__ bound,
                                                           unsigned long long* __restrict__ gfx) {
    constexpr int RMAX = SPT_MAX + P - 1;
    __shared__ unsigned long long acc[RMAX * RMAX * RMAX];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int a0 = off[t], cnt = off[t + 1] - a0;
    if (cnt == 0) return;
    const int S = (int)(long long)bound[2];
    const int rx = st.tx + P - 1, ry = st.ty + P - 1, rz = st.tz + P - 1, r3 = rx * ry * rz;
    const int x0 = (t % st.tnx) * st.tx, y0 = ((t / st.tnx) % st.tny) * st.ty, z0 = (t / (st.tnx * st.tny)) * st.tz;
    for (int k = tid; k < r3; k += SP_THREADS) acc[k] = 0ull;
    __syncthreads();
    for (int k = tid; k < cnt; k += SP_THREADS) {
        const int i = list[a0 + k];
        const double qi = q[i];
        int i0[3];
        T w[3][8];
        stencil_of(x[i], y[i], z[i], g, i0, w);
        const int so = stencil_offset(P);
        const int ox = wrap(i0[0] - so, g.nx) - x0, oy = wrap(i0[1] - so, g.ny) - y0, oz = wrap(i0[2] - so, g.nz) - z0;
        // a listed atom's origin lies in its tile by construction (the same arithmetic)
        if (ox < 0 || ox >= st.tx || oy < 0 || oy >= st.ty || oz < 0 || oz >= st.tz) continue;
#pragma unroll
        for (int c = 0; c < P; ++c) {
#pragma unroll
            for (int b = 0; b < P; ++b) {
                const T wbc = (T)qi * w[1][b] * w[2][c];
                unsigned long long* row = acc + ((size_t)(oz + c) * ry + oy + b) * rx + ox;
#pragma unroll
                for (int a = 0; a < P; ++a) {
                    const long long v = llrint(ldexp((double)(wbc * w[0][a]), S));
                    if (v != 0) atomicAdd(row + a, (unsigned long long)v);
                }
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < r3; k += SP_THREADS) {
        const unsigned long long v = acc[k];
        if (v == 0ull) continue;
        const int lz = k / (rx * ry), ly = (k / rx) % ry, lx = k % rx;
        const size_t gi = ((size_t)wrap(z0 + lz, g.nz) * g.ny + wrap(y0 + ly, g.ny)) * g.nx + wrap(x0 + lx, g.nx);
        atomicAdd(gfx + gi, v);
    }
}

// grid = fixed-point sums * 2^-S; clears the sums for the next call.
template <typename T>
__global__ __launch_bounds__(256) void k_spread_fix(unsigned long long* __restrict__ gfx, size_t G,
                                                    const unsigned long long* __restrict__ bound, T* __restrict__ grid) {
    const int S = (int)(long long)bound[2];
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < G; k += (size_t)gridDim.x * blockDim.x) {
        const long long v = (long long)gfx[k];
        grid[k] = (T)ldexp((double)v, -S);
        gfx[k] = 0ull;
    }
}

// ------------------------------------------------------------ FFT kernels
template <typename T>
__global__ __launch_bounds__(512) void k_fft_xy_r2c(const T* __restrict__ grid, cplx<T>* __restrict__ kg, int nx, int ny, FftPlan px,
                             FftPlan py, const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy,
                             const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* stwx = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* stwy = stwx + nx;
    const int nyp = (ny + 1) / 2, nkx = nx / 2 + 1, hs = ny + 1;
    cplx<T>* A = stwy + ny;
    cplx<T>* H = A + nyp * nx;
    const int tid = threadIdx.x, bd = blockDim.x, zp = blockIdx.x;
    for (int k = tid; k < nx; k += bd) stwx[k] = twx[k];
    for (int k = tid; k < ny; k += bd) stwy[k] = twy[k];
    const T* plane = grid + (size_t)zp * nx * ny;
    for (int idx = tid; idx < nyp * nx; idx += bd) {
        const int yp = idx / nx, xx = idx - yp * nx, y0 = 2 * yp;
        const T re = plane[y0 * nx + xx];
        const T im = (y0 + 1 < ny) ? plane[(y0 + 1) * nx + xx] : T(0);
        A[idx] = {re, im};
    }
    __syncthreads();
    lds_fft(A, nyp, nx, px, stwx, -1);
    for (int idx = tid; idx < nyp * nkx; idx += bd) {
        const int yp = idx / nkx, k = idx - yp * nkx;
        const cplx<T> Zk = A[yp * nx + k], Zn = A[yp * nx + (k == 0 ? 0 : nx - k)];
        const cplx<T> a = {(Zk.x + Zn.x) * T(0.5), (Zk.y - Zn.y) * T(0.5)};
        const cplx<T> d = {Zk.x - Zn.x, Zk.y + Zn.y};
        H[k * hs + 2 * yp] = a;
        if (2 * yp + 1 < ny) H[k * hs + 2 * yp + 1] = {d.y * T(0.5), -d.x * T(0.5)};
    }
    __syncthreads();
    lds_fft(H, nkx, hs, py, stwy, -1);
    cplx<T>* out = kg + (size_t)zp * ny * nkx;
    for (int idx = tid; idx < ny * nkx; idx += bd) {
        const int ky = idx / nkx, kx = idx - ky * nkx;
        out[idx] = H[kx * hs + ky];
    }
}

// mode bit 1: energy sum corner*eta*|Q|^2 into partials[blockIdx.x];
// mode bit 2: multiply by eta, inverse z transform, store back;
// mode bit 4: store the forward transform (no eta) back.
template <typename T>
__global__ __launch_bounds__(512) void k_fft_z(cplx<T>* __restrict__ kg, int nx, int ny, int nz, FftPlan pz,
                        const cplx<T>* __restrict__ twz, const T* __restrict__ eta, const T* __restrict__ etac,
                        int tx, int mode, double* __restrict__ partials, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double sred[16];
    cplx<T>* stw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* C = stw + nz;
    const int nkx = nx / 2 + 1, hs = nz + 1;
    const int ntile = (nkx + tx - 1) / tx;
    const int ky = blockIdx.x / ntile, kx0 = (blockIdx.x - ky * ntile) * tx;
    const int cnt = min(tx, nkx - kx0);
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < nz; k += bd) stw[k] = twz[k];
    for (int idx = tid; idx < nz * cnt; idx += bd) {
        const int kz = idx / cnt, c = idx - kz * cnt;
        C[c * hs + kz] = kg[((size_t)kz * ny + ky) * nkx + kx0 + c];
    }
    __syncthreads();
    lds_fft(C, cnt, hs, pz, stw, -1);
    double e = 0.0;
    for (int idx = tid; idx < cnt * nz; idx += bd) {
        const int c = idx / nz, kz = idx - c * nz, kx = kx0 + c;
        const cplx<T> v = C[c * hs + kz];
        const size_t o = ((size_t)kz * ny + ky) * nkx + kx;
        if (mode & 1) {
            const T et = eta[o];
            const double corner = (kx == 0 || kx == nx / 2) ? 0.5 : 1.0;  // MC/src/pme.cu:234
            e += corner * (double)et * ((double)v.x * v.x + (double)v.y * v.y);
        }
        if (mode & 2) {
            const T ec = etac[o];
            C[c * hs + kz] = {v.x * ec, v.y * ec};
        }
    }
    if (mode & 1) {
        e = block_sum(e, sred);
        if (tid == 0) partials[blockIdx.x] = e;
    }
    if (mode & 2) {
        __syncthreads();
        lds_fft(C, cnt, hs, pz, stw, +1);
    }
    if (mode & 6) {
        __syncthreads();
        for (int idx = tid; idx < nz * cnt; idx += bd) {
            const int kz = idx / cnt, c = idx - kz * cnt;
            kg[((size_t)kz * ny + ky) * nkx + kx0 + c] = C[c * hs + kz];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(512) void k_fft_xy_c2r(const cplx<T>* __restrict__ kg, T* __restrict__ out, int nx, int ny, FftPlan px,
                             FftPlan py, const cplx<T>* __restrict__ twx, const cplx<T>* __restrict__ twy,
                             int accumulate, RefreshGate gate) {
    fft_xy_c2r_plane<T>(kg, out, nx, ny, px, py, twx, twy, accumulate, gate, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------ split plane passes
// For xy planes larger than one workgroup's LDS (the reference's pme_spacing on
// boxes wider than ~11 nm gives planes of 100^2 to 180^2 and more), the plane
// transform runs as two passes through global memory: x rows, then y columns.
// Rows are taken in pairs (y0, y0 + 1) of one z plane, packed as one complex row.

// Forward x pass: rpb row pairs per workgroup -> half-spectrum rows kg[z][y][kx].
template <typename T>
__global__ __launch_bounds__(512) void k_fft_x_r2c(const T* __restrict__ grid, cplx<T>* __restrict__ kg, int nx, int ny,
                                                   int nz, FftPlan px, const cplx<T>* __restrict__ twx, int rpb,
                                                   const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* stwx = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = stwx + nx;
    const int nyp = (ny + 1) / 2, nkx = nx / 2 + 1, npairs = nz * nyp;
    const int p0 = blockIdx.x * rpb, cnt = min(rpb, npairs - p0);
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < nx; k += bd) stwx[k] = twx[k];
    for (int idx = tid; idx < cnt * nx; idx += bd) {
        const int r = idx / nx, xx = idx - r * nx, pr = p0 + r, z = pr / nyp, y0 = 2 * (pr - z * nyp);
        const T* row = grid + ((size_t)z * ny + y0) * nx;
        A[idx] = {row[xx], y0 + 1 < ny ? row[nx + xx] : T(0)};
    }
    __syncthreads();
    lds_fft(A, cnt, nx, px, stwx, -1);
    for (int idx = tid; idx < cnt * nkx; idx += bd) {
        const int r = idx / nkx, k = idx - r * nkx, pr = p0 + r, z = pr / nyp, y0 = 2 * (pr - z * nyp);
        const cplx<T> Zk = A[r * nx + k], Zn = A[r * nx + (k == 0 ? 0 : nx - k)];
        cplx<T>* out = kg + ((size_t)z * ny + y0) * nkx;
        out[k] = {(Zk.x + Zn.x) * T(0.5), (Zk.y - Zn.y) * T(0.5)};
        if (y0 + 1 < ny) out[nkx + k] = {(Zk.y + Zn.y) * T(0.5), -(Zk.x - Zn.x) * T(0.5)};
    }
}

// y pass over kg[z][y][kx] in place: workgroup = (z, tile of tx kx columns);
// sign -1 forward, +1 inverse.  With a gate (the potential refresh's inverse
// pass) it runs only while a pending list is not yet folded (applied[0]).
template <typename T>
__global__ __launch_bounds__(512) void k_fft_y(cplx<T>* __restrict__ kg, int nx, int ny, FftPlan py,
                                               const cplx<T>* __restrict__ twy, int tx, int sign, const int* __restrict__ gate,
                                               RefreshGate rg) {
    if (gate && *gate == 0) return;
    if (rg.pending_n && (rg.applied[0] == *rg.seq || *rg.pending_n == 0)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* stw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* C = stw + ny;
    const int nkx = nx / 2 + 1, hs = ny + 1, ntile = (nkx + tx - 1) / tx;
    const int z = blockIdx.x / ntile, kx0 = (blockIdx.x - z * ntile) * tx, cnt = min(tx, nkx - kx0);
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < ny; k += bd) stw[k] = twy[k];
    cplx<T>* base = kg + (size_t)z * ny * nkx + kx0;
    for (int idx = tid; idx < ny * cnt; idx += bd) {
        const int y = idx / cnt, c = idx - y * cnt;
        C[c * hs + y] = base[(size_t)y * nkx + c];
    }
    __syncthreads();
    lds_fft(C, cnt, hs, py, stw, sign);
    for (int idx = tid; idx < ny * cnt; idx += bd) {
        const int y = idx / cnt, c = idx - y * cnt;
        base[(size_t)y * nkx + c] = C[c * hs + y];
    }
}

// Inverse x pass: half-spectrum rows (y0, y0 + 1) -> one complex row -> the two
// real rows, stored into or accumulated onto out.  With a gate each workgroup
// applies a pending list once and records it in applied[blockIdx.x].
template <typename T>
__global__ __launch_bounds__(512) void k_fft_x_c2r(const cplx<T>* __restrict__ kg, T* __restrict__ out, int nx, int ny,
                                                   int nz, FftPlan px, const cplx<T>* __restrict__ twx, int rpb,
                                                   int accumulate, RefreshGate gate) {
    if (gate.pending_n && (gate.applied[blockIdx.x] == *gate.seq || *gate.pending_n == 0)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* stwx = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = stwx + nx;
    const int nyp = (ny + 1) / 2, nkx = nx / 2 + 1, npairs = nz * nyp;
    const int p0 = blockIdx.x * rpb, cnt = min(rpb, npairs - p0);
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < nx; k += bd) stwx[k] = twx[k];
    for (int idx = tid; idx < cnt * nx; idx += bd) {
        const int r = idx / nx, k = idx - r * nx, pr = p0 + r, z = pr / nyp, y0 = 2 * (pr - z * nyp);
        const bool hb = y0 + 1 < ny;
        const cplx<T>* in = kg + ((size_t)z * ny + y0) * nkx;
        cplx<T> a, b;
        if (k < nkx) {
            a = in[k];
            b = hb ? in[nkx + k] : cplx<T>{T(0), T(0)};
            if (k == 0 || 2 * k == nx) { a.y = T(0); b.y = T(0); }
        } else {
            const int kk = nx - k;
            a = in[kk]; a.y = -a.y;
            b = hb ? in[nkx + kk] : cplx<T>{T(0), T(0)};
            b.y = -b.y;
        }
        A[idx] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    lds_fft(A, cnt, nx, px, stwx, +1);
    for (int idx = tid; idx < cnt * nx; idx += bd) {
        const int r = idx / nx, xx = idx - r * nx, pr = p0 + r, z = pr / nyp, y0 = 2 * (pr - z * nyp);
        T* row = out + ((size_t)z * ny + y0) * nx;
        const cplx<T> v = A[idx];
        if (accumulate) {
            row[xx] += v.x;
            if (y0 + 1 < ny) row[nx + xx] += v.y;
        } else {
            row[xx] = v.x;
            if (y0 + 1 < ny) row[nx + xx] = v.y;
        }
    }
    if (gate.applied) {
        __syncthreads();
        if (tid == 0) gate.applied[blockIdx.x] = *gate.seq;
    }
}

// ------------------------------------------------------------ host side
namespace {

int good_fft_dim(int n) {  // MC/src/pme.cu:46-51
    const int NMAX = std::max(16, n * 2 + 32);
    for (int m = std::max(8, n); m <= NMAX; ++m) if (fft_len_ok(m)) return m;
    return n;
}

double alpha_from_rtol(double rc, double rtol) {  // MC/src/pme.cu:54-63
    double beta = 5.0;
    int i = 0;
    while (std::erfc(beta * rc) > rtol) { beta *= 2.0; if (++i > 60) break; }
    double low = 0.0, high = beta;
    for (int it = 0; it < 60; ++it) {
        beta = 0.5 * (low + high);
        if (std::erfc(beta * rc) > rtol) low = beta; else high = beta;
    }
    return beta;
}

int fft_threads(long long elems, int min_threads = 256) {
    long long t = (elems + 14) / 15;
    int b = min_threads;
    while (b < t && b < 512) b *= 2;
    require(t <= 512, "FFT plane too large for one workgroup", PHB_ERR_UNSUPPORTED);
    return b;
}

size_t plan_csz(const phb_pme_s* P) {
    return P->precision == PHB_FP64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
}

// One workgroup holds the whole xy plane: at most 7,680 complex elements per LDS
// transform and 160 KiB of LDS.
bool plane_fits(const phb_pme_s* P) {
    const int nx = P->nx, ny = P->ny, nkx = nx / 2 + 1;
    const long long plane_elems = std::max<long long>((long long)((ny + 1) / 2) * nx, (long long)nkx * ny);
    return plane_elems <= 7680 && P->plane_lds(plan_csz(P)) <= 160 * 1024;
}

// Launch shapes of the plane transform, fused (one workgroup per plane) or split
// (x rows, then y columns, through global memory).
void set_plane_mode(phb_pme_s* P, bool split) {
    const int nx = P->nx, ny = P->ny, nz = P->nz, nkx = nx / 2 + 1, nyp = (ny + 1) / 2;
    const size_t csz = plan_csz(P);
    if (!split) {
        require(plane_fits(P), "pmeCreate: xy plane too large for one workgroup", PHB_ERR_UNSUPPORTED);
        // planes of >= 1024 elements get 8 waves (their LDS stages are latency-bound)
        const long long plane_elems = std::max<long long>((long long)nyp * nx, (long long)nkx * ny);
        P->plane_threads = fft_threads(plane_elems, plane_elems >= 1024 ? 512 : 256);
        P->split = false;
        return;
    }
    const size_t lds_cap = 64 * 1024 / csz;   // complex elements per workgroup (occupancy)
    require(nx <= 4096 && ny <= 4096 && (size_t)2 * nx <= lds_cap && (size_t)2 * ny + 1 <= lds_cap,
            "pmeCreate: grid lines longer than the row and column passes take", PHB_ERR_UNSUPPORTED);
    P->split = true;
    // x pass: row pairs per workgroup up to 4,096 complex elements
    int rpb = (int)std::min<size_t>(4096 / nx, (lds_cap - nx) / nx);
    rpb = std::max(1, std::min(rpb, nz * nyp));
    P->x_rpb = rpb;
    P->x_blocks = (nz * nyp + rpb - 1) / rpb;
    P->x_threads = fft_threads((long long)rpb * nx);
    // y pass: up to 16 kx columns per workgroup (256-byte row segments in FP64)
    int ytx = std::min(nkx, 16);
    ytx = std::min<int>(ytx, (int)std::min<size_t>(4096 / ny, (lds_cap - ny) / (ny + 1)));
    ytx = std::max(1, ytx);
    P->y_tx = ytx;
    P->y_blocks = nz * ((nkx + ytx - 1) / ytx);
    P->y_threads = fft_threads((long long)ytx * ny);
}

}  // namespace

// MC/src/pme.cu:77-113 with the near-zero fix described in the header.
std::vector<double> spme_moduli(int n, int pme_order) {
    const int splineOrder = pme_order - 1;
    require(splineOrder >= 1, "pme_order must be >= 2");
    std::vector<double> data(splineOrder, 0.0);
    data[0] = 1.0;
    for (int k = 2; k <= splineOrder; ++k) {
        const double invk = 1.0 / k;
        for (int m = k - 1; m > 0; --m) data[m] = invk * ((k - m) * data[m - 1] + (m + 1) * data[m]);
        data[0] = invk * data[0];
    }
    std::vector<double> mod(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double sc = 0.0, ss = 0.0;
        for (int j = 0; j < splineOrder; ++j) {
            const double arg = (2.0 * M_PI * i * (j + 1)) / double(n);
            sc += data[j] * std::cos(arg);
            ss += data[j] * std::sin(arg);
        }
        double v = sc * sc + ss * ss;
        if (!(v > 0.0)) v = 1e-30;
        mod[i] = v;
    }
    if (n % 2 == 0) { const int h = n / 2; if (!(mod[h] > 0.0)) mod[h] = mod[h - 1]; }
    std::vector<double> orig = mod;
    for (int i = 0; i < n; ++i)
        if (orig[i] < 1e-7) mod[i] = 0.5 * (orig[(i + n - 1) % n] + orig[(i + 1) % n]);
    return mod;
}

}  // namespace phb

using namespace phb;

phb_pme_s::~phb_pme_s() = default;

namespace phb {

static SpreadTiles spread_tiles_of(const phb_pme_s* P) {
    return SpreadTiles{P->sp_t[0], P->sp_t[1], P->sp_t[2], P->sp_tn[0], P->sp_tn[1], P->sp_tn[2]};
}

template <typename T, int P>
static void spread_tiles(phb_pme_s* Pl, const phb_atoms_s* a) {
    hipLaunchKernelGGL((k_spread_tile<T, P>), dim3(Pl->sp_nt), dim3(SP_THREADS), 0, stream(), a->x.p, a->y.p, a->z.p,
                       a->q.p, Pl->geom, spread_tiles_of(Pl), Pl->sp_off.p, a->sp_list.p, Pl->sp_bound.p, Pl->sp_fix.p);
}

template <typename T>
static void spread_impl(phb_pme_s* P, const phb_atoms_s* a, T* grid) {
    hipStream_t st = stream();
    const size_t G = (size_t)P->nx * P->ny * P->nz;
    const SpreadTiles tl = spread_tiles_of(P);
    if (a->n > 0)
        hipLaunchKernelGGL(k_spread_bin, dim3(ceil_div(a->n, 256)), dim3(256), 0, st, a->x.p, a->y.p, a->z.p, a->q.p,
                           a->n, P->geom, tl, P->sp_count.p, a->sp_slot.p, P->sp_bound.p);
    hipLaunchKernelGGL(k_spread_scan, dim3(1), dim3(SP_SCAN), 0, st, P->sp_count.p, P->sp_nt, P->sp_off.p, P->sp_reach,
                       P->sp_bound.p);
    if (a->n > 0) {
        hipLaunchKernelGGL(k_spread_fill, dim3(ceil_div(a->n, 256)), dim3(256), 0, st, a->x.p, a->y.p, a->z.p, a->q.p,
                           a->n, P->geom, tl, P->sp_off.p, a->sp_slot.p, a->sp_list.p);
        switch (P->order) {
            case 3: spread_tiles<T, 3>(P, a); break;
            case 4: spread_tiles<T, 4>(P, a); break;
            case 5: spread_tiles<T, 5>(P, a); break;
            case 6: spread_tiles<T, 6>(P, a); break;
            case 7: spread_tiles<T, 7>(P, a); break;
            default: spread_tiles<T, 8>(P, a); break;
        }
    }
    const int nb = (int)std::min<size_t>((G + 255) / 256, 4096);
    hipLaunchKernelGGL(k_spread_fix<T>, dim3(nb), dim3(256), 0, st, P->sp_fix.p, G, P->sp_bound.p, grid);
    PHB_HIP(hipGetLastError());
}

void PmeKernels::launch_spread(phb_pme_s* P, const phb_atoms_s* a) {
    if (P->precision == PHB_FP64) spread_impl<double>(P, a, P->grid64.p);
    else spread_impl<float>(P, a, P->grid32.p);
}

template <typename T>
static void fft_forward_impl(phb_pme_s* P, const T* grid, cplx<T>* kg, const cplx<T>* twx, const cplx<T>* twy,
                             const cplx<T>* twz, const T* eta, const T* etac, int zmode, double* partials,
                             const int* gate = nullptr) {
    hipStream_t st = stream();
    const int nx = P->nx, ny = P->ny, nz = P->nz;
    if (P->split) {
        hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(P->x_blocks), dim3(P->x_threads), P->x_lds(sizeof(cplx<T>)), st, grid,
                           kg, nx, ny, nz, P->px, twx, P->x_rpb, gate);
        hipLaunchKernelGGL(k_fft_y<T>, dim3(P->y_blocks), dim3(P->y_threads), P->y_lds(sizeof(cplx<T>)), st, kg, nx,
                           ny, P->py, twy, P->y_tx, -1, gate, RefreshGate{});
    } else {
        hipLaunchKernelGGL(k_fft_xy_r2c<T>, dim3(nz), dim3(P->plane_threads), P->plane_lds(sizeof(cplx<T>)), st, grid,
                           kg, nx, ny, P->px, P->py, twx, twy, gate);
    }
    hipLaunchKernelGGL(k_fft_z<T>, dim3(P->z_blocks), dim3(P->z_threads), P->z_lds(sizeof(cplx<T>)), st, kg, nx, ny,
                       nz, P->pz, twz, eta, etac, P->z_tx, zmode, partials, gate);
    PHB_HIP(hipGetLastError());
}

template <typename T>
static void fft_inverse_impl(phb_pme_s* P, cplx<T>* kg, T* out, const cplx<T>* twx, const cplx<T>* twy,
                             int accumulate, RefreshGate gate = {}) {
    if (P->split) {   // y columns in place on kg, then x rows into out
        hipLaunchKernelGGL(k_fft_y<T>, dim3(P->y_blocks), dim3(P->y_threads), P->y_lds(sizeof(cplx<T>)), stream(), kg,
                           P->nx, P->ny, P->py, twy, P->y_tx, +1, nullptr, gate);
        hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(P->x_blocks), dim3(P->x_threads), P->x_lds(sizeof(cplx<T>)), stream(),
                           kg, out, P->nx, P->ny, P->nz, P->px, twx, P->x_rpb, accumulate, gate);
    } else {
        hipLaunchKernelGGL(k_fft_xy_c2r<T>, dim3(P->nz), dim3(P->plane_threads), P->plane_lds(sizeof(cplx<T>)),
                           stream(), kg, out, P->nx, P->ny, P->px, P->py, twx, twy, accumulate, gate);
    }
    PHB_HIP(hipGetLastError());
}

// grid -> kgrid (-> eta, energy partials) [-> phi]
void PmeKernels::launch_fft(phb_pme_s* P, int zmode, bool inverse_to_phi, bool accumulate, double* partials) {
    if (P->precision == PHB_FP64)
        launch_fft_buf(P, P->grid64.p, P->phi64.p, zmode, inverse_to_phi, accumulate, nullptr, partials);
    else launch_fft_buf(P, P->grid32.p, P->phi32.p, zmode, inverse_to_phi, accumulate, nullptr, partials);
}

void PmeKernels::launch_fft_buf(phb_pme_s* P, const double* grid, double* phi, int zmode, bool inverse_to_phi,
                                bool accumulate, const int* gate, double* partials) {
    require(P->precision == PHB_FP64, "launch_fft_buf: precision mismatch", PHB_ERR_INTERNAL);
    fft_forward_impl<double>(P, grid, P->kg64.p, P->twx64.p, P->twy64.p, P->twz64.p, P->eta64.p, P->conv_eta64(),
                             zmode, partials ? partials : P->partials.p, gate);
    if (inverse_to_phi) fft_inverse_impl<double>(P, P->kg64.p, phi, P->twx64.p, P->twy64.p, accumulate);
}

void PmeKernels::launch_c2r_buf(phb_pme_s* P, double* phi, bool accumulate, RefreshGate gate) {
    fft_inverse_impl<double>(P, P->kg64.p, phi, P->twx64.p, P->twy64.p, accumulate, gate);
}

void PmeKernels::launch_c2r_buf(phb_pme_s* P, float* phi, bool accumulate, RefreshGate gate) {
    fft_inverse_impl<float>(P, P->kg32.p, phi, P->twx32.p, P->twy32.p, accumulate, gate);
}

void PmeKernels::launch_fft_buf(phb_pme_s* P, const float* grid, float* phi, int zmode, bool inverse_to_phi,
                                bool accumulate, const int* gate, double* partials) {
    require(P->precision == PHB_FP32, "launch_fft_buf: precision mismatch", PHB_ERR_INTERNAL);
    fft_forward_impl<float>(P, grid, P->kg32.p, P->twx32.p, P->twy32.p, P->twz32.p, P->eta32.p, P->conv_eta32(),
                            zmode, partials ? partials : P->partials.p, gate);
    if (inverse_to_phi) fft_inverse_impl<float>(P, P->kg32.p, phi, P->twx32.p, P->twy32.p, accumulate);
}

void PmeKernels::launch_pair(const phb_atoms_s* a, const phb_params_s* p, const phb_csr_s* nl, const phb_box* box,
                             int mode, double alpha, double kf, int precision, double* partials, int nblocks) {
    if (mode & PAIR_LJ) require(a->max_type < p->d.ntypes, "atom type id outside the LJ parameter table");
    PairArgs A;
    A.x = a->x.p; A.y = a->y.p; A.z = a->z.p; A.q = a->q.p; A.type = a->type.p;
    A.head = nl->head.p; A.list = nl->list.p;
    A.N = a->n; A.T = p->d.ntypes;
    A.Lx = box->Lx; A.Ly = box->Ly; A.Lz = box->Lz;
    A.eps = p->d_eps.p; A.sig = p->d_sig.p; A.shift = p->d_shift.p;
    A.use_shift = p->d.lj_shift ? 1 : 0;
    A.use_cut = p->lj_cut_active ? 1 : 0;
    A.lj_rc2 = p->lj_rc * p->lj_rc;
    A.kf = kf;
    A.alpha = alpha;
    hipStream_t st = stream();
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, st, A, partials); };
    if (precision == PHB_FP64) {
        switch (mode) {
            case PAIR_LJ: go(k_pair_energy<double, PAIR_LJ>); break;
            case PAIR_COUL_CUT: go(k_pair_energy<double, PAIR_COUL_CUT>); break;
            case PAIR_EWALD: go(k_pair_energy<double, PAIR_EWALD>); break;
            case PAIR_LJ | PAIR_COUL_CUT: go(k_pair_energy<double, PAIR_LJ | PAIR_COUL_CUT>); break;
            default: go(k_pair_energy<double, PAIR_LJ | PAIR_EWALD>); break;
        }
    } else {
        switch (mode) {
            case PAIR_LJ: go(k_pair_energy<float, PAIR_LJ>); break;
            case PAIR_COUL_CUT: go(k_pair_energy<float, PAIR_COUL_CUT>); break;
            case PAIR_EWALD: go(k_pair_energy<float, PAIR_EWALD>); break;
            case PAIR_LJ | PAIR_COUL_CUT: go(k_pair_energy<float, PAIR_LJ | PAIR_COUL_CUT>); break;
            default: go(k_pair_energy<float, PAIR_LJ | PAIR_EWALD>); break;
        }
    }
    PHB_HIP(hipGetLastError());
}

int pair_blocks(int N) { return std::max(1, std::min(ceil_div(PAIR_LANES * N, 256), 2048)); }

void PmeKernels::launch_moments(const phb_atoms_s* a, double* partials, int nblocks) {
    hipLaunchKernelGGL(k_charge_moments, dim3(nblocks), dim3(256), 0, stream(), a->q.p, a->z.p, a->n, partials);
    PHB_HIP(hipGetLastError());
}

// Self, net-charge and slab terms from the charge moments (MC/src/pme.cu:425-447).
void corrections(const phb_pme_s* P, const phb_box* box, const double mom[4], double* e_self, double* e_qcorr,
                 double* e_slab) {
    const double elF = P->elF, alpha = P->alpha, V = P->vol;
    const double Q = mom[0], Q2 = mom[1];
    *e_self = -elF * alpha / std::sqrt(M_PI) * Q2;
    *e_qcorr = alpha > 0.0 ? -(M_PI * elF) * (Q * Q) / (2.0 * V * alpha * alpha) : 0.0;
    double E_slab = 0.0;
    if (P->use3dc) {
        const double dc = (2.0 * M_PI / V) * elF;
        E_slab = dc * mom[2] * mom[2];
        if (std::abs(Q) > 1e-4) E_slab -= dc * Q * (mom[3] + Q * (box->Lz * box->Lz) / 12.0);
    }
    *e_slab = E_slab;
}

}  // namespace phb

// ----------------------------------------------------------------------- ABI
namespace phb {

// The plan behind pmeCreate (MC/src/pme.cu:291-335).  cut: the plan of the
// cut-off Coulomb form for the Monte Carlo chain -- alpha 0 (real-space
// coefficient elF/r exactly, no self or net-charge term), a zero influence
// function (no reciprocal part: zero mesh kernel and potential), no slab term,
// and the smallest grid the chain's stencils accept (8^3, order 4).
std::unique_ptr<phb_pme_s> build_pme(const phb_params_desc& d, const phb_box* box, double rc, int precision,
                                     bool cut) {
    require(precision == PHB_FP64 || precision == PHB_FP32, "pmeCreate: precision must be PHB_FP64 or PHB_FP32");
    require(box->Lx > 0 && box->Ly > 0 && box->Lz > 0, "pmeCreate: box must be positive");
    require(cut || (d.pme_order >= 3 && d.pme_order <= 8), "pmeCreate: pme_order must be in [3, 8]",
            PHB_ERR_UNSUPPORTED);
    require_device();
    auto P = std::make_unique<phb_pme_s>();
    P->precision = precision;
    P->order = cut ? 4 : d.pme_order;
    P->use3dc = !cut && d.pme_3dc != 0;
    // MC/src/pme.cu:299-306
    P->alpha = cut ? 0.0 : d.ewald_alpha > 0 ? d.ewald_alpha : alpha_from_rtol(rc, d.ewald_rtol > 0 ? d.ewald_rtol : 1e-5);
    P->ewF = cut ? 0.0 : (M_PI / P->alpha) * (M_PI / P->alpha);
    P->elF = kKe / d.epsilon_r;
    P->Lx = box->Lx; P->Ly = box->Ly; P->Lz = box->Lz;
    P->vol = box->Lx * box->Ly * box->Lz;
    // MC/src/pme.cu:312-318, EXT explicit dims
    const double sp = d.pme_spacing > 0 ? d.pme_spacing : 0.12;
    int nx = cut ? 8 : d.nx > 0 ? d.nx : good_fft_dim((int)std::ceil(box->Lx / sp));
    int ny = cut ? 8 : d.ny > 0 ? d.ny : good_fft_dim((int)std::ceil(box->Ly / sp));
    int nz = cut ? 8 : d.nz > 0 ? d.nz : good_fft_dim((int)std::ceil(box->Lz / sp));
    nx = std::max(nx, P->order); ny = std::max(ny, P->order); nz = std::max(nz, P->order);
    P->nx = nx; P->ny = ny; P->nz = nz;
    P->px = make_fft_plan(nx); P->py = make_fft_plan(ny); P->pz = make_fft_plan(nz);
    P->geom = GridGeom{nx, ny, nz, P->order, box->Lx, box->Ly, box->Lz};
    const int nkx = nx / 2 + 1;
    const size_t G = (size_t)nx * ny * nz, Gk = (size_t)nkx * ny * nz;
    // launch shapes: one workgroup per xy plane where the plane fits, else row and column passes
    set_plane_mode(P.get(), !plane_fits(P.get()));
    const size_t csz = precision == PHB_FP64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
    const int ntile0 = (nkx + 7) / 8;
    P->z_tx = (nkx + ntile0 - 1) / ntile0;
    P->z_blocks = ny * ((nkx + P->z_tx - 1) / P->z_tx);
    P->z_threads = fft_threads((long long)P->z_tx * nz);
    require(P->z_lds(csz) <= 160 * 1024, "pmeCreate: z pencil too large for LDS", PHB_ERR_UNSUPPORTED);
    // influence function (MC/src/pme.cu:196-223), half spectrum, no corner weight
    std::vector<double> mX = spme_moduli(nx, P->order), mY = spme_moduli(ny, P->order), mZ = spme_moduli(nz, P->order);
    std::vector<double> eta(Gk, 0.0);
    const double Hxx = 1.0 / box->Lx, Hyy = 1.0 / box->Ly, Hzz = 1.0 / box->Lz;
    for (int kz = 0; kz < (cut ? 0 : nz); ++kz)
        for (int ky = 0; ky < ny; ++ky)
            for (int kx = 0; kx < nkx; ++kx) {
                size_t o = ((size_t)kz * ny + ky) * nkx + kx;
                if (kx == 0 && ky == 0 && kz == 0) continue;
                int mx = kx, my = (ky <= ny / 2) ? ky : ky - ny, mz = (kz <= nz / 2) ? kz : kz - nz;
                double kxv = mx * Hxx, kyv = my * Hyy, kzv = mz * Hzz;
                double m2k = kxv * kxv + kyv * kyv + kzv * kzv;
                if (m2k == 0.0) continue;
                double denom = m2k * M_PI * P->vol * mX[kx] * mY[ky] * mZ[kz];
                if (!(denom > 0.0)) continue;
                eta[o] = P->elF * std::exp(-P->ewF * m2k) / denom;
            }
    P->eta_host = eta;
    // the convolutions' copy: the reference's half weight at kx = nx/2 for odd nx (MC/src/pme.cu:234)
    std::vector<double> etac;
    if (nx % 2 == 1) {
        etac = eta;
        for (size_t o = (size_t)(nx / 2); o < Gk; o += nkx) etac[o] *= 0.5;
    }
    auto tx = twiddles(nx), ty = twiddles(ny), tz = twiddles(nz);
    if (precision == PHB_FP64) {
        P->grid64.alloc(G); P->phi64.alloc(G); P->kg64.alloc(Gk); P->eta64.alloc(Gk);
        P->eta64.upload(eta.data(), Gk);
        if (!etac.empty()) { P->etac64.alloc(Gk); P->etac64.upload(etac.data(), Gk); }
        P->twx64.alloc(nx); P->twy64.alloc(ny); P->twz64.alloc(nz);
        P->twx64.upload(tx.data(), nx); P->twy64.upload(ty.data(), ny); P->twz64.upload(tz.data(), nz);
    } else {
        std::vector<float> e32(eta.begin(), eta.end());
        auto cv = [](const std::vector<cplx<double>>& v) {
            std::vector<cplx<float>> o(v.size());
            for (size_t k = 0; k < v.size(); ++k) o[k] = {(float)v[k].x, (float)v[k].y};
            return o;
        };
        auto fx = cv(tx), fy = cv(ty), fz = cv(tz);
        P->grid32.alloc(G); P->phi32.alloc(G); P->kg32.alloc(Gk); P->eta32.alloc(Gk);
        P->eta32.upload(e32.data(), Gk);
        if (!etac.empty()) {
            std::vector<float> c32(etac.begin(), etac.end());
            P->etac32.alloc(Gk);
            P->etac32.upload(c32.data(), Gk);
        }
        P->twx32.alloc(nx); P->twy32.alloc(ny); P->twz32.alloc(nz);
        P->twx32.upload(fx.data(), nx); P->twy32.upload(fy.data(), ny); P->twz32.upload(fz.data(), nz);
    }
    P->partials.alloc(std::max<size_t>(4 * 2048, (size_t)P->z_blocks));
    P->result.alloc(16);
    {   // spread tiles; counts, fixed-point sums and bounds start cleared and each call leaves them so
        const int n3[3] = {nx, ny, nz};
        double reach = 1.0;
        for (int ax = 0; ax < 3; ++ax) {
            P->sp_t[ax] = std::min(SPT_MAX, n3[ax]);
            P->sp_tn[ax] = ceil_div(n3[ax], P->sp_t[ax]);
            // tiles whose atoms reach one point along the axis (a narrow last tile adds one; a
            // stencil longer than the axis wraps onto itself)
            const int d = ceil_div(P->order - 1, P->sp_t[ax]) + 1 + (n3[ax] % P->sp_t[ax] != 0);
            reach *= (double)d * std::max(1, ceil_div(P->order, n3[ax]));
        }
        P->sp_reach = reach;
        P->sp_nt = P->sp_tn[0] * P->sp_tn[1] * P->sp_tn[2];
        const size_t nt = (size_t)P->sp_nt;
        P->sp_count.alloc(nt);
        P->sp_off.alloc(nt + 1);
        P->sp_fix.alloc(G);
        P->sp_bound.alloc(4);
        hipStream_t st = stream();
        hipLaunchKernelGGL(k_fill<int>, dim3((unsigned)std::min<size_t>((nt + 255) / 256, 4096)), dim3(256), 0, st,
                           P->sp_count.p, nt, 0);
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3((unsigned)std::min<size_t>((G + 255) / 256, 4096)),
                           dim3(256), 0, st, P->sp_fix.p, G, 0ull);
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, st, P->sp_bound.p, (size_t)4, 0ull);
        PHB_HIP(hipGetLastError());
    }
    sync();
    return P;
}

phb_pme_s* pme_create_cut(const phb_params_desc& d, const phb_box& box) {
    return build_pme(d, &box, 0.0, PHB_FP64, true).release();
}

}  // namespace phb

extern "C" int phb_pme_create(phb_params p, const phb_box* box, double rc, int precision, phb_pme* out) {
    return guard([&] {
        require(p && box && out, "pmeCreate: null plan");
        *out = phb::build_pme(p->d, box, rc, precision, false).release();
        return PHB_OK;
    });
}

extern "C" int phb_pme_destroy(phb_pme pme) {
    return guard([&] { delete pme; return PHB_OK; });
}

extern "C" int phb_pme_plane_mode(phb_pme P, int mode, int* split) {
    return guard([&] {
        require(P != nullptr, "phb_pme_plane_mode: null plan");
        require(mode >= -1 && mode <= 1, "phb_pme_plane_mode: mode must be -1 (query), 0 (one workgroup per plane) "
                                         "or 1 (row and column passes)");
        if (mode >= 0) {
            sync();
            set_plane_mode(P, mode == 1);
        }
        if (split) *split = P->split ? 1 : 0;
        return PHB_OK;
    });
}

extern "C" int phb_pme_info(phb_pme P, int* nx, int* ny, int* nz, int* order, double* alpha) {
    return guard([&] {
        require(P != nullptr, "pme_info: null plan");
        if (nx) *nx = P->nx;
        if (ny) *ny = P->ny;
        if (nz) *nz = P->nz;
        if (order) *order = P->order;
        if (alpha) *alpha = P->alpha;
        return PHB_OK;
    });
}

namespace {

// E_real (slot 0 of the returned pair) and the reciprocal + correction terms.
// part: max(4 * pair_blocks, z_blocks) doubles, res: 8 doubles (the plan's own
// buffers, or the scratch-free entry point's).
void pme_energy_core(phb_pme_s* P, const phb_atoms_s* a, const phb_csr_s* nl, const phb_params_s* par,
                     const phb_box* box, int pair_mode, double out[8], double* part, double* res) {
    // out: [pair_lj, pair_coul, recip, Q, Q2, Mz, Qz2, -]
    hipStream_t st = stream();
    const int nbp = pair_blocks(a->n);
    if (pair_mode) {
        PmeKernels::launch_pair(a, par, nl, box, pair_mode, P->alpha, P->elF, P->precision, part, nbp);
        reduce_partials(part, nbp, 2, 2, res + 0);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, st, res, (size_t)2, 0.0);
    }
    PmeKernels::launch_spread(P, a);
    // pair partials are reduced already (stream order); reuse the buffer
    PmeKernels::launch_fft(P, 1, false, false, part);
    reduce_partials(part, P->z_blocks, 1, 1, res + 2);
    PmeKernels::launch_moments(a, part, nbp);
    reduce_partials(part, nbp, 4, 4, res + 3);
    PHB_HIP(hipMemcpyAsync(out, res, 7 * sizeof(double), hipMemcpyDeviceToHost, st));
    PHB_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" int phb_energy_lj(phb_atoms a, phb_params p, phb_csr nl, const phb_box* box, double* E) {
    return guard([&] {
        require(a && p && nl && box && E, "compute_lj_sr_energy_gpu: null argument");
        require(nl->rows == a->n, "compute_lj_sr_energy_gpu: neighbour list rows != natoms");
        if (a->n == 0 || nl->nnz == 0) { *E = 0.0; return PHB_OK; }  // MC/src/energy_gpu.cu:54
        phb::DevBuf<double> part(2 * 2048), res(2);
        const int nb = pair_blocks(a->n);
        PmeKernels::launch_pair(a, p, nl, box, PAIR_LJ, 0.0, 0.0, PHB_FP64, part.p, nb);
        reduce_partials(part.p, nb, 2, 2, res.p);
        double h[2];
        res.download(h, 2);
        sync();
        *E = h[0];
        return PHB_OK;
    });
}

extern "C" int phb_energy_coul_cut(phb_atoms a, phb_params p, phb_csr nl, const phb_box* box, double* E) {
    return guard([&] {
        require(a && p && nl && box && E, "compute_coul_sr_energy_gpu: null argument");
        require(nl->rows == a->n, "compute_coul_sr_energy_gpu: neighbour list rows != natoms");
        if (a->n == 0 || nl->nnz == 0) { *E = 0.0; return PHB_OK; }  // MC/src/energy_gpu.cu:116
        phb::DevBuf<double> part(2 * 2048), res(2);
        const int nb = pair_blocks(a->n);
        PmeKernels::launch_pair(a, p, nl, box, PAIR_COUL_CUT, 0.0, kKe / p->d.epsilon_r, PHB_FP64, part.p, nb);
        reduce_partials(part.p, nb, 2, 2, res.p);
        double h[2];
        res.download(h, 2);
        sync();
        *E = h[1];
        return PHB_OK;
    });
}

extern "C" int phb_pme_energy(phb_pme P, phb_atoms a, phb_csr nl, const phb_box* box, phb_pme_components* comps,
                              double* E) {
    return guard([&] {
        require(P && a && nl && box && E, "pmeEnergy: null argument");
        require(nl->rows == a->n, "pmeEnergy: neighbour list rows != natoms");
        if (a->n == 0) {  // MC/src/pme.cu:350-359
            if (comps) *comps = phb_pme_components{0, 0, 0, 0, 0};
            *E = 0.0;
            return PHB_OK;
        }
        // real-space only needs q, positions and the list: use a dummy LJ table
        phb_params_s dummy;
        dummy.d.ntypes = 1; dummy.d.lj_shift = 0;
        double o[8];
        pme_energy_core(P, a, nl, &dummy, box, PAIR_EWALD, o, P->partials.p, P->result.p);
        double es, eq, esl;
        corrections(P, box, o + 3, &es, &eq, &esl);
        if (comps) *comps = phb_pme_components{o[1], o[2], es, eq + esl, esl};
        *E = o[1] + o[2] + es + eq + esl;
        return PHB_OK;
    });
}

extern "C" int phb_total_energy(phb_atoms a, phb_params p, phb_csr nl, phb_pme P, const phb_box* box, double* comps,
                                double* E) {
    return guard([&] {
        require(a && p && nl && box && E, "total_energy: null argument");
        require(nl->rows == a->n, "total_energy: neighbour list rows != natoms");
        double c[6] = {0, 0, 0, 0, 0, 0};
        if (p->d.coulombtype == PHB_COUL_PME) {
            require(P != nullptr, "total_energy: PME plan is null");  // MC/src/mc.cu:54
            if (a->n > 0) {
                double o[8];
                pme_energy_core(P, a, nl, p, box, nl->nnz > 0 ? (PAIR_LJ | PAIR_EWALD) : 0, o, P->partials.p,
                                P->result.p);
                double es, eq, esl;
                corrections(P, box, o + 3, &es, &eq, &esl);
                c[0] = o[0]; c[1] = o[1]; c[2] = o[2]; c[3] = es; c[4] = eq; c[5] = esl;
            }
        } else {
            if (a->n > 0 && nl->nnz > 0) {
                phb::DevBuf<double> part(2 * 2048), res(2);
                const int nb = pair_blocks(a->n);
                PmeKernels::launch_pair(a, p, nl, box, PAIR_LJ | PAIR_COUL_CUT, 0.0, kKe / p->d.epsilon_r, PHB_FP64, part.p, nb);
                reduce_partials(part.p, nb, 2, 2, res.p);
                double h[2];
                res.download(h, 2);
                sync();
                c[0] = h[0]; c[1] = h[1];
            }
        }
        if (comps) for (int k = 0; k < 6; ++k) comps[k] = c[k];
        *E = c[0] + (c[1] + c[2] + c[3] + c[4] + c[5]);
        return PHB_OK;
    });
}

// ------------------------------------------------------------ scratch-free total energy
// The partial sums and results of phb_total_energy_noalloc live in this array
// (one copy per GPU, held until the process ends) instead of a per-call or
// per-plan allocation.  Host code never passes its name to a kernel: it takes
// the device address from hipGetSymbolAddress, once per device.
namespace {
constexpr size_t kNoallocPartials = 4 * 2048;   // >= 4 * pair_blocks(N) for every N, and the z-pass blocks
constexpr size_t kNoallocDoubles = kNoallocPartials + 16;
}  // namespace

static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;   // per device: the address of that device's copy of `scratch`

double* scratch_of_current_device() {
    int dev = 0;
    PHB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        PHB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "phb_total_energy_noalloc: no device address for the scratch array", PHB_ERR_HIP);
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" int phb_total_energy_noalloc(phb_atoms a, phb_params p, phb_csr nl, phb_pme P, const phb_box* box,
                                        double* comps, double* E) {
    return guard([&] {
        require(a && p && nl && box && E, "total_energy: null argument");
        require(nl->rows == a->n, "total_energy: neighbour list rows != natoms");
        require_device();
        double* part = scratch_of_current_device();
        double* res = part + kNoallocPartials;
        double c[6] = {0, 0, 0, 0, 0, 0};
        if (p->d.coulombtype == PHB_COUL_PME) {
            require(P != nullptr, "total_energy: PME plan is null");  // MC/src/mc.cu:54
            require((size_t)P->z_blocks <= kNoallocPartials,
                    "phb_total_energy_noalloc: the plan's grid needs more z-pass partials than the scratch array "
                    "holds (8192); use phb_total_energy");
            if (a->n > 0) {
                double o[8];
                pme_energy_core(P, a, nl, p, box, nl->nnz > 0 ? (PAIR_LJ | PAIR_EWALD) : 0, o, part, res);
                double es, eq, esl;
                corrections(P, box, o + 3, &es, &eq, &esl);
                c[0] = o[0]; c[1] = o[1]; c[2] = o[2]; c[3] = es; c[4] = eq; c[5] = esl;
            }
        } else if (a->n > 0 && nl->nnz > 0) {
            const int nb = pair_blocks(a->n);
            PmeKernels::launch_pair(a, p, nl, box, PAIR_LJ | PAIR_COUL_CUT, 0.0, kKe / p->d.epsilon_r, PHB_FP64, part,
                                    nb);
            reduce_partials(part, nb, 2, 2, res);
            double h[2];
            PHB_HIP(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
            PHB_HIP(hipStreamSynchronize(stream()));
            c[0] = h[0]; c[1] = h[1];
        }
        if (comps) for (int k = 0; k < 6; ++k) comps[k] = c[k];
        *E = c[0] + (c[1] + c[2] + c[3] + c[4] + c[5]);
        return PHB_OK;
    });
}

// ------------------------------------------------------------ stage access
extern "C" int phb_pme_spread(phb_pme P, phb_atoms a, double* grid) {
    return guard([&] {
        require(P && a && grid, "pme_spread: null argument");
        PmeKernels::launch_spread(P, a);
        const size_t G = (size_t)P->nx * P->ny * P->nz;
        if (P->precision == PHB_FP64) P->grid64.download(grid, G);
        else {
            std::vector<float> g(G);
            P->grid32.download(g.data(), G);
            sync();
            std::copy(g.begin(), g.end(), grid);
        }
        sync();
        return PHB_OK;
    });
}

namespace {
void upload_grid(phb_pme_s* P, const double* grid) {
    const size_t G = (size_t)P->nx * P->ny * P->nz;
    if (P->precision == PHB_FP64) P->grid64.upload(grid, G);
    else {
        std::vector<float> g(grid, grid + G);
        P->grid32.upload(g.data(), G);
        sync();
    }
}
}  // namespace

extern "C" int phb_pme_fft_r2c(phb_pme P, const double* grid, double* kgrid) {
    return guard([&] {
        require(P && grid && kgrid, "pme_fft_r2c: null argument");
        upload_grid(P, grid);
        const size_t Gk = (size_t)(P->nx / 2 + 1) * P->ny * P->nz;
        // z pass with mode 4: forward transform stored, no influence function
        PmeKernels::launch_fft(P, 4, false, false);
        if (P->precision == PHB_FP64) P->kg64.download(reinterpret_cast<cplx<double>*>(kgrid), Gk);
        else {
            std::vector<cplx<float>> k(Gk);
            P->kg32.download(k.data(), Gk);
            sync();
            for (size_t i = 0; i < Gk; ++i) { kgrid[2 * i] = k[i].x; kgrid[2 * i + 1] = k[i].y; }
        }
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_pme_influence(phb_pme P, double* eta) {
    return guard([&] {
        require(P && eta, "pme_influence: null argument");
        std::copy(P->eta_host.begin(), P->eta_host.end(), eta);
        return PHB_OK;
    });
}

extern "C" int phb_pme_potential(phb_pme P, const double* grid, double* phi, double* E) {
    return guard([&] {
        require(P && grid, "pme_potential: null argument");
        upload_grid(P, grid);
        PmeKernels::launch_fft(P, 3, true, false);
        reduce_partials(P->partials.p, P->z_blocks, 1, 1, P->result.p);
        const size_t G = (size_t)P->nx * P->ny * P->nz;
        if (phi) {
            if (P->precision == PHB_FP64) P->phi64.download(phi, G);
            else {
                std::vector<float> f(G);
                P->phi32.download(f.data(), G);
                sync();
                std::copy(f.begin(), f.end(), phi);
            }
        }
        double e = 0.0;
        P->result.download(&e, 1);
        sync();
        if (E) *E = e;
        return PHB_OK;
    });
}
