// pme.hpp -- SPME plan (the handle behind phb_pme) and the launchers the MC
// module reuses.
#pragma once

#include <vector>

#include "device.hpp"
#include "handles.hpp"

namespace phb {

struct GridGeom {
    int nx, ny, nz, order;
    double Lx, Ly, Lz;
};

__host__ __device__ __forceinline__ int wrap(int i, int n) { return ((i % n) + n) % n; }

// Base grid index and B-spline weights of one atom along each axis
// (MC/src/pme.cu:151-156 generalised to order p): g = (x/L)*n, i0 = floor(g),
// t = g - i0, computed in FP64; weights in T.
template <typename T>
__host__ __device__ __forceinline__ void stencil_of(double x, double y, double z, const GridGeom& g, int i0[3],
                                                    T w[3][8]) {
    const double gx = (x / g.Lx) * g.nx, gy = (y / g.Ly) * g.ny, gz = (z / g.Lz) * g.nz;
    i0[0] = (int)floor(gx); i0[1] = (int)floor(gy); i0[2] = (int)floor(gz);
    bspline_weights<T>(g.order, (T)(gx - i0[0]), w[0]);
    bspline_weights<T>(g.order, (T)(gy - i0[1]), w[1]);
    bspline_weights<T>(g.order, (T)(gz - i0[2]), w[2]);
}

std::vector<double> spme_moduli(int n, int pme_order);
int pair_blocks(int N);

}  // namespace phb

struct phb_pme_s {
    int precision = PHB_FP64;
    int order = 4;
    bool use3dc = false;
    double alpha = 0, ewF = 0, elF = 0, vol = 0, Lx = 0, Ly = 0, Lz = 0;
    int nx = 0, ny = 0, nz = 0;
    phb::FftPlan px, py, pz;
    phb::GridGeom geom{};
    int plane_threads = 256, z_tx = 8, z_blocks = 0, z_threads = 256;
    // split plane transform (planes beyond one workgroup's LDS): x pass of x_rpb
    // row pairs per workgroup, y pass of y_tx kx columns per workgroup
    bool split = false;
    int x_rpb = 1, x_blocks = 0, x_threads = 256, y_tx = 1, y_blocks = 0, y_threads = 256;
    std::vector<double> eta_host;
    phb::DevBuf<double> grid64, phi64, eta64;
    phb::DevBuf<phb::cplx<double>> kg64, twx64, twy64, twz64;
    phb::DevBuf<float> grid32, phi32, eta32;
    phb::DevBuf<phb::cplx<float>> kg32, twx32, twy32, twz32;
    // influence function of the convolutions (potential grids, mesh kernel): eta,
    // except that for odd nx the kx = nx/2 column carries half weight, as the
    // reference's energy sum counts it (MC/src/pme.cu:234), so that 0.5 Q.phi
    // equals that energy; null for even nx (eta itself)
    phb::DevBuf<double> etac64;
    phb::DevBuf<float> etac32;
    phb::DevBuf<double> partials, result;
    // spreading tiles: edges and counts along each axis, tiles whose atoms reach
    // one point, count and offset per tile, fixed-point grid, bounds of the call
    int sp_t[3] = {1, 1, 1}, sp_tn[3] = {0, 0, 0}, sp_nt = 0;
    double sp_reach = 1.0;
    phb::DevBuf<int> sp_count, sp_off;
    phb::DevBuf<unsigned long long> sp_fix, sp_bound;
    const double* conv_eta64() const { return etac64.p ? etac64.p : eta64.p; }
    const float* conv_eta32() const { return etac32.p ? etac32.p : eta32.p; }
    size_t plane_lds(size_t csz) const {
        const int nyp = (ny + 1) / 2, nkx = nx / 2 + 1;
        return ((size_t)nx + ny + (size_t)nyp * nx + (size_t)nkx * (ny + 1)) * csz;
    }
    size_t z_lds(size_t csz) const { return ((size_t)nz + (size_t)z_tx * (nz + 1)) * csz; }
    size_t x_lds(size_t csz) const { return ((size_t)nx + (size_t)x_rpb * nx) * csz; }
    size_t y_lds(size_t csz) const { return ((size_t)ny + (size_t)y_tx * (ny + 1)) * csz; }
    // entries of the per-workgroup refresh gate an inverse plane pass needs
    int plane_gate_blocks() const { return split ? x_blocks : nz; }
    ~phb_pme_s();
};

namespace phb {

// The chain's plan for coulombtype=coul_cut: alpha 0, no reciprocal part, no
// slab term, an 8^3 order-4 grid (owned by the caller; delete when done).
phb_pme_s* pme_create_cut(const phb_params_desc& d, const phb_box& box);

}  // namespace phb

namespace phb {

// Gate of the potential refresh's inverse plane pass: plane z runs only while
// accepted charge changes are pending and applied[z] differs from their list's
// sequence number, and records the number when done; workgroups share no
// state, so no fence or atomic is needed (null fields: no gate).
struct RefreshGate {
    const int* pending_n = nullptr;
    const unsigned* seq = nullptr;
    unsigned* applied = nullptr;
};

// Inverse plane pass of z-plane zp (one workgroup): C2R along y then x from the
// half spectrum kg, stored into or accumulated onto out; nplanes workgroups share
// the gate's completion count.
template <typename T>
__device__ __forceinline__ void fft_xy_c2r_plane(const cplx<T>* __restrict__ kg, T* __restrict__ out, int nx, int ny,
                                                 const FftPlan& px, const FftPlan& py, const cplx<T>* __restrict__ twx,
                                                 const cplx<T>* __restrict__ twy, int accumulate, const RefreshGate& gate,
                                                 int zp, int nplanes) {
    if (gate.pending_n && (gate.applied[zp] == *gate.seq || *gate.pending_n == 0)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* stwx = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* stwy = stwx + nx;
    const int nyp = (ny + 1) / 2, nkx = nx / 2 + 1, hs = ny + 1;
    cplx<T>* A = stwy + ny;
    cplx<T>* H = A + nyp * nx;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int k = tid; k < nx; k += bd) stwx[k] = twx[k];
    for (int k = tid; k < ny; k += bd) stwy[k] = twy[k];
    const cplx<T>* in = kg + (size_t)zp * ny * nkx;
    // all of a thread's loads are issued before the first LDS store (latency overlap)
    for (int base = tid; base < ny * nkx; base += 8 * bd) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * bd;
            if (idx < ny * nkx) v[u] = in[idx];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * bd;
            if (idx < ny * nkx) {
                const int ky = idx / nkx, kx = idx - ky * nkx;
                H[kx * hs + ky] = v[u];
            }
        }
    }
    __syncthreads();
    lds_fft(H, nkx, hs, py, stwy, +1);
    for (int idx = tid; idx < nyp * nx; idx += bd) {
        const int yp = idx / nx, k = idx - yp * nx, y0 = 2 * yp;
        const bool hb = y0 + 1 < ny;
        cplx<T> a, b;
        if (k < nkx) {
            a = H[k * hs + y0];
            b = hb ? H[k * hs + y0 + 1] : cplx<T>{T(0), T(0)};
            if (k == 0 || 2 * k == nx) { a.y = T(0); b.y = T(0); }
        } else {
            const int kk = nx - k;
            a = H[kk * hs + y0]; a.y = -a.y;
            b = hb ? H[kk * hs + y0 + 1] : cplx<T>{T(0), T(0)};
            b.y = -b.y;
        }
        A[idx] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    lds_fft(A, nyp, nx, px, stwx, +1);
    T* plane = out + (size_t)zp * nx * ny;
    for (int base = tid; base < nyp * nx; base += 4 * bd) {
        T o0[4], o1[4];
        if (accumulate) {   // every load of the chunk first
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int idx = base + u * bd;
                o0[u] = o1[u] = T(0);
                if (idx < nyp * nx) {
                    const int yp = idx / nx, xx = idx - yp * nx, y0 = 2 * yp;
                    o0[u] = plane[y0 * nx + xx];
                    if (y0 + 1 < ny) o1[u] = plane[(y0 + 1) * nx + xx];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) o0[u] = o1[u] = T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = base + u * bd;
            if (idx < nyp * nx) {
                const int yp = idx / nx, xx = idx - yp * nx, y0 = 2 * yp;
                const cplx<T> v = A[idx];
                plane[y0 * nx + xx] = o0[u] + v.x;
                if (y0 + 1 < ny) plane[(y0 + 1) * nx + xx] = o1[u] + v.y;
            }
        }
    }
    if (gate.applied) {   // this plane now holds the pending list
        __syncthreads();
        if (tid == 0) gate.applied[zp] = *gate.seq;
    }
}

struct PmeKernels {
    static void launch_spread(phb_pme_s* P, const phb_atoms_s* a);
    // forward r2c + z pass (zmode bits: 1 energy, 2 convolve + inverse z),
    // optionally the inverse plane pass into phi (store or accumulate)
    // partials: where the z pass writes its energy partials (z_blocks doubles; null: the plan's own)
    static void launch_fft(phb_pme_s* P, int zmode, bool inverse_to_phi, bool accumulate, double* partials = nullptr);
    // explicit buffers; a non-null gate makes every FFT workgroup return at once while *gate == 0
    static void launch_fft_buf(phb_pme_s* P, const double* grid, double* phi, int zmode, bool inverse_to_phi,
                               bool accumulate, const int* gate = nullptr, double* partials = nullptr);
    static void launch_fft_buf(phb_pme_s* P, const float* grid, float* phi, int zmode, bool inverse_to_phi,
                               bool accumulate, const int* gate = nullptr, double* partials = nullptr);
    static void launch_c2r_buf(phb_pme_s* P, double* phi, bool accumulate, RefreshGate gate = {});
    static void launch_c2r_buf(phb_pme_s* P, float* phi, bool accumulate, RefreshGate gate = {});
    static void launch_pair(const phb_atoms_s* a, const phb_params_s* p, const phb_csr_s* nl, const phb_box* box,
                            int mode, double alpha, double kf, int precision, double* partials, int nblocks);
    static void launch_moments(const phb_atoms_s* a, double* partials, int nblocks);
};

void corrections(const phb_pme_s* P, const phb_box* box, const double mom[4], double* e_self, double* e_qcorr,
                 double* e_slab);

}  // namespace phb
