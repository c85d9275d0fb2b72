// oracle.cpp -- CPU FP64 restatement of ph_brush's Monte Carlo energy path.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Built with
// -ffp-contract=off so that r^2 and min-image arithmetic round exactly as
// written, which the bit-exact neighbour-list comparison relies on.
//
// Each function cites the reference file:line it follows.  Pinning of each
// function (what checks it independently of this file) is listed in
// tests/test_oracle.py; every function here is pinned there.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
constexpr double kKe = 138.935458;  // MC/src/pme.cu:305, MC/src/energy_gpu.cu:78

// MC/src/neighbor.cu:15 -- dx -= L * nearbyint(dx / L)
inline double min_image(double dx, double L) { return dx - L * std::nearbyint(dx / L); }

// MC/src/neighbor.cu:17-19 -- binary search in the sorted 1-2 bond row of i
inline bool is_bonded12(int i, int j, const int* head, const int* list) {
    if (!head) return false;
    int lo = head[i], hi = head[i + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int v = list[mid];
        if (v == j) return true;
        if (v < j) lo = mid + 1; else hi = mid - 1;
    }
    return false;
}

inline double pair_r2(const or_system* s, int i, int j) {
    double dx = min_image(s->x[i] - s->x[j], s->Lx);
    double dy = min_image(s->y[i] - s->y[j], s->Ly);
    double dz = min_image(s->z[i] - s->z[j], s->Lz);
    return dx * dx + dy * dy + dz * dz;
}

// --------------------------------------------------------------------------
// FFT: own mixed-radix Cooley-Tukey (iterative radix-2 for powers of two,
// recursive smallest-prime split otherwise).  Unnormalised, sign -1 forward.
struct Plan1D {
    int n = 0;
    int sign = -1;
    std::vector<cd> w;       // w[k] = exp(sign*2*pi*i*k/n)
    std::vector<int> rev;    // bit reversal (power of two only)
    bool pow2 = false;
    Plan1D(int n_, int sign_) : n(n_), sign(sign_), w(n_) {
        for (int k = 0; k < n; ++k) {
            double a = 2.0 * M_PI * double(k) / double(n);
            w[k] = cd(std::cos(a), sign * std::sin(a));
        }
        pow2 = n > 0 && (n & (n - 1)) == 0;
        if (pow2) {
            rev.resize(n);
            int lg = 0; while ((1 << lg) < n) ++lg;
            for (int k = 0; k < n; ++k) {
                int r = 0;
                for (int b = 0; b < lg; ++b) if (k & (1 << b)) r |= 1 << (lg - 1 - b);
                rev[k] = r;
            }
        }
    }
};

int smallest_factor(int n) {
    for (int f = 2; f * f <= n; ++f) if (n % f == 0) return f;
    return n;
}

void fft_rec(const cd* in, int stride, cd* out, int n, const Plan1D& P, int wstride, cd* tmp) {
    if (n == 1) { out[0] = in[0]; return; }
    int r = smallest_factor(n), m = n / r;
    for (int q = 0; q < r; ++q) fft_rec(in + q * stride, stride * r, out + q * m, m, P, wstride * r, tmp);
    const int N0 = P.n;
    std::vector<cd> t(r);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q)
            t[q] = out[q * m + k] * P.w[(long long)q * k * wstride % N0];
        for (int s = 0; s < r; ++s) {
            cd acc = 0.0;
            for (int q = 0; q < r; ++q) acc += t[q] * P.w[(long long)q * s * m * wstride % N0];
            tmp[k + s * m] = acc;
        }
    }
    for (int k = 0; k < n; ++k) out[k] = tmp[k];
}

void fft1d(cd* a, const Plan1D& P, std::vector<cd>& scratch) {
    const int n = P.n;
    if (n <= 1) return;
    if (P.pow2) {
        for (int k = 0; k < n; ++k) if (k < P.rev[k]) std::swap(a[k], a[P.rev[k]]);
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, step = n / len;
            for (int b = 0; b < n; b += len)
                for (int k = 0; k < half; ++k) {
                    cd u = a[b + k], v = a[b + k + half] * P.w[k * step];
                    a[b + k] = u + v;
                    a[b + k + half] = u - v;
                }
        }
        return;
    }
    scratch.resize(2 * n);
    cd* in = scratch.data();
    cd* tmp = scratch.data() + n;
    std::vector<cd> out(n);
    for (int k = 0; k < n; ++k) in[k] = a[k];
    fft_rec(in, 1, out.data(), n, P, 1, tmp);
    for (int k = 0; k < n; ++k) a[k] = out[k];
}

// Transform every line of a strided 3-D complex array along one axis.
void fft_axis(cd* data, int nlines, int n, const Plan1D& P,
              const std::vector<long long>& base, long long stride) {
#pragma omp parallel
    {
        std::vector<cd> line(n), scratch;
#pragma omp for schedule(static)
        for (int l = 0; l < nlines; ++l) {
            cd* p = data + base[l];
            for (int k = 0; k < n; ++k) line[k] = p[k * stride];
            fft1d(line.data(), P, scratch);
            for (int k = 0; k < n; ++k) p[k * stride] = line[k];
        }
    }
}

// --------------------------------------------------------------------------
// MC/src/pme.cu:132-138 -- the reference's cubic weights, used verbatim at order 4.
void bspline_cubic(double t, double w[4]) {
    double t2 = t * t, t3 = t2 * t;
    w[0] = (1.0 - 3 * t + 3 * t2 - t3) / 6.0;
    w[1] = (4.0 - 6 * t2 + 3 * t3) / 6.0;
    w[2] = (1.0 + 3 * t + 3 * t2 - 3 * t3) / 6.0;
    w[3] = t3 / 6.0;
}

// EXTENSION (order != 4): cardinal B-spline M_p by the Cox-de Boor recursion
// M_{n+1}(u) = [u M_n(u) + (n+1-u) M_n(u-1)] / n, evaluated at u = t + j.
// Returns w[k] = M_p(t + p - 1 - k), the weight of grid point i0 - off + k with
// off = p - 1 - p/2; at p = 4 this is the reference stencil i0-1..i0+2.
void bspline_general(int p, double t, double* w) {
    double v[16] = {0};
    v[0] = t; v[1] = 1.0 - t;                       // M_2(t), M_2(t+1)
    for (int n = 2; n < p; ++n) {
        double nv[16] = {0};
        for (int j = 0; j <= n; ++j) {
            double u = t + j;
            double a = (j < n) ? v[j] : 0.0;
            double b = (j > 0) ? v[j - 1] : 0.0;
            nv[j] = (u * a + (n + 1 - u) * b) / n;
        }
        for (int j = 0; j <= n; ++j) v[j] = nv[j];
    }
    for (int k = 0; k < p; ++k) w[k] = v[p - 1 - k];
}

inline int stencil_offset(int p) { return p - 1 - p / 2; }

std::vector<double> lj_shift_table(const or_params* p, double lj_rc) {
    // MC/src/io.cpp:270-281 (shift evaluated at the LJ cutoff)
    const int T = p->ntypes;
    std::vector<double> tbl((size_t)T * T, 0.0);
    if (p->lj_shift) {
        for (int i = 0; i < T; ++i)
            for (int j = 0; j < T; ++j) {
                double eps = p->lj_eps[i * T + j], sig = p->lj_sig[i * T + j];
                if (eps > 0.0 && sig > 0.0) {
                    double sr = sig / lj_rc, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
                    tbl[i * T + j] = 4.0 * eps * (sr12 - sr6);
                }
            }
    }
    return tbl;
}

}  // namespace

extern "C" {

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

// MC/src/pme.cu:46-51
int or_good_fft_dim(int n) {
    const int NMAX = std::max(16, n * 2 + 32);
    auto is_good = [](int x) { for (int p : {2, 3, 5}) while (x % p == 0) x /= p; return x == 1; };
    for (int m = std::max(8, n); m <= NMAX; ++m) if (is_good(m)) return m;
    return n;
}

// MC/src/pme.cu:54-63
double or_alpha_from_rtol(double rc, double rtol) {
    double beta = 5.0; int i = 0;
    while (std::erfc(beta * rc) > rtol) { beta *= 2.0; if (++i > 60) break; }
    double low = 0.0, high = beta;
    for (int it = 0; it < 60; ++it) {
        beta = 0.5 * (low + high);
        if (std::erfc(beta * rc) > rtol) low = beta; else high = beta;
    }
    return beta;
}

void or_bspline(int order, double t, double* w) {
    if (order == 4) bspline_cubic(t, w);
    else bspline_general(order, t, w);
}

// MC/src/pme.cu:77-113.  EXTENSION: after the reference's clamp and Nyquist
// patch, any modulus below 1e-7 is replaced by the mean of its neighbours.
// Odd orders have an exact zero of the spline transform at Nyquist that the
// reference's patch misses (it only fires for values <= 0); at order 4 no
// modulus is below 1/9, so this never changes the reference result.
void or_moduli(int n, int pme_order, double* mod) {
    const int splineOrder = pme_order - 1;
    std::vector<double> data(std::max(splineOrder, 1), 0.0);
    data[0] = 1.0;
    for (int k = 2; k <= splineOrder; ++k) {
        const double invk = 1.0 / k;
        for (int m = k - 1; m > 0; --m) data[m] = invk * ((k - m) * data[m - 1] + (m + 1) * data[m]);
        data[0] = invk * data[0];
    }
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
    std::vector<double> orig(mod, mod + n);
    for (int i = 0; i < n; ++i)
        if (orig[i] < 1e-7) mod[i] = 0.5 * (orig[(i + n - 1) % n] + orig[(i + 1) % n]);
}

// MC/src/neighbor.cu:21-63 + 74-125: pairs i<j, min-image r2 <= rc2 (inclusive),
// 1-2 bonded pairs skipped, rows in ascending j.  Returns nnz; fills head/list
// when non-NULL (list only if cap >= nnz).
long long or_nlist_half(const or_system* s, double rc, int* head, int* list, long long cap) {
    const int N = s->natoms;
    const double rc2 = rc * rc;
    std::vector<int> cnt(N, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        int c = 0;
        for (int j = i + 1; j < N; ++j) {
            if (is_bonded12(i, j, s->bhead, s->blist)) continue;
            if (pair_r2(s, i, j) <= rc2) ++c;
        }
        cnt[i] = c;
    }
    std::vector<long long> off(N + 1, 0);
    for (int i = 0; i < N; ++i) off[i + 1] = off[i] + cnt[i];
    long long nnz = off[N];
    if (head) for (int i = 0; i <= N; ++i) head[i] = (int)off[i];
    if (list && cap >= nnz) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            long long w = off[i];
            for (int j = i + 1; j < N; ++j) {
                if (is_bonded12(i, j, s->bhead, s->blist)) continue;
                if (pair_r2(s, i, j) <= rc2) list[w++] = j;
            }
        }
    }
    return nnz;
}

// MC/src/energy_gpu.cu:14-46.  EXTENSION: pairs beyond lj_rc (when lj_rc < rc)
// are skipped and the shift is taken at lj_rc; lj_rc <= 0 is the reference.
double or_energy_lj(const or_system* s, const or_params* p, const int* head, const int* list) {
    const int N = s->natoms, T = p->ntypes;
    const double lj_rc = p->lj_rc > 0 ? p->lj_rc : p->rc;
    const bool use_cut = p->lj_rc > 0 && p->lj_rc < p->rc;
    const double lj_rc2 = lj_rc * lj_rc;
    std::vector<double> shift = lj_shift_table(p, lj_rc);
    double E = 0.0;
#pragma omp parallel for reduction(+ : E) schedule(static)
    for (int i = 0; i < N; ++i) {
        int ti = s->type[i];
        for (int k = head[i]; k < head[i + 1]; ++k) {
            int j = list[k];
            double r2 = pair_r2(s, i, j);
            if (use_cut && r2 > lj_rc2) continue;
            double r = std::sqrt(r2);
            int tj = s->type[j];
            double eps = p->lj_eps[ti * T + tj], sig = p->lj_sig[ti * T + tj];
            if (eps <= 0.0 || sig <= 0.0) continue;
            double sr = sig / r, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
            double e = 4.0 * eps * (sr12 - sr6);
            if (p->lj_shift) e -= shift[ti * T + tj];
            E += e;
        }
    }
    return E;
}

// MC/src/energy_gpu.cu:80-108
double or_energy_coul_cut(const or_system* s, const or_params* p, const int* head, const int* list) {
    const int N = s->natoms;
    const double kf = kKe / p->epsilon_r;
    double E = 0.0;
#pragma omp parallel for reduction(+ : E) schedule(static)
    for (int i = 0; i < N; ++i) {
        double qi = s->q[i];
        for (int k = head[i]; k < head[i + 1]; ++k) {
            int j = list[k];
            double r = std::sqrt(pair_r2(s, i, j));
            double qq = qi * s->q[j];
            if (qq == 0.0) continue;
            E += kf * qq * (1.0 / r);
        }
    }
    return E;
}

// MC/src/pme.cu:252-288 (no exclusion correction for bonded pairs, as there)
double or_energy_real_ewald(const or_system* s, const or_params* p, const int* head, const int* list) {
    const int N = s->natoms;
    const double kf = kKe / p->epsilon_r, alpha = p->ewald_alpha;
    double E = 0.0;
#pragma omp parallel for reduction(+ : E) schedule(static)
    for (int i = 0; i < N; ++i) {
        double qi = s->q[i];
        for (int k = head[i]; k < head[i + 1]; ++k) {
            int j = list[k];
            double r = std::sqrt(pair_r2(s, i, j));
            double qq = qi * s->q[j];
            if (qq == 0.0) continue;
            E += kf * qq * (std::erfc(alpha * r) / r);
        }
    }
    return E;
}

// MC/src/pme.cu:140-172, generalised to order p (EXTENSION; p = 4 is the reference).
void or_spread(const or_system* s, const or_params* p, double* grid) {
    const int nx = p->nx, ny = p->ny, nz = p->nz, P = p->pme_order, off = stencil_offset(P);
    std::fill(grid, grid + (size_t)nx * ny * nz, 0.0);
    double wx[16], wy[16], wz[16];
    for (int i = 0; i < s->natoms; ++i) {
        double gx = (s->x[i] / s->Lx) * nx; int ix = (int)std::floor(gx); double tx = gx - ix;
        double gy = (s->y[i] / s->Ly) * ny; int iy = (int)std::floor(gy); double ty = gy - iy;
        double gz = (s->z[i] / s->Lz) * nz; int iz = (int)std::floor(gz); double tz = gz - iz;
        or_bspline(P, tx, wx); or_bspline(P, ty, wy); or_bspline(P, tz, wz);
        const double qi = s->q[i];
        for (int a = 0; a < P; ++a) {
            int ixa = ((ix + a - off) % nx + nx) % nx;
            for (int b = 0; b < P; ++b) {
                int iyb = ((iy + b - off) % ny + ny) % ny;
                double wxy = qi * wx[a] * wy[b];
                for (int c = 0; c < P; ++c) {
                    int izc = ((iz + c - off) % nz + nz) % nz;
                    grid[((size_t)izc * ny + iyb) * nx + ixa] += wxy * wz[c];
                }
            }
        }
    }
}

// Forward 3-D real-to-half-complex FFT, unnormalised (the reference calls
// cufftExecD2Z with dims (nz, ny, nx), MC/src/pme.cu:334, :398-400).
void or_fft_r2c_3d(int nx, int ny, int nz, const double* in, double* outd) {
    const int nkx = nx / 2 + 1;
    cd* out = reinterpret_cast<cd*>(outd);
    Plan1D Px(nx, -1), Py(ny, -1), Pz(nz, -1);
#pragma omp parallel
    {
        std::vector<cd> row(nx), scratch;
#pragma omp for schedule(static)
        for (int zy = 0; zy < nz * ny; ++zy) {
            for (int x = 0; x < nx; ++x) row[x] = cd(in[(size_t)zy * nx + x], 0.0);
            fft1d(row.data(), Px, scratch);
            for (int kx = 0; kx < nkx; ++kx) out[(size_t)zy * nkx + kx] = row[kx];
        }
    }
    std::vector<long long> base;
    base.reserve((size_t)nz * nkx);
    for (int z = 0; z < nz; ++z) for (int kx = 0; kx < nkx; ++kx) base.push_back((long long)z * ny * nkx + kx);
    fft_axis(out, nz * nkx, ny, Py, base, nkx);
    base.clear();
    for (int y = 0; y < ny; ++y) for (int kx = 0; kx < nkx; ++kx) base.push_back((long long)y * nkx + kx);
    fft_axis(out, ny * nkx, nz, Pz, base, (long long)ny * nkx);
}

// Inverse of the above, unnormalised: out(r) = sum over the full spectrum of
// H(k) exp(+2 pi i k.r / n), with the Hermitian half implied (imaginary parts
// of the kx = 0 and kx = nx/2 planes ignored, as in a C2R transform).
void or_fft_c2r_3d(int nx, int ny, int nz, const double* ind, double* out) {
    const int nkx = nx / 2 + 1;
    std::vector<cd> h(reinterpret_cast<const cd*>(ind), reinterpret_cast<const cd*>(ind) + (size_t)nkx * ny * nz);
    Plan1D Px(nx, +1), Py(ny, +1), Pz(nz, +1);
    std::vector<long long> base;
    for (int y = 0; y < ny; ++y) for (int kx = 0; kx < nkx; ++kx) base.push_back((long long)y * nkx + kx);
    fft_axis(h.data(), ny * nkx, nz, Pz, base, (long long)ny * nkx);
    base.clear();
    for (int z = 0; z < nz; ++z) for (int kx = 0; kx < nkx; ++kx) base.push_back((long long)z * ny * nkx + kx);
    fft_axis(h.data(), nz * nkx, ny, Py, base, nkx);
#pragma omp parallel
    {
        std::vector<cd> row(nx), scratch;
#pragma omp for schedule(static)
        for (int zy = 0; zy < nz * ny; ++zy) {
            const cd* H = h.data() + (size_t)zy * nkx;
            for (int k = 0; k < nx; ++k) {
                if (k < nkx) row[k] = H[k];
                else row[k] = std::conj(H[nx - k]);
            }
            row[0] = cd(H[0].real(), 0.0);
            if (nx % 2 == 0) row[nx / 2] = cd(H[nx / 2].real(), 0.0);
            fft1d(row.data(), Px, scratch);
            for (int x = 0; x < nx; ++x) out[(size_t)zy * nx + x] = row[x].real();
        }
    }
}

// Influence function over the half spectrum, MC/src/pme.cu:196-223:
// eta = elF exp(-(pi/alpha)^2 m^2) / (pi V m^2 Bx By Bz), eta(0) = 0.
void or_influence(const or_params* p, double Lx, double Ly, double Lz, double* eta) {
    const int nx = p->nx, ny = p->ny, nz = p->nz, nkx = nx / 2 + 1;
    std::vector<double> mX(nx), mY(ny), mZ(nz);
    or_moduli(nx, p->pme_order, mX.data());
    or_moduli(ny, p->pme_order, mY.data());
    or_moduli(nz, p->pme_order, mZ.data());
    const double alpha = p->ewald_alpha;
    const double ewF = (M_PI / alpha) * (M_PI / alpha);
    const double elF = kKe / p->epsilon_r;
    const double V = Lx * Ly * Lz;
    const double Hxx = 1.0 / Lx, Hyy = 1.0 / Ly, Hzz = 1.0 / Lz;
    for (int kz = 0; kz < nz; ++kz)
        for (int ky = 0; ky < ny; ++ky)
            for (int kx = 0; kx < nkx; ++kx) {
                size_t o = ((size_t)kz * ny + ky) * nkx + kx;
                eta[o] = 0.0;
                if (kx == 0 && ky == 0 && kz == 0) continue;
                int mx = kx, my = (ky <= ny / 2) ? ky : ky - ny, mz = (kz <= nz / 2) ? kz : kz - nz;
                double kxv = mx * Hxx, kyv = my * Hyy, kzv = mz * Hzz;
                double m2k = kxv * kxv + kyv * kyv + kzv * kzv;
                if (m2k == 0.0) continue;
                double denom = m2k * M_PI * V * mX[kx] * mY[ky] * mZ[kz];
                if (!(denom > 0.0)) continue;
                eta[o] = elF * std::exp(-ewF * m2k) / denom;
            }
}

// MC/src/pme.cu:175-249 + :422: E = 0.5 * sum corner * 2 Re(eta Q conj Q),
// corner = 0.5 on the kx = 0 and kx = nx/2 planes.
double or_recip_energy(const or_params* p, double Lx, double Ly, double Lz, const double* kgrid) {
    const int nx = p->nx, ny = p->ny, nz = p->nz, nkx = nx / 2 + 1;
    std::vector<double> eta((size_t)nkx * ny * nz);
    or_influence(p, Lx, Ly, Lz, eta.data());
    const cd* K = reinterpret_cast<const cd*>(kgrid);
    double E = 0.0;
    const long long tot = (long long)nkx * ny * nz;
#pragma omp parallel for reduction(+ : E) schedule(static)
    for (long long o = 0; o < tot; ++o) {
        int kx = (int)(o % nkx);
        double corner = (kx == 0 || kx == nx / 2) ? 0.5 : 1.0;
        cd old = K[o];
        double nxr = old.real() * eta[o], nyi = old.imag() * eta[o];
        E += corner * (2.0 * (nxr * old.real() + nyi * old.imag()));
    }
    return 0.5 * E;
}

// MC/src/pme.cu:344-457 (real space, spread, FFT, reciprocal, self, qcorr, slab)
double or_pme_energy(const or_system* s, const or_params* p, const int* head, const int* list, double* comps) {
    const int N = s->natoms;
    double E_real = or_energy_real_ewald(s, p, head, list);
    std::vector<double> grid((size_t)p->nx * p->ny * p->nz);
    std::vector<double> kg((size_t)2 * (p->nx / 2 + 1) * p->ny * p->nz);
    or_spread(s, p, grid.data());
    or_fft_r2c_3d(p->nx, p->ny, p->nz, grid.data(), kg.data());
    double E_recip = or_recip_energy(p, s->Lx, s->Ly, s->Lz, kg.data());

    const double elF = kKe / p->epsilon_r, alpha = p->ewald_alpha;
    const double V = s->Lx * s->Ly * s->Lz;
    double Q = 0.0, Q2 = 0.0;
    for (int i = 0; i < N; ++i) { Q += s->q[i]; Q2 += s->q[i] * s->q[i]; }
    double E_self = -elF * alpha / std::sqrt(M_PI) * Q2;
    double E_qcorr = -(M_PI * elF) * (Q * Q) / (2.0 * V * alpha * alpha);
    double E_slab = 0.0;
    if (p->pme_3dc) {
        const double dc = (2.0 * M_PI / V) * elF;
        double Mz = 0.0;
        for (int i = 0; i < N; ++i) Mz += s->q[i] * s->z[i];
        E_slab = dc * Mz * Mz;
        if (std::abs(Q) > 1e-4) {
            double sqz2 = 0.0;
            for (int i = 0; i < N; ++i) sqz2 += s->q[i] * s->z[i] * s->z[i];
            E_slab -= dc * Q * (sqz2 + Q * (s->Lz * s->Lz) / 12.0);
        }
    }
    if (comps) { comps[0] = E_real; comps[1] = E_recip; comps[2] = E_self; comps[3] = E_qcorr; comps[4] = E_slab; }
    return E_real + E_recip + E_self + E_qcorr + E_slab;
}

// MC/src/mc.cu:43-60
double or_total_energy(const or_system* s, const or_params* p, const int* head, const int* list, double* comps) {
    double c[5] = {0, 0, 0, 0, 0};
    double E_lj = or_energy_lj(s, p, head, list);
    double E_c = 0.0;
    if (p->coulombtype == 0) { E_c = or_energy_coul_cut(s, p, head, list); c[0] = E_c; }
    else E_c = or_pme_energy(s, p, head, list, c);
    if (comps) { comps[0] = E_lj; for (int k = 0; k < 5; ++k) comps[k + 1] = c[k]; }
    return E_lj + E_c;
}

// Fixed-mu GC protonation + ion-swap chain, the inner move rule of
// phase_gc_mu_control (MC/src/mc.cu:379-503) with mu_eff held constant and no
// target stop (EXTENSION for titration at fixed pH).  Pools are built as in
// build_plan (MC/src/mc.cu:75-100) and phase_gc_mu_control (:332-341), each trial
// recomputes the total energy (MC/src/mc.cu:425-426) and pools are updated by
// erase(find)+push_back (:454-466).  Returns 0, or 1 when no pair is valid.
int or_gc_run(const or_system* s0, const or_params* p, const or_mc_params* m,
              const int* head, const int* list, long long ntrials, unsigned seed,
              or_mc_record* rec, double* q_out, int* type_out, double* E_out, long long* n_done) {
    const int N = s0->natoms;
    std::vector<double> q(s0->q, s0->q + N);
    std::vector<int> type(s0->type, s0->type + N);
    or_system s = *s0;
    s.q = q.data(); s.type = type.data();

    std::vector<int> nh_n, nh_p, w, cl;
    for (int i = 0; i < N; ++i) {
        int t = type[i] + 1;
        if (t == m->nh_n_type) nh_n.push_back(i);
        else if (t == m->nh_p_type) nh_p.push_back(i);
        else if (t == m->w_type) w.push_back(i);
    }
    double z_max_site = 0.0;
    for (int i : nh_n) z_max_site = std::max(z_max_site, s.z[i]);
    for (int i : nh_p) z_max_site = std::max(z_max_site, s.z[i]);
    const double z_cut = z_max_site + m->w_z_extra;
    { std::vector<int> f; for (int j : w) if (s.z[j] < z_cut) f.push_back(j); w.swap(f); }
    for (int j = 0; j < N; ++j) if (type[j] + 1 == m->cl_type && s.z[j] < z_cut) cl.push_back(j);

    double E = or_total_energy(&s, p, head, list, nullptr);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0), u01(0.0, 1.0);
    auto choose_one = [&](const std::vector<int>& pool) {
        std::uniform_int_distribution<int> d(0, (int)pool.size() - 1);
        return pool[d(rng)];
    };
    int status = 0;
    long long t = 0;
    for (; t < ntrials; ++t) {
        long long a = (long long)nh_n.size(), b = (long long)nh_p.size();
        long long c = (long long)w.size(), d = (long long)cl.size();
        long long w_fwd = a * c, w_bwd = b * d, S_before = a * c + b * d;
        bool can_fwd = w_fwd > 0, can_bwd = w_bwd > 0, forward;
        if (can_fwd && can_bwd) {
            double sum = (double)(w_fwd + w_bwd);
            forward = (u01(rng) < (double)w_fwd / sum);
        } else if (can_fwd) forward = true;
        else if (can_bwd) forward = false;
        else { status = 1; break; }
        int i, j;
        if (forward) { i = choose_one(nh_n); j = choose_one(w); }
        else { i = choose_one(nh_p); j = choose_one(cl); }
        const int oti = type[i], otj = type[j];
        const double oqi = q[i], oqj = q[j];
        if (forward) { type[i] = m->nh_p_type - 1; q[i] = 1.0; type[j] = m->cl_type - 1; q[j] = -1.0; }
        else { type[i] = m->nh_n_type - 1; q[i] = 0.0; type[j] = m->w_type - 1; q[j] = 0.0; }
        double E_trial = or_total_energy(&s, p, head, list, nullptr);
        double dE = E_trial - E;
        int sg = forward ? +1 : -1;
        long long S_after = forward ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
        double logH = 0.0;
        if (S_before > 0 && S_after > 0) logH = std::log((double)S_before) - std::log((double)S_after);
        double expo = (double)(-m->beta * (dE - sg * m->mu_eff)) + logH;
        expo = std::clamp(expo, -40.0, 40.0);
        double u = std::numeric_limits<double>::quiet_NaN();
        bool accept = expo >= 0.0;
        if (!accept) { u = uni(rng); accept = std::exp(expo) > u; }
        if (rec) {
            if (rec->fwd) rec->fwd[t] = forward ? 1 : 0;
            if (rec->i) rec->i[t] = i;
            if (rec->j) rec->j[t] = j;
            if (rec->acc) rec->acc[t] = accept ? 1 : 0;
            if (rec->dE) rec->dE[t] = dE;
            if (rec->expo) rec->expo[t] = expo;
            if (rec->u) rec->u[t] = u;
        }
        if (accept) {
            E = E_trial;
            if (forward) {
                nh_n.erase(std::find(nh_n.begin(), nh_n.end(), i)); nh_p.push_back(i);
                w.erase(std::find(w.begin(), w.end(), j)); cl.push_back(j);
            } else {
                nh_p.erase(std::find(nh_p.begin(), nh_p.end(), i)); nh_n.push_back(i);
                cl.erase(std::find(cl.begin(), cl.end(), j)); w.push_back(j);
            }
        } else {
            type[i] = oti; q[i] = oqi; type[j] = otj; q[j] = oqj;
        }
    }
    if (q_out) std::copy(q.begin(), q.end(), q_out);
    if (type_out) std::copy(type.begin(), type.end(), type_out);
    if (E_out) *E_out = E;
    if (n_done) *n_done = t;
    return status;
}

void or_mt_draws(unsigned seed, long long n, unsigned* out) {
    std::mt19937 rng(seed);
    for (long long k = 0; k < n; ++k) out[k] = (unsigned)rng();
}

// MC/src/mc.cu:62-65: uniform_int_distribution<int>(0, hi) per draw.
void or_uniform_int_draws(unsigned seed, long long n, const int* range_hi, int* out) {
    std::mt19937 rng(seed);
    for (long long k = 0; k < n; ++k) {
        std::uniform_int_distribution<int> d(0, range_hi[k]);
        out[k] = d(rng);
    }
}

// MC/src/mc.cu:368-369: uniform_real_distribution<double>(0, 1).
void or_canonical_draws(unsigned seed, long long n, double* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> d(0.0, 1.0);
    for (long long k = 0; k < n; ++k) out[k] = d(rng);
}

// The proposal stream's mix: hi[k] >= 0 draws uniform_int_distribution<int>(0, hi[k])
// (MC/src/mc.cu:62-65), hi[k] < 0 draws uniform_real_distribution<double>(0, 1)
// (MC/src/mc.cu:368-369), from one std::mt19937 resumed from its stream form
// (624 words + index, as main's engine between phases, MC/src/main.cu:180-206);
// state_after: the engine in the same form after the draws.
void or_mixed_draws(const unsigned* state, long long n, const int* hi, unsigned long long* out, unsigned* state_after) {
    std::mt19937 rng;
    {
        std::stringstream ss;
        for (int k = 0; k < 625; ++k) ss << state[k] << ' ';
        ss >> rng;
    }
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    for (long long k = 0; k < n; ++k) {
        if (hi[k] >= 0) {
            std::uniform_int_distribution<int> d(0, hi[k]);
            out[k] = (unsigned long long)(unsigned)d(rng);
        } else {
            const double v = u01(rng);
            std::memcpy(&out[k], &v, sizeof v);
        }
    }
    std::stringstream ss;
    ss << rng;
    for (int k = 0; k < 625; ++k) ss >> state_after[k];
}

// ---------------------------------------------------------------------------
// Delayed-acceptance surrogate (EXTENSION; no reference counterpart).
void or_desc_thresholds(double r_desc, double* th) {
    for (int k = 0; k < 12; ++k) {
        const double rk = (double)k * r_desc / 12.0;
        th[k] = rk * rk;
    }
}

void or_da_descriptors(const or_system* s, const or_desc_params* dp, int n, const int* i, const int* j, int* d,
                       double* rij) {
    double th[12];
    or_desc_thresholds(dp->r_desc, th);
    const double rc2 = dp->r_desc * dp->r_desc;
    const int N = s->natoms;
#pragma omp parallel for schedule(static)
    for (int m = 0; m < n; ++m) {
        int* out = d + (size_t)m * 48;
        for (int c = 0; c < 48; ++c) out[c] = 0;
        const int a = i[m], b = j[m];
        for (int which = 0; which < 2; ++which) {
            const int ctr = which ? b : a, sg = which ? -1 : 1;
            for (int k = 0; k < N; ++k) {
                if (k == a || k == b) continue;
                const int sp = dp->species_of_type[s->type[k]];
                if (sp < 0) continue;
                const double r2 = pair_r2(s, ctr, k);
                if (!(r2 <= rc2)) continue;
                int sh = 0;
                while (sh < 11 && r2 >= th[sh + 1]) ++sh;
                out[sh * 4 + sp] += sg;
            }
        }
        if (rij) rij[m] = std::sqrt(pair_r2(s, a, b));
    }
}

// Phase 1 as main() runs it (MC/src/main.cu:180-197): build_plan's checks
// (MC/src/mc.cu:67-122), then apply_direct_protonation (MC/src/mc.cu:124-157)
// when the target exceeds direct_pct_threshold, else phase_gc_mu_control
// (MC/src/mc.cu:320-546): mu from the target fraction, blocks of gc_block
// trials, an epoch ends between blocks once Nacc_target moves were accepted or
// Amax trials made, then mu -= kT eta logit_diff w_info; the run stops at the
// trial that reaches the target.  Pools handed to phase 2 (nh_n, nh_p, w, cl
// back to back into handover[N], sizes into handover_sizes): the direct path's
// lists, the identities in ascending atom order when the target is reached,
// else empty (the reference fills them only then).  Returns 0, or -1 with *err
// set when build_plan throws (1 no NH sites, 2 target below current, 3 not
// enough NH_N, 4 not enough W); status 1 when no valid pair was left.
int or_phase1_run(const or_system* s0, const or_params* p, const or_mc_params* m, const int* head, const int* list,
                  unsigned seed, double target_pct, int direct_pct_threshold, int gc_block, int epoch_attempt_max,
                  double nacc_frac, double mu_eta, long long max_trials, or_phase1_result* res, double* q_out,
                  int* type_out, unsigned* rng_state_out, int* handover, int* handover_sizes, int* err) {
    const int N = s0->natoms;
    std::vector<double> q(s0->q, s0->q + N);
    std::vector<int> type(s0->type, s0->type + N);
    or_system s = *s0;
    s.q = q.data(); s.type = type.data();
    *err = 0;
    *res = or_phase1_result{};
    std::vector<int> idx_nh_n, idx_nh_p, idx_w;
    for (int i = 0; i < N; ++i) {
        const int t = type[i] + 1;
        if (t == m->nh_n_type) idx_nh_n.push_back(i);
        else if (t == m->nh_p_type) idx_nh_p.push_back(i);
        else if (t == m->w_type) idx_w.push_back(i);
    }
    const int nh_total = (int)(idx_nh_n.size() + idx_nh_p.size());
    if (nh_total == 0) { *err = 1; return -1; }
    const int target_count = (int)std::floor(target_pct / 100.0 * nh_total);
    const int to_flip = target_count - (int)idx_nh_p.size();
    if (to_flip < 0) { *err = 2; return -1; }
    if ((int)idx_nh_n.size() < to_flip) { *err = 3; return -1; }
    double z_max_site = 0.0;
    for (int i : idx_nh_n) z_max_site = std::max(z_max_site, s.z[i]);
    for (int i : idx_nh_p) z_max_site = std::max(z_max_site, s.z[i]);
    const double z_cut = z_max_site + m->w_z_extra;
    { std::vector<int> f; for (int j : idx_w) if (s.z[j] < z_cut) f.push_back(j); idx_w.swap(f); }
    if ((int)idx_w.size() < to_flip) { *err = 4; return -1; }
    std::mt19937 rng(seed);
    std::vector<int> h_nn, h_np, h_w, h_cl;
    res->target_count = target_count;
    res->Nh = nh_total;
    if (target_pct > (double)direct_pct_threshold) {
        res->direct = 1;
        h_nn = idx_nh_n; h_w = idx_w;
        for (int k = 0; k < to_flip; ++k) {
            const int i = h_nn[k], j = h_w[k];
            type[i] = m->nh_p_type - 1; q[i] = 1.0;
            type[j] = m->cl_type - 1; q[j] = -1.0;
            h_np.push_back(i); h_cl.push_back(j);
        }
        h_nn.erase(h_nn.begin(), h_nn.begin() + std::max(to_flip, 0));
        h_w.erase(h_w.begin(), h_w.begin() + std::max(to_flip, 0));
        res->attempted = to_flip; res->accepted = to_flip;
        res->Nprot = (int)idx_nh_p.size() + to_flip;
        res->reached_target = 1;
        res->E = or_total_energy(&s, p, head, list, nullptr);
        res->mu_eff = m->mu_eff;
    } else {
        std::vector<int> nh_n = idx_nh_n, nh_p = idx_nh_p, w = idx_w, cl;
        for (int j = 0; j < N; ++j) if (type[j] + 1 == m->cl_type && s.z[j] < z_cut) cl.push_back(j);
        res->Nprot = (int)nh_p.size();
        double E = or_total_energy(&s, p, head, list, nullptr);
        const double beta = m->beta;
        const double fstar = std::clamp(target_pct / 100.0, 1e-6, 1.0 - 1e-6);
        double mu = (1.0 / beta) * std::log(fstar / (1.0 - fstar));
        const double mu_bound = 10.0 / beta;
        mu = std::clamp(mu, -mu_bound, mu_bound);
        bool done_all = false;
        if (res->Nprot == target_count) { res->reached_target = 1; done_all = true; }
        const int blk = gc_block > 0 ? gc_block : 2000;
        const int Amax = epoch_attempt_max > 0 ? epoch_attempt_max : 100000;
        const int R0 = std::max(0, target_count - res->Nprot);
        const double frac = nacc_frac > 0.0 ? nacc_frac : 0.10;
        const int Nacc_target = std::max(1, (int)std::ceil(frac * (double)R0));
        std::uniform_real_distribution<double> uni(0.0, 1.0), u01(0.0, 1.0);
        auto choose_one = [&](const std::vector<int>& pool) {
            std::uniform_int_distribution<int> d(0, (int)pool.size() - 1);
            return pool[d(rng)];
        };
        for (int epoch = 0; !done_all; ++epoch) {
            int acc_since = 0, att_since = 0;
            while (!done_all && acc_since < Nacc_target && att_since < Amax) {
                for (int t = 0; t < blk; ++t) {
                    if (max_trials > 0 && res->attempted >= max_trials) { done_all = true; break; }
                    const long long a = (long long)nh_n.size(), b = (long long)nh_p.size();
                    const long long c = (long long)w.size(), d = (long long)cl.size();
                    const long long w_fwd = a * c, w_bwd = b * d, S_before = a * c + b * d;
                    bool forward;
                    if (w_fwd > 0 && w_bwd > 0) forward = u01(rng) < (double)w_fwd / (double)(w_fwd + w_bwd);
                    else if (w_fwd > 0) forward = true;
                    else if (w_bwd > 0) forward = false;
                    else { res->status = 1; done_all = true; break; }
                    int i, j;
                    if (forward) { i = choose_one(nh_n); j = choose_one(w); }
                    else { i = choose_one(nh_p); j = choose_one(cl); }
                    const int oti = type[i], otj = type[j];
                    const double oqi = q[i], oqj = q[j];
                    if (forward) { type[i] = m->nh_p_type - 1; q[i] = 1.0; type[j] = m->cl_type - 1; q[j] = -1.0; }
                    else { type[i] = m->nh_n_type - 1; q[i] = 0.0; type[j] = m->w_type - 1; q[j] = 0.0; }
                    const double E_trial = or_total_energy(&s, p, head, list, nullptr);
                    const double dE = E_trial - E;
                    const int sg = forward ? 1 : -1;
                    const long long S_after = forward ? (a - 1) * (c - 1) + (b + 1) * (d + 1)
                                                      : (a + 1) * (c + 1) + (b - 1) * (d - 1);
                    double logH = 0.0;
                    if (S_before > 0 && S_after > 0) logH = std::log((double)S_before) - std::log((double)S_after);
                    double expo = (double)(-beta * (dE - sg * mu)) + logH;
                    expo = std::clamp(expo, -40.0, 40.0);
                    const bool acc = (expo >= 0.0) || (std::exp(expo) > uni(rng));
                    ++res->attempted; ++att_since;
                    if (acc) {
                        ++res->accepted; ++acc_since;
                        if (forward) ++res->accepted_forward; else ++res->accepted_backward;
                        E = E_trial;
                        if (forward) {
                            nh_n.erase(std::find(nh_n.begin(), nh_n.end(), i)); nh_p.push_back(i);
                            w.erase(std::find(w.begin(), w.end(), j)); cl.push_back(j);
                            ++res->Nprot;
                        } else {
                            nh_p.erase(std::find(nh_p.begin(), nh_p.end(), i)); nh_n.push_back(i);
                            cl.erase(std::find(cl.begin(), cl.end(), j)); w.push_back(j);
                            --res->Nprot;
                        }
                        if (res->Nprot == target_count) {
                            res->reached_target = 1;
                            for (int k = 0; k < N; ++k) {
                                const int tk = type[k] + 1;
                                if (tk == m->nh_p_type) h_np.push_back(k);
                                else if (tk == m->nh_n_type) h_nn.push_back(k);
                                else if (tk == m->w_type) { if (s.z[k] < z_cut) h_w.push_back(k); }
                                else if (tk == m->cl_type) { if (s.z[k] < z_cut) h_cl.push_back(k); }
                            }
                            done_all = true;
                            break;
                        }
                    } else {
                        type[i] = oti; q[i] = oqi; type[j] = otj; q[j] = oqj;
                    }
                }
            }
            if (done_all) break;
            const double f = res->Nh > 0 ? (double)res->Nprot / (double)res->Nh : 0.0;
            const double f0 = std::clamp(fstar, 1e-6, 1.0 - 1e-6), f1 = std::clamp(f, 1e-6, 1.0 - 1e-6);
            const double logit_diff = std::log(f1 / (1.0 - f1)) - std::log(f0 / (1.0 - f0));
            const double eta = mu_eta > 0.0 ? mu_eta : 0.3;
            const double w_info = Nacc_target > 0
                ? std::clamp((double)acc_since / (double)std::max(1, Nacc_target), 0.0, 1.0) : 1.0;
            const double mu_step = (1.0 / beta) * eta * logit_diff * w_info;
            mu -= mu_step;
            res->blocks_done = epoch + 1;
        }
        res->E = E;
        res->mu_eff = mu;
    }
    handover_sizes[0] = (int)h_nn.size(); handover_sizes[1] = (int)h_np.size();
    handover_sizes[2] = (int)h_w.size(); handover_sizes[3] = (int)h_cl.size();
    int* o = handover;
    for (int v : h_nn) *o++ = v;
    for (int v : h_np) *o++ = v;
    for (int v : h_w) *o++ = v;
    for (int v : h_cl) *o++ = v;
    if (q_out) std::copy(q.begin(), q.end(), q_out);
    if (type_out) std::copy(type.begin(), type.end(), type_out);
    if (rng_state_out) {
        std::ostringstream os;
        os << rng;
        std::istringstream is(os.str());
        for (int k = 0; k < 625; ++k) is >> rng_state_out[k];
    }
    return 0;
}

// phase2_swap_jitter (MC/src/mc.cu:159-315) after the phase-1 loop reached its
// target: pools rebuilt in ascending atom order from the identities
// (MC/src/mc.cu:480-494), W and Cl below z_cut (MC/src/mc.cu:177-184); the
// engine resumes from a saved state (the 624 words and the index, libstdc++'s
// stream form).  Every trial is recomputed in full, as the reference does.
int or_phase2_run(const or_system* s0, const or_params* p, const or_mc_params* m, const int* head, const int* list,
                  const unsigned* rng_state, const int* pools, const int* pool_sizes, long long max_attempts,
                  int patience, double best_eps, int* rec_atoms, signed char* rec_acc, double* rec_dE,
                  double* q_out, int* type_out, double* E_out, double* best_out, long long* attempted,
                  long long* accepted, int* mode, int* stop, unsigned* rng_after) {
    const int N = s0->natoms;
    std::vector<double> q(s0->q, s0->q + N);
    std::vector<int> type(s0->type, s0->type + N);
    or_system s = *s0;
    s.q = q.data(); s.type = type.data();
    double z_max_site = 0.0;
    for (int i = 0; i < N; ++i)
        if (type[i] + 1 == m->nh_n_type || type[i] + 1 == m->nh_p_type) z_max_site = std::max(z_max_site, s.z[i]);
    const double z_cut = z_max_site + m->w_z_extra;
    std::vector<int> nh_p, nh_n, w_cand, cl_pool;
    if (pools) {   // the phase-1 hand-over: nh_n, nh_p, w, cl back to back; W and Cl z-filtered
        const int* q0 = pools;
        nh_n.assign(q0, q0 + pool_sizes[0]); q0 += pool_sizes[0];
        nh_p.assign(q0, q0 + pool_sizes[1]); q0 += pool_sizes[1];
        for (int k = 0; k < pool_sizes[2]; ++k) if (s.z[q0[k]] < z_cut) w_cand.push_back(q0[k]);
        q0 += pool_sizes[2];
        for (int k = 0; k < pool_sizes[3]; ++k) if (s.z[q0[k]] < z_cut) cl_pool.push_back(q0[k]);
    } else {
        for (int i = 0; i < N; ++i) {
            const int t = type[i] + 1;
            if (t == m->nh_p_type) nh_p.push_back(i);
            else if (t == m->nh_n_type) nh_n.push_back(i);
            else if (t == m->w_type) { if (s.z[i] < z_cut) w_cand.push_back(i); }
            else if (t == m->cl_type) { if (s.z[i] < z_cut) cl_pool.push_back(i); }
        }
    }
    std::mt19937 rng;
    {
        std::ostringstream os;
        for (int k = 0; k < 624; ++k) os << rng_state[k] << ' ';
        os << rng_state[624];
        std::istringstream is(os.str());
        is >> rng;
    }
    double E = or_total_energy(&s, p, head, list, nullptr);
    double best_E = E;
    auto choose_one = [&](const std::vector<int>& pool) {
        std::uniform_int_distribution<int> d(0, (int)pool.size() - 1);
        return pool[d(rng)];
    };
    auto can_four = [&]() { return !nh_p.empty() && !nh_n.empty() && !w_cand.empty() && !cl_pool.empty(); };
    auto can_wc = [&]() { return !w_cand.empty() && !cl_pool.empty(); };
    bool prefer_wc = nh_n.empty();
    if (!prefer_wc && !can_four() && can_wc()) prefer_wc = true;
    int md = 0, st = 0;
    long long att = 0, acc = 0;
    if ((!prefer_wc && !can_four() && !can_wc()) || (prefer_wc && !can_wc())) st = 2;
    else {
        md = prefer_wc ? 2 : 4;
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        int since_best = 0;
        for (long long step = 0; step < max_attempts; ++step) {
            const bool mode_four = (!prefer_wc) && can_four();
            ++att;
            int at[4] = {-1, -1, -1, -1};
            int nt[4] = {0, 0, 0, 0};
            double nq[4] = {0, 0, 0, 0};
            int K;
            if (mode_four) {
                at[0] = choose_one(nh_p); at[1] = choose_one(nh_n); at[2] = choose_one(cl_pool); at[3] = choose_one(w_cand);
                nt[0] = m->nh_n_type - 1; nq[0] = 0.0;
                nt[1] = m->nh_p_type - 1; nq[1] = 1.0;
                nt[2] = m->w_type - 1; nq[2] = 0.0;
                nt[3] = m->cl_type - 1; nq[3] = -1.0;
                K = 4;
            } else {
                at[0] = choose_one(cl_pool); at[1] = choose_one(w_cand);
                nt[0] = m->w_type - 1; nq[0] = 0.0;
                nt[1] = m->cl_type - 1; nq[1] = -1.0;
                K = 2;
            }
            int ot[4];
            double oq[4];
            for (int k = 0; k < K; ++k) { ot[k] = type[at[k]]; oq[k] = q[at[k]]; type[at[k]] = nt[k]; q[at[k]] = nq[k]; }
            const double E_trial = or_total_energy(&s, p, head, list, nullptr);
            const double dE = E_trial - E;
            const double expo = std::clamp(-m->beta * dE, -40.0, 40.0);
            const bool ok = (dE <= 0.0) || (std::exp(expo) > uni(rng));
            if (rec_atoms) for (int k = 0; k < 4; ++k) rec_atoms[4 * step + k] = at[k];
            if (rec_acc) rec_acc[step] = ok ? 1 : 0;
            if (rec_dE) rec_dE[step] = dE;
            if (ok) {
                ++acc;
                E = E_trial;
                if (E < best_E - best_eps) { best_E = E; since_best = 0; }
                else ++since_best;
                if (mode_four) {
                    nh_p.erase(std::find(nh_p.begin(), nh_p.end(), at[0])); nh_n.push_back(at[0]);
                    nh_n.erase(std::find(nh_n.begin(), nh_n.end(), at[1])); nh_p.push_back(at[1]);
                    cl_pool.erase(std::find(cl_pool.begin(), cl_pool.end(), at[2])); w_cand.push_back(at[2]);
                    w_cand.erase(std::find(w_cand.begin(), w_cand.end(), at[3])); cl_pool.push_back(at[3]);
                } else {
                    cl_pool.erase(std::find(cl_pool.begin(), cl_pool.end(), at[0])); w_cand.push_back(at[0]);
                    w_cand.erase(std::find(w_cand.begin(), w_cand.end(), at[1])); cl_pool.push_back(at[1]);
                }
            } else {
                for (int k = 0; k < K; ++k) { type[at[k]] = ot[k]; q[at[k]] = oq[k]; }
                ++since_best;
            }
            if (since_best >= patience) { st = 1; break; }
        }
    }
    if (rng_after) {   // the engine after phase 2, in its stream form
        std::stringstream ss;
        ss << rng;
        for (int k = 0; k < 625; ++k) ss >> rng_after[k];
    }
    if (q_out) std::copy(q.begin(), q.end(), q_out);
    if (type_out) std::copy(type.begin(), type.end(), type_out);
    if (E_out) *E_out = E;
    if (best_out) *best_out = best_E;
    if (attempted) *attempted = att;
    if (accepted) *accepted = acc;
    if (mode) *mode = md;
    if (stop) *stop = st;
    return 0;
}

int or_da_run(const or_system* s0, const or_params* p, const or_mc_params* m, const int* head, const int* list,
              long long ntrials, unsigned seed, const double* sur_trial, const double* sur_rev_trial,
              const double* sur_pair, const double* sur_pair_sym, or_da_record* rec, double* q_out, int* type_out,
              double* E_out, long long* n_done) {
    const int N = s0->natoms;
    std::vector<double> q(s0->q, s0->q + N);
    std::vector<int> type(s0->type, s0->type + N);
    or_system s = *s0;
    s.q = q.data(); s.type = type.data();
    // pools exactly as or_gc_run (build_plan, MC/src/mc.cu:67-100)
    std::vector<int> nh_n, nh_p, w, cl;
    for (int i = 0; i < N; ++i) {
        int t = type[i] + 1;
        if (t == m->nh_n_type) nh_n.push_back(i);
        else if (t == m->nh_p_type) nh_p.push_back(i);
        else if (t == m->w_type) w.push_back(i);
    }
    double z_max_site = 0.0;
    for (int i : nh_n) z_max_site = std::max(z_max_site, s.z[i]);
    for (int i : nh_p) z_max_site = std::max(z_max_site, s.z[i]);
    const double z_cut = z_max_site + m->w_z_extra;
    { std::vector<int> f; for (int j : w) if (s.z[j] < z_cut) f.push_back(j); w.swap(f); }
    for (int j = 0; j < N; ++j) if (type[j] + 1 == m->cl_type && s.z[j] < z_cut) cl.push_back(j);

    double E = or_total_energy(&s, p, head, list, nullptr);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    // the delayed-acceptance chain's own pool order (no reference counterpart): an
    // accepted atom leaves its pool by taking the pool's last entry into its place
    auto swap_remove = [](std::vector<int>& v, int a) {
        auto it = std::find(v.begin(), v.end(), a);
        *it = v.back();
        v.pop_back();
    };
    auto choose_one = [&](const std::vector<int>& pool) {
        std::uniform_int_distribution<int> dd(0, (int)pool.size() - 1);
        return pool[dd(rng)];
    };
    int status = 0;
    long long t = 0;
    for (; t < ntrials; ++t) {
        long long a = (long long)nh_n.size(), b = (long long)nh_p.size();
        long long c = (long long)w.size(), d = (long long)cl.size();
        long long w_fwd = a * c, w_bwd = b * d, S_before = a * c + b * d;
        bool forward;
        if (w_fwd > 0 && w_bwd > 0) forward = uni(rng) < (double)w_fwd / (double)(w_fwd + w_bwd);
        else if (w_fwd > 0) forward = true;
        else if (w_bwd > 0) forward = false;
        else { status = 1; break; }
        int i, j;
        if (forward) { i = choose_one(nh_n); j = choose_one(w); }
        else { i = choose_one(nh_p); j = choose_one(cl); }
        // every trial draws both stages' uniforms, used or not, so the proposal
        // stream does not depend on the decisions
        const double u1 = uni(rng), u2 = uni(rng);
        const int sg = forward ? 1 : -1;
        const double sym = sur_pair_sym ? sur_pair_sym[(size_t)i * N + j] : 0.0;
        const double dEs = sur_trial ? sur_trial[t] : (double)sg * sur_pair[(size_t)i * N + j] + sym;
        const double dEr = sur_trial ? (sur_rev_trial ? sur_rev_trial[t] : -sur_trial[t])
                                     : -(double)sg * sur_pair[(size_t)i * N + j] + sym;
        long long S_after = forward ? (a - 1) * (c - 1) + (b + 1) * (d + 1) : (a + 1) * (c + 1) + (b - 1) * (d - 1);
        double logH = 0.0;
        if (S_before > 0 && S_after > 0) logH = std::log((double)S_before) - std::log((double)S_after);
        double e1 = (double)(-m->beta * (dEs - sg * m->mu_eff)) + logH;
        e1 = std::clamp(e1, -40.0, 40.0);
        const bool acc1 = e1 >= 0.0 || std::exp(e1) > u1;
        int stage = 0;
        double dE = std::numeric_limits<double>::quiet_NaN();
        if (acc1) {
            const int oti = type[i], otj = type[j];
            const double oqi = q[i], oqj = q[j];
            if (forward) { type[i] = m->nh_p_type - 1; q[i] = 1.0; type[j] = m->cl_type - 1; q[j] = -1.0; }
            else { type[i] = m->nh_n_type - 1; q[i] = 0.0; type[j] = m->w_type - 1; q[j] = 0.0; }
            const double E_trial = or_total_energy(&s, p, head, list, nullptr);
            dE = E_trial - E;
            double r1 = (double)(-m->beta * (dEr + sg * m->mu_eff)) - logH;
            r1 = std::clamp(r1, -40.0, 40.0);
            const double e2 = (((double)(-m->beta * (dE - sg * m->mu_eff)) + logH) + std::min(0.0, r1)) -
                              std::min(0.0, e1);
            const bool acc2 = e2 >= 0.0 || std::exp(e2) > u2;
            if (acc2) {
                stage = 1;
                E = E_trial;
                if (forward) {
                    swap_remove(nh_n, i); nh_p.push_back(i);
                    swap_remove(w, j); cl.push_back(j);
                } else {
                    swap_remove(nh_p, i); nh_n.push_back(i);
                    swap_remove(cl, j); w.push_back(j);
                }
            } else {
                stage = 2;
                type[i] = oti; q[i] = oqi; type[j] = otj; q[j] = oqj;
            }
        }
        if (rec) {
            if (rec->fwd) rec->fwd[t] = forward ? 1 : 0;
            if (rec->i) rec->i[t] = i;
            if (rec->j) rec->j[t] = j;
            if (rec->stage) rec->stage[t] = (signed char)stage;
            if (rec->dE_star) rec->dE_star[t] = dEs;
            if (rec->dE_star_rev) rec->dE_star_rev[t] = stage ? dEr : std::numeric_limits<double>::quiet_NaN();
            if (rec->dE) rec->dE[t] = dE;
        }
    }
    if (q_out) std::copy(q.begin(), q.end(), q_out);
    if (type_out) std::copy(type.begin(), type.end(), type_out);
    if (E_out) *E_out = E;
    if (n_done) *n_done = t;
    return status;
}

}  // extern "C"

