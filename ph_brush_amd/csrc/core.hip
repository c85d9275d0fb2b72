// core.hip -- error state, library stream, FFT plans, reductions and the
// atoms / CSR / params handles of the C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "device.hpp"
#include "handles.hpp"

namespace phb {

namespace {
thread_local std::string g_last_error;
hipStream_t g_stream = nullptr;
std::once_flag g_stream_once;
}  // namespace

void set_last_error(const std::string& m) { g_last_error = m; }

void require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error(PHB_ERR_UNSUPPORTED, "no GPU visible: ph_brush_amd has no CPU fallback");
}

hipStream_t stream() {
    require_device();
    std::call_once(g_stream_once, [] {
        if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) g_stream = nullptr;
    });
    if (!g_stream) throw Error(PHB_ERR_HIP, "could not create the library stream");
    return g_stream;
}

void sync() { PHB_HIP(hipStreamSynchronize(stream())); }

bool fft_len_ok(int n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5}) while (n % p == 0) n /= p;
    return n == 1;
}

FftPlan make_fft_plan(int n) {
    if (!fft_len_ok(n)) throw Error(PHB_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not 2/3/5-smooth");
    FftPlan p;
    p.n = n;
    int m = n;
    while (m % 4 == 0) { p.rad[p.nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p.rad[p.nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.nst++] = 5; m /= 5; }
    return p;
}

std::vector<cplx<double>> twiddles(int n) {
    std::vector<cplx<double>> w(n);
    for (int k = 0; k < n; ++k) {
        double a = 2.0 * M_PI * double(k) / double(n);
        w[k] = {std::cos(a), -std::sin(a)};
    }
    return w;
}

__global__ void k_reduce_partials(const double* __restrict__ partials, int n, int stride, int nsum,
                                  double* __restrict__ out) {
    __shared__ double sred[16];
    for (int c = 0; c < nsum; ++c) {
        double v = 0.0;
        // fixed assignment of partials to threads, fixed tree: deterministic
        for (int k = threadIdx.x; k < n; k += blockDim.x) v += partials[(size_t)k * stride + c];
        v = block_sum(v, sred);
        if (threadIdx.x == 0) out[c] = v;
        __syncthreads();
    }
}

void reduce_partials(const double* d_partials, int n, int stride, int nsum, double* d_out) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, stream(), d_partials, n, stride, nsum, d_out);
    PHB_HIP(hipGetLastError());
}

// Exclusive scan in three passes (deterministic: fixed segments, fixed order):
// block sums of SCAN_TILE-element tiles, one block scanning the tile sums, then
// each tile's scan from its offset.  Each thread owns SCAN_PER consecutive
// elements of its tile.
constexpr int SCAN_TPB = 256, SCAN_PER = 16, SCAN_TILE = SCAN_TPB * SCAN_PER;

__global__ __launch_bounds__(SCAN_TPB) void k_scan_tile_sums(const int* __restrict__ in, int n,
                                                             long long* __restrict__ tsum) {
    __shared__ long long s[SCAN_TPB / 64];
    const int b0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    long long v = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) v += b0 + k < n ? in[b0 + k] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot = 0;
        for (int w = 0; w < SCAN_TPB / 64; ++w) tot += s[w];
        tsum[blockIdx.x] = tot;
    }
}

// One block: exclusive scan of the tile sums in place (each thread a segment of
// consecutive tiles, the segment sums scanned across the block); out[n] = total.
__global__ __launch_bounds__(1024) void k_scan_tile_offsets(long long* __restrict__ tsum, int ntiles, int* __restrict__ out,
                                                            int n) {
    __shared__ long long part[1024];
    const int T = blockDim.x, t = threadIdx.x;
    const int seg = (ntiles + T - 1) / T;
    const int b = min(ntiles, t * seg), e = min(ntiles, b + seg);
    long long s = 0;
    for (int k = b; k < e; ++k) s += tsum[k];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < T; o <<= 1) {
        const long long add = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    if (t == T - 1) out[n] = (int)part[t];
    long long run = part[t] - s;
    for (int k = b; k < e; ++k) { const long long v = tsum[k]; tsum[k] = run; run += v; }
}

__global__ __launch_bounds__(SCAN_TPB) void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, int n,
                                                         const long long* __restrict__ toff) {
    __shared__ long long s[SCAN_TPB];
    const int b0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    long long mine = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = b0 + k < n ? in[b0 + k] : 0;
        mine += v[k];
    }
    s[threadIdx.x] = mine;
    __syncthreads();
    // Hillis-Steele inclusive scan of the thread sums (256 entries)
    for (int o = 1; o < SCAN_TPB; o <<= 1) {
        const long long add = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    long long run = toff[blockIdx.x] + s[threadIdx.x] - mine;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        if (b0 + k < n) out[b0 + k] = (int)run;
        run += v[k];
    }
}

// Rows sorted ascending: one wave per row.  The row is staged in the wave's LDS
// slice and each lane ranks its entries against it (every lane reads the same
// LDS word at a time: a broadcast); entries are distinct, so ranks are unique.
// Rows longer than the slice are ranked against global memory.
constexpr int SORT_TPB = 256, SORT_CAP = 2048;

__global__ __launch_bounds__(SORT_TPB) void k_sort_rows_lds(int N, const int* __restrict__ head,
                                                            const int* __restrict__ tmp, int* __restrict__ out) {
    __shared__ int buf[SORT_TPB / 64][SORT_CAP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * (SORT_TPB / 64) + w;
    if (i >= N) return;   // wave-uniform
    const int s = head[i], e = head[i + 1], len = e - s;
    if (len <= SORT_CAP) {
        int* row = buf[w];
        for (int a = lane; a < len; a += 64) row[a] = tmp[s + a];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0);
        for (int a = lane; a < len; a += 64) {
            const int v = row[a];
            int rank = 0;
            for (int b = 0; b < len; ++b) rank += row[b] < v ? 1 : 0;
            out[s + rank] = v;
        }
    } else {
        for (int a = s + lane; a < e; a += 64) {
            const int v = tmp[a];
            int rank = 0;
            for (int b = s; b < e; ++b) rank += tmp[b] < v ? 1 : 0;
            out[s + rank] = v;
        }
    }
}

void sort_rows(int N, const int* d_head, const int* d_tmp, int* d_out) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_sort_rows_lds, dim3(ceil_div(N, SORT_TPB / 64)), dim3(SORT_TPB), 0, stream(), N, d_head,
                       d_tmp, d_out);
    PHB_HIP(hipGetLastError());
}

void exclusive_scan(const int* d_in, int* d_out, int n) {
    const int ntiles = std::max(1, (n + SCAN_TILE - 1) / SCAN_TILE);
    DevBuf<long long> tsum(ntiles);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(ntiles), dim3(SCAN_TPB), 0, stream(), d_in, n, tsum.p);
    hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, stream(), tsum.p, ntiles, d_out, n);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_TPB), 0, stream(), d_in, d_out, n, tsum.p);
    PHB_HIP(hipGetLastError());
    sync();   // tsum is freed on return
}

}  // namespace phb

using namespace phb;

// ----------------------------------------------------------------------- ABI
extern "C" const char* phb_last_error(void) { return g_last_error.c_str(); }
extern "C" int phb_abi_version(void) { return PHB_ABI_VERSION; }

extern "C" int phb_device_count(int* n) {
    return guard([&] {
        require(n != nullptr, "phb_device_count: null");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *n = c;
        return PHB_OK;
    });
}

extern "C" int phb_atoms_create(int natoms, const double* x, const double* y, const double* z,
                                const double* q, const int* type0, phb_atoms* out) {
    return guard([&] {
        require(out && natoms >= 0, "phb_atoms_create: bad arguments");
        require(natoms == 0 || (x && y && z && q && type0), "phb_atoms_create: null array");
        int tmax = 0;
        for (int i = 0; i < natoms; ++i) {
            require(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]) && std::isfinite(q[i]),
                    "phb_atoms_create: non-finite position or charge");
            require(type0[i] >= 0, "phb_atoms_create: negative type id");
            tmax = std::max(tmax, type0[i]);
        }
        require_device();
        auto a = std::make_unique<phb_atoms_s>();
        a->n = natoms;
        a->max_type = tmax;
        a->x.alloc(natoms); a->y.alloc(natoms); a->z.alloc(natoms);
        a->q.alloc(natoms); a->type.alloc(natoms);
        a->sp_slot.alloc(natoms); a->sp_list.alloc(natoms);
        a->x.upload(x, natoms); a->y.upload(y, natoms); a->z.upload(z, natoms);
        a->q.upload(q, natoms); a->type.upload(type0, natoms);
        sync();
        *out = a.release();
        return PHB_OK;
    });
}

extern "C" int phb_atoms_set_positions(phb_atoms a, const double* x, const double* y, const double* z) {
    return guard([&] {
        require(a, "phb_atoms_set_positions: null handle");
        require(a->n == 0 || (x && y && z), "phb_atoms_set_positions: null array");
        for (int i = 0; i < a->n; ++i)
            require(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]),
                    "phb_atoms_set_positions: non-finite position");
        a->x.upload(x, a->n); a->y.upload(y, a->n); a->z.upload(z, a->n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_atoms_destroy(phb_atoms a) {
    return guard([&] { delete a; return PHB_OK; });
}

extern "C" int phb_atoms_natoms(phb_atoms a, int* n) {
    return guard([&] { require(a && n, "phb_atoms_natoms: null"); *n = a->n; return PHB_OK; });
}

namespace {
__global__ void k_set_identity(int n, const int* __restrict__ idx, const int* __restrict__ t,
                               const double* __restrict__ qv, int natoms, int* __restrict__ type,
                               double* __restrict__ q) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        int i = idx[k];
        if (i >= 0 && i < natoms) { type[i] = t[k]; q[i] = qv[k]; }
    }
}
}  // namespace

extern "C" int phb_atoms_set_identity(phb_atoms a, int n, const int* idx, const int* type0, const double* q) {
    return guard([&] {
        require(a && n >= 0 && (n == 0 || (idx && type0 && q)), "phb_atoms_set_identity: bad arguments");
        for (int k = 0; k < n; ++k) {
            require(idx[k] >= 0 && idx[k] < a->n, "set_atom_identity: atom index out of range");
            require(type0[k] >= 0, "set_atom_identity: type out of range");
            require(std::isfinite(q[k]), "set_atom_identity: non-finite charge");
        }
        for (int k = 0; k < n; ++k) a->max_type = std::max(a->max_type, type0[k]);
        if (n == 0) return PHB_OK;
        DevBuf<int> di(n), dt(n);
        DevBuf<double> dq(n);
        di.upload(idx, n); dt.upload(type0, n); dq.upload(q, n);
        hipLaunchKernelGGL(k_set_identity, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), n, di.p, dt.p, dq.p,
                           a->n, a->type.p, a->q.p);
        PHB_HIP(hipGetLastError());
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_atoms_get_identity(phb_atoms a, double* q, int* type0) {
    return guard([&] {
        require(a != nullptr, "phb_atoms_get_identity: null");
        if (q) a->q.download(q, a->n);
        if (type0) a->type.download(type0, a->n);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_csr_create(int rows, const int* head, const int* list, phb_csr* out) {
    return guard([&] {
        require(out && rows >= 0 && (rows == 0 || head), "phb_csr_create: bad arguments");
        long long nnz = rows > 0 ? head[rows] : 0;
        require(nnz >= 0 && (nnz == 0 || list), "phb_csr_create: bad nnz/list");
        for (int i = 0; i < rows; ++i) require(head[i] <= head[i + 1], "phb_csr_create: heads not monotone");
        require_device();
        auto c = std::make_unique<phb_csr_s>();
        c->rows = rows; c->nnz = nnz;
        c->head.alloc(rows + 1);
        std::vector<int> h0;
        if (rows == 0) { h0.assign(1, 0); head = h0.data(); }
        c->head.upload(head, rows + 1);
        c->list.alloc(nnz);
        c->list.upload(list, nnz);
        sync();
        *out = c.release();
        return PHB_OK;
    });
}

extern "C" int phb_csr_destroy(phb_csr c) {
    return guard([&] { delete c; return PHB_OK; });
}

extern "C" int phb_csr_shape(phb_csr c, int* rows, long long* nnz) {
    return guard([&] {
        require(c != nullptr, "phb_csr_shape: null");
        if (rows) *rows = c->rows;
        if (nnz) *nnz = c->nnz;
        return PHB_OK;
    });
}

extern "C" int phb_csr_download(phb_csr c, int* head, int* list) {
    return guard([&] {
        require(c != nullptr, "phb_csr_download: null");
        if (head) c->head.download(head, c->rows + 1);
        if (list) c->list.download(list, c->nnz);
        sync();
        return PHB_OK;
    });
}

extern "C" int phb_params_create(const phb_params_desc* d, phb_params* out) {
    return guard([&] {
        require(d && out, "phb_params_create: null");
        require(d->ntypes > 0 && d->lj_eps && d->lj_sig, "phb_params_create: LJ tables required");
        // MC/src/io.cpp:258
        require(d->rc > 0.0, "cutoff_nm must be > 0 in params");
        require(d->lj_rc >= 0.0 && d->lj_rc <= d->rc, "lj_rc must lie in [0, rc]");
        require(d->epsilon_r > 0.0, "epsilon_r must be > 0");
        require(d->coulombtype == PHB_COUL_CUT || d->coulombtype == PHB_COUL_PME, "unsupported coulombtype");
        require_device();
        auto p = std::make_unique<phb_params_s>();
        p->d = *d;
        const int T = d->ntypes;
        p->eps.assign(d->lj_eps, d->lj_eps + (size_t)T * T);
        p->sig.assign(d->lj_sig, d->lj_sig + (size_t)T * T);
        p->lj_rc = d->lj_rc > 0 ? d->lj_rc : d->rc;
        p->lj_cut_active = d->lj_rc > 0 && d->lj_rc < d->rc;
        // MC/src/io.cpp:270-281, evaluated at the LJ cutoff
        p->shift.assign((size_t)T * T, 0.0);
        if (d->lj_shift)
            for (int i = 0; i < T * T; ++i) {
                double eps = p->eps[i], sig = p->sig[i];
                if (eps > 0.0 && sig > 0.0) {
                    double sr = sig / p->lj_rc, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
                    p->shift[i] = 4.0 * eps * (sr12 - sr6);
                }
            }
        p->d.lj_eps = nullptr; p->d.lj_sig = nullptr;
        p->d_eps.alloc(T * T); p->d_sig.alloc(T * T); p->d_shift.alloc(T * T);
        p->d_eps.upload(p->eps.data(), T * T);
        p->d_sig.upload(p->sig.data(), T * T);
        p->d_shift.upload(p->shift.data(), T * T);
        sync();
        *out = p.release();
        return PHB_OK;
    });
}

extern "C" int phb_params_destroy(phb_params p) {
    return guard([&] { delete p; return PHB_OK; });
}
