// device.hpp -- HIP helpers shared by the gfx950 kernels: error checks, device
// buffers, the library stream, complex arithmetic, the LDS Stockham FFT,
// block reductions and B-spline weights.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

#define PHB_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::phb::Error(e_ == hipErrorOutOfMemory ? PHB_ERR_ALLOC : PHB_ERR_HIP,    \
                               std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__));             \
    } while (0)

namespace phb {

constexpr double kKe = 138.935458;  // kJ mol^-1 nm e^-2 (MC/src/pme.cu:305)

// Library stream: one per process, default priority, created on first use.
hipStream_t stream();
void sync();
// Throws PHB_ERR_UNSUPPORTED when no GPU is visible (the product has no CPU path).
void require_device();

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        PHB_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    void upload(const T* h, size_t count, size_t off = 0) {
        require(off + count <= n, "DevBuf::upload out of range", PHB_ERR_INTERNAL);
        if (count) PHB_HIP(hipMemcpyAsync(p + off, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
    }
    void download(T* h, size_t count, size_t off = 0) const {
        require(off + count <= n, "DevBuf::download out of range", PHB_ERR_INTERNAL);
        if (count) PHB_HIP(hipMemcpyAsync(h, p + off, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
    }
};

// ---------------------------------------------------------------- complex
template <typename T>
struct cplx { T x, y; };

template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by s*i (s = +-1)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul_si(cplx<T> a, int s) {
    return s > 0 ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// ------------------------------------------------------------ FFT plan
// Radix sequence of a 2/3/5-smooth length: 4s first, then 2, 3, 5.
struct FftPlan {
    int n = 0;
    int nst = 0;
    int rad[24] = {0};
};
FftPlan make_fft_plan(int n);  // throws PHB_ERR_UNSUPPORTED if n is not 2/3/5-smooth
bool fft_len_ok(int n);
// Twiddle table exp(-2 pi i k / n), k < n, in double.
std::vector<cplx<double>> twiddles(int n);

// Small DFTs, sign s = -1 forward, +1 inverse.
template <typename T, int R>
struct SmallDft;
template <typename T>
struct SmallDft<T, 2> {
    __device__ __forceinline__ static void run(cplx<T>* v, int) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    }
};
template <typename T>
struct SmallDft<T, 4> {
    __device__ __forceinline__ static void run(cplx<T>* v, int s) {
        cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cplx<T> a2 = cadd(v[1], v[3]), a3 = cmul_si(csub(v[1], v[3]), s);
        v[0] = cadd(a0, a2); v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3); v[3] = csub(a1, a3);
    }
};
template <typename T>
struct SmallDft<T, 3> {
    __device__ __forceinline__ static void run(cplx<T>* v, int s) {
        const T c = T(-0.5), d = T(0.86602540378443864676) * T(s);
        cplx<T> t1 = cadd(v[1], v[2]);
        cplx<T> t2 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        cplx<T> df = csub(v[1], v[2]);
        cplx<T> t3 = {-d * df.y, d * df.x};
        v[0] = cadd(v[0], t1); v[1] = cadd(t2, t3); v[2] = csub(t2, t3);
    }
};
template <typename T>
struct SmallDft<T, 5> {
    __device__ __forceinline__ static void run(cplx<T>* v, int s) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212) * T(s), s2 = T(0.58778525229247312917) * T(s);
        cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cplx<T> v0 = v[0];
        cplx<T> r1 = {v0.x + c1 * a1.x + c2 * a2.x, v0.y + c1 * a1.y + c2 * a2.y};
        cplx<T> r2 = {v0.x + c2 * a1.x + c1 * a2.x, v0.y + c2 * a1.y + c1 * a2.y};
        cplx<T> i1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};   // times i below
        cplx<T> i2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
        v[0] = {v0.x + a1.x + a2.x, v0.y + a1.y + a2.y};
        v[1] = {r1.x - i1.y, r1.y + i1.x};
        v[4] = {r1.x + i1.y, r1.y - i1.x};
        v[2] = {r2.x - i2.y, r2.y + i2.x};
        v[3] = {r2.x + i2.y, r2.y - i2.x};
    }
};

// One in-place Stockham stage over L lines of length n held in LDS
// (element k of line l at buf[l*ls + k]).  Every thread first loads all of
// its butterflies into registers, the block synchronises, then every thread
// writes its outputs: no ping-pong buffer is needed.
// Requires L*n <= MAXE*blockDim.x (checked by the host launcher).
template <typename T, int R, int MAXE>
__device__ __forceinline__ void stockham_stage(cplx<T>* buf, int L, int ls, int n, int Ns,
                                               const cplx<T>* tw, int sign) {
    constexpr int MAXB = MAXE / R;
    const int nb = n / R, total = L * nb, tstep = n / (Ns * R);
    cplx<T> v[MAXB][R];
#pragma unroll
    for (int it = 0; it < MAXB; ++it) {
        const int b = threadIdx.x + it * blockDim.x;
        if (b < total) {
            const int l = b / nb, j = b - l * nb, k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cplx<T> a = buf[l * ls + j + r * nb];
                cplx<T> w = tw[r * k * tstep];
                if (sign > 0) w.y = -w.y;
                v[it][r] = cmul(a, w);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXB; ++it) {
        const int b = threadIdx.x + it * blockDim.x;
        if (b < total) {
            const int l = b / nb, j = b - l * nb, k = j % Ns;
            SmallDft<T, R>::run(v[it], sign);
            const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[l * ls + idxD + r * Ns] = v[it][r];
        }
    }
    __syncthreads();
}

constexpr int kFftMaxE = 16;  // complex elements per thread per stage

template <typename T>
__device__ void lds_fft(cplx<T>* buf, int L, int ls, const FftPlan& pl, const cplx<T>* tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        switch (R) {
            case 4: stockham_stage<T, 4, kFftMaxE>(buf, L, ls, pl.n, Ns, tw, sign); break;
            case 2: stockham_stage<T, 2, kFftMaxE>(buf, L, ls, pl.n, Ns, tw, sign); break;
            case 3: stockham_stage<T, 3, 15>(buf, L, ls, pl.n, Ns, tw, sign); break;
            default: stockham_stage<T, 5, 15>(buf, L, ls, pl.n, Ns, tw, sign); break;
        }
        Ns *= R;
    }
}

// ------------------------------------------------------------ reductions
// Block sum of a double; result valid in thread 0.  blockDim.x <= 1024.
__device__ __forceinline__ double block_sum(double v, double* sred) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sred[wid] = v;
    __syncthreads();
    if (wid == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        v = lane < nw ? sred[lane] : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    }
    return v;
}

// Sums `n` partials in a fixed order into out[slot] (single block launch).
__global__ void k_reduce_partials(const double* __restrict__ partials, int n, int stride, int nsum,
                                  double* __restrict__ out);
// Host helper: sum partials[k*stride + c] for c < nsum into host out[nsum].
void reduce_partials(const double* d_partials, int n, int stride, int nsum, double* d_out);

// Exclusive scan of n ints into out[0..n] (out[n] = total), single block,
// deterministic; n up to a few million.
void exclusive_scan(const int* d_in, int* d_out, int n);
// Rows i < N of a CSR (head) sorted ascending from the unsorted entries tmp into out
// (entries of a row distinct).
void sort_rows(int N, const int* d_head, const int* d_tmp, int* d_out);

// ------------------------------------------------------------ geometry
// Minimum image and r^2 with the arithmetic rounded exactly as written (no
// FMA contraction; HIP compiles with -ffp-contract=fast-honor-pragmas), as
// MC/src/neighbor.cu:15 and :37 evaluate them.
__device__ __forceinline__ double min_image_exact(double dx, double L) {
#pragma clang fp contract(off)
    return dx - L * rint(dx / L);
}
__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// B-spline weights w[k] = M_p(t + p - 1 - k) for grid points i0 - off + k,
// off = p - 1 - p/2.  Order 4 uses the reference's cubic formula
// (MC/src/pme.cu:132-138) verbatim; other orders the Cox-de Boor recursion.
template <typename T>
__host__ __device__ __forceinline__ void bspline_weights(int p, T t, T* w) {
    if (p == 4) {
        T t2 = t * t, t3 = t2 * t;
        w[0] = (T(1) - 3 * t + 3 * t2 - t3) / T(6.0);
        w[1] = (T(4) - 6 * t2 + 3 * t3) / T(6.0);
        w[2] = (T(1) + 3 * t + 3 * t2 - 3 * t3) / T(6.0);
        w[3] = t3 / T(6.0);
        return;
    }
    T v[8];
    v[0] = t; v[1] = T(1) - t;
    for (int n = 2; n < p; ++n) {
        T nv[8];
        for (int j = 0; j <= n; ++j) {
            T u = t + T(j);
            T a = (j < n) ? v[j] : T(0);
            T b = (j > 0) ? v[j - 1] : T(0);
            nv[j] = (u * a + (T(n + 1) - u) * b) / T(n);
        }
        for (int j = 0; j <= n; ++j) v[j] = nv[j];
    }
    for (int k = 0; k < p; ++k) w[k] = v[p - 1 - k];
}
__host__ __device__ __forceinline__ int stencil_offset(int p) { return p - 1 - p / 2; }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace phb
