// nlist.hip -- half neighbour list from binned cells (replaces the reference's
// O(N^2) k_count_half / k_fill_half, MC/src/neighbor.cu:21-63).
//
// Predicate, identical to the reference so that the CSR is bit-exact:
//   i < j, j not 1-2 bonded to i, min-image r^2 <= rc^2 (inclusive), with
//   dx -= L*rint(dx/L) and r^2 = dx*dx + dy*dy + dz*dz rounded as written.
// Rows come out ascending in j (the reference's order, MC/src/neighbor.cu:55-61).
//
// Layout: atoms are counting-sorted into cells of width >= rc*(1+1e-9) (one
// cell along an axis shorter than three widths).  One wave per atom i walks
// the 27 neighbouring cells 64 candidates at a time; a wavefront ballot
// compacts the hits.  Count pass -> scan -> fill pass -> per-row rank sort.
#include <algorithm>
#include <memory>

#include "device.hpp"
#include "handles.hpp"

namespace {

using namespace phb;

struct CellGeom {
    int ncx, ncy, ncz;
    double Lx, Ly, Lz;
    double ix, iy, iz;   // 1/L along axes with >= 3 cells, else 0 (divide)
};

// dx - L rint(dx/L).  Along an axis with >= 3 cells (L >= 3 rc) the quotient is
// taken as dx * (1/L): it can round differently from dx / L only when dx is
// within an ulp of +-L/2, where either image lies at |dx| ~ L/2 > rc and the
// pair fails the cutoff whichever is taken, so the list is unchanged.
__device__ __forceinline__ double min_image_cell(double dx, double L, double inv) {
#pragma clang fp contract(off)
    return inv != 0.0 ? dx - L * rint(dx * inv) : dx - L * rint(dx / L);
}

__device__ __forceinline__ int cell_of(double v, double L, int nc) {
    int c = (int)floor(v / L * nc);
    return min(max(c, 0), nc - 1);
}

__global__ void k_cell_count(const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int N, CellGeom g, int* __restrict__ cell,
                             int* __restrict__ ccount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c = (cell_of(z[i], g.Lz, g.ncz) * g.ncy + cell_of(y[i], g.Ly, g.ncy)) * g.ncx + cell_of(x[i], g.Lx, g.ncx);
    cell[i] = c;
    atomicAdd(&ccount[c], 1);
}

__global__ void k_cell_fill(const int* __restrict__ cell, int N, const int* __restrict__ cstart,
                            int* __restrict__ cursor, int* __restrict__ catoms) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c = cell[i];
    int slot = atomicAdd(&cursor[c], 1);
    catoms[cstart[c] + slot] = i;
}

// Positions in cell order, so that a wave reading a cell's candidates loads
// consecutive addresses.
__global__ void k_cell_coords(const int* __restrict__ catoms, int N, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ xs,
                              double* __restrict__ ys, double* __restrict__ zs) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = catoms[k];
    xs[k] = x[i]; ys[k] = y[i]; zs[k] = z[i];
}

__device__ __forceinline__ bool bonded12(int i, int j, const int* __restrict__ bh, const int* __restrict__ bl) {
    if (!bh) return false;
    int lo = bh[i], hi = bh[i + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int v = bl[mid];
        if (v == j) return true;
        if (v < j) lo = mid + 1; else hi = mid - 1;
    }
    return false;
}

// MODE 0: count into counts[i]; MODE 1: write unsorted row into tmp at head[i].
template <int MODE>
__global__ __launch_bounds__(256) void k_pairs(const double* __restrict__ x, const double* __restrict__ y,
                                               const double* __restrict__ z, const double* __restrict__ xs,
                                               const double* __restrict__ ys, const double* __restrict__ zs, int N,
                                               CellGeom g, double rc2,
                                               const int* __restrict__ cell, const int* __restrict__ cstart,
                                               const int* __restrict__ catoms, const int* __restrict__ bh,
                                               const int* __restrict__ bl, int* __restrict__ counts,
                                               const int* __restrict__ head, int* __restrict__ tmp) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= N) return;  // wave-uniform
    const double xi = x[i], yi = y[i], zi = z[i];
    const int c = cell[i];
    const int cx = c % g.ncx, cy = (c / g.ncx) % g.ncy, cz = c / (g.ncx * g.ncy);
    const int ox = g.ncx >= 3 ? 1 : 0, oy = g.ncy >= 3 ? 1 : 0, oz = g.ncz >= 3 ? 1 : 0;
    int found = 0;
    const int base = MODE == 1 ? head[i] : 0;
    // i's 1-2 bond partners in registers (rows of up to four; longer rows are searched)
    int bn = 0, b0 = -1, b1 = -1, b2 = -1, b3 = -1;
    if (bh) {
        const int blo = bh[i];
        bn = bh[i + 1] - blo;
        if (bn > 0) b0 = bl[blo];
        if (bn > 1) b1 = bl[blo + 1];
        if (bn > 2) b2 = bl[blo + 2];
        if (bn > 3) b3 = bl[blo + 3];
    }
    // the x-adjacent cells of one (dy, dz) row are adjacent in the cell order: scan them as
    // one range (two where the row wraps), so a round of 64 candidates spans cells
    const int xa = cx - ox, xb = cx + ox;   // cells xa..xb along x, wrapped
    for (int dz = -oz; dz <= oz; ++dz) {
        const int kz = (cz + dz + g.ncz) % g.ncz;
        for (int dy = -oy; dy <= oy; ++dy) {
            const int ky = (cy + dy + g.ncy) % g.ncy;
            const int rowc = (kz * g.ncy + ky) * g.ncx;
            for (int part = 0; part < 2; ++part) {
                int lo, hi;   // cell range [lo, hi] of this part along x
                if (xa >= 0 && xb < g.ncx) { if (part) break; lo = xa; hi = xb; }
                else if (xa < 0) { if (part) { lo = 0; hi = xb; } else { lo = xa + g.ncx; hi = g.ncx - 1; } }
                else { if (part) { lo = 0; hi = xb - g.ncx; } else { lo = xa; hi = g.ncx - 1; } }
                const int s = cstart[rowc + lo], e = cstart[rowc + hi + 1];
                for (int b = s; b < e; b += 64) {
                    const int idx = b + lane;
                    bool pass = false;
                    int j = -1;
                    if (idx < e) {
                        j = catoms[idx];
                        if (j > i) {
                            const double ddx = min_image_cell(xi - xs[idx], g.Lx, g.ix);
                            const double ddy = min_image_cell(yi - ys[idx], g.Ly, g.iy);
                            const double ddz = min_image_cell(zi - zs[idx], g.Lz, g.iz);
                            pass = r2_exact(ddx, ddy, ddz) <= rc2 &&
                                   !(bn <= 4 ? (j == b0 || j == b1 || j == b2 || j == b3) : bonded12(i, j, bh, bl));
                        }
                    }
                    const unsigned long long m = __ballot(pass);
                    if (MODE == 1 && pass) {
                        const int pre = __popcll(m & ((1ull << lane) - 1ull));
                        tmp[base + found + pre] = j;
                    }
                    found += __popcll(m);
                }
            }
        }
    }
    if (MODE == 0 && lane == 0) counts[i] = found;
}

int cells_along(double L, double rc) {
    int nc = (int)std::floor(L / (rc * (1.0 + 1e-9)));
    return nc < 3 ? 1 : nc;
}

}  // namespace

namespace {

// Per-block maxima of the squared min-image displacement since the snapshot.
__global__ __launch_bounds__(256) void k_max_disp(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const double* __restrict__ x0,
                                                  const double* __restrict__ y0, const double* __restrict__ z0, int N,
                                                  double Lx, double Ly, double Lz, double* __restrict__ part) {
    __shared__ double s[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double d2 = 0.0;
    if (i < N) {
        const double dx = min_image_exact(x[i] - x0[i], Lx), dy = min_image_exact(y[i] - y0[i], Ly),
                     dz = min_image_exact(z[i] - z0[i], Lz);
        d2 = r2_exact(dx, dy, dz);
    }
    for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
    if (lane == 0) s[w] = d2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
}

// MODE 0: count the entries of the outer row i within rc; MODE 1: copy them (order kept).
template <int MODE>
__global__ void k_filter_rows(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                              int N, double Lx, double Ly, double Lz, double rc2, const int* __restrict__ oh,
                              const int* __restrict__ ol, int* __restrict__ counts, const int* __restrict__ head,
                              int* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    int n = 0;
    for (int e = oh[i]; e < oh[i + 1]; ++e) {
        const int j = ol[e];
        const double dx = min_image_exact(xi - x[j], Lx), dy = min_image_exact(yi - y[j], Ly),
                     dz = min_image_exact(zi - z[j], Lz);
        if (r2_exact(dx, dy, dz) <= rc2) {
            if (MODE == 1) list[head[i] + n] = j;
            ++n;
        }
    }
    if (MODE == 0) counts[i] = n;
}

void build_half(phb_atoms_s* a, const phb_box* box, phb_csr_s* bonds12, double rc, phb_csr_s* c) {
    const int N = a->n;
    c->rows = N;
    c->nnz = 0;
    c->head.alloc(N + 1);
    c->list.release();
    if (N == 0) {
        int zero = 0;
        c->head.upload(&zero, 1);
        sync();
        return;
    }
    {
        CellGeom g{cells_along(box->Lx, rc), cells_along(box->Ly, rc), cells_along(box->Lz, rc), box->Lx, box->Ly,
                   box->Lz, 0.0, 0.0, 0.0};
        g.ix = g.ncx >= 3 ? 1.0 / box->Lx : 0.0;
        g.iy = g.ncy >= 3 ? 1.0 / box->Ly : 0.0;
        g.iz = g.ncz >= 3 ? 1.0 / box->Lz : 0.0;
        const long long ncell = (long long)g.ncx * g.ncy * g.ncz;
        require(ncell < (1LL << 30), "build_half_neighbors_gpu: too many cells");
        hipStream_t st = stream();
        DevBuf<int> cell(N), ccount(ncell), cstart(ncell + 1), cursor(ncell), catoms(N), counts(N);
        DevBuf<double> xs(N), ys(N), zs(N);
        PHB_HIP(hipMemsetAsync(ccount.p, 0, ncell * sizeof(int), st));
        PHB_HIP(hipMemsetAsync(cursor.p, 0, ncell * sizeof(int), st));
        const int tb = 256;
        hipLaunchKernelGGL(k_cell_count, dim3(ceil_div(N, tb)), dim3(tb), 0, st, a->x.p, a->y.p, a->z.p, N, g, cell.p,
                           ccount.p);
        exclusive_scan(ccount.p, cstart.p, (int)ncell);
        hipLaunchKernelGGL(k_cell_fill, dim3(ceil_div(N, tb)), dim3(tb), 0, st, cell.p, N, cstart.p, cursor.p,
                           catoms.p);
        hipLaunchKernelGGL(k_cell_coords, dim3(ceil_div(N, tb)), dim3(tb), 0, st, catoms.p, N, a->x.p, a->y.p, a->z.p,
                           xs.p, ys.p, zs.p);
        const int* bh = bonds12 && bonds12->nnz > 0 ? bonds12->head.p : nullptr;
        const int* bl = bonds12 && bonds12->nnz > 0 ? bonds12->list.p : nullptr;
        const double rc2 = rc * rc;
        const int wpb = tb / 64;
        hipLaunchKernelGGL(k_pairs<0>, dim3(ceil_div(N, wpb)), dim3(tb), 0, st, a->x.p, a->y.p, a->z.p, xs.p, ys.p,
                           zs.p, N, g, rc2,
                           cell.p, cstart.p, catoms.p, bh, bl, counts.p, (const int*)nullptr, (int*)nullptr);
        PHB_HIP(hipGetLastError());
        exclusive_scan(counts.p, c->head.p, N);
        int total = 0;
        PHB_HIP(hipMemcpyAsync(&total, c->head.p + N, sizeof(int), hipMemcpyDeviceToHost, st));
        PHB_HIP(hipStreamSynchronize(st));
        require(total >= 0, "build_half_neighbors_gpu: neighbour count overflow", PHB_ERR_INTERNAL);
        c->nnz = total;
        c->list.alloc(total);
        if (total > 0) {
            DevBuf<int> tmp(total);
            hipLaunchKernelGGL(k_pairs<1>, dim3(ceil_div(N, wpb)), dim3(tb), 0, st, a->x.p, a->y.p, a->z.p, xs.p,
                               ys.p, zs.p, N, g, rc2, cell.p, cstart.p, catoms.p, bh, bl, (int*)nullptr, c->head.p,
                               tmp.p);
            sort_rows(N, c->head.p, tmp.p, c->list.p);
            PHB_HIP(hipGetLastError());
            sync();
        }
    }
}

}  // namespace

extern "C" int phb_nlist_build_half(phb_atoms a, const phb_box* box, phb_csr bonds12, double rc, phb_csr* out) {
    return guard([&] {
        require(a && box && out, "build_half_neighbors_gpu: null argument");
        require(rc > 0.0, "build_half_neighbors_gpu: rc must be > 0");
        require(box->Lx > 0 && box->Ly > 0 && box->Lz > 0, "build_half_neighbors_gpu: box must be positive");
        if (bonds12) require(bonds12->rows == a->n, "build_half_neighbors_gpu: bond CSR rows != natoms");
        auto c = std::make_unique<phb_csr_s>();
        build_half(a, box, bonds12, rc, c.get());
        *out = c.release();
        return PHB_OK;
    });
}

// ------------------------------------------------------------ skin (Verlet) list
struct phb_verlet_s {
    phb_box box{};
    phb_csr_s* bonds = nullptr;
    double rc = 0.0, skin = 0.0;
    phb_csr_s outer;                       // half list at rc + skin from the snapshot
    DevBuf<double> x0, y0, z0;             // positions at the last build
    DevBuf<double> part;
    long long builds = 0;
    double last_disp = 0.0;                // largest displacement seen by the last update
};

namespace {

void verlet_build(phb_verlet_s* v, phb_atoms_s* a) {
    build_half(a, &v->box, v->bonds, v->rc + v->skin, &v->outer);
    const size_t N = (size_t)a->n;
    v->x0.alloc(std::max<size_t>(N, 1)); v->y0.alloc(std::max<size_t>(N, 1)); v->z0.alloc(std::max<size_t>(N, 1));
    if (N) {
        PHB_HIP(hipMemcpyAsync(v->x0.p, a->x.p, N * sizeof(double), hipMemcpyDeviceToDevice, stream()));
        PHB_HIP(hipMemcpyAsync(v->y0.p, a->y.p, N * sizeof(double), hipMemcpyDeviceToDevice, stream()));
        PHB_HIP(hipMemcpyAsync(v->z0.p, a->z.p, N * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    }
    sync();
    ++v->builds;
}

}  // namespace

extern "C" int phb_verlet_create(phb_atoms a, const phb_box* box, phb_csr bonds12, double rc, double skin,
                                 phb_verlet* out) {
    return guard([&] {
        require(a && box && out, "phb_verlet_create: null argument");
        require(rc > 0.0 && skin >= 0.0, "phb_verlet_create: need rc > 0 and skin >= 0");
        require(box->Lx > 0 && box->Ly > 0 && box->Lz > 0, "phb_verlet_create: box must be positive");
        if (bonds12) require(bonds12->rows == a->n, "phb_verlet_create: bond CSR rows != natoms");
        auto v = std::make_unique<phb_verlet_s>();
        v->box = *box; v->bonds = bonds12; v->rc = rc; v->skin = skin;
        verlet_build(v.get(), a);
        *out = v.release();
        return PHB_OK;
    });
}

extern "C" int phb_verlet_destroy(phb_verlet v) {
    return guard([&] { delete v; return PHB_OK; });
}

extern "C" int phb_verlet_update(phb_verlet v, phb_atoms a, int* rebuilt, double* max_disp) {
    return guard([&] {
        require(v && a, "phb_verlet_update: null argument");
        require(a->n == v->outer.rows, "phb_verlet_update: atom count changed");
        const int N = a->n;
        double m2 = 0.0;
        if (N > 0) {
            const int nb = ceil_div(N, 256);
            v->part.alloc(nb);
            hipLaunchKernelGGL(k_max_disp, dim3(nb), dim3(256), 0, stream(), a->x.p, a->y.p, a->z.p, v->x0.p, v->y0.p,
                               v->z0.p, N, v->box.Lx, v->box.Ly, v->box.Lz, v->part.p);
            PHB_HIP(hipGetLastError());
            std::vector<double> h(nb);
            v->part.download(h.data(), nb);
            sync();
            for (double d : h) m2 = std::max(m2, d);
        }
        v->last_disp = std::sqrt(m2);
        // every pair now within rc was within rc + skin at the build while no atom moved more than skin / 2
        const bool need = 2.0 * v->last_disp > v->skin;
        if (need) verlet_build(v, a);
        if (rebuilt) *rebuilt = need ? 1 : 0;
        if (max_disp) *max_disp = v->last_disp;
        return PHB_OK;
    });
}

extern "C" int phb_verlet_half(phb_verlet v, phb_atoms a, phb_csr* out) {
    return guard([&] {
        require(v && a && out, "phb_verlet_half: null argument");
        require(a->n == v->outer.rows, "phb_verlet_half: atom count changed");
        const int N = a->n;
        auto c = std::make_unique<phb_csr_s>();
        c->rows = N;
        c->head.alloc(N + 1);
        if (N == 0) {
            int zero = 0;
            c->head.upload(&zero, 1);
            sync();
            *out = c.release();
            return PHB_OK;
        }
        DevBuf<int> counts(N);
        const double rc2 = v->rc * v->rc;
        hipLaunchKernelGGL(k_filter_rows<0>, dim3(ceil_div(N, 256)), dim3(256), 0, stream(), a->x.p, a->y.p, a->z.p, N,
                           v->box.Lx, v->box.Ly, v->box.Lz, rc2, v->outer.head.p, v->outer.list.p, counts.p,
                           (const int*)nullptr, (int*)nullptr);
        exclusive_scan(counts.p, c->head.p, N);
        int total = 0;
        PHB_HIP(hipMemcpyAsync(&total, c->head.p + N, sizeof(int), hipMemcpyDeviceToHost, stream()));
        sync();
        c->nnz = total;
        c->list.alloc(total);
        if (total > 0)
            hipLaunchKernelGGL(k_filter_rows<1>, dim3(ceil_div(N, 256)), dim3(256), 0, stream(), a->x.p, a->y.p,
                               a->z.p, N, v->box.Lx, v->box.Ly, v->box.Lz, rc2, v->outer.head.p, v->outer.list.p,
                               (int*)nullptr, c->head.p, c->list.p);
        PHB_HIP(hipGetLastError());
        sync();
        *out = c.release();
        return PHB_OK;
    });
}

extern "C" int phb_verlet_stats(phb_verlet v, long long* builds, long long* outer_nnz, double* last_disp) {
    return guard([&] {
        require(v, "phb_verlet_stats: null argument");
        if (builds) *builds = v->builds;
        if (outer_nnz) *outer_nnz = v->outer.nnz;
        if (last_disp) *last_disp = v->last_disp;
        return PHB_OK;
    });
}
