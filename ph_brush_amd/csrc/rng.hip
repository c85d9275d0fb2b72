// rng.hip -- the chains' device engine stream as a standalone draw sequence.
//
// The chain kernels draw their proposals from a device MT19937 with
// libstdc++'s uniform_int_distribution (Lemire's nearly divisionless method)
// and generate_canonical<double,53> (MC/src/mc.cu:62-65, :368-369,
// MC/src/main.cu:180).  A round of up to 64 draws is read at fixed stream
// positions, one lane per draw; a round in which any Lemire draw lands in the
// rejection zone is redrawn serially.  With the pool sizes of a brush the
// rejection zone is ~1e-6 of the range, so the chains almost never take that
// branch; this entry point runs the same helpers on a caller-chosen mix of
// ranges (near 2^32/3 a third of all draws are rejected) so that both
// branches are compared with std::mt19937 draw by draw.
#include <sstream>
#include <random>

#include "mc_dev.hpp"

namespace phb {

// One wave.  hi[k] >= 0: uniform_int_distribution<int>(0, hi[k]); hi[k] < 0:
// generate_canonical<double,53> (the double's bits).  counts[0]: rounds redrawn
// serially.
__global__ void __launch_bounds__(64) k_rng_draws(const unsigned* __restrict__ state, long long n,
                                                  const int* __restrict__ hi, unsigned long long* __restrict__ out,
                                                  unsigned* __restrict__ state_after,
                                                  long long* __restrict__ counts) {
    __shared__ RngLds R;
    const int lane = threadIdx.x;
    for (int k = lane; k < 624; k += 64) {
        R.cur[k] = state[k];
        R.out[0][k] = mt_temper(state[k]);
    }
    __syncthreads();
    int blk = 0, pos = (int)state[624];
    long long serial = 0;
    for (long long r0 = 0; r0 < n; r0 += 64) {
        const int m = (int)(n - r0 < 64 ? n - r0 : 64);
        const bool act = lane < m;
        const int h = act ? hi[r0 + lane] : 0;
        const unsigned long long canon = __ballot(act && h < 0);
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int off = lane + __popcll(canon & below);
        const int total = m + __popcll(canon);
        rng_ensure(R, blk, pos + total, lane);
        bool ok = true;
        unsigned long long v = 0;
        if (act) {
            const int q = pos + off;
            if (h < 0) {
                v = (unsigned long long)__double_as_longlong(canonical_at(R, q));
            } else {
                unsigned u = 0;
                ok = index_at(R, q, (unsigned)h + 1u, u);
                v = u;
            }
        }
        if (__all(ok)) {
            if (act) out[r0 + lane] = v;
            pos += total;
        } else {
            ++serial;
            for (int k = 0; k < m; ++k) {
                const int hk = hi[r0 + k];
                unsigned long long w;
                if (hk < 0) w = (unsigned long long)__double_as_longlong(draw_canonical(R, blk, pos, lane));
                else w = draw_index(R, blk, pos, lane, (unsigned)hk + 1u);
                if (lane == 0) out[r0 + k] = w;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // engine state in its stream form, as the chains write it back
    const bool in_cur = pos >= blk * 624;
    const unsigned* st = in_cur ? R.cur : R.prev;
    for (int k = lane; k < 624; k += 64) state_after[k] = st[k];
    if (lane == 0) {
        state_after[624] = (unsigned)(in_cur ? pos - blk * 624 : pos - (blk - 1) * 624);
        counts[0] = serial;
    }
}

}  // namespace phb

using namespace phb;

extern "C" int phb_rng_seed_state(unsigned seed, unsigned* state) {
    return guard([&] {
        require(state != nullptr, "phb_rng_seed_state: null state");
        std::mt19937 eng(seed);
        std::stringstream ss;
        ss << eng;
        for (int k = 0; k < 625; ++k) ss >> state[k];
        return PHB_OK;
    });
}

extern "C" int phb_rng_draws(const unsigned* state, long long n, const int* hi, unsigned long long* out,
                             unsigned* state_after, long long* serial_rounds) {
    return guard([&] {
        require(state && (n == 0 || (hi && out)), "phb_rng_draws: null argument");
        require(n >= 0, "phb_rng_draws: negative count");
        require(state[624] <= 624u, "phb_rng_draws: engine index beyond 624");
        require_device();
        DevBuf<unsigned> d_state(625), d_after(625);
        DevBuf<int> d_hi(std::max<long long>(n, 1));
        DevBuf<unsigned long long> d_out(std::max<long long>(n, 1));
        DevBuf<long long> d_cnt(1);
        d_state.upload(state, 625);
        if (n) d_hi.upload(hi, (size_t)n);
        hipLaunchKernelGGL(k_rng_draws, dim3(1), dim3(64), 0, stream(), d_state.p, n, d_hi.p, d_out.p, d_after.p,
                           d_cnt.p);
        PHB_HIP(hipGetLastError());
        unsigned after[625];
        long long cnt = 0;
        if (n) d_out.download(out, (size_t)n);
        d_after.download(after, 625);
        d_cnt.download(&cnt, 1);
        sync();
        if (state_after) std::copy(after, after + 625, state_after);
        if (serial_rounds) *serial_rounds = cnt;
        return PHB_OK;
    });
}
