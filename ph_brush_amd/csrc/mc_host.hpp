// mc_host.hpp -- host object behind phb_mc and the host steps shared by the
// plain and delayed-acceptance chains.
#pragma once

#include <utility>
#include <vector>

#include "mc_dev.hpp"

using phb::DevBuf;
using phb::McCtl;
using phb::McDev;
using phb::stream;

struct phb_mc_s {
    phb_atoms_s* atoms = nullptr;
    phb_params_s* params = nullptr;
    phb_csr_s* neigh = nullptr;
    phb_pme_s* pme = nullptr;
    phb_pme_s* own_pme = nullptr;   // the coul_cut plan this object made for itself
    phb_box box{};
    phb_mc_desc d{};
    int nsites = 0, nions = 0;
    DevBuf<int> fhead, fnbr, sbase;
    DevBuf<double> kreal, r2, sw, phi, gk, gf, dgrid, selfw;
    DevBuf<float> kreal32, gf32, phi32;   // the FP32 form's copies
    DevBuf<int> pbuf;
    bool lds_pools = false;
    size_t pool_lds_bytes = 0;
    DevBuf<McCtl> ctl;
    DevBuf<unsigned> applied;   // per refresh workgroup: the pending list it folded
    DevBuf<int> rd_tiles;       // direct refresh: the tiles some atom's stencil reaches (tile 0 always)
    int rd_ntiles = 0;
    DevBuf<signed char> rec_fwd, rec_acc;
    DevBuf<int> rec_i, rec_j;
    DevBuf<double> rec_dE, rec_u;
    McDev dev{};
    long long refreshes = 0, recomputes = 0, launches = 0, next_recompute = 0, trials_seen = 0;
    // profiling: event triples around (chain launch, refresh) pairs and recomputes
    bool profiling = false;
    bool phase_clock = false;   // with profiling: the delayed-acceptance chain's in-kernel phase clocks
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<std::pair<size_t, int>> marks;   // (first event index, kind: 0 chain+refresh, 1 recompute)
    double chain_ms = 0, refresh_ms = 0, recompute_ms = 0;
    double z_cut = 0.0;        // pool z filter of W and Cl (max site z + w_z_extra)
    // pools phase 1 hands to phase 2 (nh_n, nh_p, w, cl), when it ran
    std::vector<int> handover[4];
    bool has_handover = false;
    std::vector<phb_phase1_event> p1_events;   // what the last phase-1 run did, block by block
    // delayed-acceptance state (descriptor lists and histograms) and phase-2 state, owned here
    void* da = nullptr;
    void (*da_free)(void*) = nullptr;
    void* p2 = nullptr;
    void (*p2_free)(void*) = nullptr;
    ~phb_mc_s() {
        if (da_free) da_free(da);
        if (p2_free) p2_free(p2);
        delete own_pme;
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

namespace phb {
// Launches to queue before reading the control block again.
int launch_batch(const McCtl& c, long long remaining, int pending_max);
// Full (both directions) neighbour rows, ascending, from a half list; returns the entry count.
long long build_full_rows(int N, const phb_csr_s* half, DevBuf<int>& head, DevBuf<int>& nbr);
McCtl read_ctl(phb_mc_s* M);
// Direct-refresh tile list of the object's tile shape (M->dev.rd_ty x rd_tz).
void build_rd_tiles(phb_mc_s* M);
// Direct-refresh tile shape (folds the pending changes first; rebuilds the tile list).
void set_refresh_tile(phb_mc_s* M, int ty, int tz);
// Full recompute: phi = G Q from scratch, total energy, moments, E reset.
void full_recompute(phb_mc_s* M);
// ntrials more trials of the plain chain (refreshes, drift resets); ends early
// at a non-zero status (1 no valid pair, 4 nprot reached its target)
void run_plain(phb_mc_s* M, long long ntrials);
// Folds the pending accepted charge changes into phi (k-space refresh).
void refresh(phb_mc_s* M);
bool refresh_direct_allowed(const phb_mc_s* M);   // grid fits the direct refresh
bool use_direct_refresh(const phb_mc_s* M);
void fill_stats(phb_mc_s* M, const McCtl& c, phb_mc_stats* s);
void launch_probe(phb_mc_s* M, int n, const int* i, const int* j, const int* f, double* out);
}  // namespace phb
