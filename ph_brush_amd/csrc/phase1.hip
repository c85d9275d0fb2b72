// phase1.hip -- phase 1 as main() runs it (MC/src/main.cu:180-197) around the
// device chain: build_plan's checks (MC/src/mc.cu:67-122), then either
// apply_direct_protonation (MC/src/mc.cu:124-157) when the target exceeds
// direct_pct_threshold, or phase_gc_mu_control's loop (MC/src/mc.cu:320-546):
// mu from the target fraction, blocks of gc_block trials of the device chain,
// an epoch ending between blocks once Nacc_target moves were accepted or Amax
// trials made, then mu -= kT eta logit_diff w_info; the chain stops at the trial
// whose acceptance reaches the target (the device checks nprot before every
// round).  The pools phase 2 starts from are kept on the MC object.
#include <algorithm>
#include <cmath>
#include <sstream>
#include <string>
#include <vector>

#include "mc_host.hpp"

namespace phb {

__global__ void k_p1_control(McCtl* C, int nprot_target, int clear_status) {
    if (threadIdx.x == 0) {
        C->nprot_target = nprot_target;
        if (clear_status && C->status == 4) C->status = 0;
    }
}

__global__ void k_p1_pools(McCtl* C, const int* __restrict__ sizes, int nprot) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) C->sizes[k] = sizes[k];
        C->nprot = nprot;
    }
}

}  // namespace phb

using namespace phb;

namespace {

struct Ident {
    std::vector<int> type;
    std::vector<double> z;
};

Ident identities(phb_mc_s* M) {
    const int N = M->atoms->n;
    Ident h;
    h.type.resize(N);
    h.z.resize(N);
    M->atoms->type.download(h.type.data(), N);
    M->atoms->z.download(h.z.data(), N);
    sync();
    return h;
}

// the chain's pools (build_plan order) from the identities, and its counters
void rebuild_chain_pools(phb_mc_s* M, const Ident& h) {
    const int N = M->atoms->n, ns = M->nsites, ni = M->nions;
    std::vector<int> nh_n, nh_p, w, cl;
    for (int i = 0; i < N; ++i) {
        const int t = h.type[i] + 1;
        if (t == M->d.nh_n_type) nh_n.push_back(i);
        else if (t == M->d.nh_p_type) nh_p.push_back(i);
        else if (t == M->d.w_type) { if (h.z[i] < M->z_cut) w.push_back(i); }
        else if (t == M->d.cl_type) { if (h.z[i] < M->z_cut) cl.push_back(i); }
    }
    require((int)(nh_n.size() + nh_p.size()) == ns && (int)(w.size() + cl.size()) == ni,
            "phase 1: pool membership differs from the chain's pools", PHB_ERR_STATE);
    std::vector<int> buf(std::max(ns + ni, 1), -1);
    for (size_t k = 0; k < nh_n.size(); ++k) buf[k] = nh_n[k];
    for (size_t k = 0; k < nh_p.size(); ++k) buf[ns - 1 - k] = nh_p[k];
    for (size_t k = 0; k < w.size(); ++k) buf[ns + k] = w[k];
    for (size_t k = 0; k < cl.size(); ++k) buf[ns + ni - 1 - k] = cl[k];
    M->pbuf.upload(buf.data(), buf.size());
    const int sz[4] = {(int)nh_n.size(), (int)nh_p.size(), (int)w.size(), (int)cl.size()};
    DevBuf<int> d(4);
    d.upload(sz, 4);
    hipLaunchKernelGGL(k_p1_pools, dim3(1), dim3(64), 0, stream(), M->ctl.p, d.p, (int)nh_p.size());
    PHB_HIP(hipGetLastError());
    sync();
}

// the phase-1 hand-over when the target is reached (MC/src/mc.cu:480-494)
void handover_from_identities(phb_mc_s* M, const Ident& h) {
    for (auto& v : M->handover) v.clear();
    for (int i = 0; i < M->atoms->n; ++i) {
        const int t = h.type[i] + 1;
        if (t == M->d.nh_p_type) M->handover[1].push_back(i);
        else if (t == M->d.nh_n_type) M->handover[0].push_back(i);
        else if (t == M->d.w_type) { if (h.z[i] < M->z_cut) M->handover[2].push_back(i); }
        else if (t == M->d.cl_type) { if (h.z[i] < M->z_cut) M->handover[3].push_back(i); }
    }
    M->has_handover = true;
}

}  // namespace

// The counters and state every phase-1 event carries, relative to the run's start.
// Records a phase-1 event and hands it to the caller's callback, if any (the
// callback runs on this thread, between blocks, and must not throw).
static void emit(phb_mc_s* M, const phb_phase1_desc* d, const phb_phase1_event& e) {
    M->p1_events.push_back(e);
    if (d->on_event) d->on_event(&e, d->user);
}

static phb_phase1_event event_of(const McCtl& c, long long a0, long long acc0, long long f0, long long b0, int epoch,
                                 int nh, int target_count, double mu_over_kT) {
    phb_phase1_event e{};
    e.epoch = epoch;
    e.attempted = c.trials_done - a0;
    e.accepted = c.accepted - acc0;
    e.fwd_total = c.acc_fwd - f0;
    e.bwd_total = c.acc_bwd - b0;
    e.nprot = c.nprot;
    e.nh = nh;
    e.target_count = target_count;
    e.remaining = std::max(0, target_count - c.nprot);
    e.mu_over_kT = mu_over_kT;
    e.E = c.E;
    return e;
}

extern "C" int phb_phase1_run(phb_mc M, const phb_phase1_desc* d, phb_phase1_result* out) {
    return guard([&] {
        require(M && d && out, "phb_phase1_run: null argument");
        McCtl c = read_ctl(M);
        require(c.status == 0, "phb_phase1_run: the chain is stopped (status != 0)", PHB_ERR_STATE);
        *out = phb_phase1_result{};
        M->p1_events.clear();
        // ---- build_plan (MC/src/mc.cu:67-122)
        Ident h = identities(M);
        const int N = M->atoms->n;
        std::vector<int> idx_nh_n, idx_nh_p, idx_w;
        for (int i = 0; i < N; ++i) {
            const int t = h.type[i] + 1;
            if (t == M->d.nh_n_type) idx_nh_n.push_back(i);
            else if (t == M->d.nh_p_type) idx_nh_p.push_back(i);
            else if (t == M->d.w_type) idx_w.push_back(i);
        }
        const int nh_total = (int)(idx_nh_n.size() + idx_nh_p.size());
        require(nh_total > 0, "[plan] No NH sites found.", PHB_ERR_STATE);
        const int target_count = (int)std::floor(d->target_pct / 100.0 * nh_total);
        const int to_flip = target_count - (int)idx_nh_p.size();
        require(to_flip >= 0, "[plan] Target < current; de-protonation not supported.", PHB_ERR_STATE);
        require((int)idx_nh_n.size() >= to_flip, "[plan] Not enough NH_N to flip.", PHB_ERR_STATE);
        { std::vector<int> f; for (int j : idx_w) if (h.z[j] < M->z_cut) f.push_back(j); idx_w.swap(f); }
        if ((int)idx_w.size() < to_flip) {
            std::ostringstream oss;
            oss << "[plan] Not enough W below z<" << M->z_cut << " nm: need " << to_flip << ", have " << idx_w.size();
            const std::string msg = oss.str();
            require(false, msg.c_str(), PHB_ERR_STATE);
        }
        out->Nh = nh_total;
        out->target_count = target_count;
        const long long a0 = c.trials_done, acc0 = c.accepted, f0 = c.acc_fwd, b0 = c.acc_bwd;
        if (d->target_pct > (double)d->direct_pct_threshold) {
            // ---- apply_direct_protonation: the first to_flip NH_N and W of the plan
            std::vector<int> idx, ty;
            std::vector<double> q;
            for (int k = 0; k < to_flip; ++k) {
                idx.push_back(idx_nh_n[k]); ty.push_back(M->d.nh_p_type - 1); q.push_back(1.0);
                idx.push_back(idx_w[k]); ty.push_back(M->d.cl_type - 1); q.push_back(-1.0);
            }
            if (!idx.empty()) {
                refresh(M);
                const int rc = phb_atoms_set_identity(M->atoms, (int)idx.size(), idx.data(), ty.data(), q.data());
                const std::string msg = std::string("apply_direct_protonation: ") + phb_last_error();
                require(rc == PHB_OK, msg.c_str(), PHB_ERR_STATE);
            }
            for (auto& v : M->handover) v.clear();
            M->handover[0].assign(idx_nh_n.begin() + to_flip, idx_nh_n.end());
            M->handover[2].assign(idx_w.begin() + to_flip, idx_w.end());
            for (int k = 0; k < to_flip; ++k) { M->handover[1].push_back(idx_nh_n[k]); M->handover[3].push_back(idx_w[k]); }
            M->has_handover = true;
            h = identities(M);
            rebuild_chain_pools(M, h);
            full_recompute(M);   // potential grid, moments and energy of the new identities
            c = read_ctl(M);
            out->direct = 1;
            if (to_flip > 0) {
                phb_phase1_event e{};
                e.kind = PHB_P1_DIRECT;
                e.attempted = e.accepted = to_flip;
                e.nprot = c.nprot; e.nh = nh_total; e.target_count = target_count;
                e.mu_over_kT = M->d.beta * M->dev.mu;
                e.E = c.E;
                emit(M, d, e);
            }
            out->attempted = to_flip;
            out->accepted = to_flip;
            out->Nprot = c.nprot;
            out->reached_target = 1;
            out->E = c.E;
            out->mu_eff = M->dev.mu;
            return PHB_OK;
        }
        // ---- phase_gc_mu_control (MC/src/mc.cu:320-546)
        const double beta = M->d.beta;
        const double fstar = std::clamp(d->target_pct / 100.0, 1e-6, 1.0 - 1e-6);
        double mu = (1.0 / beta) * std::log(fstar / (1.0 - fstar));
        const double mu_bound = 10.0 / beta;
        mu = std::clamp(mu, -mu_bound, mu_bound);
        M->dev.mu = mu;
        M->d.mu_eff = mu;
        for (auto& v : M->handover) v.clear();
        M->has_handover = true;       // empty unless the target is reached below
        int nprot = c.nprot;
        if (nprot != target_count) {
            const int blk = d->gc_block > 0 ? d->gc_block : 2000;
            const int Amax = d->epoch_attempt_max > 0 ? d->epoch_attempt_max : 100000;
            const int R0 = std::max(0, target_count - nprot);
            const double frac = d->nacc_frac > 0.0 ? d->nacc_frac : 0.10;
            const int Nacc_target = std::max(1, (int)std::ceil(frac * (double)R0));
            const double eta = d->mu_eta > 0.0 ? d->mu_eta : 0.3;
            hipLaunchKernelGGL(k_p1_control, dim3(1), dim3(64), 0, stream(), M->ctl.p, target_count, 0);
            PHB_HIP(hipGetLastError());
            bool done_all = false;
            for (int epoch = 0; !done_all; ++epoch) {
                long long acc_since = 0, att_since = 0;
                while (!done_all && acc_since < Nacc_target && att_since < Amax) {
                    long long n = blk;
                    if (d->max_trials > 0) n = std::min<long long>(n, d->max_trials - (c.trials_done - a0));
                    if (n <= 0) { done_all = true; break; }
                    const McCtl before = c;
                    run_plain(M, n);
                    c = read_ctl(M);
                    acc_since += c.accepted - before.accepted;
                    att_since += c.trials_done - before.trials_done;
                    {
                        phb_phase1_event e = event_of(c, a0, acc0, f0, b0, epoch, nh_total, target_count, beta * mu);
                        e.kind = c.status == 4 ? PHB_P1_REACHED : c.status != 0 ? PHB_P1_NO_PAIRS : PHB_P1_CHUNK;
                        e.acc_since = acc_since; e.att_since = att_since;
                        e.block_attempted = c.trials_done - before.trials_done;
                        e.block_accepted = c.accepted - before.accepted;
                        e.step_in_block = (int)e.block_attempted;
                        e.r0 = R0; e.nacc_target = Nacc_target; e.amax = Amax; e.gc_block = blk;
                        emit(M, d, e);
                    }
                    if (c.status == 4) {
                        out->reached_target = 1;
                        handover_from_identities(M, identities(M));
                        done_all = true;
                    } else if (c.status != 0) {
                        done_all = true;   // 1: no valid pair to flip (MC/src/mc.cu:397-401)
                    }
                }
                if (done_all) break;
                const double f = nh_total > 0 ? (double)c.nprot / (double)nh_total : 0.0;
                const double ff0 = std::clamp(fstar, 1e-6, 1.0 - 1e-6), ff1 = std::clamp(f, 1e-6, 1.0 - 1e-6);
                const double logit_diff = std::log(ff1 / (1.0 - ff1)) - std::log(ff0 / (1.0 - ff0));
                const double w_info =
                    std::clamp((double)acc_since / (double)std::max(1, Nacc_target), 0.0, 1.0);
                const double mu_step = (1.0 / beta) * eta * logit_diff * w_info;
                mu -= mu_step;
                M->dev.mu = mu;
                M->d.mu_eff = mu;
                out->blocks_done = epoch + 1;
                {
                    phb_phase1_event e = event_of(c, a0, acc0, f0, b0, epoch, nh_total, target_count, beta * mu);
                    e.kind = PHB_P1_MU_UPDATE;
                    e.acc_since = acc_since; e.att_since = att_since;
                    e.early_stop = att_since >= Amax && acc_since < Nacc_target;
                    e.r0 = R0; e.nacc_target = Nacc_target; e.amax = Amax; e.gc_block = blk;
                    e.f = f;
                    e.w_info = w_info;
                    emit(M, d, e);
                }
            }
            hipLaunchKernelGGL(k_p1_control, dim3(1), dim3(64), 0, stream(), M->ctl.p, -1, 1);
            PHB_HIP(hipGetLastError());
            sync();
            c = read_ctl(M);
        } else {
            out->reached_target = 1;  // the reference returns before filling the hand-over (MC/src/mc.cu:346-353)
            phb_phase1_event e = event_of(c, a0, acc0, f0, b0, 0, nh_total, target_count, beta * mu);
            e.kind = PHB_P1_AT_TARGET;
            emit(M, d, e);
        }
        out->attempted = c.trials_done - a0;
        out->accepted = c.accepted - acc0;
        out->accepted_forward = c.acc_fwd - f0;
        out->accepted_backward = c.acc_bwd - b0;
        out->Nprot = c.nprot;
        out->status = c.status;
        out->E = c.E;
        out->mu_eff = mu;
        return PHB_OK;
    });
}

extern "C" int phb_phase1_handover(phb_mc M, int sizes[4], int* lists, long long cap) {
    return guard([&] {
        require(M && sizes, "phb_phase1_handover: null argument");
        long long total = 0;
        for (int k = 0; k < 4; ++k) {
            sizes[k] = M->has_handover ? (int)M->handover[k].size() : 0;
            total += sizes[k];
        }
        if (!lists) return PHB_OK;
        require(cap >= total, "phb_phase1_handover: lists capacity below the hand-over size");
        long long o = 0;
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < sizes[k]; ++j) lists[o++] = M->handover[k][j];
        return PHB_OK;
    });
}

extern "C" int phb_phase1_events(phb_mc M, phb_phase1_event* ev, int cap, int* n) {
    return guard([&] {
        require(M && n && cap >= 0, "phb_phase1_events: bad arguments");
        *n = (int)M->p1_events.size();
        if (ev) std::copy(M->p1_events.begin(), M->p1_events.begin() + std::min(cap, *n), ev);
        return PHB_OK;
    });
}
