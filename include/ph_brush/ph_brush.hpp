// ph_brush.hpp -- header-only C++ layer restoring the reference's host API
// (MC/include/types.hpp, device_data.hpp, neighbor.hpp, energy.hpp, pme.hpp,
// mc.hpp) on top of the C ABI in phb.h.  Same names, argument meaning and
// error behaviour: functions return by value after a device sync and throw
// std::runtime_error carrying the library's message (the reference's CUDA
// checks throw the same type, e.g. MC/src/energy_gpu.cu:7-10).
//
// The move API keeps the reference's free functions with their exact
// signatures (MC/include/mc.hpp:50-81): build_plan, apply_direct_protonation,
// phase_gc_mu_control (mutating p.mu_eff and advancing the caller's
// std::mt19937) and phase2_swap_jitter, plus main's upload_atoms_to_device /
// upload_params_to_device (MC/src/main.cu:18-66).  Each move function runs its
// chain on the device and writes the identities back into the host System, as
// the reference's set_atom_identity keeps its host mirror.
//
// Differences a caller sees:
//   - device objects own opaque handles instead of thrust vectors;
//   - `real` is double (the FP64 plain path); the FP32 form is selected per
//     PME plan at run time;
//   - the move functions write no report.txt lines (the MC program writes them
//     from the phase-1 events and phase-2 records); result counters are 64-bit;
//   - the MCPlan given to phase_gc_mu_control must describe the system's current
//     identities (build_plan(sys, p) on the same system, as main does), since the
//     device chain builds the same pools from them; a different plan is refused;
//   - the MonteCarlo class keeps one device chain across phases (its engine,
//     potential grid and pools), for callers that run many blocks or a fixed-mu
//     titration (phase_gc_fixed_mu).
#pragma once

#include <algorithm>
#include <cmath>
#include <exception>
#include <functional>
#include <iomanip>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../phb.h"

namespace ph_brush {

using real = double;

inline void check(int rc) {
    if (rc != PHB_OK) throw std::runtime_error(phb_last_error());
}

// ---------------------------------------------------------------- host types
struct Box { real Lx{0}, Ly{0}, Lz{0}; };                                     // MC/include/types.hpp:6-8
struct Atom { int id{0}, mol{0}, type{0}; real q{0}, x{0}, y{0}, z{0}; };      // :10-16 (type 1-based)
struct Bond { int id{0}, type{0}, i{0}, j{0}; };                               // :18-23 (0-based atoms)
struct System { Box box; std::vector<Atom> atoms; std::vector<Bond> bonds; int natoms{0}, nbonds{0}, ntypes{0}; };
struct CSR { std::vector<int> head, list; };                                  // :71-74

// Params (MC/include/types.hpp:34-69) with the extension keys used by the
// device path: lj_rc, explicit grid, fixed mu.
struct Params {
    real rc{0.0};
    int ntypes{0};
    std::vector<real> lj_eps, lj_sig;
    bool lj_shift{true};
    real lj_rc{0.0};
    std::string coulombtype{"coul_cut"};
    real epsilon_r{1.0};
    real ewald_rtol{1e-5};
    real ewald_alpha{0.0};
    real pme_spacing{0.12};
    int pme_order{4};
    int nx{0}, ny{0}, nz{0};
    bool pme_3dc_enabled{false};
    real pme_3dc_zfac{3.0};
    int NH_N_type{4}, NH_P_type{5}, W_type{7}, Cl_type{8};
    int rng_seed{2025};
    real beta{1.0 / (0.008314462 * 300.0)};
    real w_z_extra_nm{1.0};
    real mu_eff{0.0};
    int patience{3000};            // phase 2 (MC/include/types.hpp:58-59)
    int max_attempts{1000000};
    real target_protonation_pct{10.0};   // phase 1 (MC/include/types.hpp:55-67)
    int direct_pct_threshold{90};
    real mu_eta{0.3};
    int gc_block_fixed{2000};
    int sgmc_epoch_attempt_max{100000};
    real sgmc_nacc_frac_round{0.10};
    int energy_interval{100};      // phase-2 report interval (MC/include/types.hpp:57)
    bool phase2_wc{true};          // run phase 2 (MC/include/types.hpp:68)
    std::vector<real> lj_shift_tbl;  // LJ(rc) per type pair when lj_shift (MC/include/types.hpp:40)
};

// build_bond_adjacency (MC/src/io.cpp:148-169)
inline CSR build_bond_adjacency(const System& sys) {
    std::vector<std::vector<int>> tmp(sys.natoms);
    for (const auto& b : sys.bonds) { tmp[b.i].push_back(b.j); tmp[b.j].push_back(b.i); }
    CSR out;
    out.head.assign(sys.natoms + 1, 0);
    for (int i = 0; i < sys.natoms; ++i) {
        auto& v = tmp[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        out.head[i + 1] = out.head[i] + (int)v.size();
    }
    for (auto& v : tmp) out.list.insert(out.list.end(), v.begin(), v.end());
    return out;
}

// ------------------------------------------------------------ device objects
template <class H, int (*Destroy)(H)>
class Handle {
public:
    Handle() = default;
    explicit Handle(H h) : h_(h) {}
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    Handle(Handle&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    Handle& operator=(Handle&& o) noexcept { if (this != &o) { reset(); h_ = o.h_; o.h_ = nullptr; } return *this; }
    ~Handle() { reset(); }
    H get() const { return h_; }
    void reset() { if (h_) Destroy(h_); h_ = nullptr; }
protected:
    H h_ = nullptr;
};

// DeviceAtoms (MC/include/device_data.hpp:5-12) + upload_atoms_to_device (MC/src/main.cu:18-41)
struct DeviceAtoms : Handle<phb_atoms, phb_atoms_destroy> {
    int natoms{0};
    DeviceAtoms() = default;
    explicit DeviceAtoms(const System& sys) : natoms(sys.natoms) {
        std::vector<double> x(natoms), y(natoms), z(natoms), q(natoms);
        std::vector<int> t(natoms);
        for (int i = 0; i < natoms; ++i) {
            const Atom& a = sys.atoms[i];
            x[i] = a.x; y[i] = a.y; z[i] = a.z; q[i] = a.q; t[i] = a.type - 1;
        }
        phb_atoms h = nullptr;
        check(phb_atoms_create(natoms, x.data(), y.data(), z.data(), q.data(), t.data(), &h));
        h_ = h;
    }
};

// DeviceCSR (MC/include/device_data.hpp:14-19)
struct DeviceCSR : Handle<phb_csr, phb_csr_destroy> {
    int rows{0};
    long long nnz{0};
    DeviceCSR() = default;
    explicit DeviceCSR(phb_csr h) : Handle(h) { check(phb_csr_shape(h, &rows, &nnz)); }
    CSR download() const {
        CSR c;
        c.head.resize(rows + 1);
        c.list.resize(nnz);
        check(phb_csr_download(h_, c.head.data(), c.list.data()));
        return c;
    }
};

// copy_csr_to_device (MC/include/neighbor.hpp:6)
inline DeviceCSR copy_csr_to_device(const CSR& host) {
    phb_csr h = nullptr;
    const int rows = host.head.empty() ? 0 : (int)host.head.size() - 1;
    check(phb_csr_create(rows, host.head.data(), host.list.data(), &h));
    return DeviceCSR(h);
}

// build_half_neighbors_gpu (MC/include/neighbor.hpp:10-13)
inline DeviceCSR build_half_neighbors_gpu(const DeviceAtoms& atoms, const Box& box, const DeviceCSR& bonds12_dev,
                                          real rc) {
    phb_box b{box.Lx, box.Ly, box.Lz};
    phb_csr h = nullptr;
    check(phb_nlist_build_half(atoms.get(), &b, bonds12_dev.get(), rc, &h));
    return DeviceCSR(h);
}

// DeviceParams (MC/include/device_data.hpp:21-36) + upload_params_to_device (MC/src/main.cu:43-66)
struct DeviceParams : Handle<phb_params, phb_params_destroy> {
    int ntypes{0};
    std::string coulombtype;
    bool pme_3dc_enabled{false};
    DeviceParams() = default;
    explicit DeviceParams(const Params& p) : ntypes(p.ntypes), coulombtype(p.coulombtype),
                                             pme_3dc_enabled(p.pme_3dc_enabled) {
        phb_params_desc d{};
        d.ntypes = p.ntypes;
        d.lj_eps = p.lj_eps.data();
        d.lj_sig = p.lj_sig.data();
        d.lj_shift = p.lj_shift ? 1 : 0;
        d.rc = p.rc;
        d.lj_rc = p.lj_rc;
        d.epsilon_r = p.epsilon_r;
        d.coulombtype = p.coulombtype == "pme" ? PHB_COUL_PME : PHB_COUL_CUT;
        d.ewald_alpha = p.ewald_alpha;
        d.ewald_rtol = p.ewald_rtol;
        d.pme_spacing = p.pme_spacing;
        d.pme_order = p.pme_order;
        d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
        d.pme_3dc = p.pme_3dc_enabled ? 1 : 0;
        if (p.coulombtype != "pme" && p.coulombtype != "coul_cut")
            throw std::runtime_error("total_energy: unsupported coulombtype");  // MC/src/mc.cu:57
        phb_params h = nullptr;
        check(phb_params_create(&d, &h));
        h_ = h;
    }
};

// upload_atoms_to_device / upload_params_to_device (static in MC/src/main.cu:18-66):
// SoA positions, charges and 0-based types; the parameter tables.
inline DeviceAtoms upload_atoms_to_device(const System& sys) {
    if (sys.natoms <= 0) return DeviceAtoms();
    return DeviceAtoms(sys);
}
inline DeviceParams upload_params_to_device(const Params& p) { return DeviceParams(p); }

// std::mt19937 <-> the library's engine state: the engine's stream form
// (624 words, then the index), which operator<< / operator>> define.
inline std::vector<unsigned> engine_words(const std::mt19937& rng) {
    std::stringstream ss;
    ss << rng;
    std::vector<unsigned> w(625);
    for (auto& x : w) ss >> x;
    return w;
}
inline void engine_from_words(std::mt19937& rng, const unsigned* w) {
    std::stringstream ss;
    for (int k = 0; k < 625; ++k) ss << w[k] << ' ';
    ss >> rng;
    if (!ss) throw std::runtime_error("engine state: not a std::mt19937 stream form");
}

// compute_lj_sr_energy_gpu / compute_coul_sr_energy_gpu (MC/include/energy.hpp:7-18)
inline real compute_lj_sr_energy_gpu(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh, const Box& box) {
    phb_box b{box.Lx, box.Ly, box.Lz};
    double e = 0;
    check(phb_energy_lj(a.get(), p.get(), neigh.get(), &b, &e));
    return e;
}
inline real compute_coul_sr_energy_gpu(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh,
                                       const Box& box) {
    phb_box b{box.Lx, box.Ly, box.Lz};
    double e = 0;
    check(phb_energy_coul_cut(a.get(), p.get(), neigh.get(), &b, &e));
    return e;
}

// PMEEnergyComponents / PMEPlan / pmeCreate / pmeDestroy / pmeEnergy (MC/include/pme.hpp:7-49)
struct PMEEnergyComponents { real real_space{0}, recip_space{0}, self_term{0}, qcorr_term{0}; };

struct PMEPlan {
    int nx{0}, ny{0}, nz{0};
    int order{4};
    real alpha{0.0};
    bool use3dc{false};
    phb_pme handle{nullptr};
};

inline void pmeCreate(PMEPlan* plan, const DeviceParams& params, const Box& box, real rc, int precision = PHB_FP64) {
    if (!plan) throw std::runtime_error("pmeCreate: null plan");  // MC/src/pme.cu:292
    phb_box b{box.Lx, box.Ly, box.Lz};
    check(phb_pme_create(params.get(), &b, rc, precision, &plan->handle));
    check(phb_pme_info(plan->handle, &plan->nx, &plan->ny, &plan->nz, &plan->order, &plan->alpha));
    plan->use3dc = params.pme_3dc_enabled;   // MC/src/pme.cu:296
}

inline void pmeDestroy(PMEPlan* plan) {
    if (!plan) return;
    if (plan->handle) phb_pme_destroy(plan->handle);
    plan->handle = nullptr;
}

inline real pmeEnergy(PMEPlan& plan, const DeviceAtoms& atoms, const DeviceCSR& neigh, const Box& box,
                      PMEEnergyComponents* components = nullptr) {
    phb_box b{box.Lx, box.Ly, box.Lz};
    phb_pme_components c{};
    double e = 0;
    check(phb_pme_energy(plan.handle, atoms.get(), neigh.get(), &b, &c, &e));
    if (components) *components = PMEEnergyComponents{c.real_space, c.recip_space, c.self_term, c.qcorr_term};
    return e;
}

// total_energy (static in MC/src/mc.cu:43-60)
inline real total_energy(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh, const Box& box,
                         PMEPlan* plan) {
    phb_box b{box.Lx, box.Ly, box.Lz};
    double e = 0;
    if (p.coulombtype == "pme" && !plan) throw std::runtime_error("total_energy: PME plan is null");
    check(phb_total_energy(a.get(), p.get(), neigh.get(), plan ? plan->handle : nullptr, &b, nullptr, &e));
    return e;
}

// ------------------------------------------------------------ Monte Carlo
// PhaseGCResult (MC/include/mc.hpp:36-48); also the summary of a fixed-mu run.
struct PhaseGCResult {
    long long attempted{0}, accepted{0}, accepted_forward{0}, accepted_backward{0};
    int blocks_done{0};
    int Nh{0}, Nprot{0};
    int target_count{0};
    bool reached_target{false};
    real E{0.0};
    real mu_eff{0.0};
};

// Phase1Result (MC/include/mc.hpp:19-27): the direct flips and the pools phase 1
// hands to phase 2.
struct Phase1Result {
    long long attempted{0}, accepted{0};
    real E{0.0};
    std::vector<int> nh_n_remaining, w_remaining, nh_p_mc, cl_mc;
};

// Phase2Result (MC/include/mc.hpp:29-34).
struct Phase2Result {
    long long attempted{0}, accepted{0};
    real E{0.0};
    real best_E{0.0};
};

// Device chain at fixed mu: the move rule of phase_gc_mu_control
// (MC/src/mc.cu:379-503) with std::mt19937(p.rng_seed) semantics.
class MonteCarlo : public Handle<phb_mc, phb_mc_destroy> {
public:
    MonteCarlo(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh, PMEPlan& plan, const Box& box,
               const Params& hp, int pending_max = 8, int spec_window = 8, int recompute_interval = 1000)
        : MonteCarlo(a, p, neigh, &plan, box, hp, pending_max, spec_window, recompute_interval) {}
    // plan may be null for coulombtype coul_cut (total_energy's other branch, MC/src/mc.cu:52-56)
    MonteCarlo(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh, PMEPlan* plan, const Box& box,
               const Params& hp, int pending_max = 8, int spec_window = 8, int recompute_interval = 1000) {
        if (p.coulombtype == "pme" && (!plan || !plan->handle)) throw std::runtime_error("total_energy: PME plan is null");
        phb_box b{box.Lx, box.Ly, box.Lz};
        phb_mc_desc d{};
        d.nh_n_type = hp.NH_N_type; d.nh_p_type = hp.NH_P_type; d.w_type = hp.W_type; d.cl_type = hp.Cl_type;
        d.beta = hp.beta; d.mu_eff = hp.mu_eff; d.w_z_extra = hp.w_z_extra_nm;
        d.precision = PHB_FP64; d.pending_max = pending_max; d.recompute_interval = recompute_interval;
        d.record_cap = 0; d.spec_window = spec_window;
        mu_ = hp.mu_eff;
        phb_mc h = nullptr;
        check(phb_mc_create(a.get(), p.get(), neigh.get(), plan ? plan->handle : nullptr, &b, &d,
                            (unsigned)hp.rng_seed, &h));
        h_ = h;
    }
    // The chain continues from (set_engine) or hands back (engine_to) a caller's engine.
    void set_engine(const std::mt19937& rng) {
        const std::vector<unsigned> w = engine_words(rng);
        check(phb_mc_set_rng_state(h_, w.data()));
    }
    void engine_to(std::mt19937& rng) const {
        unsigned w[625];
        check(phb_mc_rng_state(h_, w));
        engine_from_words(rng, w);
    }
    // Pools phase 2 starts from (a Phase1Result's lists, MC/src/mc.cu:172-184).
    void set_phase2_pools(const Phase1Result& ph1) {
        const int sz[4] = {(int)ph1.nh_n_remaining.size(), (int)ph1.nh_p_mc.size(), (int)ph1.w_remaining.size(),
                           (int)ph1.cl_mc.size()};
        std::vector<int> all;
        all.reserve((size_t)sz[0] + sz[1] + sz[2] + sz[3]);
        for (const auto* v : {&ph1.nh_n_remaining, &ph1.nh_p_mc, &ph1.w_remaining, &ph1.cl_mc})
            all.insert(all.end(), v->begin(), v->end());
        check(phb_phase2_set_pools(h_, sz, all.data()));
    }
    PhaseGCResult run(long long ntrials) {
        phb_mc_stats s{};
        check(phb_mc_run(h_, ntrials, &s));
        if (s.status == 1) throw std::runtime_error("[sgmc] no valid pairs to flip; stop.");  // MC/src/mc.cu:399
        PhaseGCResult r;
        r.attempted = s.attempted; r.accepted = s.accepted;
        r.accepted_forward = s.accepted_forward; r.accepted_backward = s.accepted_backward;
        r.Nh = s.nsites; r.Nprot = s.nprot; r.E = s.E; r.mu_eff = mu_;
        return r;
    }
    // Phase 1 as main() runs it (MC/src/main.cu:180-197): build_plan's checks,
    // then apply_direct_protonation when target_protonation_pct exceeds
    // direct_pct_threshold, else phase_gc_mu_control on this chain (mu from the
    // target, epochs of gc_block_fixed trials, stop at the target).  max_trials
    // bounds the run (0: none, as in the reference).  direct_pct_threshold < 0
    // forces the direct path, > 100 the mu-control loop.
    // on_event, when given, sees each phase-1 event as it happens (phase1_events()
    // lists them afterwards); an exception it throws is rethrown once the run returns.
    PhaseGCResult phase1(const Params& hp, long long max_trials = 0, bool* direct = nullptr,
                         const std::function<void(const phb_phase1_event&)>& on_event = {}) {
        phb_phase1_desc d{};
        d.target_pct = hp.target_protonation_pct; d.direct_pct_threshold = hp.direct_pct_threshold;
        d.gc_block = hp.gc_block_fixed; d.epoch_attempt_max = hp.sgmc_epoch_attempt_max;
        d.nacc_frac = hp.sgmc_nacc_frac_round; d.mu_eta = hp.mu_eta; d.max_trials = max_trials;
        struct Observer {
            const std::function<void(const phb_phase1_event&)>* f;
            std::exception_ptr error;
        } obs{&on_event, nullptr};
        if (on_event) {
            d.on_event = [](const phb_phase1_event* e, void* u) {
                auto* o = static_cast<Observer*>(u);
                if (o->error) return;
                try { (*o->f)(*e); } catch (...) { o->error = std::current_exception(); }
            };
            d.user = &obs;
        }
        phb_phase1_result s{};
        check(phb_phase1_run(h_, &d, &s));
        if (obs.error) std::rethrow_exception(obs.error);
        // status 1 (no valid pair to flip) ends the run and returns, as MC/src/mc.cu:397-401 does
        PhaseGCResult r;
        r.attempted = s.attempted; r.accepted = s.accepted;
        r.accepted_forward = s.accepted_forward; r.accepted_backward = s.accepted_backward;
        r.blocks_done = s.blocks_done; r.Nh = s.Nh; r.Nprot = s.Nprot; r.target_count = s.target_count;
        r.reached_target = s.reached_target != 0; r.E = s.E; r.mu_eff = s.mu_eff;
        mu_ = s.mu_eff;
        if (direct) *direct = s.direct != 0;
        return r;
    }
    // The pools the last phase-1 run handed over (empty when it did not reach its target).
    Phase1Result handover() const {
        int sz[4] = {0, 0, 0, 0};
        check(phb_phase1_handover(h_, sz, nullptr, 0));
        std::vector<int> all((size_t)sz[0] + sz[1] + sz[2] + sz[3]);
        check(phb_phase1_handover(h_, sz, all.data(), (long long)all.size()));
        Phase1Result r;
        auto it = all.begin();
        r.nh_n_remaining.assign(it, it + sz[0]); it += sz[0];
        r.nh_p_mc.assign(it, it + sz[1]); it += sz[1];
        r.w_remaining.assign(it, it + sz[2]); it += sz[2];
        r.cl_mc.assign(it, it + sz[3]);
        return r;
    }
    // What the last phase1() did, block by block (the events its reference
    // counterpart reports: direct flips, chunks, mu updates, target reached).
    std::vector<phb_phase1_event> phase1_events() const {
        int n = 0;
        check(phb_phase1_events(h_, nullptr, 0, &n));
        std::vector<phb_phase1_event> ev(n);
        check(phb_phase1_events(h_, ev.data(), n, &n));
        return ev;
    }
    // phase2_swap_jitter's loop (MC/src/mc.cu:159-315) continuing this chain's
    // engine: pools from the phase-1 hand-over (else the current identities),
    // patience and max_attempts from hp.  With a trace, the first record_cap
    // trials' outcomes are returned too (for per-interval reporting).
    struct Phase2Trace {
        int mode{0};                    // 4: 4-point, 2: W-Cl only, 0: no valid proposals
        int stop{0};                    // 1: patience reached, 2: no valid proposals
        int since_best{0};
        real E0{0.0};                   // energy phase 2 started from
        std::vector<int> atoms;         // 4 per trial (W-Cl only: jCl, jW, -1, -1)
        std::vector<signed char> acc;
        std::vector<double> dE;
    };
    // With a trace, the device keeps the records of record_cap trials at a time and
    // phase 2 runs window by window, so every trial of the run is traced.
    Phase2Result phase2(const Params& hp, Phase2Trace* trace = nullptr, long long record_cap = 1LL << 22) {
        phb_phase2_desc d{};
        d.patience = hp.patience; d.best_eps = 1e-9;
        const long long window = trace ? std::max(1LL, std::min<long long>(hp.max_attempts, record_cap)) : 0;
        d.record_cap = window;
        phb_phase2_stats s{};
        check(phb_phase2_begin(h_, &d, &s));
        const double E0 = s.E;
        const int mode = s.mode;
        if (trace) {
            trace->atoms.clear();
            trace->acc.clear();
            trace->dE.clear();
        }
        if (s.mode != 0) {
            if (!trace) {
                check(phb_phase2_run(h_, hp.max_attempts, &s));
            } else {
                long long left = hp.max_attempts;
                std::vector<int> at;
                std::vector<signed char> ac;
                std::vector<double> de;
                while (left > 0) {
                    const long long first = s.attempted;
                    check(phb_phase2_record_from(h_, first));
                    check(phb_phase2_run(h_, std::min(left, window), &s));
                    const long long n = s.attempted - first;
                    at.assign(4 * (size_t)n, -1);
                    ac.assign((size_t)n, 0);
                    de.assign((size_t)n, 0.0);
                    check(phb_phase2_record(h_, n, at.data(), ac.data(), de.data()));
                    trace->atoms.insert(trace->atoms.end(), at.begin(), at.end());
                    trace->acc.insert(trace->acc.end(), ac.begin(), ac.end());
                    trace->dE.insert(trace->dE.end(), de.begin(), de.end());
                    if (s.stop != 0 || n == 0) break;
                    left -= n;
                }
            }
        }
        Phase2Result r;
        r.attempted = s.attempted; r.accepted = s.accepted; r.E = s.E; r.best_E = s.best_E;
        if (trace) {
            trace->mode = mode;
            trace->stop = s.stop;
            trace->since_best = s.since_best;
            trace->E0 = E0;
        }
        return r;
    }
private:
    double mu_ = 0.0;
};

// MCPlan + build_plan (MC/include/mc.hpp:11-17, MC/src/mc.cu:67-122): the site
// and W pools in atom order (W below z_cut = max site z + w_z_extra_nm), the
// flips to the target, and the counts its summary reports.  Same errors.
struct MCPlan {
    std::vector<int> idx_nh_n, idx_nh_p, idx_w;
    int to_flip{0};
    real z_cut{0.0};
    int nh_total{0}, cl_total{0};
    real z_max_site{0.0};
};

inline MCPlan build_plan(const System& sys, const Params& p) {
    MCPlan out;
    for (int i = 0; i < sys.natoms; ++i) {
        const int t = sys.atoms[i].type;
        if (t == p.NH_N_type) out.idx_nh_n.push_back(i);
        else if (t == p.NH_P_type) out.idx_nh_p.push_back(i);
        else if (t == p.W_type) out.idx_w.push_back(i);
        else if (t == p.Cl_type) ++out.cl_total;
    }
    out.nh_total = (int)(out.idx_nh_n.size() + out.idx_nh_p.size());
    if (out.nh_total == 0) throw std::runtime_error("[plan] No NH sites found.");
    const int target_count = (int)std::floor(p.target_protonation_pct / 100.0 * out.nh_total);
    out.to_flip = target_count - (int)out.idx_nh_p.size();
    if (out.to_flip < 0) throw std::runtime_error("[plan] Target < current; de-protonation not supported.");
    if ((int)out.idx_nh_n.size() < out.to_flip) throw std::runtime_error("[plan] Not enough NH_N to flip.");
    for (int i : out.idx_nh_n) out.z_max_site = std::max(out.z_max_site, sys.atoms[i].z);
    for (int i : out.idx_nh_p) out.z_max_site = std::max(out.z_max_site, sys.atoms[i].z);
    out.z_cut = out.z_max_site + p.w_z_extra_nm;
    std::vector<int> below;
    for (int j : out.idx_w) if (sys.atoms[j].z < out.z_cut) below.push_back(j);
    out.idx_w.swap(below);
    if ((int)out.idx_w.size() < out.to_flip) {
        std::ostringstream m;
        m << "[plan] Not enough W below z<" << std::setprecision(3) << out.z_cut << " nm: need " << out.to_flip
          << ", have " << out.idx_w.size();
        throw std::runtime_error(m.str());
    }
    return out;
}

// The device identities (type, charge) copied back into the host System, as
// the reference's set_atom_identity keeps its host mirror (MC/src/mc.cu:19-31).
inline void download_identities(const DeviceAtoms& a, System& sys) {
    std::vector<double> q(sys.natoms);
    std::vector<int> t0(sys.natoms);
    check(phb_atoms_get_identity(a.get(), q.data(), t0.data()));
    for (int i = 0; i < sys.natoms; ++i) { sys.atoms[i].q = q[i]; sys.atoms[i].type = t0[i] + 1; }
}

inline PhaseGCResult phase_gc_fixed_mu(const DeviceAtoms& a, const DeviceParams& p, const DeviceCSR& neigh,
                                       PMEPlan& plan, const Box& box, const Params& hp, long long ntrials) {
    MonteCarlo mc(a, p, neigh, plan, box, hp);
    return mc.run(ntrials);
}

// apply_direct_protonation (MC/src/mc.cu:124-157): the first to_flip NH_N and W
// of the plan flipped at once, whatever the threshold.
inline Phase1Result apply_direct_protonation(MonteCarlo& mc, const Params& hp) {
    Params q = hp;
    q.direct_pct_threshold = -1;
    PhaseGCResult g = mc.phase1(q);
    Phase1Result r = mc.handover();
    r.attempted = g.attempted; r.accepted = g.accepted; r.E = g.E;
    return r;
}

// phase_gc_mu_control (MC/src/mc.cu:320-546) on the device chain, whatever the
// threshold; the hand-over for phase 2 stays on mc (and in *seed_out if given).
inline PhaseGCResult phase_gc_mu_control(MonteCarlo& mc, const Params& hp, Phase1Result* seed_out = nullptr,
                                         long long max_trials = 0) {
    Params q = hp;
    q.direct_pct_threshold = 1000;
    PhaseGCResult g = mc.phase1(q, max_trials);
    if (seed_out) *seed_out = mc.handover();
    return g;
}

inline Phase2Result phase2_swap_jitter(MonteCarlo& mc, const Params& hp) { return mc.phase2(hp); }

// ---------------------------------------------------------------------------
// The reference's move functions with their exact signatures (MC/include/mc.hpp:50-81).

namespace detail {
// set_atom_identity's checks (MC/src/mc.cu:33-35)
inline void check_identity(const System& sys, int i, int type1, int ntypes) {
    if (i < 0 || i >= sys.natoms) throw std::runtime_error("set_atom_identity: atom index out of range");
    if (type1 < 1 || type1 > ntypes) throw std::runtime_error("set_atom_identity: type out of range");
}
// The pools the device chain builds from sys (MC/src/mc.cu:67-100, :332-341) against the caller's plan.
inline void check_plan_matches(const System& sys, const Params& p, const MCPlan& mcplan, const char* who) {
    std::vector<int> nh_n, nh_p, w;
    real z_max = 0.0;
    for (int i = 0; i < sys.natoms; ++i) {
        const int t = sys.atoms[i].type;
        if (t == p.NH_N_type) nh_n.push_back(i);
        else if (t == p.NH_P_type) nh_p.push_back(i);
    }
    for (int i : nh_n) z_max = std::max(z_max, sys.atoms[i].z);
    for (int i : nh_p) z_max = std::max(z_max, sys.atoms[i].z);
    const real z_cut = z_max + p.w_z_extra_nm;
    for (int i = 0; i < sys.natoms; ++i)
        if (sys.atoms[i].type == p.W_type && sys.atoms[i].z < z_cut) w.push_back(i);
    if (nh_n != mcplan.idx_nh_n || nh_p != mcplan.idx_nh_p || w != mcplan.idx_w || z_cut != mcplan.z_cut)
        throw std::runtime_error(std::string(who) +
                                 ": mcplan does not describe the system's identities (use build_plan(sys, p))");
}
}  // namespace detail

// apply_direct_protonation (MC/include/mc.hpp:52-59, MC/src/mc.cu:124-157): the
// first to_flip NH_N and W of the plan become NH_P (+1) and Cl (-1) in one
// batched upload; E is the total energy after the flips.
inline Phase1Result apply_direct_protonation(System& sys, DeviceAtoms& d_atoms, const DeviceParams& d_params,
                                             const DeviceCSR& d_neigh, const Box& box, PMEPlan* plan,
                                             const Params& p, const MCPlan& mcplan) {
    Phase1Result R;
    R.nh_n_remaining = mcplan.idx_nh_n;
    R.w_remaining = mcplan.idx_w;
    const int need = mcplan.to_flip;
    if (need <= 0) {
        R.E = total_energy(d_atoms, d_params, d_neigh, box, plan);
        return R;
    }
    if ((int)R.nh_n_remaining.size() < need || (int)R.w_remaining.size() < need)
        throw std::runtime_error("[direct] pools smaller than to_flip");
    std::vector<int> idx, ty0;
    std::vector<double> q;
    for (int k = 0; k < need; ++k) {
        const int i_nh = R.nh_n_remaining[k], j_w = R.w_remaining[k];
        detail::check_identity(sys, i_nh, p.NH_P_type, d_params.ntypes);
        detail::check_identity(sys, j_w, p.Cl_type, d_params.ntypes);
        idx.push_back(i_nh); ty0.push_back(p.NH_P_type - 1); q.push_back(+1.0);
        idx.push_back(j_w); ty0.push_back(p.Cl_type - 1); q.push_back(-1.0);
        R.nh_p_mc.push_back(i_nh);
        R.cl_mc.push_back(j_w);
    }
    check(phb_atoms_set_identity(d_atoms.get(), (int)idx.size(), idx.data(), ty0.data(), q.data()));
    for (size_t k = 0; k < idx.size(); ++k) { sys.atoms[idx[k]].type = ty0[k] + 1; sys.atoms[idx[k]].q = q[k]; }
    R.nh_n_remaining.erase(R.nh_n_remaining.begin(), R.nh_n_remaining.begin() + need);
    R.w_remaining.erase(R.w_remaining.begin(), R.w_remaining.begin() + need);
    R.attempted = need;
    R.accepted = need;
    R.E = total_energy(d_atoms, d_params, d_neigh, box, plan);
    return R;
}

// phase_gc_mu_control (MC/include/mc.hpp:72-81, MC/src/mc.cu:320-550): the
// mu-control GC loop on a device chain seeded from `rng`; afterwards rng has
// advanced exactly as the reference's engine, p.mu_eff holds the final mu, sys
// the final identities, and *seed_out the ascending hand-over (with E) when the
// target was reached by a move.
inline PhaseGCResult phase_gc_mu_control(System& sys, DeviceAtoms& d_atoms, const DeviceParams& d_params,
                                         const DeviceCSR& d_neigh, const Box& box, PMEPlan* plan, Params& p,
                                         const MCPlan& mcplan, std::mt19937& rng, Phase1Result* seed_out = nullptr) {
    detail::check_plan_matches(sys, p, mcplan, "phase_gc_mu_control");
    MonteCarlo mc(d_atoms, d_params, d_neigh, plan, box, p);
    mc.set_engine(rng);
    Params q = p;
    q.direct_pct_threshold = 1000;   // the loop itself, whatever the threshold
    const PhaseGCResult r = mc.phase1(q);
    mc.engine_to(rng);
    p.mu_eff = r.mu_eff;
    download_identities(d_atoms, sys);
    if (seed_out && r.reached_target && r.attempted > 0) {
        const Phase1Result h = mc.handover();
        seed_out->nh_p_mc = h.nh_p_mc;
        seed_out->nh_n_remaining = h.nh_n_remaining;
        seed_out->w_remaining = h.w_remaining;
        seed_out->cl_mc = h.cl_mc;
        seed_out->E = r.E;
    }
    return r;
}

// phase2_swap_jitter (MC/include/mc.hpp:61-70, MC/src/mc.cu:159-318): 4-point or
// W-Cl-only moves from ph1's pools (W and Cl below mcplan.z_cut), patience and
// max_attempts from p, continuing `rng`; sys gets the final identities.
inline Phase2Result phase2_swap_jitter(System& sys, DeviceAtoms& d_atoms, const DeviceParams& d_params,
                                       const DeviceCSR& d_neigh, const Box& box, PMEPlan* plan, const Params& p,
                                       const MCPlan& mcplan, const Phase1Result& ph1, std::mt19937& rng) {
    Phase1Result pools = ph1;
    auto filter_under_z = [&](std::vector<int>& pool) {
        pool.erase(std::remove_if(pool.begin(), pool.end(),
                                  [&](int idx) { return idx < 0 || idx >= sys.natoms || sys.atoms[idx].z >= mcplan.z_cut; }),
                   pool.end());
    };
    filter_under_z(pools.w_remaining);
    filter_under_z(pools.cl_mc);
    MonteCarlo mc(d_atoms, d_params, d_neigh, plan, box, p);
    mc.set_engine(rng);
    mc.set_phase2_pools(pools);
    const Phase2Result r = mc.phase2(p);
    mc.engine_to(rng);
    download_identities(d_atoms, sys);
    return r;
}

}  // namespace ph_brush
