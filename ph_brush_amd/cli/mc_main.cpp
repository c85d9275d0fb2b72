// mc_main.cpp -- the reference's command line, `MC in.data params.in
// [--write-data out.data]` (MC/src/main.cu:68-229), on this library: read the
// LAMMPS data file and params.in, wrap positions, extend Lz for the slab
// correction, build the bond CSR and the half neighbour list, create the PME
// plan, report the energies, run phase 1 (direct flips or the mu-control GC
// loop to the target) and phase 2 (4-point / W-Cl moves to patience), writing
// phase1.data, phase2.data and report.txt in the working directory.
//
// report.txt carries the reference's lines (MC/src/main.cu:114-219,
// MC/src/mc.cu:108-120, :153-155, :186-311, :356-545) with its number
// formatting: one stream whose precision and fixed flag carry over from line to
// line exactly as in the reference.  The chains run on the GPU; the lines are
// written from what they report (phase-1 events, phase-2 trial records).
// Errors print "Error: <message>" and exit 2; usage errors exit 1.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ph_brush/io.hpp"
#include "ph_brush/ph_brush.hpp"

using namespace ph_brush;

namespace {

// report.txt, truncated at start and closed at exit (MC/src/report.cpp:12-41).
class Report {
public:
    explicit Report(const std::string& path) : out_(path, std::ios::out | std::ios::trunc) {
        if (!out_.is_open()) throw std::runtime_error("Failed to open report file: " + path);
    }
    std::ostream& operator()() { return out_; }
private:
    std::ofstream out_;
};

void plan_summary(std::ostream& r, const MCPlan& plan, const Params& p) {
    r << "\n--- Plan summary ---\n";
    r << "[NH] total=" << plan.nh_total << ", NH_N=" << plan.idx_nh_n.size() << ", NH_P=" << plan.idx_nh_p.size()
      << "\n";
    r << std::fixed << std::setprecision(3) << "[W/Cl] W_total(z-filtered)=" << plan.idx_w.size()
      << ", Cl_total=" << plan.cl_total << ", z_max_site=" << plan.z_max_site << ", z_cut=" << plan.z_cut
      << " nm\n";
    r << std::setprecision(2) << "[target] pct=" << p.target_protonation_pct << ", to_flip=" << plan.to_flip
      << " remaining=" << std::max(0, plan.to_flip) << "\n\n"
      << std::endl;
}

// The phase-1 line of one chain event (each ends with std::endl, as in the
// reference, so report.txt follows the run).
void phase1_line(std::ostream& r, const phb_phase1_event& e) {
    switch (e.kind) {
        case PHB_P1_DIRECT:
            r << std::setprecision(6) << "[direct] done: applied direct flips=" << e.attempted << " E=" << e.E
              << " remaining=" << 0 << std::endl;
            break;
        case PHB_P1_AT_TARGET:
            r << "[sgmc] already at target: Nprot=" << e.nprot << "/" << e.nh << " target=" << e.target_count
              << " mu/kT=" << e.mu_over_kT << " E=" << e.E << " remaining=0" << std::endl;
            break;
        case PHB_P1_CHUNK: {
            const double rate =
                e.block_attempted > 0 ? (double)e.block_accepted / (double)e.block_attempted : 0.0;
            r << std::setprecision(6) << "[sgmc] chunk acc=" << e.acc_since << "/" << e.nacc_target
              << " attempts=" << e.att_since << " acc_rate=" << rate << " fwd_total=" << e.fwd_total
              << " bwd_total=" << e.bwd_total << " gc_block=" << e.gc_block << " remaining=" << e.remaining
              << std::endl;
            break;
        }
        case PHB_P1_MU_UPDATE:
            r << std::setprecision(6) << "[sgmc] mu-update#" << (e.epoch + 1)
              << " cause=" << (e.early_stop ? "Amax" : "Nacc") << " mu/kT=" << e.mu_over_kT << " f=" << e.f
              << " acc=" << e.acc_since << "/" << e.nacc_target << " attempts=" << e.att_since << "/" << e.amax
              << " fwd_total=" << e.fwd_total << " bwd_total=" << e.bwd_total << " remaining=" << e.remaining
              << " R0=" << e.r0 << " w_info=" << e.w_info << std::endl;
            break;
        case PHB_P1_REACHED:
            r << std::setprecision(6) << "[sgmc] reached target in epoch=" << e.epoch
              << " step_in_block=" << e.step_in_block << " Nprot=" << e.nprot << "/" << e.nh
              << " mu/kT=" << e.mu_over_kT << " E=" << e.E << " fwd_acc=" << e.fwd_total
              << " bwd_acc=" << e.bwd_total << " remaining=0" << std::endl;
            break;
        case PHB_P1_NO_PAIRS:
            r << "[sgmc] no valid pairs to flip; stop." << std::endl;
            break;
        default:
            break;
    }
}

// The phase-2 lines (MC/src/mc.cu:186-311): pools, the stop reasons, and one
// line every energy_interval trials, from the recorded trial outcomes.
void phase2_lines(std::ostream& r, const MonteCarlo::Phase2Trace& t, const Params& p, const Phase1Result& ho,
                  const System& sys, real z_cut) {
    auto below = [&](const std::vector<int>& v) {
        long long n = 0;
        for (int i : v) n += (i >= 0 && i < sys.natoms && sys.atoms[i].z < z_cut) ? 1 : 0;
        return n;
    };
    const long long nh_p = (long long)ho.nh_p_mc.size(), nh_n = (long long)ho.nh_n_remaining.size();
    const long long cl = below(ho.cl_mc), w = below(ho.w_remaining);
    r << "[phase2] pools: NH_P=" << nh_p << " NH_N=" << nh_n << " Cl=" << cl << " W=" << w << " (z_cut=" << z_cut
      << ")\n";
    if (t.mode == 0) {
        // with NH_N empty the W-Cl move is requested; otherwise neither move exists (MC/src/mc.cu:193-207)
        if (nh_n == 0) r << "[phase2] wc-only requested but pools empty; Stop.\n";
        else r << "[phase2] no valid proposals (no 4-pt and no wc-only). Stop.\n";
        return;
    }
    const long long interval = std::max(1, p.energy_interval);
    const double best_eps = 1e-9;
    double E = t.E0, best = t.E0;
    long long accepted = 0, since_best = 0;
    const long long n = (long long)t.acc.size();
    for (long long k = 0; k < n; ++k) {
        if (t.acc[k]) {
            ++accepted;
            E += t.dE[k];
            if (E < best - best_eps) { best = E; since_best = 0; }
            else ++since_best;
        } else {
            ++since_best;
        }
        const long long step = k + 1;
        if (step % interval == 0) {
            const double rate = (double)accepted / (double)step;
            r << std::setprecision(6) << "[phase2] step=" << step << " E=" << E << " accepted=" << accepted
              << " acc=" << std::setprecision(3) << rate << " since_best=" << since_best << "/" << p.patience
              << (t.atoms[4 * k + 2] < 0 ? " mode=wc-only\n" : " mode=4-pt\n");
        }
        if (since_best >= p.patience) {
            r << "[phase2] patience reached: " << since_best << "/" << p.patience << ". Stop.\n";
            return;
        }
    }
    // a stop beyond the recorded trials
    if (t.stop == 1) r << "[phase2] patience reached: " << t.since_best << "/" << p.patience << ". Stop.\n";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "Usage: " << argv[0] << " in.data params.in [--write-data out.data]\n";
        return 1;
    }
    const std::string data_path = argv[1], params_path = argv[2];
    std::string write_data_path;
    for (int i = 3; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--write-data") {
            if (i + 1 >= argc) {
                std::cerr << "--write-data requires a path\n";
                return 1;
            }
            write_data_path = argv[++i];
        } else {
            std::cerr << "Unrecognized argument: " << arg << "\n";
            return 1;
        }
    }
    PMEPlan plan;
    try {
        Report report("report.txt");
        std::ostream& rep = report();
        System sys = read_lammps_full(data_path);
        if (sys.ntypes <= 0) {
            int tmax = 0;
            for (const auto& a : sys.atoms) tmax = std::max(tmax, a.type);
            sys.ntypes = tmax;
        }
        Params p = read_params(params_path, sys.ntypes);
        wrap_positions(sys);
        const Box write_box = sys.box;
        if (p.pme_3dc_enabled) {
            const real zfac = p.pme_3dc_zfac < 2.0 ? 2.0 : p.pme_3dc_zfac;
            const real old_lz = sys.box.Lz;
            sys.box.Lz = old_lz * zfac;
            rep << std::setprecision(3) << "[3dc] Enabling slab correction: zfac=" << zfac << ", Lz: " << old_lz
                << " -> " << sys.box.Lz << " (nm)\n";
        }
        const CSR bonds12 = build_bond_adjacency(sys);
        DeviceAtoms d_atoms(sys);
        const DeviceCSR d_bonds = copy_csr_to_device(bonds12);
        const DeviceCSR d_neigh = build_half_neighbors_gpu(d_atoms, sys.box, d_bonds, p.rc);
        const bool known_coulomb = p.coulombtype == "pme" || p.coulombtype == "coul_cut";
        Params p_dev = p;
        if (!known_coulomb) p_dev.coulombtype = "coul_cut";   // LJ is reported before the coulombtype error
        const DeviceParams d_params(p_dev);
        if (p.coulombtype == "pme") {
            pmeCreate(&plan, d_params, sys.box, p.rc);
            rep << "PME grid: " << plan.nx << " x " << plan.ny << " x " << plan.nz << " order=" << plan.order
                << " alpha=" << plan.alpha << " spacing≈" << sys.box.Lx / plan.nx << "\n";
        }
        rep.setf(std::ios::fixed);
        rep << "Atoms: " << sys.natoms << "\n";
        rep << std::setprecision(3) << "Box: Lx=" << sys.box.Lx << " Ly=" << sys.box.Ly << " Lz=" << sys.box.Lz
            << "\n";
        rep << std::setprecision(4) << "cutoff: " << p.rc << "\n";
        rep << "Neighbor pairs (half): " << d_neigh.nnz << "\n";
        rep << "coulombtype: " << p.coulombtype << " epsilon_r: " << p.epsilon_r << "\n";
        const real e_lj = compute_lj_sr_energy_gpu(d_atoms, d_params, d_neigh, sys.box);
        rep << std::setprecision(6) << "E_LJ (SR) = " << e_lj << "\n";
        real e_coul = 0.0;
        if (p.coulombtype == "coul_cut") {
            e_coul = compute_coul_sr_energy_gpu(d_atoms, d_params, d_neigh, sys.box);
            rep << std::setprecision(6) << "E_Coulomb (SR, cutoff) = " << e_coul << "\n";
        } else if (p.coulombtype == "pme") {
            PMEEnergyComponents comps;
            e_coul = pmeEnergy(plan, d_atoms, d_neigh, sys.box, &comps);
            rep << std::setprecision(6) << "E_Coulomb (PME) = " << e_coul << "  [real=" << comps.real_space
                << ", k=" << comps.recip_space << ", self=" << comps.self_term << ", qcorr=" << comps.qcorr_term
                << "]\n";
        } else {
            pmeDestroy(&plan);
            std::cerr << "Error: Unsupported coulombtype: " << p.coulombtype << "\n";
            return 2;
        }
        const real e_total = e_lj + e_coul;
        rep << std::setprecision(6) << "E_total = " << e_total << "\n";

        const MCPlan mcplan = build_plan(sys, p);
        plan_summary(rep, mcplan, p);
        MonteCarlo mc(d_atoms, d_params, d_neigh, plan, sys.box, p);
        const PhaseGCResult r1 = mc.phase1(p, 0, nullptr, [&rep](const phb_phase1_event& e) { phase1_line(rep, e); });
        const Phase1Result ho = mc.handover();
        download_identities(d_atoms, sys);
        write_lammps_full("phase1.data", sys, write_box);
        rep << "[write] phase1.data written\n";

        real e_phase2 = r1.E;
        if (p.phase2_wc) {
            MonteCarlo::Phase2Trace trace;
            const Phase2Result r2 = mc.phase2(p, &trace);
            phase2_lines(rep, trace, p, ho, sys, mcplan.z_cut);
            e_phase2 = r2.E;
            download_identities(d_atoms, sys);
        } else {
            rep << "[phase2] skipped (phase2_wc=false)\n";
        }
        write_lammps_full("phase2.data", sys, write_box);
        rep << "[write] phase2.data written\n";
        rep << std::setprecision(6) << "[summary] E0/after phase1/after phase2 = " << e_total << " / " << r1.E
            << " / " << e_phase2 << "\n";
        if (!write_data_path.empty()) {
            write_lammps_full(write_data_path, sys, write_box);
            rep << "[write] Wrote updated data file to " << write_data_path << "\n";
        }
        pmeDestroy(&plan);
    } catch (const std::exception& e) {
        pmeDestroy(&plan);
        std::cerr << "Error: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
