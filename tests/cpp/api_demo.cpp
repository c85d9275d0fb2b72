// Exercises the reference-style C++ API (include/ph_brush/ph_brush.hpp) the
// way main() drives the reference (MC/src/main.cu:102-206): read a system,
// upload, build the neighbour list, create the PME plan, report energies, run a
// fixed-mu chain (or, with a target percentage, phase 1's mu-control loop with
// at most ntrials trials) and then phase 2.  Input: a text file written by the test
//   natoms nbonds Lx Ly Lz
//   x y z q type1          (natoms lines)
//   i j                    (nbonds lines, 0-based)
// Output: "key value" lines.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "ph_brush/ph_brush.hpp"

using namespace ph_brush;

int main(int argc, char** argv) {
    if (argc < 3) { std::cerr << "usage: api_demo system.txt ntrials [target_pct]\n"; return 1; }
    try {
        std::ifstream in(argv[1]);
        if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
        System sys;
        in >> sys.natoms >> sys.nbonds >> sys.box.Lx >> sys.box.Ly >> sys.box.Lz;
        sys.atoms.resize(sys.natoms);
        for (int i = 0; i < sys.natoms; ++i) {
            Atom& a = sys.atoms[i];
            a.id = i + 1;
            in >> a.x >> a.y >> a.z >> a.q >> a.type;
        }
        sys.bonds.resize(sys.nbonds);
        for (auto& b : sys.bonds) in >> b.i >> b.j;
        Params p;
        p.ntypes = 8;
        p.lj_eps.assign(64, 1.0);
        p.lj_sig.assign(64, 1.0);
        p.rc = 3.5; p.lj_rc = 1.122462048309373; p.lj_shift = true;
        p.coulombtype = "pme"; p.epsilon_r = 138.935458 / 3.0; p.ewald_alpha = 0.9; p.pme_order = 5;
        p.nx = 16; p.ny = 16; p.nz = 48; p.pme_3dc_enabled = true;
        p.beta = 1.0; p.mu_eff = 0.0; p.w_z_extra_nm = 1e6; p.rng_seed = 1;

        CSR bonds12 = build_bond_adjacency(sys);
        DeviceAtoms d_atoms(sys);
        DeviceCSR d_bonds = copy_csr_to_device(bonds12);
        DeviceCSR d_neigh = build_half_neighbors_gpu(d_atoms, sys.box, d_bonds, p.rc);
        DeviceParams d_params(p);
        PMEPlan plan;
        pmeCreate(&plan, d_params, sys.box, p.rc);
        PMEEnergyComponents comps;
        const real E_lj = compute_lj_sr_energy_gpu(d_atoms, d_params, d_neigh, sys.box);
        const real E_coul = pmeEnergy(plan, d_atoms, d_neigh, sys.box, &comps);
        std::printf("grid %d %d %d %d\n", plan.nx, plan.ny, plan.nz, plan.order);
        std::printf("nnz %lld\n", d_neigh.nnz);
        std::printf("E_lj %.17g\nE_coul %.17g\nE_real %.17g\nE_recip %.17g\nE_self %.17g\nE_qcorr %.17g\n", E_lj,
                    E_coul, comps.real_space, comps.recip_space, comps.self_term, comps.qcorr_term);
        MonteCarlo mc(d_atoms, d_params, d_neigh, plan, sys.box, p);
        PhaseGCResult r;
        if (argc > 3) {
            p.target_protonation_pct = std::atof(argv[3]);
            p.gc_block_fixed = 200;
            p.sgmc_epoch_attempt_max = 1000;
            Phase1Result ho;
            r = phase_gc_mu_control(mc, p, &ho, std::atoll(argv[2]));
            std::printf("blocks %d\nreached %d\nmu %.17g\nhandover %zu %zu %zu %zu\n", r.blocks_done,
                        (int)r.reached_target, r.mu_eff, ho.nh_n_remaining.size(), ho.nh_p_mc.size(),
                        ho.w_remaining.size(), ho.cl_mc.size());
        } else {
            r = mc.run(std::atoll(argv[2]));
        }
        std::printf("attempted %lld\naccepted %lld\nnprot %d\nE_after %.17g\n", r.attempted, r.accepted, r.Nprot, r.E);
        p.patience = 200;
        p.max_attempts = 400;
        // traced in windows of 97 device records (the MC program's per-interval lines read the trace)
        MonteCarlo::Phase2Trace trace;
        Phase2Result r2 = mc.phase2(p, &trace, 97);
        std::printf("p2_attempted %lld\np2_accepted %lld\np2_E %.17g\np2_best %.17g\n", r2.attempted, r2.accepted,
                    r2.E, r2.best_E);
        long long tacc = 0;
        double tde = 0.0;
        for (size_t k = 0; k < trace.acc.size(); ++k) if (trace.acc[k]) { ++tacc; tde += trace.dE[k]; }
        std::printf("trace_n %zu\ntrace_acc %lld\ntrace_E %.17g\n", trace.acc.size(), tacc, trace.E0 + tde);
        pmeDestroy(&plan);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";  // MC/src/main.cu:224-227
        return 2;
    }
    return 0;
}
