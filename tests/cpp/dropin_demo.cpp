// A caller written against the reference's host API (MC/include/*.hpp), typed
// as main() makes its calls (MC/src/main.cu:102-206): upload the system and
// parameters, build the neighbour list, create the PME plan, report energies,
// seed std::mt19937 from rng_seed, build_plan, then phase 1 (direct flips or
// phase_gc_mu_control with the engine and Params& it mutates) and phase 2 with
// the same engine.  Only the header and namespace differ from a reference
// caller.  Input: the text system of api_demo.cpp plus the target percentage.
// Output: "key value" lines (identities, engine states).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>

#include "ph_brush/ph_brush.hpp"

using namespace ph_brush;

static void print_engine(const char* key, const std::mt19937& rng) {
    const std::vector<unsigned> w = engine_words(rng);
    std::printf("%s", key);
    for (unsigned x : w) std::printf(" %u", x);
    std::printf("\n");
}

int main(int argc, char** argv) {
    if (argc < 4) { std::cerr << "usage: dropin_demo system.txt target_pct direct_threshold\n"; return 1; }
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
        sys.ntypes = 8;
        Params p;
        p.ntypes = 8;
        p.lj_eps.assign(64, 1.0);
        p.lj_sig.assign(64, 1.0);
        p.rc = 3.5; p.lj_rc = 1.122462048309373; p.lj_shift = true;
        p.coulombtype = "pme"; p.epsilon_r = 138.935458 / 3.0; p.ewald_alpha = 0.9; p.pme_order = 5;
        p.nx = 16; p.ny = 16; p.nz = 48; p.pme_3dc_enabled = true;
        p.beta = 1.0; p.w_z_extra_nm = 1e6; p.rng_seed = 5;
        p.target_protonation_pct = std::atof(argv[2]);
        p.direct_pct_threshold = std::atoi(argv[3]);
        p.gc_block_fixed = 200; p.sgmc_epoch_attempt_max = 1000;
        p.patience = 100000; p.max_attempts = 600;

        CSR bonds12 = build_bond_adjacency(sys);
        DeviceAtoms d_atoms = upload_atoms_to_device(sys);
        DeviceCSR d_bonds = copy_csr_to_device(bonds12);
        DeviceCSR d_neigh = build_half_neighbors_gpu(d_atoms, sys.box, d_bonds, p.rc);
        DeviceParams d_params = upload_params_to_device(p);
        PMEPlan plan;
        pmeCreate(&plan, d_params, sys.box, p.rc);
        std::printf("use3dc %d\n", (int)plan.use3dc);
        const real E_lj = compute_lj_sr_energy_gpu(d_atoms, d_params, d_neigh, sys.box);
        PMEEnergyComponents comps;
        const real E_coul = pmeEnergy(plan, d_atoms, d_neigh, sys.box, &comps);
        std::printf("E_total %.17g\n", E_lj + E_coul);
        PMEPlan* planPtr = &plan;

        std::mt19937 rng(static_cast<unsigned>(p.rng_seed));
        MCPlan mcplan = build_plan(sys, p);

        Phase1Result ph1;
        const bool use_direct_phase1 = p.target_protonation_pct > static_cast<real>(p.direct_pct_threshold);
        if (use_direct_phase1) {
            ph1 = apply_direct_protonation(sys, d_atoms, d_params, d_neigh, sys.box, planPtr, p, mcplan);
            std::printf("direct 1\n");
        } else {
            PhaseGCResult gc_result =
                phase_gc_mu_control(sys, d_atoms, d_params, d_neigh, sys.box, planPtr, p, mcplan, rng, &ph1);
            ph1.attempted = gc_result.attempted;
            ph1.accepted = gc_result.accepted;
            ph1.E = gc_result.E;
            std::printf("direct 0\nblocks %d\nreached %d\nfwd %lld\nbwd %lld\n", gc_result.blocks_done,
                        (int)gc_result.reached_target, gc_result.accepted_forward, gc_result.accepted_backward);
        }
        std::printf("p1_attempted %lld\np1_accepted %lld\np1_E %.17g\nmu_eff %.17g\n", ph1.attempted, ph1.accepted,
                    ph1.E, p.mu_eff);
        std::printf("handover %zu %zu %zu %zu\n", ph1.nh_n_remaining.size(), ph1.nh_p_mc.size(),
                    ph1.w_remaining.size(), ph1.cl_mc.size());
        std::printf("types1");
        for (const Atom& a : sys.atoms) std::printf(" %d", a.type);
        std::printf("\n");
        print_engine("engine1", rng);

        Phase2Result ph2 = phase2_swap_jitter(sys, d_atoms, d_params, d_neigh, sys.box, planPtr, p, mcplan, ph1, rng);
        std::printf("p2_attempted %lld\np2_accepted %lld\np2_E %.17g\np2_best %.17g\n", ph2.attempted, ph2.accepted,
                    ph2.E, ph2.best_E);
        std::printf("types2");
        for (const Atom& a : sys.atoms) std::printf(" %d", a.type);
        std::printf("\n");
        print_engine("engine2", rng);
        pmeDestroy(&plan);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";  // MC/src/main.cu:224-227
        return 2;
    }
    return 0;
}
