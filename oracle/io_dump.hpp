// io_dump.hpp -- test infrastructure: prints what a file-format API made of a
// LAMMPS data file and a params file, in one canonical text form, so that two
// implementations of that API (the reference's own io.cpp, built into
// oracle/_ref, and this library's C++ layer) can be compared byte for byte.
// Steps, as the reference's main() takes them (MC/src/main.cu:102-108):
// read_lammps_full, ntypes from the largest type when the header has none,
// read_params, wrap_positions, write_lammps_full (box of the data file).
// An exception prints "error: <message>" and exits 2, after whatever the
// steps before it printed.
#pragma once

#include <algorithm>
#include <cstdio>
#include <exception>
#include <string>

template <class SystemT, class ParamsT, class Api>
int io_dump_main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s in.data params.in out.data\n", argv[0]);
        return 1;
    }
    try {
        SystemT sys = Api::read_data(argv[1]);
        std::printf("natoms %d\nnbonds %d\nntypes %d\n", sys.natoms, sys.nbonds, sys.ntypes);
        std::printf("box %.17g %.17g %.17g\n", (double)sys.box.Lx, (double)sys.box.Ly, (double)sys.box.Lz);
        for (const auto& a : sys.atoms)
            std::printf("atom %d %d %d %.17g %.17g %.17g %.17g\n", a.id, a.mol, a.type, (double)a.q, (double)a.x,
                        (double)a.y, (double)a.z);
        for (const auto& b : sys.bonds) std::printf("bond %d %d %d %d\n", b.id, b.type, b.i, b.j);
        if (sys.ntypes <= 0) {
            int tmax = 0;
            for (const auto& a : sys.atoms) tmax = std::max(tmax, a.type);
            sys.ntypes = tmax;
        }
        std::fflush(stdout);
        ParamsT p = Api::read_params(argv[2], sys.ntypes);
        std::printf("rc %.17g\nntypes %d\nlj_shift %d\ncoulombtype %s\n", (double)p.rc, p.ntypes, (int)p.lj_shift,
                    p.coulombtype.c_str());
        std::printf("epsilon_r %.17g\newald_rtol %.17g\newald_alpha %.17g\npme_spacing %.17g\npme_order %d\n",
                    (double)p.epsilon_r, (double)p.ewald_rtol, (double)p.ewald_alpha, (double)p.pme_spacing,
                    p.pme_order);
        std::printf("pme_3dc %d %.17g\n", (int)p.pme_3dc_enabled, (double)p.pme_3dc_zfac);
        std::printf("types %d %d %d %d\n", p.NH_N_type, p.NH_P_type, p.W_type, p.Cl_type);
        std::printf("target_protonation_pct %.17g\nrng_seed %d\nenergy_interval %d\npatience %d\nmax_attempts %d\n",
                    (double)p.target_protonation_pct, p.rng_seed, p.energy_interval, p.patience, p.max_attempts);
        std::printf("beta %.17g\nw_z_extra_nm %.17g\ndirect_pct_threshold %d\nmu_eff %.17g\nmu_eta %.17g\n",
                    (double)p.beta, (double)p.w_z_extra_nm, p.direct_pct_threshold, (double)p.mu_eff,
                    (double)p.mu_eta);
        std::printf("gc_block_fixed %d\nsgmc_epoch_attempt_max %d\nsgmc_nacc_frac_round %.17g\nphase2_wc %d\n",
                    p.gc_block_fixed, p.sgmc_epoch_attempt_max, (double)p.sgmc_nacc_frac_round, (int)p.phase2_wc);
        for (size_t k = 0; k < p.lj_eps.size(); ++k)
            std::printf("lj %zu %.17g %.17g %.17g\n", k, (double)p.lj_eps[k], (double)p.lj_sig[k],
                        k < p.lj_shift_tbl.size() ? (double)p.lj_shift_tbl[k] : 0.0);
        std::fflush(stdout);
        Api::wrap(sys);
        for (const auto& a : sys.atoms) std::printf("wrapped %.17g %.17g %.17g\n", (double)a.x, (double)a.y, (double)a.z);
        Api::write_data(argv[3], sys);
        std::printf("written\n");
    } catch (const std::exception& e) {
        std::fflush(stdout);
        std::printf("error: %s\n", e.what());
        return 2;
    }
    return 0;
}
