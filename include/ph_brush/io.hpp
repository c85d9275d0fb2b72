// io.hpp -- the reference's file-format API (MC/include/io.hpp) on top of the C
// ABI in phb.h: read_lammps_full, read_params, wrap_positions,
// write_lammps_full (build_bond_adjacency is in ph_brush.hpp).  Same names,
// argument meaning and errors (std::runtime_error with the reference's message).
#pragma once

#include <string>
#include <vector>

#include "ph_brush.hpp"

namespace ph_brush {

// read_lammps_full (MC/include/io.hpp:7)
inline System read_lammps_full(const std::string& path) {
    phb_data h = nullptr;
    phb_data_info info{};
    check(phb_data_read(path.c_str(), &h, &info));
    const int n = info.natoms, nb = info.nbond_rows;
    std::vector<int> id(n), mol(n), type(n), bid(nb), bty(nb), bi(nb), bj(nb);
    std::vector<double> q(n), x(n), y(n), z(n);
    const int rc = phb_data_get(h, id.data(), mol.data(), type.data(), q.data(), x.data(), y.data(), z.data(),
                                bid.data(), bty.data(), bi.data(), bj.data());
    phb_data_destroy(h);
    check(rc);
    System s;
    s.natoms = info.natoms;
    s.nbonds = info.nbonds;
    s.ntypes = info.ntypes;
    s.box = Box{info.box.Lx, info.box.Ly, info.box.Lz};
    s.atoms.resize(n);
    for (int i = 0; i < n; ++i) s.atoms[i] = Atom{id[i], mol[i], type[i], q[i], x[i], y[i], z[i]};
    s.bonds.resize(nb);
    for (int k = 0; k < nb; ++k) s.bonds[k] = Bond{bid[k], bty[k], bi[k], bj[k]};
    return s;
}

// read_params (MC/include/io.hpp:11); lj_cutoff_nm and pme_grid fill the
// extension fields lj_rc and nx, ny, nz.
inline Params read_params(const std::string& path, int ntypes) {
    phb_params_file f{};
    const size_t t2 = ntypes > 0 ? (size_t)ntypes * ntypes : 0;
    std::vector<double> eps(t2), sig(t2), shift(t2);
    check(phb_params_read(path.c_str(), ntypes, &f, eps.data(), sig.data(), shift.data()));
    Params p;
    p.rc = f.rc;
    p.ntypes = f.ntypes;
    p.lj_eps = eps;
    p.lj_sig = sig;
    p.lj_shift = f.lj_shift != 0;
    p.lj_shift_tbl = shift;
    p.lj_rc = f.lj_rc;
    p.coulombtype = f.coulombtype == PHB_COUL_PME ? "pme" : f.coulombtype == PHB_COUL_CUT ? "coul_cut" : "unknown";
    p.epsilon_r = f.epsilon_r;
    p.ewald_rtol = f.ewald_rtol;
    p.ewald_alpha = f.ewald_alpha;
    p.pme_spacing = f.pme_spacing;
    p.pme_order = f.pme_order;
    p.nx = f.nx; p.ny = f.ny; p.nz = f.nz;
    p.pme_3dc_enabled = f.pme_3dc != 0;
    p.pme_3dc_zfac = f.pme_3dc_zfac;
    p.NH_N_type = f.nh_n_type; p.NH_P_type = f.nh_p_type; p.W_type = f.w_type; p.Cl_type = f.cl_type;
    p.target_protonation_pct = f.target_protonation_pct;
    p.rng_seed = f.rng_seed;
    p.energy_interval = f.energy_interval;
    p.patience = f.patience;
    p.max_attempts = f.max_attempts;
    p.beta = f.beta;
    p.w_z_extra_nm = f.w_z_extra_nm;
    p.direct_pct_threshold = f.direct_pct_threshold;
    p.mu_eff = 0.0;
    p.mu_eta = f.mu_eta;
    p.gc_block_fixed = f.gc_block_fixed;
    p.sgmc_epoch_attempt_max = f.sgmc_epoch_attempt_max;
    p.sgmc_nacc_frac_round = f.sgmc_nacc_frac_round;
    p.phase2_wc = f.phase2_wc != 0;
    return p;
}

// wrap_positions (MC/include/io.hpp:14)
inline void wrap_positions(System& sys) {
    const int n = (int)sys.atoms.size();
    std::vector<double> x(n), y(n), z(n);
    for (int i = 0; i < n; ++i) { x[i] = sys.atoms[i].x; y[i] = sys.atoms[i].y; z[i] = sys.atoms[i].z; }
    phb_box b{sys.box.Lx, sys.box.Ly, sys.box.Lz};
    check(phb_wrap_positions(n, x.data(), y.data(), z.data(), &b));
    for (int i = 0; i < n; ++i) { sys.atoms[i].x = x[i]; sys.atoms[i].y = y[i]; sys.atoms[i].z = z[i]; }
}

// write_lammps_full (MC/include/io.hpp:17)
inline void write_lammps_full(const std::string& path, const System& sys, const Box& boxForOutput) {
    const int n = (int)sys.atoms.size(), nb = (int)sys.bonds.size();
    std::vector<int> id(n), mol(n), type(n), bid(nb), bty(nb), bi(nb), bj(nb);
    std::vector<double> q(n), x(n), y(n), z(n);
    for (int i = 0; i < n; ++i) {
        const Atom& a = sys.atoms[i];
        id[i] = a.id; mol[i] = a.mol; type[i] = a.type; q[i] = a.q; x[i] = a.x; y[i] = a.y; z[i] = a.z;
    }
    for (int k = 0; k < nb; ++k) {
        const Bond& b = sys.bonds[k];
        bid[k] = b.id; bty[k] = b.type; bi[k] = b.i; bj[k] = b.j;
    }
    phb_data_view v{sys.natoms, sys.nbonds, sys.ntypes, n, nb, id.data(), mol.data(), type.data(), q.data(),
                    x.data(), y.data(), z.data(), bid.data(), bty.data(), bi.data(), bj.data()};
    phb_box b{boxForOutput.Lx, boxForOutput.Ly, boxForOutput.Lz};
    check(phb_data_write(path.c_str(), &v, &b));
}

}  // namespace ph_brush
