// handles.hpp -- definitions behind the opaque handles of the C ABI.
#pragma once

#include <vector>

#include "device.hpp"

struct phb_atoms_s {
    int n = 0;
    int max_type = 0;                 // largest type id uploaded (host-side bound check)
    phb::DevBuf<double> x, y, z, q;   // SoA, wrapped positions
    phb::DevBuf<int> type;            // 0-based
    phb::DevBuf<int> sp_slot, sp_list;   // spreading: each atom's slot in its tile, the tiles' atom lists
};

struct phb_csr_s {
    int rows = 0;
    long long nnz = 0;
    phb::DevBuf<int> head, list;
};

struct phb_params_s {
    phb_params_desc d{};
    std::vector<double> eps, sig, shift;   // host copies, T*T
    double lj_rc = 0.0;
    bool lj_cut_active = false;
    phb::DevBuf<double> d_eps, d_sig, d_shift;
};
