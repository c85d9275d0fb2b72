// io.cpp -- the reference's on-disk formats, host code: LAMMPS "full" data
// files (read and write), params.in, position wrapping and the Ewald splitting
// parameter from a real-space tolerance (MC/include/io.hpp, MC/src/io.cpp).
//
// A file the reference reads gives the same System / Params here, field by
// field and bit by bit, and a file written here is byte-identical to the one
// the reference writes from the same System: the line rules below are the
// format's, so they follow the reference's reader rather than LAMMPS's own.
// Numbers are converted by the C++ library routines the format is defined by
// (std::stoi / std::stod for tokens, stream extraction for header fields), and
// their exceptions surface with the library's message ("stoi", "stod").
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"

// Host arithmetic here must round exactly as written (wrapping, erfinv).
#pragma clang fp contract(off)

using phb::Error;
using phb::guard;
using phb::require;

struct phb_data_s {
    int natoms = 0, nbonds = 0, ntypes = 0;
    phb_box box{0, 0, 0};
    std::vector<int> id, mol, type;
    std::vector<double> q, x, y, z;
    std::vector<int> bond_id, bond_type, bond_i, bond_j;
};

namespace {

bool space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

std::string trimmed(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && space(s[a])) ++a;
    while (b > a && space(s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool begins(const std::string& s, const char* p) { return s.compare(0, std::char_traits<char>::length(p), p) == 0; }
bool holds(const std::string& s, const char* p) { return s.find(p) != std::string::npos; }

std::vector<std::string> tokens(const std::string& s) {
    std::vector<std::string> t;
    std::istringstream in(s);
    for (std::string w; in >> w;) t.push_back(w);
    return t;
}

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// Runs a parsing step, turning the library's conversion exceptions into the
// status the C ABI returns, with the library's own message.
template <class F>
auto converting(F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        throw Error(PHB_ERR_ARG, e.what());
    } catch (const std::out_of_range& e) {
        throw Error(PHB_ERR_ARG, e.what());
    }
}

// ------------------------------------------------------------ data files
// Header fields are read by stream extraction from the start of the line, so a
// line whose first word is not a number yields 0 (C++11 extraction), as in the
// reference (MC/src/io.cpp:87-103).
enum class HeaderField { none, atoms, bonds, atom_types, xbounds, ybounds, zbounds, body };

HeaderField classify_header(const std::string& s) {
    if (holds(s, " atoms") && !holds(s, "atom types")) return HeaderField::atoms;
    if (holds(s, " bonds") && !holds(s, "bond types")) return HeaderField::bonds;
    if (holds(s, " atom types")) return HeaderField::atom_types;
    if (holds(s, "xlo xhi")) return HeaderField::xbounds;
    if (holds(s, "ylo yhi")) return HeaderField::ybounds;
    if (holds(s, "zlo zhi")) return HeaderField::zbounds;
    if (begins(s, "Atoms")) return HeaderField::body;
    return HeaderField::none;
}

enum class Section { other, atoms, bonds };

// read_lammps_full (MC/src/io.cpp:79-146): header counts and box lengths from
// the lines before "Atoms", then the Atoms rows (>= 7 words: id mol type q x
// y z) and the Bonds rows (>= 4 words: id type atom-id atom-id; rows naming an
// unknown atom id are dropped).  Angles / Dihedrals / Impropers / Velocities
// end a section; every other line outside a section is skipped.
void read_data(const std::string& path, phb_data_s& D) {
    std::ifstream in(path);
    if (!in) throw Error(PHB_ERR_ARG, "Failed to open LAMMPS data: " + path);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::string line;
    while (std::getline(in, line)) {
        const std::string s = trimmed(line);
        if (s.empty()) continue;
        const HeaderField f = classify_header(s);
        if (f == HeaderField::body) break;
        std::istringstream v(s);
        switch (f) {
            case HeaderField::atoms: v >> D.natoms; break;
            case HeaderField::bonds: v >> D.nbonds; break;
            case HeaderField::atom_types: v >> D.ntypes; break;
            case HeaderField::xbounds: v >> lo[0] >> hi[0]; break;
            case HeaderField::ybounds: v >> lo[1] >> hi[1]; break;
            case HeaderField::zbounds: v >> lo[2] >> hi[2]; break;
            default: break;
        }
    }
    if (D.natoms <= 0) throw Error(PHB_ERR_ARG, "No atoms in header");
    D.box = phb_box{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};

    in.clear();
    in.seekg(0);
    std::unordered_map<int, int> index_of;
    index_of.reserve((size_t)D.natoms * 2);
    Section sec = Section::other;
    while (std::getline(in, line)) {
        const std::string s = trimmed(line);
        if (s.empty()) continue;
        if (begins(s, "Atoms")) { sec = Section::atoms; continue; }
        if (begins(s, "Bonds")) { sec = Section::bonds; continue; }
        if (begins(s, "Angles") || begins(s, "Dihedrals") || begins(s, "Impropers") || begins(s, "Velocities")) {
            sec = Section::other;
            continue;
        }
        if (sec == Section::atoms) {
            const auto t = tokens(s);
            if (t.size() < 7) continue;
            const int id = std::stoi(t[0]), mol = std::stoi(t[1]), ty = std::stoi(t[2]);
            const double q = std::stod(t[3]), x = std::stod(t[4]), y = std::stod(t[5]), z = std::stod(t[6]);
            index_of[id] = (int)D.id.size();
            D.id.push_back(id); D.mol.push_back(mol); D.type.push_back(ty);
            D.q.push_back(q); D.x.push_back(x); D.y.push_back(y); D.z.push_back(z);
        } else if (sec == Section::bonds) {
            const auto t = tokens(s);
            if (t.size() < 4) continue;
            const int id = std::stoi(t[0]), ty = std::stoi(t[1]);
            const int a = std::stoi(t[2]), b = std::stoi(t[3]);
            const auto ia = index_of.find(a), ib = index_of.find(b);
            if (ia == index_of.end() || ib == index_of.end()) continue;
            D.bond_id.push_back(id); D.bond_type.push_back(ty);
            D.bond_i.push_back(ia->second); D.bond_j.push_back(ib->second);
        }
    }
    if ((int)D.id.size() != D.natoms) throw Error(PHB_ERR_ARG, "Atom count mismatch in Atoms section");
}

// write_lammps_full (MC/src/io.cpp:326-368): fixed notation, 12 decimals for
// the box and 8 for charges and positions; the box starts at 0; the header
// counts are the System's; "N bond types" only when a bond type is positive.
void write_data(const std::string& path, const phb_data_view& v, const phb_box& box) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw Error(PHB_ERR_ARG, "Failed to open output data file: " + path);
    int bond_types = 0;
    for (int k = 0; k < v.nbond_rows; ++k) bond_types = std::max(bond_types, v.bond_type[k]);
    std::fprintf(f, "# Generated by lj-only MC\n\n%d atoms\n%d bonds\n\n%d atom types\n", v.natoms, v.nbonds,
                 v.ntypes);
    if (bond_types > 0) std::fprintf(f, "%d bond types\n", bond_types);
    std::fprintf(f, "\n%.12f %.12f xlo xhi\n%.12f %.12f ylo yhi\n%.12f %.12f zlo zhi\n\n", 0.0, box.Lx, 0.0, box.Ly,
                 0.0, box.Lz);
    std::fprintf(f, "Atoms # full\n\n");
    for (int i = 0; i < v.natoms_rows; ++i)
        std::fprintf(f, "%d %d %d %.8f %.8f %.8f %.8f\n", v.id[i], v.mol[i], v.type[i], v.q[i], v.x[i], v.y[i],
                     v.z[i]);
    std::fprintf(f, "\n");
    if (v.nbond_rows > 0) {
        std::fprintf(f, "Bonds\n\n");
        for (int k = 0; k < v.nbond_rows; ++k)
            std::fprintf(f, "%d %d %d %d\n", v.bond_id[k], v.bond_type[k], v.id[v.bond_i[k]], v.id[v.bond_j[k]]);
        std::fprintf(f, "\n");
    }
    const bool ok = std::ferror(f) == 0;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) throw Error(PHB_ERR_ARG, "Failed to write output data file: " + path);
}

// ------------------------------------------------------------ params.in
// Inverse error function by Winitzki's closed-form approximation (a = 0.147;
// S. Winitzki, "A handy approximation for the error function and its
// inverse", 2008) refined by three Newton steps on erf(v) - x.  The reference
// derives alpha the same way (MC/src/io.cpp:46-77), so the same rtol gives the
// same alpha to the last bit.
double erfinv_winitzki_newton(double x) {
    if (x <= -1.0) return -std::numeric_limits<double>::infinity();
    if (x >= 1.0) return std::numeric_limits<double>::infinity();
    const double a = 0.147;
    const double l = std::log(1.0 - x * x);
    const double t = 2.0 / (M_PI * a) + l / 2.0;
    double v = std::sqrt(std::max(0.0, std::sqrt(t * t - l / a) - t));
    if (x < 0) v = -v;
    const double two_over_sqrt_pi = 2.0 / std::sqrt(M_PI);
    for (int k = 0; k < 3; ++k) v -= (std::erf(v) - x) / (two_over_sqrt_pi * std::exp(-v * v));
    return v;
}

// alpha = erfc^-1(rtol) / rc with rtol <= 0 read as 1e-5 and clamped to
// [1e-12, 0.5] (MC/src/io.cpp:65-77).
double alpha_from_tolerance(double rc, double rtol) {
    if (rtol <= 0.0) rtol = 1e-5;
    const double y = std::clamp(rtol, 1e-12, 0.5);
    return erfinv_winitzki_newton(1.0 - y) / rc;
}

bool word_true(const std::string& v) { return v == "1" || v == "true" || v == "yes" || v == "on"; }

// phase2_wc's strict boolean (MC/src/io.cpp:24-40).
bool strict_bool(const std::string& raw) {
    const std::string s = trimmed(raw), t = lower(s);
    if (t == "1" || t == "true" || t == "yes" || t == "on") return true;
    if (t == "0" || t == "false" || t == "no" || t == "off") return false;
    throw Error(PHB_ERR_ARG, "invalid boolean: " + s);
}

struct ParamsOut {
    phb_params_file P{};
    std::vector<double> eps, sig, shift;
};

// Defaults of Params (MC/include/types.hpp:34-69) as read_params starts them
// (MC/src/io.cpp:172): no cutoff, empty LJ tables, shift on, coul_cut, eps_r 1.
void param_defaults(phb_params_file& P, int ntypes) {
    P = phb_params_file{};
    P.rc = 0.0;
    P.ntypes = ntypes;
    P.lj_shift = 1;
    P.coulombtype = PHB_COUL_CUT;
    P.epsilon_r = 1.0;
    P.ewald_rtol = 1e-5;
    P.ewald_alpha = 0.0;
    P.pme_spacing = 0.12;
    P.pme_order = 4;
    P.pme_3dc = 0;
    P.pme_3dc_zfac = 3.0;
    P.nh_n_type = 4; P.nh_p_type = 5; P.w_type = 7; P.cl_type = 8;
    P.target_protonation_pct = 10.0;
    P.rng_seed = 2025;
    P.energy_interval = 100;
    P.patience = 3000;
    P.max_attempts = 1000000;
    P.beta = 1.0 / (0.008314462 * 300.0);
    P.w_z_extra_nm = 1.0;
    P.direct_pct_threshold = 90;
    P.mu_eta = 0.3;
    P.gc_block_fixed = 2000;
    P.sgmc_epoch_attempt_max = 100000;
    P.sgmc_nacc_frac_round = 0.10;
    P.phase2_wc = 1;
    P.lj_rc = 0.0;
    P.nx = P.ny = P.nz = 0;
}

using Setter = std::function<void(phb_params_file&, const std::string&)>;

// key = value entries (MC/src/io.cpp:184-248), keys after '-' -> '_'; unknown
// keys are ignored.  lj_cutoff_nm and pme_grid are this library's extensions
// (a separate LJ cutoff; an explicit nx ny nz grid) and are absent from the
// reference's files, which therefore read the same here.
const std::unordered_map<std::string, Setter>& setters() {
    static const std::unordered_map<std::string, Setter> table = [] {
        std::unordered_map<std::string, Setter> m;
        auto real_key = [&](std::initializer_list<const char*> names, double phb_params_file::*field) {
            for (const char* n : names) m[n] = [field](phb_params_file& P, const std::string& v) { P.*field = std::stod(v); };
        };
        auto int_key = [&](std::initializer_list<const char*> names, int phb_params_file::*field) {
            for (const char* n : names) m[n] = [field](phb_params_file& P, const std::string& v) { P.*field = std::stoi(v); };
        };
        real_key({"cutoff_nm"}, &phb_params_file::rc);
        real_key({"epsilon_r"}, &phb_params_file::epsilon_r);
        real_key({"ewald_rtol"}, &phb_params_file::ewald_rtol);
        real_key({"ewald_alpha"}, &phb_params_file::ewald_alpha);
        real_key({"pme_spacing"}, &phb_params_file::pme_spacing);
        int_key({"pme_order"}, &phb_params_file::pme_order);
        int_key({"nh_n_type"}, &phb_params_file::nh_n_type);
        int_key({"nh_p_type"}, &phb_params_file::nh_p_type);
        int_key({"water_type", "w_type"}, &phb_params_file::w_type);
        int_key({"cl_type"}, &phb_params_file::cl_type);
        real_key({"target_pct", "target_protonation_pct"}, &phb_params_file::target_protonation_pct);
        int_key({"rng_seed"}, &phb_params_file::rng_seed);
        int_key({"energy_interval"}, &phb_params_file::energy_interval);
        int_key({"patience"}, &phb_params_file::patience);
        int_key({"max_attempts"}, &phb_params_file::max_attempts);
        real_key({"beta"}, &phb_params_file::beta);
        real_key({"w_z_extra_nm"}, &phb_params_file::w_z_extra_nm);
        int_key({"direct_pct_threshold", "direct_protonation_pct_threshold"}, &phb_params_file::direct_pct_threshold);
        real_key({"mu_eta"}, &phb_params_file::mu_eta);
        int_key({"gc_block_fixed"}, &phb_params_file::gc_block_fixed);
        int_key({"sgmc_epoch_attempt_max", "epoch_attempt_max"}, &phb_params_file::sgmc_epoch_attempt_max);
        real_key({"sgmc_nacc_frac_round", "nacc_frac_round"}, &phb_params_file::sgmc_nacc_frac_round);
        real_key({"lj_cutoff_nm"}, &phb_params_file::lj_rc);
        m["lj_shift"] = [](phb_params_file& P, const std::string& v) { P.lj_shift = word_true(lower(v)) ? 1 : 0; };
        m["coulombtype"] = [](phb_params_file& P, const std::string& v) {
            const std::string w = lower(v);
            P.coulombtype = (w == "coul_cut" || w == "coulcut") ? PHB_COUL_CUT : w == "pme" ? PHB_COUL_PME
                                                                                         : PHB_COUL_UNKNOWN;
        };
        // "pme_3dc = flag [zfac]": a missing flag reads as yes; a zfac that is
        // present but not a number reads as 0 (stream extraction), and any zfac
        // below 2 becomes 2 (MC/src/io.cpp:203-213).
        m["pme_3dc"] = [](phb_params_file& P, const std::string& v) {
            std::istringstream in(v);
            std::string flag;
            double zfac = P.pme_3dc_zfac;
            if (!(in >> flag)) flag = "yes";
            if (in.good()) in >> zfac;
            P.pme_3dc = word_true(lower(flag)) ? 1 : 0;
            P.pme_3dc_zfac = zfac < 2.0 ? 2.0 : zfac;
        };
        for (const char* n : {"phase2_wc", "phase2_WC"})
            m[n] = [](phb_params_file& P, const std::string& v) { P.phase2_wc = strict_bool(v) ? 1 : 0; };
        m["pme_grid"] = [](phb_params_file& P, const std::string& v) {
            std::istringstream in(v);
            int g[3] = {0, 0, 0};
            in >> g[0] >> g[1] >> g[2];
            require(g[0] > 0 && g[1] > 0 && g[2] > 0, "pme_grid needs three positive sizes");
            P.nx = g[0]; P.ny = g[1]; P.nz = g[2];
        };
        return m;
    }();
    return table;
}

// read_params (MC/src/io.cpp:171-301).
void read_params_file(const std::string& path, int ntypes, ParamsOut& out) {
    require(ntypes >= 0, "read_params: ntypes must not be negative");   // 0 fails the type checks below
    phb_params_file& P = out.P;
    param_defaults(P, ntypes);
    const size_t T2 = (size_t)ntypes * ntypes;
    out.eps.assign(T2, 0.0);
    out.sig.assign(T2, 0.0);
    std::ifstream in(path);
    if (!in) throw Error(PHB_ERR_ARG, "Failed to open params: " + path);
    const auto& table = setters();
    std::string line;
    while (std::getline(in, line)) {
        const std::string s = trimmed(line);
        if (s.empty() || s[0] == '#') continue;
        const size_t eq = s.find('=');
        std::string key;
        if (eq != std::string::npos) {
            key = trimmed(s.substr(0, eq));
            std::replace(key.begin(), key.end(), '-', '_');
        }
        if (!key.empty()) {
            const auto it = table.find(key);
            if (it != table.end()) it->second(P, trimmed(s.substr(eq + 1)));
        } else if (begins(s, "lj ")) {
            // "lj ti tj eps sig": symmetric entries of the 1-based type pair
            std::istringstream w(s);
            std::string tag;
            int ti = 0, tj = 0;
            double e = 0, g = 0;
            w >> tag >> ti >> tj >> e >> g;
            if (ti < 1 || tj < 1 || ti > ntypes || tj > ntypes)
                throw Error(PHB_ERR_ARG, "LJ type index out of bounds");
            const size_t a = (size_t)(ti - 1) * ntypes + (tj - 1), b = (size_t)(tj - 1) * ntypes + (ti - 1);
            out.eps[a] = out.eps[b] = e;
            out.sig[a] = out.sig[b] = g;
        }
    }
    if (P.rc <= 0.0) throw Error(PHB_ERR_ARG, "cutoff_nm must be > 0 in params");
    if (P.pme_spacing <= 0.0) P.pme_spacing = 0.12;
    if (P.pme_order <= 0) P.pme_order = 4;
    if (P.coulombtype == PHB_COUL_PME) {
        if (P.ewald_alpha <= 0.0) P.ewald_alpha = alpha_from_tolerance(P.rc, P.ewald_rtol);
    } else {
        P.ewald_alpha = 0.0;
    }
    // LJ(rc) per type pair for the shifted potential (MC/src/io.cpp:269-281)
    out.shift.assign(T2, 0.0);
    if (P.lj_shift) {
        for (size_t k = 0; k < T2; ++k) {
            const double e = out.eps[k], g = out.sig[k];
            if (e > 0.0 && g > 0.0) {
                const double sr = g / P.rc, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
                out.shift[k] = 4.0 * e * (sr12 - sr6);
            }
        }
    }
    if (P.direct_pct_threshold < 10 || P.direct_pct_threshold > 100)
        throw Error(PHB_ERR_ARG, "direct_pct_threshold must be between 10 and 100");
    if (P.gc_block_fixed <= 0) P.gc_block_fixed = 2000;
    if (P.sgmc_epoch_attempt_max <= 0) P.sgmc_epoch_attempt_max = 100000;
    if (P.sgmc_nacc_frac_round <= 0.0) P.sgmc_nacc_frac_round = 0.10;
    if (P.mu_eta <= 0.0) P.mu_eta = 0.3;
    const std::pair<const char*, int> types[4] = {
        {"nh_n_type", P.nh_n_type}, {"nh_p_type", P.nh_p_type}, {"w_type", P.w_type}, {"cl_type", P.cl_type}};
    for (const auto& t : types) {
        if (t.second < 1 || t.second > ntypes) {
            std::ostringstream m;
            m << t.first << " (" << t.second << ") must be between 1 and " << ntypes;
            throw Error(PHB_ERR_ARG, m.str());
        }
    }
}

// wrap_positions (MC/src/io.cpp:303-316): w = r - floor(r/L) L, then folded
// into [0, L) once from either side.
double wrap_one(double r, double L) {
    double w = r - std::floor(r / L) * L;
    if (w < 0) w += L;
    if (w >= L) w -= L;
    return w;
}

}  // namespace

extern "C" int phb_data_read(const char* path, phb_data* out, phb_data_info* info) {
    return guard([&] {
        require(path && out, "phb_data_read: null argument");
        *out = nullptr;
        auto D = std::make_unique<phb_data_s>();
        converting([&] { read_data(path, *D); });
        if (info) {
            *info = phb_data_info{D->natoms, D->nbonds, D->ntypes, (int)D->bond_id.size(), D->box};
        }
        *out = D.release();
        return PHB_OK;
    });
}

extern "C" int phb_data_get(phb_data d, int* id, int* mol, int* type, double* q, double* x, double* y, double* z,
                            int* bond_id, int* bond_type, int* bond_i, int* bond_j) {
    return guard([&] {
        require(d, "phb_data_get: null handle");
        auto put = [](const auto& v, auto* dst) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        put(d->id, id); put(d->mol, mol); put(d->type, type);
        put(d->q, q); put(d->x, x); put(d->y, y); put(d->z, z);
        put(d->bond_id, bond_id); put(d->bond_type, bond_type); put(d->bond_i, bond_i); put(d->bond_j, bond_j);
        return PHB_OK;
    });
}

extern "C" int phb_data_destroy(phb_data d) {
    delete d;
    return PHB_OK;
}

extern "C" int phb_data_write(const char* path, const phb_data_view* v, const phb_box* box) {
    return guard([&] {
        require(path && v && box, "phb_data_write: null argument");
        require(v->natoms_rows >= 0 && v->nbond_rows >= 0, "phb_data_write: negative row count");
        require(v->natoms_rows == 0 || (v->id && v->mol && v->type && v->q && v->x && v->y && v->z),
                "phb_data_write: null atom column");
        require(v->nbond_rows == 0 || (v->bond_id && v->bond_type && v->bond_i && v->bond_j),
                "phb_data_write: null bond column");
        for (int k = 0; k < v->nbond_rows; ++k)
            require(v->bond_i[k] >= 0 && v->bond_i[k] < v->natoms_rows && v->bond_j[k] >= 0 &&
                        v->bond_j[k] < v->natoms_rows,
                    "phb_data_write: bond atom index out of range");
        write_data(path, *v, *box);
        return PHB_OK;
    });
}

extern "C" int phb_params_read(const char* path, int ntypes, phb_params_file* out, double* lj_eps, double* lj_sig,
                               double* lj_shift_tbl) {
    return guard([&] {
        require(path && out, "phb_params_read: null argument");
        ParamsOut r;
        converting([&] { read_params_file(path, ntypes, r); });
        *out = r.P;
        if (lj_eps) std::copy(r.eps.begin(), r.eps.end(), lj_eps);
        if (lj_sig) std::copy(r.sig.begin(), r.sig.end(), lj_sig);
        if (lj_shift_tbl) std::copy(r.shift.begin(), r.shift.end(), lj_shift_tbl);
        return PHB_OK;
    });
}

extern "C" int phb_wrap_positions(int n, double* x, double* y, double* z, const phb_box* box) {
    return guard([&] {
        require(n >= 0 && box, "phb_wrap_positions: bad argument");
        require(n == 0 || (x && y && z), "phb_wrap_positions: null coordinate array");
        for (int i = 0; i < n; ++i) {
            x[i] = wrap_one(x[i], box->Lx);
            y[i] = wrap_one(y[i], box->Ly);
            z[i] = wrap_one(z[i], box->Lz);
        }
        return PHB_OK;
    });
}

extern "C" int phb_ewald_alpha_from_rtol(double rc, double rtol, double* alpha) {
    return guard([&] {
        require(alpha, "phb_ewald_alpha_from_rtol: null output");
        *alpha = alpha_from_tolerance(rc, rtol);
        return PHB_OK;
    });
}
