// gen_brush.cpp -- deterministic synthetic polyelectrolyte-brush generator.
//
// The reference's example inputs (LAMMPS "full" data files of PDMAEMA brushes)
// are not distributed, so workloads are generated: chains grafted on a square
// lattice grow as self-avoiding random walks, every titratable bead gets one
// counter-particle (Cl- when protonated, neutral placeholder W otherwise, which
// is what the reference's ion-swap move needs: MC/src/mc.cu:384-388), every
// permanently charged bead gets a Cl- counterion, and salt pairs are placed
// uniformly in the box.  Counter-particles sit in a shell of 1.0-2.5 sigma
// (widened to 4 and 6 sigma when crowded) around their own bead (local electroneutrality of a dense brush: positions
// are frozen, so a layer of beads with its counterions spread over the whole
// brush height would keep a capacitor-like field of hundreds of kT per sigma);
// a bead without room nearby gets its counter-particle anywhere in the brush
// layer.  Generator RNG: std::mt19937_64 seeded with spec->seed, draws via
// std::uniform_real_distribution<double>.
#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

struct CellGrid {
    double Lx, Ly, Lz, cell;
    int ncx, ncy, ncz;
    std::vector<std::vector<int>> cells;
    const std::vector<double>& X;
    const std::vector<double>& Y;
    const std::vector<double>& Z;
    CellGrid(double lx, double ly, double lz, double c, const std::vector<double>& x,
             const std::vector<double>& y, const std::vector<double>& z)
        : Lx(lx), Ly(ly), Lz(lz), X(x), Y(y), Z(z) {
        ncx = std::max(1, (int)std::floor(Lx / c));
        ncy = std::max(1, (int)std::floor(Ly / c));
        ncz = std::max(1, (int)std::floor(Lz / c));
        cell = c;
        cells.resize((size_t)ncx * ncy * ncz);
    }
    int cx(double v) const { int c = (int)std::floor(v / Lx * ncx); return std::min(std::max(c, 0), ncx - 1); }
    int cy(double v) const { int c = (int)std::floor(v / Ly * ncy); return std::min(std::max(c, 0), ncy - 1); }
    int cz(double v) const { int c = (int)std::floor(v / Lz * ncz); return std::min(std::max(c, 0), ncz - 1); }
    void insert(int idx) { cells[((size_t)cz(Z[idx]) * ncy + cy(Y[idx])) * ncx + cx(X[idx])].push_back(idx); }
    bool clear_of(double x, double y, double z, double dmin) const {
        const double d2 = dmin * dmin;
        int ix = cx(x), iy = cy(y), iz = cz(z);
        for (int dz = -1; dz <= 1; ++dz) {
            int kz = iz + dz;
            if (kz < 0 || kz >= ncz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int ky = ((iy + dy) % ncy + ncy) % ncy;
                for (int dx = -1; dx <= 1; ++dx) {
                    int kx = ((ix + dx) % ncx + ncx) % ncx;
                    for (int j : cells[((size_t)kz * ncy + ky) * ncx + kx]) {
                        double rx = x - X[j]; rx -= Lx * std::nearbyint(rx / Lx);
                        double ry = y - Y[j]; ry -= Ly * std::nearbyint(ry / Ly);
                        double rz = z - Z[j];
                        if (rx * rx + ry * ry + rz * rz < d2) return false;
                    }
                }
            }
        }
        return true;
    }
};

}  // namespace

extern "C" int phb_gen_brush(const phb_brush_spec* sp, int cap, double* x, double* y, double* z,
                             double* q, int* type1, int* mol, int* bond_i, int* bond_j, int bond_cap,
                             int* natoms, int* nbonds) {
    return phb::guard([&]() -> int {
        if (!sp || !natoms || !nbonds) throw phb::Error(PHB_ERR_ARG, "phb_gen_brush: null argument");
        if (sp->chains_x < 0 || sp->chains_y < 0 || sp->chain_len < 0 || sp->n_salt < 0 ||
            sp->n_perm < 0 || sp->n_perm > sp->chain_len)
            throw phb::Error(PHB_ERR_ARG, "phb_gen_brush: negative counts or n_perm > chain_len");
        if (!(sp->Lx > 0 && sp->Ly > 0 && sp->Lz > 0 && sp->min_dist > 0 && sp->bond > 0))
            throw phb::Error(PHB_ERR_ARG, "phb_gen_brush: box, bond and min_dist must be > 0");
        const int nch = sp->chains_x * sp->chains_y;
        const long long n_beads = (long long)nch * sp->chain_len;
        const long long n_titr = (long long)nch * (sp->chain_len - sp->n_perm);
        const long long n_perm = (long long)nch * sp->n_perm;
        const long long total = n_beads + n_perm + n_titr + 2LL * sp->n_salt;
        const long long total_bonds = (long long)nch * std::max(0, sp->chain_len - 1);
        *natoms = (int)total;
        *nbonds = (int)total_bonds;
        if (!x) return PHB_OK;  // size query
        if (total > cap || total_bonds > bond_cap)
            throw phb::Error(PHB_ERR_ARG, "phb_gen_brush: output capacity too small");

        std::mt19937_64 rng(sp->seed);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        // chain beads occupy ids [0, n_beads) in chain-major order; ions follow
        std::vector<double> X(total), Y(total), Z(total), Q(total);
        std::vector<int> T(total), M(total);
        CellGrid grid(sp->Lx, sp->Ly, sp->Lz, std::max(sp->min_dist, 1.0), X, Y, Z);
        auto cell_of = [&](int id) -> std::vector<int>& {
            return grid.cells[((size_t)grid.cz(Z[id]) * grid.ncy + grid.cy(Y[id])) * grid.ncx + grid.cx(X[id])];
        };
        auto remove_id = [&](int id) {
            auto& c = cell_of(id);
            c.erase(std::remove(c.begin(), c.end(), id), c.end());
        };
        const int len = sp->chain_len;
        // All chains grow together, one bead per chain per round, so that no
        // chain can crowd the graft point of a chain grown after it.
        for (int c = 0; c < nch; ++c) {
            const int id = c * len;
            if (len == 0) break;
            X[id] = sp->graft_spacing * (c % sp->chains_x + 0.5);
            Y[id] = sp->graft_spacing * (c / sp->chains_x + 0.5);
            X[id] -= sp->Lx * std::floor(X[id] / sp->Lx);
            Y[id] -= sp->Ly * std::floor(Y[id] / sp->Ly);
            Z[id] = sp->graft_z;
            if (!grid.clear_of(X[id], Y[id], Z[id], sp->min_dist))
                throw phb::Error(PHB_ERR_STATE, "phb_gen_brush: graft points closer than min_dist");
            grid.cells[((size_t)grid.cz(Z[id]) * grid.ncy + grid.cy(Y[id])) * grid.ncx + grid.cx(X[id])].push_back(id);
        }
        // Step directions: azimuth uniform, cos(theta) uniform on [0, 1] (a walk
        // stretched away from the grafting surface, as excluded volume and
        // charge stretch a dense brush); a trapped bead may also step sideways
        // or back (cos(theta) on [-0.5, 1]).
        auto try_place = [&](int id) {
            for (int tries = 0; tries < 400; ++tries) {
                double ct = tries < 200 ? U(rng) : 1.5 * U(rng) - 0.5, ph = 2.0 * M_PI * U(rng);
                double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
                double nx = X[id - 1] + sp->bond * st * std::cos(ph);
                double ny = Y[id - 1] + sp->bond * st * std::sin(ph);
                double nz = Z[id - 1] + sp->bond * ct;
                if (nz < 0.0 || nz >= sp->Lz) continue;
                nx -= sp->Lx * std::floor(nx / sp->Lx);
                ny -= sp->Ly * std::floor(ny / sp->Ly);
                if (nx >= sp->Lx) nx = 0.0;
                if (ny >= sp->Ly) ny = 0.0;
                if (!grid.clear_of(nx, ny, nz, sp->min_dist)) continue;
                X[id] = nx; Y[id] = ny; Z[id] = nz;
                cell_of(id).push_back(id);
                return true;
            }
            return false;
        };
        for (int b = 1; b < len; ++b)
            for (int c = 0; c < nch; ++c) {
                const int id0 = c * len;
                int placed_to = b - 1;   // beads id0..id0+placed_to are in the grid
                int backtracks = 0;
                while (placed_to < b) {
                    if (try_place(id0 + placed_to + 1)) { ++placed_to; continue; }
                    if (++backtracks > 2000)
                        throw phb::Error(PHB_ERR_STATE, "phb_gen_brush: chain " + std::to_string(c) +
                                                            " could not be grown past bead " + std::to_string(placed_to) +
                                                            " (box too dense)");
                    // step back, further the longer the chain stays trapped (never the graft bead)
                    const int back = std::min(placed_to, 1 + backtracks / 20);
                    for (int k = 0; k < back; ++k) remove_id(id0 + placed_to - k);
                    placed_to -= back;
                }
            }
        int nb = 0;
        std::vector<char> site_prot;  // protonation state of each titratable bead, in bead order
        std::vector<int> perm_ids, site_ids;
        for (int c = 0; c < nch; ++c)
            for (int b = 0; b < len; ++b) {
                const int id = c * len + b;
                M[id] = c + 1;
                if (b > 0) { bond_i[nb] = id - 1; bond_j[nb] = id; ++nb; }
                if (b < sp->n_perm) { Q[id] = 1.0; T[id] = sp->type_perm; perm_ids.push_back(id); }
                else {
                    bool prot = U(rng) < sp->p_prot;
                    site_prot.push_back(prot ? 1 : 0);
                    site_ids.push_back(id);
                    Q[id] = prot ? 1.0 : 0.0;
                    T[id] = prot ? sp->type_nh_p : sp->type_nh_n;
                }
            }
        size_t next = (size_t)n_beads;
        auto push = [&](double px, double py, double pz, double pq, int pt, int pm) {
            const int id = (int)next++;
            X[id] = px; Y[id] = py; Z[id] = pz; Q[id] = pq; T[id] = pt; M[id] = pm;
            cell_of(id).push_back(id);
        };
        // Fallback region for counter-particles: the brush layer z < (highest bead + 2);
        // salt fills the whole data box.
        double z_top = 0.0;
        for (long long i = 0; i < n_beads; ++i) z_top = std::max(z_top, Z[i]);
        const double z_layer = std::min(sp->Lz, z_top + 2.0);
        auto place_free = [&](double qv, int tv, int mv, double zmax) {
            for (int tries = 0; tries < 100000; ++tries) {
                double px = U(rng) * sp->Lx, py = U(rng) * sp->Ly, pz = U(rng) * zmax;
                if (grid.clear_of(px, py, pz, sp->min_dist)) { push(px, py, pz, qv, tv, mv); return; }
            }
            throw phb::Error(PHB_ERR_STATE, "phb_gen_brush: no free site for an ion (box too dense)");
        };
        // Shell of radius [1.0, 2.5] sigma around bead `id`, isotropic, above the
        // grafting plane.  Among up to 8 free candidates the one that keeps the
        // running bead -> counter-particle dipole sum_k (z_k - z_bead_k) nearest
        // zero is taken: rejection in the dense interior would otherwise push
        // counter-particles outward and leave a net dipole layer.
        double dip = 0.0;
        auto place_near = [&](int id, double qv, int tv, int mv) {
            double best[3] = {0, 0, 0}, best_cost = 0.0;
            int found = 0;
            for (int tries = 0; tries < 3000 && found < 8; ++tries) {
                const double rmax = tries < 1000 ? 2.5 : (tries < 2000 ? 4.0 : 6.0);   // widen when crowded
                const double r = 1.0 + (rmax - 1.0) * U(rng), ct = 2.0 * U(rng) - 1.0, ph = 2.0 * M_PI * U(rng);
                const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
                double px = X[id] + r * st * std::cos(ph), py = Y[id] + r * st * std::sin(ph);
                const double pz = Z[id] + r * ct;
                if (pz < 0.0 || pz >= sp->Lz) continue;
                px -= sp->Lx * std::floor(px / sp->Lx);
                py -= sp->Ly * std::floor(py / sp->Ly);
                if (px >= sp->Lx) px = 0.0;
                if (py >= sp->Ly) py = 0.0;
                if (!grid.clear_of(px, py, pz, sp->min_dist)) continue;
                const double cost = std::fabs(dip + (pz - Z[id]));
                if (found == 0 || cost < best_cost) { best[0] = px; best[1] = py; best[2] = pz; best_cost = cost; }
                ++found;
            }
            if (found) {
                dip += best[2] - Z[id];
                push(best[0], best[1], best[2], qv, tv, mv);
            } else {
                place_free(qv, tv, mv, z_layer);
            }
        };
        const int ion_mol = nch + 1;
        for (int id : perm_ids) place_near(id, -1.0, sp->type_cl, ion_mol);
        for (size_t k = 0; k < site_ids.size(); ++k)
            place_near(site_ids[k], site_prot[k] ? -1.0 : 0.0, site_prot[k] ? sp->type_cl : sp->type_w, ion_mol);
        for (int k = 0; k < sp->n_salt; ++k) {
            place_free(+1.0, sp->type_cation, ion_mol, sp->Lz);
            place_free(-1.0, sp->type_cl, ion_mol, sp->Lz);
        }
        for (long long i = 0; i < total; ++i) {
            x[i] = X[i]; y[i] = Y[i]; z[i] = Z[i];
            if (q) q[i] = Q[i];
            if (type1) type1[i] = T[i];
            if (mol) mol[i] = M[i];
        }
        return PHB_OK;
    });
}
