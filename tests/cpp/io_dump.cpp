// io_dump.cpp -- oracle/io_dump.hpp over this library's file-format layer
// (include/ph_brush/io.hpp), for comparison with the same dump made by the
// reference's own io.cpp.
#include "ph_brush/io.hpp"
#include "io_dump.hpp"

struct LibApi {
    static ph_brush::System read_data(const char* p) { return ph_brush::read_lammps_full(p); }
    static ph_brush::Params read_params(const char* p, int ntypes) { return ph_brush::read_params(p, ntypes); }
    static void wrap(ph_brush::System& s) { ph_brush::wrap_positions(s); }
    static void write_data(const char* p, const ph_brush::System& s) { ph_brush::write_lammps_full(p, s, s.box); }
};

int main(int argc, char** argv) { return io_dump_main<ph_brush::System, ph_brush::Params, LibApi>(argc, argv); }
