// ref_io_dump.cpp -- test infrastructure: io_dump.hpp over the reference's own
// file-format code (MC/src/io.cpp with MC/include/io.hpp, types.hpp and
// precision.hpp, FP64 build), compiled from the reference sources by
// oracle/Makefile into oracle/_ref/ref_io_dump.
#include "io.hpp"
#include "io_dump.hpp"

struct RefApi {
    static System read_data(const char* p) { return read_lammps_full(p); }
    static Params read_params(const char* p, int ntypes) { return ::read_params(p, ntypes); }
    static void wrap(System& s) { wrap_positions(s); }
    static void write_data(const char* p, const System& s) { write_lammps_full(p, s, s.box); }
};

int main(int argc, char** argv) { return io_dump_main<System, Params, RefApi>(argc, argv); }
