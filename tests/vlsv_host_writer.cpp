// Writes a VLSV file with the library's host-side writer
// (rhybrid_amd/csrc/rh_vlsv.hpp, the code rh_write_vlsv uses): the grid part
// of every file, then cell variables of known values on a 6x5x4 grid:
//   cellRhoQi[n]    = 1e-13 * (1 + n)                 (FP64 scalar)
//   cellB[n][l]     = 1e-9 * (n + 0.25 * l)           (FP64 3-vector)
//   innerFlagField  = 1 where n % 3 == 0, else 0      (one byte per cell)
//   cellNPles[n]    = n % 7                           (FP64 scalar)
// at time 12.5 s, timestep 250, dt 0.05 s; usage: vlsv_host_writer FILE
#include <cstdio>
#include <vector>

#include "../rhybrid_amd/csrc/rh_vlsv.hpp"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s FILE\n", argv[0]);
        return 2;
    }
    const rh::VlsvGrid g{6, 5, 4, 2.0e5, -6.0e5, -5.0e5, -4.0e5};
    const uint64_t nc = (uint64_t)g.nx * g.ny * g.nz;
    std::vector<double> rho(nc), B(nc * 3), np(nc);
    std::vector<uint8_t> flag(nc);
    for (uint64_t n = 0; n < nc; ++n) {
        rho[n] = 1e-13 * (1.0 + (double)n);
        for (int l = 0; l < 3; ++l) B[3 * n + l] = 1e-9 * ((double)n + 0.25 * l);
        flag[n] = n % 3 == 0 ? 1 : 0;
        np[n] = (double)(n % 7);
    }
    try {
        rh::VlsvWriter W(argv[1]);
        rh::vlsv_write_grid(W, g, 12.5, 250, 0.05);
        rh::vlsv_write_cell_var(W, nc, "cellRhoQi", rho.data(), 1, "float", 8);
        rh::vlsv_write_cell_var(W, nc, "cellB", B.data(), 3, "float", 8);
        rh::vlsv_write_cell_var(W, nc, "innerFlagField", flag.data(), 1, "uint", 1);
        rh::vlsv_write_cell_var(W, nc, "cellNPles", np.data(), 1, "float", 8);
        W.close();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
