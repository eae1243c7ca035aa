// Writes one legacy VTK file through neon_amd::vtk::writeFile (the writer
// behind Field::ioToVtk) so the CPU tests can compare it byte for byte with
// neon_amd.export.write_vtk.  Only header-inline code runs: no library call.
//   vtk_writer <prefix> <nx> <ny> <nz> <card> <double|float|int> <binary> <domain> <node>
// Value of component k at (x, y, z): (x + 10 y + 100 z) / 7 + k / 1000,
// rounded to the export type (truncated toward zero for int).
#include "neon_amd.hpp"

#include <cstdlib>
#include <iostream>

template <typename V>
static void run(char** a)
{
    using neon_amd::index_3d;
    const index_3d dim(std::atoi(a[2]), std::atoi(a[3]), std::atoi(a[4]));
    const int      card = std::atoi(a[5]);
    std::vector<V> v(dim.rMulTyped() * size_t(card));
    size_t         i = 0;
    for (int z = 0; z < dim.z; z++)
        for (int y = 0; y < dim.y; y++)
            for (int x = 0; x < dim.x; x++)
                for (int k = 0; k < card; k++)
                    v[i++] = V(double(x + 10 * y + 100 * z) / 7.0 + double(k) / 1000.0);
    neon_amd::vtk::writeFile(a[1], "field", v, dim, card, std::atoi(a[8]) != 0,
                             std::atoi(a[7]) ? neon_amd::IoFileType::BINARY : neon_amd::IoFileType::ASCII,
                             std::atoi(a[9]) != 0);
}

int main(int argc, char** argv)
{
    if (argc != 10) {
        std::cerr << "usage: vtk_writer prefix nx ny nz card type binary domain node\n";
        return 2;
    }
    const std::string t = argv[6];
    if (t == "double") run<double>(argv);
    else if (t == "float") run<float>(argv);
    else if (t == "int") run<int32_t>(argv);
    else return 2;
    return 0;
}
