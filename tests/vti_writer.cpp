// Writes one XML ImageData file through neon_amd::ioToVTI so the CPU tests
// can compare it byte for byte with neon_amd.export.write_vti.  Only
// header-inline code runs: no library call.
//   vti_writer <path> <nx> <ny> <nz> <double|float|int> <binary mask>
// (nx, ny, nz) is the node box.  Three inputs: "u" (3 components, voxels),
// "rho" (1 component, nodes), "w" (2 components, voxels); bit i of the mask
// makes input i binary.  Value of component k at (x, y, z):
// (x + 10 y + 100 z) / 7 + k / 1000 - 3, rounded to the value type
// (truncated toward zero for int).
#include "neon_amd.hpp"

#include <cstdlib>
#include <iostream>

template <typename V>
static int run(char** a)
{
    using namespace neon_amd;
    const index_3d node(std::atoi(a[2]), std::atoi(a[3]), std::atoi(a[4]));
    const index_3d vox(node.x - 1, node.y - 1, node.z - 1);
    const int      mask = std::atoi(a[6]);
    auto           f = [](const index_3d& i, int k) {
        return V(double(i.x + 10 * i.y + 100 * i.z) / 7.0 + double(k) / 1000.0 - 3.0);
    };
    auto mode = [&](int bit) { return (mask >> bit) & 1 ? IoFileType::BINARY : IoFileType::ASCII; };
    std::vector<VtiInputData_t<V>> in = {{f, 3, "u", false, mode(0)}, {f, 1, "rho", true, mode(1)},
                                         {f, 2, "w", false, mode(2)}};
    ioToVTI<V>(in, a[1], node, vox, double_3d(0.5, 1.0, 2.0), double_3d(-1.0, 0.0, 3.25));
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 7) {
        std::cerr << "usage: vti_writer path nx ny nz type mask\n";
        return 2;
    }
    const std::string t = argv[5];
    try {
        if (t == "double") return run<double>(argv);
        if (t == "float") return run<float>(argv);
        if (t == "int") return run<int32_t>(argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
    return 2;
}
