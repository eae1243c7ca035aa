// Multi-resolution lid-driven cavity on the Neon-shaped C++ API of
// neon_amd.hpp: the flow of apps/lbmMultiRes (lbmMultiRes.cu,
// lidDrivenCavity.h, lbmMultiRes.h, postProcess.h) with its command line, on
// one MI355X.
//
//   lbm_multires [--deviceType gpu] [--deviceId 0] [--numIter N]
//                [--problemType lid] [--dataType float|double] [--re R]
//                [--scale S] [--benchmark|--visual] [--freq F]
//                [--storeFine|--storeCoarse|--collisionFusedStore] [--fusedFinest]
//                [--streamFusedExpl|--streamFusedCoal|--streamFuseAll]
//                [--sliceX/Y/Z i] [--thetaX/Y/Z a] [--collision kbc|bgk]
//
// As in the reference: the lid problem's grid for --scale (edge and level
// bands, lidDrivenCavity.h:95-214; default scale 2), ulb 0.04, clength the
// top level's dimension, nu = ulb clength / Re, omega = 1 / (3 nu + 1/2)
// (:240-244); MLUPS counts level l's leaves 2^(depth-1-l) times per
// iteration, EMLUPS the whole box at level-0 resolution (lbmMultiRes.h:
// 385-396); the JSON report MultiResLBM_P<scale>_<algorithm>_<F|D> with the
// reference's members and date suffix (:439-491, Report.cpp:94-121); the
// centre lines NeonMultiResLBM_<iter>_{X,Y}.dat that verifyLidDrivenCavity
// writes after the run (postProcess.h:404-463), velocities from fout.  The
// reference chooses KBC D3Q27 or BGK D3Q19 when it is compiled
// (lbmMultiRes.cu:3-4, KBC by default); here --collision chooses.  The store
// and stream options select kernel fusions that reorder the same arithmetic
// (with the command line's choices the store is always the fine one): they
// are recorded in the report, and this build runs its own fused schedule for
// all of them.  Not part of this build: --deviceType cpu, the sphere and mesh
// problems, --vtk / --binary / --gui output.
#include "neon_amd.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

using namespace neon_amd;

struct Params
{
    std::string deviceType = "gpu";
    std::string problemType = "lid";
    int         freq = 100;
    int         Re = 100;
    int         deviceId = 0;
    int         numIter = 1000;
    bool        benchmark = true;
    bool        fineInitStore = true;
    bool        streamFusedExpl = false;
    bool        streamFusedCoal = false;
    bool        streamFuseAll = true;
    bool        collisionFusedStore = true;
    bool        fusedFinest = true;
    int         scale = 2;
    std::string dataType = "float";
    std::string collision = "kbc";
    std::string commandLine;
};

[[noreturn]] void usage(const std::string& why)
{
    std::cerr << "lbm_multires: " << why << "\n"
              << "usage: lbm_multires [--deviceType gpu] [--deviceId 0] [--numIter N] [--problemType lid]\n"
                 "       [--dataType float|double] [--re R] [--scale 0..11] [--benchmark|--visual] [--freq F]\n"
                 "       [--storeFine|--storeCoarse|--collisionFusedStore] [--fusedFinest]\n"
                 "       [--streamFusedExpl|--streamFusedCoal|--streamFuseAll] [--collision kbc|bgk]\n";
    std::exit(2);
}

Params parse(int argc, char** argv)
{
    Params p;
    p.commandLine = argv[0];
    for (int i = 1; i < argc; i++) p.commandLine += std::string(" ") + argv[i];
    auto need = [&](int& i) -> std::string {
        if (i + 1 >= argc) usage(std::string("missing value after ") + argv[i]);
        return argv[++i];
    };
    auto integer = [&](const std::string& flag, const std::string& v) -> int {
        try {
            size_t used = 0;
            const int x = std::stoi(v, &used);
            if (used != v.size()) throw std::invalid_argument(v);
            return x;
        } catch (const std::exception&) {
            usage("bad value for " + flag + ": " + v);
        }
    };
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--deviceType") p.deviceType = need(i);
        else if (a == "--deviceId") p.deviceId = integer(a, need(i));
        else if (a == "--numIter") p.numIter = integer(a, need(i));
        else if (a == "--problemType") p.problemType = need(i);
        else if (a == "--dataType") p.dataType = need(i);
        else if (a == "--re") p.Re = integer(a, need(i));
        else if (a == "--scale") p.scale = integer(a, need(i));
        else if (a == "--freq") p.freq = integer(a, need(i));
        else if (a == "--sliceX" || a == "--sliceY" || a == "--sliceZ") integer(a, need(i));
        else if (a == "--thetaX" || a == "--thetaY" || a == "--thetaZ") need(i);
        else if (a == "--benchmark") p.benchmark = true;
        else if (a == "--visual") p.benchmark = false;
        else if (a == "--storeFine") p.fineInitStore = true;
        else if (a == "--storeCoarse") p.fineInitStore = false;
        else if (a == "--collisionFusedStore") p.collisionFusedStore = true;
        else if (a == "--fusedFinest") p.fusedFinest = true;
        else if (a == "--streamFusedExpl") p.streamFusedExpl = true;
        else if (a == "--streamFusedCoal") p.streamFusedCoal = true;
        else if (a == "--streamFuseAll") p.streamFuseAll = true;
        else if (a == "--collision") p.collision = need(i);
        else if (a == "--vtk" || a == "--binary" || a == "--gui" || a == "--meshFile")
            usage(a + " output / input is not part of this build");
        else usage("unknown option " + a);
    }
    if (p.deviceType != "gpu") usage("only --deviceType gpu is part of this build");
    if (p.problemType != "lid") usage("only --problemType lid is part of this build");
    if (p.dataType != "float" && p.dataType != "double") usage("--dataType must be float or double");
    if (p.collision != "kbc" && p.collision != "bgk") usage("--collision must be kbc or bgk");
    if (p.numIter < 0 || p.Re <= 0 || p.freq <= 0) usage("bad numIter, re or freq");
    return p;
}

// the lid problem's grids (lidDrivenCavity.h:95-214): edge and levelSDF
struct LidGrid
{
    int                 edge;
    std::vector<double> sdf;
};

LidGrid lidGrid(int scale)
{
    switch (scale) {
        case 0: return {48, {0, -8.0 / 24.0, -16.0 / 24.0, -1.0}};
        case 1: return {160, {0, -16.0 / 80.0, -32.0 / 80.0, -1.0}};
        case 2: return {240, {0, -24.0 / 120.0, -80.0 / 120.0, -1.0}};
        case 3: return {320, {0, -32.0 / 160.0, -64.0 / 160.0, -1.0}};
        case 4: return {480, {0, -48.0 / 240.0, -96.0 / 240.0, -1.0}};
        case 5: return {512, {0, -64.0 / 256.0, -112.0 / 256.0, -1.0}};
        case 6: return {160, {0, -32.0 / 160.0, -64.0 / 160.0, -128.0 / 160.0, -1.0}};
        case 7: return {240, {0, -32.0 / 120.0, -80.0 / 120.0, -112.0 / 120.0, -1.0}};
        case 8: return {320, {0, -32.0 / 160.0, -64.0 / 160.0, -112.0 / 160.0, -1.0}};
        case 9: return {480, {0, -48.0 / 240.0, -96.0 / 240.0, -160.0 / 240.0, -1.0}};
        case 10:
        case 11: return {512, {0, -120.0 / 512.0, -200.0 / 512.0, -400.0 / 512.0, -1.0}};
        default: return {192, {0, -36.0 / 96.0, -72.0 / 96.0, -1.0}};  // the app's grid before the scale switch
    }
}

// sdfCube (util.h:65-81) in binary32: signed distance to the faces of the
// box [0, dim] mapped to [-1, 1]^3
float sdfCube(const index_3d& id, const index_3d& dim)
{
    const float p[3] = {(float(id.x) - float(dim.x) * 0.5f) / (float(dim.x) * 0.5f),
                        (float(id.y) - float(dim.y) * 0.5f) / (float(dim.y) * 0.5f),
                        (float(id.z) - float(dim.z) * 0.5f) / (float(dim.z) * 0.5f)};
    float d[3], m = -1e30f, len2 = 0.f;
    for (int a = 0; a < 3; a++) {
        d[a] = std::abs(p[a]) - 1.0f;
        m = std::max(m, d[a]);
        const float c = std::max(d[a], 0.0f);
        len2 += c * c;
    }
    return std::min(m, 0.0f) + std::sqrt(len2);
}

struct Lattice
{
    std::vector<std::array<int, 3>> c;
    std::vector<double>             w;
};

Lattice lattice(bool kbc)
{
    Lattice L;
    if (kbc) {  // lattice.h:14-75
        for (int x : {0, -1, 1})
            for (int y : {0, -1, 1})
                for (int z : {0, -1, 1}) {
                    L.c.push_back({x, y, z});
                    const int n = (x != 0) + (y != 0) + (z != 0);
                    L.w.push_back(n == 0 ? 8.0 / 27.0 : n == 1 ? 2.0 / 27.0 : n == 2 ? 1.0 / 54.0 : 1.0 / 216.0);
                }
    } else {  // lattice.h:108-173, the cavity benchmark's D3Q19
        lbm::D3Q19 d;
        for (int q = 0; q < 19; q++) {
            L.c.push_back({d.c_vect[q].x, d.c_vect[q].y, d.c_vect[q].z});
            L.w.push_back(d.t_vect[q]);
        }
    }
    return L;
}

std::string reportTimeSuffix()
{
    const std::time_t t = std::time(nullptr);
    const std::tm     tm = *std::localtime(&t);
    std::ostringstream oss;
    oss << std::put_time(&tm, "_D%d_%m_%Y__T%H_%M_%S");
    return oss.str();
}

std::string dateString()
{
    const std::time_t t = std::time(nullptr);
    const std::tm     tm = *std::localtime(&t);
    std::ostringstream oss;
    oss << std::put_time(&tm, "%a %d:%m:%Y %H:%M:%S");
    return oss.str();
}

template <typename T>
void writeCentreLines(const mGrid& grid, const Lattice& L, T ulb, int iteration)
{
    // verifyLidDrivenCavity (postProcess.h:404-463) over the velocities of
    // postProcess (:15-90) computed from fout
    std::vector<std::pair<T, T>> xPosVal, yPosVal;
    const index_3d dim = grid.getDimension();
    const T        scale = T(1.0 / double(ulb));
    const int      Q = grid.populations();
    for (int level = 0; level < grid.getDepth(); level++) {
        const index_3d d = grid.levelDimension(level);
        const auto     st = grid.state(level);
        const auto     f = grid.template populations<T>(level, true);
        const size_t   n = d.rMulTyped();
        for (int z = 0; z < d.z; z++)
            for (int y = 0; y < d.y; y++)
                for (int x = 0; x < d.x; x++) {
                    const index_3d id(x << level, y << level, z << level);
                    const bool     onY = id.x == dim.x / 2 && id.z == dim.z / 2;
                    const bool     onX = id.y == dim.y / 2 && id.z == dim.z / 2;
                    if (!onX && !onY) continue;
                    const size_t  i = size_t(x) + size_t(d.x) * (size_t(y) + size_t(d.y) * size_t(z));
                    const uint8_t s = st[i];
                    if (!(s & 1u) || (s & 2u)) continue;  // leaves only
                    const int cls = (s >> 2) & 3;
                    T         u[3] = {0, 0, 0};
                    if (cls == NX_CELL_BULK) {
                        T r = 0;
                        for (int q = 0; q < Q; q++) r += f[size_t(q) * n + i];
                        for (int q = 0; q < Q; q++)
                            for (int a = 0; a < 3; a++) u[a] += f[size_t(q) * n + i] * T(L.c[q][a]);
                        for (int a = 0; a < 3; a++) u[a] /= r;
                    } else if (cls == NX_CELL_MOVINGWALL) {
                        for (int q = 0; q < Q; q++)
                            for (int a = 0; a < 3; a++)
                                if (L.c[q][a] == -1 && L.c[q][(a + 1) % 3] == 0 && L.c[q][(a + 2) % 3] == 0)
                                    u[a] = T(double(f[size_t(q) * n + i]) / (6. * L.w[q]));
                    }
                    if (onY) yPosVal.push_back({T(double(id.y) / double(dim.y)), u[0] * scale});
                    if (onX) xPosVal.push_back({T(double(id.x) / double(dim.x)), u[1] * scale});
                }
    }
    auto byPos = [](const std::pair<T, T>& a, const std::pair<T, T>& b) { return a.first < b.first; };
    std::sort(xPosVal.begin(), xPosVal.end(), byPos);
    std::sort(yPosVal.begin(), yPosVal.end(), byPos);
    std::ostringstream suffix;
    suffix << std::setw(4) << std::setfill('0') << iteration;
    auto write = [](const std::vector<std::pair<T, T>>& pv, const std::string& name) {
        std::ofstream file(name);
        for (const auto& v : pv) file << v.first << " " << v.second << "\n";
    };
    write(yPosVal, "NeonMultiResLBM_" + suffix.str() + "_Y.dat");
    write(xPosVal, "NeonMultiResLBM_" + suffix.str() + "_X.dat");
}

template <typename T>
int lidDrivenCavity(const Backend& backend, const Params& params)
{
    const LidGrid  lg = lidGrid(params.scale);
    const int      depth = int(lg.sdf.size()) - 1;
    const index_3d gridDim(lg.edge, lg.edge, lg.edge);
    const index_3d box(gridDim.x - 1, gridDim.y - 1, gridDim.z - 1);
    std::vector<std::function<bool(const index_3d&)>> activeCellLambda(static_cast<size_t>(depth), nullptr);
    for (int i = 0; i < depth; i++) {
        const float hi = float(lg.sdf[size_t(i)]), lo = float(lg.sdf[size_t(i) + 1]);
        activeCellLambda[size_t(i)] = [=](const index_3d& id) {
            const float sdf = sdfCube(id, box);
            return sdf <= hi && sdf > lo;
        };
    }
    const bool  kbc = params.collision == "kbc";
    const mGrid grid = mGrid::make<T>(backend, gridDim, activeCellLambda, kbc ? mGrid::Collision::KBC
                                                                              : mGrid::Collision::BGK);
    // LBM problem (lidDrivenCavity.h:240-244)
    const T      ulb = 0.04;
    const T      clength = T(grid.getDimension(depth - 1).x);
    const T      visclb = ulb * clength / static_cast<T>(params.Re);
    const T      omega = 1.0 / (3. * visclb + 0.5);
    const Lattice L = lattice(kbc);

    mGrid g = grid;
    g.initLidDrivenCavity(double(ulb));
    const std::vector<uint64_t> leaves = g.numLeafVoxels();
    backend.syncAll();
    const auto start = std::chrono::high_resolution_clock::now();
    g.run(double(omega), params.numIter);
    backend.syncAll();
    const auto   stop = std::chrono::high_resolution_clock::now();
    const double us = double(std::chrono::duration_cast<std::chrono::microseconds>(stop - start).count());

    double MLUPS = 0;
    for (int l = 0; l < depth; l++)
        MLUPS += double(params.numIter) * std::pow(2, double(depth - 1 - l)) * double(leaves[size_t(l)]);
    MLUPS /= us;
    const double effNumIter = double(params.numIter) * double(1 << (depth - 1));
    const double effMLUPS = effNumIter * double(gridDim.x) * double(gridDim.y) * double(gridDim.y) / us;
    std::cout << "MLUPS = " << std::fixed << std::setprecision(8) << MLUPS << "\n"
              << "Time = " << us << " (microseconds)\n"
              << "Effective MLUPS = " << effMLUPS << ", Effective numActiveVoxels = " << gridDim.rMulTyped() << "\n";
    std::cout.unsetf(std::ios::floatfield);

    // report (lbmMultiRes.h:397-491)
    std::string algo = params.collisionFusedStore ? "CH" : (params.fineInitStore ? "C-H" : "C-h");
    algo += params.streamFusedExpl ? "-SE-O" : params.streamFusedCoal ? "-SO-E" : params.streamFuseAll ? "-SEO" : "-S-E-O";
    if (params.fusedFinest) algo += "+";
    const std::string typeName = std::is_same_v<T, float> ? "F" : "D";
    auto b = [](bool v) { return v ? "true" : "false"; };
    std::ostringstream vel;
    vel << "(" << double(ulb) << ", 0, 0)";
    std::ofstream r("MultiResLBM_P" + std::to_string(params.scale) + "_" + algo + "_" + typeName + reportTimeSuffix() +
                    ".json");
    r << std::setprecision(17) << "{\n"
      << "    \"Record Name\": \"Lid Driven Cavity MultiRes LBM\",\n"
      << "    \"date\": \"" << dateString() << "\",\n"
      << "    \"command_line\": \"" << params.commandLine << "\",\n"
      << "    \"DeviceType\": \"gpu\",\n"
      << "    \"Device\": \"" << backend.toString() << "\",\n"
      << "    \"Grid Size X\": " << gridDim.x << ",\n    \"Grid Size Y\": " << gridDim.y << ",\n"
      << "    \"Grid Size Z\": " << gridDim.z << ",\n    \"Depth\": " << depth << ",\n"
      << "    \"DataType\": \"" << typeName << "\",\n    \"ProblemScale\": " << params.scale << ",\n"
      << "    \"problemType\": \"" << params.problemType << "\",\n    \"omega\": " << double(omega) << ",\n"
      << "    \"Re\": " << params.Re << ",\n    \"velocity\": \"" << vel.str() << "\",\n"
      << "    \"clength\": " << double(clength) << ",\n    \"visclb\": " << double(visclb) << ",\n"
      << "    \"Collision\": \"" << (kbc ? "KBC" : "BGK") << "\",\n"
      << "    \"fineInitStore\": " << b(params.fineInitStore) << ",\n"
      << "    \"streamFusedExpl\": " << b(params.streamFusedExpl) << ",\n"
      << "    \"streamFusedCoal\": " << b(params.streamFusedCoal) << ",\n"
      << "    \"streamFuseAll\": " << b(params.streamFuseAll) << ",\n"
      << "    \"collisionFusedStore\": " << b(params.collisionFusedStore) << ",\n"
      << "    \"fusedFinest\": " << b(params.fusedFinest) << ",\n"
      << "    \"Algorithm\": \"" << algo << "\",\n"
      << "    \"Time (microsecond)\": " << us << ",\n    \"MLUPS\": " << MLUPS << ",\n"
      << "    \"NumIter\": " << params.numIter << ",\n    \"NumActiveVoxels\": [";
    for (int l = 0; l < depth; l++) r << (l ? ", " : "") << leaves[size_t(l)];
    r << "],\n    \"EMLUPS\": " << effMLUPS << ",\n    \"ENumIter\": " << effNumIter << ",\n"
      << "    \"ENumVoxels\": " << gridDim.rMulTyped() << "\n}\n";

    // the app runs its lid problem with verify on (lidDrivenCavity.h:256-266)
    writeCentreLines<T>(g, L, ulb, params.numIter);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const Params params = parse(argc, argv);
    try {
        Backend backend({params.deviceId}, Runtime::stream);
        if (params.dataType == "float") return lidDrivenCavity<float>(backend, params);
        return lidDrivenCavity<double>(backend, params);
    } catch (const NeonException& e) {
        std::cerr << "NeonException: " << e.what() << std::endl;
        return 1;
    }
}
