// Multi-resolution lid-driven cavity on the Neon-shaped C++ API of
// neon_amd.hpp: the flow of apps/lbmMultiRes (lbmMultiRes.cu,
// lidDrivenCavity.h, lbmMultiRes.h, postProcess.h) with its command line, on
// one MI355X.
//
//   lbm_multires [--deviceType gpu] [--deviceId 0] [--numIter N]
//                [--problemType lid|sphere] [--dataType float|double] [--re R]
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
// are recorded in the report, and this program runs its own fused schedule for
// all of them.  --problemType sphere runs the flow over a sphere
// (flowOverShape.h:115-190); --visual --vtk / --binary write postProcess's
// files (postProcess.h:94-227).  Not supported here: --deviceType cpu, the
// mesh problem, the --gui viewer.
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
    bool        vtk = false;
    bool        binary = false;
    int         slice[3] = {-1, -1, 1};  // --sliceX/Y/Z: 1 draws the leaves on that mid-plane
    std::string dataType = "float";
    std::string collision = "kbc";
    std::string commandLine;
};

[[noreturn]] void usage(const std::string& why)
{
    std::cerr << "lbm_multires: " << why << "\n"
              << "usage: lbm_multires [--deviceType gpu] [--deviceId 0] [--numIter N] [--problemType lid|sphere]\n"
                 "       [--dataType float|double] [--re R] [--scale 0..11] [--benchmark|--visual] [--freq F]\n"
                 "       [--vtk] [--binary] [--sliceX S] [--sliceY S] [--sliceZ S]\n"
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
        else if (a == "--sliceX") p.slice[0] = integer(a, need(i));
        else if (a == "--sliceY") p.slice[1] = integer(a, need(i));
        else if (a == "--sliceZ") p.slice[2] = integer(a, need(i));
        else if (a == "--vtk") p.vtk = true;
        else if (a == "--binary") p.binary = true;
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
        else if (a == "--gui" || a == "--meshFile")
            usage(a + " output / input is not supported by this program");
        else usage("unknown option " + a);
    }
    if (p.deviceType != "gpu") usage("only --deviceType gpu is supported by this program");
    if (p.problemType != "lid" && p.problemType != "sphere")
        usage("--problemType lid or sphere (the mesh problem is not supported by this program)");
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

// velocity of a leaf as postProcess computes it (postProcess.h:15-90): bulk
// -> sum c f / rho, moving wall or inlet -> f_q / (6 w_q) along the axis of
// the population with c = -e_a, other cells 0
template <typename T>
void leafVelocity(const Lattice& L, int Q, int cls, const T* f, size_t n, size_t i, T u[3])
{
    u[0] = u[1] = u[2] = T(0);
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
}

// The --vtk file of postProcess (postProcess.h:94-151) over the leaves that
// initVisualization draws (:232-331): every leaf of every level, or with a
// slice set only those whose voxel reaches the domain's mid-plane on that
// axis; each a VTK voxel of side 2^level at its level-0 origin, with its
// velocity and level.
template <typename T>
void writeVtk(const mGrid& grid, const Lattice& L, const int slice[3], const std::string& fileName)
{
    struct Leaf
    {
        int      o[3];
        int      spacing;
        T        u[3];
        int      level;
    };
    constexpr double tiny = 1e-7;
    const index_3d   dim0 = grid.levelDimension(0);
    const int        Q = grid.populations();
    const bool       sliced = slice[0] == 1 || slice[1] == 1 || slice[2] == 1;
    std::vector<Leaf> leaves;
    for (int level = 0; level < grid.getDepth(); level++) {
        const index_3d d = grid.levelDimension(level);
        const double   vs[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
        const auto     st = grid.state(level);
        const auto     f = grid.template populations<T>(level, true);
        const size_t   n = d.rMulTyped();
        for (int z = 0; z < d.z; z++)
            for (int y = 0; y < d.y; y++)
                for (int x = 0; x < d.x; x++) {
                    const size_t  i = size_t(x) + size_t(d.x) * (size_t(y) + size_t(d.y) * size_t(z));
                    const uint8_t s = st[i];
                    if (!(s & 1u) || (s & 2u)) continue;  // leaves only
                    const int o[3] = {x << level, y << level, z << level};
                    bool      draw = !sliced;
                    const int d0[3] = {dim0.x, dim0.y, dim0.z};
                    for (int a = 0; a < 3 && !draw; a++) {
                        const double loc = double(o[a]) / double(d0[a]);
                        if (slice[a] == 1 && loc - tiny <= 0.5 && loc + vs[a] >= 0.5 - tiny) draw = true;
                    }
                    if (!draw) continue;
                    Leaf lf{{o[0], o[1], o[2]}, 1 << level, {}, level};
                    leafVelocity<T>(L, Q, (s >> 2) & 3, f.data(), n, i, lf.u);
                    leaves.push_back(lf);
                }
    }
    std::ofstream file(fileName + ".vtk");
    file << "# vtk DataFile Version 2.0\n"
         << "mGrid\n"
         << "ASCII\n"
         << "DATASET UNSTRUCTURED_GRID\n"
         << "POINTS " << leaves.size() * 8 << " float \n";
    // corner order per leaf: (0,0,0) (1,0,0) (1,0,1) (0,0,1) (0,1,0) (1,1,0) (1,1,1) (0,1,1)
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                     {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
    for (const Leaf& lf : leaves)
        for (const auto& c : corner)
            file << float(lf.o[0] + c[0] * lf.spacing) << " " << float(lf.o[1] + c[1] * lf.spacing) << " "
                 << float(lf.o[2] + c[2] * lf.spacing) << "\n";
    file << "CELLS " << leaves.size() << " " << leaves.size() * 9 << " \n";
    // the reference's point order for its VTK voxels (cell type 11): x
    // fastest, then z, then y (VTK's voxel order with y and z exchanged;
    // the same cube)
    static const int voxelOrder[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    for (size_t k = 0; k < leaves.size(); k++) {
        file << "8 ";
        for (int j : voxelOrder) file << k * 8 + size_t(j) << " ";
        file << "\n";
    }
    file << "CELL_TYPES " << leaves.size() << " \n";
    for (size_t k = 0; k < leaves.size(); k++) file << 11 << "\n";
    file << "CELL_DATA " << leaves.size() << " \n";
    file << "VECTORS Velocity float \n";
    for (const Leaf& lf : leaves) file << lf.u[0] << " " << lf.u[1] << " " << lf.u[2] << " \n";
    file << "SCALARS Level float 1 \n";
    file << "LOOKUP_TABLE default \n";
    for (const Leaf& lf : leaves) file << lf.level << "\n";
}

// The --binary file of postProcess (postProcess.h:153-227): the velocity
// sampled on the top level's lattice -- every leaf whose level-0 origin is a
// multiple of 2^(depth-1), at origin / 2^(depth-1) -- as a 3-component voxel
// field "V" of a binary legacy VTK file named "Bin" + fileName.
template <typename T>
void writeBinary(const mGrid& grid, const Lattice& L, const std::string& fileName)
{
    const int      top = grid.getDepth() - 1;
    const index_3d dim0 = grid.levelDimension(0);
    const index_3d box(dim0.x >> top, dim0.y >> top, dim0.z >> top);
    const int      Q = grid.populations();
    const int      mask = (1 << top) - 1;
    std::vector<float> buf(box.rMulTyped() * 3, 0.f);
    for (int level = 0; level < grid.getDepth(); level++) {
        const index_3d d = grid.levelDimension(level);
        const auto     st = grid.state(level);
        const auto     f = grid.template populations<T>(level, true);
        const size_t   n = d.rMulTyped();
        for (int z = 0; z < d.z; z++)
            for (int y = 0; y < d.y; y++)
                for (int x = 0; x < d.x; x++) {
                    const int ox = x << level, oy = y << level, oz = z << level;
                    if ((ox & mask) || (oy & mask) || (oz & mask)) continue;
                    const size_t  i = size_t(x) + size_t(d.x) * (size_t(y) + size_t(d.y) * size_t(z));
                    const uint8_t s = st[i];
                    if (!(s & 1u) || (s & 2u)) continue;  // leaves only
                    T u[3];
                    leafVelocity<T>(L, Q, (s >> 2) & 3, f.data(), n, i, u);
                    const size_t at = size_t(ox >> top) + size_t(box.x) * (size_t(oy >> top) + size_t(box.y) * size_t(oz >> top));
                    for (int c = 0; c < 3; c++) buf[at * 3 + size_t(c)] = float(u[c]);
                }
    }
    vtk::writeFile("Bin" + fileName, "V", buf, box, 3, false, IoFileType::BINARY, false);
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
                    T u[3];
                    leafVelocity<T>(L, Q, (s >> 2) & 3, f.data(), n, i, u);
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

struct Problem
{
    std::string         record;  // the report's record name
    index_3d            gridDim;
    int                 depth = 0;
    double              clength = 0;
    std::array<double, 3> velocity{};  // lid or inlet
    bool                verify = false;  // centre-line files (the lid problem)
};

// runNonUniformLBM (lbmMultiRes.h:266-491): time numIter iterations, print
// and report MLUPS / EMLUPS, and for the lid problem write the centre lines
template <typename T>
int runNonUniformLBM(const Backend& backend, const Params& params, const Problem& pb, mGrid& g, const Lattice& L)
{
    const int      depth = pb.depth;
    const index_3d gridDim = pb.gridDim;
    const T        ulb = T(pb.velocity[0]);
    const T        clength = T(pb.clength);
    const T        visclb = ulb * clength / static_cast<T>(params.Re);
    const T        omega = 1.0 / (3. * visclb + 0.5);

    const std::vector<uint64_t> leaves = g.numLeafVoxels();
    const bool                  vtk = !params.benchmark && params.vtk;
    const bool                  bin = !params.benchmark && params.binary;
    auto fileName = [](int t) {
        std::ostringstream suffix;
        suffix << "Velocity_" << std::setw(4) << std::setfill('0') << t;
        return suffix.str();
    };
    backend.syncAll();
    const auto start = std::chrono::high_resolution_clock::now();
    if (vtk || bin) {
        // the loop of lbmMultiRes.h:361-376: after iteration t (t + 1 steps)
        // with t % freq == 0 and t != 0, the files of iteration t
        int done = 0;
        for (int t = params.freq; t < params.numIter; t += params.freq) {
            g.run(double(omega), t + 1 - done);
            done = t + 1;
            backend.syncAll();
            if (vtk) writeVtk<T>(g, L, params.slice, fileName(t));
            if (bin) writeBinary<T>(g, L, fileName(t));
        }
        g.run(double(omega), params.numIter - done);
    } else {
        g.run(double(omega), params.numIter);
    }
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
    vel << "(" << pb.velocity[0] << ", " << pb.velocity[1] << ", " << pb.velocity[2] << ")";
    std::ofstream r("MultiResLBM_P" + std::to_string(params.scale) + "_" + algo + "_" + typeName + reportTimeSuffix() +
                    ".json");
    r << std::setprecision(17) << "{\n"
      << "    \"Record Name\": \"" << pb.record << "\",\n"
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
      << "    \"Collision\": \"" << (params.collision == "kbc" ? "KBC" : "BGK") << "\",\n"
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

    if (vtk) writeVtk<T>(g, L, params.slice, fileName(params.numIter));  // lbmMultiRes.h:490-499
    if (bin) writeBinary<T>(g, L, fileName(params.numIter));
    if (pb.verify) writeCentreLines<T>(g, L, ulb, params.numIter);
    return 0;
}

template <typename T>
int lidDrivenCavity(const Backend& backend, const Params& params)
{
    const LidGrid  lg = lidGrid(params.scale);
    Problem        pb;
    pb.record = "Lid Driven Cavity MultiRes LBM";
    pb.depth = int(lg.sdf.size()) - 1;
    pb.gridDim = index_3d(lg.edge, lg.edge, lg.edge);
    pb.velocity = {double(T(0.04)), 0., 0.};
    pb.verify = true;  // lidDrivenCavity.h:256-266
    const index_3d box(pb.gridDim.x - 1, pb.gridDim.y - 1, pb.gridDim.z - 1);
    std::vector<std::function<bool(const index_3d&)>> activeCellLambda(static_cast<size_t>(pb.depth), nullptr);
    for (int i = 0; i < pb.depth; i++) {
        const float hi = float(lg.sdf[size_t(i)]), lo = float(lg.sdf[size_t(i) + 1]);
        activeCellLambda[size_t(i)] = [=](const index_3d& id) {
            const float sdf = sdfCube(id, box);
            return sdf <= hi && sdf > lo;
        };
    }
    const bool kbc = params.collision == "kbc";
    mGrid      g = mGrid::make<T>(backend, pb.gridDim, activeCellLambda, kbc ? mGrid::Collision::KBC
                                                                        : mGrid::Collision::BGK);
    pb.clength = double(T(g.getDimension(pb.depth - 1).x));  // lidDrivenCavity.h:241
    g.initLidDrivenCavity(pb.velocity[0]);
    return runNonUniformLBM<T>(backend, params, pb, g, lattice(kbc));
}

// flowOverSphere (flowOverShape.h:115-190): a 136s x 96s x 136s channel,
// level boxes, a sphere of radius 8s at (52s, 52s, 68s), the inlet at x = 0
template <typename T>
int flowOverSphere(const Backend& backend, const Params& params)
{
    const int s = params.scale;
    if (s < 1) usage("--scale must be at least 1 for the sphere");
    Problem pb;
    pb.record = "Sphere MultiRes LBM";
    pb.depth = 3;
    pb.gridDim = index_3d(136 * s, 96 * s, 136 * s);
    const int sx = 52 * s, sy = 52 * s, sz = 68 * s, sr = 8 * s;
    pb.velocity = {double(T(0.04)), 0., 0.};
    pb.clength = double(T(sr / (1 << (pb.depth - 1))));
    auto box = [](int x0, int x1, int y0, int y1, int z0, int z1) {
        return [=](const index_3d& i) { return i.x >= x0 && i.x < x1 && i.y >= y0 && i.y < y1 && i.z >= z0 && i.z < z1; };
    };
    const std::vector<std::function<bool(const index_3d&)>> activeCellLambda = {
        box(40 * s, 96 * s, 40 * s, 64 * s, 40 * s, 96 * s), box(24 * s, 112 * s, 24 * s, 72 * s, 24 * s, 112 * s),
        [](const index_3d&) { return true; }};
    const bool kbc = params.collision == "kbc";
    mGrid      g = mGrid::make<T>(backend, pb.gridDim, activeCellLambda, kbc ? mGrid::Collision::KBC
                                                                        : mGrid::Collision::BGK);
    // cell classes (flowOverShape.h:56-83), from each voxel's level-0 origin
    std::vector<std::vector<uint8_t>> cls(size_t(pb.depth));
    for (int l = 0; l < pb.depth; l++) {
        const index_3d d = g.levelDimension(l);
        auto&          c = cls[size_t(l)];
        c.assign(d.rMulTyped(), NX_CELL_BULK);
        size_t i = 0;
        for (int z = 0; z < d.z; z++)
            for (int y = 0; y < d.y; y++)
                for (int x = 0; x < d.x; x++, i++) {
                    const int gx = x << l, gy = y << l, gz = z << l;
                    if (gx == 0) c[i] = NX_CELL_MOVINGWALL;  // inlet
                    const T dx = T(sx - gx), dy = T(sy - gy), dz = T(sz - gz);
                    if ((dx * dx + dy * dy + dz * dz) < T(sr) * T(sr)) c[i] = NX_CELL_BOUNCEBACK;
                    if (gy == 0 || gy == pb.gridDim.y - (1 << l) || gz == 0 || gz == pb.gridDim.z - (1 << l))
                        c[i] = NX_CELL_BOUNCEBACK;
                }
    }
    g.initCells(cls, pb.velocity.data());
    return runNonUniformLBM<T>(backend, params, pb, g, lattice(kbc));
}

}  // namespace

int main(int argc, char** argv)
{
    const Params params = parse(argc, argv);
    try {
        Backend backend({params.deviceId}, Runtime::stream);
        if (params.problemType == "sphere")
            return params.dataType == "float" ? flowOverSphere<float>(backend, params)
                                              : flowOverSphere<double>(backend, params);
        if (params.dataType == "float") return lidDrivenCavity<float>(backend, params);
        return lidDrivenCavity<double>(backend, params);
    } catch (const NeonException& e) {
        std::cerr << "NeonException: " << e.what() << std::endl;
        return 1;
    }
}
