// Lid-driven cavity benchmark on the Neon-shaped C++ API of neon_amd.hpp:
// the flow of benchmarks/lbm-lid-driven-cavity-flow (app.cpp,
// RunCavityTwoPop.cu, Config.cpp, Metrics.h, Report.cpp) with its command
// line, on one MI355X.
//
//   lbm_cavity --deviceType gpu --deviceIds 0 [--grid dGrid|bGrid]
//              [--domain-size N] [--warmup-iter W] [--max-iter M]
//              [--repetitions R] [--report-filename F]
//              [--computeFP double|float|float-fast]
//              [--storageFP double|float|float-dev|half-dev]
//              [--sOCC|--nOCC] [--put|--get] [--huLattice|--huGrid]
//              [--benchmark|--visual] [--vti] [--Re R] [--ulb U]
//              [--stagedNeighbourhood]
//
// As in the reference: nu = ulb (N-2) / Re, omega = 1 / (3 nu + 1/2)
// (Config.cpp:105-111); MLUPS = N^3 iterations / microseconds over the timed
// loop (Metrics.h:31-50); visual mode writes the centre lines every 100
// iterations (RunCavityTwoPop.cu:86-151), and with --vti also the legacy VTK
// files u_<iter>.vtk and rho_<iter>.vtk that the reference writes there
// unconditionally (:105-106; its --vti flag, Config.cpp:90, is only
// recorded).  float-fast, float-dev, half-dev, --Re, --ulb and
// --stagedNeighbourhood (bGrid: the step with the block neighbourhood staged
// in LDS, bGrid::setStagedNeighbourhood) are additions;
// eGrid and --deviceType cpu are not supported and are rejected.
#include "neon_amd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

using namespace neon_amd;

struct Config
{
    double           Re = 100.;
    double           ulb = 0.04;
    int              N = 160;
    bool             benchmark = true;
    int              benchIniIter = 1000;
    int              benchMaxIter = 2000;
    int              repetitions = 1;
    std::string      deviceType = "gpu";
    std::vector<int> devices;
    std::string      reportFile = "lbm-lid-driven-cavity-flow";
    std::string      gridType = "dGrid";
    bool             stagedNeighbourhood = false;  // bGrid: the LDS-staged step
    std::string      occ = "none", transferMode = "get", stencilSemantic = "streaming";
    bool             vti = false;
    std::string      computeType = "double";
    std::string      storeType = "double";
    double           nu = 0, omega = 0, dx = 0, dt = 0;

    void setLbmParameters()
    {
        nu = ulb * double(N - 2) / Re;
        omega = 1. / (3. * nu + 0.5);
        dx = 1. / double(N - 2);
        dt = dx * ulb;
    }
};

[[noreturn]] void usage(const std::string& why)
{
    std::cerr << "lbm_cavity: " << why << "\n"
              << "usage: lbm_cavity --deviceType gpu --deviceIds 0 [--grid dGrid|bGrid] [--domain-size N]\n"
                 "       [--warmup-iter W] [--max-iter M] [--repetitions R] [--report-filename F]\n"
                 "       [--computeFP double|float|float-fast] [--storageFP double|float|float-dev|half-dev]\n"
                 "       [--sOCC|--nOCC] [--put|--get] [--huLattice|--huGrid] [--benchmark|--visual]\n"
                 "       [--Re R] [--ulb U] [--stagedNeighbourhood]\n";
    std::exit(2);
}

Config parse(int argc, char** argv)
{
    Config c;
    bool   haveType = false, haveIds = false;
    auto   need = [&](int& i) -> std::string {
        if (i + 1 >= argc) usage(std::string("missing value after ") + argv[i]);
        return argv[++i];
    };
    auto num = [&](const std::string& flag, const std::string& v) -> double {
        try {
            size_t used = 0;
            const double x = std::stod(v, &used);
            if (used != v.size()) throw std::invalid_argument(v);
            return x;
        } catch (const std::exception&) {
            usage("bad value for " + flag + ": " + v);
        }
    };
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--deviceType") {
            c.deviceType = need(i);
            haveType = true;
        } else if (a == "--deviceIds") {
            haveIds = true;
            while (i + 1 < argc && argv[i + 1][0] != '-') c.devices.push_back(int(num(a, argv[++i])));
            if (c.devices.empty()) usage("--deviceIds needs at least one id");
        } else if (a == "--grid") {
            c.gridType = need(i);
        } else if (a == "--stagedNeighbourhood") {
            c.stagedNeighbourhood = true;
        } else if (a == "--domain-size") {
            c.N = int(num(a, need(i)));
        } else if (a == "--warmup-iter") {
            c.benchIniIter = int(num(a, need(i)));
        } else if (a == "--max-iter") {
            c.benchMaxIter = int(num(a, need(i)));
        } else if (a == "--repetitions") {
            c.repetitions = int(num(a, need(i)));
        } else if (a == "--report-filename") {
            c.reportFile = need(i);
        } else if (a == "--computeFP") {
            c.computeType = need(i);
        } else if (a == "--storageFP") {
            c.storeType = need(i);
        } else if (a == "--sOCC") {
            c.occ = "standard";
        } else if (a == "--nOCC") {
            c.occ = "none";
        } else if (a == "--put") {
            c.transferMode = "put";
        } else if (a == "--get") {
            c.transferMode = "get";
        } else if (a == "--huLattice") {
            c.stencilSemantic = "streaming";
        } else if (a == "--huGrid") {
            c.stencilSemantic = "standard";
        } else if (a == "--benchmark") {
            c.benchmark = true;
        } else if (a == "--visual") {
            c.benchmark = false;
        } else if (a == "--vti") {
            c.vti = true;
        } else if (a == "--Re") {
            c.Re = num(a, need(i));
        } else if (a == "--ulb") {
            c.ulb = num(a, need(i));
        } else {
            usage("unknown option " + a);
        }
    }
    if (!haveType || !haveIds) usage("--deviceType and --deviceIds are required");
    if (c.deviceType != "gpu") usage("only --deviceType gpu is supported by this program");
    if (c.devices.size() != 1) usage("one process drives one GPU: pass exactly one device id");
    if (c.gridType != "dGrid" && c.gridType != "bGrid") usage("--grid must be dGrid or bGrid (eGrid is not supported by this program)");
    if (c.stagedNeighbourhood && (c.gridType != "bGrid" || c.storeType == "double"))
        usage("--stagedNeighbourhood needs --grid bGrid and 2- or 4-byte storage");
    if (c.N < 3 || c.benchMaxIter < 0 || c.benchIniIter < 0 || c.repetitions < 1) usage("bad sizes");
    c.setLbmParameters();
    return c;
}

std::string toString(const Config& c)
{
    std::stringstream s;
    s << ".................. Re " << c.Re << "\n"
      << "................. ulb " << c.ulb << "\n"
      << "................... N " << c.N << "\n"
      << "........... benchmark " << c.benchmark << "\n"
      << "........ benchIniIter " << c.benchIniIter << "\n"
      << "........ benchMaxIter " << c.benchMaxIter << "\n"
      << ".......... deviceType " << c.deviceType << "\n"
      << "............. devices " << c.devices[0] << "\n"
      << "............ gridType " << c.gridType << "\n"
      << "......... computeType " << c.computeType << "\n"
      << "........... storeType " << c.storeType << "\n"
      << "................. occ " << c.occ << "\n"
      << "....... transfer Mode " << c.transferMode << "\n"
      << "... transfer Semantic " << c.stencilSemantic << "\n"
      << "................... nu " << c.nu << "\n"
      << "................ omega " << c.omega << "\n"
      << "................... dx " << c.dx << "\n"
      << "................... dt " << c.dt << "\n";
    return s.str();
}

struct Report
{
    std::vector<double> mlups, loopTime, gridInit, setup;
};

// Host-side fill of RunCavityTwoPop.cu:168-206 for double / float storage; the
// deviation storages are filled on the device (nx_d3q19_fill_*).
template <typename T>
constexpr bool kHostFill = std::is_same_v<T, double> || std::is_same_v<T, float>;

template <typename Grid, typename StorageFP, typename ComputeFP>
void run(const Config& config, Report& report)
{
    using PopulationField = typename Grid::template Field<StorageFP, 19>;
    using CellTypeField = typename Grid::template Field<CellType, 1>;
    using Moment = std::conditional_t<std::is_same_v<StorageFP, double>, double, float>;
    using Clock = std::chrono::high_resolution_clock;
    auto us = [](Clock::time_point a) {
        return double(std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - a).count());
    };

    Backend     bk(config.devices, Runtime::stream);
    lbm::D3Q19  lattice;
    const index_3d dim(config.N, config.N, config.N);
    const double ulid[3] = {1., 0., 0.};

    bk.syncAll();
    auto  start = Clock::now();
    Grid  grid(bk, dim, [](const index_3d&) { return true; });
    if constexpr (std::is_same_v<Grid, bGrid>) {
        if (config.stagedNeighbourhood) grid.setStagedNeighbourhood(true);
    }
    PopulationField pop0 = grid.template newField<StorageFP, 19>("Population", 19, StorageFP{});
    PopulationField pop1 = grid.template newField<StorageFP, 19>("Population", 19, StorageFP{});
    CellTypeField   flag = grid.template newField<CellType, 1>("Material", 1, CellType{});
    typename Grid::template Field<Moment, 1> rho;
    typename Grid::template Field<Moment, 3> u;
    if (!config.benchmark) {
        rho = grid.template newField<Moment, 1>("rho", 1, Moment(0));
        u = grid.template newField<Moment, 3>("u", 3, Moment(0));
    }
    const StencilSemantic sem = config.stencilSemantic == "standard" ? StencilSemantic::standard : StencilSemantic::streaming;
    const Occ             occ = config.occ == "standard" ? Occ::standard : Occ::none;
    const TransferMode    tm = config.transferMode == "put" ? TransferMode::put : TransferMode::get;
    lbm::LbmIterationD3Q19<PopulationField, ComputeFP> iteration(sem, occ, tm, pop0, pop1, flag, config.omega);
    bk.syncAll();
    report.gridInit.push_back(us(start));
    std::cout << "Metrics: \n    Grid Init: " << std::setprecision(4) << report.gridInit.back() << " microseconds\n";

    bk.syncAll();
    start = Clock::now();
    {
        flag.forEachActiveCell([&](const index_3d& idx, const int&, CellType& f) {
            f.classification = CellType::bulk;
            f.wallNghBitflag = 0;
            if (idx.x == 0 || idx.x == dim.x - 1 || idx.y == 0 || idx.y == dim.y - 1 || idx.z == 0 ||
                idx.z == dim.z - 1) {
                f.classification = CellType::bounceBack;
                if (idx.y == dim.y - 1) f.classification = CellType::movingWall;
            }
        });
        flag.updateDeviceData(Backend::mainStreamIdx);
        for (PopulationField* p : {&iteration.getInput(), &iteration.getOutput()}) {
            if constexpr (kHostFill<StorageFP>) {
                p->forEachActiveCell([&](const index_3d& idx, const int& k, StorageFP& val) {
                    val = StorageFP(lattice.t_vect[k]);
                    if (idx.x == 0 || idx.x == dim.x - 1 || idx.y == 0 || idx.y == dim.y - 1 || idx.z == 0 ||
                        idx.z == dim.z - 1) {
                        if (idx.y == dim.y - 1) {
                            const index_3d& c = lattice.c_vect[k];
                            val = StorageFP(-6. * lattice.t_vect[k] * config.ulb *
                                            (c.x * ulid[0] + c.y * ulid[1] + c.z * ulid[2]));
                        } else {
                            val = StorageFP(0);
                        }
                    }
                });
                p->updateDeviceData(Backend::mainStreamIdx);
            } else {
                bk.sync(Backend::mainStreamIdx);
                const uint32_t* fl = static_cast<const uint32_t*>(flag.devPtr());
                const auto      d = grid.desc(StorageOf<StorageFP>::value, ComputeOf<ComputeFP>::value);
                if constexpr (std::is_same_v<Grid, dGrid>)
                    check(nx_d3q19_fill_dense(&d, p->devPtr(), fl, config.ulb, ulid, bk.stream()), "fill");
                else
                    check(nx_d3q19_fill_block(&d, p->devPtr(), fl, config.ulb, ulid, bk.stream()), "fill");
            }
        }
        auto mask = lbm::LbmContainers<PopulationField, ComputeFP>::computeWallNghMask(flag, flag);
        mask.run(Backend::mainStreamIdx);
        bk.syncAll();
    }
    report.setup.push_back(us(start));
    std::cout << "Metrics: \n    Problem Setup: " << std::setprecision(4) << report.setup.back() << " microseconds\n";

    auto exportRhoAndU = [&](int iterationId) {
        if (iterationId % 100 != 0) return;
        auto c = lbm::LbmContainers<PopulationField, ComputeFP>::computeRhoAndU(iteration.getInput(), flag, rho, u);
        c.run(Backend::mainStreamIdx);
        u.updateHostData(Backend::mainStreamIdx);
        rho.updateHostData(Backend::mainStreamIdx);
        bk.syncAll();
        std::string id = std::to_string(iterationId);
        id = std::string(5 - std::min<size_t>(5, id.size()), '0') + id;
        std::vector<std::pair<double, double>> xPosVal, yPosVal;
        const double scale = 1.0 / ulid[0];
        u.forEachActiveCell([&](const index_3d& p, const int& card, Moment& val) {
            if (p.x == dim.x / 2 && p.z == dim.z / 2 && card == 0)
                yPosVal.push_back({double(p.y) / double(dim.y), val * scale});
            if (p.y == dim.y / 2 && p.z == dim.z / 2 && card == 1)
                xPosVal.push_back({double(p.x) / double(dim.x), val * scale});
        });
        std::sort(xPosVal.begin(), xPosVal.end());
        std::sort(yPosVal.begin(), yPosVal.end());
        auto write = [](const std::vector<std::pair<double, double>>& pv, const std::string& name) {
            std::ofstream f(name);
            for (auto& v : pv) f << v.first << " " << v.second << "\n";
        };
        if (config.vti) {
            u.ioToVtk("u_" + id, "u", false);
            rho.ioToVtk("rho_" + id, "rho", false);
        }
        write(yPosVal, "NeonUniformLBM_" + id + "_Y.dat");
        write(xPosVal, "NeonUniformLBM_" + id + "_X.dat");
    };

    bk.syncAll();
    start = Clock::now();
    int clockIter = 0;
    for (int t = 0; t < config.benchMaxIter; ++t) {
        if (!config.benchmark) exportRhoAndU(t);
        if (config.benchmark && t == config.benchIniIter) {
            std::cout << "Warm up completed (" << t << " iterations ).\nStarting benchmark step ("
                      << config.benchMaxIter - config.benchIniIter << " iterations)." << std::endl;
            bk.syncAll();
            start = Clock::now();
            clockIter = 0;
        }
        iteration.run();
        ++clockIter;
    }
    std::cout << "Iterations completed" << std::endl;
    bk.syncAll();
    const double elapsed = us(start);
    const double mlups = double(size_t(config.N) * config.N * config.N) * clockIter / elapsed;
    report.loopTime.push_back(elapsed);
    report.mlups.push_back(mlups);
    std::cout << "Metrics: \n     time: " << std::setprecision(4) << elapsed << " microseconds\n"
              << "    MLUPS: " << std::setprecision(4) << mlups << " MLUPS" << std::endl;
}

template <typename Grid, typename StorageFP>
void byCompute(const Config& c, Report& r)
{
    if (c.computeType == "double") return run<Grid, StorageFP, double>(c, r);
    if (c.computeType == "float") return run<Grid, StorageFP, float>(c, r);
    if (c.computeType == "float-fast") return run<Grid, StorageFP, FastFloat>(c, r);
    usage("--computeFP must be double, float or float-fast");
}

template <typename Grid>
void byStorage(const Config& c, Report& r)
{
    if (c.storeType == "double") return byCompute<Grid, double>(c, r);
    if (c.storeType == "float") return byCompute<Grid, float>(c, r);
    if (c.storeType == "float-dev") return run<Grid, F32Dev, FastFloat>(c, r);
    if (c.storeType == "half-dev") return run<Grid, F16Dev, FastFloat>(c, r);
    usage("--storageFP must be double, float, float-dev or half-dev");
}

std::string jsonList(const std::vector<double>& v)
{
    std::stringstream s;
    s << "[";
    for (size_t i = 0; i < v.size(); i++) s << (i ? ", " : "") << std::setprecision(10) << v[i];
    return s.str() + "]";
}

}  // namespace

int main(int argc, char** argv)
{
    const Config config = parse(argc, argv);
    std::cout << "--------------- Parameters ---------------\n" << toString(config)
              << "-------------------------------------------\n";
    Report report;
    try {
        for (int i = 0; i < config.repetitions; i++) {
            if (config.gridType == "dGrid")
                byStorage<dGrid>(config, report);
            else
                byStorage<bGrid>(config, report);
        }
    } catch (const NeonException& e) {
        std::cerr << "NeonException: " << e.what() << std::endl;
        return 1;
    }
    std::ofstream f(config.reportFile + ".json");
    f << "{\n  \"name\": \"lbm-lid-driven-cavity-flow\",\n"
      << "  \"Re\": " << config.Re << ",\n  \"ulb\": " << config.ulb << ",\n  \"N\": " << config.N << ",\n"
      << "  \"benchmark\": " << (config.benchmark ? "true" : "false") << ",\n"
      << "  \"benchIniIter\": " << config.benchIniIter << ",\n  \"benchMaxIter\": " << config.benchMaxIter << ",\n"
      << "  \"repetitions\": " << config.repetitions << ",\n"
      << "  \"deviceType\": \"" << config.deviceType << "\",\n  \"devices\": [" << config.devices[0] << "],\n"
      << "  \"gridType\": \"" << config.gridType << "\",\n"
      << "  \"stagedNeighbourhood\": " << (config.stagedNeighbourhood ? "true" : "false") << ",\n"
      << "  \"computeType\": \"" << config.computeType << "\",\n  \"storeType\": \"" << config.storeType << "\",\n"
      << "  \"occ\": \"" << config.occ << "\",\n  \"transferMode\": \"" << config.transferMode << "\",\n"
      << "  \"transferSemantic\": \"" << config.stencilSemantic << "\",\n"
      << "  \"nu\": " << std::setprecision(17) << config.nu << ",\n  \"omega\": " << config.omega << ",\n"
      << "  \"dx\": " << config.dx << ",\n  \"dt\": " << config.dt << ",\n"
      << "  \"MLUPS\": " << jsonList(report.mlups) << ",\n"
      << "  \"Time\": " << jsonList(report.loopTime) << ",\n"
      << "  \"time_unit\": \"microseconds\",\n"
      << "  \"NeonGridInitTime\": " << jsonList(report.gridInit) << ",\n"
      << "  \"ProblemSetupTime\": " << jsonList(report.setup) << "\n}\n";
    return 0;
}
