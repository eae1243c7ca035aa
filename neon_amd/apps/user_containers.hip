// user_containers.hip -- a Neon user's own containers on the dense grid,
// written against include/neon_amd.hpp the way they are written against
// Neon's API (grid.newContainer, Loader::load, partition getNghData;
// dGrid_imp.h:218-251, Loader_imp.h:102-188, dPartition.h:174-190) and
// compiled with hipcc:
//   1. the D3Q19 pull-stream + bounce-back + BGK step of the cavity
//      benchmark (LbmTools.h:78-282) as user code, run for n steps beside the
//      library's LbmIterationD3Q19 from the same start: the binary64
//      populations must agree bit for bit;
//   2. a 7-point average over the box with a STENCIL load (missing
//      neighbours replaced by the cell itself), checked against the same
//      arithmetic on the host.
// Built with -ffp-contract=off, as the library is, so that both evaluate
// the same binary64 expression trees.  Exit status 0 when both hold.
//
//   user_containers [N] [steps]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "neon_amd.hpp"

using namespace neon_amd;
using PopField = dGrid::Field<double, 19>;
using FlagField = dGrid::Field<CellType, 1>;

namespace {

constexpr int kQ = 19;
__device__ constexpr int kCx[kQ] = {-1, 0, 0, -1, -1, -1, -1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
__device__ constexpr int kCy[kQ] = {0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 0, 1, 0, 1, -1, 0, 0, 1, 1};
__device__ constexpr int kCz[kQ] = {0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 1, 0, 0, 1, -1, 1, -1};

// the step as a user's container: fin -> fout
Container userIteration(const dGrid& grid, const PopField& fIn, const FlagField& flag, double omega, PopField& fOut)
{
    return grid.newContainer("UserLbmIteration", [&, omega](Loader& L) {
        auto pin = L.load(fIn, Pattern::STENCIL);
        auto ct = L.load(flag);
        auto pout = L.load(fOut);
        return [=] __device__(const dIndex& i) mutable {
            const uint32_t w = ct(i, 0);
            if (cellClassification(w) != CellType::bulk) return;
            const uint32_t wall = cellWallNghBitflag(w);
            double f[kQ];
#pragma unroll
            for (int q = 0; q < kQ; q++) {
                if (q == 9) {
                    f[9] = pin(i, 9);
                    continue;
                }
                const int o = q < 9 ? q + 10 : q - 10;
                const int8_3d up(int8_t(-kCx[q]), int8_t(-kCy[q]), int8_t(-kCz[q]));
                if ((wall >> q) & 1u) {
                    const NghData<double> nb = pin.getNghData(i, up, o, 0.0);
                    f[q] = pin(i, o) + nb.mData;
                } else {
                    f[q] = pin.getNghData(i, up, q, 0.0).mData;
                }
            }
            const double xm = f[0] + f[3] + f[4] + f[5] + f[6];
            const double xp = f[10] + f[13] + f[14] + f[15] + f[16];
            const double x0 = f[9] + f[1] + f[2] + f[7] + f[8] + f[11] + f[12] + f[17] + f[18];
            const double ym = f[1] + f[3] + f[7] + f[8] + f[14];
            const double yp = f[4] + f[11] + f[13] + f[17] + f[18];
            const double zm = f[2] + f[5] + f[7] + f[16] + f[18];
            const double zp = f[6] + f[8] + f[12] + f[15] + f[17];
            const double rho = xm + xp + x0;
            const double u0 = (xp - xm) / rho, u1 = (yp - ym) / rho, u2 = (zp - zm) / rho;
            const double usqr = 1.5 * (u0 * u0 + u1 * u1 + u2 * u2);
            const double ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const double wk = k < 3 ? 1. / 18. : 1. / 36.;
                const double eq = rho * wk * (1. - 3. * ck[k] + 4.5 * ck[k] * ck[k] - usqr);
                const double eqo = eq + rho * wk * 6. * ck[k];
                pout(i, k) = (1. - omega) * f[k] + omega * eq;
                pout(i, k + 10) = (1. - omega) * f[k + 10] + omega * eqo;
            }
            pout(i, 9) = (1. - omega) * f[9] + omega * (rho * (1. / 3.) * (1. - usqr));
        };
    });
}

void cavitySetup(const index_3d& dim, double ulb, FlagField& flag, PopField* pops[4])
{
    lbm::D3Q19 lat;
    flag.forEachActiveCell([&](const index_3d& idx, const int&, CellType& f) {
        f.classification = CellType::bulk;
        f.wallNghBitflag = 0;
        if (idx.x == 0 || idx.x == dim.x - 1 || idx.y == 0 || idx.y == dim.y - 1 || idx.z == 0 || idx.z == dim.z - 1)
            f.classification = idx.y == dim.y - 1 ? CellType::movingWall : CellType::bounceBack;
    });
    flag.updateDeviceData(Backend::mainStreamIdx);
    for (int k = 0; k < 4; k++) {
        pops[k]->forEachActiveCell([&](const index_3d& idx, const int& q, double& v) {
            v = lat.t_vect[q];
            if (idx.x == 0 || idx.x == dim.x - 1 || idx.y == 0 || idx.y == dim.y - 1 || idx.z == 0 ||
                idx.z == dim.z - 1)
                v = idx.y == dim.y - 1 ? -6. * lat.t_vect[q] * ulb * lat.c_vect[q].x : 0.;
        });
        pops[k]->updateDeviceData(Backend::mainStreamIdx);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 40;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 30;
    if (N < 4 || steps < 0 || steps % 2) {
        std::fprintf(stderr, "usage: user_containers [N >= 4] [even steps]\n");
        return 2;
    }
    try {
        Backend bk({0}, Runtime::stream);
        const index_3d dim(N, N, N);
        dGrid grid(bk, dim);
        const double ulb = 0.05, omega = 1.0 / (3.0 * ulb * (N - 2) / 100.0 + 0.5);
        PopField a0 = grid.newField<double, 19>("libIn", 19, 0.0), a1 = grid.newField<double, 19>("libOut", 19, 0.0);
        PopField b0 = grid.newField<double, 19>("userIn", 19, 0.0), b1 = grid.newField<double, 19>("userOut", 19, 0.0);
        FlagField flag = grid.newField<CellType, 1>("Material", 1, CellType{});
        PopField* pops[4] = {&a0, &a1, &b0, &b1};
        cavitySetup(dim, ulb, flag, pops);
        lbm::LbmContainers<PopField, double>::computeWallNghMask(flag, flag).run(Backend::mainStreamIdx);
        bk.syncAll();

        // 1. library iteration vs the user's container, same start
        lbm::LbmIterationD3Q19<PopField, double> lib(StencilSemantic::standard, Occ::none, TransferMode::get, a0, a1,
                                                     flag, omega);
        Container user[2] = {userIteration(grid, b0, flag, omega, b1), userIteration(grid, b1, flag, omega, b0)};
        for (int t = 0; t < steps; t++) {
            lib.run();
            user[t & 1].run(Backend::mainStreamIdx);
        }
        bk.syncAll();
        lib.getInput().updateHostData(Backend::mainStreamIdx);
        b0.updateHostData(Backend::mainStreamIdx);
        bk.syncAll();
        size_t differ = 0, total = 0;
        double maxd = 0;
        const size_t cells = dim.rMulTyped();
        for (size_t c = 0; c < cells; c++)
            for (int q = 0; q < kQ; q++) {
                const double x = lib.getInput().hostRef(c, q), y = b0.hostRef(c, q);
                total++;
                if (std::memcmp(&x, &y, sizeof(double)) != 0) {
                    differ++;
                    maxd = std::max(maxd, std::abs(x - y));
                }
            }
        std::printf("user LBM container vs library, %d^3 cavity, %d steps: %zu of %zu populations differ (max %.3e)\n",
                    N, steps, differ, total, maxd);

        // 2. 7-point average with a STENCIL load
        auto s = grid.newField<double, 1>("s", 1, 0.0), out = grid.newField<double, 1>("avg", 1, 0.0);
        std::mt19937_64 rng(7);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        s.forEachActiveCell([&](const index_3d&, const int&, double& v) { v = U(rng); });
        s.updateDeviceData(Backend::mainStreamIdx);
        const auto& sc = s;
        Container avg = grid.newContainer("Average7", [&](Loader& L) {
            auto p = L.load(sc, Pattern::STENCIL);
            auto o = L.load(out);
            return [=] __device__(const dIndex& i) mutable {
                const double c = p(i, 0);
                double acc = c;
                acc += p.getNghData<-1, 0, 0>(i, 0, c).mData;
                acc += p.getNghData<1, 0, 0>(i, 0, c).mData;
                acc += p.getNghData<0, -1, 0>(i, 0, c).mData;
                acc += p.getNghData<0, 1, 0>(i, 0, c).mData;
                acc += p.getNghData<0, 0, -1>(i, 0, c).mData;
                acc += p.getNghData<0, 0, 1>(i, 0, c).mData;
                o(i, 0) = acc / 7.0;
            };
        });
        bool refused = false;
        try {
            grid.newContainer("BadStencil", [&](Loader& L) {
                auto p = L.load(out, Pattern::STENCIL);  // a field the container writes
                return [=] __device__(const dIndex& i) mutable { p(i, 0) = 0; };
            });
        } catch (const NeonException&) {
            refused = true;
        }
        avg.run(Backend::mainStreamIdx);
        out.updateHostData(Backend::mainStreamIdx);
        bk.syncAll();
        size_t bad = 0;
        auto at = [&](int x, int y, int z, double self) {
            if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) return self;
            return s.hostRef(size_t(x) + size_t(N) * (size_t(y) + size_t(N) * size_t(z)), 0);
        };
        for (int z = 0; z < N; z++)
            for (int y = 0; y < N; y++)
                for (int x = 0; x < N; x++) {
                    const size_t c = size_t(x) + size_t(N) * (size_t(y) + size_t(N) * size_t(z));
                    const double v = s.hostRef(c, 0);
                    double acc = v;
                    acc += at(x - 1, y, z, v);
                    acc += at(x + 1, y, z, v);
                    acc += at(x, y - 1, z, v);
                    acc += at(x, y + 1, z, v);
                    acc += at(x, y, z - 1, v);
                    acc += at(x, y, z + 1, v);
                    const double want = acc / 7.0, got = out.hostRef(c, 0);
                    if (std::memcmp(&want, &got, sizeof(double)) != 0) bad++;
                }
        std::printf("7-point average container: %zu of %zu cells differ from the host; STENCIL load of a written field %s\n",
                    bad, cells, refused ? "refused" : "accepted");
        const bool ok = differ == 0 && bad == 0 && refused;
        std::printf("%s\n", ok ? "user containers OK" : "user containers FAILED");
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "user_containers: %s\n", e.what());
        return 1;
    }
}
